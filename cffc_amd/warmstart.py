"""Training data, training and evaluation of the learned warm start for the
steady bump channel (192 x 64 cells in 12 x 4 blocks of 16^2, h ~ U[0.04,
0.10], M ~ U[0.3, 0.6]).

* Data: the cases are solved together as disjoint parts of one multi-block
  mesh (one launch per stage covers every case), each from the freestream
  to its own L2 density residual <= tol (cffc_run_steady).
* Training: cffc_amd.unet.UNet on the GPU.
* Evaluation: the held-out cases are solved twice to the same tolerance, from
  the freestream and from the network's prediction, both as a batch; the
  steps every case needs and the wall time of each batch are compared.
"""
import time

import numpy as np

from . import abi, euler2d, steady, unet
from . import mesh as M


def solve_cases(cases, tol=1e-10, max_steps=200000, freeze_level=steady.STEADY_FREEZE_LEVEL, chunk=512,
                initial=None, log=None):
    """Steady states of the (h, M) cases (batches of up to chunk cases).
    initial: optional list of initial state arrays (default freestream).
    Returns (states, steps per case, seconds per batch)."""
    states, steps, secs = [], [], []
    for s in range(0, len(cases), chunk):
        part = cases[s:s + chunk]
        batch = []
        for q, (h, mach) in enumerate(part):
            m, W0 = steady.bump_case(h, mach)
            if initial is not None:
                W0 = initial[s + q]
            batch.append((m, W0))
        r = steady.solve_batch(batch, tol=tol, max_steps=max_steps, freeze_level=freeze_level)
        if log:
            log("cases %d-%d: rc=%d steps_done=%d max part steps=%d %.1fs" %
                (s, s + len(part) - 1, r.rc, r.steps_done, int(r.part_steps.max()), r.seconds))
        if r.rc != 0 or np.any(r.part_steps < 0):
            raise RuntimeError("steady solve of cases %d.. did not converge (rc=%d, %d parts left)" %
                               (s, r.rc, int(np.sum(r.part_steps < 0))))
        states += r.states
        steps += r.part_steps.tolist()
        secs.append(r.seconds)
    return states, np.array(steps), secs


def dataset(cases, states, norm):
    m0 = M.bump_channel(cases[0][0])
    X = np.stack([unet.features(M.bump_channel(h), h, mach) for h, mach in cases])
    Y = np.stack([norm.encode(unet.state_to_image(m0, W), mach) for (h, mach), W in zip(cases, states)])
    return X, Y


def fit_scale(cases, states):
    """Per-channel scale of the freestream deviations (their RMS)."""
    m0 = M.bump_channel(cases[0][0])
    d = np.concatenate([unet.Normaliser((1, 1, 1, 1)).encode(unet.state_to_image(m0, W), mach).reshape(-1, 4)
                        for (h, mach), W in zip(cases, states)])
    return np.sqrt(np.mean(d * d, axis=0)) + 1e-6


def predict_states(net, norm, cases):
    """Network prediction of the steady state of every (h, M) case."""
    m0 = M.bump_channel(cases[0][0])
    X = np.stack([unet.features(M.bump_channel(h), h, mach) for h, mach in cases])
    Y = net.predict(X)
    return [unet.image_to_state(m0, norm.decode(Y[q], mach)) for q, (h, mach) in enumerate(cases)]


def evaluate(net, norm, cases, tol=1e-10, max_steps=200000, freeze_level=steady.STEADY_FREEZE_LEVEL, log=None):
    """Cold (freestream) and warm (predicted) solves of the held-out cases,
    each set as one batch.  Returns a dict with the steps of every case and
    the wall time of each batch (the warm one including inference)."""
    _, cold_steps, cold_s = solve_cases(cases, tol, max_steps, freeze_level, log=log)
    t0 = time.perf_counter()
    init = predict_states(net, norm, cases)
    t_inf = time.perf_counter() - t0
    _, warm_steps, warm_s = solve_cases(cases, tol, max_steps, freeze_level, initial=init, log=log)
    return dict(cases=np.asarray(cases).tolist(), cold_steps=cold_steps.tolist(), warm_steps=warm_steps.tolist(),
                cold_seconds=float(sum(cold_s)), warm_seconds=float(sum(warm_s)) + t_inf, inference_seconds=t_inf,
                step_ratio_median=float(np.median(cold_steps / np.maximum(warm_steps, 1))),
                step_ratio_total=float(cold_steps.sum() / max(warm_steps.sum(), 1)))
