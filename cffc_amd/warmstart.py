"""Training data, training and evaluation of the learned warm start for the
steady bump channel (192 x 64 cells in 12 x 4 blocks of 16^2, h ~ U[0.04,
0.10], M ~ U[0.3, 0.6]).

* Data: many cases are solved at once as disjoint parts of one multi-block
  mesh (one launch per stage covers every case), from the freestream to an
  L2 density residual <= tol over all parts together, which bounds every
  part's residual by tol.
* Training: cffc_amd.unet.UNet on the GPU.
* Evaluation: every held-out case is solved twice on its own, from the
  freestream and from the network's prediction, to the same tolerance; the
  steps and wall time to reach it are compared.
"""
import time

import numpy as np

from . import abi, euler2d, steady, unet
from . import mesh as M


def solve_batch(cases, tol=1e-10, max_steps=200000, freeze_level=1e-5, chunk=64, log=None):
    """Steady states (state arrays) of the (h, M) cases, solved chunk at a time."""
    states = []
    for s in range(0, len(cases), chunk):
        part = cases[s:s + chunk]
        meshes, W0s = zip(*[steady.bump_case(h, mach) for h, mach in part])
        m, starts = M.concat(list(meshes))
        W0 = np.concatenate(W0s)
        g = euler2d.Euler2DQuadBlocks(m, steady.bump_scheme())
        g.set_W(W0)
        t0 = time.perf_counter()
        r = steady.solve_steady(g, tol=tol, max_steps=max_steps,
                                freeze_level=freeze_level)
        W = g.get_W()
        g.close()
        if log:
            log("cases %d-%d: rc=%d steps=%d l2=%.3e %.1fs" % (s, s + len(part) - 1, r.rc, r.steps, r.l2,
                                                                  time.perf_counter() - t0))
        if r.rc != 0 or r.l2 > tol:
            raise RuntimeError("steady solve of cases %d.. did not converge (rc=%d, l2=%.3e)" % (s, r.rc, r.l2))
        n = meshes[0].ncell
        for q in range(len(part)):
            states.append(W[q * n:(q + 1) * n].copy())
    return states


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


def evaluate(net, norm, cases, tol=1e-10, max_steps=200000, freeze_level=1e-5, log=None):
    """Per held-out case: (steps, seconds) cold and warm; seconds of the warm
    solve include the network inference."""
    rows = []
    for q, (h, mach) in enumerate(cases):
        m, W0 = steady.bump_case(h, mach)
        g = euler2d.Euler2DQuadBlocks(m, steady.bump_scheme())
        g.set_W(W0)
        cold = steady.solve_steady(g, tol=tol, max_steps=max_steps, freeze_level=freeze_level)
        t0 = time.perf_counter()
        Xq = unet.features(m, h, mach)[None]
        Wimg = norm.decode(net.predict(Xq)[0], mach)
        Wp = unet.image_to_state(m, Wimg)
        g.set_W(Wp)
        t_inf = time.perf_counter() - t0
        warm = steady.solve_steady(g, tol=tol, max_steps=max_steps, freeze_level=freeze_level)
        g.close()
        rows.append(dict(h=h, mach=mach, cold_steps=cold.steps, cold_s=cold.seconds, cold_rc=cold.rc,
                         cold_l2=cold.l2, warm_steps=warm.steps, warm_s=warm.seconds + t_inf, warm_rc=warm.rc,
                         warm_l2=warm.l2, inference_s=t_inf))
        if log:
            r = rows[-1]
            log("held-out %d h=%.3f M=%.3f: cold %d steps %.2fs, warm %d steps %.2fs" %
                (q, h, mach, r["cold_steps"], r["cold_s"], r["warm_steps"], r["warm_s"]))
    return rows
