"""Batched steady serving of bump-channel cases: many cases of uneven size in
one stream of launches (cffc_run_steady), each converging on its own, cold
(freestream) or warm (initial state predicted by the U-Net and resampled to
the case's mesh).

Cases: (h, M) as in the training set, meshes of 96k x 32k cells for
k = 1..4 (96 x 32 to 384 x 128, width/height 3) in blocks of 32 x 32 cells.
Reported: solves/s over the batch and the per-case latency from the start
of the batch to the step at which the case reached the tolerance (device
wall clock).
"""
import time

import numpy as np

from . import steady, unet
from . import mesh as M

BLOCK = 32


def serving_cases(n, seed):
    """(h, M, k) triples: h ~ U[0.04, 0.10], M ~ U[0.3, 0.6], k ~ U{1..4}."""
    hm = steady.sample_cases(n, seed)
    k = np.random.default_rng(seed + 1000).integers(1, 5, n)
    return [(float(h), float(m), int(kk)) for (h, m), kk in zip(hm, k)]


def case_mesh(h, mach, k):
    return steady.bump_case(h, mach, nbi=3 * k, nbj=k, ni=BLOCK, nj=BLOCK)


def warm_states(net, norm, cases):
    """Predicted initial state of every case: the network's 192 x 64 image,
    resampled to the case's mesh."""
    X = np.stack([unet.features(M.bump_channel(h), h, m) for h, m, _ in cases])
    Y = net.predict(X)
    out = []
    for q, (h, m, k) in enumerate(cases):
        img = norm.decode(Y[q], m)
        mesh, _ = case_mesh(h, m, k)
        out.append(unet.image_to_state(mesh, unet.resample_image(img, 32 * k, 96 * k)))
    return out


def serve(cases, initial=None, tol=1e-10, max_steps=300000, freeze_level=steady.STEADY_FREEZE_LEVEL):
    """Solve all cases in one batch.  Returns a dict of the batch figures."""
    batch = []
    for q, (h, m, k) in enumerate(cases):
        mesh, W0 = case_mesh(h, m, k)
        batch.append((mesh, W0 if initial is None else initial[q]))
    r = steady.solve_batch(batch, tol=tol, max_steps=max_steps, freeze_level=freeze_level)
    lat = r.part_seconds
    ok = np.isfinite(lat)
    return dict(rc=r.rc, n=len(cases), converged=int(ok.sum()), steps_done=r.steps_done,
                wall_seconds=r.seconds, solves_per_s=len(cases) / r.seconds if r.seconds > 0 else None,
                p50_s=float(np.percentile(lat[ok], 50)) if ok.any() else None,
                p99_s=float(np.percentile(lat[ok], 99)) if ok.any() else None,
                steps=r.part_steps.tolist(), latency_s=[float(x) for x in lat],
                cells=int(sum(mm.ncell for mm, _ in batch)))


def serve_cold_warm(net, norm, cases, **kw):
    cold = serve(cases, **kw)
    t0 = time.perf_counter()
    init = warm_states(net, norm, cases)
    t_inf = time.perf_counter() - t0
    warm = serve(cases, initial=init, **kw)
    warm["inference_seconds"] = t_inf
    return cold, warm
