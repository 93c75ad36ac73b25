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


def serve(cases, initial=None, tol=1e-10, max_steps=400000, freeze_level=steady.STEADY_FREEZE_LEVEL, mixed=True,
          switch_level=3e-4, segment=20000, log=None, stall_window=0):
    """Solve all cases in one batch.  mixed: FP32 flux down to switch_level,
    then FP64 to tol (steady.solve_batch_fp32_then_fp64); else FP64 only.
    stall_window > 0: the FP64 phase also freezes a case's limiters after
    that many steps without its L2 going 1% below its best.  Returns a dict
    of the batch figures."""
    batch = []
    for q, (h, m, k) in enumerate(cases):
        mesh, W0 = case_mesh(h, m, k)
        batch.append((mesh, W0 if initial is None else initial[q]))
    t0 = time.perf_counter()
    if mixed:
        first, r = steady.solve_batch_fp32_then_fp64(batch, tol=tol, switch_level=switch_level, max_steps=max_steps,
                                                     freeze_level=freeze_level, segment=segment, log=log,
                                                     stall_window=stall_window)
        if r is None:
            return dict(rc=first.rc, n=len(cases), error="first phase failed")
        s1 = np.where(first.part_steps > 0, first.part_steps, first.steps_done)
        steps = s1 + np.where(r.part_steps > 0, r.part_steps, 0)
        lat = first.seconds + r.part_seconds
        phases = dict(fp32_steps=first.part_steps.tolist(), fp32_seconds=first.seconds,
                      fp64_steps=r.part_steps.tolist(), fp64_seconds=r.seconds)
    else:
        r = steady.solve_batch(batch, tol=tol, max_steps=max_steps, freeze_level=freeze_level, segment=segment,
                               log=log, stall_window=stall_window)
        steps, lat, phases = r.part_steps, r.part_seconds, {}
    wall = time.perf_counter() - t0
    ok = np.isfinite(lat) & (r.part_steps > 0)
    return dict(rc=r.rc, n=len(cases), converged=int(ok.sum()), wall_seconds=wall,
                solves_per_s=len(cases) / wall if wall > 0 else None,
                p50_s=float(np.percentile(lat[ok], 50)) if ok.any() else None,
                p99_s=float(np.percentile(lat[ok], 99)) if ok.any() else None,
                steps=[int(x) for x in steps], latency_s=[float(x) for x in lat],
                cells=int(sum(mm.ncell for mm, _ in batch)), precision="fp32+fp64" if mixed else "fp64",
                **phases)


def serve_multigrid(cases, initial=None, tol=1e-10, max_cycles=20000, freeze_level=steady.STEADY_FREEZE_LEVEL,
                    levels=4, cfl=1.0, segment=500, per_case=True, log=None):
    """Solve all cases in one batch with the FAS multigrid solver (one
    mesh of disjoint parts; FP64).  The single-grid criterion (L2 of
    (CFL*dt)*R <= tol per case) is converted to the unscaled residual R
    with the case's median of CFL*dt, and so is the freeze level.

    per_case: every case is held to its own converted tolerance on the L2
    norm of R over its own cells (cffc_mg_set_parts); the batch stops when
    the last case has met its tolerance, and a case's time to convergence is
    the batch's wall clock at the cycle where it first met it.  Otherwise
    the batch stops on one L2 norm over all cases against the strictest
    case's conversion (the largest median of CFL*dt), and every case
    finishes with the batch."""
    from . import euler2d, multigrid
    batch = []
    for q, (h, m, k) in enumerate(cases):
        mesh, W0 = case_mesh(h, m, k)
        batch.append((mesh, W0 if initial is None else initial[q]))
    mesh, starts = M.concat([b[0] for b in batch])
    W0 = np.concatenate([b[1] for b in batch])
    sg = steady.bump_scheme()
    f = euler2d.Euler2DQuadBlocks(mesh, sg)
    f.set_W(W0)
    f.Send_All_Messages_and_BCs()
    f.CFL(set_global=False)
    dt = f.get_dt()
    f.close()
    offs = np.concatenate([[0], np.cumsum([b[0].ncell for b in batch])])
    scales = np.array([float(np.median(sg.cfl_number * dt[offs[q]:offs[q + 1]])) for q in range(len(batch))])
    scale = float(scales.max())
    t0 = time.perf_counter()
    g = multigrid.FASMultigrid(mesh, steady.bump_scheme(cfl=cfl),
                               multigrid.MGParams(levels=levels, absolute_tolerance=tol / scale))
    g.set_W(W0)
    g.initialize()
    if per_case:
        g.set_parts(list(starts) + [mesh.nblk], tol / scales, freeze_level / scales)
    t_setup = time.perf_counter() - t0
    # Execute in segments, one progress line each; a diverged batch stops.
    # clock: (regular cycles done, seconds since t0) after each segment
    cycles, parts, rc, clock = 0, [], 0, [(0, t_setup)]
    while cycles < max_cycles:
        rc, n, nr = g.execute(min(segment, max_cycles - cycles), freeze_level=freeze_level / scale)
        cycles += n
        clock.append((cycles, time.perf_counter() - t0))
        parts.append(nr[1:] if parts else nr)
        last = nr[-1, 1] if len(nr) else float("nan")
        if per_case:
            l2, done = g.part_status()
            if log:
                log("  multigrid: %d cycles, %d/%d cases converged, R %.3e, %.1f s" %
                    (cycles, int((done >= 0).sum()), len(cases), last, time.perf_counter() - t0))
        elif log:
            log("  multigrid: %d cycles, R %.3e, %.1f s" % (cycles, last, time.perf_counter() - t0))
        if rc != 0 or n < segment or not np.isfinite(last) or last > 1e12:
            break
    norms = np.concatenate(parts) if parts else np.zeros((0, 3))
    wall = time.perf_counter() - t0
    out = dict(rc=rc, n=len(cases), wall_seconds=wall, solves_per_s=len(cases) / wall if wall > 0 else None,
               cycles=cycles, setup_seconds=t_setup, levels=levels, cfl=cfl,
               final_R=float(norms[-1, 1]) if len(norms) else None, cells=int(mesh.ncell), precision="fp64",
               solver="FAS multigrid, V cycle")
    if per_case:
        l2, done = g.part_status()
        g.close()
        conv = (done >= 0) & (rc == 0)
        cy, tt = np.array([c for c, _ in clock], float), np.array([t for _, t in clock])
        # wall clock at the cycle where each case first met its tolerance,
        # along the batch's cycle clock (linear within a segment)
        lat = np.where(conv, np.interp(np.maximum(done, 0), cy, tt), np.inf)
        out.update(converged=int(conv.sum()), stopping="per case", case_cycles=done.tolist(),
                   tol_R=(tol / scales).tolist(), case_final_R=l2.tolist(),
                   p50_s=float(np.percentile(lat, 50)) if conv.all() else None,
                   p99_s=float(np.percentile(lat, 99)) if conv.all() else None, latency_s=lat.tolist())
        if log:
            log("multigrid: %d cycles, %d/%d cases converged (cycles to convergence %d .. %d), %.2f s "
                "(set-up %.2f s)" % (cycles, int(conv.sum()), len(cases), done.min(), done.max(), wall, t_setup))
    else:
        g.close()
        ok = rc == 0 and len(norms) and norms[-1, 1] < tol / scale
        if log:
            log("multigrid: %d cycles, R %.3e (tol %.3e), %.2f s (set-up %.2f s)" %
                (cycles, norms[-1, 1] if len(norms) else -1, tol / scale, wall, t_setup))
        out.update(converged=len(cases) if ok else 0, stopping="batch norm", tol_R=tol / scale,
                   p50_s=wall if ok else None, p99_s=wall if ok else None, latency_s=[wall] * len(cases))
    return out


def serve_cold_warm(net, norm, cases, **kw):
    cold = serve(cases, **kw)
    t0 = time.perf_counter()
    init = warm_states(net, norm, cases)
    t_inf = time.perf_counter() - t0
    warm = serve(cases, initial=init, **kw)
    warm["inference_seconds"] = t_inf
    return cold, warm


def main(argv=None):
    import argparse
    import json
    from . import warmstart
    ap = argparse.ArgumentParser(prog="python -m cffc_amd.serving")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--model", default=None)
    ap.add_argument("--fp64", action="store_true", help="FP64 only (default: FP32 flux, FP64 finish)")
    ap.add_argument("--max-steps", type=int, default=400000)
    ap.add_argument("--out", default=None)
    ap.add_argument("--only", choices=["both", "cold", "warm"], default="both",
                    help="run only the cold or only the warm batch (warm needs --model)")
    ap.add_argument("--solver", choices=["single", "multigrid"], default="single",
                    help="single-grid batched steady runs, or FAS multigrid over the batch")
    ap.add_argument("--levels", type=int, default=4, help="multigrid levels")
    ap.add_argument("--pick", default=None, metavar="I,J,...",
                    help="serve only these cases of the n drawn (indices into the seed's list)")
    ap.add_argument("--stall-window", type=int, default=0,
                    help="single grid: freeze a case's limiters after this many FP64 steps without a 1%% lower L2")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    cases = serving_cases(a.n, a.seed)
    if a.pick:
        cases = [cases[int(i)] for i in a.pick.split(",")]
    log("%d cases, %d cells" % (len(cases), sum(96 * k * 32 * k for _, _, k in cases)))
    if a.solver == "multigrid":
        out = {}
        if a.only in ("both", "cold"):
            out["cold"] = serve_multigrid(cases, levels=a.levels, log=log)
        if a.only in ("both", "warm") and a.model:
            net, norm = warmstart.load_model(a.model)
            t0 = time.perf_counter()
            init = warm_states(net, norm, cases)
            t_inf = time.perf_counter() - t0
            out["warm"] = serve_multigrid(cases, initial=init, levels=a.levels, log=log)
            out["warm"]["inference_seconds"] = t_inf
        for tag, r in out.items():
            log("%s (multigrid): %d/%d converged, %.2f solves/s, wall %.1f s, %d cycles" %
                (tag, r["converged"], r["n"], r["solves_per_s"] or 0, r["wall_seconds"], r["cycles"]))
        if a.out:
            with open(a.out, "w") as fo:
                json.dump({"cases": cases, **out}, fo)
        return
    kw = dict(mixed=not a.fp64, max_steps=a.max_steps, log=log, stall_window=a.stall_window)
    cold = warm = None
    if a.only == "warm" or (a.only == "both" and a.model):
        if not a.model:
            raise SystemExit("--only warm needs --model")
        net, norm = warmstart.load_model(a.model)
        if a.only == "both":
            cold, warm = serve_cold_warm(net, norm, cases, **kw)
        else:
            t0 = time.perf_counter()
            init = warm_states(net, norm, cases)
            t_inf = time.perf_counter() - t0
            warm = serve(cases, initial=init, **kw)
            warm["inference_seconds"] = t_inf
    else:
        cold = serve(cases, **kw)
    for tag, r in (("cold", cold), ("warm", warm)):
        if r:
            log("%s: %d/%d converged, %.2f solves/s, p50 %.2f s, p99 %.2f s, wall %.1f s" %
                (tag, r.get("converged", 0), r["n"], r.get("solves_per_s") or 0, r.get("p50_s") or 0,
                 r.get("p99_s") or 0, r.get("wall_seconds", 0)))
    if a.out:
        with open(a.out, "w") as f:
            json.dump({"cases": cases, "cold": cold, "warm": warm}, f)


if __name__ == "__main__":
    main()
