"""Benchmark: cell-stage updates/s of the 2D Euler explicit path on one MI355X.

Workload (the configuration the metric is quoted on): 64 x 64 blocks of
64 x 64 cells (16,777,216 cells) on [0,1]^2, the 4-quadrant Riemann problem
with a seeded 1% i.i.d. density perturbation, Roe flux, least-squares
reconstruction, Venkatakrishnan limiter, RK4, CFL 0.5, global time step,
constant-extrapolation boundaries.  Both precisions are run from the same
initial state:
  * fp64:  the reference's arithmetic, bit for bit with the CPU oracle
    (reported as `value`, `dtype` f64, and `roofline`);
  * mixed: FP32 reconstruction/limiter/Roe flux, FP64 state/update/dt/norms
    (sub-record `mixed`, with its own roofline), and its L1/Linf error
    against FP64 after 500 steps (`mixed_vs_fp64`).
A "step" is one RK4 time step (4 stages, each an exchange/BC kernel plus the
fused residual/update kernel); cell-stage updates = cells x 4 x steps.  The
timed steps are replayed from a captured hipGraph with the state resident in
HBM.

roofline: 162 algorithmic bytes per cell-stage (FP64 state read/written once
per stage, averaged over the four RK4 stages) x cells per launch / the fused
stage kernel's average duration, measured with HIP event pairs on the
context stream in a separate run of the same workload.  traffic: HBM bytes
per launch of that kernel and tile shape from rocprofv3 FETCH_SIZE /
WRITE_SIZE runs (profiles/), or null for a tile that was not profiled.

Further legs: the quadtree workload (resolution changes), the steady bump
cases cold vs warm (single grid and FAS multigrid), batched serving of 256
cases, and cpu_baseline: the CPU oracle (a C restatement of the reference
algorithm, -O2, no FMA), one pthread per core over a contiguous block
partition like the reference's one-MPI-rank-per-core decomposition, timed
on bounded samples of the same workload on 1 core and on N cores.

python bench.py [--gpus 1] [--steps 500] [--warmup 20]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
BYTES_PER_CELL_STAGE = 162.0   # algorithmic FP64 state traffic per cell-stage


def make_case(nb, nc, seed=2026):
    from cffc_amd import mesh as M
    m = M.cartesian(nb, nb, nc, nc)
    W0 = M.riemann_4quadrant(m)
    W0 = M.density_perturbation(m, W0, amplitude=0.01, seed=seed)
    m.set_wo_from_initial(W0)
    return m, W0


def autotune(m, W0, precision):
    """The library's tile autotuner on this workload (a few steps per
    instantiated tile shape, state restored).  Returns (tile, ms per step)."""
    from cffc_amd import abi, euler2d
    s = abi.Scheme(flux_precision=abi.FLUX_FP32 if precision == "mixed" else abi.FLUX_FP64)
    g = euler2d.Euler2DQuadBlocks(m, s)
    g.set_W(W0)
    tile, ms = g.autotune(nsteps=3)
    g.close()
    return tile, ms


def gpu_run(m, W0, precision, steps, warmup, tile=None):
    from cffc_amd import abi, euler2d
    s = abi.Scheme(flux_precision=abi.FLUX_FP32 if precision == "mixed" else abi.FLUX_FP64)
    g = euler2d.Euler2DQuadBlocks(m, s)
    if tile:
        g.set_tile(*tile)
    g.set_W(W0)
    if warmup:
        g.run_steps(warmup, graph=True)
    g.synchronize()
    t0 = time.perf_counter()
    rc, n, norms, dts = g.run_steps(steps, graph=True)
    g.synchronize()
    t1 = time.perf_counter()
    W = g.get_W()
    return g, rc, t1 - t0, norms, W


def kernel_time(m, W0, precision, steps, tile=None):
    """Average fused stage-kernel duration (ms) from HIP event pairs."""
    from cffc_amd import abi, euler2d
    s = abi.Scheme(flux_precision=abi.FLUX_FP32 if precision == "mixed" else abi.FLUX_FP64)
    g = euler2d.Euler2DQuadBlocks(m, s)
    if tile:
        g.set_tile(*tile)
    g.set_W(W0)
    g.run_steps(2)
    g.kernel_timing(True)
    g.run_steps(steps)
    t = g.kernel_times()
    g.kernel_timing(False)
    g.close()
    return t


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(threads):
    """The CPU oracle on bounded samples of the workload: 1 thread on 8 x 8
    blocks of 64^2 cells (262,144 cells, 8 steps), and `threads` threads on
    32 x 32 blocks of 64^2 (4,194,304 cells, 32 steps)."""
    import oracle
    from cffc_amd import abi
    oracle.build()

    def rate(nb, nsteps, nth):
        m, W0 = make_case(nb, 64)
        o = oracle.Oracle(m, abi.Scheme())
        o.set_W(W0)
        t0 = time.perf_counter()
        rc, n, _, _ = o.run(nsteps, nthreads=nth)
        t1 = time.perf_counter()
        del o
        return m.ncell * 4 * n / (t1 - t0), t1 - t0, n, m.ncell

    r1, t1, n1, c1 = rate(8, 8, 1)
    rn, tn, nn, cn = rate(32, 32, threads)
    return dict(value=rn, unit="cell-stage updates/s", cores=threads, kind="port",
                cpu_model=cpu_model(), one_core=r1, n_core=rn, n_core_speedup=rn / r1,
                sample="CPU oracle, same IC family and scheme: %d threads on 32x32 blocks of 64^2 cells "
                       "(%d cells), %d RK4 steps, %.1f s wall; 1 thread on 8x8 blocks of 64^2 (%d cells), "
                       "%d steps, %.1f s" % (threads, cn, nn, tn, c1, n1, t1))


def error_leg(m, W0, steps, tile64, tile32):
    """L1 / Linf of the mixed path against FP64 after `steps` graph-replayed
    steps from the same state (independent of the timed run's length)."""
    Ws = {}
    for prec, tile in (("fp64", tile64), ("mixed", tile32)):
        g, rc, dt, norms, W = gpu_run(m, W0, prec, steps, 0, tile)
        g.close()
        Ws[prec] = W
    d = np.abs(Ws["mixed"] - Ws["fp64"])
    print("[mixed vs fp64] %d steps: L1 %s, Linf %s" % (steps, d.mean(axis=0), d.max(axis=0)),
          file=sys.stderr, flush=True)
    return {"steps": steps, "L1": d.mean(axis=0).tolist(), "Linf": d.max(axis=0).tolist(),
            "vars": ["rho", "u", "v", "p"]}


def serving_leg(ncases, seed=5):
    """Batched steady serving: `ncases` bump cases of uneven size (96x32 to
    384x128 cells in blocks of 32^2) solved as one batch by the FAS
    multigrid solver to the single-grid criterion 1e-10, cold (freestream)
    and warm (the U-Net's prediction, resampled to each mesh).  Every case
    is held to its own converted tolerance on its own cells; the batch ends
    with the last case, and each case's latency is the batch clock at the
    cycle where it converged."""
    from cffc_amd import serving, warmstart
    cases = serving.serving_cases(ncases, seed)
    log = lambda s: print("[serving] " + s, file=sys.stderr, flush=True)
    out = {"cases": ncases, "seed": seed, "solver": "FAS multigrid, 4 levels, V cycle, CFL 1.0, FP64",
           "meshes": "96x32 .. 384x128 cells, blocks of 32^2", "tol": 1e-10}
    out["cold"] = serving.serve_multigrid(cases, log=log)
    model = os.path.join(ROOT, "cffc_amd", "data", "unet_bump.npz")
    if os.path.exists(model):
        net, norm = warmstart.load_model(model)
        t0 = time.perf_counter()
        init = serving.warm_states(net, norm, cases)
        t_inf = time.perf_counter() - t0
        out["warm"] = serving.serve_multigrid(cases, initial=init, log=log)
        out["warm"]["inference_seconds"] = t_inf
        out["warm"]["wall_seconds_with_inference"] = out["warm"]["wall_seconds"] + t_inf
    for k in ("cold", "warm"):
        if k in out:
            for key in ("latency_s", "case_cycles", "tol_R", "case_final_R"):
                out[k].pop(key, None)
    return out


def steady_leg(ncases, seed=7):
    """Steady bump cases (192 x 64 cells) from the freestream and from the
    trained U-Net's prediction, each set solved as one batch to the L2
    density residual 1e-10 (FP64)."""
    from cffc_amd import steady, warmstart
    model = os.path.join(ROOT, "cffc_amd", "data", "unet_bump.npz")
    if not os.path.exists(model):
        return {"skipped": "no trained model at cffc_amd/data/unet_bump.npz"}
    net, norm = warmstart.load_model(model)
    cases = steady.sample_cases(ncases, seed)
    cold = steady.solve_batch([steady.bump_case(h, m) for h, m in cases], tol=1e-10, max_steps=400000)
    t0 = time.perf_counter()
    init = warmstart.predict_states(net, norm, cases)
    t_inf = time.perf_counter() - t0
    warm = steady.solve_batch([(steady.bump_case(h, m)[0], W) for (h, m), W in zip(cases, init)], tol=1e-10,
                              max_steps=400000)
    cs, ws = cold.part_steps, warm.part_steps
    return {"cases": ncases, "seed": seed, "tol": 1e-10, "mesh": "192x64 cells, 12x4 blocks of 16^2",
            "scheme": "5-stage optimal smoothing, CFL 1.5, local dt, freeze 1e-5, FP64",
            "cold_steps": cs.tolist(), "warm_steps": ws.tolist(),
            "cold_batch_s": cold.seconds, "warm_batch_s": warm.seconds + t_inf, "inference_s": t_inf,
            "step_ratio_total": float(cs.sum() / max(ws.sum(), 1)),
            "time_ratio_batch": float(cold.seconds / (warm.seconds + t_inf)),
            "median_step_ratio": float(np.median(cs / np.maximum(ws, 1))),
            "target_ratio": 3.0,
            "all_converged": bool(np.all(cs > 0) and np.all(ws > 0))}


def multigrid_leg(ncases, seed=7, levels=4, cfl=1.0, cold_batch_s=None, graph=True, fmg_cycles=0):
    """The same held-out steady bump cases, cold, solved together (one mesh
    of disjoint parts) by the FAS multigrid solver (FASMultigrid2D: V cycle,
    HLLE first-order coarse levels, 5-stage optimal smoothing).  Multigrid
    stops on the L2 norm of the unscaled density residual R (the
    reference's Absolute_Convergence_Tolerance); the single-grid runs stop
    on the L2 norm of (CFL*dt)*R at 1e-10, each case on its own.  So does
    multigrid here: every case is held to 1e-10 divided by its own median
    CFL*dt of the single-grid scheme at the initial state, on the norm over
    its own cells (cffc_mg_set_parts), and the batch ends with the last
    case; the freeze level is converted the same way.  A third run starts
    with the full-multigrid start-up at fmg_cycles cycles per coarse level
    (the reference's Euler2D examples set Full_Multigrid = 50,
    examples/Euler2D/missile/missile.in:86), stopping on the batch norm
    against the median over all cells."""
    import dataclasses
    from cffc_amd import euler2d, multigrid, steady
    from cffc_amd import mesh as M
    cases = steady.sample_cases(ncases, seed)
    parts = [steady.bump_case(h, mach) for h, mach in cases]
    m, starts = M.concat([p[0] for p in parts])
    W0 = np.concatenate([p[1] for p in parts])
    sg = steady.bump_scheme()
    f = euler2d.Euler2DQuadBlocks(m, sg)
    f.set_W(W0)
    f.Send_All_Messages_and_BCs()
    f.CFL(set_global=False)
    cdt = sg.cfl_number * f.get_dt()
    f.close()
    scale = float(np.median(cdt))
    offs = np.concatenate([[0], np.cumsum([p[0].ncell for p in parts])])
    scales = np.array([float(np.median(cdt[offs[q]:offs[q + 1]])) for q in range(ncases)])
    tol = 1e-10 / scale
    freeze = steady.STEADY_FREEZE_LEVEL / scale
    p = multigrid.MGParams(levels=levels, absolute_tolerance=tol)

    def solve(W, fmg=0):
        g = multigrid.FASMultigrid(m, steady.bump_scheme(cfl=cfl), dataclasses.replace(p, fmg_cycles=fmg))
        g.set_graph(graph)
        g.set_W(W)
        g.initialize()
        if not fmg:
            g.set_parts(list(starts) + [m.nblk], 1e-10 / scales, steady.STEADY_FREEZE_LEVEL / scales)
        t0 = time.perf_counter()
        rc, cycles, norms = g.execute(20000, freeze_level=freeze)
        dt = time.perf_counter() - t0
        if fmg:
            conv = bool(rc == 0 and norms[-1, 1] < tol)
            case_cycles = None
        else:
            l2, done = g.part_status()
            conv = bool(rc == 0 and (done >= 0).all())
            case_cycles = done
        g.close()
        return rc, cycles, norms, dt, conv, case_cycles

    rc, cycles, norms, dt, conv, cc = solve(W0)
    out = {"cases": ncases, "seed": seed, "levels": levels, "cycle": "V", "cfl": cfl, "graph": graph,
           "mesh": "%d x (192x64 cells, 12x4 blocks of 16^2), coarsened to 2^2-cell blocks" % ncases,
           "stopping": "per case", "tol_R_range": [float((1e-10 / scales).min()), float((1e-10 / scales).max())],
           "rc": rc, "cycles": cycles, "case_cycles_p50_max": [int(np.median(cc)), int(cc.max())],
           "seconds": dt, "ms_per_cycle": 1e3 * dt / max(cycles, 1), "converged": conv}
    if cold_batch_s:
        out["cold_single_grid_batch_s"] = cold_batch_s
        out["time_ratio_vs_cold_single_grid"] = cold_batch_s / dt
    if fmg_cycles > 0:
        rcf, cf, nf, dtf, convf, _ = solve(W0, fmg_cycles)
        out["fmg"] = {"fmg_cycles": fmg_cycles, "rc": rcf, "cycles": cf, "seconds": dtf, "stopping": "batch norm",
                      "scale_median_cfl_dt": scale, "tol_R": tol, "final_R": float(nf[-1, 1]), "converged": convf}
        if cold_batch_s:
            out["fmg"]["time_ratio_vs_cold_single_grid"] = cold_batch_s / dtf
    # the same from the U-Net's predicted states (warm multigrid)
    model = os.path.join(ROOT, "cffc_amd", "data", "unet_bump.npz")
    if os.path.exists(model):
        from cffc_amd import warmstart
        net, norm = warmstart.load_model(model)
        t0 = time.perf_counter()
        Ww = np.concatenate(warmstart.predict_states(net, norm, cases))
        t_inf = time.perf_counter() - t0
        rcw, cw, nw, dtw, convw, ccw = solve(Ww)
        out["warm"] = {"rc": rcw, "cycles": cw, "seconds": dtw + t_inf, "inference_s": t_inf,
                       "case_cycles_p50_max": [int(np.median(ccw)), int(ccw.max())], "converged": convw}
        if cold_batch_s:
            out["warm"]["time_ratio_vs_cold_single_grid"] = cold_batch_s / (dtw + t_inf)
    return out


def amr_case(seed=3, with_quadtree=False):
    """The uneven-block quadtree workload: [0,1]^2 in 128 x 128 root blocks
    of 16^2 cells, refined twice where blocks meet the circle r = 0.3 about
    (0.5, 0.5), 2:1 balanced across faces and corners (20,044 equal blocks
    at levels 0/1/2), interior nodes disturbed by 0.1 h; then every aligned
    2x2 or 4x4 group of same-level blocks merged into one block (2,956 blocks
    of 16^2, 32^2 and 64^2 cells, 5,128,192 cells); reflecting walls; the
    cylindrical explosion (density ratio 8, pressure ratio 10) across that
    circle.  with_quadtree: also the equal-block quadtree and the merge's
    cell map (merged cell k is quadtree cell cmap[k])."""
    from cffc_amd import mesh as M
    q = M.quadtree_box(nroot=128, n=16, max_level=2, refine=M.circle_refinement(0.3, band=0.0),
                       jitter=0.1, seed=seed)
    m, cmap = M.merge_quadtree(q)
    W0 = M.cylindrical_explosion(m, radius=0.3)
    if with_quadtree:
        return m, W0, q, cmap
    return m, W0


def cell_areas(m):
    """Interior cell areas from the nodes, in quadCentroid's operation order
    (Math.h:311-350), so that they equal the solver's areas bit for bit."""
    out = []
    for b in range(m.nblk):
        x, y = m.block_nodes(b)
        n = int(m.ni[b])
        X0x, X0y = x[2:n + 2, 2:n + 2], y[2:n + 2, 2:n + 2]
        X1x, X1y = x[2:n + 2, 3:n + 3], y[2:n + 2, 3:n + 3]
        X2x, X2y = x[3:n + 3, 3:n + 3], y[3:n + 3, 3:n + 3]
        X3x, X3y = x[3:n + 3, 2:n + 2], y[3:n + 3, 2:n + 2]
        a = (X3x * X0y - X0x * X3y) + (X0x * X1y - X1x * X0y)
        a = a + (X1x * X2y - X2x * X1y)
        a = a + (X2x * X3y - X3x * X2y)
        out.append((0.5 * a).ravel())
    return np.concatenate(out)


def amr_leg(steps, warmup=10, autotune_tiles=True, fixed_tile=(16, 8)):
    """Throughput on the quadtree mesh (resolution changes every stage:
    restriction, prolongation, flux correction, deferred tile updates), both
    precisions, graph-replayed; mass drift and mixed-vs-FP64 difference."""
    t0 = time.perf_counter()
    m, W0 = amr_case()
    tmesh = time.perf_counter() - t0
    A = cell_areas(m)
    sizes = {"%dx%d" % (k, k): int(v) for k, v in zip(*np.unique(m.ni, return_counts=True))}
    per_level = {str(l): {"%dx%d" % (k, k): int(v) for k, v in zip(*np.unique(m.ni[m.level == l], return_counts=True))}
                 for l in range(int(m.level.max()) + 1)}
    out = {"workload": "euler2d_cylindrical_explosion_uneven_quadtree_%d_blocks_16_32_64" % m.nblk,
           "blocks": int(m.nblk), "blocks_by_size": sizes, "blocks_by_level_and_size": per_level,
           "cells": int(m.ncell), "steps": steps, "launches_per_stage": "1 residual/update kernel for all blocks",
           "scheme": "RK4/Roe/LSQ/Venkatakrishnan/CFL0.5/global-dt", "mesh_seconds": tmesh}
    Ws = {}
    for prec in ("mixed", "fp64"):
        tile, tune = fixed_tile, {}
        if autotune_tiles:
            tile, tune = autotune(m, W0, prec)
        g, rc, dt, norms, W = gpu_run(m, W0, prec, steps, warmup, tile)
        U = g.get_U()
        g.close()
        kt = kernel_time(m, W0, prec, 20, tile)
        stage_ms = kt["stage_ms"] / max(kt["stage_launches"], 1)
        rate = m.ncell * 4 * steps / dt
        rho0 = np.sum(A * W0[:, 0])
        out[prec] = {"value": rate, "unit": "cell-stage updates/s", "ms_per_step": dt / steps * 1e3,
                     "tile": list(tile), "stage_kernel_ms": stage_ms, "rc": rc,
                     "mass_drift_rel": float(abs(np.sum(A * U[:, 0]) - rho0) / rho0),
                     "autotune_ms_per_step": {"%dx%d" % k: round(v, 4) for k, v in tune.items()}}
        Ws[prec] = W
        print("[amr %s] tile %dx%d: %.3f s for %d steps: %.3e cell-stage/s; stage kernel %.4f ms"
              % (prec, tile[0], tile[1], dt, steps, rate, stage_ms), file=sys.stderr, flush=True)
    d = np.abs(Ws["mixed"] - Ws["fp64"])
    out["mixed_vs_fp64"] = {"L1": d.mean(axis=0).tolist(), "Linf": d.max(axis=0).tolist(),
                            "vars": ["rho", "u", "v", "p"], "steps": warmup + steps}
    return out


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        try:
            return json.load(open(p))
        except Exception:
            return None
    return None


def unet_probe(nsteps, batch=16):
    """Training steps of the warm-start U-Net (BF16 MFMA forward, backward and
    Adam kernels) on random 192 x 64 inputs: steps/s and the convolutions'
    algorithmic FLOP rate (3 x forward: forward, data and weight gradients)
    against MI355X's dense BF16 MFMA peak."""
    from cffc_amd import unet
    net = unet.UNet(64, 192, max_batch=batch, seed=1)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((4 * batch, 64, 192, unet.NIN)).astype(np.float32)
    Y = rng.standard_normal((4 * batch, 64, 192, unet.NOUT)).astype(np.float32)
    net.set_dataset(X, Y)
    fwd = 0.0
    for l in range(net.nlayers):
        ca, cb, co, lv = net.layer_shape(l)
        fwd += 2.0 * 9 * (ca + cb) * co * (64 >> lv) * (192 >> lv)
    net.train(3, batch)   # warm-up
    t0 = time.perf_counter()
    net.train(nsteps, batch)
    dt = time.perf_counter() - t0
    net.close()
    flops = 3.0 * fwd * batch * nsteps / dt
    print(json.dumps({"unet_probe_steps": nsteps, "batch": batch, "seconds": dt, "steps_per_s": nsteps / dt,
                      "conv_tflops": flops / 1e12, "bf16_peak_tflops": 2500.0, "frac": flops / 2.5e15,
                      "forward_gflop_per_sample": fwd / 1e9}))


def copy_probe(nbytes):
    """A plain device copy of a known size (torch's copy kernel), run under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to calibrate those counters."""
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--blocks", type=int, default=64, help="blocks per direction")
    ap.add_argument("--cells", type=int, default=64, help="cells per block per direction")
    ap.add_argument("--timing-steps", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="both", choices=["both", "mixed", "fp64"])
    ap.add_argument("--no-autotune", action="store_true", help="keep the default 16x8 tile")
    ap.add_argument("--tile", default=None, metavar="TIxTJ", help="use this tile shape (no autotuning)")
    ap.add_argument("--error-steps", type=int, default=500,
                    help="steps of the mixed-vs-FP64 error runs (0: skip)")
    ap.add_argument("--c5-cases", type=int, default=256,
                    help="serving cases of uneven size solved by FAS multigrid, cold and warm (0: skip)")
    ap.add_argument("--steady-cases", type=int, default=64,
                    help="held-out steady bump cases solved cold and warm (0: skip)")
    ap.add_argument("--unet-probe", type=int, default=0, metavar="STEPS",
                    help="only time STEPS training steps of the warm-start U-Net (MFMA profiling)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain copy kernel of BYTES bytes (counter calibration)")
    ap.add_argument("--amr-steps", type=int, default=500,
                    help="steps of the uneven-block quadtree (resolution-change) workload (0: skip)")
    ap.add_argument("--only-amr", action="store_true", help="run only the quadtree workload and print it")
    ap.add_argument("--mg-cases", type=int, default=64, help="steady cases solved by FAS multigrid (0: skip)")
    ap.add_argument("--mg-eager", action="store_true",
                    help="multigrid cycles launched one by one instead of replayed as a graph (profiling)")
    ap.add_argument("--only-steady", action="store_true",
                    help="run only the steady legs (cold/warm single grid, multigrid) and print them")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the whole mesh on a single MI355X "
              "(multi-GPU block sharding is a design note only); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.unet_probe:
        unet_probe(args.unet_probe)
        return
    if args.only_amr:
        fixed = tuple(int(x) for x in args.tile.lower().split("x")) if args.tile else (16, 8)
        print(json.dumps(amr_leg(args.amr_steps or 200, autotune_tiles=not args.no_autotune and not args.tile,
                                 fixed_tile=fixed)))
        return
    if args.only_steady:
        o = {"steady_warm_vs_cold": steady_leg(args.steady_cases)} if args.steady_cases > 0 else {}
        if args.mg_cases > 0:
            o["steady_multigrid"] = multigrid_leg(args.mg_cases, cold_batch_s=o.get("steady_warm_vs_cold", {}).get(
                "cold_batch_s"), graph=not args.mg_eager)
        print(json.dumps(o))
        return
    m, W0 = make_case(args.blocks, args.cells)
    ncell = m.ncell
    res = {}
    precs = ["mixed", "fp64"] if args.precision == "both" else [args.precision]
    for prec in precs:
        tile, tune = (16, 8), {}   # the library's default tile
        if args.tile:
            tile = tuple(int(x) for x in args.tile.lower().split("x"))
        elif not args.no_autotune:
            tile, tune = autotune(m, W0, prec)
            print("[%s] autotuned tile %dx%d; ms/step %s" % (prec, tile[0], tile[1],
                  {"%dx%d" % k: round(v, 3) for k, v in tune.items()}), file=sys.stderr, flush=True)
        g, rc, dt, norms, W = gpu_run(m, W0, prec, args.steps, args.warmup, tile)
        g.close()
        kt = kernel_time(m, W0, prec, args.timing_steps, tile)
        stage_ms = kt["stage_ms"] / max(kt["stage_launches"], 1)
        exch_ms = kt["exchange_ms"] / max(kt["exchange_launches"], 1)
        res[prec] = dict(rc=rc, seconds=dt, rate=ncell * 4 * args.steps / dt, W=W, tile=tile,
                         stage_ms=stage_ms, exchange_ms=exch_ms, l2_last=float(norms[-1, 1]) if len(norms) else None)
        print("[%s] %.3f s for %d steps: %.3e cell-stage/s; stage kernel %.3f ms, exchange %.3f ms"
              % (prec, dt, args.steps, res[prec]["rate"], stage_ms, exch_ms), file=sys.stderr, flush=True)

    def roofline(prec):
        r = res[prec]
        achieved = BYTES_PER_CELL_STAGE * ncell / (r["stage_ms"] * 1e-3) / 1e9
        kname = "stage_kernel<%s,%d,%d,fused>" % ("float" if prec == "mixed" else "double", *r["tile"])
        t = load_traffic("stage_kernel_%s_riemann4q_%dx%d_blocks_of_%dsq_traffic.json"
                         % (prec, args.blocks, args.blocks, args.cells))
        # per-launch HBM bytes measured for this kernel and tile shape (None: tile not profiled)
        t = t.get("kernels", {}).get(kname) if t else None
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": t.get("bytes_per_launch") if t else None,
                "kernel": kname, "kernel_avg_ms": r["stage_ms"]}

    head_prec = "fp64" if "fp64" in res else "mixed"
    head = res[head_prec]
    out = {
        "metric": "cell-stage updates/s",
        "value": head["rate"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head_prec == "fp64" else "fp32-flux/fp64-state",
        "data": "synthetic: 4-quadrant Riemann IC + seeded 1% density perturbation",
        "config": {"workload": "euler2d_riemann4q_64x64_blocks_of_64x64", "cells": ncell,
                   "tile": list(head["tile"]),
                   "blocks": args.blocks * args.blocks, "block_cells": [args.cells, args.cells],
                   "scheme": "RK4/Roe/LSQ/Venkatakrishnan/CFL0.5/global-dt",
                   "parallelism": "single-gpu"},
        "roofline": roofline(head_prec),
    }
    if "mixed" in res and head_prec != "mixed":
        r = res["mixed"]
        out["mixed"] = {"value": r["rate"], "unit": "cell-stage updates/s", "dtype": "fp32-flux/fp64-state",
                        "ms_per_step": r["seconds"] / args.steps * 1e3, "tile": list(r["tile"]),
                        "roofline": roofline("mixed")}
    if "fp64" in res and "mixed" in res and args.error_steps > 0:
        out["mixed_vs_fp64"] = error_leg(m, W0, args.error_steps, res["fp64"]["tile"], res["mixed"]["tile"])
    if args.amr_steps > 0:
        try:
            out["amr"] = amr_leg(args.amr_steps)
        except Exception as e:  # reported, never required for the throughput line
            out["amr"] = {"error": str(e)}
    if args.steady_cases > 0:
        try:
            out["steady_warm_vs_cold"] = steady_leg(args.steady_cases)
        except Exception as e:  # reported, never required for the throughput line
            out["steady_warm_vs_cold"] = {"error": str(e)}
    if args.mg_cases > 0:
        try:
            out["steady_multigrid"] = multigrid_leg(args.mg_cases, cold_batch_s=out.get(
                "steady_warm_vs_cold", {}).get("cold_batch_s"), graph=not args.mg_eager)
        except Exception as e:  # reported, never required for the throughput line
            out["steady_multigrid"] = {"error": str(e)}
    if args.c5_cases > 0:
        try:
            out["serving"] = serving_leg(args.c5_cases)
        except Exception as e:  # reported, never required for the throughput line
            out["serving"] = {"error": str(e)}
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_threads)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
