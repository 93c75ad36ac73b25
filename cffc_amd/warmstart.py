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
import argparse
import json
import time

import numpy as np

from . import abi, euler2d, steady, unet
from . import mesh as M


def solve_cases(cases, tol=1e-10, max_steps=200000, freeze_level=steady.STEADY_FREEZE_LEVEL, chunk=512,
                initial=None, log=None, segment=10000):
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
        r = steady.solve_batch(batch, tol=tol, max_steps=max_steps, freeze_level=freeze_level, segment=segment,
                               log=log)
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


def solve_single(h, mach, W0=None, tol=1e-10, max_steps=300000, freeze_level=steady.STEADY_FREEZE_LEVEL):
    """One case on its own context: (steps, seconds of the solve)."""
    m, Wf = steady.bump_case(h, mach)
    g = euler2d.Euler2DQuadBlocks(m, steady.bump_scheme())
    try:
        g.set_W(Wf if W0 is None else W0)
        g.synchronize()
        r = steady.solve_steady(g, tol=tol, max_steps=max_steps, freeze_level=freeze_level, history=False)
    finally:
        g.close()
    return r.steps, r.seconds


def evaluate(net, norm, cases, tol=1e-10, max_steps=200000, freeze_level=steady.STEADY_FREEZE_LEVEL, log=None,
             singles=0):
    """Cold (freestream) and warm (predicted) solves of the held-out cases,
    each set as one batch; the first `singles` cases are also solved alone,
    cold and warm, for the time of one solve.  Returns a dict with the steps
    of every case, the wall time of each batch (the warm one including
    inference) and the single-case times."""
    cold_states, cold_steps, cold_s = solve_cases(cases, tol, max_steps, freeze_level, log=log)
    t0 = time.perf_counter()
    init = predict_states(net, norm, cases)
    t_inf = time.perf_counter() - t0
    # prediction error on the held-out cases, in the training loss's units
    m0 = M.bump_channel(cases[0][0])
    err = [np.mean((norm.encode(unet.state_to_image(m0, Wp), mach) -
                    norm.encode(unet.state_to_image(m0, Wc), mach)) ** 2)
           for (h, mach), Wp, Wc in zip(cases, init, cold_states)]
    _, warm_steps, warm_s = solve_cases(cases, tol, max_steps, freeze_level, initial=init, log=log)
    single = []
    for q in range(min(singles, len(cases))):
        h, mach = cases[q]
        cs, ct = solve_single(h, mach, None, tol, max_steps, freeze_level)
        t1 = time.perf_counter()
        Wp = predict_states(net, norm, [cases[q]])[0]
        ti = time.perf_counter() - t1
        ws, wt = solve_single(h, mach, Wp, tol, max_steps, freeze_level)
        single.append(dict(h=float(h), mach=float(mach), cold_steps=cs, cold_s=ct, warm_steps=ws, warm_s=wt + ti,
                           inference_s=ti))
        if log:
            log("single case %d: cold %d steps %.2f s, warm %d steps %.2f s (inference %.3f s)" %
                (q, cs, ct, ws, wt + ti, ti))
    return dict(cases=np.asarray(cases).tolist(), cold_steps=cold_steps.tolist(), warm_steps=warm_steps.tolist(),
                cold_seconds=float(sum(cold_s)), warm_seconds=float(sum(warm_s)) + t_inf, inference_seconds=t_inf,
                step_ratio_median=float(np.median(cold_steps / np.maximum(warm_steps, 1))),
                step_ratio_total=float(cold_steps.sum() / max(warm_steps.sum(), 1)), single=single,
                heldout_mse=float(np.mean(err)))


# ---- persistence -------------------------------------------------------------
def save_model(path, net, norm, meta=None):
    """Weights of every layer, the target scale and free-form metadata."""
    arrs = {"scale": norm.scale, "H": net.H, "W": net.W}
    for l in range(net.nlayers):
        w, b = net.get_weights(l)
        arrs["w%d" % l], arrs["b%d" % l] = w, b
    if meta is not None:
        arrs["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(path, **arrs)


def load_model(path, max_batch=16):
    z = np.load(path)
    net = unet.UNet(H=int(z["H"]), W=int(z["W"]), max_batch=max_batch)
    for l in range(net.nlayers):
        net.set_weights(l, z["w%d" % l], z["b%d" % l])
    return net, unet.Normaliser(z["scale"])


def save_dataset(path, cases, states):
    """Cases and their steady states as freestream deviations (float16
    images, unit scale): the network's targets before scaling."""
    m0 = M.bump_channel(cases[0][0])
    unit = unet.Normaliser((1.0, 1.0, 1.0, 1.0))
    Y = np.stack([unit.encode(unet.state_to_image(m0, W), mach) for (h, mach), W in zip(cases, states)])
    np.savez_compressed(path, cases=np.asarray(cases), dev=Y.astype(np.float16))


def load_dataset(path):
    z = np.load(path)
    return z["cases"], z["dev"].astype(np.float32)


def training_arrays(cases, dev, norm):
    X = np.stack([unet.features(M.bump_channel(h), h, mach) for h, mach in cases])
    Y = (dev / norm.scale.astype(np.float32)).astype(np.float32)
    return X, Y


def main(argv=None):
    """generate: solve seeded training cases; train: fit the U-Net; evaluate:
    cold vs warm solves of held-out cases (all on the GPU)."""
    ap = argparse.ArgumentParser(prog="python -m cffc_amd.warmstart")
    ap.add_argument("cmd", choices=["generate", "train", "evaluate"])
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-steps", type=int, default=300000)
    ap.add_argument("--data", default="bump_train.npz", help="dataset file (train: comma-separated list)")
    ap.add_argument("--model", default="unet_bump.npz")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--singles", type=int, default=0, help="evaluate: also solve this many cases alone")
    ap.add_argument("--out", default=None, help="JSON report")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    report = {"cmd": a.cmd}
    if a.cmd == "generate":
        cases = steady.sample_cases(a.n, a.seed)
        t0 = time.perf_counter()
        states, steps, secs = solve_cases(cases, a.tol, a.max_steps, log=log)
        save_dataset(a.data, cases, states)
        report.update(n=a.n, seed=a.seed, tol=a.tol, seconds=time.perf_counter() - t0, steps=steps.tolist(),
                      batch_seconds=secs)
        log("generated %d cases: steps median %d max %d, %.1f s" % (a.n, int(np.median(steps)), int(steps.max()),
                                                                    report["seconds"]))
    elif a.cmd == "train":
        parts = [load_dataset(f) for f in a.data.split(",")]
        cases = np.concatenate([c for c, _ in parts])
        dev = np.concatenate([d for _, d in parts])
        log("training set: %d cases from %s" % (len(cases), a.data))
        norm = unet.Normaliser(np.sqrt(np.mean(dev.reshape(-1, 4).astype(np.float64) ** 2, axis=0)) + 1e-6)
        X, Y = training_arrays(cases, dev, norm)
        net = unet.UNet(H=X.shape[1], W=X.shape[2], max_batch=a.batch, seed=a.seed)
        net.set_dataset(X, Y)
        t0 = time.perf_counter()
        losses = net.train(a.steps, a.batch, lr=a.lr, seed=a.seed, log_every=max(a.steps // 20, 1))
        dt = time.perf_counter() - t0
        save_model(a.model, net, norm, meta={"cases": len(cases), "steps": a.steps, "batch": a.batch, "lr": a.lr})
        report.update(train_seconds=dt, loss_first=float(losses[:10].mean()), loss_last=float(losses[-50:].mean()),
                      scale=norm.scale.tolist())
        log("trained %d steps in %.1f s: loss %.4e -> %.4e" % (a.steps, dt, report["loss_first"],
                                                               report["loss_last"]))
    else:
        net, norm = load_model(a.model)
        cases = steady.sample_cases(a.n, a.seed)
        r = evaluate(net, norm, cases, a.tol, a.max_steps, log=log, singles=a.singles)
        report.update(r)
        log("held-out %d: cold steps median %d, warm %d; step ratio median %.2f total %.2f; batch time cold %.1fs "
            "warm %.1fs; prediction MSE %.3e" % (a.n, int(np.median(r["cold_steps"])), int(np.median(r["warm_steps"])),
                                                 r["step_ratio_median"], r["step_ratio_total"], r["cold_seconds"],
                                                 r["warm_seconds"], r["heldout_mse"]))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(report, f)


if __name__ == "__main__":
    main()
