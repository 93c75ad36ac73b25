"""Write tests/golden/state_kat.txt: known-answer vectors for the reference's
state and CFD routines (FluxRoe_n, RoeAverage, Limiter_Venkatakrishnan,
HartenFixPos/Neg, Reflect, BC_Characteristic_Pressure, W(U), U(W), Fx,
Runge_Kutta, MultiStage_Optimally_Smoothing, FluxHLLE_n and
Bilinear_Interpolation).

Inputs are generated here from a fixed seed (plus hand-picked edge cases: the
limiter dead band at |uQuad-u0| = TOLER, the Harten switch, sonic and
supersonic faces, all four BC_Characteristic_Pressure branches).  Outputs
come from the reference's own src_2D/Euler2D/Euler2DState.cc and
src_2D/CFD/CFD.cc, compiled from source by `make -C oracle ref` and driven by
oracle/ref_kat_driver.cc.  Doubles are stored in hex-float form, bit exact.

Run in the build container (needs /root/reference):
    make -C oracle ref && python tests/golden/make_state_kat.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_kat")
OUT = os.path.join(HERE, "state_kat.txt")
TOLER = 1e-7


def rand_W(rng, n, supersonic=False):
    d = rng.uniform(0.05, 3.0, n)
    p = rng.uniform(0.02, 3.0, n)
    a = np.sqrt(1.4 * p / d)
    scale = 3.0 if supersonic else 1.2
    u = rng.uniform(-scale, scale, n) * a
    v = rng.uniform(-scale, scale, n) * a
    return np.stack([d, u, v, p], axis=1)


def unit(rng, n):
    th = rng.uniform(0, 2 * np.pi, n)
    return np.stack([np.cos(th), np.sin(th)], axis=1)


def fmt(tag, vals):
    return tag + " " + " ".join(float(x).hex() for x in vals)


def cases():
    rng = np.random.default_rng(20261202)
    lines = []
    # Roe flux: random pairs, near-equal pairs (Harten active), supersonic, axis normals
    n = 600
    Wl, Wr = rand_W(rng, n), rand_W(rng, n)
    nn = unit(rng, n)
    for k in range(n):
        lines.append(fmt("ROE_N", list(Wl[k]) + list(Wr[k]) + list(nn[k])))
    Wl = rand_W(rng, 200)
    Wr = Wl * (1 + rng.uniform(-1e-3, 1e-3, Wl.shape))
    nn = unit(rng, 200)
    for k in range(200):
        lines.append(fmt("ROE_N", list(Wl[k]) + list(Wr[k]) + list(nn[k])))
    Wl, Wr = rand_W(rng, 100, True), rand_W(rng, 100, True)
    for k in range(100):
        lines.append(fmt("ROE_N", list(Wl[k]) + list(Wr[k]) + [1.0, 0.0]))
        lines.append(fmt("ROE_N", list(Wl[k]) + list(Wr[k]) + [0.0, -1.0]))
    # sonic rarefaction: left subsonic, right supersonic in +x (Harten fix on wave 1/4)
    for k in range(50):
        d = rng.uniform(0.5, 2.0)
        p = rng.uniform(0.5, 2.0)
        a = np.sqrt(1.4 * p / d)
        lines.append(fmt("ROE_N", [d, 0.9 * a, 0.0, p, 0.8 * d, 1.1 * a, 0.0, 0.8 * p, 1.0, 0.0]))
        lines.append(fmt("ROE_N", [d, -1.1 * a, 0.1, p, 1.2 * d, -0.9 * a, -0.1, 1.3 * p, 1.0, 0.0]))
    # identical states: F_Roe(W,W) = Fn(W)
    W = rand_W(rng, 50)
    nn = unit(rng, 50)
    for k in range(50):
        lines.append(fmt("ROE_N", list(W[k]) + list(W[k]) + list(nn[k])))
    # Roe average
    Wl, Wr = rand_W(rng, 200), rand_W(rng, 200)
    for k in range(200):
        lines.append(fmt("ROE_AVG", list(Wl[k]) + list(Wr[k])))
    # Venkatakrishnan limiter: random, plus dead-band edges
    for k in range(600):
        u0 = rng.uniform(-2, 2)
        umin = u0 - abs(rng.normal(0, 1))
        umax = u0 + abs(rng.normal(0, 1))
        uq = u0 + rng.normal(0, 1, 4) * rng.choice([1e-9, 1e-7, 1e-3, 0.3, 1.0, 3.0], 4)
        lines.append(fmt("VENKAT", list(uq) + [u0, umin, umax]))
    for off in [TOLER, -TOLER, np.nextafter(TOLER, 1), np.nextafter(-TOLER, -1), 0.0, 2 * TOLER]:
        u0 = 0.0
        lines.append(fmt("VENKAT", [u0 + off, u0 - 0.5, u0 + 0.5, u0 + off * 3, u0, -1.0, 1.0]))
    # y = 1 and y = 2 exactly
    lines.append(fmt("VENKAT", [1.0, 0.0, 0.0, 0.0, 0.0, -1.0, 1.0]))
    lines.append(fmt("VENKAT", [0.5, 0.0, 0.0, 0.0, 0.0, -1.0, 1.0]))
    lines.append(fmt("VENKAT", [0.25, -0.25, 0.0, 0.0, 0.0, -1.0, 1.0]))
    # Harten scalar
    for k in range(300):
        ll = rng.normal(0, 1)
        lr = ll + rng.choice([1e-9, 1e-6, 0.1, 1.0]) * rng.normal(0, 1)
        la = rng.uniform(min(ll, lr) - 0.5, max(ll, lr) + 0.5)
        lines.append(fmt("HARTEN", [la, ll, lr]))
    # Reflect
    W = rand_W(rng, 200)
    nn = unit(rng, 200)
    for k in range(200):
        lines.append(fmt("REFLECT", list(W[k]) + list(nn[k])))
    # BC_Characteristic_Pressure: all four Mach branches
    for k in range(400):
        Wi = rand_W(rng, 1)[0]
        Wo = rand_W(rng, 1)[0]
        nvec = unit(rng, 1)[0]
        a = np.sqrt(1.4 * Wi[3] / Wi[0])
        mach = [1.5, 0.4, -0.4, -1.5][k % 4]
        un = mach * a
        ut = rng.normal(0, 0.3) * a
        Wi[1] = un * nvec[0] - ut * nvec[1]
        Wi[2] = un * nvec[1] + ut * nvec[0]
        lines.append(fmt("BCCHARP", list(Wi) + list(Wo) + list(nvec)))
    # conversions and x-flux
    W = rand_W(rng, 200)
    for k in range(200):
        lines.append(fmt("U_OF_W", list(W[k])))
        lines.append(fmt("FX", list(W[k])))
        U = [W[k][0], W[k][0] * W[k][1], W[k][0] * W[k][2], W[k][3] / 0.4 + 0.5 * W[k][0] * (W[k][1] ** 2 + W[k][2] ** 2)]
        lines.append(fmt("W_OF_U", U))
    for nst in (1, 2, 4, 5):
        for i in range(1, nst + 1):
            lines.append(fmt("RK", [i, nst]))
    for nst in (1, 2, 3, 4, 5):
        for lim in (0, 9):
            for i in range(1, nst + 1):
                lines.append(fmt("OPTSM", [i, nst, lim]))
    # HLLE flux: random pairs, supersonic both ways (upwind branches), near-equal
    Wl, Wr = rand_W(rng, 400), rand_W(rng, 400)
    nn = unit(rng, 400)
    for k in range(400):
        lines.append(fmt("HLLE_N", list(Wl[k]) + list(Wr[k]) + list(nn[k])))
    Wl, Wr = rand_W(rng, 100, True), rand_W(rng, 100, True)
    for k in range(100):
        lines.append(fmt("HLLE_N", list(Wl[k]) + list(Wr[k]) + [1.0, 0.0]))
        lines.append(fmt("HLLE_N", list(Wl[k]) + list(Wr[k]) + [0.0, -1.0]))
    W = rand_W(rng, 50)
    nn = unit(rng, 50)
    for k in range(50):
        lines.append(fmt("HLLE_N", list(W[k]) + list(W[k] * (1 + rng.uniform(-1e-6, 1e-6, 4))) + list(nn[k])))
    # Bilinear_Interpolation (scalar states): points inside, on the edges and
    # outside skewed quadrilaterals, plus parallelograms (a == 0 branch)
    for k in range(600):
        base = np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 1.0], [1.0, 0.0]])
        quad = base + rng.normal(0, 0.15, (4, 2)) * rng.choice([0.0, 1.0], p=[0.2, 0.8])
        if k % 5 == 0:
            quad[2] = quad[1] + quad[3] - quad[0]   # parallelogram
        zeta, eta = rng.uniform(-0.3, 1.3, 2) if k % 3 == 0 else rng.uniform(0, 1, 2)
        if k % 11 == 0:
            zeta = float(rng.choice([0.0, 1.0]))
        p0 = (quad[0] * (1 - zeta) * (1 - eta) + quad[3] * zeta * (1 - eta) + quad[2] * zeta * eta +
              quad[1] * (1 - zeta) * eta)
        U = rng.normal(0, 1, 4)
        lines.append(fmt("BILIN", list(U) + list(quad.ravel()) + list(p0)))
    # axisymmetric source S(W, X) (own stream: the vectors above are unchanged)
    rng2 = np.random.default_rng(2024)
    W = rand_W(rng2, 300)
    for k in range(300):
        x, y = rng2.uniform(-2.0, 2.0), rng2.uniform(0.01, 3.0) * (1.0 if k % 10 else -1.0)
        lines.append(fmt("AXISRC", list(W[k]) + [x, y]))
    return lines


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -C oracle ref")
    inp = "\n".join(cases()) + "\n"
    res = subprocess.run([DRIVER], input=inp, capture_output=True, text=True, check=True)
    with open(OUT, "w") as f:
        f.write("# reference known-answer vectors, generated by tests/golden/make_state_kat.py\n")
        f.write(res.stdout)
    print("wrote", OUT, res.stdout.count("\n"), "lines")


if __name__ == "__main__":
    main()
