"""What the learned warm start gains over the plain solver (GPU).

A U-Net is trained on the GPU, by the library's own BF16 MFMA forward and
backward kernels, on steady bump states that the GPU solver produced; the
held-out cases are then solved from the freestream and from the network's
prediction to the same tolerance.  The solver decides on the prediction: both
runs must reach the tolerance, and the warm runs must need fewer steps.
Reduced sizes (96 x 32 cells, tol 1e-8) keep the test to about a minute;
the full 192 x 64, 512-case, 1e-10 study is run by
`python -m cffc_amd.warmstart`.
"""
import os

import numpy as np
import pytest

from cffc_amd import steady, unet, warmstart
from cffc_amd import mesh as M

pytestmark = pytest.mark.gpu

SMALL = dict(nbi=6, nbj=2, ni=16, nj=16)   # 96 x 32 cells
TOL = 1e-8
MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cffc_amd", "data",
                     "unet_bump.npz")


def _solve(cases, initial=None):
    batch = []
    for q, (h, m) in enumerate(cases):
        mesh, W0 = steady.bump_case(h, m, **SMALL)
        batch.append((mesh, W0 if initial is None else initial[q]))
    r = steady.solve_batch(batch, tol=TOL, max_steps=100000)
    assert r.rc == 0
    return r


def test_trained_network_reduces_steps():
    train = steady.sample_cases(48, 11)
    held = steady.sample_cases(8, 12)
    r = _solve(train)
    assert np.all(r.part_steps > 0), "training solves must converge"
    m0, _ = steady.bump_case(0.07, 0.45, **SMALL)
    unit = unet.Normaliser((1.0, 1.0, 1.0, 1.0))
    dev = np.stack([unit.encode(unet.state_to_image(m0, W), mach) for (h, mach), W in zip(train, r.states)])
    norm = unet.Normaliser(np.sqrt(np.mean(dev.reshape(-1, 4).astype(np.float64) ** 2, axis=0)) + 1e-6)
    X = np.stack([unet.features(steady.bump_case(h, m, **SMALL)[0], h, m) for h, m in train])
    Y = (dev / norm.scale).astype(np.float32)
    net = unet.UNet(H=32, W=96, max_batch=16, seed=3)
    net.set_dataset(X, Y)
    losses = net.train(600, 16, lr=2e-3, seed=3)
    assert losses[-20:].mean() < 0.2 * losses[:5].mean()
    Xh = np.stack([unet.features(steady.bump_case(h, m, **SMALL)[0], h, m) for h, m in held])
    Yh = net.predict(Xh)
    init = [unet.image_to_state(m0, norm.decode(Yh[q], mach)) for q, (h, mach) in enumerate(held)]
    cold = _solve(held)
    warm = _solve(held, init)
    print("cold steps", cold.part_steps.tolist(), "warm steps", warm.part_steps.tolist())
    assert np.all(cold.part_steps > 0) and np.all(warm.part_steps > 0)
    assert warm.part_steps.sum() < cold.part_steps.sum()
    # the converged states agree: the network only changes where the solver
    # starts.  Not to round-off: the limiters are frozen at the step where
    # each run passes 1e-5, from different states, and the frozen values
    # enter the steady discretisation (measured: 2.5e-4 at most)
    for a, b in zip(cold.states, warm.states):
        assert np.max(np.abs(a - b)) < 2e-3


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained model file")
def test_committed_model_on_held_out_cases():
    """The trained 192 x 64 model on held-out seeds at the full tolerance."""
    net, norm = warmstart.load_model(MODEL)
    held = steady.sample_cases(8, 777)
    init = warmstart.predict_states(net, norm, held)
    cold = steady.solve_batch([steady.bump_case(h, m) for h, m in held], tol=1e-10, max_steps=400000)
    warm = steady.solve_batch([(steady.bump_case(h, m)[0], W) for (h, m), W in zip(held, init)], tol=1e-10,
                              max_steps=400000)
    print("cold steps", cold.part_steps.tolist(), "warm steps", warm.part_steps.tolist())
    assert cold.rc == 0 and warm.rc == 0
    assert np.all(cold.part_steps > 0) and np.all(warm.part_steps > 0)
    assert warm.part_steps.sum() < cold.part_steps.sum()
