"""Learned warm start for steady bump-channel solves (include/cffc_unet.h).

A 4-level U-Net maps per-cell features of a bump case -- cell centre, height
above the wall, the bump profile, its height h and the inflow Mach number --
to the converged steady state, expressed as a normalised deviation from the
freestream.  Training data are steady states produced by the GPU solver
itself (cffc_amd.steady); training, the backward pass and inference run on
the GPU in libcffc_hip.so's BF16 MFMA kernels.
"""
import ctypes

import numpy as np

from . import euler2d
from . import mesh as M

NIN, NOUT = 8, 4


def _lib():
    L = euler2d.load()
    if not getattr(L, "_unet_bound", False):
        vp, i32, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        pf = ctypes.POINTER(ctypes.c_float)
        pi32 = ctypes.POINTER(ctypes.c_int32)
        L.cffc_unet_create.argtypes = [i32, i32, i32, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.cffc_unet_destroy.argtypes = [vp]
        L.cffc_unet_destroy.restype = None
        L.cffc_unet_num_layers.argtypes = [vp]
        L.cffc_unet_layer_shape.argtypes = [vp, i32, pi32]
        L.cffc_unet_get_weights.argtypes = [vp, i32, pf, pf]
        L.cffc_unet_set_weights.argtypes = [vp, i32, pf, pf]
        L.cffc_unet_set_dataset.argtypes = [vp, i32, pf, pf]
        L.cffc_unet_train_step.argtypes = [vp, pi32, i32, f, pf]
        L.cffc_unet_loss_grad.argtypes = [vp, pi32, i32, pf, i32, pf, pf]
        L.cffc_unet_predict.argtypes = [vp, pf, i32, pf]
        L.cffc_unet_mfma_layout_probe.argtypes = [pf]
        L._unet_bound = True
    return L


def _pf(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _pi(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _check(rc, what):
    if rc < 0:
        raise euler2d.CffcError("%s failed (%d)" % (what, rc))
    return rc


def mfma_layout_probe():
    out = np.zeros(256, dtype=np.float32)
    _check(_lib().cffc_unet_mfma_layout_probe(_pf(out)), "cffc_unet_mfma_layout_probe")
    return out.reshape(64, 4)


class UNet:
    def __init__(self, H=64, W=192, max_batch=16, seed=0):
        L = _lib()
        h = ctypes.c_void_p()
        _check(L.cffc_unet_create(H, W, max_batch, seed, ctypes.byref(h)), "cffc_unet_create")
        self.h, self.H, self.W, self.max_batch = h, H, W, max_batch
        self.nlayers = L.cffc_unet_num_layers(h)

    def close(self):
        if getattr(self, "h", None):
            _lib().cffc_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layer_shape(self, l):
        s = np.zeros(4, dtype=np.int32)
        _check(_lib().cffc_unet_layer_shape(self.h, l, _pi(s)), "layer_shape")
        return tuple(int(x) for x in s)  # (CinA, CinB, Cout, level)

    def get_weights(self, l):
        CA, CB, Co, _ = self.layer_shape(l)
        w = np.zeros((Co, 9 * (CA + CB)), dtype=np.float32)
        b = np.zeros(Co, dtype=np.float32)
        _check(_lib().cffc_unet_get_weights(self.h, l, _pf(w), _pf(b)), "get_weights")
        return w, b

    def set_weights(self, l, w, b):
        w = np.ascontiguousarray(w, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        _check(_lib().cffc_unet_set_weights(self.h, l, _pf(w), _pf(b)), "set_weights")

    def set_dataset(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        assert X.shape[1:] == (self.H, self.W, NIN) and Y.shape[1:] == (self.H, self.W, NOUT)
        _check(_lib().cffc_unet_set_dataset(self.h, X.shape[0], _pf(X), _pf(Y)), "set_dataset")
        self.nsamples = X.shape[0]

    def train_step(self, idx, lr):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = np.zeros(1, dtype=np.float32)
        _check(_lib().cffc_unet_train_step(self.h, _pi(idx), len(idx), lr, _pf(loss)), "train_step")
        return float(loss[0])

    def loss_grad(self, idx, layer):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        CA, CB, Co, _ = self.layer_shape(layer)
        gw = np.zeros((Co, 9 * (CA + CB)), dtype=np.float32)
        gb = np.zeros(Co, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(_lib().cffc_unet_loss_grad(self.h, _pi(idx), len(idx), _pf(loss), layer, _pf(gw), _pf(gb)),
               "loss_grad")
        return float(loss[0]), gw, gb

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(X.shape[:3] + (NOUT,), dtype=np.float32)
        for s in range(0, X.shape[0], self.max_batch):
            xb = np.ascontiguousarray(X[s:s + self.max_batch])
            ob = np.zeros(xb.shape[:3] + (NOUT,), dtype=np.float32)
            _check(_lib().cffc_unet_predict(self.h, _pf(xb), xb.shape[0], _pf(ob)), "predict")
            out[s:s + xb.shape[0]] = ob
        return out

    def train(self, nsteps, batch, lr=2e-3, seed=0, log_every=0):
        """Adam with a cosine learning-rate decay over random mini-batches."""
        rng = np.random.default_rng(seed)
        losses = []
        for it in range(nsteps):
            idx = rng.choice(self.nsamples, size=min(batch, self.nsamples), replace=False)
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * it / nsteps))
            losses.append(self.train_step(idx, lr_t))
            if log_every and it % log_every == 0:
                print("  step %d loss %.4e" % (it, losses[-1]), flush=True)
        return np.array(losses)


# ---- bump-case features and targets -------------------------------------------
def image_index(mesh):
    """Map the (NJ, NI) image of the structured bump mesh to state-array rows."""
    nbi, nbj, ni, nj = mesh.info["nbi"], mesh.info["nbj"], mesh.info["ni"], mesh.info["nj"]
    J, I = np.meshgrid(np.arange(nbj * nj), np.arange(nbi * ni), indexing="ij")
    b = (J // nj) * nbi + (I // ni)
    return b * (ni * nj) + (J % nj) * ni + (I % ni)


def features(mesh, h, mach):
    """(NJ, NI, 8) input image of one bump case."""
    idx = image_index(mesh)
    c = mesh.cell_centres()[idx]
    x, y = c[..., 0], c[..., 1]
    R = mesh.info["R"]
    yb = np.where((x > 1.0) & (x < 2.0), np.sqrt(np.maximum(R * R - (x - 1.5) ** 2, 0.0)) - (R - h), 0.0)
    slope = np.where((x > 1.0) & (x < 2.0), -(x - 1.5) / np.sqrt(np.maximum(R * R - (x - 1.5) ** 2, 1e-12)), 0.0)
    X = np.zeros(idx.shape + (NIN,), dtype=np.float32)
    X[..., 0] = x / 3.0
    X[..., 1] = y
    X[..., 2] = (y - yb) / (1.0 - yb)
    X[..., 3] = yb / 0.1
    X[..., 4] = slope
    X[..., 5] = (h - 0.07) / 0.03
    X[..., 6] = (mach - 0.45) / 0.15
    X[..., 7] = 1.0
    return X


class Normaliser:
    """Targets are deviations from the freestream scaled per channel:
    ((rho-rho_inf)/rho_inf, (u-u_inf)/a_inf, v/a_inf, (p-p_inf)/p_inf)/s."""

    def __init__(self, scale=(0.1, 0.1, 0.1, 0.1)):
        self.scale = np.asarray(scale, dtype=np.float64)

    def encode(self, W_img, mach):
        Wf = np.asarray(M.freestream(mach))
        a = np.sqrt(1.4 * Wf[3] / Wf[0])
        d = np.stack([(W_img[..., 0] - Wf[0]) / Wf[0], (W_img[..., 1] - Wf[1]) / a, W_img[..., 2] / a,
                      (W_img[..., 3] - Wf[3]) / Wf[3]], axis=-1)
        return (d / self.scale).astype(np.float32)

    def decode(self, Y_img, mach):
        Wf = np.asarray(M.freestream(mach))
        a = np.sqrt(1.4 * Wf[3] / Wf[0])
        d = np.asarray(Y_img, dtype=np.float64) * self.scale
        return np.stack([Wf[0] * (1 + d[..., 0]), Wf[1] + a * d[..., 1], a * d[..., 2], Wf[3] * (1 + d[..., 3])],
                        axis=-1)


def image_to_state(mesh, img):
    """(NJ, NI, 4) image -> state array rows."""
    out = np.empty((mesh.ncell, 4))
    out[image_index(mesh).ravel()] = img.reshape(-1, 4)
    return out


def state_to_image(mesh, W):
    return np.asarray(W)[image_index(mesh)]


def resample_image(img, NJ, NI):
    """Bilinear resampling of a cell-centred (nj, ni, C) bump image to
    (NJ, NI): the bump meshes are uniform in x and in the normalised height,
    so cell (I, J) of any of them sits at the same fractional position."""
    nj, ni = img.shape[:2]
    jf = np.clip((np.arange(NJ) + 0.5) * nj / NJ - 0.5, 0, nj - 1)
    if_ = np.clip((np.arange(NI) + 0.5) * ni / NI - 0.5, 0, ni - 1)
    j0 = np.minimum(np.floor(jf).astype(int), nj - 2)
    i0 = np.minimum(np.floor(if_).astype(int), ni - 2)
    wj = (jf - j0)[:, None, None]
    wi = (if_ - i0)[None, :, None]
    a = img[j0][:, i0]
    b = img[j0][:, i0 + 1]
    c = img[j0 + 1][:, i0]
    d = img[j0 + 1][:, i0 + 1]
    return (1 - wj) * ((1 - wi) * a + wi * b) + wj * ((1 - wi) * c + wi * d)
