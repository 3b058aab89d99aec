"""Learned start vectors for trlan() on 2D Anderson lattices.

A 4-layer 3x3 convolutional net (BF16 MFMA kernels in the library,
include/trlnet.h) maps the on-site potential eps (scaled by 1/6) of an
L x L periodic lattice to a start vector.  It is trained on states the
solver itself produced: for each training lattice the batched solver
(trlan_batch) finds the 8 lowest eigenpairs, and the target is the density
amplitude sqrt(sum_k psi_k^2) of those 8 states, so a start vector close to
it overlaps every wanted eigenvector.  The loss is 1 - cos^2(output,
target).  At solve time the output goes in as the start vector evec(:, 0)
with info->guess = 1; a non-finite or zero output falls back to the random
start inside trlan() (the reference's rule, SRC/trlcore.c:1716-1726).
"""
import ctypes
import os

import numpy as np

from . import trlan as T
from . import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "trlnet_anderson2d.npy")
NEV = 8
SCALE = 1.0 / 6.0

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float


def _lib():
    L = T.lib()
    if not getattr(L, "_trlnet_bound", False):
        L.trlnet_param_count.restype = _I
        L.trlnet_create.argtypes = [ctypes.POINTER(_P), ctypes.c_ulonglong]
        L.trlnet_destroy.argtypes = [_P]
        L.trlnet_get_params.argtypes = [_P, _P, _I]
        L.trlnet_set_params.argtypes = [_P, _P, _I]
        L.trlnet_forward.argtypes = [_P, _I, _I, _I, _P, _P]
        L.trlnet_forward_many.argtypes = [_P, _I, _P, _P, _P, _P, _P]
        L.trlnet_forward_many_device.argtypes = [_P, _I, _P, _P, _P, _P, ctypes.POINTER(_P)]
        L.trlnet_forward_many_device_f64.argtypes = [_P, _I, _P, _P, _P, _P, ctypes.c_double,
                                                     ctypes.POINTER(_P)]
        L.trlnet_loss_grad.argtypes = [_P, _I, _I, _I, _P, _P, ctypes.POINTER(_F), _P]
        L.trlnet_train_step.argtypes = [_P, _I, _I, _I, _P, _P, _F, ctypes.POINTER(_F)]
        for f in ("trlnet_create", "trlnet_destroy", "trlnet_get_params", "trlnet_set_params",
                  "trlnet_forward", "trlnet_forward_many", "trlnet_forward_many_device",
                  "trlnet_forward_many_device_f64",
                  "trlnet_loss_grad", "trlnet_train_step"):
            getattr(L, f).restype = _I
        L._trlnet_bound = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Net:
    """The device net (trlnet)."""

    def __init__(self, seed=0, params=None):
        L = _lib()
        self.h = _P()
        rc = L.trlnet_create(ctypes.byref(self.h), seed)
        if rc != 0:
            raise RuntimeError(f"trlnet_create failed ({rc})")
        self.nparam = L.trlnet_param_count()
        if params is not None:
            self.set_params(params)

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")

    def params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        self._chk(_lib().trlnet_get_params(self.h, p.ctypes.data, self.nparam), "get_params")
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.nparam
        self._chk(_lib().trlnet_set_params(self.h, p.ctypes.data, self.nparam), "set_params")

    def forward(self, x):
        """x: (nimg, h, w) -> (nimg, h, w)."""
        x = _f32(x)
        n, h, w = x.shape
        out = np.zeros_like(x)
        self._chk(_lib().trlnet_forward(self.h, n, h, w, x.ctypes.data, out.ctypes.data), "forward")
        return out

    def forward_flat(self, dims, flat):
        """dims: (ngroups, 3) int array of (nimg, h, w); flat: the groups'
        images back to back (float32) -> outputs in the same layout, one
        device call for all (trlnet_forward_many)."""
        dims = np.ascontiguousarray(dims, dtype=np.int32)
        flat = _f32(flat)
        out = np.empty_like(flat)
        ni, hh, ww = (np.ascontiguousarray(dims[:, k]) for k in range(3))
        self._chk(_lib().trlnet_forward_many(self.h, len(dims), ni.ctypes.data, hh.ctypes.data,
                                             ww.ctypes.data, flat.ctypes.data, out.ctypes.data),
                  "forward_many")
        return out

    def forward_flat_device(self, dims, flat):
        """forward_flat with the outputs left on the device: returns the
        device address of the outputs (FP32, same layout), valid until the
        next call on this net (trlnet_forward_many_device)."""
        dims = np.ascontiguousarray(dims, dtype=np.int32)
        flat = _f32(flat)
        ni, hh, ww = (np.ascontiguousarray(dims[:, k]) for k in range(3))
        dptr = _P()
        self._chk(_lib().trlnet_forward_many_device(self.h, len(dims), ni.ctypes.data,
                                                    hh.ctypes.data, ww.ctypes.data,
                                                    flat.ctypes.data, ctypes.byref(dptr)),
                  "forward_many_device")
        return dptr.value

    def forward_images_device(self, dims, images, scale):
        """forward_flat_device on fl32(image * scale) of FP64 images (one
        array per image, in group order), converted by the library into its
        pinned staging on several threads (trlnet_forward_many_device_f64)."""
        dims = np.ascontiguousarray(dims, dtype=np.int32)
        imgs = [np.ascontiguousarray(a, dtype=np.float64) for a in images]
        ptrs = (_P * len(imgs))(*[a.ctypes.data for a in imgs])
        ni, hh, ww = (np.ascontiguousarray(dims[:, k]) for k in range(3))
        dptr = _P()
        self._chk(_lib().trlnet_forward_many_device_f64(self.h, len(dims), ni.ctypes.data,
                                                        hh.ctypes.data, ww.ctypes.data,
                                                        ctypes.cast(ptrs, _P), float(scale),
                                                        ctypes.byref(dptr)),
                  "forward_many_device_f64")
        return dptr.value

    def forward_many(self, groups):
        """groups: list of (nimg, h, w) arrays -> list of outputs."""
        xs = [_f32(g) for g in groups]
        out = self.forward_flat([g.shape for g in xs], np.concatenate([g.ravel() for g in xs]))
        res, o = [], 0
        for g in xs:
            res.append(out[o:o + g.size].reshape(g.shape))
            o += g.size
        return res

    def loss_grad(self, x, t):
        x, t = _f32(x), _f32(t)
        n, h, w = x.shape
        loss = _F()
        g = np.zeros(self.nparam, dtype=np.float32)
        self._chk(_lib().trlnet_loss_grad(self.h, n, h, w, x.ctypes.data, t.ctypes.data,
                                          ctypes.byref(loss), g.ctypes.data), "loss_grad")
        return loss.value, g

    def train_step(self, x, t, lr):
        x, t = _f32(x), _f32(t)
        n, h, w = x.shape
        loss = _F()
        self._chk(_lib().trlnet_train_step(self.h, n, h, w, x.ctypes.data, t.ctypes.data, lr,
                                           ctypes.byref(loss)), "train_step")
        return loss.value

    def close(self):
        if self.h:
            _lib().trlnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lattice(L, Wd, seed):
    """(CSR, potential (L, L)) of one periodic 2D Anderson model."""
    (rp, ci, v), d = W.anderson2d_csr(L, Wd, seed=seed)
    return (rp, ci, v), d.reshape(L, L)


def solver_targets(models, maxlan=32, tol=1e-8, mixed=True, seed=0, kind="density"):
    """Targets made from the NEV lowest states of each model [(csr,
    potential)], found by the batched solver from random starts:
    kind="density": sqrt(sum_k psi_k^2); kind="signed": sum_k psi_k with
    each psi_k's largest-magnitude entry made positive.  Returns (targets
    list, solve results)."""
    ops, probs = [], []
    for j, ((rp, ci, v), d) in enumerate(models):
        A = T.Csr(rp, ci, v, mixed=mixed)
        ops.append(A)
        n = len(rp) - 1
        probs.append((A, -1, NEV, maxlan, tol, 1, 100000, NEV, W.start_vector(n, seed=seed + j)))
    rc, res, _ = T.solve_batch(probs)
    for A in ops:
        A.close()
    if rc != 0:
        raise RuntimeError(f"trlan_batch failed ({rc})")
    tg = []
    for (csr, d), r in zip(models, res):
        if r["nec"] < NEV:
            raise RuntimeError("a training solve did not converge")
        X = r["evec"][:, :NEV]
        if kind == "signed":
            piv = np.abs(X).argmax(axis=0)
            sg = np.sign(X[piv, np.arange(NEV)])
            tg.append((X * sg).sum(axis=1).reshape(d.shape))
        else:
            tg.append(np.sqrt((X * X).sum(axis=1)).reshape(d.shape))
    return tg, res


def train(net, X, Y, epochs=20, batch=64, lr=2e-3, seed=0, log=None):
    """Adam over (X, Y) of equal-size images; returns the loss per epoch."""
    rng = np.random.default_rng(seed)
    n = len(X)
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot = 0.0
        for b in range(0, n, batch):
            idx = perm[b:b + batch]
            tot += net.train_step(X[idx], Y[idx], lr) * len(idx)
        hist.append(tot / n)
        if log:
            log(f"epoch {ep} loss {hist[-1]:.4f}")
    return hist


def load_net(path=None):
    """The trained net shipped with the package, or None."""
    path = path or WEIGHTS
    if not os.path.exists(path):
        return None
    return Net(params=np.load(path, allow_pickle=False))


def _flat_inputs(potentials):
    """(groups of lattice indices by shape, dims, flat scaled float32 input)."""
    by_shape = {}
    for i, d in enumerate(potentials):
        by_shape.setdefault(d.shape, []).append(i)
    groups = list(by_shape.values())
    total = sum(potentials[i].size for idx in groups for i in idx)
    flat = np.empty(total, dtype=np.float32)
    dims, o = [], 0
    for idx in groups:
        h, w = potentials[idx[0]].shape
        dims.append((len(idx), h, w))
        for i in idx:
            np.multiply(potentials[i].ravel(), SCALE, out=flat[o:o + h * w], casting="unsafe")
            o += h * w
    return groups, dims, flat


def start_vectors_device(net, potentials):
    """Device addresses of the start vectors (FP32, one per lattice, in the
    order of `potentials`) for trl_batch_job.dstart: the net's outputs stay
    on the device (valid until the net's next call)."""
    by_shape = {}
    for i, d in enumerate(potentials):
        by_shape.setdefault(d.shape, []).append(i)
    groups = list(by_shape.values())
    dims = [(len(idx),) + potentials[idx[0]].shape for idx in groups]
    base = net.forward_images_device(dims, [potentials[i] for idx in groups for i in idx], SCALE)
    out = [None] * len(potentials)
    o = 0
    for idx in groups:
        for i in idx:
            out[i] = base + 4 * o
            o += potentials[i].size
    return out


def start_vectors(net, potentials):
    """Start vectors (flattened, float32) for lattices with these
    potentials: one device call for all of them, images of equal size
    together."""
    out = [None] * len(potentials)
    groups, dims, flat = _flat_inputs(potentials)
    y = net.forward_flat(dims, flat)
    o = 0
    for idx in groups:
        for i in idx:
            n = potentials[i].size
            out[i] = y[o:o + n]  # float32; trlan() callers copy it into evec(:, nec)
            o += n
    return out
