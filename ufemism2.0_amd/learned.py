"""Learned Picard initial guess: host side of the mesh message-passing network
(libufe_hip.so, ufe_mpnn_*).  Forward, backward and Adam run on the device
in hand-written kernels; this module only moves arrays and assembles the
solver-generated training set.

Training data are velocity fields the engine itself computes: each member of
the channel ensemble generator (uneven resolution, A and beta_sq drawn
log-uniformly) is solved to the nonlinear tolerance, and its per-triangle
features (ufe_mpnn_features) are paired with the converged velocity as
(log10 speed in m/yr, unit flow direction x, y).
The members form one disjoint graph, so a training step is one full-batch
forward/backward pass.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

N_FEAT = 16
N_OUT = 16
N_TARGET = 3
MIN_SPEED = 1e-3  # m/yr, floor of the log-speed target


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MpnnError(RuntimeError):
    pass


class Mpnn:
    """One network on one device."""

    def __init__(self, device: int = 0, weights: np.ndarray | None = None):
        self.lib = L.lib()
        self.n_weights = self.lib.ufe_mpnn_num_weights()
        self._h = C.c_void_p()
        w = None if weights is None else _f32(weights)
        if w is not None and w.size != self.n_weights:
            raise MpnnError(f"expected {self.n_weights} weights, got {w.size}")
        self._check(self.lib.ufe_mpnn_create(C.byref(self._h), device,
                                             _fp(w) if w is not None else C.cast(None, C.POINTER(C.c_float))))

    def _check(self, rc):
        if rc < 0:
            raise MpnnError(f"ufe_mpnn error {rc}: {self.lib.ufe_last_error(None).decode()}")
        return rc

    def weights(self) -> np.ndarray:
        w = np.zeros(self.n_weights, np.float32)
        self._check(self.lib.ufe_mpnn_get_weights(self._h, _fp(w)))
        return w

    def forward(self, X, nb) -> np.ndarray:
        X, nb = _f32(X), L.i32(nb)
        Y = np.zeros((X.shape[0], N_OUT), np.float32)
        self._check(self.lib.ufe_mpnn_forward(self._h, X.shape[0], _fp(X), L.ptr(nb, L.ip), _fp(Y)))
        return Y

    def loss_grad(self, X, nb, T):
        X, nb, T = _f32(X), L.i32(nb), _f32(T)
        loss = C.c_float(0)
        g = np.zeros(self.n_weights, np.float32)
        self._check(self.lib.ufe_mpnn_loss_grad(self._h, X.shape[0], _fp(X), L.ptr(nb, L.ip), _fp(T),
                                                C.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, X, nb, T, steps: int, lr: float = 1e-3) -> np.ndarray:
        X, nb, T = _f32(X), L.i32(nb), _f32(T)
        hist = np.zeros(max(steps, 1), np.float32)
        self._check(self.lib.ufe_mpnn_train(self._h, X.shape[0], _fp(X), L.ptr(nb, L.ip), _fp(T), steps,
                                            C.c_float(lr), _fp(hist)))
        return hist[:steps]

    def initial_guess(self, solver):
        """Sets the solver's velocity to the prediction for its uploaded inputs."""
        self._check(self.lib.ufe_mpnn_initial_guess(solver._ctx, self._h))

    def close(self):
        if self._h:
            self.lib.ufe_mpnn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def flatten_params(p: dict) -> np.ndarray:
    """Weights in the library's layout: We, be, (W_l, b_l) for l = 0..2, Wd, bd."""
    order = ["We", "be"] + [k for l in range(3) for k in (f"W{l}", f"b{l}")] + ["Wd", "bd"]
    return np.concatenate([np.asarray(p[k], np.float32).ravel() for k in order])


def unflatten_grad(g: np.ndarray, like: dict) -> dict:
    order = ["We", "be"] + [k for l in range(3) for k in (f"W{l}", f"b{l}")] + ["Wd", "bd"]
    out, o = {}, 0
    for k in order:
        n = like[k].size
        out[k] = g[o:o + n].reshape(like[k].shape)
        o += n
    return out


def features(solver) -> np.ndarray:
    """Per-triangle features of the solver's uploaded inputs (caller order)."""
    nT = solver.md.mesh.nT
    X = np.zeros((nT, N_FEAT), np.float32)
    rc = solver.lib.ufe_mpnn_features(solver._ctx, _fp(X))
    if rc < 0:
        raise MpnnError(solver.lib.ufe_last_error(solver._ctx).decode())
    return X


def solver_samples(workloads, solver_params: dict, stress: str = "DIVA"):
    """(features, neighbour lists, targets, solve info) of each workload, the
    targets being the engine's converged velocities as (log10 speed, direction)."""
    from .velocity import VelocitySolver
    out = []
    for w in workloads:
        inp = w.inputs()
        s = VelocitySolver(w.descriptor(), 0, **solver_params)
        s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
        s.set_velocity(None, None)
        if stress == "DIVA":
            s.reset_diva_state()
        info = s.run(stress)
        u, v = s.get_velocity()
        X = features(s)
        s.close()
        out.append((X, np.asarray(w.mesh.TriC, np.int32), targets(u, v), info))
    return out


def targets(u, v):
    """(log10 speed, direction x, direction y) per triangle."""
    sp = np.hypot(u, v)
    d = np.maximum(sp, 1e-12)
    return np.stack([np.log10(np.maximum(sp, MIN_SPEED)), u / d, v / d], 1).astype(np.float32)


def union(samples):
    """Disjoint union of graphs: stacked features/targets, offset neighbours."""
    Xs, nbs, Ts, off = [], [], [], 0
    for X, nb, T, *_ in samples:
        Xs.append(X)
        nbs.append(np.where(nb >= 0, nb + off, -1))
        Ts.append(T)
        off += X.shape[0]
    return np.concatenate(Xs), np.concatenate(nbs).astype(np.int32), np.concatenate(Ts)
