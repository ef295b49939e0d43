"""Learned constitutive surrogate: training on the device and its use in the
step (no counterpart in the reference).

The network replaces the plastic return of projectPlastic
(src/three/constitutive.js:52-78) for rate-independent J2 materials without
hardening (model 4; model 10 with H = 0): its inputs are the trial elastic
log strain (6 components) and the temperature, its outputs the deviatoric
Kirchhoff stress of the returned state. Samples come from the engine's own
plain return map; training runs the library's FP32 forward/backward kernels
and Adam; inference in the step runs BF16 MFMAs with FP32 accumulation.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import lib as L

NPARAM = L.MLP_NPARAM


@dataclass
class TrainConfig:
    samples: int = 4_000_000
    batch: int = 4096
    epochs: int = 150
    seed: int = 7
    lr: float = 2e-3
    lr_final: float = 1e-5
    strain_sigma: float = 0.02
    T_lo: float = 300.0
    T_hi: float = 900.0
    dt: float = 1e-4
    eps_max: float = 0.0  # <= 0: 4 sigma
    near_yield_fraction: float = 0.7

    def to_c(self):
        c = L.MpmMlpTrainConfig()
        for k in ("samples", "batch", "epochs", "seed", "lr", "lr_final", "strain_sigma", "T_lo", "T_hi", "dt",
                  "eps_max", "near_yield_fraction"):
            setattr(c, k, getattr(self, k))
        return c


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def default_norm(material, cfg: TrainConfig):
    lib = L.load()
    n = L.MpmMlpNorm()
    m, c = material.to_c(), cfg.to_c()
    L.check(lib.mpm_mlp_default_norm(C.byref(m), C.byref(c), C.byref(n)))
    return n


def samples(material, cfg: TrainConfig, n, seed=None, norm=None):
    """n (inputs [n,7], targets [n,6]) pairs drawn on the device from the
    plain return map with cfg's strain and temperature distribution."""
    lib = L.load()
    norm = norm or default_norm(material, cfg)
    X, Y = np.zeros((n, 7), np.float32), np.zeros((n, 6), np.float32)
    m, c = material.to_c(), cfg.to_c()
    L.check(lib.mpm_mlp_samples(C.byref(m), C.byref(c), C.byref(norm), n, cfg.seed if seed is None else seed,
                                _fp(X), _fp(Y)))
    return X, Y


def train(material, cfg: TrainConfig, norm=None):
    """Train on the device; returns (params [NPARAM] FP32, norm, per-epoch
    mean squared error)."""
    lib = L.load()
    norm = norm or default_norm(material, cfg)
    P = np.zeros(NPARAM, np.float32)
    hist = np.zeros(cfg.epochs, np.float32)
    m, c = material.to_c(), cfg.to_c()
    L.check(lib.mpm_mlp_train(C.byref(m), C.byref(c), C.byref(norm), _fp(P), _fp(hist)))
    return P, norm, hist


def loss_grad(params, X, Y):
    """(loss, gradient [NPARAM]) of the mean squared error by the training
    kernels."""
    lib = L.load()
    P, X, Y = _f32(params, NPARAM), _f32(X, (-1, 7)), _f32(Y, (-1, 6))
    g = np.zeros(NPARAM, np.float32)
    loss = C.c_float()
    L.check(lib.mpm_mlp_loss_grad(_fp(P), len(X), _fp(X), _fp(Y), _fp(g), C.byref(loss)))
    return float(loss.value), g


def infer(params, X):
    """Outputs [n,6] of the BF16 MFMA inference used inside the step."""
    lib = L.load()
    P, X = _f32(params, NPARAM), _f32(X, (-1, 7))
    Y = np.zeros((len(X), 6), np.float32)
    L.check(lib.mpm_mlp_infer(_fp(P), len(X), _fp(X), _fp(Y)))
    return Y


def return_batch(material, params, norm, F, ep, T, dt):
    """The step's surrogate return on a batch of trial states: (F_out [n,9],
    ep_out [n], tau [n,6], branch [n]) with branch 0 network, 1 elastic,
    2 plain return."""
    lib = L.load()
    F = _f32(F, (-1, 9))
    n = len(F)
    ep = _f32(np.broadcast_to(np.asarray(ep, np.float32), (n,)))
    T = _f32(np.broadcast_to(np.asarray(T, np.float32), (n,)))
    P = _f32(params, NPARAM)
    Fo, epo, tau = np.zeros((n, 9), np.float32), np.zeros(n, np.float32), np.zeros((n, 6), np.float32)
    br = np.zeros(n, np.int32)
    m = material.to_c()
    L.check(lib.mpm_surrogate_return_batch(C.byref(m), _fp(P), C.byref(norm), n, float(dt), _fp(F), _fp(ep), _fp(T),
                                           _fp(Fo), _fp(epo), _fp(tau), br.ctypes.data_as(C.POINTER(C.c_int32))))
    return Fo, epo, tau, br
