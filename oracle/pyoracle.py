"""ctypes binding of oracle/libts_oracle.so -- TEST INFRASTRUCTURE ONLY (the checker)."""
import ctypes as C
import os
import subprocess

import numpy as np

from turbostructred_amd.capi import Config, StepInfo, BCConfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libts_oracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    P = C.POINTER
    D = P(C.c_double)
    lib.tso_create.argtypes = [P(Config), C.c_int]
    lib.tso_create.restype = C.c_void_p
    lib.tso_destroy.argtypes = [C.c_void_p]
    lib.tso_state_size.argtypes = [C.c_void_p]
    lib.tso_state_size.restype = C.c_size_t
    lib.tso_set_state.argtypes = [C.c_void_p, D, C.c_size_t]
    lib.tso_get_state.argtypes = [C.c_void_p, D, C.c_size_t]
    lib.tso_set_time.argtypes = [C.c_void_p, C.c_double, C.c_longlong]
    lib.tso_step.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, P(StepInfo), D]
    lib.tso_grid.argtypes = [C.c_void_p, C.c_int, D, D, C.c_int]
    lib.tso_last_residual.argtypes = [C.c_void_p, D, D, C.c_size_t]
    lib.tso_flux.argtypes = [C.c_int, C.c_double, C.c_double, D, D, D, D, D]
    lib.tso_star_state.argtypes = [C.c_double] * 7 + [D, D, D, D, P(C.c_int)]
    lib.tso_exact_sample.argtypes = [C.c_double] * 8 + [D]
    lib.tso_reconstruct.argtypes = [C.c_int, C.c_int, C.c_int, D, D, D, D, D]
    lib.tso_bc_image.argtypes = [P(BCConfig), C.c_double, D, D, D]
    lib.tso_velocity_gradients.argtypes = [C.c_void_p, D, C.c_size_t]
    lib.tso_lsq_gradient.argtypes = [C.c_int, D, C.c_double, D, D, C.c_int, D]
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class Oracle:
    def __init__(self, cfg, threads=1):
        self.lib = load()
        self.cfg = cfg
        self.h = self.lib.tso_create(C.byref(cfg), threads)
        if not self.h:
            raise ValueError("oracle rejected the configuration")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.tso_destroy(self.h)
            self.h = None

    def state_size(self):
        return self.lib.tso_state_size(self.h)

    def set_state(self, aos):
        a = _arr(aos).ravel()
        assert self.lib.tso_set_state(self.h, _dp(a), a.size) == 0

    def get_state(self):
        a = np.empty(self.state_size())
        assert self.lib.tso_get_state(self.h, _dp(a), a.size) == 0
        return a

    def set_time(self, t, it):
        self.lib.tso_set_time(self.h, t, it)

    def step(self, max_steps, t_max=1e300, next_snapshot=0.0, log=False):
        info = StepInfo()
        dts = np.zeros(max_steps)
        rc = self.lib.tso_step(self.h, max_steps, t_max, next_snapshot, C.byref(info), _dp(dts))
        if rc != 0:
            raise RuntimeError(f"oracle step failed rc={rc}")
        if log:
            return info, dts[: info.steps_done]
        return info

    def grid(self, axis, n):
        c = np.empty(n)
        h = np.empty(n)
        assert self.lib.tso_grid(self.h, axis, _dp(c), _dp(h), n) == 0
        return c, h

    def last_residual(self):
        n = self.state_size() // 5
        r = np.empty(n * 5)
        sr = np.empty(n)
        assert self.lib.tso_last_residual(self.h, _dp(r), _dp(sr), n) == 0
        return r, sr

    def velocity_gradients(self):
        n = self.state_size() // 5
        g = np.empty(n * 9)
        assert self.lib.tso_velocity_gradients(self.h, _dp(g), g.size) == 0
        return g.reshape(n, 3, 3)


def flux(riemann, gamma, UL, UR, n, eps=0.0):
    lib = load()
    F = np.empty(5)
    lam = np.empty(1)
    UL, UR, n = _arr(UL), _arr(UR), _arr(n)
    lib.tso_flux(riemann, gamma, eps, _dp(UL), _dp(UR), _dp(n), _dp(F), _dp(lam))
    return F, lam[0]


def star_state(gamma, rl, ul, pl, rr, ur, pr):
    lib = load()
    o = [np.empty(1) for _ in range(4)]
    it = C.c_int()
    vac = lib.tso_star_state(gamma, rl, ul, pl, rr, ur, pr, *[_dp(x) for x in o], C.byref(it))
    return dict(p=o[0][0], u=o[1][0], rl=o[2][0], rr=o[3][0], vacuum=bool(vac), iters=it.value)


def exact_sample(gamma, rl, ul, pl, rr, ur, pr, s):
    lib = load()
    out = np.empty(3)
    lib.tso_exact_sample(gamma, rl, ul, pl, rr, ur, pr, s, _dp(out))
    return out


def reconstruct(recon, limiter, ndims, h, center, neighbours):
    lib = load()
    h = _arr(list(h) + [1.0] * (3 - len(h)))
    c = _arr(center)
    nb = _arr(neighbours).ravel()
    g = np.empty(15)
    lim = np.empty(5)
    lib.tso_reconstruct(recon, limiter, ndims, _dp(h), _dp(c), _dp(nb), _dp(g), _dp(lim))
    return g.reshape(5, 3), lim


def bc_image(bc, gamma, U, n):
    lib = load()
    out = np.empty(5)
    U, n = _arr(U), _arr(n)
    lib.tso_bc_image(C.byref(bc), gamma, _dp(U), _dp(n), _dp(out))
    return out


def lsq_gradient(ndims, point, value, points, values):
    lib = load()
    point, pts, vals = _arr(point), _arr(points).reshape(-1, 3), _arr(values)
    g = np.empty(3)
    lib.tso_lsq_gradient(ndims, _dp(point), float(value), _dp(pts), _dp(vals), len(vals), _dp(g))
    return g
