"""ctypes mirror of include/ts_fvm.h and the loader of libts_fvm.so.

The library is the product: a HIP/gfx950 build of the explicit finite-volume step.  Loading
fails loudly when it has not been built; there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libts_fvm.so")

TS_OK = 0
TS_ERR_INVALID = -1
TS_ERR_HIP = -2
TS_ERR_UNPHYSICAL = -3
TS_ERR_UNSUPPORTED = -4
TS_ERR_NOMEM = -5
TS_ERR_BUSY = -6

# RPSolver (src/Methods/MethodConfiguration.h:12-16) plus the HLLC extension
RIEMANN_GODUNOV, RIEMANN_ROE, RIEMANN_NONE, RIEMANN_HLLC = 0, 1, 2, 3
# Reconstruction (MethodConfiguration.h:5-10)
RECON_PWC, RECON_LINEAR, RECON_ENO2, RECON_LINEAR_LIM = 0, 1, 2, 3
# LimiterType (MethodConfiguration.h:18-21)
LIMITER_BJ, LIMITER_VENKAT = 0, 1
# BoundaryConditionType (BoundaryConditionConfiguration.h:12-19)
BC_WALL, BC_SYMMETRY, BC_MOVING_WALL, BC_SUBSONIC_INLET, BC_SUBSONIC_OUTLET, BC_NATURAL = range(6)
FP64, FP32_RESIDUAL = 0, 1
MAX_BC = 8

KIND_DT, KIND_STAGE1, KIND_STAGE2, KIND_STAGE3, KIND_ALL, KIND_NORMS, KIND_VISCOUS = 0, 1, 2, 3, 4, 5, 6


class BCConfig(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("velocity", C.c_double * 3),
        ("pstatic", C.c_double),
        ("density", C.c_double),
    ]


class Config(C.Structure):
    _fields_ = [
        ("ndims", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("ghost", C.c_int),
        ("lx", C.c_double),
        ("ly", C.c_double),
        ("lz", C.c_double),
        ("periodic", C.c_int * 3),
        ("coord", C.POINTER(C.c_double) * 3),
        ("h", C.POINTER(C.c_double) * 3),
        ("gamma", C.c_double),
        ("cfl", C.c_double),
        ("eps", C.c_double),
        ("riemann", C.c_int),
        ("recon", C.c_int),
        ("limiter", C.c_int),
        ("rk_stages", C.c_int),
        ("precision", C.c_int),
        ("n_bc", C.c_int),
        ("bc", BCConfig * MAX_BC),
        ("side_marker", C.c_int * 6),
        ("face_marker", C.POINTER(C.c_uint8) * 6),
        ("device", C.c_int),
        ("viscous", C.c_int),
        ("viscosity", C.c_double),
        ("external_force", C.c_int),
        ("sigma", C.c_double * 3),
        ("external_accel", C.c_int),
        ("accel", C.c_double * 3),
    ]


class StepInfo(C.Structure):
    _fields_ = [
        ("time", C.c_double),
        ("dt", C.c_double),
        ("iteration", C.c_longlong),
        ("residual_norm", C.c_double * 5),
        ("steps_done", C.c_int),
    ]


class BatchConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("hbm_budget_bytes", C.c_size_t),
        ("max_active", C.c_int),
        ("use_graphs", C.c_int),
        ("autotune_tiles", C.c_int),
    ]


class TileTrial(C.Structure):
    _fields_ = [
        ("ndims", C.c_int), ("riemann", C.c_int), ("recon_variant", C.c_int), ("precision", C.c_int),
        ("load_bucket", C.c_int), ("tile_param", C.c_int), ("passes", C.c_int),
        ("ns_per_cell_step", C.c_double), ("chosen", C.c_int),
    ]


def make_config(ndims=1, nx=1, ny=1, nz=1, lx=1.0, ly=1.0, lz=1.0, periodic=(0, 0, 0),
                gamma=1.4, cfl=0.4, eps=0.0, riemann=RIEMANN_ROE, recon=RECON_PWC,
                limiter=LIMITER_BJ, rk_stages=1, precision=FP64, bcs=None, side_marker=None,
                face_marker=None, device=0, coord=None, h=None, viscous=False, viscosity=0.0,
                sigma=None, accel=None):
    """Build a Config.  bcs: list of dicts {type, velocity, pstatic, density} (marker table);
    default is one Natural condition on every side.  face_marker: optional list of 6 uint8
    numpy arrays (or None) -- kept alive on the returned object."""
    c = Config()
    c.ndims, c.nx, c.ny, c.nz, c.ghost = ndims, nx, ny if ndims > 1 else 1, nz if ndims > 2 else 1, 1
    c.lx, c.ly, c.lz = lx, ly, lz
    for a in range(3):
        c.periodic[a] = int(periodic[a])
    c.gamma, c.cfl, c.eps = gamma, cfl, eps
    c.riemann, c.recon, c.limiter, c.rk_stages, c.precision = riemann, recon, limiter, rk_stages, precision
    if bcs is None:
        bcs = [dict(type=BC_NATURAL)]
    c.n_bc = len(bcs)
    for m, b in enumerate(bcs):
        c.bc[m].type = b.get("type", BC_NATURAL)
        v = b.get("velocity", (0.0, 0.0, 0.0))
        for a in range(3):
            c.bc[m].velocity[a] = v[a]
        c.bc[m].pstatic = b.get("pstatic", 0.0)
        c.bc[m].density = b.get("density", 0.0)
    sm = side_marker if side_marker is not None else (0,) * 6
    for s in range(6):
        c.side_marker[s] = sm[s]
    keep = []
    if face_marker is not None:
        for s in range(6):
            if face_marker[s] is not None:
                arr = np.ascontiguousarray(face_marker[s], dtype=np.uint8)
                keep.append(arr)
                c.face_marker[s] = arr.ctypes.data_as(C.POINTER(C.c_uint8))
    for name, arrs in (("coord", coord), ("h", h)):
        if arrs is not None:
            for a in range(3):
                if arrs[a] is not None:
                    arr = np.ascontiguousarray(arrs[a], dtype=np.float64)
                    keep.append(arr)
                    getattr(c, name)[a] = arr.ctypes.data_as(C.POINTER(C.c_double))
    c.device = device
    c.viscous, c.viscosity = int(viscous), float(viscosity)
    c.external_force = int(sigma is not None)
    c.external_accel = int(accel is not None)
    for a in range(3):
        c.sigma[a] = sigma[a] if sigma is not None else 0.0
        c.accel[a] = accel[a] if accel is not None else 0.0
    c._keep = keep
    return c


def state_shape(cfg):
    """(nZAll, nYAll, nXAll, 5) of the reference AoS state array."""
    g = cfg.ghost
    nxa = cfg.nx + 2 * g
    nya = cfg.ny + 2 * g if cfg.ndims > 1 else 1
    nza = cfg.nz + 2 * g if cfg.ndims > 2 else 1
    return (nza, nya, nxa, 5)


def interior(cfg, aos):
    """View of the interior cells of an AoS state array, shape (nz, ny, nx, 5)."""
    a = np.asarray(aos).reshape(state_shape(cfg))
    g = cfg.ghost
    sx = slice(g, g + cfg.nx)
    sy = slice(g, g + cfg.ny) if cfg.ndims > 1 else slice(0, 1)
    sz = slice(g, g + cfg.nz) if cfg.ndims > 2 else slice(0, 1)
    return a[sz, sy, sx, :]


_lib = None


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def load(path=LIB_PATH):
    """Load libts_fvm.so.  Raises if it has not been built: the HIP path is the product and
    there is no CPU fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP library first (python __graft_entry__.py build "
            "or make -C turbostructred_amd/csrc)")
    lib = C.CDLL(path)
    P = C.POINTER
    lib.ts_create.argtypes = [P(Config), P(C.c_void_p)]
    lib.ts_destroy.argtypes = [C.c_void_p]
    lib.ts_last_error.argtypes = [C.c_void_p]
    lib.ts_last_error.restype = C.c_char_p
    lib.ts_state_size.argtypes = [C.c_void_p]
    lib.ts_state_size.restype = C.c_size_t
    lib.ts_set_state.argtypes = [C.c_void_p, P(C.c_double), C.c_size_t]
    lib.ts_get_state.argtypes = [C.c_void_p, P(C.c_double), C.c_size_t]
    lib.ts_set_time.argtypes = [C.c_void_p, C.c_double, C.c_longlong]
    lib.ts_step.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, P(StepInfo)]
    lib.ts_enable_residual_norms.argtypes = [C.c_void_p, C.c_int]
    lib.ts_get_cells.argtypes = [C.c_void_p, P(C.c_int), C.c_size_t, P(C.c_double)]
    lib.ts_enable_dt_log.argtypes = [C.c_void_p, C.c_longlong]
    lib.ts_get_dt_log.argtypes = [C.c_void_p, P(C.c_double), C.c_longlong, P(C.c_longlong)]
    lib.ts_enable_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    lib.ts_get_kernel_timing.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_longlong)]
    lib.ts_memory_footprint.argtypes = [C.c_void_p, P(C.c_size_t)]
    lib.ts_algorithmic_bytes_per_cell.argtypes = [C.c_void_p, C.c_int, P(C.c_double)]
    lib.ts_batch_create.argtypes = [P(BatchConfig), P(C.c_void_p)]
    lib.ts_batch_destroy.argtypes = [C.c_void_p]
    lib.ts_batch_last_error.argtypes = [C.c_void_p]
    lib.ts_batch_last_error.restype = C.c_char_p
    lib.ts_batch_submit.argtypes = [C.c_void_p, P(Config), P(C.c_double), C.c_int, C.c_double,
                                    P(C.c_longlong)]
    lib.ts_batch_run.argtypes = [C.c_void_p, C.c_int, P(C.c_int)]
    lib.ts_batch_poll.argtypes = [C.c_void_p, C.c_longlong, P(C.c_int), P(C.c_double),
                                  P(StepInfo)]
    lib.ts_batch_start_server.argtypes = [C.c_void_p]
    lib.ts_batch_stop_server.argtypes = [C.c_void_p]
    lib.ts_autotune.argtypes = [P(Config), P(C.c_int), P(C.c_double)]
    lib.ts_step_noalloc.argtypes = [P(Config), C.c_void_p, C.c_double, C.c_longlong, C.c_int, C.c_double,
                                    C.c_double, P(StepInfo), C.c_void_p]
    lib.ts_autotune_state.argtypes = [P(Config), P(C.c_double), C.c_size_t, C.c_int, P(C.c_int), P(C.c_double),
                                      C.c_int, P(C.c_int), P(C.c_int)]
    lib.ts_batch_tuning.argtypes = [C.c_void_p, P(TileTrial), C.c_int, P(C.c_int)]
    lib.ts_eos_eval.argtypes = [C.c_double, C.c_int, P(C.c_double), P(C.c_double), C.c_size_t, P(C.c_double), C.c_int]
    _lib = lib
    return lib


EXPORTED = [
    "ts_create", "ts_destroy", "ts_last_error", "ts_state_size", "ts_set_state", "ts_get_state",
    "ts_set_time", "ts_step", "ts_get_cells", "ts_enable_residual_norms", "ts_enable_dt_log", "ts_get_dt_log",
    "ts_enable_kernel_timing", "ts_get_kernel_timing", "ts_memory_footprint",
    "ts_algorithmic_bytes_per_cell", "ts_batch_create", "ts_batch_destroy",
    "ts_batch_last_error", "ts_batch_submit", "ts_batch_run", "ts_batch_poll",
    "ts_batch_start_server", "ts_batch_stop_server", "ts_autotune", "ts_step_noalloc",
    "ts_eos_eval", "ts_autotune_state", "ts_batch_tuning",
]


class TSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ts error {code}: {msg}")
        self.code = code


class Solver:
    """Thin owner of one ts_ctx (one HIP stream, device-resident state)."""

    def __init__(self, cfg):
        self.lib = load()
        self.cfg = cfg
        h = C.c_void_p()
        rc = self.lib.ts_create(C.byref(cfg), C.byref(h))
        if rc != TS_OK:
            self.h = None
            msg = self.lib.ts_last_error(None)  # this thread's failed ts_create
            raise TSError(rc, msg.decode() if msg else "ts_create failed")
        self.h = h

    def _check(self, rc):
        if rc != TS_OK:
            msg = self.lib.ts_last_error(self.h)
            raise TSError(rc, msg.decode() if msg else "")

    def close(self):
        if self.h:
            self.lib.ts_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def state_size(self):
        return self.lib.ts_state_size(self.h)

    def set_state(self, aos):
        a = np.ascontiguousarray(aos, dtype=np.float64).ravel()
        self._check(self.lib.ts_set_state(self.h, _dp(a), a.size))

    def get_state(self):
        a = np.empty(self.state_size(), dtype=np.float64)
        self._check(self.lib.ts_get_state(self.h, _dp(a), a.size))
        return a

    def set_time(self, t, it):
        self._check(self.lib.ts_set_time(self.h, t, it))

    def step(self, max_steps, t_max=1e300, next_snapshot=0.0):
        info = StepInfo()
        self._check(self.lib.ts_step(self.h, max_steps, t_max, next_snapshot, C.byref(info)))
        return info

    def get_cells(self, ijk):
        """Conserved values (n, 5) of interior cells given as (n, 3) reference indices (i, j, k)."""
        idx = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        out = np.empty((len(idx), 5))
        self._check(self.lib.ts_get_cells(self.h, idx.ctypes.data_as(C.POINTER(C.c_int)), len(idx), _dp(out)))
        return out

    def enable_residual_norms(self, enable=True):
        self._check(self.lib.ts_enable_residual_norms(self.h, int(enable)))

    def enable_dt_log(self, cap):
        self._check(self.lib.ts_enable_dt_log(self.h, cap))

    def dt_log(self, cap):
        out = np.empty(cap, dtype=np.float64)
        n = C.c_longlong()
        self._check(self.lib.ts_get_dt_log(self.h, _dp(out), cap, C.byref(n)))
        return out[: n.value]

    def enable_kernel_timing(self, on=True):
        self._check(self.lib.ts_enable_kernel_timing(self.h, int(on)))

    def kernel_timing(self, kind):
        ms = C.c_double()
        n = C.c_longlong()
        self._check(self.lib.ts_get_kernel_timing(self.h, kind, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def algorithmic_bytes_per_cell(self, kind):
        b = C.c_double()
        self._check(self.lib.ts_algorithmic_bytes_per_cell(self.h, kind, C.byref(b)))
        return b.value

    def footprint(self):
        b = C.c_size_t()
        self._check(self.lib.ts_memory_footprint(self.h, C.byref(b)))
        return b.value


def autotune(cfg):
    """Time the candidate tile heights for cfg (ts_autotune): returns (rows per tile, ms/step)."""
    lib = load()
    ty, ms = C.c_int(), C.c_double()
    rc = lib.ts_autotune(C.byref(cfg), C.byref(ty), C.byref(ms))
    if rc != TS_OK:
        raise TSError(rc, "ts_autotune failed")
    return ty.value, ms.value


EOS_PRESSURE, EOS_SOUND_SPEED, EOS_INTERNAL_ENERGY, EOS_DP_DE, EOS_DP_DRHO = range(5)


def eos_eval(gamma, method, ro, x, device=0):
    """The device ideal-gas EOS functor (ts_eos_eval) over arrays: method EOS_* of (ro, x)."""
    lib = load()
    ro = np.ascontiguousarray(ro, dtype=np.float64).ravel()
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    if ro.size != x.size:
        raise ValueError("ro and x differ in length")
    out = np.empty_like(ro)
    rc = lib.ts_eos_eval(gamma, method, _dp(ro), _dp(x), ro.size, _dp(out), device)
    if rc != TS_OK:
        raise TSError(rc, lib.ts_last_error(None).decode())
    return out


def autotune_state(cfg, aos=None, steps=5, cap=32):
    """Time the candidate tile heights on a given state (ts_autotune_state): returns
    (chosen rows per tile, [(rows per tile, kernel ms per step), ...])."""
    lib = load()
    rows = (C.c_int * cap)()
    ms = (C.c_double * cap)()
    n, ch = C.c_int(), C.c_int()
    if aos is None:
        p, nd = None, 0
    else:
        a = np.ascontiguousarray(aos, dtype=np.float64).ravel()
        p, nd = _dp(a), a.size
    rc = lib.ts_autotune_state(C.byref(cfg), p, nd, steps, rows, ms, cap, C.byref(n), C.byref(ch))
    if rc != TS_OK:
        raise TSError(rc, "ts_autotune_state failed")
    return ch.value, [(rows[i], ms[i]) for i in range(n.value)]


class Batch:
    """Owner of one ts_batch: many independent solves advanced together (continuous batching)."""

    def __init__(self, hbm_budget_bytes=8 << 30, max_active=1 << 20, use_graphs=True, device=0,
                 autotune_tiles=False):
        self.lib = load()
        bc = BatchConfig()
        bc.device = device
        bc.hbm_budget_bytes = hbm_budget_bytes
        bc.max_active = max_active
        bc.use_graphs = int(use_graphs)
        bc.autotune_tiles = int(autotune_tiles)
        h = C.c_void_p()
        rc = self.lib.ts_batch_create(C.byref(bc), C.byref(h))
        if rc != TS_OK:
            raise TSError(rc, "ts_batch_create failed")
        self.h = h
        self._cfgs = {}

    def _check(self, rc):
        if rc != TS_OK:
            msg = self.lib.ts_batch_last_error(self.h)
            raise TSError(rc, msg.decode() if msg else "")

    def close(self):
        if self.h:
            self.lib.ts_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, cfg, aos, max_steps, t_end=0.0):
        a = np.ascontiguousarray(aos, dtype=np.float64).ravel()
        t = C.c_longlong()
        self._check(self.lib.ts_batch_submit(self.h, C.byref(cfg), _dp(a), max_steps, t_end, C.byref(t)))
        self._cfgs[t.value] = cfg
        return t.value

    def run(self, n_steps):
        pending = C.c_int()
        self._check(self.lib.ts_batch_run(self.h, n_steps, C.byref(pending)))
        return pending.value

    def poll(self, ticket):
        """(done, state or None, StepInfo or None); raises TSError for a failed solve"""
        cfg = self._cfgs[ticket]
        out = np.empty(int(np.prod(state_shape(cfg))), dtype=np.float64)
        done = C.c_int()
        info = StepInfo()
        rc = self.lib.ts_batch_poll(self.h, ticket, C.byref(done), _dp(out), C.byref(info))
        if done.value:
            del self._cfgs[ticket]
        self._check(rc)
        if not done.value:
            return False, None, None
        return True, out, info

    def start_server(self):
        self._check(self.lib.ts_batch_start_server(self.h))

    def stop_server(self):
        self._check(self.lib.ts_batch_stop_server(self.h))

    def tuning(self, cap=256):
        """The autotuner's timed trials (ts_batch_tuning): list of dicts."""
        arr = (TileTrial * cap)()
        n = C.c_int()
        self._check(self.lib.ts_batch_tuning(self.h, arr, cap, C.byref(n)))
        return [{f: getattr(arr[i], f) for f, _ in TileTrial._fields_} for i in range(n.value)]
