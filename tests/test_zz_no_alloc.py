"""ts_step_noalloc (the allocation-free step over a caller's device state) against ts_step on a
context, through the C ABI: bit-identical states and step information, on two calls in a row
(the second on a smaller problem that reuses the same fixed scratch region), for forward Euler
(odd step count: the result comes back from the second buffer), SSP-RK3 FP64 and the FP32
residual form; a problem larger than the region is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from turbostructred_amd import capi
import util

pytestmark = pytest.mark.gpu


def hip_runtime():
    """The HIP runtime the library itself linked (found in this process's mappings), so that the
    device buffers below come from the same runtime instance."""
    lib = capi.load()
    with open("/proc/self/maps") as f:
        paths = sorted({line.split()[-1] for line in f if "libamdhip64" in line and "/" in line})
    assert paths, "the library did not map libamdhip64"
    hip = C.CDLL(paths[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return lib, hip


H2D, D2H = 1, 2


def soa_interior(cfg, aos):
    """Interior of an AoS state as 5 planes of nx*ny*nz values (x fastest)."""
    it = capi.interior(cfg, aos)  # (nz, ny, nx, 5)
    return np.ascontiguousarray(np.moveaxis(it, -1, 0)).ravel()


def by_context(cfg, U, steps, t0=0.0, it0=0):
    s = capi.Solver(cfg)
    s.set_state(U)
    if t0 or it0:
        s.set_time(t0, it0)
    info = s.step(steps)
    out = soa_interior(cfg, s.get_state())
    s.close()
    return out, info


def by_noalloc(lib, hip, cfg, U, steps, t0=0.0, it0=0):
    host = soa_interior(cfg, U)
    d = C.c_void_p()
    assert hip.hipMalloc(C.byref(d), host.nbytes) == 0
    try:
        assert hip.hipMemcpy(d, host.ctypes.data, host.nbytes, H2D) == 0
        info = capi.StepInfo()
        rc = lib.ts_step_noalloc(C.byref(cfg), d, t0, it0, steps, 1e300, 0.0, C.byref(info), None)
        assert rc == capi.TS_OK, (rc, lib.ts_last_error(None))
        out = np.empty_like(host)
        assert hip.hipMemcpy(out.ctypes.data, d, out.nbytes, D2H) == 0
    finally:
        assert hip.hipFree(d) == 0
    return out, info


CASES = [
    # (name, config keywords, state maker, steps, start time/iteration)
    ("2d_hllc_bj_rk3_fp64", dict(ndims=2, nx=200, ny=136, cfl=0.45, riemann=capi.RIEMANN_HLLC,
                                 recon=capi.RECON_LINEAR_LIM, rk_stages=3), util.quadrants, 40, (0.0, 0)),
    ("1d_roe_pwc_euler_fp64", dict(ndims=1, nx=301, cfl=0.4, riemann=capi.RIEMANN_ROE, recon=capi.RECON_PWC,
                                   rk_stages=1), util.sod, 37, (0.5, 12)),
    ("2d_hllc_bj_rk3_fp32res", dict(ndims=2, nx=96, ny=70, cfl=0.45, riemann=capi.RIEMANN_HLLC,
                                    recon=capi.RECON_LINEAR_LIM, rk_stages=3, precision=capi.FP32_RESIDUAL),
     util.quadrants, 25, (0.0, 0)),
]


@pytest.mark.parametrize("first,second", [(0, 1), (0, 2), (2, 1)])
def test_noalloc_equals_context_two_calls(first, second):
    lib, hip = hip_runtime()
    for idx in (first, second):  # the second problem is smaller than the first
        name, kw, make, steps, (t0, it0) = CASES[idx]
        print(f"case {name}", flush=True)
        cfg = capi.make_config(**kw)
        U = make(cfg)
        a, ia = by_context(cfg, U, steps, t0, it0)
        b, ib = by_noalloc(lib, hip, cfg, U, steps, t0, it0)
        assert np.array_equal(a, b), (name, np.abs(a - b).max())
        assert (ia.time, ia.dt, ia.iteration, ia.steps_done) == (ib.time, ib.dt, ib.iteration, ib.steps_done), name


def test_noalloc_refuses_a_problem_beyond_the_region():
    lib, hip = hip_runtime()
    cfg = capi.make_config(ndims=2, nx=1024, ny=1024, cfl=0.45, riemann=capi.RIEMANN_HLLC,
                           recon=capi.RECON_LINEAR_LIM, rk_stages=3)  # 80 B per cell of stage arrays
    d = C.c_void_p()
    n = cfg.nx * cfg.ny * 5 * 8
    assert hip.hipMalloc(C.byref(d), n) == 0
    try:
        info = capi.StepInfo()
        rc = lib.ts_step_noalloc(C.byref(cfg), d, 0.0, 0, 5, 1e300, 0.0, C.byref(info), None)
    finally:
        assert hip.hipFree(d) == 0
    assert rc == capi.TS_ERR_INVALID
    assert b"scratch" in lib.ts_last_error(None)
