"""CPU-side checks of the C ABI: the library loads, exports every function declared in
include/ts_fvm.h, and rejects bad configurations without touching a GPU."""
import ctypes as C
import os
import re

import pytest

from turbostructred_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ts_fvm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\s*\*)\s+(ts_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    names = declared_functions()
    assert len(names) >= 20
    assert sorted(capi.EXPORTED) == names


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_the_header(tmp_path):
    """offsetof/sizeof of every field, from a C compile of the header, against ctypes"""
    structs = {"ts_config": capi.Config, "ts_bc_config": capi.BCConfig, "ts_step_info": capi.StepInfo,
               "ts_batch_config": capi.BatchConfig, "ts_tile_trial": capi.TileTrial}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ts_fvm.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split("\n")
    for line in filter(None, out):
        cname, field, val = line.split()
        cls = structs[cname]
        got = C.sizeof(cls) if field == "sizeof" else getattr(cls, field).offset
        assert got == int(val), (cname, field, got, val)


@pytest.mark.parametrize("bad", [
    dict(ndims=4), dict(nx=1), dict(gamma=1.0), dict(cfl=0.0), dict(riemann=7), dict(recon=9),
    dict(rk_stages=2), dict(precision=5),
])
def test_invalid_configs_rejected_before_device_use(bad):
    lib = capi.load()
    kw = dict(ndims=1, nx=10)
    kw.update(bad)
    cfg = capi.make_config(**{k: v for k, v in kw.items() if k in ("ndims", "nx", "gamma", "cfl", "riemann", "recon", "rk_stages", "precision")})
    h = C.c_void_p()
    rc = lib.ts_create(C.byref(cfg), C.byref(h))
    assert rc in (capi.TS_ERR_INVALID, capi.TS_ERR_UNSUPPORTED)
    assert not h.value


@pytest.mark.parametrize("bad,text", [
    (dict(precision=5), "unknown precision"), (dict(riemann=7), "unknown Riemann solver"),
    (dict(nx=1), "at least 2 cells"), (dict(gamma=1.0), "gamma must exceed 1"),
])
def test_failed_create_is_readable_through_a_null_context(bad, text):
    """ts_create returns no context on failure; ts_last_error(NULL) gives this thread's message
    (and the Python binding raises with it)"""
    lib = capi.load()
    kw = dict(ndims=1, nx=10)
    kw.update(bad)
    cfg = capi.make_config(**kw)
    h = C.c_void_p()
    assert lib.ts_create(C.byref(cfg), C.byref(h)) != capi.TS_OK
    msg = lib.ts_last_error(None).decode()
    assert msg.startswith("ts_create:") and text in msg, msg
    with pytest.raises(capi.TSError, match=text):
        capi.Solver(cfg)


def test_cell_count_limit_of_the_plane_offsets():
    """a state plane is addressed with 32-bit byte offsets: 2^29 cells and more are refused"""
    lib = capi.load()
    cfg = capi.make_config(ndims=3, nx=1024, ny=1024, nz=512)
    h = C.c_void_p()
    assert lib.ts_create(C.byref(cfg), C.byref(h)) == capi.TS_ERR_UNSUPPORTED
    assert "2^29" in lib.ts_last_error(None).decode()


@pytest.mark.parametrize("kw,code", [
    (dict(ndims=3, nx=4, ny=4, nz=4, precision=capi.FP32_RESIDUAL, viscous=True, viscosity=0.1),
     capi.TS_ERR_UNSUPPORTED),
    (dict(ndims=2, nx=8, ny=8, precision=capi.FP32_RESIDUAL, sigma=(1.0, 0, 0)), capi.TS_ERR_UNSUPPORTED),
    (dict(ndims=1, nx=8, viscosity=-1.0), capi.TS_ERR_INVALID),
    (dict(ndims=1, nx=8, viscous=True, viscosity=float("nan")), capi.TS_ERR_INVALID),
    (dict(ndims=1, nx=8, accel=(float("inf"), 0, 0)), capi.TS_ERR_INVALID),
])
def test_viscous_configs_outside_the_built_forms_rejected(kw, code):
    """the viscous path and the body forces are FP64; bad coefficients are refused"""
    lib = capi.load()
    cfg = capi.make_config(**kw)
    h = C.c_void_p()
    assert lib.ts_create(C.byref(cfg), C.byref(h)) == code
    assert not h.value
