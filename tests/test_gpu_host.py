"""The C++ host program path: a shock-tube driver written like the reference's own
(src/Tests/tests.h:13-80, src/Tests/1DTests/ToroTests.h) over the HipExplicitRungeKuttaFVM class,
run as a separate process, against the oracle driven with the same grid, stop rule and snapshot
clamps (kernel.h:1701-1850, ERK:1433).  Bar: state and time to 1e-12 relative, same step count,
same number of snapshot stops."""
import json
import os
import subprocess

import numpy as np
import pytest

from turbostructred_amd import capi
from oracle import pyoracle
import util

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "turbostructred_amd", "host", "shock_tube_driver")
TORO = json.load(open(os.path.join(ROOT, "tests", "golden", "toro_star_table.json")))["tests"]


def run_driver(tmp_path, left, right, x0, t_end, nx, recon, riemann, rk, compress, save_time, norms):
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is missing: build it (make -C turbostructred_amd/csrc)")
    out = str(tmp_path / "out.bin")
    args = [DRIVER, *map(repr, left), *map(repr, right), repr(x0), repr(t_end), str(nx), str(recon),
            str(riemann), str(rk), str(compress), repr(save_time), str(norms), out]
    print("driver:", " ".join(args[1:-1]), flush=True)
    subprocess.run(args, check=True, timeout=120)
    raw = open(out, "rb").read()
    hdr = np.frombuffer(raw[:16], dtype=np.int32)
    nxa, snaps, iters = int(hdr[0]), int(hdr[1]), int(hdr[2])
    f = np.frombuffer(raw[16:], dtype=np.float64)
    st = f[:7]
    X, H = f[7:7 + nxa], f[7 + nxa:7 + 2 * nxa]
    U = f[7 + 2 * nxa:].reshape(1, 1, nxa, 5)
    return dict(iters=iters, snaps=snaps, time=st[0], dt=st[1], norms=st[2:7], X=X.copy(), H=H.copy(), U=U)


def run_oracle(cfg, U, t_end, save_time):
    """Kernel::Run with SaveSolutionTime (kernel.h:1701-1850) on the oracle."""
    o = pyoracle.Oracle(cfg)
    o.set_state(U)
    nxt, snaps, it = save_time, 0, 0
    while True:
        info, _ = o.step(1 << 20, t_end, nxt if save_time > 0 else 0.0, log=True)
        it += info.steps_done
        if save_time > 0 and info.time == nxt:
            snaps += 1
            nxt += save_time
        if info.steps_done == 0 or info.time >= t_end:
            break
    return o, info, snaps, it


CASES = [  # test, nx, recon, riemann, rk, compress, save_time, norms
    ("1", 100, capi.RECON_LINEAR_LIM, capi.RIEMANN_ROE, 1, 1, 0.05, 1),   # tests.h:13-80 grid, snapshots
    ("1", 400, capi.RECON_PWC, capi.RIEMANN_GODUNOV, 1, 0, 0.0, 0),
    ("2", 300, capi.RECON_PWC, capi.RIEMANN_GODUNOV, 3, 0, 0.05, 1),     # near-vacuum: first order
    ("3", 300, capi.RECON_ENO2, capi.RIEMANN_HLLC, 3, 0, 0.0, 0),
    ("4", 200, capi.RECON_PWC, capi.RIEMANN_ROE, 1, 1, 0.01, 0),
    ("5", 250, capi.RECON_LINEAR_LIM, capi.RIEMANN_HLLC, 3, 0, 0.0, 1),
]


@pytest.mark.parametrize("case", CASES, ids=[f"toro{c[0]}-n{c[1]}-r{c[2]}-f{c[3]}-rk{c[4]}-c{c[5]}" for c in CASES])
def test_driver_matches_oracle(tmp_path, case):
    t, nx, recon, riemann, rk, compress, save_time, norms = case
    T = TORO[t]
    d = run_driver(tmp_path, T["left"], T["right"], T["x0"], T["t"], nx, recon, riemann, rk, compress,
                   save_time, norms)
    cfg = capi.make_config(ndims=1, nx=nx, cfl=0.4, riemann=riemann, recon=recon, rk_stages=rk,
                           bcs=[dict(type=capi.BC_NATURAL)], coord=[d["X"], None, None], h=[d["H"], None, None])
    xc = d["X"][1:-1]
    left = xc < T["x0"]
    rho = np.where(left, T["left"][0], T["right"][0])
    u = np.where(left, T["left"][1], T["right"][1])
    p = np.where(left, T["left"][2], T["right"][2])
    U = np.zeros((1, 1, nx + 2, 5))
    U[0, 0, 1:-1] = util.prim2cons(rho, u, 0 * u, 0 * u, p, 1.4)
    o, info, snaps, it = run_oracle(cfg, U, T["t"], save_time)
    assert d["iters"] == it and d["snaps"] == snaps, (d["iters"], it, d["snaps"], snaps)
    assert abs(d["time"] - info.time) <= 1e-12 * info.time
    assert d["time"] >= T["t"]
    err = util.normwise_rel(capi.interior(cfg, d["U"]), capi.interior(cfg, o.get_state()))
    assert np.all(err <= 1e-12), err
    if norms:
        r = np.array(info.residual_norm)
        assert np.all(np.abs(d["norms"] - r) <= 1e-12 * np.abs(r).max())


def read_sol(path, nx):
    """A one-rank .sol file (kernel.h:992-1029): double time, int rank, then 5 doubles per interior
    cell in k, j, i order."""
    raw = open(path, "rb").read()
    assert len(raw) == 8 + 4 + nx * 5 * 8, len(raw)
    t = np.frombuffer(raw[:8], dtype=np.float64)[0]
    rank = np.frombuffer(raw[8:12], dtype=np.int32)[0]
    return t, int(rank), np.frombuffer(raw[12:], dtype=np.float64).reshape(nx, 5)


def test_save_solution_file_matches_oracle(tmp_path):
    """SaveSolution of the C++ host class writes the reference's one-rank .sol layout; read back
    independently, it holds the oracle's time and state."""
    T = TORO["1"]
    nx = 200
    sol = str(tmp_path / "a.sol")
    out = str(tmp_path / "out.bin")
    subprocess.run([DRIVER, *map(repr, T["left"]), *map(repr, T["right"]), repr(T["x0"]), repr(T["t"]), str(nx),
                    str(capi.RECON_LINEAR_LIM), str(capi.RIEMANN_HLLC), "3", "0", "0.0", "0", out, sol, "-"],
                   check=True, timeout=120)
    t, rank, cells = read_sol(sol, nx)
    cfg = capi.make_config(ndims=1, nx=nx, cfl=0.4, riemann=capi.RIEMANN_HLLC, recon=capi.RECON_LINEAR_LIM,
                           rk_stages=3, bcs=[dict(type=capi.BC_NATURAL)])
    xc = (np.arange(nx) + 0.5) / nx
    left = xc < T["x0"]
    U = np.zeros((1, 1, nx + 2, 5))
    U[0, 0, 1:-1] = util.prim2cons(np.where(left, T["left"][0], T["right"][0]), np.where(left, T["left"][1], T["right"][1]),
                                   0 * xc, 0 * xc, np.where(left, T["left"][2], T["right"][2]), 1.4)
    o, info, _, _ = run_oracle(cfg, U, T["t"], 0.0)
    assert rank == 0 and abs(t - info.time) <= 1e-12 * info.time
    err = util.normwise_rel(cells.reshape(1, 1, nx, 5), capi.interior(cfg, o.get_state()))
    assert np.all(err <= 1e-12), err


def test_restart_from_solution_file_is_seamless(tmp_path):
    """LoadSolution (kernel.h:1031-1079) then Run: a run stopped at T/2, saved and restarted to T
    ends bitwise where the uninterrupted run (snapshot-clamped at T/2) ends."""
    T = TORO["3"]
    nx, half = 240, T["t"] / 2
    common = [*map(repr, T["left"]), *map(repr, T["right"]), repr(T["x0"])]
    method = [str(nx), str(capi.RECON_LINEAR_LIM), str(capi.RIEMANN_HLLC), "3", "0"]
    a_sol, b_sol, c_sol = (str(tmp_path / f"{n}.sol") for n in "abc")
    out = str(tmp_path / "out.bin")
    subprocess.run([DRIVER, *common, repr(T["t"]), *method, repr(half), "0", out, a_sol, "-"], check=True, timeout=120)
    subprocess.run([DRIVER, *common, repr(half), *method, repr(half), "0", out, b_sol, "-"], check=True, timeout=120)
    subprocess.run([DRIVER, *common, repr(T["t"]), *method, repr(half), "0", out, c_sol, b_sol], check=True, timeout=120)
    ta, _, a = read_sol(a_sol, nx)
    tb, _, _ = read_sol(b_sol, nx)
    tc, _, c = read_sol(c_sol, nx)
    assert tb == half and ta == tc
    np.testing.assert_array_equal(a, c)


def test_history_and_sensors(tmp_path):
    """history.dat (kernel.h:1646-1657, 1823-1831) every 5 iterations and the sensors of
    src/Sensors/*.h every 7, written by the C++ host class from device reads, against the oracle
    stepped to the same iterations (values as the reference prints them: 6 significant digits)."""
    T = TORO["1"]
    nx = 200
    out = str(tmp_path / "out.bin")
    subprocess.run([DRIVER, *map(repr, T["left"]), *map(repr, T["right"]), repr(T["x0"]), repr(T["t"]), str(nx),
                    str(capi.RECON_LINEAR_LIM), str(capi.RIEMANN_HLLC), "3", "0", "0.0", "0", out, "-", "-",
                    str(tmp_path)], check=True, timeout=120)
    cfg = capi.make_config(ndims=1, nx=nx, cfl=0.4, riemann=capi.RIEMANN_HLLC, recon=capi.RECON_LINEAR_LIM,
                           rk_stages=3, bcs=[dict(type=capi.BC_NATURAL)])
    xc = (np.arange(nx) + 0.5) / nx
    left = xc < T["x0"]
    U = np.zeros((1, 1, nx + 2, 5))
    U[0, 0, 1:-1] = util.prim2cons(np.where(left, T["left"][0], T["right"][0]), np.where(left, T["left"][1], T["right"][1]),
                                   0 * xc, 0 * xc, np.where(left, T["left"][2], T["right"][2]), 1.4)
    o = pyoracle.Oracle(cfg)
    o.set_state(U)
    hist, cells = {}, {}
    while True:
        info, _ = o.step(1, T["t"], 0.0, log=True)
        it = info.iteration
        if it % 5 == 0:
            hist[it] = (info.time, np.array(info.residual_norm))
        if it % 7 == 0:
            cells[it] = (info.time, capi.interior(cfg, o.get_state())[0, 0].copy())
        if info.time >= T["t"]:
            break

    def rel(a, b):
        return abs(a - b) <= 1e-5 * max(abs(b), 1e-300) + 1e-12

    lines = open(tmp_path / "history.dat").read().strip().split("\n")
    assert lines[0].startswith('VARIABLES = "time" "roR" "rouR" "roeR" "iter"')
    assert len(lines) - 1 == len(hist)
    for ln in lines[1:]:
        t, r0, r1, r4, it = ln.split()
        ht, hn = hist[int(it)]
        assert rel(float(t), ht) and rel(float(r0), hn[0]) and rel(float(r1), hn[1]) and rel(float(r4), hn[4]), ln

    g1 = 0.4
    prs = lambda c: g1 * (c[:, 4] - 0.5 * c[:, 1] ** 2 / c[:, 0])  # noqa: E731
    cl = open(tmp_path / "cell.dat").read().strip().split("\n")
    assert cl[0].startswith('TITLE = "X=') and len(cl) - 1 == len(cells)
    for ln in cl[1:]:
        it, val, t = ln.split()
        ct, st = cells[int(it)]
        assert rel(float(val), st[nx // 2, 0]) and rel(float(t), ct), ln
    for ln in open(tmp_path / "multi.dat").read().strip().split("\n")[1:]:
        it, r, u, p, t = ln.split()
        st = cells[int(it)][1]
        assert rel(float(r), st[0, 0]) and rel(float(u), st[0, 1] / st[0, 0]) and rel(float(p), prs(st[:1])[0]), ln
    for name, second in (("pmax1.dat", False), ("pmax2.dat", True)):
        for ln in open(tmp_path / name).read().strip().split("\n"):
            it, x, t = ln.split()
            st = cells[int(it)][1]
            p = prs(st)
            i = int(np.argmax(p))
            if not second:
                want = xc[i]
            elif i == 0:
                want = 0.0
            elif i == nx - 1:
                want = 1.0
            else:
                # vertex of the parabola through (x_{i-1}, x_i, x_{i+1}) (MValuePosXSensor2.h:47-63)
                h = 1.0 / nx
                a = (p[i + 1] - 2 * p[i] + p[i - 1]) / (2 * h * h)
                b = (p[i + 1] - p[i - 1]) / (2 * h)
                want = xc[i] - b / (2 * a)
            assert abs(float(x) - want) <= 1e-5, (name, ln, want)


POIS = os.path.join(ROOT, "turbostructred_amd", "host", "poiseuille_driver")


@pytest.mark.parametrize("init,rk", [(1, 1), (2, 3)])
def test_poiseuille_driver_matches_oracle_and_centre_velocity(tmp_path, init, rk):
    """RunSingleExperiment of PoiseuilleChanel.h as a C++ program over the host class: compressed
    y-axis, wall markers, viscous fluxes and Sigma from KernelConfiguration.  Started 10% above
    (or below) the exact profile, it relaxes to u_c = sigma Ly^2 / (8 mu); the state matches the
    oracle run on the program's own grid."""
    if not os.path.exists(POIS):
        raise RuntimeError(f"{POIS} is missing: build it (make -C turbostructred_amd/csrc)")
    nx, ny, mu, sigma, pin, t_end, q = 4, 24, 0.1, 2.0, 100.0, 2.5, 1.02
    out = str(tmp_path / "pois.bin")
    args = [POIS, str(nx), str(ny), repr(mu), repr(sigma), repr(pin), repr(t_end), repr(q), str(init), str(rk), out]
    print("driver:", " ".join(args[1:-1]), flush=True)
    subprocess.run(args, check=True, timeout=120)
    raw = open(out, "rb").read()
    nxa, nya, iters, _ = np.frombuffer(raw[:16], dtype=np.int32)
    f = np.frombuffer(raw[16:], dtype=np.float64)
    t, dt, uc = f[:3]
    o = 3
    X, HX = f[o:o + nxa], f[o + nxa:o + 2 * nxa]
    o += 2 * nxa
    Y, HY = f[o:o + nya], f[o + nya:o + 2 * nya]
    o += 2 * nya
    Ud = f[o:].reshape(1, nya, nxa, 5)
    assert t >= t_end and uc == sigma * 0.25 / (8 * mu)
    cfg = capi.make_config(ndims=2, nx=nx, ny=ny, lx=1.0, ly=0.5, periodic=(1, 0, 0), cfl=0.45, eps=0.05,
                           riemann=capi.RIEMANN_ROE, recon=capi.RECON_LINEAR_LIM, rk_stages=rk,
                           bcs=[dict(type=capi.BC_WALL)], viscous=True, viscosity=mu, sigma=(sigma, 0, 0),
                           coord=[X.copy(), Y.copy(), None], h=[HX.copy(), HY.copy(), None])
    scale = {0: 1.0, 1: 1.1, 2: 0.9}[init]
    yy = 1.0 - np.abs(2 * Y / 0.5 - 1.0)
    u = scale * uc * yy * (2.0 - yy)
    U0 = np.zeros((1, nya, nxa, 5))
    U0[..., 0] = 1.0
    U0[..., 1] = u[None, :, None]
    U0[..., 4] = (pin / 0.4 + 0.5 * u * u)[None, :, None]
    ref = pyoracle.Oracle(cfg)
    ref.set_state(U0)
    info = ref.step(10_000_000, t_end)
    assert info.iteration == iters and abs(info.time - t) <= 1e-12 * t
    a, b = capi.interior(cfg, Ud.ravel()), capi.interior(cfg, ref.get_state())
    assert np.all(util.normwise_rel(a, b) <= 1e-10)
    mid = Ud[0, nya // 2 - 1:nya // 2 + 1, nxa // 2]
    assert abs(np.mean(mid[:, 1] / mid[:, 0]) / uc - 1) < 0.01


WRITER = r'''
#include <cstdio>
#include <vector>
#include "ts_kernel.hpp"
// writes the Tecplot file of a 1-D state: argv = nx time grid.bin state.bin out.dat
int main(int argc, char** argv) {
    int nx = std::atoi(argv[1]);
    double t = std::atof(argv[2]);
    std::vector<double> g(2 * (nx + 2)), U(5 * (nx + 2));
    FILE* f = std::fopen(argv[3], "rb"); if (std::fread(g.data(), 8, g.size(), f) != g.size()) return 1; std::fclose(f);
    f = std::fopen(argv[4], "rb"); if (std::fread(U.data(), 8, U.size(), f) != U.size()) return 1; std::fclose(f);
    tsh::TecplotView v;
    v.nDims = 1; v.nX = nx; v.nXAll = nx + 2;
    v.X = g.data(); v.hx = g.data() + nx + 2; v.values = U.data(); v.gamma = 1.4; v.time = t;
    tsh::WriteTecplotSolution(argv[5], v);
    return 0;
}
'''


def _tecplot_of(tmp_path, name, X, H, U, t):
    """the host writer fed a state computed elsewhere (here: the oracle's)"""
    src, exe = tmp_path / "writer.cpp", tmp_path / "writer"
    if not exe.exists():
        src.write_text(WRITER)
        pkg = os.path.join(ROOT, "turbostructred_amd")
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "turbostructred_amd", "host"), str(src),
                        "-o", str(exe), "-L", pkg, "-lts_fvm", f"-Wl,-rpath,{pkg}"], check=True)
    nx = len(X) - 2
    np.concatenate([X, H]).astype(np.float64).tofile(tmp_path / "g.bin")
    np.ascontiguousarray(U, dtype=np.float64).ravel().tofile(tmp_path / "u.bin")
    out = tmp_path / name
    subprocess.run([str(exe), str(nx), repr(t), str(tmp_path / "g.bin"), str(tmp_path / "u.bin"), str(out)], check=True)
    return out.read_text()


def _same_tecplot(a, b, rel=1e-5, floor=1e-12):
    """byte-identical, except that a printed 6-digit value may fall on the other side of a
    rounding boundary, or be a quantity that is zero up to round-off (a velocity of 1e-14 in the
    undisturbed state): such lines must agree numerically to the printed precision, with an
    absolute floor for the round-off values"""
    la, lb = a.split("\n"), b.split("\n")
    assert len(la) == len(lb)
    differ = 0
    for x, y in zip(la, lb):
        if x == y:
            continue
        differ += 1
        fx, fy = np.array(x.split(), dtype=float), np.array(y.split(), dtype=float)
        np.testing.assert_allclose(fx, fy, rtol=rel, atol=floor)
    return differ


def test_reference_sod_driver_as_configured(tmp_path):
    """shock_tube_driver --reference-sod runs RunSODTestRoe1D (tests.h:13-80) exactly: compressed
    100-cell grid, Linear2psLim<limVenkatar>, Roe Eps 0.05, CFL 0.5, forward Euler, MaxTime 0.25
    with SaveSolutionTime 0.25.  Against the oracle: step count, time and state at 1e-12.  Its
    Tecplot files (init.dat, init2.dat after the .sol save/load, and the snapshot
    "solution, t = 0.25.dat") equal the same writer fed the oracle's states."""
    out = tmp_path / "out.bin"
    print("driver: --reference-sod", flush=True)
    r = subprocess.run([DRIVER, "--reference-sod", str(tmp_path), str(out)], check=True, timeout=120,
                       capture_output=True, text=True)
    assert "Iteration = 100;" in r.stdout  # ResidualOutputIters = 100
    raw = open(out, "rb").read()
    hdr = np.frombuffer(raw[:16], dtype=np.int32)
    nxa, snaps, iters = int(hdr[0]), int(hdr[1]), int(hdr[2])
    f = np.frombuffer(raw[16:], dtype=np.float64)
    time_, X, H = f[0], f[7:7 + nxa].copy(), f[7 + nxa:7 + 2 * nxa].copy()
    Ug = f[7 + 2 * nxa:].reshape(1, 1, nxa, 5)
    assert nxa == 102 and snaps == 1 and time_ == 0.25
    cfg = capi.make_config(ndims=1, nx=100, cfl=0.5, eps=0.05, riemann=capi.RIEMANN_ROE, recon=capi.RECON_LINEAR_LIM,
                           limiter=capi.LIMITER_VENKAT, coord=[X, None, None], h=[H, None, None])
    left = X < 0.5  # tests.h:85, at the cell centres of the compressed grid (ghosts too: ignored)
    U0 = util.prim2cons(np.where(left, 1.0, 0.125), 0 * X, 0 * X, 0 * X, np.where(left, 1.0, 0.1), 1.4)
    U0 = U0.reshape(1, 1, nxa, 5)
    U0[0, 0, 0] = U0[0, 0, -1] = 0.0  # the host mirror's ghost entries before the first download
    o = pyoracle.Oracle(cfg)
    o.set_state(U0)
    info = o.step(1 << 20, 0.25, 0.25)
    assert info.time == 0.25 and info.steps_done == iters
    err = util.normwise_rel(capi.interior(cfg, Ug), capi.interior(cfg, o.get_state()))
    assert np.all(err <= 1e-12), err
    init = _tecplot_of(tmp_path, "o_init.dat", X, H, U0, 0.0)
    assert (tmp_path / "init.dat").read_text() == init
    assert (tmp_path / "init2.dat").read_text() == init
    fin = _tecplot_of(tmp_path, "o_fin.dat", X, H, o.get_state(), 0.25)
    n = _same_tecplot((tmp_path / "solution, t = 0.25.dat").read_text(), fin)
    print(f"snapshot lines differing in the last printed digit: {n}")
    assert n <= 10  # of 103 lines: the digits at rounding boundaries and the round-off velocities
