"""FP32-residual form against FP64 (both on the GPU): face states, fluxes and residuals in float,
state, stage combination and dt reduction in double.  Bar: density L1 <= 1e-5 against the FP64
solve of the same problem at the same final time (Linf bounds stated per case)."""
import numpy as np
import pytest

from turbostructred_amd import capi, workloads
import util

pytestmark = pytest.mark.gpu


def solve(cfg, U, steps=None, t_end=None):
    s = capi.Solver(cfg)
    s.set_state(U)
    info = s.step(10 ** 6, t_end, t_end) if t_end is not None else s.step(steps)
    return capi.interior(cfg, s.get_state())[..., 0], info


def fp64_twin(cfg, **over):
    kw = dict(ndims=cfg.ndims, nx=cfg.nx, ny=cfg.ny, nz=cfg.nz, lx=cfg.lx, ly=cfg.ly, lz=cfg.lz,
              periodic=tuple(cfg.periodic), gamma=cfg.gamma, cfl=cfg.cfl, riemann=cfg.riemann, recon=cfg.recon,
              limiter=cfg.limiter, rk_stages=cfg.rk_stages, precision=capi.FP64)
    kw.update(over)
    return capi.make_config(**kw)


def test_shock_tubes_fp32_residual():
    """Batched-workload tubes (exact Godunov, limited linear, SSP-RK3) run to t_end: L1 <= 1e-5,
    Linf <= 1e-2 (a shock cell may round differently)."""
    for cfg, U, t_end in workloads.shock_tubes(12, seed=99, precision=capi.FP32_RESIDUAL):
        a, ia = solve(cfg, U, t_end=t_end)
        b, ib = solve(fp64_twin(cfg), U, t_end=t_end)
        assert abs(ia.time - t_end) <= 1e-12 * t_end and abs(ib.time - t_end) <= 1e-12 * t_end
        d = np.abs(a - b)
        print(f"tube n={cfg.nx}: density L1 {d.mean():.3e} Linf {d.max():.3e}")
        assert d.mean() <= 1e-5 and d.max() <= 1e-2, (cfg.nx, d.mean(), d.max())


@pytest.mark.parametrize("riemann,recon", [(capi.RIEMANN_HLLC, capi.RECON_LINEAR_LIM),
                                           (capi.RIEMANN_ROE, capi.RECON_PWC)])
def test_quadrants_fp32_residual(riemann, recon):
    """2-D four-quadrant problem, 256^2, 150 SSP-RK3 steps: L1 <= 1e-6, Linf <= 1e-4.  (Roe has
    no positivity fallback, and limited linear on conserved variables gives negative face
    pressures on this problem: Roe runs it first order, as the reference would have to.)"""
    cfg = capi.make_config(ndims=2, nx=256, ny=256, cfl=0.45, riemann=riemann, recon=recon,
                           rk_stages=3, precision=capi.FP32_RESIDUAL)
    U = util.quadrants(cfg)
    a, ia = solve(cfg, U, steps=150)
    b, ib = solve(fp64_twin(cfg), U, steps=150)
    assert abs(ia.time - ib.time) <= 1e-5 * ib.time
    d = np.abs(a - b)
    print(f"quadrants riemann={riemann} recon={recon}: density L1 {d.mean():.3e} Linf {d.max():.3e}")
    assert d.mean() <= 1e-6 and d.max() <= 1e-4, (d.mean(), d.max())


def test_taylor_green_fp32_residual():
    """3-D periodic Taylor-Green vortex (M = 0.1), 48^3, 40 steps: L1 <= 1e-6, Linf <= 1e-5."""
    n = 48
    cfg = capi.make_config(ndims=3, nx=n, ny=n, nz=n, lx=2 * np.pi, ly=2 * np.pi, lz=2 * np.pi,
                           periodic=(1, 1, 1), cfl=0.45, riemann=capi.RIEMANN_HLLC, recon=capi.RECON_LINEAR_LIM,
                           rk_stages=3, precision=capi.FP32_RESIDUAL)
    g = cfg.gamma

    def f(x, y, z):
        u = np.sin(x) * np.cos(y) * np.cos(z)
        v = -np.cos(x) * np.sin(y) * np.cos(z)
        p = 1 / (g * 0.01) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2) / 16
        return util.prim2cons(1 + 0 * x, u, v, 0 * x, p, g)

    U = util.fill_state(cfg, f)
    a, ia = solve(cfg, U, steps=40)
    b, ib = solve(fp64_twin(cfg), U, steps=40)
    d = np.abs(a - b)
    print(f"taylor-green 48^3: density L1 {d.mean():.3e} Linf {d.max():.3e}")
    assert d.mean() <= 1e-6 and d.max() <= 1e-5, (d.mean(), d.max())
