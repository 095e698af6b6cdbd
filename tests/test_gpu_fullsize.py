"""GPU checks at the full sizes of the benchmark workloads, through properties that do not need
an oracle run of that size:

* 2048^2 four-quadrant Riemann problem (Lax-Liu configuration 3, limited linear BJ + HLLC,
  SSP-RK3, FP64): the configuration is symmetric about the diagonal (rho(x, y) = rho(y, x)),
  every state stays physical, and cells the waves cannot have reached keep their initial state
  (a uniform neighbourhood has exactly zero flux divergence; only the SSP stage combinations
  3/4 U + 1/4 U and 1/3 U + 2/3 U may round by an ulp);
* 256^3 periodic Taylor-Green vortex (limited linear BJ + HLLC, SSP-RK3), FP64 and FP32-residual:
  mass, momentum and energy totals are conserved (every face flux is evaluated once and enters
  its two cells with opposite signs), and the total momentum stays zero.
"""
import numpy as np
import pytest

from turbostructred_amd import capi
import util

pytestmark = pytest.mark.gpu


def test_c2_2048_symmetry_positivity_and_far_field():
    n, steps = 2048, 20
    cfg = capi.make_config(ndims=2, nx=n, ny=n, cfl=0.45, riemann=capi.RIEMANN_HLLC, recon=capi.RECON_LINEAR_LIM,
                           limiter=capi.LIMITER_BJ, rk_stages=3, bcs=[dict(type=capi.BC_NATURAL)])
    U0 = util.quadrants(cfg)
    s = capi.Solver(cfg)
    s.set_state(U0)
    info = s.step(steps)
    assert info.steps_done == steps
    U = capi.interior(cfg, s.get_state()).reshape(n, n, 5)
    I0 = capi.interior(cfg, U0.ravel()).reshape(n, n, 5)
    rho = U[..., 0]
    assert np.all(np.isfinite(U))
    p = 0.4 * (U[..., 4] - 0.5 * (U[..., 1] ** 2 + U[..., 2] ** 2 + U[..., 3] ** 2) / rho)
    assert rho.min() > 0 and p.min() > 0
    # rho(x, y) = rho(y, x) to round-off; the momenta swap
    np.testing.assert_allclose(rho, rho.T, rtol=1e-12)
    np.testing.assert_allclose(U[..., 1], U[..., 2].T, rtol=1e-12, atol=1e-12)
    # the fastest signal (|u| + c <= 3.5 here) travels less than 3.5 * t; beyond that distance
    # from the interfaces at x = 0.5 and y = 0.5, plus three cells of stencil per stage, the state
    # is the initial one up to the rounding of the stage combinations
    h = 1.0 / n
    reach = 3.5 * info.time + 3 * steps * 3 * h
    c = (np.arange(n) + 0.5) * h
    far = (np.abs(c - 0.5) > reach)
    mask = far[:, None] & far[None, :]
    assert mask.sum() > 0.5 * n * n
    np.testing.assert_allclose(U[mask], I0[mask], rtol=4e-15, atol=0)


@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-13), (capi.FP32_RESIDUAL, 1e-6)])
def test_c3_256_tgv_conservation(precision, tol):
    n, steps = 256, 10
    L = 2 * np.pi
    cfg = capi.make_config(ndims=3, nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, periodic=(1, 1, 1), cfl=0.45,
                           riemann=capi.RIEMANN_HLLC, recon=capi.RECON_LINEAR_LIM, rk_stages=3, precision=precision)
    g = cfg.gamma
    c = (np.arange(n) + 0.5) * (L / n)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    xi = np.random.Generator(np.random.Philox(11)).uniform(-1, 1, (n, n, n))
    rho = 1 + 1e-3 * xi
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    p = 1 / (g * 0.01) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2) / 16
    U = np.zeros((n + 2, n + 2, n + 2, 5))
    I = U[1:-1, 1:-1, 1:-1]
    I[..., 0] = rho
    I[..., 1] = rho * u
    I[..., 2] = rho * v
    I[..., 4] = p / (g - 1) + 0.5 * rho * (u * u + v * v)
    del x, y, z, xi, u, v, p
    # per-variable sums over contiguous copies: numpy sums those pairwise (a reduction along
    # axis 0 of an (N, 5) array adds the rows one by one, with an error near N eps |sum|)
    def totals(a):
        a = a.reshape(-1, 5)
        return np.array([np.ascontiguousarray(a[:, q]).sum() for q in range(5)])

    tot0 = totals(I)
    scale = totals(np.abs(I))
    s = capi.Solver(cfg)
    s.set_state(U)
    del U, I
    info = s.step(steps)
    assert info.steps_done == steps
    Un = capi.interior(cfg, s.get_state()).reshape(-1, 5)
    assert np.all(np.isfinite(Un))
    tot = totals(Un)
    # mass and energy totals against their own size; the momentum totals (zero up to round-off:
    # w is zero everywhere) against the total momentum magnitude
    mom = scale[1] + scale[2] + scale[3]
    for v_ in range(5):
        ref = mom if v_ in (1, 2, 3) else scale[v_]
        assert abs(tot[v_] - tot0[v_]) <= tol * ref, (v_, tot[v_], tot0[v_], ref)
