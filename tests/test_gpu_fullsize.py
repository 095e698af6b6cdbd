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

and against the oracle at those sizes (the oracle running on the host's cores):

* the same 2048^2 problem over 20 steps, and the 256^3 Taylor-Green vortex over 3 steps, FP64:
  dt of every step and the state agree to 1e-12 (normwise per variable, as in the parity matrix),
  and every cell agrees with its own magnitude (elementwise relative, with an absolute floor only
  of 1e-3 of the largest momentum for the momentum components, which vanish by symmetry on
  lines and planes of the domain) to 1e-11.  Measured: dt within 5e-16, normwise 1.3e-15,
  elementwise 1.4e-15 for rho and E and 9.6e-13 for the momenta (that is, 1e-15 of the largest
  momentum where the floor applies).

and the per-tile dt reduction at its largest tile count: a 2^22-cell 1-D grid (69906 strips),
forward Euler and SSP-RK3, against the oracle over 6 steps (measured: dt within 3e-16,
normwise 2e-15);

and for the batched form at its full size:

* the 4096 shock tubes of the batched workload (N ~ logU[64, 4096], exact Godunov, FP32
  residual, t_end in-domain) all resident in one batch: a spread of the tubes, the smallest and
  the largest among them, end bitwise equal (state, time and step count) to the same tube solved
  alone, so sharing the persistent launches with 4095 others changes nothing;
* the served form on the serving workload's requests (2-D four-quadrant problems, side
  ~ logU[64, 1024], 50-500 steps, FP32 residual): 40 requests submitted to the server thread
  while it runs end bitwise equal to each request solved alone with ts_step.
"""
import os

import numpy as np
import pytest

from turbostructred_amd import capi, workloads
from oracle import pyoracle
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


def oracle_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def against_oracle(cfg, U, steps):
    """GPU and oracle over `steps` steps from U: (gpu interior, oracle interior, gpu dts, oracle dts)."""
    g = capi.Solver(cfg)
    g.set_state(U)
    g.enable_dt_log(steps + 4)
    gi = g.step(steps)
    dg = g.dt_log(steps + 4)
    a = capi.interior(cfg, g.get_state()).reshape(-1, 5).copy()
    g.close()
    r = pyoracle.Oracle(cfg, threads=oracle_threads())
    r.set_state(U)
    ri, dr = r.step(steps, log=True)
    b = capi.interior(cfg, r.get_state()).reshape(-1, 5).copy()
    assert gi.steps_done == ri.steps_done == steps
    assert abs(gi.time - ri.time) <= 1e-12 * ri.time
    return a, b, np.array(dg), np.array(dr)


def check_full(a, b, dg, dr, elem_tol):
    assert len(dg) == len(dr)
    np.testing.assert_array_less(np.abs(dg - dr), 1e-12 * np.abs(dr))
    nw = util.normwise_rel(a, b)
    mom = np.abs(b[:, 1:4]).max()
    ew = util.elementwise_rel(a, b, [0.0, 1e-3 * mom, 1e-3 * mom, 1e-3 * mom, 0.0])
    print(f"dt max rel {np.max(np.abs(dg - dr) / np.abs(dr)):.2e}; normwise {nw}; elementwise {ew}")
    assert np.all(nw <= 1e-12), nw
    assert np.all(ew <= elem_tol), ew


def test_c2_2048_matches_oracle():
    """2048^2 Lax-Liu 3, HLLC + BJ, SSP-RK3, CFL 0.45, 20 steps (ERK:1440-1491 each), with the
    tiling the library picks at this size."""
    n, steps = 2048, 20
    cfg = capi.make_config(ndims=2, nx=n, ny=n, cfl=0.45, riemann=capi.RIEMANN_HLLC, recon=capi.RECON_LINEAR_LIM,
                           limiter=capi.LIMITER_BJ, rk_stages=3, bcs=[dict(type=capi.BC_NATURAL)])
    a, b, dg, dr = against_oracle(cfg, util.quadrants(cfg), steps)
    check_full(a, b, dg, dr, elem_tol=1e-11)


def test_c3_256_matches_oracle():
    """256^3 periodic Taylor-Green vortex, HLLC + BJ, SSP-RK3, CFL 0.45, 3 steps."""
    n, steps = 256, 3
    L = 2 * np.pi
    cfg = capi.make_config(ndims=3, nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, periodic=(1, 1, 1), cfl=0.45,
                           riemann=capi.RIEMANN_HLLC, recon=capi.RECON_LINEAR_LIM, rk_stages=3)
    g = cfg.gamma

    def f(x, y, z):
        u = np.sin(x) * np.cos(y) * np.cos(z)
        v = -np.cos(x) * np.sin(y) * np.cos(z)
        p = 1 / (g * 0.01) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2) / 16
        return util.prim2cons(1 + 0 * x, u, v, 0 * x, p, g)

    a, b, dg, dr = against_oracle(cfg, util.fill_state(cfg, f), steps)
    check_full(a, b, dg, dr, elem_tol=1e-11)


@pytest.mark.parametrize("rk", [1, 3])
def test_1d_4m_cells_dt_reduction_matches_oracle(rk):
    """A 1-D grid of 2^22 cells (69906 strips, so 69906 per-tile minima of V/SR per dt pass or
    fused stage 1, reduced by one workgroup), with a smooth profile whose fastest cell is unique,
    limited linear BJ + HLLC, forward Euler (separate dt pass) or SSP-RK3 (fused): dt of every
    step and the state against the oracle."""
    n, steps = 1 << 22, 6
    cfg = capi.make_config(ndims=1, nx=n, cfl=0.45, riemann=capi.RIEMANN_HLLC, recon=capi.RECON_LINEAR_LIM,
                           limiter=capi.LIMITER_BJ, rk_stages=rk, bcs=[dict(type=capi.BC_NATURAL)])
    g = cfg.gamma

    def f(x, y, z):
        rho = 1.0 + 0.5 * np.sin(2 * np.pi * 3 * x)
        u = 0.3 * np.exp(-((x - 0.6173) / 0.05) ** 2)
        p = 1.0 + 0.25 * np.cos(2 * np.pi * 5 * x)
        return util.prim2cons(rho, u, 0 * x, 0 * x, p, g)

    a, b, dg, dr = against_oracle(cfg, util.fill_state(cfg, f), steps)
    check_full(a, b, dg, dr, elem_tol=1e-11)


def test_c4_4096_tube_batch_equals_single_solves():
    tubes = workloads.shock_tubes(4096, seed=2024, precision=capi.FP32_RESIDUAL)
    b = capi.Batch(hbm_budget_bytes=4 << 30)
    tickets = [b.submit(cfg, U, 10**6, te) for cfg, U, te in tubes]
    passes = 0
    while b.run(32) > 0:
        passes += 1
        assert passes < 5000, "the batch did not drain"
    sizes = np.array([cfg.nx for cfg, _, _ in tubes])
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes))} | set(range(0, 4096, 512)))
    got = {}
    for i, t in enumerate(tickets):
        done, st, info = b.poll(t)
        assert done
        if i in pick:
            got[i] = (st, info)
    b.close()
    for i in pick:
        cfg, U, te = tubes[i]
        s = capi.Solver(cfg)
        s.set_state(U)
        ainfo = s.step(10**6, te, te)
        alone = s.get_state()
        s.close()
        st, info = got[i]
        assert info.steps_done == ainfo.steps_done and info.time == ainfo.time, (i, cfg.nx)
        np.testing.assert_array_equal(capi.interior(cfg, st), capi.interior(cfg, alone))


def test_c5_served_requests_equal_single_solves():
    import time
    reqs = workloads.riemann_2d(40, seed=7, precision=capi.FP32_RESIDUAL)
    b = capi.Batch(hbm_budget_bytes=8 << 30)
    b.start_server()
    try:
        tickets = []
        for cfg, U, steps in reqs:
            tickets.append(b.submit(cfg, U, steps))
            time.sleep(0.001)
        got = {}
        deadline = time.time() + 120
        while len(got) < len(tickets) and time.time() < deadline:
            for t in tickets:
                if t not in got:
                    try:
                        d, st, info = b.poll(t)
                    except capi.TSError:  # a request whose state turns unphysical
                        got[t] = None
                        continue
                    if d:
                        got[t] = (st, info)
            time.sleep(0.002)
        assert len(got) == len(tickets), "the server did not finish the requests"
    finally:
        b.stop_server()
        b.close()
    assert sum(v is not None for v in got.values()) >= 30
    for (cfg, U, steps), t in zip(reqs, tickets):
        s = capi.Solver(cfg)
        s.set_state(U)
        if got[t] is None:  # fails alone as well
            with pytest.raises(capi.TSError):
                s.step(steps)
            s.close()
            continue
        info = s.step(steps)
        alone = s.get_state()
        s.close()
        st, binfo = got[t]
        assert binfo.steps_done == info.steps_done and binfo.time == info.time, (cfg.nx, steps)
        np.testing.assert_array_equal(capi.interior(cfg, st), capi.interior(cfg, alone))
