"""Synthetic workloads of the batched and served forms (input generation only).

All draws come from numpy's Philox counter-based generator with a recorded seed, so every
request can be regenerated exactly.  States are returned in the reference AoS layout
(ghosts included, kernel.h:88).
"""
import numpy as np

from . import capi


def prim2cons(rho, u, v, w, p, gamma):
    return np.stack([rho, rho * u, rho * v, rho * w, p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)], axis=-1)


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def exact_max_speed(gamma, rl, ul, pl, rr, ur, pr, iters=60):
    """Largest wave speed of the exact solution of 1-D Riemann problems (vectorised Newton for
    p*, Toro sec. 4.3; input generation for the t_end rule only)."""
    g = gamma
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)

    def f(p, r, pk, ak):
        A, B = 2 / ((g + 1) * r), (g - 1) / (g + 1) * pk
        shock = p > pk
        fs = (p - pk) * np.sqrt(A / (p + B))
        ds = np.sqrt(A / (B + p)) * (1 - 0.5 * (p - pk) / (B + p))
        fr = 2 * ak / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)
        dr = (p / pk) ** (-(g + 1) / (2 * g)) / (r * ak)
        return np.where(shock, fs, fr), np.where(shock, ds, dr)

    p = np.maximum(0.5 * (pl + pr), 1e-8)
    for _ in range(iters):
        fL, dL = f(p, rl, pl, al)
        fR, dR = f(p, rr, pr, ar)
        p = np.maximum(p - (fL + fR + ur - ul) / (dL + dR), 1e-12)
    sl = np.where(p > pl, ul - al * np.sqrt((g + 1) / (2 * g) * p / pl + (g - 1) / (2 * g)), ul - al)
    sr = np.where(p > pr, ur + ar * np.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g)), ur + ar)
    return np.maximum(np.abs(sl), np.abs(sr))


def shock_tubes(count=4096, seed=2024, precision=capi.FP32_RESIDUAL, riemann=capi.RIEMANN_GODUNOV):
    """Independent 1-D Riemann problems: N ~ logU[64, 4096] cells, rho ~ U[0.1,2], u ~ U[-1,1],
    p ~ U[0.1,10], gamma in {1.4, 1.67}, diaphragm ~ U[0.3, 0.7]; t_end = 0.9 min(x0, 1-x0)/S_max
    so that every wave stays inside.  Returns a list of (cfg, aos_state, t_end)."""
    r = _rng(seed)
    n = np.exp(r.uniform(np.log(64), np.log(4096), count)).astype(int)
    rl, rr = r.uniform(0.1, 2, count), r.uniform(0.1, 2, count)
    ul, ur = r.uniform(-1, 1, count), r.uniform(-1, 1, count)
    pl, pr = r.uniform(0.1, 10, count), r.uniform(0.1, 10, count)
    gam = np.where(r.uniform(0, 1, count) < 0.5, 1.4, 1.67)
    x0 = r.uniform(0.3, 0.7, count)
    smax = np.concatenate([exact_max_speed(g, rl[m], ul[m], pl[m], rr[m], ur[m], pr[m])
                           for g in (1.4, 1.67) for m in [gam == g]])
    order = np.concatenate([np.nonzero(gam == g)[0] for g in (1.4, 1.67)])
    s_all = np.empty(count)
    s_all[order] = smax
    out = []
    for i in range(count):
        t_end = 0.9 * min(x0[i], 1 - x0[i]) / float(s_all[i])
        cfg = capi.make_config(ndims=1, nx=int(n[i]), cfl=0.45, gamma=float(gam[i]), riemann=riemann,
                               recon=capi.RECON_LINEAR_LIM, rk_stages=3, precision=precision)
        N = int(n[i])
        x = (np.arange(N + 2) - 0.5) / N
        left = x < x0[i]
        U = prim2cons(np.where(left, rl[i], rr[i]), np.where(left, ul[i], ur[i]), 0 * x, 0 * x,
                      np.where(left, pl[i], pr[i]), gam[i])
        out.append((cfg, np.ascontiguousarray(U.reshape(1, 1, N + 2, 5)), t_end))
    return out


def riemann_2d(count, seed=7, precision=capi.FP32_RESIDUAL, lmin=64, lmax=1024):
    """2-D four-quadrant Riemann problems: side L ~ logU[lmin, lmax], quadrant states
    rho ~ U[0.1,2], p ~ U[0.05,2], |u|,|v| <= 1.5; steps ~ U{50..500}.  HLLC + limited linear +
    SSP-RK3, transmissive.  Returns a list of (cfg, aos_state, steps)."""
    r = _rng(seed)
    L = np.exp(r.uniform(np.log(lmin), np.log(lmax), count)).astype(int)
    steps = r.integers(50, 501, count)
    out = []
    for i in range(count):
        n = int(L[i])
        cfg = capi.make_config(ndims=2, nx=n, ny=n, cfl=0.45, riemann=capi.RIEMANN_HLLC,
                               recon=capi.RECON_LINEAR_LIM, rk_stages=3, precision=precision)
        q = r.uniform(size=(4, 4))
        rho = 0.1 + 1.9 * q[:, 0]
        p = 0.05 + 1.95 * q[:, 1]
        u = -1.5 + 3 * q[:, 2]
        v = -1.5 + 3 * q[:, 3]
        c = (np.arange(n + 2) - 0.5) / n
        X, Y = np.meshgrid(c, c)
        quad = (X >= 0.5).astype(int) + 2 * (Y >= 0.5).astype(int)
        U = prim2cons(rho[quad], u[quad], v[quad], 0 * X, p[quad], 1.4)
        out.append((cfg, np.ascontiguousarray(U.reshape(1, n + 2, n + 2, 5)), int(steps[i])))
    return out
