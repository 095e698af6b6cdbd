"""Host-side pieces of the C++ class that need no GPU (turbostructred_amd/host/ts_kernel.hpp):
Gauss-quadrature cell averages (NumericQuadrature.h, kernel.h:149-175), the ideal-gas equation of
state (idealGasEOS.h) and the Tecplot / slice writers (kernel.h:1081-1632).  A probe program built
with the host compiler calls them on small synthetic grids; the checks are library routines
(mpmath's Legendre polynomials), closed forms, thermodynamic identities and the file layouts
written out independently here."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "turbostructred_amd", "host")

PROBE = r'''
#include <cstdio>
#include <cmath>
#include "ts_kernel.hpp"
using namespace tsh;
int main(int argc, char** argv) {
    std::string dir = argv[1];
    // quadrature rules
    for (int n : {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 32}) {
        NumericQuadrature q(n, 1);
        std::printf("Q %d", n);
        for (int i = 0; i < n; i++) std::printf(" %.17g %.17g", q.Nodes()[i], q.Weights()[i]);
        std::printf("\n");
    }
    try { NumericQuadrature bad(13, 1); std::printf("BAD accepted\n"); } catch (const std::invalid_argument&) { std::printf("BAD refused\n"); }
    // cell averages: 1-D monomials, 2-D / 3-D products, constants
    const double xc = 0.3, yc = -0.7, zc = 1.1, hx = 0.25, hy = 0.5, hz = 0.125;
    for (int n : {1, 2, 3, 4, 5, 7, 10, 12}) {
        for (int p = 0; p <= 7; p++) {
            NumericQuadrature q(n, 1);
            auto f = [p](const Vector& r) { return std::vector<double>{std::pow(r.x, p)}; };
            double avg = q.IntegrateGroap(xc, 0, 0, hx, 1, 1, f, 1)[0] / hx;
            std::printf("A1 %d %d %.17g\n", n, p, avg);
        }
        for (int d = 2; d <= 3; d++) {
            NumericQuadrature q(n, d);
            auto f = [](const Vector& r) { return std::vector<double>{1.0, r.x * r.y, r.y, r.z * r.x, r.y * r.y}; };
            double V = hx * hy * (d > 2 ? hz : 1.0);
            std::valarray<double> a = q.IntegrateGroap(xc, yc, zc, hx, hy, d > 2 ? hz : 1.0, f, 5) / V;
            std::printf("A%d %d %.17g %.17g %.17g %.17g %.17g\n", d, n, a[0], a[1], a[2], a[3], a[4]);
        }
    }
    // equation of state
    IdealGasEOS eos(1.4);
    for (double ro : {0.2, 1.0, 3.5})
        for (double e : {0.1, 2.5}) {
            double p = eos.GetPressure(ro, e);
            std::printf("E %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", ro, e, p, eos.GetSoundSpeed(ro, e),
                        eos.GetInternalEnergy(ro, p), eos.GetPressureEnergyDerivative(ro, e),
                        eos.GetPressureDensityDerivative(ro, e));
        }
    // Tecplot writers on small grids: state U = (rho, rho u, rho v, rho w, rho E) from cell index
    for (int nd = 1; nd <= 3; nd++) {
        int n[3] = {4, nd > 1 ? 3 : 1, nd > 2 ? 2 : 1};
        int na[3] = {n[0] + 2, nd > 1 ? n[1] + 2 : 1, nd > 2 ? n[2] + 2 : 1};
        std::vector<double> X[3], H[3];
        for (int a = 0; a < nd; a++) {
            std::vector<BlockNode> b(1);
            BuildAxis(b, n[a], 1.0 + a, 1, X[a], H[a]);
        }
        std::vector<double> U((size_t)na[0] * na[1] * na[2] * 5, 0.0);
        for (int k = 0; k < na[2]; k++)
            for (int j = 0; j < na[1]; j++)
                for (int i = 0; i < na[0]; i++) {
                    double* u = &U[(((size_t)k * na[1] + j) * na[0] + i) * 5];
                    double ro = 1.0 + 0.1 * i + 0.01 * j + 0.001 * k;
                    u[0] = ro; u[1] = ro * 0.3 * (i - 2); u[2] = ro * 0.2 * j; u[3] = -ro * 0.1 * k;
                    u[4] = 2.5 + 0.05 * i + 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / ro;
                }
        TecplotView v;
        v.nDims = nd; v.nX = n[0]; v.nY = n[1]; v.nZ = n[2]; v.nXAll = na[0]; v.nYAll = na[1];
        v.X = X[0].data(); v.hx = H[0].data();
        if (nd > 1) v.Y = X[1].data(), v.hy = H[1].data();
        if (nd > 2) v.Z = X[2].data(), v.hz = H[2].data();
        v.values = U.data(); v.gamma = 1.4; v.molarMass = 0.029; v.time = 0.125;
        WriteTecplotSolution(dir + "/sol" + std::to_string(nd) + ".dat", v);
        WriteTecplotSlice(dir + "/s2d" + std::to_string(nd) + ".dat", v, -1, nd > 1 ? 2 : -1, -1, false);
        WriteTecplotSlice(dir + "/s1d" + std::to_string(nd) + ".dat", v, 3, -1, nd > 2 ? 1 : -1, true);
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def probe():
    d = tempfile.mkdtemp()
    src = os.path.join(d, "probe.cpp")
    exe = os.path.join(d, "probe")
    open(src, "w").write(PROBE)
    pkg = os.path.join(ROOT, "turbostructred_amd")  # the sensor classes' virtuals reference the ABI
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", HOST, src, "-o", exe, "-L", pkg, "-lts_fvm",
                    f"-Wl,-rpath,{pkg}"], check=True)
    out = subprocess.run([exe, d], check=True, capture_output=True, text=True).stdout
    rows = [line.split() for line in out.split("\n") if line]
    return d, rows


def _legendre_rule(n):
    """Gauss-Legendre nodes and weights to 30 digits from mpmath (roots of mpmath's P_n found by
    findroot from numpy's nodes, weights 2 / ((1 - x^2) P_n'(x)^2)), rounded to double"""
    import mpmath as mp
    with mp.workdps(30):
        xs = [mp.findroot(lambda t: mp.legendre(n, t), mp.mpf(x0)) for x0 in np.polynomial.legendre.leggauss(n)[0]]
        ws = [2 / ((1 - x * x) * mp.diff(lambda t: mp.legendre(n, t), x) ** 2) for x in xs]
        return np.array([float(x) for x in xs]), np.array([float(w) for w in ws])


def test_gauss_rules_match_the_textbook_rule(probe):
    """nodes and weights of every order the reference tabulates (NumericQuadrature.h:130-196)
    against the Gauss-Legendre rule evaluated to 30 digits with mpmath (to an ulp); order 10 keeps
    the reference's fall-through into the 12-point half table (its weights sum to 2 - 2 w_1 of the
    12-point rule)"""
    _, rows = probe
    for r in rows:
        if r[0] != "Q":
            continue
        n = int(r[1])
        vals = np.array(r[2:], dtype=float).reshape(n, 2)
        x, w = vals[:, 0], vals[:, 1]
        if n != 10:
            xr, wr = _legendre_rule(n)
            np.testing.assert_allclose(x, xr, rtol=2.3e-16, atol=1e-17)
            np.testing.assert_allclose(w, wr, rtol=2.3e-16)
        else:
            x12, w12 = _legendre_rule(12)
            r12, v12 = x12[6:], w12[6:]  # nonnegative roots ascending and their weights
            # the mirror fill of a 6-entry half table into 10 slots (NumericQuadrature.h:204-211)
            wx = [v12[5], v12[4], v12[3], v12[2], v12[0], v12[0], v12[2], v12[3], v12[4], v12[5]]
            xx = [-r12[5], -r12[4], -r12[3], -r12[2], r12[0], -r12[0], r12[2], r12[3], r12[4], r12[5]]
            np.testing.assert_array_equal(w, wx)
            np.testing.assert_array_equal(x, xx)
            assert abs(w.sum() - (2 - 2 * w12[7])) < 1e-15
    assert ["BAD", "refused"] in rows


def test_cell_averages_closed_forms(probe):
    """1-D: the n-point rule integrates x^p exactly for p <= 2n - 1; order 10 scales every
    average by (1 - w_1)^d of the 12-point rule.  2-D/3-D: the diagonal sampling of IntegrateGroap
    (y and z take the x index) gives avg(x y) = xc yc + hx hy / 12 and avg(y^2) = yc^2 + hy^2/12
    for n >= 2, and constants and linear functions exactly."""
    _, rows = probe
    xc, yc, zc, hx, hy, hz = 0.3, -0.7, 1.1, 0.25, 0.5, 0.125
    w1_12 = np.polynomial.legendre.leggauss(12)[1][7]
    a, b = xc - hx / 2, xc + hx / 2
    for r in rows:
        if r[0] == "A1":
            n, p, avg = int(r[1]), int(r[2]), float(r[3])
            exact = (b ** (p + 1) - a ** (p + 1)) / ((p + 1) * hx)
            if n == 10:
                # the quirky rule is symmetric, so odd moments of the node set stay exact up to
                # the missing weight; only the constant is checked exactly
                if p == 0:
                    assert avg == pytest.approx(1 - w1_12, rel=1e-14)
            elif p <= 2 * n - 1:
                assert avg == pytest.approx(exact, rel=1e-13), (n, p)
        elif r[0] in ("A2", "A3"):
            n = int(r[1])
            one, xy, y, zx, yy = map(float, r[2:])
            d = 2 if r[0] == "A2" else 3
            # per axis the order-10 weights sum to 2 - 2 w_1 against a Jacobian of h/2
            scale = (1 - w1_12) ** d if n == 10 else 1.0
            assert one == pytest.approx(scale, rel=1e-14)
            if n in (1, 10):
                continue
            assert y == pytest.approx(yc, rel=1e-14)
            assert xy == pytest.approx(xc * yc + hx * hy / 12, rel=1e-13)
            assert yy == pytest.approx(yc * yc + hy * hy / 12, rel=1e-13)
            if r[0] == "A3":
                assert zx == pytest.approx(zc * xc + hz * hx / 12, rel=1e-13)


def test_ideal_gas_eos_identities(probe):
    """p = (g-1) rho e; c^2 = dp/drho + p/rho^2 dp/de = g p / rho; e(rho, p(rho, e)) = e"""
    _, rows = probe
    n = 0
    for r in rows:
        if r[0] != "E":
            continue
        ro, e, p, c, e2, dpde, dpdr = map(float, r[1:])
        assert p == pytest.approx(0.4 * ro * e, rel=1e-15)
        assert c * c == pytest.approx(dpdr + p / ro**2 * dpde, rel=1e-14)
        assert e2 == pytest.approx(e, rel=1e-15)
        n += 1
    assert n == 6


def _state(nd):
    n = [4, 3 if nd > 1 else 1, 2 if nd > 2 else 1]
    na = [n[0] + 2, n[1] + 2 if nd > 1 else 1, n[2] + 2 if nd > 2 else 1]
    U = np.zeros((na[2], na[1], na[0], 5))
    for k in range(na[2]):
        for j in range(na[1]):
            for i in range(na[0]):
                ro = 1.0 + 0.1 * i + 0.01 * j + 0.001 * k
                u = [ro * 0.3 * (i - 2), ro * 0.2 * j, -ro * 0.1 * k]
                U[k, j, i] = [ro, *u, 2.5 + 0.05 * i + 0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2) / ro]
    axes = []
    for a in range(3):
        if a < nd:
            h = (1.0 + a) / n[a]
            axes.append((np.array([(-h + 0.5 * h) + h * q for q in range(n[a] + 2)]), np.full(n[a] + 2, h)))
        else:
            axes.append((np.zeros(1), np.ones(1)))
    return n, U, axes


def _e(x):
    return "%e" % x


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_tecplot_solution_layout(probe, nd):
    """SaveSolutionToTecplot: the 1-D point list, and the 2-D/3-D FEBLOCK zone -- header, node
    coordinates at the cells' minus faces, seven cell-centred blocks, connectivity from 1 -- as
    kernel.h:1081-1416 writes them with one rank"""
    d, _ = probe
    got = open(os.path.join(d, f"sol{nd}.dat")).read()
    n, U, axes = _state(nd)
    g, cond = 1.4, 0.4 * 0.029 / 8.3144598
    L = []
    if nd == 1:
        L.append('VARIABLES = "X" "ro" "u" "v" "w" "P" "e" ')
        for i in range(1, n[0] + 1):
            ro, m1, m2, m3, E = U[0, 0, i]
            u, v, w = m1 / ro, m2 / ro, m3 / ro
            e = E / ro - 0.5 * (u * u + v * v + w * w)
            L.append(" ".join(_e(q) for q in (axes[0][0][i], ro, u, v, w, 0.4 * ro * e, e)) + " ")
    else:
        names = '"X" "Y" ' if nd == 2 else '"X" "Y" "Z" '
        L.append("VARIABLES = " + names + '"Rho" "u" "v" "w" "P" "e""T"')
        L.append('ZONE T="%e"' % 0.125)
        nodes = (n[0] + 1) * (n[1] + 1) * (n[2] + 1 if nd > 2 else 1)
        L.append(f"N={nodes}, E={n[0] * n[1] * (n[2] if nd > 2 else 1)}, F=FEBLOCK, ET=" +
                 ("QUADRILATERAL" if nd == 2 else "BRICK"))
        L.append("VARLOCATION = (" + ", ".join(["NODAL"] * nd + ["CELLCENTERED"] * 7) + ")")
        ks = range(1, n[2] + 2) if nd > 2 else [0]
        for a in range(nd):
            for k in ks:
                for j in range(1, n[1] + 2):
                    for i in range(1, n[0] + 2):
                        idx = (i, j, k)[a]
                        L.append(_e(axes[a][0][idx] - 0.5 * axes[a][1][idx]))
        kc = range(1, n[2] + 1) if nd > 2 else [0]
        cells = [U[k, j, i] for k in kc for j in range(1, n[1] + 1) for i in range(1, n[0] + 1)]
        cols = [[] for _ in range(7)]
        for c in cells:
            ro, m1, m2, m3, E = c
            if nd == 2:
                u, v, w = m1 / ro, m2 / ro, m3 / ro
                e = E / ro
                e = e - 0.5 * (u * u + v * v + w * w)
                P = 0.4 * (ro * e)
            else:
                ke = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / ro
                u, v, w = m1 / ro, m2 / ro, m3 / ro
                e = E - ke
                P = 0.4 * e
                e = e / ro
            for q, val in enumerate((ro, u, v, w, P, e, cond * e)):
                cols[q].append(val)
        for col in cols:
            L.extend(_e(x) for x in col)
        nx1, ny1 = n[0] + 1, n[1] + 1
        for k in (range(n[2]) if nd > 2 else [0]):
            for j in range(n[1]):
                for i in range(n[0]):
                    bl = nx1 * ny1 * k + nx1 * j + i + 1
                    if nd == 2:
                        L.append(f"{bl} {bl + nx1} {bl + nx1 + 1} {bl + 1}")
                    else:
                        bb = bl + nx1 * ny1
                        L.append(f"{bb + 1} {bb + nx1 + 1} {bb + nx1} {bb} {bl + 1} {bl + nx1 + 1} {bl + nx1} {bl}")
    assert got == "\n".join(L) + "\n"


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_tecplot_slice_layouts(probe, nd):
    """Save2DSliceToTecplot (no temperature) and SaveSliceToTecplot (temperature and the fixed
    coordinates) in the POINT layout of kernel.h:1416-1632"""
    d, _ = probe
    n, U, axes = _state(nd)
    for name, (I, J, K), temp in ((f"s2d{nd}.dat", (-1, 2 if nd > 1 else -1, -1), False),
                                  (f"s1d{nd}.dat", (3, -1, 1 if nd > 2 else -1), True)):
        got = open(os.path.join(d, name)).read()
        head = "VARIABLES = " + "".join(f'"{c}" ' for c, v in zip("XYZ", (I, J, K)) if v == -1)
        head += '"ro" "u" "v" "w" "P" "e" '
        if temp:
            head += '"T" ' + "".join(f'"{c}_cnst" ' for c, v in zip("xyz", (I, J, K)) if v != -1)
        L = [head, 'ZONE T="1"',
             f"I={n[0] if I == -1 else 1}J={n[1] if J == -1 else 1}K={n[2] if K == -1 else 1}F=POINT",
             "DT=(" + " ".join(["SINGLE"] * (12 if temp else 9)) + ")"]
        jr = range(1, n[1] + 1) if nd > 1 else [0]
        kr = range(1, n[2] + 1) if nd > 2 else [0]
        for i in range(1, n[0] + 1):
            if I != -1 and i != I:
                continue
            for j in jr:
                if J != -1 and j != J:
                    continue
                for k in kr:
                    if K != -1 and k != K:
                        continue
                    pos = (axes[0][0][i], axes[1][0][j] if nd > 1 else 0.0, axes[2][0][k] if nd > 2 else 0.0)
                    ro, m1, m2, m3, E = U[k, j, i]
                    u, v, w = m1 / ro, m2 / ro, m3 / ro
                    e = E / ro - 0.5 * (u * u + v * v + w * w)
                    vals = [p for p, q in zip(pos, (I, J, K)) if q == -1] + [ro, u, v, w, 0.4 * ro * e, e]
                    if temp:
                        vals += [0.4 * 0.029 * e / 8.3144598] + [p for p, q in zip(pos, (I, J, K)) if q != -1]
                    L.append(" ".join(_e(x) for x in vals) + " ")
        assert got == "\n".join(L) + "\n", name
