/*
 * ts_oracle.cpp -- TEST INFRASTRUCTURE ONLY: the CPU FP64 restatement of TurboStructured's
 * explicit finite-volume step, used to check the HIP library.  It is never linked into, or
 * called by, libts_fvm.so.
 *
 * It keeps the reference's data layout (AoS, ghost layers, serial index of grid.h:64-67), its
 * per-face accumulation order and its dt rule, written as plain loops without the reference's
 * per-face heap allocations.  Citations are to files under the reference's src/ directory.
 *
 * Parity status per function (how each is pinned, see tests/test_oracle_*.py):
 *   roe_flux           consistency, rotation, upwinding, Rankine-Hugoniot, dt0 closed form,
 *                      Sod conservation invariants; Harten's entropy fix (harten) by the closed
 *                      form of a resting and a slowly moving contact at Eps = 0.05 and 0.2
 *   exact star state   Toro's star table held by the reference (ToroTests.h:148-191),
 *                      f(p*) = 0 residual, vacuum closed form
 *   hllc_flux          consistency, contact preservation, supersonic upwinding, rotation,
 *                      Toro tests 1-5: positive density and pressure, L1 density error falling
 *                      at a first-order rate (extension: no reference)
 *   reconstructions    linear exactness, ENO smallest-slope rule, limiter bounds, BJ / Venkat
 *                      closed forms on 3-cell stencils
 *   boundary images    per-type closed forms (BCGeneral.h:186-268)
 *   whole step         Sod invariants (mass/energy conservation, momentum growth 0.9*dt),
 *                      first-order convergence to the exact solution, direction invariance
 *                      x/y/z, periodic conservation, diagonal symmetry, second-order
 *                      convergence on the smooth Liska-Wendroff solution
 *                      (ExactEulerSolution.h:30-44; orders 2.5 and 2.3 at 16/32/64)
 *   viscous terms      least squares vs numpy SVD and the reference's Eigen JacobiSVD, wall
 *                      dummy gradients, Couette and stress-divergence closed forms, viscous SR,
 *                      force/acceleration sources, Poiseuille centre velocity
 *   residual norms     first Sod step: norm = 2 h |F*| of the diaphragm face (kernel.h:326)
 * The exact star state is pinned against the mathematics (f(p*) = 0 to round-off) and Toro's
 * 5-digit table; the reference's own p* comes from ALGLIB's tolerance-limited LM
 * (GodunovSolverPerfectGasEOS.h:157-185), so agreement with the reference itself below ~1e-5
 * relative is parity unpinned.
 * The compiled reference cannot be built in this image without stand-ins for its MSVC MPI
 * header, so no trajectory of the reference itself pins the whole step.
 */
#include "ts_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

constexpr int NV = 5;

/* ---------------------------------------------------------------------------------------- */
/* Riemann solvers                                                                          */
/* ---------------------------------------------------------------------------------------- */

/* Physical flux F(U).n  (RoeSolverPerfectGasEOS.h:15-32) */
void euler_flux(double g, const double* U, const double* n, double* F) {
    double ro = U[0];
    double vx = U[1] / ro, vy = U[2] / ro, vz = U[3] / ro;
    double roE = U[4];
    double p = (g - 1.0) * (roE - ro * (vx * vx + vy * vy + vz * vz) / 2.0);
    double vn = vx * n[0] + vy * n[1] + vz * n[2];
    F[0] = ro * vn;
    F[1] = ro * vn * vx + n[0] * p;
    F[2] = ro * vn * vy + n[1] * p;
    F[3] = ro * vn * vz + n[2] * p;
    F[4] = vn * (roE + p);
}

double harten(double z, double e) { /* RoeSolverPerfectGasEOS.h:35-40 */
    z = std::fabs(z);
    if (z < e) z = ((z * z) / e + e) * 0.5;
    return z;
}

/* Roe-Pike flux with Harten's fix (RoeSolverPerfectGasEOS.h:48-134) */
void roe_flux(double g, double eps, const double* UL, const double* UR, const double* fn,
              double* F, double* lam) {
    double FL[NV], FR[NV];
    euler_flux(g, UL, fn, FL);
    euler_flux(g, UR, fn, FR);
    for (int i = 0; i < NV; i++) F[i] = FL[i] + FR[i];
    double ro_l = UL[0], ro_r = UR[0];
    double vl[3] = {UL[1] / ro_l, UL[2] / ro_l, UL[3] / ro_l};
    double vr[3] = {UR[1] / ro_r, UR[2] / ro_r, UR[3] / ro_r};
    double e_l = UL[4] / ro_l, e_r = UR[4] / ro_r;
    double k = 0.5 * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]);
    double h_l = (e_l - k) * g + k;
    k = 0.5 * (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
    double h_r = (e_r - k) * g + k;
    double sl = std::sqrt(ro_l), sr = std::sqrt(ro_r);
    double ql = sl / (sl + sr);
    double qr = sr / (sl + sr);
    double v[3] = {ql * vl[0] + qr * vr[0], ql * vl[1] + qr * vr[1], ql * vl[2] + qr * vr[2]};
    double h = ql * h_l + qr * h_r;
    double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double phi2 = 0.5 * (g - 1) * vv;
    double dn = 1.0;
    double c = std::sqrt((g - 1) * h - phi2);
    double uw = v[0] * fn[0] + v[1] * fn[1] + v[2] * fn[2];
    double eig_max = std::fabs(uw) + c * dn;
    double AA1 = harten(uw, eps * eig_max);
    double AA3 = harten(uw + c * dn, eps * eig_max);
    double AA4 = harten(uw - c * dn, eps * eig_max);
    double Eig1 = AA1;
    double Eiga = (AA3 - AA4) * 0.5 / (dn * c);
    double Eigb = (AA3 + AA4) * 0.5 - AA1;
    double Qa[NV] = {0, fn[0], fn[1], fn[2], uw};
    double Qb[NV] = {1, v[0], v[1], v[2], h};
    double vvr = v[0] * vr[0] + v[1] * vr[1] + v[2] * vr[2];
    double vvl = v[0] * vl[0] + v[1] * vl[1] + v[2] * vl[2];
    double R1 = phi2 * ro_r - (g - 1) * ro_r * (vvr - e_r);
    double D1 = (ro_r * fn[0]) * (vr[0] - v[0]) + (ro_r * fn[1]) * (vr[1] - v[1]) +
                (ro_r * fn[2]) * (vr[2] - v[2]);
    double R2 = phi2 * ro_l - (g - 1) * ro_l * (vvl - e_l);
    double D2 = (ro_l * fn[0]) * (vl[0] - v[0]) + (ro_l * fn[1]) * (vl[1] - v[1]) +
                (ro_l * fn[2]) * (vl[2] - v[2]);
    double C2 = 1.0 / (c * c);
    double DN2 = 1.0 / (dn * dn);
    for (int i = 0; i < NV; i++) {
        F[i] += (Eig1 * UL[i] + Eiga * (R2 * Qa[i] + D2 * Qb[i]) +
                 Eigb * (R2 * Qb[i] * C2 + D2 * Qa[i] * DN2) - Eig1 * UR[i] -
                 Eiga * (R1 * Qa[i] + D1 * Qb[i]) - Eigb * (R1 * Qb[i] * C2 + D1 * Qa[i] * DN2));
        F[i] *= 0.5;
    }
    *lam = eig_max;
}

/* HLLC (Toro, Riemann Solvers, 3rd ed., sec. 10.4, eqs. 10.37-10.39) with the Davis wave
   speed bounds (eq. 10.48).  Extension: the reference has no HLLC.  Its dt eigenvalue is
   max(|S_L|, |S_R|). */
void hllc_flux(double g, const double* UL, const double* UR, const double* fn, double* F,
               double* lam) {
    int d = fn[0] != 0.0 ? 0 : (fn[1] != 0.0 ? 1 : 2);
    double rl = UL[0], rr = UR[0];
    double il = 1.0 / rl, ir = 1.0 / rr;
    double vl[3] = {UL[1] * il, UL[2] * il, UL[3] * il};
    double vr[3] = {UR[1] * ir, UR[2] * ir, UR[3] * ir};
    double pl = (g - 1.0) * (UL[4] - 0.5 * rl * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]));
    double pr = (g - 1.0) * (UR[4] - 0.5 * rr * (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]));
    double al = std::sqrt(g * pl * il), ar = std::sqrt(g * pr * ir);
    double ul = vl[d], ur = vr[d];
    double SL = std::min(ul - al, ur - ar);
    double SR = std::max(ul + al, ur + ar);
    double ml = rl * (SL - ul), mr = rr * (SR - ur);
    double Ss = (pr - pl + ul * ml - ur * mr) / (ml - mr);
    *lam = std::max(std::fabs(SL), std::fabs(SR));
    /* choose the side of the contact the face sits on */
    bool left = Ss >= 0.0;
    double ik = left ? il : ir;
    const double* U = left ? UL : UR;
    const double* vel = left ? vl : vr;
    double S = left ? SL : SR;
    double un = left ? ul : ur, p = left ? pl : pr, r = left ? rl : rr;
    double Fk[NV];
    Fk[0] = r * un;
    Fk[1] = Fk[0] * vel[0];
    Fk[2] = Fk[0] * vel[1];
    Fk[3] = Fk[0] * vel[2];
    Fk[1 + d] += p;
    Fk[4] = un * (U[4] + p);
    bool supersonic = left ? (S >= 0.0) : (S <= 0.0);
    if (supersonic) {
        for (int i = 0; i < NV; i++) F[i] = Fk[i];
        return;
    }
    double m = r * (S - un);
    double coef = m / (S - Ss);
    double Us[NV];
    Us[0] = coef;
    Us[1] = coef * vel[0];
    Us[2] = coef * vel[1];
    Us[3] = coef * vel[2];
    Us[1 + d] = coef * Ss;
    Us[4] = coef * (U[4] * ik + (Ss - un) * (Ss + p / m));
    for (int i = 0; i < NV; i++) F[i] = Fk[i] + S * (Us[i] - U[i]);
}

/* Exact Riemann solver: star state by Newton-Raphson on f(p) = fL(p) + fR(p) + du
   (Toro sec. 4.3, initial guess of Toro's adaptive scheme sec. 4.3.2).  The reference obtains
   p* with ALGLIB's bounded Levenberg-Marquardt instead (GodunovSolverPerfectGasEOS.h:143-185,
   stopping at epsg=1e-10), so the two agree to about 1e-10 relative, not to round-off. */
struct Star {
    double p, u, rl, rr;
    bool shock_l, shock_r, vacuum;
    double SL, SHL, STL, SR, SHR, STR, smax;
    int iters;
};

double fk(double g, double r, double pk, double ak, double p, double* dfdp) {
    if (p > pk) { /* shock, GodunovSolverPerfectGasEOS.h:62-67 */
        double A = 2 / ((g + 1.0) * r);
        double B = pk * (g - 1.0) / (g + 1.0);
        double q = std::sqrt(A / (p + B));
        *dfdp = q * (1.0 - 0.5 * (p - pk) / (B + p));
        return (p - pk) * q;
    }
    /* rarefaction, GodunovSolverPerfectGasEOS.h:68-73 */
    double pr = p / pk;
    double res = std::pow(pr, (g - 1.0) / (2 * g)) - 1.0;
    res *= 2 * ak / (g - 1.0);
    *dfdp = std::pow(pr, -(g + 1.0) / (2 * g)) / (r * ak);
    return res;
}

Star star_state(double g, double rl, double ul, double pl, double rr, double ur, double pr) {
    Star s{};
    double al = std::sqrt(g * pl / rl), ar = std::sqrt(g * pr / rr);
    double du = ur - ul;
    s.vacuum = (2.0 / (g - 1.0)) * (al + ar) <= du;
    double p;
    int it = 0;
    if (s.vacuum) {
        p = 0.0;
    } else {
        /* initial guess (Toro sec. 9.5.1 / 4.3.2) */
        double pmin = std::min(pl, pr), pmax = std::max(pl, pr);
        double ppv = 0.5 * (pl + pr) - 0.125 * du * (rl + rr) * (al + ar);
        ppv = std::max(0.0, ppv);
        if (pmax / pmin <= 2.0 && pmin <= ppv && ppv <= pmax) {
            p = ppv;
        } else if (ppv < pmin) {
            double z = (g - 1.0) / (2.0 * g);
            double plz = std::pow(pl, z), prz = std::pow(pr, z);
            double t = (al + ar - 0.5 * (g - 1.0) * du) / (al / plz + ar / prz);
            p = std::pow(t, 1.0 / z);
        } else {
            double A = 2.0 / ((g + 1.0) * rl), B = (g - 1.0) / (g + 1.0) * pl;
            double gl = std::sqrt(A / (ppv + B));
            A = 2.0 / ((g + 1.0) * rr);
            B = (g - 1.0) / (g + 1.0) * pr;
            double gr = std::sqrt(A / (ppv + B));
            p = (gl * pl + gr * pr - du) / (gl + gr);
        }
        if (!(p > 0.0)) p = 1e-12 * pmin;
        for (it = 1; it <= 60; it++) {
            double dl, dr;
            double f = fk(g, rl, pl, al, p, &dl) + fk(g, rr, pr, ar, p, &dr) + du;
            double pn = p - f / (dl + dr);
            if (!(pn > 0.0)) pn = 0.01 * p;
            double ch = std::fabs(pn - p) / (pn + p);
            p = pn;
            if (ch <= 1e-15) break;
        }
    }
    s.iters = it;
    s.p = p;
    double d1, d2;
    double f1 = fk(g, rl, pl, al, p, &d1), f2 = fk(g, rr, pr, ar, p, &d2);
    /* GodunovSolverPerfectGasEOS.h:191-265 */
    s.u = 0.5 * (ul + ur) + 0.5 * (f2 - f1);
    double C = (g - 1.0) / (g + 1.0);
    s.smax = 0;
    double prl = p / pl;
    if (p > pl) {
        s.shock_l = true;
        s.rl = (prl + C) / (C * prl + 1.0) * rl;
        s.SL = ul - al * std::sqrt((g + 1) * prl / (2 * g) + (g - 1) / (2 * g));
        s.smax = std::max(s.smax, std::fabs(s.SL));
    } else {
        s.shock_l = false;
        s.rl = rl * std::pow(prl, 1.0 / g);
        s.SHL = ul - al;
        s.STL = s.u - al * std::pow(prl, (g - 1) / (2 * g));
        s.smax = std::max(s.smax, std::fabs(s.SHL));
    }
    double prr = p / pr;
    if (p > pr) {
        s.shock_r = true;
        s.rr = (prr + C) / (C * prr + 1.0) * rr;
        s.SR = ur + ar * std::sqrt((g + 1) * prr / (2 * g) + (g - 1) / (2 * g));
        s.smax = std::max(s.smax, std::fabs(s.SR));
    } else {
        s.shock_r = false;
        s.rr = rr * std::pow(prr, 1.0 / g);
        s.SHR = ur + ar;
        s.STR = s.u + ar * std::pow(prr, (g - 1) / (2 * g));
        s.smax = std::max(s.smax, std::fabs(s.SHR));
    }
    if (s.vacuum) { /* tails of the two fans bound the vacuum (Toro sec. 4.6) */
        s.STL = ul + 2.0 * al / (g - 1.0);
        s.STR = ur - 2.0 * ar / (g - 1.0);
    }
    return s;
}

/* Sample the exact solution at x/t = S (GodunovSolverPerfectGasEOS.h:299-413; the vacuum
   branch follows Toro sec. 4.6 where the reference would divide 0 by 0). */
void exact_sample(double g, double rl, double ul, double pl, double rr, double ur, double pr,
                  const Star& st, double S, double* ro, double* u, double* p, bool* left_side) {
    bool left = st.vacuum ? (S <= 0.5 * (st.STL + st.STR)) : (st.u >= S);
    *left_side = left;
    if (left) {
        if (st.shock_l && !st.vacuum) {
            if (st.SL >= S) { *ro = rl; *u = ul; *p = pl; }
            else { *ro = st.rl; *u = st.u; *p = st.p; }
        } else {
            if (st.SHL > S) { *ro = rl; *u = ul; *p = pl; }
            else if (S > st.STL) {
                if (st.vacuum) { *ro = 0; *u = 0; *p = 0; }
                else { *ro = st.rl; *u = st.u; *p = st.p; }
            } else {
                double al = std::sqrt(g * pl / rl);
                double CL = 2.0 / (g + 1) + (g - 1) * (ul - S) / ((g + 1) * al);
                *ro = std::pow(CL, 2.0 / (g - 1)) * rl;
                double uu = al + 0.5 * (g - 1) * ul + S;
                uu *= 2 / (g + 1);
                *u = uu;
                *p = std::pow(CL, 2.0 * g / (g - 1)) * pl;
            }
        }
    } else {
        if (st.shock_r && !st.vacuum) {
            if (st.SR <= S) { *ro = rr; *u = ur; *p = pr; }
            else { *ro = st.rr; *u = st.u; *p = st.p; }
        } else {
            if (st.SHR < S) { *ro = rr; *u = ur; *p = pr; }
            else if (S < st.STR) {
                if (st.vacuum) { *ro = 0; *u = 0; *p = 0; }
                else { *ro = st.rr; *u = st.u; *p = st.p; }
            } else {
                double ar = std::sqrt(g * pr / rr);
                double CR = 2.0 / (g + 1) - (g - 1) * (ur - S) / ((g + 1) * ar);
                *ro = std::pow(CR, 2.0 / (g - 1)) * rr;
                double uu = -ar + 0.5 * (g - 1) * ur + S;
                uu *= 2 / (g + 1);
                *u = uu;
                *p = std::pow(CR, 2.0 * g / (g - 1)) * pr;
            }
        }
    }
}

/* Exact Godunov flux (GodunovSolverPerfectGasEOS.h:276-433) */
void godunov_flux(double g, const double* UL, const double* UR, const double* fn, double* F,
                  double* lam) {
    double roL = UL[0], roR = UR[0];
    auto pres = [g](const double* c) {
        double ro = c[0];
        double vx = c[1] / ro, vy = c[2] / ro, vz = c[3] / ro;
        double E = c[4] / ro;
        return (g - 1.0) * ro * (E - (vx * vx + vy * vy + vz * vz) / 2.0);
    };
    double pL = pres(UL), pR = pres(UR);
    double velL[3] = {UL[1] / UL[0], UL[2] / UL[0], UL[3] / UL[0]};
    double velR[3] = {UR[1] / UR[0], UR[2] / UR[0], UR[3] / UR[0]};
    double uL = velL[0] * fn[0] + velL[1] * fn[1] + velL[2] * fn[2];
    double uR = velR[0] * fn[0] + velR[1] * fn[1] + velR[2] * fn[2];
    Star st = star_state(g, roL, uL, pL, roR, uR, pR);
    double ro, u, p;
    bool left;
    exact_sample(g, roL, uL, pL, roR, uR, pR, st, 0.0, &ro, &u, &p, &left);
    const double* vel = left ? velL : velR;
    double vn = vel[0] * fn[0] + vel[1] * fn[1] + vel[2] * fn[2];
    double Vf[3] = {vel[0] + (u - vn) * fn[0], vel[1] + (u - vn) * fn[1],
                    vel[2] + (u - vn) * fn[2]};
    double Uf[NV];
    if (ro == 0.0) {
        for (int i = 0; i < NV; i++) F[i] = 0.0;
    } else {
        Uf[0] = ro;
        Uf[1] = ro * Vf[0];
        Uf[2] = ro * Vf[1];
        Uf[3] = ro * Vf[2];
        Uf[4] = ro * (p / (ro * (g - 1.0)) + 0.5 * (Vf[0] * Vf[0] + Vf[1] * Vf[1] + Vf[2] * Vf[2]));
        euler_flux(g, Uf, fn, F);
    }
    *lam = st.smax;
}

void flux(int riemann, double g, double eps, const double* UL, const double* UR,
          const double* fn, double* F, double* lam) {
    if (riemann == TS_RIEMANN_ROE) roe_flux(g, eps, UL, UR, fn, F, lam);
    else if (riemann == TS_RIEMANN_HLLC) hllc_flux(g, UL, UR, fn, F, lam);
    else godunov_flux(g, UL, UR, fn, F, lam);
}

/* ---------------------------------------------------------------------------------------- */
/* Boundary images (BCGeneral.h:186-268)                                                    */
/* ---------------------------------------------------------------------------------------- */
void bc_image(const ts_bc_config& bc, double g, const double* U, const double* n, double* out) {
    switch (bc.type) {
    case TS_BC_SYMMETRY: {
        double dot = n[0] * U[1] + n[1] * U[2] + n[2] * U[3];
        for (int i = 0; i < NV; i++) out[i] = U[i];
        out[1] -= 2.0 * (dot * n[0]);
        out[2] -= 2.0 * (dot * n[1]);
        out[3] -= 2.0 * (dot * n[2]);
        break;
    }
    case TS_BC_WALL:
        out[0] = U[0];
        out[1] = U[1] * (-1);
        out[2] = U[2] * (-1);
        out[3] = U[3] * (-1);
        out[4] = U[4];
        break;
    case TS_BC_MOVING_WALL: {
        double rho = U[0];
        double vin[3] = {U[1] / rho, U[2] / rho, U[3] / rho};
        double rhoe = U[4] - 0.5 * rho * (vin[0] * vin[0] + vin[1] * vin[1] + vin[2] * vin[2]);
        double vd[3] = {2.0 * bc.velocity[0] - vin[0], 2.0 * bc.velocity[1] - vin[1],
                        2.0 * bc.velocity[2] - vin[2]};
        out[0] = rho;
        out[1] = rho * vd[0];
        out[2] = rho * vd[1];
        out[3] = rho * vd[2];
        out[4] = rhoe + 0.5 * rho * (vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
        break;
    }
    case TS_BC_SUBSONIC_INLET: {
        const double* V = bc.velocity;
        out[0] = bc.density;
        out[1] = bc.density * V[0];
        out[2] = bc.density * V[1];
        out[3] = bc.density * V[2];
        out[4] = bc.pstatic / (g - 1) + 0.5 * bc.density * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
        break;
    }
    default: /* Natural, SubsonicOutlet */
        for (int i = 0; i < NV; i++) out[i] = U[i];
    }
}

/* Least-squares gradient at `point` from the value there and at m stencil points
   (GradientComputer.h:19-53): row i is point - p_i (ndims columns), right-hand side
   value - value_i.  The reference solves it with Eigen's JacobiSVD; for the full-rank stencils
   used here that is the unique least-squares solution, found here through the normal equations
   and a Cholesky factorisation. */
void lsq_gradient(int nd, const double* point, double value, const double (*pts)[3], const double* vals,
                  int m, double* grad) {
    double M[3][3] = {{0}}, b[3] = {0, 0, 0};
    for (int i = 0; i < m; i++) {
        double r[3];
        for (int a = 0; a < nd; a++) r[a] = point[a] - pts[i][a];
        double rhs = value - vals[i];
        for (int a = 0; a < nd; a++) {
            for (int c = 0; c < nd; c++) M[a][c] += r[a] * r[c];
            b[a] += r[a] * rhs;
        }
    }
    double L[3][3] = {{0}};
    for (int a = 0; a < nd; a++) {
        for (int c = 0; c <= a; c++) {
            double s = M[a][c];
            for (int q = 0; q < c; q++) s -= L[a][q] * L[c][q];
            L[a][c] = (a == c) ? std::sqrt(s) : s / L[c][c];
        }
    }
    double y[3];
    for (int a = 0; a < nd; a++) {
        double s = b[a];
        for (int q = 0; q < a; q++) s -= L[a][q] * y[q];
        y[a] = s / L[a][a];
    }
    for (int a = nd - 1; a >= 0; a--) {
        double s = y[a];
        for (int q = a + 1; q < nd; q++) s -= L[q][a] * grad[q];
        grad[a] = s / L[a][a];
    }
    for (int a = nd; a < 3; a++) grad[a] = 0.0;
}

/* ---------------------------------------------------------------------------------------- */
/* Reconstruction objects (src/Reconstruction, utility/Limiters.h)                           */
/* ---------------------------------------------------------------------------------------- */
struct Rec {
    double v[NV];
    double g[NV][3];
    double lim[NV];
};

/* value at the offset (px,py,pz) from the centre: SampleSolution.  With `positive` set (the
   HLLC and exact-Godunov fluxes, which need a real sound speed) a face state with rho <= 0 or
   2 rho E <= |m|^2 falls back to the cell value; the reference's Roe path never does this. */
inline void sample(int recon, const Rec& r, const double* pt, double* out, bool positive) {
    for (int i = 0; i < NV; i++) {
        double dot = r.g[i][0] * pt[0] + r.g[i][1] * pt[1] + r.g[i][2] * pt[2];
        if (recon == TS_RECON_PWC) out[i] = r.v[i];
        else if (recon == TS_RECON_LINEAR_LIM) out[i] = r.v[i] + r.lim[i] * dot;
        else out[i] = r.v[i] + dot;
    }
    if (positive && recon != TS_RECON_PWC) {
        double q = 2.0 * out[0] * out[4] - (out[1] * out[1] + out[2] * out[2] + out[3] * out[3]);
        if (!(out[0] > 0.0 && q > 0.0))
            for (int i = 0; i < NV; i++) out[i] = r.v[i];
    }
}

/* nb[6][5] neighbours x-,x+,y-,y+,z-,z+; c* their centre coordinates along the axis;
   hc the cell's own sizes; eps2 the Venkatakrishnan parameter of the cell */
void reconstruct(int recon, int limiter, int nd, const double* v, const double (*nb)[NV],
                 const double* xc, const double (*nbc)[3], const double* hc, double eps2,
                 Rec& r) {
    for (int i = 0; i < NV; i++) {
        r.v[i] = v[i];
        r.g[i][0] = r.g[i][1] = r.g[i][2] = 0.0;
        r.lim[i] = 1.0;
    }
    if (recon == TS_RECON_PWC) return;
    for (int a = 0; a < nd; a++) {
        const double* m = nb[2 * a];
        const double* p = nb[2 * a + 1];
        double xm = nbc[2 * a][a], xp = nbc[2 * a + 1][a], x0 = xc[a];
        for (int i = 0; i < NV; i++) {
            if (recon == TS_RECON_ENO2) { /* ENO2PointsStencil.h:21-58 */
                double gl = (v[i] - m[i]) / (x0 - xm);
                double gr = (p[i] - v[i]) / (xp - x0);
                r.g[i][a] = (std::fabs(gl) < std::fabs(gr)) ? gl : gr;
            } else { /* Linear2PointsStencil.h:66-97 */
                r.g[i][a] = (p[i] - m[i]) / (xp - xm);
            }
        }
    }
    if (recon != TS_RECON_LINEAR_LIM) return;
    /* projections to the face centres (Linear2psLim.h:74-90): (v + g.p) - v */
    double proj[6][NV];
    for (int a = 0; a < nd; a++) {
        for (int s = 0; s < 2; s++) {
            double pt[3] = {0, 0, 0};
            pt[a] = (s == 0 ? -0.5 : 0.5) * hc[a];
            for (int i = 0; i < NV; i++) {
                double dot = r.g[i][0] * pt[0] + r.g[i][1] * pt[1] + r.g[i][2] * pt[2];
                double fv = r.v[i] + dot;
                proj[2 * a + s][i] = fv - r.v[i];
            }
        }
    }
    /* Limiter::ComputeLimiterValues (Limiters.h:28-67) */
    for (int i = 0; i < NV; i++) {
        double umax = v[i], umin = v[i];
        for (int n = 0; n < 2 * nd; n++) {
            double val = nb[n][i];
            if (umax < val) umax = val;
            if (umin > val) umin = val;
        }
        if (umax - umin < 3.0 * std::numeric_limits<double>::min()) continue;
        double lim_min = 1.0;
        for (int j = 0; j < 2 * nd; j++) {
            double lv = 1.0;
            double d2 = proj[j][i];
            if (d2 > 0 || d2 < 0) {
                double um = (d2 > 0) ? umax : umin;
                if (limiter == TS_LIMITER_BJ) { /* Limiters.h:80-82 */
                    lv = std::min(1.0, (um - v[i]) / d2);
                } else { /* Limiters.h:102-115 */
                    double un = 0.5 * (umax - umin);
                    double umn = um / un, uin = v[i] / un, dn = d2 / un;
                    double dm = umn - uin;
                    lv = ((dm * dm + eps2) + 2 * dn * dm) /
                         (dm * dm + 2 * dn * dn + dm * dn + eps2);
                }
            }
            if (lim_min > lv) lim_min = lv;
        }
        r.lim[i] = lim_min;
    }
}

/* ---------------------------------------------------------------------------------------- */
/* The solver                                                                               */
/* ---------------------------------------------------------------------------------------- */
template <class F>
void parallel_for(int nthreads, int n, F&& f) {
    if (nthreads <= 1 || n < 2 * nthreads) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    int chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++) {
        int b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&f, b, e] {
            for (int i = b; i < e; i++) f(i);
        });
    }
    for (auto& x : th) x.join();
}

}  // namespace

struct tso_ctx {
    ts_config cfg;
    int nd, n[3], gl[3], na[3];
    std::vector<double> C[3], H[3];
    std::vector<double> vol;
    std::vector<double> U, R, SR, U0, Ust;
    std::vector<double> G; /* velocity gradients, 9 per cell (ghosts included): [u,v,w][x,y,z] */
    std::vector<Rec> rec;
    int rl[3], rn[3]; /* recon layers and extents */
    double time = 0;
    long long iter = 0;
    int nthreads = 1;
    double K = 5.0; /* Constants::VENKATAKR_KOEFF, utility/Constants.h:11 */

    size_t idx(int i, int j, int k) const { return ((size_t)k * na[1] + j) * na[0] + i; }
    size_t ridx(int i, int j, int k) const { return ((size_t)k * rn[1] + j) * rn[0] + i; }

    int marker(int side, int a, int b) const {
        /* face (a,b): a = fast transverse index, b = slow, interior-based (0-based) */
        const uint8_t* fm = cfg.face_marker[side];
        if (fm) {
            int nfast = (side < 2) ? n[1] : n[0];
            return fm[(size_t)b * nfast + a];
        }
        return cfg.side_marker[side];
    }

    void build_grid() {
        /* grid.h:83-121 and 124-218 (uniform branch): running sums of the step */
        double L[3] = {cfg.lx, cfg.ly, cfg.lz};
        for (int a = 0; a < 3; a++) {
            C[a].assign(na[a], 0.0);
            H[a].assign(na[a], 0.0);
            if (a >= nd) {
                H[a][0] = 1.0; /* grid.h:343-344 */
                continue;
            }
            if (cfg.coord[a] && cfg.h[a]) {
                for (int i = 0; i < na[a]; i++) {
                    C[a][i] = cfg.coord[a][i];
                    H[a][i] = cfg.h[a][i];
                }
                continue;
            }
            double stp = L[a] / n[a];
            double xc = -(gl[a] * stp) + 0.5 * stp;
            int c = 0;
            for (; c < gl[a]; c++) {
                C[a][c] = xc;
                H[a][c] = stp;
                xc += stp;
            }
            for (int i = 0; i < n[a]; i++) {
                C[a][c + i] = xc;
                H[a][c + i] = stp;
                xc += 0.5 * stp * (1.0 + 1.0);
                stp *= 1.0;
            }
            c += n[a];
            for (int i = 0; i < gl[a]; i++) {
                C[a][c + i] = xc;
                H[a][c + i] = stp;
                xc += stp;
            }
        }
        vol.resize((size_t)na[0] * na[1] * na[2]);
        for (int i = 0; i < na[0]; i++)
            for (int j = 0; j < na[1]; j++)
                for (int k = 0; k < na[2]; k++) vol[idx(i, j, k)] = H[0][i] * H[1][j] * H[2][k];
    }

    /* ExchangeValues (periodic self-wrap on one rank, kernel.h:334-690) followed by
       ComputeDummyCellValues (kernel.h:693-915) */
    void fill_ghosts(std::vector<double>& V) {
        double g = cfg.gamma;
        /* transverse axes of axis a, and their interior extents */
        auto transverse = [&](int a, int& ja, int& ka, int& nj, int& nk) {
            ja = (a == 0) ? 1 : 0;
            ka = (a == 2) ? 1 : 2;
            nj = ja < nd ? n[ja] : 1;
            nk = ka < nd ? n[ka] : 1;
        };
        for (int a = 0; a < nd; a++) {
            if (!cfg.periodic[a]) continue;
            int ja, ka, nj, nk;
            transverse(a, ja, ka, nj, nk);
            int lo = gl[a], hi = gl[a] + n[a] - 1;
            for (int l = 1; l <= gl[a]; l++)
                for (int kk = 0; kk < nk; kk++)
                    for (int jj = 0; jj < nj; jj++) {
                        int D[3], S[3];
                        D[ja] = S[ja] = jj + gl[ja];
                        D[ka] = S[ka] = kk + gl[ka];
                        D[a] = hi + l; S[a] = lo + l - 1;  /* minus-direction exchange */
                        for (int v = 0; v < NV; v++)
                            V[idx(D[0], D[1], D[2]) * NV + v] = V[idx(S[0], S[1], S[2]) * NV + v];
                        D[a] = lo - l; S[a] = hi - l + 1;  /* plus-direction exchange */
                        for (int v = 0; v < NV; v++)
                            V[idx(D[0], D[1], D[2]) * NV + v] = V[idx(S[0], S[1], S[2]) * NV + v];
                    }
        }
        for (int a = 0; a < nd; a++) {
            if (cfg.periodic[a]) continue;
            double nl[3] = {0, 0, 0};
            nl[a] = -1.0; /* faceNormalL is used on both sides (kernel.h:743,760) */
            int ja, ka, nj, nk;
            transverse(a, ja, ka, nj, nk);
            int lo = gl[a], hi = gl[a] + n[a] - 1;
            for (int kk = 0; kk < nk; kk++)
                for (int jj = 0; jj < nj; jj++)
                    for (int l = 1; l <= gl[a]; l++)
                        for (int side = 0; side < 2; side++) {
                            int I[3], S[3];
                            I[ja] = S[ja] = jj + gl[ja];
                            I[ka] = S[ka] = kk + gl[ka];
                            if (side == 0) { I[a] = lo - l; S[a] = lo + l - 1; }
                            else { I[a] = hi + l; S[a] = hi - l + 1; }
                            int m = marker(2 * a + side, jj, kk);
                            bc_image(cfg.bc[m], g, &V[idx(S[0], S[1], S[2]) * NV], nl,
                                     &V[idx(I[0], I[1], I[2]) * NV]);
                        }
        }
    }

    /* ComputeSolutionReconstruction (ERK:879-925) + periodic ExchangeReconstructions */
    void reconstruct_all(const std::vector<double>& V) {
        parallel_for(nthreads, n[2] * n[1], [&](int row) {
            int kk = row / n[1], jj = row % n[1];
            for (int ii = 0; ii < n[0]; ii++) {
                int i = ii + gl[0], j = jj + gl[1], k = kk + gl[2];
                double nb[6][NV];
                double nbc[6][3] = {{0}};
                for (int a = 0; a < nd; a++)
                    for (int s = 0; s < 2; s++) {
                        int J[3] = {i, j, k};
                        J[a] += (s == 0) ? -1 : 1;
                        for (int v = 0; v < NV; v++) nb[2 * a + s][v] = V[idx(J[0], J[1], J[2]) * NV + v];
                        nbc[2 * a + s][0] = C[0][J[0]];
                        nbc[2 * a + s][1] = C[1][J[1]];
                        nbc[2 * a + s][2] = C[2][J[2]];
                    }
                double xc[3] = {C[0][i], C[1][j], C[2][k]};
                double hc[3] = {H[0][i], H[1][j], H[2][k]};
                double eps2 = 0;
                if (cfg.recon == TS_RECON_LINEAR_LIM && cfg.limiter == TS_LIMITER_VENKAT) {
                    /* limVenkatar::SendCellInfo (Limiters.h:93-96) */
                    double dx = std::pow(hc[0] * hc[1] * hc[2], 1.0 / nd);
                    eps2 = std::pow(K * dx, 3.0);
                }
                reconstruct(cfg.recon, cfg.limiter, nd, &V[idx(i, j, k) * NV], nb, xc, nbc, hc,
                            eps2, rec[ridx(ii + 1, jj + rl[1], kk + rl[2])]);
            }
        });
        /* ExchangeReconstructions (ERK:928-1147): only periodic self-wrap matters on one rank */
        for (int a = 0; a < nd; a++) {
            if (!cfg.periodic[a]) continue;
            int ja = (a == 0) ? 1 : 0, ka = (a == 2) ? 1 : 2;
            int nj = ja < nd ? n[ja] : 1, nk = ka < nd ? n[ka] : 1;
            for (int kk = 0; kk < nk; kk++)
                for (int jj = 0; jj < nj; jj++) {
                    int D[3], S[3];
                    D[ja] = S[ja] = jj + rl[ja];
                    D[ka] = S[ka] = kk + rl[ka];
                    D[a] = n[a] + 1; S[a] = 1;
                    rec[ridx(D[0], D[1], D[2])] = rec[ridx(S[0], S[1], S[2])];
                    D[a] = 0; S[a] = n[a];
                    rec[ridx(D[0], D[1], D[2])] = rec[ridx(S[0], S[1], S[2])];
                }
        }
    }

    double vel(const std::vector<double>& V, size_t c, int d) const { return V[c * NV + 1 + d] / V[c * NV]; }

    /* ComputeVelocityGradients (ERK:79-132): central differences in the interior
       (ComputeGradientNonUniformStructured, kernel.h:918-938), the periodic exchange
       (ExchangeGradients, ERK:135-300) and the dummy-cell gradients of the walls
       (ComputeDummyCellGradients, ERK:303-665). V has its ghost layers filled. */
    void velocity_gradients(const std::vector<double>& V) {
        G.assign(vol.size() * 9, 0.0);
        for (int k = gl[2]; k < gl[2] + n[2]; k++)
            for (int j = gl[1]; j < gl[1] + n[1]; j++)
                for (int i = gl[0]; i < gl[0] + n[0]; i++) {
                    int I[3] = {i, j, k};
                    size_t c = idx(i, j, k);
                    for (int a = 0; a < nd; a++) {
                        int P[3] = {i, j, k}, M[3] = {i, j, k};
                        P[a]++;
                        M[a]--;
                        size_t cp = idx(P[0], P[1], P[2]), cm = idx(M[0], M[1], M[2]);
                        double dX = C[a][I[a] + 1] - C[a][I[a] - 1];
                        for (int d = 0; d < 3; d++) G[c * 9 + d * 3 + a] = (vel(V, cp, d) - vel(V, cm, d)) / dX;
                    }
                }
        for (int a = 0; a < nd; a++) {
            int ja = (a == 0) ? 1 : 0, ka = (a == 2) ? 1 : 2;
            int nj = ja < nd ? n[ja] : 1, nk = ka < nd ? n[ka] : 1;
            int lo = gl[a], hi = gl[a] + n[a] - 1;
            for (int kk = 0; kk < nk; kk++)
                for (int jj = 0; jj < nj; jj++)
                    for (int side = 0; side < 2; side++) {
                        int I[3], Gc[3];
                        I[ja] = Gc[ja] = jj + gl[ja];
                        I[ka] = Gc[ka] = kk + gl[ka];
                        I[a] = side ? hi : lo;
                        Gc[a] = side ? hi + 1 : lo - 1;
                        size_t c = idx(I[0], I[1], I[2]), gc = idx(Gc[0], Gc[1], Gc[2]);
                        if (cfg.periodic[a]) {
                            /* single-rank exchange: the ghost takes the gradient of the wrapped cell */
                            int S[3] = {I[0], I[1], I[2]};
                            S[a] = side ? lo : hi;
                            size_t sc = idx(S[0], S[1], S[2]);
                            for (int q = 0; q < 9; q++) G[gc * 9 + q] = G[sc * 9 + q];
                            continue;
                        }
                        /* face centre, boundary value (BCGeneral.h:90-94) and the T-stencil of
                           Stencil::InternalBorderStencil (Stencil.h:28-61) plus the dummy cell */
                        double fc[3] = {C[0][I[0]], C[1][I[1]], C[2][I[2]]};
                        fc[a] = side ? C[a][I[a]] + 0.5 * H[a][I[a]] : C[a][I[a]] - 0.5 * H[a][I[a]];
                        double nrm[3] = {0, 0, 0};
                        nrm[a] = side ? 1.0 : -1.0;
                        int m = marker(2 * a + side, jj, kk);
                        double dummy[NV], Vf[NV];
                        bc_image(cfg.bc[m], cfg.gamma, &V[c * NV], nrm, dummy);
                        for (int v = 0; v < NV; v++) Vf[v] = 0.5 * (V[c * NV + v] + dummy[v]);
                        int cells[6][3];
                        int ns = 0;
                        auto push = [&](int di, int dj, int dk) {
                            cells[ns][0] = I[0] + di;
                            cells[ns][1] = I[1] + dj;
                            cells[ns][2] = I[2] + dk;
                            ns++;
                        };
                        push(0, 0, 0);
                        if (a == 0) {
                            if (nd > 1) { push(0, 1, 0); push(0, -1, 0); }
                            if (nd > 2) { push(0, 0, 1); push(0, 0, -1); }
                        } else if (a == 1) {
                            push(1, 0, 0); push(-1, 0, 0);
                            if (nd > 2) { push(0, 0, 1); push(0, 0, -1); }
                        } else {
                            push(1, 0, 0); push(-1, 0, 0); push(0, 1, 0); push(0, -1, 0);
                        }
                        cells[ns][0] = Gc[0];
                        cells[ns][1] = Gc[1];
                        cells[ns][2] = Gc[2];
                        ns++;
                        double pts[6][3], vals[6];
                        for (int q = 0; q < ns; q++)
                            for (int e = 0; e < 3; e++) pts[q][e] = C[e][cells[q][e]];
                        for (int d = 0; d < 3; d++) {
                            for (int q = 0; q < ns; q++) vals[q] = vel(V, idx(cells[q][0], cells[q][1], cells[q][2]), d);
                            double gf[3];
                            lsq_gradient(nd, fc, Vf[1 + d] / Vf[0], pts, vals, ns, gf);
                            for (int e = 0; e < 3; e++) G[gc * 9 + d * 3 + e] = 2.0 * gf[e] - G[c * 9 + d * 3 + e];
                        }
                    }
        }
    }

    /* viscous flux through the face between cells cL and cR = cL + e_a, the right cell at index
       i along a (ComputeViscousFluxes, ERK:668-874): face-normal derivatives from the two cell
       values, tangential ones from the distance-weighted cell gradients */
    void viscous_flux(const std::vector<double>& V, int a, size_t cL, size_t cR, int i, double* fv) const {
        const double mu = cfg.viscosity;
        const double s_visc = (-2.0 / 3.0) * mu;
        double dx = C[a][i] - C[a][i - 1];
        double l = H[a][i - 1] / dx;
        double r = H[a][i] / dx;
        double D[3][3]; /* D[d][e] = d(vel_d)/dx_e at the face */
        double uf[3];
        for (int d = 0; d < 3; d++) {
            for (int e = 0; e < 3; e++) D[d][e] = 0.5 * (r * G[cL * 9 + d * 3 + e] + l * G[cR * 9 + d * 3 + e]);
            double vl = vel(V, cL, d), vr = vel(V, cR, d);
            D[d][a] = (vr - vl) / dx;
            uf[d] = 0.5 * (l * vr + r * vl);
        }
        double tau_d = s_visc * (D[0][0] + D[1][1] + D[2][2]);
        double t[3]; /* row a of the stress tensor */
        for (int d = 0; d < 3; d++)
            t[d] = (d == a) ? tau_d + 2 * mu * D[a][a] : mu * (D[std::min(a, d)][std::max(a, d)] + D[std::max(a, d)][std::min(a, d)]);
        fv[0] = 0.0;
        fv[1] = t[0];
        fv[2] = t[1];
        fv[3] = t[2];
        fv[4] = uf[0] * t[0] + uf[1] * t[1] + uf[2] * t[2];
    }

    /* ProcessExternalForces / ProcessExternalAcceleration (kernel.h:941-989), after the sweeps */
    void sources(const std::vector<double>& V, std::vector<double>& Rv) const {
        if (!cfg.external_force && !cfg.external_accel) return;
        for (int k = gl[2]; k < gl[2] + n[2]; k++)
            for (int j = gl[1]; j < gl[1] + n[1]; j++)
                for (int i = gl[0]; i < gl[0] + n[0]; i++) {
                    size_t c = idx(i, j, k);
                    double volume = H[0][i] * H[1][j] * H[2][k];
                    double u = vel(V, c, 0), v = vel(V, c, 1), w = vel(V, c, 2);
                    double* r = &Rv[c * NV];
                    if (cfg.external_force) {
                        const double* s = cfg.sigma;
                        r[1] += volume * s[0];
                        r[2] += volume * s[1];
                        r[3] += volume * s[2];
                        r[4] += (s[0] * u + s[1] * v + s[2] * w) * volume;
                    }
                    if (cfg.external_accel) {
                        const double* g = cfg.accel;
                        double ro = V[c * NV];
                        r[1] += ro * volume * g[0];
                        r[2] += ro * volume * g[1];
                        r[3] += ro * volume * g[2];
                        r[4] += (ro * g[0] * u + ro * g[1] * v + ro * g[2] * w) * volume;
                    }
                }
    }

    /* ComputeResidual (ERK:1150-1424) */
    void residual(const std::vector<double>& V, std::vector<double>& Rv, std::vector<double>& SRv) {
        std::fill(Rv.begin(), Rv.end(), 0.0);
        std::fill(SRv.begin(), SRv.end(), 0.0);
        double g = cfg.gamma;
        for (int a = 0; a < nd; a++) {
            int ja = (a == 0) ? 1 : 0, ka = (a == 2) ? 1 : 2;
            int nj = ja < nd ? n[ja] : 1, nk = ka < nd ? n[ka] : 1;
            double fn[3] = {0, 0, 0};
            fn[a] = 1.0;
            parallel_for(nthreads, nk * nj, [&](int line) {
                int kk = line / nj, jj = line % nj;
                int T[3];
                T[ja] = jj + gl[ja];
                T[ka] = kk + gl[ka];
                /* face area (ERK:1187-1190, 1268-1269, 1349) */
                double fS;
                if (a == 0) fS = (nd == 1) ? 1.0 : (nd == 2 ? H[1][T[1]] : H[1][T[1]] * H[2][T[2]]);
                else if (a == 1) fS = (nd == 3) ? H[0][T[0]] * H[2][T[2]] : H[0][T[0]];
                else fS = H[0][T[0]] * H[1][T[1]];
                double fl[NV] = {0}, fr[NV];
                int RT[3];
                RT[ja] = jj + rl[ja];
                RT[ka] = kk + rl[ka];
                int lo = gl[a], hi = gl[a] + n[a] - 1;
                for (int i = lo; i <= hi + 1; i++) {
                    double hlv = 0.5 * H[a][i - 1];
                    double hrv = -0.5 * H[a][i];
                    double pl[3] = {0, 0, 0}, pr[3] = {0, 0, 0};
                    pl[a] = hlv;
                    pr[a] = hrv;
                    double UL[NV], UR[NV];
                    const bool pos = cfg.riemann != TS_RIEMANN_ROE;
                    int ri = i - lo; /* recon index of the left cell */
                    if (!cfg.periodic[a] && i == lo) {
                        RT[a] = 1;
                        sample(cfg.recon, rec[ridx(RT[0], RT[1], RT[2])], pr, UR, pos);
                        int m = marker(2 * a, jj, kk);
                        bc_image(cfg.bc[m], g, UR, fn, UL);
                    } else if (!cfg.periodic[a] && i == hi + 1) {
                        RT[a] = n[a];
                        sample(cfg.recon, rec[ridx(RT[0], RT[1], RT[2])], pl, UL, pos);
                        int m = marker(2 * a + 1, jj, kk);
                        bc_image(cfg.bc[m], g, UL, fn, UR);
                    } else {
                        RT[a] = ri;
                        sample(cfg.recon, rec[ridx(RT[0], RT[1], RT[2])], pl, UL, pos);
                        RT[a] = ri + 1;
                        sample(cfg.recon, rec[ridx(RT[0], RT[1], RT[2])], pr, UR, pos);
                    }
                    double lam;
                    flux(cfg.riemann, g, cfg.eps, UL, UR, fn, fr, &lam);
                    int I[3] = {T[0], T[1], T[2]};
                    if (cfg.viscous) {
                        /* fr -= fvisc (ERK:1221-1222) */
                        I[a] = i - 1;
                        size_t cL = idx(I[0], I[1], I[2]);
                        I[a] = i;
                        double fv[NV];
                        viscous_flux(V, a, cL, idx(I[0], I[1], I[2]), i, fv);
                        for (int v = 0; v < NV; v++) fr[v] -= fv[v];
                    }
                    I[a] = i;
                    SRv[idx(I[0], I[1], I[2])] += fS * lam;
                    if (i > lo) {
                        I[a] = i - 1;
                        size_t id = idx(I[0], I[1], I[2]);
                        for (int v = 0; v < NV; v++) Rv[id * NV + v] -= (fr[v] - fl[v]) * fS;
                        SRv[id] += fS * lam;
                        /* viscous spectral-radius term (ERK:1239-1248): gas viscosity, Prandtl
                           number 1, whether or not the viscous fluxes are on */
                        double roFace = std::sqrt(UL[0] * UR[0]);
                        double vsr = std::max(4.0 / (3.0 * roFace), g / roFace);
                        vsr *= cfg.viscosity / 1.0;
                        vsr *= fS * fS;
                        vsr /= vol[id];
                        SRv[id] += vsr;
                    }
                    for (int v = 0; v < NV; v++) fl[v] = fr[v];
                }
            });
        }
        sources(V, Rv);
    }

    double compute_dt(const std::vector<double>& SRv, double next_snap) {
        /* ComputeTimeStep (ERK:1427-1437): min over every cell, ghosts included */
        double m = std::numeric_limits<double>::infinity();
        bool first = true;
        for (size_t c = 0; c < SRv.size(); c++) {
            double q = vol[c] / SRv[c];
            if (first) { m = q; first = false; }
            else if (q < m) m = q;
        }
        double dt = m * cfg.cfl;
        if (next_snap > 0 && next_snap < time + dt) dt = next_snap - time;
        return dt;
    }

    void stage(const std::vector<double>& Vin, std::vector<double>& Rv, std::vector<double>& SRv) {
        std::vector<double>& Vm = const_cast<std::vector<double>&>(Vin);
        fill_ghosts(Vm);
        reconstruct_all(Vin);
        if (cfg.viscous) velocity_gradients(Vin);
        residual(Vin, Rv, SRv);
    }

    int step(double next_snap, double* dt_out, double* norms) {
        /* stage 1 = the reference's IterationStep (ERK:1440-1491) */
        stage(U, R, SR);
        double dt = compute_dt(SR, next_snap);
        if (!(dt > 0) || !std::isfinite(dt)) return TS_ERR_UNPHYSICAL;
        size_t nc = vol.size();
        if (cfg.rk_stages == 1) {
            /* UpdateSolution (kernel.h:319-331) */
            for (int v = 0; v < NV; v++) {
                double s = 0;
                for (size_t c = 0; c < nc; c++) {
                    double q = dt * (R[c * NV + v] / vol[c]);
                    U[c * NV + v] += q;
                    s += vol[c] * std::fabs(R[c * NV + v]);
                }
                if (norms) norms[v] = s;
            }
        } else {
            /* SSP-RK3 (Shu-Osher), extension; dt fixed from U^n */
            U0 = U;
            for (int v = 0; v < NV; v++) {
                double s = 0;
                for (size_t c = 0; c < nc; c++) s += vol[c] * std::fabs(R[c * NV + v]);
                if (norms) norms[v] = s;
            }
            parallel_for(nthreads, (int)nc, [&](int c) {
                for (int v = 0; v < NV; v++) U[(size_t)c * NV + v] += dt * (R[(size_t)c * NV + v] / vol[c]);
            });
            std::vector<double> R2(R.size()), SR2(SR.size());
            stage(U, R2, SR2);
            parallel_for(nthreads, (int)nc, [&](int c) {
                for (int v = 0; v < NV; v++) {
                    size_t q = (size_t)c * NV + v;
                    double t = U[q] + dt * (R2[q] / vol[c]);
                    U[q] = 0.75 * U0[q] + 0.25 * t;
                }
            });
            stage(U, R2, SR2);
            parallel_for(nthreads, (int)nc, [&](int c) {
                for (int v = 0; v < NV; v++) {
                    size_t q = (size_t)c * NV + v;
                    double t = U[q] + dt * (R2[q] / vol[c]);
                    U[q] = (1.0 / 3.0) * U0[q] + (2.0 / 3.0) * t;
                }
            });
        }
        time += dt;
        iter++;
        *dt_out = dt;
        return TS_OK;
    }
};

extern "C" {

tso_ctx* tso_create(const ts_config* cfg, int n_threads) {
    if (!cfg || cfg->ndims < 1 || cfg->ndims > 3 || cfg->ghost != 1) return nullptr;
    tso_ctx* o = new tso_ctx();
    o->cfg = *cfg;
    o->nd = cfg->ndims;
    o->n[0] = cfg->nx;
    o->n[1] = cfg->ndims > 1 ? cfg->ny : 1;
    o->n[2] = cfg->ndims > 2 ? cfg->nz : 1;
    for (int a = 0; a < 3; a++) {
        o->gl[a] = a < o->nd ? cfg->ghost : 0;
        o->na[a] = o->n[a] + 2 * o->gl[a];
        o->rl[a] = a < o->nd ? 1 : 0;
        o->rn[a] = o->n[a] + 2 * o->rl[a];
        if (a >= o->nd) o->cfg.periodic[a] = 1; /* grid.h:92-95 */
    }
    o->nthreads = std::max(1, n_threads);
    o->build_grid();
    size_t nc = (size_t)o->na[0] * o->na[1] * o->na[2];
    o->U.assign(nc * NV, 0.0);
    o->R.assign(nc * NV, 0.0);
    o->SR.assign(nc, 0.0);
    o->rec.assign((size_t)o->rn[0] * o->rn[1] * o->rn[2], Rec{});
    return o;
}

void tso_destroy(tso_ctx* o) { delete o; }

size_t tso_state_size(const tso_ctx* o) { return o->U.size(); }

int tso_set_state(tso_ctx* o, const double* aos, size_t n) {
    if (n != o->U.size()) return TS_ERR_INVALID;
    std::memcpy(o->U.data(), aos, n * sizeof(double));
    return TS_OK;
}

int tso_get_state(tso_ctx* o, double* aos, size_t n) {
    if (n != o->U.size()) return TS_ERR_INVALID;
    std::memcpy(aos, o->U.data(), n * sizeof(double));
    return TS_OK;
}

int tso_set_time(tso_ctx* o, double time, long long iteration) {
    o->time = time;
    o->iter = iteration;
    return TS_OK;
}

int tso_step(tso_ctx* o, int max_steps, double t_max, double next_snapshot_time,
             ts_step_info* out, double* dt_log) {
    int done = 0;
    double dt = 0, norms[NV] = {0};
    for (; done < max_steps; done++) {
        int rc = o->step(next_snapshot_time, &dt, norms);
        if (rc) return rc;
        if (dt_log) dt_log[done] = dt;
        if (o->time >= t_max) { done++; break; }
        if (next_snapshot_time > 0 && o->time == next_snapshot_time) { done++; break; }
    }
    if (out) {
        out->time = o->time;
        out->dt = dt;
        out->iteration = o->iter;
        for (int v = 0; v < NV; v++) out->residual_norm[v] = norms[v];
        out->steps_done = done;
    }
    return TS_OK;
}

int tso_grid(const tso_ctx* o, int axis, double* coord, double* h, int n) {
    if (axis < 0 || axis > 2 || n != o->na[axis]) return TS_ERR_INVALID;
    for (int i = 0; i < n; i++) {
        coord[i] = o->C[axis][i];
        h[i] = o->H[axis][i];
    }
    return TS_OK;
}

int tso_last_residual(const tso_ctx* o, double* residual, double* spectral_radius, size_t n) {
    if (n != o->SR.size()) return TS_ERR_INVALID;
    std::memcpy(residual, o->R.data(), n * NV * sizeof(double));
    std::memcpy(spectral_radius, o->SR.data(), n * sizeof(double));
    return TS_OK;
}

void tso_flux(int riemann, double gamma, double eps, const double* UL, const double* UR,
              const double* n, double* F, double* max_eig) {
    flux(riemann, gamma, eps, UL, UR, n, F, max_eig);
}

int tso_star_state(double gamma, double rl, double ul, double pl, double rr, double ur,
                   double pr, double* pstar, double* ustar, double* rstarl, double* rstarr,
                   int* iterations) {
    Star s = star_state(gamma, rl, ul, pl, rr, ur, pr);
    *pstar = s.p;
    *ustar = s.u;
    *rstarl = s.rl;
    *rstarr = s.rr;
    if (iterations) *iterations = s.iters;
    return s.vacuum ? 1 : 0;
}

void tso_exact_sample(double gamma, double rl, double ul, double pl, double rr, double ur,
                      double pr, double s, double* out3) {
    Star st = star_state(gamma, rl, ul, pl, rr, ur, pr);
    bool left;
    exact_sample(gamma, rl, ul, pl, rr, ur, pr, st, s, &out3[0], &out3[1], &out3[2], &left);
}

void tso_reconstruct(int recon, int limiter, int ndims, const double* h, const double* center,
                     const double* neighbours, double* grads, double* lim) {
    double nb[6][NV];
    double nbc[6][3] = {{0}};
    for (int n = 0; n < 2 * ndims; n++) {
        for (int v = 0; v < NV; v++) nb[n][v] = neighbours[n * NV + v];
        int a = n / 2;
        nbc[n][a] = (n % 2 == 0 ? -1.0 : 1.0) * h[a];
    }
    double xc[3] = {0, 0, 0};
    double hc[3] = {h[0], ndims > 1 ? h[1] : 1.0, ndims > 2 ? h[2] : 1.0};
    double eps2 = std::pow(5.0 * std::pow(hc[0] * hc[1] * hc[2], 1.0 / ndims), 3.0);
    Rec r;
    reconstruct(recon, limiter, ndims, center, nb, xc, nbc, hc, eps2, r);
    for (int v = 0; v < NV; v++) {
        for (int a = 0; a < 3; a++) grads[v * 3 + a] = r.g[v][a];
        lim[v] = r.lim[v];
    }
}

void tso_bc_image(const ts_bc_config* bc, double gamma, const double* U, const double* n,
                  double* out) {
    bc_image(*bc, gamma, U, n, out);
}

int tso_velocity_gradients(tso_ctx* o, double* grads, size_t n) {
    if (n != o->vol.size() * 9) return TS_ERR_INVALID;
    std::vector<double> V = o->U;
    o->fill_ghosts(V);
    o->velocity_gradients(V);
    std::memcpy(grads, o->G.data(), n * sizeof(double));
    return TS_OK;
}

void tso_lsq_gradient(int ndims, const double* point, double value, const double* points,
                      const double* values, int m, double* grad) {
    lsq_gradient(ndims, point, value, reinterpret_cast<const double (*)[3]>(points), values, m, grad);
}

}  // extern "C"
