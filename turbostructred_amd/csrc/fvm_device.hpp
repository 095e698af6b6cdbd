// fvm_device.hpp -- device-side physics of the explicit finite-volume step (gfx950).
//
// Point-wise pieces used by the fused stage kernels: boundary images, cell reconstruction and
// the Riemann fluxes, templated on the arithmetic type R (double, or float for the
// FP32-residual form).  Each function cites the reference routine whose result it reproduces;
// the arithmetic is written for CDNA (fused multiply-add, reciprocal instead of division where
// the result stays within a few ulp), not transcribed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsd {

constexpr int NV = 5;

// ---- boundary conditions -------------------------------------------------------------------
// Values of BoundaryConditionType (src/BoundaryConditions/BoundaryConditionConfiguration.h:12-19)
enum : int { BC_WALL = 0, BC_SYMMETRY = 1, BC_MOVING_WALL = 2, BC_SUBSONIC_INLET = 3,
             BC_SUBSONIC_OUTLET = 4, BC_NATURAL = 5 };

struct BCDev {
    int type;
    double vel[3];
    double p;
    double rho;
};

// ---- fast reciprocal / square root ---------------------------------------------------------
// Hardware estimates refined by Newton steps (double) or used directly (float).  Results are
// within an ulp or two of the correctly rounded value; the parity tolerance is 1e-12.  Only for
// operands well inside the normal range (densities, sound speeds, mass fluxes): the estimate
// overflows for subnormal divisors.
__device__ inline double frcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double fdiv(double a, double b) {
    double r = frcp(b);
    double q = a * r;
    double e = __builtin_fma(-b, q, a);
    return __builtin_fma(e, r, q);
}
__device__ inline float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// sqrt from the reciprocal-square-root estimate with two Newton-Raphson corrections (no
// rescaling: for normal-range arguments; 0 maps to 0, negative arguments to NaN)
__device__ inline double fsqrt(double x) {
    double r = __builtin_amdgcn_rsq(x);
    double y = x * r;
    double h = 0.5 * r;
    double e = __builtin_fma(-h, y, 0.5);
    y = __builtin_fma(y, e, y);
    h = __builtin_fma(h, e, h);
    double d = __builtin_fma(-y, y, x);
    y = __builtin_fma(d, h, y);
    d = __builtin_fma(-y, y, x);
    y = __builtin_fma(d, h, y);
    return x == 0.0 ? 0.0 : y;
}
__device__ inline float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// one-instruction maximum / minimum in the arithmetic type (v_max_f64 / v_max_f32, no
// conversion to double for float operands)
__device__ inline double vmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ inline float vmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ inline double vmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ inline float vmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ inline double vabs(double a) { return __builtin_fabs(a); }
__device__ inline float vabs(float a) { return __builtin_fabsf(a); }

// smallest divisor the limiter uses: well inside the normal range, where the reciprocal
// estimate and its Newton steps are accurate
template <typename R> __device__ inline R tiny_divisor();
template <> __device__ inline double tiny_divisor<double>() { return 1e-290; }
template <> __device__ inline float tiny_divisor<float>() { return 1e-35f; }

// ---- equation of state (generalEOS.h:6-23; ideal gas, idealGasEOS.h:7-46) -------------------
// The reference's GeneralEOS interface as a device functor: pressure, sound speed and specific
// internal energy, and the two pressure derivatives, as functions of the density and the
// specific internal energy e (these five follow the reference's formulas operation for
// operation, with IEEE division and square root).  The kernels work on conserved variables, so
// the functor also gives the volumetric forms they evaluate: the pressure of an internal energy
// density rho e, the sound speed from p and 1/rho, and rho e of a pressure -- each the
// reference's method with the density factor folded in (for an ideal gas
// GetPressure(rho, rho e / rho) = (gamma - 1) rho e, with no division).  HLLC, its dt
// eigenvalues, the boundary images and the exact solver's pressures and face energy go through
// it.  The Roe flux and the exact solver's wave relations are the reference's perfect-gas
// formulas (RoeSolverPerfectGasEOS.h, GodunovSolverPerfectGasEOS.h) and use gamma directly.
__device__ inline double ieee_sqrt(double x) { return __builtin_sqrt(x); }
__device__ inline float ieee_sqrt(float x) { return __builtin_sqrtf(x); }

template <typename R>
struct IdealGasEOS {
    R gamma;
    __device__ explicit IdealGasEOS(R g) : gamma(g) {}
    // GeneralEOS (generalEOS.h:9-21) as IdealGasEOS implements it (idealGasEOS.h:22-44)
    __device__ R GetPressure(R ro, R e) const { return (gamma - R(1)) * ro * e; }
    __device__ R GetSoundSpeed(R ro, R e) const {
        const R pressure = (gamma - R(1)) * ro * e;
        return ieee_sqrt(gamma * pressure / ro);
    }
    __device__ R GetInternalEnergy(R ro, R pressure) const { return pressure / ((gamma - R(1)) * ro); }
    __device__ R GetPressureEnergyDerivative(R ro, R) const { return (gamma - R(1)) * ro; }
    __device__ R GetPressureDensityDerivative(R, R e) const { return (gamma - R(1)) * e; }
    // volumetric forms of the kernels (rho_e = rho e; inv_ro = 1/rho from the caller's reciprocal)
    __device__ R PressureOfEnergyDensity(R rho_e) const { return (gamma - R(1)) * rho_e; }
    __device__ R SoundSpeedOfPressure(R p, R inv_ro) const { return fsqrt(gamma * p * inv_ro); }
    __device__ R EnergyDensityOfPressure(R p) const { return p / (gamma - R(1)); }
};

// Image state of a boundary condition across a face normal to `axis` (getDummyValues,
// BCGeneral.h:186-268).  Every condition in the reference is even in the normal's sign, so the
// ghost fill (normal -e_axis on both sides, kernel.h:743,760) and the boundary-face states
// (normal +e_axis, ExplicitRungeKuttaFVM.h:1202,1208) share this one function.
template <typename R>
__device__ __forceinline__ void bc_image(const BCDev& bc, R gamma, int axis, const R (&u)[NV], R (&o)[NV]) {
    switch (bc.type) {
    case BC_SYMMETRY:
        // U - 2 (n.m) n: only the normal momentum changes sign
        o[0] = u[0]; o[1] = u[1]; o[2] = u[2]; o[3] = u[3]; o[4] = u[4];
        o[1 + axis] = -u[1 + axis];
        break;
    case BC_WALL:
        o[0] = u[0]; o[1] = -u[1]; o[2] = -u[2]; o[3] = -u[3]; o[4] = u[4];
        break;
    case BC_MOVING_WALL: {
        R rho = u[0];
        R ir = R(1) / rho;
        R vx = u[1] * ir, vy = u[2] * ir, vz = u[3] * ir;
        R rhoe = u[4] - R(0.5) * rho * (vx * vx + vy * vy + vz * vz);
        R dx = R(2) * R(bc.vel[0]) - vx, dy = R(2) * R(bc.vel[1]) - vy, dz = R(2) * R(bc.vel[2]) - vz;
        o[0] = rho; o[1] = rho * dx; o[2] = rho * dy; o[3] = rho * dz;
        o[4] = rhoe + R(0.5) * rho * (dx * dx + dy * dy + dz * dz);
        break;
    }
    case BC_SUBSONIC_INLET: {
        R r = R(bc.rho);
        R vx = R(bc.vel[0]), vy = R(bc.vel[1]), vz = R(bc.vel[2]);
        o[0] = r; o[1] = r * vx; o[2] = r * vy; o[3] = r * vz;
        o[4] = IdealGasEOS<R>(gamma).EnergyDensityOfPressure(R(bc.p)) + R(0.5) * r * (vx * vx + vy * vy + vz * vz);
        break;
    }
    default:  // Natural, SubsonicOutlet: copy
        o[0] = u[0]; o[1] = u[1]; o[2] = u[2]; o[3] = u[3]; o[4] = u[4];
    }
}

template <typename R> __device__ inline R tiny_spread();
template <> __device__ inline double tiny_spread<double>() { return 3.0 * 2.2250738585072014e-308; }
template <> __device__ inline float tiny_spread<float>() { return 3.0f * 1.17549435e-38f; }

// ---- reconstruction ------------------------------------------------------------------------
// Values of Reconstruction / LimiterType (src/Methods/MethodConfiguration.h:5-21)
enum : int { RC_PWC = 0, RC_LINEAR = 1, RC_ENO2 = 2, RC_LINLIM = 3 };
enum : int { LIM_BJ = 0, LIM_VENKAT = 1 };

// Geometry of one cell along its ND axes: half sizes and slope denominators.
template <int ND, typename R>
struct CellGeo {
    R hh[ND];   // 0.5 * h (face sample offset, Linear2psLim.h:77-89, ERK:1194-1195)
    R rc[ND];   // 1 / (x_{i+1} - x_{i-1})  (Linear2PointsStencil.h:72)
    R rl[ND];   // 1 / (x_i - x_{i-1})      (ENO2PointsStencil.h:28)
    R rr[ND];   // 1 / (x_{i+1} - x_i)      (ENO2PointsStencil.h:29)
    R eps2;     // Venkatakrishnan (K dx)^3 (Limiters.h:93-96)
};

// Reconstruct one cell and return its face states: face[2a] on the minus face of axis a,
// face[2a+1] on the plus face.  nb[2a], nb[2a+1] are the axis neighbours (x-,x+,y-,y+,z-,z+),
// the stencil of ComputeSolutionReconstruction (ERK:897-916).
// With POS set (the HLLC and exact-Godunov paths, which need a real sound speed), a face state
// with rho <= 0 or 2 rho E <= |m|^2 falls back to the cell value.
template <int ND, int RECON, int LIM, typename R, bool POS = false>
__device__ inline void reconstruct(const R (&c)[NV], const R (&nb)[2 * ND][NV],
                                   const CellGeo<ND, R>& geo, R (&face)[2 * ND][NV]) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
        if constexpr (RECON == RC_PWC) {  // PiecewiseConstant.h:15
#pragma unroll
            for (int f = 0; f < 2 * ND; f++) face[f][v] = c[v];
        } else if constexpr (RECON == RC_LINLIM && LIM == LIM_BJ) {
            // Limited linear with Barth-Jespersen (Linear2psLim.h:64-96, Limiters.h:28-67,80-82).
            // Along each axis the two face projections are +-d_a, d_a = g_a h_a / 2 with the
            // central slope g_a (Linear2PointsStencil.h:66-97); one is bounded by umax and the
            // other by umin, so the minimum over the faces of min(1, (u_m - u_i)/d2) is
            // min(umax - c, c - umin) / max_a |d_a| when that is below 1.  The projection is
            // formed as (u_+ - u_-) (h_a / (2 (x_+ - x_-))) rather than the reference's
            // (c + g.p) - c: the two differ by an ulp of c, and so do the face values.  A
            // variable without spread has d_a = 0 and keeps its cell value, as the reference's
            // spread test (Limiters.h:50) gives.
            R d[ND], dmax = R(0);
#pragma unroll
            for (int a = 0; a < ND; a++) {
                d[a] = (nb[2 * a + 1][v] - nb[2 * a][v]) * (geo.rc[a] * geo.hh[a]);
                dmax = vmax(dmax, vabs(d[a]));
            }
            R umax = c[v], umin = c[v];
#pragma unroll
            for (int n = 0; n < 2 * ND; n++) {
                umax = vmax(umax, nb[n][v]);
                umin = vmin(umin, nb[n][v]);
            }
            const R hi = umax - c[v], lo = c[v] - umin;
            const R m = vmin(hi, lo);
            R lim = R(1);
            // a divisor below the floor (a variable constant up to round-off) is raised to it:
            // the reciprocal estimate stays in range, the quotient below 1, and the face values
            // move by less than the floor
            if (m < dmax) lim = fdiv(m, vmax(dmax, tiny_divisor<R>()));
#pragma unroll
            for (int a = 0; a < ND; a++) {
                const R t = lim * d[a];
                face[2 * a][v] = c[v] - t;
                face[2 * a + 1][v] = c[v] + t;
            }
        } else {
            R g[ND];
#pragma unroll
            for (int a = 0; a < ND; a++) {
                if constexpr (RECON == RC_ENO2) {  // ENO2PointsStencil.h:21-58
                    R gl = (c[v] - nb[2 * a][v]) * geo.rl[a];
                    R gr = (nb[2 * a + 1][v] - c[v]) * geo.rr[a];
                    g[a] = (__builtin_fabs(gl) < __builtin_fabs(gr)) ? gl : gr;
                } else {  // Linear2PointsStencil.h:66-97
                    g[a] = (nb[2 * a + 1][v] - nb[2 * a][v]) * geo.rc[a];
                }
            }
            R lim = R(1);
            if constexpr (RECON == RC_LINLIM) {
                // Limiter::ComputeLimiterValues (Limiters.h:28-67).  The minimum over the
                // faces of min(1, l_j) is found by comparing fractions A/D, then divided once.
                R umax = c[v], umin = c[v];
#pragma unroll
                for (int n = 0; n < 2 * ND; n++) {
                    umax = __builtin_fmax(umax, nb[n][v]);
                    umin = __builtin_fmin(umin, nb[n][v]);
                }
                if (!(umax - umin < tiny_spread<R>())) {
                    {  // Venkatakrishnan (Barth-Jespersen takes the branch above)
                        R bestA = R(1), bestD = R(1);
                        R inv_un = frcp(R(0.5) * (umax - umin));
#pragma unroll
                        for (int a = 0; a < ND; a++) {
#pragma unroll
                            for (int s = 0; s < 2; s++) {
                                R pt = s ? geo.hh[a] : -geo.hh[a];
                                R d2;
                                {
#pragma clang fp contract(off)
                                    R fv = c[v] + g[a] * pt;
                                    d2 = fv - c[v];
                                }
                                if (d2 > R(0) || d2 < R(0)) {
                                    // Venkatakrishnan (Limiters.h:102-115), normalised by
                                    // (umax - umin)/2
                                    R um = (d2 > R(0)) ? umax : umin;
                                    R dm = um * inv_un - c[v] * inv_un;
                                    R dn = d2 * inv_un;
                                    R A = (dm * dm + geo.eps2) + R(2) * dn * dm;
                                    R D = dm * dm + R(2) * dn * dn + dm * dn + geo.eps2;
                                    if (A * bestD < bestA * D) { bestA = A; bestD = D; }
                                }
                            }
                        }
                        lim = (bestA == bestD) ? R(1) : bestA / bestD;
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < ND; a++) {
                R d = g[a] * geo.hh[a];
                if constexpr (RECON == RC_LINLIM) {  // Linear2psLim.h:17-21 (Venkatakrishnan)
                    face[2 * a][v] = c[v] + lim * (g[a] * (-geo.hh[a]));
                    face[2 * a + 1][v] = c[v] + lim * d;
                } else {  // Linear2PointsStencil.h:26-30
                    face[2 * a][v] = c[v] - d;
                    face[2 * a + 1][v] = c[v] + d;
                }
            }
        }
    }
    if constexpr (POS && RECON != RC_PWC) {
        bool bad[2 * ND];
        bool any = false;
#pragma unroll
        for (int f = 0; f < 2 * ND; f++) {
            R q = R(2) * face[f][0] * face[f][4] -
                  (face[f][1] * face[f][1] + face[f][2] * face[f][2] + face[f][3] * face[f][3]);
            bad[f] = !(face[f][0] > R(0) && q > R(0));
            any = any || bad[f];
        }
        // rare: a branch the wavefront skips when none of its lanes needs the fallback
        if (__builtin_expect(any, 0)) {
#pragma unroll
            for (int f = 0; f < 2 * ND; f++)
                if (bad[f]) {
#pragma unroll
                    for (int v = 0; v < NV; v++) face[f][v] = c[v];
                }
        }
    }
}

// ---- Riemann solvers -------------------------------------------------------------------------
enum : int { FX_GODUNOV = 0, FX_ROE = 1, FX_HLLC = 3 };

// Internal energy density E - rho|v|^2/2 from the conserved state.  In the FP32 form the
// difference is redone in double when it is a small part of E (near vacuum, where the float
// subtraction loses most of its digits).
__device__ inline double internal_energy(double E, double rho_ke) { return E - rho_ke; }
__device__ inline float internal_energy(float E, float rho_ke, const float (&U)[NV]) {
    float e = E - rho_ke;
    if (e < 0.01f * E) {
        double m2 = double(U[1]) * U[1] + double(U[2]) * U[2] + double(U[3]) * U[3];
        e = float(double(U[4]) - 0.5 * m2 / double(U[0]));
    }
    return e;
}
__device__ inline double internal_energy(double E, double rho_ke, const double (&)[NV]) {
    return E - rho_ke;
}

// Roe-Pike flux with Harten's entropy fix (RoeSolverPerfectGasEOS.h:48-134) across a face with
// normal +e_D.  lam receives |u.n| + c of the Roe average (RoeSolverPerfectGasEOS.h:88,128).
template <int D, typename R>
__device__ inline void roe_flux(R gamma, R eps, const R (&UL)[NV], const R (&UR)[NV], R (&F)[NV],
                                R& lam) {
    const R gm1 = gamma - R(1);
    R il = frcp(UL[0]), ir = frcp(UR[0]);
    R vl[3] = {UL[1] * il, UL[2] * il, UL[3] * il};
    R vr[3] = {UR[1] * ir, UR[2] * ir, UR[3] * ir};
    R kl = R(0.5) * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]);
    R kr = R(0.5) * (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
    R el = UL[4] * il, er = UR[4] * ir;
    // physical fluxes (RoeSolverPerfectGasEOS.h:15-32)
    R pl = gm1 * internal_energy(UL[4], UL[0] * kl, UL);
    R pr = gm1 * internal_energy(UR[4], UR[0] * kr, UR);
    R hl = (el - kl) * gamma + kl;
    R hr = (er - kr) * gamma + kr;
    R sl = fsqrt(UL[0]), sr = fsqrt(UR[0]);
    R isum = frcp(sl + sr);
    R ql = sl * isum, qr = sr * isum;
    R v[3] = {ql * vl[0] + qr * vr[0], ql * vl[1] + qr * vr[1], ql * vl[2] + qr * vr[2]};
    R h = ql * hl + qr * hr;
    R vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    R phi2 = R(0.5) * gm1 * vv;
    R c = fsqrt(gm1 * h - phi2);
    R uw = v[D];
    R eig = __builtin_fabs(uw) + c;
    R e = eps * eig;
    R a1 = __builtin_fabs(uw), a3 = __builtin_fabs(uw + c), a4 = __builtin_fabs(uw - c);
    if (a1 < e) a1 = (a1 * a1 / e + e) * R(0.5);  // Harten (RoeSolverPerfectGasEOS.h:35-40)
    if (a3 < e) a3 = (a3 * a3 / e + e) * R(0.5);
    if (a4 < e) a4 = (a4 * a4 / e + e) * R(0.5);
    R ic = frcp(c);
    R Eig1 = a1;
    R Eiga = (a3 - a4) * R(0.5) * ic;
    R Eigb = (a3 + a4) * R(0.5) - a1;
    R R1 = phi2 * UR[0] - gm1 * UR[0] * ((v[0] * vr[0] + v[1] * vr[1] + v[2] * vr[2]) - er);
    R R2 = phi2 * UL[0] - gm1 * UL[0] * ((v[0] * vl[0] + v[1] * vl[1] + v[2] * vl[2]) - el);
    R D1 = UR[0] * (vr[D] - v[D]);
    R D2 = UL[0] * (vl[D] - v[D]);
    R C2 = ic * ic;
    // dissipation (RoeSolverPerfectGasEOS.h:96-125) with Qa = (0, n, u.n), Qb = (1, v, h)
    R vnl = vl[D], vnr = vr[D];
    R FL[NV], FR[NV];
    FL[0] = UL[0] * vnl; FR[0] = UR[0] * vnr;
    FL[1] = FL[0] * vl[0]; FR[1] = FR[0] * vr[0];
    FL[2] = FL[0] * vl[1]; FR[2] = FR[0] * vr[1];
    FL[3] = FL[0] * vl[2]; FR[3] = FR[0] * vr[2];
    FL[1 + D] += pl; FR[1 + D] += pr;
    FL[4] = vnl * (UL[4] + pl); FR[4] = vnr * (UR[4] + pr);
    R Qb[NV] = {R(1), v[0], v[1], v[2], h};
#pragma unroll
    for (int i = 0; i < NV; i++) {
        R qa = (i == 4) ? uw : ((i == 1 + D) ? R(1) : R(0));
        R diss = Eig1 * (UL[i] - UR[i]) + Eiga * ((R2 - R1) * qa + (D2 - D1) * Qb[i]) +
                 Eigb * ((R2 - R1) * Qb[i] * C2 + (D2 - D1) * qa);
        F[i] = R(0.5) * (FL[i] + FR[i] + diss);
    }
    lam = eig;
}

// HLLC flux (Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, 3rd ed.,
// sec. 10.4, eqs. 10.37-10.39) with Davis wave-speed bounds (eq. 10.48); lam = max(|SL|,|SR|).
// Extension with no counterpart in the reference.  Branch-free: the upwind side of the contact
// is selected first, then its flux and star state are evaluated once.
template <int D, typename R>
__device__ inline void hllc_flux(R gamma, const R (&UL)[NV], const R (&UR)[NV], R (&F)[NV],
                                 R& lam) {
    const IdealGasEOS<R> eos(gamma);
    R il = frcp(UL[0]), ir = frcp(UR[0]);
    R vl[3] = {UL[1] * il, UL[2] * il, UL[3] * il};
    R vr[3] = {UR[1] * ir, UR[2] * ir, UR[3] * ir};
    R pl = eos.PressureOfEnergyDensity(internal_energy(UL[4], R(0.5) * UL[0] * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]), UL));
    R pr = eos.PressureOfEnergyDensity(internal_energy(UR[4], R(0.5) * UR[0] * (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]), UR));
    R al = eos.SoundSpeedOfPressure(pl, il), ar = eos.SoundSpeedOfPressure(pr, ir);
    R ul = vl[D], ur = vr[D];
    R SL = __builtin_fmin(ul - al, ur - ar);
    R SR = __builtin_fmax(ul + al, ur + ar);
    R ml = UL[0] * (SL - ul), mr = UR[0] * (SR - ur);
    R Ss = fdiv(pr - pl + ul * ml - ur * mr, ml - mr);
    lam = __builtin_fmax(__builtin_fabs(SL), __builtin_fabs(SR));
    bool left = Ss >= R(0);
    R U[NV], vel[3];
#pragma unroll
    for (int i = 0; i < NV; i++) U[i] = left ? UL[i] : UR[i];
#pragma unroll
    for (int i = 0; i < 3; i++) vel[i] = left ? vl[i] : vr[i];
    R S = left ? SL : SR;
    R un = left ? ul : ur, p = left ? pl : pr, ik = left ? il : ir;
    R m = left ? ml : mr;
    R Fk[NV];
    Fk[0] = U[0] * un;
    Fk[1] = Fk[0] * vel[0];
    Fk[2] = Fk[0] * vel[1];
    Fk[3] = Fk[0] * vel[2];
    Fk[1 + D] += p;
    Fk[4] = un * (U[4] + p);
    bool supersonic = left ? (S >= R(0)) : (S <= R(0));
    // supersonic side: the star correction is multiplied by zero (and its operands made finite),
    // which gives Fk exactly with two selects instead of five
    // m / (S - Ss) and p / m from one reciprocal of m (S - Ss), which is positive on a
    // subsonic side (S < 0 <= Ss with m = rho (S - u) < 0, or the mirror case)
    const R dS = S - Ss;
    const R rmd = frcp(supersonic ? R(1) : m * dS);
    R coef = supersonic ? R(0) : m * m * rmd;
    R Sc = supersonic ? R(0) : S;
    R Us[NV];
    Us[0] = coef;
    Us[1] = coef * vel[0];
    Us[2] = coef * vel[1];
    Us[3] = coef * vel[2];
    Us[1 + D] = coef * Ss;
    R pm = supersonic ? R(0) : p * dS * rmd;
    Us[4] = coef * (U[4] * ik + (Ss - un) * (Ss + pm));
#pragma unroll
    for (int i = 0; i < NV; i++) F[i] = Fk[i] + Sc * (Us[i] - U[i]);
}

// ---- exact Riemann solver (always double) -----------------------------------------------------
// Powers x^y = exp(y ln x) of the exact solver's bases (pressure ratios, the two-rarefaction
// guess, the fan's C_K), all non-negative, with positive exponents, evaluated in line:
//   ln x: x = m 2^e with m in [1/sqrt2, sqrt2), ln m = 2 atanh s = 2 s (1 + s^2/3 + s^4/5 + ...)
//         with s = (m - 1)/(m + 1), |s| <= 0.172, summed to s^21 (the next term is below 1e-18);
//   exp t: t = k ln2 + r with k = rint(t / ln2), |r| <= 0.347, ln2 split in a high part with 32
//         significant bits (k ln2_hi exact) and the rest, and e^r by its Taylor series to r^13
//         (remainder below 5e-18), scaled by 2^k.
// On 2e6 bases in [1e-12, 1e4] with the solver's exponents this is within 2e-15 of the correctly
// rounded power for y <= 1 and 1.5e-14 for y up to 7 (the rounding of t = y ln x), the same as
// exp(y ln x) with the library's ln and exp, in 101 VALU instructions against their 139 (and 223
// for the library's correctly rounded pow).  One out-of-line copy: inlined at each call site the
// kernels outgrew the instruction cache.
// x = 0 gives 0; x < 0 or NaN gives NaN; x = +inf gives +inf.
static __device__ __noinline__ double dpow(double x, double y) {
    if (!(x > 0.0)) return x == 0.0 ? 0.0 : __builtin_nan("");
    if (x == __builtin_inf()) return x;
    constexpr double LN2_HI = 0.6931471803691238, LN2_LO = 1.9082149292705877e-10;
    int e;
    double m = __builtin_frexp(x, &e);
    if (m < 0.7071067811865476) {
        m *= 2.0;
        e -= 1;
    }
    const double s = fdiv(m - 1.0, m + 1.0);
    const double w = s * s;
    double p = 1.0 / 21.0;
    p = __builtin_fma(p, w, 1.0 / 19.0);
    p = __builtin_fma(p, w, 1.0 / 17.0);
    p = __builtin_fma(p, w, 1.0 / 15.0);
    p = __builtin_fma(p, w, 1.0 / 13.0);
    p = __builtin_fma(p, w, 1.0 / 11.0);
    p = __builtin_fma(p, w, 1.0 / 9.0);
    p = __builtin_fma(p, w, 1.0 / 7.0);
    p = __builtin_fma(p, w, 1.0 / 5.0);
    p = __builtin_fma(p, w, 1.0 / 3.0);
    p = __builtin_fma(p, w, 1.0);
    const double lnm = 2.0 * s * p;
    const double de = double(e);
    const double lnx = de * LN2_HI + __builtin_fma(de, LN2_LO, lnm);
    const double t = y * lnx;
    const double k = __builtin_rint(t * 1.4426950408889634);
    const double r = __builtin_fma(-k, LN2_LO, __builtin_fma(-k, LN2_HI, t));
    // 1/j! for j = 13 .. 2, then 1, 1
    double q = 1.6059043836821613e-10;
    q = __builtin_fma(q, r, 2.08767569878681e-09);
    q = __builtin_fma(q, r, 2.505210838544172e-08);
    q = __builtin_fma(q, r, 2.755731922398589e-07);
    q = __builtin_fma(q, r, 2.7557319223985893e-06);
    q = __builtin_fma(q, r, 2.48015873015873e-05);
    q = __builtin_fma(q, r, 0.0001984126984126984);
    q = __builtin_fma(q, r, 0.001388888888888889);
    q = __builtin_fma(q, r, 0.008333333333333333);
    q = __builtin_fma(q, r, 0.041666666666666664);
    q = __builtin_fma(q, r, 0.16666666666666666);
    q = __builtin_fma(q, r, 0.5);
    q = __builtin_fma(q, r, 1.0);
    q = __builtin_fma(q, r, 1.0);
    return __builtin_ldexp(q, int(k));
}

// The side-K constants of f_K (GodunovSolverPerfectGasEOS.h:58-97), formed once per face
// instead of once per Newton iteration.
struct SideDev {
    double pk, ak;
    double A, B;  // 2 / ((g+1) rho_K), p_K (g-1)/(g+1): the shock branch
    double ipk;   // 1 / p_K
    double ira;   // 1 / (rho_K a_K)
    double c2a;   // 2 a_K / (g-1)
};
// cA = 2/(g+1), cB = (g-1)/(g+1), c2 = 2/(g-1)
__device__ inline SideDev side_dev(double cA, double cB, double c2, double r, double pk, double ak) {
    SideDev s;
    s.pk = pk;
    s.ak = ak;
    s.A = cA * frcp(r);
    s.B = pk * cB;
    s.ipk = frcp(pk);
    s.ira = frcp(r * ak);
    s.c2a = c2 * ak;
    return s;
}

// f_K(p) and its derivative (GodunovSolverPerfectGasEOS.h:58-97; derivative Toro eq. 4.37);
// z = (g-1)/2g.  The reciprocals and the square root are the refined estimates (a few ulps).
// `pzk` = (p/pK)^z, used on the rarefaction side.
__device__ inline double fk_dev(const SideDev& s, double p, double pzk, double& d, double& pz) {
    if (p > s.pk) {
        const double ipb = frcp(p + s.B);
        const double q = fsqrt(s.A * ipb);
        d = q * (1.0 - 0.5 * (p - s.pk) * ipb);
        pz = 0.0;
        return (p - s.pk) * q;
    }
    // rarefaction: the power (p/pK)^z gives f, f' = (p/pK)^(z-1)/(r aK) and the tail speed of
    // the fan (the sampler reuses pz)
    const double pr = p * s.ipk;
    pz = pzk;
    d = pz * frcp(pr) * s.ira;
    return (pz - 1.0) * s.c2a;
}
// (p/pL)^z and (p/pR)^z with one power: (p/pR)^z = (p/pL)^z (pL/pR)^z, the last factor formed
// once per face; skipped when neither side is a rarefaction at p
__device__ inline void side_powers(const SideDev& l, const SideDev& r, double z, double rho_lr, double p, double& pzl,
                                   double& pzr) {
    pzl = pzr = 0.0;
    if (!(p > l.pk) || !(p > r.pk)) {
        pzl = dpow(p * l.ipk, z);
        pzr = pzl * rho_lr;
    }
}

// Star pressure and velocity, wave speeds and smax; the star densities are formed by the sampler,
// for the sampled side only.
struct StarDev {
    double p, u, SL, SHL, STL, SR, SHR, STR, smax;
    bool shock_l, shock_r, vacuum;
};

// Star state by Newton-Raphson from Toro's adaptive initial guess (Toro sec. 4.3, 9.5.1); the
// reference uses ALGLIB's bounded Levenberg-Marquardt for the same root
// (GodunovSolverPerfectGasEOS.h:143-185).  Divergent lanes simply run a masked loop.
__device__ inline StarDev star_state_dev(double g, double rl, double ul, double pl, double rr,
                                         double ur, double pr) {
    StarDev s;
    double pe = -1.0, fle = 0.0, fre = 0.0, zle = 0.0, zre = 0.0;  // last Newton evaluation
    double al = __builtin_sqrt(g * pl / rl), ar = __builtin_sqrt(g * pr / rr);
    double du = ur - ul;
    s.vacuum = (2.0 / (g - 1.0)) * (al + ar) <= du;
    const double z = (g - 1.0) / (2.0 * g);
    const double cA = 2.0 / (g + 1.0), cB = (g - 1.0) / (g + 1.0), c2 = 2.0 / (g - 1.0);
    const SideDev sl = side_dev(cA, cB, c2, rl, pl, al), sr = side_dev(cA, cB, c2, rr, pr, ar);
    const double rho_lr = dpow(pl * sr.ipk, z);  // (pL/pR)^z
    double p = 0.0;
    // Inputs that are not a physical state (zero, negative or NaN density or pressure) give a
    // NaN star pressure straight away.  Such inputs reach the solver on halo lanes, which are
    // fed by a cross-lane shift and whose flux is discarded; iterating on them to the cap would
    // hold the whole wavefront.  (For a face whose flux is used they give NaN either way.)
    const bool physical = rl > 0.0 && pl > 0.0 && rr > 0.0 && pr > 0.0;
    if (!physical) {
        p = __builtin_nan("");
    } else if (!s.vacuum) {
        double pmin = __builtin_fmin(pl, pr), pmax = __builtin_fmax(pl, pr);
        double ppv = 0.5 * (pl + pr) - 0.125 * du * (rl + rr) * (al + ar);
        ppv = __builtin_fmax(0.0, ppv);
        if (pmax / pmin <= 2.0 && pmin <= ppv && ppv <= pmax) {
            p = ppv;
        } else if (ppv < pmin) {
            double plz = dpow(pl, z), prz = dpow(pr, z);
            double t = (al + ar - 0.5 * (g - 1.0) * du) / (al / plz + ar / prz);
            p = dpow(t, 1.0 / z);
        } else {
            double A = 2.0 / ((g + 1.0) * rl), B = (g - 1.0) / (g + 1.0) * pl;
            double gl = __builtin_sqrt(A / (ppv + B));
            A = 2.0 / ((g + 1.0) * rr);
            B = (g - 1.0) / (g + 1.0) * pr;
            double gr = __builtin_sqrt(A / (ppv + B));
            p = (gl * pl + gr * pr - du) / (gl + gr);
        }
        if (!(p > 0.0)) p = 1e-12 * pmin;
        for (int it = 1; it <= 60; it++) {
            double dl, dr, zl, zr;
            double ql, qr;
            side_powers(sl, sr, z, rho_lr, p, ql, qr);
            double fl = fk_dev(sl, p, ql, dl, zl), fr = fk_dev(sr, p, qr, dr, zr);
            pe = p, fle = fl, fre = fr, zle = zl, zre = zr;
            double f = fl + fr + du;
            double pn = p - fdiv(f, dl + dr);
            if (!(pn > 0.0)) pn = 0.01 * p;
            // relative change |pn - p| / (pn + p) <= 1e-15, without the division (pn + p > 0)
            const bool done = __builtin_fabs(pn - p) <= 1e-15 * (pn + p);
            p = pn;
            if (done) break;
        }
    }
    s.p = p;
    double f1, f2, pzl, pzr;
    if (p == pe) {  // Newton stopped on a fixed point (e.g. equal states): f_L, f_R are known
        f1 = fle, f2 = fre, pzl = zle, pzr = zre;
    } else {
        double d1, d2;
        double ql, qr;
        side_powers(sl, sr, z, rho_lr, p, ql, qr);
        f1 = fk_dev(sl, p, ql, d1, pzl);
        f2 = fk_dev(sr, p, qr, d2, pzr);
    }
    s.u = 0.5 * (ul + ur) + 0.5 * (f2 - f1);
    s.smax = 0.0;
    double prl = p / pl;
    s.SL = s.SHL = s.STL = s.SR = s.SHR = s.STR = 0.0;
    if (p > pl) {
        s.shock_l = true;
        s.SL = ul - al * __builtin_sqrt((g + 1) * prl / (2 * g) + (g - 1) / (2 * g));
        s.smax = __builtin_fmax(s.smax, __builtin_fabs(s.SL));
    } else {
        s.shock_l = false;
        s.SHL = ul - al;
        s.STL = s.u - al * pzl;  // (p/pL)^((g-1)/2g), from f_L
        s.smax = __builtin_fmax(s.smax, __builtin_fabs(s.SHL));
    }
    double prr = p / pr;
    if (p > pr) {
        s.shock_r = true;
        s.SR = ur + ar * __builtin_sqrt((g + 1) * prr / (2 * g) + (g - 1) / (2 * g));
        s.smax = __builtin_fmax(s.smax, __builtin_fabs(s.SR));
    } else {
        s.shock_r = false;
        s.SHR = ur + ar;
        s.STR = s.u + ar * pzr;
        s.smax = __builtin_fmax(s.smax, __builtin_fabs(s.SHR));
    }
    if (s.vacuum) {  // the fan tails bound the vacuum (Toro sec. 4.6)
        s.STL = ul + 2.0 * al / (g - 1.0);
        s.STR = ur - 2.0 * ar / (g - 1.0);
    }
    return s;
}

// Exact Godunov flux: the solution sampled at x/t = 0 (GodunovSolverPerfectGasEOS.h:276-433),
// tangential velocity from the upwind side of the contact.  Computed in double for both
// arithmetic forms.
template <int D, typename R>
__device__ inline void godunov_flux(R gammaR, const R (&ULr)[NV], const R (&URr)[NV], R (&F)[NV],
                                    R& lam) {
    const double g = double(gammaR);
    const IdealGasEOS<double> eos(g);
    double UL[NV], UR[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) { UL[i] = double(ULr[i]); UR[i] = double(URr[i]); }
    double il = 1.0 / UL[0], ir = 1.0 / UR[0];
    double vL[3] = {UL[1] * il, UL[2] * il, UL[3] * il};
    double vR[3] = {UR[1] * ir, UR[2] * ir, UR[3] * ir};
    double pL = eos.GetPressure(UL[0], UL[4] * il - (vL[0] * vL[0] + vL[1] * vL[1] + vL[2] * vL[2]) / 2.0);
    double pR = eos.GetPressure(UR[0], UR[4] * ir - (vR[0] * vR[0] + vR[1] * vR[1] + vR[2] * vR[2]) / 2.0);
    double uL = vL[D], uR = vR[D];
    StarDev st = star_state_dev(g, UL[0], uL, pL, UR[0], uR, pR);
    double ro, u, p;
    const double S = 0.0;
    bool left = st.vacuum ? (S <= 0.5 * (st.STL + st.STR)) : (st.u >= S);
    // The side of the contact the sample lies on, then one of four cases for that side K:
    // 0 its initial state, 1 its star state, 2 inside its rarefaction fan, 3 vacuum.  The
    // conditions are the reference's with the right side mirrored by sgn = -1 (exact), so one
    // fan evaluation serves both sides.
    const double sgn = left ? 1.0 : -1.0;
    const bool shockK = left ? st.shock_l : st.shock_r;
    const double SK = left ? st.SL : st.SR, SHK = left ? st.SHL : st.SHR, STK = left ? st.STL : st.STR;
    const double rK = left ? UL[0] : UR[0], uK = left ? uL : uR, pK = left ? pL : pR;
    int cs;
    if (shockK && !st.vacuum) cs = (sgn * (SK - S) >= 0.0) ? 0 : 1;
    else if (sgn * (SHK - S) > 0.0) cs = 0;
    else if (sgn * (S - STK) > 0.0) cs = st.vacuum ? 3 : 1;
    else cs = 2;
    if (cs == 0) { ro = rK; u = uK; p = pK; }
    else if (cs == 1) {
        // star density behind a shock (Toro eq. 4.50) or a rarefaction (eq. 4.53)
        const double prK = st.p / pK;
        if (shockK) {
            const double C = (g - 1.0) / (g + 1.0);
            ro = (prK + C) / (C * prK + 1.0) * rK;
        } else {
            ro = rK * dpow(prK, 1.0 / g);
        }
        u = st.u;
        p = st.p;
    }
    else if (cs == 3) { ro = 0; u = 0; p = 0; }
    else {
        // inside the fan (Toro eq. 4.56): C^(2/(g-1)) scales the density, and the pressure by
        // C^(2g/(g-1)) = C^(2/(g-1)) C^2
        const double aK = __builtin_sqrt(g * pK / rK);
        const double C = 2.0 / (g + 1) + sgn * ((g - 1) * (uK - S) / ((g + 1) * aK));
        const double w = dpow(C, 2.0 / (g - 1));
        ro = w * rK;
        u = (sgn * aK + 0.5 * (g - 1) * uK + S) * (2 / (g + 1));
        p = w * C * C * pK;
    }
    const double* vel = left ? vL : vR;
    double Vf[3] = {vel[0], vel[1], vel[2]};
    Vf[D] = vel[D] + (u - vel[D]);
    double Fd[NV];
    if (ro == 0.0) {
#pragma unroll
        for (int i = 0; i < NV; i++) Fd[i] = 0.0;
    } else {
        // Uface and F(Uface) (GodunovSolverPerfectGasEOS.h:416-427)
        double U0 = ro, U1 = ro * Vf[0], U2 = ro * Vf[1], U3 = ro * Vf[2];
        double U4 = ro * (eos.GetInternalEnergy(ro, p) + 0.5 * (Vf[0] * Vf[0] + Vf[1] * Vf[1] + Vf[2] * Vf[2]));
        double iro = 1.0 / U0;
        double vx = U1 * iro, vy = U2 * iro, vz = U3 * iro;
        double pp = eos.PressureOfEnergyDensity(U4 - U0 * (vx * vx + vy * vy + vz * vz) / 2.0);
        double vv[3] = {vx, vy, vz};
        double vn = vv[D];
        Fd[0] = U0 * vn;
        Fd[1] = U0 * vn * vx;
        Fd[2] = U0 * vn * vy;
        Fd[3] = U0 * vn * vz;
        Fd[1 + D] += pp;
        Fd[4] = vn * (U4 + pp);
    }
#pragma unroll
    for (int i = 0; i < NV; i++) F[i] = R(Fd[i]);
    lam = R(st.smax);
}

// Face eigenvalue only (the dt pass needs no flux).
template <int FLUX, int D, typename R>
__device__ inline R face_eigen(R gamma, const R (&UL)[NV], const R (&UR)[NV]) {
    if constexpr (FLUX == FX_ROE) {
        // |u.n| + c of the Roe average (RoeSolverPerfectGasEOS.h:60-88)
        const R gm1 = gamma - R(1);
        R il = frcp(UL[0]), ir = frcp(UR[0]);
        R vl[3] = {UL[1] * il, UL[2] * il, UL[3] * il};
        R vr[3] = {UR[1] * ir, UR[2] * ir, UR[3] * ir};
        R kl = R(0.5) * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]);
        R kr = R(0.5) * (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
        R hl = (UL[4] * il - kl) * gamma + kl;
        R hr = (UR[4] * ir - kr) * gamma + kr;
        R sl = fsqrt(UL[0]), sr = fsqrt(UR[0]);
        R isum = frcp(sl + sr);
        R ql = sl * isum, qr = sr * isum;
        R v[3] = {ql * vl[0] + qr * vr[0], ql * vl[1] + qr * vr[1], ql * vl[2] + qr * vr[2]};
        R h = ql * hl + qr * hr;
        R phi2 = R(0.5) * gm1 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        return __builtin_fabs(v[D]) + fsqrt(gm1 * h - phi2);
    } else if constexpr (FLUX == FX_HLLC) {
        const IdealGasEOS<R> eos(gamma);
        R il = frcp(UL[0]), ir = frcp(UR[0]);
        R ul = UL[1 + D] * il, ur = UR[1 + D] * ir;
        R vl[3] = {UL[1] * il, UL[2] * il, UL[3] * il};
        R vr[3] = {UR[1] * ir, UR[2] * ir, UR[3] * ir};
        R pl = eos.PressureOfEnergyDensity(internal_energy(UL[4], R(0.5) * UL[0] * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]), UL));
        R pr = eos.PressureOfEnergyDensity(internal_energy(UR[4], R(0.5) * UR[0] * (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]), UR));
        R al = eos.SoundSpeedOfPressure(pl, il), ar = eos.SoundSpeedOfPressure(pr, ir);
        R SL = __builtin_fmin(ul - al, ur - ar);
        R SR = __builtin_fmax(ul + al, ur + ar);
        return __builtin_fmax(__builtin_fabs(SL), __builtin_fabs(SR));
    } else {
        R F[NV], lam;
        godunov_flux<D, R>(gamma, UL, UR, F, lam);
        return lam;
    }
}

template <int FLUX, int D, typename R>
__device__ inline void face_flux(R gamma, R eps, const R (&UL)[NV], const R (&UR)[NV], R (&F)[NV],
                                 R& lam) {
    if constexpr (FLUX == FX_ROE) roe_flux<D, R>(gamma, eps, UL, UR, F, lam);
    else if constexpr (FLUX == FX_HLLC) hllc_flux<D, R>(gamma, UL, UR, F, lam);
    else godunov_flux<D, R>(gamma, UL, UR, F, lam);
}

}  // namespace tsd
