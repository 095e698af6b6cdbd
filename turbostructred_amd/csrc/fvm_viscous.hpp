// fvm_viscous.hpp -- viscous pre-pass of one stage: velocity gradients of the
// interior cells, then the viscous flux of every face, written for the stage kernel to subtract
// from its convective fluxes.
//
//   vis_gradients   ComputeGradientNonUniformStructured (kernel.h:918-938): central differences
//                   (u(i+1) - u(i-1)) / (x(i+1) - x(i-1)) per axis, one thread per cell, ghost
//                   neighbours from the periodic wrap or the boundary image.
//   vis_face_flux   ComputeViscousFluxes (ERK:668-874), one thread per face, one launch per axis:
//                   normal derivatives from the two cell values, tangential ones from the
//                   distance-weighted cell gradients.  A dummy cell's gradient is computed where it
//                   is needed, by the one face that uses it: the periodic partner's gradient
//                   (ExchangeGradients, ERK:135-300), or at a boundary 2 g_face - g_in, with g_face
//                   the least-squares gradient at the face centre over the boundary cell, its
//                   transverse neighbours and the dummy cell, anchored at the face value
//                   (U_in + U_dummy) / 2 (ComputeDummyCellGradients, ERK:303-665,
//                   GradientComputer.h:19-53, Stencil.h:28-61, BCGeneral.h:90-94).
// Both are HBM-bound elementwise passes (no reuse beyond L2): gradients read 5 and write 3 nd
// doubles per cell, the face pass reads 5 + 3 nd doubles per cell side and writes 4 per face.
#pragma once

#include "fvm_kernels.hpp"

namespace tsd {

__device__ inline bool vis_gate(const KArgs& a) { return a.v.gate_go ? a.ctl->go != 0 : a.ctl->active != 0; }

// state of a cell given by interior indices, at most one of them out of range by one: an
// interior cell, or a dummy cell (periodic partner or the boundary image of the adjacent cell)
template <int DIM>
__device__ inline void vis_load(const KArgs& a, int i, int j, int k, double (&u)[NV]) {
    const GridArgs& g = a.g;
    int I[3] = {i, j, k};
    int axis = -1, side = 0;
#pragma unroll
    for (int e = 0; e < DIM; e++) {
        const int n = g.n[e];
        if (I[e] < 0 || I[e] >= n) {
            side = I[e] < 0 ? 0 : 1;
            if (g.periodic[e]) {
                I[e] = I[e] < 0 ? I[e] + n : I[e] - n;
            } else {
                axis = e;
                I[e] = I[e] < 0 ? 0 : n - 1;
            }
        }
    }
    const unsigned c = (unsigned(I[2]) * unsigned(g.n[1]) + unsigned(I[1])) * unsigned(g.n[0]) + unsigned(I[0]);
    const double* U = (const double*)a.uin;
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = ld_plane(U, a.plane, v, c);
    if (axis >= 0) {
        // face index of a boundary side: the transverse indices, faster one first
        const int fidx = axis == 0 ? I[2] * g.n[1] + I[1] : (axis == 1 ? I[2] * g.n[0] + I[0] : I[1] * g.n[0] + I[0]);
        double t[NV];
        bc_image(g.bc[bc_marker(g, 2 * axis + side, fidx)], g.gamma, axis, u, t);
#pragma unroll
        for (int v = 0; v < NV; v++) u[v] = t[v];
    }
}

template <int DIM>
__global__ void __launch_bounds__(256) vis_gradients(KArgs a) {
    if (!vis_gate(a)) return;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.plane) return;
    const GridArgs& g = a.g;
    const int nx = g.n[0], ny = g.n[1];
    const int i = int(c % nx), j = int((c / nx) % ny), k = int(c / ((size_t)nx * ny));
    const int I[3] = {i, j, k};
#pragma unroll
    for (int e = 0; e < DIM; e++) {
        double up[NV], um[NV];
        vis_load<DIM>(a, i + (e == 0), j + (e == 1), k + (e == 2), up);
        vis_load<DIM>(a, i - (e == 0), j - (e == 1), k - (e == 2), um);
        const double dX = a.v.x[e][I[e] + 2] - a.v.x[e][I[e]];
#pragma unroll
        for (int d = 0; d < 3; d++)
            st_plane(a.v.grad, a.plane, d * DIM + e, unsigned(c), (up[1 + d] / up[0] - um[1 + d] / um[0]) / dX);
    }
}

// least squares through the normal equations (the same loops as the CPU restatement)
template <int DIM, int NP>
__device__ inline void vis_lsq(const double (&pt)[3], double value, const double (&pts)[NP][3], const double (&vals)[NP],
                               double (&grad)[3]) {
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, b[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < NP; q++) {
        double r[3];
#pragma unroll
        for (int e = 0; e < DIM; e++) r[e] = pt[e] - pts[q][e];
        const double rhs = value - vals[q];
#pragma unroll
        for (int e = 0; e < DIM; e++) {
#pragma unroll
            for (int f = 0; f < DIM; f++) M[e][f] += r[e] * r[f];
            b[e] += r[e] * rhs;
        }
    }
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int e = 0; e < DIM; e++) {
#pragma unroll
        for (int f = 0; f <= e; f++) {
            double s = M[e][f];
#pragma unroll
            for (int q = 0; q < f; q++) s -= L[e][q] * L[f][q];
            L[e][f] = (e == f) ? __builtin_sqrt(s) : s / L[f][f];
        }
    }
    double y[3];
#pragma unroll
    for (int e = 0; e < DIM; e++) {
        double s = b[e];
#pragma unroll
        for (int q = 0; q < e; q++) s -= L[e][q] * y[q];
        y[e] = s / L[e][e];
    }
#pragma unroll
    for (int e = DIM - 1; e >= 0; e--) {
        double s = y[e];
#pragma unroll
        for (int q = e + 1; q < DIM; q++) s -= L[q][e] * grad[q];
        grad[e] = s / L[e][e];
    }
#pragma unroll
    for (int e = DIM; e < 3; e++) grad[e] = 0.0;
}

// velocity gradient G[d][e] of a cell with interior indices I, of which index I[ax] may be -1 or
// n (a dummy cell across the faces normal to ax)
template <int DIM, int ax>
__device__ inline void vis_cell_grad(const KArgs& a, const int (&I)[3], double (&G)[3][3]) {
    const GridArgs& g = a.g;
    const int n = g.n[ax];
    const bool ghost = I[ax] < 0 || I[ax] >= n;
    int B[3] = {I[0], I[1], I[2]};
    if (ghost) B[ax] = g.periodic[ax] ? (I[ax] < 0 ? n - 1 : 0) : (I[ax] < 0 ? 0 : n - 1);
    const unsigned c = (unsigned(B[2]) * unsigned(g.n[1]) + unsigned(B[1])) * unsigned(g.n[0]) + unsigned(B[0]);
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int e = 0; e < 3; e++) G[d][e] = e < DIM ? ld_plane(a.v.grad, a.plane, d * DIM + e, c) : 0.0;
    if (!ghost || g.periodic[ax]) return;
    const int side = I[ax] < 0 ? 0 : 1;
    // face centre
    double fc[3];
#pragma unroll
    for (int e = 0; e < 3; e++) fc[e] = e < DIM ? a.v.x[e][B[e] + 1] : 0.0;
    const double hb = g.h[ax][B[ax] + 1];
    fc[ax] = side ? a.v.x[ax][B[ax] + 1] + 0.5 * hb : a.v.x[ax][B[ax] + 1] - 0.5 * hb;
    // stencil: the boundary cell (slot 0), +-1 along the first transverse axis (1, 2), +-1 along
    // the second (3, 4), the dummy cell (last) -- 2 DIM points for every axis, in the order of
    // Stencil::InternalBorderStencil, with constant slots so that the arrays stay in registers
    constexpr int NP = 2 * DIM;
    constexpr int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
    int S[NP][3];
#pragma unroll
    for (int q = 0; q < NP; q++) {
        const int e = (q == 1 || q == 2) ? t1 : ((q == 3 || q == 4) ? t2 : -1);
        const int d = (q == 1 || q == 3) ? 1 : -1;
#pragma unroll
        for (int c = 0; c < 3; c++) S[q][c] = (q == NP - 1) ? I[c] : B[c] + (c == e ? d : 0);
    }
    double pts[NP][3], vel[NP][3];
    double uin[NV], udum[NV];
#pragma unroll
    for (int q = 0; q < NP; q++) {
        double u[NV];
        vis_load<DIM>(a, S[q][0], S[q][1], S[q][2], u);
        if (q == 0) {
#pragma unroll
            for (int v = 0; v < NV; v++) uin[v] = u[v];
        }
        if (q == NP - 1) {
#pragma unroll
            for (int v = 0; v < NV; v++) udum[v] = u[v];
        }
#pragma unroll
        for (int e = 0; e < 3; e++) pts[q][e] = e < DIM ? a.v.x[e][S[q][e] + 1] : 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) vel[q][d] = u[1 + d] / u[0];
    }
    double Vf[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) Vf[v] = 0.5 * (uin[v] + udum[v]);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double vals[NP];
#pragma unroll
        for (int q = 0; q < NP; q++) vals[q] = vel[q][d];
        double gf[3];
        vis_lsq<DIM, NP>(fc, Vf[1 + d] / Vf[0], pts, vals, gf);
#pragma unroll
        for (int e = 0; e < 3; e++) G[d][e] = 2.0 * gf[e] - G[d][e];
    }
}

template <int DIM, int ax>
__global__ void __launch_bounds__(256) vis_face_flux(KArgs a) {
    if (!vis_gate(a)) return;
    const unsigned f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.v.nface[ax]) return;
    const GridArgs& g = a.g;
    const unsigned m0 = unsigned(g.n[0]) + (ax == 0), m1 = unsigned(g.n[1]) + (ax == 1);
    int R[3] = {int(f % m0), int((f / m0) % m1), int(f / (m0 * m1))};
    int L[3] = {R[0], R[1], R[2]};
    L[ax] -= 1;
    double UL[NV], UR[NV], GL[3][3], GR[3][3];
    vis_load<DIM>(a, L[0], L[1], L[2], UL);
    vis_load<DIM>(a, R[0], R[1], R[2], UR);
    vis_cell_grad<DIM, ax>(a, L, GL);
    vis_cell_grad<DIM, ax>(a, R, GR);
    const double* X = a.v.x[ax];
    const double* H = g.h[ax];
    const int ir = R[ax] + 1;  // index of the right cell in the ghost-inclusive arrays
    const double mu = a.v.mu;
    const double s_visc = (-2.0 / 3.0) * mu;
    const double dx = X[ir] - X[ir - 1];
    const double l = H[ir - 1] / dx;
    const double r = H[ir] / dx;
    double D[3][3], uf[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int e = 0; e < 3; e++) D[d][e] = 0.5 * (r * GL[d][e] + l * GR[d][e]);
        const double vl = UL[1 + d] / UL[0], vr = UR[1 + d] / UR[0];
        D[d][ax] = (vr - vl) / dx;
        uf[d] = 0.5 * (l * vr + r * vl);
    }
    const double tau_d = s_visc * (D[0][0] + D[1][1] + D[2][2]);
    double t[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int lo = d < ax ? d : ax, hi = d < ax ? ax : d;  // constants after unrolling
        t[d] = (d == ax) ? tau_d + 2 * mu * D[ax][ax] : mu * (D[lo][hi] + D[hi][lo]);
    }
    double* F = a.v.fv[ax];
    const size_t P = a.v.nface[ax];
    st_plane(F, P, 0, f, t[0]);
    st_plane(F, P, 1, f, t[1]);
    st_plane(F, P, 2, f, t[2]);
    st_plane(F, P, 3, f, uf[0] * t[0] + uf[1] * t[1] + uf[2] * t[2]);
}

}  // namespace tsd
