// fvm_kernels3d.hpp -- fused finite-volume stage kernel for 3-D grids (gfx950).
//
// A workgroup is ZW = 8 wavefronts, one per z-plane k = k0-1+w; wavefronts 1..6 produce output
// (ZOUT = 6 planes per workgroup) and the two edge wavefronts only supply reconstructions (each
// loads the one extra centre row its outer z-neighbour needs).  Inside a plane the scheme is the
// 2-D one: 64 lanes along x (60 output columns) and a three-row window marching along y, with
// the values carried from row to row in per-wavefront LDS slots.
// z-neighbours travel through LDS once per row: the centre row of the state, then the z+ face
// state of each plane, then the z-face flux, which is read one row later after the next row's
// first barrier -- two barriers per row.  Each z-face flux is evaluated once, by the wavefront
// above it.  The residual of a cell is accumulated x, then y,
// then z (ERK:1232,1312,1392) and its spectral radius x-, x+, y-, y+, z-, z+ (ERK:1225-1248,
// 1305-1328, 1385-1408).  VIS instances add the viscous face fluxes of the pre-pass, the
// viscous spectral-radius term after each plus face (the z-face part travels through LDS with
// the z eigenvalue) and the body-force sources.
#pragma once

#include "fvm_kernels.hpp"

namespace tsd {

template <int RECON, int LIM, int FLUX, int MODE, typename R, typename TIn, typename TOut, int VIS = 0>
__global__ void __launch_bounds__(64 * ZW) fvm_stage3d(KArgs a) {
    using MT = ModeTraits<MODE>;
    StepCtl* ctl = a.ctl;
    __shared__ R lds_u[ZW][NV][64];  // centre row of every plane
    __shared__ R lds_z[ZW][NV][64];  // z+ face state of every plane
    __shared__ R lds_h[ZW][NV][64];  // flux through the z-face below every plane
    __shared__ R lds_l[ZW][64];      // eigenvalue of that face
    __shared__ double lds_vz[VIS ? ZW : 1][64];  // its viscous spectral-radius term before / V
    __shared__ double red[ZW];
    if constexpr (MT::sr) {
        if (!ctl->active) {
            if (blockIdx.x == 0 && threadIdx.x == 0) ctl->go = 0;
            return;
        }
    } else {
        if (!ctl->go) return;
        if constexpr (MODE == MODE_RN) {
            if (ctl->active) return;  // only in the step that ends the ts_step call
        }
    }
    const GridArgs& g = a.g;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar plane indices
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int nyb = (ny + a.ty - 1) / a.ty;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);  // one tile per workgroup
    const int strip = tile % a.nstrips;
    const int yb = (tile / a.nstrips) % nyb;
    const int zb = tile / (a.nstrips * nyb);
    const int k = zb * ZOUT - 1 + w;  // this wavefront's plane (may be a ghost plane)
    int ks, kside;
    map_index(k, nz, g.periodic[2], ks, kside);
    const bool out_plane = w >= 1 && w <= ZOUT && k < nz;

    const int xg = strip * STRIP - 2 + lane;
    int xs, xside;
    map_index(xg, nx, g.periodic[0], xs, xside);
    const bool x_out = lane >= 2 && lane < 2 + STRIP && xg < nx;
    const bool out_lane = out_plane && x_out;
    const bool ghost_x_lane = out_plane && xg == nx && lane >= 2 && lane <= 62;

    const R gamma = R(g.gamma), eps = R(g.eps);
    const double dt = MT::sr ? 0.0 : ctl->dt;

    CellGeo<3, R> geo;
    const double hx = g.h[0][xs + 1], ihx = g.ih[0][xs + 1];
    const double hz = g.h[2][ks + 1], ihz = g.ih[2][ks + 1];
    geo.hh[0] = R(0.5 * hx);
    geo.rc[0] = R(g.rc[0][xs]);
    geo.rl[0] = R(g.rl[0][xs]);
    geo.rr[0] = R(g.rr[0][xs]);
    geo.hh[2] = R(0.5 * hz);
    geo.rc[2] = R(g.rc[2][ks]);
    geo.rl[2] = R(g.rl[2][ks]);
    geo.rr[2] = R(g.rr[2][ks]);
    geo.eps2 = R(g.eps2);
    // per-cell Venkatakrishnan eps^2 on non-uniform grids (see GridArgs::ev)
    const bool ev_cells = RECON == RC_LINLIM && LIM == LIM_VENKAT && g.ev[0] != nullptr;
    const double ev_xz = ev_cells ? g.ev[0][xs] * g.ev[2][ks] : 0.0;

    const TIn* __restrict__ uin = (const TIn*)a.uin;
    const TOut* __restrict__ lin = (const TOut*)a.lin;
    TOut* uo = (TOut*)a.uout;
    const size_t P = a.plane;
    // FP32-residual stages 2 and 3 with the increment form (see stage_tile)
    constexpr bool INC_F2 = MODE == MODE_F2 && sizeof(TOut) == 4, INC_S3 = MODE == MODE_S3 && sizeof(TIn) == 4;
    const bool inc = (INC_F2 && a.dout != nullptr) || (INC_S3 && a.din != nullptr);
    // the centre L(U^n) (stage 2) or increment (stage 3) is loaded with the centre U^n
    using IncT = typename std::conditional<MODE == MODE_F2, TOut, TIn>::type;
    const IncT* incsrc = (MODE == MODE_F2) ? (const IncT*)a.lin : (const IncT*)a.din;

    // Stencil input at (xg, row r, plane kk), in two steps: the loads (issue_cell), and the
    // value formed from them with its ghost image (complete_cell).  Prefetched cells are
    // completed one iteration after their loads are issued, so that no arithmetic on the loaded
    // values makes the wavefront wait for them early.
    struct RawCell {
        TIn u[NV];
        TOut l[MODE == MODE_F2 ? NV : 1];
        int rs, yside, kks, zside;
    };
    auto issue_cell = [&](int r, int kk, RawCell& q) {
        map_index(r, ny, g.periodic[1], q.rs, q.yside);
        map_index(kk, nz, g.periodic[2], q.kks, q.zside);
        const unsigned off = (unsigned(q.kks) * unsigned(ny) + unsigned(q.rs)) * unsigned(nx) + unsigned(xs);
#pragma unroll
        for (int v = 0; v < NV; v++) q.u[v] = ld_plane(uin, P, v, off);
        if constexpr (MODE == MODE_F2) {
#pragma unroll
            for (int v = 0; v < NV; v++) q.l[v] = ld_plane(lin, P, v, off);
        }
    };
    auto complete_cell = [&](const RawCell& q, R (&u)[NV]) {
        if constexpr (MODE == MODE_F2) {
#pragma unroll
            for (int v = 0; v < NV; v++) u[v] = R(double(q.u[v]) + dt * double(q.l[v]));
        } else {
#pragma unroll
            for (int v = 0; v < NV; v++) u[v] = R(q.u[v]);
        }
        // one call per axis, so that the axis is a constant inside bc_image (a run-time axis
        // would index the state arrays dynamically and put them in scratch memory)
        R t[NV];
        bool img = true;
        if (xside != 0) bc_image(g.bc[bc_marker(g, xside < 0 ? 0 : 1, q.kks * ny + q.rs)], gamma, 0, u, t);
        else if (q.yside != 0) bc_image(g.bc[bc_marker(g, q.yside < 0 ? 2 : 3, q.kks * nx + xs)], gamma, 1, u, t);
        else if (q.zside != 0) bc_image(g.bc[bc_marker(g, q.zside < 0 ? 4 : 5, q.rs * nx + xs)], gamma, 2, u, t);
        else img = false;
        if (img) {
#pragma unroll
            for (int v = 0; v < NV; v++) u[v] = t[v];
        }
    };
    auto load_cell = [&](int r, int kk, R (&u)[NV]) {
        RawCell q;
        issue_cell(r, kk, q);
        complete_cell(q, u);
    };

    double block_min = __builtin_inf();
    double rn[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int y0 = yb * a.ty;
    const int yend = (y0 + a.ty < ny) ? y0 + a.ty : ny;
    R Um[NV], U0[NV], Up[NV];
    RawCell Pn;             // prefetched stencil row, completed one iteration later
    R Zx[NV];               // centre row of the z-neighbour no wavefront holds (edge wavefronts)
    RawCell Zq;             // its prefetch
    double Nc[NV];          // U^n of the row finalized in this iteration (SSP stages 2, 3)
    IncT Dc[NV];            // its L(U^n) or increment (FP32-residual stages 2, 3)
    R ytop[NV];             // +y face state of the previous row
    // carried to the next row in LDS, per wavefront and lane (no barrier needed): the y-flux
    // through the pending row's bottom face and the x and z parts of its residual
    __shared__ R carry[ZW][3][NV][64];
    R (&gbot)[NV][64] = carry[w][0];
    R (&dfx)[NV][64] = carry[w][1];
    R (&dhz)[NV][64] = carry[w][2];
    double srp = 0.0;       // x terms + bottom y term of the pending row's spectral radius
    R lzb = R(0), lzt = R(0);  // z-face eigenvalues of the pending row
    double vzt = 0.0;          // viscous spectral-radius term (times V) of its top z-face
    R lam_gb = R(0);
    load_cell(y0 - 2, k, Um);
    load_cell(y0 - 1, k, U0);
    issue_cell(y0, k, Pn);
#pragma unroll
    for (int v = 0; v < NV; v++) { Nc[v] = 0.0; dfx[v][lane] = dhz[v][lane] = gbot[v][lane] = R(0); }
    const bool edge = (w == 0) || (w == ZW - 1);
    const int zoff = (w == 0) ? -1 : 1;  // the z-neighbour an edge wavefront loads itself
    if (edge) issue_cell(y0 - 1, k + zoff, Zq);
    // the z-flux below this plane for the previous row is still in lds_h[w] when it is read
    R lamzc = R(0), Szc = R(0);
    R Szb = R(0);  // z-face area of the pending row (its y size differs from the current row's)
    bool prev_out = false;
    const int nit = yend - y0 + 2;
    for (int t = 0; t < nit; t++) {
        const int rc = y0 - 1 + t;  // row being reconstructed
        const bool row_out = rc >= y0 && rc < yend;
        complete_cell(Pn, Up);
        if (t + 1 < nit) issue_cell(rc + 2, k, Pn);
        if (edge) {
            complete_cell(Zq, Zx);
            if (t + 1 < nit) issue_cell(rc + 1, k + zoff, Zq);
        }
        if constexpr (MT::use_un) {
            // U^n of the row finalized in this iteration (read earlier as a stencil row)
            if (out_plane && t >= 2) {
                const unsigned offn = (unsigned(ks) * unsigned(ny) + unsigned(rc - 1)) * unsigned(nx) + unsigned(xs);
#pragma unroll
                for (int v = 0; v < NV; v++) Nc[v] = ld_plane(a.un, P, v, offn);
                if (inc) {
#pragma unroll
                    for (int v = 0; v < NV; v++) Dc[v] = ld_plane(incsrc, P, v, offn);
                }
            }
        }
        int rcs, rcside;
        map_index(rc, ny, g.periodic[1], rcs, rcside);
        const double hy = g.h[1][rcs + 1];
        geo.hh[1] = R(0.5 * hy);
        geo.rc[1] = R(g.rc[1][rcs]);
        geo.rl[1] = R(g.rl[1][rcs]);
        geo.rr[1] = R(g.rr[1][rcs]);
        if (ev_cells) geo.eps2 = R(ev_xz * g.ev[1][rcs]);

        // (1) share the centre row, reconstruct
#pragma unroll
        for (int v = 0; v < NV; v++) lds_u[w][v][lane] = U0[v];
        __syncthreads();
        // (3) one row late: the flux through the face above the previous row, which the
        // wavefront above wrote before this barrier; it is read before the next barrier, after
        // which lds_h is written again (two barriers per row instead of three)
        if (out_plane && prev_out) {
            lzb = lamzc;
            Szb = Szc;
            lzt = lds_l[w + 1][lane];
            if constexpr (VIS) vzt = lds_vz[w + 1][lane];
            if constexpr (MT::flux) {
#pragma unroll
                for (int v = 0; v < NV; v++) dhz[v][lane] = (lds_h[w + 1][v][lane] - lds_h[w][v][lane]) * Szc;
            }
        }
        R face[6][NV];
        {
            R nb[6][NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                nb[0][v] = dpp_from_left(U0[v]);
                nb[1][v] = dpp_from_right(U0[v]);
                nb[2][v] = Um[v];
                nb[3][v] = Up[v];
                nb[4][v] = (w == 0) ? Zx[v] : lds_u[w - 1][v][lane];
                nb[5][v] = (w == ZW - 1) ? Zx[v] : lds_u[w + 1][v][lane];
            }
            reconstruct<3, RECON, LIM, R, FLUX != FX_ROE>(U0, nb, geo, face);
        }
        // (2) share the z+ face state; the wavefront above each z-face evaluates it
#pragma unroll
        for (int v = 0; v < NV; v++) lds_z[w][v][lane] = face[5][v];
        __syncthreads();
        const R Sz = R(hx * hy);  // face area of a z-face (ERK:1349)
        R H[NV], lamz = R(0);
#pragma unroll
        for (int v = 0; v < NV; v++) H[v] = R(0);
        if (w >= 1 && row_out) {
            R ZL[NV], ZR[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) { ZL[v] = lds_z[w - 1][v][lane]; ZR[v] = face[4][v]; }
            if (!g.periodic[2]) {
                if (k == 0) bc_image(g.bc[bc_marker(g, 4, rcs * nx + xs)], gamma, 2, ZR, ZL);
                if (k == nz) bc_image(g.bc[bc_marker(g, 5, rcs * nx + xs)], gamma, 2, ZL, ZR);
            }
            if constexpr (MT::flux) face_flux<FLUX, 2, R>(gamma, eps, ZL, ZR, H, lamz);
            else lamz = face_eigen<FLUX, 2, R>(gamma, ZL, ZR);
            if constexpr (VIS && MT::flux) {
                if (a.v.viscous) {
                    const int kf = k < 0 ? 0 : (k > nz ? nz : k);
                    const unsigned fi = (unsigned(kf) * unsigned(ny) + unsigned(rcs)) * unsigned(nx) + unsigned(xs);
#pragma unroll
                    for (int v = 1; v < NV; v++) H[v] -= R(ld_plane(a.v.fv[2], a.v.nface[2], v - 1, fi));
                }
            }
            if constexpr (VIS && MT::sr) {
                // max(4/(3 rho_f), gamma/rho_f) mu Sz^2; the plane below divides by its volume
                const double ro = __builtin_sqrt(double(ZL[0]) * double(ZR[0]));
                const double va = 4.0 / (3.0 * ro), vb = g.gamma / ro;
                double vz = (va < vb) ? vb : va;
                vz *= a.v.mu / 1.0;
                vz *= double(Sz) * double(Sz);
                lds_vz[w][lane] = vz;
            }
            if constexpr (MT::flux) {
#pragma unroll
                for (int v = 0; v < NV; v++) lds_h[w][v][lane] = H[v];
            }
            lds_l[w][lane] = lamz;
        }
        lamzc = lamz;
        Szc = Sz;
        prev_out = row_out;
        if constexpr (MT::sr) {
            // top ghost plane: only the face below it (ERK:1385)
            if (row_out && k == nz && w >= 1 && x_out)
                block_min = __builtin_fmin(block_min, dt_candidate(hx * hy * g.h[2][nz + 1], double(Sz * lamz)));
        }

        if (out_plane && t >= 1) {
            // y-face between rows rc-1 and rc
            R GL[NV], GR[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) { GL[v] = ytop[v]; GR[v] = face[2][v]; }
            if (!g.periodic[1]) {
                if (rc == 0) bc_image(g.bc[bc_marker(g, 2, ks * nx + xs)], gamma, 1, GR, GL);
                if (rc == ny) bc_image(g.bc[bc_marker(g, 3, ks * nx + xs)], gamma, 1, GL, GR);
            }
            const R Sy = R(hx * hz);  // face area of a y-face (ERK:1269)
            R G[NV], lam;
            if constexpr (MT::flux) face_flux<FLUX, 1, R>(gamma, eps, GL, GR, G, lam);
            else lam = face_eigen<FLUX, 1, R>(gamma, GL, GR);
            if constexpr (VIS && MT::flux) {
                if (a.v.viscous) {
                    const unsigned fi = (unsigned(ks) * unsigned(ny + 1) + unsigned(rc)) * unsigned(nx) + unsigned(xs);
#pragma unroll
                    for (int v = 1; v < NV; v++) G[v] -= R(ld_plane(a.v.fv[1], a.v.nface[1], v - 1, fi));
                }
            }
            if constexpr (MT::sr) {
                if (t >= 2) {
                    double sr;
                    {
#pragma clang fp contract(off)
                        sr = srp + double(Sy * lam);
                    }
                    if constexpr (VIS)
                        sr = sr + viscous_sr(double(GL[0]), double(GR[0]), g.gamma, a.v.mu, double(Sy), hx * g.h[1][rc] * hz);
                    {
#pragma clang fp contract(off)
                        sr = sr + double(Szb * lzb);
                        sr = sr + double(Szb * lzt);
                        if constexpr (VIS) sr = sr + vzt / (hx * g.h[1][rc] * hz);
                    }
                    if (out_lane) block_min = __builtin_fmin(block_min, dt_candidate(hx * g.h[1][rc] * hz, sr));
                }
                if (rc == ny && out_lane)  // top ghost row (ERK:1305)
                    block_min = __builtin_fmin(block_min, dt_candidate(hx * g.h[1][ny + 1] * hz, double(Sy * lam)));
                lam_gb = lam;
            }
            if constexpr (MT::flux) {
                if (t >= 2) {
                    // finalize row rc-1 (held in Um)
                    const int rf = rc - 1;
                    const double ivol = ihx * g.ih[1][rf + 1] * ihz;
                    const unsigned off = (unsigned(ks) * unsigned(ny) + unsigned(rf)) * unsigned(nx) + unsigned(xs);
                    R resv[NV];
#pragma unroll
                    for (int v = 0; v < NV; v++) resv[v] = (dfx[v][lane] - (G[v] - gbot[v][lane]) * Sy) - dhz[v][lane];
                    if constexpr (VIS) add_sources(a.v, Um, hx * g.h[1][rf + 1] * hz, resv);
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        R res = resv[v];
                        if constexpr (MODE == MODE_RN) {
                            if (out_lane) rn[v] += hx * g.h[1][rf + 1] * hz * __builtin_fabs(double(res));
                        } else if constexpr (MODE == MODE_F1) {
                            if (out_lane) st_plane(uo, P, v, off, TOut(double(res) * ivol));
                        } else if (inc && MODE == MODE_F2) {
                            // FP32-residual stage 2: the increment dt/4 (L(U^n) + L(U1)) in float
                            // beside U2 (stage 3 accumulates U^n + 2/3 (D + dt L(U2)) in double)
                            const double d = 0.25 * dt * (double(Dc[v]) + double(res) * ivol);
                            if (out_lane) {
                                st_plane((TOut*)a.dout, P, v, off, TOut(d));
                                st_plane(uo, P, v, off, TOut(Nc[v] + d));
                            }
                        } else if (inc && MODE == MODE_S3) {
                            const double d = double(Dc[v]);
                            const double q = Nc[v] + (2.0 / 3.0) * (d + dt * (double(res) * ivol));
                            if (out_lane) st_plane(uo, P, v, off, TOut(q));
                        } else {
                            double q = double(Um[v]) + dt * (double(res) * ivol);
                            if constexpr (MT::use_un) q = MT::ca * Nc[v] + MT::cb * q;
                            if (out_lane) st_plane(uo, P, v, off, TOut(q));
                        }
                    }
                }
#pragma unroll
                for (int v = 0; v < NV; v++) gbot[v][lane] = G[v];
            }
            if (row_out) {
                // x-faces of row rc
                R UL[NV], UR[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    UL[v] = dpp_from_left(face[1][v]);
                    UR[v] = face[0][v];
                }
                if (!g.periodic[0]) {
                    if (xg == 0) bc_image(g.bc[bc_marker(g, 0, ks * ny + rcs)], gamma, 0, UR, UL);
                    if (xg == nx) bc_image(g.bc[bc_marker(g, 1, ks * ny + rcs)], gamma, 0, UL, UR);
                }
                const R Sx = R(hy * hz);  // face area of an x-face (ERK:1190)
                R F[NV], lamx;
                if constexpr (MT::flux) face_flux<FLUX, 0, R>(gamma, eps, UL, UR, F, lamx);
                else lamx = face_eigen<FLUX, 0, R>(gamma, UL, UR);
                if constexpr (VIS && MT::flux) {
                    if (a.v.viscous) {
                        const unsigned fi = (unsigned(ks) * unsigned(ny) + unsigned(rcs)) * unsigned(nx + 1) +
                                            unsigned(xg < 0 ? 0 : (xg > nx ? nx : xg));
#pragma unroll
                        for (int v = 1; v < NV; v++) F[v] -= R(ld_plane(a.v.fv[0], a.v.nface[0], v - 1, fi));
                    }
                }
                if constexpr (MT::sr) {
                    R lam_r = dpp_from_right(lamx);
                    {
#pragma clang fp contract(off)
                        srp = double(Sx * lamx) + double(Sx * lam_r);
                    }
                    if constexpr (VIS) {
                        const double rl = dpp_from_right(double(UL[0])), rr = dpp_from_right(double(UR[0]));
                        srp = srp + viscous_sr(rl, rr, g.gamma, a.v.mu, double(Sx), hx * hy * hz);
                    }
                    {
#pragma clang fp contract(off)
                        srp = srp + double(Sy * lam_gb);
                    }
                    if (ghost_x_lane)  // right ghost column (ERK:1225)
                        block_min = __builtin_fmin(block_min, dt_candidate(g.h[0][nx + 1] * hy * hz, double(Sx * lamx)));
                }
                if constexpr (MT::flux) {
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        R Fr = dpp_from_right(F[v]);
                        dfx[v][lane] = -((Fr - F[v]) * Sx);
                    }
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
            ytop[v] = face[3][v];
            Um[v] = U0[v];
            U0[v] = Up[v];
        }
    }
    if constexpr (MT::sr) {
        // the workgroup's minimum, stored for its tile; dt_finalize_kernel follows the launch
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) block_min = __builtin_fmin(block_min, __shfl_xor(block_min, o));
        if (lane == 0) red[w] = block_min;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = red[0];
            for (int q = 1; q < ZW; q++) m = __builtin_fmin(m, red[q]);
            st_plane(a.dt_part, 0, 0, unsigned(tile), m);
        }
    }
    if constexpr (MODE == MODE_RN) {
        __shared__ double rred[ZW * NV];
        rn_reduce_finalize(ctl, a.rn_part, rn, rred, lane, w, ZW);
    }
}

}  // namespace tsd
