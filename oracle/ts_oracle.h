/*
 * ts_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * C interface of the CPU FP64 restatement of TurboStructured's explicit finite-volume step.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker.  The product (libts_fvm.so) never links or calls it.
 */
#ifndef TS_ORACLE_H
#define TS_ORACLE_H

#include <stddef.h>
#include "../include/ts_fvm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tso_ctx tso_ctx;

tso_ctx* tso_create(const ts_config* cfg, int n_threads);
void tso_destroy(tso_ctx* o);
size_t tso_state_size(const tso_ctx* o);
int tso_set_state(tso_ctx* o, const double* aos, size_t n);
int tso_get_state(tso_ctx* o, double* aos, size_t n);
int tso_set_time(tso_ctx* o, double time, long long iteration);
/* same contract as ts_step; dt_log (may be NULL) receives the dt of each executed step */
int tso_step(tso_ctx* o, int max_steps, double t_max, double next_snapshot_time,
             ts_step_info* out, double* dt_log);
/* grid arrays as built by the oracle (ghosts included) */
int tso_grid(const tso_ctx* o, int axis, double* coord, double* h, int n);
/* spectral radius and residual of the last forward-Euler / first stage, reference layout */
int tso_last_residual(const tso_ctx* o, double* residual, double* spectral_radius, size_t n);

/* point-wise pieces, for unit tests: U* are conserved 5-vectors, n a unit axis normal */
void tso_flux(int riemann, double gamma, double eps, const double* UL, const double* UR,
              const double* n, double* F, double* max_eig);
/* exact Riemann problem star state (Newton); returns 1 if the data generate vacuum */
int tso_star_state(double gamma, double rl, double ul, double pl, double rr, double ur,
                   double pr, double* pstar, double* ustar, double* rstarl, double* rstarr,
                   int* iterations);
/* sample the exact solution of a 1-D Riemann problem at speed s = x/t: rho, u, p */
void tso_exact_sample(double gamma, double rl, double ul, double pl, double rr, double ur,
                      double pr, double s, double* out3);
/* reconstruction of one interior cell from its 2*ndims axis neighbours (order x-,x+,y-,y+,z-,z+)
   on a uniform grid with spacings h[3]; returns vals, grads[5][3] (row major) and lim[5] */
void tso_reconstruct(int recon, int limiter, int ndims, const double* h, const double* center,
                     const double* neighbours, double* grads, double* lim);
/* boundary image of a state (getDummyValues) */
void tso_bc_image(const ts_bc_config* bc, double gamma, const double* U, const double* n,
                  double* out);
/* velocity gradients of the current state (ghost layers filled first), 9 per cell of the
   reference layout: grads[cell*9 + d*3 + e] = d(vel_d)/dx_e, ghosts included */
int tso_velocity_gradients(tso_ctx* o, double* grads, size_t n);
/* least-squares gradient (GradientComputer::ExecuteByLeastSquares): points is m x 3 */
void tso_lsq_gradient(int ndims, const double* point, double value, const double* points,
                      const double* values, int m, double* grad);

#ifdef __cplusplus
}
#endif
#endif
