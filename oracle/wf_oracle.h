/*
 * wf_oracle.h -- FP64 CPU oracle of the wildfire time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C/OpenMP restatement of the
 * reference's CUDA time step (src/cu/solver.cu, pde.cu, turbulence.cu,
 * pressure.cu, functions.cu, ibm.cu) in the reference's own data layout.  It is
 * the checker of the HIP library and the timed CPU baseline; nothing in the
 * product links, loads or calls it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg use it.
 *
 * Layout: scalar fields IDX(i,j,k) = k + Nz*(j + Ny*i); state y = [u|v|w|T|Y],
 * Y compact at Nz_Y_max levels; spectral buffers FFTWIDX(i,j,k) = j + (Ny-1)*i +
 * (Nx-1)*(Ny-1)*k over the (Nx-1) x (Ny-1) x (Nz-1) unique pressure nodes.
 */
#ifndef WF_ORACLE_H
#define WF_ORACLE_H

#include "../include/wildfires.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

/* Workspace for one parameter set (stage buffers, turbulence planes, spectra). */
oracle_ctx *oracle_create(const Parameters *P);
void oracle_destroy(oracle_ctx *ctx);

/* IBM mask restated from src/cu/ibm.cu:57-95 (bit-exact integer outputs). */
void oracle_ibm(const Parameters *P, const double *topography, int *cut_nodes, int *Nz_Y, double *z_ibm,
                int *Nz_Y_max);

/* src/cu/pde.cu:139-787: RHS of u,v,w,T (and Y where k < Nz_Y) plus the 18
 * turbulence planes and the u_tau plane (R_turb has 18*N + Nx*Ny entries). */
void oracle_rhs(const Parameters *P, const double *R_old, double *R_new, double *R_turb);
/* src/cu/turbulence.cu:255-454: SGS terms added to R_new from R_turb. */
void oracle_turbulence(const Parameters *P, const double *R_turb, double *R_new);
/* src/cu/pde.cu:14-99: cut-node, top and dead-node conditions on a state-shaped vector. */
void oracle_boundary_conditions(const Parameters *P, double *R_new);
/* src/cu/pde.cu:910-917 */
void oracle_phi(oracle_ctx *ctx, const double *R_old, double *R_new);
/* src/cu/pde.cu:101-137 */
void oracle_bounds(const Parameters *P, double *y);
/* src/cu/pde.cu:789-908 */
void oracle_velocity_correction(const Parameters *P, double *y_np1, const double *y_n, const double *p);
/* src/cu/pressure.cu:506-606 with src/cu/functions.cu:431-450: pressure solve,
 * iterative when variable_density == 1.  *iterations gets the reference's
 * reported count (src/cu/pressure.cu:601); *solves the number of solves run. */
void oracle_pressure(oracle_ctx *ctx, const double *y_np1, const double *y_n, double *p, double *error,
                     int *iterations, int *solves);
/* One step n (src/cu/solver.cu:178-265 without I/O): y_n and p updated in place.
 * Returns 0, or -1 for an unknown method. */
int oracle_step(oracle_ctx *ctx, double *y_n, double *p, int n);
/* Steps step0 .. step0+n_steps-1. */
int oracle_run(oracle_ctx *ctx, double *y_n, double *p, int step0, int n_steps);
/* Sum of pressure solves over the steps run so far. */
long oracle_pressure_solves(const oracle_ctx *ctx);

/* Learned subgrid closure (restating wildfires_amd/csrc/hip/wf_closure.h):
 * params is the WF_CLOSURE_NPARAMS block of include/wildfires.h, kept by
 * pointer (NULL: plain Smagorinsky).  One closure per process. */
void oracle_set_closure(const float *params);
/* The network alone on n rows of the 12 raw inputs (see wf_solver_set_closure). */
void oracle_closure_eval(const float *params, const double *features, long n, double *nu_t, int *learned);
/* Training loss weight * sum (softplus(z) - nu_target/out_scale)^2 over n rows
 * and its gradient (a parameter-block-shaped float array; the normalisation,
 * range and out_scale entries are 0), restating the GPU trainer. */
void oracle_closure_gradient(const float *params, const double *features, const double *nu_target, long n,
                             double weight, float *grad, double *loss);

/* In-place complex DFT of length n over interleaved (re,im) doubles at the given
 * element stride; sign -1 forward, +1 inverse, no normalisation. */
void oracle_fft(double *data, int n, int stride, int sign);
/* Tridiagonal coefficients a (Nz-2), b (Nz-1), c (Nz-2) of one (r,s) system
 * (src/cu/pressure.cu:356-390). */
void oracle_tridiagonal_coefficients(const Parameters *P, int r, int s, double *a, double *b, double *c);
/* Thomas solve of one system with real a,b,c and complex right-hand side d
 * (interleaved), result in x (src/cu/pressure.cu:418-443); n = Nz-1 unknowns. */
void oracle_thomas(int n, const double *a, const double *b, const double *c, const double *d, double *x);

/* Number of OpenMP threads the oracle uses. */
int oracle_threads(void);

#ifdef __cplusplus
}
#endif

#endif
