/*
 * fvoracle.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain-C, OpenMP CPU restatement of the reference's explicit residual
 * path (happy12/FiniteVolume), used by tests/ as the parity checker for the
 * HIP library and by bench.py's cpu_baseline leg as a timed CPU port.  The
 * product (finitevolume_amd, libfvk) never links or loads it.
 *
 * Takes the same structs as the product's C ABI (include/fvk.h).  Every
 * function cites the reference file:line whose algorithm and order of
 * operations it restates; the extensions (second order, RK, local dt) have
 * no reference code and are defined here and in the product identically.
 */
#ifndef FVORACLE_H_INCLUDED
#define FVORACLE_H_INCLUDED

#include <stdint.h>

#include "fvk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fvo_solver fvo_solver;

int fvo_create(const fvk_mesh_desc* mesh, const fvk_bc* bcs, int32_t n_bcs, const fvk_params* params,
               const fvk_initial_condition* ic, fvo_solver** out);
void fvo_destroy(fvo_solver* s);
int fvo_step(fvo_solver* s, int32_t nsteps, double* residual_hist, double* dt_hist);
int fvo_residual(const fvo_solver* s, double* norms);
int fvo_set_field(fvo_solver* s, const double* U4, const double* turb);
int fvo_get_field(const fvo_solver* s, double* U4, double* turb);
int fvo_set_cfl(fvo_solver* s, double cfl);
int fvo_residual_eval(fvo_solver* s, double* R);
int fvo_gradient(fvo_solver* s, int32_t scheme, const double* phi, const double* boundary_phi, double* grad);
int fvo_wall_distance(fvo_solver* s, double* dist);
int fvo_wall_distance_faces(const fvk_mesh_desc* mesh, const int32_t* wall_faces, int64_t n_wall, double* dist);
int fvo_wall_forces(fvo_solver* s, double rho_ref, double vx_ref, double vy_ref, double length_ref, double* cl,
                    double* cd);
int fvo_wall_reports(fvo_solver* s, const int32_t* wall_faces, int64_t n, const fvk_wall_reference* ref,
                     int32_t pressure_only, fvk_wall_sample* samples, fvk_wall_force_report* reports, int32_t cap,
                     int32_t* n_reports);
/* limited primitive gradients of the last residual evaluation (order 2):
 * [8*n_cells] per cell (drho/dx, drho/dy, du/dx, du/dy, dv/dx, dv/dy, dp/dx, dp/dy) and
 * the limiter values [4*n_cells] */
int fvo_last_gradients(const fvo_solver* s, double* grad, double* phi);

/* standalone physics */
void fvo_flux(int32_t scheme, int64_t n, const double* UL, const double* UR, const double* nxy, double gamma,
              double tol, int32_t max_iter, double* out);
void fvo_sa_source(int64_t n, const double* nut, const double* nu, const double* omega, const double* dist,
                   const double* grad_nut, const fvk_sa_constants* c, double* out, double* eddy);
void fvo_sst_source(int64_t n, const double* rho, const double* k, const double* omega, const double* nu,
                    const double* dist, const double* gu, const double* gv, const double* gk, const double* gw,
                    int32_t variant, const fvk_sst_constants* c, int32_t kato, double* out);
double fvo_next_cfl(double current, double cfl_max, const double* hist, int32_t n_hist, double cfl_min,
                    double growth, double shrink, int32_t window, double divergence_threshold);
double fvo_venkatakrishnan(double d1, double d2, double eps2);

#ifdef __cplusplus
}
#endif

#endif
