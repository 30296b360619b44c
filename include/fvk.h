/*
 * fvk.h -- C ABI of the MI355X (gfx950) explicit finite-volume residual
 * library.  Plain pointers and sizes only; no C++ or torch types cross it.
 *
 * The entry points replace the public API of the reference solver classes
 * (happy12/FiniteVolume), one opaque fvk_solver per solver object:
 *
 *   fvk_create        <- EulerFVMSolver::EulerFVMSolver        include/EulerFVMSolver.h:107-112
 *                        NavierStokesFVMSolver::...            include/NavierStokesFVMSolver.h:118-121
 *                        RANSTurbulenceSASolver::...           include/RANSTurbulenceSASolver.h:107-111
 *                        RANSTurbulenceSSTSolver::...          src/RANSTurbulenceSSTSolver.cpp:40-46
 *                        UnstructuredFVMSolver::...            include/UnstructuredFVMSolver.h:34-35
 *                        AdvectionDiffusionFVMSolver::...      src/AdvectionDiffusionFVMSolver.cpp:16-31
 *   fvk_step          <- *::step() / *::run(int)               src/EulerFVMSolver.cpp:58-121
 *   fvk_residual      <- *::residual(), nut_residual(),
 *                        k_residual(), omega_residual(),
 *                        UnstructuredFVMSolver::residual_norm() include/EulerFVMSolver.h:141
 *   fvk_get_field     <- *::field(), nut_field(), k_field(),
 *                        omega_field()                         include/EulerFVMSolver.h:129
 *   fvk_set_field     <- *::set_field(...)                     include/EulerFVMSolver.h:151-155
 *   fvk_set_cfl       <- *::set_cfl(double)                    include/EulerFVMSolver.h:160
 *   fvk_gas_gamma     <- *::gas_gamma()                        include/EulerFVMSolver.h:134
 *   fvk_gradient      <- GradientCalculator::compute           src/GradientReconstruction.cpp:83-150
 *   fvk_wall_distance <- compute_wall_distance                 src/WallDistance.cpp:26-44
 *
 * Entry points with no counterpart in the reference (extensions that
 * reduce to the reference path when switched off, see fvk_params):
 *   fvk_residual_eval, fvk_residual_vjp, fvk_wall_forces, fvk_wall_forces_vjp,
 *   fvk_adjoint_solve_forces, fvk_batch_*.
 *
 * Conventions (same as the reference):
 *  - cell and face indices are the caller's (reference) order; the library
 *    renumbers internally and maps back on every copy in or out;
 *  - a NaN/Inf residual is NOT an error: it is returned in the residual
 *    norms so the caller's divergence check stays in charge
 *    (src/main.cpp:65-69, 3496);
 *  - every function returns FVK_OK (0) or a nonzero status; fvk_last_error()
 *    gives a message for the last nonzero status on the calling thread.
 */
#ifndef FVK_H_INCLUDED
#define FVK_H_INCLUDED

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVK_ABI_VERSION 1

enum fvk_status {
    FVK_OK = 0,
    FVK_ERR_INVALID = 1,     /* bad argument, size mismatch, unknown enum value */
    FVK_ERR_HIP = 2,         /* a HIP runtime call failed */
    FVK_ERR_NOMEM = 3,       /* device or host allocation failed */
    FVK_ERR_UNSUPPORTED = 4, /* option combination not built */
    FVK_ERR_DIVERGED = 5     /* an iterative solve left the finite numbers */
};

/* Which solver class the handle stands in for. */
enum fvk_equation {
    FVK_EQ_EULER = 0,          /* EulerFVMSolver */
    FVK_EQ_NAVIER_STOKES = 1,  /* NavierStokesFVMSolver */
    FVK_EQ_RANS_SA = 2,        /* RANSTurbulenceSASolver */
    FVK_EQ_RANS_SST = 3,       /* RANSTurbulenceSSTSolver */
    FVK_EQ_DIFFUSION = 4,      /* UnstructuredFVMSolver (scalar diffusion) */
    FVK_EQ_ADVECTION_DIFFUSION = 5 /* AdvectionDiffusionFVMSolver (scalar, uniform velocity) */
};

/* NumericalFluxScheme, include/EulerFVMSolver.h:40-44 (same order). */
enum fvk_flux_scheme { FVK_FLUX_RUSANOV = 0, FVK_FLUX_HLLC = 1, FVK_FLUX_EXACT = 2 };

/* GradientScheme, include/GradientReconstruction.h (same order). */
enum fvk_gradient_scheme { FVK_GRAD_GREEN_GAUSS = 0, FVK_GRAD_LEAST_SQUARES = 1 };

/* EulerBoundaryType (include/EulerFVMSolver.h:17-21) and NSBoundaryType
 * (include/NavierStokesFVMSolver.h) share codes: 0 = wall (slip for Euler,
 * no-slip for NS/RANS), 1 = farfield, 2 = outflow.  For FVK_EQ_DIFFUSION the
 * codes are BoundaryType (include/UnstructuredMesh.h): 0 = Dirichlet,
 * 1 = Neumann, and fvk_bc.value is the prescribed value/flux. */
enum fvk_bc_type { FVK_BC_WALL = 0, FVK_BC_FARFIELD = 1, FVK_BC_OUTFLOW = 2 };
enum fvk_diffusion_bc_type { FVK_BC_DIRICHLET = 0, FVK_BC_NEUMANN = 1 };

/* SSTLimiterVariant, include/SSTKOmega.h (same order). */
enum fvk_sst_variant { FVK_SST_VORTICITY = 0, FVK_SST_STRAIN_RATE = 1 };

/* EulerICMode, include/EulerFVMSolver.h:32-35 (same order). */
enum fvk_ic_mode { FVK_IC_FREESTREAM = 0, FVK_IC_TWO_REGION = 1, FVK_IC_NONE = 2 };

/* Extension switches (no reference code; defaults reproduce the reference). */
/* FVK_LIMITER_FIRST_ORDER sets every slope to zero (phi = 0): the second-order
 * path then reproduces the first-order (reference) step up to rounding (its
 * fluxes take primitive face states). */
enum fvk_limiter { FVK_LIMITER_NONE = 0, FVK_LIMITER_VENKATAKRISHNAN = 1, FVK_LIMITER_FIRST_ORDER = 2 };

/*
 * Mesh, in the reference's UnstructuredMesh layout (include/UnstructuredMesh.h:
 * Node :20, Face :28, Cell :39) flattened to arrays.  Face
 * geometry (normal, length, midpoint) is computed by the library from the
 * node coordinates, as the reference solver constructors do
 * (UnstructuredMesh::compute_geometry, include/UnstructuredMesh.h:89-105);
 * cell volumes and centroids are taken as given (the reference's mesh
 * reader fills them, src/MeshReader.cpp:56-76).
 */
typedef struct fvk_mesh_desc {
    int64_t n_nodes, n_cells, n_faces;
    int32_t n_patches;             /* number of boundary patches (bcs[] length) */
    int32_t reserved0;
    const double* node_xy;         /* [2*n_nodes]  x,y */
    const int32_t* face_nodes;     /* [2*n_faces]  node1,node2 (cell_left's CCW walk) */
    const int32_t* face_cells;     /* [2*n_faces]  cell_left, cell_right (-1 = boundary) */
    const int32_t* face_patch;     /* [n_faces]    patch id (-1 = interior) */
    const int32_t* cell_face_ptr;  /* [n_cells+1]  CSR offsets into cell_faces */
    const int32_t* cell_faces;     /* Cell::faces, in the reference order */
    const double* cell_volume;     /* [n_cells] */
    const double* cell_xy;         /* [2*n_cells]  centroid */
    /* Optional (NULL = none): translation added to cell_right's centroid to
     * place it across the face as seen from cell_left (periodic face pairs).
     * Only gradient/reconstruction geometry uses it. [2*n_faces] */
    const double* face_shift;
} fvk_mesh_desc;

/* Per-patch boundary condition (EulerBoundaryCondition, NSBoundaryCondition,
 * RANSBoundaryConditionSA/SST, BoundaryPatch), indexed by patch id. */
typedef struct fvk_bc {
    int32_t type;                  /* fvk_bc_type, or fvk_diffusion_bc_type */
    int32_t is_isothermal_wall;    /* NS/RANS no-slip wall: 0 adiabatic, 1 isothermal */
    double wall_u, wall_v;         /* NS/RANS no-slip wall velocity */
    double wall_temperature;       /* NS/RANS isothermal wall */
    double farfield_state[4];      /* conserved rho, rho_u, rho_v, E */
    double farfield_nut;           /* SA farfield nut */
    double farfield_k, farfield_omega; /* SST farfield k, omega */
    double value;                  /* diffusion: Dirichlet value or Neumann flux */
} fvk_bc;

/* SAModelConstants, include/SpalartAllmaras.h:13-23. */
typedef struct fvk_sa_constants {
    double cb1, cb2, sigma, kappa, cw2, cw3, cv1, cv2, cv3;
} fvk_sa_constants;

/* SSTModelConstants, include/SSTKOmega.h. */
typedef struct fvk_sst_constants {
    double beta_star, kappa, a1, sigma_k1, sigma_k2, sigma_omega1, sigma_omega2, beta1, beta2;
} fvk_sst_constants;

typedef struct fvk_params {
    int32_t equation;              /* fvk_equation */
    int32_t flux_scheme;           /* fvk_flux_scheme */
    int32_t gradient_scheme;       /* fvk_gradient_scheme (NS/RANS) */
    int32_t exact_riemann_max_iter;
    double gamma, gas_constant, mu, prandtl, prandtl_t, cfl;
    double exact_riemann_tol;
    fvk_sa_constants sa;
    fvk_sst_constants sst;
    int32_t sst_variant;           /* fvk_sst_variant */
    int32_t kato_launder;
    double alpha, fixed_dt;        /* FVK_EQ_DIFFUSION: diffusivity, time step */
    /* ---- extensions; the defaults (1, NONE, 1, 0) are the reference method ---- */
    int32_t order;                 /* 1: first order (reference); 2: MUSCL with LSQ primitive gradients (Euler) */
    int32_t limiter;               /* fvk_limiter (order 2 only) */
    double venkat_k;               /* Venkatakrishnan K; eps^2 = (K*sqrt(V))^3 */
    int32_t rk_stages;             /* 1: forward Euler (reference); 2: SSP-RK2 (Heun); 3: SSP-RK3 (Shu-Osher) */
    int32_t local_dt;              /* 0: one global CFL dt (reference); 1: per-cell local dt */
    /* ---- execution tuning (same results bit for bit; 0 = library default) ---- */
    int32_t kernel_variant;        /* Euler: 1 face sweep + per-cell gather, 2 per-cell fused flux and update,
                                      3 one fused pass per cell tile with its halo staged in LDS (0: 1);
                                      NS/RANS: 1 gradient, face and cell passes, 3 one fused pass per tile
                                      computing its gradients (cells of at most 16 faces); a global dt adds
                                      a dt pass before the first stage's tile pass;
                                      batches also 4: variant 3 as one persistent launch per stage */
    int32_t block_size;            /* threads per workgroup of the Euler kernels: 64, 128, 256 or 512,
                                      for variant 3 any multiple of 64 up to 512 (0: 512; variant 3: 256) */
    int32_t tile_cells;            /* variant 3: most cells of one tile, at most block_size (0: block_size) */
    int32_t reserved1;
    /* FVK_EQ_ADVECTION_DIFFUSION: the uniform advection velocity (u_adv,
     * v_adv of src/AdvectionDiffusionFVMSolver.cpp:16-20); alpha, fixed_dt
     * and gradient_scheme as above */
    double advection_u, advection_v;
} fvk_params;

/* Initial condition (EulerInitialCondition, include/EulerFVMSolver.h:59-67,
 * plus the turbulence initial values and the diffusion initial blob). */
typedef struct fvk_initial_condition {
    int32_t mode;                  /* fvk_ic_mode */
    int32_t reserved0;
    double rho, u, v, p;
    double rho_l, u_l, v_l, p_l;
    double rho_r, u_r, v_r, p_r;
    double x0;
    double initial_nut, initial_k, initial_omega;
    double initial_value, initial_radius; /* diffusion */
} fvk_initial_condition;

typedef struct fvk_solver fvk_solver;

/* Reference defaults: gamma 1.4, R 287-free (1.0), mu 0, Pr 0.72, Pr_t 0.9,
 * cfl 0.5, Rusanov, least squares, exact tol 1e-6 / 20 iterations, SA and SST
 * constants of include/SpalartAllmaras.h:13-23 and include/SSTKOmega.h,
 * first order, forward Euler, global dt. */
void fvk_params_default(fvk_params* p);
void fvk_initial_condition_default(fvk_initial_condition* ic);

/* Number of conserved variables per cell for an equation: 4 (Euler, NS),
 * 5 (SA: + nut), 6 (SST: + k, omega), 1 (diffusion, advection-diffusion). */
int fvk_num_vars(int32_t equation);

int fvk_create(const fvk_mesh_desc* mesh, const fvk_bc* bcs, int32_t n_bcs, const fvk_params* params,
               const fvk_initial_condition* ic, fvk_solver** out);
void fvk_destroy(fvk_solver* s);

/* Advance nsteps steps (one reference step() each).  residual_hist, if not
 * NULL, receives nsteps * fvk_num_vars rows of L2 residual norms; dt_hist,
 * if not NULL, receives the nsteps time steps (for local_dt: the minimum
 * local dt).  One device->host copy at the end of the call.  The call
 * returns once those rows are final: with three or more RK stages the last
 * step's later stages may still be running.  Every other call on the solver
 * waits for them before it reads or writes the device. */
int fvk_step(fvk_solver* s, int32_t nsteps, double* residual_hist, double* dt_hist);

/* Waits for all work queued on the solver's stream (what fvk_step or
 * fvk_wall_force_reports_noalloc left running).  No reference counterpart
 * (the reference's step() returns with the state updated); for callers
 * that time the device or hand its stream's results to other code. */
int fvk_synchronize(fvk_solver* s);

/* Last step's residual norms [fvk_num_vars]: rho, rho_u, rho_v, E (, nut | k, omega);
 * all zero before the first step and after fvk_set_field. */
int fvk_residual(const fvk_solver* s, double* norms);

/* Conserved field in reference cell order.  U4 is AoS [4*n_cells]
 * (EulerState rho, rho_u, rho_v, E; diffusion: phi [n_cells]); turb is
 * NULL or AoS [n_turb*n_cells]: nut (SA) or k, omega (SST).  set_field with
 * turb == NULL keeps the turbulence field, like the reference's
 * set_field(U) overloads. */
int fvk_set_field(fvk_solver* s, const double* U4, const double* turb, int64_t n_cells);
int fvk_get_field(const fvk_solver* s, double* U4, double* turb, int64_t n_cells);
/* The same two calls under the names of a state-vector binding: the nvars
 * doubles of each cell AoS, [nvars*n_cells] (Euler/NS: 4; SA: 5 with nut
 * last; SST: 6 with k, omega last; diffusion: 1). */
int fvk_set_state(fvk_solver* s, const double* U_aos, int64_t n_cells);
int fvk_get_state(const fvk_solver* s, double* U_aos, int64_t n_cells);

int fvk_set_cfl(fvk_solver* s, double cfl);
double fvk_gas_gamma(const fvk_solver* s);

/* ---- the driver contract (src/main.cpp:3400-3564 and its NS/SA/SST and
 * diffusion twins) ---- */

/* next_cfl, src/CflRamp.cpp:7-27 (CflRampParams fields as arguments). */
double fvk_next_cfl(double current_cfl, double cfl_max, const double* history, int32_t n_history, double cfl_min,
                    double growth_factor, double shrink_factor, int32_t window, double divergence_threshold);

/* CheckpointEquation, include/Checkpoint.h (same tags). */
enum fvk_checkpoint_equation {
    FVK_CKPT_DIFFUSION = 0, FVK_CKPT_EULER = 1, FVK_CKPT_ADVECTION_DIFFUSION = 2, FVK_CKPT_NAVIER_STOKES = 3,
    FVK_CKPT_RANS_SA = 4, FVK_CKPT_RANS_SST = 5
};

/* Checkpoint::write, src/Checkpoint.cpp: the 40-byte header ("FVMCKPT\0",
 * version 2, the solver's equation tag, cell count, step_index,
 * build_number), then the field in reference cell order: EulerState AoS
 * (4 doubles per cell; diffusion: phi), then nut (SA) or k then omega
 * (SST), one array each.  Byte for byte the reference's file. */
int fvk_checkpoint_write(fvk_solver* s, const char* path, int64_t step_index, uint64_t build_number);

/* Reads a checkpoint into the solver with the reference driver's resume
 * rules: diffusion needs a diffusion checkpoint (Checkpoint::read); a flow
 * solver takes any flow-field checkpoint's EulerState payload
 * (Checkpoint::read_flow_field) and the turbulence fields only from one of
 * its own model, else it keeps its own (set_field(U)).  Header, equation
 * and cell-count mismatches are errors with the reference's messages. */
int fvk_checkpoint_read(fvk_solver* s, const char* path, int64_t* step_index, uint64_t* build_number,
                        int32_t* source_equation);

/* The run loop of the reference driver (CaseInput fields it reads). */
typedef struct fvk_run_config {
    int64_t nsteps;                /* last step index to run ("nsteps") */
    int32_t cfl_ramp;              /* 1: cfl_mode = ramp, from ramp_cfl_min up to cfl */
    int32_t ramp_window;           /* CflRampParams::window (15) */
    double cfl;                    /* "cfl": the fixed CFL, or the ramp's ceiling */
    double ramp_cfl_min, ramp_growth, ramp_shrink, ramp_divergence_threshold; /* 0.05, 1.5, 0.5, 2.0 */
    double tolerance[6];           /* residual_tolerance_{rho, rho_u, rho_v, E, nut | k, omega}
                                      (diffusion: [0] = residual_tolerance); < 0: not checked */
    const char* residual_file;     /* NULL or "": no residual CSV */
    int32_t residual_interval;     /* a CSV row every N steps (1) */
    int32_t output_precision;      /* significant digits of every CSV double (6) */
    const char* checkpoint_file;   /* NULL or "": none; resumed from when it exists, written at the end */
    uint64_t build_number;         /* written into the checkpoint header */
} fvk_run_config;

typedef struct fvk_run_result {
    int64_t resume_start;          /* first step index run: 1, or the checkpoint's step + 1 */
    int64_t last_step;             /* last completed step index */
    int32_t status;                /* 0: ran to nsteps, 1: converged, 2: diverged (NaN/Inf residual) */
    int32_t resumed_equation;      /* the checkpoint's equation tag, -1 when nothing was resumed */
    double residual[6];            /* residual norms of the last step */
    double cfl;                    /* CFL after the last step */
} fvk_run_result;

void fvk_run_config_default(fvk_run_config* c);

/* The reference driver's loop on the device solver (src/main.cpp:3491-3557;
 * NS :3739-3830, SA :4018-4110, SST :4343-4440, diffusion :3139-3190):
 * auto-resume from checkpoint_file, then per step: one step, the NaN/Inf
 * bit test on every residual component, the CFL ramp on the density
 * residual, the CSV row (std::setprecision(output_precision), flushed),
 * the convergence test on every set tolerance; at the end a checkpoint
 * unless the run diverged.  Divergence is a status, not an error. */
int fvk_run(fvk_solver* s, const fvk_run_config* cfg, fvk_run_result* result);

/* Launch shape the solver runs its Euler stage with (no counterpart in the
 * reference: it reports what the defaults or a tuner chose, and the tile
 * layout of variant 3). */
typedef struct fvk_launch_shape {
    int32_t kernel_variant;        /* 1, 2, 3 or 4 */
    int32_t block_size;            /* threads per workgroup */
    int32_t tile_cells;            /* variant 3: most owned cells of a tile (else 0) */
    int32_t n_tiles;               /* variant 3: number of tiles (else 0) */
    int32_t tile_max_staged;       /* variant 3: most owned + ring-1 cells of a tile */
    int32_t tile_max_ring2;        /* variant 3, order 2: most ring-2 cells of a tile */
    int32_t tile_max_faces;        /* variant 3: most faces of a tile */
    int32_t persistent_grid;       /* variant 4: workgroups of the persistent launch (else 0) */
    int64_t tile_lds_bytes;        /* variant 3: dynamic LDS per workgroup */
} fvk_launch_shape;
int fvk_get_launch_shape(const fvk_solver* s, fvk_launch_shape* shape);

/* GradientCalculator::compute for this solver's mesh and scheme: phi
 * [n_cells], boundary_phi [n_faces] (reference face order; only boundary
 * entries read) -> grad [2*n_cells] (dphidx, dphidy). */
int fvk_gradient(fvk_solver* s, int32_t scheme, const double* phi, const double* boundary_phi, double* grad);

/* Wall distance of every cell to the solver's no-slip wall faces
 * (RANS constructors, src/RANSTurbulenceSASolver.cpp:67), [n_cells]. */
int fvk_wall_distance(fvk_solver* s, double* dist);
/* Standalone: distance of every cell centroid to the given faces. */
int fvk_wall_distance_faces(const fvk_mesh_desc* mesh, const int32_t* wall_faces, int64_t n_wall, double* dist);

/* Extension: one residual evaluation R(U) (net flux in, per cell, the
 * quantity whose L2 norm residual() reports; for order 2 with the limiter
 * and gradients recomputed) without updating the state.  R is AoS
 * [nvars*n_cells]. */
int fvk_residual_eval(fvk_solver* s, double* R);

/* Extension: discrete adjoint (vector-Jacobian product) of R(U, X) for the
 * Euler path: given lambda [4*n_cells], returns U_bar = (dR/dU)^T lambda
 * [4*n_cells] and, if X_bar != NULL, (dR/dX)^T lambda [2*n_nodes] with X the
 * node coordinates (face geometry from the face's nodes; cell volume and
 * centroid from the shoelace formulas over the polygon that the cell's
 * faces close).  Gradients and limiters are recomputed.  Rusanov and HLLC
 * fluxes; X_bar is not available for periodic meshes. */
int fvk_residual_vjp(fvk_solver* s, const double* lambda, double* U_bar, double* X_bar);

/* Extension: pressure lift and drag coefficients on the wall patches
 * (src/WallTraction.cpp:163-216, pressure part: p of the wall cell times
 * A n, drag along (dir_x, dir_y), lift along it rotated +90 degrees,
 * normalised by 0.5 rho_ref |V|^2 L_ref). */
int fvk_wall_forces(fvk_solver* s, double rho_ref, double vx_ref, double vy_ref, double length_ref,
                    double* cl, double* cd);

/* ---- wall diagnostics on the device (src/WallTraction.cpp:95-216) ---- */

/* WallReferenceQuantities, include/WallTraction.h:99-106 (same fields). */
typedef struct fvk_wall_reference {
    double rho_ref, velocity_x_ref, velocity_y_ref, p_ref, length_ref, moment_reference_x, moment_reference_y;
} fvk_wall_reference;

/* WallFaceSample, include/WallTraction.h:36-45, plus the three per-sample
 * coefficients of skin_friction_coefficient, pressure_coefficient and
 * wall_y_plus (src/WallTraction.cpp:145-161). */
typedef struct fvk_wall_sample {
    int32_t face_index;            /* reference face id */
    int32_t patch_id;
    double x_mid, y_mid, p, rho, effective_viscosity, tau_wall, tx, ty, y_wall_normal;
    double cf, cp, y_plus;
} fvk_wall_sample;

/* WallForceReport, include/WallTraction.h:122-128 (patch_id -1: the total). */
typedef struct fvk_wall_force_report {
    int32_t patch_id;
    int32_t reserved0;
    double friction_drag, pressure_drag, total_drag, cd_friction, cd_pressure, cd_total, lift, cl, moment, cm;
} fvk_wall_force_report;

/* compute_wall_traction_samples (src/NavierStokesFVMSolver.cpp:333-352,
 * RANSTurbulenceSASolver.cpp:353-374, RANSTurbulenceSSTSolver.cpp:435-464)
 * followed by compute_wall_traction (src/WallTraction.cpp:95-146) for the
 * boundary faces wall_faces[n_wall] (reference face ids, in the caller's
 * order), on the device from the solver's current state: cell gradients of
 * u and v with the solver's scheme and boundary fields, the Jasak-corrected
 * wall gradient, the effective viscosity (mu, mu + rho nu_t for SA and SST),
 * tau_wall, y_wall_normal and the coefficients against `ref`.  An Euler
 * solver has no viscous stress: its samples carry the pressure with
 * effective_viscosity = tau_wall = 0 (the reference wires wall forces only
 * for no-slip walls).  An interior face in the list is an error. */
int fvk_wall_traction(fvk_solver* s, const int32_t* wall_faces, int64_t n_wall, const fvk_wall_reference* ref,
                      fvk_wall_sample* samples);

/* compute_wall_forces (src/WallTraction.cpp:163-216) over the samples of
 * fvk_wall_traction: one report per patch that has a sample, by ascending
 * patch id, then the domain total (patch_id -1); each sum runs over the
 * samples in the caller's order, as the reference's does.  reports has room
 * for cap entries (n_patches + 1 always suffices); *n_reports receives the
 * count.  samples may be NULL. */
int fvk_wall_force_reports(fvk_solver* s, const int32_t* wall_faces, int64_t n_wall, const fvk_wall_reference* ref,
                           fvk_wall_sample* samples, fvk_wall_force_report* reports, int32_t cap,
                           int32_t* n_reports);

/* No counterpart in the reference: fvk_wall_force_reports for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It allocates nothing: wall_faces
 * (reference face ids), samples (may be NULL), reports [cap >= n_bcs + 1]
 * and n_reports are the caller's DEVICE buffers, and the kernels' scratch
 * is a fixed device array of the library, one per GPU, which holds at most
 * FVK_NOALLOC_MAX_WALL_FACES wall faces (larger n_wall is an error).  The
 * work is queued on the solver's stream and the call returns without
 * waiting: the outputs are final after fvk_synchronize.  Every entry of the
 * scratch a call reads it wrote first, so nothing carries over between
 * calls; but calls that overlap on one GPU (two solvers on two host threads)
 * share the scratch and race.  An entry of wall_faces that is not a boundary
 * face gives a sample with face_index -1 and NaN values and is left out of
 * the sums. */
#define FVK_NOALLOC_MAX_WALL_FACES 65536
int fvk_wall_force_reports_noalloc(fvk_solver* s, const int32_t* wall_faces, int64_t n_wall,
                                   const fvk_wall_reference* ref, fvk_wall_sample* samples,
                                   fvk_wall_force_report* reports, int32_t cap, int32_t* n_reports);

/* Extension: gradient of J = w_cl Cl + w_cd Cd (the pressure forces of
 * fvk_wall_forces) with respect to the state, U_bar [4*n_cells], and the
 * node coordinates, X_bar [2*n_nodes]; either may be NULL.  Overwrites both. */
int fvk_wall_forces_vjp(fvk_solver* s, double rho_ref, double vx_ref, double vy_ref, double length_ref, double w_cl,
                        double w_cd, double* U_bar, double* X_bar);

/* Extension: the shape sensitivity of J = w_cl Cl + w_cd Cd through the
 * flow, by one adjoint solve at the current state (a converged steady
 * state of the solver's residual R):
 *   (dR/dU)^T lambda = -dJ/dU,     dJ/dX = dJ/dX|_U + lambda^T dR/dX.
 * lambda is iterated on the device in pseudo time, d lambda/d tau =
 * P ((dR/dU)^T lambda + dJ/dU), with SSP-RK3 and P = dt_c / V_c the
 * primal's local time step at CFL `cfl` (the transposed operator has the
 * primal's spectrum, so the primal's stable CFL serves), from lambda = 0,
 * for at most max_iter iterations, stopping once the adjoint residual
 * |(dR/dU)^T lambda + dJ/dU| has dropped by `tol` (checked every
 * check_every iterations).  dJ_dX [2*n_nodes]; lambda_out [4*n_cells]
 * (optional); report (optional) the iterations and the first and final
 * residual.  Euler, Rusanov or HLLC, no periodic faces.  A non-finite
 * residual returns FVK_ERR_DIVERGED. */
typedef struct fvk_adjoint_report {
    int32_t iterations;
    double residual_first, residual_final;
} fvk_adjoint_report;
int fvk_adjoint_solve_forces(fvk_solver* s, double rho_ref, double vx_ref, double vy_ref, double length_ref,
                             double w_cl, double w_cd, double cfl, int32_t max_iter, double tol, int32_t check_every,
                             double* dJ_dX, double* lambda_out, fvk_adjoint_report* report);

/* Device time accumulated while timing is on, measured with HIP events on
 * the solver's stream: mode 1, an event pair around each launch (per-kernel
 * times; the events themselves add a few microseconds to each launch);
 * mode 2, one pair around all the launches of each fvk_step call, reported
 * as "stages" with a count of steps x RK stages; 0 turns timing off.
 * names is a '\n'-separated list written into a caller buffer. */
int fvk_set_kernel_timing(fvk_solver* s, int32_t enable);
int fvk_kernel_times(const fvk_solver* s, char* names, int64_t names_cap, double* total_ms, int64_t* count,
                     int32_t cap, int32_t* n_kernels);

/* Device memory the solver holds, bytes. */
int64_t fvk_device_bytes(const fvk_solver* s);

/* A plain device copy of n doubles (bandwidth calibration for counters). */
int fvk_copy_probe(int64_t n_doubles, int32_t reps, double* ms_per_copy);

/* The device division the flux kernels use (a refined reciprocal without
 * the range scaling and special-case fixup of the full IEEE sequence) and
 * the compiler's IEEE x / y, on n caller pairs (host arrays).  Both agree
 * bit for bit when x == 0 or when |x|, |y| and |x / y| lie in
 * [2^-900, 2^900]; a zero divisor gives NaN instead of +-inf.  No reference
 * counterpart: it checks that precondition. */
int fvk_divide_probe(int64_t n, const double* x, const double* y, double* q_fast, double* q_ieee);

/* Extension (no reference counterpart): independent cases stepped together.
 * The cases' meshes are the disconnected parts of one device mesh, each a
 * contiguous range of cells; every kernel launch covers all cases, the
 * residual norms are reduced per case, and a case can be frozen (its work is
 * skipped, its state kept).  With SSP-RK stages a stage is one launch of the
 * fused tile kernel over the active cases' tiles: a workgroup per tile
 * (kernel_variant 3, the default) or, with kernel_variant 4, a persistent
 * grid of the workgroups that fit on the device pulling tiles from a queue.  Euler only, local time steps, no periodic
 * faces, each case at least one workgroup of cells.  bcs holds the cases'
 * boundary conditions one after the other (n_bcs[k] for case k, whose patch
 * ids index its own); ics[k] is case k's initial condition; params are
 * shared. */
typedef struct fvk_batch fvk_batch;
int fvk_batch_create(int32_t n_cases, const fvk_mesh_desc* meshes, const fvk_bc* bcs, const int32_t* n_bcs,
                     const fvk_params* params, const fvk_initial_condition* ics, fvk_batch** out);
void fvk_batch_destroy(fvk_batch* b);
/* nsteps steps of every active case; case_residuals (may be NULL) receives
 * [nsteps][n_cases][4] L2 norms (a frozen case's rows are 0). */
int fvk_batch_step(fvk_batch* b, int32_t nsteps, double* case_residuals);
/* active[n_cases]: 1 steps the case, 0 freezes it. */
int fvk_batch_set_active(fvk_batch* b, const int32_t* active);
/* Replay each step from a captured hipGraph (1) or launch it directly (0). */
int fvk_batch_set_graph(fvk_batch* b, int32_t enable);
/* Case k's state, AoS [4 * n_cells of case k] in its own cell order. */
int fvk_batch_get_field(fvk_batch* b, int32_t k, double* U);
/* The solver over the union mesh (kernel timing, device bytes). */
fvk_solver* fvk_batch_solver(fvk_batch* b);

/* ---- mesh ingest (src/MeshReader.cpp, docs/fvmesh-format.md) ---- */

/* A mesh read from a file; owns the arrays fvk_mesh_file_desc points into. */
typedef struct fvk_mesh_file fvk_mesh_file;

/* MeshReader::read (src/MeshReader.cpp:963-988): ".msh" (Gmsh ASCII 2.2 or
 * 4.x, by $MeshFormat) or ".fvmesh" (FVMESH 1.0).  Faces, their orientation
 * and order, Cell::faces, patch numbering and names, and cell volumes and
 * centroids are those the reference's reader builds
 * (build_cells_faces_patches, :317-456); a mesh it rejects (untagged
 * boundary edges, inverted, degenerate or self-intersecting cells, zero-
 * length faces, malformed sections) is an error here with its message. */
int fvk_mesh_read(const char* path, fvk_mesh_file** out);
void fvk_mesh_file_destroy(fvk_mesh_file* m);
/* The mesh as fvk_create takes it (pointers into the handle's arrays). */
int fvk_mesh_file_desc(const fvk_mesh_file* m, fvk_mesh_desc* desc);
/* Cell::node_ids as CSR: ptr [n_cells + 1], nodes [ptr[n_cells]]. */
int fvk_mesh_file_cell_nodes(const fvk_mesh_file* m, const int32_t** ptr, const int32_t** nodes);
/* BoundaryPatch::name of patch id `patch`, NULL out of range. */
const char* fvk_mesh_file_patch_name(const fvk_mesh_file* m, int32_t patch);

/* Mesh-generation helper (host only, no reference counterpart): Bridson
 * Poisson-disk samples of radius `radius` in [x0,x1) x [y0,y1), outside a
 * symmetric NACA 4-digit section of the given thickness (chord 1, leading
 * edge at the origin) dilated by `margin`, after the n_seed caller points.
 * Deterministic for `seed`.  out_xy [2*cap]; the count is *n_out. */
int fvk_poisson_disk(double x0, double y0, double x1, double y1, double radius, uint64_t seed, const double* seed_xy,
                     int64_t n_seed, double thickness, double margin, double* out_xy, int64_t cap, int64_t* n_out);

const char* fvk_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FVK_H_INCLUDED */
