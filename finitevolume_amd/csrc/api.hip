// C ABI of the fvk library (include/fvk.h): solver lifetime, state
// transfer in reference cell order, stepping, and the standalone entry
// points.  One HIP stream per solver; kernels are in kernels_*.hip.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "norms.h"
#include "solver.h"

namespace fvk {

int launch_euler_residual_eval(fvk_solver* s, double* R_dev);
int launch_visc_residual_eval(fvk_solver* s, double* R_dev);
int launch_diffusion_residual_eval(fvk_solver* s, double* R_dev);
int launch_wall_distance_segments(const double* cx, const double* cy, int nc, const double* seg, int nw, double* out,
                                  hipStream_t st);

namespace {
thread_local std::string g_err;
}

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) return set_error(FVK_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return FVK_OK;
}

void* dev_alloc(fvk_solver* s, size_t bytes) {
    void* p = nullptr;
    size_t b = (bytes + 255) & ~size_t(255);
    if (hipMalloc(&p, b) != hipSuccess) return nullptr;
    s->allocations.push_back(p);
    s->device_bytes += (int64_t)b;
    return p;
}

// pinned host memory mapped into the device's address space; *dev receives
// the device view (nullptr on failure)
void* host_mapped_alloc(fvk_solver* s, size_t bytes, void** dev) {
    void* h = nullptr;
    *dev = nullptr;
    if (hipHostMalloc(&h, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) return nullptr;
    s->host_allocations.push_back(h);
    std::memset(h, 0, bytes);
    if (hipHostGetDevicePointer(dev, h, 0) != hipSuccess) {
        *dev = nullptr;
        return nullptr;
    }
    return h;
}

KernelTimer* timer_named(fvk_solver* s, const char* name) {
    for (auto& k : s->timers)
        if (k.name == name) return &k;
    s->timers.push_back(KernelTimer{});
    s->timers.back().name = name;
    return &s->timers.back();
}

// records the start event of a pair covering `weight` launches or stages on
// the solver's stream; the caller records *ev_stop after them
void timer_start(fvk_solver* s, const char* name, int64_t weight, hipEvent_t* ev_stop) {
    *ev_stop = nullptr;
    KernelTimer* t = timer_named(s, name);
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return;
    if (hipEventCreate(&b) != hipSuccess) { hipEventDestroy(a); return; }
    hipEventRecord(a, s->stream);
    t->start.push_back(a);
    t->stop.push_back(b);
    t->weight.push_back(weight);
    *ev_stop = b;
}

// timing mode 1: an event pair around one launch
void timer_begin(fvk_solver* s, const char* name, hipEvent_t* ev_stop) {
    *ev_stop = nullptr;
    if (s->timing != 1) return;
    timer_start(s, name, 1, ev_stop);
}

void timer_end(fvk_solver* s, hipEvent_t ev_stop) {
    if (ev_stop) hipEventRecord(ev_stop, s->stream);
}

static void timers_collect(fvk_solver* s) {
    for (auto& t : s->timers) {
        for (size_t i = 0; i < t.start.size(); ++i) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, t.start[i], t.stop[i]) == hipSuccess) {
                t.total_ms += ms;
                t.count += t.weight[i];
            }
            hipEventDestroy(t.start[i]);
            hipEventDestroy(t.stop[i]);
        }
        t.start.clear();
        t.stop.clear();
        t.weight.clear();
    }
}

}  // namespace fvk

using namespace fvk;

// the two scalar-field solvers (UnstructuredFVMSolver, AdvectionDiffusionFVMSolver)
static bool scalar_eq(int eq) { return eq == FVK_EQ_DIFFUSION || eq == FVK_EQ_ADVECTION_DIFFUSION; }

static int nvars_eq(int eq) {
    return eq == FVK_EQ_RANS_SA ? 5 : (eq == FVK_EQ_RANS_SST ? 6 : (scalar_eq(eq) ? 1 : 4));
}

extern "C" {

void fvk_params_default(fvk_params* p) {
    std::memset(p, 0, sizeof *p);
    p->equation = FVK_EQ_EULER;
    p->flux_scheme = FVK_FLUX_RUSANOV;
    p->gradient_scheme = FVK_GRAD_LEAST_SQUARES;
    p->exact_riemann_max_iter = 20;
    p->gamma = 1.4;
    p->gas_constant = 1.0;
    p->mu = 0.0;
    p->prandtl = 0.72;
    p->prandtl_t = 0.9;
    p->cfl = 0.5;
    p->exact_riemann_tol = 1e-6;
    p->sa = fvk_sa_constants{0.1355, 0.622, 2.0 / 3.0, 0.41, 0.3, 2.0, 7.1, 0.7, 0.9};
    p->sst = fvk_sst_constants{0.09, 0.41, 0.31, 0.85, 1.0, 0.5, 0.856, 0.075, 0.0828};
    p->sst_variant = FVK_SST_VORTICITY;
    p->order = 1;
    p->limiter = FVK_LIMITER_NONE;
    p->venkat_k = 5.0;
    p->rk_stages = 1;
    p->local_dt = 0;
}

void fvk_initial_condition_default(fvk_initial_condition* ic) {
    std::memset(ic, 0, sizeof *ic);
    ic->mode = FVK_IC_NONE;
    ic->rho = 1.0; ic->p = 1.0;
    ic->rho_l = 1.0; ic->p_l = 1.0;
    ic->rho_r = 0.125; ic->p_r = 0.1;
}

int fvk_num_vars(int32_t equation) { return nvars_eq(equation); }

const char* fvk_last_error(void) { return fvk::g_err.c_str(); }

void fvk_destroy(fvk_solver* s) {
    if (!s) return;
    if (s->stream) hipStreamSynchronize(s->stream);
    timers_collect(s);
    fvk::adjoint_release(s);
    for (void* p : s->allocations) hipFree(p);
    for (void* p : s->host_allocations) hipHostFree(p);
    if (s->row_ready) hipEventDestroy(s->row_ready);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"

int fvk::upload_state(fvk_solver* s, const double* U4, const double* turb, bool keep_turb) {
    const int64_t nc = s->nc, ld = s->ld;
    const int nv = s->nv;
    std::vector<double> h((size_t)nv * ld, 0.0);
    if (!U4) return set_error(FVK_ERR_INVALID, "state pointer is NULL");
    if (scalar_eq(s->p.equation)) {
        for (int64_t i = 0; i < nc; ++i) h[i] = U4[s->cell_new2old[i]];
    } else if (s->p.equation == FVK_EQ_EULER) {
        // Euler: one double4 per cell
        for (int64_t i = 0; i < nc; ++i) {
            const double* u = U4 + 4 * (int64_t)s->cell_new2old[i];
            for (int q = 0; q < 4; ++q) h[4 * i + q] = u[q];
        }
    } else {
        for (int64_t i = 0; i < nc; ++i) {
            const double* u = U4 + 4 * (int64_t)s->cell_new2old[i];
            for (int q = 0; q < 4; ++q) h[q * ld + i] = u[q];
        }
    }
    int nt = nv > 4 ? nv - 4 : 0;
    if (nt && (!turb || keep_turb)) {
        // keep the device turbulence rows
        hipError_t e = hipMemcpyAsync(h.data() + 4 * ld, s->U[s->cur] + 4 * ld, sizeof(double) * nt * ld,
                                      hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return hip_check(e, "read turbulence field");
    } else if (nt) {
        for (int64_t i = 0; i < nc; ++i) {
            const double* t = turb + nt * (int64_t)s->cell_new2old[i];
            for (int q = 0; q < nt; ++q) h[(4 + q) * ld + i] = t[q];
        }
    }
    hipError_t e = hipMemcpyAsync(s->U[s->cur], h.data(), sizeof(double) * nv * ld, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return hip_check(e, "upload state");
}

extern "C" {

int fvk_create(const fvk_mesh_desc* mesh, const fvk_bc* bcs, int32_t n_bcs, const fvk_params* params,
               const fvk_initial_condition* ic, fvk_solver** out) {
    fvk::TraceRange tr("fvk_create");
    return fvk::create_solver(mesh, bcs, n_bcs, params, ic, nullptr, 0, out);
}

}  // extern "C"

namespace fvk {

int create_solver(const fvk_mesh_desc* mesh, const fvk_bc* bcs, int32_t n_bcs, const fvk_params* params,
                  const fvk_initial_condition* ic, const int32_t* cell_case, int32_t n_cases, fvk_solver** out) {
    if (!mesh || !params || !out) return set_error(FVK_ERR_INVALID, "NULL argument to fvk_create");
    *out = nullptr;
    if (params->equation < FVK_EQ_EULER || params->equation > FVK_EQ_ADVECTION_DIFFUSION)
        return set_error(FVK_ERR_INVALID, "unknown equation %d", params->equation);
    if (params->flux_scheme < 0 || params->flux_scheme > 2)
        return set_error(FVK_ERR_INVALID, "unknown flux scheme %d", params->flux_scheme);
    if (params->gradient_scheme < 0 || params->gradient_scheme > 1)
        return set_error(FVK_ERR_INVALID, "unknown gradient scheme %d", params->gradient_scheme);
    if (params->order != 1 && params->order != 2)
        return set_error(FVK_ERR_INVALID, "order must be 1 or 2 (got %d)", params->order);
    if (params->order == 2 && params->equation != FVK_EQ_EULER)
        return set_error(FVK_ERR_UNSUPPORTED, "second-order reconstruction is built for the Euler equations only");
    if (params->rk_stages < 1 || params->rk_stages > 3)
        return set_error(FVK_ERR_INVALID, "rk_stages must be 1, 2 or 3 (got %d)", params->rk_stages);
    if (scalar_eq(params->equation) && (params->rk_stages != 1 || params->local_dt))
        return set_error(FVK_ERR_UNSUPPORTED, "the scalar solvers are forward Euler with a fixed dt");
    if (n_bcs < 0 || n_bcs > kMaxPatches) return set_error(FVK_ERR_INVALID, "n_bcs out of range (%d)", n_bcs);
    if (n_bcs > 0 && !bcs) return set_error(FVK_ERR_INVALID, "bcs is NULL");

    if (params->kernel_variant < 0 || params->kernel_variant > 3)
        return set_error(FVK_ERR_INVALID, "kernel_variant must be 0, 1, 2 or 3 (got %d; the persistent tile launch, "
                         "4, is for batches: fvk_batch_create)", params->kernel_variant);
    if (params->kernel_variant == 3 && (scalar_eq(params->equation) || (cell_case && params->equation != FVK_EQ_EULER)))
        return set_error(FVK_ERR_UNSUPPORTED,
                         "the fused tile kernels (variant 3) are built for Euler, Navier-Stokes and RANS cases "
                         "and for batched Euler cases");
    {
        const int b = params->block_size;
        const bool pow2 = b == 64 || b == 128 || b == 256 || b == 512;
        // the tile kernels take any whole number of wavefronts (their passes then fit the tile)
        const bool waves = params->kernel_variant == 3 && b >= 64 && b <= 512 && b % 64 == 0;
        if (b != 0 && !pow2 && !waves)
            return set_error(FVK_ERR_INVALID,
                             "block_size must be 0, 64, 128, 256 or 512, or a multiple of 64 up to 512 for "
                             "kernel_variant 3 (got %d)", b);
    }

    fvk_solver* s = new (std::nothrow) fvk_solver;
    if (!s) return set_error(FVK_ERR_NOMEM, "host allocation failed");
    s->p = *params;
    s->variant = params->kernel_variant == 0 ? 1 : params->kernel_variant;
    s->block = params->block_size != 0 ? params->block_size : (s->variant == 3 ? 256 : 512);
    if (params->tile_cells < 0 || params->tile_cells > s->block) {
        delete s;
        return set_error(FVK_ERR_INVALID, "tile_cells must be 0..block_size (got %d)", params->tile_cells);
    }
    s->tile_cells = params->tile_cells != 0 ? params->tile_cells : s->block;
    s->nv = nvars_eq(params->equation);
    s->n_bcs = n_bcs;
    s->bcs.assign(bcs, bcs + n_bcs);
    if (cell_case) {
        s->n_cases = n_cases;
        s->cell_case_old.assign(cell_case, cell_case + mesh->n_cells);
    }
    for (int i = 0; i < n_bcs; ++i) {
        int t = bcs[i].type;
        if (t < 0 || t > 2) {
            delete s;
            return set_error(FVK_ERR_INVALID, "boundary condition %d has unknown type %d", i, t);
        }
    }
    int rc = hip_check(hipGetDevice(&s->device), "hipGetDevice");
    if (rc == FVK_OK) rc = noalloc_setup(s->device);
    if (rc == FVK_OK) rc = hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (rc == FVK_OK) rc = hip_check(hipEventCreateWithFlags(&s->row_ready, hipEventDisableTiming), "hipEventCreate");
    if (rc == FVK_OK) rc = build_layout(s, mesh);
    if (rc != FVK_OK) {
        std::string msg = fvk::g_err;
        fvk_destroy(s);
        fvk::g_err = msg;
        return rc;
    }
    if (params->equation == FVK_EQ_EULER && s->mesh.deg > kMaxEulerDegree) {
        int deg = s->mesh.deg;
        fvk_destroy(s);
        return set_error(FVK_ERR_UNSUPPORTED, "a cell has %d faces; the Euler kernels take at most %d", deg,
                         kMaxEulerDegree);
    }
    // reference face id -> device face id, for entry points that take
    // reference face lists in device memory
    s->d_face_old2new = static_cast<int*>(dev_alloc(s, sizeof(int) * std::max<int64_t>(s->nf, 1)));
    if (!s->d_face_old2new ||
        hip_check(hipMemcpy(s->d_face_old2new, s->face_old2new.data(), sizeof(int) * s->nf, hipMemcpyHostToDevice),
                  "face map upload") != FVK_OK) {
        std::string msg = fvk::g_err.empty() ? std::string("device allocation failed") : fvk::g_err;
        fvk_destroy(s);
        return set_error(FVK_ERR_NOMEM, "%s", msg.c_str());
    }
    const int64_t pad = 256;
    s->ld = (s->nc + pad - 1) / pad * pad;
    s->ldf = (s->nf + pad - 1) / pad * pad;
    const int nv = s->nv;
    // Euler reads neighbour states while writing its own: never in place
    int nbuf = params->rk_stages > 1 ? 3 : (params->equation == FVK_EQ_EULER ? 2 : 1);
    for (int b = 0; b < 3; ++b) s->U[b] = nullptr;
    for (int b = 0; b < nbuf; ++b) s->U[b] = static_cast<double*>(dev_alloc(s, sizeof(double) * nv * s->ld));
    for (int b = nbuf; b < 3; ++b) s->U[b] = s->U[0];
    const bool euler = params->equation == FVK_EQ_EULER;
    // the fused Euler variant has no face-flux array: each cell evaluates its faces' fluxes
    if (!euler || s->variant == 1) {
        s->flux = static_cast<double*>(dev_alloc(s, sizeof(double) * nv * s->ldf));
        s->dtf = static_cast<double*>(dev_alloc(s, sizeof(double) * s->ldf));
    }
    s->dtc = static_cast<double*>(dev_alloc(s, sizeof(double) * s->ld));
    const bool advdiff = params->equation == FVK_EQ_ADVECTION_DIFFUSION;
    int ng = euler ? (params->order == 2 ? 8 : 0)
                   : (params->equation == FVK_EQ_DIFFUSION ? 0 : (advdiff ? 2 : 10));
    if (ng) s->grad = static_cast<double*>(dev_alloc(s, sizeof(double) * ng * s->ld));
    if (params->equation == FVK_EQ_RANS_SA || params->equation == FVK_EQ_RANS_SST)
        s->aux = static_cast<double*>(dev_alloc(s, sizeof(double) * 4 * s->ld));
    if (advdiff) s->aux = static_cast<double*>(dev_alloc(s, sizeof(double) * s->ldf));  // boundary phi
    // launch grids (workgroups of 64 threads or more, or one per tile plus a
    // spare) are rounded up to a multiple of the 8 XCDs
    int64_t max_blocks = std::max((std::max(s->nc, s->nf) + 63) / 64, (int64_t)s->mesh.ntiles + 1) + 8;
    // partial sums: up to six residual rows and the dt row
    s->partials = static_cast<double*>(dev_alloc(s, sizeof(double) * 7 * max_blocks));
    s->dtmin = static_cast<unsigned long long*>(dev_alloc(s, sizeof(unsigned long long) * kDtWords));
    s->hist_cap = 1024;
    s->res_hist_host = static_cast<double*>(host_mapped_alloc(s, sizeof(double) * 6 * s->hist_cap,
                                                              reinterpret_cast<void**>(&s->res_hist)));
    s->dt_hist_host = static_cast<double*>(host_mapped_alloc(s, sizeof(double) * s->hist_cap,
                                                             reinterpret_cast<void**>(&s->dt_hist)));
    double* wd = static_cast<double*>(dev_alloc(s, sizeof(double) * s->ld));
    if (!s->U[0] || ((!euler || s->variant == 1) && (!s->flux || !s->dtf)) || !s->dtc || (ng && !s->grad) || !s->partials || !s->dtmin ||
        !s->res_hist || !s->dt_hist || !wd || (nbuf == 3 && (!s->U[1] || !s->U[2]))) {
        fvk_destroy(s);
        return set_error(FVK_ERR_NOMEM, "device allocation failed");
    }
    s->mesh.wdist = wd;
    std::vector<unsigned long long> init(kDtWords);
    double dm = DBL_MAX;
    std::memcpy(&init[0], &dm, 8);
    for (auto& w : init) w = init[0];
    rc = hip_check(hipMemcpy(s->dtmin, init.data(), sizeof(unsigned long long) * init.size(), hipMemcpyHostToDevice),
                   "init dt slots");
    if (rc == FVK_OK) rc = hip_check(hipMemset(s->U[0], 0, sizeof(double) * nv * s->ld), "init state");
    if (rc == FVK_OK && s->grad) rc = hip_check(hipMemset(s->grad, 0, sizeof(double) * ng * s->ld), "init gradients");
    // wall distance to the no-slip walls (RANS constructors, src/RANSTurbulenceSASolver.cpp:67)
    if (rc == FVK_OK) {
        std::vector<double> inf((size_t)s->ld, INFINITY);
        rc = hip_check(hipMemcpy(wd, inf.data(), sizeof(double) * s->ld, hipMemcpyHostToDevice), "init wall distance");
        if (rc == FVK_OK && (params->equation == FVK_EQ_RANS_SA || params->equation == FVK_EQ_RANS_SST)) {
            std::vector<int> wf;
            for (int64_t f = 0; f < s->nf; ++f)
                if (mesh->face_cells[2 * f + 1] == -1 && s->bcs[mesh->face_patch[f]].type == FVK_BC_WALL) wf.push_back((int)f);
            if (!wf.empty()) {
                int nw = (int)wf.size();
                std::vector<double> seg(4 * (size_t)nw);
                for (int k = 0; k < nw; ++k) {
                    const double* a = mesh->node_xy + 2 * mesh->face_nodes[2 * wf[k]];
                    const double* b = mesh->node_xy + 2 * mesh->face_nodes[2 * wf[k] + 1];
                    seg[k] = a[0]; seg[nw + k] = a[1]; seg[2 * nw + k] = b[0]; seg[3 * nw + k] = b[1];
                }
                double* dseg = nullptr;
                rc = hip_check(hipMalloc(&dseg, sizeof(double) * seg.size()), "wall segments");
                if (rc == FVK_OK) rc = hip_check(hipMemcpy(dseg, seg.data(), sizeof(double) * seg.size(), hipMemcpyHostToDevice), "wall segments");
                if (rc == FVK_OK) rc = launch_wall_distance_segments(s->mesh.cx, s->mesh.cy, (int)s->nc, dseg, nw, wd, s->stream);
                if (rc == FVK_OK) rc = hip_check(hipStreamSynchronize(s->stream), "wall distance");
                if (dseg) hipFree(dseg);
            }
        }
    }
    if (rc != FVK_OK) {
        std::string msg = fvk::g_err;
        fvk_destroy(s);
        fvk::g_err = msg;
        return rc;
    }
    // initial condition (solver constructors, e.g. src/EulerFVMSolver.cpp:24-37)
    std::vector<double> U4((size_t)(scalar_eq(s->p.equation) ? 1 : 4) * s->nc, 0.0);
    std::vector<double> T((size_t)(nv > 4 ? nv - 4 : 1) * s->nc, 0.0);
    if (ic) {
        double g = params->gamma;
        auto prim = [g](double* o, double rho, double u, double v, double p) {
            o[0] = rho; o[1] = rho * u; o[2] = rho * v; o[3] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v);
        };
        for (int64_t c = 0; c < s->nc; ++c) {
            double x = mesh->cell_xy[2 * c], y = mesh->cell_xy[2 * c + 1];
            if (scalar_eq(params->equation)) {
                // src/UnstructuredFVMSolver.cpp:13-21 (sqrt), src/AdvectionDiffusionFVMSolver.cpp:22-31 (hypot)
                const double r = params->equation == FVK_EQ_DIFFUSION ? std::sqrt(x * x + y * y) : std::hypot(x, y);
                if (ic->mode != FVK_IC_NONE && r < ic->initial_radius) U4[c] = ic->initial_value;
                continue;
            }
            if (ic->mode == FVK_IC_FREESTREAM) prim(&U4[4 * c], ic->rho, ic->u, ic->v, ic->p);
            else if (ic->mode == FVK_IC_TWO_REGION) {
                if (x < ic->x0) prim(&U4[4 * c], ic->rho_l, ic->u_l, ic->v_l, ic->p_l);
                else prim(&U4[4 * c], ic->rho_r, ic->u_r, ic->v_r, ic->p_r);
            }
            if (nv == 5) T[c] = ic->initial_nut;
            if (nv == 6) { T[2 * c] = ic->initial_k; T[2 * c + 1] = ic->initial_omega; }
        }
    }
    // the set-up above used the null stream; the solver's stream is non-blocking
    rc = hip_check(hipDeviceSynchronize(), "solver set-up");
    if (rc == FVK_OK) rc = upload_state(s, U4.data(), nv > 4 ? T.data() : nullptr, false);
    if (rc != FVK_OK) {
        std::string msg = fvk::g_err;
        fvk_destroy(s);
        fvk::g_err = msg;
        return rc;
    }
    *out = s;
    return FVK_OK;
}

}  // namespace fvk

// The launch parameters of compute_wall_forces' reference quantities
// (src/WallTraction.cpp:168-172), derived on the host as the reference does.
void fvk::wall_launch_ref(WallLaunch& w, const fvk_wall_reference& ref) {
    const double v_mag = std::hypot(ref.velocity_x_ref, ref.velocity_y_ref);
    w.dir_x = ref.velocity_x_ref / v_mag;
    w.dir_y = ref.velocity_y_ref / v_mag;
    w.lift_x = -w.dir_y;
    w.lift_y = w.dir_x;
    w.q = 0.5 * ref.rho_ref * v_mag * v_mag * ref.length_ref;
    w.q2 = w.q * ref.length_ref;
    w.mx = ref.moment_reference_x;
    w.my = ref.moment_reference_y;
    w.rho_ref = ref.rho_ref;
    w.v2 = ref.velocity_x_ref * ref.velocity_x_ref + ref.velocity_y_ref * ref.velocity_y_ref;
    w.p_ref = ref.p_ref;
}

// Wall samples and force reports of host face lists, through a workspace
// allocated for the call.
int fvk::wall_reports(fvk_solver* s, const int32_t* wf, int64_t n, const fvk_wall_reference* ref, int pressure_only,
                      fvk_wall_sample* samples, fvk_wall_force_report* reports, int32_t cap, int32_t* n_reports) {
    if (scalar_eq(s->p.equation)) return set_error(FVK_ERR_UNSUPPORTED, "no wall forces for a scalar field");
    if (n < 0 || n > (int64_t)1 << 30 || cap < 0) return set_error(FVK_ERR_INVALID, "bad wall face count");
    for (int64_t i = 0; i < n; ++i) {
        if (wf[i] < 0 || wf[i] >= s->nf) return set_error(FVK_ERR_INVALID, "wall face %d out of range", wf[i]);
        if (s->face_old2new[wf[i]] < s->nfi)
            return set_error(FVK_ERR_INVALID, "face %d is an interior face, not a wall face", wf[i]);
    }
    if (int rc = settle(s); rc != FVK_OK) return rc;
    const int np = s->n_bcs;
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    const size_t b_faces = sizeof(int) * nn, b_contrib = sizeof(double) * 5 * nn, b_patch = sizeof(int) * nn;
    const size_t b_samples = samples ? sizeof(fvk_wall_sample) * nn : 0;
    const size_t b_rep = sizeof(fvk_wall_force_report) * (size_t)(np + 1);
    const size_t align = 256;
    auto up = [&](size_t b) { return (b + align - 1) / align * align; };
    const size_t total = up(b_faces) + up(b_contrib) + up(b_patch) + up(b_samples) + up(b_rep) + up(sizeof(int));
    char* ws = nullptr;
    int rc = hip_check(hipMalloc(&ws, total), "wall diagnostics workspace");
    if (rc != FVK_OK) return rc;
    char* p = ws;
    auto take = [&](size_t b) { char* q = p; p += up(b); return q; };
    WallLaunch w{};
    int* d_faces = reinterpret_cast<int*>(take(b_faces));
    w.contrib = reinterpret_cast<double*>(take(b_contrib));
    w.patch = reinterpret_cast<int*>(take(b_patch));
    w.samples = samples ? reinterpret_cast<fvk_wall_sample*>(take(b_samples)) : nullptr;
    w.reports = reinterpret_cast<fvk_wall_force_report*>(take(b_rep));
    w.n_reports = reinterpret_cast<int*>(take(sizeof(int)));
    w.wall_faces = d_faces;
    w.face_old2new = s->d_face_old2new;
    w.n = (int)n;
    w.n_patches = np;
    w.pressure_only = pressure_only;
    w.cap = np + 1;
    wall_launch_ref(w, *ref);
    if (n > 0) rc = hip_check(hipMemcpyAsync(d_faces, wf, sizeof(int) * n, hipMemcpyHostToDevice, s->stream), "copy");
    if (rc == FVK_OK) rc = launch_wall_diagnostics(s, w);
    int nrep = 0;
    if (rc == FVK_OK)
        rc = hip_check(hipMemcpyAsync(&nrep, w.n_reports, sizeof(int), hipMemcpyDeviceToHost, s->stream), "copy");
    if (rc == FVK_OK) rc = hip_check(hipStreamSynchronize(s->stream), "wall diagnostics");
    if (rc == FVK_OK && nrep > cap) rc = set_error(FVK_ERR_INVALID, "%d reports do not fit cap %d", nrep, cap);
    if (rc == FVK_OK && nrep > 0)
        rc = hip_check(hipMemcpy(reports, w.reports, sizeof(fvk_wall_force_report) * nrep, hipMemcpyDeviceToHost),
                       "copy");
    if (rc == FVK_OK && samples && n > 0)
        rc = hip_check(hipMemcpy(samples, w.samples, sizeof(fvk_wall_sample) * n, hipMemcpyDeviceToHost), "copy");
    if (rc == FVK_OK) *n_reports = nrep;
    hipFree(ws);
    return rc;
}

int fvk::settle(const fvk_solver* s) {
    if (!s->inflight) return FVK_OK;
    s->inflight = false;
    return hip_check(hipStreamSynchronize(s->stream), "step");
}

extern "C" {

int fvk_set_field(fvk_solver* s, const double* U4, const double* turb, int64_t n_cells) {
    if (!s) return set_error(FVK_ERR_INVALID, "NULL solver");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    if (n_cells != s->nc)
        return set_error(FVK_ERR_INVALID, "set_field: %lld cells given, mesh has %lld", (long long)n_cells, (long long)s->nc);
    int rc = upload_state(s, U4, turb, turb == nullptr);
    if (rc == FVK_OK) std::memset(s->last_res, 0, sizeof s->last_res);
    return rc;
}

int fvk_get_field(const fvk_solver* sc, double* U4, double* turb, int64_t n_cells) {
    fvk_solver* s = const_cast<fvk_solver*>(sc);
    if (!s || !U4) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    if (n_cells != s->nc)
        return set_error(FVK_ERR_INVALID, "get_field: %lld cells given, mesh has %lld", (long long)n_cells, (long long)s->nc);
    const int64_t nc = s->nc, ld = s->ld;
    const int nv = s->nv;
    std::vector<double> h((size_t)nv * ld);
    hipError_t e = hipMemcpyAsync(h.data(), s->U[s->cur], sizeof(double) * nv * ld, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_check(e, "download state");
    if (scalar_eq(s->p.equation)) {
        for (int64_t i = 0; i < nc; ++i) U4[s->cell_new2old[i]] = h[i];
        return FVK_OK;
    }
    const bool aos = s->p.equation == FVK_EQ_EULER;
    for (int64_t i = 0; i < nc; ++i) {
        double* u = U4 + 4 * (int64_t)s->cell_new2old[i];
        for (int q = 0; q < 4; ++q) u[q] = aos ? h[4 * i + q] : h[q * ld + i];
    }
    int nt = nv > 4 ? nv - 4 : 0;
    if (turb && nt) {
        for (int64_t i = 0; i < nc; ++i) {
            double* t = turb + nt * (int64_t)s->cell_new2old[i];
            for (int q = 0; q < nt; ++q) t[q] = h[(4 + q) * ld + i];
        }
    }
    return FVK_OK;
}

int fvk_set_state(fvk_solver* s, const double* U_aos, int64_t n_cells) {
    if (!s || !U_aos) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    if (n_cells != s->nc)
        return set_error(FVK_ERR_INVALID, "set_state: %lld cells given, mesh has %lld", (long long)n_cells, (long long)s->nc);
    const int nv = s->nv;
    if (nv <= 4) return fvk_set_field(s, U_aos, nullptr, n_cells);
    const int nt = nv - 4;
    std::vector<double> U(4 * (size_t)n_cells), T((size_t)nt * n_cells);
    for (int64_t c = 0; c < n_cells; ++c) {
        for (int q = 0; q < 4; ++q) U[4 * c + q] = U_aos[nv * c + q];
        for (int q = 0; q < nt; ++q) T[nt * c + q] = U_aos[nv * c + 4 + q];
    }
    return fvk_set_field(s, U.data(), T.data(), n_cells);
}

int fvk_get_state(const fvk_solver* s, double* U_aos, int64_t n_cells) {
    if (!s || !U_aos) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    if (n_cells != s->nc)
        return set_error(FVK_ERR_INVALID, "get_state: %lld cells given, mesh has %lld", (long long)n_cells, (long long)s->nc);
    const int nv = s->nv;
    if (nv <= 4) return fvk_get_field(s, U_aos, nullptr, n_cells);
    const int nt = nv - 4;
    std::vector<double> U(4 * (size_t)n_cells), T((size_t)nt * n_cells);
    int rc = fvk_get_field(s, U.data(), T.data(), n_cells);
    if (rc != FVK_OK) return rc;
    for (int64_t c = 0; c < n_cells; ++c) {
        for (int q = 0; q < 4; ++q) U_aos[nv * c + q] = U[4 * c + q];
        for (int q = 0; q < nt; ++q) U_aos[nv * c + 4 + q] = T[nt * c + q];
    }
    return FVK_OK;
}

int fvk_step(fvk_solver* s, int32_t nsteps, double* residual_hist, double* dt_hist) {
    fvk::TraceRange tr("fvk_step");
    if (!s) return set_error(FVK_ERR_INVALID, "NULL solver");
    if (nsteps < 0) return set_error(FVK_ERR_INVALID, "nsteps < 0");
    const int nv = s->nv;
    const int stages = scalar_eq(s->p.equation) ? 1 : s->p.rk_stages;
    const double* hr = s->res_hist_host;
    const double* hd = s->dt_hist_host;
    int done = 0;
    while (done < nsteps) {
        int chunk = std::min(s->hist_cap, nsteps - done);
        bool early = false;
        // timing mode 2: one event pair around the chunk's launches, counted
        // as chunk x stages stages (no event between the kernels)
        hipEvent_t region_stop = nullptr;
        if (s->timing == 2) fvk::timer_start(s, "stages", (int64_t)chunk * stages, &region_stop);
        for (int i = 0; i < chunk; ++i) {
            int rc = FVK_OK;
            if (s->p.equation == FVK_EQ_DIFFUSION) {
                rc = launch_diffusion_step(s, 1, i);
            } else if (s->p.equation == FVK_EQ_ADVECTION_DIFFUSION) {
                rc = launch_advdiff_step(s, 1, i);
            } else {
                for (int st = 0; st < stages && rc == FVK_OK; ++st) {
                    // the history index is the step within this chunk; the dt slot follows the global count
                    int64_t step = s->step_count;
                    if (s->p.equation == FVK_EQ_EULER) rc = launch_euler_stage(s, st, stages, 0, 1, step);
                    else rc = launch_visc_stage(s, st, stages, 1, step);
                    // the chunk's last row is final after stage 1's launches, which reduce
                    // stage 0's norms; the host need not wait for the later stages
                    if (rc == FVK_OK && i == chunk - 1 && st == 1 && stages > 2 && !s->timing &&
                        s->n_cases == 0 && s->fin_nblocks == 0) {
                        rc = hip_check(hipEventRecord(s->row_ready, s->stream), "step event");
                        early = rc == FVK_OK;
                    }
                }
            }
            if (rc != FVK_OK) return rc;
            s->step_count++;
            (void)i;
        }
        {
            int rc = flush_pending_finalize(s);
            if (rc != FVK_OK) return rc;
        }
        if (region_stop) hipEventRecord(region_stop, s->stream);
        // the finalize kernels wrote the rows into mapped host memory: one sync per chunk
        hipError_t e = early ? hipEventSynchronize(s->row_ready) : hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return hip_check(e, "step");
        s->inflight = early;
        if (s->timing) timers_collect(s);
        // rows were written at index (global step % hist_cap)
        for (int i = 0; i < chunk; ++i) {
            int64_t gstep = s->step_count - chunk + i;
            int64_t idx = scalar_eq(s->p.equation) ? i : gstep % s->hist_cap;
            if (residual_hist) std::memcpy(residual_hist + (int64_t)(done + i) * nv, &hr[idx * nv], sizeof(double) * nv);
            if (dt_hist) dt_hist[done + i] = hd[idx];
            if (i == chunk - 1) std::memcpy(s->last_res, &hr[idx * nv], sizeof(double) * nv);
        }
        done += chunk;
    }
    return FVK_OK;
}

int fvk_synchronize(fvk_solver* s) {
    if (!s) return set_error(FVK_ERR_INVALID, "NULL solver");
    s->inflight = false;
    return hip_check(hipStreamSynchronize(s->stream), "synchronize");
}

int fvk_residual(const fvk_solver* s, double* norms) {
    if (!s || !norms) return set_error(FVK_ERR_INVALID, "NULL argument");
    std::memcpy(norms, s->last_res, sizeof(double) * s->nv);
    return FVK_OK;
}

int fvk_set_cfl(fvk_solver* s, double cfl) {
    if (!s) return set_error(FVK_ERR_INVALID, "NULL solver");
    s->p.cfl = cfl;
    return FVK_OK;
}

double fvk_gas_gamma(const fvk_solver* s) { return s ? s->p.gamma : NAN; }

int fvk_get_launch_shape(const fvk_solver* s, fvk_launch_shape* shape) {
    if (!s || !shape) return set_error(FVK_ERR_INVALID, "NULL argument");
    fvk_launch_shape o{};
    o.kernel_variant = s->variant == 3 && s->tile_queue ? 4 : s->variant;  // 4: a persistent batch
    o.block_size = s->block;
    if (s->variant == 3) {
        o.tile_cells = s->tile_cells;
        o.n_tiles = s->mesh.ntiles;
        o.tile_max_staged = s->mesh.tmax_g;
        o.tile_max_ring2 = s->mesh.tmax_2;
        o.tile_max_faces = s->mesh.tmax_f;
        o.tile_lds_bytes = (int64_t)s->tile_lds;
        o.persistent_grid = s->tile_queue ? s->persist_grid : 0;
    }
    *shape = o;
    return FVK_OK;
}

int64_t fvk_device_bytes(const fvk_solver* s) { return s ? s->device_bytes : 0; }

int fvk_residual_eval(fvk_solver* s, double* R) {
    fvk::TraceRange tr("fvk_residual_eval");
    if (!s || !R) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    const int nv = s->nv;
    double* d = nullptr;
    int rc = hip_check(hipMalloc((void**)&d, sizeof(double) * nv * s->ld), "residual buffer");
    if (rc != FVK_OK) return rc;
    if (s->p.equation == FVK_EQ_EULER) rc = launch_euler_residual_eval(s, d);
    else if (s->p.equation == FVK_EQ_DIFFUSION) rc = launch_diffusion_residual_eval(s, d);
    else if (s->p.equation == FVK_EQ_ADVECTION_DIFFUSION) rc = launch_advdiff_residual_eval(s, d);
    else rc = launch_visc_residual_eval(s, d);
    std::vector<double> h((size_t)nv * s->ld);
    if (rc == FVK_OK) rc = hip_check(hipMemcpyAsync(h.data(), d, sizeof(double) * nv * s->ld, hipMemcpyDeviceToHost, s->stream), "residual copy");
    int rs = hip_check(hipStreamSynchronize(s->stream), "residual eval");
    hipFree(d);
    if (rc == FVK_OK) rc = rs;
    if (rc != FVK_OK) return rc;
    for (int64_t i = 0; i < s->nc; ++i) {
        int64_t c = s->cell_new2old[i];
        for (int q = 0; q < nv; ++q) R[c * nv + q] = s->p.equation == FVK_EQ_EULER ? h[nv * i + q] : h[q * s->ld + i];
    }
    return FVK_OK;
}

int fvk_gradient(fvk_solver* s, int32_t scheme, const double* phi, const double* bphi, double* grad) {
    if (!s || !phi || !bphi || !grad) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (scheme < 0 || scheme > 1) return set_error(FVK_ERR_INVALID, "unknown gradient scheme %d", scheme);
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    const int64_t nc = s->nc, nf = s->nf;
    std::vector<double> hp(nc), hb(nf), hg(2 * nc);
    for (int64_t i = 0; i < nc; ++i) hp[i] = phi[s->cell_new2old[i]];
    for (int64_t i = 0; i < nf; ++i) hb[i] = bphi[s->face_new2old[i]];
    double *dp = nullptr, *db = nullptr, *dg = nullptr;
    int rc = hip_check(hipMalloc(&dp, sizeof(double) * nc), "gradient buffers");
    if (rc == FVK_OK) rc = hip_check(hipMalloc(&db, sizeof(double) * nf), "gradient buffers");
    if (rc == FVK_OK) rc = hip_check(hipMalloc(&dg, sizeof(double) * 2 * nc), "gradient buffers");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(dp, hp.data(), sizeof(double) * nc, hipMemcpyHostToDevice), "gradient upload");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(db, hb.data(), sizeof(double) * nf, hipMemcpyHostToDevice), "gradient upload");
    if (rc == FVK_OK) rc = launch_gradient(s, scheme, dp, db, dg);
    if (rc == FVK_OK) rc = hip_check(hipStreamSynchronize(s->stream), "gradient");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(hg.data(), dg, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost), "gradient download");
    if (dp) hipFree(dp);
    if (db) hipFree(db);
    if (dg) hipFree(dg);
    if (rc != FVK_OK) return rc;
    for (int64_t i = 0; i < nc; ++i) {
        int64_t c = s->cell_new2old[i];
        grad[2 * c] = hg[2 * i];
        grad[2 * c + 1] = hg[2 * i + 1];
    }
    return FVK_OK;
}

int fvk_wall_distance(fvk_solver* s, double* dist) {
    if (!s || !dist) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    std::vector<double> h(s->nc);
    int rc = hip_check(hipMemcpy(h.data(), s->mesh.wdist, sizeof(double) * s->nc, hipMemcpyDeviceToHost), "wall distance");
    if (rc != FVK_OK) return rc;
    for (int64_t i = 0; i < s->nc; ++i) dist[s->cell_new2old[i]] = h[i];
    return FVK_OK;
}

int fvk_wall_distance_faces(const fvk_mesh_desc* mesh, const int32_t* wall_faces, int64_t n_wall, double* dist) {
    if (!mesh || !dist || (n_wall > 0 && !wall_faces)) return set_error(FVK_ERR_INVALID, "NULL argument");
    const int64_t nc = mesh->n_cells;
    if (n_wall == 0) {
        for (int64_t c = 0; c < nc; ++c) dist[c] = INFINITY;
        return FVK_OK;
    }
    int nw = (int)n_wall;
    std::vector<double> seg(4 * (size_t)nw), cx(nc), cy(nc);
    for (int k = 0; k < nw; ++k) {
        if (wall_faces[k] < 0 || wall_faces[k] >= mesh->n_faces) return set_error(FVK_ERR_INVALID, "wall face out of range");
        const double* a = mesh->node_xy + 2 * mesh->face_nodes[2 * wall_faces[k]];
        const double* b = mesh->node_xy + 2 * mesh->face_nodes[2 * wall_faces[k] + 1];
        seg[k] = a[0]; seg[nw + k] = a[1]; seg[2 * nw + k] = b[0]; seg[3 * nw + k] = b[1];
    }
    for (int64_t c = 0; c < nc; ++c) { cx[c] = mesh->cell_xy[2 * c]; cy[c] = mesh->cell_xy[2 * c + 1]; }
    double *dseg = nullptr, *dcx = nullptr, *dcy = nullptr, *dout = nullptr;
    int rc = hip_check(hipMalloc(&dseg, sizeof(double) * seg.size()), "wall distance buffers");
    if (rc == FVK_OK) rc = hip_check(hipMalloc(&dcx, sizeof(double) * nc), "wall distance buffers");
    if (rc == FVK_OK) rc = hip_check(hipMalloc(&dcy, sizeof(double) * nc), "wall distance buffers");
    if (rc == FVK_OK) rc = hip_check(hipMalloc(&dout, sizeof(double) * nc), "wall distance buffers");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(dseg, seg.data(), sizeof(double) * seg.size(), hipMemcpyHostToDevice), "upload");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(dcx, cx.data(), sizeof(double) * nc, hipMemcpyHostToDevice), "upload");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(dcy, cy.data(), sizeof(double) * nc, hipMemcpyHostToDevice), "upload");
    if (rc == FVK_OK) rc = launch_wall_distance_segments(dcx, dcy, (int)nc, dseg, nw, dout, nullptr);
    if (rc == FVK_OK) rc = hip_check(hipDeviceSynchronize(), "wall distance");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(dist, dout, sizeof(double) * nc, hipMemcpyDeviceToHost), "download");
    hipFree(dseg); hipFree(dcx); hipFree(dcy); hipFree(dout);
    return rc;
}

int fvk_residual_vjp(fvk_solver* s, const double* lambda, double* U_bar, double* X_bar) {
    fvk::TraceRange tr("fvk_residual_vjp");
    if (!s || !lambda || !U_bar) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    return adjoint_vjp(s, lambda, U_bar, X_bar);
}

int fvk_wall_forces(fvk_solver* s, double rho_ref, double vx_ref, double vy_ref, double length_ref, double* cl,
                    double* cd) {
    if (!s || !cl || !cd) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (scalar_eq(s->p.equation)) return set_error(FVK_ERR_UNSUPPORTED, "no pressure for a scalar field");
    // src/WallTraction.cpp:163-216 over pressure-only samples (A p_cell n) of
    // every wall face, in reference face order
    std::vector<int32_t> wf;
    for (int64_t f = 0; f < s->nf; ++f) {
        const int i = s->face_old2new[f];
        if (i >= s->nfi && s->bcs[s->face_patch_old[f]].type == FVK_BC_WALL) wf.push_back((int32_t)f);
    }
    fvk_wall_reference ref{rho_ref, vx_ref, vy_ref, 0.0, length_ref, 0.0, 0.0};
    std::vector<fvk_wall_force_report> rep(s->n_bcs + 1);
    int32_t nrep = 0;
    int rc = wall_reports(s, wf.data(), (int64_t)wf.size(), &ref, 1, nullptr, rep.data(), (int32_t)rep.size(), &nrep);
    if (rc != FVK_OK) return rc;
    *cl = nrep > 0 ? rep[nrep - 1].cl : 0.0;
    *cd = nrep > 0 ? rep[nrep - 1].cd_total : 0.0;
    return FVK_OK;
}

int fvk_wall_traction(fvk_solver* s, const int32_t* wall_faces, int64_t n_wall, const fvk_wall_reference* ref,
                      fvk_wall_sample* samples) {
    if (!s || !ref || (n_wall > 0 && (!wall_faces || !samples))) return set_error(FVK_ERR_INVALID, "NULL argument");
    std::vector<fvk_wall_force_report> rep(s->n_bcs + 1);
    int32_t nrep = 0;
    return wall_reports(s, wall_faces, n_wall, ref, 0, samples, rep.data(), (int32_t)rep.size(), &nrep);
}

int fvk_wall_force_reports(fvk_solver* s, const int32_t* wall_faces, int64_t n_wall, const fvk_wall_reference* ref,
                           fvk_wall_sample* samples, fvk_wall_force_report* reports, int32_t cap,
                           int32_t* n_reports) {
    if (!s || !ref || !n_reports || (n_wall > 0 && !wall_faces) || (cap > 0 && !reports))
        return set_error(FVK_ERR_INVALID, "NULL argument");
    return wall_reports(s, wall_faces, n_wall, ref, 0, samples, reports, cap, n_reports);
}

int fvk_wall_forces_vjp(fvk_solver* s, double rho_ref, double vx_ref, double vy_ref, double length_ref, double w_cl,
                        double w_cd, double* U_bar, double* X_bar) {
    if (!s || (!U_bar && !X_bar)) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    if (scalar_eq(s->p.equation)) return set_error(FVK_ERR_UNSUPPORTED, "no pressure for a scalar field");
    // J = w_cl Cl + w_cd Cd = sum over wall faces of p_cell (A n) . e, with
    // A n = (ey, -ex) for the edge e = node2 - node1 and
    // e = (w_cl lift_dir + w_cd drag_dir) / (0.5 rho_ref |V|^2 L_ref)
    std::vector<double> U(4 * (size_t)s->nc);
    int rc = fvk_get_field(s, U.data(), nullptr, s->nc);
    if (rc != FVK_OK) return rc;
    const double vm = std::hypot(vx_ref, vy_ref);
    const double dx = vx_ref / vm, dy = vy_ref / vm, lx = -dy, ly = dx;
    const double q = 0.5 * rho_ref * vm * vm * length_ref;
    const double ex = (w_cl * lx + w_cd * dx) / q, ey = (w_cl * ly + w_cd * dy) / q;
    const double g = s->p.gamma, gm = g - 1.0;
    if (U_bar) std::fill(U_bar, U_bar + 4 * s->nc, 0.0);
    if (X_bar) std::fill(X_bar, X_bar + 2 * s->nn, 0.0);
    std::vector<int> dl(s->nf), dr(s->nf);
    rc = hip_check(hipMemcpy(dl.data(), s->mesh.fl, sizeof(int) * s->nf, hipMemcpyDeviceToHost), "face copy");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(dr.data(), s->mesh.fr, sizeof(int) * s->nf, hipMemcpyDeviceToHost), "face copy");
    if (rc != FVK_OK) return rc;
    const double* X = s->node_xy.data();
    for (int64_t i = 0; i < s->nf; ++i) {
        if (dr[i] >= 0 || s->bcs[-1 - dr[i]].type != FVK_BC_WALL) continue;
        const int n1 = s->face_nodes_dev[2 * i], n2 = s->face_nodes_dev[2 * i + 1];
        const double ax = X[2 * n2] - X[2 * n1], ay = X[2 * n2 + 1] - X[2 * n1 + 1];
        const double an = ay * ex - ax * ey;  // (A n) . e
        const int c = s->cell_new2old[dl[i]];
        const double* u = &U[4 * (size_t)c];
        const double uu = u[1] / u[0], vv = u[2] / u[0];
        if (U_bar) {
            // p = (g - 1) (E - 0.5 (ru^2 + rv^2) / r)
            U_bar[4 * (size_t)c] += an * gm * 0.5 * (uu * uu + vv * vv);
            U_bar[4 * (size_t)c + 1] += -an * gm * uu;
            U_bar[4 * (size_t)c + 2] += -an * gm * vv;
            U_bar[4 * (size_t)c + 3] += an * gm;
        }
        if (X_bar) {
            const double p = gm * (u[3] - 0.5 * u[0] * (uu * uu + vv * vv));
            X_bar[2 * n2 + 1] += p * ex;
            X_bar[2 * n1 + 1] -= p * ex;
            X_bar[2 * n2] -= p * ey;
            X_bar[2 * n1] += p * ey;
        }
    }
    return FVK_OK;
}

int fvk_adjoint_solve_forces(fvk_solver* s, double rho_ref, double vx_ref, double vy_ref, double length_ref,
                             double w_cl, double w_cd, double cfl, int32_t max_iter, double tol, int32_t check_every,
                             double* dJ_dX, double* lambda_out, fvk_adjoint_report* report) {
    fvk::TraceRange tr("fvk_adjoint_solve_forces");
    if (!s || !dJ_dX) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    std::vector<double> g(4 * (size_t)s->nc), jx(2 * (size_t)s->nn);
    int rc = fvk_wall_forces_vjp(s, rho_ref, vx_ref, vy_ref, length_ref, w_cl, w_cd, g.data(), jx.data());
    if (rc != FVK_OK) return rc;
    rc = fvk::adjoint_solve(s, g.data(), cfl, max_iter, tol, check_every, lambda_out, dJ_dX, report);
    if (rc != FVK_OK) return rc;
    for (size_t i = 0; i < jx.size(); ++i) dJ_dX[i] += jx[i];
    return FVK_OK;
}

int fvk_set_kernel_timing(fvk_solver* s, int32_t enable) {
    if (!s) return set_error(FVK_ERR_INVALID, "NULL solver");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    hipStreamSynchronize(s->stream);
    timers_collect(s);
    if (enable < 0 || enable > 2) return set_error(FVK_ERR_INVALID, "timing mode %d (0, 1 or 2)", enable);
    s->timing = enable;
    if (enable) {
        for (auto& t : s->timers) { t.total_ms = 0.0; t.count = 0; }
    }
    return FVK_OK;
}

int fvk_kernel_times(const fvk_solver* sc, char* names, int64_t names_cap, double* total_ms, int64_t* count,
                     int32_t cap, int32_t* n_kernels) {
    fvk_solver* s = const_cast<fvk_solver*>(sc);
    if (!s) return set_error(FVK_ERR_INVALID, "NULL solver");
    if (int rc = fvk::settle(s); rc != FVK_OK) return rc;
    hipStreamSynchronize(s->stream);
    timers_collect(s);
    std::string all;
    int n = 0;
    for (auto& t : s->timers) {
        if (n < cap) {
            if (total_ms) total_ms[n] = t.total_ms;
            if (count) count[n] = t.count;
        }
        all += t.name;
        all += '\n';
        ++n;
    }
    if (n_kernels) *n_kernels = n;
    if (names && names_cap > 0) {
        std::strncpy(names, all.c_str(), (size_t)names_cap - 1);
        names[names_cap - 1] = '\0';
    }
    return FVK_OK;
}

int fvk_copy_probe(int64_t n_doubles, int32_t reps, double* ms_per_copy) {
    if (n_doubles <= 0 || reps <= 0 || !ms_per_copy) return set_error(FVK_ERR_INVALID, "bad copy probe arguments");
    double *a = nullptr, *b = nullptr;
    int rc = hip_check(hipMalloc(&a, sizeof(double) * n_doubles), "probe alloc");
    if (rc == FVK_OK) rc = hip_check(hipMalloc(&b, sizeof(double) * n_doubles), "probe alloc");
    if (rc == FVK_OK) rc = hip_check(hipMemset(a, 0, sizeof(double) * n_doubles), "probe init");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == FVK_OK) rc = hip_check(hipEventCreate(&e0), "event");
    if (rc == FVK_OK) rc = hip_check(hipEventCreate(&e1), "event");
    if (rc == FVK_OK) rc = launch_copy_probe(b, a, n_doubles, nullptr);
    if (rc == FVK_OK) rc = hip_check(hipEventRecord(e0, nullptr), "event");
    for (int r = 0; r < reps && rc == FVK_OK; ++r) rc = launch_copy_probe(b, a, n_doubles, nullptr);
    if (rc == FVK_OK) rc = hip_check(hipEventRecord(e1, nullptr), "event");
    if (rc == FVK_OK) rc = hip_check(hipEventSynchronize(e1), "event");
    float ms = 0.0f;
    if (rc == FVK_OK) rc = hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
    if (rc == FVK_OK) *ms_per_copy = ms / reps;
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (a) hipFree(a);
    if (b) hipFree(b);
    return rc;
}

int fvk_divide_probe(int64_t n, const double* x, const double* y, double* q_fast, double* q_ieee) {
    if (n < 0 || (n > 0 && (!x || !y || !q_fast || !q_ieee))) return set_error(FVK_ERR_INVALID, "bad probe arguments");
    if (n == 0) return FVK_OK;
    const size_t b = sizeof(double) * (size_t)n;
    double* d = nullptr;
    int rc = hip_check(hipMalloc(&d, 4 * b), "divide probe buffers");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(d, x, b, hipMemcpyHostToDevice), "divide probe copy");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(d + n, y, b, hipMemcpyHostToDevice), "divide probe copy");
    if (rc == FVK_OK) rc = launch_div_probe(d, d + n, d + 2 * n, d + 3 * n, n, nullptr);
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(q_fast, d + 2 * n, b, hipMemcpyDeviceToHost), "divide probe copy");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(q_ieee, d + 3 * n, b, hipMemcpyDeviceToHost), "divide probe copy");
    if (d) hipFree(d);
    return rc;
}

}  // extern "C"
