// Test infrastructure only (the oracle): a C entry layer that drives the
// reference solver classes of happy12/FiniteVolume, compiled from the
// reference's own sources (see oracle/Makefile), so tests and fixture
// generation can run the reference on in-memory meshes.  Nothing in the
// product links or loads this.
//
// Every entry point takes the same fvk_mesh_desc / fvk_bc / fvk_params
// structs as the product's C ABI (include/fvk.h) and builds the reference's
// UnstructuredMesh (include/UnstructuredMesh.h:20-109) from them.

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fvk.h"

#include "AdvectionDiffusionFVMSolver.h"
#include "CflRamp.h"
#include "Checkpoint.h"
#include "EulerFVMSolver.h"
#include "ExactRiemannFlux.h"
#include "GradientReconstruction.h"
#include "MeshReader.h"
#include "HllcFlux.h"
#include "NavierStokesFVMSolver.h"
#include "RANSTurbulenceSASolver.h"
#include "RANSTurbulenceSSTSolver.h"
#include "RusanovFlux.h"
#include "SSTKOmega.h"
#include "SpalartAllmaras.h"
#include "UnstructuredFVMSolver.h"
#include "WallDistance.h"
#include "WallTraction.h"

namespace {

UnstructuredMesh build_mesh(const fvk_mesh_desc* d, const fvk_bc* bcs, int n_bcs) {
    UnstructuredMesh m;
    m.nodes.resize((size_t)d->n_nodes);
    for (int64_t i = 0; i < d->n_nodes; ++i) m.nodes[i] = Node{d->node_xy[2 * i], d->node_xy[2 * i + 1]};
    m.faces.resize((size_t)d->n_faces);
    for (int64_t f = 0; f < d->n_faces; ++f) {
        Face& face = m.faces[f];
        face.node1 = d->face_nodes[2 * f];
        face.node2 = d->face_nodes[2 * f + 1];
        face.cell_left = d->face_cells[2 * f];
        face.cell_right = d->face_cells[2 * f + 1];
        face.patch_id = d->face_patch[f];
        face.nx = face.ny = face.area = face.x_mid = face.y_mid = 0.0;
    }
    m.cells.resize((size_t)d->n_cells);
    for (int64_t c = 0; c < d->n_cells; ++c) {
        Cell& cell = m.cells[c];
        cell.volume = d->cell_volume[c];
        cell.x_centroid = d->cell_xy[2 * c];
        cell.y_centroid = d->cell_xy[2 * c + 1];
        cell.faces.assign(d->cell_faces + d->cell_face_ptr[c], d->cell_faces + d->cell_face_ptr[c + 1]);
    }
    m.patches.resize((size_t)d->n_patches);
    for (int p = 0; p < d->n_patches; ++p) {
        m.patches[p].name = "patch_" + std::to_string(p);
        if (bcs && p < n_bcs) {
            m.patches[p].type = bcs[p].type == FVK_BC_NEUMANN ? BoundaryType::Neumann : BoundaryType::Dirichlet;
            m.patches[p].value = bcs[p].value;
        }
    }
    return m;
}

NumericalFluxScheme flux_of(int s) {
    return s == FVK_FLUX_HLLC ? NumericalFluxScheme::HLLC
                              : (s == FVK_FLUX_EXACT ? NumericalFluxScheme::Exact : NumericalFluxScheme::Rusanov);
}
GradientScheme grad_of(int s) { return s == FVK_GRAD_GREEN_GAUSS ? GradientScheme::GreenGauss : GradientScheme::LeastSquares; }

EulerInitialCondition ic_of(const fvk_initial_condition* ic) {
    EulerInitialCondition e;
    if (!ic) return e;
    e.mode = ic->mode == FVK_IC_TWO_REGION ? EulerICMode::TwoRegion : EulerICMode::Freestream;
    e.rho = ic->rho; e.u = ic->u; e.v = ic->v; e.p = ic->p;
    e.rho_l = ic->rho_l; e.u_l = ic->u_l; e.v_l = ic->v_l; e.p_l = ic->p_l;
    e.rho_r = ic->rho_r; e.u_r = ic->u_r; e.v_r = ic->v_r; e.p_r = ic->p_r;
    e.x0 = ic->x0;
    return e;
}

EulerState st(const double* s) { return EulerState{s[0], s[1], s[2], s[3]}; }

NSBoundaryCondition ns_bc(const fvk_bc& b) {
    NSBoundaryCondition n;
    n.type = b.type == FVK_BC_FARFIELD ? NSBoundaryType::Farfield
                                       : (b.type == FVK_BC_OUTFLOW ? NSBoundaryType::Outflow : NSBoundaryType::NoSlipWall);
    n.wall_u = b.wall_u;
    n.wall_v = b.wall_v;
    n.is_isothermal_wall = b.is_isothermal_wall != 0;
    n.wall_temperature = b.wall_temperature;
    n.farfield_state = st(b.farfield_state);
    return n;
}

SAModelConstants sa_of(const fvk_sa_constants& c) {
    SAModelConstants s;
    s.cb1 = c.cb1; s.cb2 = c.cb2; s.sigma = c.sigma; s.kappa = c.kappa; s.cw2 = c.cw2; s.cw3 = c.cw3;
    s.cv1 = c.cv1; s.cv2 = c.cv2; s.cv3 = c.cv3;
    return s;
}
SSTModelConstants sst_of(const fvk_sst_constants& c) {
    SSTModelConstants s;
    s.beta_star = c.beta_star; s.kappa = c.kappa; s.a1 = c.a1; s.sigma_k1 = c.sigma_k1; s.sigma_k2 = c.sigma_k2;
    s.sigma_omega1 = c.sigma_omega1; s.sigma_omega2 = c.sigma_omega2; s.beta1 = c.beta1; s.beta2 = c.beta2;
    return s;
}

struct RefSolver {
    int eq = 0;
    int64_t n = 0;
    double gamma = 1.4;
    UnstructuredMesh mesh;  // the solver's mesh with compute_geometry() done (wall diagnostics)
    std::unique_ptr<EulerFVMSolver> euler;
    std::unique_ptr<NavierStokesFVMSolver> ns;
    std::unique_ptr<RANSTurbulenceSASolver> sa;
    std::unique_ptr<RANSTurbulenceSSTSolver> sst;
    std::unique_ptr<UnstructuredFVMSolver> diff;
    std::unique_ptr<AdvectionDiffusionFVMSolver> adv;
};

int nvars(int eq) {
    return eq == FVK_EQ_RANS_SA ? 5 : (eq == FVK_EQ_RANS_SST ? 6 : (eq >= FVK_EQ_DIFFUSION ? 1 : 4));
}

}  // namespace

extern "C" {

int fvr_create(const fvk_mesh_desc* d, const fvk_bc* bcs, int32_t n_bcs, const fvk_params* p,
               const fvk_initial_condition* ic, void** out) {
    UnstructuredMesh mesh = build_mesh(d, bcs, n_bcs);
    auto* r = new RefSolver;
    r->eq = p->equation;
    r->n = d->n_cells;
    r->gamma = p->gamma;
    r->mesh = mesh;
    r->mesh.compute_geometry();
    EulerInitialCondition e = ic_of(ic);
    switch (p->equation) {
        case FVK_EQ_EULER: {
            std::vector<EulerBoundaryCondition> b(n_bcs);
            for (int i = 0; i < n_bcs; ++i) {
                b[i].type = bcs[i].type == FVK_BC_FARFIELD ? EulerBoundaryType::Farfield
                            : (bcs[i].type == FVK_BC_OUTFLOW ? EulerBoundaryType::Outflow : EulerBoundaryType::Wall);
                b[i].farfield_state = st(bcs[i].farfield_state);
            }
            r->euler.reset(new EulerFVMSolver(mesh, b, p->gamma, p->cfl, e, flux_of(p->flux_scheme),
                                              p->exact_riemann_tol, p->exact_riemann_max_iter));
            break;
        }
        case FVK_EQ_NAVIER_STOKES: {
            std::vector<NSBoundaryCondition> b(n_bcs);
            for (int i = 0; i < n_bcs; ++i) b[i] = ns_bc(bcs[i]);
            r->ns.reset(new NavierStokesFVMSolver(mesh, b, p->gamma, p->gas_constant, p->mu, p->prandtl, p->cfl, e,
                                                  flux_of(p->flux_scheme), grad_of(p->gradient_scheme),
                                                  p->exact_riemann_tol, p->exact_riemann_max_iter));
            break;
        }
        case FVK_EQ_RANS_SA: {
            std::vector<RANSBoundaryConditionSA> b(n_bcs);
            for (int i = 0; i < n_bcs; ++i) {
                b[i].ns = ns_bc(bcs[i]);
                b[i].farfield_nut = bcs[i].farfield_nut;
            }
            r->sa.reset(new RANSTurbulenceSASolver(mesh, b, p->gamma, p->gas_constant, p->mu, p->prandtl, p->prandtl_t,
                                                   p->cfl, e, ic ? ic->initial_nut : 0.0, flux_of(p->flux_scheme),
                                                   grad_of(p->gradient_scheme), p->exact_riemann_tol,
                                                   p->exact_riemann_max_iter, sa_of(p->sa)));
            break;
        }
        case FVK_EQ_RANS_SST: {
            std::vector<RANSBoundaryConditionSST> b(n_bcs);
            for (int i = 0; i < n_bcs; ++i) {
                b[i].ns = ns_bc(bcs[i]);
                b[i].farfield_k = bcs[i].farfield_k;
                b[i].farfield_omega = bcs[i].farfield_omega;
            }
            r->sst.reset(new RANSTurbulenceSSTSolver(
                mesh, b, p->gamma, p->gas_constant, p->mu, p->prandtl, p->prandtl_t, p->cfl, e,
                ic ? ic->initial_k : 0.0, ic ? ic->initial_omega : 0.0, flux_of(p->flux_scheme),
                grad_of(p->gradient_scheme), p->exact_riemann_tol, p->exact_riemann_max_iter,
                p->sst_variant == FVK_SST_STRAIN_RATE ? SSTLimiterVariant::StrainRate : SSTLimiterVariant::Vorticity,
                sst_of(p->sst), p->kato_launder != 0));
            break;
        }
        case FVK_EQ_DIFFUSION:
            r->diff.reset(new UnstructuredFVMSolver(mesh, p->alpha, p->fixed_dt, ic ? ic->initial_value : 0.0,
                                                    ic ? ic->initial_radius : 0.0));
            break;
        case FVK_EQ_ADVECTION_DIFFUSION:
            r->adv.reset(new AdvectionDiffusionFVMSolver(mesh, p->alpha, p->advection_u, p->advection_v, p->fixed_dt,
                                                         grad_of(p->gradient_scheme), ic ? ic->initial_value : 0.0,
                                                         ic ? ic->initial_radius : 0.0));
            break;
        default:
            delete r;
            return FVK_ERR_INVALID;
    }
    *out = r;
    return FVK_OK;
}

void fvr_destroy(void* h) { delete static_cast<RefSolver*>(h); }

int fvr_set_field(void* h, const double* U4, const double* turb) {
    auto* r = static_cast<RefSolver*>(h);
    if (r->eq == FVK_EQ_DIFFUSION) {
        r->diff->set_field(std::vector<double>(U4, U4 + r->n));
        return FVK_OK;
    }
    if (r->eq == FVK_EQ_ADVECTION_DIFFUSION) {
        r->adv->set_field(std::vector<double>(U4, U4 + r->n));
        return FVK_OK;
    }
    std::vector<EulerState> U((size_t)r->n);
    for (int64_t c = 0; c < r->n; ++c) U[c] = st(U4 + 4 * c);
    switch (r->eq) {
        case FVK_EQ_EULER: r->euler->set_field(U); break;
        case FVK_EQ_NAVIER_STOKES: r->ns->set_field(U); break;
        case FVK_EQ_RANS_SA:
            if (turb) r->sa->set_field(U, std::vector<double>(turb, turb + r->n));
            else r->sa->set_field(U);
            break;
        case FVK_EQ_RANS_SST:
            if (turb) {
                std::vector<double> k((size_t)r->n), w((size_t)r->n);
                for (int64_t c = 0; c < r->n; ++c) { k[c] = turb[2 * c]; w[c] = turb[2 * c + 1]; }
                r->sst->set_field(U, k, w);
            } else {
                r->sst->set_field(U);
            }
            break;
    }
    return FVK_OK;
}

int fvr_get_field(void* h, double* U4, double* turb) {
    auto* r = static_cast<RefSolver*>(h);
    if (r->eq == FVK_EQ_DIFFUSION || r->eq == FVK_EQ_ADVECTION_DIFFUSION) {
        const std::vector<double>& phi = r->eq == FVK_EQ_DIFFUSION ? r->diff->field() : r->adv->field();
        std::memcpy(U4, phi.data(), sizeof(double) * r->n);
        return FVK_OK;
    }
    const std::vector<EulerState>* U = nullptr;
    switch (r->eq) {
        case FVK_EQ_EULER: U = &r->euler->field(); break;
        case FVK_EQ_NAVIER_STOKES: U = &r->ns->field(); break;
        case FVK_EQ_RANS_SA: U = &r->sa->field(); break;
        case FVK_EQ_RANS_SST: U = &r->sst->field(); break;
    }
    for (int64_t c = 0; c < r->n; ++c) {
        U4[4 * c] = (*U)[c].rho;
        U4[4 * c + 1] = (*U)[c].rho_u;
        U4[4 * c + 2] = (*U)[c].rho_v;
        U4[4 * c + 3] = (*U)[c].E;
    }
    if (turb && r->eq == FVK_EQ_RANS_SA) std::memcpy(turb, r->sa->nut_field().data(), sizeof(double) * r->n);
    if (turb && r->eq == FVK_EQ_RANS_SST) {
        for (int64_t c = 0; c < r->n; ++c) {
            turb[2 * c] = r->sst->k_field()[c];
            turb[2 * c + 1] = r->sst->omega_field()[c];
        }
    }
    return FVK_OK;
}

static void residual_of(RefSolver* r, double* out) {
    if (r->eq == FVK_EQ_DIFFUSION || r->eq == FVK_EQ_ADVECTION_DIFFUSION) {
        out[0] = r->eq == FVK_EQ_DIFFUSION ? r->diff->residual_norm() : r->adv->residual_norm();
        return;
    }
    const EulerResidualNorms* n = nullptr;
    switch (r->eq) {
        case FVK_EQ_EULER: n = &r->euler->residual(); break;
        case FVK_EQ_NAVIER_STOKES: n = &r->ns->residual(); break;
        case FVK_EQ_RANS_SA: n = &r->sa->residual(); out[4] = r->sa->nut_residual(); break;
        case FVK_EQ_RANS_SST:
            n = &r->sst->residual();
            out[4] = r->sst->k_residual();
            out[5] = r->sst->omega_residual();
            break;
    }
    out[0] = n->rho; out[1] = n->rho_u; out[2] = n->rho_v; out[3] = n->E;
}

int fvr_step(void* h, int32_t nsteps, double* res_hist) {
    auto* r = static_cast<RefSolver*>(h);
    int nv = nvars(r->eq);
    for (int s = 0; s < nsteps; ++s) {
        switch (r->eq) {
            case FVK_EQ_EULER: r->euler->step(); break;
            case FVK_EQ_NAVIER_STOKES: r->ns->step(); break;
            case FVK_EQ_RANS_SA: r->sa->step(); break;
            case FVK_EQ_RANS_SST: r->sst->step(); break;
            case FVK_EQ_DIFFUSION: r->diff->step(); break;
            case FVK_EQ_ADVECTION_DIFFUSION: r->adv->step(); break;
        }
        if (res_hist) residual_of(r, res_hist + (int64_t)s * nv);
    }
    return FVK_OK;
}

void fvr_set_cfl(void* h, double cfl) {
    auto* r = static_cast<RefSolver*>(h);
    switch (r->eq) {
        case FVK_EQ_EULER: r->euler->set_cfl(cfl); break;
        case FVK_EQ_NAVIER_STOKES: r->ns->set_cfl(cfl); break;
        case FVK_EQ_RANS_SA: r->sa->set_cfl(cfl); break;
        case FVK_EQ_RANS_SST: r->sst->set_cfl(cfl); break;
    }
}

// Wall samples and force reports: the solvers' compute_wall_traction_samples
// (NS, SA, SST) or, for Euler and pressure_only, pressure samples built as
// compute_wall_traction fills them with zero viscosity (p of the wall cell,
// tau = 0); then the coefficient functions and compute_wall_forces.
int fvr_wall_reports(void* h, const int32_t* wf, int64_t n, const fvk_wall_reference* rq, int32_t pressure_only,
                     fvk_wall_sample* samples, fvk_wall_force_report* reports, int32_t cap, int32_t* n_reports) {
    auto* r = static_cast<RefSolver*>(h);
    std::vector<int> faces(wf, wf + n);
    std::vector<WallFaceSample> ws;
    if (!pressure_only && r->eq == FVK_EQ_NAVIER_STOKES) ws = r->ns->compute_wall_traction_samples(faces);
    else if (!pressure_only && r->eq == FVK_EQ_RANS_SA) ws = r->sa->compute_wall_traction_samples(faces);
    else if (!pressure_only && r->eq == FVK_EQ_RANS_SST) ws = r->sst->compute_wall_traction_samples(faces);
    else {
        const std::vector<EulerState>* U = nullptr;
        switch (r->eq) {
            case FVK_EQ_EULER: U = &r->euler->field(); break;
            case FVK_EQ_NAVIER_STOKES: U = &r->ns->field(); break;
            case FVK_EQ_RANS_SA: U = &r->sa->field(); break;
            case FVK_EQ_RANS_SST: U = &r->sst->field(); break;
            default: return FVK_ERR_UNSUPPORTED;
        }
        for (int f : faces) {
            const Face& face = r->mesh.faces[f];
            WallFaceSample smp;
            smp.face_index = f;
            smp.x_mid = face.x_mid;
            smp.y_mid = face.y_mid;
            smp.p = pressure((*U)[face.cell_left], r->gamma);
            smp.rho = (*U)[face.cell_left].rho;
            smp.effective_viscosity = 0.0;
            smp.tau_wall = 0.0;
            smp.tx = -face.ny;
            smp.ty = face.nx;
            smp.y_wall_normal = face_normal_distance(r->mesh, face);
            ws.push_back(smp);
        }
    }
    WallReferenceQuantities ref;
    ref.rho_ref = rq->rho_ref;
    ref.velocity_x_ref = rq->velocity_x_ref;
    ref.velocity_y_ref = rq->velocity_y_ref;
    ref.p_ref = rq->p_ref;
    ref.length_ref = rq->length_ref;
    ref.moment_reference_x = rq->moment_reference_x;
    ref.moment_reference_y = rq->moment_reference_y;
    for (size_t i = 0; samples && i < ws.size(); ++i) {
        const WallFaceSample& w = ws[i];
        fvk_wall_sample& o = samples[i];
        o.face_index = w.face_index;
        o.patch_id = r->mesh.faces[w.face_index].patch_id;
        o.x_mid = w.x_mid; o.y_mid = w.y_mid; o.p = w.p; o.rho = w.rho;
        o.effective_viscosity = w.effective_viscosity; o.tau_wall = w.tau_wall;
        o.tx = w.tx; o.ty = w.ty; o.y_wall_normal = w.y_wall_normal;
        o.cf = skin_friction_coefficient(w, ref);
        o.cp = pressure_coefficient(w, ref);
        o.y_plus = wall_y_plus(w);
    }
    std::vector<WallForceReport> rep = compute_wall_forces(r->mesh, ws, ref);
    for (size_t i = 0; i < rep.size() && (int32_t)i < cap; ++i) {
        const WallForceReport& w = rep[i];
        fvk_wall_force_report& o = reports[i];
        o.patch_id = w.patch_id;
        o.reserved0 = 0;
        o.friction_drag = w.friction_drag; o.pressure_drag = w.pressure_drag; o.total_drag = w.total_drag;
        o.cd_friction = w.cd_friction; o.cd_pressure = w.cd_pressure; o.cd_total = w.cd_total;
        o.lift = w.lift; o.cl = w.cl; o.moment = w.moment; o.cm = w.cm;
    }
    *n_reports = (int32_t)rep.size();
    return (int32_t)rep.size() > cap ? FVK_ERR_INVALID : FVK_OK;
}

// Checkpoint::write of the solver's field with the solver's own tag (the
// reference driver's closing write, src/main.cpp:3553 and twins).
int fvr_checkpoint_write(void* h, const char* path, int64_t step, uint64_t build) {
    auto* r = static_cast<RefSolver*>(h);
    bool ok = false;
    switch (r->eq) {
        case FVK_EQ_DIFFUSION:
            ok = Checkpoint::write(path, CheckpointEquation::Diffusion, step, build, r->diff->field());
            break;
        case FVK_EQ_ADVECTION_DIFFUSION:
            ok = Checkpoint::write(path, CheckpointEquation::AdvectionDiffusion, step, build, r->adv->field());
            break;
        case FVK_EQ_EULER: ok = Checkpoint::write(path, CheckpointEquation::Euler, step, build, r->euler->field()); break;
        case FVK_EQ_NAVIER_STOKES:
            ok = Checkpoint::write(path, CheckpointEquation::NavierStokes, step, build, r->ns->field());
            break;
        case FVK_EQ_RANS_SA:
            ok = Checkpoint::write(path, CheckpointEquation::RANS_SA, step, build, r->sa->field(), r->sa->nut_field());
            break;
        case FVK_EQ_RANS_SST:
            ok = Checkpoint::write(path, CheckpointEquation::RANS_SST, step, build, r->sst->field(), r->sst->k_field(),
                                   r->sst->omega_field());
            break;
    }
    return ok ? FVK_OK : FVK_ERR_INVALID;
}

// Checkpoint::read of a file with the given tag and cell count: U [4n]
// (diffusion: phi [n]) and turb [nt n] (nut, or k then omega, AoS per cell).
int fvr_checkpoint_read(const char* path, int32_t tag, int64_t n, int64_t* step, uint64_t* build, double* U,
                        double* turb) {
    long long st = 0;
    unsigned long long bd = 0;
    bool ok = false;
    auto cp = [&](const std::vector<EulerState>& v) {
        for (int64_t c = 0; c < n; ++c) {
            U[4 * c] = v[c].rho; U[4 * c + 1] = v[c].rho_u; U[4 * c + 2] = v[c].rho_v; U[4 * c + 3] = v[c].E;
        }
    };
    CheckpointEquation eq = static_cast<CheckpointEquation>(tag);
    if (eq == CheckpointEquation::Diffusion || eq == CheckpointEquation::AdvectionDiffusion) {
        std::vector<double> phi;
        ok = Checkpoint::read(path, eq, (size_t)n, st, bd, phi);
        if (ok) std::memcpy(U, phi.data(), sizeof(double) * n);
    } else if (eq == CheckpointEquation::RANS_SA) {
        std::vector<EulerState> v;
        std::vector<double> nut;
        ok = Checkpoint::read(path, eq, (size_t)n, st, bd, v, nut);
        if (ok) { cp(v); std::memcpy(turb, nut.data(), sizeof(double) * n); }
    } else if (eq == CheckpointEquation::RANS_SST) {
        std::vector<EulerState> v;
        std::vector<double> k, w;
        ok = Checkpoint::read(path, eq, (size_t)n, st, bd, v, k, w);
        if (ok) {
            cp(v);
            for (int64_t c = 0; c < n; ++c) { turb[2 * c] = k[c]; turb[2 * c + 1] = w[c]; }
        }
    } else {
        std::vector<EulerState> v;
        ok = Checkpoint::read(path, eq, (size_t)n, st, bd, v);
        if (ok) cp(v);
    }
    *step = st;
    *build = bd;
    return ok ? FVK_OK : FVK_ERR_INVALID;
}

// MeshReader::read into a heap mesh: counts [nodes, cells, faces, patches,
// sum of cell face counts]; fvr_mesh_arrays copies its arrays out.
int fvr_mesh_read(const char* path, void** out, int64_t* counts) {
    auto* m = new UnstructuredMesh;
    if (!MeshReader::read(path, *m)) {
        delete m;
        return FVK_ERR_INVALID;
    }
    int64_t nnz = 0;
    for (const Cell& c : m->cells) nnz += (int64_t)c.faces.size();
    counts[0] = (int64_t)m->nodes.size();
    counts[1] = (int64_t)m->cells.size();
    counts[2] = (int64_t)m->faces.size();
    counts[3] = (int64_t)m->patches.size();
    counts[4] = nnz;
    *out = m;
    return FVK_OK;
}

int fvr_mesh_arrays(void* h, double* node_xy, int32_t* face_nodes, int32_t* face_cells, int32_t* face_patch,
                    int32_t* cell_ptr, int32_t* cell_faces, int32_t* cell_nodes, double* cell_volume, double* cell_xy,
                    char* names, int64_t names_cap) {
    auto* m = static_cast<UnstructuredMesh*>(h);
    for (size_t i = 0; i < m->nodes.size(); ++i) { node_xy[2 * i] = m->nodes[i].x; node_xy[2 * i + 1] = m->nodes[i].y; }
    for (size_t f = 0; f < m->faces.size(); ++f) {
        const Face& fc = m->faces[f];
        face_nodes[2 * f] = fc.node1; face_nodes[2 * f + 1] = fc.node2;
        face_cells[2 * f] = fc.cell_left; face_cells[2 * f + 1] = fc.cell_right;
        face_patch[f] = fc.patch_id;
    }
    int64_t k = 0;
    cell_ptr[0] = 0;
    for (size_t c = 0; c < m->cells.size(); ++c) {
        const Cell& cl = m->cells[c];
        for (size_t j = 0; j < cl.faces.size(); ++j) {
            cell_faces[k] = cl.faces[j];
            cell_nodes[k] = cl.node_ids[j];
            ++k;
        }
        cell_ptr[c + 1] = (int32_t)k;
        cell_volume[c] = cl.volume;
        cell_xy[2 * c] = cl.x_centroid; cell_xy[2 * c + 1] = cl.y_centroid;
    }
    std::string all;
    for (const BoundaryPatch& p : m->patches) all += p.name + "\n";
    if ((int64_t)all.size() + 1 > names_cap) return FVK_ERR_INVALID;
    std::memcpy(names, all.c_str(), all.size() + 1);
    return FVK_OK;
}

void fvr_mesh_free(void* h) { delete static_cast<UnstructuredMesh*>(h); }

// ---- standalone functions -------------------------------------------------

int fvr_flux(int32_t scheme, int64_t n, const double* UL, const double* UR, const double* nxy, double gamma,
             double tol, int32_t max_iter, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        EulerState L = st(UL + 4 * i), R = st(UR + 4 * i);
        double nx = nxy[2 * i], ny = nxy[2 * i + 1];
        EulerState F;
        if (scheme == FVK_FLUX_HLLC) F = hllc_flux(L, R, nx, ny, gamma);
        else if (scheme == FVK_FLUX_EXACT) F = exact_riemann_flux(L, R, nx, ny, gamma, tol, max_iter);
        else F = rusanov_flux(L, R, nx, ny, gamma);
        out[4 * i] = F.rho; out[4 * i + 1] = F.rho_u; out[4 * i + 2] = F.rho_v; out[4 * i + 3] = F.E;
    }
    return FVK_OK;
}

int fvr_star_pressure(int64_t n, const double* prim /*rho_L,u_L,p_L,rho_R,u_R,p_R*/, double gamma, double tol,
                      int32_t max_iter, double* p_star, double* u_star) {
    for (int64_t i = 0; i < n; ++i) {
        const double* q = prim + 6 * i;
        double cL = std::sqrt(gamma * q[2] / q[0]), cR = std::sqrt(gamma * q[5] / q[3]);
        double ps = ExactRiemannDetail::solve_star_pressure(q[0], q[1], q[2], cL, q[3], q[4], q[5], cR, gamma, tol,
                                                            max_iter);
        p_star[i] = ps;
        u_star[i] = 0.5 * (q[1] + q[4]) + 0.5 * (ExactRiemannDetail::pressure_function(ps, q[3], q[5], cR, gamma) -
                                                  ExactRiemannDetail::pressure_function(ps, q[0], q[2], cL, gamma));
    }
    return FVK_OK;
}

int fvr_geometry(const fvk_mesh_desc* d, double* mid, double* area, double* nxy) {
    UnstructuredMesh m = build_mesh(d, nullptr, 0);
    m.compute_geometry();
    for (int64_t f = 0; f < d->n_faces; ++f) {
        const Face& face = m.faces[f];
        mid[2 * f] = face.x_mid; mid[2 * f + 1] = face.y_mid;
        area[f] = face.area;
        nxy[2 * f] = face.nx; nxy[2 * f + 1] = face.ny;
    }
    return FVK_OK;
}

int fvr_gradient(const fvk_mesh_desc* d, int32_t scheme, const double* phi, const double* bphi, double* grad) {
    UnstructuredMesh m = build_mesh(d, nullptr, 0);
    m.compute_geometry();
    GradientCalculator calc(m, grad_of(scheme));
    std::vector<double> p(phi, phi + d->n_cells), b(bphi, bphi + d->n_faces);
    std::vector<Gradient2> g = calc.compute(m, p, b);
    for (int64_t c = 0; c < d->n_cells; ++c) { grad[2 * c] = g[c].dphidx; grad[2 * c + 1] = g[c].dphidy; }
    return FVK_OK;
}

// face_gradient + corrected_face_gradient_vector + face_normal_distance for
// every face, with the caller's cell gradients (src/GradientReconstruction.cpp:153-200).
int fvr_face_gradient(const fvk_mesh_desc* d, const double* phi, const double* bphi, const double* grad,
                      double* dphidn, double* gvec, double* normal_distance) {
    UnstructuredMesh m = build_mesh(d, nullptr, 0);
    m.compute_geometry();
    for (int64_t f = 0; f < d->n_faces; ++f) {
        const Face& face = m.faces[f];
        int cl = face.cell_left, cr = face.cell_right;
        bool b = cr == -1;
        Gradient2 gl{grad[2 * cl], grad[2 * cl + 1]};
        Gradient2 gr = b ? Gradient2{} : Gradient2{grad[2 * cr], grad[2 * cr + 1]};
        double phi_R = b ? bphi[f] : phi[cr];
        FaceGradient fg = face_gradient(m, face, phi[cl], phi_R, gl, gr, bphi[f]);
        Gradient2 v = corrected_face_gradient_vector(fg, face.nx, face.ny);
        dphidn[f] = fg.dphidn;
        gvec[2 * f] = v.dphidx; gvec[2 * f + 1] = v.dphidy;
        normal_distance[f] = face_normal_distance(m, face);
    }
    return FVK_OK;
}

int fvr_wall_distance(const fvk_mesh_desc* d, const int32_t* wall_faces, int64_t n_wall, double* out) {
    UnstructuredMesh m = build_mesh(d, nullptr, 0);
    m.compute_geometry();
    std::vector<int> w(wall_faces, wall_faces + n_wall);
    std::vector<double> dist = compute_wall_distance(m, w);
    std::memcpy(out, dist.data(), sizeof(double) * d->n_cells);
    return FVK_OK;
}

int fvr_sa_source(int64_t n, const double* nut, const double* nu, const double* omega, const double* dist,
                  const double* grad_nut, const fvk_sa_constants* c, double* out /* [3n] P, D, CD */,
                  double* eddy /* [n] */) {
    SAModelConstants k = sa_of(*c);
    for (int64_t i = 0; i < n; ++i) {
        SASourceTerms s = compute_sa_source_terms(nut[i], nu[i], omega[i], dist[i],
                                                  Gradient2{grad_nut[2 * i], grad_nut[2 * i + 1]}, k);
        out[3 * i] = s.production; out[3 * i + 1] = s.destruction; out[3 * i + 2] = s.cross_diffusion;
        eddy[i] = sa_eddy_viscosity(nut[i], nu[i], k);
    }
    return FVK_OK;
}

int fvr_sst_source(int64_t n, const double* rho, const double* k, const double* omega, const double* nu,
                   const double* dist, const double* gu, const double* gv, const double* gk, const double* gw,
                   int32_t variant, const fvk_sst_constants* c, int32_t kato, double* out /* [10n] */) {
    SSTModelConstants kc = sst_of(*c);
    SSTLimiterVariant v = variant == FVK_SST_STRAIN_RATE ? SSTLimiterVariant::StrainRate : SSTLimiterVariant::Vorticity;
    for (int64_t i = 0; i < n; ++i) {
        SSTSourceTerms s = compute_sst_source_terms(
            rho[i], k[i], omega[i], nu[i], dist[i], Gradient2{gu[2 * i], gu[2 * i + 1]},
            Gradient2{gv[2 * i], gv[2 * i + 1]}, Gradient2{gk[2 * i], gk[2 * i + 1]},
            Gradient2{gw[2 * i], gw[2 * i + 1]}, v, kc, kato != 0);
        double* o = out + 10 * i;
        o[0] = s.k_production; o[1] = s.k_destruction; o[2] = s.omega_production; o[3] = s.omega_destruction;
        o[4] = s.cross_diffusion; o[5] = s.nu_t; o[6] = s.F1; o[7] = s.F2; o[8] = s.S; o[9] = s.Omega;
    }
    return FVK_OK;
}

double fvr_next_cfl(double current, double cfl_max, const double* hist, int32_t n_hist, double cfl_min,
                    double growth, double shrink, int32_t window, double divergence_threshold) {
    CflRampParams p;
    p.cfl_min = cfl_min; p.growth_factor = growth; p.shrink_factor = shrink; p.window = window;
    p.divergence_threshold = divergence_threshold;
    return next_cfl(current, cfl_max, std::vector<double>(hist, hist + n_hist), p);
}

}  // extern "C"
