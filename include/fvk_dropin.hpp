// C++ drop-ins for the reference's solver classes over the fvk C ABI.
//
// A reference build that adds this header (with fvk.h) to its include path
// and links libfvk.so can construct HipEulerFVMSolver wherever it
// constructs EulerFVMSolver, and HipNavierStokesFVMSolver for
// NavierStokesFVMSolver: the constructors take the same arguments and the
// classes have the same public members (include/EulerFVMSolver.h:107-160,
// include/NavierStokesFVMSolver.h:118-140), so a case-file switch selects
// the GPU path with no other change to the driver.  It needs the reference's
// own headers (UnstructuredMesh.h, EulerFVMSolver.h, NavierStokesFVMSolver.h,
// GradientReconstruction.h) and nothing of its sources.
//
// The mesh is handed over as the reader built it: node order, face order and
// orientation, each cell's face order and the reader's cell volumes and
// centroids are kept (they fix the order of every sum).  The library maps
// its internal renumbering back on every copy, so field() is in mesh order.
//
// Extensions of fvk (second order, SSP-RK, local time steps, kernel shapes)
// are off unless a constructor's last argument switches them on; with them
// off the residual histories are the reference's (first-order paths bit for
// bit).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "EulerFVMSolver.h"
#include "GradientReconstruction.h"
#include "NavierStokesFVMSolver.h"
#include "UnstructuredMesh.h"
#include "fvk.h"

namespace fvk_dropin {

inline void check(int rc) {
    if (rc != FVK_OK) throw std::runtime_error(std::string("fvk: ") + fvk_last_error());
}

// The reference mesh flattened into fvk_mesh_desc arrays (owned here).
class MeshArrays {
public:
    explicit MeshArrays(const UnstructuredMesh& mesh, int32_t n_patches) {
        xy_.reserve(2 * mesh.nodes.size());
        for (const Node& n : mesh.nodes) {
            xy_.push_back(n.x);
            xy_.push_back(n.y);
        }
        for (const Face& f : mesh.faces) {
            fn_.push_back(f.node1);
            fn_.push_back(f.node2);
            fc_.push_back(f.cell_left);
            fc_.push_back(f.cell_right);
            fp_.push_back(f.patch_id);
        }
        ptr_.push_back(0);
        for (const Cell& c : mesh.cells) {
            for (int f : c.faces) cf_.push_back(f);
            ptr_.push_back((int32_t)cf_.size());
            vol_.push_back(c.volume);
            cxy_.push_back(c.x_centroid);
            cxy_.push_back(c.y_centroid);
        }
        d_ = fvk_mesh_desc{};
        d_.n_nodes = (int64_t)mesh.nodes.size();
        d_.n_cells = (int64_t)mesh.cells.size();
        d_.n_faces = (int64_t)mesh.faces.size();
        d_.n_patches = n_patches;
        d_.node_xy = xy_.data();
        d_.face_nodes = fn_.data();
        d_.face_cells = fc_.data();
        d_.face_patch = fp_.data();
        d_.cell_face_ptr = ptr_.data();
        d_.cell_faces = cf_.data();
        d_.cell_volume = vol_.data();
        d_.cell_xy = cxy_.data();
    }
    const fvk_mesh_desc* desc() const { return &d_; }

private:
    std::vector<double> xy_, vol_, cxy_;
    std::vector<int32_t> fn_, fc_, fp_, ptr_, cf_;
    fvk_mesh_desc d_;
};

inline fvk_initial_condition initial_condition(const EulerInitialCondition& ic) {
    fvk_initial_condition q;
    fvk_initial_condition_default(&q);
    q.mode = ic.mode == EulerICMode::Freestream ? FVK_IC_FREESTREAM : FVK_IC_TWO_REGION;
    q.rho = ic.rho; q.u = ic.u; q.v = ic.v; q.p = ic.p;
    q.rho_l = ic.rho_l; q.u_l = ic.u_l; q.v_l = ic.v_l; q.p_l = ic.p_l;
    q.rho_r = ic.rho_r; q.u_r = ic.u_r; q.v_r = ic.v_r; q.p_r = ic.p_r;
    q.x0 = ic.x0;
    return q;
}

inline void set_farfield(fvk_bc& b, const EulerState& ff) {
    b.farfield_state[0] = ff.rho;
    b.farfield_state[1] = ff.rho_u;
    b.farfield_state[2] = ff.rho_v;
    b.farfield_state[3] = ff.E;
}

// State shared by the conserved-variable solvers: the library handle, the
// last step's norms and a lazily refreshed host copy of the field.
class ConservedSolver {
public:
    ConservedSolver(const ConservedSolver&) = delete;
    ConservedSolver& operator=(const ConservedSolver&) = delete;
    virtual ~ConservedSolver() { fvk_destroy(s_); }

    // The library's defaults (the reference method).  A modified copy passed
    // as a constructor's `extensions` argument switches fvk's extensions and
    // execution tuning on; the constructor sets the reference arguments'
    // fields over it.
    static fvk_params defaults() {
        fvk_params p;
        fvk_params_default(&p);
        return p;
    }

    void step() {
        double res[6] = {0, 0, 0, 0, 0, 0};
        check(fvk_step(s_, 1, res, nullptr));
        norms_.rho = res[0];
        norms_.rho_u = res[1];
        norms_.rho_v = res[2];
        norms_.E = res[3];
        stale_ = true;
    }
    // As the reference's run(): total_steps steps, each with its residual.
    void run(int total_steps) {
        for (int i = 0; i < total_steps; ++i) step();
    }
    const std::vector<EulerState>& field() const {
        if (stale_) {
            std::vector<double> u(4 * n_);
            check(fvk_get_field(s_, u.data(), nullptr, (int64_t)n_));
            host_.resize(n_);
            for (size_t i = 0; i < n_; ++i) host_[i] = EulerState{u[4 * i], u[4 * i + 1], u[4 * i + 2], u[4 * i + 3]};
            stale_ = false;
        }
        return host_;
    }
    double gas_gamma() const { return fvk_gas_gamma(s_); }
    // A NaN in the norms is returned, not thrown: the driver's divergence
    // check stays in charge (src/main.cpp:56-69).
    const EulerResidualNorms& residual() const { return norms_; }
    void set_field(const std::vector<EulerState>& new_U) {
        if (new_U.size() != n_) throw std::invalid_argument("set_field: one state per mesh cell");
        std::vector<double> u(4 * n_);
        for (size_t i = 0; i < n_; ++i) {
            u[4 * i] = new_U[i].rho;
            u[4 * i + 1] = new_U[i].rho_u;
            u[4 * i + 2] = new_U[i].rho_v;
            u[4 * i + 3] = new_U[i].E;
        }
        check(fvk_set_field(s_, u.data(), nullptr, (int64_t)n_));
        host_ = new_U;
        stale_ = false;
        norms_ = EulerResidualNorms{};
    }
    void set_cfl(double new_cfl) { check(fvk_set_cfl(s_, new_cfl)); }
    // The library handle, for the extensions' entry points (fvk.h).
    fvk_solver* handle() const { return s_; }

protected:
    ConservedSolver(const UnstructuredMesh& mesh, const std::vector<fvk_bc>& bcs, const fvk_params& p,
                    const fvk_initial_condition& q)
        : n_(mesh.cells.size()) {
        MeshArrays a(mesh, (int32_t)bcs.size());
        check(fvk_create(a.desc(), bcs.data(), (int32_t)bcs.size(), &p, &q, &s_));
    }

private:
    fvk_solver* s_ = nullptr;
    size_t n_;
    EulerResidualNorms norms_;
    mutable std::vector<EulerState> host_;
    mutable bool stale_ = true;
};

inline int32_t flux_of(NumericalFluxScheme f) {
    return f == NumericalFluxScheme::Rusanov ? FVK_FLUX_RUSANOV
                                             : (f == NumericalFluxScheme::HLLC ? FVK_FLUX_HLLC : FVK_FLUX_EXACT);
}

// EulerFVMSolver's constructor and public API (include/EulerFVMSolver.h:107-160).
class HipEulerFVMSolver : public ConservedSolver {
public:
    HipEulerFVMSolver(const UnstructuredMesh& input_mesh, const std::vector<EulerBoundaryCondition>& boundary_conditions,
                      double gamma, double cfl, const EulerInitialCondition& initial_condition,
                      NumericalFluxScheme flux_scheme, double exact_riemann_tol, int exact_riemann_max_iter,
                      const fvk_params& extensions = defaults())
        : ConservedSolver(input_mesh, bcs_of(boundary_conditions),
                          params_of(gamma, cfl, flux_scheme, exact_riemann_tol, exact_riemann_max_iter, extensions),
                          fvk_dropin::initial_condition(initial_condition)) {}

private:
    static std::vector<fvk_bc> bcs_of(const std::vector<EulerBoundaryCondition>& in) {
        std::vector<fvk_bc> b(in.size());
        for (size_t i = 0; i < in.size(); ++i) {
            b[i] = fvk_bc{};
            b[i].type = (int32_t)in[i].type;  // Wall 0, Farfield 1, Outflow 2 in both
            set_farfield(b[i], in[i].farfield_state);
        }
        return b;
    }
    static fvk_params params_of(double gamma, double cfl, NumericalFluxScheme flux, double tol, int max_iter,
                                fvk_params p) {
        p.equation = FVK_EQ_EULER;
        p.flux_scheme = flux_of(flux);
        p.gamma = gamma;
        p.cfl = cfl;
        p.exact_riemann_tol = tol;
        p.exact_riemann_max_iter = max_iter;
        return p;
    }
};

// NavierStokesFVMSolver's constructor and its stepping API
// (include/NavierStokesFVMSolver.h:118-140).
class HipNavierStokesFVMSolver : public ConservedSolver {
public:
    HipNavierStokesFVMSolver(const UnstructuredMesh& input_mesh,
                             const std::vector<NSBoundaryCondition>& boundary_conditions, double gamma,
                             double gas_constant, double mu, double prandtl, double cfl,
                             const EulerInitialCondition& initial_condition, NumericalFluxScheme flux_scheme,
                             GradientScheme gradient_scheme, double exact_riemann_tol, int exact_riemann_max_iter,
                             const fvk_params& extensions = defaults())
        : ConservedSolver(input_mesh, bcs_of(boundary_conditions),
                          params_of(gamma, gas_constant, mu, prandtl, cfl, flux_scheme, gradient_scheme,
                                    exact_riemann_tol, exact_riemann_max_iter, extensions),
                          fvk_dropin::initial_condition(initial_condition)) {}

private:
    static std::vector<fvk_bc> bcs_of(const std::vector<NSBoundaryCondition>& in) {
        std::vector<fvk_bc> b(in.size());
        for (size_t i = 0; i < in.size(); ++i) {
            b[i] = fvk_bc{};
            b[i].type = (int32_t)in[i].type;  // NoSlipWall 0, Farfield 1, Outflow 2
            b[i].wall_u = in[i].wall_u;
            b[i].wall_v = in[i].wall_v;
            b[i].is_isothermal_wall = in[i].is_isothermal_wall ? 1 : 0;
            b[i].wall_temperature = in[i].wall_temperature;
            set_farfield(b[i], in[i].farfield_state);
        }
        return b;
    }
    static fvk_params params_of(double gamma, double gas_constant, double mu, double prandtl, double cfl,
                                NumericalFluxScheme flux, GradientScheme grad, double tol, int max_iter,
                                fvk_params p) {
        p.equation = FVK_EQ_NAVIER_STOKES;
        p.flux_scheme = flux_of(flux);
        p.gradient_scheme = grad == GradientScheme::GreenGauss ? FVK_GRAD_GREEN_GAUSS : FVK_GRAD_LEAST_SQUARES;
        p.gamma = gamma;
        p.gas_constant = gas_constant;
        p.mu = mu;
        p.prandtl = prandtl;
        p.cfl = cfl;
        p.exact_riemann_tol = tol;
        p.exact_riemann_max_iter = max_iter;
        return p;
    }
};

}  // namespace fvk_dropin

using fvk_dropin::HipEulerFVMSolver;
using fvk_dropin::HipNavierStokesFVMSolver;
