// Runs the C++ drop-ins (include/fvk_dropin.hpp) on a mesh and case read
// from a text file, and prints the per-step residual norms and the final
// field, for tests/test_gpu_dropin.py to compare with the library's own
// binding.  Built by `make -C oracle dropin` against the reference's headers.
//
// Input (whitespace separated):
//   n_nodes n_faces n_cells n_patches
//   n_nodes x { x y }
//   n_faces x { node1 node2 cell_left cell_right patch_id }
//   n_cells x { volume x_c y_c n_faces f... n_nodes v... }
//   equation(0 Euler, 1 NS) flux gamma cfl gas_constant mu prandtl gradient(0 GG, 1 LSQ) exact_tol exact_iter
//   n_patches x { type rho rho_u rho_v E wall_u wall_v isothermal wall_T }
//   ic_mode rho u v p rho_l u_l v_l p_l rho_r u_r v_r p_r x0
//   steps
#include <cstdio>
#include <iostream>
#include <memory>

#include "fvk_dropin.hpp"

int main() {
    size_t nn, nf, nc, np;
    std::cin >> nn >> nf >> nc >> np;
    UnstructuredMesh mesh;
    mesh.nodes.resize(nn);
    for (auto& n : mesh.nodes) std::cin >> n.x >> n.y;
    mesh.faces.resize(nf);
    for (auto& f : mesh.faces) std::cin >> f.node1 >> f.node2 >> f.cell_left >> f.cell_right >> f.patch_id;
    mesh.cells.resize(nc);
    for (auto& c : mesh.cells) {
        size_t k;
        std::cin >> c.volume >> c.x_centroid >> c.y_centroid >> k;
        c.faces.resize(k);
        for (auto& f : c.faces) std::cin >> f;
        std::cin >> k;
        c.node_ids.resize(k);
        for (auto& v : c.node_ids) std::cin >> v;
    }
    int eq, flux, grad, iters;
    double gamma, cfl, R, mu, pr, tol;
    std::cin >> eq >> flux >> gamma >> cfl >> R >> mu >> pr >> grad >> tol >> iters;
    std::vector<EulerBoundaryCondition> ebc(np);
    std::vector<NSBoundaryCondition> nbc(np);
    for (size_t i = 0; i < np; ++i) {
        int type, iso;
        EulerState ff;
        double wu, wv, wt;
        std::cin >> type >> ff.rho >> ff.rho_u >> ff.rho_v >> ff.E >> wu >> wv >> iso >> wt;
        ebc[i].type = (EulerBoundaryType)type;
        ebc[i].farfield_state = ff;
        nbc[i].type = (NSBoundaryType)type;
        nbc[i].farfield_state = ff;
        nbc[i].wall_u = wu;
        nbc[i].wall_v = wv;
        nbc[i].is_isothermal_wall = iso != 0;
        nbc[i].wall_temperature = wt;
    }
    EulerInitialCondition ic;
    int mode;
    std::cin >> mode >> ic.rho >> ic.u >> ic.v >> ic.p >> ic.rho_l >> ic.u_l >> ic.v_l >> ic.p_l >> ic.rho_r >> ic.u_r >>
        ic.v_r >> ic.p_r >> ic.x0;
    ic.mode = mode == 0 ? EulerICMode::Freestream : EulerICMode::TwoRegion;
    int steps;
    std::cin >> steps;
    if (!std::cin) {
        std::fprintf(stderr, "dropin_check: malformed input\n");
        return 2;
    }
    const NumericalFluxScheme fs =
        flux == 0 ? NumericalFluxScheme::Rusanov : (flux == 1 ? NumericalFluxScheme::HLLC : NumericalFluxScheme::Exact);
    try {
        std::unique_ptr<fvk_dropin::ConservedSolver> s;
        if (eq == 0)
            s.reset(new HipEulerFVMSolver(mesh, ebc, gamma, cfl, ic, fs, tol, iters));
        else
            s.reset(new HipNavierStokesFVMSolver(mesh, nbc, gamma, R, mu, pr, cfl, ic, fs,
                                                 grad == 0 ? GradientScheme::GreenGauss : GradientScheme::LeastSquares,
                                                 tol, iters));
        for (int i = 0; i < steps; ++i) {
            s->step();
            const EulerResidualNorms& r = s->residual();
            std::printf("R %.17g %.17g %.17g %.17g\n", r.rho, r.rho_u, r.rho_v, r.E);
        }
        for (const EulerState& u : s->field()) std::printf("U %.17g %.17g %.17g %.17g\n", u.rho, u.rho_u, u.rho_v, u.E);
        std::printf("gamma %.17g\n", s->gas_gamma());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "dropin_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
