// Navier-Stokes, RANS (Spalart-Allmaras, SST k-omega) and scalar diffusion
// residual kernels for gfx950 (FP64).  Same three-pass structure as the
// Euler path:
//   k_visc_grad  per cell: u, v, T (+ nut | k, omega) gradients (LSQ or
//                Green-Gauss, GradientCalculator::compute), then the
//                turbulence source terms of the cell
//   k_visc_face  per face: inviscid flux minus viscous flux (Jasak-corrected
//                face gradients), turbulence advection/diffusion, dt candidate
//   k_visc_cell  per cell: gather in Cell::faces order + volumetric sources +
//                stage update + residual partials
// or (kernel variant 3, local or no dt) one fused pass over cell tiles,
// k_visc_tile, which does all three per tile in LDS.
#include <cfloat>

#include "norms.h"
#include "physics.h"
#include "solver.h"

namespace fvk {

constexpr int kVB = 256;

struct ViscArgs {
    MeshView m;
    const double* U;   // [nv][ld] stage state: rho, rho_u, rho_v, E (, nut | k, omega)
    double* Unew;
    const double* U0;
    double* G;         // [ld][2 nf] gradients, cell-major (a cell's in one 16-byte-aligned run):
                       //   u(0,1) v(2,3) T(4,5) nut|k(6,7) omega(8,9)
    double* aux;       // [4][ld]  SA: src; SST: k src, omega src, F1, nu_t
    double* flux;      // [nv][ldf]
    double* dtf;
    double* dtc;
    double* R;
    unsigned long long* dtmin;
    double* partials;
    double* res_hist;
    double* dt_hist;
    int64_t ld, ldf;
    int gscheme, maxit, sst_variant, kato;
    double gamma, Rgas, mu, pr, prt, cfl, tol, alpha, fixed_dt;
    fvk_sa_constants sa;
    fvk_sst_constants sst;
    double a, b;
    int face_dt, dt_src, first, want_res, slot;
    int64_t hist_idx;
    // the previous stage's norms, reduced by a spare workgroup of the
    // stage's first launch (gradient or tile; fin_nblocks == 0: none)
    int fin_nblocks, fin_slot, fin_nv;
    int fin_dt_row;    // the step's dt is the minimum of the partials' row fin_nv (local dt)
    int64_t fin_hist_idx;
};

enum : int { EQ_NS = 1, EQ_SA = 2, EQ_SST = 3 };

__device__ __forceinline__ State ldS(const double* U, int64_t ld, int c) {
    return State{U[c], U[ld + c], U[2 * ld + c], U[3 * ld + c]};
}

// ---- Spalart-Allmaras, src/SpalartAllmaras.cpp:8-69 ----
__device__ __forceinline__ double sa_fv1(double nut, double nu, const fvk_sa_constants& c) {
    if (nut <= 0.0) return 0.0;
    double chi = fdiv(nut, nu), chi3 = chi * chi * chi;
    return fdiv(chi3, chi3 + c.cv1 * c.cv1 * c.cv1);
}
__device__ __forceinline__ double sa_eddy(double nut, double nu, const fvk_sa_constants& c) {
    return nut * sa_fv1(nut, nu, c);
}
// The SA source's powers, the largest single cost of the SA stage as pow
// calls: x^6 in three multiplications for the reference's pow(x, 6.0), and
// y^(1/6) as cbrt(sqrt(y)) for its pow(y, 1.0 / 6.0); each within a few ulp
// (the device's pow is not the host's to the last bit either, so the SA
// path is held to the oracle within a tolerance, not bit for bit)
__device__ __forceinline__ double pow6(double x) {
    const double x2 = x * x;
    return x2 * x2 * x2;
}
__device__ __forceinline__ double root6(double y) { return cbrt(sqrt(y)); }
__device__ double sa_source(double nut, double nu, double omega, double d, double gx, double gy,
                            const fvk_sa_constants& c) {
    double chi = (nut > 0.0) ? fdiv(nut, nu) : 0.0;
    double fv1 = sa_fv1(nut, nu, c);
    double fv2 = 1.0 - fdiv(chi, 1.0 + chi * fv1);
    double d2 = d * d;
    double sbar = fdiv(nut, c.kappa * c.kappa * d2) * fv2;
    double st;
    if (sbar >= -c.cv2 * omega) st = omega + sbar;
    else st = omega + (omega * (c.cv2 * c.cv2 * omega + c.cv3 * sbar)) / ((c.cv3 - 2.0 * c.cv2) * omega - sbar);
    double P = c.cb1 * st * nut;
    double D = 0.0;
    if (nut > 0.0) {
        double sts = smax(st, 1e-10);
        double r = smin(fdiv(nut, sts * c.kappa * c.kappa * d2), 10.0);
        double gg = r + c.cw2 * (pow6(r) - r);
        const double cw36 = pow6(c.cw3);
        double fw = gg * root6(fdiv(1.0 + cw36, pow6(gg) + cw36));
        double nod = fdiv(nut, d);
        double cw1 = fdiv(c.cb1, c.kappa * c.kappa) + fdiv(1.0 + c.cb2, c.sigma);
        D = cw1 * fw * nod * nod;
    }
    double CD = (fdiv(c.cb2, c.sigma)) * (gx * gx + gy * gy);
    return P - D + CD;
}

// ---- SST k-omega, src/SSTKOmega.cpp:8-101; out: k src, omega src, F1, nu_t ----
__device__ void sst_source(double rho, double k, double w, double nu, double d, const double* gu, const double* gv,
                           const double* gk, const double* gw, int variant, int kato, const fvk_sst_constants& c,
                           double* out) {
    double s11 = gu[0], s22 = gv[1], s12 = 0.5 * (gu[1] + gv[0]);
    double S = sqrt(2.0 * (s11 * s11 + s22 * s22 + 2.0 * s12 * s12));
    double Om = fabs(gv[0] - gu[1]);
    if (k <= 0.0 || w <= 0.0) {
        // all source terms, F1 and nu_t stay zero (compute_sst_source_terms early return)
        out[0] = 0.0 - 0.0;
        out[1] = 0.0 - 0.0 + 0.0;
        out[2] = 0.0;
        out[3] = 0.0;
        return;
    }
    double d2 = d * d;
    double gdot = gk[0] * gw[0] + gk[1] * gw[1];
    double CD = smax(2.0 * rho * c.sigma_omega2 * gdot / w, 1e-20);
    double t1 = smax(sqrt(k) / (c.beta_star * w * d), 500.0 * nu / (d2 * w));
    double t2 = 4.0 * rho * c.sigma_omega2 * k / (CD * d2);
    double a1 = smin(t1, t2);
    double a14 = a1 * a1 * a1 * a1;
    double F1 = tanh(a14);
    double a2 = smax(2.0 * sqrt(k) / (c.beta_star * w * d), 500.0 * nu / (d2 * w));
    double F2 = tanh(a2 * a2);
    double lim = (variant == FVK_SST_VORTICITY) ? Om : S;
    double nut = c.a1 * k / smax(c.a1 * w, lim * F2);
    double prs = kato ? (S * Om) : (S * S);
    double mut = rho * nut;
    double bsrwk = c.beta_star * rho * w * k;
    double clip = (variant == FVK_SST_VORTICITY) ? 20.0 : 10.0;
    double kp = smin(mut * prs, clip * bsrwk);
    double kd = bsrwk;
    double g1 = fdiv(c.beta1, c.beta_star) - c.sigma_omega1 * c.kappa * c.kappa / sqrt(c.beta_star);
    double g2 = fdiv(c.beta2, c.beta_star) - c.sigma_omega2 * c.kappa * c.kappa / sqrt(c.beta_star);
    double gam = F1 * g1 + (1.0 - F1) * g2;
    double beta = F1 * c.beta1 + (1.0 - F1) * c.beta2;
    double wp = gam * rho * prs;
    double wd = beta * rho * w * w;
    double cdw = 2.0 * (1.0 - F1) * rho * c.sigma_omega2 * gdot / w;
    out[0] = kp - kd;
    out[1] = wp - wd + cdw;
    out[2] = F1;
    out[3] = nut;
}

// boundary field values of a boundary face (build_boundary_fields,
// src/NavierStokesFVMSolver.cpp:55-84; build_boundary_nut,
// src/RANSTurbulenceSASolver.cpp:98-115; build_boundary_k_omega,
// src/RANSTurbulenceSSTSolver.cpp:101-139).  b[0..2] = u, v, T; b[3] =
// nut | k; b[4] = omega; b[5], b[6] = rho k, rho omega.
template <int EQ>
__device__ __forceinline__ void boundary_values(const ViscArgs& a, const DevBC& bc, int f, const State& L, double uL,
                                                double vL, double TL, double t0, double t1, double* b) {
    if (bc.type == BC_WALL) {
        b[0] = bc.wall_u;
        b[1] = bc.wall_v;
        b[2] = bc.isothermal ? bc.wall_T : TL;
    } else {
        b[0] = uL; b[1] = vL; b[2] = TL;
    }
    if (EQ == EQ_SA) b[3] = bc.type == BC_WALL ? 0.0 : (bc.type == BC_FARFIELD ? bc.ff_nut : t0);
    if (EQ == EQ_SST) {
        if (bc.type == BC_WALL) {
            b[3] = 0.0;
            double nu1 = fdiv(a.mu, L.r);
            double d1 = a.m.fdc[f];
            b[4] = 60.0 * nu1 / (a.sst.beta1 * d1 * d1);
            b[5] = 0.0;
            b[6] = L.r * b[4];
        } else if (bc.type == BC_FARFIELD) {
            b[3] = bc.ff_k; b[4] = bc.ff_w;
            b[5] = bc.ff[0] * bc.ff_k; b[6] = bc.ff[0] * bc.ff_w;
        } else {
            b[3] = t0; b[4] = t1;
            b[5] = L.r * t0; b[6] = L.r * t1;
        }
    }
}

// NS / RANS ghost state, src/NavierStokesFVMSolver.cpp:265-284
__device__ __forceinline__ State visc_ghost(const DevBC& bc, const State& L, double g) {
    if (bc.type == BC_FARFIELD) return State{bc.ff[0], bc.ff[1], bc.ff[2], bc.ff[3]};
    if (bc.type == BC_OUTFLOW) return L;
    return ghost_noslip_wall(L, bc.wall_u, bc.wall_v, g);
}

// ---------------------------------------------------------------------------
// gradients (GradientCalculator::compute, src/GradientReconstruction.cpp:83-150)
// of NF fields at once, plus per-cell turbulence sources.  The per-cell pass
// and the viscous tile kernel share these steps, so both give the same bits.
// ---------------------------------------------------------------------------
// One slot's term of the sums.  Least squares: the per-entry weight times
// offset (geo = egeo: {w dx, w dy, ..}, periodic shift included) times the
// difference to the neighbour (or boundary) value.  Green-Gauss: the face
// value (distance-weighted, or the boundary value) times the normal pointing
// away from the cell times the length (geo = {the cell's weight, n_x, n_y, A}).
template <int NF, int GS>
__device__ __forceinline__ void grad_slot(double (&sx)[NF], double (&sy)[NF], const double (&pc)[NF],
                                          const double (&pn)[NF], bool bnd, const double4& geo) {
    if (GS == FVK_GRAD_LEAST_SQUARES) {
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            const double dphi = pn[q] - pc[q];
            sx[q] += geo.x * dphi;
            sy[q] += geo.y * dphi;
        }
    } else {
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            const double pf = bnd ? pn[q] : geo.x * pc[q] + (1.0 - geo.x) * pn[q];
            sx[q] += pf * geo.y * geo.w;
            sy[q] += pf * geo.z * geo.w;
        }
    }
}

// the gradients from the sums: iv = {LSQ inverse i0, i1, i2, .} (least
// squares) or {volume, ...} (Green-Gauss)
template <int NF, int GS>
__device__ __forceinline__ void grad_finish(const double (&sx)[NF], const double (&sy)[NF], const double4& iv,
                                            double (&gr)[2 * NF]) {
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        if (GS == FVK_GRAD_LEAST_SQUARES) {
            gr[2 * q] = iv.x * sx[q] + iv.y * sy[q];
            gr[2 * q + 1] = iv.y * sx[q] + iv.z * sy[q];
        } else {
            gr[2 * q] = fdiv(sx[q], iv.x);
            gr[2 * q + 1] = fdiv(sy[q], iv.x);
        }
    }
}

// a cell's turbulence sources from its density, fields and gradients:
// SA's source in o[0]; SST's k and omega sources, F1 and nu_t in o[0..3]
template <int EQ, int NF>
__device__ __forceinline__ void visc_sources(const ViscArgs& a, double rho, double t0c, double t1c, double wd,
                                             const double (&gr)[2 * NF], double (&o)[4]) {
    if (EQ == EQ_SA) {
        const double nul = fdiv(a.mu, rho);
        const double om = fabs(gr[2] - gr[1]);
        o[0] = sa_source(t0c, nul, om, wd, gr[6], gr[7], a.sa);
    }
    if (EQ == EQ_SST) {
        const double nul = fdiv(a.mu, rho);
        sst_source(rho, t0c, t1c, nul, wd, gr, gr + 2, gr + 6, gr + 2 * NF - 2, a.sst_variant, a.kato, a.sst, o);
    }
}

// PERIODIC selects nothing any more (the per-entry LSQ data carry the
// periodic shift); the launcher keeps the two instances
template <int EQ, int GS, bool PERIODIC>
__global__ __launch_bounds__(kVB) void k_visc_grad(ViscArgs a) {
    constexpr int NF = EQ == EQ_SST ? 5 : (EQ == EQ_SA ? 4 : 3);
    const MeshView& m = a.m;
    // a pending finalize takes the first logical workgroup, dispatched first so
    // that its one-workgroup reduction overlaps the cells; the cells follow it
    int lb = logical_block();
    if (a.fin_nblocks > 0) {
        if (lb == 0) {
            __shared__ double fsm[kVB];
            finalize_norms(a.partials, a.fin_nblocks, a.fin_nv, a.res_hist, a.dt_hist, a.fin_hist_idx, a.dtmin,
                           a.fin_slot, -1.0, fsm, true, a.fin_dt_row != 0);
            return;
        }
        --lb;
    }
    int c = lb * kVB + threadIdx.x;
    if (c >= m.nc) return;
    const double g = a.gamma;
    State Uc = ldS(a.U, a.ld, c);
    double pc[NF];
    pc[0] = fdiv(Uc.ru, Uc.r);
    pc[1] = fdiv(Uc.rv, Uc.r);
    pc[2] = temperature(Uc, g, a.Rgas);
    double t0c = 0.0, t1c = 0.0;
    if (EQ == EQ_SA) { t0c = a.U[4 * a.ld + c]; pc[3] = t0c; }
    if (EQ == EQ_SST) { t0c = a.U[4 * a.ld + c]; t1c = a.U[5 * a.ld + c]; pc[3] = t0c; pc[NF - 1] = t1c; }
    double sx[NF], sy[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) { sx[q] = 0.0; sy[q] = 0.0; }
    // the cell's faces in Cell::faces order from the ELL (neighbour and face in
    // one load); LSQ takes the per-entry weight times offset (egeo, periodic
    // shift included), Green-Gauss the face's weight and normal
    const int64_t ldm = m.ld;
    for (int k = 0; k < m.deg; ++k) {
        const int2 en = m.ell[k * ldm + c];
        if (en.x < 0) break;  // ELL padding (after the cell's faces)
        const int f = en.x >> 1;
        const int cr = en.y;  // neighbour, or -1 - patch on a boundary
        double pn[NF];
        if (cr < 0) {
            double bv[7];
            boundary_values<EQ>(a, m.bc[-1 - cr], f, Uc, pc[0], pc[1], pc[2], t0c, t1c, bv);
#pragma unroll
            for (int q = 0; q < NF; ++q) pn[q] = bv[q];
        } else {
            const int o = cr;
            State Uo = ldS(a.U, a.ld, o);
            pn[0] = fdiv(Uo.ru, Uo.r);
            pn[1] = fdiv(Uo.rv, Uo.r);
            pn[2] = temperature(Uo, g, a.Rgas);
            if (EQ == EQ_SA) pn[3] = a.U[4 * a.ld + o];
            if (EQ == EQ_SST) { pn[3] = a.U[4 * a.ld + o]; pn[NF - 1] = a.U[5 * a.ld + o]; }
        }
        double4 geo;
        if (GS == FVK_GRAD_LEAST_SQUARES) geo = m.egeo[k * ldm + c];  // {w dx, w dy, ...}, w = 1/(dx^2 + dy^2)
        else {
            const bool owner = en.x & 1;
            geo = double4{owner ? m.fwl[f] : m.fwr[f], owner ? m.fnx[f] : -m.fnx[f], owner ? m.fny[f] : -m.fny[f],
                          m.fa[f]};
        }
        grad_slot<NF, GS>(sx, sy, pc, pn, cr < 0, geo);
    }
    double gr[2 * NF];
    if (GS == FVK_GRAD_LEAST_SQUARES)
        grad_finish<NF, GS>(sx, sy, double4{m.lsq[c], m.lsq[m.nc + c], m.lsq[2 * m.nc + c], 0.0}, gr);
    else
        grad_finish<NF, GS>(sx, sy, double4{m.vol[c], 0.0, 0.0, 0.0}, gr);
    double2* gc = reinterpret_cast<double2*>(a.G + (int64_t)c * (2 * NF));
#pragma unroll
    for (int q = 0; q < NF; ++q) gc[q] = double2{gr[2 * q], gr[2 * q + 1]};
    if (EQ != EQ_NS) {
        double o4[4];
        visc_sources<EQ, NF>(a, Uc.r, t0c, t1c, m.wdist[c], gr, o4);
        if (EQ == EQ_SA) a.aux[c] = o4[0];
        if (EQ == EQ_SST) {
#pragma unroll
            for (int q = 0; q < 4; ++q) a.aux[q * a.ld + c] = o4[q];
        }
    }
}

// The geometry of one face that the viscous face body reads: unit normal,
// length, the Jasak distance d.n and correction vector k, the left cell's
// distance weight (face_normal_geometry, src/GradientReconstruction.cpp:28-40)
struct VFaceGeo {
    double nx, ny, A, dc, wl, kx, ky;
};
__device__ __forceinline__ VFaceGeo face_geo(const MeshView& m, int f) {
    return VFaceGeo{m.fnx[f], m.fny[f], m.fa[f], m.fdc[f], m.fwl[f], m.fkx[f], m.fky[f]};
}
// from a viscous tile's face record: r0 = {n_x, n_y, A, w_left}, r1 = {w_right, d.n, k_x, k_y}
__device__ __forceinline__ VFaceGeo face_geo(const double4& r0, const double4& r1) {
    return VFaceGeo{r0.x, r0.y, r0.z, r1.y, r0.w, r1.z, r1.w};
}

// Jasak corrected face gradient, src/GradientReconstruction.cpp:153-185, 195-200
struct FG {
    double dn, gx, gy;
};
__device__ __forceinline__ FG face_grad(const VFaceGeo& fg, bool b, double phiL, double phiRv, double gLx,
                                        double gLy, double gRx, double gRy) {
    double direct = fdiv(phiRv - phiL, fg.dc);
    FG o;
    if (b) { o.gx = gLx; o.gy = gLy; }
    else {
        double wL = fg.wl;
        o.gx = wL * gLx + (1.0 - wL) * gRx;
        o.gy = wL * gLy + (1.0 - wL) * gRy;
    }
    o.dn = direct + o.gx * fg.kx + o.gy * fg.ky;
    return o;
}
__device__ __forceinline__ FG face_grad(const MeshView& m, int f, bool b, double phiL, double phiRv, double gLx,
                                        double gLy, double gRx, double gRy) {
    return face_grad(face_geo(m, f), b, phiL, phiRv, gLx, gLy, gRx, gRy);
}
__device__ __forceinline__ void corrected(const FG& fg, double nx, double ny, double& gx, double& gy) {
    double avg = fg.gx * nx + fg.gy * ny;
    double delta = fg.dn - avg;
    gx = fg.gx + delta * nx;
    gy = fg.gy + delta * ny;
}

// The compute_dt candidate of one face (NS src/NavierStokesFVMSolver.cpp:231-258,
// SA src/RANSTurbulenceSASolver.cpp:313-346, SST src/RANSTurbulenceSSTSolver.cpp:380-429)
// from its two states (R unused on a boundary face, b), their velocities and
// turbulence fields, the normal and the face-normal distance len
template <int EQ>
__device__ __forceinline__ double visc_dt_candidate(const ViscArgs& a, bool b, const State& L, const State& R,
                                                    double uL, double vL, double uR, double vR, double t0L,
                                                    double t1L, double t0R, double t1R, double nx, double ny,
                                                    double len) {
    const double g = a.gamma;
    double cand;
    double VnL = uL * nx + vL * ny;
    double Smax = fabs(VnL) + sound_speed(L, g);
    double rmin = L.r;
    double ntmax = 0.0, wmax = 0.0;
    if (EQ == EQ_SA) ntmax = sa_eddy(t0L, fdiv(a.mu, L.r), a.sa);
    if (EQ == EQ_SST) { ntmax = (t0L > 0.0 && t1L > 0.0) ? fdiv(t0L, t1L) : 0.0; wmax = t1L; }
    if (!b) {
        double VnR = uR * nx + vR * ny;
        Smax = smax(Smax, fabs(VnR) + sound_speed(R, g));
        rmin = smin(rmin, R.r);
        if (EQ == EQ_SA) ntmax = smax(ntmax, sa_eddy(t0R, fdiv(a.mu, R.r), a.sa));
        if (EQ == EQ_SST) {
            ntmax = smax(ntmax, (t0R > 0.0 && t1R > 0.0) ? fdiv(t0R, t1R) : 0.0);
            wmax = smax(wmax, t1R);
        }
    }
    double nu = (a.mu > 0.0) ? fdiv(a.mu, rmin) : 0.0;
    if (EQ == EQ_NS) {
        double St = Smax + fdiv(2.0 * nu, len);
        cand = fdiv(a.cfl * len, St);
    } else {
        double nue = nu + ntmax;
        double St = Smax + fdiv(2.0 * nue, len);
        cand = fdiv(a.cfl * len, St);
        if (EQ == EQ_SST) {
            double coef = smax(a.sst.beta_star, smax(a.sst.beta1, a.sst.beta2));
            double dts = fdiv(a.cfl, coef * smax(wmax, 1e-30));
            cand = smin(cand, dts);
        }
    }
    return cand;
}

// ---------------------------------------------------------------------------
// face pass: NavierStokesFVMSolver::step pass 1 (src/NavierStokesFVMSolver.cpp:119-189),
// RANSTurbulenceSASolver (src/RANSTurbulenceSASolver.cpp:157-259),
// RANSTurbulenceSSTSolver (src/RANSTurbulenceSSTSolver.cpp:183-307) and the
// compute_dt candidates (NS :231-258, SA :313-346, SST :380-429)
// ---------------------------------------------------------------------------
// Per-cell field access of one face's two sides (0 = left, 1 = right):
// u(side, q) conserved field q (turbulence fields from q = 4), g(side, q)
// gradient component q, aux(side, q) per-cell auxiliary field q.  The face
// and tile kernels read them from HBM or LDS.
template <int EQ, int SCHEME, class Acc>
__device__ __forceinline__ double visc_face_body(const ViscArgs& a, int f, int cr, const VFaceGeo& fgeo,
                                                 const Acc& acc, double (&out)[6]) {
    const MeshView& m = a.m;
    double cand = DBL_MAX;
    {
        const double g = a.gamma;
        bool b = cr < 0;
        const DevBC* bc = b ? &m.bc[-1 - cr] : nullptr;
        double nx = fgeo.nx, ny = fgeo.ny, A = fgeo.A;
        State L{acc.u(0, 0), acc.u(0, 1), acc.u(0, 2), acc.u(0, 3)};
        State R = b ? L : State{acc.u(1, 0), acc.u(1, 1), acc.u(1, 2), acc.u(1, 3)};
        State Rs = b ? visc_ghost(*bc, L, g) : R;
        State inv = num_flux<SCHEME>(L, Rs, nx, ny, g, a.tol, a.maxit);
        double uL = fdiv(L.ru, L.r), vL = fdiv(L.rv, L.r), TL = temperature(L, g, a.Rgas);
        double uR = 0.0, vR = 0.0, TR = 0.0;
        if (!b) { uR = fdiv(R.ru, R.r); vR = fdiv(R.rv, R.r); TR = temperature(R, g, a.Rgas); }
        double t0L = 0.0, t1L = 0.0, t0R = 0.0, t1R = 0.0;
        if (EQ == EQ_SA || EQ == EQ_SST) { t0L = acc.u(0, 4); if (!b) t0R = acc.u(1, 4); }
        if (EQ == EQ_SST) { t1L = acc.u(0, 5); if (!b) t1R = acc.u(1, 5); }
        double bv[7] = {0, 0, 0, 0, 0, 0, 0};
        if (b) boundary_values<EQ>(a, *bc, f, L, uL, vL, TL, t0L, t1L, bv);
        double vx = 0.0, vy = 0.0, vE = 0.0, adv0 = 0.0, adv1 = 0.0, dif0 = 0.0, dif1 = 0.0;
        double Vn = 0.0;
        if (EQ != EQ_NS) {
            double uRg = fdiv(Rs.ru, Rs.r), vRg = fdiv(Rs.rv, Rs.r);
            Vn = 0.5 * ((uL + uRg) * nx + (vL + vRg) * ny);
        }
        if (EQ == EQ_SA) {
            double nR = b ? bv[3] : t0R;
            double up = (Vn >= 0.0) ? t0L : nR;
            adv0 = Vn * up * A;
        }
        if (EQ == EQ_SST) {
            double rkL = L.r * t0L, rkR = b ? bv[5] : R.r * t0R;
            adv0 = Vn * ((Vn >= 0.0) ? rkL : rkR) * A;
            double rwL = L.r * t1L, rwR = b ? bv[6] : R.r * t1R;
            adv1 = Vn * ((Vn >= 0.0) ? rwL : rwR) * A;
        }
        if (!b || bc->type == BC_WALL) {
            double gLu0 = acc.g(0, 0), gLu1 = acc.g(0, 1), gLv0 = acc.g(0, 2), gLv1 = acc.g(0, 3);
            double gLT0 = acc.g(0, 4), gLT1 = acc.g(0, 5);
            double gRu0 = 0, gRu1 = 0, gRv0 = 0, gRv1 = 0, gRT0 = 0, gRT1 = 0;
            if (!b) {
                gRu0 = acc.g(1, 0); gRu1 = acc.g(1, 1); gRv0 = acc.g(1, 2); gRv1 = acc.g(1, 3);
                gRT0 = acc.g(1, 4); gRT1 = acc.g(1, 5);
            }
            FG fu = face_grad(fgeo, b, uL, b ? bv[0] : uR, gLu0, gLu1, gRu0, gRu1);
            FG fv = face_grad(fgeo, b, vL, b ? bv[1] : vR, gLv0, gLv1, gRv0, gRv1);
            FG fT = face_grad(fgeo, b, TL, b ? bv[2] : TR, gLT0, gLT1, gRT0, gRT1);
            double gux, guy, gvx, gvy, gTx, gTy;
            corrected(fu, nx, ny, gux, guy);
            corrected(fv, nx, ny, gvx, gvy);
            corrected(fT, nx, ny, gTx, gTy);
            double cp = g * a.Rgas / (g - 1.0);
            double mu_eff = a.mu, k_eff = a.mu * cp / a.pr;
            double tdvf = 0.0, nulL = 0.0, nulR = 0.0, ntL = 0.0, ntR = 0.0;
            if (EQ == EQ_SA) {
                nulL = fdiv(a.mu, L.r);
                ntL = sa_eddy(t0L, nulL, a.sa);
                nulR = b ? nulL : fdiv(a.mu, R.r);
                ntR = b ? sa_eddy(bv[3], nulR, a.sa) : sa_eddy(t0R, nulR, a.sa);
                double rR = b ? L.r : R.r;
                tdvf = 0.5 * (L.r * ntL + rR * ntR);
                mu_eff = a.mu + tdvf;
                k_eff = cp * a.mu / a.pr + cp * tdvf / a.prt;
            }
            if (EQ == EQ_SST) {
                ntL = acc.aux(0, 3);
                ntR = b ? 0.0 : acc.aux(1, 3);
                double rR = b ? L.r : R.r;
                tdvf = 0.5 * (L.r * ntL + rR * ntR);
                mu_eff = a.mu + tdvf;
                k_eff = cp * a.mu / a.pr + cp * tdvf / a.prt;
            }
            double uf = b ? bv[0] : 0.5 * (uL + uR);
            double vf = b ? bv[1] : 0.5 * (vL + vR);
            double div = gux + gvy;
            double txx = mu_eff * (2.0 * gux - (2.0 / 3.0) * div);
            double tyy = mu_eff * (2.0 * gvy - (2.0 / 3.0) * div);
            double txy = mu_eff * (guy + gvx);
            double qx = -k_eff * gTx, qy = -k_eff * gTy;
            if (b && !bc->isothermal) { qx = 0.0; qy = 0.0; }
            vx = txx * nx + txy * ny;
            vy = txy * nx + tyy * ny;
            vE = (uf * txx + vf * txy - qx) * nx + (uf * txy + vf * tyy - qy) * ny;
            if (EQ == EQ_SA) {
                double nlf = 0.5 * (nulL + nulR);
                double ntf = 0.5 * (ntL + ntR);
                FG fn = face_grad(fgeo, b, t0L, b ? bv[3] : t0R, acc.g(0, 6), acc.g(0, 7),
                                  b ? 0.0 : acc.g(1, 6), b ? 0.0 : acc.g(1, 7));
                double dif = (nlf + ntf) / a.sa.sigma;
                dif0 = -dif * fn.dn * A;
            }
            if (EQ == EQ_SST) {
                double F1L = acc.aux(0, 2);
                double F1R = b ? F1L : acc.aux(1, 2);
                double F1f = b ? F1L : 0.5 * (F1L + F1R);
                double sk = F1f * a.sst.sigma_k1 + (1.0 - F1f) * a.sst.sigma_k2;
                double sw = F1f * a.sst.sigma_omega1 + (1.0 - F1f) * a.sst.sigma_omega2;
                double mek = a.mu + sk * tdvf, mew = a.mu + sw * tdvf;
                FG fk = face_grad(fgeo, b, t0L, b ? bv[3] : t0R, acc.g(0, 6), acc.g(0, 7),
                                  b ? 0.0 : acc.g(1, 6), b ? 0.0 : acc.g(1, 7));
                FG fw = face_grad(fgeo, b, t1L, b ? bv[4] : t1R, acc.g(0, 8), acc.g(0, 9),
                                  b ? 0.0 : acc.g(1, 8), b ? 0.0 : acc.g(1, 9));
                dif0 = -mek * fk.dn * A;
                dif1 = -mew * fw.dn * A;
            }
        }
        out[0] = A * (inv.r - 0.0);
        out[1] = A * (inv.ru - vx);
        out[2] = A * (inv.rv - vy);
        out[3] = A * (inv.E - vE);
        if (EQ == EQ_SA) out[4] = adv0 + dif0;
        if (EQ == EQ_SST) { out[4] = adv0 + dif0; out[5] = adv1 + dif1; }
        if (a.face_dt) cand = visc_dt_candidate<EQ>(a, b, L, R, uL, vL, uR, vR, t0L, t1L, t0R, t1R, nx, ny, fgeo.dc);
    }
    return cand;
}

// global dt beside the fused tile pass (its first stage): the faces'
// compute_dt candidates as the face pass forms them.  Each workgroup takes a
// contiguous range of faces and leaves one atomic minimum; the grid is
// bounded, since device-wide atomics on one address are serialised across
// the XCDs.
constexpr int kVDtGridMax = 4 * 256;
template <int EQ>
__global__ __launch_bounds__(kVB) void k_visc_dt(ViscArgs a) {
    const MeshView& m = a.m;
    const int64_t ld = a.ld;
    const int64_t per = ((int64_t)m.nf + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)logical_block() * per;
    const int64_t hi = lo + per < m.nf ? lo + per : m.nf;
    double cand = DBL_MAX;
    for (int64_t f = lo + threadIdx.x; f < hi; f += kVB) {
        const int cl = m.fl[f], cr = m.fr[f];
        const bool b = cr < 0;
        const State L = ldS(a.U, ld, cl);
        const State R = b ? L : ldS(a.U, ld, cr);
        const double uL = fdiv(L.ru, L.r), vL = fdiv(L.rv, L.r);
        double uR = 0.0, vR = 0.0;
        if (!b) {
            uR = fdiv(R.ru, R.r);
            vR = fdiv(R.rv, R.r);
        }
        double t0L = 0.0, t1L = 0.0, t0R = 0.0, t1R = 0.0;
        if (EQ == EQ_SA || EQ == EQ_SST) {
            t0L = a.U[4 * ld + cl];
            if (!b) t0R = a.U[4 * ld + cr];
        }
        if (EQ == EQ_SST) {
            t1L = a.U[5 * ld + cl];
            if (!b) t1R = a.U[5 * ld + cr];
        }
        cand = smin(cand, visc_dt_candidate<EQ>(a, b, L, R, uL, vL, uR, vR, t0L, t1L, t0R, t1R, m.fnx[f], m.fny[f],
                                                m.fdc[f]));
    }
    __shared__ double sm[kVB];
    sm[threadIdx.x] = cand;
    __syncthreads();
    for (int s2 = kVB / 2; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2) sm[threadIdx.x] = smin(sm[threadIdx.x], sm[threadIdx.x + s2]);
        __syncthreads();
    }
    if (threadIdx.x == 0) dt_atomic_min(a.dtmin, a.slot, sm[0]);
}

// global field access for the face sweep (NG gradients per cell, cell-major)
template <int NG>
struct GlobalSides {
    const double *U, *G, *X;
    int64_t ld;
    int c[2];
    __device__ double u(int s, int q) const { return U[q * ld + c[s]]; }
    __device__ double g(int s, int q) const {
        const double2 v = reinterpret_cast<const double2*>(G + (int64_t)c[s] * NG)[q >> 1];
        return (q & 1) ? v.y : v.x;
    }
    __device__ double aux(int s, int q) const { return X[q * ld + c[s]]; }
};

template <int EQ, int SCHEME>
__global__ __launch_bounds__(kVB) void k_visc_face(ViscArgs a) {
    constexpr int NV = EQ == EQ_SST ? 6 : (EQ == EQ_SA ? 5 : 4);
    const MeshView& m = a.m;
    int f = xcd_index();
    double cand = DBL_MAX;
    if (f < m.nf) {
        const int cl = m.fl[f], cr = m.fr[f];
        const GlobalSides<2 * (NV - 1)> acc{a.U, a.G, a.aux, a.ld, {cl, cr < 0 ? cl : cr}};
        double out[6];
        cand = visc_face_body<EQ, SCHEME>(a, f, cr, face_geo(m, f), acc, out);
#pragma unroll
        for (int q = 0; q < NV; ++q) a.flux[q * a.ldf + f] = out[q];
        if (a.face_dt && a.dt_src == 2) a.dtf[f] = cand;
    }
    if (a.face_dt && a.dt_src != 2) {  // local dt: the cell pass reduces the step's dt from its cells
        __shared__ double sm[kVB];
        sm[threadIdx.x] = cand;
        __syncthreads();
        for (int s = kVB / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) sm[threadIdx.x] = smin(sm[threadIdx.x], sm[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) dt_atomic_min(a.dtmin, a.slot, sm[0]);
    }
}

// deterministic block sum of the squared residual components into
// partials[q * gridDim.x + blockIdx.x]; a one-block finalize kernel reduces
// them (fixed order) once the kernel has ended.  with_dt: the block minimum of
// dt goes to row NV (the step's dt, reduced by the finalize instead of by
// device-wide atomics)
template <int NV>
__device__ __forceinline__ void vblock_partials(const double (&v)[NV], double* partials, bool with_dt = false,
                                                double dt = DBL_MAX) {
    __shared__ double sm[NV + 1][kVB];
#pragma unroll
    for (int q = 0; q < NV; ++q) sm[q][threadIdx.x] = v[q];
    sm[NV][threadIdx.x] = dt;
    __syncthreads();
    for (int s = kVB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int q = 0; q < NV; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + s];
            if (with_dt) sm[NV][threadIdx.x] = smin(sm[NV][threadIdx.x], sm[NV][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[q * gridDim.x + blockIdx.x] = sm[q][0];
        if (with_dt) partials[NV * gridDim.x + blockIdx.x] = sm[NV][0];
    }
}

// ---------------------------------------------------------------------------
// cell pass: gather (Cell::faces order), sources (SA :279, SST :323-325),
// update (SST: rho k, rho omega divided by the new density, :335-353)
// ---------------------------------------------------------------------------
// sources, then the stage update (or the residual output) of cell c from
// its gathered face sum r and, for local dt, the minimum of its faces' dt
// the update's inputs of cell c, loaded apart from the arithmetic so that a
// kernel can issue them early: volume, source terms (AUX: from the gradient
// pass's output; else the caller sets them), dt (unless it is the cell's own
// new local dt) and the RK base state
template <int NV>
struct ViscUpdateInputs {
    double V, s0, s1, dt;
    double u0[NV];
};
template <int EQ, int NV, bool AUX = true>
__device__ __forceinline__ ViscUpdateInputs<NV> load_visc_update_inputs(const ViscArgs& a, int c) {
    const int64_t ld = a.ld;
    ViscUpdateInputs<NV> x{};
    x.V = a.m.vol[c];
    if (AUX && EQ == EQ_SA) x.s0 = a.aux[c];
    if (AUX && EQ == EQ_SST) { x.s0 = a.aux[c]; x.s1 = a.aux[ld + c]; }
    if (!a.R) {
        if (a.dt_src == 1) x.dt = dt_read(a.dtmin, a.slot);
        else if (a.dt_src == 0) x.dt = a.dtc[c];
#pragma unroll
        for (int q = 0; q < NV; ++q) x.u0[q] = a.first ? 0.0 : a.U0[q * ld + c];
    }
    return x;
}

// sources and the stage update of cell c from its residual r and its state u
// (read before any write: the last RK stage writes the buffer that holds the
// base state, whose values x already holds)
template <int EQ, int NV>
__device__ __forceinline__ void visc_cell_finish_with(const ViscArgs& a, int c, const ViscUpdateInputs<NV>& x,
                                                      const double (&u)[NV], double (&r)[NV], double dmin) {
    const int64_t ld = a.ld;
    const double V = x.V;
    if (EQ == EQ_SA) r[4] += V * x.s0;
    if (EQ == EQ_SST) { r[4] += V * x.s0; r[5] += V * x.s1; }
    if (a.R) {
#pragma unroll
        for (int q = 0; q < NV; ++q) a.R[q * ld + c] = r[q];
        return;
    }
    double dt = x.dt;
    if (a.dt_src == 2) { dt = dmin; a.dtc[c] = dmin; }
    double h = fdiv(dt, V);
    double rho_old = u[0];
    double rho_new = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double un = a.first ? u[q] + h * r[q] : a.a * x.u0[q] + a.b * (u[q] + h * r[q]);
        if (q == 0) rho_new = un;
        a.Unew[q * ld + c] = un;
    }
#pragma unroll
    for (int q = 4; q < NV; ++q) {
        double t = u[q];
        if (EQ == EQ_SST) {
            double cur = rho_old * t;
            double nw = a.first ? cur + h * r[q] : a.a * (x.u0[0] * x.u0[q]) + a.b * (cur + h * r[q]);
            a.Unew[q * ld + c] = fdiv(nw, rho_new);
        } else {
            a.Unew[q * ld + c] = a.first ? t + h * r[q] : a.a * x.u0[q] + a.b * (t + h * r[q]);
        }
    }
}

template <int EQ, int NV>
__device__ __forceinline__ void visc_cell_finish(const ViscArgs& a, int c, double (&r)[NV], double dmin) {
    const ViscUpdateInputs<NV> x = load_visc_update_inputs<EQ, NV>(a, c);
    double u[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) u[q] = a.U[q * a.ld + c];
    visc_cell_finish_with<EQ, NV>(a, c, x, u, r, dmin);
}

template <int EQ>
__global__ __launch_bounds__(kVB) void k_visc_cell(ViscArgs a) {
    constexpr int NV = EQ == EQ_SST ? 6 : (EQ == EQ_SA ? 5 : 4);
    const MeshView& m = a.m;
    int c = xcd_index();
    double r[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) r[q] = 0.0;
    const int64_t ldf = a.ldf;
    double dmin = DBL_MAX;
    if (c < m.nc) {
        for (int k = 0; k < m.deg; ++k) {  // Cell::faces order, from the ELL (coalesced across cells)
            const int e = m.ell[k * m.ld + c].x;
            if (e < 0) break;  // ELL padding (after the cell's faces)
            int f = e >> 1;
            if (e & 1) {
#pragma unroll
                for (int q = 0; q < NV; ++q) r[q] = r[q] - a.flux[q * ldf + f];
            } else {
#pragma unroll
                for (int q = 0; q < NV; ++q) r[q] += a.flux[q * ldf + f];
            }
            if (a.dt_src == 2) dmin = smin(dmin, a.dtf[f]);
        }
        visc_cell_finish<EQ, NV>(a, c, r, dmin);
    }
    if (!a.want_res) return;
    double sq[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) sq[q] = r[q] * r[q];
    vblock_partials<NV>(sq, a.partials, a.face_dt && a.dt_src == 2, dmin);
}

// ---------------------------------------------------------------------------
// fused gradient + face + cell pass over the solver's cell tiles (kernel
// variant 3, local or no dt).  A tile stages the states of its owned cells
// and two halo rings; it computes the gradients (and the turbulence sources,
// SST's F1 and nu_t) of its owned and ring-1 cells into LDS, the fluxes of
// its faces into LDS, then each owned cell gathers them in Cell::faces order
// and is updated.  Same operations as k_visc_grad + k_visc_face +
// k_visc_cell, so the same bits; a face on a tile's edge is evaluated by both
// tiles, a ring-1 cell's gradient by every tile it borders.
// ---------------------------------------------------------------------------
struct TileSides {
    const double *U, *G, *X;
    int MG;
    int c[2];
    __device__ double u(int s, int q) const { return U[q * MG + c[s]]; }
    __device__ double g(int s, int q) const { return G[q * MG + c[s]]; }
    __device__ double aux(int s, int q) const { return X[(q - 2) * MG + c[s]]; }  // SST: F1 (2), nu_t (3)
};

// gradient and sources of local cell j of a viscous tile, from the
// primitives P ([NF][ML]) and states U ([NV][MG]) in LDS: en and geo hold
// its D slots' packed ELL entries and gradient geometry (-1: no slot), iv
// its per-cell gradient data, wd its wall distance.  Writes G (and SST's F1,
// nu_t to X); an owned cell's sources go to s0, s1.
template <int EQ, int D, int GS>
__device__ __forceinline__ void visc_tile_grad(const ViscArgs& a, const double* U, const double* P, double* G,
                                               double* X, int MG, int ML, int f0, int j, bool owned,
                                               const int (&en)[D], const double4 (&geo)[D], const double4& iv,
                                               double wd, double& s0, double& s1) {
    constexpr int NF = EQ == EQ_SST ? 5 : (EQ == EQ_SA ? 4 : 3);
    const MeshView& m = a.m;
    double pc[NF], sx[NF], sy[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        pc[q] = P[q * ML + j];
        sx[q] = 0.0;
        sy[q] = 0.0;
    }
    const double t0c = NF > 3 ? pc[3 % NF] : 0.0, t1c = NF > 4 ? pc[4 % NF] : 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        if (en[k] == -1) continue;  // ELL padding (after the cell's faces)
        const bool bnd = tell_is_boundary(en[k]);
        double pn[NF];
        if (bnd) {
            // the face and its patch from the tile's list (owned cell) or the mesh (ring 1)
            const int x = tell_index(en[k]);
            int f, code;
            if (en[k] & 1) {
                f = x;
                code = m.fr[x];
            } else {
                const int4 fd = m.tface[f0 + x];
                f = fd.x;
                code = fd.z;
            }
            const State Uc{U[j], U[MG + j], U[2 * MG + j], U[3 * MG + j]};
            double bv[7];
            boundary_values<EQ>(a, m.bc[-1 - code], f, Uc, pc[0], pc[1], pc[2], t0c, t1c, bv);
#pragma unroll
            for (int q = 0; q < NF; ++q) pn[q] = bv[q];
        } else {
            const int l = tell_nb(en[k]);
#pragma unroll
            for (int q = 0; q < NF; ++q) pn[q] = P[q * ML + l];
        }
        grad_slot<NF, GS>(sx, sy, pc, pn, bnd, geo[k]);
    }
    double gr[2 * NF];
    grad_finish<NF, GS>(sx, sy, iv, gr);
#pragma unroll
    for (int q = 0; q < 2 * NF; ++q) G[q * MG + j] = gr[q];
    if (EQ == EQ_SA && owned) {
        double o4[4];
        visc_sources<EQ, NF>(a, U[j], t0c, t1c, wd, gr, o4);
        s0 = o4[0];
    }
    if (EQ == EQ_SST) {
        double o4[4];
        visc_sources<EQ, NF>(a, U[j], t0c, t1c, wd, gr, o4);
        X[j] = o4[2];
        X[MG + j] = o4[3];
        if (owned) {
            s0 = o4[0];
            s1 = o4[1];
        }
    }
}

// The gradient geometry of a viscous tile's slot entry en (tgk: the slot's
// per-slot copy): least squares, or a slot whose face is not a tile face,
// reads the copy; a Green-Gauss slot with a tile face forms it from the face's
// record, {the cell's distance weight, the normal pointing away from the
// cell, the length}, the values the copy holds.  No slot (-1): zeros.
__device__ __forceinline__ double4 visc_slot_geo(const MeshView& m, bool lsq, int en, const double4* tgk, int f0) {
    if (en == -1) return double4{0.0, 0.0, 0.0, 0.0};
    const int lf = tell_lf(en);
    if (lsq || lf < 0) return *tgk;
    const double4* r = m.tvf + 2 * (int64_t)(f0 + lf);
    const double4 r0 = r[0];
    if (tell_owner(en)) return double4{r0.w, r0.x, r0.y, r0.z};
    return double4{r[1].x, -r0.x, -r0.y, r0.z};
}

constexpr int kVTileMax = 512;
// D: slots per cell the gradient holds in registers (at least the mesh width)
template <int EQ, int SCHEME, int D>
__global__ __launch_bounds__(kVTileMax) void k_visc_tile(ViscArgs a) {
    constexpr int NV = EQ == EQ_SST ? 6 : (EQ == EQ_SA ? 5 : 4);
    constexpr int NF = NV - 1, NG = 2 * NF, NX = EQ == EQ_SST ? 2 : 0;
    extern __shared__ double lds[];
    const MeshView& m = a.m;
    int t = logical_block();
    // a pending finalize (the previous stage's norms) takes the first logical
    // workgroup, dispatched first so that its one-workgroup reduction overlaps
    // the tiles; the tiles follow it
    if (a.fin_nblocks > 0) {
        if (t == 0) {
            finalize_norms(a.partials, a.fin_nblocks, a.fin_nv, a.res_hist, a.dt_hist, a.fin_hist_idx, a.dtmin,
                           a.fin_slot, -1.0, lds, true, a.fin_dt_row != 0);
            return;
        }
        --t;
    }
    const int bs = blockDim.x, tid = threadIdx.x;
    const int MG = m.tmax_g, ML = m.tmax_l, MF = m.tmax_f, W = m.twidth;
    const int64_t ld = a.ld;
    double* U = lds;            // [NV][MG] states of the owned and ring-1 cells
    double* G = U + NV * MG;    // [NG][MG] their gradients
    double* X = G + NG * MG;    // [NX][MG] SST: their F1 and nu_t
    double* P = X + NX * MG;    // [NF][ML] u, v, T (and the turbulence fields) of every staged cell
    double* F = P;              // [NV][MF] face fluxes and [MF] dt candidates, over P once the gradients are done
    double* Dt = F + NV * MF;
    int c0 = 0, n = 0, hp = 0, h1 = 0, nh = 0, f0 = 0, nft = 0;
    if (t < m.ntiles) {
        const int4 d0 = m.tdesc[2 * t], d1 = m.tdesc[2 * t + 1];
        c0 = d0.x;
        n = d0.y;
        hp = d0.z;
        h1 = d0.w;
        nh = d1.x;
        f0 = d1.y;
        nft = d1.z;
    }
    const int ng = n + h1;  // owned and ring-1 cells (with gradients); ring 2 follows up to n + nh
    const double g = a.gamma;
    const bool lsq = a.gscheme == FVK_GRAD_LEAST_SQUARES;
    const int* te = m.tell + (int64_t)t * W * MG;
    const double4* tg = m.tegeo + (int64_t)t * W * MG;
    const double4* tc = m.tclsq + (int64_t)t * MG;

    // local cell tid's device id, and its gradient inputs, issued before the
    // staging so that their latency overlaps it (local cells from bs on,
    // ring cells only, are read where they are used); a slot's geometry only
    // if the slot is a face, so the padding of cells with fewer faces than
    // slots is not fetched, and for Green-Gauss from its tile face's record
    // where it has one (the face pass reads the same records)
    const int gc0 = tid < n ? c0 + tid : (tid < n + nh ? m.thalo[hp + tid - n] : 0);
    int en0[D];
    double4 geo0[D], iv0{};
    double wd0 = 0.0;
    if (tid < ng) {
#pragma unroll
        for (int k = 0; k < D; ++k) en0[k] = k < W ? te[k * MG + tid] : -1;
#pragma unroll
        for (int k = 0; k < D; ++k) geo0[k] = visc_slot_geo(m, lsq, en0[k], tg + k * MG + tid, f0);
        iv0 = tc[tid];
        if (EQ != EQ_NS) wd0 = m.wdist[gc0];
    }

    // 1. states of the owned and ring-1 cells, primitives of every staged cell
    for (int j = tid; j < n + nh; j += bs) {
        const int gc = j == tid ? gc0 : m.thalo[hp + j - n];  // j > tid: j >= bs >= n, a ring cell
        double u[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) u[q] = a.U[q * ld + gc];
        if (j < ng) {
#pragma unroll
            for (int q = 0; q < NV; ++q) U[q * MG + j] = u[q];
        }
        const State sc{u[0], u[1], u[2], u[3]};
        P[j] = fdiv(sc.ru, sc.r);
        P[ML + j] = fdiv(sc.rv, sc.r);
        P[2 * ML + j] = temperature(sc, g, a.Rgas);
#pragma unroll
        for (int q = 3; q < NF; ++q) P[q * ML + j] = u[q + 1];
    }
    __syncthreads();

    // 2. gradients of the owned and ring-1 cells in Cell::faces order, and
    // their sources (thread tid's owned cell tid keeps its own in registers)
    double s0 = 0.0, s1 = 0.0;
    if (tid < ng) {
        if (lsq)
            visc_tile_grad<EQ, D, FVK_GRAD_LEAST_SQUARES>(a, U, P, G, X, MG, ML, f0, tid, tid < n, en0, geo0, iv0,
                                                          wd0, s0, s1);
        else
            visc_tile_grad<EQ, D, FVK_GRAD_GREEN_GAUSS>(a, U, P, G, X, MG, ML, f0, tid, tid < n, en0, geo0, iv0,
                                                        wd0, s0, s1);
    }
    for (int j = tid + bs; j < ng; j += bs) {
        int en[D];
        double4 geo[D];
#pragma unroll
        for (int k = 0; k < D; ++k) en[k] = k < W ? te[k * MG + j] : -1;
#pragma unroll
        for (int k = 0; k < D; ++k) geo[k] = visc_slot_geo(m, lsq, en[k], tg + k * MG + j, f0);
        const double wd = EQ == EQ_SST ? m.wdist[m.thalo[hp + j - n]] : 0.0;
        double d0, d1;
        if (lsq)
            visc_tile_grad<EQ, D, FVK_GRAD_LEAST_SQUARES>(a, U, P, G, X, MG, ML, f0, j, false, en, geo, tc[j], wd,
                                                          d0, d1);
        else
            visc_tile_grad<EQ, D, FVK_GRAD_GREEN_GAUSS>(a, U, P, G, X, MG, ML, f0, j, false, en, geo, tc[j], wd,
                                                        d0, d1);
    }
    __syncthreads();

    // the update's inputs of owned cell tid, loaded while the faces are computed
    ViscUpdateInputs<NV> upd{};
    if (tid < n) upd = load_visc_update_inputs<EQ, NV, false>(a, c0 + tid);
    upd.s0 = s0;
    upd.s1 = s1;

    // 3. fluxes (and dt candidates) of the tile's faces
    double cand_min = DBL_MAX;
    for (int e = tid; e < nft; e += bs) {
        const int4 fd = m.tface[f0 + e];
        const double4* r = m.tvf + 2 * (int64_t)(f0 + e);
        const TileSides acc{U, G, X, MG, {fd.y, fd.z < 0 ? fd.y : fd.z}};
        double out[6];
        const double cand = visc_face_body<EQ, SCHEME>(a, fd.x, fd.z, face_geo(r[0], r[1]), acc, out);
#pragma unroll
        for (int q = 0; q < NV; ++q) F[q * MF + e] = out[q];
        if (a.face_dt) {
            Dt[e] = cand;
            cand_min = smin(cand_min, cand);
        }
    }
    __syncthreads();

    // 4. gather in Cell::faces order (owner minus), sources and update of the owned cells
    double r[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) r[q] = 0.0;
    if (tid < n) {
        const int* te = m.tell + (int64_t)t * W * MG;
        double dmin = DBL_MAX;
        for (int k = 0; k < W; ++k) {
            const int en = te[k * MG + tid];
            if (en == -1) continue;  // ELL padding (after the cell's faces)
            const int lf = tell_lf(en);
            if (tell_owner(en)) {
#pragma unroll
                for (int q = 0; q < NV; ++q) r[q] = r[q] - F[q * MF + lf];
            } else {
#pragma unroll
                for (int q = 0; q < NV; ++q) r[q] += F[q * MF + lf];
            }
            if (a.dt_src == 2) dmin = smin(dmin, Dt[lf]);
        }
        double u[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) u[q] = U[q * MG + tid];
        visc_cell_finish_with<EQ, NV>(a, c0 + tid, upd, u, r, dmin);
    }

    // 5. block reductions in the (now free) LDS: residual partials, dt minimum
    int p2 = 64;  // the largest power of two not above the workgroup size
    while (2 * p2 <= bs) p2 *= 2;
    if (a.want_res) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NV; ++q) lds[q * bs + tid] = r[q] * r[q];
        __syncthreads();
        if (tid < bs - p2) {
#pragma unroll
            for (int q = 0; q < NV; ++q) lds[q * bs + tid] += lds[q * bs + tid + p2];
        }
        __syncthreads();
        for (int s2 = p2 / 2; s2 > 0; s2 >>= 1) {
            if (tid < s2) {
#pragma unroll
                for (int q = 0; q < NV; ++q) lds[q * bs + tid] += lds[q * bs + tid + s2];
            }
            __syncthreads();
        }
        if (tid == 0) {
#pragma unroll
            for (int q = 0; q < NV; ++q) a.partials[q * gridDim.x + blockIdx.x] = lds[q * bs];
        }
    }
    if (a.face_dt) {
        __syncthreads();
        lds[tid] = cand_min;
        __syncthreads();
        if (tid < bs - p2) lds[tid] = smin(lds[tid], lds[tid + p2]);
        __syncthreads();
        for (int s2 = p2 / 2; s2 > 0; s2 >>= 1) {
            if (tid < s2) lds[tid] = smin(lds[tid], lds[tid + s2]);
            __syncthreads();
        }
        // the step's dt (local dt: only its history), reduced by the finalize
        // from one entry per workgroup instead of by device-wide atomics
        if (tid == 0) a.partials[NV * gridDim.x + blockIdx.x] = lds[0];
    }
}

// ---------------------------------------------------------------------------
// scalar diffusion, UnstructuredFVMSolver::step (src/UnstructuredFVMSolver.cpp:44-92)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kVB) void k_diff_face(ViscArgs a) {
    const MeshView& m = a.m;
    int f = xcd_index();
    if (f >= m.nf) return;
    int cl = m.fl[f], cr = m.fr[f];
    double phiL = a.U[cl];
    double F;
    if (cr >= 0) {
        double dx = m.cx[cr] - m.cx[cl], dy = m.cy[cr] - m.cy[cl];
        double d = sqrt(dx * dx + dy * dy);
        double gn = (a.U[cr] - phiL) / d;
        F = -a.alpha * gn * m.fa[f];
    } else {
        const DevBC& bc = m.bc[-1 - cr];
        if (bc.type == 1) {
            F = bc.value * m.fa[f];
        } else {
            double dx = m.fmx[f] - m.cx[cl], dy = m.fmy[f] - m.cy[cl];
            double d = sqrt(dx * dx + dy * dy);
            double gn = (bc.value - phiL) / d;
            F = -a.alpha * gn * m.fa[f];
        }
    }
    a.flux[f] = F;
}

__global__ __launch_bounds__(kVB) void k_diff_cell(ViscArgs a) {
    const MeshView& m = a.m;
    int c = xcd_index();
    double r[1] = {0.0};
    if (c < m.nc) {
        for (int k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
            int e = m.cent[k];
            if (e & 1) r[0] = r[0] - a.flux[e >> 1];
            else r[0] += a.flux[e >> 1];
        }
        if (a.R) a.R[c] = r[0];
        else a.Unew[c] = a.U[c] + (fdiv(a.fixed_dt, m.vol[c])) * r[0];
    }
    if (!a.want_res) return;
    double sq[1] = {r[0] * r[0]};
    vblock_partials<1>(sq, a.partials);
}

// ---------------------------------------------------------------------------
// advection-diffusion (AdvectionDiffusionFVMSolver::step,
// src/AdvectionDiffusionFVMSolver.cpp:42-104): boundary values, the cell
// gradients of phi (k_scalar_grad), the face fluxes, then the diffusion
// solver's cell pass (k_diff_cell)
// ---------------------------------------------------------------------------

// build_boundary_phi (src/AdvectionDiffusionFVMSolver.cpp:33-40): the
// Dirichlet value, else the cell's own phi; a.aux holds [ldf] face values
__global__ __launch_bounds__(kVB) void k_advdiff_bphi(ViscArgs a) {
    const MeshView& m = a.m;
    const int f = m.nfi + xcd_index();
    if (f >= m.nf) return;
    const DevBC& bc = m.bc[-1 - m.fr[f]];
    a.aux[f] = bc.type == 0 ? bc.value : a.U[m.fl[f]];
}

// the face pass (:50-72) and boundary_flux (:117-138): upwind advection
// with the uniform velocity plus Jasak-corrected diffusion; a.G holds the
// cell gradients [2][nc] (x then y, per cell pairs)
__global__ __launch_bounds__(kVB) void k_advdiff_face(ViscArgs a, double u_adv, double v_adv) {
    const MeshView& m = a.m;
    const int f = xcd_index();
    if (f >= m.nf) return;
    const int cl = m.fl[f], cr = m.fr[f];
    const double phiL = a.U[cl];
    const double A = m.fa[f];
    const double Vn = u_adv * m.fnx[f] + v_adv * m.fny[f];
    const double* g = a.G;
    double F;
    if (cr >= 0) {
        const double phiR = a.U[cr];
        const FG fg = face_grad(m, f, false, phiL, phiR, g[2 * cl], g[2 * cl + 1], g[2 * cr], g[2 * cr + 1]);
        const double diffusive = -a.alpha * fg.dn * A;
        const double up = (Vn >= 0.0) ? phiL : phiR;
        F = Vn * up * A + diffusive;
    } else {
        const DevBC& bc = m.bc[-1 - cr];
        if (bc.type == 1) {
            F = Vn * phiL * A + bc.value * A;
        } else {
            const FG fg = face_grad(m, f, true, phiL, bc.value, g[2 * cl], g[2 * cl + 1], 0.0, 0.0);
            const double diffusive = -a.alpha * fg.dn * A;
            const double up = (Vn >= 0.0) ? phiL : bc.value;
            F = Vn * up * A + diffusive;
        }
    }
    a.flux[f] = F;
}

// ---------------------------------------------------------------------------
// standalone gradient (fvk_gradient): one scalar field with caller-given
// boundary values (device face order)
// ---------------------------------------------------------------------------
template <int GS>
__global__ __launch_bounds__(kVB) void k_scalar_grad(MeshView m, const double* phi, const double* bphi, double* grad,
                                                     int periodic) {
    int c = xcd_index();
    if (c >= m.nc) return;
    double pc = phi[c], cx = m.cx[c], cy = m.cy[c];
    double sx = 0.0, sy = 0.0;
    for (int k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
        int e = m.cent[k];
        int f = e >> 1;
        bool owner = e & 1;
        int cr = m.fr[f];
        double pn, qx, qy;
        if (cr < 0) { pn = bphi[f]; qx = m.fmx[f]; qy = m.fmy[f]; }
        else {
            int o = owner ? cr : m.fl[f];
            pn = phi[o];
            qx = m.cx[o];
            qy = m.cy[o];
            if (periodic) {
                if (owner) { qx = qx + m.fsx[f]; qy = qy + m.fsy[f]; }
                else { qx = qx - m.fsx[f]; qy = qy - m.fsy[f]; }
            }
        }
        if (GS == FVK_GRAD_LEAST_SQUARES) {
            double dx = qx - cx, dy = qy - cy;
            double w = 1.0 / (dx * dx + dy * dy);
            double dphi = pn - pc;
            sx += w * dx * dphi;
            sy += w * dy * dphi;
        } else {
            double wc = owner ? m.fwl[f] : m.fwr[f];
            double nx = owner ? m.fnx[f] : -m.fnx[f];
            double ny = owner ? m.fny[f] : -m.fny[f];
            double pf = (cr < 0) ? pn : wc * pc + (1.0 - wc) * pn;
            sx += pf * nx * m.fa[f];
            sy += pf * ny * m.fa[f];
        }
    }
    if (GS == FVK_GRAD_LEAST_SQUARES) {
        double i0 = m.lsq[c], i1 = m.lsq[m.nc + c], i2 = m.lsq[2 * m.nc + c];
        grad[2 * c] = i0 * sx + i1 * sy;
        grad[2 * c + 1] = i1 * sx + i2 * sy;
    } else {
        grad[2 * c] = fdiv(sx, m.vol[c]);
        grad[2 * c + 1] = fdiv(sy, m.vol[c]);
    }
}

// compute_wall_distance (src/WallDistance.cpp:13-43): brute force over the
// wall segments, staged through LDS in tiles shared by the block's cells
__global__ __launch_bounds__(kVB) void k_wall_distance(const double* cx, const double* cy, int nc,
                                                       const double* seg /* [4][nw] x1 y1 x2 y2 */, int nw,
                                                       double* out) {
    __shared__ double s[4][kVB];
    int c = xcd_index();
    double px = c < nc ? cx[c] : 0.0, py = c < nc ? cy[c] : 0.0;
    double best = INFINITY;
    for (int base = 0; base < nw; base += kVB) {
        int j = base + threadIdx.x;
        if (j < nw) {
#pragma unroll
            for (int q = 0; q < 4; ++q) s[q][threadIdx.x] = seg[q * nw + j];
        }
        __syncthreads();
        int n = min(kVB, nw - base);
        for (int t = 0; t < n; ++t) {
            double x1 = s[0][t], y1 = s[1][t], x2 = s[2][t], y2 = s[3][t];
            double dx = x2 - x1, dy = y2 - y1;
            double l2 = dx * dx + dy * dy;
            double tt = (l2 > 0.0) ? ((px - x1) * dx + (py - y1) * dy) / l2 : 0.0;
            tt = (tt < 0.0) ? 0.0 : ((1.0 < tt) ? 1.0 : tt);
            double ex = px - (x1 + tt * dx), ey = py - (y1 + tt * dy);
            best = smin(best, ex * ex + ey * ey);
        }
        __syncthreads();
    }
    if (c < nc) out[c] = sqrt(best);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {

ViscArgs make_vargs(fvk_solver* s) {
    ViscArgs a{};
    a.m = s->mesh;
    a.G = s->grad;
    a.aux = s->aux;
    a.flux = s->flux;
    a.dtf = s->dtf;
    a.dtc = s->dtc;
    a.dtmin = s->dtmin;
    a.partials = s->partials;
    a.res_hist = s->res_hist;
    a.dt_hist = s->dt_hist;
    a.ld = s->ld;
    a.ldf = s->ldf;
    a.gscheme = s->p.gradient_scheme;
    a.maxit = s->p.exact_riemann_max_iter;
    a.sst_variant = s->p.sst_variant;
    a.kato = s->p.kato_launder;
    a.gamma = s->p.gamma;
    a.Rgas = s->p.gas_constant;
    a.mu = s->p.mu;
    a.pr = s->p.prandtl;
    a.prt = s->p.prandtl_t;
    a.cfl = s->p.cfl;
    a.tol = s->p.exact_riemann_tol;
    a.alpha = s->p.alpha;
    a.fixed_dt = s->p.fixed_dt;
    a.sa = s->p.sa;
    a.sst = s->p.sst;
    return a;
}

template <int EQ>
void launch_grad_eq(fvk_solver* s, int grid, const ViscArgs& a) {
    bool ls = s->p.gradient_scheme == FVK_GRAD_LEAST_SQUARES;
    if (s->periodic) {
        if (ls) k_visc_grad<EQ, FVK_GRAD_LEAST_SQUARES, true><<<grid, kVB, 0, s->stream>>>(a);
        else k_visc_grad<EQ, FVK_GRAD_GREEN_GAUSS, true><<<grid, kVB, 0, s->stream>>>(a);
    } else {
        if (ls) k_visc_grad<EQ, FVK_GRAD_LEAST_SQUARES, false><<<grid, kVB, 0, s->stream>>>(a);
        else k_visc_grad<EQ, FVK_GRAD_GREEN_GAUSS, false><<<grid, kVB, 0, s->stream>>>(a);
    }
}

template <int EQ, int D>
void launch_tile_d(fvk_solver* s, int grid, const ViscArgs& a) {
    const size_t lds = s->tile_lds;
    auto go = [&](auto kernel) {
        allow_lds(reinterpret_cast<const void*>(kernel), lds);
        kernel<<<grid, s->block, lds, s->stream>>>(a);
    };
    if (s->p.flux_scheme == FLUX_HLLC) go(k_visc_tile<EQ, FLUX_HLLC, D>);
    else if (s->p.flux_scheme == FLUX_EXACT) go(k_visc_tile<EQ, FLUX_EXACT, D>);
    else go(k_visc_tile<EQ, FLUX_RUSANOV, D>);
}
template <int EQ>
void launch_tile_eq(fvk_solver* s, int grid, const ViscArgs& a) {
    const int w = s->mesh.twidth;
    if (w <= 4) launch_tile_d<EQ, 4>(s, grid, a);
    else if (w <= 9) launch_tile_d<EQ, 9>(s, grid, a);
    else launch_tile_d<EQ, kMaxEulerDegree>(s, grid, a);
}

template <int EQ>
void launch_face_eq(fvk_solver* s, int grid, const ViscArgs& a) {
    if (s->p.flux_scheme == FLUX_HLLC) k_visc_face<EQ, FLUX_HLLC><<<grid, kVB, 0, s->stream>>>(a);
    else if (s->p.flux_scheme == FLUX_EXACT) k_visc_face<EQ, FLUX_EXACT><<<grid, kVB, 0, s->stream>>>(a);
    else k_visc_face<EQ, FLUX_RUSANOV><<<grid, kVB, 0, s->stream>>>(a);
}

void run_visc(fvk_solver* s, ViscArgs& a) {
    const int gc = grid_for(s->nc, kVB);
    const int gf = grid_for(s->nf, kVB);
    hipEvent_t ev = nullptr;
    int nparts = gc;
    if (s->variant == 3) {
        // the fused tile pass updates its cells in the launch that evaluates
        // their faces: a global dt is formed first, by its own pass over the
        // faces (first stage; later stages read the step's slot)
        ViscArgs at = a;
        if (a.dt_src == 1 && a.face_dt && !a.R) {
            const int gd = std::min(gf, kVDtGridMax);
            timer_begin(s, "visc_dt", &ev);
            switch (s->p.equation) {
                case FVK_EQ_NAVIER_STOKES: k_visc_dt<EQ_NS><<<gd, kVB, 0, s->stream>>>(a); break;
                case FVK_EQ_RANS_SA: k_visc_dt<EQ_SA><<<gd, kVB, 0, s->stream>>>(a); break;
                default: k_visc_dt<EQ_SST><<<gd, kVB, 0, s->stream>>>(a); break;
            }
            timer_end(s, ev);
            at.face_dt = 0;
        }
        // a pending finalize takes the first logical workgroup
        const int gt = (s->mesh.ntiles + (a.fin_nblocks > 0 ? 1 : 0) + kXcds - 1) / kXcds * kXcds;
        timer_begin(s, "visc_tile", &ev);
        switch (s->p.equation) {
            case FVK_EQ_NAVIER_STOKES: launch_tile_eq<EQ_NS>(s, gt, at); break;
            case FVK_EQ_RANS_SA: launch_tile_eq<EQ_SA>(s, gt, at); break;
            default: launch_tile_eq<EQ_SST>(s, gt, at); break;
        }
        timer_end(s, ev);
        nparts = gt;
    } else {
        // a pending finalize takes the first logical workgroup of one more XCD round
        const int gg = gc + (a.fin_nblocks > 0 ? kXcds : 0);
        timer_begin(s, "visc_grad", &ev);
        switch (s->p.equation) {
            case FVK_EQ_NAVIER_STOKES: launch_grad_eq<EQ_NS>(s, gg, a); break;
            case FVK_EQ_RANS_SA: launch_grad_eq<EQ_SA>(s, gg, a); break;
            default: launch_grad_eq<EQ_SST>(s, gg, a); break;
        }
        timer_end(s, ev);
        timer_begin(s, "visc_face", &ev);
        switch (s->p.equation) {
            case FVK_EQ_NAVIER_STOKES: launch_face_eq<EQ_NS>(s, gf, a); break;
            case FVK_EQ_RANS_SA: launch_face_eq<EQ_SA>(s, gf, a); break;
            default: launch_face_eq<EQ_SST>(s, gf, a); break;
        }
        timer_end(s, ev);
        timer_begin(s, "visc_cell", &ev);
        switch (s->p.equation) {
            case FVK_EQ_NAVIER_STOKES: k_visc_cell<EQ_NS><<<gc, kVB, 0, s->stream>>>(a); break;
            case FVK_EQ_RANS_SA: k_visc_cell<EQ_SA><<<gc, kVB, 0, s->stream>>>(a); break;
            default: k_visc_cell<EQ_SST><<<gc, kVB, 0, s->stream>>>(a); break;
        }
        timer_end(s, ev);
    }
    if (a.want_res) {
        if (s->p.rk_stages >= 2 && !a.R) {
            // the next stage's first launch reduces this stage's partials
            s->fin_nblocks = nparts;
            s->fin_slot = a.slot;
            s->fin_hist_idx = a.hist_idx;
            s->fin_rearm = 1;
            s->fin_dt_row = a.face_dt && a.dt_src == 2;
            return;
        }
        timer_begin(s, "visc_finalize", &ev);
        launch_norm_finalize(s, nparts, s->nv, a.hist_idx, a.slot, -1.0, a.face_dt && a.dt_src == 2);
        timer_end(s, ev);
    }
}

}  // namespace

int launch_visc_stage(fvk_solver* s, int stage, int stages, int want_res, int64_t step) {
    static const double RA[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0.5, 0}, {0, 0.75, 1.0 / 3.0}};
    static const double RB[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0.5, 0}, {1, 0.25, 2.0 / 3.0}};
    ViscArgs a = make_vargs(s);
    const int A = s->cur;
    int in = stage == 0 ? A : (A + stage) % 3;
    int out = (stage == stages - 1) ? A : (A + stage + 1) % 3;
    a.U = s->U[in];
    a.Unew = s->U[out];
    a.U0 = s->U[A];
    a.face_dt = stage == 0;
    a.dt_src = s->p.local_dt ? (stage == 0 ? 2 : 0) : 1;
    a.first = stage == 0;
    a.a = RA[stages][stage];
    a.b = RB[stages][stage];
    a.want_res = want_res && stage == 0;
    a.slot = (int)(step & 1);
    a.hist_idx = step % s->hist_cap;
    if (s->fin_nblocks > 0) {  // the previous stage's norms, folded into this stage's first launch
        a.fin_nblocks = s->fin_nblocks;
        a.fin_slot = s->fin_slot;
        a.fin_hist_idx = s->fin_hist_idx;
        a.fin_nv = s->nv;
        a.fin_dt_row = s->fin_dt_row;
        s->fin_nblocks = 0;
    }
    run_visc(s, a);
    return hip_check(hipGetLastError(), "viscous stage launch");
}

int launch_visc_residual_eval(fvk_solver* s, double* R_dev) {
    ViscArgs a = make_vargs(s);
    a.U = s->U[s->cur];
    a.R = R_dev;
    run_visc(s, a);
    return hip_check(hipGetLastError(), "viscous residual launch");
}

int launch_diffusion_step(fvk_solver* s, int want_res, int64_t step) {
    ViscArgs a = make_vargs(s);
    a.U = s->U[s->cur];
    a.Unew = s->U[s->cur];
    a.want_res = want_res;
    a.hist_idx = step % s->hist_cap;
    const int gc = grid_for(s->nc, kVB);
    const int gf = grid_for(s->nf, kVB);
    hipEvent_t ev = nullptr;
    timer_begin(s, "diffusion_face", &ev);
    k_diff_face<<<gf, kVB, 0, s->stream>>>(a);
    timer_end(s, ev);
    timer_begin(s, "diffusion_cell", &ev);
    k_diff_cell<<<gc, kVB, 0, s->stream>>>(a);
    timer_end(s, ev);
    if (want_res) launch_norm_finalize(s, gc, 1, a.hist_idx, 0, s->p.fixed_dt);
    return hip_check(hipGetLastError(), "diffusion launch");
}

int launch_diffusion_residual_eval(fvk_solver* s, double* R_dev) {
    ViscArgs a = make_vargs(s);
    a.U = s->U[s->cur];
    a.R = R_dev;
    const int gc = grid_for(s->nc, kVB);
    const int gf = grid_for(s->nf, kVB);
    k_diff_face<<<gf, kVB, 0, s->stream>>>(a);
    k_diff_cell<<<gc, kVB, 0, s->stream>>>(a);
    return hip_check(hipGetLastError(), "diffusion residual launch");
}

namespace {
void advdiff_pass(fvk_solver* s, ViscArgs& a) {
    const int gc = grid_for(s->nc, kVB);
    const int gf = grid_for(s->nf, kVB);
    const int gb = grid_for(std::max<int64_t>(s->nf - s->nfi, 1), kVB);
    hipEvent_t ev = nullptr;
    timer_begin(s, "advdiff_grad", &ev);
    k_advdiff_bphi<<<gb, kVB, 0, s->stream>>>(a);
    if (s->p.gradient_scheme == FVK_GRAD_LEAST_SQUARES)
        k_scalar_grad<FVK_GRAD_LEAST_SQUARES><<<gc, kVB, 0, s->stream>>>(s->mesh, a.U, a.aux, a.G, s->periodic);
    else
        k_scalar_grad<FVK_GRAD_GREEN_GAUSS><<<gc, kVB, 0, s->stream>>>(s->mesh, a.U, a.aux, a.G, s->periodic);
    timer_end(s, ev);
    timer_begin(s, "advdiff_face", &ev);
    k_advdiff_face<<<gf, kVB, 0, s->stream>>>(a, s->p.advection_u, s->p.advection_v);
    timer_end(s, ev);
    timer_begin(s, "advdiff_cell", &ev);
    k_diff_cell<<<gc, kVB, 0, s->stream>>>(a);
    timer_end(s, ev);
}
}  // namespace

int launch_advdiff_step(fvk_solver* s, int want_res, int64_t step) {
    ViscArgs a = make_vargs(s);
    a.U = s->U[s->cur];
    a.Unew = s->U[s->cur];
    a.want_res = want_res;
    a.hist_idx = step % s->hist_cap;
    advdiff_pass(s, a);
    if (want_res) launch_norm_finalize(s, grid_for(s->nc, kVB), 1, a.hist_idx, 0, s->p.fixed_dt);
    return hip_check(hipGetLastError(), "advection-diffusion launch");
}

int launch_advdiff_residual_eval(fvk_solver* s, double* R_dev) {
    ViscArgs a = make_vargs(s);
    a.U = s->U[s->cur];
    a.R = R_dev;
    advdiff_pass(s, a);
    return hip_check(hipGetLastError(), "advection-diffusion residual launch");
}

int launch_gradient(fvk_solver* s, int scheme, const double* phi, const double* bphi_dev, double* grad) {
    const int gc = grid_for(s->nc, kVB);
    if (scheme == FVK_GRAD_LEAST_SQUARES)
        k_scalar_grad<FVK_GRAD_LEAST_SQUARES><<<gc, kVB, 0, s->stream>>>(s->mesh, phi, bphi_dev, grad, s->periodic);
    else
        k_scalar_grad<FVK_GRAD_GREEN_GAUSS><<<gc, kVB, 0, s->stream>>>(s->mesh, phi, bphi_dev, grad, s->periodic);
    return hip_check(hipGetLastError(), "gradient launch");
}

// ---------------------------------------------------------------------------
// wall diagnostics: the samples of compute_wall_traction
// (src/WallTraction.cpp:95-143) with the solvers' inputs
// (src/NavierStokesFVMSolver.cpp:333-352, src/RANSTurbulenceSASolver.cpp:353-374,
// src/RANSTurbulenceSSTSolver.cpp:435-464), one thread per wall face, then the
// per-patch sums of compute_wall_forces (src/WallTraction.cpp:163-216) in the
// caller's sample order by one workgroup
// ---------------------------------------------------------------------------
namespace {

// SST F2 and eddy viscosity with the reference's guards
// (src/SSTKOmega.cpp sst_F2, sst_eddy_viscosity)
__device__ double sst_F2_ref(double k, double w, double nu, double d, const fvk_sst_constants& c) {
    if (k <= 0.0 || w <= 0.0) return 0.0;
    const double d2 = d * d;
    const double arg2 = smax(2.0 * sqrt(k) / (c.beta_star * w * d), 500.0 * nu / (d2 * w));
    return tanh(arg2 * arg2);
}
__device__ double sst_eddy_ref(double k, double w, double S, double Om, double F2, int variant,
                               const fvk_sst_constants& c) {
    if (k <= 0.0) return 0.0;
    const double lim = (variant == FVK_SST_VORTICITY) ? Om : S;
    return c.a1 * k / smax(c.a1 * w, lim * F2);
}

constexpr int kWallBlock = 256;

template <int EQ>
__global__ __launch_bounds__(kVB) void k_wall_samples(ViscArgs a, WallLaunch w) {
    const int i = blockIdx.x * kVB + threadIdx.x;
    if (i >= w.n) return;
    const MeshView& m = a.m;
    const int fref = w.wall_faces[i];
    const bool valid = fref >= 0 && fref < m.nf && m.fr[w.face_old2new[fref]] < 0;
    double* ct = w.contrib;
    if (!valid) {
        // not a boundary face: reported as face -1 with NaN values, left out of the sums
        w.patch[i] = -1;
        for (int q = 0; q < 5; ++q) ct[q * w.n + i] = 0.0;
        if (w.samples) {
            fvk_wall_sample o{};
            o.face_index = -1;
            o.patch_id = -1;
            const double nan = __builtin_nan("");
            o.x_mid = o.y_mid = o.p = o.rho = o.effective_viscosity = o.tau_wall = nan;
            o.tx = o.ty = o.y_wall_normal = o.cf = o.cp = o.y_plus = nan;
            w.samples[i] = o;
        }
        return;
    }
    const int f = w.face_old2new[fref];
    const int cl = m.fl[f];
    const int patch = -1 - m.fr[f];
    const double g = a.gamma;
    // the Euler solver keeps its state as one double4 per cell, the others as [nv][ld]
    const State U = w.aos ? [&] {
        const double4 q = reinterpret_cast<const double4*>(a.U)[cl];
        return State{q.x, q.y, q.z, q.w};
    }() : ldS(a.U, a.ld, cl);
    const double u = U.ru / U.r, v = U.rv / U.r;
    const double p = pressure(U, g);
    const double nx = m.fnx[f], ny = m.fny[f];
    double mu = 0.0, tau = 0.0;
    if (EQ != 0 && !w.pressure_only) {
        const double* gc = a.G + (int64_t)cl * w.ng;  // u (0, 1), v (2, 3) of the cell
        double bv[7];
        boundary_values<EQ>(a, m.bc[patch], f, U, u, v, temperature(U, g, a.Rgas),
                            EQ == EQ_NS ? 0.0 : a.U[4 * a.ld + cl], EQ == EQ_SST ? a.U[5 * a.ld + cl] : 0.0, bv);
        // face_gradient with the wall value, then the corrected vector
        // (src/GradientReconstruction.cpp:153-200)
        double gux, guy, gvx, gvy;
        corrected(face_grad(m, f, true, u, bv[0], gc[0], gc[1], 0.0, 0.0), nx, ny, gux, guy);
        corrected(face_grad(m, f, true, v, bv[1], gc[2], gc[3], 0.0, 0.0), nx, ny, gvx, gvy);
        mu = a.mu;
        if (EQ == EQ_SA) {
            mu = a.mu + U.r * sa_eddy(a.U[4 * a.ld + cl], a.mu / U.r, a.sa);
        } else if (EQ == EQ_SST) {
            const double s11 = gc[0], s22 = gc[3], s12 = 0.5 * (gc[1] + gc[2]);
            const double S = sqrt(2.0 * (s11 * s11 + s22 * s22 + 2.0 * s12 * s12));
            const double Om = fabs(gc[2] - gc[1]);
            const double k = a.U[4 * a.ld + cl], om = a.U[5 * a.ld + cl];
            const double nul = a.mu / U.r;
            const double F2 = sst_F2_ref(k, om, nul, m.wdist[cl], a.sst);
            mu = a.mu + U.r * sst_eddy_ref(k, om, S, Om, F2, a.sst_variant, a.sst);
        }
        const double div_v = gux + gvy;
        const double txx = mu * (2.0 * gux - (2.0 / 3.0) * div_v);
        const double tyy = mu * (2.0 * gvy - (2.0 / 3.0) * div_v);
        const double txy = mu * (guy + gvx);
        const double trx = txx * nx + txy * ny;
        const double try_ = txy * nx + tyy * ny;
        tau = -(trx * (-ny) + try_ * nx);
    }
    const double tx = -ny, ty = nx;
    const double xm = m.fmx[f], ym = m.fmy[f];
    const double A = m.fa[f];
    // compute_wall_forces, per sample (src/WallTraction.cpp:180-200)
    const double fx = A * tau * tx, fy = A * tau * ty;
    const double px = A * p * nx, py = A * p * ny;
    const double rx = xm - w.mx, ry = ym - w.my;
    const double dfx = fx + px, dfy = fy + py;
    const double fd = fx * w.dir_x + fy * w.dir_y;
    const double pd = px * w.dir_x + py * w.dir_y;
    ct[i] = fd;
    ct[w.n + i] = pd;
    ct[2 * w.n + i] = fd + pd;
    ct[3 * w.n + i] = (fx + px) * w.lift_x + (fy + py) * w.lift_y;
    ct[4 * w.n + i] = rx * dfy - ry * dfx;
    w.patch[i] = patch;
    if (w.samples) {
        fvk_wall_sample o;
        o.face_index = fref;
        o.patch_id = patch;
        o.x_mid = xm;
        o.y_mid = ym;
        o.p = p;
        o.rho = U.r;
        o.effective_viscosity = mu;
        o.tau_wall = tau;
        o.tx = tx;
        o.ty = ty;
        o.y_wall_normal = m.fdc[f];
        // skin_friction_coefficient, pressure_coefficient, wall_y_plus (src/WallTraction.cpp:145-161)
        o.cf = tau / (0.5 * w.rho_ref * w.v2);
        o.cp = (p - w.p_ref) / (0.5 * w.rho_ref * w.v2);
        const double u_tau = sqrt(fabs(tau) / U.r);
        o.y_plus = U.r * u_tau * o.y_wall_normal / mu;
        w.samples[i] = o;
    }
}

// one workgroup: per-patch and total sums in sample order (thread 0, from
// LDS chunks the workgroup stages), then the reports by ascending patch id
// and the total last (the reference's std::map order after its partition)
__global__ __launch_bounds__(kWallBlock) void k_wall_reduce(WallLaunch w) {
    extern __shared__ double wsm[];
    const int np = w.n_patches;
    double* acc = wsm;                        // [np][5]
    double* chunk = wsm + 5 * np;             // [5][kWallBlock]
    int* cp = reinterpret_cast<int*>(chunk + 5 * kWallBlock);  // [kWallBlock] patch of each staged sample
    int* cnt = cp + kWallBlock;               // [np]
    for (int q = threadIdx.x; q < 5 * np; q += kWallBlock) acc[q] = 0.0;
    for (int q = threadIdx.x; q < np; q += kWallBlock) cnt[q] = 0;
    double tot[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int ntot = 0;
    __syncthreads();
    for (int base = 0; base < w.n; base += kWallBlock) {
        const int i = base + threadIdx.x;
        if (i < w.n) {
#pragma unroll
            for (int q = 0; q < 5; ++q) chunk[q * kWallBlock + threadIdx.x] = w.contrib[q * w.n + i];
            cp[threadIdx.x] = w.patch[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int e = min(kWallBlock, w.n - base);
            for (int j = 0; j < e; ++j) {
                const int pj = cp[j];
                if (pj < 0 || pj >= np) continue;
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    const double x = chunk[q * kWallBlock + j];
                    acc[pj * 5 + q] += x;
                    tot[q] += x;
                }
                cnt[pj] += 1;
                ++ntot;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    int k = 0;
    auto emit = [&](int pid, const double* s) {
        fvk_wall_force_report r{};
        r.patch_id = pid;
        r.friction_drag = s[0];
        r.pressure_drag = s[1];
        r.total_drag = s[2];
        r.lift = s[3];
        r.moment = s[4];
        r.cd_friction = r.friction_drag / w.q;
        r.cd_pressure = r.pressure_drag / w.q;
        r.cd_total = r.total_drag / w.q;
        r.cl = r.lift / w.q;
        r.cm = r.moment / w.q2;
        if (k < w.cap) w.reports[k] = r;
        ++k;
    };
    for (int pid = 0; pid < np; ++pid)
        if (cnt[pid] > 0) emit(pid, acc + 5 * pid);
    if (ntot > 0) emit(-1, tot);
    *w.n_reports = k;
}

}  // namespace

int launch_wall_diagnostics(fvk_solver* s, const WallLaunch& w0) {
    WallLaunch w = w0;
    const int eq = s->p.equation;
    ViscArgs a = make_vargs(s);
    a.U = s->U[s->cur];
    w.ng = eq == FVK_EQ_RANS_SST ? 10 : (eq == FVK_EQ_RANS_SA ? 8 : 6);
    w.aos = eq == FVK_EQ_EULER;
    if (eq != FVK_EQ_EULER && !w.pressure_only) {
        // cell gradients of the current state (the stage's gradient pass, its
        // turbulence sources included, with no finalize folded in)
        a.fin_nblocks = 0;
        const int gc = grid_for(s->nc, kVB);
        if (eq == FVK_EQ_NAVIER_STOKES) launch_grad_eq<EQ_NS>(s, gc, a);
        else if (eq == FVK_EQ_RANS_SA) launch_grad_eq<EQ_SA>(s, gc, a);
        else launch_grad_eq<EQ_SST>(s, gc, a);
    }
    if (w.n > 0) {
        const int grid = (w.n + kVB - 1) / kVB;
        if (eq == FVK_EQ_EULER || w.pressure_only) k_wall_samples<0><<<grid, kVB, 0, s->stream>>>(a, w);
        else if (eq == FVK_EQ_NAVIER_STOKES) k_wall_samples<EQ_NS><<<grid, kVB, 0, s->stream>>>(a, w);
        else if (eq == FVK_EQ_RANS_SA) k_wall_samples<EQ_SA><<<grid, kVB, 0, s->stream>>>(a, w);
        else k_wall_samples<EQ_SST><<<grid, kVB, 0, s->stream>>>(a, w);
    }
    const size_t lds = sizeof(double) * (5 * (size_t)w.n_patches + 5 * kWallBlock) +
                       sizeof(int) * ((size_t)kWallBlock + w.n_patches);
    k_wall_reduce<<<1, kWallBlock, lds, s->stream>>>(w);
    return hip_check(hipGetLastError(), "wall diagnostics launch");
}

int launch_wall_distance_segments(const double* cx, const double* cy, int nc, const double* seg, int nw, double* out,
                                  hipStream_t st) {
    int grid = grid_for(nc, kVB);
    k_wall_distance<<<grid, kVB, 0, st>>>(cx, cy, nc, seg, nw, out);
    return hip_check(hipGetLastError(), "wall distance launch");
}

}  // namespace fvk
