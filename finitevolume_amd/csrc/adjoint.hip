// Discrete adjoint (vector-jacobian product) of the Euler residual
//
//   R_c(U, X) = sum over the faces f of c, in Cell::faces order, of
//               -A_f F*(U_L^f, U_R^f, n_f) (c owns f) or +A_f F*(...) (c is its neighbour)
//
// (src/EulerFVMSolver.cpp:61-104) with, for second order, MUSCL states from
// the LSQ gradients of (rho, u, v, p) limited by Venkatakrishnan.  Given
// lambda, the kernels return U_bar = (dR/dU)^T lambda and X_bar =
// (dR/dX)^T lambda, X the node coordinates: face length and normal come from
// the face's nodes, face midpoints from their average, cell centroid and
// volume (through the limiter's eps^2) from the shoelace formulas over the
// cell polygon.  The reference has no adjoint; the derivative is that of
// the forward kernels' own operations.
//
// Passes (every sum is a gather in a fixed order: no atomics):
//   grad     forward limited gradients G (the primal's first pass)
//   face     per face: adjoint of A F*(recon L, recon R or ghost), giving
//            each side's state adjoint and the adjoint of its four MUSCL
//            increments, and the face edge-vector adjoint
//   cell     per cell: gathers its face sides; reverses limiter and LSQ
//            (recomputed), leaving per ELL slot the adjoint of the
//            neighbour's primitive state and geometry adjoints
//   gather   per cell: adds the neighbours' slot contributions that point
//            back to it, maps primitive to conservative adjoints
//            and, for node coordinates, reverses the cell polygon's
//            shoelace centroid and area and its faces' edge and midpoint
//            adjoints onto its vertices
//   nodes    (node coordinates) per node: gathers its cells' vertex adjoints
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.h"
#include "solver.h"

namespace fvk {

struct AdjointWork {
    int D = 0;
    bool geo = false;
    // per face side: the adjoint of the side's face state (second order: the
    // primitive face state; first order: the conserved state), and fok, the
    // sides whose face state was reconstructed (bit 0 left, bit 1 right)
    double4 *lam = nullptr, *ful = nullptr, *fur = nullptr;
    unsigned char* fok = nullptr;
    double4 *ubp = nullptr, *S = nullptr, *ubar = nullptr;
    double2* fe = nullptr;
    int* bslot = nullptr;
    // node coordinates
    double4* cb = nullptr;
    double2 *mb = nullptr, *pb = nullptr, *xcv = nullptr, *xbar = nullptr, *node_xy = nullptr;
    int *poly = nullptr, *pslot = nullptr;
    int *cvptr = nullptr, *cvent = nullptr;
    // adjoint solve: dJ/dU, the RK stage base, the pseudo-time scaling, residual partials
    double4 *sg = nullptr, *slam0 = nullptr;
    double *sP = nullptr, *spart = nullptr;
    double4* lsqm = nullptr;  // per cell: LSQ normal matrix m00, m01, m11 and its determinant
    double2* lsqd = nullptr;  // per ELL slot: offset (dx, dy) to the LSQ neighbour point
};

void adjoint_release(fvk_solver* s) {
    delete s->adj;  // device buffers are in the solver's allocation list
    s->adj = nullptr;
}

struct AdjArgs {
    MeshView m;
    const double4* U;
    const double4* G;
    const double4* lam;
    double4 *ful, *fur, *ubp, *S, *ubar, *cb;
    unsigned char* fok;
    double2 *fe, *mb, *pb, *xcv, *xbar;
    const int* bslot;
    const int *poly, *pslot;
    const int *cvptr, *cvent;
    const double2* node_xy;
    const double4* lsqm;
    const double2* lsqd;
    double gamma, venk_k;
    int limiter, order, grad, nn;
};

constexpr int kAdjBlock = 256;

__device__ __forceinline__ State st4(const double4& u) { return State{u.x, u.y, u.z, u.w}; }
__device__ __forceinline__ double4 d4(const State& s) { return double4{s.r, s.ru, s.rv, s.E}; }

// ---------------------------------------------------------------------------
// face pass
// ---------------------------------------------------------------------------
// The adjoint of one face: its two states (R unused on a boundary face,
// cr < 0 the boundary code), their limited gradients, the face geometry and
// the lambdas of its cells.  Writes both sides' state and MUSCL-increment
// adjoints and, for node coordinates (WX), the edge-vector adjoint; without
// it the geometry adjoints are dead code.
template <int SCHEME, int ORDER, bool WX>
__device__ __forceinline__ void adj_face_body(const AdjArgs& a, int f, int cr, const State& L, const State& R,
                                              const double4& gl0, const double4& gl1, const double4& gr0,
                                              const double4& gr1, const double4& ro, const double4& fg,
                                              const double4& ll, const double4& lr) {
    const MeshView& m = a.m;
    const double g = a.gamma;
    const double nx = fg.x, ny = fg.y, A = fg.z;
    const bool interior = cr >= 0;
    const State Fb = interior ? State{lr.x - ll.x, lr.y - ll.y, lr.z - ll.z, lr.w - ll.w}
                              : State{-ll.x, -ll.y, -ll.z, -ll.w};
    const State Hb{A * Fb.r, A * Fb.ru, A * Fb.rv, A * Fb.E};
    double nxb = 0.0, nyb = 0.0, Ab;
    const DevBC* bc = interior ? nullptr : &m.bc[-1 - cr];
    if (ORDER == 2) {
        // primitive face states, as the forward face pass forms them
        const Prim Lf = face_state_w(L, gl0, gl1, ro.x, ro.y, g);
        const Prim Rf = interior ? face_state_w(R, gr0, gr1, ro.z, ro.w, g)
                                 : euler_ghost_w(bc->type, bc->ff, Lf, nx, ny, g);
        const State H = num_flux_w<SCHEME>(Lf, Rf, nx, ny, g, 0.0, 0);
        Ab = Fb.r * H.r + Fb.ru * H.ru + Fb.rv * H.rv + Fb.E * H.E;
        Prim Lfb{0, 0, 0, 0}, Rfb{0, 0, 0, 0};
        num_flux_w_rev<SCHEME>(Lf, Rf, nx, ny, g, Hb, Lfb, Rfb, nxb, nyb);
        if (!interior) euler_ghost_w_rev(bc->type, Lf, nx, ny, Rfb, Lfb, nxb, nyb);
        // each side's primitive face-state adjoint: the cell pass forms from
        // it both the state adjoint (through W(U), one conversion per cell)
        // and, where the side was reconstructed, the increments' adjoint
        a.ful[f] = double4{Lfb.r, Lfb.u, Lfb.v, Lfb.p};
        int ok = face_state_w_reconstructs(L, gl0, gl1, ro.x, ro.y, g) ? 1 : 0;
        if (interior) {
            a.fur[f] = double4{Rfb.r, Rfb.u, Rfb.v, Rfb.p};
            if (face_state_w_reconstructs(R, gr0, gr1, ro.z, ro.w, g)) ok |= 2;
        }
        a.fok[f] = (unsigned char)ok;
    } else {
        const State Rf = interior ? R : euler_ghost_state(bc->type, bc->ff, L, nx, ny, g);
        const State H = num_flux<SCHEME>(L, Rf, nx, ny, g, 0.0, 0);
        Ab = Fb.r * H.r + Fb.ru * H.ru + Fb.rv * H.rv + Fb.E * H.E;
        State Lfb{0, 0, 0, 0}, Rfb{0, 0, 0, 0};
        num_flux_rev<SCHEME>(L, Rf, nx, ny, g, Hb, Lfb, Rfb, nxb, nyb);
        if (!interior) euler_ghost_rev(bc->type, L, nx, ny, g, Rfb, Lfb, nxb, nyb);
        a.ful[f] = d4(Lfb);
        if (interior) a.fur[f] = d4(Rfb);
    }
    if (WX) {
        double exb = 0.0, eyb = 0.0;
        edge_rev(nx, ny, A, nxb, nyb, Ab, exb, eyb);
        a.fe[f] = double2{exb, eyb};
    }
}

// Without the geometry adjoints the pass fits 128 VGPRs, four wavefronts per
// SIMD, with no spills; with them it needs more and runs at three.
template <int SCHEME, int ORDER, bool WX>
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(WX ? 1 : 4))) void k_adj_face(AdjArgs a) {
    const MeshView& m = a.m;
    const int f = blockIdx.x * kAdjBlock + threadIdx.x;
    if (f >= m.nf) return;
    const int cl = m.fl[f], cr = m.fr[f];
    const double4 fg = m.fgeo[f];
    const State L = st4(a.U[cl]);
    double4 gl0{0, 0, 0, 0}, gl1{0, 0, 0, 0}, gr0{0, 0, 0, 0}, gr1{0, 0, 0, 0}, ro{0, 0, 0, 0};
    if (ORDER == 2) {
        ro = m.frec[f];
        gl0 = a.G[2 * cl];
        gl1 = a.G[2 * cl + 1];
    }
    State R{0, 0, 0, 0};
    const double4 ll = a.lam[cl];
    double4 lr{0, 0, 0, 0};
    if (cr >= 0) {
        R = st4(a.U[cr]);
        if (ORDER == 2) {
            gr0 = a.G[2 * cr];
            gr1 = a.G[2 * cr + 1];
        }
        lr = a.lam[cr];
    }
    adj_face_body<SCHEME, ORDER, WX>(a, f, cr, L, R, gl0, gl1, gr0, gr1, ro, fg, ll, lr);
}

// ---------------------------------------------------------------------------
// cell pass: face sides, then the reverse of the limited LSQ gradient
// (recomputed exactly as the forward gradient kernel evaluates it)
// ---------------------------------------------------------------------------
// The cell reverse of cell c (device index): its ELL entries en (.x the
// face entry; .y the neighbour, whose primitives nbr(en.y, w) returns, or a
// boundary code), the entries' geometry eg, its state and limited gradients.
template <int D, bool WX, class Nbr>
__device__ __forceinline__ void adj_cell_body(const AdjArgs& a, int c, const int2 (&en)[D], const double4 (&eg)[D],
                                              const State& Uc, const double4& G0, const double4& G1, Nbr nbr) {
    const MeshView& m = a.m;
    const double g = a.gamma;
    const int64_t ld = m.ld;
    double Gb[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    double rb[D][2];
#pragma unroll
    for (int k = 0; k < D; ++k) rb[k][0] = rb[k][1] = 0.0;
    // the face sides of this cell, summed in ws (first order: conserved-state
    // adjoints; second order: primitive face-state adjoints, converted once
    // below; an increment's adjoint is the side's where the face state was
    // reconstructed, else zero)
    double ws[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        if (en[k].x < 0) continue;
        const int f = en[k].x >> 1;
        const bool owner = en[k].x & 1;
        const double4 u = owner ? a.ful[f] : a.fur[f];
        ws[0] += u.x; ws[1] += u.y; ws[2] += u.z; ws[3] += u.w;
        if (a.order == 2) {
            const bool rec = (a.fok[f] >> (owner ? 0 : 1)) & 1;
            const double4 d = rec ? u : double4{0.0, 0.0, 0.0, 0.0};
            const double rx = eg[k].z, ry = eg[k].w;
            Gb[0][0] += d.x * rx; Gb[0][1] += d.x * ry;
            Gb[1][0] += d.y * rx; Gb[1][1] += d.y * ry;
            Gb[2][0] += d.z * rx; Gb[2][1] += d.z * ry;
            Gb[3][0] += d.w * rx; Gb[3][1] += d.w * ry;
            if (WX) {
                rb[k][0] += d.x * G0.x + d.y * G0.z + d.z * G1.x + d.w * G1.z;
                rb[k][1] += d.x * G0.y + d.y * G0.w + d.z * G1.y + d.w * G1.w;
            }
        }
    }
    double Cbx = 0.0, Cby = 0.0, Vb = 0.0;
    double nbx[D], nby[D], Pbx[D], Pby[D];
#pragma unroll
    for (int k = 0; k < D; ++k) nbx[k] = nby[k] = Pbx[k] = Pby[k] = 0.0;
    State Ub{0, 0, 0, 0};
    if (a.grad) {
        // ---- forward recompute ----
        const Prim wc = to_prim(Uc, g);
        const double wcq[4] = {wc.r, wc.u, wc.v, wc.p};
        double wn[D][4];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            wn[k][0] = wn[k][1] = wn[k][2] = wn[k][3] = 0.0;
            if (en[k].x < 0) continue;
            if (en[k].y < 0) {
                const double4 fg = m.fgeo[en[k].x >> 1];
                const DevBC& bc = m.bc[-1 - en[k].y];
                const State Gs = euler_ghost_state(bc.type, bc.ff, Uc, fg.x, fg.y, g);
                const double wg0 = Gs.r, wg1 = fdiv(Gs.ru, Gs.r), wg2 = fdiv(Gs.rv, Gs.r), wg3 = pressure(Gs, g);
                wn[k][0] = 0.5 * (wcq[0] + wg0);
                wn[k][1] = 0.5 * (wcq[1] + wg1);
                wn[k][2] = 0.5 * (wcq[2] + wg2);
                wn[k][3] = 0.5 * (wcq[3] + wg3);
            } else {
                nbr(en[k].y, wn[k]);
            }
        }
        double bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0}, wmin[4], wmax[4];
        int amin[4] = {-1, -1, -1, -1}, amax[4] = {-1, -1, -1, -1};
#pragma unroll
        for (int q = 0; q < 4; ++q) wmin[q] = wmax[q] = wcq[q];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (en[k].x < 0) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double dphi = wn[k][q] - wcq[q];
                bx[q] += eg[k].x * dphi;
                by[q] += eg[k].y * dphi;
                if (wn[k][q] < wmin[q]) { wmin[q] = wn[k][q]; amin[q] = k; }
                if (wmax[q] < wn[k][q]) { wmax[q] = wn[k][q]; amax[q] = k; }
            }
        }
        const double4 iv = m.clsq[c];
        const double eps2 = iv.w;
        double gx[4], gy[4], lim[4];
        int kl[4] = {-1, -1, -1, -1};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            gx[q] = iv.x * bx[q] + iv.y * by[q];
            gy[q] = iv.y * bx[q] + iv.z * by[q];
            lim[q] = 1.0;
        }
        if (a.limiter == FVK_LIMITER_VENKATAKRISHNAN) {
            VenkatMin vm[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) vm[q] = VenkatMin{1.0, 1.0};
#pragma unroll
            for (int k = 0; k < D; ++k) {
                if (en[k].x < 0) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (venkat_pick(vm[q], wmin[q], wmax[q], wcq[q], gx[q] * eg[k].z + gy[q] * eg[k].w, eps2))
                        kl[q] = k;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) lim[q] = venkat_value(vm[q]);
        }
        // ---- reverse: G = lim g ----
        double gbx[4], gby[4], wcb[4] = {ws[0], ws[1], ws[2], ws[3]}, wnb[D][4], eps2b = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) wnb[k][0] = wnb[k][1] = wnb[k][2] = wnb[k][3] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double limb = Gb[q][0] * gx[q] + Gb[q][1] * gy[q];
            gbx[q] = lim[q] * Gb[q][0];
            gby[q] = lim[q] * Gb[q][1];
            // lim = venkatakrishnan(delta, d2, eps2) of the minimising face
#pragma unroll
            for (int k = 0; k < D; ++k) {
                if (k != kl[q]) continue;
                const double rx = eg[k].z, ry = eg[k].w;
                const double d2 = gx[q] * rx + gy[q] * ry;
                const bool up = d2 > 0.0;
                const double d1 = (up ? wmax[q] : wmin[q]) - wcq[q];
                double p1, p2, pe;
                venkatakrishnan_grad(d1, d2, eps2, p1, p2, pe);
                const double d1b = limb * p1, d2b = limb * p2;
                eps2b += limb * pe;
                gbx[q] += d2b * rx;
                gby[q] += d2b * ry;
                if (WX) {
                    rb[k][0] += d2b * gx[q];
                    rb[k][1] += d2b * gy[q];
                }
                wcb[q] -= d1b;
                const int ae = up ? amax[q] : amin[q];
                if (ae < 0) wcb[q] += d1b;
#pragma unroll
                for (int j = 0; j < D; ++j)
                    if (j == ae) wnb[j][q] += d1b;
            }
        }
        // ---- reverse: g = P b, b = sum_k (w d)_k (W_k - W_c) ----
        double bbx[4], bby[4], i0b = 0.0, i1b = 0.0, i2b = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            bbx[q] = iv.x * gbx[q] + iv.y * gby[q];
            bby[q] = iv.y * gbx[q] + iv.z * gby[q];
            if (WX) {
                i0b += gbx[q] * bx[q];
                i1b += gbx[q] * by[q] + gby[q] * bx[q];
                i2b += gby[q] * by[q];
            }
        }
        double wdxb[D], wdyb[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            wdxb[k] = wdyb[k] = 0.0;
            if (en[k].x < 0) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double t = eg[k].x * bbx[q] + eg[k].y * bby[q];
                wnb[k][q] += t;
                wcb[q] -= t;
                if (WX) {
                    wdxb[k] += bbx[q] * (wn[k][q] - wcq[q]);
                    wdyb[k] += bby[q] * (wn[k][q] - wcq[q]);
                }
            }
        }
        // ---- the neighbour values: interior -> the neighbour's own primitive
        // adjoint (gathered by that cell), boundary -> 0.5 (W_c + W(ghost(U_c))) ----
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (en[k].x < 0) continue;
            if (en[k].y >= 0) {
                a.S[k * ld + c] = double4{wnb[k][0], wnb[k][1], wnb[k][2], wnb[k][3]};
            } else {
                const double4 fg = m.fgeo[en[k].x >> 1];
                const DevBC& bc = m.bc[-1 - en[k].y];
                const State Gs = euler_ghost_state(bc.type, bc.ff, Uc, fg.x, fg.y, g);
#pragma unroll
                for (int q = 0; q < 4; ++q) wcb[q] += 0.5 * wnb[k][q];
                State Gsb{0, 0, 0, 0};
                prim_rev(Gs, Prim{0.5 * wnb[k][0], 0.5 * wnb[k][1], 0.5 * wnb[k][2], 0.5 * wnb[k][3]}, Gsb, g);
                euler_ghost_rev(bc.type, Uc, fg.x, fg.y, g, Gsb, Ub, nbx[k], nby[k]);
            }
        }
        prim_rev(Uc, Prim{wcb[0], wcb[1], wcb[2], wcb[3]}, Ub, g);
        if (WX) {
            // The stored inverse is P = (m11, -m01, m00) / det with det =
            // m00 m11 - m01^2 (the set-up's formula): reversed operation by
            // operation.  (-P Pb P would be the derivative of the exact
            // inverse; for the thin, nearly collinear stencils of stretched
            // triangles det keeps few digits and P is not that inverse.)
            const double4 mm = a.lsqm[c];
            const double idet = fdiv(1.0, mm.w);
            const double detb = -(i0b * iv.x + i1b * iv.y + i2b * iv.z) * idet;
            const double M00 = i2b * idet + detb * mm.z;
            const double M11 = i0b * idet + detb * mm.x;
            const double Ms = 0.5 * (-i1b * idet - 2.0 * detb * mm.y);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                if (en[k].x < 0) continue;
                // offset d to the neighbour point and weight w = 1/|d|^2, as the set-up formed them
                const double2 dd = a.lsqd[k * ld + c];
                const double dx = dd.x, dy = dd.y;
                const double w = fdiv(1.0, dx * dx + dy * dy);
                const double sq = M00 * dx * dx + 2.0 * Ms * dx * dy + M11 * dy * dy;
                const double wq = wdxb[k] * dx + wdyb[k] * dy;
                const double w2 = w * w;
                Pbx[k] = 2.0 * w * (M00 * dx + Ms * dy) - 2.0 * w2 * sq * dx + wdxb[k] * w - 2.0 * w2 * wq * dx;
                Pby[k] = 2.0 * w * (Ms * dx + M11 * dy) - 2.0 * w2 * sq * dy + wdyb[k] * w - 2.0 * w2 * wq * dy;
                Cbx -= Pbx[k];
                Cby -= Pby[k];
            }
            // eps^2 = (K sqrt(V))^3
            const double V = m.cgeo[c].z, K = a.venk_k;
            Vb = eps2b * 1.5 * K * K * K * sqrt(V);
        }
    } else if (a.order == 2) {
        prim_rev(Uc, Prim{ws[0], ws[1], ws[2], ws[3]}, Ub, g);
    } else {
        Ub = State{ws[0], ws[1], ws[2], ws[3]};
    }
    a.ubp[c] = d4(Ub);
    if (WX) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (en[k].x < 0) continue;
            // face midpoint - centroid offsets
            Cbx -= rb[k][0];
            Cby -= rb[k][1];
            if (en[k].y >= 0) {
                a.pb[k * ld + c] = double2{Pbx[k], Pby[k]};
                a.mb[k * ld + c] = double2{rb[k][0], rb[k][1]};
            } else {
                // the LSQ point of a boundary face is its midpoint; the slot's
                // pb holds the adjoint of the face normal (ghost state) instead
                a.pb[k * ld + c] = double2{nbx[k], nby[k]};
                a.mb[k * ld + c] = double2{rb[k][0] + Pbx[k], rb[k][1] + Pby[k]};
            }
        }
        a.cb[c] = double4{Cbx, Cby, Vb, 0.0};
    }
}

template <int D, bool WX>
__global__ __launch_bounds__(kAdjBlock) void k_adj_cell(AdjArgs a) {
    const MeshView& m = a.m;
    const int c = blockIdx.x * kAdjBlock + threadIdx.x;
    if (c >= m.nc) return;
    const int64_t ld = m.ld;
    int2 en[D];
    double4 eg[D];
#pragma unroll
    for (int k = 0; k < D; ++k) en[k] = m.ell[k * ld + c];
#pragma unroll
    for (int k = 0; k < D; ++k) eg[k] = en[k].x >= 0 ? m.egeo[k * ld + c] : double4{0.0, 0.0, 0.0, 0.0};
    double4 G0{0, 0, 0, 0}, G1{0, 0, 0, 0};
    if (a.order == 2) {
        G0 = a.G[2 * c];
        G1 = a.G[2 * c + 1];
    }
    const double g = a.gamma;
    adj_cell_body<D, WX>(a, c, en, eg, st4(a.U[c]), G0, G1, [&](int o, double (&w)[4]) {
        const State Uo = st4(a.U[o]);
        w[0] = Uo.r;
        w[1] = fdiv(Uo.ru, Uo.r);
        w[2] = fdiv(Uo.rv, Uo.r);
        w[3] = pressure(Uo, g);
    });
}

// ---------------------------------------------------------------------------
// gather pass: the neighbours' slot contributions that point back to this
// cell; then (node coordinates) the cell polygon: shoelace centroid and
// area reversed to its vertices, plus the edge-vector and midpoint adjoints
// of its faces (each face's edge adjoint is added by its owner, along the
// owner's directed edge node1 -> node2)
// ---------------------------------------------------------------------------
// The state part of cell c's gather: its own face sides (ubp) plus the
// neighbours' slot contributions, mapped from primitive to conservative
// adjoints; WX: the neighbours' LSQ-point adjoints summed into cbx, cby.
template <int D, bool WX>
__device__ __forceinline__ State adj_gather_state(const AdjArgs& a, int c, double& cbx, double& cby) {
    const MeshView& m = a.m;
    const int64_t ld = m.ld;
    State Ub = st4(a.ubp[c]);
    if (a.grad) {
        int2 en[D];
        int bs[D];
#pragma unroll
        for (int k = 0; k < D; ++k) en[k] = m.ell[k * ld + c];
#pragma unroll
        for (int k = 0; k < D; ++k) bs[k] = a.bslot[k * ld + c];
        Prim wb{0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (en[k].x < 0 || en[k].y < 0) continue;
            const int64_t src = (int64_t)bs[k] * ld + en[k].y;
            const double4 s4 = a.S[src];
            wb.r += s4.x; wb.u += s4.y; wb.v += s4.z; wb.p += s4.w;
            if (WX) {
                const double2 p = a.pb[src];
                cbx += p.x;
                cby += p.y;
            }
        }
        prim_rev(st4(a.U[c]), wb, Ub, a.gamma);
    }
    return Ub;
}

template <int D, bool WX>
__global__ __launch_bounds__(kAdjBlock) void k_adj_gather(AdjArgs a) {
    const MeshView& m = a.m;
    const int c = blockIdx.x * kAdjBlock + threadIdx.x;
    if (c >= m.nc) return;
    const int64_t ld = m.ld;
    double cbx = 0.0, cby = 0.0;
    const State Ub = adj_gather_state<D, WX>(a, c, cbx, cby);
    a.ubar[c] = d4(Ub);
    if (!WX) return;
    int v[D], ps[D];
    int nv = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        v[k] = a.poly[k * ld + c];
        ps[k] = a.pslot[k * ld + c];
        if (v[k] >= 0) nv = k + 1;
    }
    const double2 base = a.node_xy[v[0]];
    double x[D], y[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        x[k] = y[k] = 0.0;
        if (k < nv) {
            const double2 p = a.node_xy[v[k]];
            x[k] = p.x - base.x;
            y[k] = p.y - base.y;
        }
    }
    double A2 = 0.0, Sx = 0.0, Sy = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        if (i >= nv) continue;
        const double xj = (i + 1 < nv) ? x[(i + 1) % D] : x[0];
        const double yj = (i + 1 < nv) ? y[(i + 1) % D] : y[0];
        const double cr = x[i] * yj - xj * y[i];
        A2 += cr;
        Sx += (x[i] + xj) * cr;
        Sy += (y[i] + yj) * cr;
    }
    const double i3A2 = fdiv(1.0, 3.0 * A2);
    const double cx = Sx * i3A2, cy = Sy * i3A2;
    const double4 cb = a.cb[c];
    const double Cbx = cb.x + cbx, Cby = cb.y + cby;
    const double Sxb = Cbx * i3A2, Syb = Cby * i3A2;
    const double A2b = -(Cbx * cx + Cby * cy) * (3.0 * i3A2) + 0.5 * (A2 > 0.0 ? 1.0 : -1.0) * cb.z;
    double xb[D], yb[D];
#pragma unroll
    for (int k = 0; k < D; ++k) xb[k] = yb[k] = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        if (i >= nv) continue;
        const bool wrap = i + 1 >= nv;
        const double xj = wrap ? x[0] : x[(i + 1) % D];
        const double yj = wrap ? y[0] : y[(i + 1) % D];
        const double cr = x[i] * yj - xj * y[i];
        const double crb = A2b + Sxb * (x[i] + xj) + Syb * (y[i] + yj);
        double xjb = Sxb * cr - crb * y[i];
        double yjb = Syb * cr + crb * x[i];
        xb[i] += Sxb * cr + crb * yj;
        yb[i] += Syb * cr - crb * xj;
        // the face of this edge (vertex i -> vertex i+1): midpoint and edge vector
        const int k = ps[i];
        const int2 e = m.ell[k * ld + c];
        const double2 m2 = a.mb[k * ld + c];
        xb[i] += 0.5 * m2.x;
        yb[i] += 0.5 * m2.y;
        xjb += 0.5 * m2.x;
        yjb += 0.5 * m2.y;
        if (e.x & 1) {
            const int f = e.x >> 1;
            const double2 fe = a.fe[f];
            double exb = fe.x, eyb = fe.y;
            if (e.y < 0 && a.grad) {
                const double2 nb = a.pb[k * ld + c];
                const double4 fg = m.fgeo[f];
                edge_rev(fg.x, fg.y, fg.z, nb.x, nb.y, 0.0, exb, eyb);
            }
            xb[i] -= exb;
            yb[i] -= eyb;
            xjb += exb;
            yjb += eyb;
        }
        if (wrap) { xb[0] += xjb; yb[0] += yjb; }
        else {
#pragma unroll
            for (int j = 0; j < D; ++j)
                if (j == i + 1) { xb[j] += xjb; yb[j] += yjb; }
        }
    }
#pragma unroll
    for (int k = 0; k < D; ++k)
        if (k < nv) a.xcv[k * ld + c] = double2{xb[k], yb[k]};
}

// set-up (node coordinates): per cell the LSQ normal matrix and its
// determinant, per slot the offset to the LSQ point, formed exactly as the
// host set-up of the stored inverse forms them (same operands, same order,
// IEEE division), so that the reverse of the inverse's formula sees its
// operands' own bits
template <int D>
__global__ __launch_bounds__(kAdjBlock) void k_adj_lsq_setup(MeshView m, double4* lsqm, double2* lsqd) {
    const int c = blockIdx.x * kAdjBlock + threadIdx.x;
    if (c >= m.nc) return;
    const int64_t ld = m.ld;
    const double4 cc = m.cgeo[c];
    double m00 = 0.0, m01 = 0.0, m11 = 0.0;
    for (int k = 0; k < D; ++k) {
        const int2 e = m.ell[k * ld + c];
        if (e.x < 0) continue;
        double qx, qy;
        if (e.y < 0) {
            const double2 mid = m.fmid2[e.x >> 1];
            qx = mid.x;
            qy = mid.y;
        } else {
            const double4 o = m.cgeo[e.y];
            qx = o.x;
            qy = o.y;
        }
        const double dx = qx - cc.x, dy = qy - cc.y;
        const double w = 1.0 / (dx * dx + dy * dy);
        m00 += w * dx * dx;
        m01 += w * dx * dy;
        m11 += w * dy * dy;
        lsqd[k * ld + c] = double2{dx, dy};
    }
    lsqm[c] = double4{m00, m01, m11, m00 * m11 - m01 * m01};
}

// per node: the vertex contributions of the cells around it
__global__ __launch_bounds__(kAdjBlock) void k_adj_nodes(AdjArgs a) {
    const int n = blockIdx.x * kAdjBlock + threadIdx.x;
    if (n >= a.nn) return;
    const int e0 = a.cvptr[n], e1 = a.cvptr[n + 1];
    double xb = 0.0, yb = 0.0;
    // eight incidences in flight at a time
    for (int e = e0; e < e1; e += 8) {
        int idx[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) idx[j] = e + j < e1 ? a.cvent[e + j] : -1;
        double2 t[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j] = idx[j] >= 0 ? a.xcv[idx[j]] : double2{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            xb += t[j].x;
            yb += t[j].y;
        }
    }
    a.xbar[n] = double2{xb, yb};
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

int grid_of(int64_t n) { return (int)((n + kAdjBlock - 1) / kAdjBlock); }

template <typename T>
T* alloc_n(fvk_solver* s, int64_t n) {
    return static_cast<T*>(dev_alloc(s, sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
}

template <typename T>
int upload_vec(fvk_solver* s, T* dst, const std::vector<T>& v) {
    return hip_check(hipMemcpy(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "adjoint maps upload");
}

int width_of(int deg) { return deg <= 3 ? 3 : deg <= 4 ? 4 : deg <= 9 ? 9 : kMaxEulerDegree; }

// state maps: the slot of the same face in the neighbour's list
int build_state_maps(fvk_solver* s, AdjointWork* w, std::vector<int2>& ell) {
    const int D = w->D;
    const int64_t ld = s->mesh.ld, nc = s->nc;
    ell.assign((size_t)s->mesh.deg * ld, int2{-1, -1});
    int rc = hip_check(hipMemcpy(ell.data(), s->mesh.ell, sizeof(int2) * ell.size(), hipMemcpyDeviceToHost),
                       "adjoint ELL download");
    if (rc != FVK_OK) return rc;
    const int deg = s->mesh.deg;
    std::vector<int> bslot((size_t)D * ld, -1);
    for (int64_t c = 0; c < nc; ++c) {
        for (int k = 0; k < deg; ++k) {
            const int2 e = ell[(size_t)k * ld + c];
            if (e.x < 0 || e.y < 0) continue;
            int found = -1;
            for (int j = 0; j < deg; ++j)
                if (ell[(size_t)j * ld + e.y].x == (e.x ^ 1)) { found = j; break; }
            if (found < 0) return set_error(FVK_ERR_INVALID, "adjoint: face %d is missing from its neighbour's list", e.x >> 1);
            bslot[(size_t)k * ld + c] = found;
        }
    }
    return upload_vec(s, w->bslot, bslot);
}

// node-coordinate maps: each cell's polygon (its directed edges chained:
// node1 -> node2 on the faces it owns, reversed on the others) with the ELL
// slot of the face leaving each vertex, and node -> (cell, vertex) incidences
int build_geo_maps(fvk_solver* s, AdjointWork* w, const std::vector<int2>& ell) {
    const int D = w->D, deg = s->mesh.deg;
    const int64_t ld = s->mesh.ld, nc = s->nc, nn = s->nn;
    std::vector<int> poly((size_t)D * ld, -1), pslot((size_t)D * ld, 0);
    const std::vector<int>& fn = s->face_nodes_dev;
    for (int64_t c = 0; c < nc; ++c) {
        int m = 0;
        int from[kMaxEulerDegree], to[kMaxEulerDegree], slot[kMaxEulerDegree];
        for (int k = 0; k < deg; ++k) {
            const int2 e = ell[(size_t)k * ld + c];
            if (e.x < 0) continue;
            const int f = e.x >> 1;
            const bool owner = e.x & 1;
            from[m] = owner ? fn[2 * f] : fn[2 * f + 1];
            to[m] = owner ? fn[2 * f + 1] : fn[2 * f];
            slot[m] = k;
            ++m;
        }
        int cur = 0;
        bool used[kMaxEulerDegree] = {false};
        used[0] = true;
        poly[c] = from[0];
        pslot[c] = slot[0];
        for (int i = 1; i < m; ++i) {
            int nxt = -1;
            for (int j = 0; j < m; ++j)
                if (!used[j] && from[j] == to[cur]) { nxt = j; break; }
            if (nxt < 0)
                return set_error(FVK_ERR_INVALID, "adjoint: the faces of cell %d do not form a closed polygon",
                                 s->cell_new2old[c]);
            used[nxt] = true;
            poly[(size_t)i * ld + c] = from[nxt];
            pslot[(size_t)i * ld + c] = slot[nxt];
            cur = nxt;
        }
        if (to[cur] != from[0])
            return set_error(FVK_ERR_INVALID, "adjoint: the faces of cell %d do not form a closed polygon",
                             s->cell_new2old[c]);
    }
    std::vector<int> cvptr(nn + 1, 0);
    for (size_t i = 0; i < poly.size(); ++i)
        if (poly[i] >= 0) cvptr[poly[i] + 1]++;
    for (int64_t n = 0; n < nn; ++n) cvptr[n + 1] += cvptr[n];
    std::vector<int> cvent(std::max(cvptr[nn], 1));
    {
        std::vector<int> fill(cvptr.begin(), cvptr.end() - 1);
        for (size_t i = 0; i < poly.size(); ++i)
            if (poly[i] >= 0) cvent[fill[poly[i]]++] = (int)i;
    }
    std::vector<double2> xy(nn);
    for (int64_t n = 0; n < nn; ++n) xy[n] = double2{s->node_xy[2 * n], s->node_xy[2 * n + 1]};
    w->mb = alloc_n<double2>(s, (int64_t)D * ld);
    w->pb = alloc_n<double2>(s, (int64_t)D * ld);
    w->cb = alloc_n<double4>(s, ld);
    w->xcv = alloc_n<double2>(s, (int64_t)D * ld);
    w->xbar = alloc_n<double2>(s, nn);
    w->node_xy = alloc_n<double2>(s, nn);
    w->poly = alloc_n<int>(s, (int64_t)D * ld);
    w->pslot = alloc_n<int>(s, (int64_t)D * ld);
    w->cvptr = alloc_n<int>(s, nn + 1);
    w->cvent = alloc_n<int>(s, (int64_t)cvent.size());
    w->lsqm = alloc_n<double4>(s, ld);
    w->lsqd = alloc_n<double2>(s, (int64_t)D * ld);
    if (!w->mb || !w->pb || !w->cb || !w->xcv || !w->xbar || !w->node_xy || !w->poly || !w->pslot || !w->cvptr ||
        !w->cvent || !w->lsqm || !w->lsqd)
        return set_error(FVK_ERR_NOMEM, "adjoint: device allocation failed");
    int rc = upload_vec(s, w->poly, poly);
    if (rc == FVK_OK) rc = upload_vec(s, w->pslot, pslot);
    if (rc == FVK_OK) rc = upload_vec(s, w->cvptr, cvptr);
    if (rc == FVK_OK) rc = upload_vec(s, w->cvent, cvent);
    if (rc == FVK_OK) rc = upload_vec(s, w->node_xy, xy);
    if (rc == FVK_OK && s->p.order == 2) {
        const int gc = grid_of(nc);
        switch (D) {
            case 3: k_adj_lsq_setup<3><<<gc, kAdjBlock, 0, s->stream>>>(s->mesh, w->lsqm, w->lsqd); break;
            case 4: k_adj_lsq_setup<4><<<gc, kAdjBlock, 0, s->stream>>>(s->mesh, w->lsqm, w->lsqd); break;
            case 9: k_adj_lsq_setup<9><<<gc, kAdjBlock, 0, s->stream>>>(s->mesh, w->lsqm, w->lsqd); break;
            default: k_adj_lsq_setup<kMaxEulerDegree><<<gc, kAdjBlock, 0, s->stream>>>(s->mesh, w->lsqm, w->lsqd);
        }
        rc = hip_check(hipGetLastError(), "adjoint LSQ set-up launch");
        if (rc == FVK_OK) rc = hip_check(hipStreamSynchronize(s->stream), "adjoint LSQ set-up");
    }
    if (rc == FVK_OK) w->geo = true;
    return rc;
}

int ensure_work(fvk_solver* s, bool want_x) {
    if (!s->adj) {
        AdjointWork* w = new (std::nothrow) AdjointWork;
        if (!w) return set_error(FVK_ERR_NOMEM, "host allocation failed");
        s->adj = w;
        const int64_t ld = s->mesh.ld, nf = s->nf;
        w->D = width_of(s->mesh.deg);
        w->lam = alloc_n<double4>(s, ld);
        w->ful = alloc_n<double4>(s, nf);
        w->fur = alloc_n<double4>(s, nf);
        w->fok = alloc_n<unsigned char>(s, nf);
        w->ubp = alloc_n<double4>(s, ld);
        w->S = alloc_n<double4>(s, (int64_t)w->D * ld);
        w->ubar = alloc_n<double4>(s, ld);
        w->fe = alloc_n<double2>(s, nf);
        w->bslot = alloc_n<int>(s, (int64_t)w->D * ld);
        if (!w->lam || !w->ful || !w->fur || !w->fok || !w->ubp || !w->S || !w->ubar || !w->fe ||
            !w->bslot) {
            adjoint_release(s);
            return set_error(FVK_ERR_NOMEM, "adjoint: device allocation failed");
        }
        std::vector<int2> ell;
        int rc = build_state_maps(s, w, ell);
        if (rc == FVK_OK) rc = hip_check(hipMemset(w->S, 0, sizeof(double4) * (size_t)w->D * ld), "adjoint init");
        if (rc == FVK_OK && want_x) rc = build_geo_maps(s, w, ell);
        if (rc != FVK_OK) {
            adjoint_release(s);
            return rc;
        }
        return FVK_OK;
    }
    if (want_x && !s->adj->geo) {
        std::vector<int2> ell((size_t)s->mesh.deg * s->mesh.ld);
        int rc = hip_check(hipMemcpy(ell.data(), s->mesh.ell, sizeof(int2) * ell.size(), hipMemcpyDeviceToHost),
                           "adjoint ELL download");
        if (rc == FVK_OK) rc = build_geo_maps(s, s->adj, ell);
        return rc;
    }
    return FVK_OK;
}

template <int D>
void launch_cells(fvk_solver* s, const AdjArgs& a, bool wx) {
    hipEvent_t ev = nullptr;
    const int gc = grid_of(s->nc);
    timer_begin(s, "adj_cell", &ev);
    if (wx) k_adj_cell<D, true><<<gc, kAdjBlock, 0, s->stream>>>(a);
    else k_adj_cell<D, false><<<gc, kAdjBlock, 0, s->stream>>>(a);
    timer_end(s, ev);
    timer_begin(s, "adj_gather", &ev);
    if (wx) k_adj_gather<D, true><<<gc, kAdjBlock, 0, s->stream>>>(a);
    else k_adj_gather<D, false><<<gc, kAdjBlock, 0, s->stream>>>(a);
    timer_end(s, ev);
    if (wx) {
        timer_begin(s, "adj_nodes", &ev);
        k_adj_nodes<<<grid_of(s->nn), kAdjBlock, 0, s->stream>>>(a);
        timer_end(s, ev);
    }
}

template <int ORDER>
void launch_faces(fvk_solver* s, const AdjArgs& a, bool wx) {
    const int gf = grid_of(s->nf);
    if (s->p.flux_scheme == FLUX_HLLC) {
        if (wx) k_adj_face<FLUX_HLLC, ORDER, true><<<gf, kAdjBlock, 0, s->stream>>>(a);
        else k_adj_face<FLUX_HLLC, ORDER, false><<<gf, kAdjBlock, 0, s->stream>>>(a);
    } else {
        if (wx) k_adj_face<FLUX_RUSANOV, ORDER, true><<<gf, kAdjBlock, 0, s->stream>>>(a);
        else k_adj_face<FLUX_RUSANOV, ORDER, false><<<gf, kAdjBlock, 0, s->stream>>>(a);
    }
}

}  // namespace

namespace {

int check_supported(fvk_solver* s, bool wx) {
    if (s->p.equation != FVK_EQ_EULER)
        return set_error(FVK_ERR_UNSUPPORTED, "the residual VJP is built for the Euler equations");
    if (s->p.flux_scheme == FLUX_EXACT)
        return set_error(FVK_ERR_UNSUPPORTED, "no adjoint of the exact Riemann flux (use Rusanov or HLLC)");
    if (wx && s->periodic)
        return set_error(FVK_ERR_UNSUPPORTED, "node-coordinate adjoint of a periodic mesh (its cells are not closed polygons)");
    return FVK_OK;
}

AdjArgs make_args(fvk_solver* s) {
    AdjointWork* w = s->adj;
    AdjArgs a{};
    a.m = s->mesh;
    a.U = reinterpret_cast<const double4*>(s->U[s->cur]);
    a.G = reinterpret_cast<const double4*>(s->grad);
    a.lam = w->lam;
    a.ful = w->ful; a.fur = w->fur; a.fok = w->fok;
    a.ubp = w->ubp; a.S = w->S; a.ubar = w->ubar; a.fe = w->fe; a.bslot = w->bslot;
    a.mb = w->mb; a.pb = w->pb; a.cb = w->cb; a.xcv = w->xcv; a.xbar = w->xbar;
    a.poly = w->poly; a.pslot = w->pslot; a.cvptr = w->cvptr;
    a.cvent = w->cvent; a.node_xy = w->node_xy;
    a.lsqm = w->lsqm; a.lsqd = w->lsqd;
    a.gamma = s->p.gamma;
    a.venk_k = s->p.venkat_k;
    a.limiter = s->p.limiter;
    a.order = s->p.order == 2 ? 2 : 1;
    a.grad = a.order == 2 && s->p.limiter != FVK_LIMITER_FIRST_ORDER;
    a.nn = (int)s->nn;
    return a;
}

// the VJP's passes over lambda already in w->lam (device order): U_bar into
// w->ubar and, with wx, X_bar into w->xbar.  The forward gradients the
// passes read (s->grad) must be those of the current state.
int run_passes(fvk_solver* s, bool wx) {
    AdjointWork* w = s->adj;
    const AdjArgs a = make_args(s);
    hipEvent_t ev = nullptr;
    timer_begin(s, "adj_face", &ev);
    if (a.order == 2) launch_faces<2>(s, a, wx);
    else launch_faces<1>(s, a, wx);
    timer_end(s, ev);
    switch (w->D) {
        case 3: launch_cells<3>(s, a, wx); break;
        case 4: launch_cells<4>(s, a, wx); break;
        case 9: launch_cells<9>(s, a, wx); break;
        default: launch_cells<kMaxEulerDegree>(s, a, wx); break;
    }
    return hip_check(hipGetLastError(), "adjoint launch");
}

int upload_cells(fvk_solver* s, double4* dst, const double* src) {
    const int64_t nc = s->nc, ld = s->mesh.ld;
    std::vector<double4> h(ld, double4{0, 0, 0, 0});
    for (int64_t i = 0; i < nc; ++i) {
        const double* l = src + 4 * (int64_t)s->cell_new2old[i];
        h[i] = double4{l[0], l[1], l[2], l[3]};
    }
    int rc = hip_check(hipMemcpyAsync(dst, h.data(), sizeof(double4) * ld, hipMemcpyHostToDevice, s->stream),
                       "adjoint upload");
    if (rc == FVK_OK) rc = hip_check(hipStreamSynchronize(s->stream), "adjoint upload");
    return rc;
}

int download_cells(fvk_solver* s, double* dst, const double4* src) {
    const int64_t nc = s->nc, ld = s->mesh.ld;
    std::vector<double4> h(ld);
    int rc = hip_check(hipMemcpyAsync(h.data(), src, sizeof(double4) * ld, hipMemcpyDeviceToHost, s->stream),
                       "adjoint copy");
    int rs = hip_check(hipStreamSynchronize(s->stream), "adjoint copy");
    if (rc == FVK_OK) rc = rs;
    if (rc != FVK_OK) return rc;
    for (int64_t i = 0; i < nc; ++i) {
        double* o = dst + 4 * (int64_t)s->cell_new2old[i];
        o[0] = h[i].x; o[1] = h[i].y; o[2] = h[i].z; o[3] = h[i].w;
    }
    return FVK_OK;
}

int download_nodes(fvk_solver* s, double* dst) {
    std::vector<double2> hx(s->nn);
    int rc = hip_check(hipMemcpyAsync(hx.data(), s->adj->xbar, sizeof(double2) * s->nn, hipMemcpyDeviceToHost,
                                      s->stream), "adjoint copy");
    int rs = hip_check(hipStreamSynchronize(s->stream), "adjoint copy");
    if (rc == FVK_OK) rc = rs;
    if (rc != FVK_OK) return rc;
    for (int64_t n = 0; n < s->nn; ++n) { dst[2 * n] = hx[n].x; dst[2 * n + 1] = hx[n].y; }
    return FVK_OK;
}

// ---- adjoint solve: pseudo-time iteration of  dlambda/dtau = P (J^T lambda + g) ----

// P_c = dt_c / V_c with dt_c the primal's local time step at the current
// state (the minimum over the cell's faces of CFL len / (|V.n| + c)): the
// transpose system is iterated with the primal's own preconditioner
template <int D>
__global__ __launch_bounds__(kAdjBlock) void k_adj_precond(MeshView m, const double4* U, double gamma, double cfl,
                                                          double* P) {
    const int c = blockIdx.x * kAdjBlock + threadIdx.x;
    if (c >= m.nc) return;
    const int64_t ld = m.ld;
    double d = DBL_MAX;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int2 e = m.ell[k * ld + c];
        if (e.x < 0) continue;
        const int f = e.x >> 1;
        const double nx = m.fnx[f], ny = m.fny[f];
        auto wave = [&](int i) {
            const State u = st4(U[i]);
            const double vn = fdiv(u.ru, u.r) * nx + fdiv(u.rv, u.r) * ny;
            return fabs(vn) + sound_speed(u, gamma);
        };
        const int l = m.fl[f], r = m.fr[f];
        double smax = wave(l), len;
        if (r >= 0) {
            smax = fmax(smax, wave(r));
            len = fdiv(fmin(m.vol[l], m.vol[r]), m.fa[f]);
        } else {
            len = fdiv(m.vol[l], m.fa[f]);
        }
        d = fmin(d, fdiv(cfl * len, smax));
    }
    P[c] = fdiv(d, m.vol[c]);
}

// one SSP-RK stage: lam <- a lam0 + b (lam + P (ubar + g)); with partials,
// per-block sums of |ubar + g|^2 (the adjoint residual of the stage's lam)
__global__ __launch_bounds__(kAdjBlock) void k_adj_rk(int nc, const double4* lam0, double4* lam, const double4* ubar,
                                                     const double4* g, const double* P, double a, double b,
                                                     double* partials) {
    __shared__ double red[kAdjBlock];
    const int c = blockIdx.x * kAdjBlock + threadIdx.x;
    double sq = 0.0;
    if (c < nc) {
        const double4 u = ubar[c], gg = g[c], l = lam[c], l0 = lam0[c];
        const double4 r{u.x + gg.x, u.y + gg.y, u.z + gg.z, u.w + gg.w};
        const double pc = P[c];
        lam[c] = double4{a * l0.x + b * (l.x + pc * r.x), a * l0.y + b * (l.y + pc * r.y),
                         a * l0.z + b * (l.z + pc * r.z), a * l0.w + b * (l.w + pc * r.w)};
        sq = r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w;
    }
    if (!partials) return;
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s2 = kAdjBlock / 2; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// k_adj_gather (state only) and k_adj_rk in one pass, for the adjoint
// solve: the same operations in the same order, without the U_bar round
// trip through memory.  copy (the first stage): lam0 takes lam first, as the
// solve's copy before the stage did.
template <int D>
__global__ __launch_bounds__(kAdjBlock) void k_adj_gather_rk(AdjArgs a, double4* lam0, double4* lam, const double4* g,
                                                             const double* P, double ca, double cb, int copy,
                                                             double* partials) {
    __shared__ double red[kAdjBlock];
    const int c = blockIdx.x * kAdjBlock + threadIdx.x;
    double sq = 0.0;
    if (c < a.m.nc) {
        double cbx = 0.0, cby = 0.0;
        const State Ub = adj_gather_state<D, false>(a, c, cbx, cby);
        const double4 u = d4(Ub), gg = g[c], l = lam[c];
        const double4 l0 = copy ? l : lam0[c];
        if (copy) lam0[c] = l;
        const double4 r{u.x + gg.x, u.y + gg.y, u.z + gg.z, u.w + gg.w};
        const double pc = P[c];
        lam[c] = double4{ca * l0.x + cb * (l.x + pc * r.x), ca * l0.y + cb * (l.y + pc * r.y),
                         ca * l0.z + cb * (l.z + pc * r.z), ca * l0.w + cb * (l.w + pc * r.w)};
        sq = r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w;
    }
    if (!partials) return;
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s2 = kAdjBlock / 2; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// one pseudo-time stage of the adjoint solve: the state-only face and cell
// passes, then the gather fused with the update of lam
int run_stage_rk(fvk_solver* s, double4* lam0, const double4* g, const double* P, double ca, double cb, bool copy,
                 double* partials) {
    AdjointWork* w = s->adj;
    const AdjArgs a = make_args(s);
    hipEvent_t ev = nullptr;
    timer_begin(s, "adj_face", &ev);
    if (a.order == 2) launch_faces<2>(s, a, false);
    else launch_faces<1>(s, a, false);
    timer_end(s, ev);
    const int gc = grid_of(s->nc);
    timer_begin(s, "adj_cell", &ev);
    switch (w->D) {
        case 3: k_adj_cell<3, false><<<gc, kAdjBlock, 0, s->stream>>>(a); break;
        case 4: k_adj_cell<4, false><<<gc, kAdjBlock, 0, s->stream>>>(a); break;
        case 9: k_adj_cell<9, false><<<gc, kAdjBlock, 0, s->stream>>>(a); break;
        default: k_adj_cell<kMaxEulerDegree, false><<<gc, kAdjBlock, 0, s->stream>>>(a); break;
    }
    timer_end(s, ev);
    timer_begin(s, "adj_gather_rk", &ev);
    const int cp = copy ? 1 : 0;
    switch (w->D) {
        case 3: k_adj_gather_rk<3><<<gc, kAdjBlock, 0, s->stream>>>(a, lam0, w->lam, g, P, ca, cb, cp, partials); break;
        case 4: k_adj_gather_rk<4><<<gc, kAdjBlock, 0, s->stream>>>(a, lam0, w->lam, g, P, ca, cb, cp, partials); break;
        case 9: k_adj_gather_rk<9><<<gc, kAdjBlock, 0, s->stream>>>(a, lam0, w->lam, g, P, ca, cb, cp, partials); break;
        default:
            k_adj_gather_rk<kMaxEulerDegree><<<gc, kAdjBlock, 0, s->stream>>>(a, lam0, w->lam, g, P, ca, cb, cp,
                                                                           partials);
    }
    timer_end(s, ev);
    return hip_check(hipGetLastError(), "adjoint solve launch");
}

}  // namespace

int adjoint_vjp(fvk_solver* s, const double* lambda, double* U_bar, double* X_bar) {
    const bool wx = X_bar != nullptr;
    int rc = check_supported(s, wx);
    if (rc != FVK_OK) return rc;
    rc = ensure_work(s, wx);
    if (rc != FVK_OK) return rc;
    AdjointWork* w = s->adj;
    rc = upload_cells(s, w->lam, lambda);
    if (rc != FVK_OK) return rc;
    if (s->p.order == 2) {
        rc = launch_euler_gradients(s);
        if (rc != FVK_OK) return rc;
    }
    rc = run_passes(s, wx);
    if (rc != FVK_OK) return rc;
    rc = download_cells(s, U_bar, w->ubar);
    if (rc == FVK_OK && wx) rc = download_nodes(s, X_bar);
    return rc;
}

int adjoint_solve(fvk_solver* s, const double* g_host, double cfl, int max_iter, double tol, int check_every,
                  double* lambda_out, double* X_bar, fvk_adjoint_report* rep) {
    const bool wx = X_bar != nullptr;
    int rc = check_supported(s, wx);
    if (rc != FVK_OK) return rc;
    if (max_iter < 0 || !(cfl > 0.0) || check_every < 1) return set_error(FVK_ERR_INVALID, "adjoint solve arguments");
    rc = ensure_work(s, wx);
    if (rc != FVK_OK) return rc;
    AdjointWork* w = s->adj;
    const int64_t nc = s->nc, ld = s->mesh.ld;
    const int gc = grid_of(nc);
    if (!w->sg) {
        w->sg = alloc_n<double4>(s, ld);
        w->slam0 = alloc_n<double4>(s, ld);
        w->sP = alloc_n<double>(s, ld);
        w->spart = alloc_n<double>(s, gc);
        if (!w->sg || !w->slam0 || !w->sP || !w->spart)
            return set_error(FVK_ERR_NOMEM, "adjoint solve: device allocation failed");
    }
    double4 *g = w->sg, *lam0 = w->slam0;
    double *P = w->sP, *part = w->spart;
    rc = upload_cells(s, g, g_host);
    if (rc == FVK_OK) rc = hip_check(hipMemsetAsync(w->lam, 0, sizeof(double4) * ld, s->stream), "adjoint init");
    if (rc == FVK_OK && s->p.order == 2) rc = launch_euler_gradients(s);
    if (rc != FVK_OK) return rc;
    const double4* U = reinterpret_cast<const double4*>(s->U[s->cur]);
    switch (w->D) {
        case 3: k_adj_precond<3><<<gc, kAdjBlock, 0, s->stream>>>(s->mesh, U, s->p.gamma, cfl, P); break;
        case 4: k_adj_precond<4><<<gc, kAdjBlock, 0, s->stream>>>(s->mesh, U, s->p.gamma, cfl, P); break;
        case 9: k_adj_precond<9><<<gc, kAdjBlock, 0, s->stream>>>(s->mesh, U, s->p.gamma, cfl, P); break;
        default: k_adj_precond<kMaxEulerDegree><<<gc, kAdjBlock, 0, s->stream>>>(s->mesh, U, s->p.gamma, cfl, P);
    }
    rc = hip_check(hipGetLastError(), "adjoint solve launch");
    if (rc != FVK_OK) return rc;
    // SSP-RK3 in pseudo time (the primal's integrator and local step)
    const double A[3] = {0.0, 0.75, 1.0 / 3.0}, B[3] = {1.0, 0.25, 2.0 / 3.0};
    std::vector<double> hp(gc);
    double r0 = -1.0, rn = -1.0;
    int it = 0;
    for (; it < max_iter; ++it) {
        const bool check = it % check_every == 0 || it + 1 == max_iter;
        for (int st = 0; st < 3 && rc == FVK_OK; ++st)
            rc = run_stage_rk(s, lam0, g, P, A[st], B[st], st == 0, st == 0 && check ? part : nullptr);
        if (rc != FVK_OK) return rc;
        if (check) {
            rc = hip_check(hipMemcpyAsync(hp.data(), part, sizeof(double) * gc, hipMemcpyDeviceToHost, s->stream),
                           "adjoint solve residual");
            if (rc == FVK_OK) rc = hip_check(hipStreamSynchronize(s->stream), "adjoint solve");
            if (rc != FVK_OK) return rc;
            double sq = 0.0;
            for (double v : hp) sq += v;
            rn = std::sqrt(sq);
            if (r0 < 0.0) r0 = rn;
            if (!std::isfinite(rn)) return set_error(FVK_ERR_DIVERGED, "adjoint solve diverged (non-finite residual)");
            if (rn <= tol * r0) { ++it; break; }
        }
    }
    // the residual of the final lambda, and its X adjoint
    rc = run_passes(s, wx);
    if (rc != FVK_OK) return rc;
    k_adj_rk<<<gc, kAdjBlock, 0, s->stream>>>((int)nc, w->lam, lam0, w->ubar, g, P, 1.0, 0.0, part);
    rc = hip_check(hipGetLastError(), "adjoint solve launch");
    if (rc == FVK_OK)
        rc = hip_check(hipMemcpyAsync(hp.data(), part, sizeof(double) * gc, hipMemcpyDeviceToHost, s->stream),
                       "adjoint solve residual");
    if (rc == FVK_OK) rc = hip_check(hipStreamSynchronize(s->stream), "adjoint solve");
    if (rc != FVK_OK) return rc;
    double sq = 0.0;
    for (double v : hp) sq += v;
    rn = std::sqrt(sq);
    if (rep) {
        rep->iterations = it;
        rep->residual_first = r0 < 0.0 ? rn : r0;
        rep->residual_final = rn;
    }
    if (lambda_out) rc = download_cells(s, lambda_out, w->lam);
    if (rc == FVK_OK && wx) rc = download_nodes(s, X_bar);
    return rc;
}

}  // namespace fvk
