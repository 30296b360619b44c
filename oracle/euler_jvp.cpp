// TEST INFRASTRUCTURE ONLY (the oracle): forward-mode derivative (Jacobian-
// vector product) of the Euler residual R(U, X), for checking the library's
// hand-written discrete adjoint by the dot-product identity
// <lambda, J v> = <J^T lambda, v>.  The product never links or loads it.
//
// The residual is restated once more, generic in its number type, and run on
// dual numbers (value, derivative): the derivative along (vU, vX) is then
// exact to rounding for the branch each value takes.  It follows the same
// path as the oracle's fvoracle.c: first-order faces as
// src/EulerFVMSolver.cpp:61-104 (Rusanov include/RusanovFlux.h:33-43, HLLC
// include/HllcFlux.h:35-76, ghost states src/EulerFVMSolver.cpp:174-193), and
// the second-order extension (LSQ gradients of rho, u, v, p with the
// boundary neighbour at the face midpoint, the Venkatakrishnan limiter with
// its minimum by cross-multiplication, MUSCL primitive face states with the
// cell-state fallback, and the fluxes on primitive states).  Geometry is a function of the node coordinates X: face length,
// normal and midpoint from the face's nodes (include/UnstructuredMesh.h:
// 89-105), cell area and centroid from the shoelace formulas over the cell's
// polygon (src/MeshReader.cpp:32-76), evaluated relative to its first
// vertex.
#include <cmath>
#include <cstdint>
#include <vector>

#include "fvk.h"

namespace {

// a value and its derivative, in a real type F (double, or long double for
// an estimate of the double evaluation's own rounding)
template <class F> struct DualT {
    F v = 0.0, d = 0.0;
    DualT() = default;
    DualT(F x) : v(x), d(0.0) {}  // NOLINT: constants promote
    DualT(F x, F dx) : v(x), d(dx) {}
};
template <class F> inline DualT<F> operator+(DualT<F> a, DualT<F> b) { return {a.v + b.v, a.d + b.d}; }
template <class F> inline DualT<F> operator-(DualT<F> a, DualT<F> b) { return {a.v - b.v, a.d - b.d}; }
template <class F> inline DualT<F> operator-(DualT<F> a) { return {-a.v, -a.d}; }
template <class F> inline DualT<F> operator*(DualT<F> a, DualT<F> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <class F> inline DualT<F> operator/(DualT<F> a, DualT<F> b) {
    const F q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
// mixed forms with a double constant
template <class F> inline DualT<F> operator+(double a, DualT<F> b) { return DualT<F>(a) + b; }
template <class F> inline DualT<F> operator-(double a, DualT<F> b) { return DualT<F>(a) - b; }
template <class F> inline DualT<F> operator*(double a, DualT<F> b) { return DualT<F>(a) * b; }
template <class F> inline DualT<F> operator/(double a, DualT<F> b) { return DualT<F>(a) / b; }
template <class F> inline DualT<F> operator+(DualT<F> a, double b) { return a + DualT<F>(b); }
template <class F> inline DualT<F> operator-(DualT<F> a, double b) { return a - DualT<F>(b); }
template <class F> inline DualT<F> operator*(DualT<F> a, double b) { return a * DualT<F>(b); }
template <class F> inline DualT<F> operator/(DualT<F> a, double b) { return a / DualT<F>(b); }
template <class F> inline DualT<F>& operator+=(DualT<F>& a, DualT<F> b) { return a = a + b; }
template <class F> inline bool operator<(DualT<F> a, DualT<F> b) { return a.v < b.v; }
template <class F> inline bool operator>(DualT<F> a, DualT<F> b) { return a.v > b.v; }
template <class F> inline bool operator<=(DualT<F> a, DualT<F> b) { return a.v <= b.v; }
template <class F> inline bool operator>=(DualT<F> a, DualT<F> b) { return a.v >= b.v; }
template <class F> inline bool operator==(DualT<F> a, DualT<F> b) { return a.v == b.v; }
template <class F> inline bool operator!=(DualT<F> a, DualT<F> b) { return a.v != b.v; }
template <class F> inline DualT<F> sqrt(DualT<F> a) {
    using std::sqrt;
    const F r = sqrt(a.v);
    return {r, a.d / (F(2.0) * r)};
}
template <class F> inline DualT<F> fabs(DualT<F> a) { return a.v < F(0.0) ? -a : a; }

// std::min / std::max semantics (the first argument on a tie)
template <class T> inline T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> inline T smax(T a, T b) { return (a < b) ? b : a; }

template <class T> struct S4 { T r, ru, rv, E; };

template <class T> T pressure(const S4<T>& s, double g) {
    T u = s.ru / s.r, v = s.rv / s.r;
    return (g - 1.0) * (s.E - 0.5 * s.r * (u * u + v * v));
}
template <class T> S4<T> from_prim(T rho, T u, T v, T p, double g) {
    return {rho, rho * u, rho * v, p / (g - 1.0) + 0.5 * rho * (u * u + v * v)};
}
template <class T> T sound(const S4<T>& s, double g) { return sqrt(g * pressure(s, g) / s.r); }
template <class T> S4<T> phys_flux(const S4<T>& s, T nx, T ny, double g) {
    T u = s.ru / s.r, v = s.rv / s.r;
    T Vn = u * nx + v * ny;
    T p = pressure(s, g);
    return {s.r * Vn, s.ru * Vn + p * nx, s.rv * Vn + p * ny, (s.E + p) * Vn};
}
template <class T> S4<T> rusanov(const S4<T>& L, const S4<T>& R, T nx, T ny, double g) {
    T VnL = (L.ru * nx + L.rv * ny) / L.r, VnR = (R.ru * nx + R.rv * ny) / R.r;
    T Smax = smax(fabs(VnL) + sound(L, g), fabs(VnR) + sound(R, g));
    S4<T> FL = phys_flux(L, nx, ny, g), FR = phys_flux(R, nx, ny, g);
    T h = 0.5 * Smax;
    return {0.5 * (FL.r + FR.r) - h * (R.r - L.r), 0.5 * (FL.ru + FR.ru) - h * (R.ru - L.ru),
            0.5 * (FL.rv + FR.rv) - h * (R.rv - L.rv), 0.5 * (FL.E + FR.E) - h * (R.E - L.E)};
}
template <class T> S4<T> hllc(const S4<T>& L, const S4<T>& R, T nx, T ny, double g) {
    T tx = -ny, ty = nx;
    T rL = L.r, rR = R.r;
    T uL = L.ru / rL, vL = L.rv / rL, uR = R.ru / rR, vR = R.rv / rR;
    T VnL = uL * nx + vL * ny, VnR = uR * nx + vR * ny;
    T VtL = uL * tx + vL * ty, VtR = uR * tx + vR * ty;
    T pL = pressure(L, g), pR = pressure(R, g);
    T cL = sound(L, g), cR = sound(R, g);
    T SL = smin(VnL - cL, VnR - cR), SR = smax(VnL + cL, VnR + cR);
    if (SL >= T(0.0)) return phys_flux(L, nx, ny, g);
    if (SR <= T(0.0)) return phys_flux(R, nx, ny, g);
    T Ss = (pR - pL + rL * VnL * (SL - VnL) - rR * VnR * (SR - VnR)) / (rL * (SL - VnL) - rR * (SR - VnR));
    const bool left = Ss >= T(0.0);
    T SK = left ? SL : SR, rK = left ? rL : rR, VnK = left ? VnL : VnR, VtK = left ? VtL : VtR, pK = left ? pL : pR;
    const S4<T>& UK = left ? L : R;
    T factor = rK * (SK - VnK) / (SK - Ss);
    T mn = factor * Ss, mt = factor * VtK;
    T Es = factor * (UK.E / rK + (Ss - VnK) * (Ss + pK / (rK * (SK - VnK))));
    S4<T> F = phys_flux(UK, nx, ny, g);
    return {F.r + SK * (factor - UK.r), F.ru + SK * ((mn * nx - mt * ny) - UK.ru),
            F.rv + SK * ((mn * ny + mt * nx) - UK.rv), F.E + SK * (Es - UK.E)};
}

// second order: primitive face states W = (rho, u, v, p) and the fluxes on
// them (fvoracle.c's hllc_w / rusanov_w, the same operations)
template <class T> struct P4 { T r, u, v, p; };
template <class T> T specific_energy(const P4<T>& w, double g) {
    return w.p / ((g - 1.0) * w.r) + 0.5 * (w.u * w.u + w.v * w.v);
}
template <class T> S4<T> phys_flux_w(const P4<T>& w, T E, T Vn, T nx, T ny) {
    T m = w.r * Vn;
    return {m, m * w.u + w.p * nx, m * w.v + w.p * ny, (E + w.p) * Vn};
}
template <class T> S4<T> rusanov_w(const P4<T>& L, const P4<T>& R, T nx, T ny, double g) {
    T VnL = L.u * nx + L.v * ny, VnR = R.u * nx + R.v * ny;
    T Smax = smax(fabs(VnL) + sqrt(g * L.p / L.r), fabs(VnR) + sqrt(g * R.p / R.r));
    T EL = L.r * specific_energy(L, g), ER = R.r * specific_energy(R, g);
    S4<T> FL = phys_flux_w(L, EL, VnL, nx, ny), FR = phys_flux_w(R, ER, VnR, nx, ny);
    T h = 0.5 * Smax;
    return {0.5 * (FL.r + FR.r) - h * (R.r - L.r), 0.5 * (FL.ru + FR.ru) - h * (R.r * R.u - L.r * L.u),
            0.5 * (FL.rv + FR.rv) - h * (R.r * R.v - L.r * L.v), 0.5 * (FL.E + FR.E) - h * (ER - EL)};
}
template <class T> S4<T> hllc_w(const P4<T>& L, const P4<T>& R, T nx, T ny, double g) {
    T VnL = L.u * nx + L.v * ny, VnR = R.u * nx + R.v * ny;
    T cL = sqrt(g * L.p / L.r), cR = sqrt(g * R.p / R.r);
    T SL = smin(VnL - cL, VnR - cR), SR = smax(VnL + cL, VnR + cR);
    if (SL >= T(0.0)) return phys_flux_w(L, L.r * specific_energy(L, g), VnL, nx, ny);
    if (SR <= T(0.0)) return phys_flux_w(R, R.r * specific_energy(R, g), VnR, nx, ny);
    T mL = L.r * (SL - VnL), mR = R.r * (SR - VnR);
    T Ss = (R.p - L.p + mL * VnL - mR * VnR) / (mL - mR);
    const bool left = Ss >= T(0.0);
    const P4<T>& K = left ? L : R;
    T SK = left ? SL : SR, VnK = left ? VnL : VnR, mK = left ? mL : mR;
    T VtK = K.v * nx - K.u * ny;
    T eK = specific_energy(K, g), EK = K.r * eK;
    T factor = mK / (SK - Ss);
    T mn = factor * Ss, mt = factor * VtK;
    T Es = factor * (eK + (Ss - VnK) * (Ss + K.p / mK));
    S4<T> F = phys_flux_w(K, EK, VnK, nx, ny);
    return {F.r + SK * (factor - K.r), F.ru + SK * ((mn * nx - mt * ny) - K.r * K.u),
            F.rv + SK * ((mn * ny + mt * nx) - K.r * K.v), F.E + SK * (Es - EK)};
}

struct Mesh {
    int64_t nn, nc, nf;
    const int32_t *fnode, *fcell, *fpatch, *cptr, *cface, *nptr, *nodes;
    const fvk_bc* bcs;
    const fvk_params* p;
};

template <class T> S4<T> ghost(const Mesh& m, int64_t f, const S4<T>& L, T nx, T ny) {
    const fvk_bc& bc = m.bcs[m.fpatch[f]];
    const double g = m.p->gamma;
    if (bc.type == FVK_BC_FARFIELD)
        return {T(bc.farfield_state[0]), T(bc.farfield_state[1]), T(bc.farfield_state[2]), T(bc.farfield_state[3])};
    if (bc.type == FVK_BC_OUTFLOW) return L;
    T u = L.ru / L.r, v = L.rv / L.r;
    T Vn = u * nx + v * ny;
    return from_prim(L.r, u - 2.0 * Vn * nx, v - 2.0 * Vn * ny, pressure(L, g), g);
}

template <class T> void prim(const S4<T>& s, double g, T* w) {
    w[0] = s.r;
    w[1] = s.ru / s.r;
    w[2] = s.rv / s.r;
    w[3] = pressure(s, g);
}

template <class T> P4<T> ghost_w(const Mesh& m, int64_t f, const P4<T>& L, T nx, T ny) {
    const fvk_bc& bc = m.bcs[m.fpatch[f]];
    if (bc.type == FVK_BC_FARFIELD) {
        const S4<T> ff{T(bc.farfield_state[0]), T(bc.farfield_state[1]), T(bc.farfield_state[2]),
                       T(bc.farfield_state[3])};
        T w[4];
        prim(ff, m.p->gamma, w);
        return {w[0], w[1], w[2], w[3]};
    }
    if (bc.type == FVK_BC_OUTFLOW) return L;
    T Vn = L.u * nx + L.v * ny;
    return {L.r, L.u - 2.0 * Vn * nx, L.v - 2.0 * Vn * ny, L.p};
}

template <class T> void residual(const Mesh& m, const T* U, const T* X, T* R) {
    const fvk_params& P = *m.p;
    const double g = P.gamma;
    const int64_t nc = m.nc, nf = m.nf;
    // face geometry
    std::vector<T> fnx(nf), fny(nf), fa(nf), fmx(nf), fmy(nf);
    for (int64_t f = 0; f < nf; ++f) {
        const T* a = X + 2 * m.fnode[2 * f];
        const T* b = X + 2 * m.fnode[2 * f + 1];
        fmx[f] = 0.5 * (a[0] + b[0]);
        fmy[f] = 0.5 * (a[1] + b[1]);
        T dx = b[0] - a[0], dy = b[1] - a[1];
        fa[f] = sqrt(dx * dx + dy * dy);
        fnx[f] = dy / fa[f];
        fny[f] = -dx / fa[f];
    }
    // cell polygon: shoelace area and centroid relative to the first vertex
    std::vector<T> cx(nc), cy(nc), vol(nc);
    for (int64_t c = 0; c < nc; ++c) {
        const int32_t b0 = m.nptr[c], n = m.nptr[c + 1] - b0;
        const T* v0 = X + 2 * m.nodes[b0];
        T A2 = 0.0, Sx = 0.0, Sy = 0.0;
        for (int i = 0; i < n; ++i) {
            const T* pi = X + 2 * m.nodes[b0 + i];
            const T* pj = X + 2 * m.nodes[b0 + (i + 1) % n];
            T xi = pi[0] - v0[0], yi = pi[1] - v0[1], xj = pj[0] - v0[0], yj = pj[1] - v0[1];
            T cr = xi * yj - xj * yi;
            A2 += cr;
            Sx += (xi + xj) * cr;
            Sy += (yi + yj) * cr;
        }
        cx[c] = v0[0] + Sx / (3.0 * A2);
        cy[c] = v0[1] + Sy / (3.0 * A2);
        vol[c] = fabs(0.5 * A2);
    }
    auto state = [&](int64_t c) { return S4<T>{U[4 * c], U[4 * c + 1], U[4 * c + 2], U[4 * c + 3]}; };
    std::vector<T> G((size_t)8 * nc, T(0.0));
    if (P.order == 2) {
        for (int64_t c = 0; c < nc; ++c) {
            const S4<T> Uc = state(c);
            T wc[4];
            prim(Uc, g, wc);
            T bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0}, wmin[4], wmax[4];
            T m00 = 0.0, m01 = 0.0, m11 = 0.0;
            for (int q = 0; q < 4; ++q) wmin[q] = wmax[q] = wc[q];
            for (int k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
                const int f = m.cface[k];
                const int cl = m.fcell[2 * f], cr = m.fcell[2 * f + 1];
                T wn[4], qx, qy;
                if (cr == -1) {
                    T wg[4];
                    prim(ghost(m, f, Uc, fnx[f], fny[f]), g, wg);
                    for (int q = 0; q < 4; ++q) wn[q] = 0.5 * (wc[q] + wg[q]);
                    qx = fmx[f];
                    qy = fmy[f];
                } else {
                    const int o = cl == (int)c ? cr : cl;
                    prim(state(o), g, wn);
                    qx = cx[o];
                    qy = cy[o];
                }
                T dx = qx - cx[c], dy = qy - cy[c];
                T w = 1.0 / (dx * dx + dy * dy);
                m00 += w * dx * dx;
                m01 += w * dx * dy;
                m11 += w * dy * dy;
                for (int q = 0; q < 4; ++q) {
                    T dphi = wn[q] - wc[q];
                    bx[q] += w * dx * dphi;
                    by[q] += w * dy * dphi;
                    wmin[q] = smin(wmin[q], wn[q]);
                    wmax[q] = smax(wmax[q], wn[q]);
                }
            }
            T det = m00 * m11 - m01 * m01;
            T i0 = m11 / det, i1 = -m01 / det, i2 = m00 / det;
            T gx[4], gy[4], lim[4];
            for (int q = 0; q < 4; ++q) {
                gx[q] = i0 * bx[q] + i1 * by[q];
                gy[q] = i1 * bx[q] + i2 * by[q];
                lim[q] = 1.0;
            }
            if (P.limiter == FVK_LIMITER_VENKATAKRISHNAN) {
                T kh = P.venkat_k * sqrt(vol[c]);
                T eps2 = kh * kh * kh;
                T bn[4] = {1.0, 1.0, 1.0, 1.0}, bd[4] = {1.0, 1.0, 1.0, 1.0};
                for (int k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
                    const int f = m.cface[k];
                    T rx = fmx[f] - cx[c], ry = fmy[f] - cy[c];
                    for (int q = 0; q < 4; ++q) {
                        T d2 = gx[q] * rx + gy[q] * ry;
                        T d1 = (d2 > T(0.0) ? wmax[q] : wmin[q]) - wc[q];
                        T d11 = d1 * d1;
                        T num = (d11 + eps2) + 2.0 * d1 * d2;
                        T den = d11 + 2.0 * d2 * d2 + d1 * d2 + eps2;
                        if (d2 != T(0.0) && num * bd[q] < bn[q] * den) {
                            bn[q] = num;
                            bd[q] = den;
                        }
                    }
                }
                for (int q = 0; q < 4; ++q) lim[q] = bn[q] / bd[q];
            } else if (P.limiter == FVK_LIMITER_FIRST_ORDER) {
                for (int q = 0; q < 4; ++q) lim[q] = 0.0;
            }
            for (int q = 0; q < 4; ++q) {
                G[8 * c + 2 * q] = lim[q] * gx[q];
                G[8 * c + 2 * q + 1] = lim[q] * gy[q];
            }
        }
    }
    auto recon = [&](int64_t c, T rx, T ry) {
        const T* gr = &G[8 * c];
        T d0 = gr[0] * rx + gr[1] * ry, d1 = gr[2] * rx + gr[3] * ry;
        T d2 = gr[4] * rx + gr[5] * ry, d3 = gr[6] * rx + gr[7] * ry;
        T w[4];
        prim(state(c), g, w);
        T rho = w[0] + d0, pp = w[3] + d3;
        if (!(rho > T(0.0)) || !(pp > T(0.0))) return P4<T>{w[0], w[1], w[2], w[3]};
        return P4<T>{rho, w[1] + d1, w[2] + d2, pp};
    };
    std::vector<S4<T>> F(nf);
    for (int64_t f = 0; f < nf; ++f) {
        const int cl = m.fcell[2 * f], cr = m.fcell[2 * f + 1];
        S4<T> Fs;
        if (P.order == 2) {
            const P4<T> L = recon(cl, fmx[f] - cx[cl], fmy[f] - cy[cl]);
            const P4<T> Rs = cr != -1 ? recon(cr, fmx[f] - cx[cr], fmy[f] - cy[cr]) : ghost_w(m, f, L, fnx[f], fny[f]);
            Fs = P.flux_scheme == FVK_FLUX_HLLC ? hllc_w(L, Rs, fnx[f], fny[f], g) : rusanov_w(L, Rs, fnx[f], fny[f], g);
        } else {
            const S4<T> L = state(cl);
            const S4<T> Rs = cr != -1 ? state(cr) : ghost(m, f, L, fnx[f], fny[f]);
            Fs = P.flux_scheme == FVK_FLUX_HLLC ? hllc(L, Rs, fnx[f], fny[f], g) : rusanov(L, Rs, fnx[f], fny[f], g);
        }
        const T A = fa[f];
        F[f] = {A * Fs.r, A * Fs.ru, A * Fs.rv, A * Fs.E};
    }
    for (int64_t c = 0; c < nc; ++c) {
        T r[4] = {0, 0, 0, 0};
        for (int k = m.cptr[c]; k < m.cptr[c + 1]; ++k) {
            const int f = m.cface[k];
            const S4<T>& o = F[f];
            if (m.fcell[2 * f] == (int)c) {
                r[0] = r[0] - o.r; r[1] = r[1] - o.ru; r[2] = r[2] - o.rv; r[3] = r[3] - o.E;
            } else {
                r[0] += o.r; r[1] += o.ru; r[2] += o.rv; r[3] += o.E;
            }
        }
        for (int q = 0; q < 4; ++q) R[4 * c + q] = r[q];
    }
}

template <class F>
int jvp(const fvk_mesh_desc* d, const int32_t* node_ptr, const int32_t* nodes, const fvk_bc* bcs, const fvk_params* p,
        const double* U, const double* X, const double* vU, const double* vX, double* R, double* JV) {
    if (p->equation != FVK_EQ_EULER || p->flux_scheme == FVK_FLUX_EXACT || d->face_shift) return FVK_ERR_UNSUPPORTED;
    using D = DualT<F>;
    Mesh m{d->n_nodes, d->n_cells, d->n_faces, d->face_nodes, d->face_cells, d->face_patch, d->cell_face_ptr,
           d->cell_faces, node_ptr, nodes, bcs, p};
    std::vector<D> u((size_t)4 * m.nc), x((size_t)2 * m.nn), r((size_t)4 * m.nc);
    for (int64_t i = 0; i < 4 * m.nc; ++i) u[i] = D(U[i], vU ? vU[i] : 0.0);
    for (int64_t i = 0; i < 2 * m.nn; ++i) x[i] = D(X[i], vX ? vX[i] : 0.0);
    residual<D>(m, u.data(), x.data(), r.data());
    for (int64_t i = 0; i < 4 * m.nc; ++i) {
        if (R) R[i] = (double)r[i].v;
        JV[i] = (double)r[i].d;
    }
    return FVK_OK;
}

}  // namespace

extern "C" {

// R(U, X) and its derivative along (vU, vX): R [4 nc], JV [4 nc].  The mesh
// arrays are fvk_mesh_desc's plus the cell polygons (node_ptr [nc+1],
// nodes); Euler only, Rusanov or HLLC, no periodic faces.

// R(U, X) and its derivative along (vU, vX): R [4 nc], JV [4 nc].  The mesh
// arrays are fvk_mesh_desc's plus the cell polygons (node_ptr [nc+1],
// nodes); Euler only, Rusanov or HLLC, no periodic faces.
int fvj_residual_jvp(const fvk_mesh_desc* d, const int32_t* node_ptr, const int32_t* nodes, const fvk_bc* bcs,
                     const fvk_params* p, const double* U, const double* X, const double* vU, const double* vX,
                     double* R, double* JV) {
    return jvp<double>(d, node_ptr, nodes, bcs, p, U, X, vU, vX, R, JV);
}

// The same evaluated in long double (64-bit significand): its difference
// from fvj_residual_jvp estimates the rounding of the double evaluation,
// which is large where the residual is ill-conditioned (the LSQ normal
// matrix of a thin cell whose stencil points are nearly collinear).
int fvj_residual_jvp_ld(const fvk_mesh_desc* d, const int32_t* node_ptr, const int32_t* nodes, const fvk_bc* bcs,
                        const fvk_params* p, const double* U, const double* X, const double* vU, const double* vX,
                        double* R, double* JV) {
    return jvp<long double>(d, node_ptr, nodes, bcs, p, U, X, vU, vX, R, JV);
}

}  // extern "C"
