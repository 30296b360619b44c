/*
 * fvoracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).  See fvoracle.h.
 *
 * CPU restatement of happy12/FiniteVolume's explicit residual path in plain
 * C (OpenMP over faces and cells).  It follows the reference's order of
 * floating-point operations so that, built with -ffp-contract=off and
 * without -ffast-math, it reproduces the reference bit for bit; each
 * function cites what it restates.  The extensions (second order MUSCL +
 * Venkatakrishnan limiter, SSP Runge-Kutta, local time stepping, periodic
 * face pairs, adjoint objective) have no reference code: they are specified
 * here, and the HIP library implements the same specification.
 *
 * Parity pins (tests/): against the reference built from its own sources
 * (oracle/_ref) and the golden fixtures under tests/golden; for the
 * extensions, against their reductions to the reference path and closed
 * forms.
 */
#include "fvoracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double r, ru, rv, E;
} st4;

typedef struct {
    double x, y;
} vec2;

struct fvo_solver {
    fvk_params p;
    int64_t nn, nc, nf;
    int32_t n_bcs;
    fvk_bc* bcs;
    /* mesh */
    double* node_xy;
    int32_t *fnode, *fcell, *fpatch, *cptr, *cface;
    double *vol, *cxy, *shift;
    /* geometry (UnstructuredMesh::compute_geometry) */
    double *fmid, *farea, *fnx, *fny;
    /* gradient precompute (GradientCalculator ctor) */
    double* lsq_inv; /* 3 per cell */
    double* wall_dist;
    /* state */
    double* U;    /* AoS 4 per cell */
    double* turb; /* nut (SA) or k,omega (SST), AoS */
    double res[6];
    /* scratch for the extended path */
    double* grad8; /* limited primitive gradients 8 per cell */
    double* lim4;  /* limiter values 4 per cell */
};

/* std::min / std::max semantics: min(a,b) = (b < a) ? b : a; max(a,b) = (a < b) ? b : a */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* the two scalar-field solvers (UnstructuredFVMSolver, AdvectionDiffusionFVMSolver) */
static int scalar_of(int eq) { return eq == FVK_EQ_DIFFUSION || eq == FVK_EQ_ADVECTION_DIFFUSION; }

static int nvars_of(int eq) {
    return eq == FVK_EQ_RANS_SA ? 5 : (eq == FVK_EQ_RANS_SST ? 6 : (scalar_of(eq) ? 1 : 4));
}

/* ------------------------------------------------------------------------
 * EulerState physics: include/EulerState.h:45-101
 * ---------------------------------------------------------------------- */
static inline st4 from_primitive(double rho, double u, double v, double p, double g) {
    st4 s;
    s.r = rho;
    s.ru = rho * u;
    s.rv = rho * v;
    s.E = p / (g - 1.0) + 0.5 * rho * (u * u + v * v);
    return s;
}
static inline double pressure(st4 s, double g) {
    double u = s.ru / s.r, v = s.rv / s.r;
    return (g - 1.0) * (s.E - 0.5 * s.r * (u * u + v * v));
}
static inline double temperature(st4 s, double g, double R) { return pressure(s, g) / (s.r * R); }
static inline double sound_speed(st4 s, double g) { return sqrt(g * pressure(s, g) / s.r); }
static inline st4 phys_flux(st4 s, double nx, double ny, double g) {
    double u = s.ru / s.r, v = s.rv / s.r;
    double Vn = u * nx + v * ny;
    double p = pressure(s, g);
    st4 f;
    f.r = s.r * Vn;
    f.ru = s.ru * Vn + p * nx;
    f.rv = s.rv * Vn + p * ny;
    f.E = (s.E + p) * Vn;
    return f;
}
static inline st4 ld4(const double* a) {
    st4 s = {a[0], a[1], a[2], a[3]};
    return s;
}
static inline void sv4(double* a, st4 s) {
    a[0] = s.r; a[1] = s.ru; a[2] = s.rv; a[3] = s.E;
}

/* Rusanov: include/RusanovFlux.h:33-43 */
static st4 rusanov(st4 L, st4 R, double nx, double ny, double g) {
    double VnL = (L.ru * nx + L.rv * ny) / L.r;
    double VnR = (R.ru * nx + R.rv * ny) / R.r;
    double Smax = smax(fabs(VnL) + sound_speed(L, g), fabs(VnR) + sound_speed(R, g));
    st4 FL = phys_flux(L, nx, ny, g), FR = phys_flux(R, nx, ny, g);
    double h = 0.5 * Smax;
    st4 o;
    o.r = 0.5 * (FL.r + FR.r) - h * (R.r - L.r);
    o.ru = 0.5 * (FL.ru + FR.ru) - h * (R.ru - L.ru);
    o.rv = 0.5 * (FL.rv + FR.rv) - h * (R.rv - L.rv);
    o.E = 0.5 * (FL.E + FR.E) - h * (R.E - L.E);
    return o;
}

/* HLLC: include/HllcFlux.h:35-76 (Davis speeds, Toro 10.37-10.39) */
static st4 hllc(st4 L, st4 R, double nx, double ny, double g) {
    double tx = -ny, ty = nx;
    double rL = L.r, rR = R.r;
    double uL = L.ru / rL, vL = L.rv / rL, uR = R.ru / rR, vR = R.rv / rR;
    double VnL = uL * nx + vL * ny, VnR = uR * nx + vR * ny;
    double VtL = uL * tx + vL * ty, VtR = uR * tx + vR * ty;
    double pL = pressure(L, g), pR = pressure(R, g);
    double cL = sound_speed(L, g), cR = sound_speed(R, g);
    double SL = smin(VnL - cL, VnR - cR);
    double SR = smax(VnL + cL, VnR + cR);
    if (SL >= 0.0) return phys_flux(L, nx, ny, g);
    if (SR <= 0.0) return phys_flux(R, nx, ny, g);
    double Ss = (pR - pL + rL * VnL * (SL - VnL) - rR * VnR * (SR - VnR)) / (rL * (SL - VnL) - rR * (SR - VnR));
    int left = Ss >= 0.0;
    double SK = left ? SL : SR, rK = left ? rL : rR, VnK = left ? VnL : VnR, VtK = left ? VtL : VtR;
    double pK = left ? pL : pR;
    st4 UK = left ? L : R;
    double factor = rK * (SK - VnK) / (SK - Ss);
    double mn = factor * Ss, mt = factor * VtK;
    double Es = factor * (UK.E / rK + (Ss - VnK) * (Ss + pK / (rK * (SK - VnK))));
    st4 us = {factor, mn * nx - mt * ny, mn * ny + mt * nx, Es};
    st4 F = phys_flux(UK, nx, ny, g);
    st4 o = {F.r + SK * (us.r - UK.r), F.ru + SK * (us.ru - UK.ru), F.rv + SK * (us.rv - UK.rv),
             F.E + SK * (us.E - UK.E)};
    return o;
}

/* Exact Riemann: include/ExactRiemannFlux.h:28-234 (Toro ch. 4) */
static double pfun(double p, double rK, double pK, double cK, double g) {
    if (p > pK) {
        double A = 2.0 / ((g + 1.0) * rK), B = (g - 1.0) / (g + 1.0) * pK;
        return (p - pK) * sqrt(A / (p + B));
    }
    return (2.0 * cK / (g - 1.0)) * (pow(p / pK, (g - 1.0) / (2.0 * g)) - 1.0);
}
static double pfun_d(double p, double rK, double pK, double cK, double g) {
    if (p > pK) {
        double A = 2.0 / ((g + 1.0) * rK), B = (g - 1.0) / (g + 1.0) * pK;
        return sqrt(A / (p + B)) * (1.0 - 0.5 * (p - pK) / (p + B));
    }
    return (1.0 / (rK * cK)) * pow(p / pK, -(g + 1.0) / (2.0 * g));
}
static double guess_p(double rL, double uL, double pL, double cL, double rR, double uR, double pR, double cR,
                      double g) {
    double ppv = smax(1e-6, 0.5 * (pL + pR) - 0.125 * (uR - uL) * (rL + rR) * (cL + cR));
    double AL = 2.0 / ((g + 1.0) * rL), BL = (g - 1.0) / (g + 1.0) * pL;
    double AR = 2.0 / ((g + 1.0) * rR), BR = (g - 1.0) / (g + 1.0) * pR;
    double gL = sqrt(AL / (ppv + BL)), gR = sqrt(AR / (ppv + BR));
    return smax(1e-6, (gL * pL + gR * pR - (uR - uL)) / (gL + gR));
}
static double star_p(double rL, double uL, double pL, double cL, double rR, double uR, double pR, double cR, double g,
                     double tol, int maxit) {
    double p = guess_p(rL, uL, pL, cL, rR, uR, pR, cR, g);
    for (int it = 0; it < maxit; ++it) {
        double f = pfun(p, rL, pL, cL, g) + pfun(p, rR, pR, cR, g) + (uR - uL);
        double fd = pfun_d(p, rL, pL, cL, g) + pfun_d(p, rR, pR, cR, g);
        double pn = p - f / fd;
        if (pn < 1e-10) pn = 1e-10;
        double ch = 2.0 * fabs(pn - p) / (pn + p);
        p = pn;
        if (ch < tol) break;
    }
    return p;
}
static void fan_left(double rL, double uL, double pL, double cL, double g, double S, double* ro, double* uo,
                     double* po) {
    double cf = (2.0 / (g + 1.0)) * (cL + 0.5 * (g - 1.0) * (uL - S));
    *uo = (2.0 / (g + 1.0)) * (cL + 0.5 * (g - 1.0) * uL + S);
    *ro = rL * pow(cf / cL, 2.0 / (g - 1.0));
    *po = pL * pow(cf / cL, 2.0 * g / (g - 1.0));
}
static void fan_right(double rR, double uR, double pR, double cR, double g, double S, double* ro, double* uo,
                      double* po) {
    double cf = (2.0 / (g + 1.0)) * (cR - 0.5 * (g - 1.0) * (uR - S));
    *uo = (2.0 / (g + 1.0)) * (-cR + 0.5 * (g - 1.0) * uR + S);
    *ro = rR * pow(cf / cR, 2.0 / (g - 1.0));
    *po = pR * pow(cf / cR, 2.0 * g / (g - 1.0));
}
static void samp_left(double rL, double uL, double pL, double cL, double g, double ps, double us, double S,
                      double* ro, double* uo, double* po) {
    if (ps > pL) {
        double SLs = uL - cL * sqrt((g + 1.0) / (2.0 * g) * (ps / pL) + (g - 1.0) / (2.0 * g));
        if (S <= SLs) { *ro = rL; *uo = uL; *po = pL; return; }
        double q = ps / pL;
        *ro = rL * (q + (g - 1.0) / (g + 1.0)) / ((g - 1.0) / (g + 1.0) * q + 1.0);
        *uo = us; *po = ps;
    } else {
        double csL = cL * pow(ps / pL, (g - 1.0) / (2.0 * g));
        double SHL = uL - cL, STL = us - csL;
        if (S <= SHL) { *ro = rL; *uo = uL; *po = pL; }
        else if (S >= STL) { *ro = rL * pow(ps / pL, 1.0 / g); *uo = us; *po = ps; }
        else fan_left(rL, uL, pL, cL, g, S, ro, uo, po);
    }
}
static void samp_right(double rR, double uR, double pR, double cR, double g, double ps, double us, double S,
                       double* ro, double* uo, double* po) {
    if (ps > pR) {
        double SRs = uR + cR * sqrt((g + 1.0) / (2.0 * g) * (ps / pR) + (g - 1.0) / (2.0 * g));
        if (S >= SRs) { *ro = rR; *uo = uR; *po = pR; return; }
        double q = ps / pR;
        *ro = rR * (q + (g - 1.0) / (g + 1.0)) / ((g - 1.0) / (g + 1.0) * q + 1.0);
        *uo = us; *po = ps;
    } else {
        double csR = cR * pow(ps / pR, (g - 1.0) / (2.0 * g));
        double SHR = uR + cR, STR = us + csR;
        if (S >= SHR) { *ro = rR; *uo = uR; *po = pR; }
        else if (S <= STR) { *ro = rR * pow(ps / pR, 1.0 / g); *uo = us; *po = ps; }
        else fan_right(rR, uR, pR, cR, g, S, ro, uo, po);
    }
}
static st4 exact_flux(st4 L, st4 R, double nx, double ny, double g, double tol, int maxit) {
    double tx = -ny, ty = nx;
    double rL = L.r, rR = R.r;
    double uxL = L.ru / rL, uyL = L.rv / rL, uxR = R.ru / rR, uyR = R.rv / rR;
    double uL = uxL * nx + uyL * ny, vL = uxL * tx + uyL * ty;
    double uR = uxR * nx + uyR * ny, vR = uxR * tx + uyR * ty;
    double pL = pressure(L, g), pR = pressure(R, g);
    double cL = sound_speed(L, g), cR = sound_speed(R, g);
    double rs, us, ps, vs;
    if (2.0 * (cL + cR) / (g - 1.0) <= uR - uL) {
        double SLv = uL + 2.0 * cL / (g - 1.0), SRv = uR - 2.0 * cR / (g - 1.0);
        if (0.0 <= uL - cL) { rs = rL; us = uL; ps = pL; vs = vL; }
        else if (0.0 <= SLv) { fan_left(rL, uL, pL, cL, g, 0.0, &rs, &us, &ps); vs = vL; }
        else if (0.0 < SRv) { st4 z = {0.0, 0.0, 0.0, 0.0}; return z; }
        else if (0.0 <= uR + cR) { fan_right(rR, uR, pR, cR, g, 0.0, &rs, &us, &ps); vs = vR; }
        else { rs = rR; us = uR; ps = pR; vs = vR; }
    } else {
        double pst = star_p(rL, uL, pL, cL, rR, uR, pR, cR, g, tol, maxit);
        double ust = 0.5 * (uL + uR) + 0.5 * (pfun(pst, rR, pR, cR, g) - pfun(pst, rL, pL, cL, g));
        if (0.0 <= ust) { samp_left(rL, uL, pL, cL, g, pst, ust, 0.0, &rs, &us, &ps); vs = vL; }
        else { samp_right(rR, uR, pR, cR, g, pst, ust, 0.0, &rs, &us, &ps); vs = vR; }
    }
    double uxs = us * nx - vs * ny, uys = us * ny + vs * nx;
    return phys_flux(from_primitive(rs, uxs, uys, ps, g), nx, ny, g);
}

static inline st4 num_flux(const fvk_params* p, st4 L, st4 R, double nx, double ny) {
    switch (p->flux_scheme) {
        case FVK_FLUX_HLLC: return hllc(L, R, nx, ny, p->gamma);
        case FVK_FLUX_EXACT: return exact_flux(L, R, nx, ny, p->gamma, p->exact_riemann_tol, p->exact_riemann_max_iter);
        default: return rusanov(L, R, nx, ny, p->gamma);
    }
}

void fvo_flux(int32_t scheme, int64_t n, const double* UL, const double* UR, const double* nxy, double gamma,
              double tol, int32_t max_iter, double* out) {
    fvk_params p;
    memset(&p, 0, sizeof p);
    p.flux_scheme = scheme; p.gamma = gamma; p.exact_riemann_tol = tol; p.exact_riemann_max_iter = max_iter;
    for (int64_t i = 0; i < n; ++i) sv4(out + 4 * i, num_flux(&p, ld4(UL + 4 * i), ld4(UR + 4 * i), nxy[2 * i], nxy[2 * i + 1]));
}

/* ------------------------------------------------------------------------
 * mesh geometry
 * ---------------------------------------------------------------------- */

/* UnstructuredMesh::compute_geometry, include/UnstructuredMesh.h:89-105 */
static void compute_geometry(fvo_solver* s) {
    for (int64_t f = 0; f < s->nf; ++f) {
        const double* n1 = s->node_xy + 2 * s->fnode[2 * f];
        const double* n2 = s->node_xy + 2 * s->fnode[2 * f + 1];
        s->fmid[2 * f] = 0.5 * (n1[0] + n2[0]);
        s->fmid[2 * f + 1] = 0.5 * (n1[1] + n2[1]);
        double dx = n2[0] - n1[0], dy = n2[1] - n1[1];
        s->farea[f] = sqrt(dx * dx + dy * dy);
        s->fnx[f] = dy / s->farea[f];
        s->fny[f] = -dx / s->farea[f];
    }
}

/* position of cell_right's centroid in cell_left's frame (periodic shift; none in the reference) */
static inline vec2 right_centroid(const fvo_solver* s, int64_t f) {
    int cr = s->fcell[2 * f + 1];
    vec2 v = {s->cxy[2 * cr], s->cxy[2 * cr + 1]};
    if (s->shift) { v.x = v.x + s->shift[2 * f]; v.y = v.y + s->shift[2 * f + 1]; }
    return v;
}
/* position of the neighbour across face f seen from cell c (centroid, or face midpoint on a boundary) */
static inline vec2 neighbour_pos(const fvo_solver* s, int64_t f, int c) {
    int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
    vec2 v;
    if (cr == -1) { v.x = s->fmid[2 * f]; v.y = s->fmid[2 * f + 1]; return v; }
    if (cl == c) return right_centroid(s, f);
    v.x = s->cxy[2 * cl]; v.y = s->cxy[2 * cl + 1];
    if (s->shift) { v.x = v.x - s->shift[2 * f]; v.y = v.y - s->shift[2 * f + 1]; }
    return v;
}
/* face midpoint in the frame of cell c */
static inline vec2 mid_in(const fvo_solver* s, int64_t f, int c) {
    vec2 v = {s->fmid[2 * f], s->fmid[2 * f + 1]};
    if (s->shift && s->fcell[2 * f + 1] == c && s->fcell[2 * f] != c) {
        v.x = v.x - s->shift[2 * f]; v.y = v.y - s->shift[2 * f + 1];
    }
    return v;
}

/* GradientCalculator ctor, LSQ precompute: src/GradientReconstruction.cpp:45-80 */
static void lsq_precompute(fvo_solver* s) {
    for (int64_t c = 0; c < s->nc; ++c) {
        double m00 = 0.0, m01 = 0.0, m11 = 0.0;
        for (int k = s->cptr[c]; k < s->cptr[c + 1]; ++k) {
            int f = s->cface[k];
            vec2 q = neighbour_pos(s, f, (int)c);
            double dx = q.x - s->cxy[2 * c], dy = q.y - s->cxy[2 * c + 1];
            double w = 1.0 / (dx * dx + dy * dy);
            m00 += w * dx * dx;
            m01 += w * dx * dy;
            m11 += w * dy * dy;
        }
        double det = m00 * m11 - m01 * m01;
        s->lsq_inv[3 * c] = m11 / det;
        s->lsq_inv[3 * c + 1] = -m01 / det;
        s->lsq_inv[3 * c + 2] = m00 / det;
    }
}

/* GradientCalculator::compute: src/GradientReconstruction.cpp:83-150.
 * phi strided by `ps` (value of cell c at phi[c*ps]); bphi indexed by face. */
static void gradient(const fvo_solver* s, int scheme, const double* phi, int ps, const double* bphi, double* grad) {
    if (scheme == FVK_GRAD_GREEN_GAUSS) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < s->nc; ++c) {
            double sx = 0.0, sy = 0.0;
            double cx = s->cxy[2 * c], cy = s->cxy[2 * c + 1];
            for (int k = s->cptr[c]; k < s->cptr[c + 1]; ++k) {
                int f = s->cface[k];
                int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
                double pf;
                if (cr == -1) {
                    pf = bphi[f];
                } else {
                    int other = (cl == (int)c) ? cr : cl;
                    /* distances in cell_left's frame (identical to the reference without periodic shift) */
                    vec2 pc, po;
                    if (cl == (int)c) { pc.x = cx; pc.y = cy; po = right_centroid(s, f); }
                    else { pc = right_centroid(s, f); po.x = s->cxy[2 * cl]; po.y = s->cxy[2 * cl + 1]; }
                    double dc = hypot(pc.x - s->fmid[2 * f], pc.y - s->fmid[2 * f + 1]);
                    double dO = hypot(po.x - s->fmid[2 * f], po.y - s->fmid[2 * f + 1]);
                    double wc = dO / (dc + dO);
                    pf = wc * phi[c * ps] + (1.0 - wc) * phi[(int64_t)other * ps];
                }
                double nx = s->fnx[f], ny = s->fny[f];
                if (cl != (int)c) { nx = -nx; ny = -ny; }
                sx += pf * nx * s->farea[f];
                sy += pf * ny * s->farea[f];
            }
            grad[2 * c] = sx / s->vol[c];
            grad[2 * c + 1] = sy / s->vol[c];
        }
        return;
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s->nc; ++c) {
        double bx = 0.0, by = 0.0;
        for (int k = s->cptr[c]; k < s->cptr[c + 1]; ++k) {
            int f = s->cface[k];
            int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
            vec2 q = neighbour_pos(s, f, (int)c);
            double pn = (cr == -1) ? bphi[f] : phi[(int64_t)((cl == (int)c) ? cr : cl) * ps];
            double dx = q.x - s->cxy[2 * c], dy = q.y - s->cxy[2 * c + 1];
            double w = 1.0 / (dx * dx + dy * dy);
            double dphi = pn - phi[c * ps];
            bx += w * dx * dphi;
            by += w * dy * dphi;
        }
        const double* inv = s->lsq_inv + 3 * c;
        grad[2 * c] = inv[0] * bx + inv[1] * by;
        grad[2 * c + 1] = inv[1] * bx + inv[2] * by;
    }
}

/* face_normal_geometry + face_gradient + corrected_face_gradient_vector +
 * face_normal_distance: src/GradientReconstruction.cpp:22-40, 153-200 */
typedef struct {
    double dist, cos_theta, dhx, dhy;
} fgeom;
static fgeom face_normal_geometry(const fvo_solver* s, int64_t f) {
    int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
    vec2 R;
    if (cr == -1) { R.x = s->fmid[2 * f]; R.y = s->fmid[2 * f + 1]; }
    else R = right_centroid(s, f);
    double dx = R.x - s->cxy[2 * cl], dy = R.y - s->cxy[2 * cl + 1];
    fgeom g;
    g.dist = hypot(dx, dy);
    g.dhx = dx / g.dist;
    g.dhy = dy / g.dist;
    g.cos_theta = s->fnx[f] * g.dhx + s->fny[f] * g.dhy;
    return g;
}
static double face_normal_distance(const fvo_solver* s, int64_t f) {
    fgeom g = face_normal_geometry(s, f);
    return g.dist * g.cos_theta;
}
typedef struct {
    double dphidn, gx, gy;
} facegrad;
static facegrad face_gradient(const fvo_solver* s, int64_t f, double phiL, double phiR, const double* gL,
                              const double* gR, double bval) {
    int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
    int b = cr == -1;
    vec2 R;
    if (b) { R.x = s->fmid[2 * f]; R.y = s->fmid[2 * f + 1]; }
    else R = right_centroid(s, f);
    double phiRv = b ? bval : phiR;
    fgeom g = face_normal_geometry(s, f);
    double direct = (phiRv - phiL) / (g.dist * g.cos_theta);
    double kx = s->fnx[f] - g.dhx / g.cos_theta;
    double ky = s->fny[f] - g.dhy / g.cos_theta;
    facegrad o;
    if (b) {
        o.gx = gL[0]; o.gy = gL[1];
    } else {
        double dL = hypot(s->cxy[2 * cl] - s->fmid[2 * f], s->cxy[2 * cl + 1] - s->fmid[2 * f + 1]);
        double dR = hypot(R.x - s->fmid[2 * f], R.y - s->fmid[2 * f + 1]);
        double wL = dR / (dL + dR);
        o.gx = wL * gL[0] + (1.0 - wL) * gR[0];
        o.gy = wL * gL[1] + (1.0 - wL) * gR[1];
    }
    o.dphidn = direct + o.gx * kx + o.gy * ky;
    return o;
}
static void corrected_vector(facegrad fg, double nx, double ny, double* gx, double* gy) {
    double avg = fg.gx * nx + fg.gy * ny;
    double delta = fg.dphidn - avg;
    *gx = fg.gx + delta * nx;
    *gy = fg.gy + delta * ny;
}

/* compute_wall_distance: src/WallDistance.cpp:13-43 */
static double seg_dist_sq(double px, double py, double x1, double y1, double x2, double y2) {
    double dx = x2 - x1, dy = y2 - y1;
    double l2 = dx * dx + dy * dy;
    double t = (l2 > 0.0) ? ((px - x1) * dx + (py - y1) * dy) / l2 : 0.0;
    t = (t < 0.0) ? 0.0 : ((1.0 < t) ? 1.0 : t);
    double ex = px - (x1 + t * dx), ey = py - (y1 + t * dy);
    return ex * ex + ey * ey;
}
static void wall_distance_impl(int64_t nc, const double* cxy, const double* node_xy, const int32_t* fnode,
                               const int32_t* wf, int64_t nw, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
        double m = INFINITY;
        for (int64_t k = 0; k < nw; ++k) {
            const double* a = node_xy + 2 * fnode[2 * wf[k]];
            const double* b = node_xy + 2 * fnode[2 * wf[k] + 1];
            m = smin(m, seg_dist_sq(cxy[2 * c], cxy[2 * c + 1], a[0], a[1], b[0], b[1]));
        }
        out[c] = sqrt(m);
    }
}

/* ------------------------------------------------------------------------
 * turbulence models
 * ---------------------------------------------------------------------- */

/* src/SpalartAllmaras.cpp:8-69 */
static double sa_cw1(const fvk_sa_constants* c) { return c->cb1 / (c->kappa * c->kappa) + (1.0 + c->cb2) / c->sigma; }
static double sa_fv1(double nut, double nu, const fvk_sa_constants* c) {
    if (nut <= 0.0) return 0.0;
    double chi = nut / nu, chi3 = chi * chi * chi;
    return chi3 / (chi3 + c->cv1 * c->cv1 * c->cv1);
}
static double sa_eddy(double nut, double nu, const fvk_sa_constants* c) { return nut * sa_fv1(nut, nu, c); }
static void sa_source(double nut, double nu, double omega, double d, double gx, double gy, const fvk_sa_constants* c,
                      double* P, double* D, double* CD) {
    double chi = (nut > 0.0) ? nut / nu : 0.0;
    double fv1 = sa_fv1(nut, nu, c);
    double fv2 = 1.0 - chi / (1.0 + chi * fv1);
    double d2 = d * d;
    double sbar = (nut / (c->kappa * c->kappa * d2)) * fv2;
    double st;
    if (sbar >= -c->cv2 * omega) st = omega + sbar;
    else st = omega + omega * (c->cv2 * c->cv2 * omega + c->cv3 * sbar) / ((c->cv3 - 2.0 * c->cv2) * omega - sbar);
    *P = c->cb1 * st * nut;
    *D = 0.0;
    if (nut > 0.0) {
        double sts = smax(st, 1e-10);
        double r = smin(nut / (sts * c->kappa * c->kappa * d2), 10.0);
        double gg = r + c->cw2 * (pow(r, 6.0) - r);
        double fw = gg * pow((1.0 + pow(c->cw3, 6.0)) / (pow(gg, 6.0) + pow(c->cw3, 6.0)), 1.0 / 6.0);
        double nod = nut / d;
        *D = sa_cw1(c) * fw * nod * nod;
    }
    *CD = (c->cb2 / c->sigma) * (gx * gx + gy * gy);
}
void fvo_sa_source(int64_t n, const double* nut, const double* nu, const double* omega, const double* dist,
                   const double* grad_nut, const fvk_sa_constants* c, double* out, double* eddy) {
    for (int64_t i = 0; i < n; ++i) {
        sa_source(nut[i], nu[i], omega[i], dist[i], grad_nut[2 * i], grad_nut[2 * i + 1], c, out + 3 * i,
                  out + 3 * i + 1, out + 3 * i + 2);
        eddy[i] = sa_eddy(nut[i], nu[i], c);
    }
}

/* src/SSTKOmega.cpp:8-101 */
typedef struct {
    double kp, kd, wp, wd, cd, nut, F1, F2, S, Om;
} sst_src;
static double sst_gamma1(const fvk_sst_constants* c) {
    return c->beta1 / c->beta_star - c->sigma_omega1 * c->kappa * c->kappa / sqrt(c->beta_star);
}
static double sst_gamma2(const fvk_sst_constants* c) {
    return c->beta2 / c->beta_star - c->sigma_omega2 * c->kappa * c->kappa / sqrt(c->beta_star);
}
static sst_src sst_source(double rho, double k, double w, double nu, double d, const double* gu, const double* gv,
                          const double* gk, const double* gw, int variant, const fvk_sst_constants* c, int kato) {
    sst_src r;
    memset(&r, 0, sizeof r);
    double s11 = gu[0], s22 = gv[1], s12 = 0.5 * (gu[1] + gv[0]);
    r.S = sqrt(2.0 * (s11 * s11 + s22 * s22 + 2.0 * s12 * s12));
    r.Om = fabs(gv[0] - gu[1]);
    if (k <= 0.0 || w <= 0.0) return r;
    double d2 = d * d;
    double gdot = gk[0] * gw[0] + gk[1] * gw[1];
    /* F1 */
    double CD = smax(2.0 * rho * c->sigma_omega2 * gdot / w, 1e-20);
    double t1 = smax(sqrt(k) / (c->beta_star * w * d), 500.0 * nu / (d2 * w));
    double t2 = 4.0 * rho * c->sigma_omega2 * k / (CD * d2);
    double a1 = smin(t1, t2);
    double a14 = a1 * a1 * a1 * a1;
    r.F1 = tanh(a14);
    /* F2 */
    double a2 = smax(2.0 * sqrt(k) / (c->beta_star * w * d), 500.0 * nu / (d2 * w));
    r.F2 = tanh(a2 * a2);
    double lim = (variant == FVK_SST_VORTICITY) ? r.Om : r.S;
    r.nut = c->a1 * k / smax(c->a1 * w, lim * r.F2);
    double prs = kato ? (r.S * r.Om) : (r.S * r.S);
    double mut = rho * r.nut;
    double bsrwk = c->beta_star * rho * w * k;
    double clip = (variant == FVK_SST_VORTICITY) ? 20.0 : 10.0;
    r.kp = smin(mut * prs, clip * bsrwk);
    r.kd = bsrwk;
    double gam = r.F1 * sst_gamma1(c) + (1.0 - r.F1) * sst_gamma2(c);
    double beta = r.F1 * c->beta1 + (1.0 - r.F1) * c->beta2;
    r.wp = gam * rho * prs;
    r.wd = beta * rho * w * w;
    r.cd = 2.0 * (1.0 - r.F1) * rho * c->sigma_omega2 * gdot / w;
    return r;
}
void fvo_sst_source(int64_t n, const double* rho, const double* k, const double* omega, const double* nu,
                    const double* dist, const double* gu, const double* gv, const double* gk, const double* gw,
                    int32_t variant, const fvk_sst_constants* c, int32_t kato, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        sst_src r = sst_source(rho[i], k[i], omega[i], nu[i], dist[i], gu + 2 * i, gv + 2 * i, gk + 2 * i, gw + 2 * i,
                               variant, c, kato);
        double* o = out + 10 * i;
        o[0] = r.kp; o[1] = r.kd; o[2] = r.wp; o[3] = r.wd; o[4] = r.cd; o[5] = r.nut; o[6] = r.F1; o[7] = r.F2;
        o[8] = r.S; o[9] = r.Om;
    }
}

/* next_cfl: src/CflRamp.cpp:7-27 */
double fvo_next_cfl(double current, double cfl_max, const double* hist, int32_t n_hist, double cfl_min,
                    double growth, double shrink, int32_t window, double divergence_threshold) {
    if (n_hist < window) return current;
    double trend = log10(hist[n_hist - 1]) - log10(hist[0]);
    const double dead = log10(1.1);
    double next = current;
    if (trend > divergence_threshold) next = cfl_min;
    else if (trend < 0.0) next = current * growth;
    else if (trend > dead) next = current * shrink;
    return (next < cfl_min) ? cfl_min : ((cfl_max < next) ? cfl_max : next);
}

/* ------------------------------------------------------------------------
 * boundary states
 * ---------------------------------------------------------------------- */

/* EulerFVMSolver::ghost_state (src/EulerFVMSolver.cpp:174-193) and
 * NavierStokesFVMSolver::ghost_state (src/NavierStokesFVMSolver.cpp:265-284,
 * also RANS SA :416-431 and SST) */
static st4 ghost(const fvo_solver* s, int64_t f, st4 L) {
    const fvk_bc* bc = &s->bcs[s->fpatch[f]];
    double g = s->p.gamma;
    if (bc->type == FVK_BC_FARFIELD) return ld4(bc->farfield_state);
    if (bc->type == FVK_BC_OUTFLOW) return L;
    double u = L.ru / L.r, v = L.rv / L.r;
    if (s->p.equation == FVK_EQ_EULER) {
        double Vn = u * s->fnx[f] + v * s->fny[f];
        double ug = u - 2.0 * Vn * s->fnx[f], vg = v - 2.0 * Vn * s->fny[f];
        return from_primitive(L.r, ug, vg, pressure(L, g), g);
    }
    double pL = pressure(L, g);
    return from_primitive(L.r, 2.0 * bc->wall_u - u, 2.0 * bc->wall_v - v, pL, g);
}

/* ------------------------------------------------------------------------
 * time step candidates per face (the reference's compute_dt bodies)
 * ---------------------------------------------------------------------- */

/* EulerFVMSolver::compute_dt per-face candidate, src/EulerFVMSolver.cpp:138-158 */
static double dt_face_euler(const fvo_solver* s, const double* U, int64_t f) {
    double g = s->p.gamma;
    int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
    st4 L = ld4(U + 4 * cl);
    double uL = L.ru / L.r, vL = L.rv / L.r;
    double VnL = uL * s->fnx[f] + vL * s->fny[f];
    double Smax = fabs(VnL) + sound_speed(L, g);
    double len;
    if (cr != -1) {
        st4 R = ld4(U + 4 * cr);
        double uR = R.ru / R.r, vR = R.rv / R.r;
        double VnR = uR * s->fnx[f] + vR * s->fny[f];
        Smax = smax(Smax, fabs(VnR) + sound_speed(R, g));
        len = smin(s->vol[cl], s->vol[cr]) / s->farea[f];
    } else {
        len = s->vol[cl] / s->farea[f];
    }
    return s->p.cfl * len / Smax;
}

/* NavierStokesFVMSolver::compute_dt (src/NavierStokesFVMSolver.cpp:231-258),
 * RANSTurbulenceSASolver::compute_dt (src/RANSTurbulenceSASolver.cpp:313-346),
 * RANSTurbulenceSSTSolver::compute_dt (src/RANSTurbulenceSSTSolver.cpp:380-429) */
static double dt_face_viscous(const fvo_solver* s, const double* U, const double* turb, int64_t f) {
    const fvk_params* p = &s->p;
    double g = p->gamma;
    int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
    st4 L = ld4(U + 4 * cl);
    double uL = L.ru / L.r, vL = L.rv / L.r;
    double VnL = uL * s->fnx[f] + vL * s->fny[f];
    double Smax = fabs(VnL) + sound_speed(L, g);
    double rmin = L.r;
    double ntmax = 0.0, wmax = 0.0;
    if (p->equation == FVK_EQ_RANS_SA) ntmax = sa_eddy(turb[cl], p->mu / L.r, &p->sa);
    if (p->equation == FVK_EQ_RANS_SST) {
        double k = turb[2 * cl], w = turb[2 * cl + 1];
        ntmax = (k > 0.0 && w > 0.0) ? k / w : 0.0;
        wmax = w;
    }
    if (cr != -1) {
        st4 R = ld4(U + 4 * cr);
        double uR = R.ru / R.r, vR = R.rv / R.r;
        double VnR = uR * s->fnx[f] + vR * s->fny[f];
        Smax = smax(Smax, fabs(VnR) + sound_speed(R, g));
        rmin = smin(rmin, R.r);
        if (p->equation == FVK_EQ_RANS_SA) ntmax = smax(ntmax, sa_eddy(turb[cr], p->mu / R.r, &p->sa));
        if (p->equation == FVK_EQ_RANS_SST) {
            double k = turb[2 * cr], w = turb[2 * cr + 1];
            ntmax = smax(ntmax, (k > 0.0 && w > 0.0) ? k / w : 0.0);
            wmax = smax(wmax, w);
        }
    }
    double len = face_normal_distance(s, f);
    double nu = (p->mu > 0.0) ? p->mu / rmin : 0.0;
    if (p->equation == FVK_EQ_NAVIER_STOKES) {
        double St = Smax + 2.0 * nu / len;
        return p->cfl * len / St;
    }
    double nue = nu + ntmax;
    double St = Smax + 2.0 * nue / len;
    double dtd = p->cfl * len / St;
    if (p->equation == FVK_EQ_RANS_SA) return dtd;
    double coef = smax(p->sst.beta_star, smax(p->sst.beta1, p->sst.beta2));
    double dts = p->cfl / (coef * smax(wmax, 1e-30));
    return smin(dtd, dts);
}

static double dt_face(const fvo_solver* s, const double* U, const double* turb, int64_t f) {
    return s->p.equation == FVK_EQ_EULER ? dt_face_euler(s, U, f) : dt_face_viscous(s, U, turb, f);
}

/* ------------------------------------------------------------------------
 * Euler residual, first order (reference) and second order (extension)
 * ---------------------------------------------------------------------- */

typedef struct {
    double r, u, v, p;
} prim4;
static inline prim4 to_prim(st4 s, double g) {
    prim4 w = {s.r, s.ru / s.r, s.rv / s.r, pressure(s, g)};
    return w;
}

/* Venkatakrishnan limiter function (extension): phi = ((d1^2 + eps2) + 2 d1 d2) /
 * (d1^2 + 2 d2^2 + d1 d2 + eps2) for the unlimited increment d2 and the
 * allowed one d1 (same sign). */
double fvo_venkatakrishnan(double d1, double d2, double eps2) {
    double d11 = d1 * d1;
    return ((d11 + eps2) + 2.0 * d1 * d2) / (d11 + 2.0 * d2 * d2 + d1 * d2 + eps2);
}

/* Extension (no reference code): limited least-squares gradients of the
 * primitive variables (rho, u, v, p).  Boundary faces contribute the face
 * value 0.5 (W_c + W_ghost) at the face midpoint; gradients use the
 * reference's LSQ weights and precomputed inverse (src/GradientReconstruction.cpp:45-80,
 * 121-148); the Venkatakrishnan limiter takes eps^2 = (K sqrt(V))^3, the
 * neighbour min/max including the cell itself, and phi = min(1, min over faces). */
static void limited_gradients(fvo_solver* s, const double* U) {
    double g = s->p.gamma;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s->nc; ++c) {
        prim4 Wc = to_prim(ld4(U + 4 * c), g);
        double wc[4] = {Wc.r, Wc.u, Wc.v, Wc.p};
        double bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0};
        double wmin[4], wmax[4];
        for (int q = 0; q < 4; ++q) wmin[q] = wmax[q] = wc[q];
        for (int k = s->cptr[c]; k < s->cptr[c + 1]; ++k) {
            int f = s->cface[k];
            int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
            double wn[4];
            if (cr == -1) {
                st4 G = ghost(s, f, ld4(U + 4 * c));
                prim4 Wg = to_prim(G, g);
                wn[0] = 0.5 * (Wc.r + Wg.r); wn[1] = 0.5 * (Wc.u + Wg.u);
                wn[2] = 0.5 * (Wc.v + Wg.v); wn[3] = 0.5 * (Wc.p + Wg.p);
            } else {
                int o = (cl == (int)c) ? cr : cl;
                prim4 Wo = to_prim(ld4(U + 4 * o), g);
                wn[0] = Wo.r; wn[1] = Wo.u; wn[2] = Wo.v; wn[3] = Wo.p;
            }
            vec2 qp = neighbour_pos(s, f, (int)c);
            double dx = qp.x - s->cxy[2 * c], dy = qp.y - s->cxy[2 * c + 1];
            double w = 1.0 / (dx * dx + dy * dy);
            for (int q = 0; q < 4; ++q) {
                double dphi = wn[q] - wc[q];
                bx[q] += w * dx * dphi;
                by[q] += w * dy * dphi;
                wmin[q] = smin(wmin[q], wn[q]);
                wmax[q] = smax(wmax[q], wn[q]);
            }
        }
        const double* inv = s->lsq_inv + 3 * c;
        double gx[4], gy[4], lim[4];
        for (int q = 0; q < 4; ++q) {
            gx[q] = inv[0] * bx[q] + inv[1] * by[q];
            gy[q] = inv[1] * bx[q] + inv[2] * by[q];
            lim[q] = 1.0;
        }
        if (s->p.limiter == FVK_LIMITER_VENKATAKRISHNAN) {
            double kh = s->p.venkat_k * sqrt(s->vol[c]);
            double eps2 = kh * kh * kh;
            /* phi_q = min(1, min over faces with d2 != 0 of num/den): d1 d2 >= 0
             * makes num >= 0 and den > 0, so the smallest ratio is selected by
             * cross-multiplication (a strict decrease replaces it) and divided once */
            double bn[4] = {1.0, 1.0, 1.0, 1.0}, bd[4] = {1.0, 1.0, 1.0, 1.0};
            for (int k = s->cptr[c]; k < s->cptr[c + 1]; ++k) {
                int f = s->cface[k];
                vec2 m = mid_in(s, f, (int)c);
                double rx = m.x - s->cxy[2 * c], ry = m.y - s->cxy[2 * c + 1];
                for (int q = 0; q < 4; ++q) {
                    double d2 = gx[q] * rx + gy[q] * ry;
                    double d1 = (d2 > 0.0 ? wmax[q] : wmin[q]) - wc[q];
                    double d11 = d1 * d1;
                    double num = (d11 + eps2) + 2.0 * d1 * d2;
                    double den = d11 + 2.0 * d2 * d2 + d1 * d2 + eps2;
                    if (d2 != 0.0 && num * bd[q] < bn[q] * den) { bn[q] = num; bd[q] = den; }
                }
            }
            for (int q = 0; q < 4; ++q) lim[q] = bn[q] / bd[q];
        }
        if (s->p.limiter == FVK_LIMITER_FIRST_ORDER)
            for (int q = 0; q < 4; ++q) lim[q] = 0.0;
        for (int q = 0; q < 4; ++q) {
            s->grad8[8 * c + 2 * q] = lim[q] * gx[q];
            s->grad8[8 * c + 2 * q + 1] = lim[q] * gy[q];
            s->lim4[4 * c + q] = lim[q];
        }
    }
}

/* ------------------------------------------------------------------------
 * Extension: second-order faces on primitive face states.  A MUSCL face
 * state is W = (rho, u, v, p); the fluxes take it as it is and form
 * conserved quantities only where the flux uses them (equal, in exact
 * arithmetic, to Rusanov/HLLC of from_primitive(W)).
 * ---------------------------------------------------------------------- */

/* MUSCL face state of cell c at offset (rx, ry): W + grad.r per primitive,
 * or W itself where the reconstructed density or pressure is not positive. */
static prim4 reconstruct_w(const fvo_solver* s, const double* U, int c, double rx, double ry) {
    prim4 W = to_prim(ld4(U + 4 * c), s->p.gamma);
    const double* gr = s->grad8 + 8 * c;
    double d0 = gr[0] * rx + gr[1] * ry, d1 = gr[2] * rx + gr[3] * ry;
    double d2 = gr[4] * rx + gr[5] * ry, d3 = gr[6] * rx + gr[7] * ry;
    double rho = W.r + d0, p = W.p + d3;
    if (!(rho > 0.0) || !(p > 0.0)) return W;
    prim4 o = {rho, W.u + d1, W.v + d2, p};
    return o;
}

/* specific total energy E / rho of a primitive state */
static inline double specific_energy(prim4 w, double g) {
    return w.p / ((g - 1.0) * w.r) + 0.5 * (w.u * w.u + w.v * w.v);
}

/* physical flux of a primitive state with total energy E and normal velocity Vn */
static inline st4 phys_flux_w(prim4 w, double E, double Vn, double nx, double ny) {
    double m = w.r * Vn;
    st4 f = {m, m * w.u + w.p * nx, m * w.v + w.p * ny, (E + w.p) * Vn};
    return f;
}

static st4 rusanov_w(prim4 L, prim4 R, double nx, double ny, double g) {
    double VnL = L.u * nx + L.v * ny, VnR = R.u * nx + R.v * ny;
    double Smax = smax(fabs(VnL) + sqrt(g * L.p / L.r), fabs(VnR) + sqrt(g * R.p / R.r));
    double EL = L.r * specific_energy(L, g), ER = R.r * specific_energy(R, g);
    st4 FL = phys_flux_w(L, EL, VnL, nx, ny), FR = phys_flux_w(R, ER, VnR, nx, ny);
    double h = 0.5 * Smax;
    st4 o = {0.5 * (FL.r + FR.r) - h * (R.r - L.r), 0.5 * (FL.ru + FR.ru) - h * (R.r * R.u - L.r * L.u),
             0.5 * (FL.rv + FR.rv) - h * (R.r * R.v - L.r * L.v), 0.5 * (FL.E + FR.E) - h * (ER - EL)};
    return o;
}

/* HLLC on primitive states (Davis speeds, Toro 10.37-10.39); m_K =
 * rho_K (S_K - Vn_K) is the mass flux through the outer wave */
static st4 hllc_w(prim4 L, prim4 R, double nx, double ny, double g) {
    double VnL = L.u * nx + L.v * ny, VnR = R.u * nx + R.v * ny;
    double cL = sqrt(g * L.p / L.r), cR = sqrt(g * R.p / R.r);
    double SL = smin(VnL - cL, VnR - cR);
    double SR = smax(VnL + cL, VnR + cR);
    if (SL >= 0.0) return phys_flux_w(L, L.r * specific_energy(L, g), VnL, nx, ny);
    if (SR <= 0.0) return phys_flux_w(R, R.r * specific_energy(R, g), VnR, nx, ny);
    double mL = L.r * (SL - VnL), mR = R.r * (SR - VnR);
    double Ss = (R.p - L.p + mL * VnL - mR * VnR) / (mL - mR);
    int left = Ss >= 0.0;
    prim4 K = left ? L : R;
    double SK = left ? SL : SR, VnK = left ? VnL : VnR, mK = left ? mL : mR;
    double VtK = K.v * nx - K.u * ny;
    double eK = specific_energy(K, g), EK = K.r * eK;
    double factor = mK / (SK - Ss);
    double mn = factor * Ss, mt = factor * VtK;
    double Es = factor * (eK + (Ss - VnK) * (Ss + K.p / mK));
    st4 F = phys_flux_w(K, EK, VnK, nx, ny);
    st4 o = {F.r + SK * (factor - K.r), F.ru + SK * ((mn * nx - mt * ny) - K.r * K.u),
             F.rv + SK * ((mn * ny + mt * nx) - K.r * K.v), F.E + SK * (Es - EK)};
    return o;
}

static st4 num_flux_w(const fvk_params* p, prim4 L, prim4 R, double nx, double ny) {
    double g = p->gamma;
    switch (p->flux_scheme) {
        case FVK_FLUX_HLLC: return hllc_w(L, R, nx, ny, g);
        case FVK_FLUX_EXACT:
            return num_flux(p, from_primitive(L.r, L.u, L.v, L.p, g), from_primitive(R.r, R.u, R.v, R.p, g), nx, ny);
        default: return rusanov_w(L, R, nx, ny, g);
    }
}

/* the Euler ghost state of a primitive face state: farfield -> the
 * prescribed state, outflow -> copy, wall -> the normal velocity reflected */
static prim4 ghost_w(const fvo_solver* s, int64_t f, prim4 L) {
    const fvk_bc* bc = &s->bcs[s->fpatch[f]];
    if (bc->type == FVK_BC_FARFIELD) return to_prim(ld4(bc->farfield_state), s->p.gamma);
    if (bc->type == FVK_BC_OUTFLOW) return L;
    double Vn = L.u * s->fnx[f] + L.v * s->fny[f];
    prim4 o = {L.r, L.u - 2.0 * Vn * s->fnx[f], L.v - 2.0 * Vn * s->fny[f], L.p};
    return o;
}

/* Face flux for the Euler path.  First order: the face pass of
 * EulerFVMSolver::step (src/EulerFVMSolver.cpp:64-82).  Second order: the
 * same face loop over MUSCL primitive face states. */
static st4 euler_face_flux(const fvo_solver* s, const double* U, int64_t f) {
    int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
    st4 F;
    if (s->p.order == 2) {
        prim4 L = reconstruct_w(s, U, cl, s->fmid[2 * f] - s->cxy[2 * cl], s->fmid[2 * f + 1] - s->cxy[2 * cl + 1]);
        prim4 R;
        if (cr != -1) {
            vec2 rc = right_centroid(s, f);
            R = reconstruct_w(s, U, cr, s->fmid[2 * f] - rc.x, s->fmid[2 * f + 1] - rc.y);
        } else {
            R = ghost_w(s, f, L);
        }
        F = num_flux_w(&s->p, L, R, s->fnx[f], s->fny[f]);
    } else {
        st4 L = ld4(U + 4 * cl);
        st4 R = (cr != -1) ? ld4(U + 4 * cr) : ghost(s, f, L);
        F = num_flux(&s->p, L, R, s->fnx[f], s->fny[f]);
    }
    double A = s->farea[f];
    st4 o = {A * F.r, A * F.ru, A * F.rv, A * F.E};
    return o;
}

/* ------------------------------------------------------------------------
 * viscous residuals: NS (src/NavierStokesFVMSolver.cpp:95-189),
 * SA (src/RANSTurbulenceSASolver.cpp:118-259), SST (src/RANSTurbulenceSSTSolver.cpp:142-307)
 * ---------------------------------------------------------------------- */

/* residual R (AoS nvars per cell) of state (U, turb): face pass + gather,
 * including volumetric sources; the cell pass of each reference step()
 * without the update. */
static void residual(fvo_solver* s, const double* U, const double* turb, double* R) {
    const fvk_params* p = &s->p;
    const int nv = nvars_of(p->equation);
    const int64_t nc = s->nc, nf = s->nf;
    double* ff = (double*)malloc(sizeof(double) * nf * nv);
    double g = p->gamma;

    if (p->equation == FVK_EQ_DIFFUSION) {
        /* UnstructuredFVMSolver::step face pass + boundary_flux, src/UnstructuredFVMSolver.cpp:44-71, 113-129 */
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < nf; ++f) {
            int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
            double phiL = U[cl];
            if (cr != -1) {
                double dx = s->cxy[2 * cr] - s->cxy[2 * cl], dy = s->cxy[2 * cr + 1] - s->cxy[2 * cl + 1];
                double d = sqrt(dx * dx + dy * dy);
                double gn = (U[cr] - phiL) / d;
                ff[f] = -p->alpha * gn * s->farea[f];
            } else {
                const fvk_bc* bc = &s->bcs[s->fpatch[f]];
                if (bc->type == FVK_BC_NEUMANN) {
                    ff[f] = bc->value * s->farea[f];
                } else {
                    double dx = s->fmid[2 * f] - s->cxy[2 * cl], dy = s->fmid[2 * f + 1] - s->cxy[2 * cl + 1];
                    double d = sqrt(dx * dx + dy * dy);
                    double gn = (bc->value - phiL) / d;
                    ff[f] = -p->alpha * gn * s->farea[f];
                }
            }
        }
    } else if (p->equation == FVK_EQ_ADVECTION_DIFFUSION) {
        /* AdvectionDiffusionFVMSolver::step face pass + boundary_flux,
         * src/AdvectionDiffusionFVMSolver.cpp:42-72, 117-138, with the
         * boundary values of build_boundary_phi (:33-40) */
        double* bphi = calloc(nf, sizeof(double));
        double* gr = malloc(sizeof(double) * 2 * nc);
        for (int64_t f = 0; f < nf; ++f) {
            if (s->fcell[2 * f + 1] != -1) continue;
            const fvk_bc* bc = &s->bcs[s->fpatch[f]];
            bphi[f] = bc->type == FVK_BC_DIRICHLET ? bc->value : U[s->fcell[2 * f]];
        }
        gradient(s, p->gradient_scheme, U, 1, bphi, gr);
        const double zero2[2] = {0.0, 0.0};
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < nf; ++f) {
            int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
            double phiL = U[cl];
            double Vn = p->advection_u * s->fnx[f] + p->advection_v * s->fny[f];
            double A = s->farea[f];
            if (cr != -1) {
                double phiR = U[cr];
                facegrad fg = face_gradient(s, f, phiL, phiR, gr + 2 * cl, gr + 2 * cr, 0.0);
                double diffusive = -p->alpha * fg.dphidn * A;
                double up = (Vn >= 0.0) ? phiL : phiR;
                ff[f] = Vn * up * A + diffusive;
            } else {
                const fvk_bc* bc = &s->bcs[s->fpatch[f]];
                if (bc->type == FVK_BC_NEUMANN) {
                    ff[f] = Vn * phiL * A + bc->value * A;
                } else {
                    facegrad fg = face_gradient(s, f, phiL, 0.0, gr + 2 * cl, zero2, bc->value);
                    double diffusive = -p->alpha * fg.dphidn * A;
                    double up = (Vn >= 0.0) ? phiL : bc->value;
                    ff[f] = Vn * up * A + diffusive;
                }
            }
        }
        free(bphi);
        free(gr);
    } else if (p->equation == FVK_EQ_EULER) {
        if (p->order == 2) limited_gradients(s, U);
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < nf; ++f) sv4(ff + 4 * f, euler_face_flux(s, U, f));
    } else {
        /* primitives, boundary fields, gradients, sources */
        double cp = g * p->gas_constant / (g - 1.0);
        double *u = malloc(sizeof(double) * nc), *v = malloc(sizeof(double) * nc), *T = malloc(sizeof(double) * nc);
        double *nul = malloc(sizeof(double) * nc), *nut_t = malloc(sizeof(double) * nc);
        double *bu = calloc(nf, sizeof(double)), *bv = calloc(nf, sizeof(double)), *bT = calloc(nf, sizeof(double));
        double *gu = malloc(sizeof(double) * 2 * nc), *gv = malloc(sizeof(double) * 2 * nc),
               *gT = malloc(sizeof(double) * 2 * nc);
        double *b1 = NULL, *b2 = NULL, *b3 = NULL, *b4 = NULL, *g1 = NULL, *g2 = NULL, *src = NULL;
        sst_src* sst = NULL;
        int sa = p->equation == FVK_EQ_RANS_SA, sstm = p->equation == FVK_EQ_RANS_SST;
        for (int64_t c = 0; c < nc; ++c) {
            st4 Uc = ld4(U + 4 * c);
            u[c] = Uc.ru / Uc.r;
            v[c] = Uc.rv / Uc.r;
            T[c] = temperature(Uc, g, p->gas_constant);
            nul[c] = p->mu / Uc.r;
            if (sa) nut_t[c] = sa_eddy(turb[c], nul[c], &p->sa);
        }
        /* build_boundary_fields, src/NavierStokesFVMSolver.cpp:55-84 */
        for (int64_t f = 0; f < nf; ++f) {
            int cl = s->fcell[2 * f];
            if (s->fcell[2 * f + 1] != -1) continue;
            const fvk_bc* bc = &s->bcs[s->fpatch[f]];
            if (bc->type == FVK_BC_WALL) {
                bu[f] = bc->wall_u;
                bv[f] = bc->wall_v;
                bT[f] = bc->is_isothermal_wall ? bc->wall_temperature : T[cl];
            } else {
                bu[f] = u[cl]; bv[f] = v[cl]; bT[f] = T[cl];
            }
        }
        gradient(s, p->gradient_scheme, u, 1, bu, gu);
        gradient(s, p->gradient_scheme, v, 1, bv, gv);
        gradient(s, p->gradient_scheme, T, 1, bT, gT);
        if (sa) {
            /* build_boundary_nut, src/RANSTurbulenceSASolver.cpp:98-115 */
            b1 = calloc(nf, sizeof(double));
            for (int64_t f = 0; f < nf; ++f) {
                if (s->fcell[2 * f + 1] != -1) continue;
                const fvk_bc* bc = &s->bcs[s->fpatch[f]];
                b1[f] = bc->type == FVK_BC_WALL ? 0.0 : (bc->type == FVK_BC_FARFIELD ? bc->farfield_nut : turb[s->fcell[2 * f]]);
            }
            g1 = malloc(sizeof(double) * 2 * nc);
            gradient(s, p->gradient_scheme, turb, 1, b1, g1);
            src = malloc(sizeof(double) * nc);
            for (int64_t c = 0; c < nc; ++c) {
                double om = fabs(gv[2 * c] - gu[2 * c + 1]);
                double P, D, CD;
                sa_source(turb[c], nul[c], om, s->wall_dist[c], g1[2 * c], g1[2 * c + 1], &p->sa, &P, &D, &CD);
                src[c] = P - D + CD;
            }
        }
        if (sstm) {
            /* build_boundary_k_omega, src/RANSTurbulenceSSTSolver.cpp:101-139 */
            double *kk = malloc(sizeof(double) * nc), *ww = malloc(sizeof(double) * nc);
            for (int64_t c = 0; c < nc; ++c) { kk[c] = turb[2 * c]; ww[c] = turb[2 * c + 1]; }
            b1 = calloc(nf, sizeof(double)); b2 = calloc(nf, sizeof(double));
            b3 = calloc(nf, sizeof(double)); b4 = calloc(nf, sizeof(double));
            for (int64_t f = 0; f < nf; ++f) {
                if (s->fcell[2 * f + 1] != -1) continue;
                int cl = s->fcell[2 * f];
                const fvk_bc* bc = &s->bcs[s->fpatch[f]];
                if (bc->type == FVK_BC_WALL) {
                    b1[f] = 0.0;
                    double nu1 = p->mu / U[4 * cl];
                    double d1 = face_normal_distance(s, f);
                    b2[f] = 60.0 * nu1 / (p->sst.beta1 * d1 * d1);
                    b3[f] = 0.0;
                    b4[f] = U[4 * cl] * b2[f];
                } else if (bc->type == FVK_BC_FARFIELD) {
                    b1[f] = bc->farfield_k; b2[f] = bc->farfield_omega;
                    b3[f] = bc->farfield_state[0] * bc->farfield_k;
                    b4[f] = bc->farfield_state[0] * bc->farfield_omega;
                } else {
                    b1[f] = kk[cl]; b2[f] = ww[cl];
                    b3[f] = U[4 * cl] * kk[cl]; b4[f] = U[4 * cl] * ww[cl];
                }
            }
            g1 = malloc(sizeof(double) * 2 * nc);
            g2 = malloc(sizeof(double) * 2 * nc);
            gradient(s, p->gradient_scheme, kk, 1, b1, g1);
            gradient(s, p->gradient_scheme, ww, 1, b2, g2);
            sst = malloc(sizeof(sst_src) * nc);
            for (int64_t c = 0; c < nc; ++c)
                sst[c] = sst_source(U[4 * c], kk[c], ww[c], nul[c], s->wall_dist[c], gu + 2 * c, gv + 2 * c, g1 + 2 * c,
                                    g2 + 2 * c, p->sst_variant, &p->sst, p->kato_launder);
            free(kk);
            free(ww);
        }
        /* face pass */
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < nf; ++f) {
            int cl = s->fcell[2 * f], cr = s->fcell[2 * f + 1];
            int b = cr == -1;
            st4 L = ld4(U + 4 * cl);
            st4 Rs = b ? ghost(s, f, L) : ld4(U + 4 * cr);
            st4 inv = num_flux(p, L, Rs, s->fnx[f], s->fny[f]);
            double nx = s->fnx[f], ny = s->fny[f], A = s->farea[f];
            double vx = 0.0, vy = 0.0, vE = 0.0; /* viscous rho_u, rho_v, E */
            double t_adv1 = 0.0, t_adv2 = 0.0, t_dif1 = 0.0, t_dif2 = 0.0;
            double Vn = 0.0;
            if (sa || sstm) {
                double uRg = Rs.ru / Rs.r, vRg = Rs.rv / Rs.r;
                Vn = 0.5 * ((u[cl] + uRg) * nx + (v[cl] + vRg) * ny);
            }
            if (sa) {
                double nR = b ? b1[f] : turb[cr];
                double up = (Vn >= 0.0) ? turb[cl] : nR;
                t_adv1 = Vn * up * A;
            }
            if (sstm) {
                double rkL = U[4 * cl] * turb[2 * cl];
                double rkR = b ? b3[f] : U[4 * cr] * turb[2 * cr];
                t_adv1 = Vn * ((Vn >= 0.0) ? rkL : rkR) * A;
                double rwL = U[4 * cl] * turb[2 * cl + 1];
                double rwR = b ? b4[f] : U[4 * cr] * turb[2 * cr + 1];
                t_adv2 = Vn * ((Vn >= 0.0) ? rwL : rwR) * A;
            }
            int visc = !b || s->bcs[s->fpatch[f]].type == FVK_BC_WALL;
            if (visc) {
                double zero[2] = {0.0, 0.0};
                facegrad fu = face_gradient(s, f, u[cl], b ? bu[f] : u[cr], gu + 2 * cl, b ? zero : gu + 2 * cr, bu[f]);
                facegrad fv = face_gradient(s, f, v[cl], b ? bv[f] : v[cr], gv + 2 * cl, b ? zero : gv + 2 * cr, bv[f]);
                facegrad fT = face_gradient(s, f, T[cl], b ? bT[f] : T[cr], gT + 2 * cl, b ? zero : gT + 2 * cr, bT[f]);
                double gux, guy, gvx, gvy, gTx, gTy;
                corrected_vector(fu, nx, ny, &gux, &guy);
                corrected_vector(fv, nx, ny, &gvx, &gvy);
                corrected_vector(fT, nx, ny, &gTx, &gTy);
                double mu_eff = p->mu, k_eff = p->mu * cp / p->prandtl;
                double tdvf = 0.0, nulR = 0.0, ntR = 0.0;
                if (sa) {
                    nulR = b ? nul[cl] : nul[cr];
                    ntR = b ? sa_eddy(b1[f], nulR, &p->sa) : nut_t[cr];
                    double rR = b ? U[4 * cl] : U[4 * cr];
                    tdvf = 0.5 * (U[4 * cl] * nut_t[cl] + rR * ntR);
                    mu_eff = p->mu + tdvf;
                    k_eff = cp * p->mu / p->prandtl + cp * tdvf / p->prandtl_t;
                }
                if (sstm) {
                    double ntL = sst[cl].nut, ntRr = b ? 0.0 : sst[cr].nut;
                    double rR = b ? U[4 * cl] : U[4 * cr];
                    tdvf = 0.5 * (U[4 * cl] * ntL + rR * ntRr);
                    mu_eff = p->mu + tdvf;
                    k_eff = cp * p->mu / p->prandtl + cp * tdvf / p->prandtl_t;
                }
                double uf = b ? bu[f] : 0.5 * (u[cl] + u[cr]);
                double vf = b ? bv[f] : 0.5 * (v[cl] + v[cr]);
                double div = gux + gvy;
                double txx = mu_eff * (2.0 * gux - (2.0 / 3.0) * div);
                double tyy = mu_eff * (2.0 * gvy - (2.0 / 3.0) * div);
                double txy = mu_eff * (guy + gvx);
                double qx = -k_eff * gTx, qy = -k_eff * gTy;
                if (b && !s->bcs[s->fpatch[f]].is_isothermal_wall) { qx = 0.0; qy = 0.0; }
                vx = txx * nx + txy * ny;
                vy = txy * nx + tyy * ny;
                vE = (uf * txx + vf * txy - qx) * nx + (uf * txy + vf * tyy - qy) * ny;
                if (sa) {
                    double nlf = 0.5 * (nul[cl] + nulR);
                    double ntf = 0.5 * (nut_t[cl] + ntR);
                    double zero2[2] = {0.0, 0.0};
                    facegrad fn = face_gradient(s, f, turb[cl], b ? b1[f] : turb[cr], g1 + 2 * cl,
                                                b ? zero2 : g1 + 2 * cr, b1[f]);
                    double dif = (nlf + ntf) / p->sa.sigma;
                    t_dif1 = -dif * fn.dphidn * A;
                }
                if (sstm) {
                    double F1R = b ? sst[cl].F1 : sst[cr].F1;
                    double F1f = b ? sst[cl].F1 : 0.5 * (sst[cl].F1 + F1R);
                    double sk = F1f * p->sst.sigma_k1 + (1.0 - F1f) * p->sst.sigma_k2;
                    double sw = F1f * p->sst.sigma_omega1 + (1.0 - F1f) * p->sst.sigma_omega2;
                    double mek = p->mu + sk * tdvf, mew = p->mu + sw * tdvf;
                    double zero2[2] = {0.0, 0.0};
                    facegrad fk = face_gradient(s, f, turb[2 * cl], b ? b1[f] : turb[2 * cr], g1 + 2 * cl,
                                                b ? zero2 : g1 + 2 * cr, b1[f]);
                    facegrad fw = face_gradient(s, f, turb[2 * cl + 1], b ? b2[f] : turb[2 * cr + 1], g2 + 2 * cl,
                                                b ? zero2 : g2 + 2 * cr, b2[f]);
                    t_dif1 = -mek * fk.dphidn * A;
                    t_dif2 = -mew * fw.dphidn * A;
                }
            }
            double* o = ff + (int64_t)nv * f;
            o[0] = A * (inv.r - 0.0);
            o[1] = A * (inv.ru - vx);
            o[2] = A * (inv.rv - vy);
            o[3] = A * (inv.E - vE);
            if (sa) o[4] = t_adv1 + t_dif1;
            if (sstm) { o[4] = t_adv1 + t_dif1; o[5] = t_adv2 + t_dif2; }
        }
        /* volumetric sources are added in the gather below */
        if (sa || sstm) {
            double* vs = malloc(sizeof(double) * nc * 2);
            for (int64_t c = 0; c < nc; ++c) {
                if (sa) vs[2 * c] = s->vol[c] * src[c];
                if (sstm) {
                    vs[2 * c] = s->vol[c] * (sst[c].kp - sst[c].kd);
                    vs[2 * c + 1] = s->vol[c] * (sst[c].wp - sst[c].wd + sst[c].cd);
                }
            }
            /* gather with sources */
#pragma omp parallel for schedule(static)
            for (int64_t c = 0; c < nc; ++c) {
                double r[6] = {0, 0, 0, 0, 0, 0};
                for (int k = s->cptr[c]; k < s->cptr[c + 1]; ++k) {
                    int f = s->cface[k];
                    const double* o = ff + (int64_t)nv * f;
                    if (s->fcell[2 * f] == (int)c) for (int q = 0; q < nv; ++q) r[q] = r[q] - o[q];
                    else for (int q = 0; q < nv; ++q) r[q] += o[q];
                }
                r[4] += vs[2 * c];
                if (sstm) r[5] += vs[2 * c + 1];
                for (int q = 0; q < nv; ++q) R[(int64_t)nv * c + q] = r[q];
            }
            free(vs);
            free(ff);
            free(u); free(v); free(T); free(nul); free(nut_t); free(bu); free(bv); free(bT);
            free(gu); free(gv); free(gT); free(b1); free(b2); free(b3); free(b4); free(g1); free(g2);
            free(src); free(sst);
            return;
        }
        free(u); free(v); free(T); free(nul); free(nut_t); free(bu); free(bv); free(bT);
        free(gu); free(gv); free(gT);
    }
    /* gather: the cell pass of the reference step()s (e.g. src/EulerFVMSolver.cpp:89-104),
     * summing each cell's faces in Cell::faces order, minus for the owner */
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
        double r[6] = {0, 0, 0, 0, 0, 0};
        for (int k = s->cptr[c]; k < s->cptr[c + 1]; ++k) {
            int f = s->cface[k];
            const double* o = ff + (int64_t)nv * f;
            if (s->fcell[2 * f] == (int)c) for (int q = 0; q < nv; ++q) r[q] = r[q] - o[q];
            else for (int q = 0; q < nv; ++q) r[q] += o[q];
        }
        for (int q = 0; q < nv; ++q) R[(int64_t)nv * c + q] = r[q];
    }
    free(ff);
}

/* ------------------------------------------------------------------------
 * step: dt from the step's initial state, then rk_stages residual
 * evaluations and updates.  rk_stages = 1 with a global dt is the
 * reference step() (e.g. src/EulerFVMSolver.cpp:58-112); the SST update
 * advances rho k, rho omega and divides by the new density
 * (src/RANSTurbulenceSSTSolver.cpp:335-353).
 * ---------------------------------------------------------------------- */

static const double RK_A[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0.5, 0}, {0, 0.75, 1.0 / 3.0}};
static const double RK_B[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0.5, 0}, {1, 0.25, 2.0 / 3.0}};

static int one_step(fvo_solver* s, double* res_out, double* dt_out) {
    const fvk_params* p = &s->p;
    const int nv = nvars_of(p->equation);
    const int nt = nv > 4 ? nv - 4 : 0;
    const int64_t nc = s->nc, nf = s->nf;
    int stages = p->rk_stages < 1 ? 1 : (p->rk_stages > 3 ? 3 : p->rk_stages);
    int diff = scalar_of(p->equation);
    int sstm = p->equation == FVK_EQ_RANS_SST;
    int su = diff ? 1 : 4; /* doubles per cell in U */

    /* dt: global minimum over faces (compute_dt) or per-cell minimum over the cell's faces */
    double* dtc = malloc(sizeof(double) * nc);
    double dtg = DBL_MAX;
    if (diff) {
        dtg = p->fixed_dt;
        for (int64_t c = 0; c < nc; ++c) dtc[c] = dtg;
    } else {
        double* dtf = malloc(sizeof(double) * nf);
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < nf; ++f) dtf[f] = dt_face(s, s->U, s->turb, f);
        for (int64_t f = 0; f < nf; ++f) dtg = smin(dtg, dtf[f]);
        if (p->local_dt) {
            for (int64_t c = 0; c < nc; ++c) {
                double d = DBL_MAX;
                for (int k = s->cptr[c]; k < s->cptr[c + 1]; ++k) d = smin(d, dtf[s->cface[k]]);
                dtc[c] = d;
            }
        } else {
            for (int64_t c = 0; c < nc; ++c) dtc[c] = dtg;
        }
        free(dtf);
    }
    if (dt_out) *dt_out = dtg;

    double* U0 = malloc(sizeof(double) * nc * su);
    memcpy(U0, s->U, sizeof(double) * nc * su);
    double* T0 = NULL;  /* conserved turbulence at the step start: nut or rho k, rho omega */
    if (nt) {
        T0 = malloc(sizeof(double) * nc * nt);
        for (int64_t c = 0; c < nc; ++c)
            for (int q = 0; q < nt; ++q) T0[nt * c + q] = sstm ? s->U[4 * c] * s->turb[nt * c + q] : s->turb[nt * c + q];
    }
    double* R = malloc(sizeof(double) * nc * nv);
    for (int st = 0; st < stages; ++st) {
        residual(s, s->U, s->turb, R);
        if (st == 0) {
            double sq[6] = {0, 0, 0, 0, 0, 0};
            for (int64_t c = 0; c < nc; ++c)
                for (int q = 0; q < nv; ++q) sq[q] += R[nv * c + q] * R[nv * c + q];
            for (int q = 0; q < nv; ++q) res_out[q] = sqrt(sq[q]);
        }
        double a = RK_A[stages][st], b = RK_B[stages][st];
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < nc; ++c) {
            double h = dtc[c] / s->vol[c];
            double* Uc = s->U + su * c;
            double rho_old = diff ? 0.0 : Uc[0];
            if (st == 0) {
                for (int q = 0; q < su; ++q) Uc[q] += h * R[nv * c + q];
            } else {
                for (int q = 0; q < su; ++q) Uc[q] = a * U0[su * c + q] + b * (Uc[q] + h * R[nv * c + q]);
            }
            if (nt) {
                double* tc = s->turb + nt * c;
                for (int q = 0; q < nt; ++q) {
                    double cur = sstm ? rho_old * tc[q] : tc[q];
                    double nw = (st == 0) ? cur + h * R[nv * c + 4 + q] : a * T0[nt * c + q] + b * (cur + h * R[nv * c + 4 + q]);
                    tc[q] = sstm ? nw / Uc[0] : nw;
                }
            }
        }
    }
    free(R); free(U0); free(T0); free(dtc);
    return FVK_OK;
}

/* ------------------------------------------------------------------------
 * public
 * ---------------------------------------------------------------------- */

int fvo_create(const fvk_mesh_desc* d, const fvk_bc* bcs, int32_t n_bcs, const fvk_params* params,
               const fvk_initial_condition* ic, fvo_solver** out) {
    fvo_solver* s = calloc(1, sizeof(fvo_solver));
    s->p = *params;
    s->nn = d->n_nodes; s->nc = d->n_cells; s->nf = d->n_faces;
    s->n_bcs = n_bcs;
    s->bcs = malloc(sizeof(fvk_bc) * (n_bcs > 0 ? n_bcs : 1));
    if (n_bcs > 0) memcpy(s->bcs, bcs, sizeof(fvk_bc) * n_bcs);
#define DUP(dst, src, n, T)                   \
    do {                                      \
        dst = malloc(sizeof(T) * (n));        \
        memcpy(dst, src, sizeof(T) * (n));    \
    } while (0)
    DUP(s->node_xy, d->node_xy, 2 * s->nn, double);
    DUP(s->fnode, d->face_nodes, 2 * s->nf, int32_t);
    DUP(s->fcell, d->face_cells, 2 * s->nf, int32_t);
    DUP(s->fpatch, d->face_patch, s->nf, int32_t);
    DUP(s->cptr, d->cell_face_ptr, s->nc + 1, int32_t);
    DUP(s->cface, d->cell_faces, s->cptr[s->nc], int32_t);
    DUP(s->vol, d->cell_volume, s->nc, double);
    DUP(s->cxy, d->cell_xy, 2 * s->nc, double);
    if (d->face_shift) DUP(s->shift, d->face_shift, 2 * s->nf, double);
#undef DUP
    s->fmid = malloc(sizeof(double) * 2 * s->nf);
    s->farea = malloc(sizeof(double) * s->nf);
    s->fnx = malloc(sizeof(double) * s->nf);
    s->fny = malloc(sizeof(double) * s->nf);
    compute_geometry(s);
    s->lsq_inv = malloc(sizeof(double) * 3 * s->nc);
    lsq_precompute(s);
    s->grad8 = calloc(8 * s->nc, sizeof(double));
    s->lim4 = calloc(4 * s->nc, sizeof(double));
    int nv = nvars_of(params->equation);
    int su = scalar_of(params->equation) ? 1 : 4;
    s->U = calloc(su * s->nc, sizeof(double));
    if (nv > 4) s->turb = calloc((nv - 4) * s->nc, sizeof(double));
    /* wall distance to no-slip walls (RANS constructors) */
    s->wall_dist = malloc(sizeof(double) * s->nc);
    {
        int32_t* wf = malloc(sizeof(int32_t) * (s->nf + 1));
        int64_t nw = 0;
        for (int64_t f = 0; f < s->nf; ++f)
            if (s->fcell[2 * f + 1] == -1 && s->fpatch[f] >= 0 && s->fpatch[f] < n_bcs && s->bcs[s->fpatch[f]].type == FVK_BC_WALL)
                wf[nw++] = (int32_t)f;
        if (nw == 0) for (int64_t c = 0; c < s->nc; ++c) s->wall_dist[c] = INFINITY;
        else if (params->equation == FVK_EQ_RANS_SA || params->equation == FVK_EQ_RANS_SST)
            wall_distance_impl(s->nc, s->cxy, s->node_xy, s->fnode, wf, nw, s->wall_dist);
        else for (int64_t c = 0; c < s->nc; ++c) s->wall_dist[c] = INFINITY;
        free(wf);
    }
    /* initial condition (constructors, e.g. src/EulerFVMSolver.cpp:24-37) */
    if (ic && ic->mode != FVK_IC_NONE) {
        double g = params->gamma;
        for (int64_t c = 0; c < s->nc; ++c) {
            if (params->equation == FVK_EQ_DIFFUSION) {
                /* src/UnstructuredFVMSolver.cpp:13-21 */
                double x = s->cxy[2 * c], y = s->cxy[2 * c + 1];
                s->U[c] = (sqrt(x * x + y * y) < ic->initial_radius) ? ic->initial_value : 0.0;
                continue;
            }
            if (params->equation == FVK_EQ_ADVECTION_DIFFUSION) {
                /* src/AdvectionDiffusionFVMSolver.cpp:22-31 (hypot) */
                s->U[c] = (hypot(s->cxy[2 * c], s->cxy[2 * c + 1]) < ic->initial_radius) ? ic->initial_value : 0.0;
                continue;
            }
            st4 v;
            if (ic->mode == FVK_IC_FREESTREAM) v = from_primitive(ic->rho, ic->u, ic->v, ic->p, g);
            else if (s->cxy[2 * c] < ic->x0) v = from_primitive(ic->rho_l, ic->u_l, ic->v_l, ic->p_l, g);
            else v = from_primitive(ic->rho_r, ic->u_r, ic->v_r, ic->p_r, g);
            sv4(s->U + 4 * c, v);
            if (nv == 5) s->turb[c] = ic->initial_nut;
            if (nv == 6) { s->turb[2 * c] = ic->initial_k; s->turb[2 * c + 1] = ic->initial_omega; }
        }
    } else if (ic && !scalar_of(params->equation)) {
        for (int64_t c = 0; c < s->nc; ++c) {
            if (nv == 5) s->turb[c] = ic->initial_nut;
            if (nv == 6) { s->turb[2 * c] = ic->initial_k; s->turb[2 * c + 1] = ic->initial_omega; }
        }
    }
    *out = s;
    return FVK_OK;
}

void fvo_destroy(fvo_solver* s) {
    if (!s) return;
    free(s->bcs); free(s->node_xy); free(s->fnode); free(s->fcell); free(s->fpatch); free(s->cptr); free(s->cface);
    free(s->vol); free(s->cxy); free(s->shift); free(s->fmid); free(s->farea); free(s->fnx); free(s->fny);
    free(s->lsq_inv); free(s->wall_dist); free(s->U); free(s->turb); free(s->grad8); free(s->lim4);
    free(s);
}

int fvo_step(fvo_solver* s, int32_t nsteps, double* residual_hist, double* dt_hist) {
    int nv = nvars_of(s->p.equation);
    for (int i = 0; i < nsteps; ++i) {
        double dt;
        one_step(s, s->res, &dt);
        if (residual_hist) memcpy(residual_hist + (int64_t)i * nv, s->res, sizeof(double) * nv);
        if (dt_hist) dt_hist[i] = dt;
    }
    return FVK_OK;
}

int fvo_residual(const fvo_solver* s, double* norms) {
    memcpy(norms, s->res, sizeof(double) * nvars_of(s->p.equation));
    return FVK_OK;
}

int fvo_set_field(fvo_solver* s, const double* U4, const double* turb) {
    int nv = nvars_of(s->p.equation);
    int su = scalar_of(s->p.equation) ? 1 : 4;
    memcpy(s->U, U4, sizeof(double) * su * s->nc);
    if (turb && nv > 4) memcpy(s->turb, turb, sizeof(double) * (nv - 4) * s->nc);
    memset(s->res, 0, sizeof s->res);
    return FVK_OK;
}

int fvo_get_field(const fvo_solver* s, double* U4, double* turb) {
    int nv = nvars_of(s->p.equation);
    int su = scalar_of(s->p.equation) ? 1 : 4;
    memcpy(U4, s->U, sizeof(double) * su * s->nc);
    if (turb && nv > 4) memcpy(turb, s->turb, sizeof(double) * (nv - 4) * s->nc);
    return FVK_OK;
}

int fvo_set_cfl(fvo_solver* s, double cfl) {
    s->p.cfl = cfl;
    return FVK_OK;
}

int fvo_residual_eval(fvo_solver* s, double* R) {
    residual(s, s->U, s->turb, R);
    return FVK_OK;
}

int fvo_gradient(fvo_solver* s, int32_t scheme, const double* phi, const double* bphi, double* grad) {
    gradient(s, scheme, phi, 1, bphi, grad);
    return FVK_OK;
}

int fvo_wall_distance(fvo_solver* s, double* dist) {
    memcpy(dist, s->wall_dist, sizeof(double) * s->nc);
    return FVK_OK;
}

int fvo_wall_distance_faces(const fvk_mesh_desc* d, const int32_t* wf, int64_t nw, double* dist) {
    if (nw == 0) {
        for (int64_t c = 0; c < d->n_cells; ++c) dist[c] = INFINITY;
        return FVK_OK;
    }
    wall_distance_impl(d->n_cells, d->cell_xy, d->node_xy, d->face_nodes, wf, nw, dist);
    return FVK_OK;
}

/* SST F2 and eddy viscosity with the guards of src/SSTKOmega.cpp (sst_F2,
 * sst_eddy_viscosity), as compute_wall_traction_samples calls them */
static double sst_F2_ref(double k, double w, double nu, double d, const fvk_sst_constants* c) {
    if (k <= 0.0 || w <= 0.0) return 0.0;
    double d2 = d * d;
    double a2 = smax(2.0 * sqrt(k) / (c->beta_star * w * d), 500.0 * nu / (d2 * w));
    return tanh(a2 * a2);
}
static double sst_eddy_ref(double k, double w, double S, double Om, double F2, int variant,
                           const fvk_sst_constants* c) {
    if (k <= 0.0) return 0.0;
    double lim = (variant == FVK_SST_VORTICITY) ? Om : S;
    return c->a1 * k / smax(c->a1 * w, lim * F2);
}

/* Wall samples and force reports.
 * Samples: compute_wall_traction (src/WallTraction.cpp:95-143) with the
 * inputs of compute_wall_traction_samples (src/NavierStokesFVMSolver.cpp:333-352,
 * src/RANSTurbulenceSASolver.cpp:353-374, src/RANSTurbulenceSSTSolver.cpp:435-464);
 * Euler, or pressure_only, gives pressure-only samples (mu_eff = tau = 0);
 * coefficients: src/WallTraction.cpp:145-161.
 * Reports: compute_wall_forces (src/WallTraction.cpp:163-216), patches in
 * ascending id then the total. */
int fvo_wall_reports(fvo_solver* s, const int32_t* wf, int64_t n, const fvk_wall_reference* ref, int32_t pressure_only,
                     fvk_wall_sample* samples, fvk_wall_force_report* reports, int32_t cap, int32_t* n_reports) {
    const fvk_params* p = &s->p;
    const int64_t nc = s->nc, nf = s->nf;
    const double g = p->gamma;
    const int eq = p->equation;
    if (scalar_of(eq)) return FVK_ERR_UNSUPPORTED;
    for (int64_t i = 0; i < n; ++i)
        if (wf[i] < 0 || wf[i] >= nf || s->fcell[2 * wf[i] + 1] != -1) return FVK_ERR_INVALID;
    const int visc = eq != FVK_EQ_EULER && !pressure_only;
    double *u = NULL, *v = NULL, *bu = NULL, *bv = NULL, *gu = NULL, *gv = NULL;
    if (visc) {
        u = malloc(sizeof(double) * nc); v = malloc(sizeof(double) * nc);
        bu = calloc(nf, sizeof(double)); bv = calloc(nf, sizeof(double));
        gu = malloc(sizeof(double) * 2 * nc); gv = malloc(sizeof(double) * 2 * nc);
        for (int64_t c = 0; c < nc; ++c) {
            st4 Uc = ld4(s->U + 4 * c);
            u[c] = Uc.ru / Uc.r;
            v[c] = Uc.rv / Uc.r;
        }
        /* build_boundary_fields, src/NavierStokesFVMSolver.cpp:55-84 (u, v) */
        for (int64_t f = 0; f < nf; ++f) {
            if (s->fcell[2 * f + 1] != -1) continue;
            const fvk_bc* bc = &s->bcs[s->fpatch[f]];
            int cl = s->fcell[2 * f];
            if (bc->type == FVK_BC_WALL) { bu[f] = bc->wall_u; bv[f] = bc->wall_v; }
            else { bu[f] = u[cl]; bv[f] = v[cl]; }
        }
        gradient(s, p->gradient_scheme, u, 1, bu, gu);
        gradient(s, p->gradient_scheme, v, 1, bv, gv);
    }
    /* compute_wall_forces' reference directions, :168-172 */
    double v_mag = hypot(ref->velocity_x_ref, ref->velocity_y_ref);
    double dir_x = ref->velocity_x_ref / v_mag, dir_y = ref->velocity_y_ref / v_mag;
    double lift_x = -dir_y, lift_y = dir_x;
    double q = 0.5 * ref->rho_ref * v_mag * v_mag * ref->length_ref;
    double q2 = q * ref->length_ref;
    double v2 = ref->velocity_x_ref * ref->velocity_x_ref + ref->velocity_y_ref * ref->velocity_y_ref;
    int np = s->n_bcs;
    double* acc = calloc((size_t)(np + 1) * 5, sizeof(double));
    int* cnt = calloc((size_t)(np + 1), sizeof(int));
    const double zero2[2] = {0.0, 0.0};
    for (int64_t i = 0; i < n; ++i) {
        int f = wf[i], cl = s->fcell[2 * f], patch = s->fpatch[f];
        st4 Uc = ld4(s->U + 4 * cl);
        double pc = pressure(Uc, g);
        double nx = s->fnx[f], ny = s->fny[f];
        double mu = 0.0, tau = 0.0;
        if (visc) {
            double gux, guy, gvx, gvy;
            corrected_vector(face_gradient(s, f, u[cl], 0.0, gu + 2 * cl, zero2, bu[f]), nx, ny, &gux, &guy);
            corrected_vector(face_gradient(s, f, v[cl], 0.0, gv + 2 * cl, zero2, bv[f]), nx, ny, &gvx, &gvy);
            mu = p->mu;
            if (eq == FVK_EQ_RANS_SA) {
                mu = p->mu + Uc.r * sa_eddy(s->turb[cl], p->mu / Uc.r, &p->sa);
            } else if (eq == FVK_EQ_RANS_SST) {
                const double* a = gu + 2 * cl;
                const double* b = gv + 2 * cl;
                double s11 = a[0], s22 = b[1], s12 = 0.5 * (a[1] + b[0]);
                double S = sqrt(2.0 * (s11 * s11 + s22 * s22 + 2.0 * s12 * s12));
                double Om = fabs(b[0] - a[1]);
                double k = s->turb[2 * cl], om = s->turb[2 * cl + 1];
                double F2 = sst_F2_ref(k, om, p->mu / Uc.r, s->wall_dist[cl], &p->sst);
                mu = p->mu + Uc.r * sst_eddy_ref(k, om, S, Om, F2, p->sst_variant, &p->sst);
            }
            double div_v = gux + gvy;
            double txx = mu * (2.0 * gux - (2.0 / 3.0) * div_v);
            double tyy = mu * (2.0 * gvy - (2.0 / 3.0) * div_v);
            double txy = mu * (guy + gvx);
            double trx = txx * nx + txy * ny;
            double try_ = txy * nx + tyy * ny;
            tau = -(trx * (-ny) + try_ * nx);
        }
        double tx = -ny, ty = nx;
        double A = s->farea[f];
        double xm = s->fmid[2 * f], ym = s->fmid[2 * f + 1];
        if (samples) {
            fvk_wall_sample* o = samples + i;
            o->face_index = f;
            o->patch_id = patch;
            o->x_mid = xm; o->y_mid = ym;
            o->p = pc; o->rho = Uc.r;
            o->effective_viscosity = mu;
            o->tau_wall = tau;
            o->tx = tx; o->ty = ty;
            o->y_wall_normal = face_normal_distance(s, f);
            o->cf = tau / (0.5 * ref->rho_ref * v2);
            o->cp = (pc - ref->p_ref) / (0.5 * ref->rho_ref * v2);
            double u_tau = sqrt(fabs(tau) / Uc.r);
            o->y_plus = Uc.r * u_tau * o->y_wall_normal / mu;
        }
        double fx = A * tau * tx, fy = A * tau * ty;
        double px = A * pc * nx, py = A * pc * ny;
        double rx = xm - ref->moment_reference_x, ry = ym - ref->moment_reference_y;
        double dfx = fx + px, dfy = fy + py;
        double mom = rx * dfy - ry * dfx;
        double fd = fx * dir_x + fy * dir_y;
        double pd = px * dir_x + py * dir_y;
        double lift = (fx + px) * lift_x + (fy + py) * lift_y;
        int keys[2] = {patch, np};
        for (int k = 0; k < 2; ++k) {
            double* a = acc + 5 * keys[k];
            a[0] += fd; a[1] += pd; a[2] += fd + pd; a[3] += lift; a[4] += mom;
            cnt[keys[k]] += 1;
        }
    }
    int k = 0;
    for (int pid = 0; pid <= np; ++pid) {
        if (cnt[pid] == 0) continue;
        if (k < cap) {
            fvk_wall_force_report* r = reports + k;
            const double* a = acc + 5 * pid;
            memset(r, 0, sizeof *r);
            r->patch_id = pid == np ? -1 : pid;
            r->friction_drag = a[0]; r->pressure_drag = a[1]; r->total_drag = a[2]; r->lift = a[3]; r->moment = a[4];
            r->cd_friction = a[0] / q; r->cd_pressure = a[1] / q; r->cd_total = a[2] / q;
            r->cl = a[3] / q; r->cm = a[4] / q2;
        }
        ++k;
    }
    *n_reports = k;
    free(acc); free(cnt);
    free(u); free(v); free(bu); free(bv); free(gu); free(gv);
    return k > cap ? FVK_ERR_INVALID : FVK_OK;
}

/* pressure lift and drag over every wall face in face order: the total of
 * fvo_wall_reports' pressure-only samples */
int fvo_wall_forces(fvo_solver* s, double rho_ref, double vx, double vy, double L, double* cl, double* cd) {
    int32_t* wf = malloc(sizeof(int32_t) * (size_t)(s->nf > 0 ? s->nf : 1));
    int64_t n = 0;
    for (int64_t f = 0; f < s->nf; ++f)
        if (s->fcell[2 * f + 1] == -1 && s->bcs[s->fpatch[f]].type == FVK_BC_WALL) wf[n++] = (int32_t)f;
    fvk_wall_reference ref = {rho_ref, vx, vy, 0.0, L, 0.0, 0.0};
    fvk_wall_force_report* rep = malloc(sizeof(fvk_wall_force_report) * (size_t)(s->n_bcs + 1));
    int32_t nrep = 0;
    int rc = fvo_wall_reports(s, wf, n, &ref, 1, NULL, rep, s->n_bcs + 1, &nrep);
    *cl = nrep > 0 ? rep[nrep - 1].cl : 0.0;
    *cd = nrep > 0 ? rep[nrep - 1].cd_total : 0.0;
    free(wf);
    free(rep);
    return rc;
}

int fvo_last_gradients(const fvo_solver* s, double* grad, double* phi) {
    if (grad) memcpy(grad, s->grad8, sizeof(double) * 8 * s->nc);
    if (phi) memcpy(phi, s->lim4, sizeof(double) * 4 * s->nc);
    return FVK_OK;
}
