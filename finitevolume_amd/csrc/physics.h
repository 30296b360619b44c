// Device physics for the explicit residual: ideal-gas EOS, the three
// numerical fluxes, boundary ghost states and the Venkatakrishnan limiter.
//
// Every function evaluates the same sequence of IEEE double operations as
// the reference (cited per function) so that, with FMA contraction off
// (-ffp-contract=off) and correctly rounded divide/sqrt, first-order
// results match the reference bit for bit.  min/max follow std::min /
// std::max (NaN in the second argument is ignored), not fmin/fmax, except
// for the limiter's extremes (venkat_extend), where the two give the same
// limiter.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace fvk {

#define FVK_HD __host__ __device__ __forceinline__

struct State {
    double r, ru, rv, E;
};

FVK_HD double smin(double a, double b) { return (b < a) ? b : a; }

// x / y.  On the device: the reciprocal refined by two Newton steps and the
// quotient corrected once (Markstein), which is the sequence the compiler
// emits for x / y minus its range scaling (v_div_scale) and special-case
// fixup (v_div_fixup).  For finite nonzero operands whose exponents and
// quotient stay far from the overflow and denormal limits, as every state
// and geometry value here does, the scaling is the identity and the fixup
// returns its input, so the bits are those of the IEEE quotient, in 8
// instructions instead of 11.  Host code divides.
FVK_HD double fdiv(double x, double y) {
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(y);
    double e = fma(-y, r, 1.0);
    r = fma(r, e, r);
    e = fma(-y, r, 1.0);
    r = fma(r, e, r);
    const double q = x * r;
    const double rem = fma(-y, q, x);
    return fma(rem, r, q);
#else
    return x / y;
#endif
}
FVK_HD double smax(double a, double b) { return (a < b) ? b : a; }

// from_primitive, include/EulerState.h:45-52
FVK_HD State from_primitive(double rho, double u, double v, double p, double g) {
    State s;
    s.r = rho;
    s.ru = rho * u;
    s.rv = rho * v;
    s.E = fdiv(p, g - 1.0) + 0.5 * rho * (u * u + v * v);
    return s;
}

// pressure, include/EulerState.h:58-62 (velocity re-divided from momentum)
FVK_HD double pressure(const State& s, double g) {
    double u = fdiv(s.ru, s.r), v = fdiv(s.rv, s.r);
    return (g - 1.0) * (s.E - 0.5 * s.r * (u * u + v * v));
}

// sound_speed, include/EulerState.h:80-82
FVK_HD double sound_speed(const State& s, double g) { return sqrt(fdiv(g * pressure(s, g), s.r)); }

// temperature, include/EulerState.h:73-75
FVK_HD double temperature(const State& s, double g, double R) { return fdiv(pressure(s, g), s.r * R); }

// primitive variables of a state: rho, u, v, p (u = rho_u / rho, p from
// pressure(); the bits of the reference's own conversions)
struct Prim {
    double r, u, v, p;
};

FVK_HD Prim to_prim(const State& s, double g) { return Prim{s.r, fdiv(s.ru, s.r), fdiv(s.rv, s.r), pressure(s, g)}; }

// flux, include/EulerState.h:89-101
FVK_HD State phys_flux(const State& s, double nx, double ny, double g) {
    double u = fdiv(s.ru, s.r), v = fdiv(s.rv, s.r);
    double Vn = u * nx + v * ny;
    double p = pressure(s, g);
    return State{s.r * Vn, s.ru * Vn + p * nx, s.rv * Vn + p * ny, (s.E + p) * Vn};
}

// rusanov_flux, include/RusanovFlux.h:33-43
FVK_HD State rusanov(const State& L, const State& R, double nx, double ny, double g) {
    double VnL = fdiv(L.ru * nx + L.rv * ny, L.r);
    double VnR = fdiv(R.ru * nx + R.rv * ny, R.r);
    double Smax = smax(fabs(VnL) + sound_speed(L, g), fabs(VnR) + sound_speed(R, g));
    State FL = phys_flux(L, nx, ny, g), FR = phys_flux(R, nx, ny, g);
    double h = 0.5 * Smax;
    return State{0.5 * (FL.r + FR.r) - h * (R.r - L.r), 0.5 * (FL.ru + FR.ru) - h * (R.ru - L.ru),
                 0.5 * (FL.rv + FR.rv) - h * (R.rv - L.rv), 0.5 * (FL.E + FR.E) - h * (R.E - L.E)};
}

// hllc_flux, include/HllcFlux.h:35-76: Davis wave speeds, supersonic
// early-outs, contact speed (Toro 10.37), star state (10.39), F_K + S_K (U*_K - U_K)
FVK_HD State hllc(const State& L, const State& R, double nx, double ny, double g) {
    double tx = -ny, ty = nx;
    double rL = L.r, rR = R.r;
    double uL = fdiv(L.ru, rL), vL = fdiv(L.rv, rL), uR = fdiv(R.ru, rR), vR = fdiv(R.rv, rR);
    double VnL = uL * nx + vL * ny, VnR = uR * nx + vR * ny;
    double VtL = uL * tx + vL * ty, VtR = uR * tx + vR * ty;
    double pL = pressure(L, g), pR = pressure(R, g);
    double cL = sound_speed(L, g), cR = sound_speed(R, g);
    double SL = smin(VnL - cL, VnR - cR);
    double SR = smax(VnL + cL, VnR + cR);
    if (SL >= 0.0) return phys_flux(L, nx, ny, g);
    if (SR <= 0.0) return phys_flux(R, nx, ny, g);
    double Ss = fdiv(pR - pL + rL * VnL * (SL - VnL) - rR * VnR * (SR - VnR), rL * (SL - VnL) - rR * (SR - VnR));
    bool left = Ss >= 0.0;
    double SK = left ? SL : SR, rK = left ? rL : rR, VnK = left ? VnL : VnR, VtK = left ? VtL : VtR;
    double pK = left ? pL : pR;
    State UK = left ? L : R;
    double factor = fdiv(rK * (SK - VnK), SK - Ss);
    double mn = factor * Ss, mt = factor * VtK;
    double Es = factor * (fdiv(UK.E, rK) + (Ss - VnK) * (Ss + fdiv(pK, rK * (SK - VnK))));
    State F = phys_flux(UK, nx, ny, g);
    return State{F.r + SK * (factor - UK.r), F.ru + SK * ((mn * nx - mt * ny) - UK.ru),
                 F.rv + SK * ((mn * ny + mt * nx) - UK.rv), F.E + SK * (Es - UK.E)};
}

// ---- exact Riemann solver, include/ExactRiemannFlux.h:28-234 (Toro ch. 4) ----
FVK_HD double er_f(double p, double rK, double pK, double cK, double g) {
    if (p > pK) {
        double A = 2.0 / ((g + 1.0) * rK), B = (g - 1.0) / (g + 1.0) * pK;
        return (p - pK) * sqrt(A / (p + B));
    }
    return (2.0 * cK / (g - 1.0)) * (pow(p / pK, (g - 1.0) / (2.0 * g)) - 1.0);
}
FVK_HD double er_fd(double p, double rK, double pK, double cK, double g) {
    if (p > pK) {
        double A = 2.0 / ((g + 1.0) * rK), B = (g - 1.0) / (g + 1.0) * pK;
        return sqrt(A / (p + B)) * (1.0 - 0.5 * (p - pK) / (p + B));
    }
    return (1.0 / (rK * cK)) * pow(p / pK, -(g + 1.0) / (2.0 * g));
}
FVK_HD void er_fan(double rK, double uK, double pK, double cK, double g, double S, double sgn, double* ro,
                   double* uo, double* po) {
    // sgn = +1: left fan (Toro 4.56); sgn = -1: right fan (4.63)
    double cf = (2.0 / (g + 1.0)) * (cK + sgn * 0.5 * (g - 1.0) * (uK - S));
    *uo = (2.0 / (g + 1.0)) * (sgn * cK + 0.5 * (g - 1.0) * uK + S);
    *ro = rK * pow(cf / cK, 2.0 / (g - 1.0));
    *po = pK * pow(cf / cK, 2.0 * g / (g - 1.0));
}

FVK_HD State exact_flux(const State& L, const State& R, double nx, double ny, double g, double tol, int maxit) {
    double tx = -ny, ty = nx;
    double rL = L.r, rR = R.r;
    double uxL = L.ru / rL, uyL = L.rv / rL, uxR = R.ru / rR, uyR = R.rv / rR;
    double uL = uxL * nx + uyL * ny, vL = uxL * tx + uyL * ty;
    double uR = uxR * nx + uyR * ny, vR = uxR * tx + uyR * ty;
    double pL = pressure(L, g), pR = pressure(R, g);
    double cL = sound_speed(L, g), cR = sound_speed(R, g);
    double rs, us, ps, vs;
    if (2.0 * (cL + cR) / (g - 1.0) <= uR - uL) {
        // vacuum generated (Toro 4.7)
        double SLv = uL + 2.0 * cL / (g - 1.0), SRv = uR - 2.0 * cR / (g - 1.0);
        if (0.0 <= uL - cL) { rs = rL; us = uL; ps = pL; vs = vL; }
        else if (0.0 <= SLv) { er_fan(rL, uL, pL, cL, g, 0.0, 1.0, &rs, &us, &ps); vs = vL; }
        else if (0.0 < SRv) { return State{0.0, 0.0, 0.0, 0.0}; }
        else if (0.0 <= uR + cR) { er_fan(rR, uR, pR, cR, g, 0.0, -1.0, &rs, &us, &ps); vs = vR; }
        else { rs = rR; us = uR; ps = pR; vs = vR; }
    } else {
        // two-shock-seeded Newton iteration for p* (Toro 4.46/4.47, 4.85/4.86)
        double ppv = smax(1e-6, 0.5 * (pL + pR) - 0.125 * (uR - uL) * (rL + rR) * (cL + cR));
        double AL = 2.0 / ((g + 1.0) * rL), BL = (g - 1.0) / (g + 1.0) * pL;
        double AR = 2.0 / ((g + 1.0) * rR), BR = (g - 1.0) / (g + 1.0) * pR;
        double gL = sqrt(AL / (ppv + BL)), gR = sqrt(AR / (ppv + BR));
        double p = smax(1e-6, (gL * pL + gR * pR - (uR - uL)) / (gL + gR));
        for (int it = 0; it < maxit; ++it) {
            double f = er_f(p, rL, pL, cL, g) + er_f(p, rR, pR, cR, g) + (uR - uL);
            double fd = er_fd(p, rL, pL, cL, g) + er_fd(p, rR, pR, cR, g);
            double pn = p - f / fd;
            if (pn < 1e-10) pn = 1e-10;
            double ch = 2.0 * fabs(pn - p) / (pn + p);
            p = pn;
            if (ch < tol) break;
        }
        double pst = p;
        double ust = 0.5 * (uL + uR) + 0.5 * (er_f(pst, rR, pR, cR, g) - er_f(pst, rL, pL, cL, g));
        if (0.0 <= ust) {
            vs = vL;
            if (pst > pL) {
                double SLs = uL - cL * sqrt((g + 1.0) / (2.0 * g) * (pst / pL) + (g - 1.0) / (2.0 * g));
                if (0.0 <= SLs) { rs = rL; us = uL; ps = pL; }
                else {
                    double q = pst / pL;
                    rs = rL * (q + (g - 1.0) / (g + 1.0)) / ((g - 1.0) / (g + 1.0) * q + 1.0);
                    us = ust; ps = pst;
                }
            } else {
                double csL = cL * pow(pst / pL, (g - 1.0) / (2.0 * g));
                double SHL = uL - cL, STL = ust - csL;
                if (0.0 <= SHL) { rs = rL; us = uL; ps = pL; }
                else if (0.0 >= STL) { rs = rL * pow(pst / pL, 1.0 / g); us = ust; ps = pst; }
                else er_fan(rL, uL, pL, cL, g, 0.0, 1.0, &rs, &us, &ps);
            }
        } else {
            vs = vR;
            if (pst > pR) {
                double SRs = uR + cR * sqrt((g + 1.0) / (2.0 * g) * (pst / pR) + (g - 1.0) / (2.0 * g));
                if (0.0 >= SRs) { rs = rR; us = uR; ps = pR; }
                else {
                    double q = pst / pR;
                    rs = rR * (q + (g - 1.0) / (g + 1.0)) / ((g - 1.0) / (g + 1.0) * q + 1.0);
                    us = ust; ps = pst;
                }
            } else {
                double csR = cR * pow(pst / pR, (g - 1.0) / (2.0 * g));
                double SHR = uR + cR, STR = ust + csR;
                if (0.0 >= SHR) { rs = rR; us = uR; ps = pR; }
                else if (0.0 <= STR) { rs = rR * pow(pst / pR, 1.0 / g); us = ust; ps = pst; }
                else er_fan(rR, uR, pR, cR, g, 0.0, -1.0, &rs, &us, &ps);
            }
        }
    }
    double uxs = us * nx - vs * ny, uys = us * ny + vs * nx;
    return phys_flux(from_primitive(rs, uxs, uys, ps, g), nx, ny, g);
}

// Flux scheme codes (NumericalFluxScheme order, include/EulerFVMSolver.h:40-44)
enum : int { FLUX_RUSANOV = 0, FLUX_HLLC = 1, FLUX_EXACT = 2 };

template <int SCHEME>
FVK_HD State num_flux(const State& L, const State& R, double nx, double ny, double g, double tol, int maxit) {
    if constexpr (SCHEME == FLUX_HLLC) return hllc(L, R, nx, ny, g);
    else if constexpr (SCHEME == FLUX_EXACT) return exact_flux(L, R, nx, ny, g, tol, maxit);
    else return rusanov(L, R, nx, ny, g);
}

// Boundary codes carried per boundary face.
enum : int { BC_WALL = 0, BC_FARFIELD = 1, BC_OUTFLOW = 2 };

// Euler slip-wall ghost, src/EulerFVMSolver.cpp:185-192: reflect the normal velocity.
FVK_HD State ghost_slip_wall(const State& L, double nx, double ny, double g) {
    double u = fdiv(L.ru, L.r), v = fdiv(L.rv, L.r);
    double Vn = u * nx + v * ny;
    return from_primitive(L.r, u - 2.0 * Vn * nx, v - 2.0 * Vn * ny, pressure(L, g), g);
}

// No-slip wall ghost, src/NavierStokesFVMSolver.cpp:280-283: mirror both
// velocity components about the wall velocity.
FVK_HD State ghost_noslip_wall(const State& L, double wu, double wv, double g) {
    double u = fdiv(L.ru, L.r), v = fdiv(L.rv, L.r);
    double p = pressure(L, g);
    return from_primitive(L.r, 2.0 * wu - u, 2.0 * wv - v, p, g);
}

// Euler ghost state by boundary type, src/EulerFVMSolver.cpp:174-193:
// farfield -> the prescribed conserved state, outflow -> copy, wall -> slip mirror
FVK_HD State euler_ghost_state(int type, const double* ff, const State& L, double nx, double ny, double g) {
    if (type == BC_FARFIELD) return State{ff[0], ff[1], ff[2], ff[3]};
    if (type == BC_OUTFLOW) return L;
    return ghost_slip_wall(L, nx, ny, g);
}

// ---- second-order faces on primitive face states (extension) ----
// A MUSCL face state is primitive, W = (rho, u, v, p).  The fluxes below
// take it as it is and form conserved quantities only where the flux uses
// them, instead of converting W to conserved variables and dividing them
// back.  Each equals its conserved counterpart of from_primitive(W) in exact
// arithmetic; the rounding is this formulation's own.

// specific total energy E / rho of a primitive state
FVK_HD double specific_energy(const Prim& w, double g) {
    return fdiv(w.p, (g - 1.0) * w.r) + 0.5 * (w.u * w.u + w.v * w.v);
}

// physical flux of a primitive state with total energy E and normal velocity Vn
FVK_HD State phys_flux_w(const Prim& w, double E, double Vn, double nx, double ny) {
    const double m = w.r * Vn;
    return State{m, m * w.u + w.p * nx, m * w.v + w.p * ny, (E + w.p) * Vn};
}

// rusanov() on primitive states
FVK_HD State rusanov_w(const Prim& L, const Prim& R, double nx, double ny, double g) {
    const double VnL = L.u * nx + L.v * ny, VnR = R.u * nx + R.v * ny;
    const double Smax = smax(fabs(VnL) + sqrt(fdiv(g * L.p, L.r)), fabs(VnR) + sqrt(fdiv(g * R.p, R.r)));
    const double EL = L.r * specific_energy(L, g), ER = R.r * specific_energy(R, g);
    const State FL = phys_flux_w(L, EL, VnL, nx, ny), FR = phys_flux_w(R, ER, VnR, nx, ny);
    const double h = 0.5 * Smax;
    return State{0.5 * (FL.r + FR.r) - h * (R.r - L.r), 0.5 * (FL.ru + FR.ru) - h * (R.r * R.u - L.r * L.u),
                 0.5 * (FL.rv + FR.rv) - h * (R.r * R.v - L.r * L.v), 0.5 * (FL.E + FR.E) - h * (ER - EL)};
}

// hllc() on primitive states: Davis wave speeds, supersonic early-outs,
// contact speed (Toro 10.37), star state (10.39), F_K + S_K (U*_K - U_K);
// m_K = rho_K (S_K - Vn_K) is the mass flux through the outer wave
FVK_HD State hllc_w(const Prim& L, const Prim& R, double nx, double ny, double g) {
    const double VnL = L.u * nx + L.v * ny, VnR = R.u * nx + R.v * ny;
    const double cL = sqrt(fdiv(g * L.p, L.r)), cR = sqrt(fdiv(g * R.p, R.r));
    const double SL = smin(VnL - cL, VnR - cR);
    const double SR = smax(VnL + cL, VnR + cR);
    if (SL >= 0.0) return phys_flux_w(L, L.r * specific_energy(L, g), VnL, nx, ny);
    if (SR <= 0.0) return phys_flux_w(R, R.r * specific_energy(R, g), VnR, nx, ny);
    const double mL = L.r * (SL - VnL), mR = R.r * (SR - VnR);
    const double Ss = fdiv(R.p - L.p + mL * VnL - mR * VnR, mL - mR);
    const bool left = Ss >= 0.0;
    const Prim K = left ? L : R;
    const double SK = left ? SL : SR, VnK = left ? VnL : VnR, mK = left ? mL : mR;
    const double VtK = K.v * nx - K.u * ny;
    const double eK = specific_energy(K, g), EK = K.r * eK;
    const double factor = fdiv(mK, SK - Ss);
    const double mn = factor * Ss, mt = factor * VtK;
    const double Es = factor * (eK + (Ss - VnK) * (Ss + fdiv(K.p, mK)));
    const State F = phys_flux_w(K, EK, VnK, nx, ny);
    return State{F.r + SK * (factor - K.r), F.ru + SK * ((mn * nx - mt * ny) - K.r * K.u),
                 F.rv + SK * ((mn * ny + mt * nx) - K.r * K.v), F.E + SK * (Es - EK)};
}

template <int SCHEME>
FVK_HD State num_flux_w(const Prim& L, const Prim& R, double nx, double ny, double g, double tol, int maxit) {
    if constexpr (SCHEME == FLUX_HLLC) return hllc_w(L, R, nx, ny, g);
    else if constexpr (SCHEME == FLUX_EXACT)
        return exact_flux(from_primitive(L.r, L.u, L.v, L.p, g), from_primitive(R.r, R.u, R.v, R.p, g), nx, ny, g,
                          tol, maxit);
    else return rusanov_w(L, R, nx, ny, g);
}

// euler_ghost_state() of a primitive face state: farfield -> the prescribed
// state, outflow -> copy, wall -> the normal velocity reflected
FVK_HD Prim euler_ghost_w(int type, const double* ff, const Prim& L, double nx, double ny, double g) {
    if (type == BC_FARFIELD) return to_prim(State{ff[0], ff[1], ff[2], ff[3]}, g);
    if (type == BC_OUTFLOW) return L;
    const double Vn = L.u * nx + L.v * ny;
    return Prim{L.r, L.u - 2.0 * Vn * nx, L.v - 2.0 * Vn * ny, L.p};
}

// MUSCL face state from the cell's primitives w and the increments d_q =
// G_q . r: w + d, or w itself where the reconstructed density or pressure
// is not positive
FVK_HD Prim reconstruct_w(const Prim& w, double d0, double d1, double d2, double d3) {
    const double rho = w.r + d0, p = w.p + d3;
    if (!(rho > 0.0) || !(p > 0.0)) return w;
    return Prim{rho, w.u + d1, w.v + d2, p};
}

// reconstruct_w() of a cell with conserved state Uc and limited gradients
// g0 = (rho_x, rho_y, u_x, u_y), g1 = (v_x, v_y, p_x, p_y) at offset r
FVK_HD Prim face_state_w(const State& Uc, const double4& g0, const double4& g1, double rx, double ry, double gam) {
    const double d0 = g0.x * rx + g0.y * ry;
    const double d1 = g0.z * rx + g0.w * ry;
    const double d2 = g1.x * rx + g1.y * ry;
    const double d3 = g1.z * rx + g1.w * ry;
    return reconstruct_w(to_prim(Uc, gam), d0, d1, d2, d3);
}

// Venkatakrishnan limiter function (extension, no reference code):
// ((d1^2 + eps2) + 2 d1 d2) / (d1^2 + 2 d2^2 + d1 d2 + eps2)
FVK_HD double venkatakrishnan(double d1, double d2, double eps2) {
    double d11 = d1 * d1;
    return fdiv((d11 + eps2) + 2.0 * d1 * d2, d11 + 2.0 * d2 * d2 + d1 * d2 + eps2);
}

// phi = min(1, min over the cell's faces with d2 != 0 of venkatakrishnan(d1,
// d2, eps2)), d1 the allowed increment on d2's side (wmax - wc for d2 > 0,
// wmin - wc for d2 < 0).  Then d1 d2 >= 0, so the numerator is >= 0 and the
// denominator > 0, and the smallest ratio is found by cross-multiplication:
// one division per variable instead of one per face.  Start from {1, 1};
// venkat_pick returns true when the face becomes the minimiser (a strict
// decrease, as smin keeps the earlier face on a tie).
struct VenkatMin {
    double num, den;
};
FVK_HD bool venkat_pick(VenkatMin& b, double wlo, double whi, double wc, double d2, double eps2) {
    const double d1 = (d2 > 0.0 ? whi : wlo) - wc;
    const double d11 = d1 * d1;
    const double num = (d11 + eps2) + 2.0 * d1 * d2;
    const double den = d11 + 2.0 * d2 * d2 + d1 * d2 + eps2;
    const bool take = d2 != 0.0 && num * b.den < b.num * den;
    if (take) b = VenkatMin{num, den};
    return take;
}
FVK_HD double venkat_value(const VenkatMin& b) { return fdiv(b.num, b.den); }

// The same pick with the allowed increments kept as extremes of the rounded
// differences dphi = wn - wc: dlo = min(0, min dphi), dhi = max(0, max dphi).
// Rounding is monotone, so the extreme of the rounded differences is the
// rounded difference of the extreme, wmin - wc or wmax - wc, up to the sign
// of a zero, which reaches neither numerator nor denominator.  A face with
// d2 == 0 has num == den and never undercuts the minimum (which starts at
// {1, 1} and keeps num <= den), so it needs no test.  On the device the
// extremes of arithmetic results are single fmin/fmax instructions.
FVK_HD void venkat_extend(double& dlo, double& dhi, double dphi) {
    dlo = fmin(dlo, dphi);
    dhi = fmax(dhi, dphi);
}
FVK_HD void venkat_pick_d(VenkatMin& b, double dlo, double dhi, double d2, double eps2) {
    const double d1 = d2 > 0.0 ? dhi : dlo;
    const double d11 = d1 * d1;
    const double num = (d11 + eps2) + 2.0 * d1 * d2;
    const double den = d11 + 2.0 * d2 * d2 + d1 * d2 + eps2;
    if (num * b.den < b.num * den) b = VenkatMin{num, den};
}

}  // namespace fvk
