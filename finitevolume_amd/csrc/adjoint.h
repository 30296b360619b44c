// Reverse-mode (adjoint) counterparts of the Euler residual's building
// blocks in physics.h, for the discrete adjoint of R(U, X).  There is no
// reference code for the adjoint: each function below is the exact
// derivative of the forward function of the same name, by hand, in reverse
// order of its operations.  Branches (min/max, supersonic early outs,
// positivity fallbacks) take the derivative of the branch the forward
// evaluation took.
//
// Convention: the *_rev functions ADD into their adjoint outputs.
//
// The functions are __host__ __device__ so the same code can be checked
// against finite differences on the CPU.
#pragma once

#include "physics.h"

namespace fvk {

// primitive variables of a state: rho, u, v, p (u = rho_u / rho, p from pressure())
struct Prim {
    double r, u, v, p;
};

FVK_HD Prim to_prim(const State& s, double g) { return Prim{s.r, fdiv(s.ru, s.r), fdiv(s.rv, s.r), pressure(s, g)}; }

// adjoint of to_prim: sb += (dW/dU)^T wb
// (reverse-only arithmetic: one reciprocal of rho instead of five divisions)
FVK_HD void prim_rev(const State& s, const Prim& wb, State& sb, double g) {
    const double ir = fdiv(1.0, s.r);
    const double u = s.ru * ir, v = s.rv * ir;
    const double gm = g - 1.0;
    // p = (g - 1) (E - 0.5 r (u^2 + v^2))
    sb.E += gm * wb.p;
    double rb = wb.r - gm * 0.5 * (u * u + v * v) * wb.p;
    const double ub = wb.u - gm * s.r * u * wb.p;
    const double vb = wb.v - gm * s.r * v * wb.p;
    // u = ru / r, v = rv / r
    sb.ru += ub * ir;
    sb.rv += vb * ir;
    rb -= (ub * u + vb * v) * ir;
    sb.r += rb;
}

// adjoint of from_primitive(w): wb += (dU/dW)^T ob
FVK_HD void from_prim_rev(const Prim& w, const State& ob, Prim& wb, double g) {
    wb.r += ob.r + ob.ru * w.u + ob.rv * w.v + ob.E * 0.5 * (w.u * w.u + w.v * w.v);
    wb.u += ob.ru * w.r + ob.E * w.r * w.u;
    wb.v += ob.rv * w.r + ob.E * w.r * w.v;
    wb.p += ob.E * (1.0 / (g - 1.0));
}

// adjoint of phys_flux(s, nx, ny)
FVK_HD void phys_flux_rev(const State& s, double nx, double ny, double g, const State& Fb, State& sb, double& nxb,
                          double& nyb) {
    const double u = fdiv(s.ru, s.r), v = fdiv(s.rv, s.r);
    const double Vn = u * nx + v * ny;
    const double p = pressure(s, g);
    const double Vnb = Fb.r * s.r + Fb.ru * s.ru + Fb.rv * s.rv + Fb.E * (s.E + p);
    sb.r += Fb.r * Vn;
    sb.ru += Fb.ru * Vn;
    sb.rv += Fb.rv * Vn;
    sb.E += Fb.E * Vn;
    const double pb = Fb.ru * nx + Fb.rv * ny + Fb.E * Vn;
    nxb += Fb.ru * p + Vnb * u;
    nyb += Fb.rv * p + Vnb * v;
    prim_rev(s, Prim{0.0, Vnb * nx, Vnb * ny, pb}, sb, g);
}

// adjoint of c = sound_speed(s) = sqrt(g p / rho), as primitive adjoints
FVK_HD void sound_speed_rev(double r, double c, double cb, Prim& wb, double g) {
    const double t = cb / (2.0 * c * r);  // dc/dp = g / (2 c r), dc/drho = -c / (2 r) = -(c^2 / g) dc/dp
    wb.p += t * g;
    wb.r -= t * c * c;
}

// adjoint of rusanov(L, R, nx, ny)
FVK_HD void rusanov_rev(const State& L, const State& R, double nx, double ny, double g, const State& Ob, State& Lb,
                        State& Rb, double& nxb, double& nyb) {
    const double VnL = (L.ru * nx + L.rv * ny) / L.r;
    const double VnR = (R.ru * nx + R.rv * ny) / R.r;
    const double cL = sound_speed(L, g), cR = sound_speed(R, g);
    const double a = fabs(VnL) + cL, b = fabs(VnR) + cR;
    const bool takeR = a < b;  // smax(a, b)
    const double Smax = takeR ? b : a;
    const double h = 0.5 * Smax;
    const State Fh{0.5 * Ob.r, 0.5 * Ob.ru, 0.5 * Ob.rv, 0.5 * Ob.E};
    phys_flux_rev(L, nx, ny, g, Fh, Lb, nxb, nyb);
    phys_flux_rev(R, nx, ny, g, Fh, Rb, nxb, nyb);
    const double hb = -(Ob.r * (R.r - L.r) + Ob.ru * (R.ru - L.ru) + Ob.rv * (R.rv - L.rv) + Ob.E * (R.E - L.E));
    const double Sb = 0.5 * hb;
    Rb.r -= h * Ob.r; Rb.ru -= h * Ob.ru; Rb.rv -= h * Ob.rv; Rb.E -= h * Ob.E;
    Lb.r += h * Ob.r; Lb.ru += h * Ob.ru; Lb.rv += h * Ob.rv; Lb.E += h * Ob.E;
    const State& K = takeR ? R : L;
    State& Kb = takeR ? Rb : Lb;
    const double VnK = takeR ? VnR : VnL, cK = takeR ? cR : cL;
    const double Vnb = VnK >= 0.0 ? Sb : -Sb;
    const double iK = fdiv(1.0, K.r);
    Kb.ru += Vnb * nx * iK;
    Kb.rv += Vnb * ny * iK;
    Kb.r -= Vnb * VnK * iK;
    nxb += Vnb * K.ru * iK;
    nyb += Vnb * K.rv * iK;
    Prim wb{0.0, 0.0, 0.0, 0.0};
    sound_speed_rev(K.r, cK, Sb, wb, g);
    prim_rev(K, wb, Kb, g);
}

// adjoint of hllc(L, R, nx, ny)
FVK_HD void hllc_rev(const State& L, const State& R, double nx, double ny, double g, const State& Ob, State& Lb,
                     State& Rb, double& nxb, double& nyb) {
    const double tx = -ny, ty = nx;
    const double rL = L.r, rR = R.r;
    const double uL = fdiv(L.ru, rL), vL = fdiv(L.rv, rL), uR = fdiv(R.ru, rR), vR = fdiv(R.rv, rR);
    const double VnL = uL * nx + vL * ny, VnR = uR * nx + vR * ny;
    const double VtL = uL * tx + vL * ty, VtR = uR * tx + vR * ty;
    const double pL = pressure(L, g), pR = pressure(R, g);
    const double cL = sound_speed(L, g), cR = sound_speed(R, g);
    const double a1 = VnL - cL, a2 = VnR - cR;
    const bool slR = a2 < a1;  // smin
    const double SL = slR ? a2 : a1;
    const double b1 = VnL + cL, b2 = VnR + cR;
    const bool srR = b1 < b2;  // smax
    const double SR = srR ? b2 : b1;
    if (SL >= 0.0) { phys_flux_rev(L, nx, ny, g, Ob, Lb, nxb, nyb); return; }
    if (SR <= 0.0) { phys_flux_rev(R, nx, ny, g, Ob, Rb, nxb, nyb); return; }
    const double num = pR - pL + rL * VnL * (SL - VnL) - rR * VnR * (SR - VnR);
    const double den = rL * (SL - VnL) - rR * (SR - VnR);
    const double Ss = fdiv(num, den);
    const bool left = Ss >= 0.0;
    const double SK = left ? SL : SR, rK = left ? rL : rR, VnK = left ? VnL : VnR, VtK = left ? VtL : VtR;
    const double pK = left ? pL : pR;
    const State& UK = left ? L : R;
    const double D = rK * (SK - VnK);
    const double E2 = SK - Ss;
    const double factor = fdiv(D, E2);
    const double mn = factor * Ss, mt = factor * VtK;
    const double T = Ss + fdiv(pK, D);
    const double Q = fdiv(UK.E, rK) + (Ss - VnK) * T;
    const double Es = factor * Q;

    // out = F(UK) + SK (U* - UK)
    State UKb{0.0, 0.0, 0.0, 0.0};
    phys_flux_rev(UK, nx, ny, g, Ob, UKb, nxb, nyb);
    double SKb = Ob.r * (factor - UK.r) + Ob.ru * ((mn * nx - mt * ny) - UK.ru) +
                 Ob.rv * ((mn * ny + mt * nx) - UK.rv) + Ob.E * (Es - UK.E);
    double factorb = Ob.r * SK;
    UKb.r -= Ob.r * SK; UKb.ru -= Ob.ru * SK; UKb.rv -= Ob.rv * SK; UKb.E -= Ob.E * SK;
    const double mnb = SK * (Ob.ru * nx + Ob.rv * ny);
    const double mtb = SK * (-Ob.ru * ny + Ob.rv * nx);
    nxb += SK * (Ob.ru * mn + Ob.rv * mt);
    nyb += SK * (-Ob.ru * mt + Ob.rv * mn);
    const double Esb = SK * Ob.E;
    // Es = factor Q, Q = E_K / r_K + (Ss - Vn_K) T, T = Ss + p_K / D
    factorb += Esb * Q;
    const double Qb = Esb * factor;
    const double irK = fdiv(1.0, rK);
    UKb.E += Qb * irK;
    double rKb = -Qb * UK.E * irK * irK;
    double Ssb = Qb * T;
    double VnKb = -Qb * T;
    const double Tb = Qb * (Ss - VnK);
    Ssb += Tb;
    const double iD = fdiv(1.0, D);
    const double pKb = Tb * iD;
    double Db = -Tb * pK * iD * iD;
    // mn = factor Ss, mt = factor Vt_K
    factorb += mnb * Ss;
    Ssb += mnb * factor;
    factorb += mtb * VtK;
    const double VtKb = mtb * factor;
    // factor = D / (SK - Ss)
    const double iE2 = fdiv(1.0, E2);
    Db += factorb * iE2;
    const double E2b = -factorb * factor * iE2;
    SKb += E2b;
    Ssb -= E2b;
    // D = r_K (SK - Vn_K)
    rKb += Db * (SK - VnK);
    SKb += Db * rK;
    VnKb -= Db * rK;
    // Ss = num / den
    const double iden = fdiv(1.0, den);
    const double numb = Ssb * iden;
    const double denb = -Ssb * Ss * iden;
    double pLb = -numb, pRb = numb;
    double rLb = numb * VnL * (SL - VnL) + denb * (SL - VnL);
    double rRb = -numb * VnR * (SR - VnR) - denb * (SR - VnR);
    double VnLb = numb * rL * (SL - 2.0 * VnL) - denb * rL;
    double VnRb = -numb * rR * (SR - 2.0 * VnR) + denb * rR;
    double SLb = numb * rL * VnL + denb * rL;
    double SRb = -numb * rR * VnR - denb * rR;
    double VtLb = 0.0, VtRb = 0.0;
    if (left) { SLb += SKb; rLb += rKb; VnLb += VnKb; VtLb += VtKb; pLb += pKb; }
    else { SRb += SKb; rRb += rKb; VnRb += VnKb; VtRb += VtKb; pRb += pKb; }
    double cLb = 0.0, cRb = 0.0;
    if (slR) { VnRb += SLb; cRb -= SLb; } else { VnLb += SLb; cLb -= SLb; }
    if (srR) { VnRb += SRb; cRb += SRb; } else { VnLb += SRb; cLb += SRb; }
    Prim wLb{rLb, VnLb * nx + VtLb * tx, VnLb * ny + VtLb * ty, pLb};
    Prim wRb{rRb, VnRb * nx + VtRb * tx, VnRb * ny + VtRb * ty, pRb};
    sound_speed_rev(rL, cL, cLb, wLb, g);
    sound_speed_rev(rR, cR, cRb, wRb, g);
    nxb += VnLb * uL + VnRb * uR;
    nyb += VnLb * vL + VnRb * vR;
    const double txb = VtLb * uL + VtRb * uR, tyb = VtLb * vL + VtRb * vR;
    nyb -= txb;
    nxb += tyb;
    prim_rev(L, wLb, Lb, g);
    prim_rev(R, wRb, Rb, g);
    State& Kb = left ? Lb : Rb;
    Kb.r += UKb.r; Kb.ru += UKb.ru; Kb.rv += UKb.rv; Kb.E += UKb.E;
}

template <int SCHEME>
FVK_HD void num_flux_rev(const State& L, const State& R, double nx, double ny, double g, const State& Ob, State& Lb,
                         State& Rb, double& nxb, double& nyb) {
    static_assert(SCHEME != FLUX_EXACT, "no adjoint of the exact Riemann flux");
    if constexpr (SCHEME == FLUX_HLLC) hllc_rev(L, R, nx, ny, g, Ob, Lb, Rb, nxb, nyb);
    else rusanov_rev(L, R, nx, ny, g, Ob, Lb, Rb, nxb, nyb);
}

// adjoint of ghost_slip_wall(L, n)
FVK_HD void ghost_slip_wall_rev(const State& L, double nx, double ny, double g, const State& Gb, State& Lb,
                                double& nxb, double& nyb) {
    const double u = fdiv(L.ru, L.r), v = fdiv(L.rv, L.r);
    const double Vn = u * nx + v * ny;
    const Prim wg{L.r, u - 2.0 * Vn * nx, v - 2.0 * Vn * ny, pressure(L, g)};
    Prim wgb{0.0, 0.0, 0.0, 0.0};
    from_prim_rev(wg, Gb, wgb, g);
    const double Vnb = -2.0 * (wgb.u * nx + wgb.v * ny);
    nxb += -2.0 * Vn * wgb.u + Vnb * u;
    nyb += -2.0 * Vn * wgb.v + Vnb * v;
    prim_rev(L, Prim{wgb.r, wgb.u + Vnb * nx, wgb.v + Vnb * ny, wgb.p}, Lb, g);
}

// adjoint of the Euler ghost state (farfield: constant; outflow: copy; wall: slip mirror)
FVK_HD void euler_ghost_rev(int type, const State& L, double nx, double ny, double g, const State& Gb, State& Lb,
                            double& nxb, double& nyb) {
    if (type == BC_FARFIELD) return;
    if (type == BC_OUTFLOW) {
        Lb.r += Gb.r; Lb.ru += Gb.ru; Lb.rv += Gb.rv; Lb.E += Gb.E;
        return;
    }
    ghost_slip_wall_rev(L, nx, ny, g, Gb, Lb, nxb, nyb);
}

// adjoint of the MUSCL face state: out = from_primitive(W(U) + d) with the
// increments d_q = G_q . r, or U itself when the reconstructed density or
// pressure is not positive.  Returns db (the adjoint of the four increments)
// and adds dU.  Where the increments are all zero the forward returns U
// exactly, which is the same function, so the general derivative applies.
FVK_HD void reconstruct_rev(const State& Uc, const double4& g0, const double4& g1, double rx, double ry, double gam,
                            const State& Ob, State& Ucb, double (&db)[4]) {
    const double d0 = g0.x * rx + g0.y * ry;
    const double d1 = g0.z * rx + g0.w * ry;
    const double d2 = g1.x * rx + g1.y * ry;
    const double d3 = g1.z * rx + g1.w * ry;
    const double u = fdiv(Uc.ru, Uc.r), v = fdiv(Uc.rv, Uc.r), p = pressure(Uc, gam);
    const double rho = Uc.r + d0, pp = p + d3;
    if (!(rho > 0.0) || !(pp > 0.0)) {
        Ucb.r += Ob.r; Ucb.ru += Ob.ru; Ucb.rv += Ob.rv; Ucb.E += Ob.E;
        db[0] = db[1] = db[2] = db[3] = 0.0;
        return;
    }
    Prim wb{0.0, 0.0, 0.0, 0.0};
    from_prim_rev(Prim{rho, u + d1, v + d2, pp}, Ob, wb, gam);
    db[0] = wb.r; db[1] = wb.u; db[2] = wb.v; db[3] = wb.p;
    prim_rev(Uc, wb, Ucb, gam);
}

// partial derivatives of venkatakrishnan(d1, d2, eps2)
FVK_HD void venkatakrishnan_grad(double d1, double d2, double eps2, double& p1, double& p2, double& pe) {
    const double N = d1 * d1 + eps2 + 2.0 * d1 * d2;
    const double D = d1 * d1 + 2.0 * d2 * d2 + d1 * d2 + eps2;
    const double iD2 = fdiv(1.0, D * D);
    p1 = ((2.0 * d1 + 2.0 * d2) * D - N * (2.0 * d1 + d2)) * iD2;
    p2 = (2.0 * d1 * D - N * (4.0 * d2 + d1)) * iD2;
    pe = (D - N) * iD2;
}

// adjoint of the face length and unit normal of an edge e = node2 - node1:
// A = |e|, n = (e_y, -e_x) / A.  Adds the adjoint of e.
FVK_HD void edge_rev(double nx, double ny, double A, double nxb, double nyb, double Ab, double& exb, double& eyb) {
    // e = A (-ny, nx)
    const double iA = fdiv(1.0, A);
    const double At = Ab - (nxb * nx + nyb * ny) * iA;
    exb += -nyb * iA - At * ny;
    eyb += nxb * iA + At * nx;
}

}  // namespace fvk
