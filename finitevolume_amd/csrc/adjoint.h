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

// ---- second order: adjoints of the fluxes on primitive face states ----

FVK_HD void add_to(Prim& a, const Prim& b) {
    a.r += b.r; a.u += b.u; a.v += b.v; a.p += b.p;
}

// adjoint of e = specific_energy(w) = p / ((g - 1) rho) + (u^2 + v^2) / 2
FVK_HD void specific_energy_rev(const Prim& w, double eb, Prim& wb, double g) {
    const double t = fdiv(eb, (g - 1.0) * w.r);  // de/dp
    wb.p += t;
    wb.r -= t * fdiv(w.p, w.r);
    wb.u += eb * w.u;
    wb.v += eb * w.v;
}

// adjoint of phys_flux_w(w, E, Vn, nx, ny): adds wb, Eb, Vnb, nxb, nyb
FVK_HD void phys_flux_w_rev(const Prim& w, double E, double Vn, double nx, double ny, const State& Fb, Prim& wb,
                            double& Eb, double& Vnb, double& nxb, double& nyb) {
    const double m = w.r * Vn;
    const double mb = Fb.r + Fb.ru * w.u + Fb.rv * w.v;
    wb.u += Fb.ru * m;
    wb.v += Fb.rv * m;
    wb.p += Fb.ru * nx + Fb.rv * ny + Fb.E * Vn;
    nxb += Fb.ru * w.p;
    nyb += Fb.rv * w.p;
    Eb += Fb.E * Vn;
    Vnb += Fb.E * (E + w.p) + mb * w.r;
    wb.r += mb * Vn;
}

// adjoint of Vn = u nx + v ny
FVK_HD void normal_velocity_rev(const Prim& w, double nx, double ny, double Vnb, Prim& wb, double& nxb, double& nyb) {
    wb.u += Vnb * nx;
    wb.v += Vnb * ny;
    nxb += Vnb * w.u;
    nyb += Vnb * w.v;
}

// adjoint of the flux of one primitive state (an HLLC supersonic early out)
FVK_HD void phys_flux_state_w_rev(const Prim& w, double nx, double ny, double g, const State& Ob, Prim& wb, double& nxb,
                                  double& nyb) {
    const double e = specific_energy(w, g), Vn = w.u * nx + w.v * ny;
    double Eb = 0.0, Vnb = 0.0;
    phys_flux_w_rev(w, w.r * e, Vn, nx, ny, Ob, wb, Eb, Vnb, nxb, nyb);
    wb.r += Eb * e;
    specific_energy_rev(w, Eb * w.r, wb, g);
    normal_velocity_rev(w, nx, ny, Vnb, wb, nxb, nyb);
}

// adjoint of rusanov_w(L, R, nx, ny)
FVK_HD void rusanov_w_rev(const Prim& L, const Prim& R, double nx, double ny, double g, const State& Ob, Prim& Lb,
                          Prim& Rb, double& nxb, double& nyb) {
    const double VnL = L.u * nx + L.v * ny, VnR = R.u * nx + R.v * ny;
    const double cL = sqrt(fdiv(g * L.p, L.r)), cR = sqrt(fdiv(g * R.p, R.r));
    const double a = fabs(VnL) + cL, b = fabs(VnR) + cR;
    const bool takeR = a < b;  // smax(a, b)
    const double h = 0.5 * (takeR ? b : a);
    const double eL = specific_energy(L, g), eR = specific_energy(R, g);
    const double EL = L.r * eL, ER = R.r * eR;
    // out = (F_L + F_R) / 2 - h (U_R - U_L)
    const State Fh{0.5 * Ob.r, 0.5 * Ob.ru, 0.5 * Ob.rv, 0.5 * Ob.E};
    double ELb = 0.0, ERb = 0.0, VnLb = 0.0, VnRb = 0.0;
    phys_flux_w_rev(L, EL, VnL, nx, ny, Fh, Lb, ELb, VnLb, nxb, nyb);
    phys_flux_w_rev(R, ER, VnR, nx, ny, Fh, Rb, ERb, VnRb, nxb, nyb);
    const double hb = -(Ob.r * (R.r - L.r) + Ob.ru * (R.r * R.u - L.r * L.u) + Ob.rv * (R.r * R.v - L.r * L.v) +
                        Ob.E * (ER - EL));
    // U = (rho, rho u, rho v, E) of each side, weighted by -h (right) and +h (left)
    Rb.r -= h * (Ob.r + Ob.ru * R.u + Ob.rv * R.v);
    Rb.u -= h * Ob.ru * R.r;
    Rb.v -= h * Ob.rv * R.r;
    ERb -= h * Ob.E;
    Lb.r += h * (Ob.r + Ob.ru * L.u + Ob.rv * L.v);
    Lb.u += h * Ob.ru * L.r;
    Lb.v += h * Ob.rv * L.r;
    ELb += h * Ob.E;
    // E = rho e
    Lb.r += ELb * eL;
    specific_energy_rev(L, ELb * L.r, Lb, g);
    Rb.r += ERb * eR;
    specific_energy_rev(R, ERb * R.r, Rb, g);
    // Smax = |Vn_K| + c_K of the larger side
    const double Sb = 0.5 * hb;
    if (takeR) {
        VnRb += VnR >= 0.0 ? Sb : -Sb;
        sound_speed_rev(R.r, cR, Sb, Rb, g);
    } else {
        VnLb += VnL >= 0.0 ? Sb : -Sb;
        sound_speed_rev(L.r, cL, Sb, Lb, g);
    }
    normal_velocity_rev(L, nx, ny, VnLb, Lb, nxb, nyb);
    normal_velocity_rev(R, nx, ny, VnRb, Rb, nxb, nyb);
}

// adjoint of hllc_w(L, R, nx, ny)
FVK_HD void hllc_w_rev(const Prim& L, const Prim& R, double nx, double ny, double g, const State& Ob, Prim& Lb,
                       Prim& Rb, double& nxb, double& nyb) {
    const double VnL = L.u * nx + L.v * ny, VnR = R.u * nx + R.v * ny;
    const double cL = sqrt(fdiv(g * L.p, L.r)), cR = sqrt(fdiv(g * R.p, R.r));
    const double a1 = VnL - cL, a2 = VnR - cR;
    const bool slR = a2 < a1;  // smin
    const double SL = slR ? a2 : a1;
    const double b1 = VnL + cL, b2 = VnR + cR;
    const bool srR = b1 < b2;  // smax
    const double SR = srR ? b2 : b1;
    if (SL >= 0.0) { phys_flux_state_w_rev(L, nx, ny, g, Ob, Lb, nxb, nyb); return; }
    if (SR <= 0.0) { phys_flux_state_w_rev(R, nx, ny, g, Ob, Rb, nxb, nyb); return; }
    const double mL = L.r * (SL - VnL), mR = R.r * (SR - VnR);
    const double den = mL - mR;
    const double Ss = fdiv(R.p - L.p + mL * VnL - mR * VnR, den);
    const bool left = Ss >= 0.0;
    const Prim& K = left ? L : R;
    const double SK = left ? SL : SR, VnK = left ? VnL : VnR, mK = left ? mL : mR;
    const double VtK = K.v * nx - K.u * ny;
    const double eK = specific_energy(K, g), EK = K.r * eK;
    const double E2 = SK - Ss;
    const double factor = fdiv(mK, E2);
    const double mn = factor * Ss, mt = factor * VtK;
    const double imK = fdiv(1.0, mK);
    const double T = Ss + K.p * imK;
    const double Q = eK + (Ss - VnK) * T;
    const double Es = factor * Q;

    // out = F(K) + SK (U* - U_K)
    Prim Kb{0.0, 0.0, 0.0, 0.0};
    double EKb = 0.0, VnKb = 0.0;
    phys_flux_w_rev(K, EK, VnK, nx, ny, Ob, Kb, EKb, VnKb, nxb, nyb);
    double SKb = Ob.r * (factor - K.r) + Ob.ru * ((mn * nx - mt * ny) - K.r * K.u) +
                 Ob.rv * ((mn * ny + mt * nx) - K.r * K.v) + Ob.E * (Es - EK);
    double factorb = Ob.r * SK;
    Kb.r -= SK * (Ob.r + Ob.ru * K.u + Ob.rv * K.v);
    Kb.u -= SK * Ob.ru * K.r;
    Kb.v -= SK * Ob.rv * K.r;
    EKb -= SK * Ob.E;
    const double mnb = SK * (Ob.ru * nx + Ob.rv * ny);
    const double mtb = SK * (-Ob.ru * ny + Ob.rv * nx);
    nxb += SK * (Ob.ru * mn + Ob.rv * mt);
    nyb += SK * (-Ob.ru * mt + Ob.rv * mn);
    const double Esb = SK * Ob.E;
    // Es = factor Q, Q = e_K + (Ss - Vn_K) T, T = Ss + p_K / m_K
    factorb += Esb * Q;
    const double Qb = Esb * factor;
    double eKb = Qb;
    double Ssb = Qb * T;
    VnKb -= Qb * T;
    const double Tb = Qb * (Ss - VnK);
    Ssb += Tb;
    Kb.p += Tb * imK;
    double mKb = -Tb * K.p * imK * imK;
    // mn = factor Ss, mt = factor Vt_K
    factorb += mnb * Ss + mtb * VtK;
    Ssb += mnb * factor;
    const double VtKb = mtb * factor;
    // factor = m_K / (SK - Ss)
    const double iE2 = fdiv(1.0, E2);
    mKb += factorb * iE2;
    const double E2b = -factorb * factor * iE2;
    SKb += E2b;
    Ssb -= E2b;
    // E_K = rho_K e_K, e_K, Vt_K = v_K nx - u_K ny
    Kb.r += EKb * eK;
    eKb += EKb * K.r;
    specific_energy_rev(K, eKb, Kb, g);
    Kb.v += VtKb * nx;
    Kb.u -= VtKb * ny;
    nxb += VtKb * K.v;
    nyb -= VtKb * K.u;
    // Ss = num / den, num = p_R - p_L + m_L Vn_L - m_R Vn_R, den = m_L - m_R
    const double iden = fdiv(1.0, den);
    const double numb = Ssb * iden;
    const double denb = -Ssb * Ss * iden;
    Prim wLb{0.0, 0.0, 0.0, -numb}, wRb{0.0, 0.0, 0.0, numb};
    double mLb = numb * VnL + denb, mRb = -numb * VnR - denb;
    double VnLb = numb * mL, VnRb = -numb * mR;
    double SLb = 0.0, SRb = 0.0;
    if (left) { SLb += SKb; VnLb += VnKb; mLb += mKb; add_to(wLb, Kb); }
    else { SRb += SKb; VnRb += VnKb; mRb += mKb; add_to(wRb, Kb); }
    // m_L = rho_L (S_L - Vn_L), m_R = rho_R (S_R - Vn_R)
    wLb.r += mLb * (SL - VnL);
    SLb += mLb * L.r;
    VnLb -= mLb * L.r;
    wRb.r += mRb * (SR - VnR);
    SRb += mRb * R.r;
    VnRb -= mRb * R.r;
    // Davis speeds
    double cLb = 0.0, cRb = 0.0;
    if (slR) { VnRb += SLb; cRb -= SLb; } else { VnLb += SLb; cLb -= SLb; }
    if (srR) { VnRb += SRb; cRb += SRb; } else { VnLb += SRb; cLb += SRb; }
    sound_speed_rev(L.r, cL, cLb, wLb, g);
    sound_speed_rev(R.r, cR, cRb, wRb, g);
    normal_velocity_rev(L, nx, ny, VnLb, wLb, nxb, nyb);
    normal_velocity_rev(R, nx, ny, VnRb, wRb, nxb, nyb);
    add_to(Lb, wLb);
    add_to(Rb, wRb);
}

template <int SCHEME>
FVK_HD void num_flux_w_rev(const Prim& L, const Prim& R, double nx, double ny, double g, const State& Ob, Prim& Lb,
                           Prim& Rb, double& nxb, double& nyb) {
    static_assert(SCHEME != FLUX_EXACT, "no adjoint of the exact Riemann flux");
    if constexpr (SCHEME == FLUX_HLLC) hllc_w_rev(L, R, nx, ny, g, Ob, Lb, Rb, nxb, nyb);
    else rusanov_w_rev(L, R, nx, ny, g, Ob, Lb, Rb, nxb, nyb);
}

// adjoint of euler_ghost_w (farfield: constant; outflow: copy; wall: the
// normal velocity reflected)
FVK_HD void euler_ghost_w_rev(int type, const Prim& L, double nx, double ny, const Prim& Gb, Prim& Lb, double& nxb,
                              double& nyb) {
    if (type == BC_FARFIELD) return;
    if (type == BC_OUTFLOW) {
        add_to(Lb, Gb);
        return;
    }
    const double Vn = L.u * nx + L.v * ny;
    const double Vnb = -2.0 * (Gb.u * nx + Gb.v * ny);
    nxb += -2.0 * Vn * Gb.u;
    nyb += -2.0 * Vn * Gb.v;
    Lb.r += Gb.r;
    Lb.u += Gb.u;
    Lb.v += Gb.v;
    Lb.p += Gb.p;
    normal_velocity_rev(L, nx, ny, Vnb, Lb, nxb, nyb);
}

// adjoint of face_state_w: W_f = W(U) + d with the increments d_q = G_q . r,
// or W(U) itself where the reconstructed density or pressure is not
// positive.  Returns db (the adjoint of the four increments) and adds dU.
FVK_HD void face_state_w_rev(const State& Uc, const double4& g0, const double4& g1, double rx, double ry, double gam,
                             const Prim& Wfb, State& Ucb, double (&db)[4]) {
    const double d0 = g0.x * rx + g0.y * ry;
    const double d3 = g1.z * rx + g1.w * ry;
    const Prim w = to_prim(Uc, gam);
    const bool rec = (w.r + d0 > 0.0) && (w.p + d3 > 0.0);
    db[0] = rec ? Wfb.r : 0.0;
    db[1] = rec ? Wfb.u : 0.0;
    db[2] = rec ? Wfb.v : 0.0;
    db[3] = rec ? Wfb.p : 0.0;
    prim_rev(Uc, Wfb, Ucb, gam);
}

// whether face_state_w reconstructs (true) or returns W(U) itself (false)
FVK_HD bool face_state_w_reconstructs(const State& Uc, const double4& g0, const double4& g1, double rx, double ry,
                                      double gam) {
    const double d0 = g0.x * rx + g0.y * ry;
    const double d3 = g1.z * rx + g1.w * ry;
    const Prim w = to_prim(Uc, gam);
    return (w.r + d0 > 0.0) && (w.p + d3 > 0.0);
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
