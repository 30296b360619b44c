// Finite-difference check of the hand-written adjoints in adjoint.h on the
// CPU: for random inputs x, directions v and output weights w,
//   w . (f(x + h v) - f(x - h v)) / 2h   ==   (f_rev(w)) . v
// Prints one line per function: name, number of trials, max relative error.
#include <cstdio>
#include <cmath>
#include <random>

#include "adjoint.h"

using namespace fvk;

static std::mt19937_64 rng(12345);
static double uni(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

static State rand_state(double g, double mach_max) {
    double r = uni(0.3, 2.0), p = uni(0.3, 2.0);
    double c = std::sqrt(g * p / r);
    double ang = uni(0.0, 2.0 * M_PI), m = uni(0.0, mach_max);
    return from_primitive(r, m * c * std::cos(ang), m * c * std::sin(ang), p, g);
}
static State add(const State& a, const State& d, double h) {
    return State{a.r + h * d.r, a.ru + h * d.ru, a.rv + h * d.rv, a.E + h * d.E};
}
static double dot(const State& a, const State& b) { return a.r * b.r + a.ru * b.ru + a.rv * b.rv + a.E * b.E; }
static State rand_dir() { return State{uni(-1, 1), uni(-1, 1), uni(-1, 1), uni(-1, 1)}; }
static Prim prand() { return Prim{uni(-1, 1), uni(-1, 1), uni(-1, 1), uni(-1, 1)}; }
static Prim padd(const Prim& a, const Prim& d, double h) {
    return Prim{a.r + h * d.r, a.u + h * d.u, a.v + h * d.v, a.p + h * d.p};
}
static double pdot(const Prim& a, const Prim& b) { return a.r * b.r + a.u * b.u + a.v * b.v + a.p * b.p; }
static double pdot(const Prim& a, const State& b) { return a.r * b.r + a.u * b.ru + a.v * b.rv + a.p * b.E; }

struct Err {
    double max = 0.0;
    int n = 0;
    void add(double fd, double ad) {
        double e = std::fabs(fd - ad) / std::fmax(1e-8, std::fmax(std::fabs(fd), std::fabs(ad)));
        if (e > max) max = e;
        ++n;
    }
    void add_abs(double e) {
        if (e > max) max = e;
        ++n;
    }
    void print(const char* name) const { std::printf("%s %d %.3e\n", name, n, max); }
};

int main() {
    const double g = 1.4, h = 1e-6;
    const int trials = 2000;
    // ---- fluxes, both schemes, all branches (Mach up to 2.5 reaches the supersonic early outs) ----
    Err e_hllc, e_rus, e_wall, e_rec, e_prim, e_venk, e_edge, e_hllc_w, e_rus_w, e_wall_w, e_same;
    int hits[4] = {0, 0, 0, 0};
    for (int t = 0; t < trials; ++t) {
        State L = rand_state(g, 2.5), R = rand_state(g, 2.5);
        double a = uni(0, 2 * M_PI), nx = std::cos(a), ny = std::sin(a);
        State vL = rand_dir(), vR = rand_dir();
        double vnx = uni(-1, 1), vny = uni(-1, 1);
        State w = rand_dir();
        for (int scheme = 0; scheme < 2; ++scheme) {
            auto f = [&](double s) {
                State l = add(L, vL, s), r = add(R, vR, s);
                double x = nx + s * vnx, y = ny + s * vny;
                return dot(w, scheme ? hllc(l, r, x, y, g) : rusanov(l, r, x, y, g));
            };
            double fd = (f(h) - f(-h)) / (2 * h);
            State Lb{0, 0, 0, 0}, Rb{0, 0, 0, 0};
            double nxb = 0, nyb = 0;
            if (scheme) hllc_rev(L, R, nx, ny, g, w, Lb, Rb, nxb, nyb);
            else rusanov_rev(L, R, nx, ny, g, w, Lb, Rb, nxb, nyb);
            double ad = dot(Lb, vL) + dot(Rb, vR) + nxb * vnx + nyb * vny;
            (scheme ? e_hllc : e_rus).add(fd, ad);
        }
        // branch coverage of HLLC
        {
            double uL = L.ru / L.r, vLv = L.rv / L.r, uR = R.ru / R.r, vRv = R.rv / R.r;
            double VnL = uL * nx + vLv * ny, VnR = uR * nx + vRv * ny;
            double cL = sound_speed(L, g), cR = sound_speed(R, g);
            double SL = smin(VnL - cL, VnR - cR), SR = smax(VnL + cL, VnR + cR);
            if (SL >= 0) hits[0]++;
            else if (SR <= 0) hits[1]++;
            else {
                double pL = pressure(L, g), pR = pressure(R, g);
                double Ss = (pR - pL + L.r * VnL * (SL - VnL) - R.r * VnR * (SR - VnR)) /
                            (L.r * (SL - VnL) - R.r * (SR - VnR));
                hits[Ss >= 0 ? 2 : 3]++;
            }
        }
        // slip-wall ghost
        {
            auto f = [&](double s) { return dot(w, ghost_slip_wall(add(L, vL, s), nx + s * vnx, ny + s * vny, g)); };
            double fd = (f(h) - f(-h)) / (2 * h);
            State Lb{0, 0, 0, 0};
            double nxb = 0, nyb = 0;
            ghost_slip_wall_rev(L, nx, ny, g, w, Lb, nxb, nyb);
            e_wall.add(fd, dot(Lb, vL) + nxb * vnx + nyb * vny);
        }
        // MUSCL face state with small increments (the primitive face state's adjoint)
        {
            double4 g0{uni(-.2, .2), uni(-.2, .2), uni(-.2, .2), uni(-.2, .2)};
            double4 g1{uni(-.2, .2), uni(-.2, .2), uni(-.2, .2), uni(-.2, .2)};
            double4 v0{uni(-1, 1), uni(-1, 1), uni(-1, 1), uni(-1, 1)}, v1{uni(-1, 1), uni(-1, 1), uni(-1, 1), uni(-1, 1)};
            double rx = uni(-.5, .5), ry = uni(-.5, .5), vrx = uni(-1, 1), vry = uni(-1, 1);
            const Prim wfb = prand();
            auto f = [&](double s) {
                double4 a0{g0.x + s * v0.x, g0.y + s * v0.y, g0.z + s * v0.z, g0.w + s * v0.w};
                double4 a1{g1.x + s * v1.x, g1.y + s * v1.y, g1.z + s * v1.z, g1.w + s * v1.w};
                double x = rx + s * vrx, y = ry + s * vry;
                return pdot(wfb, face_state_w(add(L, vL, s), a0, a1, x, y, g));
            };
            double fd = (f(h) - f(-h)) / (2 * h);
            State Ub{0, 0, 0, 0};
            double db[4];
            face_state_w_rev(L, g0, g1, rx, ry, g, wfb, Ub, db);
            // d_q = G_q . r: dG_q = db_q r, dr = sum_q db_q G_q
            double ad = dot(Ub, vL);
            ad += db[0] * (v0.x * rx + v0.y * ry) + db[1] * (v0.z * rx + v0.w * ry) + db[2] * (v1.x * rx + v1.y * ry) +
                  db[3] * (v1.z * rx + v1.w * ry);
            ad += (db[0] * g0.x + db[1] * g0.z + db[2] * g1.x + db[3] * g1.z) * vrx +
                  (db[0] * g0.y + db[1] * g0.w + db[2] * g1.y + db[3] * g1.w) * vry;
            e_rec.add(fd, ad);
        }
        // fluxes on primitive face states, both schemes, and the primitive slip-wall ghost
        {
            const Prim WL = to_prim(L, g), WR = to_prim(R, g);
            const Prim vWL = prand(), vWR = prand();
            for (int scheme = 0; scheme < 2; ++scheme) {
                auto f = [&](double s) {
                    Prim l = padd(WL, vWL, s), r = padd(WR, vWR, s);
                    double x = nx + s * vnx, y = ny + s * vny;
                    return dot(w, scheme ? hllc_w(l, r, x, y, g) : rusanov_w(l, r, x, y, g));
                };
                double fd = (f(h) - f(-h)) / (2 * h);
                Prim Lb{0, 0, 0, 0}, Rb{0, 0, 0, 0};
                double nxb = 0, nyb = 0;
                if (scheme) hllc_w_rev(WL, WR, nx, ny, g, w, Lb, Rb, nxb, nyb);
                else rusanov_w_rev(WL, WR, nx, ny, g, w, Lb, Rb, nxb, nyb);
                (scheme ? e_hllc_w : e_rus_w).add(fd, pdot(Lb, vWL) + pdot(Rb, vWR) + nxb * vnx + nyb * vny);
            }
            const Prim gb = prand();
            auto f = [&](double s) {
                return pdot(gb, euler_ghost_w(BC_WALL, nullptr, padd(WL, vWL, s), nx + s * vnx, ny + s * vny, g));
            };
            double fd = (f(h) - f(-h)) / (2 * h);
            Prim Lb{0, 0, 0, 0};
            double nxb = 0, nyb = 0;
            euler_ghost_w_rev(BC_WALL, WL, nx, ny, gb, Lb, nxb, nyb);
            e_wall_w.add(fd, pdot(Lb, vWL) + nxb * vnx + nyb * vny);
            // the primitive fluxes equal the conserved ones of from_primitive(W)
            const State a = hllc_w(WL, WR, nx, ny, g), b = hllc(L, R, nx, ny, g);
            const State c = rusanov_w(WL, WR, nx, ny, g), d = rusanov(L, R, nx, ny, g);
            const double sc = 1.0 + std::fabs(b.r) + std::fabs(b.ru) + std::fabs(b.rv) + std::fabs(b.E);
            e_same.add_abs((std::fabs(a.r - b.r) + std::fabs(a.ru - b.ru) + std::fabs(a.rv - b.rv) +
                            std::fabs(a.E - b.E)) / sc);
            e_same.add_abs((std::fabs(c.r - d.r) + std::fabs(c.ru - d.ru) + std::fabs(c.rv - d.rv) +
                            std::fabs(c.E - d.E)) / sc);
        }
        // primitive map
        {
            Prim wb{uni(-1, 1), uni(-1, 1), uni(-1, 1), uni(-1, 1)};
            auto f = [&](double s) {
                Prim p = to_prim(add(L, vL, s), g);
                return wb.r * p.r + wb.u * p.u + wb.v * p.v + wb.p * p.p;
            };
            double fd = (f(h) - f(-h)) / (2 * h);
            State Lb{0, 0, 0, 0};
            prim_rev(L, wb, Lb, g);
            e_prim.add(fd, dot(Lb, vL));
        }
        // limiter function
        {
            double d1 = uni(-1, 1), d2 = uni(-1, 1), ep = uni(1e-4, 0.5);
            double v1 = uni(-1, 1), v2 = uni(-1, 1), ve = uni(-1, 1);
            auto f = [&](double s) { return venkatakrishnan(d1 + s * v1, d2 + s * v2, ep + s * ve); };
            double fd = (f(h) - f(-h)) / (2 * h);
            double p1, p2, pe;
            venkatakrishnan_grad(d1, d2, ep, p1, p2, pe);
            e_venk.add(fd, p1 * v1 + p2 * v2 + pe * ve);
        }
        // edge -> length and normal
        {
            double ex = uni(-1, 1), ey = uni(-1, 1), vx = uni(-1, 1), vy = uni(-1, 1);
            double wnx = uni(-1, 1), wny = uni(-1, 1), wA = uni(-1, 1);
            auto f = [&](double s) {
                double x = ex + s * vx, y = ey + s * vy, A = std::sqrt(x * x + y * y);
                return wnx * (y / A) + wny * (-x / A) + wA * A;
            };
            double fd = (f(h) - f(-h)) / (2 * h);
            double A = std::sqrt(ex * ex + ey * ey), exb = 0, eyb = 0;
            edge_rev(ey / A, -ex / A, A, wnx, wny, wA, exb, eyb);
            e_edge.add(fd, exb * vx + eyb * vy);
        }
    }
    e_hllc.print("hllc");
    e_rus.print("rusanov");
    e_wall.print("slip_wall");
    e_rec.print("face_state");
    e_hllc_w.print("hllc_w");
    e_rus_w.print("rusanov_w");
    e_wall_w.print("slip_wall_w");
    e_same.print("primitive_equals_conserved");
    e_prim.print("primitive");
    e_venk.print("venkatakrishnan");
    e_edge.print("edge");
    std::printf("hllc_branches %d %d %d %d\n", hits[0], hits[1], hits[2], hits[3]);
    return 0;
}
