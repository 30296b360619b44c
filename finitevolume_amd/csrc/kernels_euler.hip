// Euler residual kernels for gfx950 (FP64, HBM-bound; no MFMA).
//
// One residual evaluation + stage update is three launches:
//   k_euler_grad  (order 2 only)  per cell: primitive LSQ gradients + Venkatakrishnan limiter
//   k_euler_face                  per face: (reconstruct) + numerical flux x length,
//                                 plus the step's dt candidates on the first stage
//   k_euler_cell                  per cell: gather of its faces' fluxes in the reference's
//                                 Cell::faces order, RK stage update, L2 residual partials
// The face pass then cell pass split is the reference's race-free structure
// (src/EulerFVMSolver.cpp:41-57): no atomics on the state, deterministic sums.
#include <cfloat>

#include "physics.h"
#include "solver.h"

namespace fvk {

constexpr int kBlock = 256;

struct EulerArgs {
    MeshView m;
    const double* U;      // stage state [4][ld]
    double* Unew;         // updated state [4][ld] (may alias U for forward Euler)
    const double* U0;     // RK base state [4][ld]
    double* G;            // limited primitive gradients [8][ld]
    double* flux;         // [4][ldf]
    double* dtf;          // [ldf] per-face dt candidate
    double* dtc;          // [ld] per-cell dt
    double* R;            // optional residual output [4][ld] (residual_eval)
    unsigned long long* dtmin;
    double* partials;
    unsigned int* ticket;
    double* res_hist;
    double* dt_hist;
    int64_t ld, ldf;
    double gamma, cfl, tol, venk_k;
    int maxit, limiter;
    double a, b;          // RK combination U' = a U0 + b (U + dt/V R) (stage > 0)
    int face_dt;          // face pass computes dt candidates (first stage)
    int dt_src;           // cell pass dt: 0 per-cell dtc, 1 global slot, 2 local min over faces (and store)
    int first;            // first stage: U' = U + dt/V R
    int want_res;         // accumulate residual norms
    int slot;             // dt slot of this step
    int64_t hist_idx;
};

__device__ __forceinline__ State ldU(const double* U, int64_t ld, int c) {
    return State{U[c], U[ld + c], U[2 * ld + c], U[3 * ld + c]};
}

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

// ghost state for the Euler equations, src/EulerFVMSolver.cpp:174-193
__device__ __forceinline__ State euler_ghost(const DevBC& bc, const State& L, double nx, double ny, double g) {
    if (bc.type == BC_FARFIELD) return State{bc.ff[0], bc.ff[1], bc.ff[2], bc.ff[3]};
    if (bc.type == BC_OUTFLOW) return L;
    return ghost_slip_wall(L, nx, ny, g);
}

// ---------------------------------------------------------------------------
// gradients + limiter (extension): LSQ of rho, u, v, p with the reference's
// weights 1/d^2 and precomputed inverse; boundary faces contribute
// 0.5 (W_c + W_ghost) at the face midpoint; Venkatakrishnan with
// eps^2 = (K sqrt(V))^3, phi = min(1, min over faces).
// ---------------------------------------------------------------------------
template <bool PERIODIC>
__global__ __launch_bounds__(kBlock) void k_euler_grad(EulerArgs a) {
    const MeshView& m = a.m;
    int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.nc) return;
    const double g = a.gamma;
    State Uc = ldU(a.U, a.ld, c);
    double wc[4] = {Uc.r, Uc.ru / Uc.r, Uc.rv / Uc.r, pressure(Uc, g)};
    double cx = m.cx[c], cy = m.cy[c];
    double bx[4] = {0.0, 0.0, 0.0, 0.0}, by[4] = {0.0, 0.0, 0.0, 0.0};
    double wmin[4], wmax[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) wmin[q] = wmax[q] = wc[q];
    const int k0 = m.cptr[c], k1 = m.cptr[c + 1];
    for (int k = k0; k < k1; ++k) {
        int e = m.cent[k];
        int f = e >> 1;
        bool owner = e & 1;
        int cr = m.fr[f];
        double qx, qy, wn[4];
        if (cr < 0) {
            const DevBC& bc = m.bc[-1 - cr];
            State G = euler_ghost(bc, Uc, m.fnx[f], m.fny[f], g);
            double wg0 = G.r, wg1 = G.ru / G.r, wg2 = G.rv / G.r, wg3 = pressure(G, g);
            wn[0] = 0.5 * (wc[0] + wg0);
            wn[1] = 0.5 * (wc[1] + wg1);
            wn[2] = 0.5 * (wc[2] + wg2);
            wn[3] = 0.5 * (wc[3] + wg3);
            qx = m.fmx[f];
            qy = m.fmy[f];
        } else {
            int o = owner ? cr : m.fl[f];
            State Uo = ldU(a.U, a.ld, o);
            wn[0] = Uo.r;
            wn[1] = Uo.ru / Uo.r;
            wn[2] = Uo.rv / Uo.r;
            wn[3] = pressure(Uo, g);
            qx = m.cx[o];
            qy = m.cy[o];
            if (PERIODIC) {
                if (owner) { qx = qx + m.fsx[f]; qy = qy + m.fsy[f]; }
                else { qx = qx - m.fsx[f]; qy = qy - m.fsy[f]; }
            }
        }
        double dx = qx - cx, dy = qy - cy;
        double w = 1.0 / (dx * dx + dy * dy);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double dphi = wn[q] - wc[q];
            bx[q] += w * dx * dphi;
            by[q] += w * dy * dphi;
            wmin[q] = smin(wmin[q], wn[q]);
            wmax[q] = smax(wmax[q], wn[q]);
        }
    }
    const double i0 = m.lsq[c], i1 = m.lsq[m.nc + c], i2 = m.lsq[2 * m.nc + c];
    double gx[4], gy[4], lim[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        gx[q] = i0 * bx[q] + i1 * by[q];
        gy[q] = i1 * bx[q] + i2 * by[q];
        lim[q] = 1.0;
    }
    if (a.limiter == FVK_LIMITER_VENKATAKRISHNAN) {
        double kh = a.venk_k * sqrt(m.vol[c]);
        double eps2 = kh * kh * kh;
        for (int k = k0; k < k1; ++k) {
            int e = m.cent[k];
            int f = e >> 1;
            double mx = m.fmx[f], my = m.fmy[f];
            if (PERIODIC && !(e & 1)) { mx = mx - m.fsx[f]; my = my - m.fsy[f]; }
            double rx = mx - cx, ry = my - cy;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double d2 = gx[q] * rx + gy[q] * ry;
                double ph = 1.0;
                if (d2 > 0.0) ph = venkatakrishnan(wmax[q] - wc[q], d2, eps2);
                else if (d2 < 0.0) ph = venkatakrishnan(wmin[q] - wc[q], d2, eps2);
                lim[q] = smin(lim[q], ph);
            }
        }
    }
    if (a.limiter == FVK_LIMITER_FIRST_ORDER) {
#pragma unroll
        for (int q = 0; q < 4; ++q) lim[q] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a.G[(2 * q) * a.ld + c] = lim[q] * gx[q];
        a.G[(2 * q + 1) * a.ld + c] = lim[q] * gy[q];
    }
}

// MUSCL face state (extension): the cell's own state when the increment is
// exactly zero or the reconstructed density/pressure is not positive.
__device__ __forceinline__ State reconstruct(const EulerArgs& a, const State& Uc, int c, double rx, double ry) {
    const double* G = a.G;
    const int64_t ld = a.ld;
    double d0 = G[c] * rx + G[ld + c] * ry;
    double d1 = G[2 * ld + c] * rx + G[3 * ld + c] * ry;
    double d2 = G[4 * ld + c] * rx + G[5 * ld + c] * ry;
    double d3 = G[6 * ld + c] * rx + G[7 * ld + c] * ry;
    if (d0 == 0.0 && d1 == 0.0 && d2 == 0.0 && d3 == 0.0) return Uc;
    double u = Uc.ru / Uc.r, v = Uc.rv / Uc.r, p = pressure(Uc, a.gamma);
    double rho = Uc.r + d0, pp = p + d3;
    if (!(rho > 0.0) || !(pp > 0.0)) return Uc;
    return from_primitive(rho, u + d1, v + d2, pp, a.gamma);
}

// ---------------------------------------------------------------------------
// face pass: EulerFVMSolver::step pass 1 (src/EulerFVMSolver.cpp:64-82) and,
// on the first stage, the compute_dt candidate of every face (:138-158)
// ---------------------------------------------------------------------------
template <int SCHEME, int ORDER, bool PERIODIC>
__global__ __launch_bounds__(kBlock) void k_euler_face(EulerArgs a) {
    const MeshView& m = a.m;
    int f = blockIdx.x * kBlock + threadIdx.x;
    double cand = DBL_MAX;
    if (f < m.nf) {
        const double g = a.gamma;
        int cl = m.fl[f], cr = m.fr[f];
        double nx = m.fnx[f], ny = m.fny[f], A = m.fa[f];
        State L = ldU(a.U, a.ld, cl);
        State R, Lf, Rf;
        if (ORDER == 2) {
            double mx = m.fmx[f], my = m.fmy[f];
            Lf = reconstruct(a, L, cl, mx - m.cx[cl], my - m.cy[cl]);
            if (cr >= 0) {
                R = ldU(a.U, a.ld, cr);
                double rcx = m.cx[cr], rcy = m.cy[cr];
                if (PERIODIC) { rcx = rcx + m.fsx[f]; rcy = rcy + m.fsy[f]; }
                Rf = reconstruct(a, R, cr, mx - rcx, my - rcy);
            } else {
                Rf = euler_ghost(m.bc[-1 - cr], Lf, nx, ny, g);
            }
        } else {
            Lf = L;
            if (cr >= 0) { R = ldU(a.U, a.ld, cr); Rf = R; }
            else Rf = euler_ghost(m.bc[-1 - cr], L, nx, ny, g);
        }
        State F = num_flux<SCHEME>(Lf, Rf, nx, ny, g, a.tol, a.maxit);
        a.flux[f] = A * F.r;
        a.flux[a.ldf + f] = A * F.ru;
        a.flux[2 * a.ldf + f] = A * F.rv;
        a.flux[3 * a.ldf + f] = A * F.E;
        if (a.face_dt) {
            double uL = L.ru / L.r, vL = L.rv / L.r;
            double VnL = uL * nx + vL * ny;
            double Smax = fabs(VnL) + sound_speed(L, g);
            if (cr >= 0) {
                double uR = R.ru / R.r, vR = R.rv / R.r;
                double VnR = uR * nx + vR * ny;
                Smax = smax(Smax, fabs(VnR) + sound_speed(R, g));
            }
            cand = a.cfl * m.fdl[f] / Smax;
            if (a.dt_src == 2) a.dtf[f] = cand;
        }
    }
    if (a.face_dt) {
        // block minimum (std::min semantics: NaN candidates never win), one atomic per block
        __shared__ double sm[kBlock];
        sm[threadIdx.x] = cand;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) sm[threadIdx.x] = smin(sm[threadIdx.x], sm[threadIdx.x + s]);
            __syncthreads();
        }
        // positive doubles order like their bit patterns; NaN bits exceed +inf and DBL_MAX
        if (threadIdx.x == 0) atomicMin(&a.dtmin[a.slot], dbits(sm[0]));
    }
}

// deterministic block sum of NV values per thread into partials[q * gridDim.x + block];
// the last block to finish reduces the partials in a fixed order.  Returns
// true in the last block (all threads), with the totals in tot[].
template <int NV>
__device__ __forceinline__ bool block_partials(const double (&v)[NV], double* partials, unsigned int* ticket,
                                               double (&tot)[NV]) {
    __shared__ double sm[NV][kBlock];
    __shared__ bool last;
#pragma unroll
    for (int q = 0; q < NV; ++q) sm[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int q = 0; q < NV; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[q * gridDim.x + blockIdx.x] = sm[q][0];
        __threadfence();
        unsigned int t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (unsigned int i = threadIdx.x; i < gridDim.x; i += kBlock) {
#pragma unroll
        for (int q = 0; q < NV; ++q)
            acc[q] += __hip_atomic_load(&partials[q * gridDim.x + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) sm[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int q = 0; q < NV; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + s];
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) tot[q] = sm[q][0];
    if (threadIdx.x == 0) *ticket = 0u;
    return true;
}

// ---------------------------------------------------------------------------
// cell pass: EulerFVMSolver::step pass 2 (src/EulerFVMSolver.cpp:89-111),
// generalised to an SSP-RK stage and a per-cell dt
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_euler_cell(EulerArgs a) {
    const MeshView& m = a.m;
    int c = blockIdx.x * kBlock + threadIdx.x;
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < m.nc) {
        double dmin = DBL_MAX;
        const int k0 = m.cptr[c], k1 = m.cptr[c + 1];
        for (int k = k0; k < k1; ++k) {
            int e = m.cent[k];
            int f = e >> 1;
            double F0 = a.flux[f], F1 = a.flux[a.ldf + f], F2 = a.flux[2 * a.ldf + f], F3 = a.flux[3 * a.ldf + f];
            if (e & 1) {
                r[0] = r[0] - F0; r[1] = r[1] - F1; r[2] = r[2] - F2; r[3] = r[3] - F3;
            } else {
                r[0] += F0; r[1] += F1; r[2] += F2; r[3] += F3;
            }
            if (a.dt_src == 2) dmin = smin(dmin, a.dtf[f]);
        }
        if (a.R) {
#pragma unroll
            for (int q = 0; q < 4; ++q) a.R[q * a.ld + c] = r[q];
        } else {
            double dt;
            if (a.dt_src == 1) dt = __longlong_as_double((long long)a.dtmin[a.slot]);
            else if (a.dt_src == 2) { dt = dmin; a.dtc[c] = dmin; }
            else dt = a.dtc[c];
            double h = dt / m.vol[c];
            if (a.first) {
#pragma unroll
                for (int q = 0; q < 4; ++q) a.Unew[q * a.ld + c] = a.U[q * a.ld + c] + h * r[q];
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    a.Unew[q * a.ld + c] = a.a * a.U0[q * a.ld + c] + a.b * (a.U[q * a.ld + c] + h * r[q]);
            }
        }
    }
    if (!a.want_res) return;
    double sq[4] = {r[0] * r[0], r[1] * r[1], r[2] * r[2], r[3] * r[3]};
    double tot[4];
    if (block_partials<4>(sq, a.partials, a.ticket, tot) && threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) a.res_hist[a.hist_idx * 4 + q] = sqrt(tot[q]);
        a.dt_hist[a.hist_idx] = __longlong_as_double((long long)a.dtmin[a.slot]);
        a.dtmin[a.slot ^ 1] = dbits(DBL_MAX);
    }
}

// ---------------------------------------------------------------------------
// host launcher: one residual evaluation + stage update
// ---------------------------------------------------------------------------
namespace {

template <int SCHEME, int ORDER>
void launch_face(bool periodic, int grid, hipStream_t st, const EulerArgs& a) {
    if (periodic) k_euler_face<SCHEME, ORDER, true><<<grid, kBlock, 0, st>>>(a);
    else k_euler_face<SCHEME, ORDER, false><<<grid, kBlock, 0, st>>>(a);
}

template <int ORDER>
void launch_face_scheme(int scheme, bool periodic, int grid, hipStream_t st, const EulerArgs& a) {
    if (scheme == FLUX_HLLC) launch_face<FLUX_HLLC, ORDER>(periodic, grid, st, a);
    else if (scheme == FLUX_EXACT) launch_face<FLUX_EXACT, ORDER>(periodic, grid, st, a);
    else launch_face<FLUX_RUSANOV, ORDER>(periodic, grid, st, a);
}

EulerArgs make_args(fvk_solver* s) {
    EulerArgs a{};
    a.m = s->mesh;
    a.G = s->grad;
    a.flux = s->flux;
    a.dtf = s->dtf;
    a.dtc = s->dtc;
    a.dtmin = s->dtmin;
    a.partials = s->partials;
    a.ticket = s->ticket;
    a.res_hist = s->res_hist;
    a.dt_hist = s->dt_hist;
    a.ld = s->ld;
    a.ldf = s->ldf;
    a.gamma = s->p.gamma;
    a.cfl = s->p.cfl;
    a.tol = s->p.exact_riemann_tol;
    a.maxit = s->p.exact_riemann_max_iter;
    a.venk_k = s->p.venkat_k;
    a.limiter = s->p.limiter;
    return a;
}

void run_face_grad(fvk_solver* s, EulerArgs& a) {
    const int gc = (int)((s->nc + kBlock - 1) / kBlock);
    const int gf = (int)((s->nf + kBlock - 1) / kBlock);
    hipEvent_t ev = nullptr;
    if (s->p.order == 2) {
        timer_begin(s, "euler_grad", &ev);
        if (s->periodic) k_euler_grad<true><<<gc, kBlock, 0, s->stream>>>(a);
        else k_euler_grad<false><<<gc, kBlock, 0, s->stream>>>(a);
        timer_end(s, ev);
    }
    timer_begin(s, "euler_face", &ev);
    if (s->p.order == 2) launch_face_scheme<2>(s->p.flux_scheme, s->periodic, gf, s->stream, a);
    else launch_face_scheme<1>(s->p.flux_scheme, s->periodic, gf, s->stream, a);
    timer_end(s, ev);
}

}  // namespace

// stage in [0, stages); the state buffers rotate A -> B -> C -> A so that no
// copy of the step's base state is made: stage 0 reads A (= U^n) and writes
// B (or A for one stage); later stages read the previous stage's buffer and
// A as the RK base, the last writes A.
int launch_euler_stage(fvk_solver* s, int stage, int stages, int /*dt_mode_first*/, int want_res, int64_t step) {
    static const double RA[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0.5, 0}, {0, 0.75, 1.0 / 3.0}};
    static const double RB[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0.5, 0}, {1, 0.25, 2.0 / 3.0}};
    EulerArgs a = make_args(s);
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
    run_face_grad(s, a);
    const int gc = (int)((s->nc + kBlock - 1) / kBlock);
    hipEvent_t ev = nullptr;
    timer_begin(s, "euler_cell", &ev);
    k_euler_cell<<<gc, kBlock, 0, s->stream>>>(a);
    timer_end(s, ev);
    return hip_check(hipGetLastError(), "euler stage launch");
}

int launch_euler_residual_eval(fvk_solver* s, double* R_dev) {
    EulerArgs a = make_args(s);
    a.U = s->U[s->cur];
    a.face_dt = 0;
    a.dt_src = 0;
    a.want_res = 0;
    a.R = R_dev;
    run_face_grad(s, a);
    const int gc = (int)((s->nc + kBlock - 1) / kBlock);
    k_euler_cell<<<gc, kBlock, 0, s->stream>>>(a);
    return hip_check(hipGetLastError(), "euler residual launch");
}

// plain streaming copy (bandwidth calibration of the HBM byte counters)
__global__ __launch_bounds__(kBlock) void k_copy(double2* __restrict__ dst, const double2* __restrict__ src, int64_t n2) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * kBlock;
    for (; i < n2; i += stride) dst[i] = src[i];
}

int launch_copy_probe(double* dst, const double* src, int64_t n, hipStream_t st) {
    int64_t n2 = n / 2;
    int grid = 256 * 16;
    k_copy<<<grid, kBlock, 0, st>>>(reinterpret_cast<double2*>(dst), reinterpret_cast<const double2*>(src), n2);
    return hip_check(hipGetLastError(), "copy probe launch");
}

}  // namespace fvk
