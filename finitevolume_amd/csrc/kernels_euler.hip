// Euler residual kernels for gfx950 (FP64, HBM-bound; no MFMA).
//
// One residual evaluation + RK stage update, two execution variants with
// bit-identical results:
//
//  variant 1 (face sweep + gather)
//   k_e_grad    (order 2 only)  per cell: primitive LSQ gradients + Venkatakrishnan limiter
//   k_e_face    per face: numerical flux x length (one double4 per face) and, on the
//               first stage, the step's dt candidates
//   k_e_gather  per cell: the cell's face fluxes in the reference's Cell::faces
//               order, owner minus (src/EulerFVMSolver.cpp:89-104), stage update,
//               L2 residual partials
//  variant 2 (fused)
//   k_e_dt      (global dt, first stage only)  per face: compute_dt candidate
//   k_e_grad    as above
//   k_e_cell    per cell: the flux of every face of the cell, evaluated in the
//               face's own orientation (owner on the left, stored normal) so both
//               cells of a face add the same bits, then the update
//
// Per-cell kernels are instantiated for the mesh's largest face count D (3,
// 4, 9 or 16) and fully unrolled: the D face entries are loaded together, then
// the D neighbours, so a cell waits for two memory round trips rather than
// one pair per face; faces are still summed in Cell::faces order.  State and
// gradients are arrays of per-cell vectors (double4): a gathered neighbour
// costs one or two cache lines.  Blocks are remapped so that each XCD (own
// L2) sweeps one contiguous range of the RCM-ordered cells or faces.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <mutex>
#include <unordered_map>

#include "physics.h"
#include "solver.h"
#include "tile.h"
#include "norms.h"

namespace fvk {

constexpr int kMaxBlock = 512;

struct EulerArgs {
    MeshView m;
    const double4* U;     // stage state [ld]
    double4* Unew;        // updated state [ld], never the buffer read as U
    const double4* U0;    // RK base state [ld]
    double4* G;           // limited primitive gradients [2*ld]: (rho_x, rho_y, u_x, u_y), (v_x, v_y, p_x, p_y)
    double4* flux;        // [nf] face flux x length (variant 1)
    double* dtf;          // [nf] per-face dt candidate (variant 1, local dt)
    double* dtc;          // [ld] per-cell dt
    double4* R;           // optional residual output [ld] (residual_eval)
    unsigned long long* dtmin;
    double* partials;
    double* res_hist;
    double* dt_hist;
    double gamma, cfl, tol;
    int maxit, limiter;
    double a, b;          // RK combination U' = a U0 + b (U + dt/V R) (stage > 0)
    int face_dt;          // this pass evaluates the step's dt candidates
    int dt_src;           // 0 per-cell dtc, 1 global slot, 2 local min over the cell's faces (and store)
    int first;            // first stage: U' = U + dt/V R
    int want_res;         // accumulate residual norms
    int slot;             // dt slot of this step
    int64_t hist_idx;
    // batched cases (null when there is one case): case ranges in device
    // order, per-case activity, case of the first cell of each logical block
    const int* case_start;
    const int* active;
    const int* case_lo;
    double* case_partials;
    int n_cases;
    int stages;
    // finalize of the previous stage's (or step's) norms, run by one spare
    // workgroup of the tile or face kernel (fin_nblocks == 0: none); with
    // fin_rearm it also re-arms the next step's dt slot; with fin_dt_row the
    // step's dt is the minimum of the partials' fifth row (tile kernels)
    int fin_nblocks, fin_slot, fin_rearm, fin_dt_row;
    int64_t fin_hist_idx;
    int rearm_slot;       // >= 0: the gather re-arms this dt slot (forward Euler with a folded finalize)
    // batched cases on the tile kernel: the tiles of the active cases
    // (tile_list[0 .. *n_tile_list)), a workgroup per entry; each
    // tile's residual sums go to tile_partials [4][ntiles]
    const int* tile_list;
    const int* n_tile_list;
    double* tile_partials;
    int block;  // the workgroup size of this launch (a kernel argument, not a dispatch-packet read)
    unsigned int* tile_queue;  // persistent launches: per-XCD claim counters, then the workgroups done
};



// case of device cell c (binary search of the case ranges)
__device__ __forceinline__ int case_of(const EulerArgs& a, int c) {
    int lo = 0, hi = a.n_cases;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.case_start[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

// false for the cells of a frozen case
__device__ __forceinline__ bool cell_active(const EulerArgs& a, int c) {
    return a.active == nullptr || a.active[case_of(a, c)] != 0;
}

__device__ __forceinline__ State st(const double4& u) { return State{u.x, u.y, u.z, u.w}; }

// ghost state for the Euler equations, src/EulerFVMSolver.cpp:174-193
__device__ __forceinline__ State euler_ghost(const DevBC& bc, const State& L, double nx, double ny, double g) {
    return euler_ghost_state(bc.type, bc.ff, L, nx, ny, g);
}

// compute_dt candidate of one face (src/EulerFVMSolver.cpp:139-158), states in
// the face's orientation; R == nullptr on a boundary face.
__device__ __forceinline__ double dt_candidate(const State& L, const State* R, double nx, double ny, double len,
                                               double g, double cfl) {
    double uL = fdiv(L.ru, L.r), vL = fdiv(L.rv, L.r);
    double VnL = uL * nx + vL * ny;
    double Smax = fabs(VnL) + sound_speed(L, g);
    if (R) {
        double uR = fdiv(R->ru, R->r), vR = fdiv(R->rv, R->r);
        double VnR = uR * nx + vR * ny;
        Smax = smax(Smax, fabs(VnR) + sound_speed(*R, g));
    }
    return fdiv(cfl * len, Smax);
}

// dt_candidate() from the cells' primitives: the same bits, since W holds
// the results of the same divisions and pressure()
__device__ __forceinline__ double dt_candidate_w(const Prim& L, const Prim* R, double nx, double ny, double len,
                                                 double g, double cfl) {
    double Smax = fabs(L.u * nx + L.v * ny) + sqrt(fdiv(g * L.p, L.r));
    if (R) Smax = smax(Smax, fabs(R->u * nx + R->v * ny) + sqrt(fdiv(g * R->p, R->r)));
    return fdiv(cfl * len, Smax);
}

// block minimum, then one atomic per block.  Positive doubles order like
// their bit patterns and NaN bits exceed +inf and DBL_MAX, so a NaN
// candidate never wins (std::min semantics of the reference's reduction).
// the workgroup's minimum of v into the dt slot (its XCD's sub-slot)
__device__ __forceinline__ void block_min_to(double v, unsigned long long* dtmin, int slot) {
    __shared__ double sm[kMaxBlock];
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = smin(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) dt_atomic_min(dtmin, slot, sm[0]);
}

// deterministic block sum of the four squared residual components into
// partials[q * gridDim.x + blockIdx.x] (fixed tree order); k_e_finalize
// reduces them after the kernel ends, so no block needs a device-scope fence
// (a release at agent scope writes the XCD's L2 back)
__device__ __forceinline__ void block_partials4(const double (&v)[4], double* partials) {
    __shared__ double sm[4][kMaxBlock];
#pragma unroll
    for (int q = 0; q < 4; ++q) sm[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int q = 0; q < 4; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) partials[q * gridDim.x + blockIdx.x] = sm[q][0];
    }
}

// batched cases: the block's cells belong to case_lo[block] and possibly the
// next case; the squared residuals are summed separately for the two (fixed
// tree order) into case_partials[(q * 2 + side) * nblocks + logical block]
__device__ __forceinline__ void block_partials_cases(const EulerArgs& a, int side, const double (&v)[4]) {
    __shared__ double sm[4][kMaxBlock];
    const int lb = logical_block();
    for (int s2 = 0; s2 < 2; ++s2) {
#pragma unroll
        for (int q = 0; q < 4; ++q) sm[q][threadIdx.x] = side == s2 ? v[q] : 0.0;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
#pragma unroll
                for (int q = 0; q < 4; ++q) sm[q][threadIdx.x] += sm[q][threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) a.case_partials[(q * 2 + s2) * gridDim.x + lb] = sm[q][0];
        }
        __syncthreads();
    }
}

// one block per case: the case's residual norms from the block partials of
// its cell range, into row (*step mod cap) of the per-case history
constexpr int kCaseFinBlock = 256;
__global__ __launch_bounds__(kCaseFinBlock) void k_case_finalize(const double* partials, int nblocks,
                                                                 const int* case_blocks, const int* case_lo,
                                                                 double* case_res, const unsigned int* step,
                                                                 int cap, int n_cases) {
    __shared__ double sm[kCaseFinBlock];
    const int k = blockIdx.x;
    const int b0 = case_blocks[2 * k], b1 = case_blocks[2 * k + 1];
    const int64_t row = (int64_t)(*step % (unsigned int)cap);
    for (int q = 0; q < 4; ++q) {
        double acc = 0.0;
        for (int lb = b0 + (int)threadIdx.x; lb <= b1; lb += kCaseFinBlock)
            acc += partials[(q * 2 + (case_lo[lb] == k ? 0 : 1)) * nblocks + lb];
        sm[threadIdx.x] = acc;
        __syncthreads();
        for (int s = kCaseFinBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) case_res[(row * n_cases + k) * 4 + q] = sqrt(sm[0]);
        __syncthreads();
    }
}

__global__ void k_step_tick(unsigned int* step) { *step += 1u; }

constexpr int kFinBlock = 1024;
__global__ __launch_bounds__(kFinBlock) void k_e_finalize(const double* partials, int nblocks, int nv,
                                                          double* res_hist, double* dt_hist, int64_t hist_idx,
                                                          unsigned long long* dtmin, int slot,
                                                          double fixed_dt = -1.0, int rearm = 1,
                                                          int dt_row = 0) {
    __shared__ double sm[kFinBlock];
    finalize_norms(partials, nblocks, nv, res_hist, dt_hist, hist_idx, dtmin, slot, fixed_dt, sm, rearm != 0,
                   dt_row != 0);
}

// stage update of one cell: U' = U + dt/V R (src/EulerFVMSolver.cpp:103) on
// the first stage, a U0 + b (U + dt/V R) on later SSP-RK stages; or the
// residual itself for residual_eval
// (the loads split from the arithmetic, so a kernel can issue them early)
struct UpdateInputs {
    double4 u0;      // stage base state (later SSP-RK stages)
    double vol, dt;  // cell volume; dt unless it is the cell's own new local dt
};
__device__ __forceinline__ UpdateInputs load_update_inputs(const EulerArgs& a, int c) {
    UpdateInputs x{};
    if (a.R) return x;
    x.vol = a.m.vol[c];
    if (a.dt_src == 1) x.dt = dt_read(a.dtmin, a.slot);
    else if (a.dt_src == 0) x.dt = a.dtc[c];
    if (!a.first) x.u0 = a.U0[c];
    return x;
}
__device__ __forceinline__ void update_cell_with(const EulerArgs& a, int c, const UpdateInputs& x,
                                                 const double4& u, double r0, double r1, double r2, double r3,
                                                 double dmin) {
    if (a.R) {
        a.R[c] = double4{r0, r1, r2, r3};
        return;
    }
    double dt = x.dt;
    if (a.dt_src == 2) { dt = dmin; a.dtc[c] = dmin; }
    const double h = fdiv(dt, x.vol);
    if (a.first) {
        a.Unew[c] = double4{u.x + h * r0, u.y + h * r1, u.z + h * r2, u.w + h * r3};
    } else {
        const double4 u0 = x.u0;
        a.Unew[c] = double4{a.a * u0.x + a.b * (u.x + h * r0), a.a * u0.y + a.b * (u.y + h * r1),
                            a.a * u0.z + a.b * (u.z + h * r2), a.a * u0.w + a.b * (u.w + h * r3)};
    }
}
__device__ __forceinline__ void update_cell(const EulerArgs& a, int c, const double4& u, double r0, double r1,
                                            double r2, double r3, double dmin) {
    update_cell_with(a, c, load_update_inputs(a, c), u, r0, r1, r2, r3, dmin);
}

// per-block partial sums of the squared residual (L2 norms of the stage)
__device__ __forceinline__ void finish_residual(const EulerArgs& a, double r0, double r1, double r2, double r3) {
    if (!a.want_res) return;
    double sq[4] = {r0 * r0, r1 * r1, r2 * r2, r3 * r3};
    block_partials4(sq, a.partials);
}

// ---------------------------------------------------------------------------
// global dt: compute_dt over all faces (reference mode, variant 2)
// ---------------------------------------------------------------------------
// Each workgroup takes a contiguous range of faces and leaves one atomic
// minimum: the grid is bounded (dt_grid), since device-wide atomics on one
// address are serialised across the XCDs.
__global__ __launch_bounds__(kMaxBlock) void k_e_dt(EulerArgs a) {
    const MeshView& m = a.m;
    const int64_t per = ((int64_t)m.nf + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)logical_block() * per;
    const int64_t hi = lo + per < m.nf ? lo + per : m.nf;
    double cand = DBL_MAX;
    for (int64_t f = lo + threadIdx.x; f < hi; f += blockDim.x) {
        const int cl = m.fl[f], cr = m.fr[f];
        const double4 fg = m.fgeo[f];
        const State L = st(a.U[cl]);
        double c;
        if (cr >= 0) {
            const State R = st(a.U[cr]);
            c = dt_candidate(L, &R, fg.x, fg.y, fg.w, a.gamma, a.cfl);
        } else {
            c = dt_candidate(L, nullptr, fg.x, fg.y, fg.w, a.gamma, a.cfl);
        }
        cand = smin(cand, c);
    }
    block_min_to(cand, a.dtmin, a.slot);
}

// ---------------------------------------------------------------------------
// gradients + limiter (extension): LSQ of rho, u, v, p with the reference's
// weights 1/d^2 and precomputed inverse; a boundary face contributes
// 0.5 (W_c + W_ghost) at its midpoint; Venkatakrishnan with
// eps^2 = (K sqrt(V))^3, phi = min(1, min over faces); the first-order
// limiter sets phi = 0.
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kMaxBlock) void k_e_grad(EulerArgs a) {
    const MeshView& m = a.m;
    const int c = xcd_index();
    if (c >= m.nc || !cell_active(a, c)) return;
    const double g = a.gamma;
    const int64_t ld = m.ld;
    int2 en[D];
    double4 eg[D], un[D];
#pragma unroll
    for (int k = 0; k < D; ++k) en[k] = m.ell[k * ld + c];
#pragma unroll
    for (int k = 0; k < D; ++k)  // w dx, w dy, face offset (precomputed, same bits); padding slots are not read
        eg[k] = en[k].x >= 0 ? m.egeo[k * ld + c] : double4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < D; ++k)
        if (en[k].x >= 0 && en[k].y >= 0) un[k] = a.U[en[k].y];
    const State Uc = st(a.U[c]);
    const double wc[4] = {Uc.r, fdiv(Uc.ru, Uc.r), fdiv(Uc.rv, Uc.r), pressure(Uc, g)};
    double bx[4] = {0.0, 0.0, 0.0, 0.0}, by[4] = {0.0, 0.0, 0.0, 0.0};
    double dlo[4] = {0.0, 0.0, 0.0, 0.0}, dhi[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        if (en[k].x < 0) continue;  // ELL padding (always after the cell's faces)
        double wn[4];
        if (en[k].y < 0) {
            const double4 fg = m.fgeo[en[k].x >> 1];
            const State Gs = euler_ghost(m.bc[-1 - en[k].y], Uc, fg.x, fg.y, g);
            const double wg0 = Gs.r, wg1 = fdiv(Gs.ru, Gs.r), wg2 = fdiv(Gs.rv, Gs.r), wg3 = pressure(Gs, g);
            wn[0] = 0.5 * (wc[0] + wg0);
            wn[1] = 0.5 * (wc[1] + wg1);
            wn[2] = 0.5 * (wc[2] + wg2);
            wn[3] = 0.5 * (wc[3] + wg3);
        } else {
            const State Uo = st(un[k]);
            wn[0] = Uo.r;
            wn[1] = fdiv(Uo.ru, Uo.r);
            wn[2] = fdiv(Uo.rv, Uo.r);
            wn[3] = pressure(Uo, g);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double dphi = wn[q] - wc[q];
            bx[q] += eg[k].x * dphi;
            by[q] += eg[k].y * dphi;
            venkat_extend(dlo[q], dhi[q], dphi);
        }
    }
    const double4 iv = m.clsq[c];  // LSQ inverse and eps^2
    double gx[4], gy[4], lim[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        gx[q] = iv.x * bx[q] + iv.y * by[q];
        gy[q] = iv.y * bx[q] + iv.z * by[q];
        lim[q] = 1.0;
    }
    if (a.limiter == FVK_LIMITER_VENKATAKRISHNAN) {
        const double eps2 = iv.w;
        VenkatMin vm[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) vm[q] = VenkatMin{1.0, 1.0};
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (en[k].x < 0) continue;
            const double rx = eg[k].z, ry = eg[k].w;
#pragma unroll
            for (int q = 0; q < 4; ++q) venkat_pick_d(vm[q], dlo[q], dhi[q], gx[q] * rx + gy[q] * ry, eps2);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) lim[q] = venkat_value(vm[q]);
    } else if (a.limiter == FVK_LIMITER_FIRST_ORDER) {
#pragma unroll
        for (int q = 0; q < 4; ++q) lim[q] = 0.0;
    }
    a.G[2 * c] = double4{lim[0] * gx[0], lim[0] * gy[0], lim[1] * gx[1], lim[1] * gy[1]};
    a.G[2 * c + 1] = double4{lim[2] * gx[2], lim[2] * gy[2], lim[3] * gx[3], lim[3] * gy[3]};
}

// ---------------------------------------------------------------------------
// fused cell pass (variant 2): the cell's face fluxes, face by face in the
// face's orientation (EulerFVMSolver::step pass 1), gathered in Cell::faces
// order with the owner's minus sign (pass 2), then the update
// ---------------------------------------------------------------------------
template <int D, int SCHEME, int ORDER>
__global__ __launch_bounds__(kMaxBlock) void k_e_cell(EulerArgs a) {
    const MeshView& m = a.m;
    const int c = xcd_index();
    const double g = a.gamma;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    double dmin = DBL_MAX;
    if (c < m.nc) {
        const int64_t ld = m.ld;
        int2 en[D];
#pragma unroll
        for (int k = 0; k < D; ++k) en[k] = m.ell[k * ld + c];
        const double4 uc4 = a.U[c];
        const State Uc = st(uc4);
        double4 gc0{}, gc1{};
        if (ORDER == 2) {
            gc0 = a.G[2 * c];
            gc1 = a.G[2 * c + 1];
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (en[k].x < 0) continue;
            const int f = en[k].x >> 1;
            const bool owner = en[k].x & 1;
            const double4 fg = m.fgeo[f];
            State F;
            double cand = DBL_MAX;
            if (en[k].y < 0) {
                // boundary face: this cell is the owner
                if (ORDER == 2) {
                    const double4 ro = m.frec[f];
                    const Prim Lf = face_state_w(Uc, gc0, gc1, ro.x, ro.y, g);
                    const DevBC& bc = m.bc[-1 - en[k].y];
                    F = num_flux_w<SCHEME>(Lf, euler_ghost_w(bc.type, bc.ff, Lf, fg.x, fg.y, g), fg.x, fg.y, g,
                                           a.tol, a.maxit);
                } else {
                    const State Rf = euler_ghost(m.bc[-1 - en[k].y], Uc, fg.x, fg.y, g);
                    F = num_flux<SCHEME>(Uc, Rf, fg.x, fg.y, g, a.tol, a.maxit);
                }
                if (a.face_dt) cand = dt_candidate(Uc, nullptr, fg.x, fg.y, fg.w, g, a.cfl);
            } else {
                const State Uo = st(a.U[en[k].y]);
                const State& L = owner ? Uc : Uo;
                const State& R = owner ? Uo : Uc;
                if (ORDER == 2) {
                    const double4 go0 = a.G[2 * en[k].y], go1 = a.G[2 * en[k].y + 1];
                    const double4 ro = m.frec[f];
                    const Prim Lf = face_state_w(L, owner ? gc0 : go0, owner ? gc1 : go1, ro.x, ro.y, g);
                    const Prim Rf = face_state_w(R, owner ? go0 : gc0, owner ? go1 : gc1, ro.z, ro.w, g);
                    F = num_flux_w<SCHEME>(Lf, Rf, fg.x, fg.y, g, a.tol, a.maxit);
                } else {
                    F = num_flux<SCHEME>(L, R, fg.x, fg.y, g, a.tol, a.maxit);
                }
                if (a.face_dt) cand = dt_candidate(L, &R, fg.x, fg.y, fg.w, g, a.cfl);
            }
            const double A = fg.z;
            const double F0 = A * F.r, F1 = A * F.ru, F2 = A * F.rv, F3 = A * F.E;
            if (owner) {
                r0 = r0 - F0; r1 = r1 - F1; r2 = r2 - F2; r3 = r3 - F3;
            } else {
                r0 += F0; r1 += F1; r2 += F2; r3 += F3;
            }
            if (a.face_dt) dmin = smin(dmin, cand);
        }
        update_cell(a, c, uc4, r0, r1, r2, r3, dmin);
    }
    // local dt: the global minimum is still reported (dt history)
    if (a.face_dt) block_min_to(dmin, a.dtmin, a.slot);
    finish_residual(a, r0, r1, r2, r3);
}

// ---------------------------------------------------------------------------
// face sweep (variant 1): EulerFVMSolver::step pass 1 (src/EulerFVMSolver.cpp:
// 64-82) with MUSCL states for order 2, one thread per face, writing the
// face flux x length as one double4; the first stage also evaluates the
// step's dt candidates (:139-158)
// ---------------------------------------------------------------------------
template <int SCHEME, int ORDER>
__global__ __launch_bounds__(kMaxBlock) void k_e_face(EulerArgs a) {
    const MeshView& m = a.m;
    // a pending finalize takes the first logical workgroup, dispatched first so
    // that its one-workgroup reduction overlaps the faces; the faces follow it
    int lb = logical_block();
    if (a.fin_nblocks > 0) {
        if (lb == 0) {
            __shared__ double fsm[kMaxBlock];
            finalize_norms(a.partials, a.fin_nblocks, 4, a.res_hist, a.dt_hist, a.fin_hist_idx, a.dtmin,
                           a.fin_slot, -1.0, fsm, a.fin_rearm != 0, a.fin_dt_row != 0);
            return;
        }
        --lb;
    }
    const int f = lb * blockDim.x + threadIdx.x;
    const double g = a.gamma;
    double cand = DBL_MAX;
    if (f < m.nf && cell_active(a, m.fl[f])) {
        const int cl = m.fl[f], cr = m.fr[f];
        const double4 fg = m.fgeo[f];
        const State L = st(a.U[cl]);
        State R = L;
        if (cr >= 0) R = st(a.U[cr]);
        State F;
        if (ORDER == 2) {
            const double4 ro = m.frec[f];  // midpoint - centroid, left and right (precomputed, same bits)
            const Prim Lf = face_state_w(L, a.G[2 * cl], a.G[2 * cl + 1], ro.x, ro.y, g);
            Prim Rf;
            if (cr >= 0) {
                Rf = face_state_w(R, a.G[2 * cr], a.G[2 * cr + 1], ro.z, ro.w, g);
            } else {
                const DevBC& bc = m.bc[-1 - cr];
                Rf = euler_ghost_w(bc.type, bc.ff, Lf, fg.x, fg.y, g);
            }
            F = num_flux_w<SCHEME>(Lf, Rf, fg.x, fg.y, g, a.tol, a.maxit);
        } else {
            F = num_flux<SCHEME>(L, cr >= 0 ? R : euler_ghost(m.bc[-1 - cr], L, fg.x, fg.y, g), fg.x, fg.y, g, a.tol,
                                 a.maxit);
        }
        const double A = fg.z;
        a.flux[f] = double4{A * F.r, A * F.ru, A * F.rv, A * F.E};
        if (a.face_dt) {
            cand = dt_candidate(L, cr >= 0 ? &R : nullptr, fg.x, fg.y, fg.w, g, a.cfl);
            if (a.dt_src == 2) a.dtf[f] = cand;
        }
    }
    if (a.face_dt) block_min_to(cand, a.dtmin, a.slot);
}

// ---------------------------------------------------------------------------
// gather (variant 1): EulerFVMSolver::step pass 2 (src/EulerFVMSolver.cpp:
// 89-104) over the face fluxes, Cell::faces order, owner minus; stage update
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kMaxBlock) void k_e_gather(EulerArgs a) {
    const MeshView& m = a.m;
    const int c = xcd_index();
    if (a.rearm_slot >= 0 && c == 0) dt_arm(a.dtmin, a.rearm_slot);  // unused until the next step
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    int side = 0;
    bool act = true;
    if (a.active && c < m.nc) {
        const int k = case_of(a, c);
        act = a.active[k] != 0;
        side = k != a.case_lo[logical_block()];
        // a frozen case keeps its state: forward Euler writes the other buffer
        // of the pair, so it is copied; RK stages leave the base state alone
        if (!act && a.stages == 1 && !a.R) a.Unew[c] = a.U[c];
    }
    if (c < m.nc && act) {
        const int64_t ld = m.ld;
        int2 en[D];
        double4 F[D];
        double df[D];
#pragma unroll
        for (int k = 0; k < D; ++k) en[k] = m.ell[k * ld + c];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (en[k].x >= 0) {
                F[k] = a.flux[en[k].x >> 1];
                if (a.dt_src == 2) df[k] = a.dtf[en[k].x >> 1];
            }
        }
        const double4 u = a.U[c];
        double dmin = DBL_MAX;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (en[k].x < 0) continue;
            if (en[k].x & 1) {
                r0 = r0 - F[k].x; r1 = r1 - F[k].y; r2 = r2 - F[k].z; r3 = r3 - F[k].w;
            } else {
                r0 += F[k].x; r1 += F[k].y; r2 += F[k].z; r3 += F[k].w;
            }
            if (a.dt_src == 2) dmin = smin(dmin, df[k]);
        }
        update_cell(a, c, u, r0, r1, r2, r3, dmin);
    }
    if (a.active) {
        if (a.want_res) {
            const double sq[4] = {r0 * r0, r1 * r1, r2 * r2, r3 * r3};
            block_partials_cases(a, side, sq);
        }
    } else {
        finish_residual(a, r0, r1, r2, r3);
    }
}

// sum and minimum over a wavefront by a butterfly, in a fixed order (the
// callers take lane 0's value; smin ignores a NaN from the partner lane, as
// std::min ignores one in its second argument)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = smin(v, __shfl_xor(v, o, 64));
    return v;
}

// flux x length of tile face e (as k_e_face), from LDS; returns its dt candidate
template <int SCHEME, int ORDER>
__device__ __forceinline__ double tile_face(const EulerArgs& a, const TileLds& L, int e, int fd,
                                            const double4& fg) {
    const MeshView& m = a.m;
    const double g = a.gamma;
    const int MG = L.MG, MF = L.MF;
    const int l = tface_left(fd), r = tface_right(fd);
    State F;
    double cand = DBL_MAX;
    if (ORDER == 2) {
        // primitive face states: the cells' primitives from LDS plus their increments
        const double* in = L.I + e;
        const Prim Wl{L.U[l], L.W[l], L.W[MG + l], L.W[2 * MG + l]};
        const Prim Lf = reconstruct_w(Wl, in[0], in[MF], in[2 * MF], in[3 * MF]);
        Prim Wr, Rf;
        if (r >= 0) {
            Wr = Prim{L.U[r], L.W[r], L.W[MG + r], L.W[2 * MG + r]};
            Rf = reconstruct_w(Wr, in[4 * MF], in[5 * MF], in[6 * MF], in[7 * MF]);
        } else {
            const DevBC& bc = m.bc[-1 - r];
            Rf = euler_ghost_w(bc.type, bc.ff, Lf, fg.x, fg.y, g);
        }
        F = num_flux_w<SCHEME>(Lf, Rf, fg.x, fg.y, g, a.tol, a.maxit);
        if (a.face_dt) cand = dt_candidate_w(Wl, r >= 0 ? &Wr : nullptr, fg.x, fg.y, fg.w, g, a.cfl);
    } else {
        const State Ls{L.U[l], L.U[MG + l], L.U[2 * MG + l], L.U[3 * MG + l]};
        State R = Ls, Rf;
        if (r >= 0) {
            R = State{L.U[r], L.U[MG + r], L.U[2 * MG + r], L.U[3 * MG + r]};
            Rf = R;
        } else {
            Rf = euler_ghost(m.bc[-1 - r], Ls, fg.x, fg.y, g);
        }
        F = num_flux<SCHEME>(Ls, Rf, fg.x, fg.y, g, a.tol, a.maxit);
        if (a.face_dt) cand = dt_candidate(Ls, r >= 0 ? &R : nullptr, fg.x, fg.y, fg.w, g, a.cfl);
    }
    const double A = fg.z;
    L.F[e] = A * F.r;
    L.F[MF + e] = A * F.ru;
    L.F[2 * MF + e] = A * F.rv;
    L.F[3 * MF + e] = A * F.E;
    if (a.face_dt) L.Dt[e] = cand;
    return cand;
}

// ---------------------------------------------------------------------------
// fused cell-tile pass (variant 3): one workgroup per tile of graph-adjacent
// cells.  The tile's states and its halo (one ring for the fluxes, a second
// for the neighbours' gradients) are staged in LDS.  The gradients of the
// owned cells and the first ring leave each tile face's MUSCL increments in
// LDS; the fluxes of every face of the owned cells and the gather + update
// then read LDS only.  Every value is computed by the
// same operations as in the face sweep + gather kernels, so the states are
// bit-identical; faces on a tile's edge are evaluated by both tiles.  The
// mesh data of each thread's first gradient cell is loaded before the
// staging barrier, so its latency overlaps the staging; the data of its
// first faces and its update, before the barrier that precedes their use.
// ---------------------------------------------------------------------------
// WPE: the wavefronts per SIMD the kernel is held to.  At 1 (no bound) the
// mesh data of each thread's first two faces is loaded before the staging
// barrier too, which costs registers.  At 4 (at most 128 VGPRs; for tiles
// small enough that four workgroups share a CU's LDS) it is loaded once the
// thread's gradients are done, before the barrier that ends them; triangles
// with the HLLC and Rusanov fluxes fit so without spills.
// tid_: the thread index, when the caller has one (a persistent loop makes
// it opaque per tile); else threadIdx.x
template <int D, int SCHEME, int ORDER, int WPE, bool BATCH>
__device__ __forceinline__ void tile_body(const EulerArgs& a, double* lds, int t, int tid_ = -1) {
    constexpr bool PF = WPE < 4;
    const MeshView& m = a.m;
    const double g = a.gamma;
    const int bs = a.block, tid = tid_ >= 0 ? tid_ : (int)threadIdx.x;
    TileLds L;
    L.MG = m.tmax_g;
    L.M2 = m.tmax_2;
    L.MF = m.tmax_f;
    const int MG = L.MG, MF = L.MF;
    L.U = lds;
    if (ORDER == 2) {
        L.W = lds + 4 * MG;
        L.W2 = lds + 7 * MG;
        L.I = L.W2 + 4 * L.M2;
        L.F = L.I;
    } else {
        L.W = L.W2 = L.I = nullptr;
        L.F = lds + 4 * MG;
    }
    L.Dt = L.F + 4 * MF;
    // the tile's ranges: one record, one scalar load
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
    const int ng = n + h1, nl = n + nh;
    L.ng = ng;
    // the loads below are issued unconditionally, from addresses clamped into
    // their arrays, and their values used only where they belong to a cell
    // of the tile: with no branch around them the wait for the first ones
    // leaves the later ones in flight (threads past the tile's staged cells
    // read its last row again, from the cache; a spare workgroup past the
    // last tile reads the last tile's first row)
    const int ta = min(t, m.ntiles - 1), jt = min(tid, max(ng - 1, 0));
    const int* te = m.tell + (int64_t)ta * D * MG;

    // 0. loads issued before the staging barrier, so that their latency
    // overlaps it.  First the halo cells' device ids of this thread's first
    // two staged cells (the states' addresses depend on them; thalo is
    // non-empty whenever a tile has a halo); then the local ELL of cell tid
    // (its gradient, and its gather when owned), its gradient mesh data and
    // (PF) the mesh data of the first two faces of this thread
    // (up to nine slots per cell are held in registers, so their loads run
    // ahead of their use; LEAN: cells wider than that read their slots
    // where they use them, which keeps sixteen-slot cells from spilling).
    // Where cells may have fewer faces than slots (D > 3), a slot's geometry
    // is loaded only if the slot is a face: the padding is not fetched, and
    // the wait for the slots coincides with the one for the halo ids
    constexpr bool PAD = D > 3;
    int hc[2] = {0, 0};
    if (m.nthalo > 0) {
#pragma unroll
        for (int i = 0; i < 2; ++i) hc[i] = m.thalo[min(max(hp + tid + i * bs - n, 0), m.nthalo - 1)];
    }
    constexpr bool LEAN = D > 9;
    int pen[D];
    double4 peg[LEAN ? 1 : D], piv{0, 0, 0, 0};
    if (!LEAN) {
#pragma unroll
        for (int k = 0; k < D; ++k) pen[k] = te[k * MG + jt];
    }
    // tile-ordered geometry (second order only): no index load first
    const double4* tg = ORDER == 2 ? m.tegeo + (int64_t)ta * D * MG : nullptr;
    if (ORDER == 2) {
        if (!LEAN) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                if constexpr (PAD) peg[k] = pen[k] != -1 ? tg[k * MG + jt] : double4{0, 0, 0, 0};
                else peg[k] = tg[k * MG + jt];
            }
        }
        piv = m.tclsq[(int64_t)ta * MG + jt];
    }
    int pfd[2] = {0, 0};
    double4 pfg[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    if (PF) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (tid + i * bs < nft) {
                pfd[i] = m.tfc[f0 + tid + i * bs];
                pfg[i] = m.tfgeo[f0 + tid + i * bs];
            }
        }
    }

    // (keep the scheduler from moving the use of the halo ids above the loads
    // just issued, which would wait for the ids before issuing them)
    __builtin_amdgcn_sched_barrier(0);
    // the staging's loads at a raised wave priority: a workgroup starting a
    // tile issues them ahead of the arithmetic of the CU's other workgroups
    __builtin_amdgcn_s_setprio(2);

    // 1. states (and primitives) of the tile and its halo; the first two
    // passes load (unconditionally, cell 0 for a thread without a cell)
    // before they store
    auto stage_cell = [&](int j, const double4& u) { tile_stage(L, j, u, g, ORDER == 2); };
    {
        const int j0 = tid, j1 = tid + bs;
        const int g0 = j0 < n ? c0 + j0 : (j0 < nl ? hc[0] : -1);
        const int g1 = j1 < n ? c0 + j1 : (j1 < nl ? hc[1] : -1);
        const double4 u0 = a.U[max(g0, 0)], u1 = a.U[max(g1, 0)];
        if (g0 >= 0) stage_cell(j0, u0);
        if (g1 >= 0) stage_cell(j1, u1);
    }
    for (int j = tid + 2 * bs; j < nl; j += bs) stage_cell(j, a.U[j < n ? c0 + j : m.thalo[hp + j - n]]);
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();

    // 2. limited LSQ gradients of the owned cells and ring 1, and the MUSCL
    // increments of the tile faces on their sides
    if (ORDER == 2) {
        const double2* rec = m.trec ? m.trec + (int64_t)t * D * MG : nullptr;
        if constexpr (LEAN) {
            for (int j = tid; j < ng; j += bs)
                tile_grad_lean<D>(m, g, a.limiter, L, j, te + j, tg + j, j == tid ? piv : m.tclsq[(int64_t)t * MG + j],
                                  f0, rec ? rec + j : nullptr);
        } else {
            if (tid < ng) tile_grad<D>(m, g, a.limiter, L, tid, pen, peg, piv, f0, rec ? rec + tid : nullptr);
        }
        for (int j = tid + bs; !LEAN && j < ng; j += bs) {
            int en[D];
            double4 eg[D];
#pragma unroll
            for (int k = 0; k < D; ++k) en[k] = te[k * MG + j];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                if constexpr (PAD) eg[k] = en[k] != -1 ? tg[k * MG + j] : double4{0, 0, 0, 0};
                else eg[k] = tg[k * MG + j];
            }
            tile_grad<D>(m, g, a.limiter, L, j, en, eg, m.tclsq[(int64_t)t * MG + j], f0, rec ? rec + j : nullptr);
        }
    }
    if (!PF) {  // the first two faces' data, loaded while the gradients finish
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (tid + i * bs < nft) {
                pfd[i] = m.tfc[f0 + tid + i * bs];
                pfg[i] = m.tfgeo[f0 + tid + i * bs];
            }
        }
    }
    // the update's inputs of owned cell tid (LEAN: after the face pass)
    UpdateInputs upd{};
    if (!LEAN && tid < n) upd = load_update_inputs(a, c0 + tid);
    __syncthreads();

    // 3. fluxes of the tile's faces
    double cand_min = DBL_MAX;
    for (int i = 0, e = tid; e < nft; ++i, e += bs) {
        int fd = i == 0 ? pfd[0] : pfd[1];
        double4 fg = i == 0 ? pfg[0] : pfg[1];
        if (i >= 2) {
            fd = m.tfc[f0 + e];
            fg = m.tfgeo[f0 + e];
        }
        cand_min = smin(cand_min, tile_face<SCHEME, ORDER>(a, L, e, fd, fg));
    }
    __syncthreads();

    // 4. gather in Cell::faces order, owner minus, and the update (as k_e_gather)
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    if (LEAN && tid < n) {
#pragma unroll
        for (int k = 0; k < D; ++k) pen[k] = te[k * MG + tid];
        upd = load_update_inputs(a, c0 + tid);
    }
    if (tid < n) {
        double dmin = DBL_MAX;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (pen[k] == -1) continue;
            const int lf = tell_lf(pen[k]);
            const double* F = L.F + lf;
            if (tell_owner(pen[k])) {
                r0 = r0 - F[0]; r1 = r1 - F[MF]; r2 = r2 - F[2 * MF]; r3 = r3 - F[3 * MF];
            } else {
                r0 += F[0]; r1 += F[MF]; r2 += F[2 * MF]; r3 += F[3 * MF];
            }
            if (a.dt_src == 2) dmin = smin(dmin, L.Dt[lf]);
        }
        const double4 u{L.U[tid], L.U[MG + tid], L.U[2 * MG + tid], L.U[3 * MG + tid]};
        update_cell_with(a, c0 + tid, upd, u, r0, r1, r2, r3, dmin);
    }

    // 5. block reductions: a butterfly inside each wavefront, then the
    // wavefronts' values in order (a fixed order, one barrier)
    __shared__ double red[5][kMaxBlock / 64];
    const int nw = bs >> 6, w = tid >> 6, lane = tid & 63;
    if (a.want_res) {
        const double s0 = wave_sum(r0 * r0), s1 = wave_sum(r1 * r1), s2 = wave_sum(r2 * r2), s3 = wave_sum(r3 * r3);
        if (lane == 0) {
            red[0][w] = s0;
            red[1][w] = s1;
            red[2][w] = s2;
            red[3][w] = s3;
        }
    }
    const bool want_dt = !BATCH && a.face_dt;  // batched cases keep no step dt
    if (want_dt) {
        const double dmin = wave_min(cand_min);
        if (lane == 0) red[4][w] = dmin;
    }
    __syncthreads();
    if (a.want_res && tid < 4) {
        double acc = red[tid][0];
        for (int i = 1; i < nw; ++i) acc += red[tid][i];
        if (BATCH) a.tile_partials[(int64_t)tid * m.ntiles + t] = acc;
        else a.partials[tid * gridDim.x + blockIdx.x] = acc;
    }
    // the step's dt (local dt: only its history) is reduced by the finalize
    // from one entry per workgroup, not by device-wide atomics
    if (want_dt && tid == 4) {
        double acc = red[4][0];
        for (int i = 1; i < nw; ++i) acc = smin(acc, red[4][i]);
        a.partials[4 * gridDim.x + blockIdx.x] = acc;
    }
}

// batched cases: one launch per stage over the list of the active cases'
// tiles, a workgroup per list entry; the grid covers every tile of the
// batch, so the same launch (captured once in a graph) serves any set of
// active cases, and the workgroups past the list's end return at once.
// (A grid-stride loop over the list keeps the kernel arguments live across
// its iterations and spills at this register budget.)
template <int D, int SCHEME, int ORDER, int WPE>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void k_e_tile_cases(EulerArgs a) {
    extern __shared__ double lds[];
    const int it = logical_block();
    if (it >= *a.n_tile_list) return;
    // the entry is the workgroup's own: keep it (and the tile's ranges) scalar
    tile_body<D, SCHEME, ORDER, WPE, true>(a, lds, __builtin_amdgcn_readfirstlane(a.tile_list[it]));
}

// The queues of the persistent launches (kernel variant 4): the items
// (tiles, or entries of a tile list) split into eight contiguous ranges, one
// per XCD, each with its claim counter on a cache line of its own, then a
// count of the workgroups done.  A workgroup takes the items of its XCD's
// range in order (blockIdx.x mod 8, the XCD the dispatcher places it on), so
// neighbouring tiles, which share halo cells, run in one L2.
// the next item of this workgroup's range, or INT_MAX once it is drained
__device__ __forceinline__ int queue_claim(unsigned int* q, int n) {
    const int x = blockIdx.x % kXcds;
    const int per = (n + kXcds - 1) / kXcds;
    const int k = (int)atomicAdd(&q[x * kQueueStride], 1u);
    const int t = x * per + k;
    return (k < per && t < n) ? t : INT_MAX;
}

// thread 0 of every workgroup, after its last claim: the last one out resets
// the queue to zero for the next launch (no memset in a captured step)
__device__ __forceinline__ void queue_release(unsigned int* q) {
    if (atomicAdd(&q[kXcds * kQueueStride], 1u) == gridDim.x - 1) {
        for (int x = 0; x < kXcds; ++x) atomicExch(&q[x * kQueueStride], 0u);
        atomicExch(&q[kXcds * kQueueStride], 0u);
    }
}

// batched cases, one persistent launch per stage (kernel variant 4): as many
// workgroups as fit on the device at once, each taking entries of the active
// cases' tile list from its XCD's queue until the range runs dry.  A
// workgroup claims its next entry when it starts a tile, so the claim's
// latency overlaps the tile; the claims alternate between two LDS words, so
// a slow reader of one never sees the next claim.  Every workgroup leaves the
// loop once its range is drained (claims only grow).  The tiles' sums go to
// their own rows, so the results do not depend on which workgroup ran a tile.
template <int D, int SCHEME, int ORDER, int WPE>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void k_e_tile_persist(EulerArgs a) {
    extern __shared__ double lds[];
    __shared__ int claim[2];
    const int n = *a.n_tile_list;
    if (threadIdx.x == 0) claim[0] = queue_claim(a.tile_queue, n);
    __syncthreads();
    int it = claim[0];
    for (int i = 0; it < n; ++i) {
        if (threadIdx.x == 0) claim[(i + 1) & 1] = queue_claim(a.tile_queue, n);
        // the thread index made opaque per tile, so that nothing derived from
        // it is kept live across the loop
        int tid = threadIdx.x;
        __asm__ volatile("" : "+v"(tid));
        tile_body<D, SCHEME, ORDER, WPE, true>(a, lds, __builtin_amdgcn_readfirstlane(a.tile_list[it]), tid);
        __syncthreads();
        it = claim[(i + 1) & 1];
    }
    if (threadIdx.x == 0) queue_release(a.tile_queue);
}

template <int D, int SCHEME, int ORDER, int WPE>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void k_e_tile(EulerArgs a) {
    extern __shared__ double lds[];
    int t = logical_block();
    // a pending finalize (the previous stage's norms) takes the first logical
    // workgroup, dispatched first so that its one-workgroup reduction overlaps
    // the tiles; the tiles follow it
    if (a.fin_nblocks > 0) {
        if (t == 0) {
            finalize_norms(a.partials, a.fin_nblocks, 4, a.res_hist, a.dt_hist, a.fin_hist_idx, a.dtmin,
                           a.fin_slot, -1.0, lds, a.fin_rearm != 0, a.fin_dt_row != 0);
            return;
        }
        --t;
    }
    tile_body<D, SCHEME, ORDER, WPE, false>(a, lds, t);
}

// batched cases on tiles: one block per case, the case's residual norms from
// the per-tile sums of its tile range in a fixed tree order (a frozen case
// gets zero rows), into row (*step mod cap) of the per-case history
__global__ __launch_bounds__(kCaseFinBlock) void k_case_finalize_tiles(const double* tile_partials, int ntiles,
                                                                       const int* case_tiles, const int* active,
                                                                       double* case_res, const unsigned int* step,
                                                                       int cap, int n_cases) {
    __shared__ double sm[kCaseFinBlock];
    const int k = blockIdx.x;
    const int t0 = case_tiles[k], t1 = case_tiles[k + 1];
    const bool on = active[k] != 0;
    const int64_t row = (int64_t)(*step % (unsigned int)cap);
    for (int q = 0; q < 4; ++q) {
        double acc = 0.0;
        if (on)
            for (int t = t0 + (int)threadIdx.x; t < t1; t += kCaseFinBlock) acc += tile_partials[(int64_t)q * ntiles + t];
        sm[threadIdx.x] = acc;
        __syncthreads();
        for (int s = kCaseFinBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) case_res[(row * n_cases + k) * 4 + q] = sqrt(sm[0]);
        __syncthreads();
    }
}

constexpr int kCuLds = 160 * 1024;  // LDS per CU

// Raises a kernel's dynamic LDS limit to `lds` when it exceeds the default
// 64 KB (a gfx950 workgroup may use the CU's whole 160 KB), once per kernel.
void allow_lds(const void* kernel, size_t lds) {
    if (lds <= 65536) return;
    static std::mutex mu;
    static std::unordered_map<const void*, size_t> raised;
    std::lock_guard<std::mutex> lock(mu);
    size_t& cur = raised[kernel];
    if (cur >= lds) return;
    if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess) cur = lds;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
namespace {

constexpr int kDtBlock = 256;  // global-dt pass beside the tile kernels
// the global-dt pass's grid: a face per thread up to four workgroups per CU
int dt_grid(int64_t nf, int bs) { return std::min(grid_for(nf, bs), 4 * 256); }


EulerArgs make_args(fvk_solver* s) {
    EulerArgs a{};
    a.rearm_slot = -1;
    a.m = s->mesh;
    a.G = reinterpret_cast<double4*>(s->grad);
    a.flux = reinterpret_cast<double4*>(s->flux);
    a.dtf = s->dtf;
    a.dtc = s->dtc;
    a.dtmin = s->dtmin;
    a.partials = s->partials;
    a.res_hist = s->res_hist;
    a.dt_hist = s->dt_hist;
    a.gamma = s->p.gamma;
    a.cfl = s->p.cfl;
    a.tol = s->p.exact_riemann_tol;
    a.maxit = s->p.exact_riemann_max_iter;
    a.limiter = s->p.limiter;
    a.stages = s->p.rk_stages;
    a.block = s->block;
    a.tile_queue = s->tile_queue;
    if (s->n_cases > 0) {
        a.case_start = s->d_case_start;
        a.active = s->d_active;
        a.case_lo = s->d_case_lo;
        a.case_partials = s->case_partials;
        a.n_cases = s->n_cases;
    }
    return a;
}

template <int D>
struct Launch {
    static void grad(fvk_solver* s, const EulerArgs& a) {
        k_e_grad<D><<<grid_for(s->nc, s->block), s->block, 0, s->stream>>>(a);
    }
    template <int ORDER>
    static void cell_o(fvk_solver* s, const EulerArgs& a) {
        int g = grid_for(s->nc, s->block);
        switch (s->p.flux_scheme) {
            case FLUX_HLLC: k_e_cell<D, FLUX_HLLC, ORDER><<<g, s->block, 0, s->stream>>>(a); break;
            case FLUX_EXACT: k_e_cell<D, FLUX_EXACT, ORDER><<<g, s->block, 0, s->stream>>>(a); break;
            default: k_e_cell<D, FLUX_RUSANOV, ORDER><<<g, s->block, 0, s->stream>>>(a); break;
        }
    }
    static void cell(fvk_solver* s, const EulerArgs& a) {
        if (s->p.order == 2) cell_o<2>(s, a);
        else cell_o<1>(s, a);
    }
    static void gather(fvk_solver* s, const EulerArgs& a) {
        k_e_gather<D><<<grid_for(s->nc, s->block), s->block, 0, s->stream>>>(a);
    }
    template <int ORDER>
    static void tile_o(fvk_solver* s, const EulerArgs& a, int grid) {
        const size_t lds = s->tile_lds;
        // four wavefronts per SIMD when the CU's LDS holds enough workgroups for them
        const bool four = D == 3 && ORDER == 2 && s->p.flux_scheme != FLUX_EXACT &&
                          (int64_t)(kCuLds / std::max<size_t>(lds, 1)) * s->block >= 4 * 4 * 64;
        // batched cases take the tile-list instance, or its persistent form
        // (variant 4) when the solver has a tile queue
        const bool persist = a.tile_list && s->tile_queue;
        auto go = [&](auto single, auto cases, auto persistent) {
            const void* k = persist ? reinterpret_cast<const void*>(persistent)
                                    : a.tile_list ? reinterpret_cast<const void*>(cases)
                                                  : reinterpret_cast<const void*>(single);
            allow_lds(k, lds);
            if (persist && s->persist_query) {
                // the workgroups of this kernel that fit on the device at once:
                // per CU, as many as its LDS (160 KB) holds and as many as its
                // four SIMDs' registers hold (512 VGPRs per lane each, in
                // granules of 8), times the CUs
                hipFuncAttributes fa{};
                int dev = 0, cus = 0;
                if (hipFuncGetAttributes(&fa, k) != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
                    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
                    s->persist_grid = 0;
                    return;
                }
                const int vgpr = std::max(8, (fa.numRegs + 7) / 8 * 8);
                const int waves_simd = std::min(8, 512 / vgpr);
                const int by_regs = 4 * waves_simd / std::max(1, s->block / 64);
                const int by_lds = (int)(kCuLds / std::max<size_t>(1, lds + fa.sharedSizeBytes));
                s->persist_grid = (std::max(1, std::min(by_regs, by_lds)) * cus + kXcds - 1) / kXcds * kXcds;
                return;
            }
            if (persist) persistent<<<s->persist_grid, s->block, lds, s->stream>>>(a);
            else if (a.tile_list) cases<<<grid, s->block, lds, s->stream>>>(a);
            else single<<<grid, s->block, lds, s->stream>>>(a);
        };
        if constexpr (D == 3) {
            if (four) {
                if (s->p.flux_scheme == FLUX_HLLC)
                    go(k_e_tile<3, FLUX_HLLC, 2, 4>, k_e_tile_cases<3, FLUX_HLLC, 2, 4>,
                       k_e_tile_persist<3, FLUX_HLLC, 2, 4>);
                else
                    go(k_e_tile<3, FLUX_RUSANOV, 2, 4>, k_e_tile_cases<3, FLUX_RUSANOV, 2, 4>,
                       k_e_tile_persist<3, FLUX_RUSANOV, 2, 4>);
                return;
            }
        }
        switch (s->p.flux_scheme) {
            case FLUX_HLLC:
                go(k_e_tile<D, FLUX_HLLC, ORDER, 1>, k_e_tile_cases<D, FLUX_HLLC, ORDER, 1>,
                   k_e_tile_persist<D, FLUX_HLLC, ORDER, 1>);
                break;
            case FLUX_EXACT:
                go(k_e_tile<D, FLUX_EXACT, ORDER, 1>, k_e_tile_cases<D, FLUX_EXACT, ORDER, 1>,
                   k_e_tile_persist<D, FLUX_EXACT, ORDER, 1>);
                break;
            default:
                go(k_e_tile<D, FLUX_RUSANOV, ORDER, 1>, k_e_tile_cases<D, FLUX_RUSANOV, ORDER, 1>,
                   k_e_tile_persist<D, FLUX_RUSANOV, ORDER, 1>);
                break;
        }
    }
    static void tile(fvk_solver* s, const EulerArgs& a, int grid) {
        if (s->p.order == 2) tile_o<2>(s, a, grid);
        else tile_o<1>(s, a, grid);
    }
};

template <int ORDER>
void face_o(fvk_solver* s, const EulerArgs& a) {
    // a pending finalize takes the first logical workgroup (one more XCD round)
    int g = grid_for(s->nf, s->block) + (a.fin_nblocks > 0 ? kXcds : 0);
    switch (s->p.flux_scheme) {
        case FLUX_HLLC: k_e_face<FLUX_HLLC, ORDER><<<g, s->block, 0, s->stream>>>(a); break;
        case FLUX_EXACT: k_e_face<FLUX_EXACT, ORDER><<<g, s->block, 0, s->stream>>>(a); break;
        default: k_e_face<FLUX_RUSANOV, ORDER><<<g, s->block, 0, s->stream>>>(a); break;
    }
}

template <int D>
void run_stage_d(fvk_solver* s, EulerArgs& a, bool dt_pass) {
    hipEvent_t ev = nullptr;
    if (s->variant == 3 && s->n_cases > 0) {
        // batched cases: one launch over the active cases' tiles, then the
        // per-case norms
        a.tile_list = s->d_tile_list;
        a.n_tile_list = s->d_n_tile_list;
        a.tile_partials = s->tile_partials;
        a.tile_queue = s->tile_queue;
        a.fin_nblocks = 0;
        timer_begin(s, "euler_tile", &ev);
        Launch<D>::tile(s, a, (s->mesh.ntiles + kXcds - 1) / kXcds * kXcds);
        timer_end(s, ev);
        if (a.want_res) {
            timer_begin(s, "euler_finalize", &ev);
            k_case_finalize_tiles<<<s->n_cases, kCaseFinBlock, 0, s->stream>>>(
                s->tile_partials, s->mesh.ntiles, s->d_case_tiles, s->d_active, s->case_res, s->d_step, s->hist_cap,
                s->n_cases);
            k_step_tick<<<1, 1, 0, s->stream>>>(s->d_step);
            timer_end(s, ev);
        }
        return;
    }
    if (s->variant == 3) {
        if (dt_pass) {  // its block minimum needs a power-of-two workgroup
            timer_begin(s, "euler_dt", &ev);
            k_e_dt<<<dt_grid(s->nf, kDtBlock), kDtBlock, 0, s->stream>>>(a);
            timer_end(s, ev);
        }
        // a pending finalize takes a spare workgroup (logical block 0)
        const int grid = (s->mesh.ntiles + (a.fin_nblocks > 0 ? 1 : 0) + kXcds - 1) / kXcds * kXcds;
        timer_begin(s, "euler_tile", &ev);
        Launch<D>::tile(s, a, grid);
        timer_end(s, ev);
        if (a.want_res) {
            if (a.stages >= 2 && !a.R) {
                // the next stage's tile kernel reduces this stage's partials
                s->fin_nblocks = grid;
                s->fin_slot = a.slot;
                s->fin_hist_idx = a.hist_idx;
                s->fin_rearm = 1;
                s->fin_dt_row = a.face_dt;
            } else {
                timer_begin(s, "euler_finalize", &ev);
                k_e_finalize<<<1, kFinBlock, 0, s->stream>>>(a.partials, grid, 4, a.res_hist, a.dt_hist,
                                                             a.hist_idx, a.dtmin, a.slot, -1.0, 1, a.face_dt);
                timer_end(s, ev);
            }
        }
        return;
    }
    const bool split = s->variant == 1;
    if (dt_pass && !split) {
        timer_begin(s, "euler_dt", &ev);
        k_e_dt<<<dt_grid(s->nf, s->block), s->block, 0, s->stream>>>(a);
        timer_end(s, ev);
    }
    if (s->p.order == 2) {
        timer_begin(s, "euler_grad", &ev);
        Launch<D>::grad(s, a);
        timer_end(s, ev);
    }
    if (split) {
        EulerArgs af = a;
        af.face_dt = dt_pass || a.face_dt;
        timer_begin(s, "euler_face", &ev);
        if (s->p.order == 2) face_o<2>(s, af);
        else face_o<1>(s, af);
        timer_end(s, ev);
        EulerArgs ag = a;
        ag.face_dt = 0;
        ag.fin_nblocks = 0;
        // the norms of this stage go to the next face launch (next stage, or
        // for forward Euler the next step, whose dt slot the gather re-arms)
        const bool fold = a.want_res && s->n_cases == 0 && !a.R;
        if (fold && a.stages == 1) ag.rearm_slot = a.slot ^ 1;
        timer_begin(s, "euler_gather", &ev);
        Launch<D>::gather(s, ag);
        timer_end(s, ev);
        if (fold) {
            s->fin_nblocks = grid_for(s->nc, s->block);
            s->fin_slot = a.slot;
            s->fin_hist_idx = a.hist_idx;
            s->fin_rearm = a.stages >= 2;
            s->fin_dt_row = 0;
            return;
        }
    } else {
        timer_begin(s, "euler_cell", &ev);
        Launch<D>::cell(s, a);
        timer_end(s, ev);
    }
    if (a.want_res) {
        timer_begin(s, "euler_finalize", &ev);
        if (s->n_cases > 0) launch_case_finalize(s);
        else
            k_e_finalize<<<1, kFinBlock, 0, s->stream>>>(a.partials, grid_for(s->nc, s->block), 4, a.res_hist,
                                                         a.dt_hist, a.hist_idx, a.dtmin, a.slot);
        timer_end(s, ev);
    }
}

// sizes the persistent batched launch (variant 4) through the same dispatch
// the stages take, without launching anything
template <int D>
void size_persistent(fvk_solver* s) {
    EulerArgs a = make_args(s);
    a.tile_list = s->d_tile_list;
    a.n_tile_list = s->d_n_tile_list;
    a.tile_queue = s->tile_queue;
    s->persist_query = true;
    Launch<D>::tile(s, a, 0);
    s->persist_query = false;
}

void run_stage(fvk_solver* s, EulerArgs& a, bool dt_pass) {
    const int deg = s->mesh.deg;
    if (deg <= 3) run_stage_d<3>(s, a, dt_pass);
    else if (deg <= 4) run_stage_d<4>(s, a, dt_pass);
    else if (deg <= 9) run_stage_d<9>(s, a, dt_pass);
    else run_stage_d<kMaxEulerDegree>(s, a, dt_pass);
}

}  // namespace

// stage in [0, stages).  The per-cell kernels read their neighbours' stage
// state while they write their own update, so the output buffer is never the
// one read for neighbours: forward Euler alternates two buffers (A -> B, then
// B is current); RK2/RK3 rotate A -> B -> C -> A, where the last stage writes
// A, which it reads only as the cell's own RK base state.
int launch_euler_stage(fvk_solver* s, int stage, int stages, int /*unused*/, int want_res, int64_t step) {
    static const double RA[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0.5, 0}, {0, 0.75, 1.0 / 3.0}};
    static const double RB[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0.5, 0}, {1, 0.25, 2.0 / 3.0}};
    EulerArgs a = make_args(s);
    const int A = s->cur;
    int in, out;
    if (stages == 1) {
        in = A;
        out = A ^ 1;
    } else {
        in = stage == 0 ? A : (A + stage) % 3;
        out = (stage == stages - 1) ? A : (A + stage + 1) % 3;
    }
    a.U = reinterpret_cast<const double4*>(s->U[in]);
    a.Unew = reinterpret_cast<double4*>(s->U[out]);
    a.U0 = reinterpret_cast<const double4*>(s->U[A]);
    a.first = stage == 0;
    a.a = RA[stages][stage];
    a.b = RB[stages][stage];
    a.want_res = want_res && stage == 0;
    a.slot = (int)(step & 1);
    a.hist_idx = step % s->hist_cap;
    if (s->fin_nblocks > 0) {  // the previous stage's (or step's) norms, folded into this launch
        a.fin_nblocks = s->fin_nblocks;
        a.fin_slot = s->fin_slot;
        a.fin_hist_idx = s->fin_hist_idx;
        a.fin_rearm = s->fin_rearm;
        a.fin_dt_row = s->fin_dt_row;
        s->fin_nblocks = 0;
    }
    bool dt_pass = false;
    if (s->p.local_dt) {
        a.face_dt = stage == 0;
        a.dt_src = stage == 0 ? 2 : 0;
    } else {
        a.face_dt = 0;
        a.dt_src = 1;
        dt_pass = stage == 0;
    }
    run_stage(s, a, dt_pass);
    if (stages == 1) s->cur = out;
    return hip_check(hipGetLastError(), "euler stage launch");
}

int launch_euler_residual_eval(fvk_solver* s, double* R_dev) {
    EulerArgs a = make_args(s);
    a.U = reinterpret_cast<const double4*>(s->U[s->cur]);
    a.face_dt = 0;
    a.dt_src = 0;
    a.want_res = 0;
    a.R = reinterpret_cast<double4*>(R_dev);
    run_stage(s, a, false);
    return hip_check(hipGetLastError(), "euler residual launch");
}

// residual norms of nv components from nblocks block partials, the step's dt
// into the history (fixed_dt >= 0: that value) and the next dt slot re-armed
int launch_norm_finalize(fvk_solver* s, int nblocks, int nv, int64_t hist_idx, int slot, double fixed_dt,
                         bool dt_row) {
    k_e_finalize<<<1, kFinBlock, 0, s->stream>>>(s->partials, nblocks, nv, s->res_hist, s->dt_hist, hist_idx,
                                                 s->dtmin, slot, fixed_dt, 1, dt_row ? 1 : 0);
    return hip_check(hipGetLastError(), "finalize launch");
}

// a finalize still pending after the last stage launch of a call (Euler or viscous)
int flush_pending_finalize(fvk_solver* s) {
    if (s->fin_nblocks <= 0) return FVK_OK;
    k_e_finalize<<<1, kFinBlock, 0, s->stream>>>(s->partials, s->fin_nblocks, s->nv, s->res_hist, s->dt_hist,
                                                 s->fin_hist_idx, s->dtmin, s->fin_slot, -1.0, s->fin_rearm,
                                                 s->fin_dt_row);
    s->fin_nblocks = 0;
    return hip_check(hipGetLastError(), "finalize launch");
}

// per-case residual norms of the step and the step counter
int launch_case_finalize(fvk_solver* s) {
    k_case_finalize<<<s->n_cases, kCaseFinBlock, 0, s->stream>>>(s->case_partials, grid_for(s->nc, s->block),
                                                                s->d_case_blocks, s->d_case_lo, s->case_res,
                                                                s->d_step, s->hist_cap, s->n_cases);
    k_step_tick<<<1, 1, 0, s->stream>>>(s->d_step);
    return hip_check(hipGetLastError(), "case finalize launch");
}

// limited gradients of the current state into the gradient buffer (the
// first pass of the residual, shared with the adjoint)
int launch_euler_gradients(fvk_solver* s) {
    EulerArgs a = make_args(s);
    a.U = reinterpret_cast<const double4*>(s->U[s->cur]);
    hipEvent_t ev = nullptr;
    timer_begin(s, "euler_grad", &ev);
    const int deg = s->mesh.deg;
    if (deg <= 3) Launch<3>::grad(s, a);
    else if (deg <= 4) Launch<4>::grad(s, a);
    else if (deg <= 9) Launch<9>::grad(s, a);
    else Launch<kMaxEulerDegree>::grad(s, a);
    timer_end(s, ev);
    return hip_check(hipGetLastError(), "euler gradient launch");
}

// plain streaming copy (bandwidth calibration of the HBM byte counters)
constexpr int kCopyBlock = 256;
__global__ __launch_bounds__(kCopyBlock) void k_copy(double2* __restrict__ dst, const double2* __restrict__ src,
                                                     int64_t n2) {
    int64_t i = (int64_t)blockIdx.x * kCopyBlock + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * kCopyBlock;
    for (; i < n2; i += stride) dst[i] = src[i];
}

// fdiv (the flux kernels' division) beside the compiler's x / y on the
// caller's operand pairs, so both can be compared bit for bit.
__global__ void k_div_probe(const double* x, const double* y, double* q_fast, double* q_ieee, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        q_fast[i] = fdiv(x[i], y[i]);
        q_ieee[i] = x[i] / y[i];
    }
}

int set_persistent_grid(fvk_solver* s) {
    if (!s->tile_queue) return set_error(FVK_ERR_INVALID, "persistent launch without a tile queue");
    const int deg = s->mesh.deg;
    s->persist_grid = 0;
    if (deg <= 3) size_persistent<3>(s);
    else if (deg <= 4) size_persistent<4>(s);
    else if (deg <= 9) size_persistent<9>(s);
    else size_persistent<kMaxEulerDegree>(s);
    if (s->persist_grid <= 0) return set_error(FVK_ERR_INVALID, "persistent launch: no grid");
    return FVK_OK;
}

int launch_div_probe(const double* x, const double* y, double* q_fast, double* q_ieee, int64_t n, hipStream_t st) {
    const int64_t g = std::min<int64_t>((n + 255) / 256, 4096);
    k_div_probe<<<(int)std::max<int64_t>(g, 1), 256, 0, st>>>(x, y, q_fast, q_ieee, n);
    return hip_check(hipGetLastError(), "divide probe launch");
}

int launch_copy_probe(double* dst, const double* src, int64_t n, hipStream_t st) {
    int64_t n2 = n / 2;
    int grid = 256 * 16;
    k_copy<<<grid, kCopyBlock, 0, st>>>(reinterpret_cast<double2*>(dst), reinterpret_cast<const double2*>(src), n2);
    return hip_check(hipGetLastError(), "copy probe launch");
}

}  // namespace fvk
