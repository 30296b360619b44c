// Residual-norm finalize shared by the Euler and viscous kernels: one
// workgroup reduces the per-block partial sums of the squared residuals in a
// fixed order, writes the step's norms and dt into the history, and re-arms
// the next step's dt slot.  It runs either as its own one-block kernel or as
// a spare workgroup of the next launch on the stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace fvk {

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

// The global dt of a step lives in one of two slots (step parity).  Each slot
// is spread over kDtSub sub-slots a cache line apart, one per XCD: a
// workgroup takes the atomic minimum into its XCD's sub-slot, and the readers
// take the minimum of the eight (the same value as one slot would hold; the
// device-wide atomics on one address are serialised across the XCDs).
constexpr int kDtSub = 8, kDtStride = 16;  // sub-slots per slot; 64-bit words per sub-slot (128 B)
constexpr int kDtWords = 2 * kDtSub * kDtStride;
__device__ __forceinline__ unsigned long long* dt_sub(unsigned long long* dtmin, int slot, int sub) {
    return dtmin + (slot * kDtSub + sub) * kDtStride;
}
__device__ __forceinline__ double dt_read(const unsigned long long* dtmin, int slot) {
    unsigned long long v[kDtSub];
#pragma unroll
    for (int i = 0; i < kDtSub; ++i) v[i] = dtmin[(slot * kDtSub + i) * kDtStride];
    unsigned long long m = v[0];
#pragma unroll
    for (int i = 1; i < kDtSub; ++i) m = v[i] < m ? v[i] : m;  // positive doubles order as their bits
    return __longlong_as_double((long long)m);
}
__device__ __forceinline__ void dt_arm(unsigned long long* dtmin, int slot) {
#pragma unroll
    for (int i = 0; i < kDtSub; ++i) *dt_sub(dtmin, slot, i) = dbits(DBL_MAX);
}
// a workgroup's minimum into its XCD's sub-slot of the slot
__device__ __forceinline__ void dt_atomic_min(unsigned long long* dtmin, int slot, double v) {
    atomicMin(dt_sub(dtmin, slot, blockIdx.x % kDtSub), dbits(v));
}

// one block: residual norms = sqrt of the sums of the partials (fixed order
// for a given workgroup size), the step's dt into the history, and the next
// step's dt slot re-armed; sm holds blockDim.x doubles.  dt_row: the step's
// dt is the minimum of the partials' row nv (one entry per block) instead of
// the atomic minimum in the dt slot.
__device__ __forceinline__ void finalize_norms(const double* partials, int nblocks, int nv, double* res_hist,
                                               double* dt_hist, int64_t hist_idx, unsigned long long* dtmin,
                                               int slot, double fixed_dt, double* sm, bool rearm = true,
                                               bool dt_row = false) {
    const int nt = blockDim.x, tid = threadIdx.x;
    int p2 = 64;  // the largest power of two not above the workgroup size
    while (2 * p2 <= nt) p2 *= 2;
    for (int q = 0; q < nv; ++q) {
        // four independent loads in flight per thread, summed in a fixed order
        const double* p = partials + (int64_t)q * nblocks;
        double acc = 0.0;
        for (int i = tid; i < nblocks; i += 4 * nt) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = i + u * nt < nblocks ? p[i + u * nt] : 0.0;
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += v[u];
        }
        sm[tid] = acc;
        __syncthreads();
        if (tid < nt - p2) sm[tid] += sm[tid + p2];
        __syncthreads();
        for (int s = p2 / 2; s > 0; s >>= 1) {
            if (tid < s) sm[tid] += sm[tid + s];
            __syncthreads();
        }
        if (tid == 0) res_hist[hist_idx * nv + q] = sqrt(sm[0]);
        __syncthreads();
    }
    double dmin = DBL_MAX;
    if (dt_row && fixed_dt < 0.0) {
        const double* p = partials + (int64_t)nv * nblocks;
        for (int i = tid; i < nblocks; i += nt) dmin = p[i] < dmin ? p[i] : dmin;
        sm[tid] = dmin;
        __syncthreads();
        if (tid < nt - p2) sm[tid] = sm[tid + p2] < sm[tid] ? sm[tid + p2] : sm[tid];
        __syncthreads();
        for (int s = p2 / 2; s > 0; s >>= 1) {
            if (tid < s) sm[tid] = sm[tid + s] < sm[tid] ? sm[tid + s] : sm[tid];
            __syncthreads();
        }
        dmin = sm[0];
    }
    if (tid == 0) {
        if (fixed_dt >= 0.0) {
            dt_hist[hist_idx] = fixed_dt;
        } else {
            dt_hist[hist_idx] = dt_row ? dmin : dt_read(dtmin, slot);
            if (rearm) dt_arm(dtmin, slot ^ 1);
        }
    }
}

}  // namespace fvk
