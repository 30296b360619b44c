// fvk_wall_force_reports_noalloc: the wall diagnostics of
// fvk_wall_force_reports for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).  Inputs and outputs
// are the caller's device buffers; the kernels' scratch (per-sample force
// contributions and patch ids) is a file-scope device array of fixed size,
// whose address on each device is looked up once, when the first solver is
// created on that device.
#include <mutex>

#include "solver.h"

namespace fvk {
namespace {

constexpr int kNoallocMaxFaces = FVK_NOALLOC_MAX_WALL_FACES;

// [5][kNoallocMaxFaces] contributions, then [kNoallocMaxFaces] patch ids (as int)
__device__ __align__(256) double scratch[6 * kNoallocMaxFaces];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Looks up this device's copy of `scratch` once.  The host-side name of the
// array is not a device address; kernels and copies only ever get the
// pointer kept here.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return set_error(FVK_ERR_UNSUPPORTED, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return FVK_OK;
    void* ptr = nullptr;
    int rc = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc != FVK_OK) return rc;
    g_scratch_ptr[device] = static_cast<double*>(ptr);
    return FVK_OK;
}

}  // namespace fvk

extern "C" int fvk_wall_force_reports_noalloc(fvk_solver* s, const int32_t* wall_faces, int64_t n_wall,
                                              const fvk_wall_reference* ref, fvk_wall_sample* samples,
                                              fvk_wall_force_report* reports, int32_t cap, int32_t* n_reports) {
    using namespace fvk;
    if (!s || !ref || !reports || !n_reports || (n_wall > 0 && !wall_faces))
        return set_error(FVK_ERR_INVALID, "NULL argument");
    if (s->p.equation == FVK_EQ_DIFFUSION || s->p.equation == FVK_EQ_ADVECTION_DIFFUSION)
        return set_error(FVK_ERR_UNSUPPORTED, "no wall forces for a scalar field");
    if (n_wall < 0 || n_wall > kNoallocMaxFaces)
        return set_error(FVK_ERR_INVALID, "n_wall %lld outside 0..%d, the scratch capacity of the no-alloc entry point",
                         (long long)n_wall, kNoallocMaxFaces);
    if (cap < s->n_bcs + 1)
        return set_error(FVK_ERR_INVALID, "cap %d is below n_bcs + 1 = %d", cap, s->n_bcs + 1);
    double* sp = (s->device >= 0 && s->device < kMaxDevices) ? g_scratch_ptr[s->device] : nullptr;
    if (!sp) return set_error(FVK_ERR_HIP, "no scratch address for device %d", s->device);
    WallLaunch w{};
    w.wall_faces = wall_faces;
    w.face_old2new = s->d_face_old2new;
    w.n = (int)n_wall;
    w.n_patches = s->n_bcs;
    w.pressure_only = 0;
    wall_launch_ref(w, *ref);
    w.samples = samples;
    w.contrib = sp;
    w.patch = reinterpret_cast<int*>(sp + 5 * (size_t)kNoallocMaxFaces);
    w.reports = reports;
    w.cap = cap;
    w.n_reports = n_reports;
    const int rc = launch_wall_diagnostics(s, w);
    // still running: every later call on the solver that touches the device waits for it first
    s->inflight = true;
    return rc;
}
