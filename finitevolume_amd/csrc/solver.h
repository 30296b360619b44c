// Internal definitions of the fvk library: the device mesh layout, the
// solver object behind the opaque fvk_solver handle, and the kernel
// launchers.  Host code only includes this; kernels see the View structs.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <string>
#include <vector>

#include "fvk.h"

namespace fvk {

// A named range on the ROCTx marker stream for the lifetime of the object
// (rocprofv3 --marker-trace shows the library's entry points around their
// kernels; without a profiler attached the calls return at once).
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

constexpr int kMaxPatches = 1024;
constexpr int kMaxEulerDegree = 16;  // most faces of one cell the per-cell Euler kernels take
// dynamic LDS bytes one tile may use: a gfx950 workgroup may take the CU's
// 160 KB, less 1 KB for the tile kernels' static arrays (block reductions)
constexpr int kMaxTileLds = 163840 - 1024;
constexpr int kXcds = 8;             // accelerator complex dies of an MI355X, one L2 each

// Workgroups are dispatched round-robin over the XCDs; the remap below gives
// XCD x the contiguous range [x * per, (x + 1) * per) of logical blocks, so
// that neighbouring cells and faces share an L2.  The grid must be a
// multiple of kXcds (grid_for).
__device__ __forceinline__ int logical_block() {
    const int per = gridDim.x / kXcds;
    return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}
__device__ __forceinline__ int xcd_index() { return logical_block() * blockDim.x + threadIdx.x; }
inline int grid_for(int64_t n, int bs) {
    const int64_t g = (n + bs - 1) / bs;
    return (int)((g + kXcds - 1) / kXcds * kXcds);
}

// Per-patch boundary data on the device (fvk_bc reduced to what kernels read).
struct DevBC {
    int type;          // BC_WALL / BC_FARFIELD / BC_OUTFLOW (diffusion: 0 Dirichlet, 1 Neumann)
    int isothermal;
    double wall_u, wall_v, wall_T;
    double ff[4];      // farfield conserved state
    double ff_nut, ff_k, ff_w;
    double value;
};

// Read-only mesh arrays in device order (cells RCM-renumbered; faces
// interior-first, sorted by their lower cell; boundary faces grouped by
// boundary type).  Face orientation (cell_left, normal) is the caller's.
struct MeshView {
    int nc, nf, nfi;        // cells, faces, interior faces (faces [0, nfi) are interior)
    int nnz;                // CSR entries
    const double* vol;      // [nc]
    const double* cx;       // [nc] centroid x
    const double* cy;       // [nc] centroid y
    const double* lsq;      // [3][nc] LSQ inverse {m11, -m01, m00}/det (GradientCalculator ctor)
    const int* cptr;        // [nc+1]
    const int* cent;        // [nnz] (device face << 1) | (cell is the face's owner)
    const int* fl;          // [nf] cell_left
    const int* fr;          // [nf] cell_right, or -1 - patch on boundary faces
    const double* fnx;      // [nf]
    const double* fny;      // [nf]
    const double* fa;       // [nf] face length
    const double* fmx;      // [nf] face midpoint
    const double* fmy;      // [nf]
    const double* fsx;      // [nf] periodic shift of cell_right (nullptr: none)
    const double* fsy;      // [nf]
    const double* fdl;      // [nf] Euler dt length: min(V_L, V_R)/A (boundary V_L/A)
    const double* fdc;      // [nf] dist * cos(theta) = face_normal_distance (Jasak)
    const double* fkx;      // [nf] n - dhat / cos(theta)
    const double* fky;      // [nf]
    const double* fwl;      // [nf] d_R / (d_L + d_R): distance weight of cell_left at the face
    const double* fwr;      // [nf] d_L / (d_L + d_R)
    const double* wdist;    // [nc] wall distance (RANS)
    const DevBC* bc;        // [n_patches]
    // ---- gather-friendly copies for the cell-centric Euler kernels ----
    int deg;                // ELL width: the largest number of faces of a cell
    int64_t ld;             // padded cell count (ELL and per-cell array stride)
    const int2* ell;        // [width][ld] (width 3, 4, 9 or 16 >= deg) per face of the cell, in Cell::faces order:
                            //   x = (device face << 1) | owner, y = neighbour cell or -1 - patch; x = -1 pads
    const double4* cgeo;    // [ld] centroid x, centroid y, volume, Venkatakrishnan eps^2
    const double4* clsq;    // [ld] LSQ inverse {m11, -m01, m00}/det, 0
    const double4* fgeo;    // [nf] nx, ny, length, Euler dt length
    const double2* fmid2;   // [nf] face midpoint
    const double2* fsh2;    // [nf] periodic shift of cell_right (nullptr: none)
    const double4* egeo;    // [width][ld] per ELL entry: w dx, w dy (LSQ weight 1/d^2 times the offset to the
                            //   neighbour, or to the face midpoint on a boundary), face midpoint - centroid
    const double4* frec;    // [nf] MUSCL offsets: midpoint - centroid of cell_left, of cell_right (periodic-shifted)
    // ---- cell tiles of the fused Euler stage (kernel variant 3) ----
    // Tile t owns the device cells [tstart[t], tstart[t+1]) (graph-grown, at
    // most one workgroup of cells).  Its local cells are the owned ones, then
    // the first ring of neighbours, then (second order) the second ring.
    int ntiles;
    int tmax_g, tmax_l, tmax_f; // most local cells with gradients (owned + ring 1), local cells, faces in a tile
    int tmax_2;                 // most ring-2 cells of a tile
    int twidth;                 // slot width of tell
    const int4* tdesc;      // [ntiles][2] the ranges below in one record: {first cell, owned cells, first
                            //   halo entry, ring-1 cells}, {halo cells, first tile face, tile faces, 0}
    int nthalo;             // entries of thalo
    const int* tstart;      // [ntiles + 1]
    const int* thptr;       // [ntiles + 1] offsets into thalo
    const int* thalo;       // ring-1 then ring-2 cells of each tile (device ids)
    const int* th1;         // [ntiles] ring-1 count
    const int* tell;        // [ntiles][width][tmax_g] per local cell with ELL entries, slot k, packed
                            //   (tell_*): neighbour, orientation and tile face of the slot
    const int* tfptr;       // [ntiles + 1] offsets into tface
    const int4* tface;      // per tile face: {device face, local cell_left, local cell_right or -1 - patch, 0}
    const int* tfc;         // per tile face, packed (tface_pack): local cell_left in the low 16 bits,
                            //   local cell_right or -1 - patch as a signed high half
    // static geometry in tile order: the gradient data of each tile's owned
    // and ring-1 cells ([ntiles][width][tmax_g] per slot, [ntiles][tmax_g]
    // per cell; tiles that compute gradients): egeo and clsq for least
    // squares, and for the viscous Green-Gauss gradient {the cell's distance
    // weight, the normal pointing away from the cell, the face length} per
    // slot and {volume, 0, 0, 0} per cell; fgeo of each tile face (as
    // tface); on a periodic mesh, the MUSCL offset (frec) of each slot's side
    // of its face (trec, as tegeo: for cell_right it differs in rounding from
    // egeo's midpoint offset)
    const double4 *tegeo, *tclsq, *tfgeo;
    const double2* trec;
    // viscous tiles: per tile face (as tface) two records, {n_x, n_y, A,
    // left distance weight} and {right distance weight, d.n, k_x, k_y}: the
    // face pass's geometry, and the Green-Gauss geometry of every slot whose
    // face is a tile face
    const double4* tvf;
};

// Packed tile ELL slot (MeshView::tell).  -1 pads.  An interior face: bit 31
// clear; the local neighbour in bits 0-14, the orientation (set: the cell
// owns the face) in bit 15, and 1 + the slot's tile face index in bits 16-30
// (0: not a tile face, as a ring-1 cell's faces away from the owned cells).
// A boundary face, which its cell owns: bit 31 set; bits 1-30 hold the tile
// face index (bit 0 clear: an owned cell) or the device face (bit 0 set: a
// ring-1 cell, whose boundary faces are never tile faces).
__host__ __device__ inline int tell_interior(int nb, bool owner, int lf) {
    return ((lf + 1) << 16) | ((int)owner << 15) | nb;
}
__host__ __device__ inline int tell_boundary_slot(bool owned_cell, int lf, int face) {
    return (int)(0x80000000u | (owned_cell ? (unsigned)lf << 1 : ((unsigned)face << 1) | 1u));
}
__host__ __device__ inline bool tell_is_boundary(int p) { return p < -1; }
__host__ __device__ inline int tell_nb(int p) { return p & 0x7fff; }
__host__ __device__ inline bool tell_owner(int p) { return p < -1 || ((p >> 15) & 1); }
__host__ __device__ inline int tell_index(int p) { return (p & 0x7fffffff) >> 1; }  // boundary: tile or device face
__host__ __device__ inline int tell_lf(int p) {  // the slot's tile face, -1 if none
    return p >= 0 ? (p >> 16) - 1 : ((p & 1) ? -1 : tell_index(p));
}
// packed tile face cells (MeshView::tfc)
__host__ __device__ inline int tface_pack(int l, int r) { return (int)(((unsigned)r << 16) | (unsigned)(l & 0xffff)); }
__host__ __device__ inline int tface_left(int p) { return p & 0xffff; }
__host__ __device__ inline int tface_right(int p) { return p >> 16; }

struct AdjointWork;  // device buffers and index maps of the residual VJP (adjoint.hip)

struct KernelTimer {
    std::string name;
    std::vector<hipEvent_t> start, stop;  // pending pairs
    std::vector<int64_t> weight;          // launches (or stages) each pending pair covers
    double total_ms = 0.0;
    int64_t count = 0;
};

}  // namespace fvk

struct fvk_solver {
    fvk_params p;
    int nv = 4;                 // variables per cell
    int n_bcs = 0;
    int variant = 1;            // Euler kernels: 1 face sweep + gather, 2 fused per cell, 3 fused per cell tile
    int block = 512;            // Euler workgroup size
    int tile_cells = 256;       // variant 3: most cells of one tile (at most block)
    std::vector<int> tile_start; // variant 3: [ntiles + 1] device cell ranges of the tiles
    size_t tile_lds = 0;         // variant 3: dynamic LDS bytes per tile
    // a stage's norm partials waiting for the next stage's (or, forward
    // Euler, the next step's) tile or face launch to reduce them with a
    // spare workgroup (fin_nblocks == 0: nothing pending; fvk_step flushes
    // what is left); fin_rearm: that finalize re-arms the next dt slot
    int fin_nblocks = 0, fin_slot = 0, fin_rearm = 1, fin_dt_row = 0;
    int64_t fin_hist_idx = 0;
    int64_t nc = 0, nf = 0, nn = 0;
    int nfi = 0, nnz = 0;
    bool periodic = false;
    hipStream_t stream = nullptr;
    int device = 0;
    // With three or more RK stages the last step's residual row is final once
    // stage 1's launches end (only stage 0 leaves norms pending), so fvk_step
    // waits on row_ready, recorded there, and returns with the later stages
    // still running (inflight); every other call that touches the device
    // waits for the stream first (settle)
    hipEvent_t row_ready = nullptr;
    mutable bool inflight = false;

    // host copies
    std::vector<int> cell_new2old, cell_old2new;
    std::vector<int> face_new2old, face_old2new;
    std::vector<fvk_bc> bcs;
    std::vector<double> node_xy;                // for adjoint geometry
    std::vector<int> face_nodes_dev;            // [2*nf] device face order
    std::vector<int> cell_ptr_old, cell_faces_old;
    std::vector<int> face_patch_old;            // patch of each reference face (-1 interior)
    int* d_face_old2new = nullptr;              // device copy of face_old2new

    // device arrays (owned)
    std::vector<void*> allocations;
    fvk::MeshView mesh{};
    double* U[3] = {nullptr, nullptr, nullptr}; // state buffers [nv][nc_pad] (RK ping-pong: A, B, C)
    int cur = 0;                                 // which buffer holds the current state
    int64_t ld = 0;                              // leading dimension of SoA state arrays (padded nc)
    int64_t ldf = 0;                             // padded nf
    double* flux = nullptr;                      // [nv][ldf] face fluxes
    double* dtf = nullptr;                       // [ldf] per-face dt candidates (local dt)
    double* dtc = nullptr;                       // [ld] per-cell dt
    double* grad = nullptr;                      // [ng][ld] cell gradients
    double* aux = nullptr;                       // [na][ld] per-cell auxiliaries (sources, nu_t, F1 ...)
    double* partials = nullptr;                  // [nv][max_blocks]
    unsigned long long* dtmin = nullptr;         // [kDtWords] global dt (bits) per step parity, per-XCD sub-slots
    double* res_hist = nullptr;                  // [hist_cap][nv] device view of res_hist_host
    double* dt_hist = nullptr;                   // [hist_cap] device view of dt_hist_host
    // the step history lives in pinned, mapped host memory: the finalize
    // kernels write it directly and the host reads it after the stream sync
    double* res_hist_host = nullptr;
    double* dt_hist_host = nullptr;
    std::vector<void*> host_allocations;         // hipHostMalloc'd, freed by fvk_destroy
    int hist_cap = 0;
    int64_t step_count = 0;
    double last_res[6] = {0, 0, 0, 0, 0, 0};
    int64_t device_bytes = 0;

    // residual VJP workspace, built on first use
    fvk::AdjointWork* adj = nullptr;

    // batched independent cases (fvk_batch_*): every case's cells are one
    // contiguous range of the device order, so each logical block of a
    // per-cell kernel holds at most two cases
    int n_cases = 0;
    std::vector<int> cell_case_old;              // [nc] case of each caller-order cell
    std::vector<int> case_cell_start;            // [n_cases + 1] device order
    int* d_case_start = nullptr;                 // device copy
    int* d_active = nullptr;                     // [n_cases] 1: stepping, 0: frozen
    int* d_case_lo = nullptr;                    // [nblocks] case of the first cell of each logical block
    int* d_case_blocks = nullptr;                // [2 n_cases] first and last logical block of each case
    double* case_res = nullptr;                  // [hist_cap][n_cases][4] per-case residual norms
    double* case_partials = nullptr;             // [4][2][nblocks] block sums split at a case boundary
    unsigned int* d_step = nullptr;              // device step counter (row of case_res)
    // batched cases on the tile kernel: the tiles of the active cases, their
    // count, each case's tile range and per-tile residual sums
    int* d_tile_list = nullptr;
    int* d_n_tile_list = nullptr;
    int* d_case_tiles = nullptr;                 // [n_cases + 1]
    std::vector<int> case_tiles;                 // host copy
    double* tile_partials = nullptr;             // [4][ntiles]

    // timing
    int timing = 0;  // 1: an event pair around every launch; 2: one around each fvk_step call's stages
    // batched cases, kernel variant 4: one persistent tile launch per stage,
    // persist_grid workgroups pulling the tile list from tile_queue
    unsigned int* tile_queue = nullptr;
    int persist_grid = 0;
    bool persist_query = false;  // set_persistent_grid: the launcher sizes the grid instead of launching
    std::vector<fvk::KernelTimer> timers;
};

namespace fvk {

// ---- layout.cpp ----
int build_layout(fvk_solver* s, const fvk_mesh_desc* d);
void rcm_order(int n, const std::vector<int>& ptr, const std::vector<int>& adj, std::vector<int>& order);

// ---- api helpers ----
int create_solver(const fvk_mesh_desc* mesh, const fvk_bc* bcs, int32_t n_bcs, const fvk_params* params,
                  const fvk_initial_condition* ic, const int32_t* cell_case, int32_t n_cases, fvk_solver** out);
int upload_state(fvk_solver* s, const double* U4, const double* turb, bool keep_turb);
int set_error(int code, const char* fmt, ...);
int hip_check(hipError_t e, const char* what);
void* dev_alloc(fvk_solver* s, size_t bytes);

// ---- kernel launchers (kernels_*.hip) ----
int launch_euler_stage(fvk_solver* s, int stage, int stages, int dt_mode_first, int want_res, int64_t step);
// the persistent tile launches' grid (kernel variant 4): the workgroups
// that fit on the device at once
int set_persistent_grid(fvk_solver* s);
// their queue: a claim counter per XCD, kQueueStride words apart (a cache
// line each), then the count of workgroups done; kQueueWords in all
constexpr int kQueueStride = 16;
constexpr int kQueueWords = (kXcds + 1) * kQueueStride;
int launch_visc_stage(fvk_solver* s, int stage, int stages, int want_res, int64_t step);
int launch_diffusion_step(fvk_solver* s, int want_res, int64_t step);
int launch_advdiff_step(fvk_solver* s, int want_res, int64_t step);
int launch_advdiff_residual_eval(fvk_solver* s, double* R_dev);
int launch_gradient(fvk_solver* s, int scheme, const double* phi, const double* bphi_dev, double* grad);
int launch_wall_distance(fvk_solver* s, const int* wall_faces, int nw, const double* node_xy_dev,
                         const int* fnodes_dev, double* out);
int launch_residual_eval(fvk_solver* s, double* R_dev);
int launch_euler_gradients(fvk_solver* s);
int launch_case_finalize(fvk_solver* s);
int flush_pending_finalize(fvk_solver* s);
// waits for work that fvk_step left running on the solver's stream
int settle(const fvk_solver* s);
int launch_norm_finalize(fvk_solver* s, int nblocks, int nv, int64_t hist_idx, int slot, double fixed_dt,
                         bool dt_row = false);
int adjoint_vjp(fvk_solver* s, const double* lambda, double* U_bar, double* X_bar);
int adjoint_solve(fvk_solver* s, const double* g_host, double cfl, int max_iter, double tol, int check_every,
                  double* lambda_out, double* X_bar, fvk_adjoint_report* rep);
void adjoint_release(fvk_solver* s);
int launch_copy_probe(double* dst, const double* src, int64_t n, hipStream_t st);
// raises a kernel's dynamic LDS limit above the default 64 KB (once per kernel)
void allow_lds(const void* kernel, size_t lds);
// wall diagnostics (kernels_visc.hip): samples and force sums of the faces
// wall_faces[n] (reference ids, device memory); contrib [5][n] and patch [n]
// are scratch; dir, lift, q and q2 as compute_wall_forces derives them from
// the reference quantities (host std::hypot)
struct WallLaunch {
    const int* wall_faces;
    const int* face_old2new;
    int n, n_patches, ng, pressure_only, aos;
    double dir_x, dir_y, lift_x, lift_y, q, q2, mx, my, rho_ref, v2, p_ref;
    fvk_wall_sample* samples;          // [n] or nullptr
    double* contrib;
    int* patch;
    fvk_wall_force_report* reports;    // [cap]
    int cap;
    int* n_reports;
};
int launch_wall_diagnostics(fvk_solver* s, const WallLaunch& w);
// per-device address of the no-alloc entry point's scratch (noalloc.hip), once per device
int noalloc_setup(int device);
void wall_launch_ref(WallLaunch& w, const fvk_wall_reference& ref);
int wall_reports(fvk_solver* s, const int32_t* wf, int64_t n, const fvk_wall_reference* ref, int pressure_only,
                 fvk_wall_sample* samples, fvk_wall_force_report* reports, int32_t cap, int32_t* n_reports);
int launch_div_probe(const double* x, const double* y, double* q_fast, double* q_ieee, int64_t n, hipStream_t st);

// timers
void timer_begin(fvk_solver* s, const char* name, hipEvent_t* ev_stop);
void timer_end(fvk_solver* s, hipEvent_t ev_stop);

}  // namespace fvk
