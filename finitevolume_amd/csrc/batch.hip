// Batched independent cases: one solver over the union of the cases' meshes.
//
// The cases are disconnected components of one mesh, laid out so that each
// case is a contiguous range of the device cell order (RCM within it), its
// cell tiles a contiguous range of the tiles.  A stage is one launch of the
// fused tile kernel for all cases over the list of the active cases' tiles,
// so small cases that would leave most of the 256 CUs idle on their own
// share the launch, and the per-launch cost is paid once per stage instead
// of once per case.
// The residual norms are reduced per case from per-tile sums in a fixed
// order.  A converged case is frozen: its tiles leave the list and its state
// stays.  With the graph on, one step is captured in a hipGraph and
// replayed; the list lives in device memory, so freezing needs no new
// capture.  Variant 1 (face sweep + gather, block sums split at case
// boundaries) remains for forward Euler.
#include <algorithm>
#include <cstring>
#include <vector>

#include "solver.h"

struct fvk_batch {
    fvk_solver* s = nullptr;
    int n_cases = 0;
    std::vector<int64_t> cell_off;  // caller order: case k's cells are [cell_off[k], cell_off[k+1])
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;            // steps per replay
    int64_t steps = 0;
};

namespace fvk {
namespace {

void drop_graph(fvk_batch* b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int launch_step(fvk_batch* b) {
    fvk_solver* s = b->s;
    const int stages = s->p.rk_stages;
    int rc = FVK_OK;
    for (int st = 0; st < stages && rc == FVK_OK; ++st) rc = launch_euler_stage(s, st, stages, 0, 1, s->step_count);
    s->step_count++;
    return rc;
}

int build_graph(fvk_batch* b) {
    fvk_solver* s = b->s;
    drop_graph(b);
    const bool timing = s->timing;
    s->timing = false;
    b->graph_steps = s->p.rk_stages == 1 ? 2 : 1;
    const int cur = s->cur;
    const int64_t count = s->step_count;
    int rc = hip_check(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "graph capture");
    for (int i = 0; i < b->graph_steps && rc == FVK_OK; ++i) rc = launch_step(b);
    hipGraph_t g = nullptr;
    int rc2 = hip_check(hipStreamEndCapture(s->stream, &g), "graph capture");
    s->timing = timing;
    // capturing launched nothing: restore the buffer rotation and step count
    s->cur = cur;
    s->step_count = count;
    if (rc == FVK_OK) rc = rc2;
    if (rc == FVK_OK) {
        b->graph = g;
        rc = hip_check(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
    } else if (g) {
        hipGraphDestroy(g);
    }
    return rc;
}

// the tiles of the active cases, in tile order, and their count: what the
// tile launch works through (a captured graph reads both from
// device memory, so freezing cases needs no new capture)
int upload_tile_list(fvk_solver* s, const int32_t* active) {
    std::vector<int> list;
    for (int k = 0; k < s->n_cases; ++k)
        if (active[k])
            for (int t = s->case_tiles[k]; t < s->case_tiles[k + 1]; ++t) list.push_back(t);
    const int n = (int)list.size();
    int rc = FVK_OK;
    if (n > 0) rc = hip_check(hipMemcpy(s->d_tile_list, list.data(), sizeof(int) * n, hipMemcpyHostToDevice), "tile list");
    if (rc == FVK_OK) rc = hip_check(hipMemcpy(s->d_n_tile_list, &n, sizeof(int), hipMemcpyHostToDevice), "tile list");
    return rc;
}

}  // namespace
}  // namespace fvk

using namespace fvk;

int fvk_batch_create(int32_t n_cases, const fvk_mesh_desc* meshes, const fvk_bc* bcs, const int32_t* n_bcs,
                     const fvk_params* params, const fvk_initial_condition* ics, fvk_batch** out) {
    if (!meshes || !bcs || !n_bcs || !params || !ics || !out || n_cases <= 0)
        return set_error(FVK_ERR_INVALID, "NULL argument or no cases");
    *out = nullptr;
    if (params->equation != FVK_EQ_EULER)
        return set_error(FVK_ERR_UNSUPPORTED, "batched cases are built for the Euler equations");
    if (!params->local_dt)
        return set_error(FVK_ERR_UNSUPPORTED, "batched cases step with local time steps (local_dt = 1)");
    if (params->kernel_variant == 2 || params->kernel_variant > 4)
        return set_error(FVK_ERR_UNSUPPORTED,
                         "batched cases use the fused tile kernel (variant 3, or 4: persistent) or the face sweep + "
                         "gather (variant 1)");
    // the fused tile kernel by default; it leaves a frozen case's state in
    // place, which needs SSP-RK stages (forward Euler alternates buffers).
    // Variant 4 is the same kernel as one persistent launch per stage: a
    // workgroup per CU slot pulls the active cases' tiles from a queue.
    const bool persistent = params->kernel_variant == 4;
    const int variant = persistent ? 3
                                   : (params->kernel_variant != 0 ? params->kernel_variant
                                                                  : (params->rk_stages >= 2 ? 3 : 1));
    if (variant == 3 && params->rk_stages < 2)
        return set_error(FVK_ERR_UNSUPPORTED, "batched cases on the tile kernel need rk_stages >= 2");
    // ---- union of the meshes ----
    int64_t nn = 0, nc = 0, nf = 0, nnz = 0, nb = 0;
    for (int k = 0; k < n_cases; ++k) {
        const fvk_mesh_desc& d = meshes[k];
        if (d.face_shift) return set_error(FVK_ERR_UNSUPPORTED, "batched cases cannot be periodic (case %d)", k);
        if (n_bcs[k] < 0) return set_error(FVK_ERR_INVALID, "case %d: n_bcs < 0", k);
        nn += d.n_nodes;
        nc += d.n_cells;
        nf += d.n_faces;
        nnz += d.cell_face_ptr[d.n_cells];
        nb += n_bcs[k];
    }
    if (nc > (1LL << 30) || nf > (1LL << 30) || nb > kMaxPatches)
        return set_error(FVK_ERR_INVALID, "batched mesh too large (%lld cells, %lld faces, %lld patches)",
                         (long long)nc, (long long)nf, (long long)nb);
    std::vector<double> node_xy(2 * nn), vol(nc), cxy(2 * nc);
    std::vector<int32_t> fnodes(2 * nf), fcells(2 * nf), fpatch(nf), cptr(nc + 1), cfaces(nnz), ccase(nc);
    fvk_batch* b = new (std::nothrow) fvk_batch;
    if (!b) return set_error(FVK_ERR_NOMEM, "host allocation failed");
    b->n_cases = n_cases;
    b->cell_off.assign(n_cases + 1, 0);
    int64_t on = 0, oc = 0, of = 0, oz = 0, ob = 0;
    cptr[0] = 0;
    for (int k = 0; k < n_cases; ++k) {
        const fvk_mesh_desc& d = meshes[k];
        std::memcpy(&node_xy[2 * on], d.node_xy, sizeof(double) * 2 * d.n_nodes);
        std::memcpy(&vol[oc], d.cell_volume, sizeof(double) * d.n_cells);
        std::memcpy(&cxy[2 * oc], d.cell_xy, sizeof(double) * 2 * d.n_cells);
        for (int64_t f = 0; f < d.n_faces; ++f) {
            fnodes[2 * (of + f)] = (int32_t)(d.face_nodes[2 * f] + on);
            fnodes[2 * (of + f) + 1] = (int32_t)(d.face_nodes[2 * f + 1] + on);
            fcells[2 * (of + f)] = (int32_t)(d.face_cells[2 * f] + oc);
            const int32_t r = d.face_cells[2 * f + 1];
            fcells[2 * (of + f) + 1] = r < 0 ? -1 : (int32_t)(r + oc);
            fpatch[of + f] = d.face_patch[f] < 0 ? -1 : (int32_t)(d.face_patch[f] + ob);
        }
        for (int64_t c = 0; c < d.n_cells; ++c) {
            cptr[oc + c + 1] = (int32_t)(d.cell_face_ptr[c + 1] + oz);
            ccase[oc + c] = k;
        }
        for (int64_t z = 0; z < d.cell_face_ptr[d.n_cells]; ++z) cfaces[oz + z] = (int32_t)(d.cell_faces[z] + of);
        b->cell_off[k + 1] = oc + d.n_cells;
        on += d.n_nodes;
        oc += d.n_cells;
        of += d.n_faces;
        oz += d.cell_face_ptr[d.n_cells];
        ob += n_bcs[k];
    }
    std::vector<fvk_bc> allbc(bcs, bcs + nb);
    fvk_mesh_desc u{};
    u.n_nodes = nn;
    u.n_cells = nc;
    u.n_faces = nf;
    u.n_patches = (int32_t)nb;
    u.node_xy = node_xy.data();
    u.face_nodes = fnodes.data();
    u.face_cells = fcells.data();
    u.face_patch = fpatch.data();
    u.cell_face_ptr = cptr.data();
    u.cell_faces = cfaces.data();
    u.cell_volume = vol.data();
    u.cell_xy = cxy.data();
    u.face_shift = nullptr;
    fvk_params p = *params;
    p.kernel_variant = variant;
    int rc = create_solver(&u, allbc.data(), (int32_t)nb, &p, &ics[0], ccase.data(), n_cases, &b->s);
    if (rc != FVK_OK) {
        delete b;
        return rc;
    }
    fvk_solver* s = b->s;
    // ---- per-case initial state ----
    {
        std::vector<double> U(4 * nc);
        const double g = p.gamma;
        auto prim = [g](double* o, double rho, double uu, double vv, double pp) {
            o[0] = rho; o[1] = rho * uu; o[2] = rho * vv; o[3] = pp / (g - 1.0) + 0.5 * rho * (uu * uu + vv * vv);
        };
        for (int k = 0; k < n_cases; ++k) {
            const fvk_initial_condition& ic = ics[k];
            for (int64_t c = b->cell_off[k]; c < b->cell_off[k + 1]; ++c) {
                if (ic.mode == FVK_IC_TWO_REGION && cxy[2 * c] >= ic.x0) prim(&U[4 * c], ic.rho_r, ic.u_r, ic.v_r, ic.p_r);
                else if (ic.mode == FVK_IC_TWO_REGION) prim(&U[4 * c], ic.rho_l, ic.u_l, ic.v_l, ic.p_l);
                else prim(&U[4 * c], ic.rho, ic.u, ic.v, ic.p);
            }
        }
        rc = upload_state(s, U.data(), nullptr, false);
    }
    // ---- case ranges, per-block case map ----
    const int bs = s->block;
    const int64_t nblocks_raw = (nc + bs - 1) / bs;
    const int64_t nblocks = (nblocks_raw + 7) / 8 * 8;  // the launch grid (a multiple of the 8 XCDs)
    for (int k = 0; k < n_cases && rc == FVK_OK && variant == 1; ++k)
        if (s->case_cell_start[k + 1] - s->case_cell_start[k] < bs)
            rc = set_error(FVK_ERR_INVALID, "case %d has %d cells, fewer than the %d of a workgroup", k,
                           s->case_cell_start[k + 1] - s->case_cell_start[k], bs);
    std::vector<int> case_lo(nblocks, n_cases - 1), case_blocks(2 * n_cases);
    for (int64_t lb = 0; lb < nblocks_raw; ++lb) {
        const int c0 = (int)(lb * bs);
        case_lo[lb] = (int)(std::upper_bound(s->case_cell_start.begin(), s->case_cell_start.end(), c0) -
                            s->case_cell_start.begin()) - 1;
    }
    for (int k = 0; k < n_cases; ++k) {
        case_blocks[2 * k] = s->case_cell_start[k] / bs;
        case_blocks[2 * k + 1] = (s->case_cell_start[k + 1] - 1) / bs;
    }
    if (rc == FVK_OK) {
        s->d_case_start = static_cast<int*>(dev_alloc(s, sizeof(int) * (n_cases + 1)));
        s->d_active = static_cast<int*>(dev_alloc(s, sizeof(int) * n_cases));
        s->d_case_lo = static_cast<int*>(dev_alloc(s, sizeof(int) * nblocks));
        s->d_case_blocks = static_cast<int*>(dev_alloc(s, sizeof(int) * 2 * n_cases));
        s->case_partials = static_cast<double*>(dev_alloc(s, sizeof(double) * 8 * nblocks));
        s->case_res = static_cast<double*>(dev_alloc(s, sizeof(double) * 4 * n_cases * s->hist_cap));
        s->d_step = static_cast<unsigned int*>(dev_alloc(s, 256));
        if (!s->d_case_start || !s->d_active || !s->d_case_lo || !s->d_case_blocks || !s->case_partials ||
            !s->case_res || !s->d_step)
            rc = set_error(FVK_ERR_NOMEM, "batch: device allocation failed");
    }
    if (rc == FVK_OK && variant == 3) {
        // each tile lies in one case (the cases are disconnected and tiles
        // grow over faces): the cases' tile ranges
        const int nt = s->mesh.ntiles;
        s->case_tiles.assign(n_cases + 1, nt);
        s->case_tiles[0] = 0;
        int t = 0;
        for (int k = 0; k < n_cases && rc == FVK_OK; ++k) {
            s->case_tiles[k] = t;
            while (t < nt && s->tile_start[t + 1] <= s->case_cell_start[k + 1]) ++t;
            if (t < nt && s->tile_start[t] < s->case_cell_start[k + 1])
                rc = set_error(FVK_ERR_INVALID, "batch: tile %d straddles cases %d and %d", t, k, k + 1);
        }
        s->case_tiles[n_cases] = t;
        if (rc == FVK_OK && t != nt) rc = set_error(FVK_ERR_INVALID, "batch: %d tiles are in no case", nt - t);
        if (rc == FVK_OK) {
            s->d_tile_list = static_cast<int*>(dev_alloc(s, sizeof(int) * std::max(nt, 1)));
            s->d_n_tile_list = static_cast<int*>(dev_alloc(s, sizeof(int)));
            s->d_case_tiles = static_cast<int*>(dev_alloc(s, sizeof(int) * (n_cases + 1)));
            s->tile_partials = static_cast<double*>(dev_alloc(s, sizeof(double) * 4 * std::max(nt, 1)));
            if (!s->d_tile_list || !s->d_n_tile_list || !s->d_case_tiles || !s->tile_partials)
                rc = set_error(FVK_ERR_NOMEM, "batch: device allocation failed");
        }
        if (rc == FVK_OK && persistent) {
            // the queue of the persistent launch: the next list entry and the
            // workgroups done; zero here, and reset to zero by the last
            // workgroup of every launch (no memset in a captured step)
            s->tile_queue = static_cast<unsigned int*>(dev_alloc(s, kQueueWords * sizeof(unsigned int)));
            if (!s->tile_queue) rc = set_error(FVK_ERR_NOMEM, "batch: device allocation failed");
            if (rc == FVK_OK) rc = hip_check(hipMemset(s->tile_queue, 0, kQueueWords * sizeof(unsigned int)), "batch queue");
            if (rc == FVK_OK) rc = set_persistent_grid(s);
        }
        if (rc == FVK_OK)
            rc = hip_check(hipMemcpy(s->d_case_tiles, s->case_tiles.data(), sizeof(int) * (n_cases + 1),
                                     hipMemcpyHostToDevice), "batch maps");
    }
    if (rc == FVK_OK) {
        std::vector<int> ones(n_cases, 1);
        rc = hip_check(hipMemcpy(s->d_case_start, s->case_cell_start.data(), sizeof(int) * (n_cases + 1),
                                 hipMemcpyHostToDevice), "batch maps");
        if (rc == FVK_OK) rc = hip_check(hipMemcpy(s->d_active, ones.data(), sizeof(int) * n_cases, hipMemcpyHostToDevice), "batch maps");
        if (rc == FVK_OK) rc = hip_check(hipMemcpy(s->d_case_lo, case_lo.data(), sizeof(int) * nblocks, hipMemcpyHostToDevice), "batch maps");
        if (rc == FVK_OK)
            rc = hip_check(hipMemcpy(s->d_case_blocks, case_blocks.data(), sizeof(int) * 2 * n_cases, hipMemcpyHostToDevice), "batch maps");
        if (rc == FVK_OK) rc = hip_check(hipMemset(s->d_step, 0, 256), "batch maps");
        if (rc == FVK_OK) rc = hip_check(hipDeviceSynchronize(), "batch set-up");
    }
    if (rc == FVK_OK && variant == 3) rc = upload_tile_list(s, std::vector<int>(n_cases, 1).data());
    if (rc != FVK_OK) {
        fvk_destroy(s);
        delete b;
        return rc;
    }
    *out = b;
    return FVK_OK;
}

void fvk_batch_destroy(fvk_batch* b) {
    if (!b) return;
    drop_graph(b);
    fvk_destroy(b->s);
    delete b;
}

int fvk_batch_set_graph(fvk_batch* b, int32_t enable) {
    if (!b) return set_error(FVK_ERR_INVALID, "NULL batch");
    b->use_graph = enable != 0;
    if (!b->use_graph) drop_graph(b);
    return FVK_OK;
}

int fvk_batch_set_active(fvk_batch* b, const int32_t* active) {
    if (!b || !active) return set_error(FVK_ERR_INVALID, "NULL argument");
    fvk_solver* s = b->s;
    int rc = hip_check(hipStreamSynchronize(s->stream), "batch");
    if (rc == FVK_OK)
        rc = hip_check(hipMemcpy(s->d_active, active, sizeof(int) * b->n_cases, hipMemcpyHostToDevice), "batch active");
    if (rc == FVK_OK && s->d_tile_list) rc = upload_tile_list(s, active);
    return rc;
}

int fvk_batch_step(fvk_batch* b, int32_t nsteps, double* case_residuals) {
    fvk::TraceRange tr("fvk_batch_step");
    if (!b) return set_error(FVK_ERR_INVALID, "NULL batch");
    if (nsteps < 0) return set_error(FVK_ERR_INVALID, "nsteps < 0");
    fvk_solver* s = b->s;
    const int nk = b->n_cases;
    int rc = FVK_OK;
    if (b->use_graph && !b->exec) rc = build_graph(b);
    if (rc != FVK_OK) return rc;
    // a graph replays whole graph_steps; a leftover step is launched directly
    std::vector<double> h((size_t)4 * nk * s->hist_cap);
    int done = 0;
    while (done < nsteps && rc == FVK_OK) {
        const int chunk = std::min(s->hist_cap, nsteps - done);
        const int64_t first = b->steps;
        int i = 0;
        while (i < chunk && rc == FVK_OK) {
            if (b->exec && chunk - i >= b->graph_steps) {
                // a captured forward-Euler pair leaves the buffers where they were
                rc = hip_check(hipGraphLaunch(b->exec, s->stream), "graph launch");
                s->step_count += b->graph_steps;
                i += b->graph_steps;
            } else {
                if (b->exec && b->graph_steps == 2) {
                    // a single direct step would break the captured buffer pairing
                    drop_graph(b);
                }
                rc = launch_step(b);
                i += 1;
            }
        }
        b->steps += chunk;
        if (rc != FVK_OK) break;
        rc = hip_check(hipMemcpyAsync(h.data(), s->case_res, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                                      s->stream), "batch residuals");
        if (rc == FVK_OK) rc = hip_check(hipStreamSynchronize(s->stream), "batch step");
        if (rc != FVK_OK) break;
        if (case_residuals)
            for (int j = 0; j < chunk; ++j) {
                const int64_t row = (first + j) % s->hist_cap;
                std::memcpy(case_residuals + (int64_t)(done + j) * nk * 4, &h[(size_t)row * nk * 4],
                            sizeof(double) * nk * 4);
            }
        done += chunk;
    }
    return rc;
}

int fvk_batch_get_field(fvk_batch* b, int32_t k, double* U) {
    if (!b || !U) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (k < 0 || k >= b->n_cases) return set_error(FVK_ERR_INVALID, "case %d out of range", k);
    std::vector<double> all(4 * (size_t)b->s->nc);
    int rc = fvk_get_field(b->s, all.data(), nullptr, b->s->nc);
    if (rc != FVK_OK) return rc;
    std::memcpy(U, &all[4 * b->cell_off[k]], sizeof(double) * 4 * (b->cell_off[k + 1] - b->cell_off[k]));
    return FVK_OK;
}

fvk_solver* fvk_batch_solver(fvk_batch* b) { return b ? b->s : nullptr; }
