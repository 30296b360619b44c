// Device re-layout of a reference-layout mesh (host side).
//
//  * cells are renumbered by reverse Cuthill-McKee over the face graph, so
//    that neighbours sit close in memory (the reordering the reference
//    suggests but does not implement, include/UnstructuredMesh.h:10-15);
//  * faces keep their orientation (cell_left, node order, normal) and are
//    reordered: interior faces first, sorted by their lower new cell index,
//    then boundary faces grouped by boundary type;
//  * each cell's face list (CSR) keeps the reference's Cell::faces order, so
//    the per-cell sum of face fluxes adds in the same order as the
//    reference's cell pass (src/EulerFVMSolver.cpp:89-96);
//  * every state-independent quantity is computed here once, with the
//    reference's formulas evaluated in the same order: face geometry
//    (include/UnstructuredMesh.h:89-105), the least-squares inverse
//    (src/GradientReconstruction.cpp:45-80), the Jasak face geometry and
//    distance weights (:22-40, :103-106, :174-176) and the Euler dt length
//    (src/EulerFVMSolver.cpp:153-155).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>
#include <tuple>

#include "solver.h"

namespace fvk {

void rcm_order(int n, const std::vector<int>& ptr, const std::vector<int>& adj, std::vector<int>& order) {
    order.clear();
    order.reserve(n);
    std::vector<int> deg(n);
    for (int i = 0; i < n; ++i) deg[i] = ptr[i + 1] - ptr[i];
    std::vector<char> seen(n, 0);
    std::vector<int> level(n, -1);
    std::vector<int> nb;
    // components in order of their lowest-degree node
    std::vector<int> by_deg(n);
    std::iota(by_deg.begin(), by_deg.end(), 0);
    std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    std::vector<int> queue;
    queue.reserve(n);
    auto bfs_levels = [&](int root, int& last_level_min_deg) {
        // BFS from root over the unvisited component; returns the farthest
        // node with the lowest degree (pseudo-peripheral node search step)
        std::vector<int> q{root};
        std::vector<int> comp;
        level[root] = 0;
        size_t h = 0;
        int maxl = 0;
        while (h < q.size()) {
            int v = q[h++];
            comp.push_back(v);
            for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
                int w = adj[k];
                if (level[w] < 0 && !seen[w]) {
                    level[w] = level[v] + 1;
                    maxl = std::max(maxl, level[w]);
                    q.push_back(w);
                }
            }
        }
        int best = root, bd = 1 << 30;
        for (int v : comp)
            if (level[v] == maxl && deg[v] < bd) { bd = deg[v]; best = v; }
        for (int v : comp) level[v] = -1;
        last_level_min_deg = maxl;
        return best;
    };
    for (int s0 : by_deg) {
        if (seen[s0]) continue;
        int root = s0, ecc = 0;
        for (int it = 0; it < 4; ++it) {
            int e2 = 0;
            int cand = bfs_levels(root, e2);
            if (it > 0 && e2 <= ecc) break;
            ecc = e2;
            root = cand;
        }
        size_t head = order.size();
        order.push_back(root);
        seen[root] = 1;
        while (head < order.size()) {
            int v = order[head++];
            nb.clear();
            for (int k = ptr[v]; k < ptr[v + 1]; ++k)
                if (!seen[adj[k]]) { nb.push_back(adj[k]); seen[adj[k]] = 1; }
            std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
            for (int w : nb) order.push_back(w);
        }
    }
    std::reverse(order.begin(), order.end());
}

namespace {

template <typename T>
T* upload(fvk_solver* s, const std::vector<T>& h) {
    if (h.empty()) return nullptr;
    T* d = static_cast<T*>(dev_alloc(s, sizeof(T) * h.size()));
    if (!d) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

}  // namespace

// Graph-grown cell tiles.  Each tile grows breadth-first from its seed,
// taking within one breadth-first depth first the cells with the most
// neighbours already tiled (so holes fill instead of fragmenting), up to
// 7/8 of the capacity T; the next seed is the untiled cell with the most
// tiled neighbours (the lowest RCM position first).  Tiles below T/2 are
// then merged into the neighbouring tile they share most faces with, where
// it has room.  The device order becomes tile by tile, each tile's cells by
// their face count (deg), then in RCM order: a wavefront then holds cells of
// one or two face counts, so the kernels' slot loops skip the padding slots
// for whole wavefronts instead of masking them lane by lane.
void tile_order(fvk_solver* s, int n, const std::vector<int>& aptr, const std::vector<int>& adj, int T,
                const int32_t* cfp) {
    const std::vector<int>& rcm = s->cell_new2old;
    std::vector<int> pos(n);
    for (int i = 0; i < n; ++i) pos[rcm[i]] = i;
    std::vector<int> tile(n, -1), ntaken(n, 0), depth(n, -1);
    // seed queue: (-tiled neighbours, RCM position)
    using Key = std::pair<int, int>;
    std::priority_queue<Key, std::vector<Key>, std::greater<Key>> seeds;
    for (int i = 0; i < n; ++i) seeds.push({0, i});
    const int target = std::max(1, T - T / 8);
    std::vector<int> size;
    std::vector<int> touched;
    struct Item { int d, k, p, c; bool operator>(const Item& o) const { return std::tie(d, k, p) > std::tie(o.d, o.k, o.p); } };
    for (;;) {
        int seed = -1;
        while (!seeds.empty()) {
            const Key k = seeds.top();
            seeds.pop();
            const int c = rcm[k.second];
            if (tile[c] < 0 && -k.first == ntaken[c]) { seed = c; break; }
        }
        if (seed < 0) break;
        const int tid = (int)size.size();
        std::priority_queue<Item, std::vector<Item>, std::greater<Item>> front;
        touched.clear();
        depth[seed] = 0;
        touched.push_back(seed);
        front.push({0, 0, pos[seed], seed});
        int cnt = 0;
        while (!front.empty() && cnt < target) {
            const Item it = front.top();
            front.pop();
            const int c = it.c;
            if (tile[c] >= 0) continue;
            tile[c] = tid;
            ++cnt;
            for (int e = aptr[c]; e < aptr[c + 1]; ++e) {
                const int nb = adj[e];
                if (tile[nb] >= 0) continue;
                ++ntaken[nb];
                seeds.push({-ntaken[nb], pos[nb]});
                if (depth[nb] < 0) { depth[nb] = it.d + 1; touched.push_back(nb); }
                front.push({depth[nb], -ntaken[nb], pos[nb], nb});
            }
        }
        for (int c : touched) depth[c] = -1;
        size.push_back(cnt);
    }
    // merge small tiles into a neighbour with room
    const int nt = (int)size.size();
    std::vector<std::vector<int>> members(nt);
    for (int i = 0; i < n; ++i) members[tile[rcm[i]]].push_back(rcm[i]);
    std::vector<int> by_size(nt);
    std::iota(by_size.begin(), by_size.end(), 0);
    std::stable_sort(by_size.begin(), by_size.end(), [&](int a, int b) { return size[a] < size[b]; });
    std::vector<int> share(nt, 0);
    for (int t : by_size) {
        if (size[t] == 0 || size[t] >= T / 2) continue;
        std::vector<int> cand;
        for (int c : members[t])
            for (int e = aptr[c]; e < aptr[c + 1]; ++e) {
                const int u = tile[adj[e]];
                if (u == t) continue;
                if (share[u]++ == 0) cand.push_back(u);
            }
        std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return share[a] > share[b]; });
        for (int u : cand) {
            if (size[u] + size[t] <= T) {
                for (int c : members[t]) tile[c] = u;
                members[u].insert(members[u].end(), members[t].begin(), members[t].end());
                members[t].clear();
                size[u] += size[t];
                size[t] = 0;
                break;
            }
        }
        for (int u : cand) share[u] = 0;
    }
    // device order: tiles in the order of their first cell in RCM order,
    // each tile's cells by face count, then in RCM order
    std::vector<int> first(nt, n);
    for (int i = 0; i < n; ++i) first[tile[rcm[i]]] = std::min(first[tile[rcm[i]]], i);
    std::vector<int> tord;
    for (int t = 0; t < nt; ++t)
        if (size[t] > 0) tord.push_back(t);
    std::stable_sort(tord.begin(), tord.end(), [&](int a, int b) { return first[a] < first[b]; });
    std::vector<int> out;
    out.reserve(n);
    s->tile_start.clear();
    for (int t : tord) {
        s->tile_start.push_back((int)out.size());
        std::vector<int>& mm = members[t];
        std::sort(mm.begin(), mm.end(), [&](int a, int b) {
            const int da = cfp[a + 1] - cfp[a], db = cfp[b + 1] - cfp[b];
            return da != db ? da < db : pos[a] < pos[b];
        });
        out.insert(out.end(), mm.begin(), mm.end());
    }
    s->tile_start.push_back(n);
    s->cell_new2old.swap(out);
}

// Per tile: the halo rings, the tile-local copies of the ELL entries of the
// cells that need gradients, the tile's face list and each owned cell's
// slots as indices into it.  The static geometry the tile kernel reads
// (per-slot gradient geometry sgeo, per-cell gradient data cdat, face
// geometry and MUSCL offsets) is also copied in tile order, so its loads are
// contiguous and need no index loaded first.  two_rings: the tile computes
// the gradients of its owned and ring-1 cells (a second halo ring, ring-1
// slots, sgeo and cdat); rec: also each slot's MUSCL offset (frec; the
// periodic Euler tiles).  Returns the LDS bytes of a second-order Euler tile
// (the viscous tiles size theirs from the maxima).
int build_tiles(fvk_solver* s, const std::vector<int>& cptr, const std::vector<int>& cent, const std::vector<int>& fl,
                const std::vector<int>& fr, int width, bool two_rings, bool rec,
                const std::vector<double4>& sgeo, const std::vector<double4>& cdat,
                const std::vector<double4>& fgeo, const std::vector<double4>& frec, size_t* lds_bytes,
                const std::vector<double4>* vrec = nullptr) {
    const int nt = (int)s->tile_start.size() - 1;
    const int T = s->block;  // the workgroup size (tiles hold at most tile_cells <= T cells)
    const int64_t nc = s->nc, nf = s->nf;
    std::vector<int> loc(nc, -1), fmark(nf, -1);
    std::vector<int> thptr(1, 0), thalo, th1(nt), tfptr(1, 0);
    std::vector<int4> tface;
    struct TileInfo { int n, h1, h2; };
    std::vector<TileInfo> info(nt);
    int mg = 0, ml = 0, mf = 0, m2 = 0;
    // pass 1: rings and faces
    std::vector<int> cells;
    for (int t = 0; t < nt; ++t) {
        const int c0 = s->tile_start[t], n = s->tile_start[t + 1] - c0;
        if (n > T) return set_error(FVK_ERR_INVALID, "tile %d has %d cells (more than %d)", t, n, T);
        cells.clear();
        for (int j = 0; j < n; ++j) { loc[c0 + j] = j; cells.push_back(c0 + j); }
        auto add_ring = [&](int from, int to) {
            for (int j = from; j < to; ++j) {
                const int c = cells[j];
                for (int k = cptr[c]; k < cptr[c + 1]; ++k) {
                    const int f = cent[k] >> 1;
                    if (fr[f] < 0) continue;
                    const int nb = (cent[k] & 1) ? fr[f] : fl[f];
                    if (loc[nb] < 0) { loc[nb] = (int)cells.size(); cells.push_back(nb); }
                }
            }
        };
        add_ring(0, n);
        const int h1 = (int)cells.size() - n;
        // ring 1 by face count (its cells compute gradients, as the owned ones
        // do), then in the order found
        std::stable_sort(cells.begin() + n, cells.end(),
                         [&](int a, int b) { return cptr[a + 1] - cptr[a] < cptr[b + 1] - cptr[b]; });
        for (int j = n; j < n + h1; ++j) loc[cells[j]] = j;
        if (two_rings) add_ring(n, n + h1);
        const int h2 = (int)cells.size() - n - h1;
        info[t] = TileInfo{n, h1, h2};
        for (size_t j = n; j < cells.size(); ++j) thalo.push_back(cells[j]);
        thptr.push_back((int)thalo.size());
        th1[t] = h1;
        // faces of the owned cells, each once
        int nft = 0;
        for (int j = 0; j < n; ++j) {
            const int c = c0 + j;
            for (int k = cptr[c]; k < cptr[c + 1]; ++k) {
                const int f = cent[k] >> 1;
                if (fmark[f] == t) continue;
                fmark[f] = t;
                const int r = fr[f] >= 0 ? loc[fr[f]] : fr[f];
                tface.push_back(int4{f, loc[fl[f]], r, 0});
                ++nft;
            }
        }
        tfptr.push_back((int)tface.size());
        mg = std::max(mg, n + h1);
        ml = std::max(ml, n + h1 + h2);
        m2 = std::max(m2, h2);
        mf = std::max(mf, nft);
        for (int c : cells) loc[c] = -1;
    }
    // the packed face cells hold local indices and patch codes in 16 bits each,
    // the packed slots local indices and tile face indices in 15
    if (ml > 32767 || mf > 32766 || s->n_bcs > 32767)
        return set_error(FVK_ERR_INVALID, "tiles of %d staged cells or %d patches exceed the packed face format", ml,
                         s->n_bcs);
    // pass 2: tile-local ELL and face slots (need the maxima for the strides)
    std::vector<int> tell((size_t)nt * width * mg, -1);
    std::vector<double4> tegeo, tclsq;
    std::vector<double2> trec;
    if (two_rings) {
        tegeo.assign((size_t)nt * width * mg, double4{0.0, 0.0, 0.0, 0.0});
        tclsq.assign((size_t)nt * mg, double4{0.0, 0.0, 0.0, 0.0});
        if (rec) trec.assign((size_t)nt * width * mg, double2{0.0, 0.0});
    }
    const int64_t ld = s->mesh.ld;
    if (nf >= (1 << 30) - 1) return set_error(FVK_ERR_INVALID, "%lld faces exceed the packed tile format", (long long)nf);
    for (int t = 0; t < nt; ++t) {
        const int c0 = s->tile_start[t];
        const TileInfo& ti = info[t];
        cells.clear();
        for (int j = 0; j < ti.n; ++j) cells.push_back(c0 + j);
        for (int j = thptr[t]; j < thptr[t + 1]; ++j) cells.push_back(thalo[j]);
        for (size_t j = 0; j < cells.size(); ++j) loc[cells[j]] = (int)j;
        const int nft = tfptr[t + 1] - tfptr[t];
        for (int e = tfptr[t]; e < tfptr[t + 1]; ++e) fmark[tface[e].x] = e - tfptr[t];
        // without gradients the kernel reads the local entries of the owned cells only (no ring-2 halo)
        const int nell = two_rings ? ti.n + ti.h1 : ti.n;
        for (int j = 0; j < nell; ++j) {
            const int c = cells[j];
            if (two_rings) {
                tclsq[(size_t)t * mg + j] = cdat[c];
                for (int k = 0; k < cptr[c + 1] - cptr[c]; ++k)
                    tegeo[((size_t)t * width + k) * mg + j] = sgeo[(size_t)k * ld + c];
            }
            for (int k = cptr[c]; k < cptr[c + 1]; ++k) {
                const int slot = k - cptr[c];
                const int e = cent[k], f = e >> 1;
                int nb;
                if (fr[f] < 0) nb = fr[f];
                else {
                    const int g = (e & 1) ? fr[f] : fl[f];
                    nb = loc[g];
                    if (nb < 0) return set_error(FVK_ERR_INVALID, "tile %d: a neighbour is outside its halo", t);
                }
                if (nb >= ti.n + ti.h1 + ti.h2) return set_error(FVK_ERR_INVALID, "tile %d: local index out of range", t);
                // the slot's face in the tile's list: every face of an owned cell,
                // and a ring-1 cell's faces towards the owned cells
                const int lf = fmark[f];
                const bool ours = lf >= 0 && lf < nft && tface[tfptr[t] + lf].x == f;
                if (j < ti.n && !ours) return set_error(FVK_ERR_INVALID, "tile %d: face slot out of range", t);
                tell[((size_t)t * width + slot) * mg + j] =
                    nb >= 0 ? tell_interior(nb, e & 1, ours ? lf : -1) : tell_boundary_slot(j < ti.n, lf, f);
                // periodic: the MUSCL offset of the face's side (frec), which for
                // cell_right differs from the limiter offset in rounding
                if (two_rings && rec) {
                    const double4 fr4 = frec[f];
                    trec[((size_t)t * width + slot) * mg + j] =
                        (e & 1) ? double2{fr4.x, fr4.y} : double2{fr4.z, fr4.w};
                }
            }
        }
        // every face of the tile reads cells with staged states (owned or ring 1)
        for (int e = tfptr[t]; e < tfptr[t + 1]; ++e) {
            const int4 fd = tface[e];
            if (fd.y < 0 || fd.y >= ti.n + ti.h1 || fd.z >= ti.n + ti.h1 || (fd.z < 0 && -1 - fd.z >= s->n_bcs))
                return set_error(FVK_ERR_INVALID, "tile %d: face %d has a cell outside the staged states", t, fd.x);
        }
        for (int c : cells) loc[c] = -1;
    }
    MeshView& m = s->mesh;
    m.ntiles = nt;
    m.tmax_g = mg;
    m.tmax_l = ml;
    m.tmax_f = mf;
    m.tmax_2 = m2;
    m.twidth = width;
    std::vector<int4> tdesc(2 * (size_t)nt);
    for (int t = 0; t < nt; ++t) {
        const int c0 = s->tile_start[t];
        tdesc[2 * t] = int4{c0, s->tile_start[t + 1] - c0, thptr[t], th1[t]};
        tdesc[2 * t + 1] = int4{thptr[t + 1] - thptr[t], tfptr[t], tfptr[t + 1] - tfptr[t], 0};
    }
    m.tdesc = upload(s, tdesc);
    m.nthalo = (int)thalo.size();
    m.tstart = upload(s, s->tile_start);
    m.thptr = upload(s, thptr);
    m.thalo = thalo.empty() ? nullptr : upload(s, thalo);
    m.th1 = upload(s, th1);
    m.tell = upload(s, tell);
    m.tfptr = upload(s, tfptr);
    m.tface = upload(s, tface);
    std::vector<int> tfc(tface.size());
    for (size_t e = 0; e < tface.size(); ++e) tfc[e] = tface_pack(tface[e].y, tface[e].z);
    m.tfc = upload(s, tfc);
    std::vector<double4> tfgeo(tface.size());
    for (size_t e = 0; e < tface.size(); ++e) tfgeo[e] = fgeo[tface[e].x];
    m.tfgeo = upload(s, tfgeo);
    m.tvf = nullptr;
    if (vrec) {  // the viscous face records (two per device face) in tile face order
        std::vector<double4> tvf(2 * tface.size());
        for (size_t e = 0; e < tface.size(); ++e) {
            tvf[2 * e] = (*vrec)[2 * (size_t)tface[e].x];
            tvf[2 * e + 1] = (*vrec)[2 * (size_t)tface[e].x + 1];
        }
        m.tvf = upload(s, tvf);
        if (!m.tvf) return set_error(FVK_ERR_NOMEM, "device allocation for the tiles failed");
    }
    m.tegeo = two_rings ? upload(s, tegeo) : nullptr;
    m.tclsq = two_rings ? upload(s, tclsq) : nullptr;
    m.trec = two_rings && rec ? upload(s, trec) : nullptr;
    if (!m.tdesc || !m.tstart || !m.thptr || (!thalo.empty() && !m.thalo) || !m.th1 || !m.tell || !m.tfptr ||
        !m.tface || !m.tfc ||
        !m.tfgeo || (two_rings && (!m.tegeo || !m.tclsq || (rec && !m.trec))))
        return set_error(FVK_ERR_NOMEM, "device allocation for the tiles failed");
    // second order: U (4) and W (3: u, v, p) of the owned and ring-1 cells, W
    // (4) of ring 2, then the MUSCL increments (4 per side) of each tile face,
    // which each face's fluxes (4) and dt candidate (1) replace; first order:
    // U (4) of the owned and ring-1 cells, then the fluxes and dt candidates
    const size_t words = two_rings ? (size_t)7 * mg + (size_t)4 * m2 + (size_t)8 * mf
                                   : (size_t)4 * mg + (size_t)5 * mf;
    *lds_bytes = sizeof(double) * std::max(words, (size_t)4 * T);  // at least the block reductions' 4 T doubles
    return FVK_OK;
}

int build_layout(fvk_solver* s, const fvk_mesh_desc* d) {
    const int64_t nc = d->n_cells, nf = d->n_faces, nn = d->n_nodes;
    if (nc <= 0 || nf <= 0 || nn <= 0 || nc > (1LL << 30) || nf > (1LL << 30))
        return set_error(FVK_ERR_INVALID, "mesh sizes out of range (cells %lld, faces %lld, nodes %lld)",
                         (long long)nc, (long long)nf, (long long)nn);
    if (!d->node_xy || !d->face_nodes || !d->face_cells || !d->face_patch || !d->cell_face_ptr || !d->cell_faces ||
        !d->cell_volume || !d->cell_xy)
        return set_error(FVK_ERR_INVALID, "mesh descriptor has a NULL array");
    s->nc = nc;
    s->nf = nf;
    s->nn = nn;
    const int32_t* fc = d->face_cells;
    // ---- validation (the reference's reader rejects the same defects, src/MeshReader.cpp:407-443) ----
    if (d->cell_face_ptr[0] != 0) return set_error(FVK_ERR_INVALID, "cell_face_ptr[0] must be 0");
    for (int64_t c = 0; c < nc; ++c)
        if (d->cell_face_ptr[c + 1] < d->cell_face_ptr[c]) return set_error(FVK_ERR_INVALID, "cell_face_ptr not monotonic");
    const int64_t nnz = d->cell_face_ptr[nc];
    for (int64_t k = 0; k < nnz; ++k)
        if (d->cell_faces[k] < 0 || d->cell_faces[k] >= nf) return set_error(FVK_ERR_INVALID, "cell_faces entry out of range");
    for (int64_t f = 0; f < nf; ++f) {
        if (fc[2 * f] < 0 || fc[2 * f] >= nc || fc[2 * f + 1] >= nc || fc[2 * f + 1] < -1)
            return set_error(FVK_ERR_INVALID, "face %lld has an out-of-range cell", (long long)f);
        if (d->face_nodes[2 * f] < 0 || d->face_nodes[2 * f] >= nn || d->face_nodes[2 * f + 1] < 0 ||
            d->face_nodes[2 * f + 1] >= nn)
            return set_error(FVK_ERR_INVALID, "face %lld has an out-of-range node", (long long)f);
        if (fc[2 * f + 1] == -1 && (d->face_patch[f] < 0 || d->face_patch[f] >= s->n_bcs))
            return set_error(FVK_ERR_INVALID,
                             "boundary face %lld has patch %d but %d boundary conditions were given "
                             "(every boundary face must belong to a patch with a boundary condition)",
                             (long long)f, d->face_patch[f], s->n_bcs);
    }
    s->periodic = d->face_shift != nullptr;
    s->node_xy.assign(d->node_xy, d->node_xy + 2 * nn);
    s->cell_ptr_old.assign(d->cell_face_ptr, d->cell_face_ptr + nc + 1);
    s->cell_faces_old.assign(d->cell_faces, d->cell_faces + nnz);
    s->face_patch_old.assign(d->face_patch, d->face_patch + d->n_faces);

    // ---- face geometry, reference order (UnstructuredMesh::compute_geometry) ----
    std::vector<double> mx(nf), my(nf), ar(nf), nx(nf), ny(nf);
    for (int64_t f = 0; f < nf; ++f) {
        const double* n1 = d->node_xy + 2 * d->face_nodes[2 * f];
        const double* n2 = d->node_xy + 2 * d->face_nodes[2 * f + 1];
        mx[f] = 0.5 * (n1[0] + n2[0]);
        my[f] = 0.5 * (n1[1] + n2[1]);
        double dx = n2[0] - n1[0], dy = n2[1] - n1[1];
        ar[f] = std::sqrt(dx * dx + dy * dy);
        nx[f] = dy / ar[f];
        ny[f] = -dx / ar[f];
    }
    const double* sh = d->face_shift;
    auto shx = [&](int64_t f) { return sh ? sh[2 * f] : 0.0; };
    auto shy = [&](int64_t f) { return sh ? sh[2 * f + 1] : 0.0; };

    // ---- RCM over the interior-face graph ----
    std::vector<int> aptr(nc + 1, 0), adj;
    for (int64_t f = 0; f < nf; ++f)
        if (fc[2 * f + 1] >= 0 && fc[2 * f + 1] != fc[2 * f]) { aptr[fc[2 * f] + 1]++; aptr[fc[2 * f + 1] + 1]++; }
    for (int64_t c = 0; c < nc; ++c) aptr[c + 1] += aptr[c];
    adj.resize(aptr[nc]);
    {
        std::vector<int> fill(aptr.begin(), aptr.end() - 1);
        for (int64_t f = 0; f < nf; ++f) {
            int a = fc[2 * f], b = fc[2 * f + 1];
            if (b >= 0 && a != b) { adj[fill[a]++] = b; adj[fill[b]++] = a; }
        }
    }
    rcm_order((int)nc, aptr, adj, s->cell_new2old);
    if (!s->cell_case_old.empty()) {
        // batched cases: one contiguous range per case, RCM order within it
        // (the cases are disconnected, so this only reorders whole components)
        const std::vector<int>& cc = s->cell_case_old;
        std::stable_sort(s->cell_new2old.begin(), s->cell_new2old.end(), [&](int a, int b) { return cc[a] < cc[b]; });
        s->case_cell_start.assign(s->n_cases + 1, 0);
        for (int64_t c = 0; c < nc; ++c) s->case_cell_start[cc[c] + 1]++;
        for (int k = 0; k < s->n_cases; ++k) s->case_cell_start[k + 1] += s->case_cell_start[k];
    }
    if (s->variant == 3) tile_order(s, (int)nc, aptr, adj, s->tile_cells, d->cell_face_ptr);
    s->cell_old2new.assign(nc, 0);
    for (int64_t i = 0; i < nc; ++i) s->cell_old2new[s->cell_new2old[i]] = (int)i;
    const std::vector<int>& o2n = s->cell_old2new;

    // ---- face order ----
    std::vector<int64_t> key(nf);
    std::vector<int> forder(nf);
    std::iota(forder.begin(), forder.end(), 0);
    for (int64_t f = 0; f < nf; ++f) {
        int a = o2n[fc[2 * f]];
        if (fc[2 * f + 1] >= 0) {
            int b = o2n[fc[2 * f + 1]];
            key[f] = (int64_t)std::min(a, b) * (int64_t)(nc + 1) + std::max(a, b);
        } else {
            int t = s->bcs[d->face_patch[f]].type;
            key[f] = ((int64_t)(nc + 1) * (nc + 1)) + ((int64_t)t * (nc + 1) + a);
        }
    }
    std::stable_sort(forder.begin(), forder.end(), [&](int a, int b) { return key[a] < key[b]; });
    s->face_new2old = forder;
    s->face_old2new.assign(nf, 0);
    for (int64_t i = 0; i < nf; ++i) s->face_old2new[forder[i]] = (int)i;
    int nfi = 0;
    for (int64_t i = 0; i < nf; ++i)
        if (fc[2 * forder[i] + 1] >= 0) nfi = (int)i + 1;
    s->nfi = nfi;
    s->nnz = (int)nnz;

    // ---- CSR in new cell order, reference face order within a cell ----
    std::vector<int> cptr(nc + 1, 0), cent(nnz);
    for (int64_t i = 0; i < nc; ++i) {
        int c = s->cell_new2old[i];
        cptr[i + 1] = cptr[i] + (d->cell_face_ptr[c + 1] - d->cell_face_ptr[c]);
    }
    for (int64_t i = 0; i < nc; ++i) {
        int c = s->cell_new2old[i];
        int k0 = cptr[i];
        for (int k = d->cell_face_ptr[c]; k < d->cell_face_ptr[c + 1]; ++k) {
            int f = d->cell_faces[k];
            int owner = (fc[2 * f] == c) ? 1 : 0;
            cent[k0++] = (s->face_old2new[f] << 1) | owner;
        }
    }

    // ---- per-cell data ----
    std::vector<double> vol(nc), cx(nc), cy(nc), lsq(3 * nc);
    for (int64_t i = 0; i < nc; ++i) {
        int c = s->cell_new2old[i];
        vol[i] = d->cell_volume[c];
        cx[i] = d->cell_xy[2 * c];
        cy[i] = d->cell_xy[2 * c + 1];
    }
    // least-squares inverse, src/GradientReconstruction.cpp:56-79 (neighbour:
    // opposite centroid, periodic-shifted; boundary: face midpoint)
    for (int64_t i = 0; i < nc; ++i) {
        int c = s->cell_new2old[i];
        double m00 = 0.0, m01 = 0.0, m11 = 0.0;
        for (int k = d->cell_face_ptr[c]; k < d->cell_face_ptr[c + 1]; ++k) {
            int f = d->cell_faces[k];
            int cl = fc[2 * f], cr = fc[2 * f + 1];
            double qx, qy;
            if (cr == -1) { qx = mx[f]; qy = my[f]; }
            else if (cl == c) {
                qx = d->cell_xy[2 * cr]; qy = d->cell_xy[2 * cr + 1];
                if (sh) { qx = qx + shx(f); qy = qy + shy(f); }
            } else {
                qx = d->cell_xy[2 * cl]; qy = d->cell_xy[2 * cl + 1];
                if (sh) { qx = qx - shx(f); qy = qy - shy(f); }
            }
            double dx = qx - d->cell_xy[2 * c], dy = qy - d->cell_xy[2 * c + 1];
            double w = 1.0 / (dx * dx + dy * dy);
            m00 += w * dx * dx;
            m01 += w * dx * dy;
            m11 += w * dy * dy;
        }
        double det = m00 * m11 - m01 * m01;
        lsq[i] = m11 / det;
        lsq[nc + i] = -m01 / det;
        lsq[2 * nc + i] = m00 / det;
    }

    // ---- per-face data in device order ----
    std::vector<int> fl(nf), fr(nf), fnodes(2 * nf);
    std::vector<double> fnx(nf), fny(nf), fa(nf), fmx(nf), fmy(nf), fdl(nf), fdc(nf), fkx(nf), fky(nf), fwl(nf),
        fwr(nf), fsx, fsy;
    if (sh) { fsx.resize(nf); fsy.resize(nf); }
    for (int64_t i = 0; i < nf; ++i) {
        int f = forder[i];
        int cl = fc[2 * f], cr = fc[2 * f + 1];
        fl[i] = o2n[cl];
        fr[i] = cr >= 0 ? o2n[cr] : -1 - d->face_patch[f];
        fnodes[2 * i] = d->face_nodes[2 * f];
        fnodes[2 * i + 1] = d->face_nodes[2 * f + 1];
        fnx[i] = nx[f]; fny[i] = ny[f]; fa[i] = ar[f]; fmx[i] = mx[f]; fmy[i] = my[f];
        if (sh) { fsx[i] = shx(f); fsy[i] = shy(f); }
        // Euler compute_dt length, src/EulerFVMSolver.cpp:153-155
        fdl[i] = cr >= 0 ? std::min(d->cell_volume[cl], d->cell_volume[cr]) / ar[f] : d->cell_volume[cl] / ar[f];
        // face_normal_geometry, src/GradientReconstruction.cpp:28-40
        double xR, yR;
        if (cr == -1) { xR = mx[f]; yR = my[f]; }
        else {
            xR = d->cell_xy[2 * cr]; yR = d->cell_xy[2 * cr + 1];
            if (sh) { xR = xR + shx(f); yR = yR + shy(f); }
        }
        double xL = d->cell_xy[2 * cl], yL = d->cell_xy[2 * cl + 1];
        double ddx = xR - xL, ddy = yR - yL;
        double dist = std::hypot(ddx, ddy);
        double dhx = ddx / dist, dhy = ddy / dist;
        double cth = nx[f] * dhx + ny[f] * dhy;
        fdc[i] = dist * cth;
        fkx[i] = nx[f] - dhx / cth;
        fky[i] = ny[f] - dhy / cth;
        // distance weights, src/GradientReconstruction.cpp:103-106 and :174-176
        if (cr >= 0) {
            double dL = std::hypot(xL - mx[f], yL - my[f]);
            double dR = std::hypot(xR - mx[f], yR - my[f]);
            fwl[i] = dR / (dL + dR);
            fwr[i] = dL / (dR + dL);
        } else {
            fwl[i] = 1.0;
            fwr[i] = 0.0;
        }
    }
    s->face_nodes_dev = fnodes;

    // ---- upload ----
    MeshView& m = s->mesh;
    m.nc = (int)nc;
    m.nf = (int)nf;
    m.nfi = nfi;
    m.nnz = (int)nnz;
    m.vol = upload(s, vol);
    m.cx = upload(s, cx);
    m.cy = upload(s, cy);
    m.lsq = upload(s, lsq);
    m.cptr = upload(s, cptr);
    m.cent = upload(s, cent);
    m.fl = upload(s, fl);
    m.fr = upload(s, fr);
    m.fnx = upload(s, fnx);
    m.fny = upload(s, fny);
    m.fa = upload(s, fa);
    m.fmx = upload(s, fmx);
    m.fmy = upload(s, fmy);
    m.fsx = sh ? upload(s, fsx) : nullptr;
    m.fsy = sh ? upload(s, fsy) : nullptr;
    m.fdl = upload(s, fdl);
    m.fdc = upload(s, fdc);
    m.fkx = upload(s, fkx);
    m.fky = upload(s, fky);
    m.fwl = upload(s, fwl);
    m.fwr = upload(s, fwr);
    if (nf > 0 && (!m.fnx || !m.fny || !m.fa || !m.fmx || !m.fmy || !m.fdl || !m.fdc || !m.fkx || !m.fky || !m.fwl ||
                   !m.fwr))
        return set_error(FVK_ERR_NOMEM, "device allocation for the face geometry failed");
    std::vector<DevBC> hb(std::max(1, s->n_bcs));
    for (int i = 0; i < s->n_bcs; ++i) {
        const fvk_bc& b = s->bcs[i];
        DevBC& o = hb[i];
        o.type = b.type;
        o.isothermal = b.is_isothermal_wall;
        o.wall_u = b.wall_u; o.wall_v = b.wall_v; o.wall_T = b.wall_temperature;
        for (int q = 0; q < 4; ++q) o.ff[q] = b.farfield_state[q];
        o.ff_nut = b.farfield_nut; o.ff_k = b.farfield_k; o.ff_w = b.farfield_omega;
        o.value = b.value;
    }
    m.bc = upload(s, hb);
    if (!m.vol || !m.cptr || !m.cent || !m.fl || !m.bc)
        return set_error(FVK_ERR_NOMEM, "device allocation for the mesh failed");

    // ---- gather-friendly copies (cell-centric Euler kernels) ----
    const int64_t ld = (nc + 255) / 256 * 256;
    int deg = 0;
    for (int64_t i = 0; i < nc; ++i) deg = std::max(deg, cptr[i + 1] - cptr[i]);
    // the per-cell kernels are instantiated for widths 3, 4, 9 and 16 and read
    // every slot of their width: pad the slot arrays to it
    const int width = deg <= 3 ? 3 : deg <= 4 ? 4 : deg <= 9 ? 9 : std::max(deg, kMaxEulerDegree);
    std::vector<int2> ell((size_t)width * ld, int2{-1, -1});
    for (int64_t i = 0; i < nc; ++i) {
        for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
            int e = cent[k];
            int f = e >> 1;
            int nb;
            if (fr[f] < 0) nb = fr[f];
            else nb = (e & 1) ? fr[f] : fl[f];
            ell[(size_t)(k - cptr[i]) * ld + i] = int2{e, nb};
        }
    }
    // eps^2 = (K sqrt(V))^3 of the Venkatakrishnan limiter (same operation order as the kernels had)
    std::vector<double4> cgeo(ld, double4{0.0, 0.0, 1.0, 0.0}), clsq(ld, double4{0.0, 0.0, 0.0, 0.0});
    for (int64_t i = 0; i < nc; ++i) {
        double kh = s->p.venkat_k * std::sqrt(vol[i]);
        cgeo[i] = double4{cx[i], cy[i], vol[i], kh * kh * kh};
        clsq[i] = double4{lsq[i], lsq[nc + i], lsq[2 * nc + i], kh * kh * kh};
    }
    // per ELL entry: LSQ weight times offset (w = 1/(dx^2+dy^2), GradientCalculator,
    // src/GradientReconstruction.cpp:126-143) and the face midpoint offset, both
    // in the cell's own frame (periodic shift applied), evaluated in the kernels' order
    std::vector<double4> egeo((size_t)width * ld, double4{0.0, 0.0, 0.0, 0.0});
    for (int64_t i = 0; i < nc; ++i) {
        for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
            int e = cent[k];
            int f = e >> 1;
            bool owner = e & 1;
            double qx, qy;
            if (fr[f] < 0) {
                qx = fmx[f];
                qy = fmy[f];
            } else {
                int o = owner ? fr[f] : fl[f];
                qx = cx[o];
                qy = cy[o];
                if (sh) {
                    if (owner) { qx = qx + fsx[f]; qy = qy + fsy[f]; }
                    else { qx = qx - fsx[f]; qy = qy - fsy[f]; }
                }
            }
            double dx = qx - cx[i], dy = qy - cy[i];
            double w = 1.0 / (dx * dx + dy * dy);
            double mx = fmx[f], my = fmy[f];
            if (sh && !owner) { mx = mx - fsx[f]; my = my - fsy[f]; }
            egeo[(size_t)(k - cptr[i]) * ld + i] = double4{w * dx, w * dy, mx - cx[i], my - cy[i]};
        }
    }
    // per face: MUSCL offsets of both sides (midpoint - centroid; cell_right periodic-shifted)
    std::vector<double4> frec(nf);
    for (int64_t i = 0; i < nf; ++i) {
        int l = fl[i], r = fr[i];
        double rx = 0.0, ry = 0.0;
        if (r >= 0) {
            double ox = cx[r], oy = cy[r];
            if (sh) { ox = ox + fsx[i]; oy = oy + fsy[i]; }
            rx = fmx[i] - ox;
            ry = fmy[i] - oy;
        }
        frec[i] = double4{fmx[i] - cx[l], fmy[i] - cy[l], rx, ry};
    }
    std::vector<double4> fgeo(nf);
    std::vector<double2> fmid2(nf), fsh2;
    for (int64_t i = 0; i < nf; ++i) {
        fgeo[i] = double4{fnx[i], fny[i], fa[i], fdl[i]};
        fmid2[i] = double2{fmx[i], fmy[i]};
    }
    if (sh) {
        fsh2.resize(nf);
        for (int64_t i = 0; i < nf; ++i) fsh2[i] = double2{fsx[i], fsy[i]};
    }
    m.deg = deg;
    m.ld = ld;
    m.ell = upload(s, ell);
    m.cgeo = upload(s, cgeo);
    m.clsq = upload(s, clsq);
    m.fgeo = upload(s, fgeo);
    m.fmid2 = upload(s, fmid2);
    m.fsh2 = sh ? upload(s, fsh2) : nullptr;
    m.egeo = upload(s, egeo);
    m.frec = upload(s, frec);
    if (!m.ell || !m.cgeo || !m.clsq || !m.fgeo || !m.fmid2 || (sh && !m.fsh2) || !m.egeo || !m.frec)
        return set_error(FVK_ERR_NOMEM, "device allocation for the mesh failed");
    if (s->variant == 3) {
        size_t lds = 0;
        const bool euler = s->p.equation == FVK_EQ_EULER;
        int rc;
        // viscous face records: {n_x, n_y, A, w_left}, {w_right, d.n, k_x, k_y} per device face
        std::vector<double4> vrec;
        if (!euler) {
            vrec.resize(2 * (size_t)nf);
            for (int64_t i = 0; i < nf; ++i) {
                vrec[2 * i] = double4{fnx[i], fny[i], fa[i], fwl[i]};
                vrec[2 * i + 1] = double4{fwr[i], fdc[i], fkx[i], fky[i]};
            }
        }
        if (!euler && width > kMaxEulerDegree)
            return set_error(FVK_ERR_UNSUPPORTED, "viscous tiles take cells of at most %d faces (a cell has %d)",
                             kMaxEulerDegree, deg);
        if (euler) {
            const bool second = s->p.order == 2;
            rc = build_tiles(s, cptr, cent, fl, fr, width, second, second && sh != nullptr, egeo, clsq, fgeo, frec,
                             &lds);
        } else if (s->p.gradient_scheme == FVK_GRAD_LEAST_SQUARES) {
            // viscous tiles compute their gradients: least squares reads the
            // per-slot weight times offset and the LSQ inverse
            rc = build_tiles(s, cptr, cent, fl, fr, width, true, false, egeo, clsq, fgeo, frec, &lds, &vrec);
        } else {
            // Green-Gauss: per slot the cell's distance weight, the normal
            // pointing away from the cell and the face length; per cell the volume
            std::vector<double4> ggeo((size_t)width * ld, double4{0.0, 0.0, 0.0, 0.0}), gvol(ld, double4{1.0, 0.0, 0.0, 0.0});
            for (int64_t i = 0; i < nc; ++i) {
                gvol[i] = double4{vol[i], 0.0, 0.0, 0.0};
                for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
                    const int f = cent[k] >> 1;
                    const bool owner = cent[k] & 1;
                    ggeo[(size_t)(k - cptr[i]) * ld + i] =
                        double4{owner ? fwl[f] : fwr[f], owner ? fnx[f] : -fnx[f], owner ? fny[f] : -fny[f], fa[f]};
                }
            }
            rc = build_tiles(s, cptr, cent, fl, fr, width, true, false, ggeo, gvol, fgeo, frec, &lds, &vrec);
        }
        if (rc != FVK_OK) return rc;
        if (!euler) {
            // viscous tiles: state, gradients (and SST's F1, nu_t) of the owned and
            // ring-1 cells, then the primitives of every staged cell, which the
            // face fluxes and dt candidates replace once the gradients are
            // done; at least the residual reduction's nv doubles per thread
            const int nv = s->nv, nf = nv - 1, nx = s->p.equation == FVK_EQ_RANS_SST ? 2 : 0;
            const size_t words = (size_t)(nv + 2 * nf + nx) * m.tmax_g +
                                 std::max((size_t)nf * m.tmax_l, (size_t)(nv + 1) * m.tmax_f);
            lds = sizeof(double) * std::max(words, (size_t)nv * s->block);
        }
        if (lds > kMaxTileLds)
            return set_error(FVK_ERR_UNSUPPORTED, "a tile of %d cells needs %zu B of LDS (limit %d): use smaller tiles",
                             s->block, lds, kMaxTileLds);
        s->tile_lds = lds;
    }
    return FVK_OK;
}

}  // namespace fvk
