// Mesh ingest: the reference's readers for Gmsh ASCII 2.2 and 4.x and for
// FVMESH 1.0 (src/MeshReader.cpp, docs/fvmesh-format.md), producing the
// arrays of fvk_mesh_desc with the reference's face creation order, face
// orientation, Cell::faces order, patch numbering and cell geometry, so a
// case file's mesh gives the same residual bits through either reader.
//
// The reference matches shared edges with a std::map keyed by the node pair,
// one lookup per polygon edge.  Here every half-edge (cell, edge) is tagged
// with its node-pair key and its visit index in the reference's walk (cells
// in order, edges in polygon order); one sort by (key, visit) groups the
// visits of each edge, the earliest visit of a group creates the face, and
// faces are numbered by that visit: the same order the map-based walk
// produces, without a tree.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "fvk.h"
#include "solver.h"

struct fvk_mesh_file {
    std::vector<double> node_xy;
    std::vector<int32_t> face_nodes, face_cells, face_patch;
    std::vector<int32_t> cell_ptr, cell_faces, cell_node_ptr, cell_nodes;
    std::vector<double> cell_volume, cell_xy;
    std::vector<std::string> patch_names;
};

namespace {

using fvk::set_error;

// A 2D element of the file (a cell) or a 1D boundary element, as the
// reference's TempElement: node indices already mapped to 0-based
struct Element {
    int physical_id;
    std::vector<int> nodes;
};

// A cursor over the whole file with the two reading primitives the
// reference's parsers mix: whitespace-skipping token reads (operator>>) and
// getline
struct Cursor {
    const char* p;
    const char* end;
    bool fail = false;

    bool eof() const { return p >= end; }
    void skip_ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    }
    // std::getline: the characters up to the next '\n', which is consumed
    bool getline(std::string& out) {
        if (p >= end) return false;
        const char* q = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!q) q = end;
        out.assign(p, (size_t)(q - p));
        p = q < end ? q + 1 : end;
        return true;
    }
    long long read_ll() {
        skip_ws();
        if (p >= end) { fail = true; return 0; }
        char* e = nullptr;
        long long v = std::strtoll(p, &e, 10);
        if (e == p) { fail = true; return 0; }
        p = e;
        return v;
    }
    int read_int() { return (int)read_ll(); }
    double read_double() {
        skip_ws();
        if (p >= end) { fail = true; return 0.0; }
        char* e = nullptr;
        double v = std::strtod(p, &e);
        if (e == p) { fail = true; return 0.0; }
        p = e;
        return v;
    }
    std::string read_word() {
        skip_ws();
        const char* s = p;
        while (p < end && !(*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
        return std::string(s, (size_t)(p - s));
    }
};

bool starts_with(const std::string& s, const char* pre) { return s.rfind(pre, 0) == 0; }

// the physical name of a $PhysicalNames line: the text between the first
// and the last double quote, else the rest of the line after "dim id"
void physical_name_line(const std::string& line, std::map<int, std::string>& names) {
    Cursor c{line.data(), line.data() + line.size()};
    int dim = c.read_int(), id = c.read_int();
    std::string name(c.p, (size_t)(c.end - c.p));
    size_t q0 = name.find('"'), q1 = name.rfind('"');
    if (q0 != std::string::npos && q1 != std::string::npos && q1 > q0) name = name.substr(q0 + 1, q1 - q0 - 1);
    if (dim == 1) names[id] = name;
}

int element_nodes(int type) { return type == 1 ? 2 : (type == 2 ? 3 : (type == 3 ? 4 : 0)); }

// read_gmsh_v22, src/MeshReader.cpp:465-574
int read_gmsh22(Cursor& in, std::vector<double>& xy, std::vector<Element>& cells, std::vector<Element>& lines,
                std::map<int, std::string>& names) {
    std::unordered_map<long long, int> id2i;
    std::string line;
    while (in.getline(line)) {
        if (line.empty()) continue;
        if (starts_with(line, "$PhysicalNames")) {
            int count = in.read_int();
            in.getline(line);
            for (int i = 0; i < count && in.getline(line); ++i) physical_name_line(line, names);
        } else if (starts_with(line, "$Nodes")) {
            int count = in.read_int();
            xy.assign(2 * (size_t)std::max(count, 0), 0.0);
            for (int i = 0; i < count; ++i) {
                long long id = in.read_ll();
                double x = in.read_double(), y = in.read_double();
                in.read_double();
                id2i[id] = i;
                xy[2 * i] = x;
                xy[2 * i + 1] = y;
            }
            in.getline(line);
        } else if (starts_with(line, "$Elements")) {
            int count = in.read_int();
            in.getline(line);
            for (int i = 0; i < count && in.getline(line); ++i) {
                Cursor e{line.data(), line.data() + line.size()};
                long long number = e.read_ll();
                int type = e.read_int(), ntags = e.read_int();
                int phys = -1;
                for (int t = 0; t < ntags; ++t) {
                    int tag = e.read_int();
                    if (t == 0) phys = tag;
                }
                int nn = element_nodes(type);
                if (nn == 0) continue;
                Element el{phys, std::vector<int>((size_t)nn)};
                for (int k = 0; k < nn; ++k) {
                    long long nid = e.read_ll();
                    auto it = id2i.find(nid);
                    if (it == id2i.end())
                        return set_error(FVK_ERR_INVALID,
                                         "Error reading mesh: $Elements entry %lld (vertex %d) references undefined node "
                                         "id %lld", number, k, nid);
                    el.nodes[k] = it->second;
                }
                (type == 1 ? lines : cells).push_back(std::move(el));
            }
        }
    }
    return FVK_OK;
}

// read_gmsh_v41, src/MeshReader.cpp:591-782
int read_gmsh41(Cursor& in, std::vector<double>& xy, std::vector<Element>& cells, std::vector<Element>& lines,
                std::map<int, std::string>& names) {
    std::unordered_map<long long, int> id2i;
    std::map<std::pair<int, int>, std::vector<int>> entity_tags;
    std::string line;
    while (in.getline(line)) {
        if (line.empty()) continue;
        if (starts_with(line, "$PhysicalNames")) {
            int count = in.read_int();
            in.getline(line);
            for (int i = 0; i < count && in.getline(line); ++i) physical_name_line(line, names);
        } else if (starts_with(line, "$Entities")) {
            int np = in.read_int(), ncv = in.read_int(), ns = in.read_int(), nv = in.read_int();
            in.getline(line);
            for (int i = 0; i < np; ++i) in.getline(line);
            for (int dim : {1, 2}) {
                int count = dim == 1 ? ncv : ns;
                for (int i = 0; i < count && in.getline(line); ++i) {
                    Cursor e{line.data(), line.data() + line.size()};
                    int tag = e.read_int();
                    for (int k = 0; k < 6; ++k) e.read_double();
                    int ntags = e.read_int();
                    std::vector<int> tags((size_t)std::max(ntags, 0));
                    for (int t = 0; t < ntags; ++t) tags[t] = e.read_int();
                    entity_tags[{dim, tag}] = std::move(tags);
                }
            }
            for (int i = 0; i < nv; ++i) in.getline(line);
        } else if (starts_with(line, "$Nodes")) {
            long long nblocks = in.read_ll(), nnodes = in.read_ll();
            in.read_ll();
            in.read_ll();
            in.getline(line);
            xy.reserve(2 * (size_t)std::max(nnodes, 0LL));
            for (long long b = 0; b < nblocks; ++b) {
                in.read_int();
                in.read_int();
                int parametric = in.read_int();
                long long num = in.read_ll();
                in.getline(line);
                if (parametric != 0)
                    return set_error(FVK_ERR_INVALID,
                                     "Error reading mesh: parametric $Nodes coordinates are not supported");
                std::vector<long long> tags((size_t)std::max(num, 0LL));
                for (long long i = 0; i < num; ++i) tags[(size_t)i] = in.read_ll();
                in.getline(line);
                for (long long i = 0; i < num; ++i) {
                    double x = in.read_double(), y = in.read_double();
                    in.read_double();
                    id2i[tags[(size_t)i]] = (int)(xy.size() / 2);
                    xy.push_back(x);
                    xy.push_back(y);
                }
                in.getline(line);
            }
        } else if (starts_with(line, "$Elements")) {
            long long nblocks = in.read_ll();
            in.read_ll();
            in.read_ll();
            in.read_ll();
            in.getline(line);
            for (long long b = 0; b < nblocks; ++b) {
                int edim = in.read_int(), etag = in.read_int(), type = in.read_int();
                long long num = in.read_ll();
                in.getline(line);
                int nn = element_nodes(type);
                int phys = -1;
                auto ent = entity_tags.find({edim, etag});
                if (ent != entity_tags.end() && !ent->second.empty()) phys = ent->second.front();
                for (long long i = 0; i < num && in.getline(line); ++i) {
                    if (nn == 0) continue;
                    Cursor e{line.data(), line.data() + line.size()};
                    long long etg = e.read_ll();
                    Element el{phys, std::vector<int>((size_t)nn)};
                    for (int k = 0; k < nn; ++k) {
                        long long nid = e.read_ll();
                        auto it = id2i.find(nid);
                        if (it == id2i.end())
                            return set_error(FVK_ERR_INVALID,
                                             "Error reading mesh: $Elements entry %lld (entity dim %d tag %d, vertex %d) "
                                             "references undefined node id %lld", etg, edim, etag, k, nid);
                        el.nodes[k] = it->second;
                    }
                    (type == 1 ? lines : cells).push_back(std::move(el));
                }
            }
        }
    }
    return FVK_OK;
}

// MeshReader::read_fvmesh, src/MeshReader.cpp:825-961
int read_fvmesh(Cursor& in, const char* path, std::vector<double>& xy, std::vector<Element>& cells,
                std::vector<Element>& lines, std::map<int, std::string>& names) {
    std::string header, line;
    if (!in.getline(header))
        return set_error(FVK_ERR_INVALID, "Error in mesh file '%s': file is empty (missing '# FVMESH <major>.<minor>' "
                         "header)", path);
    if (!header.empty() && header.back() == '\r') header.pop_back();
    Cursor h{header.data(), header.data() + header.size()};
    std::string hash = h.read_word(), tag = h.read_word(), version = h.read_word();
    int major = -1, minor = -1;
    size_t dot = version.find('.');
    if (hash == "#" && tag == "FVMESH" && dot != std::string::npos) {
        char* e = nullptr;
        std::string a = version.substr(0, dot), b = version.substr(dot + 1);
        long ma = std::strtol(a.c_str(), &e, 10);
        bool ok = e != a.c_str();
        long mi = std::strtol(b.c_str(), &e, 10);
        ok = ok && e != b.c_str();
        if (ok) { major = (int)ma; minor = (int)mi; }
    }
    if (major != 1 || minor != 0)
        return set_error(FVK_ERR_INVALID, "Error in mesh file '%s': unsupported FVMESH version '%s' (supported: 1.0)",
                         path, version.c_str());
    std::unordered_map<std::string, int> name2id;
    int line_number = 1;
    while (in.getline(line)) {
        ++line_number;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#') continue;
        Cursor t{line.data(), line.data() + line.size()};
        std::string sec = t.read_word();
        if (sec == "POINTS") {
            int count = t.read_int();
            xy.assign(2 * (size_t)std::max(count, 0), 0.0);
            for (int i = 0; i < count; ++i) {
                xy[2 * i] = in.read_double();
                xy[2 * i + 1] = in.read_double();
            }
            in.getline(line);
        } else if (sec == "CELLS") {
            int count = t.read_int();
            const long long nn = (long long)(xy.size() / 2);
            for (int i = 0; i < count; ++i) {
                int nv = in.read_int();
                Element el{-1, std::vector<int>((size_t)std::max(nv, 0))};
                for (int k = 0; k < nv; ++k) {
                    int idx = in.read_int();
                    if (idx < 0 || idx >= nn)
                        return set_error(FVK_ERR_INVALID, "Error in mesh file '%s': CELLS entry %d (vertex %d) "
                                         "references out-of-range node index %d", path, i, k, idx);
                    el.nodes[k] = idx;
                }
                cells.push_back(std::move(el));
            }
            in.getline(line);
        } else if (sec == "BOUNDARY") {
            int count = t.read_int();
            const long long nn = (long long)(xy.size() / 2);
            for (int i = 0; i < count; ++i) {
                int a = in.read_int(), b = in.read_int();
                std::string pname = in.read_word();
                if (a < 0 || a >= nn || b < 0 || b >= nn)
                    return set_error(FVK_ERR_INVALID, "Error in mesh file '%s': BOUNDARY entry %d (patch '%s') "
                                     "references out-of-range node index %d", path, i, pname.c_str(),
                                     (a < 0 || a >= nn) ? a : b);
                auto it = name2id.find(pname);
                int pid;
                if (it == name2id.end()) {
                    pid = (int)name2id.size();
                    name2id[pname] = pid;
                    names[pid] = pname;
                } else {
                    pid = it->second;
                }
                lines.push_back(Element{pid, {a, b}});
            }
            in.getline(line);
        } else {
            return set_error(FVK_ERR_INVALID, "Error in mesh file '%s' at line %d: unrecognized section '%s'", path,
                             line_number, sec.c_str());
        }
    }
    if (xy.empty() || cells.empty())
        return set_error(FVK_ERR_INVALID, "Error in mesh file '%s': no POINTS or CELLS sections found", path);
    return FVK_OK;
}

// polygon_signed_area and compute_cell_geometry, src/MeshReader.cpp:32-76
double signed_area(const std::vector<double>& xy, const int* ids, int n) {
    double a = 0.0;
    for (int i = 0; i < n; ++i) {
        const double* p0 = &xy[2 * (size_t)ids[i]];
        const double* p1 = &xy[2 * (size_t)ids[(i + 1) % n]];
        a += p0[0] * p1[1] - p1[0] * p0[1];
    }
    return 0.5 * a;
}

// segments_properly_intersect, src/MeshReader.cpp:91-101
bool cross_properly(const double* a1, const double* a2, const double* b1, const double* b2) {
    auto side = [](const double* o, const double* p, const double* q) {
        return (p[0] - o[0]) * (q[1] - o[1]) - (p[1] - o[1]) * (q[0] - o[0]);
    };
    const double d1 = side(b1, b2, a1), d2 = side(b1, b2, a2), d3 = side(a1, a2, b1), d4 = side(a1, a2, b2);
    return ((d1 > 0.0 && d2 < 0.0) || (d1 < 0.0 && d2 > 0.0)) && ((d3 > 0.0 && d4 < 0.0) || (d3 < 0.0 && d4 > 0.0));
}

// build_cells_faces_patches, src/MeshReader.cpp:317-456, with the hard errors
// of validate_mesh_geometry (:125-296)
int build(fvk_mesh_file* m, const std::vector<Element>& cells, const std::vector<Element>& lines,
          const std::map<int, std::string>& names) {
    const int64_t nc = (int64_t)cells.size();
    const std::vector<double>& xy = m->node_xy;
    m->cell_node_ptr.assign((size_t)nc + 1, 0);
    for (int64_t c = 0; c < nc; ++c) m->cell_node_ptr[c + 1] = m->cell_node_ptr[c] + (int32_t)cells[c].nodes.size();
    const int64_t nh = m->cell_node_ptr[nc];
    m->cell_nodes.resize((size_t)nh);
    for (int64_t c = 0; c < nc; ++c)
        std::copy(cells[c].nodes.begin(), cells[c].nodes.end(), m->cell_nodes.begin() + m->cell_node_ptr[c]);
    // half-edge h = (cell, edge e) in visit order: node cell_nodes[h] -> the next
    std::vector<uint64_t> key((size_t)nh);
    std::vector<int32_t> hcell((size_t)nh);
    for (int64_t c = 0; c < nc; ++c) {
        const int32_t b0 = m->cell_node_ptr[c], n = m->cell_node_ptr[c + 1] - b0;
        for (int e = 0; e < n; ++e) {
            const uint32_t a = (uint32_t)m->cell_nodes[b0 + e], b = (uint32_t)m->cell_nodes[b0 + (e + 1) % n];
            key[b0 + e] = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
            hcell[b0 + e] = (int32_t)c;
        }
    }
    std::vector<int64_t> ord((size_t)nh);
    for (int64_t h = 0; h < nh; ++h) ord[h] = h;
    std::sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return key[x] != key[y] ? key[x] < key[y] : x < y; });
    // the first visit of each edge creates its face; later visits point at it
    std::vector<int64_t> first((size_t)nh);
    std::vector<uint8_t> creates((size_t)nh, 0);
    for (int64_t i = 0; i < nh;) {
        int64_t j = i;
        while (j < nh && key[ord[j]] == key[ord[i]]) first[ord[j++]] = ord[i];
        creates[ord[i]] = 1;
        i = j;
    }
    std::vector<int32_t> face_of_visit((size_t)nh);
    int32_t nf = 0;
    for (int64_t h = 0; h < nh; ++h)
        if (creates[h]) face_of_visit[h] = nf++;
    m->face_nodes.assign(2 * (size_t)nf, 0);
    m->face_cells.assign(2 * (size_t)nf, -1);
    m->face_patch.assign((size_t)nf, -1);
    m->cell_faces.resize((size_t)nh);
    m->cell_ptr = m->cell_node_ptr;
    for (int64_t h = 0; h < nh; ++h) {
        const int32_t c = hcell[h];
        const int32_t b0 = m->cell_node_ptr[c], n = m->cell_node_ptr[c + 1] - b0, e = (int32_t)(h - b0);
        if (creates[h]) {
            const int32_t f = face_of_visit[h];
            m->face_nodes[2 * f] = m->cell_nodes[h];                // this cell's traversal order
            m->face_nodes[2 * f + 1] = m->cell_nodes[b0 + (e + 1) % n];
            m->face_cells[2 * f] = c;
            m->cell_faces[h] = f;
        } else {
            const int32_t f = face_of_visit[first[h]];
            m->face_cells[2 * f + 1] = c;  // a later visit is the neighbour across the face
            m->cell_faces[h] = f;
        }
    }
    // cell geometry, src/MeshReader.cpp:56-76
    m->cell_volume.resize((size_t)nc);
    m->cell_xy.resize(2 * (size_t)nc);
    std::vector<double> sarea((size_t)nc);
    for (int64_t c = 0; c < nc; ++c) {
        const int32_t b0 = m->cell_node_ptr[c], n = m->cell_node_ptr[c + 1] - b0;
        const int* ids = &m->cell_nodes[b0];
        const double area = signed_area(xy, ids, n);
        double cx = 0.0, cy = 0.0;
        for (int i = 0; i < n; ++i) {
            const double* p0 = &xy[2 * (size_t)ids[i]];
            const double* p1 = &xy[2 * (size_t)ids[(i + 1) % n]];
            const double cr = p0[0] * p1[1] - p1[0] * p0[1];
            cx += (p0[0] + p1[0]) * cr;
            cy += (p0[1] + p1[1]) * cr;
        }
        cx /= (6.0 * area);
        cy /= (6.0 * area);
        sarea[c] = area;
        m->cell_volume[c] = std::fabs(area);
        m->cell_xy[2 * c] = cx;
        m->cell_xy[2 * c + 1] = cy;
    }
    // boundary patches from the 1D elements, lazily numbered by first use
    std::map<int, int> phys2patch;
    for (const Element& el : lines) {
        const uint32_t a = (uint32_t)el.nodes[0], b = (uint32_t)el.nodes[1];
        const uint64_t k = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
        auto it = std::lower_bound(ord.begin(), ord.end(), k, [&](int64_t h, uint64_t v) { return key[h] < v; });
        if (it == ord.end() || key[*it] != k) continue;  // not an edge of any cell
        int pidx;
        auto pit = phys2patch.find(el.physical_id);
        if (pit == phys2patch.end()) {
            auto nit = names.find(el.physical_id);
            m->patch_names.push_back(nit != names.end() ? nit->second : "patch_" + std::to_string(el.physical_id));
            pidx = (int)m->patch_names.size() - 1;
            phys2patch[el.physical_id] = pidx;
        } else {
            pidx = pit->second;
        }
        m->face_patch[face_of_visit[*it]] = pidx;
    }
    int untagged = 0;
    for (int32_t f = 0; f < nf; ++f) untagged += m->face_cells[2 * f + 1] == -1 && m->face_patch[f] == -1;
    if (untagged > 0)
        return set_error(FVK_ERR_INVALID, "Error: mesh has %d boundary faces with no assigned patch (untagged boundary "
                         "edges). Every boundary face must belong to a tagged patch.", untagged);
    // hard geometry errors of validate_mesh_geometry
    for (size_t i = 0; i < xy.size(); ++i)
        if (!std::isfinite(xy[i]))
            return set_error(FVK_ERR_INVALID, "Error: mesh node %zu has a non-finite coordinate", i / 2);
    for (int64_t c = 0; c < nc; ++c) {
        if (!std::isfinite(sarea[c]) || sarea[c] <= 0.0)
            return set_error(FVK_ERR_INVALID, "Error: cell %lld has a %s signed volume (%g)", (long long)c,
                             sarea[c] == 0.0 ? "degenerate (zero)" : "negative (or non-finite)", sarea[c]);
        const int32_t b0 = m->cell_node_ptr[c], n = m->cell_node_ptr[c + 1] - b0;
        const int* ids = &m->cell_nodes[b0];
        for (int i = 0; i < n; ++i)
            for (int j = i + 2; j < n; ++j) {
                if (i == 0 && j == n - 1) continue;
                if (cross_properly(&xy[2 * (size_t)ids[i]], &xy[2 * (size_t)ids[(i + 1) % n]], &xy[2 * (size_t)ids[j]],
                                   &xy[2 * (size_t)ids[(j + 1) % n]]))
                    return set_error(FVK_ERR_INVALID, "Error: cell %lld is self-intersecting -- a tangled/bowtie polygon",
                                     (long long)c);
            }
    }
    for (int32_t f = 0; f < nf; ++f) {
        const double* p = &xy[2 * (size_t)m->face_nodes[2 * f]];
        const double* q = &xy[2 * (size_t)m->face_nodes[2 * f + 1]];
        const double dx = q[0] - p[0], dy = q[1] - p[1];
        const double len = std::sqrt(dx * dx + dy * dy);
        if (!std::isfinite(len) || len <= 0.0)
            return set_error(FVK_ERR_INVALID, "Error: face %d has a degenerate (zero or non-finite) length", f);
    }
    return FVK_OK;
}

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

}  // namespace

extern "C" {

int fvk_mesh_read(const char* path, fvk_mesh_file** out) {
    if (!path || !out) return set_error(FVK_ERR_INVALID, "NULL argument");
    *out = nullptr;
    const std::string name(path);
    const bool msh = ends_with(name, ".msh"), fvm = ends_with(name, ".fvmesh");
    if (!msh && !fvm)
        return set_error(FVK_ERR_INVALID, "Error in mesh file '%s': unrecognized extension (expected .msh or .fvmesh)",
                         path);
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return set_error(FVK_ERR_INVALID, "Error: could not open mesh file '%s'", path);
    std::string buf;
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    buf.resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = sz > 0 ? std::fread(&buf[0], 1, (size_t)sz, fp) : 0;
    std::fclose(fp);
    buf.resize(got);
    fvk_mesh_file* m = new (std::nothrow) fvk_mesh_file;
    if (!m) return set_error(FVK_ERR_NOMEM, "host allocation failed");
    std::vector<Element> cells, lines;
    std::map<int, std::string> names;
    int rc = FVK_OK;
    Cursor in{buf.data(), buf.data() + buf.size()};
    if (fvm) {
        rc = read_fvmesh(in, path, m->node_xy, cells, lines, names);
    } else {
        // MeshReader::read_gmsh, src/MeshReader.cpp:784-823: the version from
        // $MeshFormat (2.2 when the data starts without one)
        double version = 2.2;
        std::string line;
        while (in.getline(line)) {
            if (starts_with(line, "$MeshFormat")) {
                version = in.read_double();
                break;
            }
            if (starts_with(line, "$PhysicalNames") || starts_with(line, "$Nodes")) break;
        }
        Cursor again{buf.data(), buf.data() + buf.size()};
        if (version == 2.2) rc = read_gmsh22(again, m->node_xy, cells, lines, names);
        else if (version >= 4.0) rc = read_gmsh41(again, m->node_xy, cells, lines, names);
        else rc = set_error(FVK_ERR_INVALID, "Error in mesh file '%s': unsupported Gmsh format version %g (supported: "
                            "2.2, 4.x ASCII)", path, version);
        if (rc == FVK_OK && (m->node_xy.empty() || cells.empty()))
            rc = set_error(FVK_ERR_INVALID,
                           "Error reading mesh: no nodes or cells were found in the $Nodes/$Elements sections");
    }
    if (rc == FVK_OK && (m->node_xy.size() / 2 > (size_t)INT32_MAX || cells.size() > (size_t)INT32_MAX))
        rc = set_error(FVK_ERR_UNSUPPORTED, "mesh too large for 32-bit indices");
    if (rc == FVK_OK) rc = build(m, cells, lines, names);
    if (rc != FVK_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return FVK_OK;
}

void fvk_mesh_file_destroy(fvk_mesh_file* m) { delete m; }

int fvk_mesh_file_desc(const fvk_mesh_file* m, fvk_mesh_desc* d) {
    if (!m || !d) return set_error(FVK_ERR_INVALID, "NULL argument");
    std::memset(d, 0, sizeof *d);
    d->n_nodes = (int64_t)(m->node_xy.size() / 2);
    d->n_cells = (int64_t)m->cell_volume.size();
    d->n_faces = (int64_t)m->face_patch.size();
    d->n_patches = (int32_t)m->patch_names.size();
    d->node_xy = m->node_xy.data();
    d->face_nodes = m->face_nodes.data();
    d->face_cells = m->face_cells.data();
    d->face_patch = m->face_patch.data();
    d->cell_face_ptr = m->cell_ptr.data();
    d->cell_faces = m->cell_faces.data();
    d->cell_volume = m->cell_volume.data();
    d->cell_xy = m->cell_xy.data();
    d->face_shift = nullptr;
    return FVK_OK;
}

int fvk_mesh_file_cell_nodes(const fvk_mesh_file* m, const int32_t** ptr, const int32_t** nodes) {
    if (!m || !ptr || !nodes) return set_error(FVK_ERR_INVALID, "NULL argument");
    *ptr = m->cell_node_ptr.data();
    *nodes = m->cell_nodes.data();
    return FVK_OK;
}

const char* fvk_mesh_file_patch_name(const fvk_mesh_file* m, int32_t patch) {
    if (!m || patch < 0 || patch >= (int32_t)m->patch_names.size()) return nullptr;
    return m->patch_names[patch].c_str();
}

}  // extern "C"
