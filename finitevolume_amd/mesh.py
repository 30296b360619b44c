"""Unstructured 2D polygon meshes in the reference's layout, plus generators.

The arrays mirror ``UnstructuredMesh`` (include/UnstructuredMesh.h:20-109):
nodes, faces (node1, node2, cell_left, cell_right, patch_id) and cells
(volume, centroid, ordered face list, ordered node list).  Face creation
order, face orientation and each cell's face order follow the rules of the
reference's mesh construction, because they fix the summation order of the
residual and therefore bitwise parity:

* generic polygons -> faces by edge matching, the first cell to walk an edge
  owns it as cell_left with its own node order, the second becomes
  cell_right (src/MeshReader.cpp:317-362);
* structured verification meshes -> faces in the order bottom, left, top
  (last row), right (last column) per cell (src/main.cpp:382-448).

Generators for the benchmark cases (periodic torus, NACA 4-digit O-grid,
triangulated box) use the same rules.  Geometry is numpy; nothing here
touches a GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class UnstructuredMesh:
    node_xy: np.ndarray            # [n_nodes, 2] float64
    face_nodes: np.ndarray         # [n_faces, 2] int32
    face_cells: np.ndarray         # [n_faces, 2] int32 (cell_right = -1 on boundary)
    face_patch: np.ndarray         # [n_faces] int32 (-1 interior)
    cell_face_ptr: np.ndarray      # [n_cells + 1] int32
    cell_faces: np.ndarray         # [nnz] int32
    cell_node_ptr: np.ndarray      # [n_cells + 1] int32
    cell_nodes: np.ndarray         # [nnz] int32
    cell_volume: np.ndarray        # [n_cells] float64
    cell_xy: np.ndarray            # [n_cells, 2] float64
    patches: List[str] = field(default_factory=list)
    face_shift: Optional[np.ndarray] = None  # [n_faces, 2] float64, periodic pairs
    relative_geometry: bool = False          # polygon geometry evaluated relative to each first vertex

    @property
    def n_nodes(self) -> int:
        return int(self.node_xy.shape[0])

    @property
    def n_cells(self) -> int:
        return int(self.cell_volume.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.face_nodes.shape[0])

    def boundary_faces(self) -> np.ndarray:
        return np.nonzero(self.face_cells[:, 1] < 0)[0]

    def faces_of(self, c: int) -> np.ndarray:
        return self.cell_faces[self.cell_face_ptr[c]:self.cell_face_ptr[c + 1]]

    def face_geometry(self):
        """Face midpoint, length and unit normal, computed exactly as
        UnstructuredMesh::compute_geometry (include/UnstructuredMesh.h:89-105):
        normal = (dy, -dx) / |edge| for the edge node1 -> node2."""
        p1 = self.node_xy[self.face_nodes[:, 0]]
        p2 = self.node_xy[self.face_nodes[:, 1]]
        mid = np.empty_like(p1)
        mid[:, 0] = 0.5 * (p1[:, 0] + p2[:, 0])
        mid[:, 1] = 0.5 * (p1[:, 1] + p2[:, 1])
        dx = p2[:, 0] - p1[:, 0]
        dy = p2[:, 1] - p1[:, 1]
        area = np.sqrt(dx * dx + dy * dy)
        nx = dy / area
        ny = -dx / area
        return mid, area, nx, ny

    def with_nodes(self, node_xy: np.ndarray) -> "UnstructuredMesh":
        """The same topology on moved nodes: cell volumes and centroids are
        recomputed by the shoelace formulas over each cell's polygon (face
        geometry follows from the nodes wherever it is evaluated)."""
        node_xy = np.ascontiguousarray(node_xy, dtype=np.float64)
        area, cxy = _polygon_geometry(node_xy, self.cell_node_ptr.astype(np.int64), self.cell_nodes.astype(np.int64),
                                      relative=self.relative_geometry)
        return UnstructuredMesh(node_xy=node_xy, face_nodes=self.face_nodes, face_cells=self.face_cells,
                                face_patch=self.face_patch, cell_face_ptr=self.cell_face_ptr,
                                cell_faces=self.cell_faces, cell_node_ptr=self.cell_node_ptr,
                                cell_nodes=self.cell_nodes, cell_volume=np.abs(area), cell_xy=cxy,
                                patches=list(self.patches), face_shift=self.face_shift,
                                relative_geometry=self.relative_geometry)

    def cell_rings(self, seeds, n_rings: int) -> np.ndarray:
        """Cells within n_rings face-neighbour steps of the seed cells (sorted)."""
        inner = self.face_cells[:, 1] >= 0
        a, b = self.face_cells[inner, 0], self.face_cells[inner, 1]
        mark = np.zeros(self.n_cells, dtype=bool)
        mark[np.asarray(seeds, dtype=np.int64)] = True
        for _ in range(n_rings):
            grow = mark.copy()
            grow[b[mark[a]]] = True
            grow[a[mark[b]]] = True
            mark = grow
        return np.nonzero(mark)[0]

    def submesh(self, cells, cut_patch: int):
        """The polygons of `cells` as a mesh of their own: boundary edges keep
        their patch, edges cut from the rest of the mesh get `cut_patch`.
        Returns (mesh, node map sub -> full)."""
        cells = np.asarray(cells, dtype=np.int64)
        ptr = self.cell_node_ptr.astype(np.int64)
        deg = ptr[cells + 1] - ptr[cells]
        idx = np.concatenate([np.arange(ptr[c], ptr[c + 1]) for c in cells])
        nodes_full = self.cell_nodes[idx].astype(np.int64)
        used, inv = np.unique(nodes_full, return_inverse=True)
        sub_ptr = np.concatenate([[0], np.cumsum(deg)])
        keep = np.zeros(self.n_cells, dtype=bool)
        keep[cells] = True
        fl, fr = self.face_cells[:, 0], self.face_cells[:, 1]
        inside_l = keep[fl]
        inside_r = np.where(fr >= 0, keep[np.maximum(fr, 0)], False)
        bnd = (inside_l & (fr < 0)) | (inside_l ^ inside_r)
        bnd &= inside_l | inside_r
        faces = np.nonzero(bnd)[0]
        patch = np.where(fr[faces] < 0, self.face_patch[faces], cut_patch)
        pos = np.searchsorted(used, self.face_nodes[faces].astype(np.int64))
        sub = from_polygons(self.node_xy[used], sub_ptr, inv, pos, patch,
                            list(self.patches) + ["cut"] * max(0, cut_patch + 1 - len(self.patches)),
                            relative_geometry=self.relative_geometry)
        return sub, used

    def validate(self) -> None:
        """Structural checks (the reference's reader rejects the same
        defects, src/MeshReader.cpp:407-443 and :125-288)."""
        nf, nc = self.n_faces, self.n_cells
        assert self.face_cells.shape == (nf, 2)
        assert np.all(self.face_cells[:, 0] >= 0) and np.all(self.face_cells[:, 0] < nc)
        assert np.all(self.face_cells[:, 1] < nc)
        bnd = self.face_cells[:, 1] < 0
        assert np.all(self.face_patch[bnd] >= 0), "untagged boundary face"
        assert np.all(self.face_patch[bnd] < max(1, len(self.patches)))
        assert np.all(self.cell_volume > 0.0) and np.all(np.isfinite(self.cell_volume))
        _, area, _, _ = self.face_geometry()
        assert np.all(area > 0.0)
        counts = np.bincount(self.cell_faces, minlength=nf)
        expect = np.where(bnd, 1, 2)
        assert np.array_equal(counts, expect), "every face must be listed by its cells"


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------

def _polygon_geometry(node_xy: np.ndarray, ptr: np.ndarray, nodes: np.ndarray,
                      relative: bool = False) -> Tuple[np.ndarray, np.ndarray]:
    """Shoelace signed area and centroid per polygon, summed vertex by vertex
    in the order of src/MeshReader.cpp:32-76 (a += x_i y_{i+1} - x_{i+1} y_i;
    c = sum (x_i + x_{i+1}) cross_i / (6 A); volume = |A|).

    relative=True evaluates the same formula on coordinates taken relative to
    each polygon's first vertex (then adds it back): identical mathematics,
    but robust for cells whose size is many orders below their distance from
    the origin (the 1e-11-chord wall cells of the O-grid)."""
    n_cells = ptr.shape[0] - 1
    deg = np.diff(ptr)
    maxdeg = int(deg.max()) if n_cells else 0
    a = np.zeros(n_cells)
    cx = np.zeros(n_cells)
    cy = np.zeros(n_cells)
    base = node_xy[nodes[ptr[:-1]]] if relative else np.zeros((n_cells, 2))
    for k in range(maxdeg):
        m = deg > k
        idx = np.nonzero(m)[0]
        i0 = nodes[ptr[idx] + k]
        i1 = nodes[ptr[idx] + (k + 1) % deg[idx]]
        x0 = node_xy[i0, 0] - base[idx, 0]
        y0 = node_xy[i0, 1] - base[idx, 1]
        x1 = node_xy[i1, 0] - base[idx, 0]
        y1 = node_xy[i1, 1] - base[idx, 1]
        cross = x0 * y1 - x1 * y0
        a[idx] = a[idx] + cross
        cx[idx] = cx[idx] + (x0 + x1) * cross
        cy[idx] = cy[idx] + (y0 + y1) * cross
    area = 0.5 * a
    cx = cx / (6.0 * area)
    cy = cy / (6.0 * area)
    if relative:
        cx = cx + base[:, 0]
        cy = cy + base[:, 1]
    return area, np.stack([cx, cy], axis=1)


def from_polygons(node_xy: np.ndarray, cell_ptr: np.ndarray, cell_nodes: np.ndarray,
                  boundary_edges: np.ndarray, boundary_patch: np.ndarray,
                  patch_names: Sequence[str], relative_geometry: bool = False) -> UnstructuredMesh:
    """Builds faces from cell polygons by edge matching, with the reference
    reader's rules (src/MeshReader.cpp:317-362): cells are walked in order,
    each cell's edges node[e] -> node[e+1]; the first visitor of an edge
    creates the face (cell_left, its node order), the second visitor becomes
    cell_right; each cell's face list is its edge walk order.  Boundary edges
    then receive their patch ids (:364-394).  Sort-based instead of the
    reference's std::map; the face numbering is identical (faces are numbered
    in order of first visit)."""
    node_xy = np.ascontiguousarray(node_xy, dtype=np.float64)
    cell_ptr = np.asarray(cell_ptr, dtype=np.int64)
    cell_nodes = np.asarray(cell_nodes, dtype=np.int64)
    n_cells = cell_ptr.shape[0] - 1
    deg = np.diff(cell_ptr)
    n_nodes = node_xy.shape[0]
    # edge walk, cell-major: edge k of cell c goes nodes[k] -> nodes[(k+1)%deg]
    e_cell = np.repeat(np.arange(n_cells, dtype=np.int64), deg)
    local = np.arange(cell_nodes.shape[0], dtype=np.int64) - np.repeat(cell_ptr[:-1], deg)
    a = cell_nodes
    b = cell_nodes[cell_ptr[e_cell] + (local + 1) % deg[e_cell]]
    key = np.minimum(a, b) * n_nodes + np.maximum(a, b)
    uniq, first, inverse, counts = np.unique(key, return_index=True, return_inverse=True, return_counts=True)
    if np.any(counts > 2):
        raise ValueError("an edge is shared by more than two cells")
    # faces numbered by order of first visit
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.shape[0])
    face_of_edge = rank[inverse]
    n_faces = uniq.shape[0]
    first_edge = first[order]                      # edge id creating face f
    face_nodes = np.stack([a[first_edge], b[first_edge]], axis=1).astype(np.int32)
    face_cells = np.full((n_faces, 2), -1, dtype=np.int32)
    face_cells[:, 0] = e_cell[first_edge]
    # second visitor
    is_second = np.ones(key.shape[0], dtype=bool)
    is_second[first_edge] = False
    sec = np.nonzero(is_second)[0]
    face_cells[face_of_edge[sec], 1] = e_cell[sec]
    cell_faces = face_of_edge.astype(np.int32)
    # boundary patches
    face_patch = np.full(n_faces, -1, dtype=np.int32)
    if boundary_edges is not None and len(boundary_edges):
        be = np.asarray(boundary_edges, dtype=np.int64)
        bkey = np.minimum(be[:, 0], be[:, 1]) * n_nodes + np.maximum(be[:, 0], be[:, 1])
        pos = np.searchsorted(uniq, bkey)
        ok = (pos < uniq.shape[0]) & (uniq[np.minimum(pos, uniq.shape[0] - 1)] == bkey)
        face_patch[rank[pos[ok]]] = np.asarray(boundary_patch, dtype=np.int32)[ok]
    area, cxy = _polygon_geometry(node_xy, cell_ptr, cell_nodes, relative=relative_geometry)
    mesh = UnstructuredMesh(
        node_xy=node_xy,
        face_nodes=face_nodes,
        face_cells=face_cells,
        face_patch=face_patch,
        cell_face_ptr=cell_ptr.astype(np.int32),
        cell_faces=cell_faces,
        cell_node_ptr=cell_ptr.astype(np.int32),
        cell_nodes=cell_nodes.astype(np.int32),
        cell_volume=np.abs(area),
        cell_xy=cxy,
        patches=list(patch_names),
        relative_geometry=relative_geometry,
    )
    return mesh


# ---------------------------------------------------------------------------
# structured verification meshes (src/main.cpp:382-560)
# ---------------------------------------------------------------------------

PATCH_BOTTOM, PATCH_LEFT, PATCH_TOP, PATCH_RIGHT = 0, 1, 2, 3


def structured(n_x: int, n_y: int, displace: Callable[[np.ndarray, np.ndarray], Tuple[np.ndarray, np.ndarray]],
               periodic: bool = False, lx: float = 1.0, ly: float = 1.0) -> UnstructuredMesh:
    """n_x by n_y quads on the reference square [0,1]^2 mapped by
    ``displace(X, Y) -> (x, y)``; faces per cell in the order bottom, left,
    top (last row), right (last column) with the cell's CCW node order, as
    build_structured_verification_mesh (src/main.cpp:382-448) builds them.
    Patches: 0 bottom, 1 left, 2 top, 3 right.

    periodic=True closes both directions instead (a torus: the bottom face of
    row 0 pairs with the top row, the left face of column 0 with the last
    column; no boundary faces).  face_shift then carries (0, -ly) / (-lx, 0)
    for the wrapping faces; displace must be the identity scaled to lx, ly."""
    hx, hy = 1.0 / n_x, 1.0 / n_y

    def nid(i, j):
        return j * (n_x + 1) + i

    I, J = np.meshgrid(np.arange(n_x + 1), np.arange(n_y + 1))
    X = I.ravel() * hx
    Y = J.ravel() * hy
    xs, ys = displace(X, Y)
    node_xy = np.stack([np.asarray(xs, dtype=np.float64), np.asarray(ys, dtype=np.float64)], axis=1)
    n_cells = n_x * n_y
    cj, ci = np.divmod(np.arange(n_cells), n_x)
    quad = np.stack([nid(ci, cj), nid(ci + 1, cj), nid(ci + 1, cj + 1), nid(ci, cj + 1)], axis=1)
    # shoelace as the structured builder writes it (a2, cx / (3 a2))
    a2 = np.zeros(n_cells)
    cx = np.zeros(n_cells)
    cy = np.zeros(n_cells)
    for k in range(4):
        p1 = node_xy[quad[:, k]]
        p2 = node_xy[quad[:, (k + 1) % 4]]
        cross = p1[:, 0] * p2[:, 1] - p2[:, 0] * p1[:, 1]
        a2 = a2 + cross
        cx = cx + (p1[:, 0] + p2[:, 0]) * cross
        cy = cy + (p1[:, 1] + p2[:, 1]) * cross
    vol = np.abs(a2) * 0.5
    cxy = np.stack([cx / (3.0 * a2), cy / (3.0 * a2)], axis=1)

    fn, fc, fp, fs = [], [], [], []
    cell_lists: List[List[int]] = [[] for _ in range(n_cells)]

    def add(n1, n2, cl, cr, patch, shift=(0.0, 0.0)):
        f = len(fn)
        fn.append((n1, n2))
        fc.append((cl, cr))
        fp.append(patch if cr == -1 else -1)
        fs.append(shift)
        cell_lists[cl].append(f)
        if cr != -1:
            cell_lists[cr].append(f)

    for j in range(n_y):
        for i in range(n_x):
            c = j * n_x + i
            if periodic:
                below = c - n_x if j > 0 else (n_y - 1) * n_x + i
                left = c - 1 if i > 0 else j * n_x + n_x - 1
                add(nid(i, j), nid(i + 1, j), c, below, PATCH_BOTTOM, (0.0, -ly) if j == 0 else (0.0, 0.0))
                add(nid(i, j + 1), nid(i, j), c, left, PATCH_LEFT, (-lx, 0.0) if i == 0 else (0.0, 0.0))
            else:
                add(nid(i, j), nid(i + 1, j), c, c - n_x if j > 0 else -1, PATCH_BOTTOM)
                add(nid(i, j + 1), nid(i, j), c, c - 1 if i > 0 else -1, PATCH_LEFT)
                if j == n_y - 1:
                    add(nid(i + 1, j + 1), nid(i, j + 1), c, -1, PATCH_TOP)
                if i == n_x - 1:
                    add(nid(i + 1, j), nid(i + 1, j + 1), c, -1, PATCH_RIGHT)
    ptr = np.zeros(n_cells + 1, dtype=np.int32)
    ptr[1:] = np.cumsum([len(l) for l in cell_lists])
    mesh = UnstructuredMesh(
        node_xy=node_xy,
        face_nodes=np.asarray(fn, dtype=np.int32),
        face_cells=np.asarray(fc, dtype=np.int32),
        face_patch=np.asarray(fp, dtype=np.int32),
        cell_face_ptr=ptr,
        cell_faces=np.asarray([f for l in cell_lists for f in l], dtype=np.int32),
        cell_node_ptr=np.arange(0, 4 * n_cells + 1, 4, dtype=np.int32),
        cell_nodes=quad.ravel().astype(np.int32),
        cell_volume=vol,
        cell_xy=cxy,
        patches=["bottom", "left", "top", "right"],
        face_shift=np.asarray(fs, dtype=np.float64) if periodic else None,
    )
    return mesh


def skewed(n: int, wobble: float) -> UnstructuredMesh:
    """build_skewed_verification_mesh (src/main.cpp:472-478): X + w sin(pi X) sin(pi Y) on both axes."""
    pi = 3.14159265358979323846

    def disp(X, Y):
        pert = wobble * np.sin(pi * X) * np.sin(pi * Y)
        return X + pert, Y + pert
    return structured(n, n, disp)


def sheared(n_x: int, n_y: int, shear: float) -> UnstructuredMesh:
    """build_sheared_verification_mesh (src/main.cpp:516-520): x = X + shear Y."""
    return structured(n_x, n_y, lambda X, Y: (X + shear * Y, Y))


def cartesian(n_x: int, n_y: int, lx: float = 1.0, ly: float = 1.0) -> UnstructuredMesh:
    return structured(n_x, n_y, lambda X, Y: (X * lx, Y * ly))


def periodic_box(n_x: int, n_y: int, lx: float = 1.0, ly: float = 1.0) -> UnstructuredMesh:
    """Doubly periodic n_x by n_y quads on [0,lx] x [0,ly]: every face interior."""
    return structured(n_x, n_y, lambda X, Y: (X * lx, Y * ly), periodic=True, lx=lx, ly=ly)


def flat_plate(n_x: int, L: float, H: float, first_cell_height: float, growth_ratio: float) -> UnstructuredMesh:
    """build_flat_plate_mesh (src/main.cpp:551-598): uniform in x, geometric
    wall-normal spacing rescaled to land on H."""
    y = [0.0]
    h = first_cell_height
    while y[-1] < H:
        y.append(y[-1] + h)
        h *= growth_ratio
    rescale = H / y[-1]
    y = np.asarray(y) * rescale
    n_y = len(y) - 1

    def disp(X, Y):
        j = np.rint(Y * n_y).astype(np.int64)
        return X * L, y[j]
    return structured(n_x, n_y, disp)


# ---------------------------------------------------------------------------
# benchmark meshes
# ---------------------------------------------------------------------------

def tri_box(n_x: int, n_y: int, lx: float = 1.0, ly: float = 1.0) -> UnstructuredMesh:
    """n_x by n_y quads on [0,lx]x[0,ly], each split along the (i,j)-(i+1,j+1)
    diagonal: 2 n_x n_y triangles.  One patch per side: 0 bottom, 1 right,
    2 top, 3 left (boundary edges listed in that order)."""
    I, J = np.meshgrid(np.arange(n_x + 1), np.arange(n_y + 1))
    node_xy = np.stack([I.ravel() * (lx / n_x), J.ravel() * (ly / n_y)], axis=1).astype(np.float64)

    def nid(i, j):
        return j * (n_x + 1) + i
    cj, ci = np.divmod(np.arange(n_x * n_y), n_x)
    a, b, c, d = nid(ci, cj), nid(ci + 1, cj), nid(ci + 1, cj + 1), nid(ci, cj + 1)
    tris = np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], axis=1).reshape(-1, 3)
    ptr = np.arange(0, 3 * tris.shape[0] + 1, 3)
    i = np.arange(n_x)
    jj = np.arange(n_y)
    edges = [np.stack([nid(i, 0), nid(i + 1, 0)], 1), np.stack([nid(n_x, jj), nid(n_x, jj + 1)], 1),
             np.stack([nid(i + 1, n_y), nid(i, n_y)], 1), np.stack([nid(0, jj + 1), nid(0, jj)], 1)]
    patches = np.concatenate([np.full(e.shape[0], k) for k, e in enumerate(edges)])
    return from_polygons(node_xy, ptr, tris.ravel(), np.concatenate(edges), patches,
                         ["bottom", "right", "top", "left"])


def naca4_surface(code: str, x: np.ndarray, upper: np.ndarray, closed_te: bool = True) -> Tuple[np.ndarray, np.ndarray]:
    """NACA 4-digit section (chord 1, leading edge at the origin).  Thickness
    0.2969 sqrt(x) - 0.1260 x - 0.3516 x^2 + 0.2843 x^3 - 0.1036 x^4 (closed
    trailing edge; -0.1015 for the open one); camber m, p from the code."""
    m = int(code[0]) / 100.0
    p = int(code[1]) / 10.0
    t = int(code[2:]) / 100.0
    a4 = -0.1036 if closed_te else -0.1015
    yt = 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 + a4 * x ** 4)
    if m == 0.0:
        return x.copy(), np.where(upper, yt, -yt)
    yc = np.where(x < p, m / p ** 2 * (2 * p * x - x ** 2), m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * x - x ** 2))
    dyc = np.where(x < p, 2 * m / p ** 2 * (p - x), 2 * m / (1 - p) ** 2 * (p - x))
    th = np.arctan(dyc)
    sgn = np.where(upper, 1.0, -1.0)
    return x - sgn * yt * np.sin(th), yc + sgn * yt * np.cos(th)


def radial_spacing(n_layers: int, r_far: float, ratio: float, first: Optional[float] = None) -> np.ndarray:
    """Normalised wall-normal stations s_0 = 0 < ... < s_n = 1 with geometric
    growth ``ratio``.  With first=None the first spacing is whatever makes
    n_layers geometric layers sum to r_far (for 512 layers at 1.05 to 50
    chords: 3.5e-11 chord).  With ``first`` given, growth is geometric from
    that spacing and capped (constant spacing) once the remaining layers
    fill the distance."""
    if first is None:
        h = np.power(ratio, np.arange(n_layers, dtype=np.float64))
    else:
        h = np.empty(n_layers)
        cur, done = first, 0.0
        for k in range(n_layers):
            remaining = n_layers - k
            if done + cur * remaining >= r_far:
                h[k:] = (r_far - done) / remaining
                break
            h[k] = cur
            done += cur
            cur *= ratio
    s = np.concatenate([[0.0], np.cumsum(h)])
    return s / s[-1]


def naca_ogrid(n_surface: int = 1024, n_layers: int = 512, code: str = "0012", r_far: float = 50.0,
               ratio: float = 1.05, first: Optional[float] = None, triangles: bool = True,
               quad_layers: int = 0) -> UnstructuredMesh:
    """O-grid around a NACA 4-digit section: n_surface cosine-clustered
    surface points (trailing edge at parameter 0), n_layers layers out to a
    circle of radius r_far chords around mid-chord, geometric wall-normal
    growth (see radial_spacing).  triangles=True splits every quad on the
    same diagonal: 2 n_surface n_layers triangles (with quad_layers > 0 the
    first quad_layers layers stay quads).  Patches: 0 wall,
    1 farfield.  Cells are numbered layer by layer from the wall."""
    th = 2.0 * np.pi * np.arange(n_surface) / n_surface
    xs = 0.5 * (1.0 + np.cos(th))
    upper = th <= np.pi
    sx, sy = naca4_surface(code, xs, upper)
    sy = np.where(np.isclose(th, 0.0) | np.isclose(th, np.pi), 0.0 if code[0] == "0" else sy, sy)
    ox = 0.5 + r_far * np.cos(th)
    oy = r_far * np.sin(th)
    s = radial_spacing(n_layers, r_far, ratio, first)
    # node (i, j) -> j * n_surface + i
    node_x = sx[None, :] + s[:, None] * (ox - sx)[None, :]
    node_y = sy[None, :] + s[:, None] * (oy - sy)[None, :]
    node_xy = np.stack([node_x.ravel(), node_y.ravel()], axis=1)

    def nid(i, j):
        return j * n_surface + (i % n_surface)
    cj, ci = np.divmod(np.arange(n_surface * n_layers), n_surface)
    # i runs trailing edge -> upper -> leading edge -> lower (clockwise seen
    # from outside along j); (i,j),(i,j+1),(i+1,j+1),(i+1,j) is counter-clockwise
    a, b, c, d = nid(ci, cj), nid(ci, cj + 1), nid(ci + 1, cj + 1), nid(ci + 1, cj)
    if triangles and quad_layers > 0:
        # hybrid: quads in the first quad_layers layers, triangle pairs beyond
        quad = cj < quad_layers
        polys = []
        for k in range(n_surface * n_layers):
            if quad[k]:
                polys.append((a[k], b[k], c[k], d[k]))
            else:
                polys.append((a[k], b[k], c[k]))
                polys.append((a[k], c[k], d[k]))
        deg = np.fromiter((len(p) for p in polys), dtype=np.int64, count=len(polys))
        cells = np.fromiter((v for p in polys for v in p), dtype=np.int64, count=int(deg.sum()))
        ptr = np.concatenate([[0], np.cumsum(deg)])
    elif triangles:
        cells = np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], axis=1).reshape(-1, 3)
        ptr = np.arange(0, 3 * cells.shape[0] + 1, 3)
    else:
        cells = np.stack([a, b, c, d], 1)
        ptr = np.arange(0, 4 * cells.shape[0] + 1, 4)
    i = np.arange(n_surface)
    wall = np.stack([nid(i + 1, 0), nid(i, 0)], 1)
    far = np.stack([nid(i, n_layers), nid(i + 1, n_layers)], 1)
    patches = np.concatenate([np.zeros(n_surface, int), np.ones(n_surface, int)])
    mesh = from_polygons(node_xy, ptr, cells.ravel(), np.concatenate([wall, far]), patches,
                         ["wall", "farfield"], relative_geometry=True)
    return mesh


# ---------------------------------------------------------------------------
# initial conditions
# ---------------------------------------------------------------------------

def from_primitive(rho, u, v, p, gamma):
    """EulerState from_primitive (include/EulerState.h:45-52), vectorised."""
    rho = np.asarray(rho, dtype=np.float64)
    out = np.empty(rho.shape + (4,))
    out[..., 0] = rho
    out[..., 1] = rho * u
    out[..., 2] = rho * v
    out[..., 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return out


def isentropic_vortex(xy: np.ndarray, gamma: float = 1.4, mach: float = 0.5, angle_deg: float = 45.0,
                      beta: float = 5.0, radius: float = 0.1, center=(0.5, 0.5)) -> np.ndarray:
    """Isentropic vortex on a uniform stream (rho_inf = p_inf = 1, T = p/rho):
    f = (1 - r^2/R^2)/2, du = -(beta/2pi)((y-y0)/R) e^f, dv = (beta/2pi)((x-x0)/R) e^f,
    dT = -(gamma-1) beta^2/(8 gamma pi^2) e^{2f}, rho = (1+dT)^{1/(gamma-1)}, p = rho^gamma.
    Returns AoS conserved states [n, 4]."""
    c_inf = math.sqrt(gamma)
    a = math.radians(angle_deg)
    u_inf = mach * c_inf * math.cos(a)
    v_inf = mach * c_inf * math.sin(a)
    dx = xy[:, 0] - center[0]
    dy = xy[:, 1] - center[1]
    f = 0.5 * (1.0 - (dx * dx + dy * dy) / (radius * radius))
    ef = np.exp(f)
    du = -(beta / (2.0 * math.pi)) * (dy / radius) * ef
    dv = (beta / (2.0 * math.pi)) * (dx / radius) * ef
    dT = -(gamma - 1.0) * beta * beta / (8.0 * gamma * math.pi ** 2) * ef * ef
    rho = np.power(1.0 + dT, 1.0 / (gamma - 1.0))
    p = np.power(rho, gamma)
    return from_primitive(rho, u_inf + du, v_inf + dv, p, gamma)
