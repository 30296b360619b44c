"""Voronoi meshes around a NACA 0012 section for the batched-airfoil workload.

Case k: N_k = exp(U[ln 2e4, ln 4e5]) cells, Mach ~ U[0.3, 0.85], angle of
attack ~ U[-4, 6] degrees, all drawn from numpy's default_rng(1000 + k).
Points: Bridson Poisson-disk samples in the box [-5, 5]^2 (radius chosen for
N_k; N ~ area / (1.597 r^2) measured for this sampler) outside the section
(chord 1, leading edge at the origin), seeded with a row of points half a
radius off the surface.  The mesh is the Voronoi diagram of those points
(scipy.spatial.Voronoi), bounded by two kinds of generator that are not
cells: points half a radius inside the section, staggered half a spacing
against the outer row (their bisectors with the outer row form the wall),
and a ring of points half a radius outside the box (their bisectors form
the farfield boundary).  So both boundaries are polylines of Voronoi edges
that follow the section and the box to within a fraction of the radius.
Cells with more than 9 faces (about one in 1e3) lose their shortest edge,
whose two vertices merge, until none is left; a sample with a cell of fewer
than 3 faces would be redrawn with the next seed of the stream.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, List, Tuple

import numpy as np

from . import mesh as M

DENSITY = 1.597  # area / (N r^2) of the Poisson-disk sampler
HALF_WIDTH = 5.0


def _sampler():
    from .solver import lib
    L = lib()
    f = L.fvk_poisson_disk
    f.argtypes = [C.c_double] * 5 + [C.c_uint64, C.c_void_p, C.c_int64, C.c_double, C.c_double, C.c_void_p,
                                     C.c_int64, C.POINTER(C.c_int64)]
    f.restype = C.c_int
    return f


def batch_case_params(k: int) -> Dict[str, float]:
    """Cell count, Mach number and angle of attack of case k."""
    rng = np.random.default_rng(1000 + k)
    n = float(np.exp(rng.uniform(math.log(2e4), math.log(4e5))))
    mach = float(rng.uniform(0.3, 0.85))
    aoa = float(rng.uniform(-4.0, 6.0))
    return {"n_target": n, "mach": mach, "aoa": aoa, "seed": 1000 + k}


def _surface(n: int, thickness: float) -> np.ndarray:
    """Closed NACA 00tt polyline, counter-clockwise (upper surface from the
    trailing edge to the leading edge, then the lower surface back)."""
    beta = np.linspace(0.0, math.pi, n)
    x = 0.5 * (1.0 - np.cos(beta))
    code = "00%02d" % int(round(thickness * 100))
    xu, yu = M.naca4_surface(code, x[::-1], np.ones(n, bool))
    xl, yl = M.naca4_surface(code, x[1:-1], np.zeros(n - 2, bool))
    return np.concatenate([np.stack([xu, yu], 1), np.stack([xl, yl], 1)])


def _resample(poly: np.ndarray, spacing: float, offset: float) -> Tuple[np.ndarray, np.ndarray]:
    """Points at arc length (j + offset) * h around the closed polyline (h the
    spacing rounded to divide the perimeter) and the outward normals there."""
    seg = np.roll(poly, -1, axis=0) - poly
    ln = np.hypot(seg[:, 0], seg[:, 1])
    cum = np.concatenate([[0.0], np.cumsum(ln)])
    total = cum[-1]
    m = max(8, int(round(total / spacing)))
    h = total / m
    s = (np.arange(m) + offset) * h
    i = np.clip(np.searchsorted(cum, s, side="right") - 1, 0, len(ln) - 1)
    t = (s - cum[i]) / ln[i]
    pts = poly[i] + t[:, None] * seg[i]
    tang = seg[i] / ln[i][:, None]
    normal = np.stack([tang[:, 1], -tang[:, 0]], 1)  # outward for a counter-clockwise polyline
    return pts, normal


def _inside(x: np.ndarray, y: np.ndarray, thickness: float, margin: float) -> np.ndarray:
    xc = np.clip(x, 0.0, 1.0)
    yt = 5.0 * thickness * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 + 0.2843 * xc ** 3 -
                            0.1036 * xc ** 4)
    return (x > 0.0) & (x < 1.0) & (np.abs(y) < yt - margin)


def _collapse_short_edges(verts: np.ndarray, cells: np.ndarray, ptr: np.ndarray, bverts: np.ndarray,
                          max_faces: int, rounds: int = 6):
    """Cells with more than max_faces edges lose their shortest edge: its two
    vertices merge at their midpoint (everywhere, so the neighbours stay
    conforming and the face between the two cells across that edge
    disappears).  Repeated until no cell exceeds max_faces.  Returns the
    updated polygons, boundary edges and whether every cell kept at least
    three edges."""
    for _ in range(rounds):
        deg = np.diff(ptr)
        big = np.nonzero(deg > max_faces)[0]
        if big.size == 0:
            return cells, ptr, bverts, bool(deg.min() >= 3)
        parent = np.arange(len(verts))
        for c in big:
            ring = cells[ptr[c]:ptr[c + 1]]
            p = verts[ring]
            ln = np.hypot(*(np.roll(p, -1, axis=0) - p).T)
            j = int(np.argmin(ln))
            va, vb = ring[j], ring[(j + 1) % len(ring)]
            while parent[va] != va:
                va = parent[va]
            while parent[vb] != vb:
                vb = parent[vb]
            if va != vb:
                verts[va] = 0.5 * (verts[va] + verts[vb])
                parent[vb] = va
        for _i in range(8):  # flatten the merge chains
            parent = parent[parent]
        cells = parent[cells]
        bverts = parent[bverts]
        # drop repeated consecutive vertices (cyclically) in each polygon
        cid = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
        nxt = np.arange(len(cells)) + 1
        nxt[ptr[1:] - 1] = ptr[:-1]
        keep = cells != cells[nxt]
        cells = cells[keep]
        ptr = np.concatenate([[0], np.cumsum(np.bincount(cid[keep], minlength=len(ptr) - 1))])
    deg = np.diff(ptr)
    return cells, ptr, bverts, bool(deg.min() >= 3 and deg.max() <= max_faces)


def naca_voronoi(n_target: float, seed: int, thickness: float = 0.12, max_faces: int = 9,
                 attempts: int = 8) -> M.UnstructuredMesh:
    from scipy.spatial import Voronoi
    hw = HALF_WIDTH
    area = (2 * hw) ** 2 - 0.6851 * thickness
    r = math.sqrt(area / (DENSITY * n_target))
    poly = _surface(1200, thickness)
    outer, nrm = _resample(poly, r, 0.0)
    outer = outer + 0.5 * r * nrm
    inner, nrm_i = _resample(poly, r, 0.5)
    inner = inner - 0.5 * r * nrm_i
    # inner generators must stay inside the section: where it is thinner than
    # a radius they move to the chord line (duplicates dropped)
    bad = ~_inside(inner[:, 0], inner[:, 1], thickness, 0.05 * r)
    inner[bad, 1] = 0.0
    inner[:, 0] = np.clip(inner[:, 0], 0.25 * r, 1.0 - 0.25 * r)
    keep_i = np.ones(len(inner), bool)
    order = np.argsort(inner[:, 0])
    last = -np.inf
    for j in order:
        if bad[j]:
            if inner[j, 0] - last < 0.5 * r:
                keep_i[j] = False
            else:
                last = inner[j, 0]
    inner = inner[keep_i]
    # ghost ring just outside the box
    g = np.arange(-hw - 1.5 * r, hw + 1.5 * r + 1e-12, r)
    ring = []
    for d in (0.5 * r, 1.5 * r):
        e = hw + d
        ring += [np.stack([g, np.full_like(g, -e)], 1), np.stack([g, np.full_like(g, e)], 1),
                 np.stack([np.full_like(g, -e), g], 1), np.stack([np.full_like(g, e), g], 1)]
    ghost = np.unique(np.round(np.concatenate(ring), 12), axis=0)
    sample = _sampler()
    for attempt in range(attempts):
        cap = int(1.5 * n_target) + 4 * len(outer) + 1024
        out = np.empty(2 * cap)
        n = C.c_int64()
        seeds = np.ascontiguousarray(outer, dtype=np.float64)
        rc = sample(-hw, -hw, hw, hw, r, seed * 1000003 + attempt, seeds.ctypes.data, len(seeds), thickness,
                    0.5 * r, out.ctypes.data, cap, C.byref(n))
        if rc != 0:
            raise RuntimeError(f"Poisson-disk sampling failed ({rc})")
        pts = out[: 2 * n.value].reshape(-1, 2)
        pts = pts[(np.abs(pts[:, 0]) < hw) & (np.abs(pts[:, 1]) < hw)]
        n_kept = len(pts)
        allp = np.concatenate([pts, inner, ghost])
        vor = Voronoi(allp)
        regions = [vor.regions[vor.point_region[i]] for i in range(n_kept)]
        deg = np.fromiter((len(rg) for rg in regions), dtype=np.int64, count=n_kept)
        if any(-1 in rg for rg in regions) or deg.min() < 3:
            continue
        cells = np.fromiter((v for rg in regions for v in rg), dtype=np.int64, count=int(deg.sum()))
        ptr = np.concatenate([[0], np.cumsum(deg)])
        verts = vor.vertices.copy()
        rp = vor.ridge_points
        one = (rp[:, 0] < n_kept) ^ (rp[:, 1] < n_kept)
        bverts = np.array([vor.ridge_vertices[i] for i in np.nonzero(one)[0]], dtype=np.int64)
        cells, ptr, bverts, ok = _collapse_short_edges(verts, cells, ptr, bverts, max_faces)
        if not ok:
            continue
        deg = np.diff(ptr)
        used, local = np.unique(cells, return_inverse=True)
        node_xy = verts[used]
        # counter-clockwise polygons
        x = node_xy[local, 0]
        y = node_xy[local, 1]
        cid = np.repeat(np.arange(n_kept), deg)
        nxt = np.arange(len(local)) + 1
        nxt[ptr[1:] - 1] = ptr[:-1]
        cross = np.bincount(cid, weights=x * y[nxt] - x[nxt] * y, minlength=n_kept)
        flip = cross < 0.0
        if np.any(flip):
            for c in np.nonzero(flip)[0]:
                local[ptr[c]:ptr[c + 1]] = local[ptr[c]:ptr[c + 1]][::-1]
        # boundary edges from the generator pairs (a kept cell and a non-cell generator)
        a, b = rp[one, 0], rp[one, 1]
        other = np.where(a < n_kept, b, a)
        patch = np.where(other < n_kept + len(inner), 0, 1)
        live = bverts[:, 0] != bverts[:, 1]
        pos = np.searchsorted(used, bverts[live])
        return M.from_polygons(node_xy, ptr, local, pos, patch[live], ["wall", "farfield"], relative_geometry=True)
    raise RuntimeError(f"no valid Voronoi mesh after {attempts} samples (seed {seed})")
