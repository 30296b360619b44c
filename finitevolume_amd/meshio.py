"""Mesh files: the library's reader (fvk_mesh_read: Gmsh ASCII 2.2 / 4.x and
FVMESH 1.0, with the reference's face order, src/MeshReader.cpp) and
writers of the same three formats for meshes built in Python.

The writers follow the layouts the reference reads (docs/fvmesh-format.md;
FvMeshWriter::write, src/FvMeshWriter.cpp, for FVMESH): cells with their
node lists in order, and one boundary edge per boundary face, named by the
face's patch.  Gmsh files carry triangles and quadrangles only.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

from . import abi
from .mesh import UnstructuredMesh


def _lib():
    from .solver import lib
    L = lib()
    if not getattr(L, "_meshio_bound", False):
        vp, ip = C.c_void_p, C.POINTER(C.c_int32)
        L.fvk_mesh_read.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.fvk_mesh_file_destroy.argtypes = [vp]
        L.fvk_mesh_file_destroy.restype = None
        L.fvk_mesh_file_desc.argtypes = [vp, C.POINTER(abi.MeshDescC)]
        L.fvk_mesh_file_cell_nodes.argtypes = [vp, C.POINTER(ip), C.POINTER(ip)]
        L.fvk_mesh_file_patch_name.argtypes = [vp, C.c_int32]
        L.fvk_mesh_file_patch_name.restype = C.c_char_p
        L._meshio_bound = True
    return L


def read_mesh(path: str) -> UnstructuredMesh:
    """MeshReader::read through the library: the mesh with the reference's
    faces, Cell::faces order, patch names and cell geometry."""
    from .solver import _check
    L = _lib()
    h = C.c_void_p()
    _check(L.fvk_mesh_read(os.fsencode(path), C.byref(h)))
    try:
        d = abi.MeshDescC()
        _check(L.fvk_mesh_file_desc(h, C.byref(d)))
        nn, nc, nf = d.n_nodes, d.n_cells, d.n_faces
        ptr, nodes = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
        _check(L.fvk_mesh_file_cell_nodes(h, C.byref(ptr), C.byref(nodes)))

        def arr(p, n, dt):
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True) if n > 0 else np.zeros(0, dt)

        cptr = arr(d.cell_face_ptr, nc + 1, np.int32)
        nnz = int(cptr[-1])
        m = UnstructuredMesh(
            node_xy=arr(d.node_xy, 2 * nn, np.float64).reshape(nn, 2),
            face_nodes=arr(d.face_nodes, 2 * nf, np.int32).reshape(nf, 2),
            face_cells=arr(d.face_cells, 2 * nf, np.int32).reshape(nf, 2),
            face_patch=arr(d.face_patch, nf, np.int32),
            cell_face_ptr=cptr,
            cell_faces=arr(d.cell_faces, nnz, np.int32),
            cell_node_ptr=arr(ptr, nc + 1, np.int32),
            cell_nodes=arr(nodes, nnz, np.int32),
            cell_volume=arr(d.cell_volume, nc, np.float64),
            cell_xy=arr(d.cell_xy, 2 * nc, np.float64).reshape(nc, 2),
            patches=[L.fvk_mesh_file_patch_name(h, i).decode() for i in range(d.n_patches)],
        )
    finally:
        L.fvk_mesh_file_destroy(h)
    return m


def _patch_names(mesh: UnstructuredMesh) -> Sequence[str]:
    n = int(mesh.face_patch.max()) + 1 if mesh.n_faces else 0
    names = list(mesh.patches) + [f"patch_{i}" for i in range(len(mesh.patches), n)]
    return [nm.replace(" ", "_") for nm in names]


def _boundary(mesh: UnstructuredMesh):
    b = np.nonzero((mesh.face_cells[:, 1] < 0) & (mesh.face_patch >= 0))[0]
    return b, mesh.face_nodes[b], mesh.face_patch[b]


def write_fvmesh(path: str, mesh: UnstructuredMesh, precision: int = 17) -> None:
    """FVMESH 1.0 (docs/fvmesh-format.md), laid out as FvMeshWriter::write
    lays it out: POINTS, CELLS, BOUNDARY (one line per tagged boundary face)."""
    names = _patch_names(mesh)
    fmt = f"%.{precision}g"
    with open(path, "w") as f:
        f.write("# FVMESH 1.0\n")
        f.write(f"POINTS {mesh.n_nodes}\n")
        f.writelines(f"{fmt % x} {fmt % y}\n" for x, y in mesh.node_xy)
        f.write(f"CELLS {mesh.n_cells}\n")
        p, nd = mesh.cell_node_ptr, mesh.cell_nodes
        f.writelines(f"{p[c + 1] - p[c]} " + " ".join(map(str, nd[p[c]:p[c + 1]])) + "\n" for c in range(mesh.n_cells))
        _, fn, fp = _boundary(mesh)
        f.write(f"BOUNDARY {len(fp)}\n")
        f.writelines(f"{a} {b} {names[q]}\n" for (a, b), q in zip(fn, fp))


def _gmsh_elements(mesh):
    p = mesh.cell_node_ptr
    nv = np.diff(p)
    if np.any((nv != 3) & (nv != 4)):
        raise ValueError("Gmsh files carry triangles and quadrangles only")
    return p, nv


def write_gmsh22(path: str, mesh: UnstructuredMesh, precision: int = 17) -> None:
    """Gmsh ASCII 2.2: physical curve groups 1.. per patch (named), the
    boundary edges as 2-node lines, the cells as triangles/quadrangles in
    physical surface group 1000; node ids 1-based."""
    names = _patch_names(mesh)
    p, nv = _gmsh_elements(mesh)
    _, fn, fp = _boundary(mesh)
    fmt = f"%.{precision}g"
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
        f.write(f"$PhysicalNames\n{len(names) + 1}\n")
        f.writelines(f'1 {i + 1} "{nm}"\n' for i, nm in enumerate(names))
        f.write('2 1000 "fluid"\n$EndPhysicalNames\n')
        f.write(f"$Nodes\n{mesh.n_nodes}\n")
        f.writelines(f"{i + 1} {fmt % x} {fmt % y} 0\n" for i, (x, y) in enumerate(mesh.node_xy))
        f.write("$EndNodes\n")
        f.write(f"$Elements\n{len(fp) + mesh.n_cells}\n")
        k = 1
        for (a, b), q in zip(fn, fp):
            f.write(f"{k} 1 2 {q + 1} {q + 1} {a + 1} {b + 1}\n")
            k += 1
        for c in range(mesh.n_cells):
            ids = " ".join(str(i + 1) for i in mesh.cell_nodes[p[c]:p[c + 1]])
            f.write(f"{k} {2 if nv[c] == 3 else 3} 2 1000 1 {ids}\n")
            k += 1
        f.write("$EndElements\n")


def write_gmsh41(path: str, mesh: UnstructuredMesh, precision: int = 17) -> None:
    """Gmsh ASCII 4.1: one curve entity per patch carrying its physical
    group, one surface entity for the cells; the nodes in one block
    (parametric 0); the elements in blocks by entity and type."""
    names = _patch_names(mesh)
    p, nv = _gmsh_elements(mesh)
    _, fn, fp = _boundary(mesh)
    fmt = f"%.{precision}g"
    npatch = len(names)
    lo, hi = mesh.node_xy.min(axis=0), mesh.node_xy.max(axis=0)
    box = f"{fmt % lo[0]} {fmt % lo[1]} 0 {fmt % hi[0]} {fmt % hi[1]} 0"
    with open(path, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
        f.write(f"$PhysicalNames\n{npatch + 1}\n")
        f.writelines(f'1 {i + 1} "{nm}"\n' for i, nm in enumerate(names))
        f.write('2 1000 "fluid"\n$EndPhysicalNames\n')
        f.write(f"$Entities\n0 {npatch} 1 0\n")
        f.writelines(f"{i + 1} {box} 1 {i + 1} 0\n" for i in range(npatch))
        f.write(f"1 {box} 1 1000 0\n$EndEntities\n")
        n = mesh.n_nodes
        f.write(f"$Nodes\n1 {n} 1 {n}\n2 1 0 {n}\n")
        f.writelines(f"{i + 1}\n" for i in range(n))
        f.writelines(f"{fmt % x} {fmt % y} 0\n" for x, y in mesh.node_xy)
        f.write("$EndNodes\n")
        # cells in mesh order: one block per run of cells of one type
        runs = []
        for c in range(mesh.n_cells):
            if runs and nv[runs[-1][-1]] == nv[c]:
                runs[-1].append(c)
            else:
                runs.append([c])
        blocks = [q for q in range(npatch) if np.any(fp == q)]
        nblk = len(blocks) + len(runs)
        f.write(f"$Elements\n{nblk} {len(fp) + mesh.n_cells} 1 {len(fp) + mesh.n_cells}\n")
        k = 1
        for q in blocks:
            sel = np.nonzero(fp == q)[0]
            f.write(f"1 {q + 1} 1 {len(sel)}\n")
            for j in sel:
                f.write(f"{k} {fn[j, 0] + 1} {fn[j, 1] + 1}\n")
                k += 1
        for run in runs:
            f.write(f"2 1 {2 if nv[run[0]] == 3 else 3} {len(run)}\n")
            for c in run:
                f.write(f"{k} " + " ".join(str(i + 1) for i in mesh.cell_nodes[p[c]:p[c + 1]]) + "\n")
                k += 1
        f.write("$EndElements\n")
