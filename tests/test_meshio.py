"""Mesh ingest without a GPU: the library's readers (fvk_mesh_read, host
code) against the reference's MeshReader::read (src/MeshReader.cpp), on
FVMESH, Gmsh 2.2 and Gmsh 4.1 files written from this project's mesh
generators: the same faces in the same order and orientation, the same
Cell::faces, patches and cell geometry, bit for bit; and the same
rejections.  Files the reference read are kept as fixtures
(tests/golden/meshes) with its arrays (tests/golden/mesh_reader.npz), so
the comparison runs where the reference build is absent."""
import os

import numpy as np
import pytest

from finitevolume_amd import mesh as M
from finitevolume_amd.build import LIB
from oracle import pyref

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")
FIELDS = ("node_xy", "face_nodes", "face_cells", "face_patch", "cell_face_ptr", "cell_faces", "cell_nodes",
          "cell_volume", "cell_xy")
FILES = ("voronoi.fvmesh", "naca_tri.msh", "naca_mixed_v41.msh", "wavy_quads.fvmesh", "wavy_quads_v22.msh")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libfvk.so is not built")


def ours(path):
    from finitevolume_amd.meshio import read_mesh
    m = read_mesh(path)
    return dict(node_xy=m.node_xy, face_nodes=m.face_nodes, face_cells=m.face_cells, face_patch=m.face_patch,
                cell_face_ptr=m.cell_face_ptr, cell_faces=m.cell_faces, cell_nodes=m.cell_nodes,
                cell_volume=m.cell_volume, cell_xy=m.cell_xy, patches=list(m.patches))


def same(a, b, label):
    for k in FIELDS:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (label, k)
    assert list(a["patches"]) == list(b["patches"]), label


@pytest.mark.parametrize("name", FILES)
def test_reader_matches_reference_fixture(name):
    g = np.load(os.path.join(HERE, "golden", "mesh_reader.npz"))
    key = name.replace(".", "_")
    want = {k: g[f"{key}_{k}"] for k in FIELDS}
    want["patches"] = [str(x) for x in g[f"{key}_patches"]]
    same(ours(os.path.join(MESHES, name)), want, name)


@pytest.mark.skipif(not pyref.available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name", FILES)
def test_reader_matches_reference_live(name):
    path = os.path.join(MESHES, name)
    same(ours(path), pyref.read_mesh(path), name)


@pytest.mark.skipif(not pyref.available(), reason="reference build (oracle/_ref) not present")
def test_round_trip_and_rejections(tmp_path):
    from finitevolume_amd import meshio
    from finitevolume_amd.solver import FVKError
    m = M.naca_ogrid(24, 6, quad_layers=2)
    for w, ext in ((meshio.write_fvmesh, ".fvmesh"), (meshio.write_gmsh22, ".msh"), (meshio.write_gmsh41, ".msh")):
        p = str(tmp_path / ("m" + ext))
        w(p, m)
        a, b = ours(p), pyref.read_mesh(p)
        same(a, b, w.__name__)
        # the written mesh is the generator's (node order, cell node lists)
        assert np.array_equal(a["node_xy"], m.node_xy) and np.array_equal(a["cell_nodes"], m.cell_nodes)
    # rejected by both readers: an untagged boundary edge, an inverted cell,
    # an out-of-range node, an unknown section, a bad version
    base = open(str(tmp_path / "m.fvmesh")).read().splitlines()
    nb = next(i for i, ln in enumerate(base) if ln.startswith("BOUNDARY"))
    cells0 = next(i for i, ln in enumerate(base) if ln.startswith("CELLS")) + 1
    inv = base[cells0].split()
    bad = {
        "untagged": base[:nb] + [f"BOUNDARY {len(base) - nb - 2}"] + base[nb + 2:],
        "inverted": base[:cells0] + [" ".join([inv[0]] + inv[1:][::-1])] + base[cells0 + 1:],
        "range": base[:cells0] + ["3 0 1 999999"] + base[cells0 + 1:],
        "section": base + ["FACES 0"],
        "version": ["# FVMESH 2.0"] + base[1:],
    }
    for what, lines in bad.items():
        p = str(tmp_path / f"bad_{what}.fvmesh")
        open(p, "w").write("\n".join(lines) + "\n")
        assert pyref.read_mesh(p) is None, what
        with pytest.raises(FVKError):
            ours(p)
