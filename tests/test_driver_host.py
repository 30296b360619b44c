"""Host-side pieces of the driver contract, without a GPU: the library's
next_cfl against the reference's outputs (tests/golden/cfl_ramp.npz), and
the reference-written checkpoint fixtures' layout (include/Checkpoint.h)."""
import os

import numpy as np
import pytest

from finitevolume_amd.build import LIB

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CKPT_HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("equation", "<u4"), ("cell_count", "<u8"),
                        ("step_index", "<i8"), ("build_number", "<u8")])


def read_checkpoint(path):
    """Parses a FVMCKPT file: (header, payload doubles)."""
    raw = open(path, "rb").read()
    h = np.frombuffer(raw[:40], dtype=CKPT_HEADER)[0]
    return h, np.frombuffer(raw[40:], dtype="<f8")


@pytest.fixture(scope="module")
def solver_mod():
    if not os.path.exists(LIB):
        pytest.skip("libfvk.so is not built")
    from finitevolume_amd import solver
    return solver


def test_next_cfl_matches_reference(solver_mod):
    g = np.load(os.path.join(GOLD, "cfl_ramp.npz"))
    for key in ("short", "falling", "flat", "mild_rise", "diverging"):
        out = np.array([solver_mod.next_cfl(cur, 10.0, g[f"seq_{key}"]) for cur in (0.05, 1.0, 9.0)])
        assert np.array_equal(out, g[f"out_{key}"]), key


def test_next_cfl_branches(solver_mod):
    # window not full: hold; clamp to [cfl_min, cfl_max]
    assert solver_mod.next_cfl(0.7, 10.0, [1.0] * 14) == 0.7
    assert solver_mod.next_cfl(9.0, 10.0, np.logspace(0, -1, 15)) == 10.0
    assert solver_mod.next_cfl(3.0, 10.0, np.logspace(0, 3, 15)) == 0.05


@pytest.mark.parametrize("name,tag,nt,ncell", [("ckpt_euler_naca.bin", 1, 4, 1024), ("ckpt_ns_wavy.bin", 3, 4, 256),
                                               ("ckpt_sa_wavy.bin", 4, 5, 256), ("ckpt_sst_wavy.bin", 5, 6, 256),
                                               ("ckpt_diffusion.bin", 0, 1, 256), ("ckpt_advdiff.bin", 2, 1, 256)])
def test_checkpoint_fixture_layout(name, tag, nt, ncell):
    h, payload = read_checkpoint(os.path.join(GOLD, name))
    assert h["magic"] == b"FVMCKPT" and h["version"] == 2
    assert h["equation"] == tag and h["cell_count"] == ncell
    assert payload.size == nt * ncell
    assert np.all(np.isfinite(payload))
