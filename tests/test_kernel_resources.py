"""Register and scratch budgets of the hot kernel instances, read from the
built code objects' metadata (no GPU needed).  A change that makes the
compiler spill one of them, or pushes it past the registers its occupancy
was sized for, slows the workload it serves without changing any result, so
no parity test would notice."""
import os

import pytest

import kernel_resources as kr

# (translation unit, kernel symbol, workload, most VGPRs): the register
# budgets are those of the occupancy each launch is sized for (512 VGPRs per
# SIMD lane shared by the wavefronts on it)
HOT = [
    ("kernels_euler", "_ZN3fvk8k_e_tileILi3ELi1ELi2ELi4EEEvNS_9EulerArgsE",
     "2nd-order HLLC tile stage on triangles, four wavefronts per SIMD", 128),
    ("kernels_euler", "_ZN3fvk8k_e_tileILi9ELi1ELi2ELi1EEEvNS_9EulerArgsE",
     "2nd-order HLLC tile stage on polygons of up to nine faces", 256),
    ("kernels_euler", "_ZN3fvk16k_e_tile_persistILi9ELi1ELi2ELi1EEEvNS_9EulerArgsE",
     "persistent batched tile stage on polygons of up to nine faces", 256),
    ("kernels_euler", "_ZN3fvk8k_e_faceILi1ELi1EEEvNS_9EulerArgsE",
     "1st-order HLLC face sweep (reference mode)", 256),
    ("kernels_visc", "_ZN3fvk11k_visc_tileILi2ELi1ELi4EEEvNS_8ViscArgsE",
     "SA RANS tile stage (HLLC, cells of up to four faces), three wavefronts per SIMD", 168),
    ("adjoint", "_ZN3fvk10k_adj_faceILi1ELi2ELb0EEEvNS_7AdjArgsE",
     "state adjoint of the 2nd-order HLLC face, four wavefronts per SIMD", 128),
    ("adjoint", "_ZN3fvk10k_adj_cellILi3ELb0EEEvNS_7AdjArgsE",
     "state adjoint of the limited LSQ gradient on triangles", 168),
]


@pytest.mark.skipif(not os.path.isdir(kr.BUILD), reason="the library's objects are not built")
@pytest.mark.parametrize("unit,symbol,what,vgpr_max", HOT, ids=[h[1] for h in HOT])
def test_hot_kernel_fits_its_budget(unit, symbol, what, vgpr_max):
    res = kr.kernel_resources(unit)
    assert symbol in res, f"{symbol} ({what}) is not in {unit}"
    r = res[symbol]
    assert r["scratch"] == 0, f"{what}: {r['scratch']} B of scratch per lane"
    assert r["vgpr"] <= vgpr_max, f"{what}: {r['vgpr']} VGPRs (budget {vgpr_max})"
