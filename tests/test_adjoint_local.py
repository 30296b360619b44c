"""The hand-written adjoints of the flux (on conserved and on primitive face
states), ghost, face-state, limiter and geometry functions
(finitevolume_amd/csrc/adjoint.h) against central finite
differences of the forward functions (physics.h), compiled for the CPU with
g++ from the same headers the kernels use.  A wrong sign, a dropped term or a
swapped operand in any adjoint gives an O(1) error here."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers are not available")
    exe = str(tmp_path_factory.mktemp("adj") / "adjoint_local_check")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(REPO, "finitevolume_amd", "csrc"), "-I" + os.path.join(REPO, "include"),
           os.path.join(HERE, "adjoint_local_check.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        parts = line.split()
        res[parts[0]] = [float(x) for x in parts[1:]]
    return res


@pytest.mark.parametrize("name,tol", [("hllc", 1e-5), ("rusanov", 1e-5), ("slip_wall", 1e-5), ("face_state", 1e-4),
                                      ("hllc_w", 1e-5), ("rusanov_w", 1e-5), ("slip_wall_w", 1e-5),
                                      ("primitive", 1e-5), ("venkatakrishnan", 1e-5), ("edge", 1e-5)])
def test_adjoint_matches_fd(results, name, tol):
    n, err = results[name]
    assert n >= 1000
    assert err < tol, f"{name}: max relative error {err:.3e}"


def test_primitive_fluxes_equal_conserved(results):
    # hllc_w / rusanov_w on W = to_prim(U) against hllc / rusanov on U: the
    # same function, so they differ by rounding only (relative to the flux)
    n, err = results["primitive_equals_conserved"]
    assert n >= 1000
    assert err < 1e-13, err


def test_hllc_branches_covered(results):
    # supersonic left, supersonic right, star left, star right
    assert all(h >= 10 for h in results["hllc_branches"])
