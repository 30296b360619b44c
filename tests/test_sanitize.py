"""AddressSanitizer and UndefinedBehaviorSanitizer on the CPU:

  * the oracle's C restatement (make -C oracle sanitize), driven through its
    Python binding over every equation set for a few steps;
  * the library's mesh reader (finitevolume_amd/csrc/meshio.cpp, the host
    code that parses untrusted files), built alone (tests/sanitize_meshio.cpp,
    host code instrumented, device code not), on the committed meshes and on
    truncated and corrupted copies of them, which it must reject with an
    error and no sanitizer report."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
MESHES = os.path.join(HERE, "golden", "meshes")


def _libasan():
    if shutil.which("gcc") is None:
        return None
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


ORACLE_RUN = r"""
import sys
sys.path.insert(0, {repo!r}); sys.path.insert(0, {tests!r})
from oracle import pyoracle
pyoracle.LIB_PATH = {lib!r}
import numpy as np
import cases
from finitevolume_amd import abi
from oracle.pyoracle import OracleSolver
runs = [cases.case_1r(16), cases.case_1r(16, flux=abi.FLUX_HLLC, order=2, limiter=abi.LIMITER_VENKATAKRISHNAN,
        rk_stages=3, local_dt=1), cases.case_sod(flux=abi.FLUX_EXACT, nx=16, ny=4), cases.case_naca(16, 8),
        cases.adjoint_naca(abi.FLUX_HLLC, 2), cases.case_ns_wavy(n=8),
        cases.case_ns_wavy(n=8, gradient=abi.GRAD_GREEN_GAUSS), cases.case_rans_wavy(abi.EQ_RANS_SA, n=8),
        cases.case_rans_wavy(abi.EQ_RANS_SST, n=8), cases.case_diffusion(8), cases.case_advdiff(n=8)]
for m, bcs, p, ic, U0 in runs:
    o = OracleSolver(m, bcs, p, ic)
    if U0 is not None:
        o.set_field(U0)
    assert np.all(np.isfinite(o.step(3)))
    o.residual_eval() if p.equation in (abi.EQ_EULER,) else None
print("OK")
"""


@pytest.mark.skipif(_libasan() is None, reason="gcc's AddressSanitizer runtime not present")
def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "sanitize"], check=True)
    lib = os.path.join(REPO, "oracle", "_build", "libfvoracle_san.so")
    env = dict(os.environ, LD_PRELOAD=_libasan(), ASAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS="2")
    code = ORACLE_RUN.format(repo=REPO, tests=HERE, lib=lib)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stderr[-4000:]


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc not present")
def test_mesh_reader_under_sanitizers(tmp_path):
    exe = str(tmp_path / "san_meshio")
    inc = [f"-I{os.path.join(REPO, 'include')}", f"-I{os.path.join(REPO, 'finitevolume_amd', 'csrc')}"]
    cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-std=c++17", "-O1", "-g",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all", *inc, os.path.join(HERE, "sanitize_meshio.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    good = sorted(os.path.join(MESHES, f) for f in os.listdir(MESHES))
    bad = []
    for path in good:
        text = open(path).read()
        name = os.path.basename(path)
        for frac in (0.05, 0.3, 0.6, 0.95):  # truncated
            p = tmp_path / f"trunc{int(frac * 100)}_{name}"
            p.write_text(text[: int(len(text) * frac)])
            bad.append(str(p))
        lines = text.splitlines()
        for k, repl in enumerate(("999999999", "-7", "nan", "x")):  # a damaged token mid-file
            ln = len(lines) // 2 + k
            toks = lines[ln].split()
            if toks:
                toks[-1] = repl
                p = tmp_path / f"tok{k}_{name}"
                p.write_text("\n".join(lines[:ln] + [" ".join(toks)] + lines[ln + 1:]) + "\n")
                bad.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([exe, *good, *bad], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    out = r.stdout.splitlines()
    assert len(out) == len(good) + len(bad)
    assert all(ln.startswith("OK ") for ln in out[: len(good)]), out[: len(good)]
    # every truncated or damaged file is rejected or read in full, never crashes
    assert sum(ln.startswith("ERR ") for ln in out[len(good):]) >= len(bad) // 2
