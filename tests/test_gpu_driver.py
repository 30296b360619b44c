"""The driver contract on the GPU: checkpoints byte-compatible with the
reference's Checkpoint (src/Checkpoint.cpp), resume from a reference-written
file and continue as the oracle does from the same state, and the run loop
of src/main.cpp:3491-3557 with its residual CSV.

Reference-written checkpoints are the golden fixtures tests/golden/ckpt_*.bin;
where the reference build (oracle/_ref) is present, the reference also
reads the files the GPU writes.
"""
import os
import sys

import numpy as np
import pytest

import cases
from finitevolume_amd import abi
from finitevolume_amd.solver import FVKError, FVKSolver
from oracle import pyoracle, pyref
from oracle.pyoracle import OracleSolver

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_driver_host import read_checkpoint  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_golden import checkpoint_cases  # noqa: E402

pytestmark = pytest.mark.gpu

CKPT = {c[0]: c for c in checkpoint_cases()}


def _fresh(name):
    _, case, _, step, build = CKPT[name]
    m, bcs, p, ic, U0 = case
    return m, bcs, p, ic, U0, step, build


def _expected_field(name, m, eq):
    h, pay = read_checkpoint(os.path.join(GOLD, name))
    n = m.n_cells
    if eq in (abi.EQ_DIFFUSION, abi.EQ_ADVECTION_DIFFUSION):
        return pay[:n], None
    U = pay[: 4 * n].reshape(n, 4)
    nt = {abi.EQ_RANS_SA: 1, abi.EQ_RANS_SST: 2}.get(eq, 0)
    T = pay[4 * n:].reshape(nt, n).T.copy() if nt else None
    return U, T


@pytest.mark.parametrize("name", list(CKPT))
def test_resume_reference_checkpoint_and_write_it_back(name, tmp_path):
    m, bcs, p, ic, U0, step, build = _fresh(name)
    g = FVKSolver(m, bcs, p, ic)
    st, bd, src = g.checkpoint_read(os.path.join(GOLD, name))
    assert (st, bd, src) == (step, build, read_checkpoint(os.path.join(GOLD, name))[0]["equation"])
    U, T = _expected_field(name, m, p.equation)
    f = g.field()
    if isinstance(f, tuple):
        assert np.array_equal(f[0], U) and np.array_equal(f[1], T)
    else:
        assert np.array_equal(f, U)
    assert np.all(g.residual() == 0.0)
    # written back: byte for byte the reference's file
    out = tmp_path / "back.bin"
    g.checkpoint_write(str(out), step, build)
    assert out.read_bytes() == open(os.path.join(GOLD, name), "rb").read()
    # and the run continues as the oracle continues from the same state
    o = OracleSolver(m, bcs, p, ic)
    o.set_field(U, T)
    rg = g.step(20)
    ro = o.step(20)
    assert cases.residual_close(rg, ro, 1e-10)
    fg, fo = g.field(), o.field()
    for a, b in zip(fg if isinstance(fg, tuple) else (fg,), fo if isinstance(fo, tuple) else (fo,)):
        assert cases.rel_state_diff(a, b, floor=1e-12) <= 1e-12


@pytest.mark.skipif(not pyref.available(), reason="reference build (oracle/_ref) not present")
def test_reference_reads_gpu_checkpoint(tmp_path):
    for name in ("ckpt_euler_naca.bin", "ckpt_sa_wavy.bin", "ckpt_sst_wavy.bin", "ckpt_diffusion.bin",
                 "ckpt_advdiff.bin"):
        m, bcs, p, ic, U0, _, _ = _fresh(name)
        g = FVKSolver(m, bcs, p, ic)
        if U0 is not None:
            g.set_field(U0)
        g.step(7)
        path = str(tmp_path / name)
        g.checkpoint_write(path, 7, 3)
        tag = read_checkpoint(path)[0]["equation"]
        st, bd, U, T = pyref.checkpoint_read(path, int(tag), m.n_cells)
        assert (st, bd) == (7, 3)
        f = g.field()
        if isinstance(f, tuple):
            assert np.array_equal(U, f[0]) and np.array_equal(T, f[1])
        else:
            assert np.array_equal(U, f)


def test_flow_field_resume_across_equations():
    """An SA solver resumes the flow field of a Navier-Stokes checkpoint and
    keeps its own nut (src/main.cpp:3940-3960); its own model's checkpoint
    brings nut along."""
    m, bcs, p, ic, _, _, _ = _fresh("ckpt_sa_wavy.bin")
    g = FVKSolver(m, bcs, p, ic)
    nut0 = g.field()[1].copy()
    _, _, src = g.checkpoint_read(os.path.join(GOLD, "ckpt_ns_wavy.bin"))
    assert src == abi.CKPT_NAVIER_STOKES
    U_ns, _ = _expected_field("ckpt_ns_wavy.bin", m, abi.EQ_NAVIER_STOKES)
    assert np.array_equal(g.field()[0], U_ns) and np.array_equal(g.field()[1], nut0)
    g.checkpoint_read(os.path.join(GOLD, "ckpt_sa_wavy.bin"))
    U_sa, T_sa = _expected_field("ckpt_sa_wavy.bin", m, abi.EQ_RANS_SA)
    assert np.array_equal(g.field()[0], U_sa) and np.array_equal(g.field()[1], T_sa)


def test_checkpoint_errors(tmp_path):
    m, bcs, p, ic, _, _, _ = _fresh("ckpt_euler_naca.bin")
    g = FVKSolver(m, bcs, p, ic)
    raw = open(os.path.join(GOLD, "ckpt_euler_naca.bin"), "rb").read()
    bad = tmp_path / "bad.bin"
    for blob, what in ((b"NOTACKPT" + raw[8:], "bad magic"), (raw[:40 + 100], "truncated"), (raw[:20], "short header")):
        bad.write_bytes(blob)
        with pytest.raises(FVKError):
            g.checkpoint_read(str(bad))
    with pytest.raises(FVKError):  # cell count of another mesh
        g.checkpoint_read(os.path.join(GOLD, "ckpt_ns_wavy.bin"))
    with pytest.raises(FVKError):  # a diffusion field is not a flow field
        FVKSolver(*cases.case_ns_wavy()[:4]).checkpoint_read(os.path.join(GOLD, "ckpt_diffusion.bin"))


def _oracle_loop(case, cfg):
    """The reference driver's loop order (src/main.cpp:3491-3534) over the
    oracle: residual rows, final CFL, status."""
    m, bcs, p, ic, U0 = case
    o = OracleSolver(m, bcs, p, ic)
    if U0 is not None:
        o.set_field(U0)
    cfl = cfg.ramp_cfl_min if cfg.cfl_ramp else cfg.cfl
    o.set_cfl(cfl)
    rows, window, status = [], [], 0
    for k in range(1, cfg.nsteps + 1):
        r = o.step(1)[0]
        if not np.all(np.isfinite(r)):
            status = 2
            break
        if cfg.cfl_ramp:
            window = (window + [r[0]])[-cfg.ramp_window:]
            cfl = pyoracle.next_cfl(cfl, cfg.cfl, np.array(window), cfg.ramp_cfl_min, cfg.ramp_growth,
                                    cfg.ramp_shrink, cfg.ramp_window, cfg.ramp_divergence_threshold)
            o.set_cfl(cfl)
        rows.append((k, r))
        tol = [cfg.tolerance[i] for i in range(len(r))]
        if any(t >= 0 for t in tol) and all(t < 0 or r[i] < t for i, t in enumerate(tol)):
            status = 1
            break
    return rows, cfl, status, o


def _parse_csv(path):
    lines = open(path).read().splitlines()
    return lines[0], [ln.split(",") for ln in lines[1:]]


@pytest.mark.parametrize("equation", [abi.EQ_EULER, abi.EQ_RANS_SA])
def test_run_loop_with_ramp_and_csv(equation, tmp_path):
    case = cases.case_naca(64, 32) if equation == abi.EQ_EULER else cases.case_rans_wavy(abi.EQ_RANS_SA)
    m, bcs, p, ic, U0 = case
    csv = str(tmp_path / "res.csv")
    cfg = abi.run_config(90, cfl=2.0 if equation == abi.EQ_EULER else 0.6, cfl_ramp=True, residual_file=csv,
                         output_precision=17, residual_interval=3, tolerance=[1e-14])
    g = FVKSolver(m, bcs, p, ic)
    if U0 is not None:
        g.set_field(U0)
    res = g.run_case(cfg)
    rows, cfl, status, o = _oracle_loop(case, cfg)
    assert (res.status, res.last_step, res.resume_start) == (status, rows[-1][0], 1)
    assert res.cfl == cfl
    head, body = _parse_csv(csv)
    want = "step,residual_rho,residual_rho_u,residual_rho_v,residual_E" + (",residual_nut" if equation else "")
    assert head == want
    expect = [r for r in rows if r[0] % 3 == 0]
    assert [int(b[0]) for b in body] == [k for k, _ in expect]
    vals = np.array([[float(x) for x in b[1:]] for b in body])
    assert cases.residual_close(vals, np.array([r for _, r in expect]), 1e-10)
    # every double printed as std::ostream << setprecision(17) prints it
    for b in body:
        assert all(x == "%.17g" % float(x) for x in b[1:])
    if equation == abi.EQ_EULER:  # first-order Euler: the same state bits as the oracle
        assert np.array_equal(g.field(), o.field())
    else:
        assert cases.rel_state_diff(g.field()[0], o.field()[0], floor=1e-12) <= 1e-12


def test_run_resumes_from_its_checkpoint(tmp_path):
    """nsteps 40 with a checkpoint, then a fresh solver with nsteps 70 on the
    same file resumes at 41 and ends where an uninterrupted 70-step run
    ends, bit for bit (fixed CFL)."""
    m, bcs, p, ic, _ = cases.case_naca(64, 32, order=2, limiter=abi.LIMITER_VENKATAKRISHNAN, rk_stages=3,
                                       local_dt=1)
    ck = str(tmp_path / "run.ckpt")
    a = FVKSolver(m, bcs, p, ic)
    r1 = a.run_case(abi.run_config(40, cfl=0.8, checkpoint_file=ck, build_number=5))
    assert (r1.status, r1.last_step, r1.resumed_equation) == (0, 40, -1)
    assert read_checkpoint(ck)[0]["step_index"] == 40
    b = FVKSolver(m, bcs, p, ic)
    r2 = b.run_case(abi.run_config(70, cfl=0.8, checkpoint_file=ck))
    assert (r2.resume_start, r2.last_step, r2.resumed_equation) == (41, 70, abi.CKPT_EULER)
    c = FVKSolver(m, bcs, p, ic)
    c.step(70)
    assert np.array_equal(b.field(), c.field())
    assert read_checkpoint(ck)[0]["step_index"] == 70


def test_run_diverges_and_converges(tmp_path):
    m, bcs, p, ic, _ = cases.case_naca(32, 16)
    g = FVKSolver(m, bcs, p, ic)
    U = g.field()
    U[3, 1] = np.nan
    g.set_field(U)
    ck = str(tmp_path / "none.ckpt")
    r = g.run_case(abi.run_config(10, cfl=0.8, checkpoint_file=ck))
    assert (r.status, r.last_step) == (abi.RUN_DIVERGED, 1)
    assert not os.path.exists(ck)  # no checkpoint after a divergence
    g2 = FVKSolver(m, bcs, p, ic)
    g2.step(3)
    r = g2.run_case(abi.run_config(10, cfl=0.8, tolerance=[1e9, 1e9, 1e9, 1e9]))
    assert (r.status, r.last_step) == (abi.RUN_CONVERGED, 1)
