"""The Fortran host program (host/sac_vac.f90): par/ini parsing, the main
loop and the ini/out/log files, as a drop-in for the serial `vac`.

CPU tests build the program with flang and check parsing and the refusal
of configurations outside the GPU path (the library then stops at device
initialisation).  GPU tests run whole problems through the program and
compare the saved snapshots with the oracle at the FP64 bar.
"""
import os
import subprocess

import numpy as np
import pytest

from sac_amd import problems, vacio
from tests.helpers import make_oracle, oracle_run, per_var_errors

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "host", "_build", "sac_vac")


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        from host import build as hb
        hb.build()
    return EXE


def _run(exe, cwd, par_text, timeout=300):
    return subprocess.run([exe], input=par_text, capture_output=True, text=True, cwd=cwd,
                          timeout=timeout)


def _setup(tmp_path, prob, fmt="ascii", out_fmt="binary", **par):
    ini = tmp_path / ("ini." + fmt)
    # the physics tag follows the first '_' of the header (src/vacio.t:858-866)
    header = prob.name.replace("_", "") + f"_mhd{prob.ndim}{prob.ndir}"
    if fmt == "ascii":
        vacio.write_ini_ascii(ini, prob, header=header)
    else:
        vacio.write_ini_binary(ini, prob, header=header)
    names = {1: "rho m1 e b1 eb rhob bg1", 2: "rho m1 m2 e b1 b2 eb rhob bg1 bg2",
             3: "rho m1 m2 m3 e b1 b2 b3 eb rhob bg1 bg2 bg3"}[prob.ndir]
    tb = [problems.BC_NAMES[int(c)] for c in prob.typeB[0]]
    vacio.write_par(tmp_path / "run.par", ini.name, "out." + out_fmt, "run.log",
                    typefileout=out_fmt, nw=prob.nw, ne=2 + 2 * prob.ndir, ndim=prob.ndim,
                    typeB=tb, courantpar=prob.courantpar, wnames=names, **par)
    return (tmp_path / "run.par").read_text()


def test_vacio_binary_roundtrip(tmp_path):
    prob = problems.stratified(n=6, nxyz=(6, 7, 8))
    vacio.write_ini_binary(tmp_path / "a.ini", prob)
    vacio.write_ini_binary(tmp_path / "b.ini", prob, fullgrid=True)
    (s,) = vacio.read_binary(tmp_path / "a.ini")
    assert s.nx == (6, 7, 8) and s.nw == 13 and s.ndim == 3
    np.testing.assert_array_equal(s.w, prob.w[(Ellipsis,) + prob.mesh_slices()])
    (f,) = vacio.read_binary(tmp_path / "b.ini")
    np.testing.assert_array_equal(f.w, prob.w)
    vacio.write_ini_ascii(tmp_path / "c.ini", prob)
    (a,) = vacio.read_ascii(tmp_path / "c.ini")
    np.testing.assert_array_equal(a.w, s.w)
    np.testing.assert_array_equal(a.x, s.x)


def test_host_parses_then_needs_device(exe, tmp_path):
    """A supported configuration is parsed through to device initialisation;
    without a GPU the library reports it and the program stops non-zero."""
    prob = problems.brio_wu(n=50)
    par = _setup(tmp_path, prob, itmax=3)
    r = _run(exe, tmp_path, par)
    text = r.stdout + r.stderr
    if r.returncode == 0:  # a GPU is visible: the run itself must have worked
        assert "Finish Advance" in text
    else:
        assert "libsac_amd" in text and "device" in text, text


@pytest.mark.parametrize("edit,msg", [
    ("dimsplit=F", "dimsplit=T is not built"),
    ("smallfix=F", "smallfix=T is not built"),
    ("typeadvance='fourstep'", "typeadvance=adams2 is not built"),
])
def test_host_refuses_unbuilt_methods(exe, tmp_path, edit, msg):
    prob = problems.stratified(n=6, nxyz=(6, 6, 6))
    par = _setup(tmp_path, prob, itmax=1)
    repl = {"dimsplit=F": "dimsplit=T", "smallfix=F": "smallfix=T",
            "typeadvance='fourstep'": "typeadvance='adams2'"}[edit]
    r = _run(exe, tmp_path, par.replace(edit, repl))
    assert r.returncode != 0
    assert msg in r.stdout + r.stderr


@pytest.mark.gpu
def test_host_briowu_ascii_in_binary_out(exe, tmp_path):
    """The program's run equals the library's own main loop bit for bit
    (same calls, same dt), and the snapshots and log land where and when
    the save settings ask."""
    from sac_amd import device
    prob = problems.brio_wu(n=200)
    steps = 10
    par = _setup(tmp_path, prob, fmt="ascii", out_fmt="binary", itmax=steps, ditsave=(1, 5))
    r = _run(exe, tmp_path, par)
    assert r.returncode == 0, r.stdout + r.stderr
    snaps = vacio.read_binary(tmp_path / "out.binary")
    assert [s.it for s in snaps] == [0, 5, 10]
    w, dts = device.run(prob, steps=steps)
    np.testing.assert_array_equal(snaps[-1].w, w[(Ellipsis,) + prob.mesh_slices()])
    assert snaps[-1].t == sum(dts)
    hdr, cols, rows = vacio.read_log(tmp_path / "run.log")
    assert rows.shape[0] == steps + 1 and rows[-1, 0] == steps
    np.testing.assert_allclose(rows[1:, 2], dts, rtol=1e-5)  # log precision: 6 digits


@pytest.mark.gpu
def test_host_briowu_fixed_dt_matches_oracle(exe, tmp_path, oracle_mod):
    """With a fixed dt (dtpar) the run is compared with the oracle itself."""
    prob = problems.brio_wu(n=200)
    steps, dt = 20, 2.0e-4
    par = _setup(tmp_path, prob, fmt="binary", out_fmt="binary", itmax=steps, dtpar=dt)
    r = _run(exe, tmp_path, par)
    assert r.returncode == 0, r.stdout + r.stderr
    snaps = vacio.read_binary(tmp_path / "out.binary")
    assert snaps[-1].it == steps
    o = make_oracle(prob, oracle_mod)
    for _ in range(steps):
        o.advance(dt)
    ref = o.get_w()
    got = prob.w.copy()
    got[(Ellipsis,) + prob.mesh_slices()] = snaps[-1].w
    errs = per_var_errors(prob, got, ref, 6)
    assert max(errs) <= 1e-12, errs


@pytest.mark.gpu
def test_host_stratified_binary_in_ascii_out(exe, tmp_path, oracle_mod):
    prob = problems.stratified(n=16, nxyz=(16, 12, 10))
    steps = 8
    par = _setup(tmp_path, prob, fmt="binary", out_fmt="ascii", itmax=steps, ditsave=(4, 8))
    r = _run(exe, tmp_path, par)
    assert r.returncode == 0, r.stdout + r.stderr
    snaps = vacio.read_ascii(tmp_path / "out.ascii")
    assert [s.it for s in snaps] == [0, 8]
    ref, rdts, _ = oracle_run(prob, oracle_mod, steps)
    refm = ref[(Ellipsis,) + prob.mesh_slices()]
    # ASCII rows carry 11 significant digits and flush |w| <= 5e-16 to zero
    for iw in range(8):
        scale = np.abs(refm[iw]).max()
        assert np.abs(snaps[-1].w[iw] - refm[iw]).max() <= 1e-10 * scale + 5e-16, iw


@pytest.mark.gpu
def test_host_threestep_divbfix_extrapolating_boundaries(exe, tmp_path, oracle_mod):
    """The drop-in program with settings beyond the plain path: Shu's
    threestep integrator, Powell's div B sources (divbfix=T, the
    reference's default) and cont1/grad1 boundaries along the vertical axis,
    free-running against the oracle with the same settings."""
    prob = problems.stratified(n=12, nxyz=(12, 10, 8), amp=3e-2)
    prob.typeB[:, 0] = problems.BC["cont1"]
    prob.typeB[:, 1] = problems.BC["grad1"]
    steps = 6
    par = _setup(tmp_path, prob, fmt="binary", out_fmt="binary", itmax=steps,
                 typeadvance="threestep", divbfix=True)
    r = _run(exe, tmp_path, par)
    assert r.returncode == 0, r.stdout + r.stderr
    snaps = vacio.read_binary(tmp_path / "out.binary")
    assert snaps[-1].it == steps
    o = make_oracle(prob, oracle_mod)
    o.set_divbfix(True)
    dts = [o.step(prob.courantpar, "threestep") for _ in range(steps)]
    assert snaps[-1].t == pytest.approx(sum(dts), rel=1e-12)
    ref = o.get_w()
    got = prob.w.copy()
    got[(Ellipsis,) + prob.mesh_slices()] = snaps[-1].w
    errs = per_var_errors(prob, got, ref, 8)
    assert max(errs) <= 1e-12, errs


@pytest.mark.gpu
@pytest.mark.parametrize("out_fmt", ["binary", "ascii"])
def test_host_async_saves_equal_sync_saves(exe, tmp_path, out_fmt):
    """Saves copy the state off the device while the next steps run
    (sac_dev_save_begin/ready/end) and write it later with the it, t and dt
    of the iteration that asked for it; the snapshot and log files are those
    of immediate downloads (SAC_SYNC_SAVES=1) byte for byte, for saves every
    step (each one still in flight when the next is due) and every 3 steps."""
    prob = problems.stratified(n=16, nxyz=(16, 12, 10))
    for dits in ((1, 1), (3, 2)):
        outs = []
        for env in ({}, {"SAC_SYNC_SAVES": "1"}):
            d = tmp_path / f"{dits[0]}_{len(env)}"
            d.mkdir()
            par = _setup(d, prob, fmt="binary", out_fmt=out_fmt, itmax=7, ditsave=dits)
            r = subprocess.run([exe], input=par, capture_output=True, text=True, cwd=d,
                               timeout=300, env=dict(os.environ, **env))
            assert r.returncode == 0, r.stdout + r.stderr
            outs.append(((d / f"out.{out_fmt}").read_bytes(), (d / "run.log").read_bytes()))
        assert outs[0] == outs[1], dits
    snaps = (vacio.read_binary if out_fmt == "binary" else vacio.read_ascii)(
        tmp_path / f"3_0/out.{out_fmt}")
    assert [s.it for s in snaps] == [0, 2, 4, 6, 7]
