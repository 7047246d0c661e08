"""The drop-in `md_hip name` executable against the `MD name` contract (MD.cpp:72-754).

CPU tests cover the argument and load-failure paths (no device call happens there); GPU
tests run whole jobs in a temporary directory and compare the files with the library
driven directly and with the oracle's step-0 energies.
"""
import os
import subprocess

import numpy as np
import pytest

import softmold_amd as sm
from conftest import gpu_available

CLI = os.path.join(os.path.dirname(sm.__file__), "md_hip")


def _run(args, cwd, timeout=300):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)


def _table(path):
    with open(path) as f:
        return np.array([[float(x) for x in line.split("\t")] for line in f if line.strip()])


def _frames(path):
    with open(path) as f:
        lines = f.read().split("\n")
    out, k = [], 0
    while k < len(lines) and lines[k].strip():
        n = int(lines[k])
        assert lines[k + 1] == "test"
        rows = np.array([[float(x) for x in ln.split("\t")] for ln in lines[k + 2:k + 2 + n]])
        out.append(rows)
        k += 2 + n
    return out


def test_cli_built():
    assert os.path.exists(CLI), "md_hip is not built"


def test_cli_usage(tmp_path):
    r = _run([], tmp_path)
    assert r.returncode == 0
    assert "usage:" in r.stderr
    r = _run(["a", "b"], tmp_path)
    assert r.returncode == 0 and "usage:" in r.stderr


def test_cli_missing_file(tmp_path):
    r = _run(["nosuchsystem"], tmp_path)
    assert r.returncode == 255
    assert "System load failure" in r.stderr


def _make(tmp_path, name, n_lipids=500, steps=40, store=20, measure=10, dlxy=0.0, tension=0.0):
    m = sm.Mpd.liposome(1234, n_lipids, 3.45)
    dt = m["deltaT"]
    m["finalTime"] = steps * dt
    m["storeInterval"] = store * dt
    m["measureInterval"] = measure * dt
    if dlxy:
        m["deltaLXY"] = dlxy
        m["tension"] = tension
    m.write(str(tmp_path / name))
    return m


@pytest.mark.gpu
def test_cli_job_matches_library(tmp_path):
    """40 steps, store every 20, measure every 10: files as MD.cpp writes them, final
    state equal to the same steps run through the library in one call."""
    if not gpu_available():
        pytest.skip("no GPU")
    m = _make(tmp_path, "job")
    d = m.system()
    print("case: md_hip job L500 40 steps", flush=True)
    r = _run(["job"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "starting main loop:" in r.stderr
    assert "Resize acceptance ratio:" in r.stderr

    pot = _table(tmp_path / "potential_job.dat")
    kin = _table(tmp_path / "kinetic_job.dat")
    size = _table(tmp_path / "size_job.dat")
    temp = _table(tmp_path / "temp_job.dat")
    np.testing.assert_allclose(pot[:, 0], [0, 0.2, 0.4, 0.6, 0.8], rtol=0, atol=1e-12)
    assert kin.shape == (5, 2) and size.shape == (5, 4) and temp.shape == (5, 2)
    # step 0: K = 3/2 N T exactly (|v| = sqrt(3T) from the generator), U from the oracle
    n = len(d["type"])
    assert kin[0, 1] == pytest.approx(1.5 * n * d["temperature"], rel=1e-6)
    import oracle
    s0 = oracle.System(d["pos"], d["type"], d["vel"], d["box"], d["fconst"], d["uconst"], d["n_types"],
                       d["chain"], d["chain_const"], cutoff=d["cutoff"])
    assert pot[0, 1] == pytest.approx(s0.pair_potential() + s0.chain_potential(), rel=2e-6)
    np.testing.assert_allclose(size[:, 1:], np.tile(d["box"], (5, 1)), rtol=1e-6)
    np.testing.assert_allclose(temp[:, 1], d["temperature"])
    # without BOND / BEND records MD writes lBond = 0/0 and zero bend sums
    # (dataExtraction.h:1009-1027), in the text x86-64 prints for that NaN
    lb = (tmp_path / "lBond_job.dat").read_text().splitlines()
    be = (tmp_path / "bend_job.dat").read_text().splitlines()
    assert [ln.split("\t") for ln in lb] == [[f"{t:g}", "-nan"] for t in (0, 0.2, 0.4, 0.6, 0.8)]
    assert [ln.split("\t")[1:] for ln in be] == [["0", "0", "0"]] * 5

    frames = _frames(tmp_path / "frames_job.xyz")
    assert len(frames) == 3    # step 0, 20, 40
    np.testing.assert_allclose(frames[0][:, 1:], d["pos"], rtol=1e-5, atol=1e-4)
    assert (frames[0][:, 0] == d["type"]).all()

    # the rewritten .mpd holds the mid-step state of step 40 (MD.cpp:373-381)
    out = sm.Mpd.read(str(tmp_path / "job"))
    assert out["initialTime"] == pytest.approx(40 * d["dt"])
    tmp = sm.Mpd.read(str(tmp_path / "job.tmp"))
    assert tmp["initialTime"] == out["initialTime"]
    o = out.system()

    # the library on the system as md_hip read it (the .mpd text keeps 15 digits, so the
    # file's state is not bit-identical to the generator's in-memory doubles)
    m0 = sm.Mpd.liposome(1234, 500, 3.45)
    m0.write(str(tmp_path / "fresh"))
    lib = sm.Solver.from_system(sm.Mpd.read(str(tmp_path / "fresh")).system())
    lib.init_forces()
    lib.run(0, 40)
    lib.step_first(40)
    pos, vel, _, box = lib.download()
    lib.close()
    np.testing.assert_allclose(o["pos"], pos, rtol=1e-14, atol=1e-12)
    np.testing.assert_allclose(o["vel"], vel, rtol=1e-14, atol=1e-12)
    np.testing.assert_allclose(frames[2][:, 1:], pos, rtol=1e-5, atol=1e-4)


@pytest.mark.gpu
def test_cli_restart_and_mc(tmp_path):
    """MC box moves (deltaLXY != 0) write resizeHist_; a restart from the stored mid-step
    .mpd continues with forces + Verlet::second (MD.cpp:274-308)."""
    if not gpu_available():
        pytest.skip("no GPU")
    _make(tmp_path, "mc", steps=64, store=32, measure=16, dlxy=0.05, tension=1.0)
    print("case: md_hip MC L500 64 steps", flush=True)
    r = _run(["mc"], tmp_path)
    assert r.returncode == 0, r.stderr
    ratio = float(r.stderr.strip().split("Resize acceptance ratio:")[1].split()[0])
    assert 0.0 <= ratio <= 1.0
    hist = _table(tmp_path / "resizeHist_mc.dat")
    assert hist.shape == (203, 3)
    assert hist[:, 1:].sum() == pytest.approx(64 // 8)   # trials at 8..64
    assert hist[0, 0] == pytest.approx(-0.05)
    size = _table(tmp_path / "size_mc.dat")
    assert size.shape == (5, 4)

    m = sm.Mpd.read(str(tmp_path / "mc"))
    assert m["initialTime"] == pytest.approx(64 * m["deltaT"])
    m["finalTime"] = 96 * m["deltaT"]
    m.write(str(tmp_path / "mc"))
    print("case: md_hip MC restart 64..96", flush=True)
    r = _run(["mc"], tmp_path)
    assert r.returncode == 0, r.stderr
    pot = _table(tmp_path / "potential_mc.dat")
    # first job: t=0,16,32,48,64; restart from 64 measures at 80, 96 (i != start)
    np.testing.assert_allclose(pot[:, 0] / m["deltaT"], [0, 16, 32, 48, 64, 80, 96], atol=1e-9)
    assert np.isfinite(pot[:, 1]).all()
    # the restarted run continues the uninterrupted one: same job in one go to t = 96
    # (the noise after the restart is drawn afresh, as the reference's is, so the
    # comparison is statistical: the vesicle is still relaxing, alike in both runs)
    _make(tmp_path, "one", steps=96, store=32, measure=16, dlxy=0.05, tension=1.0)
    print("case: md_hip MC 0..96 in one job", flush=True)
    r = _run(["one"], tmp_path)
    assert r.returncode == 0, r.stderr
    one = _table(tmp_path / "potential_one.dat")
    np.testing.assert_allclose(one[:5], pot[:5], rtol=1e-12)    # identical up to the store
    np.testing.assert_allclose(one[5:, 1], pot[5:, 1], rtol=0.02)
    # the kinetic energy after the restart follows the uninterrupted run (the generated
    # vesicle is still releasing potential energy into heat at this point, alike in both)
    kin = _table(tmp_path / "kinetic_mc.dat")
    kin1 = _table(tmp_path / "kinetic_one.dat")
    np.testing.assert_allclose(kin1[:5], kin[:5], rtol=1e-12)
    np.testing.assert_allclose(kin1[5:, 1], kin[5:, 1], rtol=0.1)


@pytest.mark.gpu
def test_cli_extents_and_kinetic_histogram_match_reference(golden_traj, tmp_path):
    """flicker_ (the particles' extent along each axis at step 0 and every measurement) and
    kEnergyDensity_ (the histogram of 0.5 |v|^2 in bins of 1e-4 over those measurements,
    dataExtraction.h:552-567,1456-1487,1512-1519) against the reference MD's own files: on
    the BOND/BEND system run for step 0 only, both files equal the reference's (the
    step-0 velocities straddle the bin edge at 4.5, so the split between two bins checks
    the FP64 bin arithmetic); on AB400 over 40 steps, the step-0 extents equal the
    reference's and the later lines keep its format (their values follow the noise)."""
    if not gpu_available():
        pytest.skip("no GPU")
    (tmp_path / "BS.mpd").write_bytes(golden_traj["BS/input_mpd"].tobytes())
    print("case: md_hip BS step 0", flush=True)
    r = _run(["BS"], tmp_path)
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(_table(tmp_path / "flicker_BS.dat"), golden_traj["BS/flicker"])
    ke = _table(tmp_path / "kEnergyDensity_BS.dat")
    assert len(ke) == int(golden_traj["BS/kEnergyDensity_lines"][0])
    np.testing.assert_array_equal(ke[ke[:, 1] != 0], golden_traj["BS/kEnergyDensity_nonzero"])
    np.testing.assert_allclose(ke[:, 0], np.arange(len(ke)) * 1e-4, rtol=1e-5, atol=1e-9)

    name = "AB400t"
    (tmp_path / f"{name}.mpd").write_bytes(golden_traj[name + "/input_mpd"].tobytes())
    print("case: md_hip AB400t 40 steps", flush=True)
    r = _run([name], tmp_path)
    assert r.returncode == 0, r.stderr
    fl = _table(tmp_path / f"flicker_{name}.dat")
    ref = golden_traj[name + "/flicker"]
    assert fl.shape == ref.shape
    np.testing.assert_array_equal(fl[0], ref[0])
    np.testing.assert_array_equal(fl[:, 0], _table(tmp_path / f"potential_{name}.dat")[:, 0])
    size = _table(tmp_path / f"size_{name}.dat")
    assert np.all(fl[:, 1:] > 0) and np.all(fl[:, 1:] <= size[:, 1:] * (1 + 1e-6))
    ke = _table(tmp_path / f"kEnergyDensity_{name}.dat")
    assert ke[:, 1].sum() == pytest.approx(1.0, abs=1e-4)
