"""CPU tests of the product's host side: the C-ABI library loads and exports every symbol
the header declares, the generators reproduce the reference's `.mpd` bytes, and the
`.mpd` reader/writer keep the reference's format and error behaviour.  No GPU calls."""
import hashlib
import json
import os
import re
import tempfile

import numpy as np
import pytest

import softmold_amd as sm
from conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "softmold_hip.h")


def test_library_exports_every_declared_symbol():
    text = open(HEADER).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sm_\w+)\s*\(", text, re.M))
    assert declared == set(sm.EXPORTS)
    L = sm.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert b"gfx950" in L.sm_version()


def _sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def _generate_and_hash(rec, path):
    kind, seed, n, rho = rec["argv"][0], int(rec["argv"][1]), int(rec["argv"][2]), float(rec["argv"][3])
    m = sm.Mpd.liposome(seed, n, rho) if kind == "liposome" else sm.Mpd.bilayer(seed, n, rho)
    nbeads = m["nParticles"]
    m.write(path)
    m.close()
    h = _sha(path + ".mpd")
    os.remove(path + ".mpd")
    return h, nbeads


@pytest.mark.parametrize("name", ["L500", "B400", "L5000", "B4000", "L2000s7", "L80000", "B1M"])
def test_generators_match_reference_bytes(name, tmp_path):
    """Byte-identical .mpd files against the reference generators' own output; B1M is the
    1M-lipid flat bilayer, whose lattice caps it at 999,698 lipids (lipidModels.h:1822-1828)."""
    rec = json.load(open(os.path.join(GOLDEN, "gen_sha256.json")))[name]
    h, nbeads = _generate_and_hash(rec, str(tmp_path / name))
    assert h == rec["sha256"]
    if name == "B1M":
        assert nbeads == rec["n_particles"] == 2999094


def test_replica_sweep_inputs_match_reference_bytes(tmp_path):
    """The 64 liposomes of the replica sweep (seeds 1000..1063, the sweep's lipid counts)
    are byte-identical to the reference generator's files."""
    from softmold_amd.sweep import sweep_counts

    recs = json.load(open(os.path.join(GOLDEN, "gen_sha256.json")))
    counts = sweep_counts()
    total = 0
    for r, n in enumerate(counts):
        rec = recs[f"S{r:02d}"]
        assert rec["argv"] == ["liposome", str(1000 + r), str(int(n)), "3.45"]
        h, nbeads = _generate_and_hash(rec, str(tmp_path / f"S{r:02d}"))
        assert h == rec["sha256"], r
        assert nbeads == rec["n_particles"]
        total += nbeads
    assert total == 2805567


def test_mpd_round_trip_is_byte_identical(golden_traj, tmp_path):
    for name in ("L500", "B400t"):
        src = tmp_path / f"{name}.mpd"
        src.write_bytes(golden_traj[name + "/final_mpd"].tobytes())
        m = sm.Mpd.read(str(tmp_path / name))
        m.write(str(tmp_path / (name + "_rt")))
        assert (tmp_path / f"{name}_rt.mpd").read_bytes() == src.read_bytes()


def test_mpd_system_view(tmp_path):
    m = sm.Mpd.bilayer(1234, 400, 3.45)
    d = m.system()
    assert d["pos"].shape == (1176, 3) and d["n_types"] == 6
    assert d["chain"].tolist() == [[0, 392, 3]]
    assert d["delta_lxy"] == 0.01 and d["tension"] == 0.0
    assert m["deltaLXY"] == 0.01 and m["nParticles"] == 1176
    m["tension"] = 0.5
    m.write(str(tmp_path / "t"))
    assert "\ntension 0.5 \n" in (tmp_path / "t.mpd").read_text()


def test_mpd_errors_mirror_reference(tmp_path):
    with pytest.raises(sm.SmError) as e:
        sm.Mpd.read(str(tmp_path / "missing"))
    assert e.value.code == sm.SM_ERR_IO
    good = sm.Mpd.liposome(1, 50, 3.45)
    good.write(str(tmp_path / "g"))
    text = (tmp_path / "g.mpd").read_text()
    (tmp_path / "bad.mpd").write_text(text.replace("cutoff 2", "cutof 2"))
    with pytest.raises(sm.SmError, match="is not a recognized command"):
        sm.Mpd.read(str(tmp_path / "bad"))
    (tmp_path / "bad2.mpd").write_text(text.replace("seed 1", "seed x"))
    with pytest.raises(sm.SmError, match="not a numerical type"):
        sm.Mpd.read(str(tmp_path / "bad2"))
    # a particle outside the box (Blob::errorChecking, system.h:452-469)
    lines = text.split("\n")
    i = next(k for k, l in enumerate(lines) if l.startswith("positions")) + 1
    f = lines[i].split()
    lines[i] = " " + " ".join([f[0], "401.5", f[2], f[3]])
    (tmp_path / "bad3.mpd").write_text("\n".join(lines))
    with pytest.raises(sm.SmError, match="out of bounds"):
        sm.Mpd.read(str(tmp_path / "bad3"))


def test_unsupported_molecules_are_refused(tmp_path):
    m = sm.Mpd.liposome(1, 50, 3.45)
    m.write(str(tmp_path / "g"))
    text = (tmp_path / "g.mpd").read_text()
    # turn the CHAIN record into a BOND record over the first two beads (supported)
    text = re.sub(r"\nmolecule \n8\t 1\n 0.7  100  1  100 \n 0  50  3 \n",
                  "\nmolecule \n6\t 1\n 0.7  100 \n 0  1 \n", text)
    (tmp_path / "b.mpd").write_text(text)
    b = sm.Mpd.read(str(tmp_path / "b"))
    d = b.system()
    assert d["bond"].tolist() == [[0, 1]] and d["bond_const"].tolist() == [[0.7, 100.0]]
    assert len(d["chain"]) == 0 and len(d["bend"]) == 0
    # a SOLID record (molecules.h:20) has no implementation here
    text = re.sub(r"\nmolecule \n6\t 1\n 0.7  100 \n 0  1 \n", "\nmolecule \n10\t 1\n 1 \n 0 \n", text)
    (tmp_path / "s.mpd").write_text(text)
    b = sm.Mpd.read(str(tmp_path / "s"))
    with pytest.raises(sm.SmError) as e:
        b.system()
    assert e.value.code == sm.SM_ERR_UNSUPPORTED


def _gamma_type_mpd(tmp_path, name, gammas):
    """A generated vesicle whose `gamma` line is replaced by a gammaType line."""
    m = sm.Mpd.liposome(4, 60, 3.45)
    m.write(str(tmp_path / "src"))
    m.close()
    text = (tmp_path / "src.mpd").read_text()
    line = "gammaType " + " ".join(f"{g:g}" for g in gammas) + " \n" if gammas else ""
    text = text.replace("gamma 1 \n", "").replace("nTypes 6 \n", "nTypes 6 \n" + line)
    (tmp_path / f"{name}.mpd").write_text(text)


def test_gamma_type_thermostat_view(tmp_path):
    """Without a scalar gamma, MD.cpp runs the gammaType thermostat (MD.cpp:135-138), whose
    compute() applies gammaType[0] to every particle (langevin.h:254) with sigma taken at the
    first temperature (langevin.h:235-241)."""
    _gamma_type_mpd(tmp_path, "t", [0.5, 2.0, 3.0, 4.0, 5.0, 6.0])
    d = sm.Mpd.read(str(tmp_path / "t")).system()
    assert d["gamma"] == 0.5 and d["sigma_temperature"] == 3.0
    d = sm.Mpd.liposome(4, 60, 3.45).system()
    assert d["gamma"] == 1.0 and d["sigma_temperature"] == 0.0


def test_no_gamma_stops_like_md(tmp_path):
    _gamma_type_mpd(tmp_path, "n", [])
    with pytest.raises(sm.SmError, match="No gamma available"):
        sm.Mpd.read(str(tmp_path / "n")).system()
    cli = os.path.join(os.path.dirname(sm.__file__), "md_hip")
    import subprocess
    r = subprocess.run([cli, "n"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout == "Error(main): No gamma available!\n"


def test_create_refuses_invalid_systems():
    """sm_create checks its input on the host before any device call, as the reference's
    Blob::errorChecking does before a run (system.h:438-580): a position outside
    [0, L] on any axis (system.h:452-469; NaN included), a type outside [0, nTypes), a CHAIN
    block or BOND record naming a particle past the end (system.h:490-535), and
    non-positive cutoff, dt, gamma or box are SM_ERR_INVALID; fewer than 3 cells along an
    axis is SM_ERR_UNSUPPORTED.  A position exactly at L passes the check."""
    d = sm.Mpd.bilayer(1234, 250, 3.45).system()
    n = len(d["type"])
    L = np.asarray(d["box"])

    def with_pos(i, axis, value):
        p = np.array(d["pos"], copy=True)
        p[i, axis] = value
        return dict(d, pos=p)

    bad_type = np.array(d["type"], copy=True)
    bad_type[7] = d["n_types"]
    bad_chain = np.array(d["chain"], copy=True)
    bad_chain[-1, 1] += 1                      # one chain more than the beads hold
    cases = {
        "x above L": (with_pos(3, 0, L[0] * (1 + 1e-12)), sm.SM_ERR_INVALID, "out of bounds"),
        "y below 0": (with_pos(4, 1, -1e-9), sm.SM_ERR_INVALID, "out of bounds"),
        "z NaN": (with_pos(5, 2, np.nan), sm.SM_ERR_INVALID, "out of bounds"),
        "type": (dict(d, type=bad_type), sm.SM_ERR_INVALID, "type outside"),
        "chain": (dict(d, chain=bad_chain), sm.SM_ERR_INVALID, "CHAIN block out of range"),
        "bond": (dict(d, bond=np.array([[0, n]], np.int32), bond_const=np.array([[0.7, 100.0]])),
                 sm.SM_ERR_INVALID, "out of range"),
        "cutoff": (dict(d, cutoff=0.0), sm.SM_ERR_INVALID, "cutoff"),
        "dt": (dict(d, dt=-0.01), sm.SM_ERR_INVALID, "dt"),
        "gamma": (dict(d, gamma=0.0), sm.SM_ERR_INVALID, "gamma"),
        "box": (dict(d, box=np.array([L[0], -L[1], L[2]])), sm.SM_ERR_INVALID, "box"),
        "cells": (dict(d, box=np.array([L[0], L[1], 5.9]), pos=np.minimum(d["pos"], [L[0], L[1], 5.9])),
                  sm.SM_ERR_UNSUPPORTED, "fewer than 3 cells"),
    }
    for name, (e, code, text) in cases.items():
        with pytest.raises(sm.SmError) as err:
            sm.Solver.from_system(e)
        assert err.value.code == code and text in str(err.value), (name, str(err.value))
    # at exactly L the check passes (here the device step may then fail for want of a GPU)
    try:
        sm.Solver.from_system(with_pos(3, 0, L[0]))
    except sm.SmError as err:
        assert err.code != sm.SM_ERR_INVALID, str(err)
