"""Observables tooling: vesicle radius, area per lipid, and the `.mpd` tension editor.

The reference computes the vesicle radius in analysis/vesicleRadius.cpp and edits the
tension with analysis/changeTension.cpp.  Their replacements here:

- the vesicle_radius and change_tension programs, host-only, run on the CPU here;
- sm_vesicle_radius, the same per-frame computation on device-resident state, checked
  on the GPU against the oracle's restatement.

The oracle's radius (orc_vesicle_radius) is pinned by closed forms (heads on a sphere), by
a hand-evaluated case of the reference's image rules, and by the liposome generator's
known geometry.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import softmold_amd as sm
from conftest import gpu_available
import oracle as orc
from softmold_amd import observables as obs

HERE = os.path.dirname(sm.__file__)
VR = os.path.join(HERE, "vesicle_radius")
CT = os.path.join(HERE, "change_tension")
HEAD = 2


def _sphere(n, R, centre):
    """n points on a sphere (golden-angle spiral)."""
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    r = np.sqrt(1.0 - z * z)
    phi = k * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([centre[0] + R * r * np.cos(phi), centre[1] + R * r * np.sin(phi), centre[2] + R * z], 1)


# ---- the oracle ------------------------------------------------------------------------

def test_oracle_radius_of_a_sphere():
    p = _sphere(2000, 7.5, (20.0, 21.0, 19.0))
    t = np.full(len(p), HEAD, np.int32)
    # the heads' centre of mass is not exactly the sphere centre for a finite spiral, so
    # compare with the mean distance from the actual centre of mass
    com = p.mean(0)
    want = np.linalg.norm(p - com, axis=1).mean()
    assert orc.vesicle_radius(p, t, [40.0, 40.0, 40.0]) == pytest.approx(want, rel=1e-12)
    assert want == pytest.approx(7.5, rel=1e-3)
    # other types are ignored
    q = np.vstack([p, [[1.0, 1.0, 1.0], [39.0, 2.0, 3.0]]])
    tq = np.concatenate([t, [3, 1]]).astype(np.int32)
    assert orc.vesicle_radius(q, tq, [40.0, 40.0, 40.0]) == pytest.approx(want, rel=1e-12)


def test_oracle_radius_centre_of_mass_images_towards_particle_0():
    """A sphere cut by the x boundary: the centre-of-mass loop images the far heads back
    next to particle 0, so the centre is right; the radius loop then sees heads within half
    a box of it on the near side only, and ADDS the image term for the wrapped ones."""
    L = 40.0
    p = _sphere(500, 5.0, (1.0, 20.0, 20.0))
    order = np.argsort(-p[:, 0])          # particle 0 on the +x side (not wrapped)
    p = p[order]
    w = p.copy()
    w[:, 0] = np.mod(w[:, 0], L)
    t = np.full(len(p), HEAD, np.int32)
    com = p.mean(0)
    # hand restatement of the radius loop: d = x - com; d += (d > L/2 ? L : d < -L/2 ? -L : 0)
    d = w - com
    d += np.where(d > L / 2, L, np.where(d < -L / 2, -L, 0.0))
    want = np.linalg.norm(d, axis=1).mean()
    got = orc.vesicle_radius(w, t, [L, L, L])
    assert got == pytest.approx(want, rel=1e-12)
    assert got > 10.0          # the added image term inflates it, as in the reference


def test_oracle_radius_hand_case():
    """Two heads at x = 0.5 and 9.5 in a box of 10: the centre-of-mass loop moves the second
    to -0.5 (centre 0); the radius loop gives |0.5| and |9.5 + 10| = 19.5, mean 10."""
    p = np.array([[0.5, 0.0, 0.0], [9.5, 0.0, 0.0]])
    assert orc.vesicle_radius(p, np.array([HEAD, HEAD], np.int32), [10.0, 10.0, 10.0]) == 10.0


def test_liposome_leaflets_match_the_generator():
    assert obs.liposome_leaflets(5000, 3.45) == (1812, 3188)
    for n in (500, 2000, 5000):
        m = sm.Mpd.liposome(1234, n, 3.45)
        t = m.system()["type"]
        m.close()
        # an inner lipid has its head first, an outer lipid last (lipidModels.h:1400,1445)
        inner = int(np.sum(t[0::3] == HEAD))
        assert obs.liposome_leaflets(n, 3.45) == (inner, n - inner)


def test_liposome_radius_from_generator_geometry():
    """Generated heads sit on two spheres: inner heads at R, outer heads at R + 5 b
    (lipidModels.h:1396-1400,1440-1445), both centred on the box centre."""
    n = 2000
    m = sm.Mpd.liposome(1234, n, 3.45)
    d = m.system()
    m.close()
    R = math.sqrt((n / 3.45) / (4.0 * math.pi))
    inner, outer = obs.liposome_leaflets(n, 3.45)
    want = (inner * R + outer * (R + 5 * 0.7)) / n
    got = orc.vesicle_radius(d["pos"], d["type"], d["box"])
    assert got == pytest.approx(want, rel=2e-3)


def test_area_per_lipid_formulas():
    assert obs.bilayer_area_per_lipid([10.0, 20.0, 40.0], 100) == pytest.approx(4.0)
    rows = np.array([[10.0, 20.0, 40.0], [5.0, 5.0, 40.0]])
    assert np.allclose(obs.bilayer_area_per_lipid(rows, 50), [8.0, 1.0])
    assert obs.vesicle_area_per_lipid(10.0, 100) == pytest.approx(4.0 * math.pi)


# ---- the programs (host only) ------------------------------------------------------------

def test_programs_built():
    assert os.path.exists(VR) and os.path.exists(CT)


def test_vesicle_radius_usage(tmp_path):
    r = subprocess.run([VR], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage:" in r.stderr


def _write_frames(path, frames, types):
    with open(path, "w") as f:
        for p in frames:
            f.write(f"{len(p)}\ntest\n")
            for t, (x, y, z) in zip(types, p):
                f.write(f"{t}\t{float(x)!r}\t{float(y)!r}\t{float(z)!r}\n")


def test_vesicle_radius_program_matches_oracle(tmp_path):
    m = sm.Mpd.liposome(77, 600, 3.45)
    m["storeInterval"] = 2.5
    m.write(str(tmp_path / "ves"))
    d = m.system()
    m.close()
    rng = np.random.default_rng(5)
    frames = [d["pos"] + rng.normal(0.0, 0.3, d["pos"].shape) for _ in range(3)]
    frames = [np.mod(p, d["box"]) for p in frames]
    _write_frames(tmp_path / "frames_ves.xyz", frames, d["type"])
    # the box changes between frames; the size file has a line per 1.25 time units
    boxes = [d["box"] * s for s in (1.0, 1.0, 1.01, 1.01, 0.99, 0.99)]
    with open(tmp_path / "size_ves.dat", "w") as f:
        for k, b in enumerate(boxes):
            f.write(f"{k * 1.25!r}\t{float(b[0])!r}\t{float(b[1])!r}\t{float(b[2])!r}\n")
    r = subprocess.run([VR, "ves"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    assert len(lines) == 3
    for k, line in enumerate(lines):
        t = 2.5 * k
        box = boxes[2 * k]          # first size line at or after the frame's time
        want = orc.vesicle_radius(frames[k], d["type"], box)
        assert line == f"{t:g} {want:g}"
    assert "Read frame 2 with time 5 tau." in r.stderr


def test_change_tension_program(tmp_path):
    m = sm.Mpd.bilayer(1234, 400, 3.45)
    m.write(str(tmp_path / "b"))
    m.close()
    r = subprocess.run([CT, "b", "0.75"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    m = sm.Mpd.read(str(tmp_path / "b"))
    assert m["tension"] == 0.75
    # the same edit through the library writes the same bytes
    m.write(str(tmp_path / "c"))
    m.close()
    assert (tmp_path / "b.mpd").read_bytes() == (tmp_path / "c.mpd").read_bytes()
    # tension 0 drops the key, as Blob::setTension does (system.h:4539-4546)
    r = subprocess.run([CT, "b", "0"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "tension" not in (tmp_path / "b.mpd").read_text()
    r = subprocess.run([CT, "b"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "name tension" in r.stderr


# ---- on the device -------------------------------------------------------------------------

@pytest.mark.gpu
def test_device_vesicle_radius_matches_oracle():
    if not gpu_available():
        pytest.skip("no GPU")
    systems = [sm.Mpd.liposome(s, n, 3.45).system() for s, n in ((3, 900), (4, 2500), (5, 1600))]
    b = sm.Solver.batch(systems)
    b.init_forces()
    for steps in (0, 40):
        if steps:
            b.run(1, steps)
        pos = b.download()[0]
        got = b.vesicle_radius()
        for r, d in enumerate(systems):
            lo, hi = b.first[r], b.first[r] + b.count[r]
            want = orc.vesicle_radius(pos[lo:hi], d["type"], d["box"])
            assert got[r] == pytest.approx(want, rel=1e-12, abs=0), (steps, r)
    s = sm.Solver.from_system(systems[1])
    s.init_forces()
    s.run(1, 40)
    assert s.vesicle_radius() == got[1]
