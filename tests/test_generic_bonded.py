"""Generic BOND and BEND records (MDcuda.cu:162-184; system.h:1881-2035, 2718-2830,
3399-3530) on the oracle and on the device.

A CHAIN of length L is, term for term, L-1 bonds (k, k+1) and L-2 bends (k, k+1, k+2) with
the chain's constants (system.h:1782-1866).  Rewriting a system's CHAIN blocks as BOND and
BEND records therefore leaves every force and energy term unchanged; only the order in
which a bead's terms are summed differs.  The CHAIN path is pinned against the reference
elsewhere, so this equivalence pins the generic path; finite differences and Newton's
third law pin it on records that no CHAIN can express (bonds across lipids).
"""
import numpy as np
import pytest

import softmold_amd as sm
from conftest import gpu_available
import oracle as O


def _records(d):
    """The CHAIN blocks of system dict d as BOND and BEND records."""
    bonds, bc, bends, ec = [], [], [], []
    for (start, nch, length), c in zip(d["chain"], d["chain_const"]):
        for k in range(start, start + nch * length, length):
            for j in range(length - 1):
                bonds.append((k + j, k + j + 1))
                bc.append((c[0], c[1]))
            for j in range(length - 2):
                bends.append((k + j, k + j + 1, k + j + 2))
                ec.append((c[2], c[3]))
    e = dict(d)
    e["chain"] = np.zeros((0, 3), np.int32)
    e["chain_const"] = np.zeros((0, 4))
    e["bond"], e["bond_const"] = np.array(bonds, np.int32), np.array(bc)
    e["bend"], e["bend_const"] = np.array(bends, np.int32), np.array(ec)
    return e


def _crosslinked(d, n_links=60, seed=3):
    """d plus BOND records between the heads of random lipid pairs (l0 1.5, k 20) and BEND
    records over random head triples: terms no CHAIN record can express."""
    rng = np.random.default_rng(seed)
    heads = np.flatnonzero(d["type"] == 2)
    e = dict(d)
    b = rng.choice(heads, size=(n_links, 2), replace=True)
    b = b[b[:, 0] != b[:, 1]]
    t = rng.choice(heads, size=(n_links // 2, 3), replace=True)
    t = t[(t[:, 0] != t[:, 1]) & (t[:, 1] != t[:, 2]) & (t[:, 0] != t[:, 2])]
    e["bond"], e["bond_const"] = b.astype(np.int32), np.tile([1.5, 20.0], (len(b), 1))
    e["bend"], e["bend_const"] = t.astype(np.int32), np.tile([0.5, 10.0], (len(t), 1))
    return e


def _oracle(d, pos=None):
    return O.System(d["pos"] if pos is None else pos, d["type"], d["vel"], d["box"], d["fconst"], d["uconst"],
                    d["n_types"], d["chain"], d["chain_const"], cutoff=d["cutoff"], dt=d["dt"],
                    gamma=d["gamma"], temp=d["temperature"], delta_lxy=d["delta_lxy"], tension=d["tension"],
                    seed=d["seed"], bonds=d.get("bond"), bond_const=d.get("bond_const"), bends=d.get("bend"),
                    bend_const=d.get("bend_const"), balls=d.get("ball"), ball_const=d.get("ball_const"))


def _deformed(d, seed=7, amp=0.15):
    rng = np.random.default_rng(seed)
    return np.mod(d["pos"] + rng.uniform(-amp, amp, d["pos"].shape), d["box"])


# ---- oracle (CPU) --------------------------------------------------------------------

def test_oracle_bond_stats_match_reference_md_output(golden_traj, tmp_path):
    """oracle.bond_stats against the reference MD's own step-0 lBond_ and bend_ lines
    (dataExtraction.h:861-930,1009-1027) for the B400 bilayer with a BOND molecule over 80
    random head pairs and a BEND molecule over 40 random head triples; heads span the
    10.8-wide box, so many differences take the >= / <= image.  MD prints 6 significant
    digits."""
    path = tmp_path / "BS.mpd"
    path.write_bytes(golden_traj["BS/input_mpd"].tobytes())
    m = O.read_mpd(str(path))
    mols = {mol["type"]: np.array(mol["bonds"]) for mol in m["molecule"]}
    bonds, bends = mols[6], mols[7]
    assert len(bonds) == 80 and len(bends) == 40
    got = O.bond_stats(m["pos"], m["size"], bonds, bends)
    lb, bd = golden_traj["BS/lBond"], golden_traj["BS/bend"]
    assert lb[0, 0] == 0.0 and bd[0, 0] == 0.0
    np.testing.assert_allclose(got, [lb[0, 1], bd[0, 1], bd[0, 2], bd[0, 3]], rtol=6e-6, atol=0)
    # the imaging matters at this size: without it the mean bond length is far larger
    raw = np.linalg.norm(m["pos"][bonds[:, 0]] - m["pos"][bonds[:, 1]], axis=1).mean()
    assert raw > 1.05 * got[0]


def test_oracle_records_equal_chains():
    d = sm.Mpd.liposome(5, 300, 3.45).system()
    pos = _deformed(d)
    a, b = _oracle(d, pos), _oracle(_records(d), pos)
    fa, fb = a.chain_force(), b.chain_force()
    assert np.abs(fa - fb).max() <= 1e-12 * np.abs(fa).max()
    assert b.chain_potential() == pytest.approx(a.chain_potential(), rel=1e-13)
    s = [1.003, 0.998, 1.0 / (1.003 * 0.998)]
    assert b.chain_dpotential(s) == pytest.approx(a.chain_dpotential(s), rel=1e-10)


def test_oracle_records_obey_newton_and_gradient():
    d = _crosslinked(sm.Mpd.liposome(5, 300, 3.45).system())
    d["chain"] = np.zeros((0, 3), np.int32)       # the generic records alone
    d["chain_const"] = np.zeros((0, 4))
    pos = _deformed(d)
    o = _oracle(d, pos)
    f = o.chain_force()
    assert np.abs(f.sum(0)).max() <= 1e-10 * np.abs(f).max()
    # -dU/dx by central differences on a few particles of the records
    h = 1e-6
    for i in (int(d["bond"][0, 0]), int(d["bend"][0, 1]), int(d["bend"][1, 2])):
        for k in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, k] += h
            p2[i, k] -= h
            du = (_oracle(d, p1).chain_potential() - _oracle(d, p2).chain_potential()) / (2 * h)
            assert -du == pytest.approx(f[i, k], rel=1e-5, abs=1e-6)


def test_oracle_bond_closed_form():
    """One bond stretched to r: U = k (r - l0)^2 / 2, |F| = k (r - l0) along the bond."""
    d = sm.Mpd.liposome(5, 300, 3.45).system()
    d = dict(d)
    d["chain"] = np.zeros((0, 3), np.int32)
    d["chain_const"] = np.zeros((0, 4))
    pos = d["pos"].copy()
    pos[1] = pos[0] + np.array([0.9, 0.0, 0.0])
    d["bond"], d["bond_const"] = np.array([[0, 1]], np.int32), np.array([[0.7, 100.0]])
    o = _oracle(d, pos)
    assert o.chain_potential() == pytest.approx(0.5 * 100.0 * 0.2 ** 2, rel=1e-12)
    f = o.chain_force()
    # the stretched bond pulls particle 0 towards particle 1 (+x) and 1 towards 0
    assert f[0] == pytest.approx([100.0 * 0.2, 0.0, 0.0], rel=1e-12, abs=1e-12)
    assert f[1] == pytest.approx([-100.0 * 0.2, 0.0, 0.0], rel=1e-12, abs=1e-12)


def test_invalid_records_are_refused():
    """Checked on the host before any device call."""
    d = sm.Mpd.liposome(5, 100, 3.45).system()
    e = dict(d, bond=np.array([[0, 0]], np.int32), bond_const=np.array([[1.0, 1.0]]))
    with pytest.raises(sm.SmError, match="repeats"):
        sm.Solver.from_system(e)
    e = dict(d, bond=np.array([[0, len(d["type"])]], np.int32), bond_const=np.array([[1.0, 1.0]]))
    with pytest.raises(sm.SmError, match="out of range"):
        sm.Solver.from_system(e)


# ---- device ----------------------------------------------------------------------------


@pytest.mark.gpu
def test_device_records_match_oracle_and_chains():
    if not gpu_available():
        pytest.skip("no GPU")
    base = sm.Mpd.liposome(9, 1500, 3.45).system()
    pos = _deformed(base)
    for name, d in (("records", _records(base)), ("crosslinked", _crosslinked(base))):
        print(f"case: {name}", flush=True)
        s = sm.Solver.from_system(d)
        s.upload(pos=pos)
        _, fc = s.compute_forces()
        o = _oracle(d, pos)
        want = o.chain_force()
        err = np.linalg.norm(fc - want, axis=1)
        scale = np.maximum(np.linalg.norm(want, axis=1), 0.05 * np.sqrt((want ** 2).sum(1).mean()))
        assert (err / scale).max() <= 1e-5, name
        U = s.measure()
        assert U[1] == pytest.approx(o.chain_potential(), rel=1e-10, abs=1e-9), name
        sc = np.array([1.002, 0.999, 1.0 / (1.002 * 0.999)])
        dup, duc = s.potential_delta(sc)
        assert duc == pytest.approx(o.chain_dpotential(sc), rel=1e-9, abs=1e-9), name
        s.close()


@pytest.mark.gpu
def test_device_records_trajectory_equals_chain_trajectory_closely():
    """Steps with CHAIN blocks and with the same terms as records: the FP64 bonded forces
    differ only in summation order, so the FP32 accelerations agree to rounding and the
    trajectories stay together over a short run."""
    if not gpu_available():
        pytest.skip("no GPU")
    d = sm.Mpd.liposome(9, 1500, 3.45).system()
    a, b = sm.Solver.from_system(d), sm.Solver.from_system(_records(d))
    for s in (a, b):
        s.init_forces()
        s.run(1, 20)
    pa, pb = a.download()[0], b.download()[0]
    dx = np.abs(pa - pb)
    dx = np.minimum(dx, d["box"] - dx)
    assert dx.max() < 1e-3
    ua, ub = a.measure(), b.measure()
    assert ub == pytest.approx(ua, rel=1e-4)


@pytest.mark.gpu
def test_batch_with_records_equals_solo():
    if not gpu_available():
        pytest.skip("no GPU")
    x = _crosslinked(sm.Mpd.bilayer(11, 400, 3.45).system())
    y = sm.Mpd.bilayer(12, 300, 3.45).system()
    x.update(delta_lxy=0.05, tension=0.5)
    y.update(delta_lxy=0.05)
    bt = sm.Solver.batch([x, y])
    bt.init_forces()
    bt.run(1, 17)
    bp = bt.download()[0]
    bm = bt.measure()
    for r, d in enumerate((x, y)):
        s = sm.Solver.from_system(d)
        s.init_forces()
        s.run(1, 17)
        lo, hi = bt.first[r], bt.first[r] + bt.count[r]
        assert np.array_equal(s.download()[0], bp[lo:hi]), r
        assert np.array_equal(s.measure(), bm[r]), r
        s.close()


@pytest.mark.gpu
def test_device_bond_stats_match_oracle():
    if not gpu_available():
        pytest.skip("no GPU")
    base = sm.Mpd.liposome(9, 1500, 3.45).system()
    chain_only = sm.Solver.from_system(base)
    st = chain_only.bond_stats()
    assert np.isnan(st[0]) and list(st[1:]) == [0.0, 0.0, 0.0]     # no records: 0/0 and sums of nothing
    chain_only.close()
    d = _crosslinked(base)
    s = sm.Solver.from_system(d)
    s.init_forces()
    s.run(1, 30)
    pos = s.download()[0]
    want = O.bond_stats(pos, d["box"], d["bond"], d["bend"])
    assert s.bond_stats() == pytest.approx(want, rel=1e-12)
    s.close()


# ---- BALL records ---------------------------------------------------------------------

def _with_ball(d, n_members=40, r0=3.0, k=50.0, seed=5):
    """d plus one BALL: a head bead as the centre and the nearest other heads as members,
    so some sit inside r0 (repelled) and some outside (no force)."""
    rng = np.random.default_rng(seed)
    heads = np.flatnonzero(d["type"] == 2)
    c = int(rng.choice(heads))
    dist = np.linalg.norm(d["pos"][heads] - d["pos"][c], axis=1)
    members = heads[np.argsort(dist)[1:n_members + 1]]
    e = dict(d)
    e["ball"] = np.array([[c, m] for m in members], np.int32)
    e["ball_const"] = np.tile([r0, k], (len(members), 1))
    return e


def test_oracle_ball_closed_form():
    """One member at r < r0: U = k (r - r0)^2 / 2 and a push apart of k (r0 - r); a member
    at r > r0 contributes nothing (harmonicHalfF / harmonicHalfP, MD.h:1164-1222)."""
    d = dict(sm.Mpd.liposome(5, 300, 3.45).system())
    d["chain"], d["chain_const"] = np.zeros((0, 3), np.int32), np.zeros((0, 4))
    pos = d["pos"].copy()
    pos[1] = pos[0] + np.array([1.0, 0.0, 0.0])
    pos[2] = pos[0] + np.array([0.0, 3.5, 0.0])
    d["ball"], d["ball_const"] = np.array([[0, 1], [0, 2]], np.int32), np.array([[2.0, 50.0], [2.0, 50.0]])
    o = _oracle(d, pos)
    assert o.chain_potential() == pytest.approx(0.5 * 50.0 * 1.0, rel=1e-12)
    f = o.chain_force()
    assert f[0] == pytest.approx([-50.0, 0.0, 0.0], rel=1e-12, abs=1e-12)
    assert f[1] == pytest.approx([50.0, 0.0, 0.0], rel=1e-12, abs=1e-12)
    assert np.all(f[2] == 0.0)
    # the box move's dU leaves BALL terms out (MD.cpp:664)
    assert o.chain_dpotential([1.01, 0.99, 1.0 / (1.01 * 0.99)]) == 0.0


def test_oracle_ball_gradient():
    d = _with_ball(sm.Mpd.liposome(5, 300, 3.45).system())
    d["chain"], d["chain_const"] = np.zeros((0, 3), np.int32), np.zeros((0, 4))
    pos = d["pos"]
    o = _oracle(d, pos)
    f = o.chain_force()
    assert np.abs(f).max() > 1.0                       # some members are inside r0
    assert np.abs(f.sum(0)).max() <= 1e-10 * np.abs(f).max()
    h = 1e-6
    for i in (int(d["ball"][0, 0]), int(d["ball"][0, 1])):
        for k in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, k] += h
            p2[i, k] -= h
            du = (_oracle(d, p1).chain_potential() - _oracle(d, p2).chain_potential()) / (2 * h)
            assert -du == pytest.approx(f[i, k], rel=1e-5, abs=1e-6)


def test_ball_molecule_from_mpd(tmp_path):
    """A BALL molecule record in a .mpd file: the view takes the first record's centre for
    every member; records with another centre are refused."""
    import re
    m = sm.Mpd.liposome(1, 50, 3.45)
    m.write(str(tmp_path / "g"))
    text = (tmp_path / "g.mpd").read_text()
    # append a BALL as a second molecule after the CHAIN record
    text2 = re.sub(r"nMolecules 1 ", "nMolecules 2 ", text)
    i = text2.index("\nmolecule \n")
    j = text2.index("\n", text2.index(" 0  50  3 ", i)) + 1
    text2 = text2[:j] + "19\t 3\n 3  50 \n 0  3 \n 0  6 \n 0  9 \n" + text2[j:]
    (tmp_path / "b.mpd").write_text(text2)
    d = sm.Mpd.read(str(tmp_path / "b")).system()
    assert d["ball"].tolist() == [[0, 3], [0, 6], [0, 9]]
    assert d["ball_const"].tolist() == [[3.0, 50.0]] * 3
    (tmp_path / "c.mpd").write_text(text2.replace(" 0  6 \n", " 1  6 \n"))
    with pytest.raises(sm.SmError, match="centre"):
        sm.Mpd.read(str(tmp_path / "c")).system()


@pytest.mark.gpu
def test_device_ball_matches_oracle():
    if not gpu_available():
        pytest.skip("no GPU")
    d = _with_ball(sm.Mpd.liposome(9, 1500, 3.45).system())
    s = sm.Solver.from_system(d)
    _, fc = s.compute_forces()
    o = _oracle(d)
    want = o.chain_force()
    err = np.linalg.norm(fc - want, axis=1)
    scale = np.maximum(np.linalg.norm(want, axis=1), 0.05 * np.sqrt((want ** 2).sum(1).mean()))
    assert (err / scale).max() <= 1e-5
    U = s.measure()
    assert U[1] == pytest.approx(o.chain_potential(), rel=1e-10, abs=1e-9)
    sc = np.array([1.002, 0.999, 1.0 / (1.002 * 0.999)])
    assert s.potential_delta(sc)[1] == pytest.approx(o.chain_dpotential(sc), rel=1e-9, abs=1e-9)
    st = s.bond_stats()
    assert np.isnan(st[0]) and list(st[1:]) == [0.0, 0.0, 0.0]     # BALL is not a bond for lBond_
    s.close()


@pytest.mark.gpu
def test_run_equals_step_phases_with_records():
    """sm_run ends each step without a box move in the fused kick + next-first-half kernel;
    stepping phase by phase never fuses.  With BOND, BEND and BALL records and box moves
    every 8 steps, both give the same state bit for bit."""
    if not gpu_available():
        pytest.skip("no GPU")
    d = _with_ball(_crosslinked(sm.Mpd.bilayer(13, 400, 3.45).system()))
    d.update(delta_lxy=0.05, tension=0.5)
    a, b = sm.Solver.from_system(d), sm.Solver.from_system(d)
    a.init_forces()
    b.init_forces()
    a.run(0, 19)
    for i in range(19):
        b.step_first(i)
        b.step_finish(i)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)
    assert a.mc_stats() == b.mc_stats()
    a.close()
    b.close()
