"""Pins the FP64 C oracle (oracle/md_oracle.c) before it is trusted as the checker.

Each oracle function is checked against at least one of: golden vectors written by the
compiled reference (tests/golden/make_golden.py), a library routine or published
known-answer vectors, closed forms, finite differences, invariants, or brute force.
"""
import os
import tempfile

import numpy as np
import pytest

import oracle as O

CUT = 2.0


@pytest.fixture(autouse=True, scope="module")
def single_thread():
    O.set_threads(1)  # the reference's single-thread arithmetic order
    yield


# step-0 fixture name -> the trajectory fixture whose input holds the same system
SRC = {"L500": "L500", "B400": "B400t", "A500": "A500", "AB400": "AB400t"}
# A500 and AB400 are L500 and B400 with two-body tables whose (2, 3) and (3, 2) rows differ
STEP0 = ["L500", "B400", "A500", "AB400"]


def _system(golden_traj, name):
    d = tempfile.mkdtemp()
    path = os.path.join(d, "in.mpd")
    with open(path, "wb") as f:
        f.write(golden_traj[name + "/input_mpd"].tobytes())
    return O.System.from_mpd(path), O.read_mpd(path)


def _final(golden_traj, name):
    d = tempfile.mkdtemp()
    path = os.path.join(d, "out.mpd")
    with open(path, "wb") as f:
        f.write(golden_traj[name + "/final_mpd"].tobytes())
    return O.read_mpd(path)


# ---------------------------------------------------------------------------------------
# random number generators
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1234, 4294967295])
def test_mt19937_rand53_matches_numpy(seed):
    m = O.mt_stream(seed)
    rs = np.random.RandomState(seed)  # init_genrand + genrand_res53
    ours = np.array([O.mt_rand53(m) for _ in range(2000)])
    np.testing.assert_array_equal(ours, rs.random_sample(2000))


def test_mt19937_published_10000th_output():
    m = O.mt_stream(5489)
    for _ in range(9999):
        O.mt_u32(m)
    assert O.mt_u32(m) == 4123659995


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox4x32_10_known_answers(ctr, key, expect):
    assert O.philox4x32_10(ctr, key) == expect


def test_psi_mapping_exact_in_fp32_and_moments():
    vals = np.array([O.philox_langevin(77, 5, p) for p in range(20000)]).ravel()
    assert np.array_equal(vals.astype(np.float32).astype(np.float64), vals)
    assert vals.min() > -1.0 and vals.max() < 1.0
    assert abs(vals.mean()) < 0.01
    assert abs(vals.var() - 1.0 / 3.0) < 0.01
    a = np.array([O.philox_langevin(77, 5, p) for p in range(100)])
    b = np.array([O.philox_langevin(77, 6, p) for p in range(100)])
    assert not np.any(a == b)  # the step enters the counter


def test_mc_uniform_moments():
    u = np.array([O.philox_mc(3, s)[1] for s in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01


# ---------------------------------------------------------------------------------------
# cells and the pair set
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", STEP0)
def test_cell_keys_match_reference(golden_step0, golden_traj, name):
    meta, arr = golden_step0
    s, _ = _system(golden_traj, SRC[name])
    keys = s.cell_keys()
    np.testing.assert_array_equal(keys, arr[f"{name}/keys"])
    nc, cs = s.cell_grid()
    assert list(nc) == [int(x) for x in meta[name]["nCells"]]
    np.testing.assert_array_equal(cs, meta[name]["cellSize"])
    # closed form, vectorised independently
    c = np.minimum((s.pos / cs).astype(np.int64), nc - 1)
    np.testing.assert_array_equal(keys, c[:, 0] + c[:, 1] * nc[0] + c[:, 2] * nc[0] * nc[1])


def _brute_pairs(pos, box):
    d = pos[:, None, :] - pos[None, :, :]
    d -= box * np.round(d / box)
    r2 = (d ** 2).sum(-1)
    i, j = np.nonzero(np.triu(r2 < CUT * CUT, 1))
    return set(zip(i.tolist(), j.tolist()))


@pytest.mark.parametrize("name", STEP0)
def test_pair_counts_match_reference_force_count(golden_step0, golden_traj, name):
    _, arr = golden_step0
    s, _ = _system(golden_traj, SRC[name])
    np.testing.assert_array_equal(s.pair_count(), arr[f"{name}/counts"])


@pytest.mark.parametrize("name", ["L500", "B400t"])
def test_active_pair_count_equals_brute_force(golden_traj, name):
    """The oracle's count of pairs with a nonzero force term (the set the device force
    kernel's own count is compared with) against an O(N^2) evaluation of its definition:
    r^2 < rc^2 and (r^2 < r_m^2 or an outer coefficient of the row nonzero)."""
    s, _ = _system(golden_traj, name)
    nT = int(round(np.sqrt(len(s.fconst) / 6)))
    F = s.fconst.reshape(nT * nT, 6)
    d = s.pos[:, None, :] - s.pos[None, :, :]
    d -= s.box * np.round(d / s.box)
    r2 = (d ** 2).sum(-1)
    row = s.type[:, None] * nT + s.type[None, :]
    outer = (F[row, 4] != 0.0) | (F[row, 5] != 0.0)
    act = (r2 < CUT * CUT) & (outer | (r2 < F[row, 0] ** 2))
    np.fill_diagonal(act, False)
    got = s.pair_count_active()
    np.testing.assert_array_equal(got, act.sum(1))
    # a strict subset of the FORCE_COUNT pairs: head pairs between r_m and rc drop out
    allc = s.pair_count()
    assert np.all(got <= allc) and got.sum() < allc.sum()


@pytest.mark.parametrize("name", ["L500", "B400t"])
def test_pair_list_equals_brute_force(golden_traj, name):
    s, _ = _system(golden_traj, name)
    pl = s.pair_list()
    assert len(pl) == len(set(map(tuple, pl.tolist())))
    assert set(map(tuple, pl.tolist())) == _brute_pairs(s.pos, s.box)


# ---------------------------------------------------------------------------------------
# forces and energies
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", STEP0)
def test_step0_forces_bit_exact_vs_reference(golden_step0, golden_traj, name):
    _, arr = golden_step0
    s, _ = _system(golden_traj, SRC[name])
    np.testing.assert_array_equal(s.pair_force(), arr[f"{name}/fpair"])
    np.testing.assert_array_equal(s.chain_force(), arr[f"{name}/fchain"])


@pytest.mark.parametrize("name", STEP0)
def test_energies_and_dpotential_match_reference(golden_step0, golden_traj, name):
    meta, _ = golden_step0
    m = meta[name]
    s, _ = _system(golden_traj, SRC[name])
    scale = meta["scale"]
    assert s.pair_potential() == m["U_pair"]
    assert s.chain_potential() == pytest.approx(m["U_chain"], rel=1e-9, abs=1e-24)
    assert s.pair_dpotential(scale) == pytest.approx(m["dU_pair"], rel=1e-12, abs=1e-12)
    assert s.chain_dpotential(scale) == pytest.approx(m["dU_chain"], rel=1e-9, abs=1e-15)


def test_pair_force_is_minus_gradient_of_potential(golden_traj):
    s, _ = _system(golden_traj, "B400t")
    f = s.pair_force()
    h = 1e-5
    rng = np.random.default_rng(1)
    for i in rng.choice(s.n, 6, replace=False):
        for d in range(3):
            p0 = s.pos[i, d]
            s.pos[i, d] = p0 + h
            up = s.pair_potential()
            s.pos[i, d] = p0 - h
            um = s.pair_potential()
            s.pos[i, d] = p0
            assert -(up - um) / (2 * h) == pytest.approx(f[i, d], rel=1e-5, abs=1e-5)


def test_chain_force_is_minus_gradient_and_sums_to_zero(golden_traj):
    s, _ = _system(golden_traj, "L500")
    rng = np.random.default_rng(2)
    s.pos += rng.normal(0, 0.05, s.pos.shape)  # bend and stretch the chains
    f = s.chain_force()
    np.testing.assert_allclose(f.sum(0), 0.0, atol=1e-9)
    h = 1e-6
    for i in rng.choice(s.n, 8, replace=False):
        for d in range(3):
            p0 = s.pos[i, d]
            s.pos[i, d] = p0 + h
            up = s.chain_potential()
            s.pos[i, d] = p0 - h
            um = s.chain_potential()
            s.pos[i, d] = p0
            assert -(up - um) / (2 * h) == pytest.approx(f[i, d], rel=1e-5, abs=1e-6)


def test_chain_closed_form():
    # one 3-bead chain: bond 1 stretched to 1.0, bond 2 at rest, a right angle
    pos = np.array([[10.0, 10.0, 10.0], [11.0, 10.0, 10.0], [11.0, 10.7, 10.0]])
    fc = np.zeros(6 * 6 * 6)
    s = O.System(pos, [2, 3, 3], np.zeros((3, 3)), [30, 30, 30], fc, fc, 6, [[0, 1, 3]],
                 [[0.7, 100.0, 1.0, 100.0]])
    # bond: 1/2 k (r - l0)^2 = 50 * 0.09; bend: 1/2 k (1 - cos 90)^2 = 50
    assert s.chain_potential() == pytest.approx(0.5 * 100 * 0.3 ** 2 + 50.0, rel=1e-14)


def test_potential_constants_continuity(golden_traj):
    _, m = _system(golden_traj, "L500")
    U, F = m["twoBodyUconst"].reshape(-1, 6), m["twoBodyFconst"].reshape(-1, 6)
    for u, f in zip(U, F):
        rm, rc = u[0], u[3]
        inner = lambda r: u[1] * (rm - r) ** 2 + u[2]
        outer = lambda r: (rc - r) ** 2 * (u[4] - (rc - r) * u[5])
        assert inner(rm) == pytest.approx(outer(rm), abs=1e-12)          # U continuous at r_m
        assert outer(rc) == 0.0                                          # U(r_c) = 0
        # F row is -dU/dr: inner 2 C1u (rm - r) == C1f (rm - r); outer -(dU/dr)
        assert f[1] == pytest.approx(2 * u[1])
        r = 1.5
        dudr = -2 * (rc - r) * u[4] + 3 * (rc - r) ** 2 * u[5]
        assert (f[4] - f[5] * (rc - r)) * (rc - r) == pytest.approx(-dudr, abs=1e-12)


def test_dpotential_equals_direct_recompute(golden_traj):
    s, _ = _system(golden_traj, "B400t")
    a = np.array([1.001, 1.001, 1.0 / 1.001 ** 2])
    du = s.pair_dpotential(a) + s.chain_dpotential(a)
    u0 = s.pair_potential() + s.chain_potential()
    s.pos *= a
    box = s.box * a
    s.s.box[0], s.s.box[1], s.s.box[2] = box
    u1 = s.pair_potential() + s.chain_potential()
    assert du == pytest.approx(u0 - u1, rel=1e-9, abs=1e-9)


# ---------------------------------------------------------------------------------------
# integrator, thermostat, the whole loop
# ---------------------------------------------------------------------------------------
def test_verlet_exact_for_constant_force():
    n = 4
    pos = np.full((n, 3), 20.0)
    vel = np.array([[0.1, 0.2, 0.3]] * n)
    fc = np.zeros(6 * 36)
    s = O.System(pos, [1, 2, 3, 0], vel, [40, 40, 40], fc, fc, 6, np.zeros((0, 3)), np.zeros((0, 4)))
    a = np.array([0.5, -0.25, 0.125])
    s.acc[:] = a
    for _ in range(10):
        O.lib().orc_verlet_first(s.ref())
        s.acc[:] = a
        O.lib().orc_verlet_second(s.ref())
    t = 10 * 0.02
    np.testing.assert_allclose(s.pos[:3], 20.0 + vel[:3] * t + 0.5 * a * t * t, rtol=1e-13)
    np.testing.assert_allclose(s.vel[:3], vel[:3] + a * t, rtol=1e-13)
    np.testing.assert_array_equal(s.pos[3], [20.0, 20.0, 20.0])  # type 0 does not move


def test_langevin_noise_moments():
    n = 30000
    fc = np.zeros(6 * 36)
    s = O.System(np.full((n, 3), 5.0), np.ones(n, np.int32), np.zeros((n, 3)), [10, 10, 10], fc, fc,
                 6, np.zeros((0, 3)), np.zeros((0, 4)), gamma=1.0, temp=3.0, dt=0.02)
    O.lib().orc_langevin_philox(s.ref(), 11, 3.0)
    sigma2 = 6 * 3.0 * 1.0 / 0.02
    assert abs(s.acc.mean()) < 0.02 * np.sqrt(sigma2)
    assert s.acc.var() == pytest.approx(sigma2 / 3, rel=0.02)


@pytest.mark.parametrize("name", ["L500", "B400t", "A500", "AB400t"])
def test_md_loop_reproduces_reference_binary(golden_traj, name):
    """The compiled MD binary (one thread) vs the oracle loop with the reference's MT19937
    streams: same state after N steps, same measured series, same MC decisions."""
    s, meta = _system(golden_traj, name)
    steps = {"L500": 20, "B400t": 40, "A500": 20, "AB400t": 40}[name]
    md = O.MD(s, rng_mode=0, mt_streams=1)
    md.init_forces()
    U0, K0 = md.measure()
    pot, kin = golden_traj[name + "/potential"], golden_traj[name + "/kinetic"]
    assert float(f"{U0:.6g}") == pot[0, 1] and float(f"{K0:.6g}") == kin[0, 1]
    assert K0 == pytest.approx(1.5 * s.n * meta["initialTemp"], rel=1e-14)  # |v| = sqrt(3T)
    measured = []
    for i in range(steps):
        md.run(i, i)
        if i % 10 == 0 and i != 0:
            measured.append((i, *md.measure()))
    O.lib().orc_verlet_first(s.ref())  # the store happens after Verlet::first of the last step
    final = _final(golden_traj, name)
    np.testing.assert_allclose(s.pos, final["pos"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(s.vel, final["vel"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(s.box, final["size"], rtol=1e-14)
    for (i, U, K), (tp, up), (tk, kk) in zip(measured, pot[1:], kin[1:]):
        assert float(f"{U:.6g}") == up and float(f"{K:.6g}") == kk
    if name + "/resizeHist" in golden_traj:
        hist = golden_traj[name + "/resizeHist"]
        accepted = int(round(hist[:, 1].sum()))
        assert md.trials() == 4 and bin(md.decisions() & 0xF).count("1") == accepted
        if name == "B400t":
            assert accepted == 2


def _owner_rows(s):
    """Per pair within the cutoff (i < j, minimum image): the row index of the particle the
    reference's half-shell traversal makes first, restated without cells -- across cells
    the one whose forward half shell (cellOpt.h:715) holds the other's cell, within a cell
    the lower index (its list runs in descending index, cellOpt.h:572-585)."""
    nc, cs = s.cell_grid()
    c = np.minimum((s.pos / cs).astype(np.int64), nc - 1)
    nb = [(1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1), (-1, 0, 1), (0, -1, 1),
          (-1, 1, 1), (1, -1, 1), (-1, -1, 1), (1, -1, 0)]
    fwd = {tuple(x) for x in nb}
    d = s.pos[:, None, :] - s.pos[None, :, :]
    d -= s.box * np.round(d / s.box)
    r2 = (d ** 2).sum(-1)
    i, j = np.nonzero(np.triu(r2 < CUT * CUT, 1))
    off = (c[j] - c[i] + 1) % nc - 1          # j's cell seen from i's, each axis in {-1, 0, 1}
    same = np.all(off == 0, axis=1)
    i_first = same | np.array([tuple(o) in fwd for o in off.tolist()])
    return i, j, i_first, d[i, j]


@pytest.mark.parametrize("name", ["A500", "AB400"])
def test_asymmetric_table_rows_follow_traversal_order(golden_step0, golden_traj, name):
    """With (2, 3) and (3, 2) rows that differ, the reference's forces are those of each
    pair's first particle's row: an O(N^2) evaluation of that rule matches the compiled
    reference's step-0 forces, and the other rule (the second particle's row) does not."""
    _, arr = golden_step0
    s, _ = _system(golden_traj, SRC[name])
    nT = int(round(np.sqrt(len(s.fconst) / 6)))
    F = s.fconst.reshape(nT * nT, 6)
    i, j, i_first, dij = _owner_rows(s)
    assert i_first.any() and (~i_first).any()
    r = np.sqrt((dij ** 2).sum(1))

    def forces(first_is_i):
        row = np.where(first_is_i, s.type[i] * nT + s.type[j], s.type[j] * nT + s.type[i])
        C = F[row]
        o = np.where(r >= C[:, 0], 3, 0)
        c0, c1, c2 = (np.take_along_axis(C, (o + k)[:, None], 1)[:, 0] for k in range(3))
        m = c0 - r
        mag = (c1 - c2 * m) * m / r
        f = np.zeros_like(s.pos)
        np.add.at(f, i, dij * mag[:, None])
        np.add.at(f, j, -dij * mag[:, None])
        return f

    want = arr[f"{name}/fpair"]
    scale = np.abs(want).max()
    np.testing.assert_allclose(forces(i_first), want, rtol=0, atol=1e-10 * scale)
    assert np.abs(forces(~i_first) - want).max() > 1e-3 * scale

