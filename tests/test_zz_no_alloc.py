"""sm_vesicle_radius_noalloc: the vesicle-radius entry point that allocates no device memory.

Its partial sums live in a fixed per-GPU device array instead of a buffer allocated per call.
Bar: bit-identical to sm_vesicle_radius (integer-free FP64 work, same kernels, same order)
on two calls in a row, the second on a smaller input, and a clear error past its replica
limit.  Marked gpu: it runs on the MI355X.
"""
import numpy as np
import pytest

import softmold_amd as sm
from conftest import gpu_available

HEAD = 2


def _tiny(d: dict, lipid: int) -> dict:
    """One lipid of a generated vesicle, moved into a 10^3 box (5 cells per axis)."""
    p = np.asarray(d["pos"]).reshape(-1, 3)[3 * lipid:3 * lipid + 3]
    out = dict(d)
    out["pos"] = p - p[0] + 5.0
    out["type"] = np.asarray(d["type"])[3 * lipid:3 * lipid + 3]
    out["vel"] = np.asarray(d["vel"]).reshape(-1, 3)[3 * lipid:3 * lipid + 3]
    out["box"] = np.array([10.0, 10.0, 10.0])
    out["chain"] = np.array([[0, 1, 3]], np.int32)
    out["chain_const"] = np.asarray(d["chain_const"]).reshape(-1, 4)[:1]
    out["delta_lxy"] = 0.0
    return out


@pytest.mark.gpu
def test_noalloc_radius_equals_allocating_entry_point():
    if not gpu_available():
        pytest.skip("no GPU")
    # first call: three replicas of 2,700 to 7,500 beads, after a few steps
    systems = [sm.Mpd.liposome(s, n, 3.45).system() for s, n in ((3, 900), (4, 2500), (5, 1600))]
    b = sm.Solver.batch(systems)
    b.init_forces()
    b.run(1, 24)
    want = b.vesicle_radius(HEAD)
    got = b.vesicle_radius(HEAD, noalloc=True)
    assert np.all(np.isfinite(want))
    assert np.array_equal(got, want)
    b.close()
    # second call, smaller input: one 1,500-bead replica (the array still holds the
    # three-replica call's values past its first 8 words)
    s = sm.Solver.from_system(sm.Mpd.liposome(7, 500, 3.45).system())
    s.init_forces()
    s.run(1, 8)
    want1 = s.vesicle_radius(HEAD)
    got1 = s.vesicle_radius(HEAD, noalloc=True)
    assert np.isfinite(want1) and got1 == want1
    s.close()


@pytest.mark.gpu
def test_noalloc_radius_rejects_more_replicas_than_its_scratch():
    if not gpu_available():
        pytest.skip("no GPU")
    d = sm.Mpd.liposome(11, 200, 3.45).system()
    systems = [_tiny(d, k % 200) for k in range(1025)]
    b = sm.Solver.batch(systems)
    with pytest.raises(sm.SmError) as e:
        b.vesicle_radius(HEAD, noalloc=True)
    assert e.value.code == sm.SM_ERR_UNSUPPORTED and "1024" in str(e.value)
    r = b.vesicle_radius(HEAD)   # the allocating entry point takes any replica count
    assert r.shape == (1025,) and np.all(r == 0.0)
    b.close()
