"""The oracle's VortexBlobs::merge_blobs (src/VortexBlobs.cpp:110-209), pinned
by closed-form cases (two-blob merges of either sign, the cancellation skip),
cases where the greedy index order and the use of the absorbing blob's
updated state decide the result, and invariants on random sets (circulation
and its first moment conserved, no surviving untouched pair under the
threshold, IDs renumbered).  The reference has no test of its own for it."""
import numpy as np
import pytest

import oracle


def _state(x, z, circ, sigma=None):
    n = len(x)
    st = {"id": np.arange(n, dtype=np.uint32), "x": np.array(x, float), "z": np.array(z, float),
          "circ": np.array(circ, float),
          "sigma": np.array(sigma if sigma is not None else [0.05] * n, float)}
    for k in ("u", "w", "uvs", "wvs"):
        st[k] = np.linspace(1.0, 2.0, n)
    return st


def test_two_same_sign_blobs_closed_form():
    # crit = |G1 G2 / (G1 + G2)| d = (2*3/5) * 0.5 = 0.6
    st = _state([0.0, 0.3], [0.0, 0.4], [2.0, 3.0], [0.1, 0.2])
    k, out = oracle.merge_blobs(0.6, st)  # not strictly below
    assert k == 0 and out["x"].size == 2 and np.array_equal(out["u"], st["u"])
    k, out = oracle.merge_blobs(0.6000001, st)
    assert k == 1 and out["x"].size == 1
    assert out["circ"][0] == 5.0
    assert out["x"][0] == pytest.approx((2 * 0.0 + 3 * 0.3) / 5, abs=1e-16)
    assert out["z"][0] == pytest.approx((2 * 0.0 + 3 * 0.4) / 5, abs=1e-16)
    assert out["sigma"][0] == pytest.approx((2 * 0.1 + 3 * 0.2) / 5, abs=1e-16)
    assert out["u"][0] == 0.0 and out["w"][0] == 0.0 and out["uvs"][0] == 0.0
    assert out["wvs"][0] == 0.0 and out["id"][0] == 0


def test_opposite_sign_blobs_closed_form():
    # G = 1, -0.5: |G1 G2 / (G1 + G2)| = 1, so crit = d = 1; the merged blob
    # sits at the first moment (x1 - 0.5 x2) / 0.5 = 2 x1 - x2, outside the pair
    st = _state([1.0, 2.0], [0.0, 0.0], [1.0, -0.5], [0.1, 0.4])
    assert oracle.merge_blobs(1.0, st)[0] == 0
    k, out = oracle.merge_blobs(1.01, st)
    assert k == 1
    assert out["circ"][0] == 0.5 and out["x"][0] == 0.0 and out["z"][0] == 0.0
    assert out["sigma"][0] == pytest.approx((1 * 0.1 + 0.5 * 0.4) / 1.5, abs=1e-16)


def test_cancelling_pair_is_skipped():
    st = _state([0.0, 1e-9], [0.0, 0.0], [1e-3, -1e-3])
    assert oracle.merge_blobs(1e6, st)[0] == 0
    # below the 1e-15 guard too
    st = _state([0.0, 1e-9], [0.0, 0.0], [1.0, -(1.0 - 1e-16)])
    assert oracle.merge_blobs(1e6, st)[0] == 0


def test_greedy_index_order_decides():
    # pairs (0,1): crit 0.5, (1,2): 0.45 (closer), threshold 0.7.  Index order
    # merges 1 into 0 first; the merged blob (0.5, G=2) is then (2/3)*1.4 from
    # blob 2, above the threshold, so blob 2 survives untouched.
    st = _state([0.0, 1.0, 1.9], [0, 0, 0], [1.0, 1.0, 1.0])
    k, out = oracle.merge_blobs(0.7, st)
    assert k == 1
    assert np.allclose(out["x"], [0.5, 1.9]) and np.array_equal(out["circ"], [2.0, 1.0])
    assert np.array_equal(out["id"], [0, 1])
    assert out["u"][1] == st["u"][2]  # survivor fields move with it


def test_absorbing_blob_tests_with_its_updated_state():
    # blob 2 is 0.8 from blob 0's original state (crit 0.8 > 0.75) but
    # (2/3)*1.1 = 0.733 from the merged (0.5, G=2): it is absorbed too
    st = _state([0.0, 1.0, 1.6], [0, 0, 0], [1.0, 1.0, 1.0])
    k, out = oracle.merge_blobs(0.75, st)
    assert k == 2 and out["x"].size == 1 and out["circ"][0] == 3.0
    assert out["x"][0] == pytest.approx((2 * 0.5 + 1.6) / 3, abs=1e-15)


def _crit(gi, xi, zi, gj, xj, zj):
    gs = gi + gj
    if abs(gs) < 1e-15:
        return np.inf
    return abs(gi * gj / gs) * np.sqrt((xi - xj) ** 2 + (zi - zj) ** 2)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_sets_invariants(seed):
    rng = np.random.default_rng(seed)
    n = 400
    x, z = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    c = rng.normal(0, 1e-3, n)
    c[::17] = 0.0  # zero-circulation blobs merge with anything
    st = _state(x, z, c, rng.uniform(0.01, 0.05, n))
    thr = 2e-5
    k, out = oracle.merge_blobs(thr, st)
    m = out["x"].size
    assert 0 < k < n and m == n - k
    assert np.array_equal(out["id"], np.arange(m))
    assert abs(out["circ"].sum() - c.sum()) <= 1e-14 * np.abs(c).sum()
    assert abs((out["circ"] * out["x"]).sum() - (c * x).sum()) <= 1e-12 * np.abs(c * x).sum()
    # untouched survivors (u unchanged) keep their order and no two of them
    # are under the threshold (a blob that absorbed others may be: it tested
    # a later blob before its last absorption moved it)
    untouched = np.nonzero(out["u"] != 0.0)[0]
    src = np.searchsorted(st["u"], out["u"][untouched])
    assert np.all(np.diff(src) > 0)
    assert np.array_equal(out["x"][untouched], x[src])
    for ia, a in enumerate(untouched):
        for b in untouched[ia + 1:]:
            assert _crit(out["circ"][a], out["x"][a], out["z"][a], out["circ"][b],
                         out["x"][b], out["z"][b]) >= thr
    # each absorbing survivor accounts for at least one of the k merges
    assert (out["u"] == 0.0).sum() <= k
    # sigma of merged blobs stays inside the inputs' range (a |G|-weighted mean)
    assert out["sigma"].min() >= st["sigma"].min() - 1e-18
    assert out["sigma"].max() <= st["sigma"].max() + 1e-18


def test_fewer_than_two_blobs():
    assert oracle.merge_blobs(1.0, _state([0.0], [0.0], [1.0]))[0] == 0
    assert oracle.merge_blobs(1.0, _state([], [], []))[0] == 0
