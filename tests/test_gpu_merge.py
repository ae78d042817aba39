"""VortexBlobs::merge_blobs on the device (dvm_merge_blobs, merge.hip) against
the oracle's restatement of src/VortexBlobs.cpp:110-209: the same merges and
bit-identical survivors (the criterion and the merged state are evaluated in
FP64 as written), on sets where many merges chain, and invariants at sizes the
oracle's serial loop would take long on."""
import numpy as np
import pytest
import torch

import oracle
from conftest import to_dev
from dvmpp_amd import sim as dsim

pytestmark = pytest.mark.gpu

FIELDS = ("x", "z", "circ", "sigma", "u", "w", "uvs", "wvs")


def _set(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    st = {"id": np.arange(n, dtype=np.uint32), "x": rng.uniform(0, spread, n),
          "z": rng.uniform(0, spread, n), "circ": rng.normal(0, 1e-3, n),
          "sigma": rng.uniform(0.01, 0.05, n)}
    if n > 40:
        st["circ"][::23] = 0.0                      # merge with anything
        st["circ"][1::29] = -st["circ"][0::29][:len(st["circ"][1::29])]  # cancelling pairs
        st["circ"][5::31] *= 1e-6                   # long reach
    for k in ("u", "w", "uvs", "wvs"):
        st[k] = rng.normal(0, 1, n)
    return st


def _device_merge(ctx, thr, st, with_ids=True):
    d = {k: torch.tensor(st[k], dtype=torch.float64, device="cuda") for k in FIELDS}
    ids = torch.tensor(st["id"].astype(np.int32), device="cuda") if with_ids else None
    print("dvm_merge_blobs n", st["x"].size, "thr", thr, flush=True)
    k = dsim.merge_blobs(ctx, thr, *[d[f] for f in FIELDS], ids=ids)
    m = st["x"].size - k
    out = {f: d[f][:m].cpu().numpy() for f in FIELDS}
    if with_ids:
        out["id"] = ids[:m].cpu().numpy().astype(np.uint32)
    return k, out


@pytest.mark.parametrize("n,thr,seed", [(0, 1.0, 0), (1, 1.0, 0), (2, 1e9, 0), (300, 2e-5, 1),
                                        (3000, 2e-6, 2), (3000, 1e-5, 3), (12000, 1e-6, 4),
                                        (5000, 0.0, 5)])
def test_merge_matches_oracle_bit_exact(gpu_ctx, n, thr, seed):
    st = _set(n, seed)
    ko, ref = oracle.merge_blobs(thr, st)
    kg, got = _device_merge(gpu_ctx, thr, st)
    assert kg == ko
    for f in FIELDS + ("id",):
        assert np.array_equal(got[f], ref[f]), f
    if n >= 3000 and thr > 0:
        assert ko > 0


def test_merge_dense_cluster_chains(gpu_ctx):
    # a dense cluster: the first blobs absorb long chains, later ones find
    # their candidates already absorbed
    st = _set(4000, 7, spread=0.05)
    ko, ref = oracle.merge_blobs(5e-6, st)
    kg, got = _device_merge(gpu_ctx, 5e-6, st, with_ids=False)
    assert kg == ko and ko > 1000
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f


def test_merge_large_set_invariants(gpu_ctx):
    st = _set(200_000, 8, spread=4.0)
    thr = 2e-6
    kg, got = _device_merge(gpu_ctx, thr, st, with_ids=False)
    m = st["x"].size - kg
    assert kg > 0 and got["x"].size == m
    c, x = st["circ"], st["x"]
    assert abs(got["circ"].sum() - c.sum()) <= 1e-13 * np.abs(c).sum()
    assert abs((got["circ"] * got["x"]).sum() - (c * x).sum()) <= 1e-12 * np.abs(c * x).sum()
    # untouched survivors are the inputs, in order
    untouched = got["u"] != 0.0
    src = np.nonzero(np.isin(st["u"], got["u"][untouched]))[0]
    assert np.array_equal(st["x"][src], got["x"][untouched])
    # a second pass with the same threshold on the result equals the oracle's
    # pass on a subset small enough for it: the first 20,000 survivors
    sub = {f: got[f][:20000].copy() for f in FIELDS}
    sub["id"] = np.arange(sub["x"].size, dtype=np.uint32)
    ko, ref = oracle.merge_blobs(thr, sub)
    kg2, got2 = _device_merge(gpu_ctx, thr, sub)
    assert kg2 == ko
    for f in FIELDS:
        assert np.array_equal(got2[f], ref[f]), f


def _pairs(npairs, seed):
    """Separated pairs of close blobs on a square lattice (spacing 1e-2, pair
    distance below 2e-4, circulations of one sign): at threshold 1e-6 each
    pair merges and nothing else does, so the result of a prefix of the set
    does not depend on the rest."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(npairs)))
    k = np.arange(npairs)
    cx, cz = (k % side) * 1e-2, (k // side) * 1e-2
    ang = rng.uniform(0, 2 * np.pi, npairs)
    d = rng.uniform(2e-5, 2e-4, npairs)
    n = 2 * npairs
    st = {"id": np.arange(n, dtype=np.uint32), "x": np.empty(n), "z": np.empty(n),
          "circ": rng.uniform(5e-4, 1.5e-3, n), "sigma": rng.uniform(0.01, 0.05, n)}
    st["x"][0::2], st["z"][0::2] = cx, cz
    st["x"][1::2], st["z"][1::2] = cx + d * np.cos(ang), cz + d * np.sin(ang)
    for f in ("u", "w", "uvs", "wvs"):
        st[f] = rng.normal(0, 1, n)
    return st


def test_merge_flags_in_global_memory_above_lds_capacity(gpu_ctx):
    # Above 1,048,576 blobs the ordered commit keeps the absorbed flags in
    # global memory instead of its LDS bitmap; the first 100,000 pairs must
    # come out the same as when merged on their own (LDS path).
    big = _pairs(550_000, 7)
    kb, out_big = _device_merge(gpu_ctx, 1e-6, big, with_ids=False)
    assert kb == 550_000
    small = {f: v[:200_000].copy() for f, v in big.items()}
    ks, out_small = _device_merge(gpu_ctx, 1e-6, small, with_ids=False)
    assert ks == 100_000
    for f in FIELDS:
        assert np.array_equal(out_big[f][:100_000], out_small[f]), f
    # each survivor is its pair, merged: circulation adds
    assert np.array_equal(out_small["circ"], small["circ"][0::2] + small["circ"][1::2])
