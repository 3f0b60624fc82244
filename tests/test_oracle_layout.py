"""Pins the oracle's ingest side (hashing, vertex classes, slot order, the two
DCSCs) against answers that do not come from the oracle: hashed ids worked
out by hand from the bucket definition, and the whole tile layout recomputed
from the records by closed_forms.expected_layout (numpy, written from the
reference's class and ordering rules, not from the oracle's code)."""
import numpy as np
import pytest

import closed_forms as CF
from test_oracle import random_records

# Hashed ids worked out by hand. SimpleBucketHasher(n, 1) at one rank
# (src/matrix/hashers.h:48-53): 128 parts, height = floor(n / 128),
# max_range = 128 * height. Bucket b holds the ids b, b + 128, b + 256, ...
# below max_range in that order, and the buckets sit end to end, so id v
# lands at (v mod 128) * height + floor(v / 128); ids at or above max_range
# keep their value (hashers.h:56, :63).
HAND = {
    # n = 300: height 2, max_range 256
    300: {0: 0, 1: 2, 2: 4, 127: 254, 128: 1, 129: 3, 200: 145, 255: 255, 256: 256, 299: 299, 300: 300},
    # n = 5000: height 39, max_range 4992; 4991 = 38 * 128 + 127
    5000: {0: 0, 1: 39, 128: 1, 129: 40, 4991: 4991, 4990: 4952, 4992: 4992, 5000: 5000},
    # n = 2^16 + 3: height 512, max_range 65536; 65535 = 511 * 128 + 127,
    # 65534 = 511 * 128 + 126
    65539: {1: 512, 128: 1, 65535: 65535, 65534: 65023, 65536: 65536, 65539: 65539},
    # n = 8 < 128: height 0, so every id keeps its value (g1)
    8: {0: 0, 1: 1, 8: 8},
}


@pytest.mark.parametrize("n", sorted(HAND))
def test_bucket_hash_hand_values(oracle, n):
    table = CF.bucket_hash_table(n)
    g = oracle.OracleGraph(np.zeros((0, 2), dtype=np.uint32), n, False, oracle.DIRECTED)
    for v, hv in HAND[n].items():
        assert table[v] == hv, (n, v)
        assert g.hash(v) == hv, (n, v)
        assert g.unhash(hv) == v, (n, v)


@pytest.mark.parametrize("n", [300, 5000, 65539, 127, 128, 129])
def test_bucket_hash_is_the_bucket_permutation(oracle, n):
    """ora_hash / ora_unhash over every id equal the bucket construction; the
    map is a permutation of [0, n], unhash inverts it, and it is the identity
    from max_range on."""
    table = CF.bucket_hash_table(n)
    assert np.array_equal(np.sort(table), np.arange(n + 1))
    g = oracle.OracleGraph(np.zeros((0, 2), dtype=np.uint32), n, False, oracle.DIRECTED)
    ids = np.arange(n + 1)
    if n > 6000:
        ids = np.concatenate([ids[:3000], ids[-3000:], np.random.default_rng(n).integers(0, n + 1, 3000)])
    got = np.array([g.hash(int(v)) for v in ids])
    assert np.array_equal(got, table[ids])
    back = np.array([g.unhash(int(h)) for h in table[ids]])
    assert np.array_equal(back, ids)
    max_range = (n // 128) * 128
    assert all(g.hash(v) == v for v in range(max_range, n + 1))


def _check_layout(oracle, rec, n, weighted, mode, hashing=1):
    exp = CF.expected_layout(rec, n, weighted, mode, hashing=bool(hashing))
    g = oracle.OracleGraph(rec, n, weighted, mode, hashing=hashing)
    assert (g.nreg, g.nsink, g.nsrc, g.ncols) == (exp["nreg"], exp["nsink"], exp["nsrc"], exp["ncols"])
    assert np.array_equal(g.hid_of(), exp["hid_of"])
    for which in (0, 1):
        got, want = g.csc(which), exp[which]
        for key in ("colptrs", "colidxs", "global_idx", "idx", "w"):
            assert np.array_equal(got[key], want[key]), (mode, which, key)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_layout_brute_force_bucket(oracle, mode):
    # n = 5000 (height 39): hashing is not the identity; duplicate records
    # carry other weights, so the first-record rule is visible in `w`.
    rec = random_records(60 + mode, 5000, 30000, weighted=True)
    _check_layout(oracle, rec, 5000, True, mode)


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_layout_brute_force_no_hashing(oracle, mode):
    rec = random_records(70 + mode, 300, 2000, weighted=True)
    _check_layout(oracle, rec, 300, True, mode, hashing=0)


def test_layout_brute_force_unweighted_zero_ids(oracle):
    # id 0 and ids at or above max_range (= 4992 for n = 5000) take part
    rec = random_records(80, 5000, 20000, zero_id=True)
    rec = np.concatenate([rec, np.array([[0, 4999], [4995, 0], [5000, 4992], [4992, 1]], np.uint32)])
    for mode in (0, 2):
        _check_layout(oracle, rec, 5000, False, mode)


def test_layout_brute_force_rmat(oracle):
    import la3_amd as L
    rec = L.rmat(13, 16, seed=11, weighted=True)
    _check_layout(oracle, rec, 1 << 13, True, 0)


def test_layout_g1_is_the_class_listing(oracle, g1):
    # at n = 8 the hash is the identity: the slot order lists the ids class
    # by class, as the fixture's class table does
    exp = CF.expected_layout(g1["records"], 8, False, 0)
    cls = g1["classes_directed"]
    assert exp["hid_of"].tolist() == cls["regular"] + cls["sink"] + cls["source"] + cls["isolated"]
    _check_layout(oracle, g1["records_w"], 8, True, 0)
