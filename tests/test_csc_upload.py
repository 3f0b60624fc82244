"""Graphs from the reference's two CSCs (la3_graph_from_csc), on a host-only
context: the CSCs exported from a loaded graph (the reference's CSC<W>
layout, csc.h:61-73) upload into a graph whose every stored array, class
count and CSC equals the original's, for all five load modes, bucket and no
hashing, weighted and not; arrays that are not a reference layout are
rejected with the first difference named."""
import numpy as np
import pytest

import la3_amd as L
from test_oracle import random_records

ARRAYS = ("col_ptr", "col_split", "col_rows", "col_w", "row_ptr", "row_split", "row_cols", "row_w",
          "id_of", "slot_of_id", "long_rows")
MODES = (L.DIRECTED, L.REVERSED, L.UNDIRECTED, L.ACYCLIC, L.REVERSED_ACYCLIC)


@pytest.fixture(scope="module")
def host_ctx():
    ctx = L.Context(-1)
    yield ctx
    ctx.close()


def _roundtrip(ctx, rec, n, weighted, mode, hashing):
    g = L.Graph.from_records(ctx, rec, n, weighted, mode, hashing)
    csc = (g.csc(0), g.csc(1))
    u = L.Graph.from_csc(ctx, n, csc, weighted, mode, hashing)
    for key in ("nslots", "nregular", "nsink", "nsource", "nisolated", "ncols", "directed", "weighted",
                "mode", "nnz", "nnz_regular", "nnz_sink"):
        assert u.info[key] == g.info[key], key
    for name in ARRAYS:
        if name.endswith("_w") and not weighted:
            continue
        assert np.array_equal(u.array(name), g.array(name)), name
    for which in (0, 1):
        a, b = g.csc(which), u.csc(which)
        for k in a:
            assert np.array_equal(a[k], b[k]), (which, k)
    return g, csc


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("weighted", [False, True])
def test_roundtrip_random(host_ctx, mode, weighted):
    rec = random_records(131, 5000, 30000, weighted=weighted)
    for hashing in (L.HASH_BUCKET, L.HASH_NONE):
        _roundtrip(host_ctx, rec, 5000, weighted, mode, hashing)


def test_roundtrip_g1_and_zero_id(host_ctx, g1):
    _roundtrip(host_ctx, g1["records_w"], 8, True, L.DIRECTED, L.HASH_BUCKET)
    rec = random_records(132, 300, 2000, zero_id=True)
    _roundtrip(host_ctx, rec, 300, False, L.UNDIRECTED, L.HASH_BUCKET)


def test_rejects_arrays_that_are_not_a_reference_layout(host_ctx):
    rec = random_records(133, 2000, 12000, weighted=True)
    g, csc = _roundtrip(host_ctx, rec, 2000, True, L.DIRECTED, L.HASH_BUCKET)
    reg, sink = csc

    def upload(r, s):
        return L.Graph.from_csc(host_ctx, 2000, (r, s), True, L.DIRECTED, L.HASH_BUCKET)

    def changed(c, key, j, v):
        d = {k: np.array(a, copy=True) for k, a in c.items()}
        d[key][j] = v
        return d

    # entries of a column out of the locator's order (two rows swapped)
    c0 = int(np.flatnonzero(np.diff(reg["colptrs"]) >= 2)[0])
    j = int(reg["colptrs"][c0])
    swapped = {k: np.array(a, copy=True) for k, a in reg.items()}
    for key in ("global_idx", "idx", "w"):
        swapped[key][[j, j + 1]] = swapped[key][[j + 1, j]]
    # a weight is data, not layout: taken as given
    heavier = changed(reg, "w", 0, int(reg["w"][0]) + 1)
    assert np.array_equal(upload(heavier, sink).csc(0)["w"], heavier["w"])
    cases = [
        (swapped, "global_idx"),
        (changed(reg, "global_idx", 3, int(reg["global_idx"][3]) + 1), "global_idx"),
    ]
    for r, what in cases:
        with pytest.raises(L.LA3Error) as e:
            upload(r, sink)
        assert what in str(e.value), str(e.value)
    bad = changed(reg, "colptrs", 1, int(reg["colptrs"][2]) + 1)  # decreasing pointers
    with pytest.raises(L.LA3Error):
        upload(bad, sink)
    with pytest.raises(L.LA3Error):  # a row id beyond n
        upload(changed(reg, "idx", 0, 5000), sink)
