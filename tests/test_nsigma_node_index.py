"""The transport's n*sigma lookup by node index (LdsSigma::nsigma).

On a geometric energy grid the kernel takes the node nearest E,
j = round((log2 E - log2 E0) * nodes per log2 unit), in float32 from the
hardware log2, and keeps interval j if E >= x_j, else interval j-1 (the last
node's cell repeats interval n-2).  That is the interval a search finds (the
largest i <= n-2 with x_i <= E) whenever every node lies well inside its own
cell, which the context's table build checks (|t_k - k| < 0.4) before it
enables the index.  These tests hold the shipped SM tables to that condition
and compare the rule, with the float32 index perturbed by 1e-4 cells either
way (its error bound is ~4e-5), with a search at the nodes, their float64
neighbours, the interval midpoints and random energies.
"""
import numpy as np
import pytest

from petite_amd.tables import MaterialTables, geometric

MATERIALS = ["graphite", "lead", "iron", "aluminum", "molybdenum"]


def _tables():
    out = []
    for ts in ("100GeV", "400GeV"):
        for mat in MATERIALS:
            try:
                mt = MaterialTables(mat, ts)
            except Exception:
                continue
            out.append((ts, mat, mt.nsigma_tab))
    return out


TABLES = _tables()


def _records(x, y):
    """Records -1 .. n-1 as the context builds them: {x_i, x_i+1, y_i, slope_i}."""
    n = len(x)
    k = np.clip(np.arange(-1, n), 0, n - 2)
    return np.stack([x[k], x[k + 1], y[k], (y[k + 1] - y[k]) / (x[k + 1] - x[k])], 1)


def _node_index_value(x, y, E, eps, shift=0):
    n = len(x)
    lnE0, inv = geometric(x)
    l2E0 = np.float32(lnE0 / np.log(2.0))
    per = np.float32(inv * np.log(2.0))
    t = ((np.log2(E.astype(np.float32)) - l2E0).astype(np.float32) * per).astype(np.float32)
    j = np.clip(np.floor(t.astype(np.float64) + 0.5 + eps).astype(np.int64), 0, n - 1)
    rec = np.concatenate([_records(x, y), _records(x, y)[-1:]])
    ra, rb = rec[j + shift], rec[j + 1 + shift]  # records j-1 and j
    va = ra[:, 3] * (E - ra[:, 0]) + ra[:, 2]
    vb = rb[:, 3] * (E - rb[:, 0]) + rb[:, 2]
    v = np.where(E >= rb[:, 0], vb, va)
    return np.where((E >= x[0]) & (E <= x[-1]), v, 0.0)


def _search_value(x, y, E):
    n = len(x)
    lo = np.clip(np.searchsorted(x, E, side="right") - 1, 0, n - 2)
    r = _records(x, y)[lo + 1]
    return np.where((E >= x[0]) & (E <= x[-1]), r[:, 3] * (E - r[:, 0]) + r[:, 2], 0.0)


def test_tables_found():
    assert {(ts, mat) for ts, mat, _ in TABLES} >= {("100GeV", "lead"), ("100GeV", "graphite"), ("400GeV", "lead")}


@pytest.mark.parametrize("ts,mat,tab", TABLES, ids=[f"{ts}-{mat}" for ts, mat, _ in TABLES])
def test_nodes_inside_their_cells(ts, mat, tab):
    for P, (x, y) in tab.items():
        x = np.asarray(x, np.float64)
        lnE0, inv = geometric(x)
        assert inv > 0.0, (ts, mat, P)
        l2E0 = np.float32(lnE0 / np.log(2.0))
        per = np.float32(inv * np.log(2.0))
        t = (np.log2(x) - np.float64(l2E0)) * np.float64(per)
        assert np.max(np.abs(t - np.arange(len(x)))) < 1e-4, (ts, mat, P)


@pytest.mark.parametrize("ts,mat,tab", TABLES, ids=[f"{ts}-{mat}" for ts, mat, _ in TABLES])
def test_node_index_equals_search(ts, mat, tab):
    rng = np.random.default_rng(7)
    for P, (x, y) in tab.items():
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        E = np.concatenate([x, np.nextafter(x, 0.0), np.nextafter(x, np.inf), np.sqrt(x[1:] * x[:-1]),
                            np.exp(rng.uniform(np.log(x[0]) - 1.0, np.log(x[-1]) + 1.0, 20000))])
        ref = _search_value(x, y, E)
        for eps in (0.0, 1e-4, -1e-4):
            got = _node_index_value(x, y, E, eps)
            bad = np.nonzero(got != ref)[0]
            assert len(bad) == 0, (ts, mat, P, eps, E[bad[:4]])


def test_wrong_records_are_caught():
    """Reading records j and j+1 instead of j-1 and j (a plausible slip in
    the record offset) gives other values on most energies: the comparison
    above would see it."""
    _, _, tab = TABLES[0]
    x, y = (np.asarray(a, np.float64) for a in tab["Brem"])
    E = np.sqrt(x[1:] * x[:-1])
    assert np.mean(_node_index_value(x, y, E, 0.0, shift=1) != _search_value(x, y, E)) > 0.5
    assert np.array_equal(_node_index_value(x, y, E, 0.0), _search_value(x, y, E))
