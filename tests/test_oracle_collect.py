"""The oracle's COLLECT_LIST / COLLECT_SET pinned on the CPU: against every
literal fixture of cpp/tests/groupby/collect_list_tests.cpp and
collect_set_tests.cpp (lists compared in row order, sets after sorting, as
the reference's helper does), and against a pure-Python statement of the
same rules on small random inputs: a list is the group's rows in row order
with nulls kept or dropped; a set holds each distinct value once (-0.0 ==
+0.0; NaNs one element or each their own), then its nulls (one, each or
none), in order of first appearance."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_cases as G

COLLECT_CASES = G.load_collect_cases()


@pytest.mark.parametrize("case", COLLECT_CASES, ids=[c["id"] for c in COLLECT_CASES])
def test_reference_collect_fixture_through_oracle(case):
    kt, kv, kvalid = case["keys"]
    vt, vv, vvalid = case["vals"]
    kind, nulls, neq, naneq = case["agg"]
    keys, lists = O.collect([O.HostColumn(kt, kv, kvalid)], O.HostColumn(vt, vv, vvalid), kind, nulls, neq, naneq)
    G.check_collect_case(case, keys[0], lists, sort_sets=True)


def _python_collect(keys, vals, valid, kind, nulls, neq, naneq):
    """Groups in order of first appearance -> list of (values, valid) in row order."""
    groups = {}
    for i, k in enumerate(keys.tolist()):
        groups.setdefault(k, []).append(i)
    out = {}
    for k, rows in groups.items():
        if kind == O.COLLECT_LIST:
            keep = [r for r in rows if nulls == 1 or valid[r]]
        else:
            keep, seen, null_seen = [], [], False
            for r in rows:
                if not valid[r]:
                    if nulls == 1 and (neq == 1 or not null_seen):
                        keep.append(r)
                    null_seen = True
                    continue
                x = vals[r].item()
                is_nan = isinstance(x, float) and math.isnan(x)
                if is_nan:
                    dup = naneq == 0 and any(isinstance(y, float) and math.isnan(y) for y in seen)
                else:
                    dup = any(y == x for y in seen if not (isinstance(y, float) and math.isnan(y)))
                if not dup:
                    keep.append(r)
                    seen.append(x)
        out[k] = (vals[keep], valid[keep])
    return out


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("vt", [O.INT8, O.INT64, O.UINT32, O.FLOAT32, O.FLOAT64])
def test_oracle_collect_vs_python(seed, vt):
    rng = np.random.default_rng(900 + 10 * seed + vt)
    n = int(rng.integers(0, 400))
    keys = rng.integers(0, 25, n).astype(np.int32)
    if vt in (O.FLOAT32, O.FLOAT64):
        pool = np.array([-1.5, -0.0, 0.0, 2.0, 7.25, np.inf, -np.inf, np.nan], dtype=O.NP[vt])
        vals = pool[rng.integers(0, len(pool), n)]
    else:
        vals = rng.integers(0, 6, n).astype(O.NP[vt])
    valid = rng.random(n) > 0.2
    for kind in (O.COLLECT_LIST, O.COLLECT_SET):
        for nulls in (0, 1):
            for neq in (0, 1):
                for naneq in (0, 1):
                    ok, lists = O.collect([O.HostColumn(O.INT32, keys)], O.HostColumn(vt, vals, valid), kind,
                                          nulls, neq, naneq)
                    want = _python_collect(keys, vals, valid, kind, nulls, neq, naneq)
                    got_keys = ok[0][0].tolist()
                    assert got_keys == list(want.keys())
                    for k, lst in zip(got_keys, lists):
                        assert G.same_list(lst, want[k]), (kind, nulls, neq, naneq, k, lst, want[k])


def test_oracle_collect_identities():
    """A set holds each value of its list once and nothing else, and its nulls
    follow the null rules; an EXCLUDE list is the INCLUDE list without its
    nulls; null keys drop their rows unless included."""
    rng = np.random.default_rng(950)
    n = 3000
    keys = rng.integers(0, 40, n).astype(np.int64)
    kvalid = rng.random(n) > 0.1
    vals = rng.integers(-5, 5, n).astype(np.int16)
    valid = rng.random(n) > 0.3
    kc, vc = O.HostColumn(O.INT64, keys, kvalid), O.HostColumn(O.INT16, vals, valid)
    _, inc = O.collect([kc], vc, O.COLLECT_LIST, 1)
    _, exc = O.collect([kc], vc, O.COLLECT_LIST, 0)
    _, st = O.collect([kc], vc, O.COLLECT_SET, 1, 0)
    _, st_u = O.collect([kc], vc, O.COLLECT_SET, 1, 1)
    for (lv, lval), (ev, eval_), (sv, sval), (uv, uval) in zip(inc, exc, st, st_u):
        assert np.array_equal(ev, lv[lval]) and eval_.all()
        assert sorted(sv[sval].tolist()) == sorted(set(lv[lval].tolist()))
        assert int((~sval).sum()) == (1 if (~lval).any() else 0)
        assert int((~uval).sum()) == int((~lval).sum())
    k_inc, lists_inc = O.collect([kc], vc, O.COLLECT_LIST, 1, include_null_keys=True)
    assert len(lists_inc) == len(inc) + 1
    assert sum(len(v) for v, _ in lists_inc) == n


def test_oracle_collect_rejects_other_kinds():
    with pytest.raises(ValueError):
        O.collect([O.HostColumn(O.INT32, [1])], O.HostColumn(O.INT32, [1]), O.SUM)


def test_sort_list_order():
    v, m = O.sort_list(np.array([np.nan, 2.0, -1.0, 0.0, 5.0]), np.array([1, 1, 0, 1, 1], bool))
    assert np.isnan(v[3]) and v[:3].tolist() == [0.0, 2.0, 5.0] and m.tolist() == [1, 1, 1, 1, 0]
