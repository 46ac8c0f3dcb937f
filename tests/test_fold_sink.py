"""Folded statistics on the host side (graph/sinks.go:17-71): the
combination of two partials, and the tee sink's checks and its mapping
onto the library's sinks. No GPU: partials here come from the oracle.

_combine_folds is pinned by the oracle's fold of the union of two value
lists (oracle/reduce_ref.py fold, itself pinned against math.fsum and
exact fractions): any split of the values, in either order, must combine
into the fold of all of them."""
import numpy as np
import pytest

from oracle import reduce_ref as R
from strata_amd import _lib, engine
from strata_amd import reduce as rd
from strata_amd.engine import EnginePanic, MemorySink, TeeSink, _combine_folds, _empty_fold, _sink_target
from strata_amd.raster import Float32Raster

F32 = np.float32


def to_fold(d: dict) -> _lib.Fold:
    f = _lib.Fold()
    f.count, f.nan, f.pos_inf, f.neg_inf = d["count"], d["nan"], d["pos_inf"], d["neg_inf"]
    f.not_neg_zero = 1 if d["not_neg_zero"] else 0
    f.min, f.max = float(d["min"]), float(d["max"])
    engine._set_limbs(f.sum, d["x1"])
    engine._set_limbs(f.sumsq, d["x2"])
    return f


def same(f: _lib.Fold, d: dict):
    assert (f.count, f.nan, f.pos_inf, f.neg_inf) == (d["count"], d["nan"], d["pos_inf"], d["neg_inf"])
    assert bool(f.not_neg_zero) == d["not_neg_zero"]
    assert _lib.limbs_to_int(f.sum) == d["x1"] and _lib.limbs_to_int(f.sumsq) == d["x2"]
    if d["count"]:
        for got, want in ((F32(f.min), d["min"]), (F32(f.max), d["max"])):
            assert (np.isnan(got) and np.isnan(want)) or R.f32_bits(got) == R.f32_bits(want)


CASES = {
    "signs": [F32(-3.5), F32(2.25), F32(-0.0), F32(0.0), F32(1e-30), F32(-7e20)],
    "negative zeros": [F32(-0.0), F32(-0.0), F32(-0.0)],
    "nan and infinities": [F32(1), F32(np.nan), F32(np.inf), F32(-np.inf), F32(-2)],
    "wide": [F32(3.4e38), F32(-3.4e38), F32(1e-45), F32(-1e-45), F32(6.5e4), F32(-1.5)],
    "one": [F32(42)],
    "none": [],
}


@pytest.mark.parametrize("name", list(CASES))
def test_combine_is_the_fold_of_the_union(name):
    vals = CASES[name]
    whole = R.fold(vals)
    for cut in range(len(vals) + 1):
        a, b = vals[:cut], vals[cut:]
        for first, second in ((a, b), (b, a)):
            f = _empty_fold()
            _combine_folds(f, to_fold(R.fold(first)))
            _combine_folds(f, to_fold(R.fold(second)))
            same(f, whole)


def test_combine_many_random_parts():
    rng = np.random.default_rng(7)
    vals = list((rng.standard_normal(200) * 10.0 ** rng.integers(-20, 20, 200)).astype(F32))
    vals[17] = F32(-0.0)
    f = _empty_fold()
    edges = sorted(rng.integers(0, len(vals), 9))
    for lo, hi in zip([0] + edges, edges + [len(vals)]):
        _combine_folds(f, to_fold(R.fold(vals[lo:hi])))
    same(f, R.fold(vals))
    # and the finished summary is the oracle's
    s = rd._summary(f)
    want = R.summary(vals)
    assert s.Count == want["Count"] and s.Sum == want["Sum"] and s.Mean == want["Mean"]


def test_negative_sums_wrap_as_twos_complement():
    f = _empty_fold()
    _combine_folds(f, to_fold(R.fold([F32(-1.0)])))
    _combine_folds(f, to_fold(R.fold([F32(-2.0)])))
    assert _lib.limbs_to_int(f.sum) == -3 << 150
    assert f.sum[5] >> 63 == 1  # the top limb carries the sign


def host(w, h, masked=True):
    return Float32Raster(np.zeros(w * h, np.float32), w, h, w,
                         np.zeros((w * h + 63) // 64 + 1, np.uint64) if masked else None, 0)


def test_tee_checks_and_targets():
    with pytest.raises(EnginePanic, match="positive"):
        TeeSink(0, 3, True)
    s = MemorySink(host(4, 3))
    with pytest.raises(EnginePanic, match="is 5×3, its sink 0 is"):
        TeeSink(5, 3, True, [s])
    with pytest.raises(EnginePanic, match="sink 1 is nil"):
        TeeSink(4, 3, True, [s, None])
    with pytest.raises(EnginePanic, match="tee sink is Masked but its sink 0 is not Masked"):
        TeeSink(4, 3, True, [MemorySink(host(4, 3, masked=False))])
    with pytest.raises(EnginePanic, match="tee sink is not Masked but its sink 0 is Masked"):
        TeeSink(4, 3, False, [s])
    only = TeeSink(4, 3, True)
    assert _sink_target(only) is None  # statistics alone write nothing
    one = TeeSink(4, 3, True, [s])
    assert _sink_target(one) is s
    two = TeeSink(4, 3, True, [s, MemorySink(host(4, 3))])
    assert isinstance(_sink_target(two), engine._TeeWrites)
    assert _sink_target(s) is s
    assert only.Size() == (4, 3) and only.Masked()
    assert rd.Folded(only).Count == 0 and np.isnan(rd.Folded(only).Mean)
