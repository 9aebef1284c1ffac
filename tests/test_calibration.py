"""The autotuner's and placement calibration's host logic (no device):
profiles are used only on an identity match, the cost fits recover known
linear models, the provider reports calibrated costs and never a CPU
estimate it did not measure; and (on the device) the library applies a
matching profile's settings to new contexts and ignores any other."""
import json
import os

import numpy as np
import pytest

from dataframe_mojo_amd import _lib, autotune, calibration, provider
from dataframe_mojo_amd.provider import PipelineFeatures

IDENT = {"device": "AMD Instinct MI355X", "arch": "gfx950", "compute_units": 256, "build_id": "0123456789abcdef"}


def test_build_id_is_the_source_hash():
    from dataframe_mojo_amd import build

    lib = _lib.load()
    assert lib.dfr_build_id().decode() == build.build_id()


def test_identity_must_match_every_field():
    prof = {"version": 2, "identity": dict(IDENT)}
    assert calibration.identity_matches(prof, IDENT)
    for k, v in [("device", "AMD GPU"), ("arch", "gfx942"), ("compute_units", 304), ("build_id", "ffff")]:
        assert not calibration.identity_matches(prof, dict(IDENT, **{k: v})), k
    assert not calibration.identity_matches({"version": 2}, IDENT)
    assert not calibration.identity_matches(prof, None)


def test_load_profile_rejects_other_versions_and_identities(tmp_path, monkeypatch):
    path = tmp_path / "p.json"
    monkeypatch.setenv(calibration.ENV, str(path))
    monkeypatch.setattr(calibration, "device_identity", lambda device: dict(IDENT))
    good = {"version": 2, "identity": dict(IDENT), "groupby": {"fixed_ns": 1e4, "ns_per_row": 0.01, "ns_per_group": 1.0}}
    path.write_text(json.dumps(good))
    assert calibration.load_profile(0) == good
    path.write_text(json.dumps(dict(good, version=1)))
    assert calibration.load_profile(0) is None
    path.write_text(json.dumps(dict(good, identity=dict(IDENT, build_id="other"))))
    assert calibration.load_profile(0) is None
    path.write_text("{not json")
    assert calibration.load_profile(0) is None
    monkeypatch.setenv(calibration.ENV, "off")
    assert calibration.load_profile(0) is None


def test_fit_linear_recovers_a_known_model():
    rng = np.random.default_rng(0)
    xs = [[float(r), float(g)] for r, g in rng.integers(1, 10**8, size=(12, 2))]
    ys = [5e4 + 0.02 * r + 3.0 * g for r, g in xs]
    c = autotune.fit_linear(xs, ys)
    assert abs(c[0] - 5e4) < 1.0 and abs(c[1] - 0.02) < 1e-9 and abs(c[2] - 3.0) < 1e-7
    # a coefficient that would come out negative is dropped, not reported
    ys2 = [1e3 + 0.5 * r for r, _ in xs]
    c2 = autotune.fit_linear(xs, [y - 1e-9 * g for y, (_, g) in zip(ys2, xs)])
    assert min(c2) >= 0.0


def test_provider_costs_follow_the_profile(monkeypatch):
    from dataframe_mojo_amd import Column, DataFrame, Series, col

    prof = {"version": 2, "identity": dict(IDENT),
            "transfer": {"to_device": [2e4, 0.04], "to_host": [3e4, 0.05], "startup_ns": 2.5e8},
            "groupby": {"fixed_ns": 1e5, "ns_per_row": 0.01, "ns_per_group": 2.0}}
    monkeypatch.setattr(calibration, "load_profile", lambda device: prof)
    df = DataFrame([Series("k", Column.int64([1, 2, 3])), Series("v", Column.int64([4, 5, 6]))])
    lf = df.filter(col("v") > 4).group_by("k").agg(col("v").sum())
    feats = PipelineFeatures(input_rows=10**6, groups=1000)
    est = provider.pipeline_estimate(lf, 0, feats)
    assert est == pytest.approx(1e5 + 0.01 * 1e6 + 2.0 * 1000)
    # no CPU executor is measured here: the planner's own CPU model applies
    assert provider.pipeline_costs(lf, 0, feats) == (-1.0, est)
    tm = provider.transfer_model(0)
    assert tm.measured() and tm.to_device.cost(1000) == pytest.approx(2e4 + 40) and tm.startup_ns == 2.5e8
    monkeypatch.setattr(calibration, "load_profile", lambda device: None)
    assert provider.pipeline_estimate(lf, 0, feats) == -1.0
    assert not provider.transfer_model(0).measured()


def test_knob_candidates_are_accepted_values():
    # every candidate the tuner tries is inside dfr_context_tune's ranges
    # (the library refuses anything else; the tuner would skip it)
    for cls, knobs in autotune.KNOBS.items():
        for key, cands in knobs:
            assert len(set(cands)) == len(cands) >= 2, (cls, key)
            if key.endswith("lds_bytes"):
                assert all(4096 <= v <= 163840 for v in cands), key
            if key in ("max_parts", "two_level_parts"):
                assert all(v & (v - 1) == 0 for v in cands), key


@pytest.mark.gpu
def test_library_applies_only_a_matching_profile(tmp_path, monkeypatch):
    from dataframe_mojo_amd.runtime import Context

    ident = autotune.identity(0)
    path = tmp_path / "profile.json"
    monkeypatch.setenv("DATAFRAME_ROCM_CALIBRATION", str(path))
    path.write_text(json.dumps({"version": 2, "identity": ident, "tuning": {"scan_unroll": 2, "batch_tile": 65536}}))
    ctx = Context(0)
    try:
        assert ctx.calibration().startswith("applied:"), ctx.calibration()
        assert ctx.tune_get("scan_unroll") == 2 and ctx.tune_get("batch_tile") == 65536
    finally:
        ctx.close()
    path.write_text(json.dumps({"version": 2, "identity": dict(ident, build_id="not-this-build"),
                                "tuning": {"scan_unroll": 2}}))
    ctx = Context(0)
    try:
        assert ctx.calibration().startswith("none:") and "build" in ctx.calibration()
        assert ctx.tune_get("scan_unroll") == 1
    finally:
        ctx.close()
    monkeypatch.setenv("DATAFRAME_ROCM_CALIBRATION", "off")
    ctx = Context(0)
    try:
        assert ctx.calibration() == "none: DATAFRAME_ROCM_CALIBRATION=off"
    finally:
        ctx.close()


class _FakeResult:
    def __init__(self, keys, aggs):
        self.keys, self.aggs = keys, aggs
        self.groups = len(keys[0][0])

    def key(self, k):
        return None, *self.keys[k]

    def agg(self, a):
        return None, *self.aggs[a]


def test_result_digest_is_order_independent_and_row_coupled():
    rng = np.random.default_rng(3)
    n = 1000
    k = rng.integers(0, 10**6, n).astype(np.int64)
    v = rng.normal(size=n)
    valid = rng.random(n) > 0.1
    ones = np.ones(n, dtype=bool)
    base = autotune._digest_result(_FakeResult([(k, ones)], [(v, valid)]), 1, 1)
    p = rng.permutation(n)
    assert autotune._digest_result(_FakeResult([(k[p], ones)], [(v[p], valid[p])]), 1, 1) == base
    # a value moved to another group's row, one bit flipped, one validity changed
    v2 = v.copy()
    v2[[0, 1]] = v2[[1, 0]]
    assert autotune._digest_result(_FakeResult([(k, ones)], [(v2, valid)]), 1, 1) != base
    v3 = v.copy()
    v3[5] = np.nextafter(v3[5], 10.0)
    assert autotune._digest_result(_FakeResult([(k, ones)], [(v3, valid)]), 1, 1) != base
    m2 = valid.copy()
    m2[7] = not m2[7]
    assert autotune._digest_result(_FakeResult([(k, ones)], [(v, m2)]), 1, 1) != base


class _FakeCtx:
    """A context whose settings are only recorded."""

    def __init__(self, defaults):
        self.values = dict(defaults)

    def tune(self, key, value):
        self.values[key] = value


def _scripted_measure(times):
    """measure() replaced by times from a script: each call pops the next
    time for the context's current settings (or the last one left)."""
    def measure(ctx, work, reps):
        key = tuple(sorted(ctx.values.items()))
        seq = times[key]
        t = seq.pop(0) if len(seq) > 1 else seq[0]
        return t, "digest"
    return measure


@pytest.mark.parametrize("confirmed", [True, False])
def test_tuned_settings_are_kept_only_when_confirmed(confirmed, monkeypatch):
    # a knob that wins during the descent is kept only if it still beats
    # the built-in settings, measured back to back at the end
    defaults = {"a": 1, "b": 10}
    monkeypatch.setattr(autotune, "_DEFAULTS", dict(defaults))
    base, win = (("a", 1), ("b", 10)), (("a", 2), ("b", 10))
    # built-in: the first measure, candidate a=1, then the confirmation
    times = {
        base: [10.0, 10.0, 10.0 if confirmed else 9.8],
        win: [9.6, 9.6 if confirmed else 9.9],
        (("a", 3), ("b", 10)): [10.1],
        (("a", 2), ("b", 20)): [10.5],
    }
    monkeypatch.setattr(autotune, "measure", _scripted_measure(times))
    ctx = _FakeCtx(defaults)
    chosen, notes = autotune.tune_class(ctx, "x", object(), 1, 0.02, knobs=[("a", [1, 2, 3]), ("b", [10, 20])])
    if confirmed:
        assert chosen == {"a": 2} and ctx.values == {"a": 2, "b": 10}
    else:
        assert chosen == {} and ctx.values == defaults
    assert notes["confirm"] is not None
