"""CPU: the data layer (cusna_amd.data) against the reference's data-engine
fixtures (tests/testthat/fixtures/data, test-data.R:32-122): parsing,
imputation, missing and structural masks, symmetry, composition change (COW:
symmetric, n = 116, ccOption 1), behaviour scalars and every period mask."""
import gzip
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

DATASETS = ["s50", "s50_beh", "s50miss", "s50str", "knecht", "cow", "synth"]


def _mat(rows):
    return np.array([[np.nan if v is None else v for v in r] for r in rows], float)


def _eq(a, gold):
    np.testing.assert_array_equal(np.asarray(a).astype(np.int64).ravel(),
                                  np.asarray(gold, dtype=np.int64).ravel())


@pytest.mark.parametrize("name", DATASETS)
def test_data_engine_matches_reference_fixture(name):
    from cusna_amd.data import build_saom_data

    with gzip.open(os.path.join(GOLDEN, "data", f"{name}.json.gz"), "rt") as f:
        g = json.load(f)
    meta = g["meta"]
    waves = [_mat(w) for w in g["waves_input"]]
    behavior = _mat(g["behavior_input"]) if g["behavior_input"] is not None else None
    obj = build_saom_data(waves, behavior=behavior, composition=g["composition_input"],
                          cc_option=meta["cc_option"])
    assert obj.symmetric == bool(meta["symmetric"])
    assert obj.n_periods == meta["n_periods"]
    assert obj.has_composition == bool(meta["has_composition"])
    for m in range(meta["n_waves"]):
        _eq(obj.imputed[m], g["imputed"][m])
        _eq(obj.miss[m], g["miss"][m])
        _eq(obj.struct[m], g["struct"][m])
        _eq(obj.rawvals[m], g["rawvals"][m])
    if g["comp_active"] is not None:
        _eq(obj.comp_active, g["comp_active"])
    for p in range(1, meta["n_periods"] + 1):
        gp = g["periods"][p - 1]
        _eq(obj.kmask_dist(p), gp["kmask_dist"])
        assert obj.observed_distance(p) == gp["observed_distance"]
        assert obj.rate_target(p) == gp["rate_target"]
        assert obj.rate_const(p) == gp["rate_const"]
        _eq(obj.rate_target_rows(p), gp["rate_target_rows"])
        _eq(obj.rate_const_rows(p), gp["rate_const_rows"])
        _eq(obj.network_less_missing(p), gp["network_less_missing"])
        _eq(obj.network_less_missing_start(p), gp["network_less_missing_start"])
        keep, force = obj.stat_keep_force(p)
        _eq(keep, gp["keep"])
        _eq(force, gp["force"])
        keep, force = obj.stat_keep_force_struct(p)
        _eq(keep, gp["keepS"])
        _eq(force, gp["forceS"])
        _eq(obj.target_network(p), gp["target_network"])
        _eq(obj.permitted(p), gp["permitted"])
    if g.get("behavior") is not None:
        bg, beh = g["behavior"], obj.beh
        assert beh.zmin == bg["zmin"] and beh.zmax == bg["zmax"]
        np.testing.assert_allclose(beh.overall_mean, bg["overall_mean"], rtol=1e-12)
        np.testing.assert_allclose(beh.range, bg["range"], rtol=1e-12)
        np.testing.assert_allclose(beh.sim_mean, bg["sim_mean"], rtol=1e-12)
        _eq(beh.imputed, bg["imputed"])
        _eq(beh.miss, bg["miss"])
        for p in range(1, meta["n_periods"] + 1):
            bp = bg["periods"][p - 1]
            _eq(obj.beh_kmask_dist(p), bp["kmask_dist"])
            assert obj.beh_observed_distance(p) == bp["observed_distance"]
            np.testing.assert_allclose(obj.beh_values_less_missing_starts(p),
                                       bp["values_less_missing_starts"], rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(obj.beh_stat_values(beh.imputed[:, p], p),
                                       bp["stat_values_target"], rtol=1e-12, atol=1e-12)


def test_composition_period_actors():
    """Active lists and inactive masks of a period for the period call."""
    from cusna_amd.data import build_saom_data

    with gzip.open(os.path.join(GOLDEN, "data", "cow.json.gz"), "rt") as f:
        g = json.load(f)
    obj = build_saom_data([_mat(w) for w in g["waves_input"]],
                          composition=g["composition_input"])
    for p in range(1, obj.n_periods + 1):
        act, nact, inact = obj.period_actors(p)
        active = np.asarray(g["comp_active"][p - 1], bool)
        assert nact == active.sum()
        assert set(act[:nact]) == set(np.nonzero(active)[0])
        _eq(inact, ~active)
