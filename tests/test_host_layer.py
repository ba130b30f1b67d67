"""CPU tests of the host layer (effects, data, targets) against the
reference's own fixtures (tests/testthat/fixtures/{effects,targets}), with the
oracle as the statistic carrier."""
import numpy as np
import pytest

from conftest import load_golden


def _mat(rows):
    return np.array([[np.nan if x is None else float(x) for x in r] for r in rows])


def _effects_inputs(g):
    from cusna_amd.data import build_behavior_data

    n = g["n"]
    inp = g["inputs"]
    v_const = np.array([float(x) for x in inp["v_const"]])
    v_change = _mat(inp["v_change"])
    v_miss = np.array([np.nan if x is None else float(x) for x in inp["v_miss"]])
    dyad_const = _mat(inp["dyad_const"])
    dyad_change = np.stack([_mat(g["dyad_change"][m]) for m in range(3)], axis=2)
    beh = build_behavior_data(v_change, n, 3)
    covs = dict(v_const=v_const, v_change=v_change, v_miss=v_miss, dyad_const=dyad_const,
                dyad_change=dyad_change)
    return n, covs, beh


EFF_CFGS = [c["name"] for c in load_golden("effects.json.gz")["configs"]]


@pytest.mark.parametrize("name", EFF_CFGS)
def test_effect_descriptors_match_reference_fixture(name):
    """.prepare_effects_v2 & friends (test-effects.R:43-92), cross-network
    effects included (their source network passes through unchanged,
    R/effects.R:277-283)."""
    from cusna_amd.effects import Effect, effect_netrefs, prepare_effects

    g = load_golden("effects.json.gz")
    cfg = [c for c in g["configs"] if c["name"] == name][0]
    n, covs, beh = _effects_inputs(g)
    effects = []
    for r in cfg["recipe"]:
        if r.get("interaction"):
            effects.append(Effect("interaction", comps=tuple(r["comps"])))
        else:
            effects.append(Effect(r["name"], covs[r["cov"]] if "cov" in r else None,
                                  parm=r.get("parm"), dyn=r.get("dyn", False),
                                  etype=r.get("etype", "eval"), net_ref=r.get("net_ref")))
    p = cfg["period"] + 1
    z_start = None
    if cfg["use_beh"]:
        z_start = np.where(beh.miss[:, p - 1], 0.0, beh.imputed[:, p - 1].astype(float))
    d = prepare_effects(effects, n, p, beh if cfg["use_beh"] else None, z_start)
    np.testing.assert_array_equal(d.type_ids, cfg["type_ids"])
    np.testing.assert_allclose(d.p1, cfg["p1"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(d.p2, cfg["p2"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(d.vmat, _mat(cfg["vmat"]), rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(d.dyn_flags, cfg["dyn_flags"])
    np.testing.assert_allclose(d.v_dyn, cfg["v_dyn"], rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(d.dyc_idx, cfg["dyc_idx"])
    np.testing.assert_array_equal(d.vmiss.ravel(), np.asarray(cfg["vmiss"]).ravel())
    for j, dm in enumerate(cfg["dymats"] or []):
        np.testing.assert_allclose(d.dymats[j], _mat(dm), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(d.wtab, _mat(cfg["wtab"]), rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(d.int_comp.ravel(),
                                  np.array([x for r in cfg["int_idx"] for x in r]))
    np.testing.assert_array_equal(d.etypes, cfg["etypes"])
    if "netrefs" in cfg:
        np.testing.assert_array_equal(effect_netrefs(effects), cfg["netrefs"])


def test_gwesp_weights_match_reference_fixture():
    from cusna_amd.effects import gwesp_weights

    g = load_golden("effects.json.gz")
    for parm, w in g["gwesp_weights"].items():
        np.testing.assert_allclose(gwesp_weights(float(parm), g["n"]), w, rtol=1e-12)


TGT_CFGS = [c["name"] for c in load_golden("targets.json.gz")["configs"]]


@pytest.mark.parametrize("name", TGT_CFGS)
def test_moment_targets_match_reference_fixture(name, oracle_lib):
    """.mom_targets over the statistic carriers (test-targets.R:35-64), s50 and
    Knecht (missing ties and behaviour); the oracle is the carrier here, the
    GPU carriers are pinned to the oracle bit for bit elsewhere."""
    from cusna_amd.data import build_saom_data
    from cusna_amd.effects import BehEffect, Effect
    from cusna_amd.targets import mom_targets

    g = load_golden("targets.json.gz")
    cfg = [c for c in g["configs"] if c["name"] == name][0]
    covs = dict(alc_const=np.array(g["inputs"]["alc_const"], float),
                alc_change=_mat(g["inputs"]["alc_change"]))
    waves = [_mat(w) for w in cfg["waves_input"]]
    behavior = _mat(cfg["behavior_input"]) if cfg["behavior_input"] is not None else None
    data = build_saom_data(waves, behavior=behavior, composition=cfg["composition_input"])
    rate_effects = None
    if cfg["rate_effects"] is not None:
        from cusna_amd.effects import RateEffect

        rate_effects = [RateEffect(r["name"], covs[r["cov"]] if "cov" in r else None)
                        for r in cfg["rate_effects"]]
    effects = [Effect(r["name"], covs[r["cov"]] if "cov" in r else None, parm=r.get("parm"),
                      dyn=r.get("dyn", False)) for r in cfg["effects"]]
    beh_effects = [BehEffect(b) for b in cfg["beh_effects"]] if cfg["beh_effects"] else None

    def net(a, **kw):
        return oracle_lib.network_stats(a, kw["type_ids"], kw["p1"], kw["p2"], kw["vmat"],
                                        kw["vmiss"], kw["dyn_flags"], kw["vb"], kw["od0"],
                                        kw["id0"], kw["wtab"], kw["symflag"])

    mt = mom_targets(data, effects, beh_effects, cfg["conditional"] in (True, "True"),
                     net_stats=net, beh_stats=oracle_lib.behavior_stats,
                     rate_effects=rate_effects)
    if o_rate := cfg["out"].get("rate_eff_targets"):
        np.testing.assert_allclose(mt.rate_eff.sum(0), o_rate, rtol=1e-12)
    o = cfg["out"]
    np.testing.assert_allclose(mt.eval, _mat(o["eval_targets"]), rtol=1e-7)
    np.testing.assert_allclose(mt.basic_rate, o["rate_targets"], rtol=1e-7)
    if o.get("beh_targets") is not None:
        np.testing.assert_allclose(mt.beh, _mat(o["beh_targets"]), rtol=1e-7)
        np.testing.assert_allclose(mt.beh_rate, o["bdists"], rtol=1e-7)
    np.testing.assert_allclose(mt.target_vector, o["target_vector"], rtol=1e-7)
