"""The four estimation set-ups of the reference's test-estimate.R:28-63
(s50, schedule of estimate.json.gz), as MomEstimator arguments."""
import numpy as np

from conftest import load_golden


def _mat(rows):
    return np.array([[np.nan if x is None else float(x) for x in r] for r in rows])


def estimate_fixture():
    return load_golden("estimate.json.gz")


def control(g, **over):
    from cusna_amd.mom import MomControl

    S = g["settings"]
    kw = dict(firstg=S["firstg"], n2=tuple(int(x) for x in S["n2"]), batch2=int(S["batch2"]),
              n1=int(S["n1"]), n3=int(S["n3"]), nD=int(S["nD"]),
              diagonalize=S["diagonalize"], seed=int(S["seed"]))
    kw.update(over)
    return MomControl(**kw)


def setup(name, g):
    """(data, effects, beh_effects, conditional) of one configuration."""
    from cusna_amd.data import build_saom_data
    from cusna_amd.effects import BehEffect, Effect

    waves = [_mat(w) for w in g["inputs"]["s50_waves"]]
    alc = _mat(g["inputs"]["alcohol"])
    base = [Effect(x) for x in ("density", "recip", "transTrip")]
    if name == "baseline":
        return build_saom_data(waves), base, None, False
    if name == "baseline_cond":
        return build_saom_data(waves), base, None, True
    if name == "covariate":
        x = alc[:, 0]
        return (build_saom_data(waves),
                base + [Effect("egoX", x), Effect("altX", x), Effect("simX", x)], None, False)
    if name == "coevol":
        eff = base + [Effect("egoX", dyn=True), Effect("altX", dyn=True),
                      Effect("simX", dyn=True)]
        bef = [BehEffect(b) for b in ("linear", "quad", "avSim")]
        return build_saom_data(waves, behavior=alc), eff, bef, False
    raise KeyError(name)


CONFIGS = ("baseline", "baseline_cond", "covariate", "coevol")
