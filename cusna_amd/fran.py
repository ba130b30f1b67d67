"""FRAN backend for RSiena's siena07() (cran/cusna R/fran.R:27-63).

``cusna_fran`` returns a closure that siena07() calls once per Robbins-Monro
step with its current theta: it simulates one chain of every period on the
device (``MomEstimator.simulate``, the mirror of ``.mom_simulate``) and
returns the period pieces in RSiena's simstats0c layout -- ``fra`` and
``sc`` as periods x parameters matrices (siena07() forms colSums(fra) and
reads sc row by row), ``ntim0`` one duration per period. The parameter layout
is the estimator's: the effect parameters when conditional, the basic rates
of the periods first when not. Seeds come from numpy's global stream (or the
``rng`` a caller passes), so a run repeats under a fixed seed, as the R
closure leaves R's stream to siena07().
"""
from __future__ import annotations

import numpy as np

from .data import build_saom_data
from .effects import Effect
from .mom import MomEstimator

# effects whose covariate is the closure's `covariate` (R/fran.R:30-31)
COV_EFFECTS = ("egoX", "altX", "simX", "sameX", "higher", "egoXaltX", "egoPlusAltX",
               "altRThresholdX")


def _field(z, name, default=None):
    if isinstance(z, dict):
        return z.get(name, default)
    return getattr(z, name, default)


class Fran:
    """The FRAN closure (R/fran.R:38-61); ``counter`` counts its calls."""

    def __init__(self, waves, effect_names, covariate=None, conditional: bool = True,
                 **target_carriers):
        data = build_saom_data([np.asarray(w, dtype=float) for w in waves])
        effects = [Effect(nm, cov=None if covariate is None or nm not in COV_EFFECTS
                          else np.asarray(covariate, dtype=float)) for nm in effect_names]
        self.conditional = bool(conditional)
        # target_carriers: the statistics functions of mom_targets (the device's
        # by default)
        self.est = MomEstimator(data, effects, conditional=self.conditional, **target_carriers)
        self.counter = 0

    def __call__(self, z, x=None, rng: np.random.Generator | None = None) -> dict:
        theta = np.asarray(_field(z, "theta"), dtype=float).ravel()
        cc = _field(z, "cconditional")
        if cc is not None and bool(cc) != self.conditional:
            raise ValueError(f"cusna_fran(conditional = {self.conditional}) does not match the "
                             f"algorithm (cond = {bool(cc)}).")
        est = self.est
        if theta.size != est.p:
            raise ValueError(f"siena07() passed {theta.size} parameters, but the cusna model "
                             f"has {est.p} ({', '.join(est.param_names)}); check `effect_names` "
                             "against the RSiena effects object.")
        self.counter += 1
        s = est.simulate(theta, 1, rng)
        fra = np.stack([pc["sp"][0] for pc in s["pieces"]])
        sc = np.stack([pc["scp"][0] for pc in s["pieces"]])
        return {"fra": fra, "sc": sc, "ntim0": s["ntim"][0].copy(), "OK": True,
                "feasible": True}


def cusna_fran(waves, effect_names, covariate=None, conditional: bool = True,
               **target_carriers) -> Fran:
    """A FRAN closure for ``alg$FRAN`` in an RSiena siena07() run (R/fran.R:27)."""
    return Fran(waves, effect_names, covariate, conditional, **target_carriers)
