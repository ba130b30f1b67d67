"""Writes the golden fixtures under tests/golden/ (run in a container that has
the reference sources; the fixtures, not the reference, travel).

  sim.json.gz       inputs and outputs of the reference's own period-simulator
                    fixture (cran/cusna tests/testthat/fixtures/sim/sim.json.gz):
                    s50 waves, alcohol covariate, three configurations x 8 chains
  ref_cases.json.gz outputs of the reference engine compiled from its source
                    (oracle/_ref/libcusna_ref.so) on synthetic inputs that the
                    reference's fixtures do not cover: cycle3/inPop/outAct and
                    every other effect id, behaviour avAlt/totSim, symmetric and
                    conditional runs, and statistics of random networks
  multi_cases.json.gz  outputs of the reference's multi-network engine and
                    cross-network statistics compiled from source, on synthetic
                    multi-network problems (every cross effect, D = 2..4,
                    symmetric networks, masks, composition, per-chain rows)
  rate_cases.json.gz   outputs of the reference engine compiled from source on
                    periods with non-basic rate effects (covariate, out-degree,
                    in-degree kinds, clamping, conditional, symmetric) and
                    composition change (active lists, restore of inactive actors)
  large_cases.json.gz  outputs of the reference engine compiled from source at
                    the benchmark configurations' sizes: 500 actors (8 chains),
                    300-actor co-evolution, 1000 actors with twelve effects, and
                    three groups of the uneven batch at 1045-1188 actors
  abi_offsets.txt   field offsets of the reference ABI structs
  targets.json.gz, effects.json.gz, estimate.json.gz, multinet.json.gz,
  data/*.json.gz (the seven data-engine datasets of test-data.R: s50, s50_beh,
  s50miss, s50str, knecht, cow, synth):
                    copies of the reference's host-layer fixtures (inputs and
                    expected outputs)
"""
import gzip
import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
REF_FIX = "/root/reference/tests/testthat/fixtures"


def dump(name, obj):
    with gzip.open(os.path.join(HERE, name), "wt") as f:
        json.dump(obj, f, separators=(",", ":"))


def main():
    import oracle as O
    from ref_cases import build_cases

    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    assert O.reference() is not None, "oracle/_ref not built"
    # 1. the reference's simulator fixture, verbatim data
    with gzip.open(os.path.join(REF_FIX, "sim", "sim.json.gz"), "rt") as f:
        dump("sim.json.gz", json.load(f))
    for src, dst in [("targets/targets.json.gz", "targets.json.gz"),
                     ("effects/effects.json.gz", "effects.json.gz"),
                     ("estimate/estimate.json.gz", "estimate.json.gz"),
                     ("multinet/multinet.json.gz", "multinet.json.gz")]:
        shutil.copyfile(os.path.join(REF_FIX, src), os.path.join(HERE, dst))
    os.makedirs(os.path.join(HERE, "data"), exist_ok=True)
    for name in ("s50", "s50_beh", "s50miss", "s50str", "knecht", "cow", "synth"):
        shutil.copyfile(os.path.join(REF_FIX, "data", f"{name}.json.gz"),
                        os.path.join(HERE, "data", f"{name}.json.gz"))
    # 2. reference-engine outputs on synthetic cases
    dump("ref_cases.json.gz", build_cases(O))
    from multi_cases import build_cross_cases, build_multi_cases
    from rate_cases import build_rate_cases

    dump("rate_cases.json.gz", build_rate_cases(O))
    from large_cases import build_large_cases

    dump("large_cases.json.gz", build_large_cases(O))

    dump("multi_cases.json.gz", {"multi": build_multi_cases(O)["cases"],
                                 "cross": build_cross_cases(O)["cases"]})
    # 3. ABI layout of the reference header
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "abi_offsets_ref")],
                         capture_output=True, text=True, check=True).stdout
    with open(os.path.join(HERE, "abi_offsets.txt"), "w") as f:
        f.write(out)


if __name__ == "__main__":
    main()
