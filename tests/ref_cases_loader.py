"""Import shim: the golden-case specifications live beside their generator."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from ref_cases import (beh_stats_inputs, case_specs, spec_problem, stats_inputs,  # noqa: E402,F401
                       stats_specs)
