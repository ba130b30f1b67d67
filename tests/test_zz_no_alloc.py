"""GPU: cusna_saom_simulate_period_noalloc, the period call that allocates no
device memory, against cusna_saom_simulate_period (the allocating entry
point) on the same inputs: identical outputs, bit for bit, on two calls in a
row (the second smaller, so it runs over what the first left in the static
array), and a clear refusal of a call larger than the array.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _assert_same(a, b, label):
    assert a.keys() == b.keys(), label
    for key in a:
        np.testing.assert_array_equal(np.asarray(a[key]), np.asarray(b[key]),
                                      err_msg=f"{label} {key}")


def _calls():
    from cusna_amd.effects import Effect
    from cusna_amd.workloads import (coevo300_problem, effect_problem, n500_problem,
                                     poisson_capped_network)

    first = n500_problem(M=64, n=500)
    # the 50-actor fit's shape at the R glue's phase-2 batch, conditional
    rng = np.random.default_rng(8)
    x = poisson_capped_network(50, 2.3, 6, 0.4, rng)
    second = effect_problem(x, [Effect("density"), Effect("recip"), Effect("transTrip")],
                            [-2.5, 2.0, 0.4], 1.0, 256, 44, conditional=True, target_dist=40)
    third = coevo300_problem(M=8)
    return [("n500 M=64", first, True), ("n50 M=256 conditional", second, True),
            ("coevo n300 M=8", third, False)]


def test_noalloc_matches_allocating_entry(gpu_lib):
    for label, prob, nets in _calls():
        print(f"case noalloc {label}", flush=True)
        got = gpu_lib.simulate_period_noalloc(prob, want_end_nets=nets)
        want = gpu_lib.simulate_period(prob, want_end_nets=nets)
        _assert_same(got, want, label)
    # the first call again, after smaller ones: nothing left over changes it
    label, prob, nets = _calls()[0]
    _assert_same(gpu_lib.simulate_period_noalloc(prob, want_end_nets=nets),
                 gpu_lib.simulate_period(prob, want_end_nets=nets), label + " again")


def test_noalloc_refuses_calls_above_the_array(gpu_lib):
    from cusna_amd.abi import CusnaError
    from cusna_amd.workloads import n500_problem

    big = n500_problem(M=4096, n=500)  # 131 MB of packed end networks alone
    with pytest.raises(CusnaError, match="status 2") as e:
        gpu_lib.simulate_period_noalloc(big)
    assert "bytes" in str(e.value)
    # a refused call leaves the entry point usable
    small = n500_problem(M=8, n=120)
    _assert_same(gpu_lib.simulate_period_noalloc(small), gpu_lib.simulate_period(small), "after")
