"""The autotuner's selection rule (dust_amd.autotune.choose): the fastest
schedule within the accuracy bound, never a faster one outside it."""
from dust_amd import autotune as AT


def test_fastest_within_tolerance():
    table = [dict(ms=9.0, err=1e-8, k=0), dict(ms=4.0, err=5e-6, k=1),
             dict(ms=6.0, err=8e-7, k=2), dict(ms=6.0, err=2e-7, k=3)]
    assert AT.choose(table, 1e-6)["k"] == 2      # 4 ms is outside the bound; tie -> earlier
    assert AT.choose(table, 1e-5)["k"] == 1
    assert AT.choose(table, 1e-9) is None
    assert AT.choose([], 1.0) is None
