"""The driver's barostat plan (MDBackend::_is_barostat_active: a volume
move tried in each MD step with probability barostat_probability,
src/Backends/MDBackend.cpp:43-60,121-126): the steps of the tries, drawn
as geometric gaps, have the per-step rate of the input; the keys the
reference requires are required."""
import numpy as np
import pytest


def test_tries_per_step_follow_the_probability():
    from oxdna_amd import driver
    inp = dict(use_barostat="true", P="1e-4", delta_L="0.5", barostat_probability="0.05")
    plan = driver.barostat_plan(inp, 7, 1000, 201000)
    steps = np.array(sorted(plan["steps"]))
    assert steps.min() > 1000 and steps.max() <= 201000
    n = len(steps)
    # binomial(200000, 0.05): mean 10000, sd 97
    assert abs(n - 10000) <= 5 * np.sqrt(200000 * 0.05 * 0.95)
    assert plan["isotropic"] and not plan["molecular"]
    # the same seed gives the same plan; another seed another
    assert driver.barostat_plan(inp, 7, 1000, 201000)["steps"] == plan["steps"]
    assert driver.barostat_plan(inp, 8, 1000, 201000)["steps"] != plan["steps"]


def test_every_step_with_probability_one_and_keys_required():
    from oxdna_amd import driver
    inp = dict(use_barostat="1", P="1e-4", delta_L="0.5", barostat_probability="1", barostat_molecular="true",
               barostat_isotropic="false")
    plan = driver.barostat_plan(inp, 1, 0, 50)
    assert plan["steps"] == set(range(1, 51)) and plan["molecular"] and not plan["isotropic"]
    assert driver.barostat_plan({}, 1, 0, 50)["steps"] == set()
    for key in ("P", "delta_L", "barostat_probability"):
        bad = dict(inp)
        del bad[key]
        with pytest.raises(KeyError):
            driver.barostat_plan(bad, 1, 0, 50)
    with pytest.raises(ValueError):
        driver.barostat_plan(dict(inp, barostat_probability="0"), 1, 0, 50)
