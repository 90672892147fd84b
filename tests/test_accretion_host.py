"""Host parts of the gas-accretion variant's operators (CPU)."""

import importlib

import numpy as np
import pytest

import oracle as O

acc = importlib.import_module("quantum-jax_amd.accretion")


@pytest.mark.parametrize("ramp", [None, (17.0, 0.8)])
def test_seed_mass_from_time_matches_oracle(ramp):
    kw = {} if ramp is None else dict(ramp_center=ramp[0], ramp_w=ramp[1])
    for t in np.linspace(10.0, 25.0, 61):
        for placed in (0.0, 1.0):
            assert acc.seed_mass_from_time(t, placed, 1e7, 13.0, **kw) == pytest.approx(
                O.seed_mass_from_time(t, placed, 1e7, 13.0, **kw), rel=1e-12, abs=0)


def test_lambda_iso():
    assert acc.LAMBDA_ISO == pytest.approx(np.exp(1.5) / 4, rel=1e-15)
