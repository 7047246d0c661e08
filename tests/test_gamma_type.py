"""The gammaType thermostat on the device (MD.cpp:135-138, langevin.h:233-262).

With a gammaType line and no scalar gamma, the reference's Langevin force uses
gammaType[0] for every particle and computes sigma once, at the first call's temperature;
a later temperature change (the tempStepInterval ramp, MD.cpp:367-371) leaves sigma as it
was.  The scalar-gamma thermostat recomputes sigma at every call.
"""
import numpy as np
import pytest

import softmold_amd as sm
from conftest import gpu_available

pytestmark = pytest.mark.gpu


def _run(d, T_after=None):
    s = sm.Solver.from_system(d)
    s.init_forces()
    s.run(1, 10)
    if T_after is not None:
        s.set_temperature(T_after)
    s.run(11, 10)
    out = s.download()[0]
    s.close()
    return out


def test_gamma_type_sigma_ignores_temperature_changes():
    if not gpu_available():
        pytest.skip("no GPU")
    d = sm.Mpd.liposome(4, 600, 3.45).system()        # delta_lxy 0: T enters through sigma only
    typed = dict(d, gamma=0.5, sigma_temperature=d["temperature"])
    scalar = dict(d, gamma=0.5)
    print("case: gammaType vs scalar gamma 0.5", flush=True)
    a, b = _run(typed), _run(scalar)
    assert np.array_equal(a, b)                        # same sigma at the first temperature
    assert np.array_equal(_run(typed, T_after=4.5), a)  # sigma kept
    assert not np.array_equal(_run(scalar, T_after=4.5), b)
