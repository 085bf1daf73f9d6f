"""CPU checks of the route workload's wind inputs (stormbird_amd/route.py): the von Karman
spectra and their discretization as the reference's GustSpectrum.discretize
(interfaces/stormbird_setup/.../gust_spectrum.py:17-37), and the apparent-wind
convention against the oracle's WindEnvironment (environment.rs:73-190)."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from stormbird_amd import route as R


@pytest.mark.parametrize("lateral", [False, True])
def test_von_karman_spectra_integrate_to_the_variance(lateral):
    sp = R.VonKarman(11.0, 0.15, 200.0, lateral=lateral)
    total, _ = quad(lambda f: float(sp.spectrum_value(f)), 0.0, np.inf, limit=400)
    assert total == pytest.approx(sp.standard_deviation ** 2, rel=2e-3)
    # at high frequency both fall off as f^(-5/3) (the inertial range)
    f1, f2 = 50.0, 100.0
    slope = math.log(sp.spectrum_value(f2) / sp.spectrum_value(f1)) / math.log(f2 / f1)
    assert slope == pytest.approx(-5.0 / 3.0, abs=0.01)


def test_discretize_follows_the_reference_recipe():
    sp = R.VonKarman(9.0, 0.15, 200.0)
    d = sp.discretize(0.001, 0.5, 500, np.random.default_rng(4))
    f = np.asarray(d["frequencies"])
    assert f.shape == (500,) and f[0] == 0.001 and f[-1] == 0.5
    assert np.allclose(np.diff(f), (0.5 - 0.001) / 499, rtol=0, atol=1e-15)
    ph = np.asarray(d["phases"])
    assert np.all((ph >= 0.0) & (ph < 2 * math.pi))
    # the harmonics carry the spectrum's variance over the band, each frequency the bin of
    # width df around it (a midpoint rule)
    df = (0.5 - 0.001) / 499
    band, _ = quad(lambda x: float(sp.spectrum_value(x)), 0.001 - df / 2, 0.5 + df / 2, limit=200)
    assert 0.5 * np.sum(np.asarray(d["amplitudes"]) ** 2) == pytest.approx(band, rel=1e-3)
    # and the signal's mean square over many periods of the lowest frequency is that variance
    t = np.arange(0.0, 20000.0, 0.25)
    a, p = np.asarray(d["amplitudes"]), ph
    x = np.zeros_like(t)
    for k in range(f.shape[0]):
        x += a[k] * np.sin(2.0 * math.pi * f[k] * t + p[k])
    assert np.mean(x * x) == pytest.approx(0.5 * np.sum(a * a), rel=0.03)
    assert R.gust_at(d, 123.4) == pytest.approx(
        sum(a[k] * math.sin(2.0 * math.pi * f[k] * 123.4 + p[k]) for k in range(f.shape[0])), abs=1e-12)


def test_apparent_wind_convention_matches_the_wind_environment(oracle):
    import oracle_csm

    m = oracle_csm.Libm("glibc")
    env = oracle_csm.WindEnvironment(R.wind_environment(), m)
    for direction, u, ship in ((0.9, 10.0, 5.0), (2.4, 13.0, 4.0)):
        c = oracle_csm.WindCondition({"direction_coming_from": direction,
                                      "velocity_variation": {"PowerModel": {"reference_velocity": u}}}, m)
        linear = oracle_csm.scale(env.zero, ship)
        # the sails' foot sits on the deck, 10 m above the water: the reference height
        want = env.steady_apparent(c, (0.0, 0.0, 0.0), linear)
        assert np.allclose(R.apparent_wind(direction, u, ship), want, rtol=0, atol=1e-12)
