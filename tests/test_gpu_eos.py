"""The device equation-of-state functor (GeneralEOS, generalEOS.h:6-23, as IdealGasEOS,
idealGasEOS.h:7-46) through ts_eos_eval.

* The five interface methods against the reference's formulas evaluated in IEEE double in the
  same operation order: bit-identical (the functor uses IEEE division and square root there).
* Identities that fix each method independently of how it is written: e(rho, p(rho, e)) = e,
  c^2 = gamma p / rho, and the two derivatives against central differences of p(rho, e).
* gamma 1.4 and 5/3.  The volumetric forms the flux kernels call ((gamma - 1) rho e and the
  sound speed from p and 1/rho) are covered by the parity tests of those fluxes.
"""
import numpy as np
import pytest

from turbostructred_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0])
def test_ideal_gas_methods(gamma):
    rng = np.random.default_rng(5)
    n = 4096
    ro = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), n))
    e = np.exp(rng.uniform(np.log(1e-2), np.log(1e4), n))
    p = (gamma - 1.0) * ro * e
    ev = lambda m, x: capi.eos_eval(gamma, m, ro, x)
    # reference formulas, same operation order (idealGasEOS.h:22-44)
    assert np.array_equal(ev(capi.EOS_PRESSURE, e), (gamma - 1.0) * ro * e)
    assert np.array_equal(ev(capi.EOS_SOUND_SPEED, e), np.sqrt(gamma * ((gamma - 1.0) * ro * e) / ro))
    assert np.array_equal(ev(capi.EOS_INTERNAL_ENERGY, p), p / ((gamma - 1.0) * ro))
    assert np.array_equal(ev(capi.EOS_DP_DE, e), (gamma - 1.0) * ro)
    assert np.array_equal(ev(capi.EOS_DP_DRHO, e), (gamma - 1.0) * e)
    # identities
    pd = ev(capi.EOS_PRESSURE, e)
    np.testing.assert_allclose(ev(capi.EOS_INTERNAL_ENERGY, pd), e, rtol=4e-16)
    np.testing.assert_allclose(ev(capi.EOS_SOUND_SPEED, e) ** 2, gamma * pd / ro, rtol=1e-15)
    h = 1e-6
    de = (capi.eos_eval(gamma, capi.EOS_PRESSURE, ro, e * (1 + h)) -
          capi.eos_eval(gamma, capi.EOS_PRESSURE, ro, e * (1 - h))) / (2 * h * e)
    np.testing.assert_allclose(ev(capi.EOS_DP_DE, e), de, rtol=1e-8)
    dr = (capi.eos_eval(gamma, capi.EOS_PRESSURE, ro * (1 + h), e) -
          capi.eos_eval(gamma, capi.EOS_PRESSURE, ro * (1 - h), e)) / (2 * h * ro)
    np.testing.assert_allclose(ev(capi.EOS_DP_DRHO, e), dr, rtol=1e-8)


def test_eos_eval_rejects_a_bad_method():
    with pytest.raises(capi.TSError):
        capi.eos_eval(1.4, 7, [1.0], [1.0])
