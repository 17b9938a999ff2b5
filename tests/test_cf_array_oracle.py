"""CORRELATION_ARRAY weights of the oracle (calculate_correlation_array_and_save,
src/hawaii.c:2639-2674), pinned by closed forms and by what Boltzmann reweighting must do.

w = exp(-E HkT / Tb / Ts (Tb - Ts)) / WTT is exp(-E HkT (1/Ts - 1/Tb)) / WTT, WTT the same factor
at E = pesmin for bound pairs below the base temperature; free pairs take no weight above it.
A swapped Tb/Ts or a dropped sign flips which side of Tb is enhanced, which the closed forms and
the ensemble-mean check below catch.
"""
import numpy as np

import oracle_lib as O

HKT = 4.3597447222071e-18 / 1.380649e-23
FREE, BOUND = 1, 2


def test_weights_are_one_at_the_base_temperature():
    E = np.linspace(-3e-4, 2e-3, 50)
    for ps in (FREE, BOUND):
        w, ns = O.cf_array_weights(E, 120.0, ps, -4e-4, [120.0])
        assert np.array_equal(w[:, 0], np.ones(50))
        assert ns[0] == 50.0


def test_free_weights_closed_form_and_no_weight_above_base():
    rng = np.random.default_rng(3)
    E = rng.uniform(0.0, 3e-3, 200)
    temps = [60.0, 100.0, 150.0]
    w, ns = O.cf_array_weights(E, 100.0, FREE, -4e-4, temps, ratios=[2.0, 3.0, 5.0])
    ref = np.exp(-E * HKT * (1.0 / 60.0 - 1.0 / 100.0))
    assert np.allclose(w[:, 0], 2.0 * ref, rtol=1e-13, atol=0)
    assert np.all(w[:, 0] < 2.0)               # colder: high-energy trajectories lose weight
    assert np.allclose(w[:, 1], 3.0)
    assert np.all(w[:, 2] == 0.0) and ns[2] == 0.0
    assert np.isclose(ns[0], ref.sum(), rtol=1e-13)    # nstar carries no ratio


def test_bound_weights_closed_form_with_the_cap():
    pesmin = -4.5e-4
    rng = np.random.default_rng(4)
    E = rng.uniform(pesmin, 0.0, 200)
    w, ns = O.cf_array_weights(E, 100.0, BOUND, pesmin, [50.0, 200.0])
    below = np.exp(-(E - pesmin) * HKT * (1.0 / 50.0 - 1.0 / 100.0))
    above = np.exp(-E * HKT * (1.0 / 200.0 - 1.0 / 100.0))
    assert np.allclose(w[:, 0], below, rtol=1e-12)
    assert np.all(w[:, 0] <= 1.0 + 1e-15)      # capped at the well bottom
    assert np.allclose(w[:, 1], above, rtol=1e-13)
    assert np.all(w[:, 1] <= 1.0)              # bound (E < 0) pairs lose weight when hotter


def test_reweighting_moves_the_boltzmann_mean():
    # energies of a two-degree-of-freedom Boltzmann ensemble at Tb: E ~ Gamma(2, Tb / HkT);
    # reweighted to Ts the mean must become 2 Ts / HkT
    rng = np.random.default_rng(5)
    Tb = 200.0
    E = rng.gamma(2.0, Tb / HKT, 400000)
    for Ts in (120.0, 160.0):
        w, ns = O.cf_array_weights(E, Tb, FREE, 0.0, [Ts])
        mean = (w[:, 0] * E).sum() / ns[0]
        assert abs(mean / (2.0 * Ts / HKT) - 1.0) < 0.01


def test_skipped_trajectories_take_no_weight():
    # the array loop's `continue` on the step limit (src/hawaii.c:2637): no weight, not in nstar
    E = np.array([1e-4, 2e-4, 3e-4, 4e-4])
    w_all, ns_all = O.cf_array_weights(E, 100.0, FREE, 0.0, [80.0])
    w, ns = O.cf_array_weights(E, 100.0, FREE, 0.0, [80.0], counted=O.array_counted([0, 4, 1, 0]))
    assert w[1, 0] == 0.0 and w[2, 0] == w_all[2, 0]      # a dipole failure still counts
    assert np.isclose(ns[0], ns_all[0] - w_all[1, 0], rtol=1e-15)
