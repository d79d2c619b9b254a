"""Observables (CPU): closed-form cases for the order parameters of
analysis.py:294-328, the rotation-rate estimator and the unwrapped MSD."""
import numpy as np
import pytest

from gcr_amd import observables as OB


def test_chain_mean_angle_and_polar_order():
    chain = np.repeat(np.arange(3), 4)
    ang = np.array([0.3, 1.2, -2.0])
    t = np.stack([np.cos(np.repeat(ang, 4)), np.sin(np.repeat(ang, 4))], axis=1)
    phi = OB.chain_mean_angle(t, chain)
    assert np.allclose(phi, ang)
    assert OB.polar_order(np.zeros(10)) == pytest.approx(1.0)
    assert OB.polar_order(np.array([0.0, np.pi])) == pytest.approx(0.0, abs=1e-15)


def test_nematic_order_closed_form():
    # aligned head-to-tail: polar 0, nematic 1, director along the chains
    phi = np.array([0.4, 0.4 + np.pi, 0.4, 0.4 + np.pi])
    S, d = OB.nematic_order(phi)
    assert S == pytest.approx(1.0) and OB.polar_order(phi) == pytest.approx(0, abs=1e-12)
    assert d == pytest.approx(0.4)
    S, _ = OB.nematic_order(np.array([0.0, np.pi / 2]))
    assert S == pytest.approx(0.0, abs=1e-12)


def test_rotation_rate_uniform_rotation():
    t = np.linspace(0, 10, 101)
    omega = 0.05
    director = np.mod(omega * t + np.pi / 2, np.pi) - np.pi / 2
    S = np.where(t > 2.0, 0.9, 0.1)
    times, rates = OB.rotation_rate(t, director, S, period=np.pi)
    assert times[0] == pytest.approx(t[np.argmax(S > 0.7) + 4])
    assert np.allclose(rates, omega)
    assert OB.rotation_rate(t, director, np.zeros_like(t)) is None


def test_unwrapped_msd():
    L = 10.0
    rng = np.random.default_rng(0)
    x0 = rng.uniform(0, L, (50, 2))
    u = OB.Unwrapper(x0, L)
    x = x0.copy()
    true = np.zeros_like(x)
    for _ in range(200):
        d = rng.normal(0, 0.3, x.shape)
        true += d
        x = np.mod(x + d, L)
        u.add(x)
    assert np.allclose(u.unwrapped - u.origin, true)
    assert u.msd() == pytest.approx(np.mean(np.sum(true ** 2, axis=1)))
    chain = np.repeat(np.arange(10), 5)
    com = np.stack([true[chain == c].mean(0) for c in range(10)])
    assert u.chain_com_msd(chain) == pytest.approx(np.mean(np.sum(com ** 2, axis=1)))
