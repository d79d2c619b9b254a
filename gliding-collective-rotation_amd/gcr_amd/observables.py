"""Observables of the collective state (host-side, from downloaded arrays).

- chain_mean_angle: a chain's azimuth from the sum of its beads' tangents
  (output_methods.py:492-510, calc_avg_phi_chain with n-fold symmetry 1);
- polar_order: |sum_c exp(i phi_c)| / C (analysis.py:315-328);
- nematic_order: order parameter and director angle from the Q-tensor sums of
  the chains' mean angles (analysis.py:294-313);
- rotation_rate: angular velocity of the director once the system is
  nematically ordered, from director-angle differences over a fixed lag
  (the `rotation_rate1` estimator of the paper notebooks);
- Unwrapper / msd: mean-square displacement from periodic positions sampled
  often enough that no bead moves half a box between samples.
"""
import numpy as np


def chain_mean_angle(tangent, chain_id, n_chains=None):
    if n_chains is None:
        n_chains = int(chain_id.max()) + 1
    sx = np.bincount(chain_id, weights=tangent[:, 0], minlength=n_chains)
    sy = np.bincount(chain_id, weights=tangent[:, 1], minlength=n_chains)
    return np.arctan2(sy, sx)


def polar_order(phi):
    return float(np.hypot(np.sum(np.cos(phi)), np.sum(np.sin(phi))) / len(phi))


def nematic_order(phi):
    """(S, director angle) as analysis.nematic_orderparameter."""
    q11 = np.sum(2 * np.cos(phi) ** 2 - 1.0)
    q12 = np.sum(2 * np.cos(phi) * np.sin(phi))
    s_full = np.hypot(q11, q12)
    director = np.arctan((s_full - q11) / q12) if q12 != 0 else 0.0
    return float(s_full / len(phi)), float(director)


def angle_difference(a, b, period=2 * np.pi):
    """a - b folded into (-period/2, period/2]. The notebook estimator folds
    modulo 2 pi (the phase of exp(i a) / exp(i b)); period = pi treats the
    director as headless."""
    d = np.mod(np.asarray(a) - np.asarray(b) + period / 2, period) - period / 2
    return np.where(d == -period / 2, period / 2, d)


def rotation_rate(t, director, S, threshold=0.7, lag=4, time_unit=1.0, period=2 * np.pi):
    """Director angular velocity after the nematic order first exceeds
    `threshold`: (angle[k+lag] - angle[k]) / (t[k+lag] - t[k]) for every k
    from there on, as the paper notebooks' `rotation_rate1` (which uses
    time_unit = 15.5). Returns (times, rates) or None if never ordered."""
    t, director, S = map(np.asarray, (t, director, S))
    idx = int(np.argmax(S > threshold))
    if not S[idx] > threshold or len(t) - idx <= lag:
        return None
    k = np.arange(idx, len(t) - lag)
    dt = (t[k + lag] - t[k]) / time_unit
    da = angle_difference(director[k + lag], director[k], period)
    return t[k + lag] / time_unit, da / dt


def remove_outliers(rates, times, threshold=2.0):
    """Drop rates whose z-score exceeds `threshold` (the notebooks' filter)."""
    rates, times = np.asarray(rates), np.asarray(times)
    sd = rates.std()
    keep = np.abs(rates - rates.mean()) <= threshold * sd if sd > 0 else np.ones(len(rates), bool)
    return times[keep], rates[keep]


class Unwrapper:
    """Accumulates periodic positions into unwrapped trajectories."""

    def __init__(self, pos0, L):
        self.L = L
        self.last = np.array(pos0, dtype=np.float64)
        self.unwrapped = self.last.copy()
        self.origin = self.last.copy()

    def add(self, pos):
        d = pos - self.last
        d -= self.L * np.round(d / self.L)
        self.unwrapped += d
        self.last = np.array(pos, dtype=np.float64)
        return self.unwrapped

    def msd(self):
        d = self.unwrapped - self.origin
        return float(np.mean(np.sum(d * d, axis=1)))

    def chain_com_msd(self, chain_id):
        """MSD of chain centres of mass (removes the beads' motion along the chain)."""
        n = int(chain_id.max()) + 1
        d = self.unwrapped - self.origin
        cnt = np.bincount(chain_id, minlength=n)
        com = np.stack([np.bincount(chain_id, weights=d[:, i], minlength=n) / cnt
                        for i in range(d.shape[1])], axis=1)
        return float(np.mean(np.sum(com * com, axis=1)))
