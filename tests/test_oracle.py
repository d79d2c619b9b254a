"""Pins of the oracle (CPU, no GPU): each oracle function is checked against
what the mathematics fixes -- closed forms, energy gradients by central finite
differences, exact recurrences, invariants, published known-answer vectors --
and the C restatement is checked against the independent NumPy restatement.
The reference ships no golden vectors or tests for this path."""
import math

import numpy as np
import pytest

import gcr_oracle_np as N
import oracle_lib as O

L2D = 20.0


def lj_energy(r, p):
    """U(r) whose -dU/dr gives the pair law of forces.py:13-15:
    U = eps ((rho/r)^12 - 2 (rho/r)^6) with p0 = 12 eps rho^6, p1 = 12 eps rho^12."""
    return p[1] / 12.0 * r ** -12 - p[0] / 6.0 * r ** -6


def straight_chains(rng, n_chains, length, L, a=0.5, jitter=0.05, dim=2, min_dist=0.9):
    """Chains along random directions, rejected while a bead comes closer than
    min_dist to a bead of an earlier chain."""
    pos = np.zeros((n_chains * length, dim))
    tang = np.zeros_like(pos)
    for c in range(n_chains):
        for _ in range(10000):
            x0 = rng.uniform(0, L, dim)
            t = rng.normal(size=dim)
            t /= np.linalg.norm(t)
            cand = np.mod(x0 - a * np.arange(length)[:, None] * t
                          + rng.normal(0, jitter, (length, dim)), L)
            if c == 0:
                break
            d = N.min_image(cand[:, None, :] - pos[None, :c * length, :], L)
            if np.sqrt((d ** 2).sum(-1)).min() >= min_dist:
                break
        pos[c * length:(c + 1) * length] = cand
        tang[c * length:(c + 1) * length] = t
    return pos, tang, np.repeat(np.arange(n_chains), length).astype(np.int64)


# ------------------------------------------------------------------ pair law

def test_wca_params_closed_form():
    p = O.wca_params(1.0, 1.0, 0.0, 1.0)
    assert np.array_equal(p, [12.0, 12.0, 0.0, 1.0])
    p = O.wca_params(1.3, 0.7, 0.1, 2.5)
    assert np.allclose(p, [12 * 1.3 ** 6 * 0.7, 12 * 1.3 ** 12 * 0.7, 0.1, 2.5 * 1.3 - 0.1])
    assert np.allclose(p, N.wca_params(1.3, 0.7, 0.1, 2.5), rtol=1e-15)


@pytest.mark.parametrize("r0,eps", [(1.0, 1.0), (1.3, 0.7)])
def test_wca_coeff_is_minus_gradient_of_lj(r0, eps):
    p = O.wca_params(r0, eps, 0.0, 2.5)
    # zero force at the potential minimum r = repel_d_max
    assert abs(O.wca_coeff(r0 * r0, p)) < 1e-12 * p[0]
    for r in [0.8 * r0, 0.95 * r0, 1.2 * r0, 2.0 * r0]:
        h = 1e-6 * r
        dUdr = (lj_energy(r + h, p) - lj_energy(r - h, p)) / (2 * h)
        # force on b1 = coeff * (x2 - x1); radial component along (x1 - x2) is -coeff * r
        f_radial = -O.wca_coeff(r * r, p) * r
        assert f_radial == pytest.approx(-dUdr, rel=1e-7)
    # repulsive inside, attractive outside the minimum
    assert O.wca_coeff(0.9 * r0 * 0.9 * r0, p) < 0 < O.wca_coeff(1.5 * r0 * 1.5 * r0, p)


def test_pymod_python_semantics():
    for a in [-1e-17, -3.5, 0.0, -0.0, 7.25, 20.0, 41.0, -20.0]:
        assert O.pymod(a, L2D) == a % L2D
        assert math.copysign(1, O.pymod(a, L2D)) == math.copysign(1, a % L2D)
    assert O.pymod(-1e-17, 20.0) == 20.0  # the edge the kernels must clamp


def pair_energy(pos, chain_id, L, p, n_nearby):
    n = len(pos)
    E = 0.0
    for i in range(n):
        for j in range(i + 1, n):
            if chain_id[i] == chain_id[j] and (n_nearby < 0 or abs(i - j) <= n_nearby):
                continue
            d = N.min_image(pos[j] - pos[i], L)
            r = np.linalg.norm(d)
            if 0 < r <= p[3]:
                E += lj_energy(r, p)
    return E


@pytest.mark.parametrize("dim,n_nearby", [(2, 4), (2, -1), (3, 2)])
def test_pair_forces_are_energy_gradient(dim, n_nearby):
    rng = np.random.default_rng(3)
    L = 6.0
    p = O.wca_params(1.0, 1.0, 0.0, 2.5 / 2 ** (1 / 6))
    chain = np.repeat(np.arange(7), 2).astype(np.int64)
    pts = []
    while len(pts) < 14:  # sequential placement: no close pair, none near the cutoff
        x = rng.uniform(0, L, dim)
        ok = True
        for y in pts:
            r = np.linalg.norm(N.min_image(x - y, L))
            if r < 0.8 or abs(r - p[3]) < 1e-2:
                ok = False
        if ok:
            pts.append(x)
    pos = np.array(pts)
    F = O.pair_forces(pos, chain, L, p, n_nearby)
    h = 1e-6
    for b in range(len(pos)):
        for i in range(dim):
            pp, pm = pos.copy(), pos.copy()
            pp[b, i] += h
            pm[b, i] -= h
            g = (pair_energy(pp, chain, L, p, n_nearby) - pair_energy(pm, chain, L, p, n_nearby)) / (2 * h)
            assert F[b, i] == pytest.approx(-g, abs=1e-6 * max(1, abs(g)))
    assert np.allclose(F.sum(0), 0, atol=1e-10)  # Newton's third law
    assert np.allclose(F, N.pair_forces(pos, chain, L, p, n_nearby), rtol=1e-12, atol=1e-12)


def test_pair_forces_exclusion_and_shift():
    # two beads on one chain at distance 0.8: excluded when |i-j| <= n_nearby
    pos = np.array([[5.0, 5.0], [5.8, 5.0]])
    same = np.array([0, 0], dtype=np.int64)
    other = np.array([0, 1], dtype=np.int64)
    p = O.wca_params(1.0, 1.0, 0.0, 1.0)
    assert np.all(O.pair_forces(pos, same, 10.0, p, 1) == 0)
    assert np.all(O.pair_forces(pos, same, 10.0, p, -1) == 0)
    assert np.any(O.pair_forces(pos, same, 10.0, p, 0) != 0)
    F = O.pair_forces(pos, other, 10.0, p, 4)
    assert F[0, 0] < 0 and F[1, 0] > 0  # repulsion along x
    # WCA_shift s: force at distance d equals the unshifted force at d + s
    ps = O.wca_params(1.0, 1.0, 0.1, 1.0)
    Fs = O.pair_forces(pos, other, 10.0, ps, 4)
    f_expect = -O.wca_coeff(0.9 ** 2, p) * 0.9  # radial magnitude at 0.8 + 0.1
    assert Fs[1, 0] == pytest.approx(f_expect, rel=1e-13)
    # periodic image: beads across the boundary interact with the minimum image
    pos = np.array([[0.1, 5.0], [9.6, 5.0]])
    F = O.pair_forces(pos, other, 10.0, p, 4)
    assert F[0, 0] > 0 and F[1, 0] < 0


def test_pair_cutoff_inclusive():
    p = O.wca_params(1.0, 1.0, 0.0, 2.0)
    pos = np.array([[1.0, 1.0], [3.0, 1.0]])  # exactly at rc = 2
    F = O.pair_forces(pos, np.array([0, 1], dtype=np.int64), 10.0, p, 4)
    assert F[0, 0] != 0  # 0 < ds <= rc^2 includes rc (forces.py:122)


# --------------------------------------------------------------- chain forces

def bond_energy(pos, chain, L, k, a):
    E = 0.0
    for b in range(len(pos) - 1):
        if chain[b] == chain[b + 1]:
            t = N.min_image(pos[b] - pos[b + 1], L)
            E += 0.5 * k * (np.linalg.norm(t) - a) ** 2
    return E


def bend_energy_terms(pos, chain, L, kappa, theta0):
    out = []
    for m in range(1, len(pos) - 1):
        if chain[m - 1] == chain[m] == chain[m + 1]:
            ti = N.min_image(pos[m] - pos[m + 1], L)
            tj = N.min_image(pos[m - 1] - pos[m], L)
            ti /= np.linalg.norm(ti)
            tj /= np.linalg.norm(tj)
            th = np.arccos(np.clip(ti @ tj, -1, 1))
            s = np.sign(-ti[0] * tj[1] + ti[1] * tj[0]) if theta0 else 0.0
            out.append((m, 0.5 * kappa * (th + theta0 * s) ** 2))
    return out


def test_hookean_bonds_are_energy_gradient():
    rng = np.random.default_rng(5)
    pos, tang, chain = straight_chains(rng, 3, 6, L2D, jitter=0.08)
    sp = np.array([37.0, 0.0, 0.0, 0.0, 0.0, 0.1])
    rl = np.array([0.5, 0.45, 0.6])
    F, _ = O.chain_forces(pos, chain, L2D, sp, rl)
    h = 1e-6
    for b in range(len(pos)):
        for i in range(2):
            pp, pm = pos.copy(), pos.copy()
            pp[b, i] += h
            pm[b, i] -= h
            # rest length is per chain (rest_lengths[chain_id])
            ep = sum(bond_energy(pp[chain == c], chain[chain == c], L2D, 37.0, rl[c]) for c in range(3))
            em = sum(bond_energy(pm[chain == c], chain[chain == c], L2D, 37.0, rl[c]) for c in range(3))
            assert F[b, i] == pytest.approx(-(ep - em) / (2 * h), abs=1e-6)
    assert np.allclose(F.sum(0), 0, atol=1e-12)


def fene_energy(pos, chain, L, k, a, tol):
    """U = -1/2 k m^2 ln(1 - s^2/m^2) per bond, s = |t| - a, m = tol * a: the
    potential whose gradient the FENE branch of forces.py:331-341 gives while
    |s| < m."""
    E = 0.0
    m = tol * a
    for b in range(len(pos) - 1):
        if chain[b] == chain[b + 1]:
            s = np.linalg.norm(N.min_image(pos[b] - pos[b + 1], L)) - a
            E += -0.5 * k * m * m * np.log(1.0 - s * s / (m * m))
    return E


def test_fene_bonds_are_energy_gradient():
    rng = np.random.default_rng(6)
    pos, tang, chain = straight_chains(rng, 3, 6, L2D, jitter=0.02)
    tol = 0.3
    sp = np.array([37.0, 0.0, 0.0, 1.0, 0.0, tol])
    rl = np.array([0.5, 0.45, 0.6])
    for c in range(3):  # every bond inside the FENE range
        d = np.linalg.norm(N.min_image(np.diff(pos[chain == c], axis=0), L2D), axis=1) - rl[c]
        assert np.all(np.abs(d) < 0.8 * tol * rl[c])
    F, _ = O.chain_forces(pos, chain, L2D, sp, rl)
    Fn, _ = N.chain_forces(pos, chain, L2D, sp, rl)
    assert np.allclose(F, Fn, rtol=1e-12, atol=1e-11)
    h = 1e-7
    for b in range(len(pos)):
        for i in range(2):
            pp, pm = pos.copy(), pos.copy()
            pp[b, i] += h
            pm[b, i] -= h
            ep = sum(fene_energy(pp[chain == c], chain[chain == c], L2D, 37.0, rl[c], tol) for c in range(3))
            em = sum(fene_energy(pm[chain == c], chain[chain == c], L2D, 37.0, rl[c], tol) for c in range(3))
            assert F[b, i] == pytest.approx(-(ep - em) / (2 * h), rel=1e-5, abs=1e-6)
    # a dimer: the FENE force is the Hookean one divided by 1 - s^2/m^2
    dim_pos = np.array([[5.0, 5.0], [5.0 + 0.56, 5.0]])
    two = np.zeros(2, dtype=np.int64)
    Ff, _ = O.chain_forces(dim_pos, two, L2D, sp, np.array([0.5]))
    Fh, _ = O.chain_forces(dim_pos, two, L2D, np.array([37.0, 0, 0, 0, 0, tol]), np.array([0.5]))
    s_, m_ = 0.06, tol * 0.5
    assert Ff[0, 0] == pytest.approx(Fh[0, 0] / (1 - s_ * s_ / (m_ * m_)), rel=1e-12)


@pytest.mark.parametrize("ell,theta0", [(1.0, 0.0), (0.5, 0.0), (1.3, 0.0), (0.7, 0.3)])
def test_bending_is_bond_length_times_energy_gradient(ell, theta0):
    """forces.py:361-386 applies kappa*theta/sin(theta) times unit-vector
    combinations with no 1/|bond| factor: for chains whose bonds all have
    length ell the bending force is ell * (-grad of 1/2 kappa (theta+theta0 s)^2)."""
    rng = np.random.default_rng(7)
    n = 7
    ang = np.cumsum(rng.uniform(-0.6, 0.6, n - 1)) + 0.3
    pos = np.zeros((n, 2))
    pos[0] = [10.0, 10.0]
    for b in range(1, n):
        pos[b] = pos[b - 1] - ell * np.array([np.cos(ang[b - 1]), np.sin(ang[b - 1])])
    chain = np.zeros(n, dtype=np.int64)
    kappa = 3.0
    sp = np.array([0.0, kappa, theta0, 0.0, 0.0, 0.1])
    F, _ = O.chain_forces(pos, chain, L2D, sp, np.array([ell]))
    # bond lengths stay ell only to first order under a displacement: the
    # gradient of the bending energy alone is what the formula scales
    h = 1e-6
    G = np.zeros_like(pos)
    for b in range(n):
        for i in range(2):
            pp, pm = pos.copy(), pos.copy()
            pp[b, i] += h
            pm[b, i] -= h
            ep = sum(e for _, e in bend_energy_terms(pp, chain, L2D, kappa, theta0))
            em = sum(e for _, e in bend_energy_terms(pm, chain, L2D, kappa, theta0))
            G[b, i] = (ep - em) / (2 * h)
    assert np.allclose(F, -ell * G, atol=1e-6 * kappa)
    assert np.allclose(F.sum(0), 0, atol=1e-12)
    # net torque vanishes for equal bond lengths
    tq = np.sum(pos[:, 0] * F[:, 1] - pos[:, 1] * F[:, 0])
    assert abs(tq) < 1e-9


def test_straight_chain_no_bending_and_tangents():
    n = 6
    t = np.array([0.6, 0.8])
    pos = np.array([[10.0, 10.0] - 0.5 * j * t for j in range(n)])
    chain = np.zeros(n, dtype=np.int64)
    sp = np.array([100.0, 10.0, 0.0, 0.0, 1.0, 0.1])
    F, T = O.chain_forces(pos, chain, L2D, sp, np.array([0.5]))
    assert np.allclose(F, 0, atol=1e-12)  # at rest length and straight
    assert np.allclose(T, t, atol=1e-15)  # tangent points towards bead 0
    # bent chain: interior tangent is the bisector (t_i + t_j)/|t_i + t_j|
    pos[3:] = pos[2] - 0.5 * np.outer(np.arange(1, n - 2), [1.0, 0.0])
    F, T = O.chain_forces(pos, chain, L2D, sp, np.array([0.5]))
    ti, tj = np.array([1.0, 0.0]), t
    bis = (ti + tj) / np.linalg.norm(ti + tj)
    assert np.allclose(T[2], bis, atol=1e-14)
    assert np.allclose(T[0], t) and np.allclose(T[-1], [1.0, 0.0])
    Fn, Tn = N.chain_forces(pos, chain, L2D, sp, np.array([0.5]))
    assert np.allclose(F, Fn, atol=1e-13) and np.allclose(T, Tn, atol=1e-15)


def test_chain_forces_c_vs_numpy_random():
    rng = np.random.default_rng(11)
    pos, tang, chain = straight_chains(rng, 5, 9, L2D, jitter=0.1)
    sp = np.array([100.0, 10.0, 0.2, 0.0, 1.0, 0.1])
    rl = rng.uniform(0.4, 0.6, 5)
    F, T = O.chain_forces(pos, chain, L2D, sp, rl)
    Fn, Tn = N.chain_forces(pos, chain, L2D, sp, rl)
    assert np.allclose(F, Fn, rtol=1e-12, atol=1e-11)
    assert np.allclose(T, Tn, atol=1e-14)


# ------------------------------------------------------------- self-propulsion

def test_self_propulsion_rotation():
    rng = np.random.default_rng(2)
    t = rng.normal(size=(8, 2))
    t /= np.linalg.norm(t, axis=1)[:, None]
    chain = np.repeat(np.arange(4), 2).astype(np.int64)
    ang = np.array([0.0, 0.2, -0.5, np.pi / 2])
    F = O.self_propel(t, chain, 1.7, ang)
    for b in range(8):
        a = ang[chain[b]]
        assert np.linalg.norm(F[b]) == pytest.approx(1.7, rel=1e-14)
        # signed angle from t to F equals v_tan_angle
        cross = t[b, 0] * F[b, 1] - t[b, 1] * F[b, 0]
        assert math.atan2(cross, t[b] @ F[b]) == pytest.approx(a, abs=1e-14)
    assert np.array_equal(F[:2], 1.7 * t[:2])  # angle 0: plain v0 t (forces.py:214-217)
    F3 = O.self_propel(np.array([[0.0, 0.6, 0.8]]), np.array([0], dtype=np.int64), 2.0, [0.3])
    assert np.allclose(F3, [[0.0, 1.2, 1.6]])  # 3D ignores the chiral angle


# ---------------------------------------------------------------------- SRK2

def abp_sys(n, L, D=0.0, Dr=0.0, v0=1.0, dim=2, wca=False):
    return dict(dim=dim, n=n, n_nearby=4, has_springs=False, has_wca=wca, has_propel=True, L=L,
                D=D, Dr=Dr, radius=1.0, wca=O.wca_params(1.0, 1.0, 0.0, 1.0),
                spring=np.zeros(6), v0=v0)


def test_srk2_free_abp_exact_drift():
    """T = 0, one ABP: F0 = F1 = v0 t, m = v0/r, so every step moves v0 dt_used t
    with dt = min(dtmax, mdps r / v0); the first predictor uses dt0."""
    L, v0, dtmax, mdps = 50.0, 2.0, 1e-3, 1e-3
    t = np.array([[0.6, -0.8]])
    x0 = np.array([[10.0, 10.0]])
    tv = np.array([1e-7, dtmax, mdps, 1.0, -1.0, 40, 40, 25])
    pos, pos1, T, F, st = O.srk2_run(abp_sys(1, L, v0=v0), x0, t, np.array([0]), np.array([0.5]),
                                     np.array([0.0]), tv, xi=np.zeros((25, 1, 4)))
    dt = min(dtmax, mdps / v0)
    assert st[0] == dt and st[2] == 25
    assert st[1] == pytest.approx(25 * dt, rel=1e-14)
    assert np.allclose(pos, x0 + v0 * 25 * dt * t, atol=1e-13)
    assert np.allclose(pos1, x0 + v0 * 24 * dt * t + v0 * dt * t, atol=1e-13)
    assert np.allclose(F, 2 * v0 * t)


def test_srk2_spring_dimer_heun_recurrence():
    """Two beads on one chain, T = 0, no WCA: the separation s obeys
    ds/dt = 2k(a - s)/r. SRK2 with the reference's sequencing is the scalar
    map  s1 = s + dt_prev 2k(a-s)/r,  m = |k(2a - s - s1)| / (2r),
    dt = min(dtmax, mdps/m),  s' = s + dt k((a-s) + (a-s1)) / r."""
    k, a, r = 50.0, 0.5, 1.3
    sysd = dict(dim=2, n=2, n_nearby=4, has_springs=True, has_wca=False, has_propel=False,
                L=20.0, D=0.0, Dr=0.0, radius=r, wca=np.zeros(4),
                spring=np.array([k, 0.0, 0.0, 0.0, 1.0, 0.1]), v0=0.0)
    s0, dtmax, mdps, dt0 = 0.9, 2e-3, 4e-3, 1e-6
    pos = np.array([[10.0 + s0, 7.0], [10.0, 7.0]])
    nsteps = 30
    tv = np.array([dt0, dtmax, mdps, 1.0, -1.0, 40, 40, nsteps])
    out = O.srk2_run(sysd, pos, np.zeros((2, 2)), np.array([0, 0]), np.array([a]),
                     np.array([0.0]), tv, xi=np.zeros((nsteps, 2, 4)))
    s, dt, dts = s0, dt0, []
    for _ in range(nsteps):
        s1 = s + dt * 2 * k * (a - s) / r
        m = 0.5 * abs(k * ((a - s) + (a - s1))) / r  # 1/2 max(F0 + F1) / r
        dt = min(dtmax, mdps / m)
        dts.append(dt)
        s = s + dt * k * ((a - s) + (a - s1)) / r
    sep = out[0][0, 0] - out[0][1, 0]
    assert sep == pytest.approx(s, rel=1e-12)
    assert out[4][0] == pytest.approx(dts[-1], rel=1e-12)
    assert out[4][1] == pytest.approx(sum(dts), rel=1e-12)
    assert len(set(np.round(dts, 12))) > 2  # the adaptive dt actually moved
    assert out[0][0, 0] + out[0][1, 0] == pytest.approx(20.0 + s0, rel=1e-14)  # centre fixed


def test_srk2_pure_diffusion_and_time_clip():
    """F = 0: x' = (x + sqrt(2 D dt_used) xi) mod L, dt = min(dtmax, mdps / sqrt(2D));
    max_dt_usable stops the loop with elapsed == max_dt_usable exactly."""
    D, L = 0.3, 10.0
    sysd = abp_sys(3, L, D=D, v0=0.0)
    rng = np.random.default_rng(0)
    xi = rng.normal(size=(10, 3, 4))
    x0 = rng.uniform(0, L, (3, 2))
    dtmax, mdps = 1e-2, 5e-3
    dt = min(dtmax, mdps / np.sqrt(2 * D))
    mdu = 3.5 * dt
    tv = np.array([1e-6, dtmax, mdps, 1.0, mdu, 40, 40, 10])
    pos, pos1, T, F, st = O.srk2_run(sysd, x0, np.tile([1.0, 0.0], (3, 1)), np.arange(3),
                                     np.ones(3), np.zeros(3), tv, xi=xi)
    assert st[2] == 4 and st[1] == pytest.approx(mdu, rel=1e-15)
    x = x0.copy()
    used = [dt, dt, dt, mdu - 3 * dt]
    for kk, du in enumerate(used):
        x = np.mod(x + np.sqrt(2 * du * D) * xi[kk, :, :2], L)
    assert np.allclose(pos, x, atol=1e-14)


def srk4_dimer(k, a, r, s0, dt0, dtmax, mdps, nsteps):
    """The SRK4 map of the spring dimer written out as a scalar recurrence
    (SRK.py:186-215, 297-321, 355-358): forces k_i = k (a - s_i) on the first
    bead, virtual separations from s with dt_prev / 2, dt_prev / 2, dt_prev,
    total k1/6 + k2/3 + k3/3 + k4/6, m = |total| / r (force prefactor 1),
    dt = min(dtmax, mdps / m), s' = s + dt 2 total / r. Returns the final
    separation, the last virtual separation and the dt sequence."""
    s, dt, dts, s3 = s0, dt0, [], s0
    for _ in range(nsteps):
        k1 = k * (a - s)
        s1 = s + 0.5 * dt * 2 * k1 / r
        k2 = k * (a - s1)
        s2 = s + 0.5 * dt * 2 * k2 / r
        k3 = k * (a - s2)
        s3 = s + dt * 2 * k3 / r
        k4 = k * (a - s3)
        tot = k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6
        dt = min(dtmax, mdps / (abs(tot) / r))
        dts.append(dt)
        s = s + dt * 2 * tot / r
    return s, s3, dts


def dimer_sys(k, r, order):
    return dict(dim=2, n=2, n_nearby=4, has_springs=True, has_wca=False, has_propel=False,
                L=20.0, D=0.0, Dr=0.0, radius=r, wca=np.zeros(4),
                spring=np.array([k, 0.0, 0.0, 0.0, 1.0, 0.1]), v0=0.0, order=order)


def test_srk4_spring_dimer_recurrence():
    """SRK_order 4 on the spring dimer equals the scalar SRK4 recurrence with
    the reference's sequencing, the adaptive dt included."""
    k, a, r = 50.0, 0.5, 1.3
    # from 0.7 the step moves from the max-distance limit (where every
    # integrator moves the same 2 mdps) to dt = dtmax
    s0, dtmax, mdps, dt0, nsteps = 0.7, 2e-3, 4e-3, 1e-6, 30
    pos = np.array([[10.0 + s0, 7.0], [10.0, 7.0]])
    tv = np.array([dt0, dtmax, mdps, 1.0, -1.0, 40, 40, nsteps])
    out = O.srk2_run(dimer_sys(k, r, 4), pos, np.zeros((2, 2)), np.array([0, 0]),
                     np.array([a]), np.array([0.0]), tv, xi=np.zeros((nsteps, 2, 4)))
    s, s3, dts = srk4_dimer(k, a, r, s0, dt0, dtmax, mdps, nsteps)
    assert out[0][0, 0] - out[0][1, 0] == pytest.approx(s, rel=1e-12)
    assert out[1][0, 0] - out[1][1, 0] == pytest.approx(s3, rel=1e-12)  # pos1: last virtual
    assert out[4][0] == pytest.approx(dts[-1], rel=1e-12)
    assert out[4][1] == pytest.approx(sum(dts), rel=1e-12)
    assert len(set(np.round(dts, 12))) > 2
    # and it is not the SRK2 trajectory
    out2 = O.srk2_run(dimer_sys(k, r, 2), pos, np.zeros((2, 2)), np.array([0, 0]),
                      np.array([a]), np.array([0.0]), tv, xi=np.zeros((nsteps, 2, 4)))
    assert abs((out2[0][0, 0] - out2[0][1, 0]) - s) > 1e-6


def test_srk4_constant_dt_is_the_rk4_polynomial():
    """With dt pinned at dtmax (a weak spring) every step multiplies the
    extension by the fourth-order Taylor polynomial of exp(-2 k dt / r)."""
    k, a, r = 2.0, 0.5, 1.0
    s0, dtmax, mdps, nsteps = 0.51, 1e-2, 1.0, 40
    pos = np.array([[10.0 + s0, 7.0], [10.0, 7.0]])
    tv = np.array([dtmax, dtmax, mdps, 1.0, -1.0, 40, 40, nsteps])
    out = O.srk2_run(dimer_sys(k, r, 4), pos, np.zeros((2, 2)), np.array([0, 0]),
                     np.array([a]), np.array([0.0]), tv, xi=np.zeros((nsteps, 2, 4)))
    assert out[4][0] == dtmax
    z = -2 * k * dtmax / r
    p4 = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    assert out[0][0, 0] - out[0][1, 0] - a == pytest.approx((s0 - a) * p4 ** nsteps, rel=1e-10)


def test_srk4_pure_diffusion_noise_amplitudes():
    """F = 0: the real update adds sqrt(2 D dt_used) xi and each virtual one
    sqrt(2 D dt_prev) xi from pos, the same xi in every stage (SRK.py:361)."""
    D, L = 0.3, 10.0
    sysd = abp_sys(3, L, D=D, v0=0.0)
    sysd["order"] = 4
    rng = np.random.default_rng(2)
    xi = rng.normal(size=(6, 3, 4))
    x0 = rng.uniform(0, L, (3, 2))
    dtmax, mdps, dt0 = 1e-2, 5e-3, 1e-6
    tv = np.array([dt0, dtmax, mdps, 1.0, -1.0, 40, 40, 6])
    pos, pos1, T, F, st = O.srk2_run(sysd, x0, np.tile([1.0, 0.0], (3, 1)), np.arange(3),
                                     np.ones(3), np.zeros(3), tv, xi=xi)
    dt = min(dtmax, mdps / np.sqrt(2 * D))
    x = x0.copy()
    for kk in range(5):
        x = np.mod(x + np.sqrt(2 * dt * D) * xi[kk, :, :2], L)
    assert np.allclose(pos1, np.mod(x + np.sqrt(2 * dt * D) * xi[5, :, :2], L), atol=1e-14)
    assert np.allclose(pos, pos1, atol=1e-14)  # dt_prev == dt_used after the first step
    # the first step's virtual stages use dt0
    p1, q1, *_ = O.srk2_run(sysd, x0, np.tile([1.0, 0.0], (3, 1)), np.arange(3), np.ones(3),
                            np.zeros(3), np.array([dt0, dtmax, mdps, 1.0, -1.0, 40, 40, 1]),
                            xi=xi[:1])
    assert np.allclose(q1, np.mod(x0 + np.sqrt(2 * dt0 * D) * xi[0, :, :2], L), atol=1e-14)
    assert np.allclose(p1, np.mod(x0 + np.sqrt(2 * dt * D) * xi[0, :, :2], L), atol=1e-14)


def test_srk4_rod_virtual_stages_build_on_each_other():
    """With the rigid-rod projection every SRK4 stage projects from the
    position its forces were evaluated at (SRK.py:364-371), so a rigid rod
    under propulsion moves dt/2 + dt/2 + dt in the virtual stages and then
    dt_used from there: (2 dt_prev + dt_used) v0 / r per step."""
    n, a, v0, dtmax = 6, 0.5, 1.0, 1e-3
    t = np.array([1.0, 0.0])
    pos = straight_rod(n, a, 2, np.array([10.0, 10.0]), t)
    sysd = dict(dim=2, n=n, n_nearby=4, has_springs=True, has_wca=False, has_propel=True,
                L=L2D, D=0.0, Dr=0.0, radius=1.0, wca=np.array([12.0, 12.0, 0.0, 1.0]),
                spring=np.array([100.0, 0.0, 0.0, 0.0, 1.0, 0.1]), v0=v0, rod=1.0, order=4)
    tv = np.array([dtmax, dtmax, 1.0, 1.0, -1.0, 1e9, 40.0, 5.0])
    cid = np.zeros(n, dtype=np.int64)
    out, *_ = O.srk2_run(sysd, pos, np.tile(t, (n, 1)), cid, np.array([a]), np.zeros(1), tv)
    assert np.allclose(out - pos, np.tile(5 * 3 * dtmax * v0 * t, (n, 1)), atol=1e-12)
    sysd["rod"] = 0.0
    out0, *_ = O.srk2_run(sysd, pos, np.tile(t, (n, 1)), cid, np.array([a]), np.zeros(1), tv)
    assert np.allclose(out0 - pos, np.tile(5 * dtmax * v0 * t, (n, 1)), atol=1e-12)


WALL_WCA = np.array([12.0, 12.0, 0.0, 2 ** (1 / 6)])  # eps 1, bead size 1, WCA cutoff


def test_walls_are_the_lj_gradient_of_the_distance_to_the_wall():
    """forces.py:405-435: inside its range a wall pushes with -dU/dx of the pair
    energy at the distance from the wall, which sits `zone` inside each face;
    the far face mirrors the near one."""
    L, zone = 20.0, 0.75
    d = np.array([0.9, 0.95, 1.0, 1.02])  # distances from the wall
    lo = np.column_stack((zone + d, np.full(4, 10.0)))
    hi = np.column_stack((L - zone - d, np.full(4, 10.0)))
    F_lo = O.walls(lo, L, WALL_WCA, 1, zone)
    F_hi = O.walls(hi, L, WALL_WCA, 1, zone)
    h = 1e-6
    want = -(lj_energy(d + h, WALL_WCA) - lj_energy(d - h, WALL_WCA)) / (2 * h)
    assert np.allclose(F_lo[:, 0], want, rtol=1e-6)
    assert np.allclose(F_hi[:, 0], -want, rtol=1e-6)
    assert np.all(F_lo[:, 1] == 0) and np.all(F_hi[:, 1] == 0)  # no wall along y
    both = O.walls(np.array([[zone + 0.95, L - zone - 0.95]]), L, WALL_WCA, 3, zone)
    assert both[0, 0] == pytest.approx(-both[0, 1], rel=1e-15) and both[0, 0] > 0


def test_walls_range_and_shift_quirks():
    """The range test compares the squared distance with wca[3] itself
    (forces.py:432), so a wall stops at sqrt(rc), not rc; WCA_shift is
    subtracted from the distance (the pair law adds it, forces.py:124-127)."""
    L, zone = 20.0, 0.75
    rc = WALL_WCA[3]
    inside, outside = np.sqrt(rc) - 1e-9, np.sqrt(rc) + 1e-9
    assert outside < rc  # within the pair cutoff, yet beyond the wall's range
    F = O.walls(np.array([[zone + inside, 5.0], [zone + outside, 5.0]]), L, WALL_WCA, 1, zone)
    assert F[0, 0] != 0 and F[1, 0] == 0
    shifted = WALL_WCA.copy()
    shifted[2] = 0.1
    F_s = O.walls(np.array([[zone + 1.05, 5.0]]), L, shifted, 1, zone)
    F_0 = O.walls(np.array([[zone + 0.95, 5.0]]), L, WALL_WCA, 1, zone)
    assert F_s[0, 0] == pytest.approx(F_0[0, 0], rel=1e-13)


def test_srk2_walls_use_the_step_start_position():
    """SRK.py:270 hands the walls pos in both stages: a lone bead at a wall
    (T = 0, no other force) follows x' = x + dt_used F(x) / r, with m = |F(x)|/r
    (1/2 (F + F)) setting dt."""
    L, zone, r = 20.0, 0.75, 1.0
    sysd = dict(dim=2, n=1, n_nearby=4, has_springs=False, has_wca=False, has_propel=False,
                L=L, D=0.0, Dr=0.0, radius=r, wca=WALL_WCA, spring=np.zeros(6), v0=0.0,
                walls=1, wall_zone=zone)
    x0 = np.array([[zone + 0.92, 4.0]])
    dtmax, mdps, n = 1e-3, 2e-3, 12
    tv = np.array([1e-6, dtmax, mdps, 1.0, -1.0, 40, 40, n])
    pos, *_ = O.srk2_run(sysd, x0, np.array([[1.0, 0.0]]), np.array([0]), np.ones(1),
                         np.zeros(1), tv, xi=np.zeros((n, 1, 4)))
    x = x0[0, 0]
    for _ in range(n):
        f = O.walls(np.array([[x, 4.0]]), L, WALL_WCA, 1, zone)[0, 0]
        dt = min(dtmax, mdps / (abs(f) / r))
        x = x + dt * f / r
    assert pos[0, 0] == pytest.approx(x, rel=1e-13) and pos[0, 1] == 4.0


def test_multi_population_equals_the_union():
    """Two populations with the same physics whose mutual law is their own
    WCA law move exactly as the one population holding both (to rounding):
    cross forces, the shared dt and the per-population updates agree with
    the single-population loop."""
    rng = np.random.default_rng(60)
    L = 7.0
    pos, tang, chain = straight_chains(rng, 6, 5, L)
    sysd = dict(dim=2, n=30, n_nearby=4, has_springs=True, has_wca=True, has_propel=True,
                L=L, D=0.02, Dr=0.0, radius=1.0,
                wca=O.wca_params(1.0, 1.0, 0.0, 2.5 / 2 ** (1 / 6)),  # LJ to 2.5 sigma
                spring=np.array([100.0, 4.0, 0.0, 0.0, 1.0, 0.1]), v0=1.0)
    rest, vt = np.full(6, 0.5), np.linspace(-0.3, 0.3, 6)
    xi = rng.normal(size=(40, 30, 4))
    tv = np.array([1e-4, 1e-3, 0.0125, 1.0, -1.0, 40, 40, 40])
    whole = O.srk2_run(sysd, pos, tang, chain, rest, vt, tv, xi=xi)
    parts = []
    for lo, hi in ((0, 3), (3, 6)):
        sl = slice(5 * lo, 5 * hi)
        parts.append(dict(sysd=dict(sysd, n=5 * (hi - lo)), pos=pos[sl], tangent=tang[sl],
                          chain_id=chain[sl] - lo, rest_lengths=rest[lo:hi], vtan=vt[lo:hi],
                          xi=np.ascontiguousarray(xi[:, sl])))
    out, st = O.multi_srk2_run(parts, {(0, 1): sysd["wca"], (1, 0): sysd["wca"]}, tv)
    got = np.vstack([o[0] for o in out])
    assert np.max(np.abs(N.min_image(got - whole[0], L))) < 1e-11
    assert st[0] == pytest.approx(whole[4][0], rel=1e-12)
    assert st[1] == pytest.approx(whole[4][1], rel=1e-12)
    # without the cross law the populations do not feel each other
    alone, _ = O.multi_srk2_run(parts, {}, tv)
    assert np.max(np.abs(N.min_image(alone[0][0] - got[:15], L))) > 1e-4


def test_multi_population_fixed_obstacle():
    """A bead pushed off a fixed bead (T = 0) by the averaged law of
    populations.py:226-232 follows the scalar Heun map; the fixed bead stays."""
    r_a, r_b = 1.0, 0.6
    wa, wb = O.wca_params(r_a, 1.0, 0.0, 1.0), O.wca_params(r_b, 1.0, 0.0, 1.0)
    w_ab = 0.5 * (wa + wb)
    mover = dict(sysd=dict(dim=2, n=1, n_nearby=4, has_springs=False, has_wca=True,
                           has_propel=False, L=L2D, D=0.0, Dr=0.0, radius=r_a, wca=wa,
                           spring=np.zeros(6), v0=0.0),
                 pos=np.array([[10.7, 5.0]]), tangent=np.array([[1.0, 0.0]]),
                 chain_id=np.zeros(1, dtype=np.int64), rest_lengths=np.ones(1),
                 vtan=np.zeros(1), xi=np.zeros((20, 1, 4)))
    wall = dict(sysd=dict(mover["sysd"], radius=r_b, wca=wb), pos=np.array([[10.0, 5.0]]),
                tangent=np.array([[1.0, 0.0]]), chain_id=np.zeros(1, dtype=np.int64),
                rest_lengths=np.ones(1), vtan=np.zeros(1), fixed=True)
    dtmax, mdps, n = 1e-3, 2e-3, 20
    tv = np.array([1e-6, dtmax, mdps, 1.0, -1.0, 40, 40, n])
    out, st = O.multi_srk2_run([mover, wall], {(0, 1): w_ab, (1, 0): w_ab}, tv)

    def f(s):  # force on the mover at separation s (the obstacle at -s)
        ds = s * s
        return 0.0 if ds > w_ab[3] ** 2 else -s * ds ** -4 * (w_ab[0] - w_ab[1] * ds ** -3)

    s, dt = 0.7, 1e-6
    for _ in range(n):
        f0 = f(s)
        s1 = s + dt * f0 / r_a
        f1 = f0 + f(s1)
        dt = min(dtmax, mdps / (0.5 * abs(f1) / r_a))
        s = s + dt * 0.5 * f1 / r_a
    assert out[0][0][0, 0] - 10.0 == pytest.approx(s, rel=1e-12)
    assert np.array_equal(out[1][0], wall["pos"])  # the fixed population never moves
    assert st[0] == pytest.approx(dt, rel=1e-12)


def test_multi_population_diffusion_per_population():
    """F = 0: each population diffuses with its own D; one dt for all from the
    largest sqrt(2 D)."""
    rng = np.random.default_rng(61)
    pops, D = [], (0.05, 0.4)
    for k in range(2):
        pops.append(dict(sysd=abp_sys(4, 10.0, D=D[k], v0=0.0), pos=rng.uniform(0, 10, (4, 2)),
                         tangent=np.tile([1.0, 0.0], (4, 1)), chain_id=np.arange(4),
                         rest_lengths=np.ones(4), vtan=np.zeros(4),
                         xi=rng.normal(size=(6, 4, 4))))
    dtmax, mdps = 1e-2, 5e-3
    out, st = O.multi_srk2_run(pops, {}, np.array([1e-6, dtmax, mdps, 1.0, -1.0, 40, 40, 6]))
    dt = min(dtmax, mdps / np.sqrt(2 * max(D)))
    for k in range(2):
        x = pops[k]["pos"].copy()
        for j in range(6):
            x = np.mod(x + np.sqrt(2 * dt * D[k]) * pops[k]["xi"][j, :, :2], 10.0)
        assert np.allclose(out[k][0], x, atol=1e-13)
    assert st[0] == dt


def sliding_sys(n, L=20.0, chunk=None, wa=None, maxdist=1.5, v0=1.0, n_nearby=4):
    return dict(dim=2, n=n, n_nearby=n_nearby, has_springs=False, has_wca=False,
                has_propel=False, L=L, D=0.0, Dr=0.0, radius=1.0,
                wca=O.wca_params(1.0, 1.0, 0.0, 1.0), spring=np.zeros(6), v0=0.0,
                sliding_maxdist=maxdist, sliding_v0=v0, chunk=chunk, where_active=wa)


def test_sliding_pushes_antiparallel_partners_by_thread_chunk():
    """forces.py:114-120: v0 t per antiparallel partner closer than the
    sliding range; in one thread chunk only the bead with the higher index
    gets it (the reference skips b2 >= b1 there), across chunks both do;
    parallel or distant partners give nothing."""
    pos = np.array([[5.0, 5.0], [5.0, 6.2], [12.0, 5.0], [12.0, 6.2], [3.0, 15.0],
                    [3.0, 16.6]])
    t = np.array([[1.0, 0.0], [-1.0, 0.0], [1.0, 0.0], [1.0, 0.0], [0.0, 1.0], [0.0, -1.0]])
    cid = np.arange(6)
    F, _ = O.force_eval(sliding_sys(6), pos, cid, np.ones(6), np.zeros(6), t)
    assert np.allclose(F[0], 0) and np.allclose(F[1], t[1])  # one chunk: the later bead
    assert np.allclose(F[2:], 0)  # parallel pair, and a pair 1.6 apart
    chunk = np.array([0, 1, 0, 1, 0, 1], dtype=np.int32)
    F2, _ = O.force_eval(sliding_sys(6, chunk=chunk), pos, cid, np.ones(6), np.zeros(6), t)
    assert np.allclose(F2[0], t[0]) and np.allclose(F2[1], t[1])
    # two partners push twice; a same-chain neighbour does not count
    pos3 = np.array([[5.0, 5.0], [5.0, 5.8], [5.6, 5.4]])
    t3 = np.array([[-1.0, 0.0], [-1.0, 0.0], [1.0, 0.0]])
    F3, _ = O.force_eval(sliding_sys(3), pos3, np.array([0, 1, 2]), np.ones(3), np.zeros(3), t3)
    assert np.allclose(F3[2], 2 * t3[2]) and np.allclose(F3[:2], 0)
    F4, _ = O.force_eval(sliding_sys(3), pos3, np.array([0, 1, 1]), np.ones(2), np.zeros(2), t3)
    assert np.allclose(F4[2], t3[2])


def test_sliding_pair_motion_and_where_active():
    """Two antiparallel ABPs in different chunks, T = 0: each moves v0 dt_used
    along its tangent per step, and where_active counts the partner in both
    stages of the step (-1 + 2); in one chunk only the second moves."""
    pos = np.array([[5.0, 5.0], [5.0, 6.0]])
    t = np.array([[1.0, 0.0], [-1.0, 0.0]])
    tv = np.array([1e-3, 1e-3, 1.0, 1.0, -1.0, 40, 40, 10])
    wa = np.zeros(2, dtype=np.int32)
    out = O.srk2_run(sliding_sys(2, chunk=np.array([0, 1], dtype=np.int32), wa=wa), pos, t,
                     np.arange(2), np.ones(2), np.zeros(2), tv, xi=np.zeros((10, 2, 4)))
    assert np.allclose(out[0], pos + 10 * 1e-3 * t, atol=1e-13)
    assert wa.tolist() == [1, 1]
    wa1 = np.zeros(2, dtype=np.int32)
    out1 = O.srk2_run(sliding_sys(2, wa=wa1), pos, t, np.arange(2), np.ones(2), np.zeros(2), tv,
                      xi=np.zeros((10, 2, 4)))
    assert np.allclose(out1[0][0], pos[0]) and np.allclose(out1[0][1], pos[1] + 1e-2 * t[1])
    assert wa1.tolist() == [-1, 1]


def test_sliding_over_verlet_list_equals_all_pairs():
    """The CPU baseline's sliding over the Verlet list (where the reference
    computes it) counts the partners the all-pairs loop counts, on chains with
    the LJ 2.5 sigma cutoff (the sliding range lies inside the list radius):
    where_active equal, positions equal to rounding (the pair sums' order
    differs between the two loops)."""
    from gcr_amd import dropin
    from gcr_amd import workloads as W
    w = W.chains("slide_lj", 2, 150, 10, phi=0.35, cutoff_factor=W.LJ_25_SIGMA, kappa=6.0,
                 T=0.01, steps=80, seed=61)
    chunk = np.ascontiguousarray(dropin.chain_chunks(3, w.chain_id)[w.chain_id], dtype=np.int32)
    p = w.phys
    runs = []
    for mode in (0, 1):
        wa = np.zeros(w.n, dtype=np.int32)
        sysd = dict(dim=2, n=w.n, n_nearby=p.n_nearby, has_springs=True, has_wca=True,
                    has_propel=False, L=p.L, D=p.D, Dr=p.Dr, radius=p.radius,
                    wca=np.asarray(p.wca), spring=np.asarray(p.spring), v0=p.v0,
                    sliding_maxdist=1.5, sliding_v0=1.0, chunk=chunk, where_active=wa)
        tv = w.time_vals.copy()
        tv[7] = 80
        out = O.srk2_run(sysd, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle, tv,
                         seed=3, pair_mode=mode, n_threads=4)
        runs.append((out, wa))
    (o0, wa0), (o1, wa1) = runs
    assert np.array_equal(wa0, wa1) and np.sum(wa0 > 0) > 20
    assert np.max(np.abs(N.min_image(o0[0] - o1[0], p.L))) < 1e-11 * p.L


def test_multi_population_cell_lists_equal_all_pairs():
    """orc_multi_srk2_run with pair_mode 1 (each population's own pairs over
    Verlet lists, the others through cell grids: the CPU baseline) against
    pair_mode 0 (every pair): two chain populations acting on each other with
    the LJ 2.5 sigma law, and fixed motors pushing chains. Positions to
    rounding (the pair sums' order differs), where_active and every motor's
    bound bead exactly."""
    from gcr_amd import workloads as W
    w = W.chains("two", 2, 60, 8, L=30.0, cutoff_factor=W.LJ_25_SIGMA, kappa=4.0, T=0.01,
                 steps=50, seed=71)
    p = w.phys
    tv = w.time_vals.copy()
    tv[7] = 50

    def chains_pop(lo, hi, seed, wa=False):
        sl = slice(lo * 8, hi * 8)
        sysd = dict(dim=2, n=(hi - lo) * 8, n_nearby=p.n_nearby, has_springs=True,
                    has_wca=True, has_propel=True, L=p.L, D=p.D, Dr=p.Dr, radius=p.radius,
                    wca=np.asarray(p.wca), spring=np.asarray(p.spring), v0=p.v0)
        if wa:
            sysd["where_active"] = np.zeros(sysd["n"], dtype=np.int32)
        return dict(sysd=sysd, pos=w.pos[sl], tangent=w.tangent[sl],
                    chain_id=np.repeat(np.arange(hi - lo), 8), rest_lengths=w.rest_lengths[lo:hi],
                    vtan=w.v_tan_angle[lo:hi], seed=seed)

    law = np.asarray(p.wca)
    runs = [O.multi_srk2_run([chains_pop(0, 30, 1), chains_pop(30, 60, 2)],
                             {(0, 1): law, (1, 0): law}, tv, pair_mode=m, n_threads=4)
            for m in (0, 1)]
    for k in range(2):
        assert np.max(np.abs(N.min_image(runs[0][0][k][0] - runs[1][0][k][0], p.L))) < 1e-10 * p.L
    rng = np.random.default_rng(72)
    nm = 400
    motors_sysd = dict(dim=2, n=nm, n_nearby=-1, has_springs=False, has_wca=False,
                       has_propel=False, L=p.L, D=0.0, Dr=0.0, radius=1.0,
                       wca=O.wca_params(1.0, 1.0, 0.0, 1.0), spring=np.zeros(6), v0=0.0)
    got = []
    for m in (0, 1):
        ch = chains_pop(0, 60, 3, wa=True)
        mo = dict(sysd=motors_sysd, pos=rng.uniform(0, p.L, (nm, 2)) if m == 0 else got[0][3],
                  tangent=np.tile([1.0, 0.0], (nm, 1)), chain_id=np.arange(nm),
                  rest_lengths=np.ones(nm), vtan=np.zeros(nm), fixed=True, seed=9,
                  motor=dict(max_dist=0.5, imparts=1.0, k_on=0.9, k_off=0.2))
        out, _ = O.multi_srk2_run([ch, mo], {(0, 1): "motor"}, tv, pair_mode=m, n_threads=4)
        got.append((out[0][0], ch["sysd"]["where_active"].copy(), mo["motor_bead"].copy(),
                    mo["pos"]))
    assert np.max(np.abs(N.min_image(got[0][0] - got[1][0], p.L))) < 1e-10 * p.L
    # fixed motors keep their bindings: a bead's count this step is -1 plus
    # the bindings made minus those released
    assert np.array_equal(got[0][1], got[1][1]) and np.sum(got[0][1] != -1) > 5
    assert np.array_equal(got[0][2], got[1][2]) and np.sum(got[0][2] >= 0) > 50


def motor_pops(bead_pos, bead_t, motor_pos, k_on, k_off, D_m=0.0, still=True, imparts=1.0,
               max_dist=0.5, L=L2D):
    nb, nm = len(bead_pos), len(motor_pos)
    beads = dict(sysd=dict(dim=2, n=nb, n_nearby=4, has_springs=False, has_wca=False,
                           has_propel=False, L=L, D=0.0, Dr=0.0, radius=1.0,
                           wca=O.wca_params(1.0, 1.0, 0.0, 1.0), spring=np.zeros(6), v0=0.0,
                           where_active=np.zeros(nb, dtype=np.int32)),
                 pos=bead_pos, tangent=bead_t, chain_id=np.arange(nb), rest_lengths=np.ones(nb),
                 vtan=np.zeros(nb), xi=np.zeros((1, nb, 4)))
    motors = dict(sysd=dict(beads["sysd"], n=nm, D=D_m, where_active=None), pos=motor_pos,
                  tangent=np.tile([1.0, 0.0], (nm, 1)), chain_id=np.arange(nm),
                  rest_lengths=np.ones(nm), vtan=np.zeros(nm), fixed=D_m == 0.0, seed=11,
                  xi=np.ones((1, nm, 4)),
                  motor=dict(max_dist=max_dist, imparts=imparts, k_on=k_on, k_off=k_off,
                             still=still))
    return beads, motors


def test_motors_bind_in_bead_order_and_push_along_the_tangent():
    """forces.py:248-283 with k_on = 1, k_off = 0: a free motor binds to the
    first bead (index order) within its range and pushes it by imparts * t once,
    when it binds (a motor that stays bound gives no push: forces.py:268-269);
    a bead out of range or later in line gets nothing."""
    bp = np.array([[5.0, 5.0], [5.3, 5.0], [9.0, 5.0]])
    bt = np.array([[0.0, 1.0], [1.0, 0.0], [1.0, 0.0]])
    beads, motors = motor_pops(bp, bt, np.array([[5.15, 5.2], [9.0, 5.8]]), 1.0, 0.0,
                               imparts=2.0)
    tv = np.array([1e-3, 1e-3, 1.0, 1.0, -1.0, 40, 40, 1])
    out, st = O.multi_srk2_run([beads, motors], {(0, 1): "motor"}, tv)
    # SRK2: x + dt (F0 + F1) / 2 with F0 = 2 t (binding) and F1 = 0 (stays bound)
    assert np.allclose(out[0][0], bp + 1e-3 * 1.0 * np.array([[0.0, 1.0], [0, 0], [0, 0]]))
    assert motors["motor_bead"].tolist() == [0, -1]  # the far motor (0.8 away) stays free
    assert beads["sysd"]["where_active"].tolist() == [0, -1, -1]  # -1 + one binding
    assert np.array_equal(out[1][0], motors["pos"])  # fixed motors


def test_motors_unbinding_moves_them_on_and_counts_against_the_bead_that_frees_them():
    """k_off = 1: a bound motor seen by the next bead in range comes free (that
    bead's where_active drops, whichever bead it pushed) and, with k_on = 1,
    binds to it: both beads are pushed once per stage."""
    bp = np.array([[5.0, 5.0], [5.3, 5.0]])
    bt = np.array([[0.0, 1.0], [0.0, -1.0]])
    beads, motors = motor_pops(bp, bt, np.array([[5.15, 5.0]]), 1.0, 1.0)
    tv = np.array([1e-3, 1e-3, 1.0, 1.0, -1.0, 40, 40, 1])
    out, _ = O.multi_srk2_run([beads, motors], {(0, 1): "motor"}, tv)
    assert np.allclose(out[0][0], bp + 1e-3 * bt)
    # from -1 each: stage 0, bead 0 binds (0); bead 1 frees it (-2) and binds it
    # (-1); stage 1, bead 0 frees (-1), binds (0); bead 1 frees (-2), binds (-1)
    assert beads["sysd"]["where_active"].tolist() == [0, -1]
    assert motors["motor_bead"].tolist() == [1]


def test_motor_draw_rates_and_still_rule():
    """The binding draw is uniform: with k_on = 0.3 about 30% of free motors
    next to a bead bind in the first stage; bound motors of a diffusing
    population stay put (b_motors_dont_diffuse_while_pushing), free ones move
    by sqrt(2 D dt) xi."""
    # 2500 beads 2 apart, each with one motor 0.3 away (and 1.7 from the next)
    g = np.arange(50) * 2.0 + 1.0
    bp = np.array([[x, y] for y in g for x in g])
    mp = bp + np.array([0.3, 0.0])
    bt = np.tile([0.0, 1.0], (len(bp), 1))
    beads, motors = motor_pops(bp, bt, mp, 0.3, 0.0, D_m=0.5, imparts=1e-3, L=100.0)
    tv = np.array([1e-3, 1e-3, 1.0, 1.0, -1.0, 40, 40, 1])
    out, _ = O.multi_srk2_run([beads, motors], {(0, 1): "motor"}, tv)
    bound = motors["motor_bead"] >= 0
    frac = bound.mean()  # bound in stage 0 or in stage 1: 1 - 0.7^2 = 0.51
    assert abs(frac - 0.51) < 0.04
    moved = np.abs(out[1][0] - mp).max(axis=1)
    assert np.all(moved[bound] == 0)
    free_step = np.sqrt(2 * 0.5 * 1e-3) * 1.0  # xi = 1 in every component
    assert np.allclose(out[1][0][~bound], mp[~bound] + free_step)


def test_rotational_diffusion_closed_form():
    """SRK.py:398-423: t' = normalize(t + pi sqrt(2 Dr dt) (xi_r - (xi_r.t) t))."""
    Dr = 0.2
    sysd = abp_sys(2, 30.0, Dr=Dr, v0=0.0)
    t = np.array([[1.0, 0.0], [0.0, 1.0]])
    xi = np.zeros((1, 2, 4))
    xi[0, :, 2:] = [[0.3, -1.1], [2.0, 0.5]]
    dtmax, mdps = 1e-3, 1.0
    tv = np.array([1e-3, dtmax, mdps, 1.0, -1.0, 40, 40, 1])
    _, _, T, _, st = O.srk2_run(sysd, np.full((2, 2), 5.0), t, np.arange(2), np.ones(2),
                                np.zeros(2), tv, xi=xi)
    g = np.pi * np.sqrt(2 * Dr * st[0])
    e0 = np.array([1.0, g * -1.1])
    e1 = np.array([g * 2.0, 1.0])
    assert np.allclose(T[0], e0 / np.linalg.norm(e0), atol=1e-15)
    assert np.allclose(T[1], e1 / np.linalg.norm(e1), atol=1e-15)


def test_srk2_c_vs_numpy_100_steps():
    rng = np.random.default_rng(1)
    pos, tang, chain = straight_chains(rng, 4, 7, L2D)
    sysd = dict(dim=2, n=28, n_nearby=4, has_springs=True, has_wca=True, has_propel=True,
                L=L2D, D=0.01, Dr=0.0075, radius=1.0, wca=O.wca_params(1.0, 1.0, 0.0, 1.0),
                spring=np.array([100.0, 10.0, 0.0, 0.0, 1.0, 0.1]), v0=1.0)
    xi = rng.normal(size=(100, 28, 4))
    tv = np.array([1e-7, 1e-3, 0.0125, 1.0, -1.0, 40, 40, 100])
    vt = np.full(4, 0.2)
    a = O.srk2_run(sysd, pos, tang, chain, np.full(4, 0.5), vt, tv, xi=xi)
    b = N.srk2_run(sysd, pos, tang, chain, np.full(4, 0.5), vt, tv, xi)
    assert np.max(np.abs(N.min_image(a[0] - b[0], L2D))) < 1e-12
    assert np.max(np.abs(N.min_image(a[1] - b[1], L2D))) < 1e-12
    assert np.allclose(a[2], b[2], atol=1e-12)
    assert np.allclose(a[3], b[3], rtol=1e-10, atol=1e-10)
    assert np.allclose(a[4], b[4], rtol=1e-12)


def test_verlet_mode_matches_brute_force():
    """The CPU-baseline pair path (cell grid + Verlet list, Newton-3 inside
    chain-aligned thread chunks) reproduces the all-pairs path."""
    rng = np.random.default_rng(4)
    pos, tang, chain = straight_chains(rng, 30, 10, 25.0, jitter=0.02)
    sysd = dict(dim=2, n=300, n_nearby=4, has_springs=True, has_wca=True, has_propel=True,
                L=25.0, D=0.01, Dr=0.0075, radius=1.0,
                wca=O.wca_params(1.0, 1.0, 0.0, 2.5 / 2 ** (1 / 6)),
                spring=np.array([100.0, 5.0, 0.0, 0.0, 1.0, 0.1]), v0=1.0)
    xi = rng.normal(size=(60, 300, 4))
    tv = np.array([1e-7, 5e-4, 0.0278, 1.0, -1.0, 40, 40, 60])
    vt = rng.uniform(-0.5, 0.5, 30)
    a = O.srk2_run(sysd, pos, tang, chain, np.full(30, 0.5), vt, tv, xi=xi, pair_mode=0)
    for thr in (1, 3):
        b = O.srk2_run(sysd, pos, tang, chain, np.full(30, 0.5), vt, tv, xi=xi, pair_mode=1,
                       n_threads=thr)
        assert np.max(np.abs(N.min_image(a[0] - b[0], 25.0))) < 1e-11
        assert np.allclose(a[4], b[4], rtol=1e-10)


# ---------------------------------------------------------------------- noise

def test_philox_known_answers():
    """Known-answer vectors published with Random123 (Salmon et al., SC'11)."""
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                            0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                    [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_bead_noise_statistics():
    g = np.array([O.bead_noise(2, 1234, b, s) for b in range(200) for s in range(50)])
    assert abs(g.mean()) < 0.02
    assert abs(g.var() - 1.0) < 0.03
    c = np.corrcoef(g.T)
    assert np.max(np.abs(c - np.eye(4))) < 0.03
    # counter-based: the same key gives the same numbers, another bead does not
    assert np.array_equal(O.bead_noise(3, 9, 5, 77), O.bead_noise(3, 9, 5, 77))
    assert not np.array_equal(O.bead_noise(3, 9, 5, 77), O.bead_noise(3, 9, 6, 77))


# ------------------------------------------------------ rigid-rod projection

def straight_rod(n, a, dim, head, t):
    return np.array([head - a * k * t for k in range(n)])


@pytest.mark.parametrize("dim", [2, 3])
def test_rod_projection_factor_zero_is_plain_update(dim):
    rng = np.random.default_rng(40)
    prev = rng.uniform(0, L2D, (12, dim))
    F = rng.normal(size=(12, dim))
    xi = rng.normal(size=(12, 2 * dim))
    cid = np.repeat([0, 1, 2], 4).astype(np.int64)
    out = O.rod_project(prev, F, xi, cid, np.array([0.5, 0.5, 0.5]), 0.0, L2D, 1e-3, 0.02, 1.3)
    want = np.mod(prev + 1e-3 / 1.3 * F + np.sqrt(2 * 0.02 * 1e-3) * xi[:, :dim], L2D)
    assert np.allclose(out, want, rtol=0, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_rod_projection_uniform_force_translates_rigidly(dim):
    """Any uniform force in 2D (the 2D torque weights by the distance from the
    centre, which sums to zero); in 3D the reference's torque is
    |r| sum_k r_hat x u_k without that weight, so only a force along the rod
    gives a pure translation there."""
    t = np.ones(dim) / np.sqrt(dim)
    prev = straight_rod(7, 0.5, dim, np.full(dim, 10.0), t)
    f = np.array([1.0, 2.0]) if dim == 2 else 1.7 * t
    F = np.tile(f, (7, 1))
    out = O.rod_project(prev, F, np.zeros((7, 2 * dim)), np.zeros(7, dtype=np.int64),
                        np.array([0.5]), 1.0, L2D, 2e-3, 0.0, 1.0)
    assert np.allclose(out - prev, 2e-3 * F, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_rod_projection_straightens_a_bent_chain(dim):
    rng = np.random.default_rng(41)
    t = rng.normal(size=dim)
    t /= np.linalg.norm(t)
    a, n = 0.5, 9
    prev = straight_rod(n, a, dim, np.full(dim, 10.0), t) + rng.normal(0, 0.02, (n, dim))
    out = O.rod_project(prev, np.zeros((n, dim)), np.zeros((n, 2 * dim)),
                        np.zeros(n, dtype=np.int64), np.array([a]), 1.0, L2D, 1e-3, 0.0, 1.0)
    # a straight rod of spacing a through the centre of mass, along head -> tail
    com = prev.mean(axis=0)
    rhat = (prev[-1] - prev[0]) / np.linalg.norm(prev[-1] - prev[0])
    want = com + rhat * a * (np.arange(n) - (n - 1) / 2)[:, None]
    assert np.allclose(out, want, atol=1e-12)


def test_srk2_rod_corrector_starts_from_predictor():
    """SRK.py:364-371 hands the projection the position the forces were
    evaluated at; in the corrector that is pos1, so a rigid rod under a
    uniform force moves (dt_prev + dt_used) v0 / r per step, twice the plain
    SRK2 displacement."""
    n, a, v0, dtmax = 6, 0.5, 1.0, 1e-3
    t = np.array([1.0, 0.0])
    pos = straight_rod(n, a, 2, np.array([10.0, 10.0]), t)
    sysd = dict(dim=2, n=n, n_nearby=4, has_springs=True, has_wca=False, has_propel=True,
                L=L2D, D=0.0, Dr=0.0, radius=1.0, wca=np.array([12.0, 12.0, 0.0, 1.0]),
                spring=np.array([100.0, 0.0, 0.0, 0.0, 1.0, 0.1]), v0=v0, rod=1.0)
    tv = np.array([dtmax, dtmax, 1.0, 1.0, -1.0, 1e9, 40.0, 5.0])
    cid = np.zeros(n, dtype=np.int64)
    out, *_ = O.srk2_run(sysd, pos, np.tile(t, (n, 1)), cid, np.array([a]), np.zeros(1), tv)
    assert np.allclose(out - pos, np.tile(5 * 2 * dtmax * v0 * t, (n, 1)), atol=1e-12)
    sysd["rod"] = 0.0
    out0, *_ = O.srk2_run(sysd, pos, np.tile(t, (n, 1)), cid, np.array([a]), np.zeros(1), tv)
    assert np.allclose(out0 - pos, np.tile(5 * dtmax * v0 * t, (n, 1)), atol=1e-12)
