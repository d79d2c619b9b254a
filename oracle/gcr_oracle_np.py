"""NumPy FP64 restatement of the reference hot path (second, independent oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/ as a checker for the C oracle and
for small GPU parity cases. The product path never imports this module.

The C restatement (gcr_oracle.c) follows the reference loops line by line;
this one is written vectorised, from the formulas, so the two share no code.
Accumulation order differs from the reference (np.add.at per term), so the
two oracles agree to rounding, not bit for bit.
"""
import numpy as np


def wca_params(repel_d_max, eps, shift, cutoff_factor):
    """populations.py:48-58."""
    s6 = repel_d_max ** 6
    fac = 12.0 * s6 * eps
    return np.array([fac, fac * s6, shift, cutoff_factor * repel_d_max - shift])


def wca_coeff(ds, p):
    """forces.py:13-15: force on b1 is coeff * (x_b2 - x_b1)."""
    return ds ** -4.0 * (p[0] - p[1] * ds ** -3.0)


def min_image(d, L):
    """One-shot correction into [-L/2, L/2) (forces.py:44-54, 103-107)."""
    d = np.where(d < -L / 2, d + L, d)
    return np.where(d >= L / 2, d - L, d)


def pair_forces(pos, chain_id, L, p, n_nearby):
    """All-pairs restatement of forces.py:57-144 over the neighbour set of
    updates.py:586-669 (exclusion rule updates.py:645-650)."""
    n, dim = pos.shape
    d = min_image(pos[None, :, :] - pos[:, None, :], L)  # d[b1, b2] = x2 - x1
    ds = np.einsum("ijk,ijk->ij", d, d)
    idx = np.arange(n)
    same = chain_id[:, None] == chain_id[None, :]
    if n_nearby < 0:
        excl = same
    else:
        excl = same & (np.abs(idx[:, None] - idx[None, :]) <= n_nearby)
    mask = (ds > 0) & (ds <= p[3] ** 2) & ~excl
    if p[2] != 0:
        mag = np.sqrt(np.where(mask, ds, 1.0))
        nm = np.maximum(0.0, mag + p[2])
        d = d * (nm / mag)[:, :, None]
        ds = nm * nm
    c = np.zeros_like(ds)
    c[mask] = wca_coeff(ds[mask], p)
    return np.einsum("ij,ijk->ik", c, d)


def chain_bonds(chain_id):
    """Indices b with b and b+1 on the same chain (consecutive beads)."""
    b = np.arange(len(chain_id) - 1)
    return b[chain_id[:-1] == chain_id[1:]]


def chain_forces(pos, chain_id, L, sp, rest_lengths):
    """Vectorised restatement of forces.py:285-402 (Hookean, or FENE when
    sp[3] == 1, forces.py:331-341). Returns (forces, tangents); tangents are
    NaN where not defined."""
    n, dim = pos.shape
    F = np.zeros_like(pos)
    T = np.full_like(pos, np.nan)
    bonds = chain_bonds(chain_id)
    t = min_image(pos[bonds] - pos[bonds + 1], L)  # bond b -> b+1, pointing to b
    ds = np.sum(t * t, axis=1)
    ok = ds != 0
    inv = np.zeros_like(ds)
    inv[ok] = ds[ok] ** -0.5
    a = rest_lengths[chain_id[bonds]]
    if sp[3] != 0:
        m = sp[5] * a
        st = np.sqrt(ds) - a
        with np.errstate(divide="ignore", invalid="ignore"):
            den = np.where((st * st != m) & (m != 0), np.abs(st * st / np.where(m != 0, m * m, 1.0) - 1.0), 0.0)
            mag = np.where(ok, -sp[0] * inv * st / den, 0.0)
    else:
        mag = np.where(ok, sp[0] * (a * inv - 1.0), 0.0)
    f = t * mag[:, None]
    np.add.at(F, bonds, f)
    np.add.at(F, bonds + 1, -f)
    u = np.where(ok[:, None], t * inv[:, None], t)  # unit bond vectors
    bond_of = -np.ones(n, dtype=np.int64)
    bond_of[bonds] = np.arange(len(bonds))
    # interior beads m: bond (m, m+1) and bond (m-1, m) both exist
    m = np.arange(n)
    has_next = bond_of >= 0
    has_prev = np.zeros(n, dtype=bool)
    has_prev[1:] = has_next[:-1]
    interior = m[has_next & has_prev]
    ti = u[bond_of[interior]]
    tj = u[bond_of[interior - 1]]
    dot = np.sum(ti * tj, axis=1)
    ok2 = (dot > -1.0) & (dot < 1.0)
    if sp[1] != 0:
        im, ti2, tj2, d2 = interior[ok2], ti[ok2], tj[ok2], dot[ok2]
        phi = np.arccos(d2)
        if sp[2] != 0:
            phi = phi + sp[2] * np.sign(-ti2[:, 0] * tj2[:, 1] + ti2[:, 1] * tj2[:, 0])
        phi = phi * sp[1] / np.sqrt(1.0 - d2 * d2)
        ff = phi[:, None] * (d2[:, None] * tj2 - ti2)
        fb = phi[:, None] * (tj2 - d2[:, None] * ti2)
        np.add.at(F, im, ff + fb)
        np.add.at(F, im - 1, -ff)
        np.add.at(F, im + 1, -fb)
    if sp[4] != 0:
        heads = m[has_next & ~has_prev]
        tails = m[has_prev & ~has_next]
        T[heads] = u[bond_of[heads]]
        T[tails] = u[bond_of[tails - 1]]
        c = np.where(ok2, (2.0 * (1.0 + dot)) ** -0.5, 0.0)
        T[interior] = np.where(ok2[:, None], c[:, None] * (ti + tj), ti)
    return F, T


def self_propel(tangent, chain_id, v0, vtan_per_chain):
    """forces.py:206-223 with the angle looked up per chain."""
    n, dim = tangent.shape
    ang = np.asarray(vtan_per_chain)[chain_id] if vtan_per_chain is not None else np.zeros(n)
    if dim != 2:
        return v0 * tangent
    cs, sn = np.cos(ang), np.sin(ang)
    out = np.empty_like(tangent)
    out[:, 0] = v0 * (cs * tangent[:, 0] - sn * tangent[:, 1])
    out[:, 1] = v0 * (sn * tangent[:, 0] + cs * tangent[:, 1])
    return out


def force_eval(sysd, pos, chain_id, rest_lengths, vtan, tangent):
    """SRK.py:228-253 order: springs (+ tangents) -> pair -> propulsion.
    Returns (forces, tangent used for propulsion)."""
    F = np.zeros_like(pos)
    tang = tangent.copy()
    if sysd["has_springs"]:
        Fs, T = chain_forces(pos, chain_id, sysd["L"], sysd["spring"], rest_lengths)
        F += Fs
        if sysd["spring"][4] != 0:
            tang = np.where(np.isnan(T), tang, T)
    if sysd["has_wca"]:
        F += pair_forces(pos, chain_id, sysd["L"], sysd["wca"], sysd["n_nearby"])
    if sysd["has_propel"]:
        F += self_propel(tang, chain_id, sysd["v0"], vtan)
    return F, tang


def srk2_run(sysd, pos, tangent, chain_id, rest_lengths, vtan, time_vals, xi):
    """SRK2 loop of SRK.py:128-423 with injected noise xi[k, n, 2*dim].
    Returns (pos, pos1, tangent, forces, status[dt, elapsed, steps, m])."""
    pos = pos.copy()
    tangent = tangent.copy()
    n, dim = pos.shape
    dt, dtmax, mdps, cheat, mdu, _, _, max_steps = time_vals
    L, D, Dr, r = sysd["L"], sysd["D"], sysd["Dr"], sysd["radius"]
    orient = sysd["has_springs"] and sysd["spring"][4] != 0
    steps, elapsed, m = 0, 0.0, 0.0
    pos1 = pos.copy()
    F = np.zeros_like(pos)
    while (steps < max_steps or max_steps < 0) and (elapsed < mdu or mdu < 0):
        z = xi[steps]
        F0, tang = force_eval(sysd, pos, chain_id, rest_lengths, vtan, tangent)
        s2 = np.sqrt(2 * dt * D)
        pos1 = np.mod(pos + (dt / r) * F0 + (s2 * z[:, :dim] if s2 != 0 else 0.0), L)
        F1, tang1 = force_eval(sysd, pos1, chain_id, rest_lengths, vtan, tangent)
        F = F0 + F1
        m = np.sqrt(2 * D) + 0.5 * np.sqrt(np.max(np.sum(F * F, axis=1))) / r
        steps += 1
        dt = dtmax if m == 0 else min(dtmax, mdps / m)
        dt_used = cheat * dt
        if mdu > 0:
            dt_used = min(dt_used, mdu - elapsed)
        elapsed += dt_used
        s2 = np.sqrt(2 * dt_used * D)
        pos = np.mod(pos + (dt_used * 0.5 / r) * F + (s2 * z[:, :dim] if s2 != 0 else 0.0), L)
        if orient:
            tangent = tang1
        if Dr > 0 and not orient:
            sr = np.sqrt(2 * dt_used * Dr) * np.pi
            nr = z[:, dim:].copy()
            proj = np.sum(nr * tangent, axis=1)
            nr -= proj[:, None] * tangent
            t = tangent + sr * nr
            tangent = t / np.sqrt(np.sum(t * t, axis=1))[:, None]
    return pos, pos1, tangent, F, np.array([dt, elapsed, steps, m])
