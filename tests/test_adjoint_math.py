"""CPU check of the transposed collision used by the adjoint kernel.

The kernel applies mu = J^T lam with the closed form
  mu_p = (1-s) [(1-w) lam_p + w (A (1 + 1.5 u^2) - 4.5 D + e_p . (3 B + 9 C - 3 A u))]
(A = sum w_q lam_q, B = sum w_q lam_q e_q, C = sum w_q lam_q (e_q.u) e_q,
D = sum w_q lam_q (e_q.u)^2).  Here J is built column by column by central differences of
the BGK + sponge collision (05:906-972: feq = w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2)),
so a dropped term, a wrong sign or a transposed index in the closed form fails."""
from __future__ import annotations

import numpy as np

E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [1, -1, 0], [-1, 1, 0], [-1, -1, 0], [1, 0, 1], [1, 0, -1], [-1, 0, 1],
              [-1, 0, -1], [0, 1, 1], [0, 1, -1], [0, -1, 1], [0, -1, -1]], float)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def collide(f, omega, sigma):
    rho = f.sum()
    u = (f[:, None] * E).sum(0) / rho
    eu = E @ u
    feq = W * rho * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (u @ u))
    return (1 - sigma) * (f - omega * (f - feq)) + sigma * W


def closed_form(f, lam, omega, sigma):
    rho = f.sum()
    u = (f[:, None] * E).sum(0) / rho
    eu = E @ u
    wl = W * lam
    A = wl.sum()
    B = (wl[:, None] * E).sum(0)
    Cv = ((wl * eu)[:, None] * E).sum(0)
    D = (wl * eu * eu).sum()
    v = 3 * B + 9 * Cv - 3 * A * u
    return (1 - sigma) * ((1 - omega) * lam + omega * (A * (1 + 1.5 * (u @ u)) - 4.5 * D + E @ v))


def test_transposed_collision_matches_numerical_jacobian():
    rng = np.random.default_rng(0)
    for trial in range(6):
        f = W * (1.0 + 0.05 * rng.standard_normal()) + 0.004 * rng.standard_normal(19)
        lam = rng.standard_normal(19)
        omega, sigma = 1.0 / rng.uniform(0.51, 1.2), rng.uniform(0.0, 0.3)
        J = np.empty((19, 19))
        h = 1e-6
        for p in range(19):
            d = np.zeros(19)
            d[p] = h
            J[:, p] = (collide(f + d, omega, sigma) - collide(f - d, omega, sigma)) / (2 * h)
        want = J.T @ lam
        got = closed_form(f, lam, omega, sigma)
        assert np.allclose(got, want, rtol=1e-7, atol=1e-9), np.abs(got - want).max()


def test_lattice_tables():
    inv = [int(np.where((E == -E[q]).all(1))[0][0]) for q in range(19)]
    assert inv == [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15]
    assert abs(W.sum() - 1) < 1e-15


# ---- NSCBC ghost fill (05:684-835) and its transpose (k_adjoint_ghost) ----------------------

CS = 1.0 / np.sqrt(3.0)


def ghost(f, axis, plus):
    """Forward NSCBC ghost of one face cell from its interior neighbour's populations f:
    ghost = feq(rho_B, u_B) + f - feq(rho_I, u_I), incoming characteristic zeroed."""
    rho = f.sum()
    u = (f[:, None] * E).sum(0) / rho
    p = (rho - 1.0) / 3.0
    if plus:
        wp, wm = u[axis] + p / CS, 0.0
    else:
        wp, wm = 0.0, u[axis] - p / CS
    ub = u.copy()
    ub[axis] = 0.5 * (wp + wm)
    rb = 1.0 + 0.5 * CS * (wp - wm) * 3.0

    def feq(r, v):
        eu = E @ v
        return W * r * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (v @ v))
    return feq(rb, ub) + f - feq(rho, u)


def ghost_transpose(f, g, axis, plus):
    """The kernel's closed form: lam_I += g + d_rho + e . d_j (see k_adjoint_ghost)."""
    rho = f.sum()
    u = (f[:, None] * E).sum(0) / rho
    sg = 1.0 if plus else -1.0
    w = u[axis] + sg * CS * (rho - 1.0)
    rb = 1.0 + sg * w / (2 * CS)
    ub = u.copy()
    ub[axis] = w / 2

    def partials(r, v):
        eu = E @ v
        dr = W * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (v @ v))
        du = (W * r)[:, None] * (3 * E + 9 * eu[:, None] * E - 3 * v[None, :])
        return dr, du
    drb, dub = partials(rb, ub)
    dri, dui = partials(rho, u)
    aB, vB = g @ drb, g @ dub
    aI, vI = g @ dri, g @ dui
    d_rho = aB * 0.5 + vB[axis] * sg * CS * 0.5 - aI
    d_u = vB - vI
    d_u[axis] = aB * sg / (2 * CS) + vB[axis] * 0.5 - vI[axis]
    d_m0 = d_rho - (d_u @ u) / rho
    d_j = d_u / rho
    return g + d_m0 + E @ d_j


def test_transposed_nscbc_ghost_matches_numerical_jacobian():
    """Every face orientation, with only the slots interior cells pull carrying adjoint; a
    swapped characteristic, a lost identity term or a wrong chain through u = j / rho fails."""
    rng = np.random.default_rng(4)
    for axis in range(3):
        for plus in (False, True):
            pulled = (E[:, axis] == (-1 if plus else 1))     # slots entering the interior
            for trial in range(3):
                f = W * (1.0 + 0.05 * rng.standard_normal()) + 0.004 * rng.standard_normal(19)
                g = rng.standard_normal(19) * pulled
                J = np.empty((19, 19))
                h = 1e-6
                for p in range(19):
                    d = np.zeros(19)
                    d[p] = h
                    J[:, p] = (ghost(f + d, axis, plus) - ghost(f - d, axis, plus)) / (2 * h)
                want = J.T @ g
                got = ghost_transpose(f, g, axis, plus)
                assert np.allclose(got, want, rtol=1e-7, atol=1e-9), (axis, plus, np.abs(got - want).max())
