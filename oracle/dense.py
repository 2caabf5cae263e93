"""TEST INFRASTRUCTURE ONLY -- an independent numpy statement of the 5D
domain-wall operator, used to cross-check the C++ oracle.

It is written from the operator's mathematical definition rather than from
the reference's loop structure:

  M = D_W(-M5) (b + c S) + (1 - S)                (dslash.cpp:490-540)
  D_- = 1 + D_W(-M5) (-c)                         (dslash.cpp:448-480)
  D_W(m) psi(x) = (4 + m) psi(x)
      - sum_mu [ U_mu(x) (1 - g_mu)/2 psi(x + mu)
               + U_mu(x - mu)^dag (1 + g_mu)/2 psi(x - mu) ]   (dslash.cpp:400-446)
  (S psi)_s = P_- psi_{s+1} + P_+ psi_{s-1},  psi_{Ls} = -m psi_0,
              psi_{-1} = -m psi_{Ls-1}, P_+- = (1 +- g5)/2     (dslash.cpp:617-635)

Fields here have shape (T, Z, Y, X, Ls, 4, 3); the helpers convert from and
to the reference's flat (V, Ls, 12) layout.  Gamma matrices come from
oracle/freefield.py (built from Pauli matrices, not from the C++ table).
"""

import numpy as np

from .freefield import cps_gammas, gamma5

AXIS = {0: 3, 1: 2, 2: 1, 3: 0}  # mu -> array axis of (T, Z, Y, X, ...)


def to_lattice(flat, total_site, ls):
    X, Y, Z, T = total_site
    return np.asarray(flat).reshape(T, Z, Y, X, ls, 4, 3)


def to_flat(lat):
    T, Z, Y, X, ls = lat.shape[:5]
    return lat.reshape(T * Z * Y * X, ls, 12)


def gauge_lattice(gauge, total_site):
    X, Y, Z, T = total_site
    return np.asarray(gauge).reshape(T, Z, Y, X, 4, 3, 3)


def wilson_d(psi, U, mass):
    g = cps_gammas()
    eye = np.eye(4, dtype=np.complex128)
    out = (4.0 + mass) * psi
    for mu in range(4):
        ax = AXIS[mu]
        Umu = U[..., mu, :, :]
        fwd = np.roll(psi, -1, axis=ax)
        Pm = 0.5 * (eye - g[mu])
        t = np.einsum("ab,...sbc->...sac", Pm, fwd)
        out = out - np.einsum("...ij,...saj->...sai", Umu, t)
        bwd = np.roll(psi, 1, axis=ax)
        Ub = np.roll(Umu, 1, axis=ax)
        Pp = 0.5 * (eye + g[mu])
        t = np.einsum("ab,...sbc->...sac", Pp, bwd)
        out = out - np.einsum("...ji,...saj->...sai", Ub.conj(), t)
    return out


def shift_s(psi, mass):
    """(S psi)_s of the 5D operator on axis -3 (the s axis)."""
    g5 = np.diag(gamma5()).real
    up = (g5 > 0)[:, None]   # P_+ : spins 0, 1
    lo = (g5 < 0)[:, None]   # P_- : spins 2, 3
    nxt = np.roll(psi, -1, axis=-3).copy()
    nxt[..., -1, :, :] *= -mass
    prv = np.roll(psi, 1, axis=-3).copy()
    prv[..., 0, :, :] *= -mass
    return np.where(lo, nxt, 0) + np.where(up, prv, 0)


def s_scale(psi, coef):
    return psi * np.asarray(coef)[:, None, None]


def apply_m(psi, U, bs, cs, mass, m5):
    Spsi = shift_s(psi, mass)
    pre = s_scale(psi, bs) + s_scale(Spsi, cs)
    return wilson_d(pre, U, -m5) + psi - Spsi


def apply_d_minus(psi, U, cs, m5):
    return psi + wilson_d(s_scale(psi, -np.asarray(cs)), U, -m5)


def dense_matrix(fn, shape):
    n = int(np.prod(shape))
    cols = []
    for k in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[k] = 1.0
        cols.append(fn(e.reshape(shape)).reshape(n))
    return np.stack(cols, axis=1)


def ls_matrices(ls, mass, m5, bs, cs):
    """Per-chirality Ls x Ls matrices of M_ee = diag(b~) - diag(c~) S with
    b~ = 1 + b(4-M5), c~ = 1 - c(4-M5) (dslash.cpp:617-621).  Returns
    (Mplus, Mminus): Mplus acts on spins 0,1 (S: s <- s-1), Mminus on
    spins 2,3 (S: s <- s+1)."""
    bt = 1.0 + np.asarray(bs) * (4.0 - m5)
    ct = 1.0 - np.asarray(cs) * (4.0 - m5)
    Sp = np.zeros((ls, ls), dtype=np.complex128)
    Sm = np.zeros((ls, ls), dtype=np.complex128)
    for s in range(ls):
        if s >= 1:
            Sp[s, s - 1] = 1.0
        else:
            Sp[0, ls - 1] = -mass
        if s <= ls - 2:
            Sm[s, s + 1] = 1.0
        else:
            Sm[ls - 1, 0] = -mass
    D = np.diag(bt)
    C = np.diag(ct)
    return D - C @ Sp, D - C @ Sm


def apply_ls_matrices(half, Mp, Mm):
    """Apply per-chirality s-matrices to a (Vh, Ls, 12) field."""
    out = np.empty_like(half)
    out[:, :, :6] = np.einsum("st,vtk->vsk", Mp, half[:, :, :6])
    out[:, :, 6:] = np.einsum("st,vtk->vsk", Mm, half[:, :, 6:])
    return out


def parity_of_sites(total_site):
    X, Y, Z, T = total_site
    t, z, y, x = np.meshgrid(np.arange(T), np.arange(Z), np.arange(Y),
                             np.arange(X), indexing="ij")
    s = (x + y + z + t).reshape(-1)
    return 2 - s % 2


def half_sites(total_site, eo):
    """Full site indices of the half lattice with parity eo (2 even, 1 odd),
    ordered by half index = full index // 2 (core.h:232-261)."""
    par = parity_of_sites(total_site)
    idx = np.nonzero(par == eo)[0]
    assert np.all(idx // 2 == np.arange(len(idx)))
    return idx


def get_half(full, total_site, eo):
    return full[half_sites(total_site, eo)]


def set_half(full, half, total_site, eo):
    full[half_sites(total_site, eo)] = half
