"""Seeded synthetic gauge fields and boundary twists (host, FP64).

Links are U = exp(i sigma H) with H = sum_a h_a T_a, h_a ~ N(0, 1) i.i.d.,
T_a = lambda_a / 2 the SU(3) generators (Tr T_a T_b = delta_ab / 2), drawn
per link from numpy's PCG64 generator seeded with `seed`.  The exponential
is taken through the eigen-decomposition of the Hermitian H, so each link is
unitary with determinant exp(i sigma Tr H) = 1 to rounding.

This plays the role of the reference's set_g_rand_color_matrix_field
(qcd-utils.cpp:71-92, whose generator is make_g_rand_anti_hermitian_matrix,
matrix-hmc.h:125-142); bit-equality with qlat's RNG stream is not a goal,
since every consumer (GPU path and oracle) reads the same stored field.
"""

import numpy as np


def gell_mann_generators():
    """T_a = lambda_a / 2, a = 1..8."""
    lam = np.zeros((8, 3, 3), dtype=np.complex128)
    lam[0][0, 1] = lam[0][1, 0] = 1
    lam[1][0, 1], lam[1][1, 0] = -1j, 1j
    lam[2][0, 0], lam[2][1, 1] = 1, -1
    lam[3][0, 2] = lam[3][2, 0] = 1
    lam[4][0, 2], lam[4][2, 0] = -1j, 1j
    lam[5][1, 2] = lam[5][2, 1] = 1
    lam[6][1, 2], lam[6][2, 1] = -1j, 1j
    lam[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    return lam / 2


def rand_gauge_links(total_site, sigma, seed, chunk=1 << 20):
    """(V, 4, 3, 3) complex128 links U = exp(i sigma H)."""
    V = int(np.prod(total_site))
    n = V * 4
    rng = np.random.Generator(np.random.PCG64(seed))
    T = gell_mann_generators()
    out = np.empty((n, 3, 3), dtype=np.complex128)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        h = rng.standard_normal((b - a, 8))
        H = np.einsum("na,aij->nij", h, T)
        w, v = np.linalg.eigh(H)
        ph = np.exp(1j * sigma * w)
        out[a:b] = np.einsum("nij,nj,nkj->nik", v, ph, v.conj())
    return out.reshape(V, 4, 3, 3)


def twist_boundary_at_boundary(gauge, total_site, lmom=-0.5, mu=3):
    """Multiply U_mu(x) with x_mu = L_mu - 1 by exp(2 pi i lmom)
    (qcd.h:331-346); lmom = -0.5 turns periodic into antiperiodic."""
    X, Y, Z, T = total_site
    g = gauge.reshape(T, Z, Y, X, 4, 3, 3)
    ph = np.exp(2j * np.pi * lmom)
    if mu == 0:
        g[:, :, :, X - 1, 0] *= ph
    elif mu == 1:
        g[:, :, Y - 1, :, 1] *= ph
    elif mu == 2:
        g[:, Z - 1, :, :, 2] *= ph
    else:
        g[T - 1, :, :, :, 3] *= ph
    return gauge


def plaquette(gauge, total_site):
    """Average Re Tr(U_mu U_nu U_mu^+ U_nu^+)/3 over sites and planes."""
    X, Y, Z, T = total_site
    g = gauge.reshape(T, Z, Y, X, 4, 3, 3)
    ax = {0: 3, 1: 2, 2: 1, 3: 0}
    tot = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            Umu = g[..., mu, :, :]
            Unu = g[..., nu, :, :]
            Unu_xmu = np.roll(Unu, -1, axis=ax[mu])
            Umu_xnu = np.roll(Umu, -1, axis=ax[nu])
            p = Umu @ Unu_xmu @ np.conj(np.swapaxes(Umu_xnu, -1, -2)) @ np.conj(np.swapaxes(Unu, -1, -2))
            tot += np.trace(p, axis1=-2, axis2=-1).real.mean() / 3
    return tot / 6
