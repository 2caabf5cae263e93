"""TEST INFRASTRUCTURE ONLY -- analytic free-field domain-wall propagator.

numpy restatement of ``free_mom_invert`` / ``free_invert``
(qlat/qlat/include/qlat/qcd-prop.h:292-359): the infinite-Ls domain-wall
propagator on a unit gauge field, applied in momentum space.  The forward
transform is sum_x f[x] exp(-i k x) (field-fft.h:18), i.e. numpy.fft.fftn;
the backward one is unnormalised and followed by 1/V, i.e. numpy.fft.ifftn.

Gamma matrices here are built independently of the C++ oracle, from Pauli
matrices in the chiral basis: gamma_k = [[0, -i sigma_k], [i sigma_k, 0]],
gamma_t = [[0, 1], [1, 0]], gamma5 = diag(1, 1, -1, -1), and CPS gammas are
(-gamma_x, gamma_y, -gamma_z, gamma_t)  (matrix.h:426-452).
"""

import numpy as np


def pauli():
    s1 = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    s2 = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
    s3 = np.array([[1, 0], [0, -1]], dtype=np.complex128)
    return s1, s2, s3


def cps_gammas():
    z = np.zeros((2, 2), dtype=np.complex128)
    one = np.eye(2, dtype=np.complex128)
    g = []
    for s in pauli():
        g.append(np.block([[z, -1j * s], [1j * s, z]]))
    g.append(np.block([[z, one], [one, z]]))
    cps = [-g[0], g[1], -g[2], g[3]]
    return cps


def gamma5():
    return np.diag([1, 1, -1, -1]).astype(np.complex128)


def free_mom_invert_factor(total_site, mass, m5=1.0, momtwist=(0, 0, 0, 0)):
    """Per-momentum 4x4 spin matrix of qcd-prop.h:292-345 (the colour part is
    the identity).  Returns an array (T, Z, Y, X, 4, 4)."""
    X, Y, Z, T = total_site
    g = cps_gammas()
    grids = np.meshgrid(np.arange(T), np.arange(Z), np.arange(Y), np.arange(X),
                        indexing="ij")
    kg = [grids[3], grids[2], grids[1], grids[0]]  # x, y, z, t
    p2 = np.zeros(grids[0].shape)
    wp = np.full(grids[0].shape, 1.0 - m5)
    pg = np.zeros(grids[0].shape + (4, 4), dtype=np.complex128)
    for i in range(4):
        L = total_site[i]
        k = kg[i].astype(np.float64)
        k = np.where(k >= (L + 1) // 2, k - L, k)  # smod (only sin/cos used)
        kk = 2.0 * np.pi * (k + momtwist[i]) / L
        ks = np.sin(kk)
        pg += ks[..., None, None] * g[i]
        p2 += ks ** 2
        wp += 2.0 * np.sin(kk / 2.0) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        # wp = 0 only at the momentum the zero test below removes (M5 = 1)
        calpha = (1.0 + wp ** 2 + p2) / 2.0 / wp
        alpha = np.arccosh(calpha)
        lwa = 1.0 - wp * np.exp(-alpha)
        eye = np.eye(4, dtype=np.complex128)
        ipgm = -1j * pg + (mass * lwa)[..., None, None] * eye
        ipgm *= (lwa / (p2 + (mass * lwa) ** 2))[..., None, None]
    zero = (p2 < 1e-10) & (lwa < 1e-10)
    ipgm[zero] = 0.0
    return ipgm


def free_invert(prop_src, total_site, mass, m5=1.0, momtwist=(0, 0, 0, 0)):
    """prop_src: (V, 12, 12) complex128 in reference layout.  Returns the
    infinite-Ls free domain-wall solution (qcd-prop.h:347-359)."""
    X, Y, Z, T = total_site
    src = np.asarray(prop_src, dtype=np.complex128).reshape(T, Z, Y, X, 12, 12)
    f = np.fft.fftn(src, axes=(0, 1, 2, 3))
    fac = free_mom_invert_factor(total_site, mass, m5, momtwist)
    f = f.reshape(T, Z, Y, X, 4, 3, 12)
    f = np.einsum("...ab,...bcj->...acj", fac, f)
    f = f.reshape(T, Z, Y, X, 12, 12)
    sol = np.fft.ifftn(f, axes=(0, 1, 2, 3))
    return sol.reshape(X * Y * Z * T, 12, 12)
