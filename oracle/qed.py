"""TEST INFRASTRUCTURE ONLY -- numpy restatement of the QED domain-wall
operator and its solvers (qlat/qlat/lib/qed.cpp:75-411):
multiply_m_dwf_qed (:263-411), cg_with_m_dwf_qed (:178-261), invert_dwf_qed
(:119-150) and the 4D <-> 5D embeddings (:75-117).

Fields are in the reference layout: a 5D field is (V, Ls, 4) complex (site
index x + X*(y + Y*(z + Z*t))), the U(1) links gf1 are (V, 4) and the Wick
phase factors one complex number per time slice.  Gamma matrices come from
freefield.py (built from Pauli matrices, independently of the HIP code).
"""

import numpy as np

from .freefield import cps_gammas, gamma5


def multiply_m_dwf_qed(f, gf1, mass, m5, ls, total_site, wick=None, dagger=False):
    """qed.cpp:263-411 (periodic in every direction; out may be in)."""
    X, Y, Z, T = total_site
    g = cps_gammas()
    g5 = gamma5()
    lat = np.asarray(f, dtype=np.complex128).reshape(T, Z, Y, X, ls, 4)
    U = np.asarray(gf1, dtype=np.complex128).reshape(T, Z, Y, X, 4)
    t = np.arange(T)
    if wick is None:
        tw = np.ones(T, np.complex128)
        twm = np.ones(T, np.complex128)
    else:
        wick = np.asarray(wick, dtype=np.complex128)
        tw, twm = wick[t], wick[(t - 1) % T]
    if dagger:
        tw, twm = np.conj(tw), np.conj(twm)
    sh = (T, 1, 1, 1, 1, 1)
    avg = (0.5 * tw + 0.5 * twm).reshape(sh)
    inv_avg = (0.5 / tw + 0.5 / twm).reshape(sh)
    tw6, twm6 = tw.reshape(sh), twm.reshape(sh)
    df = -1.0 if dagger else 1.0

    def gm(m, v):  # m applied to the spin index of v
        return v @ m.T

    out = ((4.0 - m5) * avg + inv_avg) * lat
    # fifth dimension (qed.cpp:342-371)
    out[..., :ls - 1, :] += -0.5 * avg * lat[..., 1:, :] + df * 0.5 * avg * gm(g5, lat[..., 1:, :])
    out[..., 1:, :] += -0.5 * avg * lat[..., :ls - 1, :] - df * 0.5 * avg * gm(g5, lat[..., :ls - 1, :])
    am = avg[..., 0, :]
    out[..., ls - 1, :] += 0.5 * mass * am * lat[..., 0, :] - df * 0.5 * mass * am * gm(g5, lat[..., 0, :])
    out[..., 0, :] += 0.5 * mass * am * lat[..., ls - 1, :] + df * 0.5 * mass * am * gm(g5, lat[..., ls - 1, :])
    # hopping (qed.cpp:373-408)
    for mu in range(4):
        ax = 3 - mu
        iv_p = np.roll(lat, -1, axis=ax)
        iv_m = np.roll(lat, 1, axis=ax)
        u_p = U[..., mu]
        u_m = 1.0 / np.roll(U[..., mu], 1, axis=ax)
        if dagger:
            u_p = np.conj(1.0 / u_p)
            u_m = np.conj(1.0 / u_m)
        u_p = u_p[..., None, None]
        u_m = u_m[..., None, None]
        if mu == 3:
            tf_1 = tfm_1 = 1.0
            tf_2, tfm_2 = 1.0 / tw6, 1.0 / twm6
        else:
            tf_1 = tfm_1 = tf_2 = tfm_2 = avg
        out += -(0.5 * tf_2) * u_p * iv_p - (0.5 * tfm_2) * u_m * iv_m
        out += df * 0.5 * u_p * tf_1 * gm(g[mu], iv_p) - df * 0.5 * u_m * tfm_1 * gm(g[mu], iv_m)
    return out.reshape(-1, ls, 4)


def cg_with_m_dwf_qed(src, gf1, mass, m5, ls, total_site, wick=None, dagger=False, stop_rsd=1e-8,
                      max_num_iter=50000):
    """qed.cpp:178-261 from a zero guess: M_{!d} M_{d} x = src."""

    def a(v):
        return multiply_m_dwf_qed(multiply_m_dwf_qed(v, gf1, mass, m5, ls, total_site, wick, dagger),
                                  gf1, mass, m5, ls, total_site, wick, not dagger)

    x = np.zeros_like(src)
    r = src.copy()
    p = r.copy()
    qnorm_in = np.vdot(src, src).real
    qnorm_r = np.vdot(r, r).real
    for it in range(1, max_num_iter + 1):
        ap = a(p)
        alpha = qnorm_r / np.vdot(p, ap).real
        x = x + alpha * p
        r = r - alpha * ap
        new_qnorm_r = np.vdot(r, r).real
        if new_qnorm_r <= qnorm_in * stop_rsd ** 2:
            return x, it
        p = r + (new_qnorm_r / qnorm_r) * p
        qnorm_r = new_qnorm_r
    return x, max_num_iter + 1


def fermion_field_5d_from_4d_qed(ff4d, ls, upper, lower):
    """qed.cpp:97-117."""
    f = np.zeros((ff4d.shape[0], ls, 4), np.complex128)
    f[:, upper, :2] = ff4d[:, :2]
    f[:, lower, 2:] = ff4d[:, 2:]
    return f


def fermion_field_4d_from_5d_qed(ff5d, upper, lower):
    """qed.cpp:75-95."""
    f = np.zeros((ff5d.shape[0], 4), np.complex128)
    f[:, :2] = ff5d[:, upper, :2]
    f[:, 2:] = ff5d[:, lower, 2:]
    return f


def invert_dwf_qed(src4d, gf1, mass, m5, ls, total_site, wick=None, dagger=False, stop_rsd=1e-8,
                   max_num_iter=50000):
    """qed.cpp:119-150: M out = in (M^+ out = in when dagger)."""
    src5d = fermion_field_5d_from_4d_qed(src4d, ls, 0, ls - 1)
    src5d = multiply_m_dwf_qed(src5d, gf1, mass, m5, ls, total_site, wick, not dagger)
    sol5d, it = cg_with_m_dwf_qed(src5d, gf1, mass, m5, ls, total_site, wick, dagger, stop_rsd, max_num_iter)
    return fermion_field_4d_from_5d_qed(sol5d, ls - 1, 0), it
