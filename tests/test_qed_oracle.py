"""CPU pins of the QED restatement (oracle/qed.py): with unitary U(1) links
u and no Wick phases it is the Shamir domain-wall operator with SU(3) links
u * 1 restricted to one colour (the C++ oracle's multiply_m / multiply_mdag),
its dagger is the exact adjoint also with complex Wick factors and
non-unitary links, and its CG solves the normal equations."""

import numpy as np

from oracle import oracle as O
from oracle import qed as Q

SITE = (4, 4, 2, 6)
LS = 6
V = int(np.prod(SITE))


def fields(seed):
    rng = np.random.default_rng(seed)
    u = np.exp(1j * rng.uniform(0, 2 * np.pi, (V, 4)))
    x = rng.standard_normal((V, LS, 4)) + 1j * rng.standard_normal((V, LS, 4))
    return rng, u, x


def test_qed_is_su3_shamir_on_one_colour():
    rng, u, x = fields(1)
    orc = O.OracleInverter(SITE, u[..., None, None] * np.eye(3)[None, None], mass=0.1, ls=LS, m5=1.8,
                           mobius_scale=1.0)
    x12 = np.zeros((V, LS, 12), np.complex128)
    x12[..., 0::3] = x
    for name, dag in (("m", False), ("mdag", True)):
        ref = orc.apply_full(name, x12)
        got = Q.multiply_m_dwf_qed(x, u, 0.1, 1.8, LS, SITE, None, dag)
        assert np.linalg.norm(ref[..., 0::3] - got) < 1e-13 * np.linalg.norm(got)
        assert not ref[..., 1::3].any()


def test_qed_dagger_is_adjoint_with_wick_phases_and_nonunitary_links():
    rng, u, x = fields(2)
    y = rng.standard_normal((V, LS, 4)) + 1j * rng.standard_normal((V, LS, 4))
    g = u * (1 + 0.2 * rng.standard_normal((V, 4)))
    wick = 1.1 * np.exp(1j * rng.uniform(-0.3, 0.3, SITE[3]))
    a = np.vdot(x, Q.multiply_m_dwf_qed(y, g, 0.05, 1.0, LS, SITE, wick, False))
    b = np.vdot(Q.multiply_m_dwf_qed(x, g, 0.05, 1.0, LS, SITE, wick, True), y)
    assert abs(a - b) < 1e-13 * abs(a)


def test_qed_cg_solves_normal_equations():
    rng, u, x = fields(3)
    sol, it = Q.cg_with_m_dwf_qed(x, u, 0.1, 1.0, LS, SITE, None, False, 1e-10, 2000)
    mm = Q.multiply_m_dwf_qed(Q.multiply_m_dwf_qed(sol, u, 0.1, 1.0, LS, SITE), u, 0.1, 1.0, LS, SITE, None, True)
    assert it <= 2000 and np.linalg.norm(mm - x) <= 1e-10 * np.linalg.norm(x) * 1.01
