"""CPU checks that pin the oracle (no GPU).

The oracle (oracle/dwf_oracle.cpp) is pinned by
  * the reference's own golden values (examples-py/free-invert.log.json),
  * the analytic infinite-Ls free-field propagator (qcd-prop.h:292-359),
    restated independently in numpy (oracle/freefield.py),
  * an independent numpy statement of the 5D operator (oracle/dense.py),
  * the identities printed by examples-cpp/dslash-tests/main.cpp,
  * Clifford algebra of the gamma matrices.
"""

import json
import os

import numpy as np
import pytest

from oracle import dense, freefield
from oracle import oracle as O
from tests.helpers import ZMOBIUS_OMEGA, point_src, rand_field, rand_gauge, rel, unit_gauge

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "free_invert.json")))


def qlat_check(v, ref, eps):
    """check_log_json metric (qlat_utils/utils.py:455-545)."""
    return 2 * abs(v - ref) / (abs(v) + abs(ref)) <= eps


# ------------------------------------------------------------ goldens
@pytest.fixture(scope="module")
def free_solutions():
    site = (4, 4, 4, 8)
    V = 512
    U = unit_gauge(V)
    src = point_src(V)
    inv = O.OracleInverter(site, U, mass=0.05, ls=16, m5=1.0)
    sol1, _ = inv.invert_prop(src)
    invz = O.OracleInverter(site, U, mass=0.05, ls=12, m5=1.0, omega=ZMOBIUS_OMEGA)
    solz, _ = invz.invert_prop(src)
    with np.errstate(divide="ignore", invalid="ignore"):
        sol = freefield.free_invert(src, site, 0.05, 1.0)
    return sol, sol1, solz


def test_golden_analytic_free_field(free_solutions):
    sol, _, _ = free_solutions
    v, eps = GOLD["point_source"]["sol_qnorm_analytic"]
    assert qlat_check(np.vdot(sol, sol).real, v, eps)


def test_golden_dwf_cg_point_source(free_solutions):
    sol, sol1, _ = free_solutions
    v, eps = GOLD["point_source"]["sol1_qnorm_dwf_cg"]
    assert qlat_check(np.vdot(sol1, sol1).real, v, eps)
    d = sol1 - sol
    v, eps = GOLD["point_source"]["sol_diff_qnorm"]
    assert qlat_check(np.vdot(d, d).real, v, eps)
    assert np.vdot(d, d).real < 1e-7  # free-invert.py:40


def test_golden_zmobius_point_source(free_solutions):
    sol, _, solz = free_solutions
    v, eps = GOLD["point_source"]["sol_z_qnorm_zmobius_cg"]
    assert qlat_check(np.vdot(solz, solz).real, v, eps)
    d = solz - sol
    v, eps = GOLD["point_source"]["sol_z_diff_qnorm"]
    assert qlat_check(np.vdot(d, d).real, v, eps)


def test_antiperiodic_free_field_matches_twisted_momenta():
    """Antiperiodic T by U_t(T-1) *= exp(2 pi i lmom), lmom = -1/2
    (qcd.h:331-346), is gauge-equivalent to the uniform twist U_t =
    exp(2 pi i lmom / T) that free_invert's momtwist describes: for a source
    at t = 0, S_boundary(x) = exp(2 pi i lmom t / T) S_uniform(x)."""
    site = (4, 4, 4, 8)
    V = 512
    lmom = -0.5
    U = unit_gauge(V)
    from qlattice_amd.gauge import twist_boundary_at_boundary

    twist_boundary_at_boundary(U, site, lmom, 3)
    src = point_src(V)
    inv = O.OracleInverter(site, U, mass=0.05, ls=16, m5=1.0)
    sol1, _ = inv.invert_prop(src)
    ana = freefield.free_invert(src, site, 0.05, 1.0, momtwist=(0, 0, 0, lmom))
    t = np.arange(V) // (4 * 4 * 4)
    ana = ana * np.exp(2j * np.pi * lmom * t / 8)[:, None, None]
    assert rel(sol1, ana) < 1e-5
    per = freefield.free_invert(src, site, 0.05, 1.0)
    assert rel(sol1, per) > 1e-2  # the boundary condition does matter


# ----------------------------------------------------- gamma algebra
def test_cps_gamma_clifford_algebra():
    g = freefield.cps_gammas()
    g5 = freefield.gamma5()
    eye = np.eye(4)
    for mu in range(4):
        assert np.allclose(g[mu], g[mu].conj().T)
        for nu in range(4):
            ac = g[mu] @ g[nu] + g[nu] @ g[mu]
            assert np.allclose(ac, 2 * eye * (mu == nu))
        assert np.allclose(g[mu] @ g5 + g5 @ g[mu], 0)
    # gamma5 = gx gy gz gt of the standard basis (matrix.h:448-452)
    std = [-g[0], g[1], -g[2], g[3]]
    assert np.allclose(std[0] @ std[1] @ std[2] @ std[3], g5)


# ------------------------------------------------ independent operator
SITE = (4, 4, 4, 4)
LS = 4


@pytest.fixture(scope="module")
def small():
    V = int(np.prod(SITE))
    U = rand_gauge(SITE, 0.5, 11)
    inv = O.OracleInverter(SITE, U, mass=0.07, ls=LS, m5=1.6, mobius_scale=2.0)
    shamir = O.OracleInverter(SITE, U, mass=0.07, ls=LS, m5=1.6, mobius_scale=1.0)
    invz = O.OracleInverter(SITE, U, mass=0.07, ls=12, m5=1.6, omega=ZMOBIUS_OMEGA)
    return dict(V=V, U=U, inv=inv, shamir=shamir, invz=invz)


def test_multiply_m_matches_numpy_operator(small):
    V, U, inv = small["V"], small["U"], small["inv"]
    psi = rand_field((V, LS, 12), 1)
    ref = dense.to_flat(dense.apply_m(dense.to_lattice(psi, SITE, LS), dense.gauge_lattice(U, SITE),
                                      inv.bs, inv.cs, inv.mass, inv.m5))
    for op in ("m", "m_full", "m_with_prec_sym2"):
        assert rel(inv.apply_full(op, psi), ref) < 1e-13, op


def test_zmobius_multiply_m_matches_numpy_operator(small):
    V, U, invz = small["V"], small["U"], small["invz"]
    psi = rand_field((V, 12, 12), 2)
    ref = dense.to_flat(dense.apply_m(dense.to_lattice(psi, SITE, 12), dense.gauge_lattice(U, SITE),
                                      invz.bs, invz.cs, invz.mass, invz.m5))
    assert rel(invz.apply_full("m", psi), ref) < 1e-13
    assert rel(invz.apply_full("m_full", psi), ref) < 1e-13


def test_shamir_m_dwf(small):
    V, sh = small["V"], small["shamir"]
    psi = rand_field((V, LS, 12), 3)
    assert rel(sh.apply_full("m_dwf", psi), sh.apply_full("m", psi)) < 1e-13
    with pytest.raises(RuntimeError):
        small["inv"].apply_full("m_dwf", psi)  # Moebius: dslash.cpp:342 asserts


def test_d_minus_matches_numpy(small):
    V, U, inv = small["V"], small["U"], small["inv"]
    psi = rand_field((V, LS, 12), 4)
    ref = dense.to_flat(dense.apply_d_minus(dense.to_lattice(psi, SITE, LS), dense.gauge_lattice(U, SITE),
                                            inv.cs, inv.m5))
    assert rel(inv.apply_full("d_minus", psi), ref) < 1e-13


def test_mdag_is_adjoint(small):
    V, inv = small["V"], small["inv"]
    x = rand_field((V, LS, 12), 5)
    y = rand_field((V, LS, 12), 6)
    lhs = np.vdot(x, inv.apply_full("m", y))
    rhs = np.vdot(inv.apply_full("mdag", x), y)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_half_ops_against_numpy(small):
    """m_e_e and its inverse equal the explicit per-chirality Ls x Ls
    matrices; m_e_o is the off-diagonal block of the numpy operator."""
    V, U, inv = small["V"], small["U"], small["inv"]
    Mp, Mm = dense.ls_matrices(LS, inv.mass, inv.m5, inv.bs, inv.cs)
    h = rand_field((V // 2, LS, 12), 7)
    assert rel(inv.apply_half("m_e_e", h), dense.apply_ls_matrices(h, Mp, Mm)) < 1e-13
    assert rel(inv.apply_half("m_e_e_inv", h),
               dense.apply_ls_matrices(h, np.linalg.inv(Mp), np.linalg.inv(Mm))) < 1e-12
    assert rel(inv.apply_half("mdag_e_e", h),
               dense.apply_ls_matrices(h, Mp.conj().T, Mm.conj().T)) < 1e-13
    assert rel(inv.apply_half("mdag_e_e_inv", h),
               dense.apply_ls_matrices(h, np.linalg.inv(Mp).conj().T, np.linalg.inv(Mm).conj().T)) < 1e-12
    # M_eo: odd input embedded in a full field, take the even output half
    full = np.zeros((V, LS, 12), dtype=np.complex128)
    dense.set_half(full, h, SITE, 1)
    Mfull = dense.to_flat(dense.apply_m(dense.to_lattice(full, SITE, LS), dense.gauge_lattice(U, SITE),
                                        inv.bs, inv.cs, inv.mass, inv.m5))
    assert rel(inv.apply_half("m_e_o", h, eo_in=1), dense.get_half(Mfull, SITE, 2)) < 1e-13


def test_dslash_tests_identities(small):
    """examples-cpp/dslash-tests/main.cpp:184-215."""
    V, inv = small["V"], small["invz"]
    x = rand_field((V // 2, 12, 12), 8)
    y = rand_field((V // 2, 12, 12), 9)
    assert rel(inv.apply_half("m_e_e_inv", inv.apply_half("m_e_e", x)), x) < 1e-12
    assert rel(inv.apply_half("mdag_e_e_inv", inv.apply_half("mdag_e_e", x)), x) < 1e-12
    for a, b in (("m_e_e", "mdag_e_e"), ("m_e_e_inv", "mdag_e_e_inv")):
        lhs = np.vdot(x, inv.apply_half(a, y))
        rhs = np.vdot(inv.apply_half(b, x), y)
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    for a, b in (("hop", "hop_dag"), ("m_e_o", "mdag_e_o")):
        lhs = np.vdot(x, inv.apply_half(a, y, eo_in=1))  # x even, y odd
        rhs = np.vdot(inv.apply_half(b, x, eo_in=2), y)
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    lhs = np.vdot(x, inv.apply_half("mpc_sym2", y, 1))
    rhs = np.vdot(inv.apply_half("mpcdag_sym2", x, 1), y)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    a = np.vdot(x, inv.apply_half("hermop_sym2", x, 1))
    assert a.real > 0 and abs(a.imag) < 1e-12 * a.real


def test_invert_satisfies_m_x_equals_dminus_b(small):
    """dslash-tests/main.cpp:233-242: M invert(b) = D_- b."""
    V, inv = small["V"], small["inv"]
    inv.set_params(1e-12, 200, 300)
    b = rand_field((V, LS, 12), 10)
    x, it = inv.invert_5d(b)
    r = inv.apply_full("m", x) - inv.apply_full("d_minus", b)
    assert rel(inv.apply_full("m", x), inv.apply_full("d_minus", b)) < 1e-9
    assert it > 0 and np.isfinite(r).all()


def test_cycles_restart_when_max_num_iter_small(small):
    """Non-convergence inside a cycle returns max_num_iter+1 and triggers a
    true-residual restart (dslash.h:394-398, :506-514)."""
    V, inv = small["V"], small["inv"]
    b = rand_field((V, LS, 12), 12)
    inv.set_params(1e-10, 5, 300)
    x, it = inv.invert_5d(b)
    st = inv.stats()
    inv.set_params(1e-10, 200, 300)
    x2, it2 = inv.invert_5d(b)
    assert st["cycles"] > 1
    assert rel(x, x2) < 1e-8


def test_zero_source_returns_zero(small):
    V, inv = small["V"], small["inv"]
    x, it = inv.invert_5d(np.zeros((V, LS, 12), dtype=np.complex128))
    assert it == 0 and not np.any(x)


def test_invert_4d_is_chiral_embedding(small):
    V, inv = small["V"], small["inv"]
    inv.set_params(1e-10, 200, 300)
    b4 = rand_field((V, 12), 13)
    x4, _ = inv.invert_4d(b4)
    b5 = np.zeros((V, LS, 12), dtype=np.complex128)
    b5[:, 0, :6] = b4[:, :6]
    b5[:, LS - 1, 6:] = b4[:, 6:]
    x5, _ = inv.invert_5d(b5)
    ref = np.concatenate([x5[:, LS - 1, :6], x5[:, 0, 6:]], axis=1)
    assert rel(x4, ref) < 1e-13


def test_power_iteration_bounds_hermop(small):
    V, inv = small["V"], small["inv"]
    start = rand_field((V // 2, LS, 12), 14)
    lam = inv.find_max_eigen_value_hermop_sym2(start, 60)
    x = rand_field((V // 2, LS, 12), 15)
    ax = inv.apply_half("hermop_sym2", x, 1)
    rayleigh = np.vdot(x, ax).real / np.vdot(x, x).real
    assert 0 < rayleigh <= lam * (1 + 1e-6)


def test_invert_with_guess_pins(small):
    """invert_with_cg_with_guess (dslash.h:532-573): a zero guess is
    invert(5d) itself (ratio 1, the same CG sequence); from a rough guess the
    corrected solution still satisfies M x = D_- b (dslash-tests:233-242);
    from the exact solution of a tighter solve the correction reaches the
    tolerance in fewer iterations than from zero."""
    V, inv = small["V"], small["inv"]
    inv.set_params(1e-10, 200, 300)
    b = rand_field((V, LS, 12), 16)
    x0, it0 = inv.invert_5d(b)
    xg, itg = inv.invert_5d_with_guess(b, np.zeros_like(b))
    assert itg == it0 and np.array_equal(xg, x0)
    rough = x0 + 1e-3 * rand_field(b.shape, 17)
    xr, itr = inv.invert_5d_with_guess(b, rough)
    db = inv.apply_full("d_minus", b)
    assert rel(inv.apply_full("m", xr), db) < 1e-8
    assert rel(inv.apply_full("m", rough), db) > 1e-5  # the guess alone is far from it
    assert 0 < itr < it0
    inv.set_params(1e-8, 200, 300)


def test_five_d_residual_of_invert(small):
    """The is_checking_invert quantity (dslash.h:522-528) of a solve to
    stop_rsd lies within a condition factor of stop_rsd: the CG stops on the
    preconditioned odd system, so |M x - D_- b| / |D_- b| is of the order of
    stop_rsd, not exactly it."""
    V, inv = small["V"], small["inv"]
    b = rand_field((V, LS, 12), 18)
    for stop in (1e-6, 1e-10):
        inv.set_params(stop, 200, 300)
        x, _ = inv.invert_5d(b)
        db = inv.apply_full("d_minus", b)
        r = rel(inv.apply_full("m", x), db)
        assert 1e-3 * stop < r < 1e3 * stop, (stop, r)
    inv.set_params(1e-8, 200, 300)
