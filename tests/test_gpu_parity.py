"""GPU parity: the HIP library against the CPU oracle on the same inputs.

Bar: FP64 operators within 1e-12 relative of the oracle (they differ only
in rounding: the device uses dense Ls x Ls s-matrices and half-spinor
projection where the reference loops over full spinors), FP64 solutions
within 1e-10 relative of the oracle's, the reference's free-field goldens at
their own eps, and the mixed-precision solver at its stated tolerance.
"""

import json
import os

import numpy as np
import pytest

import qlattice_amd as q
from oracle import oracle as O
from tests.helpers import ZMOBIUS_OMEGA, point_src, qlat_check, rand_field, rand_gauge, rel, unit_gauge

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "free_invert.json")))


def make_pair(site, U, *, mass, ls, m5, mobius_scale=1.0, omega=None):
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=mass, ls=ls, m5=m5, mobius_scale=mobius_scale, omega=omega)
    inv = q.InverterDomainWall(gf=gf, fa=fa)
    orc = O.OracleInverter(site, U, mass=mass, ls=ls, m5=m5, mobius_scale=mobius_scale, omega=omega)
    return geo, inv, orc


def half(geo, ls, eo, data):
    f = q.FermionField5d(geo, ls, eo)
    f.data = np.ascontiguousarray(data)
    return f


def full(geo, ls, data):
    f = q.FermionField5d(geo, ls, 0)
    f.data = np.ascontiguousarray(data)
    return f


SITE = (4, 4, 4, 8)
LS = 8


@pytest.fixture(scope="module")
def mobius():
    print("\n[gpu] mobius 4^3x8 Ls8 operators", flush=True)
    U = rand_gauge(SITE, 0.4, 21)
    return make_pair(SITE, U, mass=0.1, ls=LS, m5=1.8, mobius_scale=2.0)


@pytest.fixture(scope="module")
def zmobius():
    U = rand_gauge(SITE, 0.4, 22)
    return make_pair(SITE, U, mass=0.05, ls=12, m5=1.0, omega=ZMOBIUS_OMEGA)


HALF_OPS = [
    ("m_e_e", q.multiply_m_e_e, False),
    ("mdag_e_e", q.multiply_mdag_e_e, False),
    ("m_e_e_inv", q.multiply_m_e_e_inv, False),
    ("mdag_e_e_inv", q.multiply_mdag_e_e_inv, False),
    ("hop", q.multiply_wilson_d_e_o_no_comm, True),
    ("hop_dag", q.multiply_wilson_ddag_e_o_no_comm, True),
    ("m_e_o", q.multiply_m_e_o, True),
    ("mdag_e_o", q.multiply_mdag_e_o, True),
    ("mpc_sym2", q.multiply_mpc_sym2, True),
    ("mpcdag_sym2", q.multiply_mpcdag_sym2, True),
    ("hermop_sym2", q.multiply_hermop_sym2, True),
]


@pytest.mark.parametrize("name,fn,uses_eo", HALF_OPS, ids=[o[0] for o in HALF_OPS])
@pytest.mark.parametrize("eo", [1, 2])
def test_half_operator_matches_oracle(mobius, name, fn, uses_eo, eo):
    geo, inv, orc = mobius
    x = rand_field((geo.total_volume() // 2, LS, 12), 100 + eo)
    got = fn(None, half(geo, LS, eo, x), inv)
    ref = orc.apply_half(name, x, eo)
    assert rel(got.data, ref) < 1e-12
    if name in ("hop", "hop_dag", "m_e_o", "mdag_e_o"):
        assert got.eo == 3 - eo


@pytest.mark.parametrize("name,fn,uses_eo", HALF_OPS, ids=[o[0] for o in HALF_OPS])
def test_zmobius_half_operator_matches_oracle(zmobius, name, fn, uses_eo):
    geo, inv, orc = zmobius
    x = rand_field((geo.total_volume() // 2, 12, 12), 200)
    got = fn(None, half(geo, 12, 1, x), inv)
    assert rel(got.data, orc.apply_half(name, x, 1)) < 1e-12


@pytest.mark.parametrize("name,fn", [("m", q.multiply_m), ("mdag", q.multiply_mdag),
                                     ("d_minus", q.multiply_d_minus)])
def test_full_operator_matches_oracle(mobius, name, fn):
    geo, inv, orc = mobius
    x = rand_field((geo.total_volume(), LS, 12), 300)
    got = fn(None, full(geo, LS, x), inv)
    assert rel(got.data, orc.apply_full(name, x)) < 1e-12


def test_out_may_alias_in(mobius):
    geo, inv, orc = mobius
    x = rand_field((geo.total_volume() // 2, LS, 12), 301)
    f = half(geo, LS, 1, x.copy())
    q.multiply_hermop_sym2(f, f, inv)
    assert rel(f.data, orc.apply_half("hermop_sym2", x, 1)) < 1e-12


def test_dot_product_and_qnorm(mobius):
    geo, inv, _ = mobius
    a = half(geo, LS, 1, rand_field((geo.total_volume() // 2, LS, 12), 302))
    b = half(geo, LS, 1, rand_field((geo.total_volume() // 2, LS, 12), 303))
    d = q.dot_product(a, b, inv)
    assert abs(d - np.vdot(a.data, b.data)) < 1e-12 * abs(d)
    assert abs(q.qnorm(a, inv) - a.qnorm()) < 1e-12 * a.qnorm()


def test_cg_with_herm_sym_2_matches_oracle(mobius):
    geo, inv, orc = mobius
    b = rand_field((geo.total_volume() // 2, LS, 12), 304)
    g = 0.01 * rand_field(b.shape, 305)
    sol = half(geo, LS, 1, g.copy())
    it = q.cg_with_herm_sym_2(sol, half(geo, LS, 1, b), inv, stop_rsd=1e-10, max_num_iter=500)
    ref, it_ref = orc.cg_with_herm_sym_2(b, g, stop_rsd=1e-10, max_num_iter=500)
    assert it == it_ref
    assert rel(sol.data, ref) < 1e-10


def test_invert_5d_with_restart_cycles_matches_oracle(mobius):
    """Tiny max_num_iter forces the outer restart loop (dslash.h:495-514)."""
    geo, inv, orc = mobius
    b = rand_field((geo.total_volume(), LS, 12), 306)
    inv.set_stop_rsd(1e-10)
    inv.set_max_num_iter(7)
    orc.set_params(1e-10, 7, 300)
    try:
        x = q.FermionField5d(geo, LS, 0)
        it = q.invert(x, full(geo, LS, b), inv)
        ref, it_ref = orc.invert_5d(b)
        st, st_ref = inv.stats(), orc.stats()
    finally:
        inv.set_stop_rsd(1e-8)
        inv.set_max_num_iter(200)
        orc.set_params(1e-8, 200, 300)
    assert st["cycles"] > 1 and st["cycles"] == st_ref["cycles"]
    assert it == it_ref
    assert rel(x.data, ref) < 1e-10


def test_zero_source_gives_zero_solution(mobius):
    geo, inv, _ = mobius
    x = q.FermionField5d(geo, LS, 0)
    it = q.invert(x, q.FermionField5d(geo, LS, 0), inv)
    assert it == 0 and not np.any(x.data)


def test_power_iteration_matches_oracle(mobius):
    geo, inv, orc = mobius
    s = rand_field((geo.total_volume() // 2, LS, 12), 307)
    lam = q.find_max_eigen_value_hermop_sym2(inv, half(geo, LS, 1, s), 30)
    lam_ref = orc.find_max_eigen_value_hermop_sym2(s, 30)
    assert abs(lam - lam_ref) < 1e-10 * lam_ref


def test_invalid_arguments_raise(mobius):
    geo, inv, _ = mobius
    x = rand_field((geo.total_volume() // 2, LS, 12), 308)
    with pytest.raises(q.QlatDwfError, match="parity"):
        q._lib.load().multiply_hermop_sym2(inv.cdata, q._lib.ptr(x), q._lib.ptr(x), 3)
    with pytest.raises(q.QlatDwfError):
        geo_bad = q.Geometry((4, 4, 4, 4))
        gf = q.GaugeField(geo_bad).set_unit()
        fa = q.FermionAction(mass=0.1, ls=40, m5=1.8)
        q.InverterDomainWall(gf=gf, fa=fa)


# ------------------------------------------------- reference goldens
@pytest.fixture(scope="module")
def free_field():
    print("\n[gpu] free-field goldens 4^3x8", flush=True)
    site = (4, 4, 4, 8)
    geo = q.Geometry(site)
    gf = q.GaugeField(geo).set_unit()
    return geo, gf


def test_golden_free_field_dwf_propagator(free_field):
    geo, gf = free_field
    fa = q.FermionAction(mass=0.05, ls=16, m5=1.0)
    inv = q.InverterDomainWall(gf=gf, fa=fa)
    src = q.mk_point_src(geo, q.Coordinate([0, 0, 0, 0]))
    sol1 = inv * src
    v, eps = GOLD["point_source"]["sol1_qnorm_dwf_cg"]
    assert qlat_check(sol1.qnorm(), v, eps), sol1.qnorm()


def test_golden_free_field_zmobius_propagator(free_field):
    geo, gf = free_field
    fa = q.FermionAction(mass=0.05, ls=12, m5=1.0, omega=ZMOBIUS_OMEGA)
    inv = q.InverterDomainWall(gf=gf, fa=fa)
    sol = inv * [q.mk_point_src(geo, q.Coordinate([0, 0, 0, 0]))]
    v, eps = GOLD["point_source"]["sol_z_qnorm_zmobius_cg"]
    assert qlat_check(sol[0].qnorm(), v, eps), sol[0].qnorm()


# ------------------------------------------------ config A (FP64 solve)
def test_config_a_fp64_solve_matches_oracle():
    """8^3x16, sigma 0.4, Shamir Ls 8, M5 1.8, mass 0.1, Gaussian 4D source,
    stop 1e-10: solution within 1e-10 relative of the oracle's."""
    print("\n[gpu] config A 8^3x16 Ls8 FP64 solve", flush=True)
    site = (8, 8, 8, 16)
    U = rand_gauge(site, 0.4, 1)
    geo, inv, orc = make_pair(site, U, mass=0.1, ls=8, m5=1.8)
    inv.set_stop_rsd(1e-10)
    orc.set_params(1e-10, 200, 300)
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 2)
    sol = inv * src
    ref, it_ref = orc.invert_4d(src.data)
    assert inv.stats()["total_iter"] == it_ref
    assert rel(sol.data, ref) < 1e-10


@pytest.mark.parametrize("three", ["0", "1"])
def test_mixed_precision_solver_reaches_tolerance(monkeypatch, three):
    """FP32 inner CG with FP64 reliable updates (fused single-RHS kernels, or
    the three-phase kernels): FP64 true residual of the odd system <=
    stop_rsd, and the solution is as close to the exact one as the FP64
    solver's at the same stop_rsd."""
    monkeypatch.setenv("QLAT_DWF_THREE_PHASE", three)
    print("\n[gpu] mixed solver 8^3x16 Moebius Ls8", flush=True)
    site = (8, 8, 8, 16)
    U = rand_gauge(site, 0.3, 3)
    geo, inv, orc = make_pair(site, U, mass=0.05, ls=8, m5=1.8, mobius_scale=2.0)
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 4)
    orc.set_params(1e-12, 200, 300)
    exact, _ = orc.invert_4d(src.data)
    inv.set_stop_rsd(1e-8)
    x64 = inv * src
    e64 = rel(x64.data, exact)
    inv.set_solver("mixed")
    xmx = inv * src
    st = inv.stats()
    assert st["true_residual"] <= 1e-8
    assert st["hermop_low"] > 0 and st["reliable_updates"] > 0
    assert rel(xmx.data, exact) < 10 * e64 + 1e-9


def test_tiled_device_order_matches_oracle(monkeypatch):
    """The device site order with y and z tiles (QLAT_DWF_TILE_Y,
    QLAT_DWF_TILE_Z) is a layout choice only: every operator and the solve
    must be unchanged."""
    site = (4, 8, 4, 8)
    U = rand_gauge(site, 0.4, 31)
    for by, bz in (("2", "2"), ("4", "1"), ("8", "2")):
        monkeypatch.setenv("QLAT_DWF_TILE_Y", by)
        monkeypatch.setenv("QLAT_DWF_TILE_Z", bz)
        for spt in ("1", "4", "-4", "-104"):
            monkeypatch.setenv("QLAT_DWF_HOP_SPT32", spt)
            monkeypatch.setenv("QLAT_DWF_HOP_SPT64", spt if spt != "4" else "2")
            geo, inv, orc = make_pair(site, U, mass=0.1, ls=6, m5=1.8, mobius_scale=2.0)
            x = rand_field((geo.total_volume() // 2, 6, 12), 400)
            for name, fn, _ in HALF_OPS:
                got = fn(None, half(geo, 6, 2, x), inv)
                assert rel(got.data, orc.apply_half(name, x, 2)) < 1e-12, (by, bz, spt, name)
            y = rand_field((geo.total_volume(), 6, 12), 401)
            assert rel(q.multiply_m(None, full(geo, 6, y), inv).data, orc.apply_full("m", y)) < 1e-12
            inv.set_solver("mixed")
            src = q.FermionField4d(geo)
            src.data = rand_field((geo.total_volume(), 12), 402)
            sol = inv * src
            ref, _ = orc.invert_4d(src.data)
            assert inv.stats()["true_residual"] <= 1e-8
            assert rel(sol.data, ref) < 1e-6, (by, bz, spt)


@pytest.mark.parametrize("fused", ["1", "0"])
@pytest.mark.parametrize("ls,omega", [(8, None), (12, "z"), (16, None)])
def test_fp32_hermop_matches_oracle(monkeypatch, fused, ls, omega):
    """The FP32 operator of the mixed solver (fused hop + s-matrix kernel and
    the unfused kernels) agrees with the FP64 oracle to FP32 accuracy."""
    monkeypatch.setenv("QLAT_DWF_FUSED", fused)
    site = (4, 4, 4, 8)
    U = rand_gauge(site, 0.4, 41)
    om = ZMOBIUS_OMEGA if omega else None
    geo, inv, orc = make_pair(site, U, mass=0.05, ls=ls, m5=1.8, mobius_scale=2.0, omega=om)
    for eo in (1, 2):
        x = rand_field((geo.total_volume() // 2, ls, 12), 500 + eo)
        got = q.multiply_hermop_sym2_fp32(None, half(geo, ls, eo, x), inv)
        assert rel(got.data, orc.apply_half("hermop_sym2", x, eo)) < 2e-6


@pytest.fixture(scope="module")
def prop_case():
    print("\n[gpu] multi-RHS propagator 4^3x8 Ls8", flush=True)
    site = (4, 4, 4, 8)
    U = rand_gauge(site, 0.4, 51)
    orc = O.OracleInverter(site, U, mass=0.05, ls=8, m5=1.8, mobius_scale=2.0)
    orc.set_params(1e-12, 200, 300)
    rng = np.random.default_rng(52)
    src = rng.standard_normal((512, 12, 12)) + 1j * rng.standard_normal((512, 12, 12))
    src[:, :, 9:] = 0.0  # three zero columns
    exact, _ = orc.invert_prop(src)
    return site, U, src, exact


@pytest.mark.parametrize("tile,fused", [("12", "1"), ("5", "1"), ("1", "1"), ("12", "0")])
def test_multi_rhs_mixed_propagator(monkeypatch, prop_case, tile, fused):
    """12 columns solved as one multi-RHS block (tile 12), in uneven tiles
    (5+5+2) and column by column (tile 1): every column reaches the FP64
    true residual and matches the oracle propagator; zero columns stay zero."""
    monkeypatch.setenv("QLAT_DWF_RHS_TILE", tile)
    monkeypatch.setenv("QLAT_DWF_FUSED", fused)
    monkeypatch.setenv("QLAT_DWF_THREE_PHASE", "0")  # the FP32 fused / unfused kernels
    site, U, src, exact = prop_case
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=0.05, ls=8, m5=1.8, mobius_scale=2.0)
    inv = q.InverterDomainWall(gf=gf, fa=fa, solver="mixed")
    inv.set_stop_rsd(1e-9)
    p = q.Prop(geo)
    p.data = src
    sol = inv * p
    st = inv.stats()
    assert st["true_residual"] <= 1e-9
    assert not np.any(sol.data[:, :, 9:])
    assert rel(sol.data, exact) < 1e-6
    for j in range(9):
        assert rel(sol.data[:, :, j], exact[:, :, j]) < 1e-6
