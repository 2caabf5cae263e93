"""Fifth-dimension extents above 16 (the reference's production actions use
Ls = 24 with mobius_scale = 4, qlat_scripts/v1/rbc_ukqcd.py:487-502):
operators, the FP64 solve and the mixed solvers against the oracle at
4^3x8, Moebius and zMoebius, Ls 20 and 24."""

import numpy as np
import pytest

import qlattice_amd as q
from oracle import oracle as O
from tests.helpers import rand_field, rand_gauge, rel

pytestmark = pytest.mark.gpu

SITE = (4, 4, 4, 8)


def zmobius_omega(ls):
    """A zMoebius-like omega: real values falling with s, the last four as
    two complex-conjugate pairs (the shape of free-invert.py:49-67), kept
    above 0.1 so that M_ee stays well conditioned at these Ls (a smaller
    omega makes b_s = (1/omega_s + 1)/2 large, and the FP64 solves of the
    dense-matrix and the LDU forms of M_ee^-1 then part by more than their
    1e-10 bar through rounding alone)."""
    w = [1.1 * 0.92 ** s for s in range(ls - 4)]
    w += [complex(0.2, -0.02), complex(0.2, 0.02), complex(0.12, -0.03), complex(0.12, 0.03)]
    return w


CASES = [(20, 2.0, None), (24, 4.0, None), (20, None, "z"), (24, None, "z")]


@pytest.fixture(scope="module", params=CASES, ids=["ls20", "ls24_scale4", "ls20_z", "ls24_z"])
def pair(request):
    ls, scale, z = request.param
    print(f"\n[gpu] Ls {ls} {'zMoebius' if z else f'Moebius scale {scale}'} 4^3x8", flush=True)
    U = rand_gauge(SITE, 0.4, 700 + ls)
    om = zmobius_omega(ls) if z else None
    kw = dict(mass=0.05, ls=ls, m5=1.8, mobius_scale=scale or 1.0, omega=om)
    geo = q.Geometry(SITE)
    gf = q.GaugeField(geo)
    gf.data = U
    inv = q.InverterDomainWall(gf=gf, fa=q.FermionAction(**kw))
    orc = O.OracleInverter(SITE, U, **kw)
    return geo, ls, inv, orc


OPS = [("m_e_e", q.multiply_m_e_e), ("m_e_e_inv", q.multiply_m_e_e_inv), ("mdag_e_e_inv", q.multiply_mdag_e_e_inv),
       ("m_e_o", q.multiply_m_e_o), ("mdag_e_o", q.multiply_mdag_e_o), ("hermop_sym2", q.multiply_hermop_sym2)]


def test_operators_match_oracle(pair):
    geo, ls, inv, orc = pair
    for eo in (1, 2):
        x = rand_field((geo.total_volume() // 2, ls, 12), 710 + eo)
        f = q.FermionField5d(geo, ls, eo)
        f.data = x
        for name, fn in OPS:
            assert rel(fn(None, f, inv).data, orc.apply_half(name, x, eo)) < 1e-12, (name, eo)
    y = rand_field((geo.total_volume(), ls, 12), 713)
    full = q.FermionField5d(geo, ls, 0)
    full.data = y
    for name, fn in (("m", q.multiply_m), ("d_minus", q.multiply_d_minus)):
        assert rel(fn(None, full, inv).data, orc.apply_full(name, y)) < 1e-12, name


def test_fp64_solve_matches_oracle(pair):
    geo, ls, inv, orc = pair
    inv.set_solver("fp64")
    inv.set_stop_rsd(1e-10)
    orc.set_params(1e-10, 200, 300)
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 720)
    sol = inv * src
    ref, it_ref = orc.invert_4d(src.data)
    assert inv.stats()["total_iter"] == it_ref
    assert rel(sol.data, ref) < 1e-10
    inv.set_stop_rsd(1e-8)


def test_low_precision_operators_match_oracle(pair):
    """The operators of the mixed solvers at Ls 20 and 24: FP32 (2e-6) and
    the 16-bit storage path (three-phase kernels with 24 rows, FP16
    arithmetic in the hop phase) within its 3e-3 bound, and above 1e-5 so
    that it is the 16-bit path that ran."""
    geo, ls, inv, orc = pair
    x = rand_field((geo.total_volume() // 2, ls, 12), 740)
    f = q.FermionField5d(geo, ls, 1)
    f.data = x
    ref = orc.apply_half("hermop_sym2", x, 1)
    inv.set_solver("mixed")
    e32 = rel(q.multiply_hermop_sym2_fp32(None, f, inv).data, ref)
    inv.set_solver("mixed16")
    e16 = rel(q.multiply_hermop_sym2_fp32(None, f, inv).data, ref)
    inv.set_solver("fp64")
    assert e32 < 2e-6
    assert 1e-5 < e16 < 3e-3


@pytest.mark.parametrize("solver", ["mixed", "mixed16"])
def test_mixed_solvers_reach_tolerance(pair, solver):
    """The mixed solvers for Ls 20 and 24 on the three-phase kernels (the
    24-row instances): FP32 storage, and FP16 storage with per-site scale."""
    geo, ls, inv, orc = pair
    orc.set_params(1e-12, 200, 300)
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 730)
    exact, _ = orc.invert_4d(src.data)
    inv.set_solver(solver)
    inv.set_stop_rsd(1e-9)
    sol = inv * src
    st = inv.stats()
    inv.set_solver("fp64")
    inv.set_stop_rsd(1e-8)
    assert st["true_residual"] <= 1e-9
    assert st["spinor_bits"] == (16 if solver == "mixed16" else 32)
    assert st["hop_arith_bits"] == (16 if solver == "mixed16" else 32)
    assert rel(sol.data, exact) < 1e-6


@pytest.mark.parametrize("z", [False, True], ids=["moebius_scale4", "zmoebius"])
def test_ls24_eight_site_blocks(monkeypatch, z):
    """Ls 24 on blocks of 8 sites x 24 rows (3 waves; the s-matrix phase as
    products of 8 sites x 2 reals on the matrix cores, complex matrices
    with the partner real's operand): the FP32- and FP16-storage operators
    against the oracle and the mixed16 solve at the FP64 true residual."""
    monkeypatch.setenv("QLAT_DWF_TP_SB", "8")
    monkeypatch.setenv("QLAT_DWF_RHS_TILE", "4")
    ls = 24
    print(f"\n[gpu] Ls 24 on 8-site blocks, {'zMoebius' if z else 'Moebius scale 4'}", flush=True)
    U = rand_gauge(SITE, 0.4, 760 + z)
    kw = dict(mass=0.05, ls=ls, m5=1.8, mobius_scale=1.0 if z else 4.0, omega=zmobius_omega(ls) if z else None)
    geo = q.Geometry(SITE)
    gf = q.GaugeField(geo)
    gf.data = U
    inv = q.InverterDomainWall(gf=gf, fa=q.FermionAction(**kw))
    orc = O.OracleInverter(SITE, U, **kw)
    x = rand_field((geo.total_volume() // 2, ls, 12), 761)
    f = q.FermionField5d(geo, ls, 1)
    f.data = x
    ref = orc.apply_half("hermop_sym2", x, 1)
    inv.set_solver("mixed")
    e32 = rel(q.multiply_hermop_sym2_fp32(None, f, inv).data, ref)
    inv.set_solver("mixed16")
    e16 = rel(q.multiply_hermop_sym2_fp32(None, f, inv).data, ref)
    assert e32 < 2e-6
    assert 1e-5 < e16 < 3e-3
    orc.set_params(1e-12, 200, 300)
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 762)
    exact, _ = orc.invert_4d(src.data)
    p = q.Prop(geo)
    p.data = np.zeros(p.data.shape, dtype=np.complex128)
    p.data[:, :, 0] = src.data
    p.data[:, :, 5] = 0.5j * src.data
    inv.set_stop_rsd(1e-9)
    sol = inv * p
    st = inv.stats()
    assert st["sites_per_block"] == 8 and st["spinor_bits"] == 16
    assert st["true_residual"] <= 1e-9
    assert rel(sol.data[:, :, 0], exact) < 1e-6
    assert rel(sol.data[:, :, 5], 0.5j * exact) < 1e-6
