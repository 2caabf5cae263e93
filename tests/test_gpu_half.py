"""GPU parity of the three-phase kernels and of FP16 spinor storage with a
per-5D-site scale (solver "mixed16"): the inner operator against the FP64
oracle at the accuracy its storage allows, and solves at the stated
FP64 true-residual tolerance against the oracle's solution.
"""

import numpy as np
import pytest

import qlattice_amd as q
from oracle import oracle as O
from tests.helpers import ZMOBIUS_OMEGA, rand_field, rand_gauge, rel

pytestmark = pytest.mark.gpu


def make(site, U, *, mass, ls, mobius_scale=2.0, omega=None, solver="mixed16"):
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=mass, ls=ls, m5=1.8, mobius_scale=mobius_scale, omega=omega)
    inv = q.InverterDomainWall(gf=gf, fa=fa, solver=solver)
    orc = O.OracleInverter(site, U, mass=mass, ls=ls, m5=1.8, mobius_scale=mobius_scale, omega=omega)
    return geo, inv, orc


def half(geo, ls, eo, data):
    f = q.FermionField5d(geo, ls, eo)
    f.data = np.ascontiguousarray(data)
    return f


@pytest.mark.parametrize("ls,omega", [(8, None), (12, "z"), (16, None)])
def test_three_phase_fp32_hermop_matches_oracle(monkeypatch, ls, omega):
    """FP32 storage on the three-phase kernels: FP32 accuracy."""
    print(f"\n[gpu] three-phase FP32 hermop Ls{ls}", flush=True)
    monkeypatch.setenv("QLAT_DWF_THREE_PHASE", "1")
    site = (4, 4, 4, 8)
    U = rand_gauge(site, 0.4, 41)
    geo, inv, orc = make(site, U, mass=0.05, ls=ls, omega=ZMOBIUS_OMEGA if omega else None, solver="mixed")
    for eo in (1, 2):
        x = rand_field((geo.total_volume() // 2, ls, 12), 500 + eo)
        got = q.multiply_hermop_sym2_fp32(None, half(geo, ls, eo, x), inv)
        assert rel(got.data, orc.apply_half("hermop_sym2", x, eo)) < 2e-6


@pytest.mark.parametrize("h2", ["1", "0"], ids=["fp16_arith", "fp32_arith"])
@pytest.mark.parametrize("ls,omega", [(8, None), (12, "z"), (16, None)])
def test_fp16_storage_hermop_matches_oracle(monkeypatch, ls, omega, h2):
    """FP16 storage with a per-site scale: the input and every intermediate
    field are rounded to 11 significant bits relative to the site maximum,
    so the operator agrees with FP64 to a few 1e-4 (and is visibly not
    FP32: the error is far above FP32 rounding).  Both hop arithmetics:
    FP16 links with FP16 products accumulated in FP32 (the default), and
    FP32 links and arithmetic (QLAT_DWF_H2=0)."""
    monkeypatch.setenv("QLAT_DWF_H2", h2)
    print(f"\n[gpu] FP16-storage hermop Ls{ls} H2={h2}", flush=True)
    site = (4, 4, 4, 8)
    U = rand_gauge(site, 0.4, 41)
    geo, inv, orc = make(site, U, mass=0.05, ls=ls, omega=ZMOBIUS_OMEGA if omega else None)
    for eo in (1, 2):
        x = rand_field((geo.total_volume() // 2, ls, 12), 600 + eo)
        got = q.multiply_hermop_sym2_fp32(None, half(geo, ls, eo, x), inv)
        err = rel(got.data, orc.apply_half("hermop_sym2", x, eo))
        assert 1e-5 < err < 3e-3, err


@pytest.mark.parametrize("ls", [6, 8, 12, 16])
def test_plain_hop_block_shapes_agree_bitwise(monkeypatch, ls):
    """The plain FP16 hop pass (y = p - H b) runs on 64-site blocks by
    default and on the three-phase block shape with QLAT_DWF_PLAIN64=0: the
    same arithmetic per (site, s), so the operator is bit-identical."""
    print(f"\n[gpu] plain hop block shapes Ls{ls}", flush=True)
    site = (4, 4, 4, 8)
    U = rand_gauge(site, 0.4, 44)
    outs = []
    for v in ("1", "0"):
        monkeypatch.setenv("QLAT_DWF_PLAIN64", v)
        geo, inv, _ = make(site, U, mass=0.05, ls=ls, omega=ZMOBIUS_OMEGA if ls == 12 else None)
        x = rand_field((geo.total_volume() // 2, ls, 12), 604)
        outs.append(q.multiply_hermop_sym2_fp32(None, half(geo, ls, 1, x), inv).data)
    assert np.array_equal(outs[0], outs[1])


def test_fp16_storage_zero_and_scale_invariance():
    """A zero field stays exactly zero (scale 0), and scaling the input by a
    power of two scales the output exactly (the per-site scale absorbs it)."""
    site = (4, 4, 4, 8)
    U = rand_gauge(site, 0.4, 42)
    geo, inv, _ = make(site, U, mass=0.05, ls=8)
    z = np.zeros((geo.total_volume() // 2, 8, 12), np.complex128)
    assert not np.any(q.multiply_hermop_sym2_fp32(None, half(geo, 8, 1, z), inv).data)
    x = rand_field((geo.total_volume() // 2, 8, 12), 603)
    a = q.multiply_hermop_sym2_fp32(None, half(geo, 8, 1, x), inv).data
    b = q.multiply_hermop_sym2_fp32(None, half(geo, 8, 1, x * 2.0**-20), inv).data
    assert np.array_equal(a * 2.0**-20, b)


@pytest.mark.parametrize("solver,three,rx", [("mixed16", "0", "0"), ("mixed16", "0", "1"), ("mixed", "1", "0")])
def test_three_phase_solver_reaches_tolerance(monkeypatch, solver, three, rx):
    """Single right-hand side through the three-phase solver (FP16 Krylov
    vectors with FP32 or FP16 r and x, or all FP32): FP64 true residual <=
    stop_rsd and the solution as close to the exact one as the FP64
    solver's at the same stop_rsd."""
    print(f"\n[gpu] {solver} three-phase solve 8^3x16 Ls8 (rx16={rx})", flush=True)
    monkeypatch.setenv("QLAT_DWF_THREE_PHASE", three)
    monkeypatch.setenv("QLAT_DWF_FP16_MODE", "2" if rx == "1" else "1")
    site = (8, 8, 8, 16)
    U = rand_gauge(site, 0.3, 3)
    geo, inv, orc = make(site, U, mass=0.05, ls=8, solver="fp64")
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 4)
    orc.set_params(1e-12, 200, 300)
    exact, _ = orc.invert_4d(src.data)
    inv.set_stop_rsd(1e-8)
    e64 = rel((inv * src).data, exact)
    inv.set_solver(solver)
    sol = inv * src
    st = inv.stats()
    assert st["true_residual"] <= 1e-8
    assert st["spinor_bits"] == (16 if solver == "mixed16" else 32)
    assert st["reliable_updates"] > 0
    assert rel(sol.data, exact) < 10 * e64 + 1e-9


def test_fp16_storage_zmobius_solve():
    """Complex s-matrices (zMoebius) through the FP16 path."""
    print("\n[gpu] mixed16 zMoebius solve 4^3x8 Ls12", flush=True)
    site = (4, 4, 4, 8)
    U = rand_gauge(site, 0.4, 43)
    geo, inv, orc = make(site, U, mass=0.05, ls=12, mobius_scale=1.0, omega=ZMOBIUS_OMEGA)
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 44)
    orc.set_params(1e-12, 200, 300)
    exact, _ = orc.invert_4d(src.data)
    inv.set_stop_rsd(1e-9)
    sol = inv * src
    assert inv.stats()["true_residual"] <= 1e-9
    assert rel(sol.data, exact) < 1e-6


@pytest.fixture(scope="module")
def prop_case():
    print("\n[gpu] FP16 multi-RHS propagator 4^3x8 Ls8", flush=True)
    site = (4, 4, 4, 8)
    U = rand_gauge(site, 0.4, 51)
    orc = O.OracleInverter(site, U, mass=0.05, ls=8, m5=1.8, mobius_scale=2.0)
    orc.set_params(1e-12, 200, 300)
    rng = np.random.default_rng(52)
    src = rng.standard_normal((512, 12, 12)) + 1j * rng.standard_normal((512, 12, 12))
    src[:, :, 9:] = 0.0  # three zero columns
    exact, _ = orc.invert_prop(src)
    return site, U, src, exact


@pytest.mark.parametrize("solver,tile,three,rx", [("mixed16", "12", "0", "0"), ("mixed16", "5", "0", "0"),
                                                  ("mixed16", "1", "0", "0"), ("mixed", "12", "1", "0"),
                                                  ("mixed16", "12", "0", "1"), ("mixed16", "0", "0", "1")])
def test_three_phase_multi_rhs_propagator(monkeypatch, prop_case, solver, tile, three, rx):
    """12 columns as one block (tile 12), in uneven tiles (5+5+2), one by one
    and in the autotuned tile (0): every column reaches the FP64 true
    residual and matches the oracle; zero columns stay zero."""
    monkeypatch.setenv("QLAT_DWF_RHS_TILE", tile)
    monkeypatch.setenv("QLAT_DWF_THREE_PHASE", three)
    monkeypatch.setenv("QLAT_DWF_FP16_MODE", "2" if rx == "1" else "1")
    site, U, src, exact = prop_case
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=0.05, ls=8, m5=1.8, mobius_scale=2.0)
    inv = q.InverterDomainWall(gf=gf, fa=fa, solver=solver)
    inv.set_stop_rsd(1e-9)
    p = q.Prop(geo)
    p.data = src
    sol = inv * p
    st = inv.stats()
    assert st["true_residual"] <= 1e-9
    assert st["spinor_bits"] == (16 if solver == "mixed16" else 32)
    assert not np.any(sol.data[:, :, 9:])
    for j in range(9):
        assert rel(sol.data[:, :, j], exact[:, :, j]) < 1e-6


@pytest.mark.parametrize("solver", ["mixed16", "mixed"])
def test_graph_replay_matches_eager_loop(monkeypatch, prop_case, solver):
    """Graph-captured CG iterations (QLAT_DWF_GRAPH = K iterations per
    replay, scalars on the device, x updated every other step with two
    directions as in the host loop) against the loop with one host round
    trip per iteration (QLAT_DWF_GRAPH=0): every column does the same
    arithmetic in the same order (a column that reaches its reliable-update
    point waits frozen for the end of the replay), so the iteration and
    reliable-update counts are equal; every column at the FP64 true residual
    and on the oracle's solution, replayed many times with the FP64 updates
    between replays, from both p buffers (K odd and even)."""
    print(f"\n[gpu] graph replay vs eager loop ({solver}) 4^3x8 Ls8", flush=True)
    monkeypatch.setenv("QLAT_DWF_RHS_TILE", "12")
    monkeypatch.setenv("QLAT_DWF_THREE_PHASE", "1")
    site, U, src, exact = prop_case
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=0.05, ls=8, m5=1.8, mobius_scale=2.0)
    out = {}
    for k in ("0", "1", "3", "16"):
        monkeypatch.setenv("QLAT_DWF_GRAPH", k)
        inv = q.InverterDomainWall(gf=gf, fa=fa, solver=solver)
        inv.set_stop_rsd(1e-9)
        p = q.Prop(geo)
        p.data = src
        sol = inv * p
        out[k] = (sol.data.copy(), inv.stats())
        st = out[k][1]
        assert st["true_residual"] <= 1e-9, k
        assert not np.any(sol.data[:, :, 9:])
        for j in range(9):
            assert rel(sol.data[:, :, j], exact[:, :, j]) < 1e-6, (k, j)
    base_sol, base = out["0"]
    for k in ("1", "3", "16"):
        sol, st = out[k]
        assert st["total_iter"] == base["total_iter"], (k, st, base)
        assert st["reliable_updates"] == base["reliable_updates"], (k, st, base)
        assert rel(sol, base_sol) < 1e-7, k


@pytest.mark.parametrize("solver", ["mixed16", "mixed"])
def test_folded_x_update_matches_every_step_update(monkeypatch, prop_case, solver):
    """x updated every other CG step with two directions at once (the
    default, double-buffered p) against x updated at every step
    (QLAT_DWF_XDEFER=0): the CG recurrences do not change, so the iteration
    and reliable-update counts agree and the solutions differ only by FP32
    rounding of x."""
    print(f"\n[gpu] folded x update vs every-step update ({solver}) 4^3x8 Ls8", flush=True)
    monkeypatch.setenv("QLAT_DWF_RHS_TILE", "5")
    monkeypatch.setenv("QLAT_DWF_THREE_PHASE", "1")
    site, U, src, exact = prop_case
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=0.05, ls=8, m5=1.8, mobius_scale=2.0)
    out = {}
    for d in ("0", "1"):
        monkeypatch.setenv("QLAT_DWF_XDEFER", d)
        inv = q.InverterDomainWall(gf=gf, fa=fa, solver=solver)
        inv.set_stop_rsd(1e-9)
        p = q.Prop(geo)
        p.data = src
        sol = inv * p
        st = inv.stats()
        assert st["true_residual"] <= 1e-9, d
        for j in range(9):
            assert rel(sol.data[:, :, j], exact[:, :, j]) < 1e-6, (d, j)
        out[d] = (sol.data.copy(), st)
    (s0, st0), (s1, st1) = out["0"], out["1"]
    assert abs(st1["total_iter"] - st0["total_iter"]) <= 0.02 * st0["total_iter"], (st0, st1)
    assert abs(st1["reliable_updates"] - st0["reliable_updates"]) <= 1, (st0, st1)
    assert rel(s1, s0) < 1e-7


@pytest.mark.parametrize("solver", ["mixed16", "mixed"])
def test_cg_step_on_sixteen_site_blocks(monkeypatch, prop_case, solver):
    """The CG step on 16-site blocks (QLAT_DWF_STEP_SB=16: t = K p on the
    matrix cores) against the default 32-site step (t = K p row by row in
    FP32 on the vector units): the same solutions to FP32 rounding, every
    column at the FP64 true residual and on the oracle's solution."""
    print(f"\n[gpu] CG step on 16-site blocks ({solver}) 4^3x8 Ls8", flush=True)
    monkeypatch.setenv("QLAT_DWF_RHS_TILE", "5")
    site, U, src, exact = prop_case
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=0.05, ls=8, m5=1.8, mobius_scale=2.0)
    out = {}
    for sb in ("32", "16"):
        monkeypatch.setenv("QLAT_DWF_STEP_SB", sb)
        inv = q.InverterDomainWall(gf=gf, fa=fa, solver=solver)
        inv.set_stop_rsd(1e-9)
        p = q.Prop(geo)
        p.data = src
        sol = inv * p
        st = inv.stats()
        assert st["true_residual"] <= 1e-9, sb
        for j in range(9):
            assert rel(sol.data[:, :, j], exact[:, :, j]) < 1e-6, (sb, j)
        out[sb] = (sol.data.copy(), st)
    (s32, st32), (s16, st16) = out["32"], out["16"]
    assert abs(st16["total_iter"] - st32["total_iter"]) <= 0.02 * st32["total_iter"]
    assert rel(s16, s32) < 1e-7


@pytest.mark.parametrize("z", [False, True], ids=["moebius", "zmoebius"])
def test_ls16_eight_site_blocks(monkeypatch, z):
    """Ls 16 on blocks of 8 sites x 16 rows (2 waves; the s-matrix phase as
    matrix-core products of 8 sites x 2 reals, three float4 columns per
    wave): FP32- and FP16-storage operators against the oracle, and a
    mixed16 propagator column at the FP64 true residual."""
    monkeypatch.setenv("QLAT_DWF_TP_SB", "8")
    monkeypatch.setenv("QLAT_DWF_RHS_TILE", "4")
    ls = 16
    print(f"\n[gpu] Ls 16 on 8-site blocks {'zMoebius' if z else 'Moebius'}", flush=True)
    site = (4, 4, 4, 8)
    U = rand_gauge(site, 0.4, 45 + z)
    om = [1.1 * 0.92 ** s for s in range(ls - 4)] + [complex(0.2, -0.02), complex(0.2, 0.02),
                                                       complex(0.12, -0.03), complex(0.12, 0.03)]
    geo, inv, orc = make(site, U, mass=0.05, ls=ls, omega=om if z else None)
    x = rand_field((geo.total_volume() // 2, ls, 12), 650)
    ref = orc.apply_half("hermop_sym2", x, 1)
    inv.set_solver("mixed")
    assert rel(q.multiply_hermop_sym2_fp32(None, half(geo, ls, 1, x), inv).data, ref) < 2e-6
    inv.set_solver("mixed16")
    e16 = rel(q.multiply_hermop_sym2_fp32(None, half(geo, ls, 1, x), inv).data, ref)
    assert 1e-5 < e16 < 3e-3
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 651)
    orc.set_params(1e-12, 200, 300)
    exact, _ = orc.invert_4d(src.data)
    p = q.Prop(geo)
    p.data = np.zeros(p.data.shape, dtype=np.complex128)
    p.data[:, :, 3] = src.data
    inv.set_stop_rsd(1e-9)
    sol = inv * p
    st = inv.stats()
    assert st["sites_per_block"] == 8
    assert st["true_residual"] <= 1e-9
    assert rel(sol.data[:, :, 3], exact) < 1e-6
