"""Batched solves over several inverters (qlat_dwf_invert_5d_batch): the
multi-RHS mixed solver with one operator per column (different masses in one
tile), against the oracle's one-by-one FP64 solves; the FP64 batch path
against single solves (iteration-exact); argument checks.
"""

import numpy as np
import pytest

import qlattice_amd as q
from oracle import oracle as O
from tests.helpers import rand_field, rand_gauge, rel

pytestmark = pytest.mark.gpu

SITE = (4, 4, 4, 8)
LS = 8
MASSES = (0.03, 0.1, 0.3)


@pytest.fixture(scope="module")
def case():
    print("\n[gpu] batch 4^3x8 Ls8, 3 sources x 3 masses", flush=True)
    U = rand_gauge(SITE, 0.4, 81)
    geo = q.Geometry(SITE)
    srcs = []
    for k in range(3):
        f = q.FermionField5d(geo, LS, 0)
        f.data = rand_field((geo.total_volume(), LS, 12), 82 + k)
        srcs.append(f)
    exact = {}
    for m in MASSES:
        orc = O.OracleInverter(SITE, U, mass=m, ls=LS, m5=1.8, mobius_scale=2.0)
        orc.set_params(1e-12, 200, 300)
        for k, f in enumerate(srcs):
            exact[(m, k)] = orc.invert_5d(f.data)[0]
    return U, geo, srcs, exact


def inverters(U, geo, solver):
    gf = q.GaugeField(geo)
    gf.data = U
    out = []
    for m in MASSES:
        inv = q.InverterDomainWall(gf=gf, fa=q.FermionAction(mass=m, ls=LS, m5=1.8, mobius_scale=2.0),
                                   solver=solver)
        inv.set_stop_rsd(1e-9)
        out.append(inv)
    return out


@pytest.mark.parametrize("solver,tile", [("mixed16", "0"), ("mixed16", "4"), ("mixed", "0")])
def test_multi_mass_batch_matches_oracle(monkeypatch, case, solver, tile):
    """9 systems (3 sources x 3 masses) in one call, masses interleaved in the
    tiles: each solution within 1e-6 of the oracle's FP64 solve, the true
    residual at the stop, and lighter masses need more iterations."""
    monkeypatch.setenv("QLAT_DWF_RHS_TILE", tile)
    U, geo, srcs, exact = case
    invs = inverters(U, geo, solver)
    pairs = [(m_i, k) for k in range(3) for m_i in range(3)]
    sols, iters = q.invert_batch([invs[m_i] for m_i, _ in pairs], [srcs[k] for _, k in pairs])
    st = invs[0].stats()
    assert st["true_residual"] <= 1e-9
    if tile != "0":
        assert st["rhs_tile"] == int(tile)
    for (m_i, k), sol, it in zip(pairs, sols, iters):
        assert rel(sol.data, exact[(MASSES[m_i], k)]) < 1e-6, (m_i, k)
    by_mass = [np.mean([it for (m_i, _), it in zip(pairs, iters) if m_i == j]) for j in range(3)]
    assert by_mass[0] > by_mass[1] > by_mass[2]


def test_fp64_batch_equals_single_solves(case):
    U, geo, srcs, exact = case
    invs = inverters(U, geo, "fp64")
    sols, iters = q.invert_batch(invs, srcs)
    for inv, f, sol, it in zip(invs, srcs, sols, iters):
        single = q.FermionField5d(geo, LS, 0)
        it1 = q.invert(single, f, inv)
        assert it == it1
        assert np.array_equal(sol.data, single.data)


def test_batch_rejects_mismatched_lattices(case):
    U, geo, srcs, _ = case
    inv_a = inverters(U, geo, "mixed16")[0]
    site_b = (4, 4, 4, 4)
    geo_b = q.Geometry(site_b)
    gf_b = q.GaugeField(geo_b)
    gf_b.data = rand_gauge(site_b, 0.4, 5)
    inv_b = q.InverterDomainWall(gf=gf_b, fa=q.FermionAction(mass=0.1, ls=LS, m5=1.8), solver="mixed16")
    with pytest.raises(q.QlatDwfError, match="share the lattice"):
        q.invert_batch([inv_a, inv_b], [srcs[0], srcs[1]])
