"""Device-built propagator sources and source propagators (qcd-prop.cpp:61-530,
qcd-prop.h:202-290, :397-420).

Pins: a point source has exactly one non-zero site; the discrete Fourier
transform (numpy's FFT, an independent routine) of a momentum source is
V * delta(k = lmom) and that of a wall source's time slice is
X*Y*Z * delta(k = lmom); the U(1) source carries the given field; every
source is a unit spin-colour matrix per site.  The source propagators are
pinned by the reference's free-field golden (examples-py/free-invert.log.json)
and by the oracle's solve of the same source.
"""

import json
import os

import numpy as np
import pytest

import qlattice_amd as q
from oracle import oracle as O
from tests.helpers import qlat_check, rand_gauge, rel

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "free_invert.json")))


def site_factor(p):
    """f(x) of a unit-matrix source, after checking the unit-matrix form."""
    f = p.data[:, 0, 0].copy()
    assert np.array_equal(p.data, f[:, None, None] * np.eye(12)[None])
    return f


def lattice_array(f, site):
    """[V] in reference order (x fastest) -> array indexed [t, z, y, x]."""
    X, Y, Z, T = site
    return f.reshape(T, Z, Y, X)


@pytest.mark.parametrize("xg,value", [((1, 2, 3, 5), 1.0), ((0, 0, 0, 0), 2.5 - 0.5j), ((9, 0, 0, 0), 1.0)])
def test_point_source(xg, value):
    print(f"\n[gpu] point source {xg}", flush=True)
    site = (4, 4, 6, 8)
    geo = q.Geometry(site)
    p = q.mk_point_src(geo, q.Coordinate(xg), value)
    f = site_factor(p)
    if all(0 <= xg[m] < site[m] for m in range(4)):
        i = geo.index_from_coordinate(xg)
        assert f[i] == value and np.count_nonzero(f) == 1
    else:  # off the lattice: no element is set (qcd-prop.cpp:70-72)
        assert not f.any()


@pytest.mark.parametrize("lmom", [(0, 0, 0, 0), (1, 0, 2, 0), (-1, 3, 1, 2)])
def test_momentum_source_is_a_single_fourier_mode(lmom):
    site = (4, 6, 8, 10)
    geo = q.Geometry(site)
    p = q.set_mom_src(q.Prop(geo), geo, q.CoordinateD(lmom))
    f = lattice_array(site_factor(p), site)
    assert np.allclose(np.abs(f), 1.0, atol=1e-14)
    # sum_x f(x) exp(-i 2 pi k.x / L) = V delta(k = lmom mod L)
    F = np.fft.fftn(f)  # axes t, z, y, x
    k = tuple(int(lmom[mu]) % site[mu] for mu in (3, 2, 1, 0))
    V = np.prod(site)
    assert abs(F[k] - V) < 1e-9
    F[k] = 0
    assert np.max(np.abs(F)) < 1e-9


@pytest.mark.parametrize("tslice,lmom", [(0, (0, 0, 0, 0)), (3, (1, 2, 0, 0)), (7, (-2, 1, 3, 0))])
def test_wall_source_is_one_slice_of_a_fourier_mode(tslice, lmom):
    site = (6, 4, 8, 8)
    geo = q.Geometry(site)
    p = q.mk_wall_src(geo, tslice, q.CoordinateD(lmom))
    f = lattice_array(site_factor(p), site)
    others = np.delete(f, tslice, axis=0)
    assert not others.any()
    F = np.fft.fftn(f[tslice])  # axes z, y, x
    k = tuple(int(lmom[mu]) % site[mu] for mu in (2, 1, 0))
    assert abs(F[k] - site[0] * site[1] * site[2]) < 1e-9
    F[k] = 0
    assert np.max(np.abs(F)) < 1e-9
    # set_tslice_mom_src is the same field
    p2 = q.set_tslice_mom_src(q.Prop(geo), geo, tslice, q.CoordinateD(lmom))
    assert np.array_equal(p.data, p2.data)


def test_u1_volume_source_and_argument_errors():
    site = (4, 4, 4, 6)
    geo = q.Geometry(site)
    rng = np.random.default_rng(5)
    fu1 = q.FieldComplexD(geo, 1)
    fu1.data[:, 0] = np.exp(2j * np.pi * rng.random(geo.total_volume()))
    p = q.mk_rand_vol_u1_src(fu1)
    assert np.array_equal(site_factor(p), fu1.data[:, 0])
    with pytest.raises(q.QlatDwfError):  # qassert(lmom[3] == 0), qcd-prop.cpp:120
        q.mk_wall_src(geo, 0, q.CoordinateD((0, 0, 0, 1)))
    with pytest.raises(ValueError):
        q.set_volume_src(q.Prop(geo), geo, q.CoordinateD((1, 0, 0, 1)))
    with pytest.raises(ValueError):
        q.mk_rand_vol_u1_src(q.FieldComplexD(geo, 2))
    # set_volume_src with lmom[3] = 0 is set_mom_src
    a = q.set_volume_src(q.Prop(geo), geo, q.CoordinateD((1, 1, 0, 0)))
    b = q.set_mom_src(q.Prop(geo), geo, q.CoordinateD((1, 1, 0, 0)))
    assert np.array_equal(a.data, b.data)


def test_golden_free_field_point_src_propagator():
    """examples-py/free-invert.log.json: the 12-column point-source DWF
    propagator on the unit 4^3x8 gauge field, built and solved on the device."""
    print("\n[gpu] free-field golden via set_point_src_propagator", flush=True)
    geo = q.Geometry((4, 4, 4, 8))
    gf = q.GaugeField(geo).set_unit()
    inv = q.InverterDomainWall(gf=gf, fa=q.FermionAction(mass=0.05, ls=16, m5=1.0))
    sol = q.set_point_src_propagator(q.Prop(), inv, q.Coordinate([0, 0, 0, 0]))
    v, eps = GOLD["point_source"]["sol1_qnorm_dwf_cg"]
    assert qlat_check(sol.qnorm(), v, eps), sol.qnorm()


@pytest.mark.parametrize("kind", ["point", "wall", "mom"])
def test_source_propagators_match_oracle_and_host_path(kind):
    """set_*_src_propagator equals inv * (the same source uploaded from the
    host) and the oracle's solve of that source."""
    print(f"\n[gpu] {kind} source propagator 4^3x8 Ls8", flush=True)
    site = (4, 4, 4, 8)
    geo = q.Geometry(site)
    U = rand_gauge(site, 0.3, 17)
    gf = q.GaugeField(geo)
    gf.data = U
    inv = q.InverterDomainWall(gf=gf, fa=q.FermionAction(mass=0.1, ls=8, m5=1.8, mobius_scale=2.0))
    inv.set_stop_rsd(1e-11)
    if kind == "point":
        src = q.mk_point_src(geo, q.Coordinate([1, 2, 3, 4]), 1.5)
        sol = q.set_point_src_propagator(q.Prop(), inv, q.Coordinate([1, 2, 3, 4]), 1.5)
    elif kind == "wall":
        src = q.mk_wall_src(geo, 2, q.CoordinateD((1, 0, 1, 0)))
        sol = q.set_wall_src_propagator(q.Prop(), inv, 2, q.CoordinateD((1, 0, 1, 0)))
    else:
        src = q.set_mom_src(q.Prop(), geo, q.CoordinateD((0, 1, 1, 1)))
        sol = q.set_mom_src_propagator(q.Prop(), inv, q.CoordinateD((0, 1, 1, 1)))
    host = inv * src
    assert rel(sol.data, host.data) < 1e-13
    orc = O.OracleInverter(site, U, mass=0.1, ls=8, m5=1.8, mobius_scale=2.0)
    orc.set_params(1e-12, 200, 300)
    ref, _ = orc.invert_prop(src.data)
    assert rel(sol.data, ref) < 1e-9
