"""Device-pointer entry points (the `_device` variants of include/qlat_dwf.h)
against their host-pointer counterparts on the same inputs: the same
computation, so the results are identical."""

import numpy as np
import pytest
import torch

import qlattice_amd as q
from qlattice_amd import _lib
from tests.helpers import rand_field, rand_gauge

pytestmark = pytest.mark.gpu

SITE = (4, 4, 4, 8)


@pytest.fixture(scope="module")
def inv():
    print("\n[gpu] device entry points 4^3x8 Ls8", flush=True)
    geo = q.Geometry(SITE)
    gf = q.GaugeField(geo)
    gf.data = rand_gauge(SITE, 0.3, 91)
    return geo, q.InverterDomainWall(gf=gf, fa=q.FermionAction(mass=0.1, ls=8, m5=1.8, mobius_scale=2.0),
                                     solver="mixed16")


def test_hermop_and_invert_prop_device(inv):
    geo, inv = inv
    L = _lib.load()
    x = rand_field((geo.total_volume() // 2, 8, 12), 92)
    host = np.empty_like(x)
    L.multiply_hermop_sym2(inv.cdata, _lib.ptr(host), _lib.ptr(x), 1)
    xd = torch.from_numpy(x).cuda()
    od = torch.empty_like(xd)
    L.multiply_hermop_sym2_device(inv.cdata, _lib.dev_ptr(od), _lib.dev_ptr(xd), 1)
    assert np.array_equal(od.cpu().numpy(), host)
    src = q.mk_point_src(geo, q.Coordinate([1, 0, 2, 3]))
    sol_h = inv * src
    sd = torch.from_numpy(src.data).cuda()
    so = torch.zeros_like(sd)
    L.invert_prop(inv.cdata, _lib.dev_ptr(so), _lib.dev_ptr(sd), device=True)
    assert np.array_equal(so.cpu().numpy(), sol_h.data)


def test_source_device_entry_points(inv):
    geo, inv = inv
    L = _lib.load()
    src = dict(kind=_lib.SRC_WALL, xg=(0, 0, 0, 0), tslice=3, lmom=(1.0, 2.0, 0.0, 0.0), value=1.0, fu1=0)
    host = q.mk_wall_src(geo, 3, q.CoordinateD((1, 2, 0, 0)))
    dev = torch.zeros((geo.total_volume(), 12, 12), dtype=torch.complex128, device="cuda")
    L.make_source(SITE, **src, prop=_lib.dev_ptr(dev), device=True)
    assert np.array_equal(dev.cpu().numpy(), host.data)
    sol_h = q.set_wall_src_propagator(q.Prop(), inv, 3, q.CoordinateD((1, 2, 0, 0)))
    so = torch.zeros_like(dev)
    L.source_propagator(inv.cdata, **src, sol=_lib.dev_ptr(so), device=True)
    assert np.array_equal(so.cpu().numpy(), sol_h.data)
