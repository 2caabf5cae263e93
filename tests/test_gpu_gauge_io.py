"""A solve from a NERSC configuration written and read back by the gauge
I/O (qcd.h:143-270): the propagator on the loaded field equals the one on
the field in memory and the oracle's, here with the fifth dimension of the
reference's production actions (Ls 24, mobius_scale 4, rbc_ukqcd.py:487-502)
on the fast kernels."""

import numpy as np
import pytest

import qlattice_amd as q
from oracle import oracle as O
from tests.helpers import rand_field, rand_gauge, rel

pytestmark = pytest.mark.gpu


def test_propagator_from_a_loaded_configuration(tmp_path):
    site = (4, 4, 4, 8)
    print("\n[gpu] NERSC round trip, then a Ls 24 propagator on the loaded field", flush=True)
    # the file holds SU(3) links; the antiperiodic twist (qcd.h:331-346) is
    # applied after loading, as in the reference's scripts
    U = rand_gauge(site, 0.4, 901, antiperiodic=False)
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U.copy()
    path = str(tmp_path / "cfg.nersc")
    gf.save(path)
    g2 = q.GaugeField()
    g2.load(path)
    assert np.max(np.abs(g2.data - U)) < 1e-13
    for g in (gf, g2):
        g.twist_boundary_at_boundary(-0.5, 3)
    U = gf.data
    kw = dict(mass=0.05, ls=24, m5=1.8, mobius_scale=4.0)
    src = q.Prop(geo)
    src.data = np.zeros(src.data.shape, dtype=np.complex128)
    col = rand_field((geo.total_volume(), 12), 902)
    src.data[:, :, 2] = col
    sols = []
    for g in (gf, g2):
        inv = q.InverterDomainWall(gf=g, fa=q.FermionAction(**kw), solver="fp64")
        inv.set_stop_rsd(1e-10)
        sols.append((inv * src).data[:, :, 2].copy())
    orc = O.OracleInverter(site, U, **kw)
    orc.set_params(1e-10, 200, 300)
    ref, _ = orc.invert_4d(col)
    assert rel(sols[1], sols[0]) < 1e-10
    assert rel(sols[1], ref) < 1e-10
