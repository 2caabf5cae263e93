"""Propagator sources and source propagators (qlat qcd-prop.cpp:61-530,
qcd-prop.h:202-290, :397-420; Python surface propagator.pyx:175-234).

Every source is a unit spin-colour matrix times a site factor f(x), built on
the device by the HIP library (``qlat_dwf_make_source``):

  point        f = value at x = xg                       set_point_src
  wall         f = exp(i mom.x) on t = tslice            set_wall_src, set_tslice_mom_src
  momentum     f = exp(i mom.x) on every site            set_mom_src, set_volume_src
  U(1) volume  f = fu1(x)                                set_rand_vol_u1_src

with mom = lmom * 2 pi / total_site (lattice_mom_mult, field.h:14-22).  The
``*_src_propagator`` functions build the source on the device and solve its 12
columns as one multi-RHS problem, so the source never crosses PCIe.
"""

import numpy as np

from . import _lib
from .fields import Coordinate, CoordinateD, FieldComplexD, Prop


def _source(kind, xg=(0, 0, 0, 0), tslice=0, lmom=None, value=1.0, fu1=None):
    """The fields of a qlat_dwf_source as keyword arguments of the binding's
    make_source / source_propagator; keep is the host array fu1 points to."""
    keep = None if fu1 is None else np.ascontiguousarray(fu1, dtype=np.complex128)
    s = dict(kind=int(kind), xg=[int(v) for v in xg], tslice=int(tslice),
             lmom=[float(0.0 if lmom is None else lmom[mu]) for mu in range(4)], value=complex(value),
             fu1=0 if keep is None else keep.ctypes.data)
    return s, keep


def _fill(prop, geo, s):
    assert prop.data.shape == (geo.total_volume(), 12, 12)
    if prop.data.dtype != np.complex128 or not prop.data.flags.c_contiguous:
        prop.data = np.zeros((geo.total_volume(), 12, 12), dtype=np.complex128)
    _lib.load().make_source(geo.total_site(), s["kind"], s["xg"], s["tslice"], s["lmom"], s["value"], s["fu1"],
                            _lib.ptr(prop.data))
    return prop


def _prop_for(prop, geo):
    if prop.geo is None:
        prop.geo = geo
    if prop.data is None or prop.data.shape != (geo.total_volume(), 12, 12):
        prop.data = np.zeros((geo.total_volume(), 12, 12), dtype=np.complex128)
    return prop


# ------------------------------------------------------------------ sources

def set_point_src(prop, geo, xg, value=1.0):
    """qcd-prop.cpp:75-89: value * unit matrix at xg (zero if xg is off the lattice)."""
    s, _ = _source(_lib.SRC_POINT, xg=Coordinate(xg), value=value)
    return _fill(_prop_for(prop, geo), geo, s)


def set_wall_src(prop, geo, tslice, lmom=None):
    """qcd-prop.cpp:117-150: exp(i mom.x) unit matrix on t = tslice (lmom[3] must be 0)."""
    s, _ = _source(_lib.SRC_WALL, tslice=tslice, lmom=CoordinateD(lmom) if lmom is not None else None)
    return _fill(_prop_for(prop, geo), geo, s)


def set_tslice_mom_src(prop, geo, tslice, lmom):
    """qcd-prop.cpp:465-499 (the same field as set_wall_src)."""
    return set_wall_src(prop, geo, tslice, lmom)


def set_mom_src(prop, geo, lmom):
    """qcd-prop.cpp:275-307: exp(i mom.x) unit matrix on every site."""
    s, _ = _source(_lib.SRC_MOM, lmom=CoordinateD(lmom))
    return _fill(_prop_for(prop, geo), geo, s)


def set_volume_src(prop, geo, lmom):
    """qcd-prop.cpp:501-530: set_mom_src with lmom[3] = 0 required."""
    if CoordinateD(lmom)[3] != 0:
        raise ValueError("set_volume_src: lmom[3] must be 0 (qcd-prop.cpp:505)")
    return set_mom_src(prop, geo, lmom)


PI = 3.141592653589793  # qlat-utils core.h:41


def set_rand_vol_u1(fu1, geo, multiplicity, rs):
    """qcd-prop.cpp:236-258 (propagator.pyx:183-189): fu1 = exp(i phase) per
    site and component, the phases u_rand_gen(PI, -PI) of rs.newtype(g) at
    site g in order.  Host code (one SHA-256 per four phases)."""
    import math

    from .rng import RngState

    if not isinstance(rs, RngState):
        raise TypeError("set_rand_vol_u1: rs must be an RngState")
    V = geo.total_volume()
    data = np.empty((V, int(multiplicity)), np.complex128)
    for g in range(V):
        rst = rs.newtype(g)
        for m in range(int(multiplicity)):
            ph = rst.u_rand_gen(PI, -PI)
            data[g, m] = complex(math.cos(ph), math.sin(ph))
    fu1.geo = geo
    fu1.data = data
    return fu1


def set_rand_vol_u1_src(prop, fu1):
    """qcd-prop.cpp:260-273: fu1(x) unit matrix (fu1: FieldComplexD with multiplicity 1)."""
    if fu1.data.shape[1:] not in ((), (1,)):
        raise ValueError("set_rand_vol_u1_src: fu1.multiplicity must be 1")
    s, keep = _source(_lib.SRC_U1, fu1=fu1.data.reshape(-1))
    out = _fill(_prop_for(prop, fu1.geo), fu1.geo, s)
    del keep
    return out


def mk_point_src(geo, xg, value=1.0):
    """propagator.pyx:201-205."""
    return set_point_src(Prop(geo), geo, xg, value)


def mk_wall_src(geo, tslice, lmom=None):
    """propagator.pyx:207-211."""
    return set_wall_src(Prop(geo), geo, tslice, lmom)


def mk_rand_vol_u1_src(fu1):
    """propagator.pyx:227-235."""
    return set_rand_vol_u1_src(Prop(fu1.geo), fu1)


# ------------------------------------------------------- source propagators

def _solve(prop, inv, s):
    geo = inv.geo()
    _prop_for(prop, geo)
    if prop.data.dtype != np.complex128 or not prop.data.flags.c_contiguous:
        prop.data = np.zeros((geo.total_volume(), 12, 12), dtype=np.complex128)
    _lib.load().source_propagator(inv.cdata, s["kind"], s["xg"], s["tslice"], s["lmom"], s["value"], s["fu1"],
                                  _lib.ptr(prop.data))
    return prop


def set_point_src_propagator(prop, inv, xg, value=1.0):
    """qcd-prop.h:202-218: the solution for set_point_src, 12 columns solved together."""
    s, _ = _source(_lib.SRC_POINT, xg=Coordinate(xg), value=value)
    return _solve(prop, inv, s)


def set_wall_src_propagator(prop, inv, tslice, lmom=None):
    """qcd-prop.h:226-243."""
    s, _ = _source(_lib.SRC_WALL, tslice=tslice, lmom=CoordinateD(lmom) if lmom is not None else None)
    return _solve(prop, inv, s)


def set_tslice_mom_src_propagator(prop, inv, tslice, lmom):
    """qcd-prop.h:403-420."""
    return set_wall_src_propagator(prop, inv, tslice, lmom)


def set_mom_src_propagator(prop, inv, lmom):
    """qcd-prop.h:273-290."""
    s, _ = _source(_lib.SRC_MOM, lmom=CoordinateD(lmom))
    return _solve(prop, inv, s)
