"""Lattice geometry and field containers in the reference's memory layout.

These mirror the parts of qlat's Python surface that the domain-wall path
consumes (``q.Coordinate``, ``q.Geometry``, ``q.GaugeField``, ``q.Prop``,
``q.FermionField4d``, ``q.FieldComplexD``).  Fields are numpy arrays in the
reference's array-of-structures complex128 layout, so data moves to and from
the HIP library without reordering on the host:

  GaugeField      (V, 4, 3, 3)   core.h:1187
  FermionField4d  (V, 12)
  FermionField5d  (V, Ls, 12) full, or (V/2, Ls, 12) with eo = 1 odd / 2 even
  Prop            (V, 12, 12)    [site][row][source column]

Site index = x + X*(y + Y*(z + Z*t)) (coordinate.h:169-188).
"""

import numpy as np


class Coordinate(tuple):
    """4-vector of ints (qlat Coordinate)."""

    def __new__(cls, v=(0, 0, 0, 0)):
        v = tuple(int(a) for a in v)
        if len(v) != 4:
            raise ValueError("Coordinate needs 4 entries")
        return super().__new__(cls, v)

    def to_list(self):
        return list(self)


class CoordinateD(tuple):
    def __new__(cls, v=(0.0, 0.0, 0.0, 0.0)):
        v = tuple(float(a) for a in v)
        if len(v) != 4:
            raise ValueError("CoordinateD needs 4 entries")
        return super().__new__(cls, v)


class Geometry:
    """Single-node lattice geometry (qlat Geometry, core.h:92-275)."""

    def __init__(self, total_site):
        ts = Coordinate(total_site)
        for n in ts:
            if n < 2 or n % 2:
                raise ValueError("every lattice extent must be even and >= 2")
        self._ts = ts

    def total_site(self):
        return self._ts

    def total_volume(self):
        return int(np.prod(self._ts))

    def local_volume(self):
        return self.total_volume()

    def show(self):
        return f"Geometry({list(self._ts)})"

    def __eq__(self, other):
        return isinstance(other, Geometry) and other._ts == self._ts

    def index_from_coordinate(self, x):
        X, Y, Z, T = self._ts
        x = [int(a) % n for a, n in zip(x, self._ts)]
        return ((x[3] * Z + x[2]) * Y + x[1]) * X + x[0]

    def coordinate_from_index(self, index):
        X, Y, Z, T = self._ts
        x0 = index % X
        index //= X
        x1 = index % Y
        index //= Y
        x2 = index % Z
        x3 = index // Z
        return Coordinate((x0, x1, x2, x3))

    def parity_sites(self, eo):
        """Full site indices of parity eo (2 even, 1 odd) in half-index order."""
        X, Y, Z, T = self._ts
        t, z, y, x = np.meshgrid(np.arange(T), np.arange(Z), np.arange(Y), np.arange(X),
                                 indexing="ij")
        par = 2 - (x + y + z + t).reshape(-1) % 2
        return np.nonzero(par == eo)[0]


class _Field:
    shape_tail = ()

    def __init__(self, geo=None, data=None):
        self.geo = geo
        self.data = data

    def qnorm(self):
        """Sum of |entries|^2 (field.h:170-195)."""
        return float(np.vdot(self.data, self.data).real)

    def copy(self):
        f = self.__class__.__new__(self.__class__)
        f.__dict__.update(self.__dict__)
        f.data = self.data.copy()
        return f

    def set_zero(self):
        self.data[...] = 0

    def __iadd__(self, o):
        self.data += o.data
        return self

    def __isub__(self, o):
        self.data -= o.data
        return self

    def __imul__(self, c):
        self.data *= c
        return self

    def set_rand(self, rng, upper=1.0, lower=0.0, device=False):
        """Field.set_rand (field_types.pyx.in:194-195 -> set_u_rand,
        core.h:1584-1603): uniform reals in [lower, upper) from rng, site g
        drawing from rng.newtype(g); device=True draws them on the GPU."""
        from .rng import RngState, set_u_rand, set_u_rand_device

        if not isinstance(rng, RngState):
            raise TypeError("set_rand: rng must be an RngState")
        (set_u_rand_device if device else set_u_rand)(self.data, rng, upper, lower, self._rng_sites())
        return self

    def _rng_sites(self):
        """Global site index of each row: None for a full field, the sites of
        the parity for a half field (eo geometry, core.h:248-262)."""
        eo = getattr(self, "eo", 0)
        return None if eo == 0 else self.geo.parity_sites(eo)

    def set_rand_g(self, rng, center=0.0, sigma=1.0, device=False):
        """Field.set_rand_g (field_types.pyx.in:198-199 -> set_g_rand,
        core.h:1619-1640); device=True draws on the GPU."""
        from .rng import RngState, set_g_rand, set_g_rand_device

        if not isinstance(rng, RngState):
            raise TypeError("set_rand_g: rng must be an RngState")
        (set_g_rand_device if device else set_g_rand)(self.data, rng, center, sigma, self._rng_sites())
        return self


class GaugeField(_Field):
    """GaugeField = FieldM<ColorMatrix, 4> (core.h:1187)."""

    def __init__(self, geo=None):
        super().__init__(geo, None)
        if geo is not None:
            self.data = np.zeros((geo.total_volume(), 4, 3, 3), dtype=np.complex128)

    def set_unit(self):
        self.data[...] = 0
        for i in range(3):
            self.data[:, :, i, i] = 1.0
        return self

    def set_rand(self, seed, sigma):
        from .gauge import rand_gauge_links

        self.data = rand_gauge_links(self.geo.total_site(), sigma, seed)
        return self

    def twist_boundary_at_boundary(self, lmom=-0.5, mu=3):
        from .gauge import twist_boundary_at_boundary

        twist_boundary_at_boundary(self.data, self.geo.total_site(), lmom, mu)
        return self

    def show_info(self):
        from .gauge import plaquette

        return {"plaq": plaquette(self.data, self.geo.total_site())}

    def save(self, path):
        """NERSC format (save_gauge_field, qcd.h:143-179); returns bytes written."""
        from .gauge_io import save_gauge_field

        return save_gauge_field(self, path)

    def load(self, path):
        """NERSC format (load_gauge_field, qcd.h:181-270); returns bytes read."""
        from .gauge_io import load_gauge_field

        return load_gauge_field(self, path)


class FermionField4d(_Field):
    def __init__(self, geo=None):
        super().__init__(geo, None)
        if geo is not None:
            self.data = np.zeros((geo.total_volume(), 12), dtype=np.complex128)


class FermionField5d(_Field):
    """FermionField5d (core.h:1203), multiplicity Ls.  eo = 0 for a full
    field, 1 (odd) or 2 (even) for a half field."""

    def __init__(self, geo=None, ls=None, eo=0):
        super().__init__(geo, None)
        self.eo = eo
        self.ls = ls
        if geo is not None and ls is not None:
            nsite = geo.total_volume() if eo == 0 else geo.total_volume() // 2
            self.data = np.zeros((nsite, ls, 12), dtype=np.complex128)


class Prop(_Field):
    """Propagator4d: WilsonMatrix per site, [row = spin*3+colour][column]."""

    def __init__(self, geo=None):
        super().__init__(geo, None)
        if geo is not None:
            self.data = np.zeros((geo.total_volume(), 12, 12), dtype=np.complex128)

    def set_rand(self, rng, upper=1.0, lower=0.0, device=False):
        """With an RngState: the reference's set_u_rand (on the GPU with
        device=True).  With an integer: complex Gaussian entries from numpy's
        PCG64 seeded with it."""
        from .rng import RngState

        if isinstance(rng, RngState):
            return super().set_rand(rng, upper, lower, device)
        g = np.random.Generator(np.random.PCG64(rng))
        self.data = (g.standard_normal(self.data.shape)
                     + 1j * g.standard_normal(self.data.shape))
        return self


class FieldComplexD(_Field):
    """FieldM<ComplexD, multiplicity> (field.h): (V, multiplicity) complex128,
    e.g. the U(1) noise of set_rand_vol_u1_src."""

    def __init__(self, geo=None, multiplicity=1):
        super().__init__(geo, None)
        if geo is not None:
            self.data = np.zeros((geo.total_volume(), multiplicity), dtype=np.complex128)


def get_half_fermion(ff, eo):
    """Half field of a full 5D field (dslash.cpp:542-561)."""
    assert ff.eo == 0
    h = FermionField5d.__new__(FermionField5d)
    h.geo, h.ls, h.eo = ff.geo, ff.ls, eo
    h.data = np.ascontiguousarray(ff.data[ff.geo.parity_sites(eo)])
    return h


def set_half_fermion(ff, half, eo):
    """Write a half field into a full 5D field (dslash.cpp:563-580)."""
    assert ff.eo == 0 and half.eo == eo
    ff.data[ff.geo.parity_sites(eo)] = half.data


def set_point_src_fermion_field(ff, xg, cs, value=1.0):
    """qcd-prop.cpp:61-73."""
    ff.data[...] = 0
    ff.data[ff.geo.index_from_coordinate(xg), cs] = value
    return ff


def fermion_field_5d_from_4d(ff4d, ls, upper=0, lower=None):
    """Chiral embedding (qcd-prop.h:96-116): spins 0,1 at s=upper, spins 2,3
    at s=lower (default Ls-1)."""
    if lower is None:
        lower = ls - 1
    f = FermionField5d(ff4d.geo, ls, 0)
    f.data[:, upper, :6] = ff4d.data[:, :6]
    f.data[:, lower, 6:] = ff4d.data[:, 6:]
    return f


def fermion_field_4d_from_5d(ff5d, upper=None, lower=0):
    """qcd-prop.h:118-138 with the reference's default use upper=Ls-1,
    lower=0 (qcd-prop.h:164)."""
    if upper is None:
        upper = ff5d.ls - 1
    f = FermionField4d(ff5d.geo)
    f.data[:, :6] = ff5d.data[:, upper, :6]
    f.data[:, 6:] = ff5d.data[:, lower, 6:]
    return f
