"""CPU checks of the host layer and the C-ABI library (no GPU calls)."""

import ctypes
import os
import re

import numpy as np
import pytest

import qlattice_amd as q
from qlattice_amd import _lib
from qlattice_amd.gauge import plaquette, rand_gauge_links, twist_boundary_at_boundary

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "qlat_dwf.h")).read()
    return sorted(set(re.findall(r"\b(qlat_dwf_\w+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    # the shared object itself, opened only to list its exported symbols
    so = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(so, s), s
    assert sorted(_lib.EXPORTED) == syms
    assert _lib.load().abi_version() == 4


def test_binding_declares_every_header_symbol():
    """The Cython module declares (and so is compiled against) every entry
    point of the header."""
    text = open(os.path.join(ROOT, "qlattice_amd", "_cdwf.pyx")).read()
    declared = set(re.findall(r"\b(qlat_dwf_\w+)\s*\(", text))
    assert set(header_symbols()) <= declared


def test_null_handle_is_an_error_not_a_crash():
    L = _lib.load()
    with pytest.raises(q.QlatDwfError, match="null"):
        L.inverter_set_params(0, 1e-8, 10, 10)
    # the C entry point itself returns a status and sets the message
    so = ctypes.CDLL(_lib.LIB_PATH)
    so.qlat_dwf_inverter_set_params.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_long, ctypes.c_long]
    so.qlat_dwf_last_error.restype = ctypes.c_char_p
    assert so.qlat_dwf_inverter_set_params(None, 1e-8, 10, 10) == -1
    assert b"null" in so.qlat_dwf_last_error()


def test_checking_invert_flag_round_trip():
    L = _lib.load()
    assert L.is_checking_invert()  # on by default, like is_checking_invert() (dslash.h:11-16)
    L.set_checking_invert(False)
    assert not L.is_checking_invert()
    L.set_checking_invert(True)
    assert L.is_checking_invert()


def test_fermion_action_mobius_and_zmobius():
    fa = q.FermionAction(mass=0.01, ls=12, m5=1.8, mobius_scale=2.0)
    assert np.allclose(fa.bs, 1.5) and np.allclose(fa.cs, 0.5)
    assert fa.omega() is None and fa.mobius_scale() == 2.0
    from tests.helpers import ZMOBIUS_OMEGA

    fz = q.FermionAction(mass=0.05, ls=12, m5=1.0, omega=ZMOBIUS_OMEGA)
    assert np.allclose(fz.omega(), ZMOBIUS_OMEGA)
    assert np.allclose(fz.bs - fz.cs, 1.0)


def test_geometry_parity_and_half_index():
    geo = q.Geometry((4, 2, 2, 4))
    for eo in (1, 2):
        idx = geo.parity_sites(eo)
        assert np.all(idx // 2 == np.arange(len(idx)))  # half index = full / 2
        for i in idx[:7]:
            x = geo.coordinate_from_index(int(i))
            assert 2 - sum(x) % 2 == eo
    with pytest.raises(ValueError):
        q.Geometry((3, 4, 4, 4))


def test_rand_gauge_is_special_unitary():
    U = rand_gauge_links((4, 4, 2, 2), 0.3, 5).reshape(-1, 3, 3)
    eye = np.eye(3)
    assert np.allclose(U @ np.conj(np.swapaxes(U, 1, 2)), eye, atol=1e-13)
    assert np.allclose(np.linalg.det(U), 1.0, atol=1e-13)
    p = plaquette(rand_gauge_links((4, 4, 4, 4), 0.3, 5), (4, 4, 4, 4))
    assert 0.3 < p < 0.99
    U2 = rand_gauge_links((4, 4, 2, 2), 0.3, 5)
    assert np.array_equal(U2.reshape(-1, 3, 3), U)  # seeded


def test_twist_boundary_flips_last_time_slice_only():
    site = (2, 2, 2, 4)
    U = rand_gauge_links(site, 0.2, 1)
    W = twist_boundary_at_boundary(U.copy(), site, -0.5, 3)
    t = np.arange(16) // 8 if False else np.arange(32) // 8
    changed = ~np.isclose(W, U).all(axis=(2, 3))
    assert changed[t == 3, 3].all() and not changed[t != 3].any() and not changed[:, :3].any()
    assert np.allclose(W[t == 3, 3], -U[t == 3, 3])


def test_point_source_fermion_field_and_chiral_embedding():
    geo = q.Geometry((2, 2, 2, 2))
    ff = q.set_point_src_fermion_field(q.FermionField4d(geo), q.Coordinate([1, 0, 1, 0]), 7, 2.0)
    i = geo.index_from_coordinate((1, 0, 1, 0))
    assert ff.data[i, 7] == 2.0 and ff.qnorm() == 4.0
    f4 = q.FermionField4d(geo)
    f4.data[:] = np.arange(12)[None, :] + 1
    f5 = q.fermion_field_5d_from_4d(f4, 6)
    assert np.array_equal(f5.data[:, 0, :6], f4.data[:, :6])
    assert np.array_equal(f5.data[:, 5, 6:], f4.data[:, 6:])
    assert not f5.data[:, 1:5].any()
    back = q.fermion_field_4d_from_5d(f5, upper=0, lower=5)
    assert np.array_equal(back.data, f4.data)


def test_half_fermion_round_trip():
    geo = q.Geometry((4, 2, 2, 2))
    f = q.FermionField5d(geo, 3, 0)
    f.data[:] = np.random.default_rng(0).standard_normal(f.data.shape)
    e = q.get_half_fermion(f, 2)
    o = q.get_half_fermion(f, 1)
    g = q.FermionField5d(geo, 3, 0)
    q.set_half_fermion(g, e, 2)
    q.set_half_fermion(g, o, 1)
    assert np.array_equal(g.data, f.data)
