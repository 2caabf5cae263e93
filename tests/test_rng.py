"""RngState (rng-state.cpp:125-258) and set_u_rand (core.h:1584-1603).

Pinned by the reference's own goldens: examples-py/free-invert.py fills its
random source with Prop.set_rand(RngState("seed").split("src_r")), and the
qnorms of the analytic free propagator (oracle/freefield.py), of the Ls 16
domain-wall CG solution and of the zMobius CG solution of that source
(the oracle's solver) are reproduced at the golden tolerances.  A wrong
split string, padding, type/index format, cache order or u64 -> double
scaling changes the source and misses them by far more than 1e-10.
"""

import hashlib
import json
import os
import struct

import numpy as np
import pytest

import qlattice_amd as q
from oracle import freefield
from oracle import oracle as O
from tests.helpers import ZMOBIUS_OMEGA, qlat_check, unit_gauge

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "free_invert.json")))


def words(msg):
    return struct.unpack(">8I", hashlib.sha256(msg).digest())


def test_draws_are_sha256_of_the_message():
    w = words(b"[1]")  # a fresh state: the message is just "[index]"
    rs = q.RngState()
    assert [rs.rand_gen() for _ in range(4)] == [w[6] << 32 | w[7], w[4] << 32 | w[5], w[2] << 32 | w[3],
                                                 w[0] << 32 | w[1]]
    assert rs.index == 4 and rs.cache_avail == 0
    seed = b"[0] {seed}".ljust(64, b" ")
    sub = b"[0] {src_r}".ljust(64, b" ")
    w = words(seed + sub + b"[17,1]")
    assert q.RngState("seed").split("src_r").newtype(17).rand_gen() == w[6] << 32 | w[7]
    # a split taken after draws carries the index; after newtype the type too
    rs = q.RngState(7).newtype(3)
    rs.rand_gen()
    long = "x" * 70  # two 64-byte blocks
    w = words(b"[0] {7}".ljust(64, b" ") + f"[3,1] {{{long}}}".encode().ljust(128, b" ") + b"[1]")
    assert rs.split(long).rand_gen() == w[6] << 32 | w[7]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 9])
def test_block_draws_match_single_draws(n):
    a = q.RngState("a").newtype(5)
    b = a.copy()
    got = a.u_rand_u64(n)
    ref = [b.rand_gen() for _ in range(n)]
    assert [int(v) for v in got] == ref
    assert (a.index, a.cache, a.cache_avail) == (b.index, b.cache, b.cache_avail)
    assert a.rand_gen() == b.rand_gen()


def test_u_rand_gen_scaling():
    rs = q.RngState("u")
    u = rs.copy().rand_gen()
    assert rs.u_rand_gen(3.0, -1.0) == float(u) * 2.0 ** -64 * 4.0 - 1.0


def test_g_rand_gen_pairs_and_moments():
    rs = q.RngState("g")
    ref = rs.copy()
    x0, x1 = rs.g_rand_gen(), rs.g_rand_gen()
    ref.index += 1  # g_rand_gen counts a draw of its own before its uniform draws
    while True:  # the polar method by hand on the same stream
        v1, v2 = ref.u_rand_gen(1.0, -1.0), ref.u_rand_gen(1.0, -1.0)
        rsq = v1 * v1 + v2 * v2
        if 0 < rsq < 1:
            break
    fac = np.sqrt(-2.0 * np.log(rsq) / rsq)
    assert x0 == v2 * fac and x1 == v1 * fac
    xs = np.array([rs.g_rand_gen(1.0, 2.0) for _ in range(20000)])
    assert abs(xs.mean() - 1.0) < 0.05 and abs(xs.std() - 2.0) < 0.05


@pytest.fixture(scope="module")
def random_source():
    site = (4, 4, 4, 8)
    src = q.Prop(q.Geometry(site)).set_rand(q.RngState("seed").split("src_r"))
    assert src.data.real.min() >= 0 and src.data.real.max() < 1 and src.data.imag.min() >= 0
    return site, src.data


def test_golden_random_source_free_field(random_source):
    site, src = random_source
    sol = freefield.free_invert(src, site, 0.05, 1.0)
    v, eps = GOLD["random_source"]["sol_qnorm_analytic"]
    assert qlat_check(np.vdot(sol, sol).real, v, eps)


def test_golden_random_source_cg(random_source):
    """free-invert.py:25-45, :69-84 with the random source (oracle CG)."""
    site, src = random_source
    U = unit_gauge(512)
    sol = freefield.free_invert(src, site, 0.05, 1.0)
    sol1, _ = O.OracleInverter(site, U, mass=0.05, ls=16, m5=1.0).invert_prop(src)
    solz, _ = O.OracleInverter(site, U, mass=0.05, ls=12, m5=1.0, omega=ZMOBIUS_OMEGA).invert_prop(src)
    g = GOLD["random_source"]
    for name, val in [("sol1_qnorm_dwf_cg", np.vdot(sol1, sol1).real),
                      ("sol_diff_qnorm", np.vdot(sol1 - sol, sol1 - sol).real),
                      ("sol_z_qnorm_zmobius_cg", np.vdot(solz, solz).real),
                      ("sol_z_diff_qnorm", np.vdot(solz - sol, solz - sol).real)]:
        v, eps = g[name]
        assert qlat_check(val, v, eps), (name, val, v)


@pytest.mark.parametrize("mult", [1, 12])
def test_set_rand_vol_u1(mult):
    """qcd-prop.cpp:236-258: unit-modulus noise whose phase at site g,
    component m is the m-th u_rand_gen(PI, -PI) of rs.newtype(g)."""
    geo = q.Geometry((2, 2, 2, 4))
    rs = q.RngState("u1")
    fu1 = q.set_rand_vol_u1(q.FieldComplexD(geo, mult), geo, mult, rs.split("noise"))
    assert fu1.data.shape == (32, mult)
    assert np.allclose(np.abs(fu1.data), 1.0, atol=1e-15, rtol=0)
    for g in (0, 7, 31):
        rst = rs.split("noise").newtype(g)
        ph = [rst.u_rand_gen(np.pi, -np.pi) for _ in range(mult)]
        assert np.allclose(np.angle(fu1.data[g]), ph, atol=1e-14, rtol=0)


def test_midstate_matches_hashlib():
    """The host SHA-256 compression (used to hand a state to the GPU fill)
    continued with a final block equals hashlib's digest of the message."""
    from qlattice_amd import rng as R

    rs = q.RngState("seed").split("a" * 70).split(3)
    tail = b"[12,5]"
    bits = (len(rs._msg) + len(tail)) * 8
    block = tail + b"\x80" + bytes(55 - len(tail)) + bits.to_bytes(8, "big")
    h = R.sha256_compress(R.midstate(rs), block)
    assert h == struct.unpack(">8I", hashlib.sha256(rs._msg + tail).digest())
    assert len(rs._msg) == rs.num_bytes == 64 * 4


def test_device_fill_rejects_bad_arguments():
    """qlat_dwf_rng_fill_uniform validates its arguments before any device
    work (no GPU needed to see the errors)."""
    import ctypes

    from qlattice_amd import _lib

    L = _lib.load()
    mid = [0] * 8
    out = np.zeros(8)
    dp = _lib.addr(out, np.float64)
    with pytest.raises(q.QlatDwfError, match="n_per_site"):
        L.rng_fill(False, mid, 64, 0, 8, 0, 1.0, 0.0, dp)
    with pytest.raises(q.QlatDwfError, match="64-byte"):
        L.rng_fill(False, mid, 63, 0, 8, 1, 1.0, 0.0, dp)
    L.rng_fill(False, mid, 64, 0, 0, 1, 1.0, 0.0, dp)  # nothing to fill
    # a null midstate pointer is an error of the C entry point
    so = ctypes.CDLL(_lib.LIB_PATH)
    so.qlat_dwf_rng_fill_uniform.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    assert so.qlat_dwf_rng_fill_uniform(None, 64, 0, 8, 1, 1.0, 0.0, dp) != 0
