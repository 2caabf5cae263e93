"""RngState and the field fills built on it (qlat-utils rng-state.h:14-93,
rng-state.cpp:125-283; set_u_rand / set_g_rand, qlat core.h:1584-1640).

The reference's generator is SHA-256 over a message that grows with every
split: a state holds the SHA-256 midstate of its message (a multiple of 64
bytes), ``split(s)`` appends ``"[index] {s}"`` (``"[type,index] {s}"`` after
``newtype``) padded with spaces to a multiple of 64 bytes, and a draw that
finds no cached value hashes the message followed by ``"[index]"`` (or
``"[type,index]"``) with standard SHA-256 padding, returns the digest's last
two 32-bit words (big-endian, high word first) as a 64-bit integer and
caches words 4-5, 2-3 and 0-1 for the next three draws.  Every draw
increments the index.  hashlib's SHA-256 object keeps exactly that midstate,
so the state here carries one.

Host code: fills of test-size fields (a draw per real number); the
operator and solvers never call it.
"""

import hashlib
import math
import struct

import numpy as np

_ULONG_MAX = (1 << 64) - 1
_FAC = 1.0 / (256.0 * 256.0 * 256.0 * 256.0) / (256.0 * 256.0 * 256.0 * 256.0)


class RngState:
    """RngState (rng-state.h:31-93): RngState(), RngState(seed: str | int),
    split(sindex), newtype(type), and the draws rand_gen, u_rand_gen,
    g_rand_gen as methods."""

    def __init__(self, seed=None):
        self._sha = hashlib.sha256()
        self._msg = b""
        self.num_bytes = 0
        self.type = _ULONG_MAX
        self.index = 0
        self.cache = [0, 0, 0]
        self.cache_avail = 0
        self.gaussian = 0.0
        self.gaussian_avail = False
        if seed is not None:
            self._split_from(self, str(seed))

    def _split_from(self, rs0, sindex):
        """split_rng_state (rng-state.cpp:148-177)."""
        if rs0.type == _ULONG_MAX:
            data = f"[{rs0.index}] {{{sindex}}}"
        else:
            data = f"[{rs0.type},{rs0.index}] {{{sindex}}}"
        raw = data.encode()
        n_blocks = (len(raw) - 1) // 64 + 1
        raw = raw.ljust(n_blocks * 64, b" ")
        sha = rs0._sha.copy()
        sha.update(raw)
        self._sha = sha
        self._msg = rs0._msg + raw
        self.num_bytes = rs0.num_bytes + n_blocks * 64
        self.type = _ULONG_MAX
        self.index = 0
        self.cache = [0, 0, 0]
        self.cache_avail = 0
        self.gaussian = 0.0
        self.gaussian_avail = False

    def copy(self):
        rs = RngState.__new__(RngState)
        rs.__dict__.update(self.__dict__)
        rs._sha = self._sha.copy()
        rs.cache = list(self.cache)
        return rs

    def split(self, sindex):
        """RngState(rs0, sindex): the child stream named str(sindex)."""
        rs = RngState.__new__(RngState)
        rs._split_from(self, str(sindex))
        return rs

    def newtype(self, type_):
        """A copy whose draws hash "[type,index]" (set_type, rng-state.cpp:20-27)."""
        if self.type != _ULONG_MAX:
            raise ValueError("RngState.newtype: the type is already set")
        type_ = int(type_)
        if not 0 <= type_ < _ULONG_MAX:
            raise ValueError("RngState.newtype: type out of range")
        rs = self.copy()
        rs.type = type_
        rs.cache_avail = 0
        rs.gaussian_avail = False
        return rs

    def _hash_words(self, index):
        sha = self._sha.copy()
        if self.type == _ULONG_MAX:
            sha.update(f"[{index}]".encode())
        else:
            sha.update(f"[{self.type},{index}]".encode())
        return struct.unpack(">8I", sha.digest())

    def rand_gen(self):
        """rng-state.cpp:186-212: the next 64-bit integer."""
        self.index += 1
        if self.cache_avail > 0:
            self.cache_avail -= 1
            r = self.cache[self.cache_avail]
            self.cache[self.cache_avail] = 0
            return r
        w = self._hash_words(self.index)
        self.cache = [w[0] << 32 | w[1], w[2] << 32 | w[3], w[4] << 32 | w[5]]
        self.cache_avail = 3
        return w[6] << 32 | w[7]

    def u_rand_gen(self, upper=1.0, lower=0.0):
        """rng-state.cpp:214-220: rand_gen * 2^-64 * (upper - lower) + lower."""
        return float(self.rand_gen()) * _FAC * (upper - lower) + lower

    def g_rand_gen(self, center=0.0, sigma=1.0):
        """rng-state.cpp:222-258: polar Box-Muller on u_rand_gen(1, -1)."""
        self.index += 1
        if self.gaussian_avail:
            self.gaussian_avail = False
            return self.gaussian * sigma + center
        num_try = 1
        while True:
            v1 = self.u_rand_gen(1.0, -1.0)
            v2 = self.u_rand_gen(1.0, -1.0)
            rsq = v1 * v1 + v2 * v2
            num_try += 1
            if not (num_try < 10000 and (rsq >= 1.0 or rsq == 0)):
                break
        if num_try > 9999:
            return 1e10
        fac = math.sqrt(-2.0 * math.log(rsq) / rsq)
        self.gaussian = v1 * fac
        self.gaussian_avail = True
        return v2 * fac * sigma + center

    def u_rand_u64(self, n):
        """n consecutive rand_gen values as a uint64 array (the same stream as
        n calls of rand_gen; whole hashes at a time)."""
        out = np.empty(n, np.uint64)
        k = 0
        while k < n and self.cache_avail > 0:
            out[k] = self.rand_gen()
            k += 1
        while k < n:
            self.index += 1
            w = self._hash_words(self.index)
            vals = [w[6] << 32 | w[7], w[4] << 32 | w[5], w[2] << 32 | w[3], w[0] << 32 | w[1]]
            take = min(4, n - k)
            out[k:k + take] = vals[:take]
            k += take
            self.index += take - 1
            self.cache = [vals[3], vals[2], vals[1]]
            self.cache_avail = 4 - take
            for j in range(3 - self.cache_avail):  # drawn cache entries are cleared
                self.cache[2 - j] = 0
        return out


_SHA256_K = (
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2)
_SHA256_H0 = (0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19)


def sha256_compress(h, block):
    """One SHA-256 compression (FIPS 180-4) of a 64-byte block onto the
    state h (8 words); returns the new state."""
    m = 0xFFFFFFFF

    def rotr(x, n):
        return ((x >> n) | (x << (32 - n))) & m

    w = list(struct.unpack(">16I", block))
    for i in range(16, 64):
        s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3)
        s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10)
        w.append((w[i - 16] + s0 + w[i - 7] + s1) & m)
    a, b, c, d, e, f, g, k = h
    for i in range(64):
        t1 = (k + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + _SHA256_K[i] + w[i]) & m
        t2 = ((rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c))) & m
        a, b, c, d, e, f, g, k = (t1 + t2) & m, a, b, c, (d + t1) & m, e, f, g
    return tuple((x + y) & m for x, y in zip(h, (a, b, c, d, e, f, g, k)))


def midstate(rs):
    """The SHA-256 state of rs after its message (a whole number of blocks)."""
    h = _SHA256_H0
    for i in range(0, len(rs._msg), 64):
        h = sha256_compress(h, rs._msg[i:i + 64])
    return h


def _fill_device(gaussian, data, rs, a, b, sites=None):
    from . import _lib

    if rs.type != _ULONG_MAX:
        raise ValueError("newtype needs a state without a type")
    reals = _site_reals(data)
    V, n = reals.shape
    # the device fill draws for sites 0 .. nfill-1; rows of a half field take
    # their sites' draws from it
    nfill = V if sites is None else int(np.max(sites)) + 1
    out = np.empty((nfill, n), np.float64)
    _lib.load().rng_fill(gaussian, midstate(rs), len(rs._msg), rs.index, nfill, n, float(a), float(b),
                         _lib.addr(out, np.float64))
    reals[...] = out if sites is None else out[np.asarray(sites)]
    return data


def set_u_rand_device(data, rs, upper=1.0, lower=-1.0, sites=None):
    """set_u_rand on the GPU (qlat_dwf_rng_fill_uniform): the same numbers
    as set_u_rand, one SHA-256 block per four draws on the device."""
    return _fill_device(False, data, rs, upper, lower, sites)


def set_g_rand_device(data, rs, center=0.0, sigma=1.0, sites=None):
    """set_g_rand on the GPU (qlat_dwf_rng_fill_gaussian), one thread per
    site; equal to set_g_rand up to the last bit of the device logarithm."""
    return _fill_device(True, data, rs, center, sigma, sites)


def _site_reals(data):
    """The site-major real view of `data` that the fills write through; a
    view, never a copy, so `data` must be a C-contiguous array."""
    a = data
    if not isinstance(a, np.ndarray) or not a.flags.c_contiguous:
        raise ValueError("the RNG fills write in place: pass a C-contiguous numpy array")
    V = a.shape[0]
    if np.iscomplexobj(a):
        return a.view(np.float64).reshape(V, -1)
    return a.reshape(V, -1)


def set_u_rand(data, rs, upper=1.0, lower=-1.0, sites=None):
    """set_u_rand (core.h:1584-1603) on a (V, ...) float64 / complex128 array
    in reference site order: row i is the site with global index
    g = sites[i] (g = i when sites is None; a half field passes the global
    indices of its parity, as geo.coordinate_from_index of an eo geometry
    gives, core.h:248-262), whose reals are drawn in memory order from
    rs.newtype(g); returns data."""
    reals = _site_reals(data)
    n = reals.shape[1]
    scale = upper - lower
    for i in range(reals.shape[0]):
        g = i if sites is None else int(sites[i])
        u = rs.newtype(g).u_rand_u64(n).astype(np.float64)
        reals[i] = u * _FAC * scale + lower
    return data


def set_g_rand(data, rs, center=0.0, sigma=1.0, sites=None):
    """set_g_rand (core.h:1619-1640): Gaussian reals, row i from
    rs.newtype(sites[i]) (rs.newtype(i) when sites is None)."""
    reals = _site_reals(data)
    for i in range(reals.shape[0]):
        rsi = rs.newtype(i if sites is None else int(sites[i]))
        reals[i] = [rsi.g_rand_gen(center, sigma) for _ in range(reals.shape[1])]
    return data


def get_data_sig(data, rs):
    """Field.get_data_sig (qlat field_base.pyx:111-146): sum over the
    (V, ...) array of each number times a real u in [-1, 1), the u field
    filled by set_u_rand(rs, 1, -1) with one real per number (per complex
    number for complex data)."""
    a = np.asarray(data)
    flat = a.reshape(a.shape[0], -1)
    fu = np.zeros(flat.shape, np.float64)
    set_u_rand(fu, rs, 1.0, -1.0)
    return np.sum(flat * fu)


def get_data_sig_arr(data, rs, sig_len):
    """get_data_sig_arr (qlat_utils utils.py:437-451): get_data_sig with
    rs.split(str(i)) for i < sig_len."""
    return np.array([get_data_sig(data, rs.split(f"{i}")) for i in range(sig_len)])
