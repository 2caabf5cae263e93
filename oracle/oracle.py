"""TEST INFRASTRUCTURE ONLY -- ctypes front end of the CPU FP64 oracle.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module.  The product package never does.

The oracle is ``oracle/dwf_oracle.cpp`` (a C++ restatement of
qlat/qlat/lib/dslash.cpp and qlat/qlat/include/qlat/dslash.h), compiled by
``make -C oracle`` into ``oracle/liboracle_dwf.so``.

Array layouts are the reference's (complex128, array of structures):
  gauge      (V, 4, 3, 3)        [site][mu][row][col]           core.h:1187
  5D full    (V, Ls, 12)         [site][s][spin*3+colour]       core.h:1203
  5D half    (V/2, Ls, 12)       half index = full index // 2   core.h:232-261
  4D fermion (V, 12)
  propagator (V, 12, 12)         [site][row][source column]     qcd-prop.h:55-94
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_dwf.so")
_lib = None

HALF_OPS = {
    "m_e_e": 0,
    "mdag_e_e": 1,
    "m_e_e_inv": 2,
    "mdag_e_e_inv": 3,
    "m_e_o": 4,
    "mdag_e_o": 5,
    "mpc_sym2": 6,
    "mpcdag_sym2": 7,
    "hermop_sym2": 8,
    "hop": 9,
    "hop_dag": 10,
}

FULL_OPS = {
    "m": 0,
    "mdag": 1,
    "d_minus": 2,
    "m_full": 3,
    "m_dwf": 4,
    "m_with_prec_sym2": 5,
    "wilson_d": 6,
}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build():
    """Compile the oracle library with its Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"oracle library {_LIB_PATH} is not built; run `make -C oracle`"
            )
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_inv_create.restype = ctypes.c_void_p
        L.oracle_inv_create.argtypes = [
            _ip, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, _dp, _dp, ctypes.c_int,
        ]
        L.oracle_inv_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_inv_set_params.argtypes = [
            ctypes.c_void_p, ctypes.c_double, ctypes.c_long, ctypes.c_long,
        ]
        L.oracle_inv_stats.argtypes = [ctypes.c_void_p, _lp, _ip, _lp]
        L.oracle_apply_half.restype = ctypes.c_int
        L.oracle_apply_half.argtypes = [
            ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int,
        ]
        L.oracle_apply_full.restype = ctypes.c_int
        L.oracle_apply_full.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.oracle_cg_with_herm_sym_2.restype = ctypes.c_long
        L.oracle_cg_with_herm_sym_2.argtypes = [
            ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_long,
            ctypes.c_int,
        ]
        for name in ("oracle_invert_5d", "oracle_invert_5d_with_guess", "oracle_invert_4d", "oracle_invert_prop"):
            f = getattr(L, name)
            f.restype = ctypes.c_long
            f.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.oracle_find_max_eigen_value_hermop_sym2.restype = ctypes.c_double
        L.oracle_find_max_eigen_value_hermop_sym2.argtypes = [
            ctypes.c_void_p, _dp, ctypes.c_long,
        ]
        L.oracle_dot_product.argtypes = [_dp, _dp, ctypes.c_long, _dp, _dp]
        L.oracle_qnorm.restype = ctypes.c_double
        L.oracle_qnorm.argtypes = [_dp, ctypes.c_long]
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def mobius_bs_cs(ls, mobius_scale=1.0, omega=None):
    """b_s, c_s of FermionAction (fermion-action.h:36-52); zMoebius from
    omega (cqlat/fermion-action.cpp:17-37)."""
    if omega is None:
        bs = np.full(ls, 0.5 * mobius_scale + 0.5, dtype=np.complex128)
        cs = np.full(ls, 0.5 * mobius_scale - 0.5, dtype=np.complex128)
    else:
        om = np.asarray(omega, dtype=np.complex128)
        assert len(om) == ls
        bs = 0.5 * (1.0 / om + 1.0)
        cs = bs - 1.0
    return bs, cs


class OracleInverter:
    """CPU FP64 InverterDomainWall (dslash.h:135-197)."""

    def __init__(self, total_site, gauge, *, mass, ls, m5, mobius_scale=1.0,
                 omega=None, is_multiplying_dminus=True):
        self.total_site = tuple(int(v) for v in total_site)
        self.V = int(np.prod(self.total_site))
        self.ls = int(ls)
        self.mass = float(mass)
        self.m5 = float(m5)
        self.mobius_scale = 0.0 if omega is not None else float(mobius_scale)
        g = np.ascontiguousarray(gauge, dtype=np.complex128)
        assert g.shape == (self.V, 4, 3, 3)
        self.bs, self.cs = mobius_bs_cs(self.ls, mobius_scale, omega)
        site = (ctypes.c_int * 4)(*self.total_site)
        self._h = lib().oracle_inv_create(
            site, _ptr(g), self.ls, self.mass, self.m5, self.mobius_scale,
            _ptr(self.bs), _ptr(self.cs), int(bool(is_multiplying_dminus)),
        )

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().oracle_inv_destroy(h)
            except Exception:  # interpreter shutdown: the library may be gone
                pass
            self._h = None

    def set_params(self, stop_rsd=1e-8, max_num_iter=200,
                   max_mixed_precision_cycle=300):
        lib().oracle_inv_set_params(self._h, stop_rsd, max_num_iter,
                                    max_mixed_precision_cycle)

    def stats(self):
        ti, cy, hc = ctypes.c_long(), ctypes.c_int(), ctypes.c_long()
        lib().oracle_inv_stats(self._h, ctypes.byref(ti), ctypes.byref(cy),
                               ctypes.byref(hc))
        return dict(total_iter=ti.value, cycles=cy.value, hermop_calls=hc.value)

    def half_shape(self):
        return (self.V // 2, self.ls, 12)

    def full_shape(self):
        return (self.V, self.ls, 12)

    def apply_half(self, op, x, eo_in=1):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        assert x.shape == self.half_shape()
        out = np.empty_like(x)
        rc = lib().oracle_apply_half(self._h, HALF_OPS[op], _ptr(out), _ptr(x),
                                     int(eo_in))
        if rc != 0:
            raise RuntimeError(f"oracle_apply_half({op}) failed: {rc}")
        return out

    def apply_full(self, op, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        assert x.shape == self.full_shape()
        out = np.empty_like(x)
        rc = lib().oracle_apply_full(self._h, FULL_OPS[op], _ptr(out), _ptr(x))
        if rc != 0:
            raise RuntimeError(f"oracle_apply_full({op}) failed: {rc}")
        return out

    def cg_with_herm_sym_2(self, src, guess=None, stop_rsd=1e-8,
                           max_num_iter=50000, verbose=False):
        src = np.ascontiguousarray(src, dtype=np.complex128)
        sol = (np.zeros_like(src) if guess is None
               else np.ascontiguousarray(guess, dtype=np.complex128).copy())
        it = lib().oracle_cg_with_herm_sym_2(self._h, _ptr(sol), _ptr(src),
                                            stop_rsd, max_num_iter, int(verbose))
        return sol, it

    def invert_5d(self, src, verbose=False):
        src = np.ascontiguousarray(src, dtype=np.complex128)
        out = np.empty_like(src)
        it = lib().oracle_invert_5d(self._h, _ptr(out), _ptr(src), int(verbose))
        return out, it

    def invert_5d_with_guess(self, src, guess, verbose=False):
        """invert_with_cg_with_guess (dslash.h:532-573); returns (solution, iterations)."""
        src = np.ascontiguousarray(src, dtype=np.complex128)
        out = np.array(guess, dtype=np.complex128, order="C", copy=True)
        assert out.shape == src.shape
        it = lib().oracle_invert_5d_with_guess(self._h, _ptr(out), _ptr(src), int(verbose))
        return out, it

    def invert_4d(self, src, verbose=False):
        src = np.ascontiguousarray(src, dtype=np.complex128)
        assert src.shape == (self.V, 12)
        out = np.empty_like(src)
        it = lib().oracle_invert_4d(self._h, _ptr(out), _ptr(src), int(verbose))
        return out, it

    def invert_prop(self, src, verbose=False):
        src = np.ascontiguousarray(src, dtype=np.complex128)
        assert src.shape == (self.V, 12, 12)
        out = np.empty_like(src)
        it = lib().oracle_invert_prop(self._h, _ptr(out), _ptr(src), int(verbose))
        return out, it

    def find_max_eigen_value_hermop_sym2(self, start, max_iter=100):
        start = np.ascontiguousarray(start, dtype=np.complex128)
        return lib().oracle_find_max_eigen_value_hermop_sym2(
            self._h, _ptr(start), int(max_iter))


def qnorm(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return lib().oracle_qnorm(_ptr(a), a.size // 12)


def dot_product(a, b):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().oracle_dot_product(_ptr(a), _ptr(b), a.size // 12, ctypes.byref(re),
                             ctypes.byref(im))
    return complex(re.value, im.value)


def num_threads():
    return lib().oracle_num_threads()
