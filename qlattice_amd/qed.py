"""QED domain-wall operator and solvers on the GPU (qlat qed.h:433-463,
qed.cpp:32-411), with the reference's argument meaning:

  f5d   (V, Ls, 4) complex128   full 5D fermion field, reference site order
  f4d   (V, 4)     complex128
  gf1   (V, 4)     complex128   U(1) links (the backward link is 1 / gf1(x - mu))
  wick  (T,)       complex128   t_wick_phase_factor_vec, or None

All arithmetic runs in the HIP library (``qlat_dwf_multiply_m_dwf_qed``,
``qlat_dwf_cg_with_m_dwf_qed``, ``qlat_dwf_invert_dwf_qed``).
"""

import numpy as np

from . import _lib


def _c(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _common(total_site, ls, gf1, wick):
    ts = tuple(int(v) for v in total_site)
    V = int(np.prod(ts))
    g = _c(gf1, (V, 4))
    w = None if wick is None or len(wick) == 0 else _c(wick, (ts[3],))
    return ts, V, g, w


def _args(ts, ls, g, mass, m5, w, is_dagger):
    return (ts, int(ls), _lib.ptr(g), float(mass), float(m5), 0 if w is None else _lib.ptr(w), bool(is_dagger))


def multiply_m_dwf_qed(f_in5d, gf1, mass, m5, ls, total_site, wick=None, is_dagger=False):
    """qed.cpp:263-411: M f (M^+ f when is_dagger)."""
    ts, V, g, w = _common(total_site, ls, gf1, wick)
    f = _c(f_in5d, (V, ls, 4))
    out = np.empty_like(f)
    _lib.load().multiply_m_dwf_qed(*_args(ts, ls, g, mass, m5, w, is_dagger), _lib.ptr(out), _lib.ptr(f))
    return out


def cg_with_m_dwf_qed(f_in5d, gf1, mass, m5, ls, total_site, wick=None, is_dagger=False, stop_rsd=1e-8,
                      max_num_iter=50000, guess=None):
    """qed.cpp:178-261: solves M^+ M x = f (M M^+ when is_dagger) from guess
    (zero by default); returns (x, iterations)."""
    ts, V, g, w = _common(total_site, ls, gf1, wick)
    f = _c(f_in5d, (V, ls, 4))
    x = np.zeros_like(f) if guess is None else _c(guess, (V, ls, 4)).copy()
    it = _lib.load().cg_with_m_dwf_qed(*_args(ts, ls, g, mass, m5, w, is_dagger), float(stop_rsd),
                                       int(max_num_iter), _lib.ptr(x), _lib.ptr(f))
    return x, it


def invert_dwf_qed(f_in4d, gf1, mass, m5, ls, total_site, wick=None, is_dagger=False, stop_rsd=1e-8,
                   max_num_iter=50000):
    """qed.cpp:119-150: 4D solution of M out = in (M^+ out = in when
    is_dagger); returns (out, iterations)."""
    ts, V, g, w = _common(total_site, ls, gf1, wick)
    f = _c(f_in4d, (V, 4))
    out = np.empty_like(f)
    it = _lib.load().invert_dwf_qed(*_args(ts, ls, g, mass, m5, w, is_dagger), float(stop_rsd),
                                    int(max_num_iter), _lib.ptr(out), _lib.ptr(f))
    return out, it


def invert_qed(sp_src, gf1, mass, m5, ls, total_site, wick=None, is_dagger=False, stop_rsd=1e-8,
               max_num_iter=50000):
    """qed.cpp:32-73: a 4D spin propagator (V, 4, 4) [site][row][column],
    one invert_dwf_qed per column; returns (solution, total iterations)."""
    sp = _c(sp_src)
    sol = np.zeros_like(sp)
    total = 0
    for i in range(4):
        col, it = invert_dwf_qed(sp[:, :, i], gf1, mass, m5, ls, total_site, wick, is_dagger, stop_rsd,
                                 max_num_iter)
        sol[:, :, i] = col
        total += it
    return sol, total
