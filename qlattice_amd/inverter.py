"""Domain-wall inverter and operators on the MI355X, qlat-compatible surface.

``InverterDomainWall`` mirrors qlat's class of the same name
(qlat/qlat/inverter.py:75-129): construct with ``gf=`` and ``fa=``, apply
with ``inv * prop`` (a ``Prop`` or a list of them), and get/set
``stop_rsd``, ``max_num_iter`` and ``max_mixed_precision_cycle``.  The
module-level ``multiply_*``, ``cg_with_herm_sym_2``, ``invert`` and
``find_max_eigen_value_hermop_sym2`` functions take the arguments of the C++
overloads in qlat/qlat/include/qlat/dslash.h:218-588.

Everything is computed by the HIP library through the C ABI of
include/qlat_dwf.h, reached through the Cython binding (_cdwf.pyx); this
module only marshals numpy arrays.
"""

import numpy as np

from . import _lib
from .fermion_action import FermionAction
from .fields import (FermionField4d, FermionField5d, GaugeField, Prop, fermion_field_4d_from_5d,
                     fermion_field_5d_from_4d, get_half_fermion, set_half_fermion)

SOLVERS = {"fp64": _lib.QLAT_DWF_SOLVER_FP64, "mixed": _lib.QLAT_DWF_SOLVER_MIXED_FP32,
           "mixed16": _lib.QLAT_DWF_SOLVER_MIXED_FP16}


class Inverter:
    pass


def free_invert(prop_src, mass, m5=1.0, momtwist=None):
    """free_invert (qcd-prop.h:347-359): the infinite-Ls domain-wall
    propagator of the free field applied to prop_src, on the GPU."""
    geo = prop_src.geo
    src = np.ascontiguousarray(prop_src.data, dtype=np.complex128)
    sol = Prop(geo)
    _lib.load().free_invert(geo.total_site(), float(mass), float(m5),
                            None if momtwist is None else [float(v) for v in momtwist],
                            _lib.ptr(sol.data), _lib.ptr(src))
    return sol


class InverterDwfFreeField(Inverter):
    """inverter.py:26-71: the analytic free-field domain-wall inverter."""

    def __init__(self, *, mass, m5=1.0, momtwist=None, qtimer=None):
        self.mass = float(mass)
        self.m5 = float(m5)
        self.momtwist = momtwist
        self.timer = qtimer

    def __mul__(self, prop_src):
        if isinstance(prop_src, Prop):
            if self.timer is not None:
                self.timer.start()
            sol = free_invert(prop_src, self.mass, self.m5, self.momtwist)
            if self.timer is not None:
                self.timer.stop()
            return sol
        if isinstance(prop_src, list):
            return [self * p for p in prop_src]
        raise Exception("InverterDwfFreeField")


class InverterDomainWall(Inverter):
    """InverterDomainWall (dslash.h:135-197; inverter.py:75-129).

    solver: "fp64" runs the reference algorithm (invert_with_cg with the
    sym2 odd preconditioning and the restart cycle loop, dslash.h:440-530)
    in FP64; "mixed" runs the inner CG in FP32 with FP64 reliable updates
    (the role of a back-end's cg_with_herm_sym_2 override, grid.h:460-484);
    "mixed16" is "mixed" with the Krylov vectors stored in FP16 with a
    per-5D-site scale.
    """

    def __init__(self, *, gf, fa, qtimer=None, solver="fp64", reliable_delta=0.01):
        assert isinstance(gf, GaugeField)
        assert isinstance(fa, FermionAction)
        L = _lib.load()
        self.geo_ = gf.geo
        self.fa = fa
        self.timer = qtimer
        self._bs = np.ascontiguousarray(fa.bs, dtype=np.complex128)
        self._cs = np.ascontiguousarray(fa.cs, dtype=np.complex128)
        gauge = np.ascontiguousarray(gf.data, dtype=np.complex128)
        # the handle is the C pointer as an integer (inverter.py:88-94)
        self.cdata = L.inverter_create(gf.geo.total_site(), _lib.ptr(gauge), fa.mass(), fa.ls(), fa.m5(),
                                       fa.mobius_scale(), int(fa.is_multiplying_dminus),
                                       int(fa.is_using_zmobius), 2, _lib.ptr(self._bs), _lib.ptr(self._cs))
        self.set_solver(solver, reliable_delta)

    def __del__(self):
        h = getattr(self, "cdata", None)
        if h:
            self.cdata = None
            _lib.load().inverter_destroy(h)

    # ----------------------------------------------------------- params
    def geo(self):
        return self.geo_

    def _params(self):
        return _lib.load().inverter_get_params(self.cdata)

    def _set_params(self, s, n, c):
        _lib.load().inverter_set_params(self.cdata, s, n, c)

    def stop_rsd(self):
        return self._params()[0]

    def set_stop_rsd(self, stop_rsd):
        _, n, c = self._params()
        self._set_params(float(stop_rsd), n, c)

    def max_num_iter(self):
        return self._params()[1]

    def set_max_num_iter(self, max_num_iter):
        s, _, c = self._params()
        self._set_params(s, int(max_num_iter), c)

    def max_mixed_precision_cycle(self):
        return self._params()[2]

    def set_max_mixed_precision_cycle(self, max_mixed_precision_cycle):
        s, n, _ = self._params()
        self._set_params(s, n, int(max_mixed_precision_cycle))

    def set_solver(self, solver, reliable_delta=0.01):
        if solver not in SOLVERS:
            raise ValueError(f"solver must be one of {sorted(SOLVERS)}")
        _lib.load().inverter_set_solver(self.cdata, SOLVERS[solver], float(reliable_delta))
        self.solver = solver

    def set_low_modes(self, lm):
        """Attach low modes (InverterDomainWall::setup(gf, fa, lm),
        dslash.h:176-186); None detaches.  The FP64 solver then starts each
        restart cycle from deflate(residual) (dslash.h:496-500)."""
        L = _lib.load()
        if lm is None:
            L.inverter_set_low_modes(self.cdata, (0, 0, 0, 0), 0, 0, 0, 0, 0)
            self.lm = None
            return
        ev = np.ascontiguousarray(lm.eigen_values, dtype=np.float64)
        bases = np.ascontiguousarray(lm.bases, dtype=np.complex64)
        coefs = np.ascontiguousarray(lm.coefs, dtype=np.complex64)
        L.inverter_set_low_modes(self.cdata, lm.block_site, int(lm.n_basis), int(lm.n_vec),
                                 _lib.addr(ev, np.float64), _lib.addr(bases, np.complex64),
                                 _lib.addr(coefs, np.complex64))
        self.lm = lm

    def set_gauge(self, gf):
        """Replace the links (InverterDomainWall::setup, dslash.h:154-162)."""
        from .force import set_gauge

        set_gauge(self, gf)

    def stats(self):
        return _lib.load().inverter_get_stats(self.cdata)

    # ------------------------------------------------------------ apply
    def __mul__(self, src):
        """prop_src: Prop, FermionField4d, full FermionField5d, or a list."""
        if isinstance(src, list):
            # inverter.py:106-107 solves the list one by one; propagators
            # here go through one batched call, all their columns as one
            # multi-RHS problem
            if len(src) > 1 and all(isinstance(p, Prop) for p in src):
                return _invert_props(self, src)
            return [self * p for p in src]
        if self.timer is not None:
            self.timer.start()
        try:
            if isinstance(src, Prop):
                sol = Prop(self.geo_)
                invert(sol, src, self)
            elif isinstance(src, FermionField4d):
                sol = FermionField4d(self.geo_)
                invert(sol, src, self)
            elif isinstance(src, FermionField5d):
                sol = FermionField5d(self.geo_, self.fa.ls(), 0)
                invert(sol, src, self)
            else:
                raise Exception("InverterDomainWall")
        finally:
            if self.timer is not None:
                self.timer.stop()
        return sol


def _invert_props(inv, props):
    """Several propagators with one inverter in one batched call: every
    column of every propagator becomes a 5D source by the chiral embedding
    (qcd-prop.h:96-116) and the 12 n columns are solved together as
    multi-RHS tiles (qlat_dwf_invert_5d_batch); each solution goes back to
    4D (qcd-prop.h:118-138).  The same columns as invert(Prop) would solve
    one propagator at a time (qcd-prop.h:169-186)."""
    ls = inv.fa.ls()
    geo = inv.geo_
    V = geo.total_volume()
    srcs = []
    for p in props:
        d = np.ascontiguousarray(p.data, dtype=np.complex128).reshape(V, 12, 12)
        for j in range(12):
            col = FermionField4d(geo)
            col.data = np.ascontiguousarray(d[:, :, j])
            srcs.append(fermion_field_5d_from_4d(col, ls))
    if inv.timer is not None:
        inv.timer.start()
    try:
        sols, _ = invert_batch([inv] * len(srcs), srcs)
    finally:
        if inv.timer is not None:
            inv.timer.stop()
    out = []
    for k in range(len(props)):
        r = Prop(geo)
        data = np.empty((V, 12, 12), dtype=np.complex128)
        for j in range(12):
            data[:, :, j] = fermion_field_4d_from_5d(sols[12 * k + j]).data
        r.data = data.reshape(r.data.shape)
        out.append(r)
    return out


# ================================================================ operators

def _half_out(in_, eo):
    out = FermionField5d.__new__(FermionField5d)
    out.geo, out.ls, out.eo = in_.geo, in_.ls, eo
    out.data = np.empty_like(in_.data)
    return out


def _check_half(in_, inv):
    assert isinstance(in_, FermionField5d) and in_.eo in (1, 2)
    assert in_.data.shape == (inv.geo().total_volume() // 2, inv.fa.ls(), 12)


def _apply_half(name, out, in_, inv, eo_arg, eo_out):
    _check_half(in_, inv)
    src = np.ascontiguousarray(in_.data, dtype=np.complex128)
    res = _half_out(in_, eo_out)
    fn = getattr(_lib.load(), name)
    args = (inv.cdata, _lib.ptr(res.data), _lib.ptr(src))
    if eo_arg is not None:
        args = args + (int(eo_arg),)
    fn(*args)
    if out is not None:
        out.geo, out.ls, out.eo, out.data = res.geo, res.ls, res.eo, res.data
        return out
    return res


def multiply_m_e_e(out, in_, inv):
    """dslash.cpp:591-639 (works for _o_o as well)."""
    return _apply_half("multiply_m_e_e", out, in_, inv, None, in_.eo)


def multiply_mdag_e_e(out, in_, inv):
    """dslash.cpp:641-692."""
    return _apply_half("multiply_mdag_e_e", out, in_, inv, None, in_.eo)


def multiply_m_e_e_inv(out, in_, inv):
    """dslash.cpp:694-769."""
    return _apply_half("multiply_m_e_e_inv", out, in_, inv, None, in_.eo)


def multiply_mdag_e_e_inv(out, in_, inv):
    """dslash.cpp:771-857."""
    return _apply_half("multiply_mdag_e_e_inv", out, in_, inv, None, in_.eo)


def multiply_wilson_d_e_o_no_comm(out, in_, inv):
    """dslash.cpp:859-908: output parity 3 - in.eo."""
    return _apply_half("multiply_wilson_d_e_o", out, in_, inv, in_.eo, 3 - in_.eo)


def multiply_wilson_ddag_e_o_no_comm(out, in_, inv):
    """dslash.cpp:910-959."""
    return _apply_half("multiply_wilson_ddag_e_o", out, in_, inv, in_.eo, 3 - in_.eo)


def multiply_m_e_o(out, in_, inv):
    """dslash.cpp:961-1003 (works for _o_e as well)."""
    return _apply_half("multiply_m_e_o", out, in_, inv, in_.eo, 3 - in_.eo)


def multiply_mdag_e_o(out, in_, inv):
    """dslash.cpp:1005-1060."""
    return _apply_half("multiply_mdag_e_o", out, in_, inv, in_.eo, 3 - in_.eo)


def multiply_m_eo_eo(out, in_, inv, eo_out, eo_in):
    """dslash.cpp:1193-1211: the (eo_out, eo_in) block of M, M_ee (or M_oo)
    for equal parities and M_eo (or M_oe) otherwise."""
    assert in_.eo == eo_in and eo_out in (1, 2)
    if eo_out == eo_in:
        return multiply_m_e_e(out, in_, inv)
    return multiply_m_e_o(out, in_, inv)


def multiply_mdag_eo_eo(out, in_, inv, eo_out, eo_in):
    """dslash.cpp:1213-1231: the (eo_out, eo_in) block of M^+."""
    assert in_.eo == eo_in and eo_out in (1, 2)
    if eo_out == eo_in:
        return multiply_mdag_e_e(out, in_, inv)
    return multiply_mdag_e_o(out, in_, inv)


def multiply_mpc_sym2(out, in_, inv):
    """dslash.cpp:1102-1119 (odd <- odd, works for even <- even)."""
    return _apply_half("multiply_mpc_sym2", out, in_, inv, in_.eo, in_.eo)


def multiply_mpcdag_sym2(out, in_, inv):
    """dslash.cpp:1121-1138."""
    return _apply_half("multiply_mpcdag_sym2", out, in_, inv, in_.eo, in_.eo)


def multiply_hermop_sym2(out, in_, inv):
    """dslash.cpp:1140-1149."""
    return _apply_half("multiply_hermop_sym2", out, in_, inv, in_.eo, in_.eo)


class LowModes:
    """LowModes (dslash.h:43-85): a block-compressed eigen system of
    multiply_hermop_sym2 on the odd sites, in the reference's in-memory
    layout (compressed-eigen-io.h:175-195):

      block_site    4 even-x block extents dividing total_site
      eigen_values  (n_vec,) float64
      bases         (n_block, n_basis, block_vol/2 * Ls * 12) complex64
      coefs         (n_block, n_vec, n_basis) complex64

    eigenvector i restricted to block b = sum_j coefs[b, i, j] bases[b, j];
    blocks are ordered lexicographically (x fastest), and inside a block the
    half site (lexicographic index of x % block_site) / 2 comes first, then
    s, then spin*3+colour (convert_half_vector, commpressed-eigen-io.cpp:925-947).
    """

    def __init__(self, block_site, eigen_values, bases, coefs):
        self.block_site = tuple(int(v) for v in block_site)
        self.eigen_values = np.asarray(eigen_values, dtype=np.float64)
        self.bases = np.asarray(bases, dtype=np.complex64)
        self.coefs = np.asarray(coefs, dtype=np.complex64)
        self.n_vec = len(self.eigen_values)
        self.n_basis = self.bases.shape[1]
        assert self.coefs.shape == (self.bases.shape[0], self.n_vec, self.n_basis)


def deflate(out, in_, inv):
    """deflate(FermionField5d&, const FermionField5d&, LowModes&)
    (dslash.cpp:251-300) with the inverter's low modes: out = in + sum_i v_i
    <v_i, in> / lambda_i with the input rounded to FP32 (the reference's
    produce step accumulates into the blocked input), or 0 without low modes."""
    return _apply_half("deflate", out, in_, inv, in_.eo, in_.eo)


def multiply_hermop_sym2_fp32(out, in_, inv):
    """Mpc^+ Mpc evaluated in FP32 (the mixed solver's inner operator)."""
    return _apply_half("multiply_hermop_sym2_fp32", out, in_, inv, in_.eo, in_.eo)


def _apply_full(name, out, in_, inv):
    assert isinstance(in_, FermionField5d) and in_.eo == 0
    src = np.ascontiguousarray(in_.data, dtype=np.complex128)
    res = FermionField5d.__new__(FermionField5d)
    res.geo, res.ls, res.eo = in_.geo, in_.ls, 0
    res.data = np.empty_like(src)
    getattr(_lib.load(), name)(inv.cdata, _lib.ptr(res.data), _lib.ptr(src))
    if out is not None:
        out.geo, out.ls, out.eo, out.data = res.geo, res.ls, 0, res.data
        return out
    return res


def multiply_m(out, in_, inv):
    """dslash.cpp:1062-1080."""
    return _apply_full("multiply_m", out, in_, inv)


def multiply_mdag(out, in_, inv):
    """dslash.cpp:1082-1100."""
    return _apply_full("multiply_mdag", out, in_, inv)


def multiply_m_full(out, in_, inv):
    """multiply_m_full (dslash.cpp:490-540): M = D_W(-M5)(b + c S) + 1 - S
    on the full field, the operator of multiply_m (one implementation on
    the GPU serves both)."""
    return _apply_full("multiply_m", out, in_, inv)


def multiply_m_dwf(out, in_, inv):
    """multiply_m_dwf (dslash.cpp:329-398): the Shamir operator; like the
    reference it requires mobius_scale == 1 (:342).  Equal to multiply_m
    for that action."""
    if inv.fa.mobius_scale() != 1.0 or inv.fa.is_using_zmobius:
        raise AssertionError("multiply_m_dwf: Shamir only (mobius_scale == 1, no zMobius)")
    return _apply_full("multiply_m", out, in_, inv)


def multiply_m_with_prec_sym2(out, in_, inv):
    """multiply_m_with_prec_sym2 (dslash.cpp:1271-1296): M x through the
    sym2 factorisation, from the half-field operators in the reference's
    order (each one on the GPU):
    out_e = M_ee (M_ee^-1 M_eo x_o + x_e),
    out_o = Mpc M_oo x_o + M_oe M_ee^-1 out_e."""
    assert isinstance(in_, FermionField5d) and in_.eo == 0
    in_e = get_half_fermion(in_, 2)
    in_o = get_half_fermion(in_, 1)
    out_e = multiply_m_e_e_inv(None, multiply_m_e_o(None, in_o, inv), inv)
    out_e.data = out_e.data + in_e.data
    out_o = multiply_m_e_e(None, in_o, inv)
    out_e = multiply_m_e_e(None, out_e, inv)
    out_o = multiply_mpc_sym2(None, out_o, inv)
    tmp = multiply_m_e_o(None, multiply_m_e_e_inv(None, out_e, inv), inv)
    out_o.data = out_o.data + tmp.data
    res = FermionField5d(in_.geo, in_.ls, 0)
    res.data = np.empty_like(np.ascontiguousarray(in_.data, dtype=np.complex128))
    set_half_fermion(res, out_e, 2)
    set_half_fermion(res, out_o, 1)
    if out is not None:
        out.geo, out.ls, out.eo, out.data = res.geo, res.ls, 0, res.data
        return out
    return res


def multiply_d_minus(out, in_, inv):
    """dslash.cpp:448-488."""
    return _apply_full("multiply_d_minus", out, in_, inv)


def dot_product(ff1, ff2, inv):
    """ff1^+ ff2 (dslash.cpp:1298-1321), reduced on the GPU of `inv`."""
    a = np.ascontiguousarray(ff1.data, dtype=np.complex128)
    b = np.ascontiguousarray(ff2.data, dtype=np.complex128)
    assert a.shape == b.shape
    return _lib.load().dot_product(inv.cdata, _lib.ptr(a), _lib.ptr(b), a.size // 12)


def qnorm(ff, inv):
    """field.h:170-195, reduced on the GPU of `inv`."""
    a = np.ascontiguousarray(ff.data, dtype=np.complex128)
    return _lib.load().qnorm(inv.cdata, _lib.ptr(a), a.size // 12)


def cg_with_herm_sym_2(sol, src, inv, stop_rsd=1e-8, max_num_iter=50000):
    """dslash.cpp:1323-1333 on odd half fields; `sol` is the initial guess
    when it holds data (cg_with_f, dslash.h:341-346).  Returns iterations,
    max_num_iter + 1 if not converged."""
    _check_half(src, inv)
    b = np.ascontiguousarray(src.data, dtype=np.complex128)
    if sol.data is None or getattr(sol, "eo", None) != src.eo:
        x = np.zeros_like(b)
    else:
        x = np.ascontiguousarray(sol.data, dtype=np.complex128).copy()
    it = _lib.load().cg_with_herm_sym_2(inv.cdata, _lib.ptr(x), _lib.ptr(b), float(stop_rsd), int(max_num_iter))
    sol.geo, sol.ls, sol.eo, sol.data = src.geo, src.ls, src.eo, x
    return it


def invert(out, in_, inv):
    """invert(5d) (dslash.cpp:1335-1340), invert(4d) (dslash.cpp:1342-1347)
    or invert(Prop) (qcd-prop.h:169-186), by the type of `in_`.  Returns the
    total number of CG iterations."""
    L = _lib.load()
    src = np.ascontiguousarray(in_.data, dtype=np.complex128)
    res = np.empty_like(src)
    if isinstance(in_, Prop):
        it = L.invert_prop(inv.cdata, _lib.ptr(res), _lib.ptr(src))
    elif isinstance(in_, FermionField4d):
        it = L.invert_4d(inv.cdata, _lib.ptr(res), _lib.ptr(src))
    elif isinstance(in_, FermionField5d):
        assert in_.eo == 0
        it = L.invert_5d(inv.cdata, _lib.ptr(res), _lib.ptr(src))
        out.ls, out.eo = in_.ls, 0
    else:
        raise TypeError("invert: unsupported field type")
    out.geo = in_.geo
    out.data = res
    return it


def invert_with_cg_with_guess(out, in_, inv):
    """invert_with_cg_with_guess (dslash.h:532-573) with the inverter's
    solver on full FermionField5d: `out` holds the initial guess x0 and
    receives x0 plus the correction solved from D_- in - M x0, with stop_rsd
    scaled by |D_- in| / |D_- in - M x0|.  Returns the correction's CG
    iterations."""
    assert isinstance(in_, FermionField5d) and in_.eo == 0
    assert isinstance(out, FermionField5d) and out.eo == 0 and out.data is not None
    src = np.ascontiguousarray(in_.data, dtype=np.complex128)
    res = np.array(out.data, dtype=np.complex128, order="C", copy=True)
    assert res.shape == src.shape
    it = _lib.load().invert_5d_with_guess(inv.cdata, _lib.ptr(res), _lib.ptr(src))
    out.data = res
    return it


def invert_batch(invs, srcs):
    """Solve srcs[j] (full FermionField5d) with invs[j] for every j in one
    batched call (qlat_dwf_invert_5d_batch): the inverters share lattice, Ls
    and gauge field and may differ in mass, M5, b, c; with the mixed solvers
    the systems run as multi-RHS tiles with per-column retirement.  Solver
    and stopping parameters are those of invs[0].  Returns (solutions,
    per-system CG iterations)."""
    n = len(invs)
    assert n == len(srcs) and n > 0
    data = [np.ascontiguousarray(f.data, dtype=np.complex128) for f in srcs]
    for f in srcs:
        assert isinstance(f, FermionField5d) and f.eo == 0
    outs = [np.empty_like(d) for d in data]
    it = np.array(_lib.load().invert_5d_batch([inv.cdata for inv in invs], [_lib.ptr(o) for o in outs],
                                              [_lib.ptr(d) for d in data]), dtype=np.int64)
    sols = []
    for f, o in zip(srcs, outs):
        r = FermionField5d(f.geo, f.ls, 0)
        r.data = o
        sols.append(r)
    return sols, it


def find_max_eigen_value_hermop_sym2(inv, rs, max_iter=100):
    """find_max_eigen_value_hermop_sym2(inv, rs, max_iter) (dslash.h:586-588,
    dslash.cpp:1349-1369): power iteration for the largest eigenvalue of
    Mpc^+ Mpc.  With an RngState `rs` the start vector is the reference's
    set_u_rand(ff, rs) on the odd half field, drawn on the GPU; an odd
    FermionField5d `rs` is used as the start vector itself."""
    from .rng import RngState, midstate

    L = _lib.load()
    if isinstance(rs, RngState):
        return L.find_max_eigen_value_hermop_sym2_rng(inv.cdata, midstate(rs), len(rs._msg), rs.index,
                                                      int(max_iter))
    _check_half(rs, inv)
    a = np.ascontiguousarray(rs.data, dtype=np.complex128)
    return L.find_max_eigen_value_hermop_sym2(inv.cdata, _lib.ptr(a), int(max_iter))


def bench_kernel(inv, which, precision, iters):
    """Average launch time (ms) of one kernel, HIP events on the inverter's
    stream.  which: 0 hop, 1 hop^+, 2 s-matrix, 3 Mpc^+ Mpc, 4 copy
    (precision 64/32); 5 hop + s-matrix, 6 CG step, 7 Mpc^+ Mpc with dot on
    the three-phase solver's tile (precision 32/16)."""
    return _lib.load().bench_kernel(inv.cdata, int(which), int(precision), int(iters))
