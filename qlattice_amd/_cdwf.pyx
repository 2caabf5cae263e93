# cython: language_level=3, boundscheck=False, wraparound=False
"""Cython binding of the C ABI in include/qlat_dwf.h.

The reference's Python layer reaches its C++ library through Cython
(qlat/qlat/cqlat/inverter.cpp behind qlat/qlat/inverter.py:75-129, with
`except +` declarations in everything.pxd).  This module is the same layer
for the HIP library: every entry point of the header is declared here with
its C signature (the C compiler checks each call against the header when
this file is built), wrapped in a Python function that takes field
addresses as integers (numpy or torch storage, see _lib.addr), releases
the GIL for the call and raises QlatDwfError with the library's message
when the call fails -- the role of cqlat's EXPORT macro
(cqlat/lib.h:28-35).  Inverter handles are the C pointers as integers, as
the reference keeps its InverterDomainWall* in an int (inverter.py:88-94).
"""

from libc.stdint cimport int64_t, uint32_t, uint64_t, uintptr_t
from libc.stdlib cimport free, malloc


cdef extern from "qlat_dwf.h" nogil:
    ctypedef struct qlat_dwf_inverter:
        pass

    ctypedef struct qlat_dwf_fermion_action:
        double mass
        int ls
        double m5
        double mobius_scale
        int is_multiplying_dminus
        int is_using_zmobius
        int cg_diagonal_mee
        const double* bs
        const double* cs

    ctypedef struct qlat_dwf_solve_stats:
        long total_iter
        long cycles
        long reliable_updates
        long hermop_fp64
        long hermop_low
        double true_residual
        double seconds
        int spinor_bits
        int rhs_tile
        int sites_per_block
        double true_residual_5d
        int hop_arith_bits

    ctypedef struct qlat_dwf_source:
        int kind
        int xg[4]
        int tslice
        double lmom[4]
        double value[2]
        const double* fu1

    const char* qlat_dwf_last_error()
    int qlat_dwf_abi_version()
    int qlat_dwf_inverter_create(qlat_dwf_inverter** inv, const int* total_site, const double* gauge,
                                 const qlat_dwf_fermion_action* fa)
    int qlat_dwf_inverter_destroy(qlat_dwf_inverter* inv)
    int qlat_dwf_inverter_set_params(qlat_dwf_inverter* inv, double stop_rsd, long max_num_iter,
                                     long max_mixed_precision_cycle)
    int qlat_dwf_inverter_get_params(const qlat_dwf_inverter* inv, double* stop_rsd, long* max_num_iter,
                                     long* max_mixed_precision_cycle)
    int qlat_dwf_inverter_set_solver(qlat_dwf_inverter* inv, int solver, double reliable_delta)
    int qlat_dwf_inverter_get_stats(const qlat_dwf_inverter* inv, qlat_dwf_solve_stats* stats)
    int qlat_dwf_set_checking_invert(int on)
    int qlat_dwf_is_checking_invert()
    int qlat_dwf_multiply_m_e_e(qlat_dwf_inverter* inv, double* out, const double* inp)
    int qlat_dwf_multiply_mdag_e_e(qlat_dwf_inverter* inv, double* out, const double* inp)
    int qlat_dwf_multiply_m_e_e_inv(qlat_dwf_inverter* inv, double* out, const double* inp)
    int qlat_dwf_multiply_mdag_e_e_inv(qlat_dwf_inverter* inv, double* out, const double* inp)
    int qlat_dwf_multiply_wilson_d_e_o(qlat_dwf_inverter* inv, double* out, const double* inp, int eo_in)
    int qlat_dwf_multiply_wilson_ddag_e_o(qlat_dwf_inverter* inv, double* out, const double* inp, int eo_in)
    int qlat_dwf_multiply_m_e_o(qlat_dwf_inverter* inv, double* out, const double* inp, int eo_in)
    int qlat_dwf_multiply_mdag_e_o(qlat_dwf_inverter* inv, double* out, const double* inp, int eo_in)
    int qlat_dwf_multiply_mpc_sym2(qlat_dwf_inverter* inv, double* out, const double* inp, int eo)
    int qlat_dwf_multiply_mpcdag_sym2(qlat_dwf_inverter* inv, double* out, const double* inp, int eo)
    int qlat_dwf_multiply_hermop_sym2(qlat_dwf_inverter* inv, double* out, const double* inp, int eo)
    int qlat_dwf_multiply_hermop_sym2_fp32(qlat_dwf_inverter* inv, double* out, const double* inp, int eo)
    int qlat_dwf_multiply_m(qlat_dwf_inverter* inv, double* out, const double* inp)
    int qlat_dwf_multiply_mdag(qlat_dwf_inverter* inv, double* out, const double* inp)
    int qlat_dwf_multiply_d_minus(qlat_dwf_inverter* inv, double* out, const double* inp)
    int qlat_dwf_dot_product(qlat_dwf_inverter* inv, const double* a, const double* b, long n_wv, double* re,
                             double* im)
    int qlat_dwf_qnorm(qlat_dwf_inverter* inv, const double* a, long n_wv, double* result)
    int qlat_dwf_cg_with_herm_sym_2(qlat_dwf_inverter* inv, double* sol, const double* src, double stop_rsd,
                                    long max_num_iter, long* iter)
    int qlat_dwf_invert_5d(qlat_dwf_inverter* inv, double* out, const double* inp, long* iter)
    int qlat_dwf_invert_5d_with_guess(qlat_dwf_inverter* inv, double* out, const double* inp, long* iter)
    int qlat_dwf_invert_4d(qlat_dwf_inverter* inv, double* out, const double* inp, long* iter)
    int qlat_dwf_invert_prop(qlat_dwf_inverter* inv, double* sol, const double* src, long* iter)
    int qlat_dwf_make_source(const int* total_site, const qlat_dwf_source* src, double* prop)
    int qlat_dwf_source_propagator(qlat_dwf_inverter* inv, const qlat_dwf_source* src, double* sol, long* iter)
    int qlat_dwf_inverter_set_low_modes(qlat_dwf_inverter* inv, const int* block_site, int n_basis, int n_vec,
                                        const double* eigen_values, const float* bases, const float* coefs)
    int qlat_dwf_deflate(qlat_dwf_inverter* inv, double* out, const double* inp, int eo)
    int qlat_dwf_free_invert(const int* total_site, double mass, double m5, const double* momtwist, double* sol,
                             const double* src)
    int qlat_dwf_rng_fill_uniform(const uint32_t* midstate, uint64_t num_bytes, uint64_t index0, int64_t n_sites,
                                  int n_per_site, double upper, double lower, double* out)
    int qlat_dwf_rng_fill_gaussian(const uint32_t* midstate, uint64_t num_bytes, uint64_t index0,
                                   int64_t n_sites, int n_per_site, double center, double sigma, double* out)
    int qlat_dwf_multiply_m_dwf_qed(const int* total_site, int ls, const double* gf1, double mass, double m5,
                                    const double* wick, int is_dagger, double* out, const double* inp)
    int qlat_dwf_cg_with_m_dwf_qed(const int* total_site, int ls, const double* gf1, double mass, double m5,
                                   const double* wick, int is_dagger, double stop_rsd, long max_num_iter,
                                   double* out, const double* inp, long* iter)
    int qlat_dwf_invert_dwf_qed(const int* total_site, int ls, const double* gf1, double mass, double m5,
                                const double* wick, int is_dagger, double stop_rsd, long max_num_iter,
                                double* out4d, const double* in4d, long* iter)
    int qlat_dwf_find_max_eigen_value_hermop_sym2(qlat_dwf_inverter* inv, const double* start, long max_iter,
                                                  double* result)
    int qlat_dwf_find_max_eigen_value_hermop_sym2_rng(qlat_dwf_inverter* inv, const uint32_t* midstate,
                                                      uint64_t num_bytes, uint64_t index0, long max_iter,
                                                      double* result)
    int qlat_dwf_invert_5d_batch(qlat_dwf_inverter* const* invs, int n, double* const* sol,
                                 const double* const* src, long* iters)
    int qlat_dwf_invert_5d_batch_device(qlat_dwf_inverter* const* invs, int n, double* const* sol,
                                        const double* const* src, long* iters)
    int qlat_dwf_fermion_force(qlat_dwf_inverter* inv, const double* phi, double stop_rsd, double* force,
                               double* grad, double* action, double* action_lo, long* iter)
    int qlat_dwf_fermion_force_device(qlat_dwf_inverter* inv, const double* phi, double stop_rsd, double* force,
                                      double* grad, double* action, double* action_lo, long* iter)
    int qlat_dwf_inverter_set_gauge(qlat_dwf_inverter* inv, const double* gauge, int on_device)
    int qlat_dwf_multiply_hermop_sym2_device(qlat_dwf_inverter* inv, double* out, const double* inp, int eo)
    int qlat_dwf_multiply_hermop_sym2_device_noalloc(qlat_dwf_inverter* inv, double* out, const double* inp,
                                                     int eo)
    int qlat_dwf_invert_5d_device(qlat_dwf_inverter* inv, double* out, const double* inp, long* iter)
    int qlat_dwf_invert_5d_with_guess_device(qlat_dwf_inverter* inv, double* out, const double* inp, long* iter)
    int qlat_dwf_invert_prop_device(qlat_dwf_inverter* inv, double* sol, const double* src, long* iter)
    int qlat_dwf_make_source_device(const int* total_site, const qlat_dwf_source* src, double* prop)
    int qlat_dwf_source_propagator_device(qlat_dwf_inverter* inv, const qlat_dwf_source* src, double* sol,
                                          long* iter)
    int qlat_dwf_bench_kernel(qlat_dwf_inverter* inv, int which, int precision, int iters, double* avg_ms)


from qlattice_amd._errors import QlatDwfError


SOLVER_FP64 = 0
SOLVER_MIXED_FP32 = 1
SOLVER_MIXED_FP16 = 2
SRC_POINT, SRC_WALL, SRC_MOM, SRC_U1 = 0, 1, 2, 3

STATS_FIELDS = ("total_iter", "cycles", "reliable_updates", "hermop_fp64", "hermop_low", "true_residual",
                "seconds", "spinor_bits", "rhs_tile", "sites_per_block", "true_residual_5d", "hop_arith_bits")


cdef int _check(int rc, str what) except -1:
    if rc != 0:
        msg = qlat_dwf_last_error().decode(errors="replace")
        raise QlatDwfError(f"{what} failed (status {rc}): {msg}")
    return 0


cdef inline qlat_dwf_inverter* _inv(uintptr_t h) except NULL:
    if h == 0:
        raise QlatDwfError("null inverter handle")
    return <qlat_dwf_inverter*>h


cdef void _site4(object v, int* out) except *:
    cdef int i
    if len(v) != 4:
        raise ValueError("expected four lattice extents / coordinates")
    for i in range(4):
        out[i] = int(v[i])


def last_error():
    return qlat_dwf_last_error().decode(errors="replace")


def abi_version():
    return qlat_dwf_abi_version()


# ---------------------------------------------------------------- setup
def inverter_create(total_site, uintptr_t gauge, double mass, int ls, double m5, double mobius_scale,
                    int is_multiplying_dminus=1, int is_using_zmobius=0, int cg_diagonal_mee=2,
                    uintptr_t bs=0, uintptr_t cs=0):
    """qlat_dwf_inverter_create; bs/cs: addresses of ls complex numbers or 0."""
    cdef int ts[4]
    cdef qlat_dwf_fermion_action fa
    cdef qlat_dwf_inverter* h = NULL
    cdef int rc
    _site4(total_site, ts)
    fa.mass = mass
    fa.ls = ls
    fa.m5 = m5
    fa.mobius_scale = mobius_scale
    fa.is_multiplying_dminus = is_multiplying_dminus
    fa.is_using_zmobius = is_using_zmobius
    fa.cg_diagonal_mee = cg_diagonal_mee
    fa.bs = <const double*>bs
    fa.cs = <const double*>cs
    with nogil:
        rc = qlat_dwf_inverter_create(&h, ts, <const double*>gauge, &fa)
    _check(rc, "qlat_dwf_inverter_create")
    return <uintptr_t>h


def inverter_destroy(uintptr_t h):
    if h:
        _check(qlat_dwf_inverter_destroy(<qlat_dwf_inverter*>h), "qlat_dwf_inverter_destroy")


def inverter_set_params(uintptr_t h, double stop_rsd, long max_num_iter, long max_mixed_precision_cycle):
    _check(qlat_dwf_inverter_set_params(_inv(h), stop_rsd, max_num_iter, max_mixed_precision_cycle),
           "qlat_dwf_inverter_set_params")


def inverter_get_params(uintptr_t h):
    cdef double s = 0
    cdef long n = 0, c = 0
    _check(qlat_dwf_inverter_get_params(_inv(h), &s, &n, &c), "qlat_dwf_inverter_get_params")
    return s, n, c


def inverter_set_solver(uintptr_t h, int solver, double reliable_delta):
    _check(qlat_dwf_inverter_set_solver(_inv(h), solver, reliable_delta), "qlat_dwf_inverter_set_solver")


def inverter_get_stats(uintptr_t h):
    cdef qlat_dwf_solve_stats st
    _check(qlat_dwf_inverter_get_stats(_inv(h), &st), "qlat_dwf_inverter_get_stats")
    return {"total_iter": st.total_iter, "cycles": st.cycles, "reliable_updates": st.reliable_updates,
            "hermop_fp64": st.hermop_fp64, "hermop_low": st.hermop_low, "true_residual": st.true_residual,
            "seconds": st.seconds, "spinor_bits": st.spinor_bits, "rhs_tile": st.rhs_tile,
            "sites_per_block": st.sites_per_block, "true_residual_5d": st.true_residual_5d,
            "hop_arith_bits": st.hop_arith_bits}


def set_checking_invert(bint on):
    _check(qlat_dwf_set_checking_invert(1 if on else 0), "qlat_dwf_set_checking_invert")


def is_checking_invert():
    return qlat_dwf_is_checking_invert() != 0


def inverter_set_gauge(uintptr_t h, uintptr_t gauge, bint on_device):
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef int rc
    with nogil:
        rc = qlat_dwf_inverter_set_gauge(p, <const double*>gauge, 1 if on_device else 0)
    _check(rc, "qlat_dwf_inverter_set_gauge")


# ------------------------------------------------------------ operators
ctypedef int (*op2_t)(qlat_dwf_inverter*, double*, const double*) noexcept nogil
ctypedef int (*op3_t)(qlat_dwf_inverter*, double*, const double*, int) noexcept nogil


cdef _call2(op2_t f, str name, uintptr_t h, uintptr_t out, uintptr_t inp):
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef int rc
    with nogil:
        rc = f(p, <double*>out, <const double*>inp)
    _check(rc, name)


cdef _call3(op3_t f, str name, uintptr_t h, uintptr_t out, uintptr_t inp, int eo):
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef int rc
    with nogil:
        rc = f(p, <double*>out, <const double*>inp, eo)
    _check(rc, name)


def multiply_m_e_e(uintptr_t h, uintptr_t out, uintptr_t inp):
    _call2(qlat_dwf_multiply_m_e_e, "qlat_dwf_multiply_m_e_e", h, out, inp)


def multiply_mdag_e_e(uintptr_t h, uintptr_t out, uintptr_t inp):
    _call2(qlat_dwf_multiply_mdag_e_e, "qlat_dwf_multiply_mdag_e_e", h, out, inp)


def multiply_m_e_e_inv(uintptr_t h, uintptr_t out, uintptr_t inp):
    _call2(qlat_dwf_multiply_m_e_e_inv, "qlat_dwf_multiply_m_e_e_inv", h, out, inp)


def multiply_mdag_e_e_inv(uintptr_t h, uintptr_t out, uintptr_t inp):
    _call2(qlat_dwf_multiply_mdag_e_e_inv, "qlat_dwf_multiply_mdag_e_e_inv", h, out, inp)


def multiply_m(uintptr_t h, uintptr_t out, uintptr_t inp):
    _call2(qlat_dwf_multiply_m, "qlat_dwf_multiply_m", h, out, inp)


def multiply_mdag(uintptr_t h, uintptr_t out, uintptr_t inp):
    _call2(qlat_dwf_multiply_mdag, "qlat_dwf_multiply_mdag", h, out, inp)


def multiply_d_minus(uintptr_t h, uintptr_t out, uintptr_t inp):
    _call2(qlat_dwf_multiply_d_minus, "qlat_dwf_multiply_d_minus", h, out, inp)


def multiply_wilson_d_e_o(uintptr_t h, uintptr_t out, uintptr_t inp, int eo_in):
    _call3(qlat_dwf_multiply_wilson_d_e_o, "qlat_dwf_multiply_wilson_d_e_o", h, out, inp, eo_in)


def multiply_wilson_ddag_e_o(uintptr_t h, uintptr_t out, uintptr_t inp, int eo_in):
    _call3(qlat_dwf_multiply_wilson_ddag_e_o, "qlat_dwf_multiply_wilson_ddag_e_o", h, out, inp, eo_in)


def multiply_m_e_o(uintptr_t h, uintptr_t out, uintptr_t inp, int eo_in):
    _call3(qlat_dwf_multiply_m_e_o, "qlat_dwf_multiply_m_e_o", h, out, inp, eo_in)


def multiply_mdag_e_o(uintptr_t h, uintptr_t out, uintptr_t inp, int eo_in):
    _call3(qlat_dwf_multiply_mdag_e_o, "qlat_dwf_multiply_mdag_e_o", h, out, inp, eo_in)


def multiply_mpc_sym2(uintptr_t h, uintptr_t out, uintptr_t inp, int eo):
    _call3(qlat_dwf_multiply_mpc_sym2, "qlat_dwf_multiply_mpc_sym2", h, out, inp, eo)


def multiply_mpcdag_sym2(uintptr_t h, uintptr_t out, uintptr_t inp, int eo):
    _call3(qlat_dwf_multiply_mpcdag_sym2, "qlat_dwf_multiply_mpcdag_sym2", h, out, inp, eo)


def multiply_hermop_sym2(uintptr_t h, uintptr_t out, uintptr_t inp, int eo):
    _call3(qlat_dwf_multiply_hermop_sym2, "qlat_dwf_multiply_hermop_sym2", h, out, inp, eo)


def multiply_hermop_sym2_fp32(uintptr_t h, uintptr_t out, uintptr_t inp, int eo):
    _call3(qlat_dwf_multiply_hermop_sym2_fp32, "qlat_dwf_multiply_hermop_sym2_fp32", h, out, inp, eo)


def multiply_hermop_sym2_device(uintptr_t h, uintptr_t out, uintptr_t inp, int eo):
    _call3(qlat_dwf_multiply_hermop_sym2_device, "qlat_dwf_multiply_hermop_sym2_device", h, out, inp, eo)


def multiply_hermop_sym2_device_noalloc(uintptr_t h, uintptr_t out, uintptr_t inp, int eo):
    _call3(qlat_dwf_multiply_hermop_sym2_device_noalloc, "qlat_dwf_multiply_hermop_sym2_device_noalloc", h, out,
           inp, eo)


def deflate(uintptr_t h, uintptr_t out, uintptr_t inp, int eo):
    _call3(qlat_dwf_deflate, "qlat_dwf_deflate", h, out, inp, eo)


# -------------------------------------------------------------- BLAS-1
def dot_product(uintptr_t h, uintptr_t a, uintptr_t b, long n_wv):
    cdef double re = 0, im = 0
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef int rc
    with nogil:
        rc = qlat_dwf_dot_product(p, <const double*>a, <const double*>b, n_wv, &re, &im)
    _check(rc, "qlat_dwf_dot_product")
    return complex(re, im)


def qnorm(uintptr_t h, uintptr_t a, long n_wv):
    cdef double r = 0
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef int rc
    with nogil:
        rc = qlat_dwf_qnorm(p, <const double*>a, n_wv, &r)
    _check(rc, "qlat_dwf_qnorm")
    return r


# -------------------------------------------------------------- solvers
ctypedef int (*solve_t)(qlat_dwf_inverter*, double*, const double*, long*) noexcept nogil


cdef long _solve(solve_t f, str name, uintptr_t h, uintptr_t out, uintptr_t inp) except? -1:
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef long it = 0
    cdef int rc
    with nogil:
        rc = f(p, <double*>out, <const double*>inp, &it)
    _check(rc, name)
    return it


def cg_with_herm_sym_2(uintptr_t h, uintptr_t sol, uintptr_t src, double stop_rsd, long max_num_iter):
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef long it = 0
    cdef int rc
    with nogil:
        rc = qlat_dwf_cg_with_herm_sym_2(p, <double*>sol, <const double*>src, stop_rsd, max_num_iter, &it)
    _check(rc, "qlat_dwf_cg_with_herm_sym_2")
    return it


def invert_5d(uintptr_t h, uintptr_t out, uintptr_t inp, bint device=False):
    if device:
        return _solve(qlat_dwf_invert_5d_device, "qlat_dwf_invert_5d_device", h, out, inp)
    return _solve(qlat_dwf_invert_5d, "qlat_dwf_invert_5d", h, out, inp)


def invert_5d_with_guess(uintptr_t h, uintptr_t out, uintptr_t inp, bint device=False):
    if device:
        return _solve(qlat_dwf_invert_5d_with_guess_device, "qlat_dwf_invert_5d_with_guess_device", h, out, inp)
    return _solve(qlat_dwf_invert_5d_with_guess, "qlat_dwf_invert_5d_with_guess", h, out, inp)


def invert_4d(uintptr_t h, uintptr_t out, uintptr_t inp):
    return _solve(qlat_dwf_invert_4d, "qlat_dwf_invert_4d", h, out, inp)


def invert_prop(uintptr_t h, uintptr_t sol, uintptr_t src, bint device=False):
    if device:
        return _solve(qlat_dwf_invert_prop_device, "qlat_dwf_invert_prop_device", h, sol, src)
    return _solve(qlat_dwf_invert_prop, "qlat_dwf_invert_prop", h, sol, src)


def invert_5d_batch(handles, sols, srcs, bint device=False):
    """qlat_dwf_invert_5d_batch[_device]: handles, sols, srcs are equal-length
    sequences of inverter handles and field addresses; returns the
    per-system iteration counts."""
    cdef int n = len(handles)
    cdef int j, rc
    if len(sols) != n or len(srcs) != n:
        raise ValueError("invert_5d_batch: handles, sols and srcs differ in length")
    cdef qlat_dwf_inverter** hv = <qlat_dwf_inverter**>malloc(max(n, 1) * sizeof(void*))
    cdef double** sv = <double**>malloc(max(n, 1) * sizeof(void*))
    cdef const double** iv = <const double**>malloc(max(n, 1) * sizeof(void*))
    cdef long* it = <long*>malloc(max(n, 1) * sizeof(long))
    if hv == NULL or sv == NULL or iv == NULL or it == NULL:
        free(hv); free(sv); free(iv); free(it)
        raise MemoryError()
    try:
        for j in range(n):
            hv[j] = <qlat_dwf_inverter*><uintptr_t>handles[j]
            sv[j] = <double*><uintptr_t>sols[j]
            iv[j] = <const double*><uintptr_t>srcs[j]
            it[j] = 0
        with nogil:
            if device:
                rc = qlat_dwf_invert_5d_batch_device(hv, n, sv, iv, it)
            else:
                rc = qlat_dwf_invert_5d_batch(hv, n, sv, iv, it)
        _check(rc, "qlat_dwf_invert_5d_batch_device" if device else "qlat_dwf_invert_5d_batch")
        return [it[j] for j in range(n)]
    finally:
        free(hv); free(sv); free(iv); free(it)


def find_max_eigen_value_hermop_sym2(uintptr_t h, uintptr_t start, long max_iter):
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef double r = 0
    cdef int rc
    with nogil:
        rc = qlat_dwf_find_max_eigen_value_hermop_sym2(p, <const double*>start, max_iter, &r)
    _check(rc, "qlat_dwf_find_max_eigen_value_hermop_sym2")
    return r


def find_max_eigen_value_hermop_sym2_rng(uintptr_t h, midstate, uint64_t num_bytes, uint64_t index0,
                                         long max_iter):
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef uint32_t mid[8]
    cdef double r = 0
    cdef int rc, i
    for i in range(8):
        mid[i] = midstate[i]
    with nogil:
        rc = qlat_dwf_find_max_eigen_value_hermop_sym2_rng(p, mid, num_bytes, index0, max_iter, &r)
    _check(rc, "qlat_dwf_find_max_eigen_value_hermop_sym2_rng")
    return r


def inverter_set_low_modes(uintptr_t h, block_site, int n_basis, int n_vec, uintptr_t eigen_values,
                           uintptr_t bases, uintptr_t coefs):
    cdef int bs[4]
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef int rc
    _site4(block_site, bs)
    with nogil:
        rc = qlat_dwf_inverter_set_low_modes(p, bs, n_basis, n_vec, <const double*>eigen_values,
                                             <const float*>bases, <const float*>coefs)
    _check(rc, "qlat_dwf_inverter_set_low_modes")


def fermion_force(uintptr_t h, uintptr_t phi, double stop_rsd, uintptr_t force, uintptr_t grad=0,
                  bint device=False):
    """Returns (action, action_lo, iterations)."""
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef double S = 0, S_lo = 0
    cdef long it = 0
    cdef int rc
    with nogil:
        if device:
            rc = qlat_dwf_fermion_force_device(p, <const double*>phi, stop_rsd, <double*>force, <double*>grad,
                                               &S, &S_lo, &it)
        else:
            rc = qlat_dwf_fermion_force(p, <const double*>phi, stop_rsd, <double*>force, <double*>grad,
                                        &S, &S_lo, &it)
    _check(rc, "qlat_dwf_fermion_force_device" if device else "qlat_dwf_fermion_force")
    return S, S_lo, it


def bench_kernel(uintptr_t h, int which, int precision, int iters):
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef double ms = 0
    cdef int rc
    with nogil:
        rc = qlat_dwf_bench_kernel(p, which, precision, iters, &ms)
    _check(rc, "qlat_dwf_bench_kernel")
    return ms


# -------------------------------------------------------------- sources
cdef void _source(qlat_dwf_source* s, int kind, xg, int tslice, lmom, value, uintptr_t fu1) except *:
    cdef int i
    s.kind = kind
    s.tslice = tslice
    for i in range(4):
        s.xg[i] = int(xg[i])
        s.lmom[i] = float(lmom[i])
    s.value[0] = float(value.real)
    s.value[1] = float(value.imag)
    s.fu1 = <const double*>fu1


def make_source(total_site, int kind, xg, int tslice, lmom, value, uintptr_t fu1, uintptr_t prop,
                bint device=False):
    cdef int ts[4]
    cdef qlat_dwf_source s
    cdef int rc
    _site4(total_site, ts)
    _source(&s, kind, xg, tslice, lmom, complex(value), fu1)
    with nogil:
        if device:
            rc = qlat_dwf_make_source_device(ts, &s, <double*>prop)
        else:
            rc = qlat_dwf_make_source(ts, &s, <double*>prop)
    _check(rc, "qlat_dwf_make_source_device" if device else "qlat_dwf_make_source")


def source_propagator(uintptr_t h, int kind, xg, int tslice, lmom, value, uintptr_t fu1, uintptr_t sol,
                      bint device=False):
    cdef qlat_dwf_inverter* p = _inv(h)
    cdef qlat_dwf_source s
    cdef long it = 0
    cdef int rc
    _source(&s, kind, xg, tslice, lmom, complex(value), fu1)
    with nogil:
        if device:
            rc = qlat_dwf_source_propagator_device(p, &s, <double*>sol, &it)
        else:
            rc = qlat_dwf_source_propagator(p, &s, <double*>sol, &it)
    _check(rc, "qlat_dwf_source_propagator_device" if device else "qlat_dwf_source_propagator")
    return it


def free_invert(total_site, double mass, double m5, momtwist, uintptr_t sol, uintptr_t src):
    cdef int ts[4]
    cdef double mt[4]
    cdef const double* pmt = NULL
    cdef int rc, i
    _site4(total_site, ts)
    if momtwist is not None:
        for i in range(4):
            mt[i] = float(momtwist[i])
        pmt = mt
    with nogil:
        rc = qlat_dwf_free_invert(ts, mass, m5, pmt, <double*>sol, <const double*>src)
    _check(rc, "qlat_dwf_free_invert")


# ------------------------------------------------------------------ RNG
def rng_fill(bint gaussian, midstate, uint64_t num_bytes, uint64_t index0, int64_t n_sites, int n_per_site,
             double a, double b, uintptr_t out):
    """qlat_dwf_rng_fill_uniform (a = upper, b = lower) or
    qlat_dwf_rng_fill_gaussian (a = center, b = sigma)."""
    cdef uint32_t mid[8]
    cdef int rc, i
    for i in range(8):
        mid[i] = midstate[i]
    with nogil:
        if gaussian:
            rc = qlat_dwf_rng_fill_gaussian(mid, num_bytes, index0, n_sites, n_per_site, a, b, <double*>out)
        else:
            rc = qlat_dwf_rng_fill_uniform(mid, num_bytes, index0, n_sites, n_per_site, a, b, <double*>out)
    _check(rc, "qlat_dwf_rng_fill_gaussian" if gaussian else "qlat_dwf_rng_fill_uniform")


# ------------------------------------------------------------------ QED
def multiply_m_dwf_qed(total_site, int ls, uintptr_t gf1, double mass, double m5, uintptr_t wick,
                       bint is_dagger, uintptr_t out, uintptr_t inp):
    cdef int ts[4]
    cdef int rc
    _site4(total_site, ts)
    with nogil:
        rc = qlat_dwf_multiply_m_dwf_qed(ts, ls, <const double*>gf1, mass, m5, <const double*>wick,
                                         1 if is_dagger else 0, <double*>out, <const double*>inp)
    _check(rc, "qlat_dwf_multiply_m_dwf_qed")


def cg_with_m_dwf_qed(total_site, int ls, uintptr_t gf1, double mass, double m5, uintptr_t wick,
                      bint is_dagger, double stop_rsd, long max_num_iter, uintptr_t out, uintptr_t inp):
    cdef int ts[4]
    cdef long it = 0
    cdef int rc
    _site4(total_site, ts)
    with nogil:
        rc = qlat_dwf_cg_with_m_dwf_qed(ts, ls, <const double*>gf1, mass, m5, <const double*>wick,
                                        1 if is_dagger else 0, stop_rsd, max_num_iter, <double*>out,
                                        <const double*>inp, &it)
    _check(rc, "qlat_dwf_cg_with_m_dwf_qed")
    return it


def invert_dwf_qed(total_site, int ls, uintptr_t gf1, double mass, double m5, uintptr_t wick,
                   bint is_dagger, double stop_rsd, long max_num_iter, uintptr_t out4d, uintptr_t in4d):
    cdef int ts[4]
    cdef long it = 0
    cdef int rc
    _site4(total_site, ts)
    with nogil:
        rc = qlat_dwf_invert_dwf_qed(ts, ls, <const double*>gf1, mass, m5, <const double*>wick,
                                     1 if is_dagger else 0, stop_rsd, max_num_iter, <double*>out4d,
                                     <const double*>in4d, &it)
    _check(rc, "qlat_dwf_invert_dwf_qed")
    return it
