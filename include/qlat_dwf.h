/*
 * qlat_dwf.h -- C ABI of the MI355X (gfx950) domain-wall fermion library.
 *
 * Drop-in replacement for the even-odd Moebius domain-wall operator and its
 * CG inverter of Qlattice (qlat/qlat/lib/dslash.cpp,
 * qlat/qlat/include/qlat/dslash.h).  All entry points take plain pointers and
 * sizes; no C++ or framework types cross this boundary.
 *
 * Field layouts at the boundary are the reference's (complex128 stored as
 * interleaved re,im doubles, array of structures):
 *   gauge field  [V][4][3][3]      GaugeField, core.h:1187
 *   5D full      [V][Ls][12]       FermionField5d, core.h:1203; 12 = spin*3+colour
 *   5D half      [V/2][Ls][12]     eo half field, half index = full index / 2
 *                                  (core.h:232-261); parity eo: 2 even, 1 odd
 *   4D fermion   [V][12]           FermionField4d
 *   propagator   [V][12][12]       Propagator4d, [site][row][source column]
 * with site index x + X*(y + Y*(z + Z*t)) (coordinate.h:169-188).
 *
 * Functions without a `_device` suffix take HOST pointers: the library
 * uploads, converts to its device layout, computes on the GPU and copies the
 * result back before returning.  `_device` functions take pointers to device
 * memory holding the same layouts and leave the result in device memory.
 * Every call is ordered on the inverter's own HIP stream and synchronises it
 * before returning.
 *
 * Errors: every function returns QLAT_DWF_OK (0) or a negative status and
 * records a message retrievable with qlat_dwf_last_error() (thread local).
 * This mirrors the reference, whose Qassert throws std::string
 * (qassert.h:77-93) and whose Python layer turns it into RuntimeError
 * (cqlat/lib.h:28-35).  Non-convergence is not an error: as in cg_with_f
 * (dslash.h:394-398) the iteration count is then max_num_iter + 1.
 */
#ifndef QLAT_DWF_H
#define QLAT_DWF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QLAT_DWF_OK 0
#define QLAT_DWF_EINVAL (-1)
#define QLAT_DWF_EHIP (-2)
#define QLAT_DWF_ENOMEM (-3)
#define QLAT_DWF_ESTATE (-4)

/* Solver of the odd-site normal equations (qlat_dwf_inverter_set_solver). */
#define QLAT_DWF_SOLVER_FP64 0       /* the reference algorithm, FP64 throughout */
#define QLAT_DWF_SOLVER_MIXED_FP32 1 /* FP32 CG with FP64 reliable updates      */
#define QLAT_DWF_SOLVER_MIXED_FP16 2 /* as MIXED_FP32, Krylov vectors stored in FP16
                                        with a per-5D-site scale (ls even in
                                        [6, 16], or 20 or 24; FP32 storage
                                        otherwise)                                 */

typedef struct qlat_dwf_inverter qlat_dwf_inverter;

/* FermionAction (fermion-action.h:8-53).  bs/cs point to `ls` complex
 * numbers (re,im pairs); when NULL they are derived from mobius_scale as
 * b = 0.5*scale + 0.5, c = 0.5*scale - 0.5 (fermion-action.h:50-51). */
typedef struct {
  double mass;
  int ls;
  double m5;
  double mobius_scale;
  int is_multiplying_dminus; /* default 1 */
  int is_using_zmobius;
  int cg_diagonal_mee;       /* only 2 (sym2 preconditioning) is supported, as in dslash.h:483-521 */
  const double* bs;
  const double* cs;
} qlat_dwf_fermion_action;

/* Statistics of the last solve (accumulated over columns for propagators). */
typedef struct {
  long total_iter;        /* inner CG iterations                                 */
  long cycles;            /* outer restart cycles (FP64 reference algorithm)     */
  long reliable_updates;  /* FP64 residual replacements (mixed solver)           */
  long hermop_fp64;       /* FP64 applications of Mpc^+ Mpc                      */
  long hermop_low;        /* FP32 applications of Mpc^+ Mpc                      */
  double true_residual;   /* max over columns of |b' - A x| / |b'| in FP64 (odd) */
  double seconds;         /* wall clock of the solve call                        */
  int spinor_bits;        /* storage of the low-precision Krylov vectors (16/32),
                             0 for the FP64 solver                              */
  int rhs_tile;           /* columns per multi-RHS block of the last batched solve */
  int sites_per_block;    /* output sites per block of its hop kernels            */
  double true_residual_5d; /* max over columns of |M x - D_- b| / |D_- b| on the full
                              5D system, the is_checking_invert check of
                              invert_with_cg (dslash.h:522-528); 0 when the check
                              is off (qlat_dwf_set_checking_invert)              */
  int hop_arith_bits;     /* arithmetic of the hop phase of its low-precision
                             kernels: 16 (FP16 links, packed FP16 projection,
                             FP16 dot products into FP32 sums) or 32; 0 for the
                             FP64 solver                                        */
} qlat_dwf_solve_stats;

/* Message of the last failed call on this thread ("" if none). */
const char* qlat_dwf_last_error(void);

/* ABI version of this header (incremented on incompatible changes). */
int qlat_dwf_abi_version(void);

/* ---------------------------------------------------------------- setup */

/* InverterDomainWall::setup(gf, fa) (dslash.h:154-162) and setup_inverter
 * (dslash.h:199-216): the inverter keeps its own device copy of the gauge
 * field.  total_site: lattice extents X Y Z T (all even). */
int qlat_dwf_inverter_create(qlat_dwf_inverter** inv, const int total_site[4], const double* gauge,
                             const qlat_dwf_fermion_action* fa);

/* ~InverterDomainWall; Python InverterDomainWall.__del__ (inverter.py:92-94). */
int qlat_dwf_inverter_destroy(qlat_dwf_inverter* inv);

/* InverterDomainWall::stop_rsd(), max_num_iter(), max_mixed_precision_cycle()
 * (dslash.h:186-196); defaults 1e-8, 200, 300 (InverterParams, dslash.h:114-133). */
int qlat_dwf_inverter_set_params(qlat_dwf_inverter* inv, double stop_rsd, long max_num_iter,
                                 long max_mixed_precision_cycle);
int qlat_dwf_inverter_get_params(const qlat_dwf_inverter* inv, double* stop_rsd, long* max_num_iter,
                                 long* max_mixed_precision_cycle);

/* Choose the inner solver, the role of a back-end's cg_with_herm_sym_2
 * override (grid.h:460-484).  reliable_delta: residual reduction between
 * FP64 residual replacements of the mixed solvers (0 selects 0.01). */
int qlat_dwf_inverter_set_solver(qlat_dwf_inverter* inv, int solver, double reliable_delta);

int qlat_dwf_inverter_get_stats(const qlat_dwf_inverter* inv, qlat_dwf_solve_stats* stats);

/* is_checking_invert() (dslash.h:11-16; on by default): after each solve
 * compute |M x - D_- b| / |D_- b| on the full 5D system (invert_with_cg,
 * dslash.h:522-528) into stats.true_residual_5d.  Process-wide, like the
 * reference's static flag.  qlat_dwf_is_checking_invert returns 0 or 1. */
int qlat_dwf_set_checking_invert(int on);
int qlat_dwf_is_checking_invert(void);

/* ------------------------------------------- operators on 5D half fields
 * `out` and `in` are [V/2][Ls][12]; out may equal in ("out can be the same
 * object as in", dslash.cpp:593). */

/* multiply_m_e_e (dslash.h:244-245, :285-286; dslash.cpp:591-639) */
int qlat_dwf_multiply_m_e_e(qlat_dwf_inverter* inv, double* out, const double* in);
/* multiply_mdag_e_e (dslash.h:247-248, :288-289; dslash.cpp:641-692) */
int qlat_dwf_multiply_mdag_e_e(qlat_dwf_inverter* inv, double* out, const double* in);
/* multiply_m_e_e_inv (dslash.h:250-251, :291-292; dslash.cpp:694-769) */
int qlat_dwf_multiply_m_e_e_inv(qlat_dwf_inverter* inv, double* out, const double* in);
/* multiply_mdag_e_e_inv (dslash.h:253-254, :294-295; dslash.cpp:771-857) */
int qlat_dwf_multiply_mdag_e_e_inv(qlat_dwf_inverter* inv, double* out, const double* in);
/* multiply_wilson_d_e_o_no_comm (dslash.h:256-258; dslash.cpp:859-908):
 * input parity eo_in, output parity 3 - eo_in. */
int qlat_dwf_multiply_wilson_d_e_o(qlat_dwf_inverter* inv, double* out, const double* in, int eo_in);
/* multiply_wilson_ddag_e_o_no_comm (dslash.h:260-262; dslash.cpp:910-959) */
int qlat_dwf_multiply_wilson_ddag_e_o(qlat_dwf_inverter* inv, double* out, const double* in, int eo_in);
/* multiply_m_e_o (dslash.h:264-265, :297-298; dslash.cpp:961-1003) */
int qlat_dwf_multiply_m_e_o(qlat_dwf_inverter* inv, double* out, const double* in, int eo_in);
/* multiply_mdag_e_o (dslash.h:267-268, :300-301; dslash.cpp:1005-1060) */
int qlat_dwf_multiply_mdag_e_o(qlat_dwf_inverter* inv, double* out, const double* in, int eo_in);
/* multiply_mpc_sym2 (dslash.h:276-277, :317-318; dslash.cpp:1102-1119); eo = parity of in and out */
int qlat_dwf_multiply_mpc_sym2(qlat_dwf_inverter* inv, double* out, const double* in, int eo);
/* multiply_mpcdag_sym2 (dslash.h:279-280, :320-321; dslash.cpp:1121-1138) */
int qlat_dwf_multiply_mpcdag_sym2(qlat_dwf_inverter* inv, double* out, const double* in, int eo);
/* multiply_hermop_sym2 (dslash.h:282-283, :323-324; dslash.cpp:1140-1149) */
int qlat_dwf_multiply_hermop_sym2(qlat_dwf_inverter* inv, double* out, const double* in, int eo);

/* multiply_hermop_sym2 evaluated in FP32 (input rounded to FP32, result
 * widened to FP64).  No counterpart in the reference: it is the operator
 * the inner loop of the mixed-precision solver applies, exposed so that
 * callers (and tests) can check it at FP32 accuracy. */
int qlat_dwf_multiply_hermop_sym2_fp32(qlat_dwf_inverter* inv, double* out, const double* in, int eo);

/* ------------------------------------------- operators on 5D full fields */

/* multiply_m (dslash.h:270-271, :311-312; dslash.cpp:1062-1080) */
int qlat_dwf_multiply_m(qlat_dwf_inverter* inv, double* out, const double* in);
/* multiply_mdag (dslash.h:273-274, :314-315; dslash.cpp:1082-1100) */
int qlat_dwf_multiply_mdag(qlat_dwf_inverter* inv, double* out, const double* in);
/* multiply_d_minus (dslash.h:227-231; dslash.cpp:448-488) */
int qlat_dwf_multiply_d_minus(qlat_dwf_inverter* inv, double* out, const double* in);

/* ------------------------------------------------------------ BLAS-1 */

/* dot_product (dslash.h:329; dslash.cpp:1298-1321): a^+ b over n_wv
 * WilsonVectors (12 complex each). */
int qlat_dwf_dot_product(qlat_dwf_inverter* inv, const double* a, const double* b, long n_wv, double* re,
                         double* im);
/* qnorm (field.h:170-195) */
int qlat_dwf_qnorm(qlat_dwf_inverter* inv, const double* a, long n_wv, double* result);

/* ------------------------------------------------------------ solvers */

/* cg_with_herm_sym_2 (dslash.h:575-578; dslash.cpp:1323-1333): CG on
 * Mpc^+ Mpc over odd half fields; `sol` holds the initial guess on entry.
 * Always runs the FP64 reference algorithm. */
int qlat_dwf_cg_with_herm_sym_2(qlat_dwf_inverter* inv, double* sol, const double* src, double stop_rsd,
                                long max_num_iter, long* iter);

/* invert(FermionField5d&, const FermionField5d&, const InverterDomainWall&)
 * (dslash.h:580-581; dslash.cpp:1335-1340 -> invert_with_cg, dslash.h:440-530) */
int qlat_dwf_invert_5d(qlat_dwf_inverter* inv, double* out, const double* in, long* iter);

/* invert_with_cg_with_guess (dslash.h:532-573) with the inverter's solver:
 * `out` ([V][Ls][12], full 5D) holds the initial guess x0 on entry; the
 * correction is solved from D_- in - M x0 with stop_rsd scaled by
 * |D_- in| / |D_- in - M x0| and added to out.  iter: the correction's
 * iterations (0 when x0 solves the system exactly). */
int qlat_dwf_invert_5d_with_guess(qlat_dwf_inverter* inv, double* out, const double* in, long* iter);

/* invert(FermionField4d&, ...) (dslash.h:583-584; dslash.cpp:1342-1347 ->
 * invert_dwf, qcd-prop.h:146-167) */
int qlat_dwf_invert_4d(qlat_dwf_inverter* inv, double* out, const double* in, long* iter);

/* invert(Propagator4d&, const Propagator4d&, const Inverter&)
 * (qcd-prop.h:169-186) as reached from Python InverterDomainWall.__mul__
 * (inverter.py:96-109 -> cqlat/inverter.cpp:23-36). */
int qlat_dwf_invert_prop(qlat_dwf_inverter* inv, double* sol, const double* src, long* iter);

/* ------------------------------------------------------------- sources
 * Propagator sources built on the device (qcd-prop.cpp:61-150, :260-307,
 * :465-530).  Every kind is a unit spin-colour matrix times a site factor
 * f(x): column cs of the propagator is the fermion field with f(x) in
 * component cs (set_propagator_col_from_fermion_field, qcd-prop.h:55-68).
 * mom = lmom * 2 pi / total_site (lattice_mom_mult, field.h:14-22) and
 * phase = sum_mu mom[mu] * x[mu]. */
#define QLAT_DWF_SRC_POINT 0 /* f = value at x = xg (set_point_src, qcd-prop.cpp:61-89) */
#define QLAT_DWF_SRC_WALL 1  /* f = exp(i phase) on t = tslice, lmom[3] = 0 (set_wall_src,
                                set_tslice_mom_src: qcd-prop.cpp:117-150, :465-499) */
#define QLAT_DWF_SRC_MOM 2   /* f = exp(i phase) everywhere (set_mom_src, qcd-prop.cpp:275-307;
                                set_volume_src, :501-530, requires lmom[3] = 0) */
#define QLAT_DWF_SRC_U1 3    /* f = fu1(x), V complex values in reference site order
                                (set_rand_vol_u1_src, qcd-prop.cpp:260-273) */

typedef struct {
  int kind;
  int xg[4];          /* POINT */
  int tslice;         /* WALL */
  double lmom[4];     /* WALL, MOM */
  double value[2];    /* POINT: re, im */
  const double* fu1;  /* U1: host pointer for the host entry points, device pointer for _device */
} qlat_dwf_source;

/* The source propagator [V][12][12] on a lattice of total_site
 * (set_point_src, set_wall_src, set_mom_src, set_rand_vol_u1_src;
 * propagator.pyx:175-234 mk_point_src, mk_wall_src, mk_rand_vol_u1_src). */
int qlat_dwf_make_source(const int total_site[4], const qlat_dwf_source* src, double* prop);

/* set_point_src_propagator / set_wall_src_propagator /
 * set_mom_src_propagator / set_tslice_mom_src_propagator (qcd-prop.h:202-290,
 * :403-420): the source is built on the device and its 12 columns are solved
 * together (invert(Prop), as qlat_dwf_invert_prop); sol is [V][12][12]. */
int qlat_dwf_source_propagator(qlat_dwf_inverter* inv, const qlat_dwf_source* src, double* sol, long* iter);

/* ---------------------------------------------------------- low modes
 * LowModes (dslash.h:43-85) held by an inverter, as InverterDomainWall::setup
 * (gf, fa, lm) does (dslash.h:176-186), in the in-memory layout of the
 * reference's compressed eigen system (compressed-eigen-io.h:175-195,
 * :332-339), single node:
 *   blocks of block_site (each entry even and dividing total_site), block b =
 *   lexicographic index (x fastest) of x / block_site; inside a block the
 *   half site bidx = (lexicographic index of x % block_site) / 2, and
 *   element k = (bidx * Ls + s) * 12 + spin*3+colour (convert_half_vector,
 *   commpressed-eigen-io.cpp:925-947); block_size = block_vol/2 * Ls * 12;
 *   bases  [n_block][n_basis][block_size]  complex64
 *   coefs  [n_block][n_vec][n_basis]       complex64
 *   eigen_values [n_vec] (of multiply_hermop_sym2 on the odd sites).
 * n_vec = 0 detaches the low modes.  With low modes attached, every restart
 * cycle of the FP64 solver starts CG from deflate(residual) (dslash.h:496-500). */
int qlat_dwf_inverter_set_low_modes(qlat_dwf_inverter* inv, const int block_site[4], int n_basis, int n_vec,
                                    const double* eigen_values, const float* bases, const float* coefs);

/* deflate(FermionField5d& out, const FermionField5d& in, LowModes&)
 * (dslash.cpp:144-300) on a half field of parity eo: the input is rounded to
 * FP32 and, as in the reference (whose blocked input buffer is also the
 * accumulator of the produce step), out = in + sum_i v_i <v_i, in> / lambda_i;
 * out = 0 when no low modes are attached.  out may equal in. */
int qlat_dwf_deflate(qlat_dwf_inverter* inv, double* out, const double* in, int eo);

/* free_invert (qcd-prop.h:347-359, free_mom_invert :292-345): the
 * infinite-Ls domain-wall propagator on the unit gauge field applied to a
 * propagator [V][12][12] in momentum space (InverterDwfFreeField,
 * inverter.py:26-71); momtwist may be NULL (no twist).  sol may equal src. */
int qlat_dwf_free_invert(const int total_site[4], double mass, double m5, const double momtwist[4], double* sol,
                         const double* src);

/* set_u_rand (qlat core.h:1584-1603, RngState rng-state.cpp:148-220) on the
 * GPU: out[g * n_per_site + m] = u * 2^-64 * (upper - lower) + lower for
 * sites g < n_sites and m < n_per_site, u the m-th rand_gen draw of
 * rs.newtype(g), where rs is described by the SHA-256 midstate of its
 * message (midstate, num_bytes: a multiple of 64) and its index (index0).
 * One SHA-256 block per four draws; out is host memory.  Each call
 * allocates and frees a device buffer of n_sites * n_per_site doubles. */
int qlat_dwf_rng_fill_uniform(const uint32_t midstate[8], uint64_t num_bytes, uint64_t index0, int64_t n_sites,
                              int n_per_site, double upper, double lower, double* out);

/* set_g_rand (qlat core.h:1619-1640, g_rand_gen rng-state.cpp:222-258) on
 * the GPU with the same state description: out[g * n_per_site + m] is the
 * m-th g_rand_gen(center, sigma) of rs.newtype(g), one thread per site
 * drawing that site's n_per_site values in sequence (the reference's
 * per-site state: cached words and spare Gaussian).  The logarithm of the
 * polar method is the device's, so values can differ from the host's in the
 * last bit.  Each call allocates and frees a device buffer of
 * n_sites * n_per_site doubles. */
int qlat_dwf_rng_fill_gaussian(const uint32_t midstate[8], uint64_t num_bytes, uint64_t index0, int64_t n_sites,
                               int n_per_site, double center, double sigma, double* out);

/* ------------------------------------------------------------------ QED
 * The QED domain-wall operator and its solvers (qed.h:433-463,
 * qed.cpp:119-411): full (not even-odd) 5D fields [V][ls][4] complex128 in
 * reference site order, U(1) links gf1 [V][4] complex128 (any non-zero
 * complex values: the backward link is 1/gf1(x - mu)), optional Wick phase
 * factors wick[T] complex128 (NULL: all 1), periodic in every direction.
 * Stateless: each call moves its operands to the device and back. */
/* multiply_m_dwf_qed (qed.cpp:263-411); is_dagger selects M^+.  out != in. */
int qlat_dwf_multiply_m_dwf_qed(const int total_site[4], int ls, const double* gf1, double mass, double m5,
                                const double* wick, int is_dagger, double* out, const double* in);
/* cg_with_m_dwf_qed (qed.cpp:178-261): M^+ M out = in (M M^+ when
 * is_dagger) from the initial guess in out; iter = iterations or
 * max_num_iter + 1. */
int qlat_dwf_cg_with_m_dwf_qed(const int total_site[4], int ls, const double* gf1, double mass, double m5,
                               const double* wick, int is_dagger, double stop_rsd, long max_num_iter, double* out,
                               const double* in, long* iter);
/* invert_dwf_qed (qed.cpp:119-150): 4D fields [V][4]; M out = in (M^+ when
 * is_dagger) through the 5D solve with the chiral 4D <-> 5D embedding. */
int qlat_dwf_invert_dwf_qed(const int total_site[4], int ls, const double* gf1, double mass, double m5,
                            const double* wick, int is_dagger, double stop_rsd, long max_num_iter, double* out4d,
                            const double* in4d, long* iter);

/* The power iteration of find_max_eigen_value_hermop_sym2 (dslash.cpp:1349-1369)
 * from a caller-supplied odd half-field start vector. */
int qlat_dwf_find_max_eigen_value_hermop_sym2(qlat_dwf_inverter* inv, const double* start, long max_iter,
                                              double* result);

/* find_max_eigen_value_hermop_sym2(inv, rs, max_iter) (dslash.h:586-588;
 * dslash.cpp:1349-1369): the start vector is set_u_rand(ff, rs) on the odd
 * half field (uniform in [-1, 1), site x drawing from rs.newtype(global index
 * of x), core.h:1584-1603), drawn on the GPU; rs is described as for
 * qlat_dwf_rng_fill_uniform (SHA-256 midstate of its message, message
 * length num_bytes, index0). */
int qlat_dwf_find_max_eigen_value_hermop_sym2_rng(qlat_dwf_inverter* inv, const uint32_t midstate[8],
                                                  uint64_t num_bytes, uint64_t index0, long max_iter,
                                                  double* result);

/* ------------------------------------------------------------- batches
 * No single counterpart in the reference, which solves one right-hand side
 * at a time with one InverterDomainWall per fermion action
 * (invert(FermionField5d&, ...), dslash.h:580-581, dslash.cpp:1335-1341).
 * sol[j] = invert(src[j]) with inverter invs[j], j < n: full 5D fields in
 * reference layout ([V][Ls][12] complex128).  All inverters must live on one
 * device and share the lattice, Ls and gauge field; they may differ in mass,
 * M5 and b, c.  Solver, stop_rsd and iteration limits are those of invs[0].
 * With the mixed solvers the systems go through the multi-RHS solver in tiles
 * (autotuned width, or QLAT_DWF_RHS_TILE), each column with its own
 * operator, reliable updates and retirement; with the FP64 solver they are
 * solved one after the other.  iters[j] receives each system's CG
 * iterations; statistics are collected on invs[0]. */
int qlat_dwf_invert_5d_batch(qlat_dwf_inverter* const* invs, int n, double* const* sol, const double* const* src,
                             long* iters);
int qlat_dwf_invert_5d_batch_device(qlat_dwf_inverter* const* invs, int n, double* const* sol,
                                    const double* const* src, long* iters);

/* ---------------------------------------------------------------- force
 * No counterpart in the reference (it has no fermion force); the
 * conventions are those of its gauge force (hmc.cpp:64-75,
 * matrix-hmc.h:202-213).
 *
 * S = phi^+ (Mpc^+ Mpc)^-1 phi for an odd half field phi ([V/2][Ls][12]);
 * X = (Mpc^+ Mpc)^-1 phi is solved with the FP64 CG to stop_rsd.  Outputs
 * (all [V][4][3][3] complex128):
 *   force[x][mu] = T_A(U_mu(x) G_mu(x)),  T_A(M) = (M - M^+)/2 - Tr(M - M^+)/6,
 *   grad[x][mu]  = G_mu(x)^+  (optional, may be NULL),
 * where dS = sum Re Tr[G_mu(x) dU_mu(x)].  For U -> exp(eps T) U with T in
 * su(3), dS/deps = sum Tr[force T]; grad is the derivative with respect to
 * the complex link entries in the conjugate-Wirtinger convention
 * (dS = Re sum conj(grad) dU).  action receives S, evaluated as
 * 2 Re <phi, X> - |Mpc X|^2 (second order in the solver error) with a
 * compensated sum; action_lo (optional, may be NULL) receives S - action, the
 * low-order part of that sum, so that differences of nearby actions keep
 * their digits.  iter receives the CG iterations. */
int qlat_dwf_fermion_force(qlat_dwf_inverter* inv, const double* phi, double stop_rsd, double* force, double* grad,
                           double* action, double* action_lo, long* iter);
int qlat_dwf_fermion_force_device(qlat_dwf_inverter* inv, const double* phi, double stop_rsd, double* force,
                                  double* grad, double* action, double* action_lo, long* iter);

/* Replace the inverter's gauge field (InverterDomainWall::setup(gf, fa)
 * again, dslash.h:154-162) without rebuilding anything else; gauge is
 * [V][4][3][3] complex128 in host (on_device = 0) or device memory. */
int qlat_dwf_inverter_set_gauge(qlat_dwf_inverter* inv, const double* gauge, int on_device);

/* ------------------------------------------ device-resident variants
 * Same semantics as above; all field pointers are device memory. */
int qlat_dwf_multiply_hermop_sym2_device(qlat_dwf_inverter* inv, double* out, const double* in, int eo);

/* qlat_dwf_multiply_hermop_sym2_device that allocates nothing: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its five FP64 work fields live in a static device
 * array of 64 MiB per GPU (held until the process ends), so the lattice must
 * satisfy 5 * 192 * Ls * (V/2 rounded up to 256) <= 64 MiB (e.g. 8^3 x 16 at
 * Ls 16); larger ones are rejected with QLAT_DWF_EINVAL.  The inverter
 * itself must have been created beforehand (creation allocates).  Calls on
 * one GPU share the array: calls that overlap on one GPU (from several
 * threads) race. */
int qlat_dwf_multiply_hermop_sym2_device_noalloc(qlat_dwf_inverter* inv, double* out, const double* in, int eo);
int qlat_dwf_invert_5d_device(qlat_dwf_inverter* inv, double* out, const double* in, long* iter);
int qlat_dwf_invert_5d_with_guess_device(qlat_dwf_inverter* inv, double* out, const double* in, long* iter);
int qlat_dwf_invert_prop_device(qlat_dwf_inverter* inv, double* sol, const double* src, long* iter);
int qlat_dwf_make_source_device(const int total_site[4], const qlat_dwf_source* src, double* prop);
int qlat_dwf_source_propagator_device(qlat_dwf_inverter* inv, const qlat_dwf_source* src, double* sol,
                                      long* iter);

/* --------------------------------------------------------- measurement
 * Average duration (ms) of one launch of a kernel, timed with HIP events
 * on the inverter's stream over `iters` back-to-back launches after one
 * warm-up launch.  which: 0 hop H, 1 hop H^+, 2 dense s-matrix (M_ee^-1),
 * 3 Mpc^+ Mpc (all its kernels), 4 plain copy of one field (precision 64
 * or 32); on the multi-RHS tile of the three-phase solver (QLAT_DWF_RHS_TILE
 * columns, precision 16 = FP16 storage or 32): 5 hop + s-matrix (b = K H t),
 * 6 CG step (x += a p, p = r + b p, t = K p), 7 Mpc^+ Mpc with the CG dot.
 * Has no counterpart in the reference (its Timer, timer.h:290, is host
 * wall clock); it exists to report kernel bandwidth. */
int qlat_dwf_bench_kernel(qlat_dwf_inverter* inv, int which, int precision, int iters, double* avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* QLAT_DWF_H */
