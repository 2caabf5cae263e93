// Internal interface between the kernel file and the C-ABI host code.
#pragma once

#include "dwf_common.h"

namespace qdwf {

constexpr int kHopBlock = 256;
constexpr int kM5Block = 256;
constexpr int kBlasBlock = 256;
constexpr int kMaxBlasGrid = 2048;
constexpr int kHopM5Sites = 32;  // output half-sites per fused hop + s-matrix block
constexpr int kMaxRhs = 16;      // right-hand sides per multi-RHS launch (tile width)

// The columns a launch works on: column rhs[i] of a stack of fields spaced
// fstride slots apart, i < n.  A single field is {1, fstride, {0}}.
struct RhsSet {
  int n;
  int64_t fstride;
  int rhs[kMaxRhs];
};

template <class T>
struct HopArgs {
  LatGeo geo;
  int ls;
  int eo_out;          // parity of the output half field (2 even, 1 odd)
  int nchunk;          // set by the launcher
  T alpha;             // out = alpha * H in (+ beta_s * x)
  const void* in;      // SoA spinor, parity 3 - eo_out
  void* out;           // SoA spinor, parity eo_out
  const void* x;       // optional xpay operand (parity eo_out) or nullptr
  const cplx<T>* beta; // per-s xpay coefficients (device), used with x
  const void* links;   // 8-link layout (dwf_common.h)
  RhsSet cols;         // in/out/x are stacks of fields
};

template <class T>
struct M5Args {
  LatGeo geo;
  int ls;
  T alpha;            // out = alpha * Mat in + beta * x
  T beta;
  const cplx<T>* mat; // [2][ls][ls] (device)
  bool real_mat;      // all entries real: skip the imaginary parts
  const void* in;
  void* out;
  const void* x;      // optional
  const void* dotp;   // optional (with x): part[col*nblk + block] = Re <dotp, out>
  double* part;
  RhsSet cols;        // blockIdx.z = column index i
};

struct CgState;

// x += alpha p (x optional); p = r + beta p; t = Mat p
template <class T>
struct CgStepArgs {
  LatGeo geo;
  int ls;
  double alpha[kMaxRhs], beta[kMaxRhs];  // per column i
  RhsSet cols;                           // blockIdx.z = column index i
  const cplx<T>* mat;
  bool real_mat;
  void* x;
  void* p;
  void* r;
  void* t;
  // merged form (ap not null): r -= alpha_r ap before p is formed
  const void* ap;
  double alpha_r[kMaxRhs];
  // three-phase variant: column spacing of x and r (FP32 stacks), the
  // s-matrix of each column (columns may belong to different masses) and
  // the sites per block (16 or 32)
  int64_t xr_fstride;  // of r
  int64_t x_fstride;   // of x
  const cplx<T>* mats[kMaxRhs];
  int sb;
  // double-buffered p (three-phase variant): when p_out is not null the new
  // p goes there, and x += alpha p + alpha_prev q with q the p_out field's
  // old contents (the previous iteration's p), read before it is replaced;
  // x is not touched for a column whose two coefficients are zero
  void* p_out;
  double alpha_prev[kMaxRhs];
  // device-scalar form (graph-captured iterations): when cg is not null the
  // per-column alpha, alpha_r, beta and mode are read from it instead
  CgState* cg;
};

// Device copy of the multi-RHS CG scalars, indexed by column slot i (the
// i-th column of the launch).  k_cg_scalars turns the hermop reductions
// into alpha, beta and |r|^2, so a CG iteration needs no host round trip
// and can be captured in a graph.  mode: 0 iterating; 2 the column reached
// its reliable-update threshold this iteration (the step only adds
// alpha_q q + alpha p to x); 1 frozen (every kernel skips the column until
// the host has done the FP64 update and reset the mode).
struct CgState {
  double rr[kMaxRhs];       // |r|^2 after the last completed step
  double maxrr[kMaxRhs];    // |r|^2 at the last reliable update
  double alpha[kMaxRhs];    // alpha of the last iteration
  double rr_new[kMaxRhs];   // |r - alpha Ap|^2 of the last iteration
  double alpha_x[kMaxRhs];  // step coefficients of the next step
  double alpha_r[kMaxRhs];
  double beta[kMaxRhs];
  double alpha_q[kMaxRhs];  // x coefficient of the previous p (in the other p buffer)
  double pend[kMaxRhs];     // alpha of a direction not yet added to x
  double target;            // stop_rsd^2 (normalised columns)
  double d2;                // reliable-update factor delta^2
  int x_defer;              // x gets two directions every other step
  int mode[kMaxRhs];
  int iters[kMaxRhs];       // iterations done since the host last reset them
  int bad[kMaxRhs];         // non-finite alpha or residual
};

// FP32 fused hop + s-matrix: out = alpha * Mat (H in) + beta * x
struct HopM5Args {
  LatGeo geo;
  int ls;
  int eo_out;
  float alpha;
  float beta;
  const void* in;
  void* out;
  const void* x;  // optional
  const void* links;
  // FP16 links (h16x2 (re, im) in the 8-link layout): with FP16 spinor
  // storage, the hop phase runs in FP16 arithmetic (hop_term_h2); null keeps
  // FP32 arithmetic
  const void* links16;
  int hop_mode;  // with links16: non-zero selects FP16 arithmetic in the hop phase
  const cplx<float>* mat;  // [2][ls][ls]
  bool real_mat;
  const void* dotp;  // optional (dag xpay form): part[i*nblk + block] = Re <dotp, out>
  double* part;
  RhsSet cols;       // in/out/x/dotp are stacks of fields
  // three-phase DOT form with dotr (FP32 stack, dotr_fstride apart): also
  // part[(n + i)*nblk + block] = Re <dotr, out>, part[(2n + i)*nblk + block] = |out|^2
  const void* dotr;
  int64_t dotr_fstride;
  // three-phase variant: the s-matrix of each column (mats[0] null: plain
  // hop, no s-matrix), sites per block (8, 16 or 32), block order
  const cplx<float>* mats[kMaxRhs];
  int sb;
  int site_major;
  int plain64;  // the plain FP16-arithmetic hop on 64-site blocks (Ls <= 16)
  const CgState* cg;  // optional: blocks of columns with cg->mode != 0 do nothing
};

// FP64 fused hop + s-matrix (k_hop_m5d): out = alpha * Mat (H in) + beta * x
// on one FP64 field; with dotp (dagger xpay form) part[block] = Re <dotp, out>.
struct HopM5dArgs {
  LatGeo geo;
  int ls;
  int eo_out;
  double alpha, beta;
  const double2* in;
  double2* out;
  const double2* x;     // optional
  const double2* links; // FP64 8-link layout
  const cplx<double>* mat;  // [2][ls][ls]
  bool real_mat;
  const double2* dotp;  // optional
  double* part;
};
// false (nothing launched) when ls does not fit (4 <= ls <= 16, 16 ls a
// multiple of 64) or the form is not available
bool launch_hop_m5d(const HopM5dArgs& a, bool dag, hipStream_t st);
bool hop_m5d_ok(int ls);
int64_t hop_m5d_blocks(const LatGeo& g);

// r_i -= alpha_i ap_i with per-column partial |r_i|^2 (part[i*nblk + block])
template <class S>
struct AxpyNormArgs {
  const S* ap;
  S* r;
  int64_t n;  // slots per field
  double alpha[kMaxRhs];
  RhsSet cols;
  double* part;
};

// r_i -= alpha_i ap_i over 5D sites: ap in the solver's storage (cols.fstride
// apart), r FP32 (r_fstride float4 apart); partial |r_i|^2 per block
struct AxpyNormSiteArgs {
  const void* ap;
  void* r;
  int64_t n5;  // 5D sites per field including pads: stride * ls
  int ls;
  double alpha[kMaxRhs];
  RhsSet cols;
  int64_t r_fstride;
  double* part;
};

// Fermion force: A[p], B[p] are FP64 half fields by parity slot (0 even,
// 1 odd); links is the FP64 8-link layout; force/grad are [V][4][3][3]
// complex128 (grad optional).
struct ForceArgs {
  LatGeo geo;
  int ls;
  const void* A[2];
  const void* B[2];
  const void* links;
  void* force;
  void* grad;
};
void launch_force(const ForceArgs& a, hipStream_t st);

template <class T>
void launch_hop(const HopArgs<T>& a, bool dag, int spt, hipStream_t st);
// returns false (nothing launched) when ls is too large for the fused kernel
bool launch_hop_m5(const HopM5Args& a, bool dag, hipStream_t st);
int64_t hop_m5_blocks(const LatGeo& g);
// Three-phase kernels for FP32 or FP16 (per-site scale) spinor storage
// (half = true); false when ls does not fit them (6 <= ls <= 16, ls even).
bool three_phase_ok(int ls, int sb = 32);
// some sites-per-block choice fits ls
inline bool three_phase_ls_ok(int ls) { return three_phase_ok(ls, 32) || three_phase_ok(ls, 16); }
int64_t hop_m5s_blocks(const LatGeo& g, int sb);
bool launch_hop_m5s(const HopM5Args& a, bool dag, bool half, hipStream_t st);
// storage: 0 all FP32; 1 p, t (ap) FP16 and x, r FP32; 2 all FP16
bool launch_cg_step_s(const CgStepArgs<float>& a, int storage, hipStream_t st);
// Unit spin-colour propagator source times a site factor f(x), written in
// the reference layout [V][12][12] (kind: 0 point, 1 wall, 2 momentum,
// 3 U(1) field; mom in radians per lattice unit).
struct SrcArgs {
  int n[4];
  int kind;
  int xg[4];
  int tslice;
  double mom[4];
  double value_re, value_im;
  const double2* fu1;  // [V] complex (kind 3)
};
void launch_make_src(double2* prop, const SrcArgs& a, hipStream_t st);

// QED domain-wall operator (qed.cpp:263-411) on full 5D fields in the
// reference layout [V][ls][4] complex128, U(1) links gf [V][4], optional
// Wick phase factors per time slice.
struct QedArgs {
  int n[4];
  int ls;
  double mass, m5;
  const double2* gf;
  const double2* wick;  // [T] or null
  int dagger;
  const double2* in;
  double2* out;
};
void launch_qed_m(const QedArgs& a, hipStream_t st);

// Free-field domain-wall propagator (free_invert, qcd-prop.h:292-359) on
// propagators [V][12][12]: a DFT along direction mu (sign -1 forward,
// +1 backward, unnormalised), and the momentum-space multiply.
void launch_dft_dir(double2* out, const double2* in, const int n[4], int mu, int sign, hipStream_t st);
void launch_free_mom(double2* prop, const int n[4], double mass, double m5, const double momtwist[4], double scale,
                     hipStream_t st);

// set_u_rand (core.h:1584-1603) on the device: out[g][m] for sites g <
// n_sites and reals m < n_per_site is u * 2^-64 * scale + lower, u the m-th
// 64-bit draw of the RngState of site g (the state's SHA-256 midstate `mid`
// over its num_bytes-byte message, type g, draws from index0 + 1).
struct RngFillArgs {
  uint32_t mid[8];
  uint64_t num_bytes;
  uint64_t index0;
  int64_t n_sites;
  int n_per_site;
  double scale, lower;
  double* out;
};
void launch_rng_fill_u(const RngFillArgs& a, hipStream_t st);
// set_g_rand (core.h:1619-1640): out[g][m] = g_rand_gen(center, sigma) of the
// same per-site state, one thread per site (scale = sigma, lower = center).
void launch_rng_fill_g(const RngFillArgs& a, hipStream_t st);

// Block-compressed low modes (deflate, dslash.cpp:144-249).  hmap[b][bidx]
// is the device half-site of half site bidx of block b; field elements are
// FP64 SoA half fields of parity eo.
struct DeflArgs {
  LatGeo geo;
  int ls;
  int eo;
  int bs[4];         // block_site
  int nblock;
  int block_vol_eo;
  int64_t bsize;     // block_vol_eo * ls * 12
  int n_basis, n_vec;
  const float2* basis;  // [nblock][n_basis][bsize]
  const float2* coef;   // [nblock][n_vec][n_basis]
  const double* evals;  // [n_vec]
  const int* hmap;      // [nblock][block_vol_eo]
  const double2* in;
  double2* out;
  double2* part;        // [nblock][n_vec]
  double2* p;           // [n_vec]
};
void launch_defl_hmap(int* hmap, const DeflArgs& a, hipStream_t st);
void launch_deflate(const DeflArgs& a, hipStream_t st);

// alpha, beta, |r|^2 and modes of the n column slots from the merged hermop
// reductions red[i] = Re <p, Ap>, red[n + i] = Re <r, Ap>, red[2n + i] = |Ap|^2
void launch_cg_scalars(CgState* cg, const double* red, int n, hipStream_t st);
int launch_axpy_norm_site(const AxpyNormSiteArgs& a, int storage, hipStream_t st);
void launch_axpy_site(double alpha, const void* p, void* x, int ls, int64_t n5, int storage, hipStream_t st);
// FP64 blocked field <-> FP16 column over n5 = stride * ls 5D sites
void launch_d2h(const double2* in, uint4* out, int ls, int64_t n5, double scale, hipStream_t st);
void launch_h2d(const uint4* in, double2* out, int ls, int64_t n5, int acc, double scale, hipStream_t st);
void launch_hop_h16(const HopArgs<float>& a, bool dag, hipStream_t st);
template <class T>
void launch_m5(const M5Args<T>& a, hipStream_t st);
template <class T>
void launch_cg_step(const CgStepArgs<T>& a, hipStream_t st);
template <class S>
void launch_axpy_norm(double a, const S* ap, S* r, int64_t n, double* part, double* out, hipStream_t st);
void launch_reduce(const double* part, int nblk, int nv, double* out, hipStream_t st);
// out[i] = sum_b part[i*nblk + b], i < nrows
void launch_reduce_rows(const double* part, int nblk, int nrows, double* out, hipStream_t st);
template <class S>
int launch_axpy_norm_multi(const AxpyNormArgs<S>& a, double* out, hipStream_t st);

template <class S>
void launch_norm2(const S* a, int64_t n, double* part, double* out, hipStream_t st);
template <class S>
void launch_dot(const S* a, const S* b, int64_t n, double* part, double* out, hipStream_t st);
// out[0] + out[1] = sum 2 Re conj(phi) x - |y|^2 with a compensated (twice-FP64)
// sum; out[0] is it rounded to FP64
void launch_action_dd(const double2* phi, const double2* x, const double2* y, int64_t n, double* part, double* out,
                      hipStream_t st);
template <class S>
void launch_axpy(double a, const S* x, S* y, int64_t n, hipStream_t st);
template <class S>
void launch_xpay(const S* x, double a, S* y, int64_t n, hipStream_t st);
template <class S>
void launch_cg_update(double a, const S* p, const S* ap, S* x, S* r, int64_t n, double* part, double* out,
                      hipStream_t st);

void launch_copy(const float4* x, float4* y, int64_t n, hipStream_t st);
// FP64 <-> FP32 blocked fields: out = scale * in (d2f); out (+)= scale * in (f2d, acc = 1 adds)
void launch_d2f(const double2* in, float4* out, int ls, int64_t st, hipStream_t s, double scale = 1.0);
void launch_f2d(const float4* in, double2* out, int ls, int64_t st, int acc, hipStream_t s, double scale = 1.0);

void launch_pack5(const double2* aos, double2* soa, const LatGeo& g, int ls, int eo, int full, hipStream_t st);
void launch_unpack5(const double2* soa, double2* aos, const LatGeo& g, int ls, int eo, int full, hipStream_t st);
void launch_prop_col_to_5d(const double2* src, double2* soa, const LatGeo& g, int ls, int eo, int col,
                           hipStream_t st);
void launch_5d_to_prop_col(const double2* soa, double2* dst, const LatGeo& g, int ls, int eo, int col,
                           hipStream_t st);
// l16 (optional): FP16 copy of the links, h16x2 (re, im) per element
void launch_pack_gauge(const double2* gauge, double2* l64, float2* l32, unsigned* l16, const LatGeo& g,
                       hipStream_t st);

}  // namespace qdwf
