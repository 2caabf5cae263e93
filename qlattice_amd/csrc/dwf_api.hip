// Host side of the gfx950 domain-wall library: the C ABI of
// include/qlat_dwf.h, the inverter state, the composition of the
// even-odd operators from the kernels in dwf_kernels.hip, and the solvers.
//
// Operator factorisation used on the device (all algebra from
// qlat/qlat/lib/dslash.cpp):
//   M_ee  = diag(b~) - diag(c~) S          (dslash.cpp:591-639)
//   M_eo  = H P,  P = diag(b) + diag(c) S  (dslash.cpp:961-1003)
//   M_eo^+ = P^+ H^+                       (dslash.cpp:1005-1060)
//   Mpc   = 1 - H K H K,        K   = P M_ee^-1      (dslash.cpp:1102-1119)
//   Mpc^+ = 1 - K^+ H^+ K^+ H^+, K^+ = M_ee^-+ P^+   (dslash.cpp:1121-1138)
// where H is the Wilson hop and S the fifth-dimension shift with the mass
// corners.  Every s-operator (M_ee, its inverse, P, K, ...) is one dense
// Ls x Ls complex matrix per chirality, built here in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <memory>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/qlat_dwf.h"
#include "dwf_internal.h"

using qdwf::cplx;
using qdwf::LatGeo;
typedef std::complex<double> cd;

// Scratch of the allocation-free entry point
// (qlat_dwf_multiply_hermop_sym2_device_noalloc): the five FP64 half fields
// of one Mpc^+ Mpc application (input and output in the device layout and
// the operator's three temporaries).  Each GPU holds its own copy for the
// life of the process; host code uses only the per-device address obtained
// with hipGetSymbolAddress (g_scratch_ptr), never the array's name.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

thread_local std::string g_last_error;
// > 0 inside an allocation-free entry point: dev_alloc throws instead of
// allocating (a guard against any path that would reach hipMalloc)
thread_local int g_noalloc_depth = 0;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct MemError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define QDWF_HIP(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw HipError(std::string(#expr) + " failed: " + hipGetErrorString(e_) + " (" + \
                     __FILE__ + ":" + std::to_string(__LINE__) + ")");                     \
  } while (0)

#define QDWF_REQUIRE(cond, msg)         \
  do {                                  \
    if (!(cond)) throw ArgError(msg);   \
  } while (0)

// ------------------------------------------------------------ s-matrices
enum MatId { MAT_MEE, MAT_MEE_DAG, MAT_MEE_INV, MAT_MEE_INVDAG, MAT_P, MAT_P_DAG, MAT_K, MAT_K_DAG, MAT_NEGC, MAT_COUNT };

typedef std::vector<cd> Mat;  // ls*ls row-major

Mat mat_mul(const Mat& a, const Mat& b, int n)
{
  Mat c(n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) c[i * n + j] += a[i * n + k] * b[k * n + j];
  return c;
}
Mat mat_adj(const Mat& a, int n)
{
  Mat c(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) c[i * n + j] = std::conj(a[j * n + i]);
  return c;
}
Mat mat_inv(const Mat& a0, int n)
{
  // Gauss-Jordan elimination with partial pivoting.
  Mat a = a0, inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::abs(a[r * n + c]) > std::abs(a[piv * n + c])) piv = r;
    if (std::abs(a[piv * n + c]) == 0.0) throw ArgError("singular fifth-dimension matrix (check mass/m5/bs/cs)");
    if (piv != c)
      for (int j = 0; j < n; ++j) {
        std::swap(a[c * n + j], a[piv * n + j]);
        std::swap(inv[c * n + j], inv[piv * n + j]);
      }
    const cd d = 1.0 / a[c * n + c];
    for (int j = 0; j < n; ++j) {
      a[c * n + j] *= d;
      inv[c * n + j] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const cd f = a[r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        a[r * n + j] -= f * a[c * n + j];
        inv[r * n + j] -= f * inv[c * n + j];
      }
    }
  }
  return inv;
}

// Fifth-dimension shift S per chirality (dslash.cpp:627-634):
// chirality 0 (spins 0,1, P+): (S psi)_s = psi_{s-1}, psi_{-1} = -m psi_{Ls-1}
// chirality 1 (spins 2,3, P-): (S psi)_s = psi_{s+1}, psi_{Ls} = -m psi_0
Mat shift_matrix(int chi, int ls, double mass)
{
  Mat S(ls * ls, 0.0);
  for (int s = 0; s < ls; ++s) {
    if (chi == 0) {
      if (s >= 1) S[s * ls + s - 1] = 1.0;
      else S[0 * ls + ls - 1] = -mass;
    } else {
      if (s <= ls - 2) S[s * ls + s + 1] = 1.0;
      else S[(ls - 1) * ls + 0] = -mass;
    }
  }
  return S;
}
Mat diag_plus(const std::vector<cd>& d, const std::vector<cd>& e, const Mat& S, int ls)
{
  // diag(d) + diag(e) S
  Mat M(ls * ls, 0.0);
  for (int i = 0; i < ls; ++i) {
    M[i * ls + i] += d[i];
    for (int j = 0; j < ls; ++j) M[i * ls + j] += e[i] * S[i * ls + j];
  }
  return M;
}

// ---------------------------------------------------------------- state
enum Work64 {
  W_IN_E, W_IN_O, W_DM_E, W_DM_O, W_OUT_EP, W_IN_OP, W_OUT_OP, W_TMP, W_ITMP, W_OUT_E, W_OUT_O,
  W_CG_R, W_CG_P, W_CG_AP, W_CG_X, W_CG_T, W_OP_T1, W_OP_T2, W_OP_T3, W_OP_Y, W_G_E, W_G_O, W_COUNT64
};
enum Work32 { V_R, V_P, V_AP, V_X, V_T, V_T1, V_T2, V_T3, V_Y, W_COUNT32 };

// Tile width (columns per multi-RHS block) and sites per block of the
// three-phase solver.
struct TileChoice {
  int width, sb;
};

}  // namespace

struct qlat_dwf_inverter {
  int device = 0;
  hipStream_t stream = nullptr;
  LatGeo geo{};
  int ls = 0;
  double mass = 0, m5 = 0, mobius_scale = 1;
  bool dminus = true;
  std::vector<cd> bs, cs;
  double stop_rsd = 1e-8;
  long max_num_iter = 200;
  long max_cycle = 300;
  int solver = QLAT_DWF_SOLVER_FP64;
  double delta = 0.01;
  // hop launch shapes (launch_hop): LDS-staged links; FP64 with streaming stores
  int hop_spt64 = -102, hop_spt32 = -4;
  bool fused32 = true;  // FP32 hop + s-matrix fusion (QLAT_DWF_FUSED=0 disables)
  bool fused64 = true;  // FP64 hop + s-matrix fusion, k_hop_m5d (QLAT_DWF_FUSED64=0 disables)
  bool three_phase32 = true;   // FP32 solver on the three-phase kernels (QLAT_DWF_THREE_PHASE=0: fused
                               // FP32 kernels, single s-matrix pass per hop)
  int tp_sb = 32;              // sites per block of the three-phase kernels (QLAT_DWF_TP_SB)
  int step_sb = 32;            // sites per block of the three-phase CG step (QLAT_DWF_STEP_SB)
  bool x_defer = true;         // x updated every other CG step, two directions at once (QLAT_DWF_XDEFER)
  int site_major = 0;          // three-phase hop block order (QLAT_DWF_SITE_MAJOR)
  int fp16_mode = 1;           // FP16 solver storage mode 1 or 2 (ThreePhase; QLAT_DWF_FP16_MODE)
  TileChoice tuned[4] = {};    // autotuned tile per storage mode of the three-phase solver
  bool real_mat[16] = {};  // s-matrix MatId has only real entries (Moebius)

  double2* l64 = nullptr;
  float2* l32 = nullptr;
  unsigned* l16 = nullptr;  // FP16 links for the FP16-arithmetic hop (h2_hop)
  // FP16-storage solver: hop phase in FP16 arithmetic (FP16 links, packed
  // FP16 projection, v_dot2_f32_f16 SU(3) rows; QLAT_DWF_H2=0 selects FP32
  // arithmetic).  Measured at 24^3x64 Ls12, b = K H t over 4 columns:
  // 1.063 ms against 1.177 ms
  int h2_hop = 1;
  int plain64 = 1;  // plain FP16 hop on 64-site blocks (QLAT_DWF_PLAIN64)
  cplx<double>* mat64 = nullptr;  // [MAT_COUNT][2][ls][ls]
  cplx<float>* mat32 = nullptr;
  cplx<double>* beta64 = nullptr;  // [2][ls]: ones, c~
  cplx<float>* beta32 = nullptr;
  double2* w64[W_COUNT64] = {};
  float4* w32[W_COUNT32] = {};
  double2* stage[2] = {nullptr, nullptr};
  size_t stage_bytes = 0;
  double* part = nullptr;
  int64_t part_cap = 0;      // doubles in part
  double* red = nullptr;     // device reduction results
  // Fine-grained (coherent) pinned host memory that every reduction the
  // host reads (norms, dot products, a tile's CG scalars) is written to
  // directly (red_map_dev: its device address), so that reading one costs
  // a stream synchronisation and no copy
  double* red_map = nullptr;
  double* red_map_dev = nullptr;
  // graph-captured multi-RHS CG: iterations per graph launch (QLAT_DWF_GRAPH,
  // 0 = one host round trip per iteration) and the device / pinned CG state.
  // Off by default: measured at most 2.5% faster (8^3x16) and 1-5% slower
  // from 12^3x24 up, where a host round trip is a small part of an iteration.
  int graph_iters = 0;
  qdwf::CgState* cg_dev = nullptr;
  qdwf::CgState* cg_host = nullptr;

  // low modes (block-compressed eigen system), qlat_dwf_inverter_set_low_modes
  int lm_bs[4] = {0, 0, 0, 0};
  int lm_nbasis = 0, lm_nvec = 0, lm_nblock = 0, lm_bve = 0;
  float2* lm_basis = nullptr;
  float2* lm_coef = nullptr;
  double* lm_evals = nullptr;
  int* lm_hmap[2] = {nullptr, nullptr};  // odd, even
  double2* lm_part = nullptr;
  double2* lm_p = nullptr;

  // multi-RHS: FP32 stacks of `mcap` fields and per-column FP64 fields
  int rhs_tile = 0;  // columns per multi-RHS solve (QLAT_DWF_RHS_TILE; 0 = autotuned / 12)
  int mcap = 0;
  float4* m32[9] = {};
  uint4* mh[8] = {};            // FP16 (per-site scale) stacks of the three-phase solver
  std::vector<double2*> col64;  // 3 per column: b, x, out_e_p

  qlat_dwf_solve_stats stats{};

  // allocation-free calls: while noalloc is set, W() hands out only the
  // slices of the static scratch array listed in w64_noalloc and every other
  // buffer request throws instead of allocating
  bool noalloc = false;
  double2* w64_noalloc[W_COUNT64] = {};
};

namespace {

typedef qlat_dwf_inverter Inv;

size_t field64_bytes(const Inv& I) { return (size_t)I.ls * 12 * I.geo.stride * sizeof(double2); }
size_t field32_bytes(const Inv& I) { return (size_t)I.ls * 6 * I.geo.stride * sizeof(float4); }
int64_t n64(const Inv& I) { return (int64_t)I.ls * 12 * I.geo.stride; }
int64_t n32(const Inv& I) { return (int64_t)I.ls * 6 * I.geo.stride; }

// Device address of `scratch` on each device, looked up once per device
// (with that device current) when the first inverter is created on it.
std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;

void init_scratch(int device)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  }
  g_scratch_ptr[device] = (unsigned char*)p;
}

unsigned char* scratch_ptr(int device)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= (int)g_scratch_ptr.size() || !g_scratch_ptr[device])
    throw std::runtime_error("the allocation-free scratch of this device is not set up");
  return g_scratch_ptr[device];
}

void* dev_alloc(size_t bytes)
{
  if (g_noalloc_depth > 0) throw std::logic_error("internal: device allocation in an allocation-free call");
  void* p = nullptr;
  const hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw MemError("hipMalloc(" + std::to_string(bytes) + " bytes) failed: " + hipGetErrorString(e));
  }
  return p;
}

double2* W(Inv& I, int id)
{
  if (I.noalloc) {
    if (!I.w64_noalloc[id]) throw std::logic_error("internal: a work field outside the scratch in an allocation-free call");
    return I.w64_noalloc[id];
  }
  if (!I.w64[id]) {
    I.w64[id] = (double2*)dev_alloc(field64_bytes(I));
    QDWF_HIP(hipMemsetAsync(I.w64[id], 0, field64_bytes(I), I.stream));
  }
  return I.w64[id];
}
float4* V(Inv& I, int id)
{
  if (I.noalloc) throw std::logic_error("internal: an FP32 work field in an allocation-free call");
  if (!I.w32[id]) {
    I.w32[id] = (float4*)dev_alloc(field32_bytes(I));
    QDWF_HIP(hipMemsetAsync(I.w32[id], 0, field32_bytes(I), I.stream));
  }
  return I.w32[id];
}
enum Multi32 { M_R, M_P, M_AP, M_X, M_T, M_A, M_B, M_Y, M_P2 };

// FP32 stack buffer `id` with room for `ncol` fields.
float4* MV(Inv& I, int id, int ncol)
{
  if (I.noalloc) throw std::logic_error("internal: a multi-RHS stack in an allocation-free call");
  if (ncol > I.mcap) {
    QDWF_HIP(hipStreamSynchronize(I.stream));
    for (auto*& p : I.m32)
      if (p) {
        QDWF_HIP(hipFree(p));
        p = nullptr;
      }
    for (auto*& p : I.mh)
      if (p) {
        QDWF_HIP(hipFree(p));
        p = nullptr;
      }
    I.mcap = ncol;
  }
  if (!I.m32[id]) {
    I.m32[id] = (float4*)dev_alloc(field32_bytes(I) * I.mcap);
    QDWF_HIP(hipMemsetAsync(I.m32[id], 0, field32_bytes(I) * I.mcap, I.stream));
  }
  return I.m32[id];
}

// FP16 stacks (16-byte units per field: fieldh_units)
enum MultiH { H_P, H_AP, H_T, H_B, H_Y, H_R, H_X, H_P2 };
int64_t fieldh_units(const Inv& I) { return (I.geo.stride >> 6) * I.ls * 208; }
uint4* MH(Inv& I, int id, int ncol)
{
  MV(I, M_R, ncol);  // settles mcap (frees every stack when it grows)
  if (!I.mh[id]) {
    const size_t bytes = (size_t)fieldh_units(I) * sizeof(uint4) * I.mcap;
    I.mh[id] = (uint4*)dev_alloc(bytes);
    QDWF_HIP(hipMemsetAsync(I.mh[id], 0, bytes, I.stream));
  }
  return I.mh[id];
}

// Per-column FP64 field: kind 0 = odd-prec source b, 1 = solution x, 2 = out_e_p
// (the FP32 fused multi-RHS path; the three-phase driver recomputes out_e_p).
double2* C64(Inv& I, int col, int kind)
{
  const size_t k = (size_t)col * 3 + kind;
  if (I.col64.size() <= k) I.col64.resize(k + 1, nullptr);
  if (!I.col64[k]) {
    I.col64[k] = (double2*)dev_alloc(field64_bytes(I));
    QDWF_HIP(hipMemsetAsync(I.col64[k], 0, field64_bytes(I), I.stream));
  }
  return I.col64[k];
}

size_t stage_size(const Inv& I);

// Two device staging buffers for host-pointer entry points, each large
// enough for a 5D full field or a propagator, allocated on first use.
double2* stage(Inv& I, int k, size_t bytes)
{
  if (I.noalloc) throw std::logic_error("internal: a staging buffer in an allocation-free call");
  if (bytes > stage_size(I)) throw ArgError("field larger than the staging buffer");
  if (!I.stage[k]) {
    I.stage_bytes = stage_size(I);
    I.stage[k] = (double2*)dev_alloc(I.stage_bytes);
  }
  return I.stage[k];
}

void sync(Inv& I) { QDWF_HIP(hipStreamSynchronize(I.stream)); }
void check_launch() { QDWF_HIP(hipGetLastError()); }

// ---------------------------------------------------------- typed ops
template <class T>
struct Prec;
template <>
struct Prec<double> {
  typedef double2 slot;
  static const cplx<double>* mat(const Inv& I, int id) { return I.mat64 + (size_t)id * 2 * I.ls * I.ls; }
  static const cplx<double>* beta(const Inv& I, int which) { return I.beta64 + (size_t)which * I.ls; }
  static const void* links(const Inv& I) { return I.l64; }
  static int spt(const Inv& I) { return I.hop_spt64; }
  static int64_t n(const Inv& I) { return n64(I); }
};
template <>
struct Prec<float> {
  typedef float4 slot;
  static const cplx<float>* mat(const Inv& I, int id) { return I.mat32 + (size_t)id * 2 * I.ls * I.ls; }
  static const cplx<float>* beta(const Inv& I, int which) { return I.beta32 + (size_t)which * I.ls; }
  static const void* links(const Inv& I) { return I.l32; }
  static int spt(const Inv& I) { return I.hop_spt32; }
  static int64_t n(const Inv& I) { return n32(I); }
};

qdwf::RhsSet one_col()
{
  qdwf::RhsSet c;
  c.n = 1;
  c.fstride = 0;
  for (int i = 0; i < qdwf::kMaxRhs; ++i) c.rhs[i] = 0;
  return c;
}

// out = alpha * Mat in + beta * x; with dotp (and x) also block partials of
// Re <dotp, out> into I.part (m5_dot_blocks of them)
template <class T>
void m5(Inv& I, typename Prec<T>::slot* out, const typename Prec<T>::slot* in, int mat, T alpha = 1,
        const typename Prec<T>::slot* x = nullptr, T beta = 0, const typename Prec<T>::slot* dotp = nullptr,
        const qdwf::RhsSet* cols = nullptr)
{
  qdwf::M5Args<T> a;
  a.dotp = dotp;
  a.part = I.part;
  a.cols = cols ? *cols : one_col();
  a.geo = I.geo;
  a.ls = I.ls;
  a.alpha = alpha;
  a.beta = beta;
  a.mat = Prec<T>::mat(I, mat);
  a.real_mat = I.real_mat[mat];
  a.in = in;
  a.out = out;
  a.x = x;
  qdwf::launch_m5<T>(a, I.stream);
  check_launch();
}

// out(parity eo_out) = alpha * H in + beta_s * x   (beta: 0 ones, 1 c~)
template <class T>
void hop(Inv& I, typename Prec<T>::slot* out, const typename Prec<T>::slot* in, int eo_out, bool dag,
         T alpha = 1, const typename Prec<T>::slot* x = nullptr, int beta = 0, const qdwf::RhsSet* cols = nullptr)
{
  qdwf::HopArgs<T> a;
  a.cols = cols ? *cols : one_col();
  a.geo = I.geo;
  a.ls = I.ls;
  a.eo_out = eo_out;
  a.nchunk = 0;
  a.alpha = alpha;
  a.in = in;
  a.out = out;
  a.x = x;
  a.beta = Prec<T>::beta(I, beta);
  a.links = Prec<T>::links(I);
  qdwf::launch_hop<T>(a, dag, Prec<T>::spt(I), I.stream);
  check_launch();
}

// Temporaries of the typed composite operators.
template <class T>
struct Tmp;
template <>
struct Tmp<double> {
  static double2* t1(Inv& I) { return W(I, W_OP_T1); }
  static double2* t2(Inv& I) { return W(I, W_OP_T2); }
  static double2* t3(Inv& I) { return W(I, W_OP_T3); }
  static double2* y(Inv& I) { return W(I, W_OP_Y); }
};
template <>
struct Tmp<float> {
  static float4* t1(Inv& I) { return V(I, V_T1); }
  static float4* t2(Inv& I) { return V(I, V_T2); }
  static float4* t3(Inv& I) { return V(I, V_T3); }
  static float4* y(Inv& I) { return V(I, V_Y); }
};

// multiply_m_e_o: out(3-eo_in) = H P in
template <class T>
void op_m_e_o(Inv& I, typename Prec<T>::slot* out, const typename Prec<T>::slot* in, int eo_in)
{
  auto* t = Tmp<T>::t1(I);
  m5<T>(I, t, in, MAT_P);
  hop<T>(I, out, t, 3 - eo_in, false);
}

// multiply_mdag_e_o: out(3-eo_in) = P^+ H^+ in
template <class T>
void op_mdag_e_o(Inv& I, typename Prec<T>::slot* out, const typename Prec<T>::slot* in, int eo_in)
{
  auto* t = Tmp<T>::t1(I);
  hop<T>(I, t, in, 3 - eo_in, true);
  m5<T>(I, out, t, MAT_P_DAG);
}

// FP32 fused out = alpha * Mat (H in) + beta * x (false if not available)
bool hop_m5_32(Inv& I, float4* out, const float4* in, int eo_out, bool dag, int mat, float alpha = 1.f,
               const float4* x = nullptr, float beta = 0.f, const float4* dotp = nullptr,
               const qdwf::RhsSet* cols = nullptr)
{
  if (!I.fused32) return false;
  if (dotp && !(dag && x)) return false;
  qdwf::HopM5Args a{};
  a.dotp = dotp;
  a.part = I.part;
  a.cols = cols ? *cols : one_col();
  a.geo = I.geo;
  a.ls = I.ls;
  a.eo_out = eo_out;
  a.alpha = alpha;
  a.beta = beta;
  a.in = in;
  a.out = out;
  a.x = x;
  a.links = I.l32;
  a.mat = Prec<float>::mat(I, mat);
  a.real_mat = I.real_mat[mat];
  if (!qdwf::launch_hop_m5(a, dag, I.stream)) return false;
  check_launch();
  return true;
}

// FP64 fused out = alpha * Mat (H in) + beta * x (false if not available)
bool hop_m5_64(Inv& I, double2* out, const double2* in, int eo_out, bool dag, int mat, double alpha = 1.0,
               const double2* x = nullptr, double beta = 0.0, const double2* dotp = nullptr)
{
  if (!I.fused64) return false;
  qdwf::HopM5dArgs a{};
  a.geo = I.geo;
  a.ls = I.ls;
  a.eo_out = eo_out;
  a.alpha = alpha;
  a.beta = beta;
  a.in = in;
  a.out = out;
  a.x = x;
  a.links = I.l64;
  a.mat = Prec<double>::mat(I, mat);
  a.real_mat = I.real_mat[mat];
  a.dotp = dotp;
  a.part = I.part;
  if (!qdwf::launch_hop_m5d(a, dag, I.stream)) return false;
  check_launch();
  return true;
}

// Mpc = 1 - H K H K
template <class T>
void op_mpc(Inv& I, typename Prec<T>::slot* out, const typename Prec<T>::slot* in, int eo)
{
  auto* t1 = Tmp<T>::t1(I);
  auto* t2 = Tmp<T>::t2(I);
  m5<T>(I, t1, in, MAT_K);
  if constexpr (std::is_same<T, float>::value) {
    if (hop_m5_32(I, t2, t1, 3 - eo, false, MAT_K)) {
      hop<T>(I, out, t2, eo, false, (T)-1, in, 0);
      return;
    }
  } else {
    if (hop_m5_64(I, t2, t1, 3 - eo, false, MAT_K)) {
      hop<T>(I, out, t2, eo, false, (T)-1, in, 0);
      return;
    }
  }
  hop<T>(I, t2, t1, 3 - eo, false);
  m5<T>(I, t1, t2, MAT_K);
  hop<T>(I, out, t1, eo, false, (T)-1, in, 0);
}

// Mpc^+ = 1 - K^+ H^+ K^+ H^+
template <class T>
void op_mpcdag(Inv& I, typename Prec<T>::slot* out, const typename Prec<T>::slot* in, int eo)
{
  auto* t1 = Tmp<T>::t1(I);
  auto* t2 = Tmp<T>::t2(I);
  if constexpr (std::is_same<T, float>::value) {
    if (hop_m5_32(I, t2, in, 3 - eo, true, MAT_K_DAG)) {
      hop_m5_32(I, out, t2, eo, true, MAT_K_DAG, -1.f, in, 1.f);
      return;
    }
  } else {
    if (hop_m5_64(I, t2, in, 3 - eo, true, MAT_K_DAG)) {
      hop_m5_64(I, out, t2, eo, true, MAT_K_DAG, -1.0, in, 1.0);
      return;
    }
  }
  hop<T>(I, t1, in, 3 - eo, true);
  m5<T>(I, t2, t1, MAT_K_DAG);
  hop<T>(I, t1, t2, eo, true);
  m5<T>(I, out, t1, MAT_K_DAG, (T)-1, in, (T)1);
}

// Mpc^+ Mpc
template <class T>
void op_hermop(Inv& I, typename Prec<T>::slot* out, const typename Prec<T>::slot* in, int eo)
{
  auto* y = Tmp<T>::y(I);
  op_mpc<T>(I, y, in, eo);
  op_mpcdag<T>(I, out, y, eo);
}

// CG form of Mpc^+ Mpc: ap = A p given t1 = K p (computed by the previous
// CG step), with Re <p, ap> reduced in the same pass, for every column of
// `cols` (stacks of fields).  dots[i] receives column i's value.
template <class T>
void hermop_t1_dot(Inv& I, typename Prec<T>::slot* ap, const typename Prec<T>::slot* p,
                   const typename Prec<T>::slot* t1, int eo, const qdwf::RhsSet& cols,
                   typename Prec<T>::slot* A, typename Prec<T>::slot* B, typename Prec<T>::slot* Y, double* dots)
{
  int64_t nparts = 0;
  bool done = false;
  if constexpr (std::is_same<T, float>::value) {
    if (hop_m5_32(I, B, t1, 3 - eo, false, MAT_K, 1.f, nullptr, 0.f, nullptr, &cols)) {
      hop<T>(I, Y, B, eo, false, (T)-1, p, 0, &cols);                         // y = p - H K H K p
      hop_m5_32(I, B, Y, 3 - eo, true, MAT_K_DAG, 1.f, nullptr, 0.f, nullptr, &cols);  // u = K^+ H^+ y
      hop_m5_32(I, ap, B, eo, true, MAT_K_DAG, -1.f, Y, 1.f, p, &cols);      // ap = y - K^+ H^+ u
      nparts = qdwf::hop_m5_blocks(I.geo);
      done = true;
    }
  } else {
    if (cols.n == 1 && cols.rhs[0] == 0 && I.fused64 && qdwf::hop_m5d_ok(I.ls)) {
      hop_m5_64(I, B, t1, 3 - eo, false, MAT_K);                      // b = K H K p
      hop<T>(I, Y, B, eo, false, (T)-1, p, 0, &cols);                  // y = p - H b
      hop_m5_64(I, B, Y, 3 - eo, true, MAT_K_DAG);                     // u = K^+ H^+ y
      hop_m5_64(I, ap, B, eo, true, MAT_K_DAG, -1.0, Y, 1.0, p);       // ap = y - K^+ H^+ u
      nparts = qdwf::hop_m5d_blocks(I.geo);
      done = true;
    }
  }
  if (!done) {
    hop<T>(I, A, t1, 3 - eo, false, (T)1, nullptr, 0, &cols);
    m5<T>(I, B, A, MAT_K, (T)1, nullptr, (T)0, nullptr, &cols);
    hop<T>(I, Y, B, eo, false, (T)-1, p, 0, &cols);
    hop<T>(I, A, Y, 3 - eo, true, (T)1, nullptr, 0, &cols);
    m5<T>(I, B, A, MAT_K_DAG, (T)1, nullptr, (T)0, nullptr, &cols);
    hop<T>(I, A, B, eo, true, (T)1, nullptr, 0, &cols);
    m5<T>(I, ap, A, MAT_K_DAG, (T)-1, Y, (T)1, p, &cols);
    nparts = (I.geo.vh + qdwf::kM5Block - 1) / qdwf::kM5Block * qdwf::SpinorIO<T>::kSlots;
  }
  qdwf::launch_reduce_rows(I.part, (int)nparts, cols.n, I.red_map_dev, I.stream);
  check_launch();
  QDWF_HIP(hipStreamSynchronize(I.stream));
  for (int i = 0; i < cols.n; ++i) dots[i] = I.red_map[i];
}

// single-field form
template <class T>
double hermop_t1_dot(Inv& I, typename Prec<T>::slot* ap, const typename Prec<T>::slot* p,
                     const typename Prec<T>::slot* t1, int eo)
{
  double d = 0;
  const qdwf::RhsSet c = one_col();
  hermop_t1_dot<T>(I, ap, p, t1, eo, c, Tmp<T>::t2(I), Tmp<T>::t3(I), Tmp<T>::y(I), &d);
  return d;
}

// x_i += alpha_i p_i (x may be null); p_i = r_i + beta_i p_i; t_i = K p_i
template <class T>
void cg_step(Inv& I, typename Prec<T>::slot* x, typename Prec<T>::slot* p, const typename Prec<T>::slot* r,
             typename Prec<T>::slot* t, const double* alpha, const double* beta, const qdwf::RhsSet& cols)
{
  qdwf::CgStepArgs<T> a{};
  a.geo = I.geo;
  a.ls = I.ls;
  for (int i = 0; i < qdwf::kMaxRhs; ++i) {
    a.alpha[i] = i < cols.n ? alpha[i] : 0.0;
    a.beta[i] = i < cols.n ? beta[i] : 0.0;
  }
  a.cols = cols;
  a.mat = Prec<T>::mat(I, MAT_K);
  a.real_mat = I.real_mat[MAT_K];
  a.x = x;
  a.p = p;
  a.r = const_cast<typename Prec<T>::slot*>(r);  // read only in this form
  a.t = t;
  qdwf::launch_cg_step<T>(a, I.stream);
  check_launch();
}

template <class T>
void cg_step(Inv& I, typename Prec<T>::slot* x, typename Prec<T>::slot* p, const typename Prec<T>::slot* r,
             typename Prec<T>::slot* t, double alpha, double beta)
{
  cg_step<T>(I, x, p, r, t, &alpha, &beta, one_col());
}

// ------------------------------------------------------------ BLAS-1
double dev_norm2(Inv& I, const double2* a)
{
  qdwf::launch_norm2<double2>(a, n64(I), I.part, I.red_map_dev, I.stream);
  check_launch();
  sync(I);
  return I.red_map[0];
}
template <class S>
cd dev_dot(Inv& I, const S* a, const S* b, int64_t n)
{
  qdwf::launch_dot<S>(a, b, n, I.part, I.red_map_dev, I.stream);
  check_launch();
  sync(I);
  return cd(I.red_map[0], I.red_map[1]);
}
template <class S>
double dev_cg_update(Inv& I, double alpha, const S* p, const S* ap, S* x, S* r, int64_t n)
{
  qdwf::launch_cg_update<S>(alpha, p, ap, x, r, n, I.part, I.red_map_dev, I.stream);
  check_launch();
  sync(I);
  return I.red_map[0];
}
void copy64(Inv& I, double2* dst, const double2* src)
{
  if (dst != src) QDWF_HIP(hipMemcpyAsync(dst, src, field64_bytes(I), hipMemcpyDeviceToDevice, I.stream));
}
void zero64(Inv& I, double2* dst) { QDWF_HIP(hipMemsetAsync(dst, 0, field64_bytes(I), I.stream)); }
void zero32(Inv& I, float4* dst) { QDWF_HIP(hipMemsetAsync(dst, 0, field32_bytes(I), I.stream)); }

// --------------------------------------------------- full-lattice ops
// multiply_d_minus (dslash.cpp:448-480): D_- psi = psi + D_W(-M5)(-c psi)
//   = diag(c~) psi + H (-c psi), c~ = 1 - c (4 - M5)
void op_d_minus(Inv& I, double2* out_e, double2* out_o, const double2* in_e, const double2* in_o)
{
  double2* t1 = W(I, W_OP_T1);
  double2* t2 = W(I, W_OP_T2);
  m5<double>(I, t1, in_o, MAT_NEGC);
  m5<double>(I, t2, in_e, MAT_NEGC);
  hop<double>(I, out_e, t1, 2, false, 1.0, in_e, 1);
  hop<double>(I, out_o, t2, 1, false, 1.0, in_o, 1);
}

// multiply_m (dslash.cpp:1062-1080): out_p = M_ee in_p + H P in_{3-p}
void op_m_full(Inv& I, double2* out_e, double2* out_o, const double2* in_e, const double2* in_o)
{
  double2* t1 = W(I, W_OP_T1);
  double2* t2 = W(I, W_OP_T2);
  m5<double>(I, t1, in_o, MAT_P);
  m5<double>(I, t2, in_e, MAT_MEE);
  hop<double>(I, out_e, t1, 2, false, 1.0, t2, 0);
  m5<double>(I, t1, in_e, MAT_P);
  m5<double>(I, t2, in_o, MAT_MEE);
  hop<double>(I, out_o, t1, 1, false, 1.0, t2, 0);
}

// multiply_mdag (dslash.cpp:1082-1100): out_p = M_ee^+ in_p + P^+ H^+ in_{3-p}
void op_mdag_full(Inv& I, double2* out_e, double2* out_o, const double2* in_e, const double2* in_o)
{
  double2* t1 = W(I, W_OP_T1);
  double2* t2 = W(I, W_OP_T2);
  hop<double>(I, t1, in_o, 2, true);
  m5<double>(I, t2, in_e, MAT_MEE_DAG);
  m5<double>(I, out_e, t1, MAT_P_DAG, 1.0, t2, 1.0);
  hop<double>(I, t1, in_e, 1, true);
  m5<double>(I, t2, in_o, MAT_MEE_DAG);
  m5<double>(I, out_o, t1, MAT_P_DAG, 1.0, t2, 1.0);
}

// ------------------------------------------------------------- solvers
// cg_with_f (dslash.h:331-399) with f = multiply_hermop_sym2, FP64.
// x: initial guess (guess_zero skips the A x evaluation, which is exact).
long cg_fp64(Inv& I, double2* x, const double2* b, double stop_rsd, long max_num_iter, bool guess_zero)
{
  if (max_num_iter == 0) return 0;
  double2* r = W(I, W_CG_R);
  double2* p = W(I, W_CG_P);
  double2* ap = W(I, W_CG_AP);
  double2* t = W(I, W_CG_T);
  const int64_t n = n64(I);
  if (guess_zero) {
    copy64(I, r, b);
  } else {
    op_hermop<double>(I, ap, x, 1);
    I.stats.hermop_fp64 += 1;
    copy64(I, r, b);
    qdwf::launch_axpy<double2>(-1.0, ap, r, n, I.stream);
    check_launch();
  }
  copy64(I, p, r);
  m5<double>(I, t, p, MAT_K);
  const double qnorm_in = dev_norm2(I, b);
  double qnorm_r = dev_norm2(I, r);
  for (long iter = 1; iter <= max_num_iter; ++iter) {
    // alpha = |r|^2 / Re <p, A p> ; r -= alpha A p (dslash.h:365-374)
    const double pap = hermop_t1_dot<double>(I, ap, p, t, 1);
    I.stats.hermop_fp64 += 1;
    const double alpha = qnorm_r / pap;
    qdwf::launch_axpy_norm<double2>(alpha, ap, r, n, I.part, I.red_map_dev, I.stream);
    check_launch();
    sync(I);
    const double new_qnorm_r = I.red_map[0];
    if (!std::isfinite(new_qnorm_r)) throw std::runtime_error("CG breakdown: non-finite residual (FP64)");
    if (new_qnorm_r <= qnorm_in * stop_rsd * stop_rsd) {
      qdwf::launch_axpy<double2>(alpha, p, x, n, I.stream);
      check_launch();
      return iter;
    }
    // x += alpha p ; p = r + beta p ; t = K p (dslash.h:369-392)
    const double beta = new_qnorm_r / qnorm_r;
    cg_step<double>(I, x, p, r, t, alpha, beta);
    qnorm_r = new_qnorm_r;
  }
  return max_num_iter + 1;
}

// Mixed-precision CG on A x = b (odd half fields) from the x64 it is given:
// the Krylov recurrence runs in FP32; whenever the iterated residual has
// dropped by `delta` since the last replacement (and at convergence) the
// solution increment is folded into the FP64 x and the residual is
// recomputed as b - A x in FP64.
// Returns iterations; stops when |b - A x| <= stop_rsd |b| in FP64.
long cg_mixed(Inv& I, double2* x64, const double2* b64, double stop_rsd, long max_total_iter, double* true_rel)
{
  double2* r64 = W(I, W_CG_R);
  double2* ax64 = W(I, W_CG_AP);
  float4* r = V(I, V_R);
  float4* p = V(I, V_P);
  float4* ap = V(I, V_AP);
  float4* x = V(I, V_X);
  float4* t = V(I, V_T);
  const int64_t n = n32(I), nd = n64(I);
  const double b2 = dev_norm2(I, b64);
  const double target = stop_rsd * stop_rsd * b2;
  auto true_residual = [&]() {
    op_hermop<double>(I, ax64, x64, 1);
    I.stats.hermop_fp64 += 1;
    copy64(I, r64, b64);
    qdwf::launch_axpy<double2>(-1.0, ax64, r64, nd, I.stream);
    check_launch();
    return dev_norm2(I, r64);
  };
  double rr = true_residual();
  if (rr <= target) {
    *true_rel = std::sqrt(rr / b2);
    return 0;
  }
  qdwf::launch_d2f(r64, r, I.ls, I.geo.stride, I.stream);
  check_launch();
  QDWF_HIP(hipMemcpyAsync(p, r, field32_bytes(I), hipMemcpyDeviceToDevice, I.stream));
  zero32(I, x);
  m5<float>(I, t, p, MAT_K);
  double maxrr = rr;
  const double d2 = I.delta * I.delta;
  long iter = 0;
  while (iter < max_total_iter) {
    const double pap = hermop_t1_dot<float>(I, ap, p, t, 1);
    I.stats.hermop_low += 1;
    const double alpha = rr / pap;
    qdwf::launch_axpy_norm<float4>(alpha, ap, r, n, I.part, I.red_map_dev, I.stream);
    check_launch();
    sync(I);
    double rr_new = I.red_map[0];
    ++iter;
    if (!std::isfinite(rr_new)) throw std::runtime_error("CG breakdown: non-finite residual (FP32)");
    if (rr_new <= d2 * maxrr || rr_new <= target) {
      // reliable update: x64 += x32 + alpha p, r64 = b - A x64 in FP64
      qdwf::launch_axpy<float4>(alpha, p, x, n, I.stream);
      qdwf::launch_f2d(x, x64, I.ls, I.geo.stride, 1, I.stream);
      check_launch();
      zero32(I, x);
      const double rr_true = true_residual();
      I.stats.reliable_updates += 1;
      if (rr_true <= target) {
        *true_rel = std::sqrt(rr_true / b2);
        return iter;
      }
      qdwf::launch_d2f(r64, r, I.ls, I.geo.stride, I.stream);
      check_launch();
      rr_new = rr_true;
      maxrr = rr_true;
      cg_step<float>(I, nullptr, p, r, t, 0.0, rr_new / rr);
    } else {
      cg_step<float>(I, x, p, r, t, alpha, rr_new / rr);
    }
    rr = rr_new;
  }
  // not converged: fold in what we have
  qdwf::launch_f2d(x, x64, I.ls, I.geo.stride, 1, I.stream);
  check_launch();
  *true_rel = std::sqrt(true_residual() / b2);
  return max_total_iter + 1;
}

// Multi-RHS mixed-precision CG: A x_c = b_c for the columns c < ncol
// (ncol <= kMaxRhs) held in C64(c, 0) (b) and C64(c, 1) (x, zero on entry).
// All columns advance together through FP32 kernels that stage each link
// block once for every column; each column keeps its own alpha, beta,
// FP64 reliable updates (residual recomputed as b - A x in FP64 when its
// iterated residual dropped by delta, and at convergence) and retires when
// its FP64 true residual reaches stop_rsd.  Returns the summed iterations.
void deflate_dev(Inv& I, double2* out, const double2* in, int eo);

long cg_mixed_multi(Inv& I, int ncol, double stop_rsd, long cap)
{
  const int64_t fs = n32(I), nd = n64(I);
  float4* R = MV(I, M_R, ncol);
  float4* P = MV(I, M_P, ncol);
  float4* AP = MV(I, M_AP, ncol);
  float4* X = MV(I, M_X, ncol);
  float4* T = MV(I, M_T, ncol);
  float4* B = MV(I, M_B, ncol);
  float4* Y = MV(I, M_Y, ncol);
  float4* A = I.fused32 ? B : MV(I, M_A, ncol);  // only the unfused path needs A
  double2* r64 = W(I, W_CG_R);
  double2* ax64 = W(I, W_CG_AP);
  std::vector<double> b2(ncol), target(ncol), rr(ncol), maxrr(ncol);
  std::vector<long> iters(ncol, 0);
  std::vector<int> active;
  auto true_residual = [&](int c) {
    op_hermop<double>(I, ax64, C64(I, c, 1), 1);
    I.stats.hermop_fp64 += 1;
    copy64(I, r64, C64(I, c, 0));
    qdwf::launch_axpy<double2>(-1.0, ax64, r64, nd, I.stream);
    check_launch();
    return dev_norm2(I, r64);
  };
  for (int c = 0; c < ncol; ++c) {
    b2[c] = dev_norm2(I, C64(I, c, 0));
    target[c] = stop_rsd * stop_rsd * b2[c];
    zero64(I, C64(I, c, 1));
    if (I.lm_nvec > 0 && b2[c] > 0.0) {
      // deflated initial guess (dslash.h:496-500) in the FP64 x, r = b - A x
      deflate_dev(I, C64(I, c, 1), C64(I, c, 0), 1);
      rr[c] = maxrr[c] = true_residual(c);
      qdwf::launch_d2f(r64, R + c * fs, I.ls, I.geo.stride, I.stream);
      check_launch();
      if (rr[c] <= target[c]) {
        const double rel = std::sqrt(rr[c] / b2[c]);
        if (rel > I.stats.true_residual) I.stats.true_residual = rel;
      }
    } else {
      // x = 0: r = b exactly
      qdwf::launch_d2f(C64(I, c, 0), R + c * fs, I.ls, I.geo.stride, I.stream);
      check_launch();
      rr[c] = maxrr[c] = b2[c];
    }
    if (rr[c] > target[c]) active.push_back(c);
  }
  QDWF_HIP(hipMemcpyAsync(P, R, field32_bytes(I) * ncol, hipMemcpyDeviceToDevice, I.stream));
  QDWF_HIP(hipMemsetAsync(X, 0, field32_bytes(I) * ncol, I.stream));
  auto colset = [&](const std::vector<int>& act) {
    qdwf::RhsSet cs;
    cs.n = (int)act.size();
    cs.fstride = fs;
    for (int i = 0; i < qdwf::kMaxRhs; ++i) cs.rhs[i] = i < cs.n ? act[i] : 0;
    return cs;
  };
  {
    const qdwf::RhsSet all = colset(active);
    if (all.n) m5<float>(I, T, P, MAT_K, 1.f, nullptr, 0.f, nullptr, &all);
  }
  const double d2 = I.delta * I.delta;
  double pap[qdwf::kMaxRhs], rr_new[qdwf::kMaxRhs], alpha[qdwf::kMaxRhs], step_alpha[qdwf::kMaxRhs];
  long it = 0;
  while (!active.empty() && it < cap) {
    const qdwf::RhsSet cs = colset(active);
    hermop_t1_dot<float>(I, AP, P, T, 1, cs, A, B, Y, pap);
    I.stats.hermop_low += cs.n;
    qdwf::AxpyNormArgs<float4> an;
    an.ap = AP;
    an.r = R;
    an.n = fs;
    an.cols = cs;
    an.part = I.part;
    for (int i = 0; i < qdwf::kMaxRhs; ++i) {
      alpha[i] = i < cs.n ? rr[active[i]] / pap[i] : 0.0;
      an.alpha[i] = alpha[i];
    }
    qdwf::launch_axpy_norm_multi<float4>(an, I.red_map_dev, I.stream);
    check_launch();
    sync(I);
    for (int i = 0; i < cs.n; ++i) rr_new[i] = I.red_map[i];
    ++it;
    std::vector<int> keep;
    std::vector<double> ka, kb;
    for (int i = 0; i < cs.n; ++i) {
      const int c = active[i];
      iters[c] += 1;
      if (!std::isfinite(rr_new[i])) throw std::runtime_error("CG breakdown: non-finite residual (FP32, multi)");
      step_alpha[i] = alpha[i];
      if (rr_new[i] <= d2 * maxrr[c] || rr_new[i] <= target[c]) {
        float4* xc = X + c * fs;
        qdwf::launch_axpy<float4>(alpha[i], P + c * fs, xc, fs, I.stream);
        qdwf::launch_f2d(xc, C64(I, c, 1), I.ls, I.geo.stride, 1, I.stream);
        check_launch();
        zero32(I, xc);
        const double rr_true = true_residual(c);
        I.stats.reliable_updates += 1;
        if (rr_true <= target[c]) {
          const double rel = std::sqrt(rr_true / b2[c]);
          if (rel > I.stats.true_residual) I.stats.true_residual = rel;
          continue;  // column retired
        }
        qdwf::launch_d2f(r64, R + c * fs, I.ls, I.geo.stride, I.stream);
        check_launch();
        rr_new[i] = rr_true;
        maxrr[c] = rr_true;
        step_alpha[i] = 0.0;
      }
      keep.push_back(c);
      ka.push_back(step_alpha[i]);
      kb.push_back(rr_new[i] / rr[c]);
      rr[c] = rr_new[i];
    }
    if (!keep.empty()) {
      const qdwf::RhsSet ks = colset(keep);
      cg_step<float>(I, X, P, R, T, ka.data(), kb.data(), ks);
    }
    active.swap(keep);
  }
  long total = 0;
  for (int c = 0; c < ncol; ++c) total += iters[c];
  for (int c : active) {  // cap reached: fold in what we have
    qdwf::launch_f2d(X + c * fs, C64(I, c, 1), I.ls, I.geo.stride, 1, I.stream);
    check_launch();
    const double rel = std::sqrt(true_residual(c) / b2[c]);
    if (rel > I.stats.true_residual) I.stats.true_residual = rel;
    total += 1;  // report non-convergence like cg_with_f (max_num_iter + 1)
  }
  return total;
}

// Krylov stacks and passes of the three-phase solver.  storage 0: all
// FP32; 1: p, A p, t = K p and the operator temporaries b, y in FP16 with a
// per-5D-site scale, r and the increment x in FP32; 2: r and x in FP16 too
// (measured to double the iterations at 24^3x64: an option, not the default).
qdwf::DeflArgs defl_args(Inv& I, int eo)
{
  qdwf::DeflArgs a{};
  a.geo = I.geo;
  a.ls = I.ls;
  a.eo = eo;
  for (int mu = 0; mu < 4; ++mu) a.bs[mu] = I.lm_bs[mu];
  a.nblock = I.lm_nblock;
  a.block_vol_eo = I.lm_bve;
  a.bsize = (int64_t)I.lm_bve * I.ls * 12;
  a.n_basis = I.lm_nbasis;
  a.n_vec = I.lm_nvec;
  a.basis = I.lm_basis;
  a.coef = I.lm_coef;
  a.evals = I.lm_evals;
  a.hmap = I.lm_hmap[eo == 1 ? 0 : 1];
  a.part = I.lm_part;
  a.p = I.lm_p;
  return a;
}

// deflate (dslash.cpp:144-300) of the device half field `in` of parity eo:
// out = in + sum_i v_i <v_i, in> / lambda_i with FP32 rounding as in the
// reference, or 0 without low modes.  out may equal in.
void deflate_dev(Inv& I, double2* out, const double2* in, int eo)
{
  if (I.lm_nvec == 0) {
    zero64(I, out);
    return;
  }
  qdwf::DeflArgs a = defl_args(I, eo);
  a.in = in;
  a.out = out;
  qdwf::launch_deflate(a, I.stream);
  check_launch();
}

struct ThreePhase {
  Inv& I;
  int storage;
  int sb;           // output sites per block of the three-phase hop kernels
  int64_t fh, fr, fx;  // column strides of the p/ap/t/b/y, the r and the x stacks
  void *R, *X;
  void *P, *AP, *T, *B, *Y;
  Inv* colinv[qdwf::kMaxRhs];  // the inverter (mass) of each stack column
  qdwf::CgState* cg = nullptr;  // set while a graph iteration is captured
  // p is double-buffered: P = Pbuf[cur] is the current direction, and a
  // flipping step writes the next one into the other buffer, whose old
  // contents (the previous direction) it can still add to x first
  void* Pbuf[2] = {nullptr, nullptr};
  int cur = 0;
  // where hermop_dot and axpy_norm put their reductions: the device buffer
  // I.red (read by k_cg_scalars in graph replays), or the mapped host
  // buffer of the eager loop
  double* red_out = nullptr;
  ThreePhase(Inv& inv, int ncol, int storage_, int sb_ = 0)
      : I(inv), storage(storage_), sb(sb_ ? sb_ : inv.tp_sb), red_out(inv.red)
  {
    if (!qdwf::three_phase_ok(I.ls, sb))
      for (int alt : {16, 32, 8})
        if (qdwf::three_phase_ok(I.ls, alt)) {
          sb = alt;
          break;
        }
    if (!qdwf::three_phase_ok(I.ls, sb))
      throw ArgError("the three-phase kernels need an even ls in [6, 16], or ls 20 or 24");
    if (storage && fieldh_units(I) * 16 >= (1ll << 32))
      throw ArgError("FP16 spinor storage needs V/2 * Ls * 52 bytes below 4 GB per field");
    for (auto*& p : colinv) p = &I;
    fh = half() ? fieldh_units(I) : n32(I);
    fr = rxh() ? fieldh_units(I) : n32(I);
    fx = xh() ? fieldh_units(I) : n32(I);
    R = rxh() ? (void*)MH(I, H_R, ncol) : (void*)MV(I, M_R, ncol);
    X = xh() ? (void*)MH(I, H_X, ncol) : (void*)MV(I, M_X, ncol);
    if (half()) {
      P = MH(I, H_P, ncol);
      Pbuf[1] = MH(I, H_P2, ncol);
      AP = MH(I, H_AP, ncol);
      T = MH(I, H_T, ncol);
      B = MH(I, H_B, ncol);
      Y = MH(I, H_Y, ncol);
    } else {
      P = MV(I, M_P, ncol);
      Pbuf[1] = MV(I, M_P2, ncol);
      AP = MV(I, M_AP, ncol);
      T = MV(I, M_T, ncol);
      B = MV(I, M_B, ncol);
      Y = MV(I, M_Y, ncol);
    }
    Pbuf[0] = P;
  }
  bool half() const { return storage >= 1; }
  bool rxh() const { return storage == 2; }
  bool xh() const { return storage == 2; }
  size_t slot_bytes() const { return half() ? sizeof(uint4) : sizeof(float4); }
  size_t rslot_bytes() const { return rxh() ? sizeof(uint4) : sizeof(float4); }
  size_t xslot_bytes() const { return xh() ? sizeof(uint4) : sizeof(float4); }
  void* at(void* base, int c) const { return (void*)((char*)base + (size_t)c * fh * slot_bytes()); }
  void* rat(void* base, int c) const { return (void*)((char*)base + (size_t)c * fr * rslot_bytes()); }
  void* xat(int c) const { return (void*)((char*)X + (size_t)c * fx * xslot_bytes()); }
  int64_t n5() const { return (int64_t)I.geo.stride * I.ls; }
  // column c of r = scale * b64
  void set_r(const double2* b64, int c, double scale)
  {
    if (rxh()) qdwf::launch_d2h(b64, (uint4*)rat(R, c), I.ls, n5(), scale, I.stream);
    else qdwf::launch_d2f(b64, (float4*)rat(R, c), I.ls, I.geo.stride, I.stream, scale);
    check_launch();
  }
  // x64 += scale * column c of x
  void add_x(double2* x64, int c, double scale)
  {
    if (xh()) qdwf::launch_h2d((const uint4*)xat(c), x64, I.ls, n5(), 1, scale, I.stream);
    else qdwf::launch_f2d((const float4*)xat(c), x64, I.ls, I.geo.stride, 1, I.stream, scale);
    check_launch();
  }
  void zero_x(int c0, int n) { QDWF_HIP(hipMemsetAsync(xat(c0), 0, (size_t)fx * xslot_bytes() * n, I.stream)); }
  void zero_p(int n)
  {
    for (void* b : Pbuf) QDWF_HIP(hipMemsetAsync(b, 0, (size_t)fh * slot_bytes() * n, I.stream));
  }
  void* other() const { return Pbuf[cur ^ 1]; }
  void flip()
  {
    cur ^= 1;
    P = Pbuf[cur];
  }
  qdwf::RhsSet colset(const std::vector<int>& act) const
  {
    qdwf::RhsSet cs;
    cs.n = (int)act.size();
    cs.fstride = fh;
    for (int i = 0; i < qdwf::kMaxRhs; ++i) cs.rhs[i] = i < cs.n ? act[i] : 0;
    return cs;
  }
  // per-column s-matrices `mat` of the columns of cs; false if any is complex
  bool mats(const qdwf::RhsSet& cs, int mat, const cplx<float>* out[qdwf::kMaxRhs]) const
  {
    bool real = true;
    for (int i = 0; i < qdwf::kMaxRhs; ++i) {
      const Inv& J = *colinv[i < cs.n ? cs.rhs[i] : cs.rhs[0]];
      out[i] = Prec<float>::mat(J, mat);
      if (i < cs.n) real = real && J.real_mat[mat];
    }
    return real;
  }
  // The merged CG form (r in FP32): the hermop pass also reduces Re <r, Ap>
  // and |Ap|^2, so |r - alpha Ap|^2 is known before r is updated, and the
  // update of r moves into the CG step.
  bool merged() const { return !rxh(); }
  // [r -= alpha_r ap (alpha_r not null)]; x += alpha p (updx); p = r + beta p; t = K p.
  // flip: the new p goes into the other buffer, and x also gets alpha_prev
  // times that buffer's previous direction.
  void step(const std::vector<int>& cols, const double* alpha, const double* beta, bool updx,
            const double* alpha_r = nullptr, const double* alpha_prev = nullptr, bool flip = false)
  {
    qdwf::CgStepArgs<float> a{};
    a.geo = I.geo;
    a.ls = I.ls;
    for (int i = 0; i < qdwf::kMaxRhs; ++i) {
      a.alpha[i] = i < (int)cols.size() ? alpha[i] : 0.0;
      a.beta[i] = i < (int)cols.size() ? beta[i] : 0.0;
      a.alpha_r[i] = (alpha_r && i < (int)cols.size()) ? alpha_r[i] : 0.0;
      a.alpha_prev[i] = (flip && alpha_prev && i < (int)cols.size()) ? alpha_prev[i] : 0.0;
    }
    a.p_out = flip ? other() : nullptr;
    a.ap = alpha_r ? AP : nullptr;
    a.cols = colset(cols);
    a.xr_fstride = fr;
    a.x_fstride = fx;
    a.sb = I.step_sb;
    a.mat = nullptr;
    a.real_mat = mats(a.cols, MAT_K, a.mats);
    a.x = updx ? X : nullptr;
    a.p = P;
    a.r = R;
    a.t = T;
    if (!qdwf::launch_cg_step_s(a, storage, I.stream)) throw ArgError("the three-phase CG step needs an even ls in [6, 16]");
    check_launch();
    if (flip) this->flip();
  }
  // p = r + beta q into the current p buffer, q the other buffer's p, and
  // t = K p, for columns whose last direction is in the other buffer (x and
  // r unchanged)
  void restart_from_other(const std::vector<int>& cols, const double* beta)
  {
    qdwf::CgStepArgs<float> a{};
    a.geo = I.geo;
    a.ls = I.ls;
    for (int i = 0; i < qdwf::kMaxRhs; ++i) a.beta[i] = i < (int)cols.size() ? beta[i] : 0.0;
    a.p_out = P;
    a.cols = colset(cols);
    a.xr_fstride = fr;
    a.x_fstride = fx;
    a.sb = I.step_sb;
    a.mat = nullptr;
    a.real_mat = mats(a.cols, MAT_K, a.mats);
    a.x = nullptr;
    a.p = other();
    a.r = R;
    a.t = T;
    if (!qdwf::launch_cg_step_s(a, storage, I.stream)) throw ArgError("the three-phase CG step needs an even ls in [6, 16]");
    check_launch();
  }
  // the merged step with its scalars and column modes read from the device
  // CG state (graph-captured iterations); the new p goes to the other
  // buffer, which holds the previous p for the folded x update
  void step_dev(const std::vector<int>& cols)
  {
    qdwf::CgStepArgs<float> a{};
    a.p_out = other();
    a.geo = I.geo;
    a.ls = I.ls;
    a.cg = cg;
    a.ap = AP;
    a.cols = colset(cols);
    a.xr_fstride = fr;
    a.x_fstride = fx;
    a.sb = I.step_sb;
    a.mat = nullptr;
    a.real_mat = mats(a.cols, MAT_K, a.mats);
    a.x = X;
    a.p = P;
    a.r = R;
    a.t = T;
    if (!qdwf::launch_cg_step_s(a, storage, I.stream)) throw ArgError("the three-phase CG step needs an even ls in [6, 16]");
    check_launch();
    flip();
  }
  void hopm5(void* out, const void* in, int eo_out, bool dag, int mat, float alpha, const void* x, float beta,
             const void* dotp, const qdwf::RhsSet& cs, bool dotr = false)
  {
    qdwf::HopM5Args a{};
    a.dotp = dotp;
    a.dotr = dotr ? R : nullptr;
    a.dotr_fstride = fr;
    a.part = I.part;
    a.cols = cs;
    a.geo = I.geo;
    a.ls = I.ls;
    a.eo_out = eo_out;
    a.alpha = alpha;
    a.beta = beta;
    a.in = in;
    a.out = out;
    a.x = x;
    a.links = I.l32;
    a.links16 = half() ? I.l16 : nullptr;
    a.hop_mode = I.h2_hop;
    a.plain64 = I.plain64;
    a.sb = sb;
    a.site_major = I.site_major;
    a.cg = cg;
    a.mat = nullptr;
    if (mat >= 0) {
      a.real_mat = mats(cs, mat, a.mats);
    } else {  // plain hop
      for (auto*& m : a.mats) m = nullptr;
      a.real_mat = true;
    }
    if (!qdwf::launch_hop_m5s(a, dag, half(), I.stream))
      throw ArgError("the three-phase hop has no kernel for this Ls and block size");
    check_launch();
  }
  int64_t hop_blocks() const { return qdwf::hop_m5s_blocks(I.geo, sb); }
  // ap = Mpc^+ Mpc p on parity eo given t = K p, as y = p - H K H t,
  // ap = y - K^+ H^+ K^+ H^+ y.  I.red[i] = Re <p, ap> per column and, in
  // the merged form, I.red[n + i] = Re <r, ap>, I.red[2n + i] = |ap|^2.
  void hermop_dot(const qdwf::RhsSet& cs, int eo)
  {
    const int nv = merged() ? 3 : 1;
    const int64_t nblk = hop_blocks();
    if (nblk * cs.n * nv > I.part_cap) throw std::runtime_error("partial-sum buffer too small");
    hopm5(B, T, 3 - eo, false, MAT_K, 1.f, nullptr, 0.f, nullptr, cs);
    hopm5(Y, B, eo, false, -1, -1.f, P, 1.f, nullptr, cs);
    hopm5(B, Y, 3 - eo, true, MAT_K_DAG, 1.f, nullptr, 0.f, nullptr, cs);
    hopm5(AP, B, eo, true, MAT_K_DAG, -1.f, Y, 1.f, P, cs, merged());
    qdwf::launch_reduce_rows(I.part, (int)nblk, nv * cs.n, red_out, I.stream);
    check_launch();
    I.stats.hermop_low += cs.n;
  }
  // one CG iteration's kernels on the columns of cs (for timing)
  void iteration(const std::vector<int>& cols)
  {
    const qdwf::RhsSet cs = colset(cols);
    const std::vector<double> a(qdwf::kMaxRhs, 1e-3), b(qdwf::kMaxRhs, 0.5);
    hermop_dot(cs, 1);
    if (merged()) {
      step(cols, a.data(), b.data(), true, a.data());
    } else {
      axpy_norm(cs, a.data());
      step(cols, a.data(), b.data(), true);
    }
  }
  // r_i -= alpha_i ap_i, |r_i|^2 into I.red[i]
  void axpy_norm(const qdwf::RhsSet& cs, const double* alpha)
  {
    qdwf::AxpyNormSiteArgs an;
    an.ap = AP;
    an.r = R;
    an.n5 = n5();
    an.ls = I.ls;
    an.cols = cs;
    an.r_fstride = fr;
    an.part = I.part;
    for (int i = 0; i < qdwf::kMaxRhs; ++i) an.alpha[i] = i < cs.n ? alpha[i] : 0.0;
    const int g = qdwf::launch_axpy_norm_site(an, storage, I.stream);
    check_launch();
    qdwf::launch_reduce_rows(I.part, g, cs.n, red_out, I.stream);
    check_launch();
  }
  // column c of x += alpha * column c of the p-type stack v
  void axpy_to_x(double alpha, void* v, int c)
  {
    qdwf::launch_axpy_site(alpha, at(v, c), xat(c), I.ls, n5(), storage, I.stream);
    check_launch();
  }
};

// Storage mode of the three-phase solver for this inverter's solver.
int storage_of(const Inv& I)
{
  if (I.solver != QLAT_DWF_SOLVER_MIXED_FP16) return 0;
  return I.fp16_mode;
}

// Arithmetic of the hop phase of the three-phase kernels in a storage mode
// (the dispatch of launch_hop_m5s): FP16 for FP16 storage with the FP16
// links allocated and h2_hop on, FP32 otherwise.
int hop_arith_bits(const Inv& I, int storage) { return (storage && I.l16 && I.h2_hop) ? 16 : 32; }

// The reductions that the last kernels wrote into the mapped host buffer
// (visible to the host once the stream has drained: coherent memory).
void read_mapped(Inv& I, int n, double* out)
{
  sync(I);
  for (int i = 0; i < n; ++i) out[i] = I.red_map[i];
}

// Multi-RHS mixed-precision CG on the three-phase kernels (ThreePhase
// storage).  Arithmetic is FP32; the FP64 side is as in cg_mixed_multi:
// per-column reliable updates (x64 += x, r64 = b - A x64 in FP64) when a
// column's iterated residual dropped by delta, and retirement on the FP64
// true residual.  b64[c], x64[c]: FP64 source and solution (x64 is
// overwritten).  Returns the summed iterations.
long cg_mixed_multi_s(Inv& I, int ncol, double2* const* b64, double2* const* x64, double stop_rsd, long cap,
                      int storage, int sb = 0, Inv* const* colinv = nullptr, long* col_iters = nullptr)
{
  // Each column is solved for b / |b| (the low-precision vectors stay O(1)
  // whatever the source's scale); rr, targets and alpha are in those units.
  // Column c's operator is that of colinv[c] (default I); all of them share
  // I's stream, lattice and gauge field.
  ThreePhase tp(I, ncol, storage, sb);
  if (colinv)
    for (int c = 0; c < ncol; ++c) tp.colinv[c] = colinv[c];
  const int64_t nd = n64(I);
  double2* r64 = W(I, W_CG_R);
  double2* ax64 = W(I, W_CG_AP);
  std::vector<double> b2(ncol), bn(ncol), rr(ncol), maxrr(ncol);
  std::vector<long> iters(ncol, 0);
  std::vector<double> pend(ncol, 0.0);  // alpha of a direction not yet added to x
  std::vector<int> active;
  const double target = stop_rsd * stop_rsd;
  // |b - A x64|^2 / |b|^2 with r64 = b - A x64
  auto true_residual = [&](int c) {
    op_hermop<double>(*tp.colinv[c], ax64, x64[c], 1);
    I.stats.hermop_fp64 += 1;
    copy64(I, r64, b64[c]);
    qdwf::launch_axpy<double2>(-1.0, ax64, r64, nd, I.stream);
    check_launch();
    return dev_norm2(I, r64) / b2[c];
  };
  for (int c = 0; c < ncol; ++c) {
    b2[c] = dev_norm2(I, b64[c]);
    bn[c] = std::sqrt(b2[c]);
    zero64(I, x64[c]);
    if (b2[c] == 0.0) continue;  // x = 0 exactly
    const Inv& J = *tp.colinv[c];
    if (J.lm_nvec > 0) {
      // deflated initial guess (dslash.h:496-500), r = b - A x in FP64
      qdwf::DeflArgs a = defl_args(const_cast<Inv&>(J), 1);
      a.in = b64[c];
      a.out = x64[c];
      qdwf::launch_deflate(a, I.stream);
      check_launch();
      rr[c] = maxrr[c] = true_residual(c);
      if (rr[c] <= target) {
        const double rel = std::sqrt(rr[c]);
        if (rel > I.stats.true_residual) I.stats.true_residual = rel;
        continue;
      }
      tp.set_r(r64, c, 1.0 / bn[c]);
    } else {
      tp.set_r(b64[c], c, 1.0 / bn[c]);  // x = 0: r = b
      rr[c] = maxrr[c] = 1.0;
    }
    if (rr[c] > target) active.push_back(c);
  }
  tp.zero_p(ncol);
  tp.zero_x(0, ncol);
  const std::vector<double> zero(qdwf::kMaxRhs, 0.0);
  if (!active.empty()) tp.step(active, zero.data(), zero.data(), false);  // p = r, t = K p
  const double d2 = I.delta * I.delta;
  double dots[3 * qdwf::kMaxRhs], rr_new[qdwf::kMaxRhs], alpha[qdwf::kMaxRhs];
  const bool merged = tp.merged();
  long it = 0;
  if (merged && I.graph_iters > 0 && !active.empty()) {
    // The same iteration with its scalars on the device: K iterations
    // (hermop + reductions, k_cg_scalars, step) are captured once per active
    // set and replayed as one graph; the host reads the CG state after each
    // replay.  A column that reaches its reliable-update threshold adds
    // alpha p to x and is frozen on the device; the host then does exactly
    // the FP64 update of the loop below.
    if (!I.cg_dev) {
      QDWF_HIP(hipMalloc((void**)&I.cg_dev, sizeof(qdwf::CgState)));
      QDWF_HIP(hipHostMalloc((void**)&I.cg_host, sizeof(qdwf::CgState), hipHostMallocDefault));
    }
    qdwf::CgState& hs = *I.cg_host;
    hipGraphExec_t exec = nullptr;
    std::vector<int> cap_cols;
    int cap_k = 0, cap_cur = 0;
    auto drop = [&] {
      if (exec) (void)hipGraphExecDestroy(exec);
      exec = nullptr;
    };
    try {
      while (!active.empty() && it < cap) {
        const int n = (int)active.size();
        hs = qdwf::CgState{};
        hs.target = target;
        hs.d2 = d2;
        hs.x_defer = I.x_defer;
        for (int i = 0; i < n; ++i) {
          hs.rr[i] = rr[active[i]];
          hs.maxrr[i] = maxrr[active[i]];
          hs.pend[i] = pend[active[i]];
        }
        QDWF_HIP(hipMemcpyAsync(I.cg_dev, &hs, sizeof hs, hipMemcpyHostToDevice, I.stream));
        const int k = (int)std::min<long>(I.graph_iters, cap - it);
        // the captured kernels name the p buffers: a graph is replayed
        // only from the buffer parity it was captured at
        if (!exec || cap_cols != active || cap_k != k || cap_cur != tp.cur) {
          drop();
          const int cur0 = tp.cur;
          const qdwf::RhsSet cs = tp.colset(active);
          const qlat_dwf_solve_stats saved = I.stats;
          QDWF_HIP(hipStreamBeginCapture(I.stream, hipStreamCaptureModeThreadLocal));
          hipGraph_t g = nullptr;
          try {
            tp.cg = I.cg_dev;
            for (int j = 0; j < k; ++j) {
              tp.hermop_dot(cs, 1);
              qdwf::launch_cg_scalars(I.cg_dev, I.red, n, I.stream);
              check_launch();
              tp.step_dev(active);
            }
            tp.cg = nullptr;
          } catch (...) {
            tp.cg = nullptr;
            (void)hipStreamEndCapture(I.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          QDWF_HIP(hipStreamEndCapture(I.stream, &g));
          const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          QDWF_HIP(e);
          I.stats = saved;
          cap_cols = active;
          cap_k = k;
          cap_cur = cur0;
        } else {
          // the capture above advanced the host's buffer parity by k steps;
          // a replay advances it the same way
          for (int j = 0; j < k; ++j) tp.flip();
        }
        QDWF_HIP(hipGraphLaunch(exec, I.stream));
        QDWF_HIP(hipMemcpyAsync(&hs, I.cg_dev, sizeof hs, hipMemcpyDeviceToHost, I.stream));
        sync(I);
        it += k;
        // restarted columns whose last p is in the current p buffer (0) or
        // in the other one (1): a column frozen at step jf of the replay
        // kept its p in the buffer of that step while the k - jf later
        // steps flipped the buffers
        std::vector<int> keep, ru[2];
        std::vector<double> ru_beta[2];
        for (int i = 0; i < n; ++i) {
          const int c = active[i];
          iters[c] += hs.iters[i];
          I.stats.hermop_low += hs.iters[i];
          if (hs.bad[i])
            throw std::runtime_error(storage ? "CG breakdown: non-finite residual (FP16 storage, multi)"
                                             : "CG breakdown: non-finite residual (FP32 three-phase, multi)");
          if (hs.mode[i] == 0) {
            rr[c] = hs.rr[i];
            pend[c] = hs.pend[i];
            keep.push_back(c);
            continue;
          }
          pend[c] = 0.0;
          // reliable update (x += alpha_q q + alpha p was done by the frozen step)
          tp.add_x(x64[c], c, bn[c]);
          tp.zero_x(c, 1);
          const double rr_true = true_residual(c);
          I.stats.reliable_updates += 1;
          if (rr_true <= target) {
            const double rel = std::sqrt(rr_true);
            if (rel > I.stats.true_residual) I.stats.true_residual = rel;
            continue;  // column retired
          }
          tp.set_r(r64, c, 1.0 / bn[c]);
          const int jf = hs.iters[i] - 1;  // the step of this replay that froze the column
          const int b = (k - jf) & 1;
          ru[b].push_back(c);
          ru_beta[b].push_back(rr_true / hs.rr[i]);  // hs.rr: |r|^2 before the frozen iteration
          rr[c] = maxrr[c] = rr_true;
          keep.push_back(c);
        }
        // p = r + beta p into the current buffer, t = K p for the updated
        // columns (x, r unchanged)
        if (!ru[0].empty()) tp.step(ru[0], zero.data(), ru_beta[0].data(), true, zero.data());
        if (!ru[1].empty()) tp.restart_from_other(ru[1], ru_beta[1].data());
        active.swap(keep);
      }
    } catch (...) {
      drop();
      throw;
    }
    drop();
  }
  // the eager loop's reductions go straight to host memory
  tp.red_out = I.red_map_dev;
  while (!active.empty() && it < cap) {
    const qdwf::RhsSet cs = tp.colset(active);
    tp.hermop_dot(cs, 1);
    read_mapped(I, (merged ? 3 : 1) * cs.n, dots);
    for (int i = 0; i < cs.n; ++i) alpha[i] = rr[active[i]] / dots[i];
    if (merged) {
      // |r - alpha Ap|^2 = |r|^2 - 2 alpha Re <r, Ap> + alpha^2 |Ap|^2
      for (int i = 0; i < cs.n; ++i) {
        const double a = alpha[i];
        rr_new[i] = std::max(0.0, rr[active[i]] - 2.0 * a * dots[cs.n + i] + a * a * dots[2 * cs.n + i]);
      }
    } else {
      tp.axpy_norm(cs, alpha);
      read_mapped(I, cs.n, rr_new);
    }
    ++it;
    std::vector<int> keep;
    std::vector<double> ka, kb, kr, kq;
    for (int i = 0; i < cs.n; ++i) {
      const int c = active[i];
      iters[c] += 1;
      if (!std::isfinite(rr_new[i]) || !std::isfinite(alpha[i]))
        throw std::runtime_error(storage ? "CG breakdown: non-finite residual (FP16 storage, multi)"
                                         : "CG breakdown: non-finite residual (FP32 three-phase, multi)");
      double step_alpha = alpha[i];
      double xa = alpha[i], xq = 0.0;  // x += xa p + xq (previous p)
      if (rr_new[i] <= d2 * maxrr[c] || rr_new[i] <= target) {
        if (pend[c] != 0.0) tp.axpy_to_x(pend[c], tp.other(), c);  // the deferred previous step
        pend[c] = 0.0;
        xa = 0.0;
        tp.axpy_to_x(alpha[i], tp.P, c);
        tp.add_x(x64[c], c, bn[c]);
        tp.zero_x(c, 1);
        const double rr_true = true_residual(c);
        I.stats.reliable_updates += 1;
        if (rr_true <= target) {
          const double rel = std::sqrt(rr_true);
          if (rel > I.stats.true_residual) I.stats.true_residual = rel;
          continue;  // column retired
        }
        tp.set_r(r64, c, 1.0 / bn[c]);
        rr_new[i] = rr_true;
        maxrr[c] = rr_true;
        step_alpha = 0.0;
      } else if (I.x_defer) {
        // x gets alpha p every other step, two directions at once (p is
        // double-buffered, so the previous one is still there)
        if (pend[c] == 0.0) {
          pend[c] = alpha[i];
          xa = 0.0;
        } else {
          xq = pend[c];
          pend[c] = 0.0;
        }
      }
      keep.push_back(c);
      ka.push_back(xa);
      kq.push_back(xq);
      kb.push_back(rr_new[i] / rr[c]);
      kr.push_back(step_alpha);  // r was replaced by the true residual when 0
      rr[c] = rr_new[i];
    }
    if (!keep.empty()) tp.step(keep, ka.data(), kb.data(), true, merged ? kr.data() : nullptr, kq.data(), true);
    active.swap(keep);
  }
  long total = 0;
  for (int c = 0; c < ncol; ++c) total += iters[c];
  for (int c : active) {  // cap reached: fold in what we have
    if (pend[c] != 0.0) tp.axpy_to_x(pend[c], tp.other(), c);
    tp.add_x(x64[c], c, bn[c]);
    const double rel = std::sqrt(true_residual(c));
    if (rel > I.stats.true_residual) I.stats.true_residual = rel;
    total += 1;
    iters[c] += 1;  // like cg_with_f: max_num_iter + 1 when not converged
  }
  if (col_iters)
    for (int c = 0; c < ncol; ++c) col_iters[c] = iters[c];
  I.stats.spinor_bits = storage ? 16 : 32;
  I.stats.hop_arith_bits = hop_arith_bits(I, storage);
  return total;
}

// Tile width (columns per multi-RHS block, from {4, 8, 12, 16} up to
// max(ncol, 4)) and sites per block (16 or 32) of the three-phase solver
// with the least time per column and CG iteration, measured with HIP events
// on this lattice; cached per storage precision.
TileChoice autotune_tile(Inv& I, int storage, int ncol)
{
  TileChoice& t = I.tuned[storage];
  const int wmax = std::min(qdwf::kMaxRhs, std::max(ncol, 4));
  if (t.width && t.width <= wmax) return t;
  const qlat_dwf_solve_stats saved = I.stats;
  MV(I, M_R, wmax);  // allocate the stacks once at the largest width
  hipEvent_t e0, e1;
  QDWF_HIP(hipEventCreate(&e0));
  QDWF_HIP(hipEventCreate(&e1));
  double best = 0;
  TileChoice choice{0, 32};
  for (int sb : {32, 16, 8}) {
    if (!qdwf::three_phase_ok(I.ls, sb)) continue;
    for (int w : {4, 8, 12, 16}) {
      if (w > wmax) continue;
      ThreePhase tp(I, w, storage, sb);
      std::vector<int> cols;
      for (int c = 0; c < w; ++c) cols.push_back(c);
      const size_t lb = (size_t)tp.fh * tp.slot_bytes() * w;
      for (void* f : {tp.P, tp.T}) QDWF_HIP(hipMemsetAsync(f, 0x3f, lb, I.stream));
      QDWF_HIP(hipMemsetAsync(tp.R, 0x3f, (size_t)tp.fr * tp.rslot_bytes() * w, I.stream));
      tp.zero_x(0, w);
      tp.iteration(cols);  // warm-up
      QDWF_HIP(hipEventRecord(e0, I.stream));
      for (int k = 0; k < 2; ++k) tp.iteration(cols);
      QDWF_HIP(hipEventRecord(e1, I.stream));
      QDWF_HIP(hipEventSynchronize(e1));
      float ms = 0;
      QDWF_HIP(hipEventElapsedTime(&ms, e0, e1));
      const double per_col = ms / (2.0 * w);
      if (!choice.width || per_col < best) {
        best = per_col;
        choice = TileChoice{w, sb};
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  I.stats = saved;
  t = choice;
  return t;
}

// Tile used by the three-phase solver for ncol columns: QLAT_DWF_RHS_TILE
// (with QLAT_DWF_TP_SB) when set, else the autotuner's choice.
TileChoice tile_for(Inv& I, int storage, int ncol)
{
  if (I.rhs_tile > 0) return TileChoice{I.rhs_tile, I.tp_sb};
  return autotune_tile(I, storage, ncol);
}

bool use_three_phase(const Inv& I)
{
  return (I.solver == QLAT_DWF_SOLVER_MIXED_FP16 || I.three_phase32) && qdwf::three_phase_ls_ok(I.ls);
}

double true_residual_fp64(Inv& I, const double2* x, const double2* b)
{
  double2* r = W(I, W_CG_R);
  double2* ax = W(I, W_CG_AP);
  op_hermop<double>(I, ax, x, 1);
  I.stats.hermop_fp64 += 1;
  copy64(I, r, b);
  qdwf::launch_axpy<double2>(-1.0, ax, r, n64(I), I.stream);
  check_launch();
  const double b2 = dev_norm2(I, b);
  return b2 > 0 ? std::sqrt(dev_norm2(I, r) / b2) : 0.0;
}

// The reference's is_checking_invert() (dslash.h:11-16, default true):
// after each solve, |M x - D_- b| / |D_- b| on the full 5D system
// (dslash.h:522-528) goes into the solve statistics.
bool g_checking_invert = true;

// |M x - dm| / |dm| over both parities (multiply_m, dslash.cpp:1062-1080);
// uses W_TMP and W_ITMP.
double residual_5d(Inv& I, const double2* x_e, const double2* x_o, const double2* dm_e, const double2* dm_o)
{
  double2* re = W(I, W_TMP);
  double2* ro = W(I, W_ITMP);
  op_m_full(I, re, ro, x_e, x_o);
  qdwf::launch_axpy<double2>(-1.0, dm_e, re, n64(I), I.stream);
  qdwf::launch_axpy<double2>(-1.0, dm_o, ro, n64(I), I.stream);
  check_launch();
  const double num = dev_norm2(I, re) + dev_norm2(I, ro);
  const double den = dev_norm2(I, dm_e) + dev_norm2(I, dm_o);
  return den > 0 ? std::sqrt(num / den) : 0.0;
}

void note_residual_5d(Inv& I, double r)
{
  if (r > I.stats.true_residual_5d) I.stats.true_residual_5d = r;
}

// invert_with_cg (dslash.h:440-530) on device halves; in_e/in_o: 5D source
// halves (FP64), results in out_e/out_o.  stop_rsd < 0 selects the
// inverter's; dminus_already: the source is D_- b already (the
// dminus_multiplied_already argument, as invert_with_cg_with_guess passes it).
long invert_halves(Inv& I, const double2* in_e, const double2* in_o, double2* out_e, double2* out_o,
                   double stop_rsd = -1.0, bool dminus_already = false)
{
  if (stop_rsd < 0) stop_rsd = I.stop_rsd;
  double2* dm_e = W(I, W_DM_E);
  double2* dm_o = W(I, W_DM_O);
  if (I.dminus && !dminus_already) {
    op_d_minus(I, dm_e, dm_o, in_e, in_o);
  } else {
    copy64(I, dm_e, in_e);
    copy64(I, dm_o, in_o);
  }
  const double dm_in_qnorm = dev_norm2(I, dm_e) + dev_norm2(I, dm_o);
  if (dm_in_qnorm == 0.0) {
    zero64(I, out_e);
    zero64(I, out_o);
    return 0;
  }
  // set_odd_prec_field_sym2 (dslash.h:401-417)
  double2* out_e_p = W(I, W_OUT_EP);
  double2* in_o_p = W(I, W_IN_OP);
  double2* tmp = W(I, W_TMP);
  m5<double>(I, out_e_p, dm_e, MAT_MEE_INV);
  op_m_e_o<double>(I, tmp, out_e_p, 2);
  copy64(I, in_o_p, dm_o);
  qdwf::launch_axpy<double2>(-1.0, tmp, in_o_p, n64(I), I.stream);
  check_launch();
  op_mpcdag<double>(I, in_o_p, in_o_p, 1);
  double2* out_o_p = W(I, W_OUT_OP);
  zero64(I, out_o_p);
  long total_iter = 0;
  double true_rel = 0.0;
  if (I.solver == QLAT_DWF_SOLVER_FP64) {
    // outer restart loop (dslash.h:495-514)
    double2* itmp = W(I, W_ITMP);
    double2* x = W(I, W_CG_X);
    const double qnorm_in_o_p = dev_norm2(I, in_o_p);
    copy64(I, itmp, in_o_p);
    double qnorm_itmp = qnorm_in_o_p;
    long cycle;
    for (cycle = 1; cycle <= I.max_cycle; ++cycle) {
      // deflated initial guess when low modes are attached (dslash.h:496-500)
      const bool guess = I.lm_nvec > 0;
      if (guess) deflate_dev(I, x, itmp, 1);
      else zero64(I, x);
      const long iter =
          cg_fp64(I, x, itmp, stop_rsd * std::sqrt(qnorm_in_o_p / qnorm_itmp), I.max_num_iter, !guess);
      total_iter += iter;
      qdwf::launch_axpy<double2>(1.0, x, out_o_p, n64(I), I.stream);
      check_launch();
      if (iter <= I.max_num_iter) break;
      op_hermop<double>(I, itmp, out_o_p, 1);
      I.stats.hermop_fp64 += 1;
      qdwf::launch_xpay<double2>(in_o_p, -1.0, itmp, n64(I), I.stream);
      check_launch();
      qnorm_itmp = dev_norm2(I, itmp);
    }
    I.stats.cycles += (cycle > I.max_cycle) ? I.max_cycle : cycle;
    true_rel = true_residual_fp64(I, out_o_p, in_o_p);
  } else if (use_three_phase(I)) {
    const long cap = I.max_num_iter * I.max_cycle;
    const double saved = I.stats.true_residual;
    I.stats.true_residual = 0.0;
    total_iter = cg_mixed_multi_s(I, 1, &in_o_p, &out_o_p, stop_rsd, cap, storage_of(I));
    true_rel = I.stats.true_residual;
    I.stats.true_residual = saved;
    I.stats.cycles += 1;
  } else {
    const long cap = I.max_num_iter * I.max_cycle;
    // deflated initial guess when low modes are attached (dslash.h:496-500);
    // cg_mixed starts from the x it is given
    if (I.lm_nvec > 0) deflate_dev(I, out_o_p, in_o_p, 1);
    total_iter = cg_mixed(I, out_o_p, in_o_p, stop_rsd, cap, &true_rel);
    I.stats.cycles += 1;
    I.stats.spinor_bits = 32;
    I.stats.hop_arith_bits = 32;
  }
  I.stats.total_iter += total_iter;
  if (true_rel > I.stats.true_residual) I.stats.true_residual = true_rel;
  // restore_field_from_odd_prec_sym2 (dslash.h:419-438)
  m5<double>(I, out_o, out_o_p, MAT_MEE_INV);
  op_m_e_o<double>(I, tmp, out_o, 1);
  m5<double>(I, W(I, W_ITMP), tmp, MAT_MEE_INV);
  copy64(I, out_e, out_e_p);
  qdwf::launch_axpy<double2>(-1.0, W(I, W_ITMP), out_e, n64(I), I.stream);
  check_launch();
  if (g_checking_invert) note_residual_5d(I, residual_5d(I, out_e, out_o, dm_e, dm_o));
  return total_iter;
}

// invert_with_cg_with_guess (dslash.h:532-573) on device halves: out_e/out_o
// hold the guess on entry and the solution on exit.  The correction is
// solved from the residual D_- b - M x with stop_rsd scaled by
// |D_- b| / |D_- b - M x|.  out_e/out_o must not be W_IN_E, W_IN_O,
// W_OUT_E or W_OUT_O (the correction solve's fields); in_e/in_o may be
// W_IN_E/W_IN_O (they are read before those are reused).
long invert_halves_with_guess(Inv& I, const double2* in_e, const double2* in_o, double2* out_e, double2* out_o)
{
  double2* dm_e = W(I, W_DM_E);
  double2* dm_o = W(I, W_DM_O);
  if (I.dminus) {
    op_d_minus(I, dm_e, dm_o, in_e, in_o);
  } else {
    copy64(I, dm_e, in_e);
    copy64(I, dm_o, in_o);
  }
  const double q = dev_norm2(I, dm_e) + dev_norm2(I, dm_o);
  double2* me = W(I, W_TMP);
  double2* mo = W(I, W_ITMP);
  op_m_full(I, me, mo, out_e, out_o);
  // the residual is the source of the correction solve
  double2* re = W(I, W_OUT_E);
  double2* ro = W(I, W_OUT_O);
  QDWF_REQUIRE(re != out_e && ro != out_o && re != in_e && ro != in_o, "internal field aliasing");
  copy64(I, re, dm_e);
  copy64(I, ro, dm_o);
  qdwf::launch_axpy<double2>(-1.0, me, re, n64(I), I.stream);
  qdwf::launch_axpy<double2>(-1.0, mo, ro, n64(I), I.stream);
  check_launch();
  const double q_sub = dev_norm2(I, re) + dev_norm2(I, ro);
  if (q_sub == 0.0) return 0;  // the guess solves the system exactly
  const double ratio = std::sqrt(q / q_sub);
  double2* ce = W(I, W_IN_E);
  double2* co = W(I, W_IN_O);
  QDWF_REQUIRE(ce != out_e && co != out_o, "internal field aliasing");
  const long it = invert_halves(I, re, ro, ce, co, I.stop_rsd * ratio, true);
  qdwf::launch_axpy<double2>(1.0, ce, out_e, n64(I), I.stream);
  qdwf::launch_axpy<double2>(1.0, co, out_o, n64(I), I.stream);
  check_launch();
  return it;
}

void reset_stats(Inv& I) { std::memset(&I.stats, 0, sizeof(I.stats)); }

// Upload a host AoS buffer into a device staging buffer.
const double2* upload(Inv& I, int k, const double* host, size_t bytes)
{
  double2* d = stage(I, k, bytes);
  QDWF_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, I.stream));
  return d;
}
void download(Inv& I, double* host, const double2* dev, size_t bytes)
{
  QDWF_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, I.stream));
  sync(I);
}

size_t half_aos_bytes(const Inv& I) { return (size_t)I.geo.vh * I.ls * 12 * sizeof(double2); }
size_t full_aos_bytes(const Inv& I) { return 2 * half_aos_bytes(I); }
size_t prop_bytes(const Inv& I) { return (size_t)I.geo.vh * 2 * 144 * sizeof(double2); }
size_t stage_size(const Inv& I)
{
  const size_t a = full_aos_bytes(I), b = prop_bytes(I);
  return a > b ? a : b;
}

template <class F>
int guarded(F&& f)
{
  try {
    f();
    g_last_error.clear();
    return QLAT_DWF_OK;
  } catch (const ArgError& e) {
    g_last_error = e.what();
    return QLAT_DWF_EINVAL;
  } catch (const MemError& e) {
    g_last_error = e.what();
    return QLAT_DWF_ENOMEM;
  } catch (const HipError& e) {
    g_last_error = e.what();
    return QLAT_DWF_EHIP;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return QLAT_DWF_ESTATE;
  }
}

void check_inv(const Inv* I)
{
  if (!I) throw ArgError("null inverter handle");
  QDWF_HIP(hipSetDevice(I->device));
}
void check_eo(int eo) { QDWF_REQUIRE(eo == 1 || eo == 2, "parity eo must be 1 (odd) or 2 (even)"); }

enum HalfOp { H_MEE, H_MEE_DAG, H_MEE_INV, H_MEE_INVDAG, H_HOP, H_HOP_DAG, H_MEO, H_MDAG_EO, H_MPC, H_MPCDAG, H_HERMOP };

void apply_half_dev(Inv& I, int op, double2* out_soa, const double2* in_soa, int eo)
{
  switch (op) {
    case H_MEE: m5<double>(I, out_soa, in_soa, MAT_MEE); break;
    case H_MEE_DAG: m5<double>(I, out_soa, in_soa, MAT_MEE_DAG); break;
    case H_MEE_INV: m5<double>(I, out_soa, in_soa, MAT_MEE_INV); break;
    case H_MEE_INVDAG: m5<double>(I, out_soa, in_soa, MAT_MEE_INVDAG); break;
    case H_HOP: hop<double>(I, out_soa, in_soa, 3 - eo, false); break;
    case H_HOP_DAG: hop<double>(I, out_soa, in_soa, 3 - eo, true); break;
    case H_MEO: op_m_e_o<double>(I, out_soa, in_soa, eo); break;
    case H_MDAG_EO: op_mdag_e_o<double>(I, out_soa, in_soa, eo); break;
    case H_MPC: op_mpc<double>(I, out_soa, in_soa, eo); break;
    case H_MPCDAG: op_mpcdag<double>(I, out_soa, in_soa, eo); break;
    case H_HERMOP: op_hermop<double>(I, out_soa, in_soa, eo); break;
    default: throw ArgError("unknown operator");
  }
}

int apply_half_host(Inv* I, int op, double* out, const double* in, int eo)
{
  return guarded([&] {
    check_inv(I);
    check_eo(eo);
    QDWF_REQUIRE(out && in, "null field pointer");
    const size_t bytes = half_aos_bytes(*I);
    const double2* d_in = upload(*I, 0, in, bytes);
    double2* a = W(*I, W_IN_E);
    double2* b = W(*I, W_OUT_E);
    qdwf::launch_pack5(d_in, a, I->geo, I->ls, eo, 0, I->stream);
    check_launch();
    apply_half_dev(*I, op, b, a, eo);
    double2* d_out = stage(*I, 1, bytes);
    qdwf::launch_unpack5(b, d_out, I->geo, I->ls, eo, 0, I->stream);
    check_launch();
    download(*I, out, d_out, bytes);
  });
}

enum FullOp { F_M, F_MDAG, F_DMINUS };

void apply_full_dev(Inv& I, int op, double2* out_aos, const double2* in_aos)
{
  double2* ie = W(I, W_IN_E);
  double2* io = W(I, W_IN_O);
  double2* oe = W(I, W_OUT_E);
  double2* oo = W(I, W_OUT_O);
  qdwf::launch_pack5(in_aos, ie, I.geo, I.ls, 2, 1, I.stream);
  qdwf::launch_pack5(in_aos, io, I.geo, I.ls, 1, 1, I.stream);
  check_launch();
  switch (op) {
    case F_M: op_m_full(I, oe, oo, ie, io); break;
    case F_MDAG: op_mdag_full(I, oe, oo, ie, io); break;
    case F_DMINUS: op_d_minus(I, oe, oo, ie, io); break;
    default: throw ArgError("unknown operator");
  }
  qdwf::launch_unpack5(oe, out_aos, I.geo, I.ls, 2, 1, I.stream);
  qdwf::launch_unpack5(oo, out_aos, I.geo, I.ls, 1, 1, I.stream);
  check_launch();
}

int apply_full_host(Inv* I, int op, double* out, const double* in)
{
  return guarded([&] {
    check_inv(I);
    QDWF_REQUIRE(out && in, "null field pointer");
    const size_t bytes = full_aos_bytes(*I);
    const double2* d_in = upload(*I, 0, in, bytes);
    double2* d_out = stage(*I, 1, bytes);
    apply_full_dev(*I, op, d_out, d_in);
    download(*I, out, d_out, bytes);
  });
}

// 5D full AoS (device) -> solve -> 5D full AoS (device)
long invert_5d_dev(Inv& I, double2* out_aos, const double2* in_aos)
{
  double2* ie = W(I, W_IN_E);
  double2* io = W(I, W_IN_O);
  double2* oe = W(I, W_OUT_E);
  double2* oo = W(I, W_OUT_O);
  qdwf::launch_pack5(in_aos, ie, I.geo, I.ls, 2, 1, I.stream);
  qdwf::launch_pack5(in_aos, io, I.geo, I.ls, 1, 1, I.stream);
  check_launch();
  const long it = invert_halves(I, ie, io, oe, oo);
  qdwf::launch_unpack5(oe, out_aos, I.geo, I.ls, 2, 1, I.stream);
  qdwf::launch_unpack5(oo, out_aos, I.geo, I.ls, 1, 1, I.stream);
  check_launch();
  return it;
}

// D_- and the sym2 odd preconditioning of one 5D source (FP64 halves):
// b = odd-prec source, oep = out_e_p (set_odd_prec_field_sym2, dslash.h:401-417).
// Returns |D_- in|^2.
double prepare_column(Inv& I, const double2* in_e, const double2* in_o, double2* b, double2* oep)
{
  double2* dm_e = W(I, W_DM_E);
  double2* dm_o = W(I, W_DM_O);
  if (I.dminus) {
    op_d_minus(I, dm_e, dm_o, in_e, in_o);
  } else {
    copy64(I, dm_e, in_e);
    copy64(I, dm_o, in_o);
  }
  const double q = dev_norm2(I, dm_e) + dev_norm2(I, dm_o);
  double2* tmp = W(I, W_TMP);
  m5<double>(I, oep, dm_e, MAT_MEE_INV);  // out_e' = M_ee^-1 (D_- in)_e
  op_m_e_o<double>(I, tmp, oep, 2);
  copy64(I, b, dm_o);
  qdwf::launch_axpy<double2>(-1.0, tmp, b, n64(I), I.stream);
  check_launch();
  op_mpcdag<double>(I, b, b, 1);
  return q;
}

// out_e' = M_ee^-1 (D_- in)_e of set_odd_prec_field_sym2 alone (recomputed at
// restore time instead of being kept per column)
void out_e_p_of(Inv& I, const double2* in_e, const double2* in_o, double2* oep)
{
  double2* dm_e = W(I, W_DM_E);
  if (I.dminus) op_d_minus(I, dm_e, W(I, W_DM_O), in_e, in_o);
  else copy64(I, dm_e, in_e);
  m5<double>(I, oep, dm_e, MAT_MEE_INV);
}

// restore_field_from_odd_prec_sym2 (dslash.h:419-438)
void restore_column(Inv& I, const double2* x_o, const double2* oep, double2* out_e, double2* out_o)
{
  double2* tmp = W(I, W_TMP);
  m5<double>(I, out_o, x_o, MAT_MEE_INV);
  op_m_e_o<double>(I, tmp, out_o, 1);
  m5<double>(I, W(I, W_ITMP), tmp, MAT_MEE_INV);
  copy64(I, out_e, oep);
  qdwf::launch_axpy<double2>(-1.0, W(I, W_ITMP), out_e, n64(I), I.stream);
  check_launch();
}

// Multi-RHS driver of the mixed solvers: columns j < n, column j with the
// operator of colinv[j] (all sharing I's stream, lattice and gauge field).
// load(j, J, in_e, in_o) writes column j's 5D source halves, store(j, J,
// out_e, out_o) takes its solution halves; both run on the shared stream.
// The columns go in tiles through cg_mixed_multi_s (three-phase kernels) or
// cg_mixed_multi (FP32 fused kernels, one mass only).  Returns the summed
// iterations; col_iters[j] receives each column's.
template <class Load, class Store>
long solve_columns(Inv& I, int n, Inv* const* colinv, Load&& load, Store&& store, long* col_iters)
{
  bool several = false;
  for (int j = 0; j < n; ++j) several = several || colinv[j] != &I;
  const bool three = use_three_phase(I) || (several && qdwf::three_phase_ls_ok(I.ls));
  const int storage = storage_of(I);
  TileChoice tc{I.rhs_tile > 0 ? I.rhs_tile : 12, I.tp_sb};
  if (three) tc = tile_for(I, storage, n);
  tc.width = std::max(1, std::min(tc.width, qdwf::kMaxRhs));
  I.stats.rhs_tile = tc.width;
  I.stats.sites_per_block = three ? tc.sb : 32;
  long total = 0;
  for (int c0 = 0; c0 < n; c0 += tc.width) {
    const int w = std::min(tc.width, n - c0);
    double2* bs[qdwf::kMaxRhs];
    double2* xs[qdwf::kMaxRhs];
    for (int j = 0; j < w; ++j) {
      Inv& J = *colinv[c0 + j];
      double2* ie = W(J, W_IN_E);
      double2* io = W(J, W_IN_O);
      load(c0 + j, J, ie, io);
      bs[j] = C64(I, j, 0);
      xs[j] = C64(I, j, 1);
      const double q = prepare_column(J, ie, io, bs[j], W(J, W_OUT_EP));
      if (q == 0.0) zero64(I, bs[j]);  // dm_in == 0: solution zero (dslash.h:477-482)
    }
    if (three) {
      total += cg_mixed_multi_s(I, w, bs, xs, I.stop_rsd, I.max_num_iter * I.max_cycle, storage, tc.sb, colinv + c0,
                                col_iters ? col_iters + c0 : nullptr);
    } else {
      for (int j = 0; j < w; ++j)
        if (colinv[c0 + j] != &I) throw ArgError("batched solves over several inverters need the three-phase solver");
      const long it = cg_mixed_multi(I, w, I.stop_rsd, I.max_num_iter * I.max_cycle);
      total += it;
      if (col_iters)
        for (int j = 0; j < w; ++j) col_iters[c0 + j] = it / w;  // the FP32 fused path counts only the sum
      I.stats.spinor_bits = 32;
      I.stats.hop_arith_bits = 32;
    }
    I.stats.cycles += 1;
    for (int j = 0; j < w; ++j) {
      Inv& J = *colinv[c0 + j];
      double2* ie = W(J, W_IN_E);
      double2* io = W(J, W_IN_O);
      double2* oe = W(J, W_OUT_E);
      double2* oo = W(J, W_OUT_O);
      double2* oep = W(J, W_OUT_EP);
      load(c0 + j, J, ie, io);
      out_e_p_of(J, ie, io, oep);
      restore_column(J, xs[j], oep, oe, oo);
      if (g_checking_invert) {
        // out_e_p_of left (D_- b)_e in W_DM_E, and (D_- b)_o in W_DM_O when D_- is applied
        const double r = residual_5d(J, oe, oo, W(J, W_DM_E), J.dminus ? W(J, W_DM_O) : io);
        note_residual_5d(I, r);
      }
      store(c0 + j, J, oe, oo);
    }
  }
  I.stats.total_iter += total;
  return total;
}

// Propagator (or 4D fermion, ncol = 1 with col = -1) on device.
long invert_prop_dev(Inv& I, double2* sol, const double2* src, bool is_prop)
{
  double2* ie = W(I, W_IN_E);
  double2* io = W(I, W_IN_O);
  double2* oe = W(I, W_OUT_E);
  double2* oo = W(I, W_OUT_O);
  long total = 0;
  const int ncol = is_prop ? 12 : 1;
  if (I.solver != QLAT_DWF_SOLVER_FP64 && is_prop && I.rhs_tile != 1) {
    std::vector<Inv*> colinv(ncol, &I);
    return solve_columns(
        I, ncol, colinv.data(),
        [&](int j, Inv& J, double2* e, double2* o) {
          qdwf::launch_prop_col_to_5d(src, e, J.geo, J.ls, 2, j, J.stream);
          qdwf::launch_prop_col_to_5d(src, o, J.geo, J.ls, 1, j, J.stream);
          check_launch();
        },
        [&](int j, Inv& J, const double2* e, const double2* o) {
          qdwf::launch_5d_to_prop_col(e, sol, J.geo, J.ls, 2, j, J.stream);
          qdwf::launch_5d_to_prop_col(o, sol, J.geo, J.ls, 1, j, J.stream);
          check_launch();
        },
        nullptr);
  }
  for (int j = 0; j < ncol; ++j) {
    const int col = is_prop ? j : -1;
    qdwf::launch_prop_col_to_5d(src, ie, I.geo, I.ls, 2, col, I.stream);
    qdwf::launch_prop_col_to_5d(src, io, I.geo, I.ls, 1, col, I.stream);
    check_launch();
    total += invert_halves(I, ie, io, oe, oo);
    qdwf::launch_5d_to_prop_col(oe, sol, I.geo, I.ls, 2, col, I.stream);
    qdwf::launch_5d_to_prop_col(oo, sol, I.geo, I.ls, 1, col, I.stream);
    check_launch();
  }
  return total;
}

// Every inverter of a batch runs on the first one's stream for the call.
struct SharedStream {
  std::vector<Inv*> invs;
  std::vector<hipStream_t> saved;
  SharedStream(Inv* const* v, int n)
  {
    for (int j = 0; j < n; ++j) {
      invs.push_back(v[j]);
      saved.push_back(v[j]->stream);
    }
    for (int j = 1; j < n; ++j) v[j]->stream = v[0]->stream;
  }
  ~SharedStream()
  {
    for (size_t j = 0; j < invs.size(); ++j) invs[j]->stream = saved[j];
  }
};

void check_batch(qlat_dwf_inverter* const* invs, int n)
{
  QDWF_REQUIRE(invs && n > 0, "invalid argument");
  for (int j = 0; j < n; ++j) {
    check_inv(invs[j]);
    const Inv& A = *invs[0];
    const Inv& B = *invs[j];
    QDWF_REQUIRE(B.device == A.device, "batched inverters must share one device");
    QDWF_REQUIRE(B.ls == A.ls && B.geo.by == A.geo.by && B.geo.bz == A.geo.bz, "batched inverters must share Ls and the site order");
    for (int mu = 0; mu < 4; ++mu) QDWF_REQUIRE(B.geo.n[mu] == A.geo.n[mu], "batched inverters must share the lattice");
  }
  QDWF_HIP(hipSetDevice(invs[0]->device));
}

// sol[j] = invert_5d(src[j]) with invs[j] (device AoS full 5D fields)
long invert_5d_batch_dev(qlat_dwf_inverter* const* invs, int n, double2* const* sol, const double2* const* src,
                         long* iters)
{
  Inv& I = *invs[0];
  SharedStream shared(invs, n);
  for (int j = 0; j < n; ++j) reset_stats(*invs[j]);
  long total = 0;
  if (I.solver == QLAT_DWF_SOLVER_FP64) {
    for (int j = 0; j < n; ++j) {
      Inv& J = *invs[j];
      const double s0 = J.stop_rsd;
      const long m0 = J.max_num_iter, c0 = J.max_cycle;
      J.stop_rsd = I.stop_rsd;
      J.max_num_iter = I.max_num_iter;
      J.max_cycle = I.max_cycle;
      long it = 0;
      try {
        it = invert_5d_dev(J, sol[j], src[j]);
      } catch (...) {
        J.stop_rsd = s0;
        J.max_num_iter = m0;
        J.max_cycle = c0;
        throw;
      }
      J.stop_rsd = s0;
      J.max_num_iter = m0;
      J.max_cycle = c0;
      iters[j] = it;
      total += it;
      I.stats.total_iter += j ? it : 0;
      I.stats.hermop_fp64 += j ? J.stats.hermop_fp64 : 0;
      I.stats.true_residual = std::max(I.stats.true_residual, J.stats.true_residual);
      if (j) note_residual_5d(I, J.stats.true_residual_5d);
    }
    return total;
  }
  std::vector<Inv*> colinv(invs, invs + n);
  return solve_columns(
      I, n, colinv.data(),
      [&](int j, Inv& J, double2* e, double2* o) {
        qdwf::launch_pack5(src[j], e, J.geo, J.ls, 2, 1, J.stream);
        qdwf::launch_pack5(src[j], o, J.geo, J.ls, 1, 1, J.stream);
        check_launch();
      },
      [&](int j, Inv& J, const double2* e, const double2* o) {
        qdwf::launch_unpack5(e, sol[j], J.geo, J.ls, 2, 1, J.stream);
        qdwf::launch_unpack5(o, sol[j], J.geo, J.ls, 1, 1, J.stream);
        check_launch();
      },
      iters);
}

// Pack an AoS gauge field already in device memory into the link layouts.
void pack_gauge_dev(Inv& I, const double2* g)
{
  const size_t lcount = (size_t)2 * 8 * 9 * I.geo.stride;
  if (!I.l64) {
    I.l64 = (double2*)dev_alloc(lcount * sizeof(double2));
    I.l32 = (float2*)dev_alloc(lcount * sizeof(float2));
    I.l16 = (unsigned*)dev_alloc(lcount * sizeof(unsigned));
    QDWF_HIP(hipMemsetAsync(I.l64, 0, lcount * sizeof(double2), I.stream));
    QDWF_HIP(hipMemsetAsync(I.l32, 0, lcount * sizeof(float2), I.stream));
    QDWF_HIP(hipMemsetAsync(I.l16, 0, lcount * sizeof(unsigned), I.stream));
  }
  qdwf::launch_pack_gauge(g, I.l64, I.l32, I.l16, I.geo, I.stream);
  check_launch();
}

// Fermion force of S = phi^+ (Mpc^+ Mpc)^-1 phi, phi an odd half field in
// reference layout (device).  X = (Mpc^+ Mpc)^-1 phi is recomputed here by
// the FP64 CG (nothing is kept between calls); with
//   a = K X, c = K H a, Y = X - H c = Mpc X, w = K^+ H^+ Y
// dS = -X^+ d(Mpc^+ Mpc) X = 2 Re[Y^+ dH c + w^+ dH a], whose link
// derivative k_force evaluates with A = (w even, Y odd), B = (c even, a odd).
long fermion_force_dev(Inv& I, const double2* phi, double stop_rsd, double2* force, double2* grad, double* action,
                       double* action_lo)
{
  double2* b = W(I, W_IN_OP);
  double2* X = W(I, W_OUT_OP);
  qdwf::launch_pack5(phi, b, I.geo, I.ls, 1, 0, I.stream);
  check_launch();
  zero64(I, X);
  const long iter = cg_fp64(I, X, b, stop_rsd, I.max_num_iter * I.max_cycle, true);
  I.stats.total_iter += iter;
  I.stats.true_residual = true_residual_fp64(I, X, b);
  double2* a1 = W(I, W_IN_E);
  double2* c1 = W(I, W_IN_O);
  double2* Y = W(I, W_OUT_E);
  double2* w = W(I, W_OUT_O);
  double2* tmp = W(I, W_TMP);
  m5<double>(I, a1, X, MAT_K);               // a = K X          (odd)
  hop<double>(I, tmp, a1, 2, false);          // H a              (even)
  m5<double>(I, c1, tmp, MAT_K);              // c = K H a        (even)
  hop<double>(I, Y, c1, 1, false, -1.0, X);   // Y = X - H c      (odd)
  hop<double>(I, tmp, Y, 2, true);            // H^+ Y            (even)
  m5<double>(I, w, tmp, MAT_K_DAG);           // w = K^+ H^+ Y    (even)
  // S = 2 Re <phi, X> - |Mpc X|^2 differs from phi^+ A^-1 phi only by
  // (X - A^-1 phi)^+ A (X - A^-1 phi): second order in the solver error.
  // The sum is compensated so that S(U + dU) - S(U - dU) keeps its digits.
  qdwf::launch_action_dd(b, X, Y, n64(I), I.part, I.red_map_dev, I.stream);
  check_launch();
  sync(I);
  *action = I.red_map[0];
  if (action_lo) *action_lo = I.red_map[1];
  qdwf::ForceArgs fa;
  fa.geo = I.geo;
  fa.ls = I.ls;
  fa.A[0] = w;
  fa.A[1] = Y;
  fa.B[0] = c1;
  fa.B[1] = a1;
  fa.links = I.l64;
  fa.force = force;
  fa.grad = grad;
  qdwf::launch_force(fa, I.stream);
  check_launch();
  return iter;
}

void build_matrices(Inv& I)
{
  const int ls = I.ls;
  std::vector<cd> bt(ls), ct(ls), negc(ls);
  for (int s = 0; s < ls; ++s) {
    bt[s] = 1.0 + I.bs[s] * (4.0 - I.m5);  // dslash.cpp:617-621
    ct[s] = 1.0 - I.cs[s] * (4.0 - I.m5);
    negc[s] = -I.cs[s];
  }
  std::vector<cd> ones(ls, 1.0), zeros(ls, 0.0);
  std::vector<cd> all((size_t)MAT_COUNT * 2 * ls * ls);
  for (int chi = 0; chi < 2; ++chi) {
    const Mat S = shift_matrix(chi, ls, I.mass);
    std::vector<cd> mct(ls);
    for (int s = 0; s < ls; ++s) mct[s] = -ct[s];
    const Mat mee = diag_plus(bt, mct, S, ls);
    const Mat mee_inv = mat_inv(mee, ls);
    const Mat pm = diag_plus(I.bs, I.cs, S, ls);
    const Mat k = mat_mul(pm, mee_inv, ls);
    Mat negcm(ls * ls, 0.0);
    for (int s = 0; s < ls; ++s) negcm[s * ls + s] = negc[s];
    const Mat list[MAT_COUNT] = {mee, mat_adj(mee, ls), mee_inv, mat_adj(mee_inv, ls), pm, mat_adj(pm, ls),
                                 k, mat_adj(k, ls), negcm};
    for (int m = 0; m < MAT_COUNT; ++m)
      std::memcpy(&all[((size_t)m * 2 + chi) * ls * ls], list[m].data(), sizeof(cd) * ls * ls);
  }
  for (int m = 0; m < MAT_COUNT; ++m) {
    bool real = true;
    for (size_t i = (size_t)m * 2 * ls * ls; i < (size_t)(m + 1) * 2 * ls * ls; ++i)
      if (all[i].imag() != 0.0) real = false;
    I.real_mat[m] = real;
  }
  std::vector<cplx<double>> h64(all.size());
  std::vector<cplx<float>> h32(all.size());
  for (size_t i = 0; i < all.size(); ++i) {
    h64[i].re = all[i].real();
    h64[i].im = all[i].imag();
    h32[i].re = (float)all[i].real();
    h32[i].im = (float)all[i].imag();
  }
  I.mat64 = (cplx<double>*)dev_alloc(h64.size() * sizeof(cplx<double>));
  I.mat32 = (cplx<float>*)dev_alloc(h32.size() * sizeof(cplx<float>));
  QDWF_HIP(hipMemcpy(I.mat64, h64.data(), h64.size() * sizeof(cplx<double>), hipMemcpyHostToDevice));
  QDWF_HIP(hipMemcpy(I.mat32, h32.data(), h32.size() * sizeof(cplx<float>), hipMemcpyHostToDevice));
  std::vector<cplx<double>> b64(2 * ls);
  std::vector<cplx<float>> b32(2 * ls);
  for (int s = 0; s < ls; ++s) {
    b64[s].re = 1.0;
    b64[s].im = 0.0;
    b64[ls + s].re = ct[s].real();
    b64[ls + s].im = ct[s].imag();
    b32[s].re = 1.0f;
    b32[s].im = 0.0f;
    b32[ls + s].re = (float)ct[s].real();
    b32[ls + s].im = (float)ct[s].imag();
  }
  I.beta64 = (cplx<double>*)dev_alloc(b64.size() * sizeof(cplx<double>));
  I.beta32 = (cplx<float>*)dev_alloc(b32.size() * sizeof(cplx<float>));
  QDWF_HIP(hipMemcpy(I.beta64, b64.data(), b64.size() * sizeof(cplx<double>), hipMemcpyHostToDevice));
  QDWF_HIP(hipMemcpy(I.beta32, b32.data(), b32.size() * sizeof(cplx<float>), hipMemcpyHostToDevice));
}

void free_all(Inv* I)
{
  if (!I) return;
  (void)hipSetDevice(I->device);
  if (I->stream) (void)hipStreamSynchronize(I->stream);
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(I->l64);
  f(I->l32);
  f(I->l16);
  f(I->mat64);
  f(I->mat32);
  f(I->beta64);
  f(I->beta32);
  for (auto* p : I->w64) f(p);
  for (auto* p : I->w32) f(p);
  for (auto* p : I->stage) f(p);
  for (auto* p : I->m32) f(p);
  for (auto* p : I->mh) f(p);
  for (auto* p : I->col64) f(p);
  f(I->part);
  f(I->red);
  if (I->red_map) (void)hipHostFree(I->red_map);
  f(I->cg_dev);
  if (I->cg_host) (void)hipHostFree(I->cg_host);
  f(I->lm_basis);
  f(I->lm_coef);
  f(I->lm_evals);
  for (auto* p : I->lm_hmap) f(p);
  f(I->lm_part);
  f(I->lm_p);
  if (I->stream) (void)hipStreamDestroy(I->stream);
  delete I;
}

}  // namespace

// ================================================================== C ABI
extern "C" {

const char* qlat_dwf_last_error(void) { return g_last_error.c_str(); }

int qlat_dwf_abi_version(void) { return 4; }

int qlat_dwf_set_checking_invert(int on)
{
  g_checking_invert = on != 0;
  g_last_error.clear();
  return QLAT_DWF_OK;
}

int qlat_dwf_is_checking_invert(void) { return g_checking_invert ? 1 : 0; }

int qlat_dwf_inverter_create(qlat_dwf_inverter** out, const int total_site[4], const double* gauge,
                             const qlat_dwf_fermion_action* fa)
{
  Inv* I = nullptr;
  const int rc = guarded([&] {
    QDWF_REQUIRE(out && total_site && gauge && fa, "null argument to qlat_dwf_inverter_create");
    *out = nullptr;
    int ndev = 0;
    QDWF_HIP(hipGetDeviceCount(&ndev));
    QDWF_REQUIRE(ndev > 0, "no HIP device visible");
    int64_t V = 1;
    for (int i = 0; i < 4; ++i) {
      QDWF_REQUIRE(total_site[i] >= 2 && total_site[i] % 2 == 0,
                   "every lattice extent must be even and >= 2 (even-odd layout, core.h:232-261)");
      V *= total_site[i];
    }
    QDWF_REQUIRE(V / 2 < (1ll << 30), "lattice too large for 32-bit site indices");
    QDWF_REQUIRE((V / 2 + 255) * (int64_t)fa->ls * 12 < (1ll << 32), "lattice x Ls too large for 32-bit field indices");
    QDWF_REQUIRE(fa->ls >= 2 && fa->ls <= 32, "ls must be in [2, 32]");
    QDWF_REQUIRE(fa->cg_diagonal_mee == 2 || fa->cg_diagonal_mee == 0,
                 "only cg_diagonal_mee == 2 is supported (dslash.h:483, :520)");
    I = new Inv;
    QDWF_HIP(hipGetDevice(&I->device));
    init_scratch(I->device);
    QDWF_HIP(hipStreamCreate(&I->stream));
    for (int i = 0; i < 4; ++i) I->geo.n[i] = total_site[i];
    I->geo.xh = total_site[0] / 2;
    // device site order: tiles of `by` rows in y and `bz` planes in z, each
    // swept in t (dwf_common.h); by = Y, bz = Z is the reference order.
    // Defaults (measured at 24^3x64): by = the divisor of Y closest to 6,
    // bz = Z (z tiles cut L2 misses of the FP16 hop by 20% without making
    // it faster, and slow the FP32/FP64 hops by ~7%).
    // QLAT_DWF_TILE_Y / QLAT_DWF_TILE_Z override the defaults.
    auto tile_of = [](int n, int pref, const char* env) {
      int b = n;
      for (int d = 1; d <= n; ++d)
        if (n % d == 0 && std::abs(d - pref) < std::abs(b - pref)) b = d;
      if (const char* e = std::getenv(env)) {
        const int v = std::atoi(e);
        if (v > 0 && n % v == 0) b = v;
      }
      return b;
    };
    I->geo.by = tile_of(total_site[1], 6, "QLAT_DWF_TILE_Y");
    I->geo.bz = tile_of(total_site[2], total_site[2], "QLAT_DWF_TILE_Z");
    I->geo.vh = V / 2;
    I->geo.stride = (I->geo.vh + qdwf::kStrideAlign - 1) / qdwf::kStrideAlign * qdwf::kStrideAlign;
    I->geo.dxh = qdwf::make_fastdiv((unsigned)I->geo.xh);
    I->geo.dby = qdwf::make_fastdiv((unsigned)I->geo.by);
    I->geo.dbz = qdwf::make_fastdiv((unsigned)I->geo.bz);
    I->geo.dnt = qdwf::make_fastdiv((unsigned)total_site[3]);
    I->geo.dnzt = qdwf::make_fastdiv((unsigned)(total_site[2] / I->geo.bz));
    I->ls = fa->ls;
    I->mass = fa->mass;
    I->m5 = fa->m5;
    I->mobius_scale = fa->mobius_scale;
    I->dminus = fa->is_multiplying_dminus != 0;
    I->bs.resize(fa->ls);
    I->cs.resize(fa->ls);
    for (int s = 0; s < fa->ls; ++s) {
      if (fa->bs && fa->cs) {
        I->bs[s] = cd(fa->bs[2 * s], fa->bs[2 * s + 1]);
        I->cs[s] = cd(fa->cs[2 * s], fa->cs[2 * s + 1]);
      } else {
        I->bs[s] = 0.5 * fa->mobius_scale + 0.5;
        I->cs[s] = 0.5 * fa->mobius_scale - 0.5;
      }
    }
    // gauge upload and conversion
    const size_t gbytes = (size_t)V * 36 * sizeof(double2);
    double2* g = (double2*)dev_alloc(gbytes);
    QDWF_HIP(hipMemcpyAsync(g, gauge, gbytes, hipMemcpyHostToDevice, I->stream));
    pack_gauge_dev(*I, g);
    QDWF_HIP(hipStreamSynchronize(I->stream));
    QDWF_HIP(hipFree(g));
    build_matrices(*I);
    // hop launch shape: QLAT_DWF_HOP_SPT32 / QLAT_DWF_HOP_SPT64 (see launch_hop)
    if (const char* e = std::getenv("QLAT_DWF_HOP_SPT32")) I->hop_spt32 = std::atoi(e);
    if (const char* e = std::getenv("QLAT_DWF_HOP_SPT64")) I->hop_spt64 = std::atoi(e);
    if (const char* e = std::getenv("QLAT_DWF_FUSED")) I->fused32 = std::atoi(e) != 0;
    if (const char* e = std::getenv("QLAT_DWF_FUSED64")) I->fused64 = std::atoi(e) != 0;
    if (const char* e = std::getenv("QLAT_DWF_THREE_PHASE")) I->three_phase32 = std::atoi(e) != 0;
    if (const char* e = std::getenv("QLAT_DWF_RHS_TILE")) I->rhs_tile = std::max(0, std::min(qdwf::kMaxRhs, std::atoi(e)));
    if (const char* e = std::getenv("QLAT_DWF_TP_SB")) {
      const int v = std::atoi(e);
      I->tp_sb = (v == 8 || v == 16) ? v : 32;
    }
    if (const char* e = std::getenv("QLAT_DWF_SITE_MAJOR")) I->site_major = std::atoi(e) != 0;
    if (const char* e = std::getenv("QLAT_DWF_PLAIN64")) I->plain64 = std::atoi(e) != 0;
    if (const char* e = std::getenv("QLAT_DWF_FP16_MODE")) I->fp16_mode = std::max(1, std::min(2, std::atoi(e)));
    if (const char* e = std::getenv("QLAT_DWF_XDEFER")) I->x_defer = std::atoi(e) != 0;
    if (const char* e = std::getenv("QLAT_DWF_STEP_SB")) I->step_sb = std::atoi(e) == 16 ? 16 : 32;
    if (const char* e = std::getenv("QLAT_DWF_GRAPH")) I->graph_iters = std::max(0, std::min(64, std::atoi(e)));
    if (const char* e = std::getenv("QLAT_DWF_H2")) I->h2_hop = std::atoi(e) != 0 ? 1 : 0;
    {
      // partial sums per column: BLAS grids, one per block of the fused
      // kernels (32 sites) and of the three-phase kernels (8 sites or more,
      // up to three sums each), or one per (256-site block, slot) (s-matrix
      // kernel)
      int64_t np = qdwf::kMaxBlasGrid * 4;
      np = std::max<int64_t>(np, (I->geo.vh + 31) / 32 + 16);
      np = std::max<int64_t>(np, 3 * ((I->geo.vh + 7) / 8) + 16);
      np = std::max<int64_t>(np, (I->geo.vh + qdwf::kM5Block - 1) / qdwf::kM5Block * 12 + 16);
      I->part_cap = np * qdwf::kMaxRhs;
      I->part = (double*)dev_alloc(I->part_cap * sizeof(double));
    }
    I->red = (double*)dev_alloc(4 * qdwf::kMaxRhs * sizeof(double));
    QDWF_HIP(hipHostMalloc((void**)&I->red_map, 4 * qdwf::kMaxRhs * sizeof(double),
                           hipHostMallocMapped | hipHostMallocCoherent));
    QDWF_HIP(hipHostGetDevicePointer((void**)&I->red_map_dev, I->red_map, 0));
    *out = I;
  });
  if (rc != QLAT_DWF_OK && I) {
    const std::string msg = g_last_error;
    free_all(I);
    g_last_error = msg;
  }
  return rc;
}

int qlat_dwf_inverter_destroy(qlat_dwf_inverter* inv)
{
  return guarded([&] {
    if (inv) free_all(inv);
  });
}

int qlat_dwf_inverter_set_params(qlat_dwf_inverter* inv, double stop_rsd, long max_num_iter,
                                 long max_mixed_precision_cycle)
{
  return guarded([&] {
    QDWF_REQUIRE(inv, "null inverter handle");
    QDWF_REQUIRE(stop_rsd >= 0 && max_num_iter >= 0 && max_mixed_precision_cycle >= 1, "invalid solver parameters");
    inv->stop_rsd = stop_rsd;
    inv->max_num_iter = max_num_iter;
    inv->max_cycle = max_mixed_precision_cycle;
  });
}

int qlat_dwf_inverter_get_params(const qlat_dwf_inverter* inv, double* stop_rsd, long* max_num_iter,
                                 long* max_mixed_precision_cycle)
{
  return guarded([&] {
    QDWF_REQUIRE(inv, "null inverter handle");
    if (stop_rsd) *stop_rsd = inv->stop_rsd;
    if (max_num_iter) *max_num_iter = inv->max_num_iter;
    if (max_mixed_precision_cycle) *max_mixed_precision_cycle = inv->max_cycle;
  });
}

int qlat_dwf_inverter_set_solver(qlat_dwf_inverter* inv, int solver, double reliable_delta)
{
  return guarded([&] {
    QDWF_REQUIRE(inv, "null inverter handle");
    QDWF_REQUIRE(solver == QLAT_DWF_SOLVER_FP64 || solver == QLAT_DWF_SOLVER_MIXED_FP32 ||
                     solver == QLAT_DWF_SOLVER_MIXED_FP16,
                 "unknown solver");
    QDWF_REQUIRE(reliable_delta >= 0 && reliable_delta < 1, "reliable_delta must be in [0, 1)");
    inv->solver = solver;
    inv->delta = reliable_delta > 0 ? reliable_delta : 0.01;
  });
}

int qlat_dwf_inverter_get_stats(const qlat_dwf_inverter* inv, qlat_dwf_solve_stats* st)
{
  return guarded([&] {
    QDWF_REQUIRE(inv && st, "null argument");
    *st = inv->stats;
  });
}

int qlat_dwf_multiply_m_e_e(qlat_dwf_inverter* inv, double* out, const double* in)
{
  return apply_half_host(inv, H_MEE, out, in, 1);
}
int qlat_dwf_multiply_mdag_e_e(qlat_dwf_inverter* inv, double* out, const double* in)
{
  return apply_half_host(inv, H_MEE_DAG, out, in, 1);
}
int qlat_dwf_multiply_m_e_e_inv(qlat_dwf_inverter* inv, double* out, const double* in)
{
  return apply_half_host(inv, H_MEE_INV, out, in, 1);
}
int qlat_dwf_multiply_mdag_e_e_inv(qlat_dwf_inverter* inv, double* out, const double* in)
{
  return apply_half_host(inv, H_MEE_INVDAG, out, in, 1);
}
int qlat_dwf_multiply_wilson_d_e_o(qlat_dwf_inverter* inv, double* out, const double* in, int eo_in)
{
  return apply_half_host(inv, H_HOP, out, in, eo_in);
}
int qlat_dwf_multiply_wilson_ddag_e_o(qlat_dwf_inverter* inv, double* out, const double* in, int eo_in)
{
  return apply_half_host(inv, H_HOP_DAG, out, in, eo_in);
}
int qlat_dwf_multiply_m_e_o(qlat_dwf_inverter* inv, double* out, const double* in, int eo_in)
{
  return apply_half_host(inv, H_MEO, out, in, eo_in);
}
int qlat_dwf_multiply_mdag_e_o(qlat_dwf_inverter* inv, double* out, const double* in, int eo_in)
{
  return apply_half_host(inv, H_MDAG_EO, out, in, eo_in);
}
int qlat_dwf_multiply_mpc_sym2(qlat_dwf_inverter* inv, double* out, const double* in, int eo)
{
  return apply_half_host(inv, H_MPC, out, in, eo);
}
int qlat_dwf_multiply_mpcdag_sym2(qlat_dwf_inverter* inv, double* out, const double* in, int eo)
{
  return apply_half_host(inv, H_MPCDAG, out, in, eo);
}
int qlat_dwf_multiply_hermop_sym2(qlat_dwf_inverter* inv, double* out, const double* in, int eo)
{
  return apply_half_host(inv, H_HERMOP, out, in, eo);
}
int qlat_dwf_multiply_m(qlat_dwf_inverter* inv, double* out, const double* in)
{
  return apply_full_host(inv, F_M, out, in);
}
int qlat_dwf_multiply_mdag(qlat_dwf_inverter* inv, double* out, const double* in)
{
  return apply_full_host(inv, F_MDAG, out, in);
}
int qlat_dwf_multiply_d_minus(qlat_dwf_inverter* inv, double* out, const double* in)
{
  return apply_full_host(inv, F_DMINUS, out, in);
}

int qlat_dwf_dot_product(qlat_dwf_inverter* inv, const double* a, const double* b, long n_wv, double* re,
                         double* im)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(a && b && re && im && n_wv >= 0, "invalid argument");
    const size_t bytes = (size_t)n_wv * 12 * sizeof(double2);
    QDWF_REQUIRE(bytes <= stage_size(*inv), "dot_product: vector larger than a 5D full field");
    const double2* da = upload(*inv, 0, a, bytes);
    const double2* db = upload(*inv, 1, b, bytes);
    const cd d = dev_dot<double2>(*inv, da, db, (int64_t)n_wv * 12);
    *re = d.real();
    *im = d.imag();
  });
}

int qlat_dwf_qnorm(qlat_dwf_inverter* inv, const double* a, long n_wv, double* result)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(a && result && n_wv >= 0, "invalid argument");
    const size_t bytes = (size_t)n_wv * 12 * sizeof(double2);
    QDWF_REQUIRE(bytes <= stage_size(*inv), "qnorm: vector larger than a 5D full field");
    const double2* da = upload(*inv, 0, a, bytes);
    qdwf::launch_norm2<double2>(da, (int64_t)n_wv * 12, inv->part, inv->red_map_dev, inv->stream);
    check_launch();
    sync(*inv);
    *result = inv->red_map[0];
  });
}

int qlat_dwf_cg_with_herm_sym_2(qlat_dwf_inverter* inv, double* sol, const double* src, double stop_rsd,
                                long max_num_iter, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(sol && src && iter && max_num_iter >= 0, "invalid argument");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    const size_t bytes = half_aos_bytes(*inv);
    const double2* d_src = upload(*inv, 0, src, bytes);
    double2* b = W(*inv, W_IN_OP);
    double2* x = W(*inv, W_OUT_OP);
    qdwf::launch_pack5(d_src, b, inv->geo, inv->ls, 1, 0, inv->stream);
    const double2* d_sol = upload(*inv, 1, sol, bytes);
    qdwf::launch_pack5(d_sol, x, inv->geo, inv->ls, 1, 0, inv->stream);
    check_launch();
    *iter = cg_fp64(*inv, x, b, stop_rsd, max_num_iter, false);
    inv->stats.total_iter = *iter;
    double2* d_out = stage(*inv, 1, bytes);
    qdwf::launch_unpack5(x, d_out, inv->geo, inv->ls, 1, 0, inv->stream);
    check_launch();
    download(*inv, sol, d_out, bytes);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_invert_5d(qlat_dwf_inverter* inv, double* out, const double* in, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(out && in && iter, "invalid argument");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    const size_t bytes = full_aos_bytes(*inv);
    const double2* d_in = upload(*inv, 0, in, bytes);
    double2* d_out = stage(*inv, 1, bytes);
    *iter = invert_5d_dev(*inv, d_out, d_in);
    download(*inv, out, d_out, bytes);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

namespace {

// invert_with_cg_with_guess on full 5D AoS device fields: out holds the guess
long invert_5d_with_guess_dev(Inv& I, double2* out_aos, const double2* in_aos)
{
  double2* ie = W(I, W_IN_E);
  double2* io = W(I, W_IN_O);
  double2* ge = W(I, W_G_E);
  double2* go = W(I, W_G_O);
  qdwf::launch_pack5(in_aos, ie, I.geo, I.ls, 2, 1, I.stream);
  qdwf::launch_pack5(in_aos, io, I.geo, I.ls, 1, 1, I.stream);
  qdwf::launch_pack5(out_aos, ge, I.geo, I.ls, 2, 1, I.stream);
  qdwf::launch_pack5(out_aos, go, I.geo, I.ls, 1, 1, I.stream);
  check_launch();
  const long it = invert_halves_with_guess(I, ie, io, ge, go);
  qdwf::launch_unpack5(ge, out_aos, I.geo, I.ls, 2, 1, I.stream);
  qdwf::launch_unpack5(go, out_aos, I.geo, I.ls, 1, 1, I.stream);
  check_launch();
  return it;
}

}  // namespace

int qlat_dwf_invert_5d_with_guess(qlat_dwf_inverter* inv, double* out, const double* in, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(out && in && iter, "invalid argument");
    QDWF_REQUIRE((const double*)out != in, "out and in must be different fields (dslash.h:545)");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    const size_t bytes = full_aos_bytes(*inv);
    const double2* d_in = upload(*inv, 0, in, bytes);
    double2* d_out = (double2*)upload(*inv, 1, out, bytes);
    *iter = invert_5d_with_guess_dev(*inv, d_out, d_in);
    download(*inv, out, d_out, bytes);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_invert_5d_with_guess_device(qlat_dwf_inverter* inv, double* out, const double* in, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(out && in && iter, "invalid argument");
    QDWF_REQUIRE((const double*)out != in, "out and in must be different fields (dslash.h:545)");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    *iter = invert_5d_with_guess_dev(*inv, (double2*)out, (const double2*)in);
    sync(*inv);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_invert_4d(qlat_dwf_inverter* inv, double* out, const double* in, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(out && in && iter, "invalid argument");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    const size_t bytes = (size_t)inv->geo.vh * 2 * 12 * sizeof(double2);
    const double2* d_in = upload(*inv, 0, in, bytes);
    double2* d_out = stage(*inv, 1, bytes);
    *iter = invert_prop_dev(*inv, d_out, d_in, false);
    download(*inv, out, d_out, bytes);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_invert_prop(qlat_dwf_inverter* inv, double* sol, const double* src, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(sol && src && iter, "invalid argument");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    const size_t bytes = prop_bytes(*inv);
    const double2* d_in = upload(*inv, 0, src, bytes);
    double2* d_out = stage(*inv, 1, bytes);
    *iter = invert_prop_dev(*inv, d_out, d_in, true);
    download(*inv, sol, d_out, bytes);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

namespace {

// Kernel arguments of a source (validated as the reference does: a wall
// source needs lmom[3] = 0, qcd-prop.cpp:120; a point outside the lattice
// leaves the source zero, :70-72).  fu1 is a device pointer.
qdwf::SrcArgs src_args(const int n[4], const qlat_dwf_source& s, const double2* fu1)
{
  QDWF_REQUIRE(s.kind >= QLAT_DWF_SRC_POINT && s.kind <= QLAT_DWF_SRC_U1, "unknown source kind");
  qdwf::SrcArgs a{};
  for (int mu = 0; mu < 4; ++mu) {
    QDWF_REQUIRE(n[mu] > 0, "total_site must be positive");
    a.n[mu] = n[mu];
    a.xg[mu] = s.xg[mu];
    // lattice_mom_mult (field.h:14-17): 2 pi / total_site, with the
    // reference's double pi (qlat-utils core.h:41)
    constexpr double kPi = 3.141592653589793;
    a.mom[mu] = s.lmom[mu] * (2 * kPi / (double)n[mu]);
  }
  a.kind = s.kind;
  a.tslice = s.tslice;
  a.value_re = s.value[0];
  a.value_im = s.value[1];
  if (s.kind == QLAT_DWF_SRC_WALL) QDWF_REQUIRE(s.lmom[3] == 0.0, "wall source: lmom[3] must be 0");
  if (s.kind == QLAT_DWF_SRC_U1) {
    QDWF_REQUIRE(fu1, "U(1) source: null fu1");
    a.fu1 = fu1;
  }
  return a;
}

size_t prop_bytes_of(const int n[4]) { return (size_t)n[0] * n[1] * n[2] * n[3] * 144 * sizeof(double2); }

// Build the source into the device propagator `prop` on the inverter's
// stream; a host fu1 goes through staging buffer 1 first.
void make_source_dev(Inv& I, const qlat_dwf_source& s, double2* prop, bool host_fu1)
{
  const double2* fu1 = nullptr;
  if (s.kind == QLAT_DWF_SRC_U1) {
    QDWF_REQUIRE(s.fu1, "U(1) source: null fu1");
    fu1 = host_fu1 ? upload(I, 1, s.fu1, (size_t)I.geo.vh * 2 * sizeof(double2)) : (const double2*)s.fu1;
  }
  qdwf::launch_make_src(prop, src_args(I.geo.n, s, fu1), I.stream);
  check_launch();
}

}  // namespace

int qlat_dwf_make_source(const int total_site[4], const qlat_dwf_source* src, double* prop)
{
  return guarded([&] {
    QDWF_REQUIRE(total_site && src && prop, "invalid argument");
    const size_t bytes = prop_bytes_of(total_site);
    const size_t vbytes = bytes / 144;
    double2* d = (double2*)dev_alloc(bytes + (src->kind == QLAT_DWF_SRC_U1 ? vbytes : 0));
    try {
      double2* fu1 = nullptr;
      if (src->kind == QLAT_DWF_SRC_U1 && src->fu1) {
        fu1 = d + bytes / sizeof(double2);
        QDWF_HIP(hipMemcpy(fu1, src->fu1, vbytes, hipMemcpyHostToDevice));
      }
      qdwf::launch_make_src(d, src_args(total_site, *src, fu1), nullptr);
      check_launch();
      QDWF_HIP(hipMemcpy(prop, d, bytes, hipMemcpyDeviceToHost));
    } catch (...) {
      (void)hipFree(d);
      throw;
    }
    QDWF_HIP(hipFree(d));
  });
}

int qlat_dwf_make_source_device(const int total_site[4], const qlat_dwf_source* src, double* prop)
{
  return guarded([&] {
    QDWF_REQUIRE(total_site && src && prop, "invalid argument");
    qdwf::launch_make_src((double2*)prop, src_args(total_site, *src, (const double2*)src->fu1), nullptr);
    check_launch();
    QDWF_HIP(hipDeviceSynchronize());
  });
}

int qlat_dwf_source_propagator(qlat_dwf_inverter* inv, const qlat_dwf_source* src, double* sol, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(src && sol && iter, "invalid argument");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    const size_t bytes = prop_bytes(*inv);
    double2* d_src = stage(*inv, 0, bytes);
    make_source_dev(*inv, *src, d_src, true);  // fu1 staged in buffer 1 before it holds the solution
    double2* d_sol = stage(*inv, 1, bytes);
    *iter = invert_prop_dev(*inv, d_sol, d_src, true);
    download(*inv, sol, d_sol, bytes);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_source_propagator_device(qlat_dwf_inverter* inv, const qlat_dwf_source* src, double* sol,
                                      long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(src && sol && iter, "invalid argument");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    double2* d_src = stage(*inv, 0, prop_bytes(*inv));
    make_source_dev(*inv, *src, d_src, false);
    *iter = invert_prop_dev(*inv, (double2*)sol, d_src, true);
    sync(*inv);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

namespace {

// device memory freed at scope exit
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) : p(dev_alloc(bytes)) {}
  ~DevBuf()
  {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  double2* d() const { return (double2*)p; }
};

// A QED call's operands in device memory (null stream, synchronous copies).
struct QedCall {
  qdwf::QedArgs a{};
  int64_t vol = 0, n = 0;  // sites, complex entries of a 5D field
  std::unique_ptr<DevBuf> gf, wick;
  QedCall(const int total_site[4], int ls, const double* gf1, double mass, double m5, const double* wick_h)
  {
    QDWF_REQUIRE(total_site && gf1, "invalid argument");
    QDWF_REQUIRE(ls > 0, "ls must be positive (qed.cpp:297)");
    vol = 1;
    for (int mu = 0; mu < 4; ++mu) {
      QDWF_REQUIRE(total_site[mu] > 0, "total_site must be positive");
      a.n[mu] = total_site[mu];
      vol *= total_site[mu];
    }
    n = vol * ls * 4;
    a.ls = ls;
    a.mass = mass;
    a.m5 = m5;
    gf.reset(new DevBuf(vol * 4 * sizeof(double2)));
    QDWF_HIP(hipMemcpy(gf->p, gf1, vol * 4 * sizeof(double2), hipMemcpyHostToDevice));
    a.gf = gf->d();
    if (wick_h) {
      wick.reset(new DevBuf(total_site[3] * sizeof(double2)));
      QDWF_HIP(hipMemcpy(wick->p, wick_h, total_site[3] * sizeof(double2), hipMemcpyHostToDevice));
      a.wick = wick->d();
    }
  }
  void apply(double2* out, const double2* in, bool dagger) const
  {
    qdwf::QedArgs b = a;
    b.in = in;
    b.out = out;
    b.dagger = dagger ? 1 : 0;
    qdwf::launch_qed_m(b, nullptr);
    check_launch();
  }
};

double qed_norm2(const double2* x, int64_t n, double* part, double* red)
{
  qdwf::launch_norm2<double2>(x, n, part, red, nullptr);
  double v = 0.0;
  QDWF_HIP(hipMemcpy(&v, red, sizeof(double), hipMemcpyDeviceToHost));
  return v;
}

// cg_with_m_dwf_qed (qed.cpp:178-261): M_{!d} M_{d} x = b from the guess in x
long qed_cg(const QedCall& q, double2* x, const double2* b, bool dagger, double stop_rsd, long max_num_iter)
{
  QDWF_REQUIRE(stop_rsd >= 0.0 && max_num_iter >= 0, "invalid stop_rsd or max_num_iter");
  if (max_num_iter == 0) return 0;
  const int64_t n = q.n;
  const size_t fb = n * sizeof(double2);
  DevBuf r(fb), p(fb), ap(fb), tmp(fb), part(2 * qdwf::kMaxBlasGrid * sizeof(double)), red(2 * sizeof(double));
  auto A = [&](double2* out, const double2* in) {
    q.apply(tmp.d(), in, dagger);
    q.apply(out, tmp.d(), !dagger);
  };
  QDWF_HIP(hipMemcpy(r.p, b, fb, hipMemcpyDeviceToDevice));
  A(ap.d(), x);
  qdwf::launch_axpy<double2>(-1.0, ap.d(), r.d(), n, nullptr);
  check_launch();
  QDWF_HIP(hipMemcpy(p.p, r.p, fb, hipMemcpyDeviceToDevice));
  const double qnorm_in = qed_norm2(b, n, (double*)part.p, (double*)red.p);
  double qnorm_r = qed_norm2(r.d(), n, (double*)part.p, (double*)red.p);
  for (long iter = 1; iter <= max_num_iter; ++iter) {
    A(ap.d(), p.d());
    qdwf::launch_dot<double2>(p.d(), ap.d(), n, (double*)part.p, (double*)red.p, nullptr);
    double pap[2];
    QDWF_HIP(hipMemcpy(pap, red.p, 2 * sizeof(double), hipMemcpyDeviceToHost));
    const double alpha = qnorm_r / pap[0];
    qdwf::launch_axpy<double2>(alpha, p.d(), x, n, nullptr);
    qdwf::launch_axpy<double2>(-alpha, ap.d(), r.d(), n, nullptr);
    check_launch();
    const double new_qnorm_r = qed_norm2(r.d(), n, (double*)part.p, (double*)red.p);
    if (new_qnorm_r <= qnorm_in * stop_rsd * stop_rsd) return iter;
    qdwf::launch_xpay<double2>(r.d(), new_qnorm_r / qnorm_r, p.d(), n, nullptr);  // p = r + beta p
    check_launch();
    qnorm_r = new_qnorm_r;
  }
  return max_num_iter + 1;
}

}  // namespace

int qlat_dwf_free_invert(const int total_site[4], double mass, double m5, const double momtwist[4], double* sol,
                         const double* src)
{
  return guarded([&] {
    QDWF_REQUIRE(total_site && sol && src, "invalid argument");
    int64_t vol = 1;
    for (int mu = 0; mu < 4; ++mu) {
      QDWF_REQUIRE(total_site[mu] > 0, "total_site must be positive");
      vol *= total_site[mu];
    }
    const double tw0[4] = {0.0, 0.0, 0.0, 0.0};
    const double* tw = momtwist ? momtwist : tw0;
    const size_t bytes = (size_t)vol * 144 * sizeof(double2);
    DevBuf a(bytes), b(bytes);
    QDWF_HIP(hipMemcpy(a.p, src, bytes, hipMemcpyHostToDevice));
    // fft_complex_field(sol, true): sum_x f(x) exp(-i k x), direction by direction
    for (int mu = 0; mu < 4; ++mu) {
      qdwf::launch_dft_dir(mu % 2 ? a.d() : b.d(), mu % 2 ? b.d() : a.d(), total_site, mu, -1, nullptr);
      check_launch();
    }
    // free_mom_invert; the 1 / V of the backward transform applied here (linear)
    qdwf::launch_free_mom(a.d(), total_site, mass, m5, tw, 1.0 / (double)vol, nullptr);
    check_launch();
    for (int mu = 0; mu < 4; ++mu) {
      qdwf::launch_dft_dir(mu % 2 ? a.d() : b.d(), mu % 2 ? b.d() : a.d(), total_site, mu, +1, nullptr);
      check_launch();
    }
    QDWF_HIP(hipMemcpy(sol, a.p, bytes, hipMemcpyDeviceToHost));
  });
}

int qlat_dwf_rng_fill_uniform(const uint32_t midstate[8], uint64_t num_bytes, uint64_t index0, int64_t n_sites,
                              int n_per_site, double upper, double lower, double* out)
{
  return guarded([&] {
    QDWF_REQUIRE(midstate && out, "invalid argument");
    QDWF_REQUIRE(n_sites >= 0 && n_per_site >= 1, "need n_sites >= 0 and n_per_site >= 1");
    QDWF_REQUIRE(num_bytes % 64 == 0, "the RngState message is a whole number of 64-byte blocks");
    QDWF_REQUIRE(n_sites <= (int64_t)1 << 40 && n_per_site <= (1 << 20), "field too large");
    if (n_sites == 0) return;
    qdwf::RngFillArgs a{};
    for (int i = 0; i < 8; ++i) a.mid[i] = midstate[i];
    a.num_bytes = num_bytes;
    a.index0 = index0;
    a.n_sites = n_sites;
    a.n_per_site = n_per_site;
    a.scale = upper - lower;
    a.lower = lower;
    const size_t bytes = (size_t)n_sites * n_per_site * sizeof(double);
    DevBuf d(bytes);
    a.out = (double*)d.p;
    qdwf::launch_rng_fill_u(a, nullptr);
    check_launch();
    QDWF_HIP(hipMemcpy(out, d.p, bytes, hipMemcpyDeviceToHost));
  });
}

int qlat_dwf_rng_fill_gaussian(const uint32_t midstate[8], uint64_t num_bytes, uint64_t index0, int64_t n_sites,
                               int n_per_site, double center, double sigma, double* out)
{
  return guarded([&] {
    QDWF_REQUIRE(midstate && out, "invalid argument");
    QDWF_REQUIRE(n_sites >= 0 && n_per_site >= 1, "need n_sites >= 0 and n_per_site >= 1");
    QDWF_REQUIRE(num_bytes % 64 == 0, "the RngState message is a whole number of 64-byte blocks");
    QDWF_REQUIRE(n_sites <= (int64_t)1 << 40 && n_per_site <= (1 << 20), "field too large");
    if (n_sites == 0) return;
    qdwf::RngFillArgs a{};
    for (int i = 0; i < 8; ++i) a.mid[i] = midstate[i];
    a.num_bytes = num_bytes;
    a.index0 = index0;
    a.n_sites = n_sites;
    a.n_per_site = n_per_site;
    a.scale = sigma;
    a.lower = center;
    const size_t bytes = (size_t)n_sites * n_per_site * sizeof(double);
    DevBuf d(bytes);
    a.out = (double*)d.p;
    qdwf::launch_rng_fill_g(a, nullptr);
    check_launch();
    QDWF_HIP(hipMemcpy(out, d.p, bytes, hipMemcpyDeviceToHost));
  });
}

int qlat_dwf_multiply_m_dwf_qed(const int total_site[4], int ls, const double* gf1, double mass, double m5,
                                const double* wick, int is_dagger, double* out, const double* in)
{
  return guarded([&] {
    QDWF_REQUIRE(out && in, "null field pointer");
    const QedCall q(total_site, ls, gf1, mass, m5, wick);
    const size_t fb = q.n * sizeof(double2);
    DevBuf din(fb), dout(fb);
    QDWF_HIP(hipMemcpy(din.p, in, fb, hipMemcpyHostToDevice));
    q.apply(dout.d(), din.d(), is_dagger != 0);
    QDWF_HIP(hipMemcpy(out, dout.p, fb, hipMemcpyDeviceToHost));
  });
}

int qlat_dwf_cg_with_m_dwf_qed(const int total_site[4], int ls, const double* gf1, double mass, double m5,
                               const double* wick, int is_dagger, double stop_rsd, long max_num_iter, double* out,
                               const double* in, long* iter)
{
  return guarded([&] {
    QDWF_REQUIRE(out && in && iter, "invalid argument");
    const QedCall q(total_site, ls, gf1, mass, m5, wick);
    const size_t fb = q.n * sizeof(double2);
    DevBuf din(fb), dx(fb);
    QDWF_HIP(hipMemcpy(din.p, in, fb, hipMemcpyHostToDevice));
    QDWF_HIP(hipMemcpy(dx.p, out, fb, hipMemcpyHostToDevice));  // the initial guess
    *iter = qed_cg(q, dx.d(), din.d(), is_dagger != 0, stop_rsd, max_num_iter);
    QDWF_HIP(hipMemcpy(out, dx.p, fb, hipMemcpyDeviceToHost));
  });
}

int qlat_dwf_invert_dwf_qed(const int total_site[4], int ls, const double* gf1, double mass, double m5,
                            const double* wick, int is_dagger, double stop_rsd, long max_num_iter, double* out4d,
                            const double* in4d, long* iter)
{
  return guarded([&] {
    QDWF_REQUIRE(out4d && in4d && iter, "invalid argument");
    const QedCall q(total_site, ls, gf1, mass, m5, wick);
    const int64_t vol = q.vol;
    const size_t fb = q.n * sizeof(double2);
    // fermion_field_5d_from_4d_qed(src5d, in, ls, 0, ls - 1) on the host side of the copy
    std::vector<double2> h5((size_t)q.n, make_double2(0.0, 0.0));
    const double2* in = (const double2*)in4d;
    for (int64_t x = 0; x < vol; ++x) {
      h5[(x * ls + 0) * 4 + 0] = in[x * 4 + 0];
      h5[(x * ls + 0) * 4 + 1] = in[x * 4 + 1];
      h5[(x * ls + ls - 1) * 4 + 2] = in[x * 4 + 2];
      h5[(x * ls + ls - 1) * 4 + 3] = in[x * 4 + 3];
    }
    DevBuf s5(fb), b5(fb), x5(fb);
    QDWF_HIP(hipMemcpy(s5.p, h5.data(), fb, hipMemcpyHostToDevice));
    q.apply(b5.d(), s5.d(), is_dagger == 0);  // M_{!d} src (qed.cpp:141-142)
    QDWF_HIP(hipMemset(x5.p, 0, fb));
    *iter = qed_cg(q, x5.d(), b5.d(), is_dagger != 0, stop_rsd, max_num_iter);
    QDWF_HIP(hipMemcpy(h5.data(), x5.p, fb, hipMemcpyDeviceToHost));
    // fermion_field_4d_from_5d_qed(out, sol5d, ls, ls - 1, 0)
    double2* out = (double2*)out4d;
    for (int64_t x = 0; x < vol; ++x) {
      out[x * 4 + 0] = h5[(x * ls + ls - 1) * 4 + 0];
      out[x * 4 + 1] = h5[(x * ls + ls - 1) * 4 + 1];
      out[x * 4 + 2] = h5[(x * ls + 0) * 4 + 2];
      out[x * 4 + 3] = h5[(x * ls + 0) * 4 + 3];
    }
  });
}

int qlat_dwf_inverter_set_low_modes(qlat_dwf_inverter* inv, const int block_site[4], int n_basis, int n_vec,
                                    const double* eigen_values, const float* bases, const float* coefs)
{
  return guarded([&] {
    check_inv(inv);
    Inv& I = *inv;
    QDWF_REQUIRE(n_vec >= 0, "n_vec must be >= 0");
    // validate before the current low modes are released
    int64_t nblock = 1, bvol = 1;
    if (n_vec > 0) {
      QDWF_REQUIRE(block_site && eigen_values && bases && coefs, "invalid argument");
      QDWF_REQUIRE(n_basis >= 1 && n_basis <= 2048, "n_basis must be in [1, 2048]");
      for (int mu = 0; mu < 4; ++mu) {
        QDWF_REQUIRE(block_site[mu] > 0 && I.geo.n[mu] % block_site[mu] == 0, "block_site must divide total_site");
        nblock *= I.geo.n[mu] / block_site[mu];
        bvol *= block_site[mu];
      }
      QDWF_REQUIRE(block_site[0] % 2 == 0, "block_site[0] must be even (x pairs of opposite parity)");
      for (int i = 0; i < n_vec; ++i) QDWF_REQUIRE(eigen_values[i] != 0.0, "eigen values must be non-zero");
    }
    auto f = [](void*& p) {
      if (p) (void)hipFree(p);
      p = nullptr;
    };
    sync(I);
    f((void*&)I.lm_basis);
    f((void*&)I.lm_coef);
    f((void*&)I.lm_evals);
    for (auto*& p : I.lm_hmap) f((void*&)p);
    f((void*&)I.lm_part);
    f((void*&)I.lm_p);
    I.lm_nvec = 0;
    if (n_vec == 0) return;
    for (int mu = 0; mu < 4; ++mu) I.lm_bs[mu] = block_site[mu];
    I.lm_nblock = (int)nblock;
    I.lm_bve = (int)(bvol / 2);
    I.lm_nbasis = n_basis;
    const int64_t bsize = (int64_t)I.lm_bve * I.ls * 12;
    const size_t nb = (size_t)nblock * n_basis * bsize * sizeof(float2);
    const size_t nc = (size_t)nblock * n_vec * n_basis * sizeof(float2);
    I.lm_basis = (float2*)dev_alloc(nb);
    I.lm_coef = (float2*)dev_alloc(nc);
    I.lm_evals = (double*)dev_alloc(n_vec * sizeof(double));
    for (auto*& p : I.lm_hmap) p = (int*)dev_alloc((size_t)nblock * I.lm_bve * sizeof(int));
    I.lm_part = (double2*)dev_alloc((size_t)nblock * n_vec * sizeof(double2));
    I.lm_p = (double2*)dev_alloc(n_vec * sizeof(double2));
    QDWF_HIP(hipMemcpyAsync(I.lm_basis, bases, nb, hipMemcpyHostToDevice, I.stream));
    QDWF_HIP(hipMemcpyAsync(I.lm_coef, coefs, nc, hipMemcpyHostToDevice, I.stream));
    QDWF_HIP(hipMemcpyAsync(I.lm_evals, eigen_values, n_vec * sizeof(double), hipMemcpyHostToDevice, I.stream));
    I.lm_nvec = n_vec;
    for (int eo = 1; eo <= 2; ++eo) {
      qdwf::launch_defl_hmap(I.lm_hmap[eo - 1], defl_args(I, eo), I.stream);
      check_launch();
    }
    sync(I);
  });
}

int qlat_dwf_deflate(qlat_dwf_inverter* inv, double* out, const double* in, int eo)
{
  return guarded([&] {
    check_inv(inv);
    check_eo(eo);
    QDWF_REQUIRE(out && in, "null field pointer");
    Inv& I = *inv;
    const size_t bytes = half_aos_bytes(I);
    const double2* d_in = upload(I, 0, in, bytes);
    double2* f = W(I, W_TMP);
    qdwf::launch_pack5(d_in, f, I.geo, I.ls, eo, 0, I.stream);
    check_launch();
    deflate_dev(I, f, f, eo);
    double2* d_out = stage(I, 1, bytes);
    qdwf::launch_unpack5(f, d_out, I.geo, I.ls, eo, 0, I.stream);
    check_launch();
    download(I, out, d_out, bytes);
  });
}

namespace {

// The power iteration of find_max_eigen_value_hermop_sym2 (dslash.cpp:1349-1369)
// on the odd half field ff (device layout): normalise, then repeatedly apply
// Mpc^+ Mpc and renormalise; returns the last norm ratio.
double power_iteration(Inv& I, double2* ff, long max_iter)
{
  const int64_t n = n64(I);
  const double q = dev_norm2(I, ff);
  QDWF_REQUIRE(q > 0, "find_max_eigen_value_hermop_sym2: the start vector is zero");
  double ratio = 1.0;
  double2* tmp = W(I, W_TMP);  // zero vector: ff = 0 + scale * ff
  zero64(I, tmp);
  qdwf::launch_xpay<double2>(tmp, 1.0 / std::sqrt(q), ff, n, I.stream);
  check_launch();
  for (long i = 0; i < max_iter; ++i) {
    op_hermop<double>(I, ff, ff, 1);
    ratio = std::sqrt(dev_norm2(I, ff));
    qdwf::launch_xpay<double2>(tmp, 1.0 / ratio, ff, n, I.stream);
    check_launch();
  }
  sync(I);
  return ratio;
}

}  // namespace

int qlat_dwf_find_max_eigen_value_hermop_sym2(qlat_dwf_inverter* inv, const double* start, long max_iter,
                                              double* result)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(start && result && max_iter >= 0, "invalid argument");
    const size_t bytes = half_aos_bytes(*inv);
    const double2* d_in = upload(*inv, 0, start, bytes);
    double2* ff = W(*inv, W_IN_OP);
    qdwf::launch_pack5(d_in, ff, inv->geo, inv->ls, 1, 0, inv->stream);
    check_launch();
    *result = power_iteration(*inv, ff, max_iter);
  });
}

int qlat_dwf_find_max_eigen_value_hermop_sym2_rng(qlat_dwf_inverter* inv, const uint32_t midstate[8],
                                                  uint64_t num_bytes, uint64_t index0, long max_iter,
                                                  double* result)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(midstate && result && max_iter >= 0, "invalid argument");
    QDWF_REQUIRE(num_bytes % 64 == 0, "the RngState message is a whole number of 64-byte blocks");
    Inv& I = *inv;
    // set_u_rand(ff, rs) on the odd half field (dslash.cpp:1355-1357): site x
    // draws Ls*24 reals in [-1, 1) from rs.newtype(global index of x); the
    // draws of a full 5D field in reference layout are made and its odd
    // half packed (each site's stream is independent of the others)
    qdwf::RngFillArgs a{};
    for (int i = 0; i < 8; ++i) a.mid[i] = midstate[i];
    a.num_bytes = num_bytes;
    a.index0 = index0;
    a.n_sites = 2 * I.geo.vh;
    a.n_per_site = I.ls * 24;
    a.scale = 2.0;
    a.lower = -1.0;
    double2* full = stage(I, 0, full_aos_bytes(I));
    a.out = (double*)full;
    qdwf::launch_rng_fill_u(a, I.stream);
    check_launch();
    double2* ff = W(I, W_IN_OP);
    qdwf::launch_pack5(full, ff, I.geo, I.ls, 1, 1, I.stream);
    check_launch();
    *result = power_iteration(I, ff, max_iter);
  });
}

int qlat_dwf_multiply_hermop_sym2_fp32(qlat_dwf_inverter* inv, double* out, const double* in, int eo)
{
  return guarded([&] {
    check_inv(inv);
    check_eo(eo);
    QDWF_REQUIRE(out && in, "null field pointer");
    Inv& I = *inv;
    const size_t bytes = half_aos_bytes(I);
    const double2* d_in = upload(I, 0, in, bytes);
    double2* a = W(I, W_IN_E);
    double2* b = W(I, W_OUT_E);
    qdwf::launch_pack5(d_in, a, I.geo, I.ls, eo, 0, I.stream);
    if (use_three_phase(I)) {
      // the inner operator of the three-phase solver: in -> r -> p (storage
      // rounding), t = K p, ap = A p, ap -> FP32 -> FP64
      ThreePhase tp(I, 1, storage_of(I));
      tp.set_r(a, 0, 1.0);
      tp.zero_p(1);
      tp.zero_x(0, 1);
      const double zero[qdwf::kMaxRhs] = {};
      tp.step({0}, zero, zero, false);
      tp.hermop_dot(tp.colset({0}), eo);
      tp.axpy_to_x(1.0, tp.AP, 0);
      zero64(I, b);
      tp.add_x(b, 0, 1.0);
    } else {
      qdwf::launch_d2f(a, V(I, V_R), I.ls, I.geo.stride, I.stream);
      check_launch();
      op_hermop<float>(I, V(I, V_AP), V(I, V_R), eo);
      qdwf::launch_f2d(V(I, V_AP), b, I.ls, I.geo.stride, 0, I.stream);
    }
    double2* d_out = stage(I, 1, bytes);
    qdwf::launch_unpack5(b, d_out, I.geo, I.ls, eo, 0, I.stream);
    check_launch();
    download(I, out, d_out, bytes);
  });
}

int qlat_dwf_multiply_hermop_sym2_device(qlat_dwf_inverter* inv, double* out, const double* in, int eo)
{
  return guarded([&] {
    check_inv(inv);
    check_eo(eo);
    QDWF_REQUIRE(out && in, "null field pointer");
    double2* a = W(*inv, W_IN_E);
    double2* b = W(*inv, W_OUT_E);
    qdwf::launch_pack5((const double2*)in, a, inv->geo, inv->ls, eo, 0, inv->stream);
    check_launch();
    op_hermop<double>(*inv, b, a, eo);
    qdwf::launch_unpack5(b, (double2*)out, inv->geo, inv->ls, eo, 0, inv->stream);
    check_launch();
    sync(*inv);
  });
}

namespace {

// While alive: the inverter's work fields `ids` are consecutive slices of
// this device's scratch array, and any other buffer request or device
// allocation throws.
struct NoAllocScope {
  Inv& I;
  NoAllocScope(Inv& inv, const int* ids, int n, size_t field_bytes) : I(inv)
  {
    unsigned char* base = scratch_ptr(I.device);
    for (auto*& p : I.w64_noalloc) p = nullptr;
    for (int k = 0; k < n; ++k) I.w64_noalloc[ids[k]] = (double2*)(base + (size_t)k * field_bytes);
    I.noalloc = true;
    ++g_noalloc_depth;
  }
  ~NoAllocScope()
  {
    I.noalloc = false;
    for (auto*& p : I.w64_noalloc) p = nullptr;
    --g_noalloc_depth;
  }
};

}  // namespace

int qlat_dwf_multiply_hermop_sym2_device_noalloc(qlat_dwf_inverter* inv, double* out, const double* in, int eo)
{
  return guarded([&] {
    check_inv(inv);
    check_eo(eo);
    QDWF_REQUIRE(out && in, "null field pointer");
    Inv& I = *inv;
    // the device-layout input and output and the FP64 operator's temporaries
    // (op_hermop -> op_mpc / op_mpcdag use W_OP_T1, W_OP_T2 and W_OP_Y)
    const int ids[5] = {W_IN_E, W_OUT_E, W_OP_T1, W_OP_T2, W_OP_Y};
    const size_t fb = (field64_bytes(I) + 255) / 256 * 256;
    if (5 * fb > kNoAllocScratchBytes)
      throw ArgError("qlat_dwf_multiply_hermop_sym2_device_noalloc: this lattice needs " + std::to_string(5 * fb) +
                     " bytes of scratch (five FP64 half fields); the maximum is " +
                     std::to_string(kNoAllocScratchBytes) +
                     " bytes -- use qlat_dwf_multiply_hermop_sym2_device");
    NoAllocScope scope(I, ids, 5, fb);
    // the scratch keeps the previous call's contents: clear it, so that no
    // field (pad entries of the device layout included) carries values left
    // by an earlier call
    QDWF_HIP(hipMemsetAsync(scratch_ptr(I.device), 0, 5 * fb, I.stream));
    double2* a = W(I, W_IN_E);
    double2* b = W(I, W_OUT_E);
    qdwf::launch_pack5((const double2*)in, a, I.geo, I.ls, eo, 0, I.stream);
    check_launch();
    op_hermop<double>(I, b, a, eo);
    qdwf::launch_unpack5(b, (double2*)out, I.geo, I.ls, eo, 0, I.stream);
    check_launch();
    sync(I);
  });
}

int qlat_dwf_invert_5d_device(qlat_dwf_inverter* inv, double* out, const double* in, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(out && in && iter, "invalid argument");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    *iter = invert_5d_dev(*inv, (double2*)out, (const double2*)in);
    sync(*inv);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_invert_5d_batch_device(qlat_dwf_inverter* const* invs, int n, double* const* sol,
                                    const double* const* src, long* iters)
{
  return guarded([&] {
    check_batch(invs, n);
    QDWF_REQUIRE(sol && src && iters, "invalid argument");
    for (int j = 0; j < n; ++j) QDWF_REQUIRE(sol[j] && src[j], "null field pointer");
    const auto t0 = std::chrono::steady_clock::now();
    invert_5d_batch_dev(invs, n, (double2* const*)sol, (const double2* const*)src, iters);
    sync(*invs[0]);
    invs[0]->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_invert_5d_batch(qlat_dwf_inverter* const* invs, int n, double* const* sol, const double* const* src,
                             long* iters)
{
  return guarded([&] {
    check_batch(invs, n);
    QDWF_REQUIRE(sol && src && iters, "invalid argument");
    for (int j = 0; j < n; ++j) QDWF_REQUIRE(sol[j] && src[j], "null field pointer");
    Inv& I = *invs[0];
    const size_t bytes = full_aos_bytes(I);
    // one device copy of every source and solution for the call
    double2* buf = (double2*)dev_alloc(2 * bytes * n);
    try {
      std::vector<double2*> dsol(n);
      std::vector<const double2*> dsrc(n);
      for (int j = 0; j < n; ++j) {
        double2* d = buf + (size_t)2 * j * (bytes / sizeof(double2));
        QDWF_HIP(hipMemcpyAsync(d, src[j], bytes, hipMemcpyHostToDevice, I.stream));
        dsrc[j] = d;
        dsol[j] = d + bytes / sizeof(double2);
      }
      const auto t0 = std::chrono::steady_clock::now();
      invert_5d_batch_dev(invs, n, dsol.data(), dsrc.data(), iters);
      for (int j = 0; j < n; ++j)
        QDWF_HIP(hipMemcpyAsync(sol[j], dsol[j], bytes, hipMemcpyDeviceToHost, I.stream));
      sync(I);
      I.stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } catch (...) {
      (void)hipStreamSynchronize(I.stream);
      (void)hipFree(buf);
      throw;
    }
    QDWF_HIP(hipFree(buf));
  });
}

int qlat_dwf_invert_prop_device(qlat_dwf_inverter* inv, double* sol, const double* src, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(sol && src && iter, "invalid argument");
    reset_stats(*inv);
    const auto t0 = std::chrono::steady_clock::now();
    *iter = invert_prop_dev(*inv, (double2*)sol, (const double2*)src, true);
    sync(*inv);
    inv->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_inverter_set_gauge(qlat_dwf_inverter* inv, const double* gauge, int on_device)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(gauge, "null gauge pointer");
    Inv& I = *inv;
    if (on_device) {
      pack_gauge_dev(I, (const double2*)gauge);
    } else {
      const size_t gbytes = (size_t)I.geo.vh * 2 * 36 * sizeof(double2);
      double2* g = (double2*)dev_alloc(gbytes);
      QDWF_HIP(hipMemcpyAsync(g, gauge, gbytes, hipMemcpyHostToDevice, I.stream));
      pack_gauge_dev(I, g);
      sync(I);
      QDWF_HIP(hipFree(g));
    }
    sync(I);
  });
}

int qlat_dwf_fermion_force(qlat_dwf_inverter* inv, const double* phi, double stop_rsd, double* force, double* grad,
                           double* action, double* action_lo, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(phi && force && action && iter && stop_rsd > 0, "invalid argument");
    Inv& I = *inv;
    reset_stats(I);
    const auto t0 = std::chrono::steady_clock::now();
    const size_t hbytes = half_aos_bytes(I);
    const size_t gbytes = (size_t)I.geo.vh * 2 * 36 * sizeof(double2);
    const double2* d_phi = upload(I, 0, phi, hbytes);
    double2* d_f = (double2*)dev_alloc(gbytes);
    double2* d_g = grad ? (double2*)dev_alloc(gbytes) : nullptr;
    try {
      *iter = fermion_force_dev(I, d_phi, stop_rsd, d_f, d_g, action, action_lo);
      QDWF_HIP(hipMemcpyAsync(force, d_f, gbytes, hipMemcpyDeviceToHost, I.stream));
      if (grad) QDWF_HIP(hipMemcpyAsync(grad, d_g, gbytes, hipMemcpyDeviceToHost, I.stream));
      sync(I);
    } catch (...) {
      (void)hipFree(d_f);
      if (d_g) (void)hipFree(d_g);
      throw;
    }
    QDWF_HIP(hipFree(d_f));
    if (d_g) QDWF_HIP(hipFree(d_g));
    I.stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_fermion_force_device(qlat_dwf_inverter* inv, const double* phi, double stop_rsd, double* force,
                                  double* grad, double* action, double* action_lo, long* iter)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(phi && force && action && iter && stop_rsd > 0, "invalid argument");
    Inv& I = *inv;
    reset_stats(I);
    const auto t0 = std::chrono::steady_clock::now();
    *iter = fermion_force_dev(I, (const double2*)phi, stop_rsd, (double2*)force, (double2*)grad, action, action_lo);
    sync(I);
    I.stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

int qlat_dwf_bench_kernel(qlat_dwf_inverter* inv, int which, int precision, int iters, double* avg_ms)
{
  return guarded([&] {
    check_inv(inv);
    QDWF_REQUIRE(avg_ms && iters > 0, "invalid argument");
    QDWF_REQUIRE(which >= 0 && which <= 7, "which must be 0..7");
    QDWF_REQUIRE(which >= 5 ? (precision == 32 || precision == 16) : (precision == 64 || precision == 32),
                 "precision must be 64 or 32 (which 0..4), 32 or 16 (which 5..7)");
    Inv& I = *inv;
    std::unique_ptr<ThreePhase> tp;
    std::vector<int> cols;
    if (which >= 5) {
      // the tile the solver uses for a 12-column propagator
      const int storage = precision == 16 ? I.fp16_mode : 0;
      const TileChoice tc = tile_for(I, storage, 12);
      const int tile = std::max(1, std::min(tc.width, qdwf::kMaxRhs));
      I.stats.rhs_tile = tile;
      I.stats.sites_per_block = tc.sb;
      I.stats.hop_arith_bits = hop_arith_bits(I, storage);
      tp.reset(new ThreePhase(I, tile, storage, tc.sb));
      for (int c = 0; c < tile; ++c) cols.push_back(c);
      // finite, non-trivial input values
      const size_t lb = (size_t)tp->fh * tp->slot_bytes() * tile;
      for (void* f : {tp->P, tp->T, tp->Y, tp->B}) QDWF_HIP(hipMemsetAsync(f, 0x3f, lb, I.stream));
      QDWF_HIP(hipMemsetAsync(tp->R, 0x3f, (size_t)tp->fr * tp->rslot_bytes() * tile, I.stream));
      tp->zero_x(0, tile);
    }
    // nonzero alpha: the step then updates x as well (it skips x for alpha 0)
    const std::vector<double> al(qdwf::kMaxRhs, 1e-3), be(qdwf::kMaxRhs, 0.5);
    auto run = [&]() {
      if (which == 5) {
        tp->hopm5(tp->B, tp->T, 2, false, MAT_K, 1.f, nullptr, 0.f, nullptr, tp->colset(cols));
      } else if (which == 6) {
        tp->step(cols, al.data(), be.data(), true, tp->merged() ? al.data() : nullptr);
      } else if (which == 7) {
        tp->hermop_dot(tp->colset(cols), 1);
      } else if (precision == 64) {
        double2* a = W(I, W_IN_E);
        double2* b = W(I, W_OUT_E);
        if (which == 0) hop<double>(I, b, a, 1, false);
        else if (which == 1) hop<double>(I, b, a, 1, true);
        else if (which == 2) m5<double>(I, b, a, MAT_MEE_INV);
        else if (which == 3) op_hermop<double>(I, b, a, 1);
        else {
          qdwf::launch_copy((const float4*)a, (float4*)b, 2 * n32(I), I.stream);
          check_launch();
        }
      } else {
        float4* a = V(I, V_R);
        float4* b = V(I, V_AP);
        if (which == 0) hop<float>(I, b, a, 1, false);
        else if (which == 1) hop<float>(I, b, a, 1, true);
        else if (which == 2) m5<float>(I, b, a, MAT_MEE_INV);
        else if (which == 3) op_hermop<float>(I, b, a, 1);
        else {
          qdwf::launch_copy(a, b, n32(I), I.stream);
          check_launch();
        }
      }
    };
    // finite, non-trivial input values
    if (which < 5) {
      if (precision == 64) QDWF_HIP(hipMemsetAsync(W(I, W_IN_E), 0x3f, field64_bytes(I), I.stream));
      else QDWF_HIP(hipMemsetAsync(V(I, V_R), 0x3f, field32_bytes(I), I.stream));
    }
    run();
    hipEvent_t e0, e1;
    QDWF_HIP(hipEventCreate(&e0));
    QDWF_HIP(hipEventCreate(&e1));
    QDWF_HIP(hipEventRecord(e0, I.stream));
    for (int i = 0; i < iters; ++i) run();
    QDWF_HIP(hipEventRecord(e1, I.stream));
    QDWF_HIP(hipEventSynchronize(e1));
    float ms = 0;
    QDWF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / iters;
  });
}

}  // extern "C"
