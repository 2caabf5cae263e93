// gfx950 kernels of the domain-wall library: Wilson hopping term, dense
// fifth-dimension (s) matrices, BLAS-1 with block reductions, and layout
// conversion between the reference's array-of-structures complex128 fields
// and the device structure-of-arrays layout (dwf_common.h).
#include "dwf_internal.h"

namespace qdwf {

// ================================================================== hop
// out(h,s) = alpha * (H in)(h,s) + beta_s * x(h,s)
// (H psi)(x) = -1/2 sum_mu [ U_mu(x)(1 -+ g_mu) psi(x+mu)
//                          + U_mu(x-mu)^+ (1 +- g_mu) psi(x-mu) ]
// upper signs: DAG=false (multiply_wilson_d_e_o_no_comm, dslash.cpp:859-908),
// lower signs: DAG=true (multiply_wilson_ddag_e_o_no_comm, dslash.cpp:910-959).
// One thread = one output half-site x SPT consecutive s values; each link is
// loaded once per thread and reused across its s values.
template <class T, int SPT, bool DAG, bool XPAY>
__global__ void __launch_bounds__(kHopBlock) k_hop(HopArgs<T> a)
{
  typedef typename SpinorIO<T>::slot_t slot_t;
  const int64_t nblk = (int64_t)gridDim.x;
  const int64_t lb = xcd_logical_block(blockIdx.x, nblk);
  const int nchunk = a.nchunk;
  const int64_t hblk = lb / nchunk;
  const int chunk = (int)(lb - hblk * nchunk);
  const int64_t h = hblk * kHopBlock + threadIdx.x;
  if (h >= a.geo.vh) return;
  const int s0 = chunk * SPT;
  int x[4];
  half_coord(a.geo, h, a.eo_out, x);
  int64_t nb[8];
  neighbours(a.geo, x, nb);
  const int pslot = (a.eo_out == 2) ? 0 : 1;
  const slot_t* __restrict__ in = (const slot_t*)a.in;
  const typename LinkIO<T>::elem_t* __restrict__ links = (const typename LinkIO<T>::elem_t*)a.links;

  cplx<T> acc[SPT][12];
#pragma unroll
  for (int j = 0; j < SPT; ++j)
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[j][k] = cmk<T>(0, 0);

  // forward: sign sg = -1 (D) / +1 (D^+); backward: opposite.
  constexpr int SF = DAG ? 1 : -1;
  constexpr int SB = -SF;
#define QDWF_HOP_DIR(D, MU, SG)                                          \
  {                                                                      \
    cplx<T> W[9];                                                        \
    LinkIO<T>::load(links, a.geo.stride, pslot, (D), h, W);              \
    _Pragma("unroll") for (int j = 0; j < SPT; ++j)                      \
    {                                                                    \
      if (SPT == 1 || s0 + j < a.ls) {                                   \
        cplx<T> psi[12];                                                 \
        SpinorIO<T>::load(in, a.ls, s0 + j, nb[D], psi);                 \
        hop_term<MU, SG>(W, psi, acc[j]);                                \
      }                                                                  \
    }                                                                    \
  }
  QDWF_HOP_DIR(0, 0, SF)
  QDWF_HOP_DIR(1, 1, SF)
  QDWF_HOP_DIR(2, 2, SF)
  QDWF_HOP_DIR(3, 3, SF)
  QDWF_HOP_DIR(4, 0, SB)
  QDWF_HOP_DIR(5, 1, SB)
  QDWF_HOP_DIR(6, 2, SB)
  QDWF_HOP_DIR(7, 3, SB)
#undef QDWF_HOP_DIR

  slot_t* __restrict__ out = (slot_t*)a.out;
  const T scale = (T)(-0.5) * a.alpha;
#pragma unroll
  for (int j = 0; j < SPT; ++j) {
    const int s = s0 + j;
    if (SPT > 1 && s >= a.ls) break;
    cplx<T> v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = cmk<T>(scale * acc[j][k].re, scale * acc[j][k].im);
    if (XPAY) {
      cplx<T> xv[12];
      SpinorIO<T>::load((const slot_t*)a.x, a.ls, s, h, xv);
      const cplx<T> b = a.beta[s];
#pragma unroll
      for (int k = 0; k < 12; ++k) v[k] = cfma(v[k], b, xv[k]);
    }
    SpinorIO<T>::store(out, a.ls, s, h, v);
  }
}

// Variant with the block's links staged in LDS: a block covers 64
// consecutive output half-sites; its NW waves load the 64 x 8 links once,
// then wave w computes s = w, w+NW, ... for all 64 sites, reading the links
// from LDS.  One link load from HBM per site per launch, independent of Ls.
template <class ST, int NW, bool DAG, bool XPAY, bool NT>
__global__ void __launch_bounds__(64 * NW) k_hop_lds(HopArgs<typename SpinorIO<ST>::real> a)
{
  typedef typename SpinorIO<ST>::real T;
  typedef typename SpinorIO<ST>::slot_t slot_t;
  typedef typename LinkIO<T>::elem_t link_t;
  __shared__ link_t sl[72][64];
  const int64_t lb = xcd_logical_block(blockIdx.x, (int64_t)gridDim.x);
  const int64_t h0 = lb * 64;
  const int pslot = (a.eo_out == 2) ? 0 : 1;
  // the block's 72 x 64 link elements are contiguous (pads are zero)
  const link_t* __restrict__ links = (const link_t*)a.links + lidx(a.geo.stride, pslot, 0, h0);
  for (int i = threadIdx.x; i < 72 * 64; i += 64 * NW) sl[i >> 6][i & 63] = NT ? load_nt(&links[i]) : links[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t h = h0 + lane;
  if (h >= a.geo.vh) return;
  int x[4];
  half_coord(a.geo, h, a.eo_out, x);
  int64_t nb[8];
  neighbours(a.geo, x, nb);
  constexpr int SF = DAG ? 1 : -1;
  constexpr int SB = -SF;
  const T scale = (T)(-0.5) * a.alpha;
  const int nrs = a.cols.n * a.ls;
  for (int rs = w; rs < nrs; rs += NW) {
    const int ci = rs / a.ls, s = rs - ci * a.ls;
    const int64_t coff = (int64_t)pick(a.cols.rhs, ci) * a.cols.fstride;
    const slot_t* __restrict__ in = (const slot_t*)a.in + coff;
    slot_t* __restrict__ out = (slot_t*)a.out + coff;
    cplx<T> acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = cmk<T>(0, 0);
#define QDWF_HOP_LDS_DIR(D, MU, SG)                                \
  {                                                                \
    cplx<T> W[9];                                                  \
    _Pragma("unroll") for (int e = 0; e < 9; ++e)                  \
    {                                                              \
      const link_t q = sl[(D) * 9 + e][lane];                      \
      W[e] = cmk<T>(q.x, q.y);                                     \
    }                                                              \
    cplx<T> psi[12];                                               \
    SpinorIO<ST>::load(in, a.ls, s, nb[D], psi);                   \
    hop_term<MU, SG>(W, psi, acc);                                 \
  }
    QDWF_HOP_LDS_DIR(0, 0, SF)
    QDWF_HOP_LDS_DIR(1, 1, SF)
    QDWF_HOP_LDS_DIR(2, 2, SF)
    QDWF_HOP_LDS_DIR(3, 3, SF)
    QDWF_HOP_LDS_DIR(4, 0, SB)
    QDWF_HOP_LDS_DIR(5, 1, SB)
    QDWF_HOP_LDS_DIR(6, 2, SB)
    QDWF_HOP_LDS_DIR(7, 3, SB)
#undef QDWF_HOP_LDS_DIR
    cplx<T> v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = cmk<T>(scale * acc[k].re, scale * acc[k].im);
    if (XPAY) {
      cplx<T> xv[12];
      SpinorIO<ST>::load((const slot_t*)a.x + coff, a.ls, s, h, xv);
      const cplx<T> b = a.beta[s];
#pragma unroll
      for (int k = 0; k < 12; ++k) v[k] = cfma(v[k], b, xv[k]);
    }
    if (NT) SpinorIO<ST>::store_stream(out, a.ls, s, h, v);
    else SpinorIO<ST>::store(out, a.ls, s, h, v);
  }
}

template <class ST, int NW, bool NT>
static void launch_hop_lds(const HopArgs<typename SpinorIO<ST>::real>& a, bool dag, hipStream_t st)
{
  const dim3 grid((unsigned)((a.geo.vh + 63) / 64));
  const dim3 blk(64 * NW);
  const bool xpay = a.x != nullptr;
  if (!dag && !xpay) hipLaunchKernelGGL((k_hop_lds<ST, NW, false, false, NT>), grid, blk, 0, st, a);
  if (!dag && xpay) hipLaunchKernelGGL((k_hop_lds<ST, NW, false, true, NT>), grid, blk, 0, st, a);
  if (dag && !xpay) hipLaunchKernelGGL((k_hop_lds<ST, NW, true, false, NT>), grid, blk, 0, st, a);
  if (dag && xpay) hipLaunchKernelGGL((k_hop_lds<ST, NW, true, true, NT>), grid, blk, 0, st, a);
}

template <class T, int SPT>
static void launch_hop_spt(const HopArgs<T>& a0, bool dag, hipStream_t st)
{
  typedef typename SpinorIO<T>::slot_t slot_t;
  if (a0.cols.n != 1 || a0.cols.rhs[0] != 0) {  // one launch per column
    for (int i = 0; i < a0.cols.n; ++i) {
      HopArgs<T> b = a0;
      const int64_t off = (int64_t)a0.cols.rhs[i] * a0.cols.fstride;
      b.in = (const slot_t*)a0.in + off;
      b.out = (slot_t*)a0.out + off;
      if (a0.x) b.x = (const slot_t*)a0.x + off;
      b.cols.n = 1;
      b.cols.rhs[0] = 0;
      launch_hop_spt<T, SPT>(b, dag, st);
    }
    return;
  }
  HopArgs<T> a = a0;
  a.nchunk = (a.ls + SPT - 1) / SPT;
  const int64_t hblocks = (a.geo.vh + kHopBlock - 1) / kHopBlock;
  const dim3 grid((unsigned)(hblocks * a.nchunk));
  const bool xpay = a.x != nullptr;
  if (!dag && !xpay) hipLaunchKernelGGL((k_hop<T, SPT, false, false>), grid, dim3(kHopBlock), 0, st, a);
  if (!dag && xpay) hipLaunchKernelGGL((k_hop<T, SPT, false, true>), grid, dim3(kHopBlock), 0, st, a);
  if (dag && !xpay) hipLaunchKernelGGL((k_hop<T, SPT, true, false>), grid, dim3(kHopBlock), 0, st, a);
  if (dag && xpay) hipLaunchKernelGGL((k_hop<T, SPT, true, true>), grid, dim3(kHopBlock), 0, st, a);
}

// spt > 0: one thread per site and `spt` s values (k_hop<.., spt>);
// spt = -2 / -4: LDS-staged links, 2 or 4 waves per block (k_hop_lds);
// spt = -102 / -104: the same with streaming output stores and link loads.
template <class T>
void launch_hop(const HopArgs<T>& a, bool dag, int spt, hipStream_t st)
{
  switch (spt) {
    case -2: launch_hop_lds<T, 2, false>(a, dag, st); break;
    case -4: launch_hop_lds<T, 4, false>(a, dag, st); break;
    case -102: launch_hop_lds<T, 2, true>(a, dag, st); break;
    case -104: launch_hop_lds<T, 4, true>(a, dag, st); break;
    case 1: launch_hop_spt<T, 1>(a, dag, st); break;
    case 2: launch_hop_spt<T, 2>(a, dag, st); break;
    case 3: launch_hop_spt<T, 3>(a, dag, st); break;
    case 4: launch_hop_spt<T, 4>(a, dag, st); break;
    case 6: launch_hop_spt<T, 6>(a, dag, st); break;
    default: launch_hop_spt<T, 2>(a, dag, st); break;
  }
}

template void launch_hop<double>(const HopArgs<double>&, bool, int, hipStream_t);
template void launch_hop<float>(const HopArgs<float>&, bool, int, hipStream_t);

void launch_hop_h16(const HopArgs<float>& a, bool dag, hipStream_t st) { launch_hop_lds<Half16, 4, false>(a, dag, st); }

// ============================================================ hop + M5
// Fused out = alpha * Mat (H in) + beta * x for FP32 fields, where Mat is a
// dense Ls x Ls per-chirality s-matrix (the K = P M_ee^-1 of the Schur
// complement, or its adjoint).  Block = SB output half-sites x Ls:
// phase 1, thread (site, s) computes (H in)(site, s) with the block's links
// staged in LDS and parks it in LDS; phase 2, thread (site, slot) applies Mat
// along s, adds beta*x, writes the result once and, with DOT, accumulates
// Re <dotp, out> for the CG.  SB = 32 keeps the LDS at 55 KB (Ls = 12) so
// two blocks share a CU and one block's s-matrix phase overlaps the other's
// memory-bound hop phase.
template <int SB, int LSMAX, bool DAG, bool XPAY, bool REAL, bool DOT>
__global__ void __launch_bounds__(SB * LSMAX) __attribute__((amdgpu_waves_per_eu(3)))
k_hop_m5(HopM5Args a)
{
  __shared__ float2 sl[72][SB];
  __shared__ float4 hr[LSMAX][6][SB];
  // logical block -> (site block, column), column fastest: the blocks that
  // share a site block (and its links) run back to back on one XCD
  const int64_t lbc = xcd_logical_block(blockIdx.x, (int64_t)gridDim.x);
  const int ncol = a.cols.n;
  const int64_t lb = lbc / ncol;
  const int ci = (int)(lbc - lb * ncol);
  const int64_t nblk = (int64_t)gridDim.x / ncol;
  const int64_t h0 = lb * SB;
  const int pslot = (a.eo_out == 2) ? 0 : 1;
  // the link block holding these SB sites (64-site link blocks)
  const float2* __restrict__ links = (const float2*)a.links + lidx(a.geo.stride, pslot, 0, h0);
  for (int i = threadIdx.x; i < 72 * SB; i += blockDim.x) sl[i / SB][i % SB] = links[(i / SB) * 64 + (i % SB)];
  __syncthreads();
  const int ls = a.ls;
  const int site = threadIdx.x % SB, row = threadIdx.x / SB;
  const int64_t h = h0 + site;
  const bool active = h < a.geo.vh;
  const int64_t coff = (int64_t)pick(a.cols.rhs, ci) * a.cols.fstride;
  if (active && row < ls) {
    // neighbours recomputed per column: cheaper than keeping them live
    int x[4];
    half_coord(a.geo, h, a.eo_out, x);
    int64_t nb[8];
    neighbours(a.geo, x, nb);
    const float4* __restrict__ in = (const float4*)a.in + coff;
    constexpr int SF = DAG ? 1 : -1;
    constexpr int SB_ = -SF;
    cplx<float> acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = cmk<float>(0, 0);
#define QDWF_HOPM5_DIR(D, MU, SG)                                   \
  {                                                                 \
    cplx<float> W[9];                                               \
    _Pragma("unroll") for (int e = 0; e < 9; ++e)                   \
    {                                                               \
      const float2 q = sl[(D) * 9 + e][site];                       \
      W[e] = cmk<float>(q.x, q.y);                                  \
    }                                                               \
    cplx<float> psi[12];                                            \
    SpinorIO<float>::load(in, ls, row, nb[D], psi);                 \
    hop_term<MU, SG>(W, psi, acc);                                  \
  }
    QDWF_HOPM5_DIR(0, 0, SF)
    QDWF_HOPM5_DIR(1, 1, SF)
    QDWF_HOPM5_DIR(2, 2, SF)
    QDWF_HOPM5_DIR(3, 3, SF)
    QDWF_HOPM5_DIR(4, 0, SB_)
    QDWF_HOPM5_DIR(5, 1, SB_)
    QDWF_HOPM5_DIR(6, 2, SB_)
    QDWF_HOPM5_DIR(7, 3, SB_)
#undef QDWF_HOPM5_DIR
#pragma unroll
    for (int k = 0; k < 6; ++k)
      hr[row][k][site] = make_float4(-0.5f * acc[2 * k].re, -0.5f * acc[2 * k].im, -0.5f * acc[2 * k + 1].re,
                                     -0.5f * acc[2 * k + 1].im);
  }
  __syncthreads();
  double dot = 0.0;  // Re <dotp, out> over this thread's outputs (DOT)
  const int kp = row;  // phase 2: thread (site, slot)
  if (active && kp < 6) {
    const int chi = (kp < 3) ? 0 : 1;
    const cplx<float>* __restrict__ M = a.mat + chi * ls * ls;
    // x and dotp are prefetched for all s when that fits in registers
    constexpr bool PREF = LSMAX <= 12;
    float4 v[LSMAX];
    float4 xv[XPAY && PREF ? LSMAX : 1];
    float4 pv[DOT && PREF ? LSMAX : 1];
#pragma unroll
    for (int t = 0; t < LSMAX; ++t) {
      if (t < ls) {
        v[t] = hr[t][kp][site];
        if (PREF) {
          const int64_t o = sidx<6>(ls, t, kp, h);
          if (XPAY) xv[t] = ((const float4*)a.x)[coff + o];
          if (DOT) pv[t] = ((const float4*)a.dotp)[coff + o];
        }
      }
    }
#pragma unroll
    for (int s = 0; s < LSMAX; ++s) {
      if (s < ls) {
        float r0 = 0.f, i0 = 0.f, r1 = 0.f, i1 = 0.f;
#pragma unroll
        for (int t = 0; t < LSMAX; ++t) {
          if (t < ls) {
            const cplx<float> m = M[s * ls + t];
            r0 += m.re * v[t].x;
            i0 += m.re * v[t].y;
            r1 += m.re * v[t].z;
            i1 += m.re * v[t].w;
            if (!REAL) {
              r0 -= m.im * v[t].y;
              i0 += m.im * v[t].x;
              r1 -= m.im * v[t].w;
              i1 += m.im * v[t].z;
            }
          }
        }
        float4 res = make_float4(a.alpha * r0, a.alpha * i0, a.alpha * r1, a.alpha * i1);
        const int64_t o = coff + sidx<6>(ls, s, kp, h);
        if (XPAY) {
          const float4 xs = PREF ? xv[PREF ? s : 0] : ((const float4*)a.x)[o];
          res.x += a.beta * xs.x;
          res.y += a.beta * xs.y;
          res.z += a.beta * xs.z;
          res.w += a.beta * xs.w;
        }
        store_nt(&((float4*)a.out)[o], res);
        if (DOT) {
          const float4 ps = PREF ? pv[PREF ? s : 0] : ((const float4*)a.dotp)[o];
          dot += (double)ps.x * res.x + (double)ps.y * res.y + (double)ps.z * res.z + (double)ps.w * res.w;
        }
      }
    }
  }
  if (DOT) {
    __shared__ double red[(SB * LSMAX + 63) / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dot += __shfl_down(dot, o, 64);
    if (lane == 0) red[wv] = dot;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) t += red[i];
      a.part[ci * nblk + lb] = t;
    }
  }
}

template <int SB, int LSMAX>
static void launch_hop_m5_ls(const HopM5Args& a, bool dag, hipStream_t st)
{
  const dim3 grid((unsigned)((a.geo.vh + SB - 1) / SB * a.cols.n));
  const dim3 blk(SB * a.ls);
  const bool xpay = a.x != nullptr;
#define QDWF_L(D, X, R, DT) hipLaunchKernelGGL((k_hop_m5<SB, LSMAX, D, X, R, DT>), grid, blk, 0, st, a)
  if (a.dotp) {
    // only the x - Mat H^+ in form carries the CG dot product
    if (dag && xpay && a.real_mat) QDWF_L(true, true, true, true);
    else if (dag && xpay) QDWF_L(true, true, false, true);
    return;
  }
  if (a.real_mat) {
    if (!dag && !xpay) QDWF_L(false, false, true, false);
    if (!dag && xpay) QDWF_L(false, true, true, false);
    if (dag && !xpay) QDWF_L(true, false, true, false);
    if (dag && xpay) QDWF_L(true, true, true, false);
  } else {
    if (!dag && !xpay) QDWF_L(false, false, false, false);
    if (!dag && xpay) QDWF_L(false, true, false, false);
    if (dag && !xpay) QDWF_L(true, false, false, false);
    if (dag && xpay) QDWF_L(true, true, false, false);
  }
#undef QDWF_L
}

// Number of partial sums a DOT launch writes (one per block).
int64_t hop_m5_blocks(const LatGeo& g) { return (g.vh + kHopM5Sites - 1) / kHopM5Sites; }

bool launch_hop_m5(const HopM5Args& a, bool dag, hipStream_t st)
{
  // phase 2 needs >= 6 rows; whole waves (SB * ls multiple of 64) for the
  // block reduction
  if (a.ls < 6 || a.ls > 16 || (kHopM5Sites * a.ls) % 64 != 0) return false;
  if (a.ls <= 8) launch_hop_m5_ls<kHopM5Sites, 8>(a, dag, st);
  else if (a.ls <= 12) launch_hop_m5_ls<kHopM5Sites, 12>(a, dag, st);
  else launch_hop_m5_ls<kHopM5Sites, 16>(a, dag, st);
  return true;
}

// ================================================ FP64 hop + s-matrix
// out = alpha * Mat (H in) + beta * x on FP64 fields (one column), in the
// three phases of k_hop_m5s: a block is SB output half-sites x Ls rows;
// phase 1, thread (site, s) forms the hop of its 5D site with the block's
// links staged in LDS; phase 2, the Ls hop results of each site go through
// LDS and thread (site, s) forms row s of the dense s-matrix product;
// phase 3 adds beta * x, stores, and reduces Re <dotp, out> per block.
// The links' LDS space is reused for the hop results once every thread
// has finished phase 1.
template <int SB, int LSMAX>
__device__ __forceinline__ void stage_km_d(cplx<double> (*km)[LSMAX][LSMAX], const cplx<double>* __restrict__ M,
                                           int ls)
{
  for (int i = threadIdx.x; i < 2 * LSMAX * LSMAX; i += blockDim.x) {
    const int chi = i / (LSMAX * LSMAX), r = (i / LSMAX) % LSMAX, t = i % LSMAX;
    if (r < ls && t < ls) km[chi][r][t] = M[(chi * ls + r) * ls + t];
  }
}

template <int SB, int LSMAX, bool DAG, bool XPAY, bool REAL, bool DOT>
__global__ void __launch_bounds__(SB * LSMAX) k_hop_m5d(HopM5dArgs a)
{
  constexpr int NSLOT = (72 > LSMAX * 12) ? 72 : LSMAX * 12;
  __shared__ double2 buf[NSLOT][SB];  // links [72][SB], then hop results [LSMAX * 12][SB]
  __shared__ int nbs[8][SB];
  __shared__ cplx<double> km[2][LSMAX][LSMAX];
  const unsigned lb = (unsigned)xcd_logical_block(blockIdx.x, (int64_t)gridDim.x);
  const int64_t h0 = (int64_t)lb * SB;
  const int pslot = (a.eo_out == 2) ? 0 : 1;
  const double2* __restrict__ links = a.links + lidx(a.geo.stride, pslot, 0, h0);
  for (int i = threadIdx.x; i < 72 * SB; i += blockDim.x) buf[i / SB][i % SB] = links[(i / SB) * 64 + (i % SB)];
  for (int i = threadIdx.x; i < 8 * SB; i += blockDim.x) {  // thread (site, d): neighbour d
    const int st = i % SB, d = i / SB;
    if (h0 + st < a.geo.vh) nbs[d][st] = (int)neighbour_of(a.geo, h0 + st, a.eo_out, d);
  }
  stage_km_d<SB, LSMAX>(km, a.mat, a.ls);
  __syncthreads();
  const int ls = a.ls;
  const int site = threadIdx.x % SB, row = threadIdx.x / SB;
  const int64_t h = h0 + site;
  const bool active = h < a.geo.vh && row < ls;
  cplx<double> acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = cmk<double>(0, 0);
  if (active) {
    constexpr int SF = DAG ? 1 : -1;
    constexpr int SG = -SF;
#define QDWF_D(D, MU, SGN)                                         \
  {                                                                \
    cplx<double> W[9];                                             \
    _Pragma("unroll") for (int e = 0; e < 9; ++e)                  \
    {                                                              \
      const double2 q = buf[(D) * 9 + e][site];                    \
      W[e] = cmk<double>(q.x, q.y);                                \
    }                                                              \
    cplx<double> psi[12];                                          \
    SpinorIO<double>::load(a.in, ls, row, nbs[D][site], psi);      \
    hop_term<MU, SGN>(W, psi, acc);                                \
  }
    QDWF_D(0, 0, SF) QDWF_D(1, 1, SF) QDWF_D(2, 2, SF) QDWF_D(3, 3, SF)
    QDWF_D(4, 0, SG) QDWF_D(5, 1, SG) QDWF_D(6, 2, SG) QDWF_D(7, 3, SG)
#undef QDWF_D
  }
  const double sc = -0.5 * a.alpha;
  __syncthreads();  // every thread is done with the links: buf becomes hr
  double2 (*hr)[SB] = buf;  // hr[row * 12 + k][site]
  if (active) {
#pragma unroll
    for (int k = 0; k < 12; ++k) hr[row * 12 + k][site] = make_double2(sc * acc[k].re, sc * acc[k].im);
  }
  __syncthreads();
  double dot = 0.0;
  if (active) {
    cplx<double> v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = cmk<double>(0, 0);
#pragma unroll
    for (int t = 0; t < LSMAX; ++t) {
      if (t < ls) {
        const cplx<double> k0 = km[0][row][t], k1 = km[1][row][t];
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          const cplx<double> m = k < 6 ? k0 : k1;  // spins 0, 1: chirality 0
          const double2 q = hr[t * 12 + k][site];
          if (REAL) {
            v[k].re += m.re * q.x;
            v[k].im += m.re * q.y;
          } else {
            v[k] = cfma(v[k], m, cmk<double>(q.x, q.y));
          }
        }
      }
    }
    if (XPAY) {
      cplx<double> xv[12];
      SpinorIO<double>::load(a.x, ls, row, h, xv);
#pragma unroll
      for (int k = 0; k < 12; ++k) v[k] = cmk<double>(v[k].re + a.beta * xv[k].re, v[k].im + a.beta * xv[k].im);
    }
    SpinorIO<double>::store_stream(a.out, ls, row, h, v);
    if (DOT) {
      cplx<double> pv[12];
      SpinorIO<double>::load(a.dotp, ls, row, h, pv);
#pragma unroll
      for (int k = 0; k < 12; ++k) dot += pv[k].re * v[k].re + pv[k].im * v[k].im;
    }
  }
  if (DOT) {
    __shared__ double red[(SB * LSMAX + 63) / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dot += __shfl_down(dot, o, 64);
    if (lane == 0) red[wv] = dot;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) t += red[i];
      a.part[lb] = t;
    }
  }
}

template <int SB, int LSMAX>
static void launch_hop_m5d_ls(const HopM5dArgs& a, bool dag, hipStream_t st)
{
  const dim3 grid((unsigned)((a.geo.vh + SB - 1) / SB));
  const dim3 blk(SB * a.ls);
  const bool xpay = a.x != nullptr, dot = a.dotp != nullptr;
#define QDWF_L(D, X, R, DT) hipLaunchKernelGGL((k_hop_m5d<SB, LSMAX, D, X, R, DT>), grid, blk, 0, st, a)
  if (dot) {
    if (dag && xpay && a.real_mat) QDWF_L(true, true, true, true);
    else if (dag && xpay) QDWF_L(true, true, false, true);
    return;
  }
  if (a.real_mat) {
    if (!dag && !xpay) QDWF_L(false, false, true, false);
    if (!dag && xpay) QDWF_L(false, true, true, false);
    if (dag && !xpay) QDWF_L(true, false, true, false);
    if (dag && xpay) QDWF_L(true, true, true, false);
  } else {
    if (!dag && !xpay) QDWF_L(false, false, false, false);
    if (!dag && xpay) QDWF_L(false, true, false, false);
    if (dag && !xpay) QDWF_L(true, false, false, false);
    if (dag && xpay) QDWF_L(true, true, false, false);
  }
#undef QDWF_L
}

bool hop_m5d_ok(int ls) { return ls >= 4 && ls <= 16 && (16 * ls) % 64 == 0; }
int64_t hop_m5d_blocks(const LatGeo& g) { return (g.vh + 15) / 16; }

bool launch_hop_m5d(const HopM5dArgs& a, bool dag, hipStream_t st)
{
  if (!hop_m5d_ok(a.ls) || !a.mat) return false;
  if (a.dotp && !(dag && a.x)) return false;
  if (a.ls <= 8) launch_hop_m5d_ls<16, 8>(a, dag, st);
  else if (a.ls <= 12) launch_hop_m5d_ls<16, 12>(a, dag, st);
  else launch_hop_m5d_ls<16, 16>(a, dag, st);
  return true;
}

// The host's merged-CG scalar recurrence (alpha = |r|^2 / Re <p,Ap>,
// |r - alpha Ap|^2 = |r|^2 - 2 alpha Re <r,Ap> + alpha^2 |Ap|^2, beta =
// |r_new|^2 / |r|^2, reliable-update test |r_new|^2 <= delta^2 max|r|^2 or
// <= target) on the device, one thread per column slot.  No FMA contraction:
// the operations and their order are the host's, so both give the same
// doubles.
__global__ void __launch_bounds__(64) k_cg_scalars(CgState* __restrict__ cg, const double* __restrict__ red, int n)
{
#pragma clang fp contract(off)
  const int i = threadIdx.x;
  if (i >= n) return;
  if (cg->mode[i] != 0) {  // reached the threshold last iteration: frozen from now on
    cg->mode[i] = 1;
    return;
  }
  const double rr = cg->rr[i];
  const double alpha = rr / red[i];
  const double rr_new = fmax(0.0, rr - 2.0 * alpha * red[n + i] + alpha * alpha * red[2 * n + i]);
  cg->alpha[i] = alpha;
  cg->rr_new[i] = rr_new;
  cg->iters[i] += 1;
  if (!isfinite(rr_new) || !isfinite(alpha)) {
    cg->bad[i] = 1;
    cg->mode[i] = 1;
    return;
  }
  // x += alpha_x p + alpha_q q (q: the previous p, in the other buffer),
  // with the folded update of the host loop: a direction's alpha waits one
  // step in pend and is added together with the next one
  const double pend = cg->pend[i];
  if (rr_new <= cg->d2 * cg->maxrr[i] || rr_new <= cg->target) {
    cg->alpha_x[i] = alpha;
    cg->alpha_q[i] = pend;
    cg->pend[i] = 0.0;
    cg->mode[i] = 2;
    return;
  }
  if (!cg->x_defer) {
    cg->alpha_x[i] = alpha;
    cg->alpha_q[i] = 0.0;
  } else if (pend == 0.0) {
    cg->alpha_x[i] = 0.0;
    cg->alpha_q[i] = 0.0;
    cg->pend[i] = alpha;
  } else {
    cg->alpha_x[i] = alpha;
    cg->alpha_q[i] = pend;
    cg->pend[i] = 0.0;
  }
  cg->alpha_r[i] = alpha;
  cg->beta[i] = rr_new / rr;
  cg->rr[i] = rr_new;
}

void launch_cg_scalars(CgState* cg, const double* red, int n, hipStream_t st)
{
  hipLaunchKernelGGL(k_cg_scalars, dim3(1), dim3(64), 0, st, cg, red, n);
}

// Source propagator: thread (site, row) writes row `row` of the site's
// 12x12 matrix, f(x) on the diagonal (set_point_src, set_wall_src,
// set_mom_src, set_rand_vol_u1_src: qcd-prop.cpp:61-150, :260-307).  The
// phase is summed over mu = 0..3 in order, as the reference loop does.
__global__ void __launch_bounds__(kBlasBlock) k_make_src(double2* __restrict__ prop, SrcArgs a)
{
#pragma clang fp contract(off)
  const int64_t vol = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.n[3];
  const int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x;
  if (i >= vol * 12) return;
  const int64_t site = i / 12;
  const int row = (int)(i - site * 12);
  int x[4];
  int64_t r = site;
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    const int64_t q = r / a.n[mu];
    x[mu] = (int)(r - q * a.n[mu]);
    r = q;
  }
  x[3] = (int)r;
  double2 f = make_double2(0.0, 0.0);
  if (a.kind == 0) {
    if (x[0] == a.xg[0] && x[1] == a.xg[1] && x[2] == a.xg[2] && x[3] == a.xg[3]) f = make_double2(a.value_re, a.value_im);
  } else if (a.kind == 1 || a.kind == 2) {
    if (a.kind == 2 || x[3] == a.tslice) {
      double phase = 0.0;
#pragma unroll
      for (int mu = 0; mu < 4; ++mu) phase += a.mom[mu] * x[mu];
      double sn, cs;
      sincos(phase, &sn, &cs);
      f = make_double2(cs, sn);
    }
  } else {
    f = a.fu1[site];
  }
  double2* out = prop + i * 12;
#pragma unroll
  for (int c = 0; c < 12; ++c) out[c] = (c == row) ? f : make_double2(0.0, 0.0);
}

void launch_make_src(double2* prop, const SrcArgs& a, hipStream_t st)
{
  const int64_t n = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.n[3] * 12;
  hipLaunchKernelGGL(k_make_src, dim3((unsigned)((n + kBlasBlock - 1) / kBlasBlock)), dim3(kBlasBlock), 0, st, prop, a);
}

// ------------------------------------------------------------ QED operator
// multiply_m_dwf_qed (qed.cpp:263-411): thread (site, s) forms out(x, s)
// from in(x, s'), s' in {s-1, s, s+1, 0, Ls-1}, and the 8 neighbours at s.
// CPS gammas (matrix.h:426-452) as one non-zero per row: row r of gamma_mu
// has coefficient kQg[mu][r] (1, i, -1, -i as 0..3) in column kQc[mu][r].
__constant__ int kQc[4][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}, {2, 3, 0, 1}};
__constant__ int kQg[4][4] = {{1, 1, 3, 3}, {2, 0, 0, 2}, {1, 3, 3, 1}, {0, 0, 0, 0}};

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zdiv(double2 a, double2 b)
{
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double2 zscale(double c, double2 a) { return make_double2(c * a.x, c * a.y); }
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
// i^k * a
__device__ __forceinline__ double2 zipow(int k, double2 a)
{
  return k == 0 ? a : k == 1 ? make_double2(-a.y, a.x) : k == 2 ? make_double2(-a.x, -a.y) : make_double2(a.y, -a.x);
}

// v += c * iv + cg * (G iv), G = gamma_5 (mu < 0) or CPS gamma_mu
__device__ __forceinline__ void qed_acc(double2 v[4], double2 c, double2 cg, int mu, const double2* __restrict__ iv)
{
  double2 w[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) w[r] = iv[r];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double2 gv = mu < 0 ? (r < 2 ? w[r] : zscale(-1.0, w[r])) : zipow(kQg[mu][r], w[kQc[mu][r]]);
    const double2 a = zmul(c, w[r]), b = zmul(cg, gv);
    v[r] = make_double2(v[r].x + a.x + b.x, v[r].y + a.y + b.y);
  }
}

__global__ void __launch_bounds__(kBlasBlock) k_qed_m(QedArgs a)
{
  const int ls = a.ls;
  const int64_t vol = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.n[3];
  const int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x;
  if (i >= vol * ls) return;
  const int64_t site = i / ls;
  const int s = (int)(i - site * ls);
  int x[4];
  int64_t r = site;
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    const int64_t q = r / a.n[mu];
    x[mu] = (int)(r - q * a.n[mu]);
    r = q;
  }
  x[3] = (int)r;
  const int T = a.n[3];
  double2 tw = make_double2(1.0, 0.0), twm = make_double2(1.0, 0.0);
  if (a.wick) {
    tw = a.wick[x[3]];
    twm = a.wick[(x[3] - 1 + T) % T];
  }
  if (a.dagger) {
    tw = zconj(tw);
    twm = zconj(twm);
  }
  const double2 one = make_double2(1.0, 0.0);
  const double2 avg = make_double2(0.5 * tw.x + 0.5 * twm.x, 0.5 * tw.y + 0.5 * twm.y);
  const double2 itw = zdiv(one, tw), itwm = zdiv(one, twm);
  const double2 iavg = make_double2(0.5 * itw.x + 0.5 * itwm.x, 0.5 * itw.y + 0.5 * itwm.y);
  const double df = a.dagger ? -1.0 : 1.0;
  const double2 zero = make_double2(0.0, 0.0);
  auto at = [&](int64_t st, int ss) { return a.in + (st * ls + ss) * 4; };
  double2 v[4] = {zero, zero, zero, zero};
  // ((4 - m5) avg + inv_avg) in
  qed_acc(v, make_double2((4.0 - a.m5) * avg.x + iavg.x, (4.0 - a.m5) * avg.y + iavg.y), zero, -1, at(site, s));
  // fifth dimension (qed.cpp:342-371)
  if (s < ls - 1) qed_acc(v, zscale(-0.5, avg), zscale(df * 0.5, avg), -1, at(site, s + 1));
  if (s > 0) qed_acc(v, zscale(-0.5, avg), zscale(-df * 0.5, avg), -1, at(site, s - 1));
  if (s == ls - 1) qed_acc(v, zscale(0.5 * a.mass, avg), zscale(-df * 0.5 * a.mass, avg), -1, at(site, 0));
  if (s == 0) qed_acc(v, zscale(0.5 * a.mass, avg), zscale(df * 0.5 * a.mass, avg), -1, at(site, ls - 1));
  // hopping (qed.cpp:373-408)
  int64_t stride = 1;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    const int L = a.n[mu];
    const int64_t sp = site + (x[mu] + 1 == L ? -(int64_t)(L - 1) : 1) * stride;
    const int64_t sm = site + (x[mu] == 0 ? (int64_t)(L - 1) : -1) * stride;
    stride *= L;
    double2 up = a.gf[site * 4 + mu];
    double2 um = zdiv(one, a.gf[sm * 4 + mu]);
    if (a.dagger) {
      up = zconj(zdiv(one, up));
      um = zconj(zdiv(one, um));
    }
    double2 tf1 = avg, tfm1 = avg, tf2 = avg, tfm2 = avg;
    if (mu == 3) {
      tf1 = tfm1 = one;
      tf2 = itw;
      tfm2 = itwm;
    }
    qed_acc(v, zscale(-0.5, zmul(tf2, up)), zscale(df * 0.5, zmul(up, tf1)), mu, at(sp, s));
    qed_acc(v, zscale(-0.5, zmul(tfm2, um)), zscale(-df * 0.5, zmul(um, tfm1)), mu, at(sm, s));
  }
  double2* o = a.out + (site * ls + s) * 4;
#pragma unroll
  for (int k = 0; k < 4; ++k) o[k] = v[k];
}

void launch_qed_m(const QedArgs& a, hipStream_t st)
{
  const int64_t n = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.n[3] * a.ls;
  hipLaunchKernelGGL(k_qed_m, dim3((unsigned)((n + kBlasBlock - 1) / kBlasBlock)), dim3(kBlasBlock), 0, st, a);
}

// ------------------------------------------------------- free-field DWF
// out(.., k_mu, ..) = sum_x in(.., x_mu, ..) exp(sign 2 pi i k x / L): thread
// (site, entry) over the 144 entries of a propagator site; the phase index
// (k x) mod L keeps the argument of sincos in [0, 2 pi).
struct DftArgs {
  int n[4];
  int mu;
  int sign;
};
__global__ void __launch_bounds__(kBlasBlock) k_dft_dir(double2* __restrict__ out, const double2* __restrict__ in,
                                                        DftArgs a)
{
  const int64_t vol = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.n[3];
  const int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x;
  if (i >= vol * 144) return;
  const int64_t site = i / 144;
  const int e = (int)(i - site * 144);
  int64_t stride = 1;
  for (int m = 0; m < a.mu; ++m) stride *= a.n[m];
  const int L = a.n[a.mu];
  const int k = (int)((site / stride) % L);
  const int64_t base = site - (int64_t)k * stride;
  const double w = 2.0 * 3.141592653589793 / L;
  double re = 0.0, im = 0.0;
  for (int x = 0; x < L; ++x) {
    const int ph = (int)(((int64_t)k * x) % L);
    double sn, cs;
    sincos(w * ph, &sn, &cs);
    sn *= a.sign;
    const double2 v = in[(base + (int64_t)x * stride) * 144 + e];
    re += v.x * cs - v.y * sn;
    im += v.x * sn + v.y * cs;
  }
  out[i] = make_double2(re, im);
}

void launch_dft_dir(double2* out, const double2* in, const int n[4], int mu, int sign, hipStream_t st)
{
  DftArgs a{};
  for (int m = 0; m < 4; ++m) a.n[m] = n[m];
  a.mu = mu;
  a.sign = sign;
  const int64_t tot = (int64_t)n[0] * n[1] * n[2] * n[3] * 144;
  hipLaunchKernelGGL(k_dft_dir, dim3((unsigned)((tot + kBlasBlock - 1) / kBlasBlock)), dim3(kBlasBlock), 0, st, out,
                     in, a);
}

// free_mom_invert (qcd-prop.h:292-345): at momentum k the propagator is
// multiplied by ipgm = lwa (-i sum_i sin(k_i) gamma_i + m lwa) / (p^2 + (m lwa)^2)
// on the spin index (colour identity), or set to zero where p^2 and lwa are
// both below 1e-10; every entry is then scaled by `scale`.  Thread (site,
// colour c, column) rewrites the 4 spin rows of that colour and column.
struct FreeMomArgs {
  int n[4];
  double mass, m5, tw[4], scale;
};
__global__ void __launch_bounds__(kBlasBlock) k_free_mom(double2* __restrict__ prop, FreeMomArgs a)
{
  const int64_t vol = (int64_t)a.n[0] * a.n[1] * a.n[2] * a.n[3];
  const int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x;
  if (i >= vol * 36) return;
  const int64_t site = i / 36;
  const int cc = (int)(i - site * 36);
  const int c = cc / 12, col = cc % 12;
  int kx[4];
  int64_t r = site;
  for (int m = 0; m < 4; ++m) {
    kx[m] = (int)(r % a.n[m]);
    r /= a.n[m];
  }
  double p2 = 0.0, wp = 1.0 - a.m5, ks[4];
  for (int m = 0; m < 4; ++m) {
    const double kk = 2.0 * 3.141592653589793 * (kx[m] + a.tw[m]) / (double)a.n[m];
    ks[m] = sin(kk);
    p2 += ks[m] * ks[m];
    const double sh = sin(kk / 2.0);
    wp += 2.0 * sh * sh;
  }
  const double calpha = (1.0 + wp * wp + p2) / 2.0 / wp;
  const double alpha = acosh(calpha);
  const double lwa = 1.0 - wp * exp(-alpha);
  double2* base = prop + site * 144;
  double2 v[4];
  for (int s = 0; s < 4; ++s) v[s] = base[(s * 3 + c) * 12 + col];
  double2 o[4];
  if (1.0e-10 > p2 && 1.0e-10 > lwa) {
    for (int s = 0; s < 4; ++s) o[s] = make_double2(0.0, 0.0);
  } else {
    const double f = lwa / (p2 + (a.mass * lwa) * (a.mass * lwa));
    // pg v with CPS gammas (one non-zero per row), then ipgm v = f (-i pg v + m lwa v)
    for (int s = 0; s < 4; ++s) {
      double2 pg = make_double2(0.0, 0.0);
      for (int m = 0; m < 4; ++m) {
        const double2 g = zipow(kQg[m][s], v[kQc[m][s]]);
        pg.x += ks[m] * g.x;
        pg.y += ks[m] * g.y;
      }
      o[s] = make_double2(f * (pg.y + a.mass * lwa * v[s].x), f * (-pg.x + a.mass * lwa * v[s].y));
    }
  }
  for (int s = 0; s < 4; ++s) base[(s * 3 + c) * 12 + col] = make_double2(a.scale * o[s].x, a.scale * o[s].y);
}

void launch_free_mom(double2* prop, const int n[4], double mass, double m5, const double momtwist[4], double scale,
                     hipStream_t st)
{
  FreeMomArgs a{};
  for (int m = 0; m < 4; ++m) {
    a.n[m] = n[m];
    a.tw[m] = momtwist[m];
  }
  a.mass = mass;
  a.m5 = m5;
  a.scale = scale;
  const int64_t tot = (int64_t)n[0] * n[1] * n[2] * n[3] * 36;
  hipLaunchKernelGGL(k_free_mom, dim3((unsigned)((tot + kBlasBlock - 1) / kBlasBlock)), dim3(kBlasBlock), 0, st, prop,
                     a);
}

// ---------------------------------------------------------------- deflation
// The three steps of deflate (dslash.cpp:144-249) over blocks of block_site:
// project the FP32-rounded input on the block bases (FP32 products, FP64
// sums), inner products with the eigenvectors' coefficients summed over
// blocks and divided by the eigenvalues, and the produce step, which adds
// sum_j basis_j c_j (FP32) to the blocked input, as the reference does.

// hmap[b][bidx]: device half-site of the bidx-th half site (parity eo) of
// block b (convert_half_vector, commpressed-eigen-io.cpp:925-947).
__global__ void __launch_bounds__(kBlasBlock) k_defl_hmap(int* __restrict__ hmap, DeflArgs a)
{
  const int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x;
  if (i >= (int64_t)a.nblock * a.block_vol_eo) return;
  const int b = (int)(i / a.block_vol_eo);
  const int bidx = (int)(i - (int64_t)b * a.block_vol_eo);
  int bc[4], l[4];
  int r = b;
  for (int mu = 0; mu < 4; ++mu) {
    const int nb = a.geo.n[mu] / a.bs[mu];
    bc[mu] = r % nb;
    r /= nb;
  }
  r = 2 * bidx;  // lexicographic index in the block; bs[0] even
  for (int mu = 0; mu < 4; ++mu) {
    l[mu] = r % a.bs[mu];
    r /= a.bs[mu];
  }
  int x[4];
  for (int mu = 0; mu < 4; ++mu) x[mu] = bc[mu] * a.bs[mu] + l[mu];
  const int odd = (x[0] + x[1] + x[2] + x[3]) & 1;
  if (odd != (a.eo == 1 ? 1 : 0)) x[0] += 1;  // the other site of the x pair
  hmap[i] = (int)half_index(a.geo, x);
}

__device__ __forceinline__ unsigned defl_slot(const DeflArgs& a, int b, int64_t k)
{
  const int per = a.ls * 12;
  const int bidx = (int)(k / per);
  const int rem = (int)(k - (int64_t)bidx * per);
  const int s = rem / 12;
  return sidx<12>(a.ls, s, rem - s * 12, a.hmap[(int64_t)b * a.block_vol_eo + bidx]);
}

// one workgroup per block; n_basis doubles2 of dynamic LDS
__global__ void __launch_bounds__(kBlasBlock) k_defl_project(DeflArgs a)
{
  extern __shared__ double2 cj[];
  const int b = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = wv; j < a.n_basis; j += kBlasBlock / 64) {
    const float2* __restrict__ bj = a.basis + ((int64_t)b * a.n_basis + j) * a.bsize;
    double re = 0.0, im = 0.0;
    for (int64_t k = lane; k < a.bsize; k += 64) {
      const float2 bk = bj[k];
      const double2 vd = a.in[defl_slot(a, b, k)];
      const float vx = (float)vd.x, vy = (float)vd.y;
      // conj(b) v in FP32
      re += (double)(bk.x * vx + bk.y * vy);
      im += (double)(bk.x * vy - bk.y * vx);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      re += __shfl_down(re, o, 64);
      im += __shfl_down(im, o, 64);
    }
    if (lane == 0) cj[j] = make_double2(re, im);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < a.n_vec; i += kBlasBlock) {
    const float2* __restrict__ ci = a.coef + ((int64_t)b * a.n_vec + i) * a.n_basis;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < a.n_basis; ++j) {  // conj(coef) c, FP64
      const double cx = ci[j].x, cy = ci[j].y;
      re += cx * cj[j].x + cy * cj[j].y;
      im += cx * cj[j].y - cy * cj[j].x;
    }
    a.part[(int64_t)b * a.n_vec + i] = make_double2(re, im);
  }
}

// p[i] = sum over blocks (block order) / lambda_i
__global__ void __launch_bounds__(kBlasBlock) k_defl_reduce(DeflArgs a)
{
  for (int i = threadIdx.x; i < a.n_vec; i += kBlasBlock) {
    double re = 0.0, im = 0.0;
    for (int b = 0; b < a.nblock; ++b) {
      const double2 v = a.part[(int64_t)b * a.n_vec + i];
      re += v.x;
      im += v.y;
    }
    a.p[i] = make_double2(re / a.evals[i], im / a.evals[i]);
  }
}

// one workgroup per block; n_basis float2 of dynamic LDS
__global__ void __launch_bounds__(kBlasBlock) k_defl_produce(DeflArgs a)
{
#pragma clang fp contract(off)
  extern __shared__ float2 cf[];
  const int b = blockIdx.x;
  for (int j = threadIdx.x; j < a.n_basis; j += kBlasBlock) {
    double re = 0.0, im = 0.0;
    for (int i = 0; i < a.n_vec; ++i) {  // coef p, FP64, i in order
      const float2 c = a.coef[((int64_t)b * a.n_vec + i) * a.n_basis + j];
      const double2 pi = a.p[i];
      re += (double)c.x * pi.x - (double)c.y * pi.y;
      im += (double)c.x * pi.y + (double)c.y * pi.x;
    }
    cf[j] = make_float2((float)re, (float)im);
  }
  __syncthreads();
  for (int64_t k = threadIdx.x; k < a.bsize; k += kBlasBlock) {
    const unsigned slot = defl_slot(a, b, k);
    const double2 vd = a.in[slot];
    float vx = (float)vd.x, vy = (float)vd.y;
    for (int j = 0; j < a.n_basis; ++j) {  // v += basis_j c_j, FP32, j in order
      const float2 bk = a.basis[((int64_t)b * a.n_basis + j) * a.bsize + k];
      const float2 c = cf[j];
      vx += bk.x * c.x - bk.y * c.y;
      vy += bk.x * c.y + bk.y * c.x;
    }
    a.out[slot] = make_double2((double)vx, (double)vy);
  }
}

void launch_defl_hmap(int* hmap, const DeflArgs& a, hipStream_t st)
{
  const int64_t n = (int64_t)a.nblock * a.block_vol_eo;
  hipLaunchKernelGGL(k_defl_hmap, dim3((unsigned)((n + kBlasBlock - 1) / kBlasBlock)), dim3(kBlasBlock), 0, st, hmap, a);
}

void launch_deflate(const DeflArgs& a, hipStream_t st)
{
  hipLaunchKernelGGL(k_defl_project, dim3(a.nblock), dim3(kBlasBlock), a.n_basis * sizeof(double2), st, a);
  hipLaunchKernelGGL(k_defl_reduce, dim3(1), dim3(kBlasBlock), 0, st, a);
  hipLaunchKernelGGL(k_defl_produce, dim3(a.nblock), dim3(kBlasBlock), a.n_basis * sizeof(float2), st, a);
}

// r_i -= alpha_i ap_i over 5D sites (ap in storage ST, r in SR) with
// per-column partial |r_i|^2 of the stored r (part[i*gridDim.x + block]).
template <class ST, class SR>
__global__ void __launch_bounds__(kBlasBlock) k_axpy_norm_site(AxpyNormSiteArgs a)
{
  typedef typename SpinorIO<ST>::slot_t slot_t;
  typedef typename SpinorIO<SR>::slot_t rslot_t;
  const int ci = blockIdx.y;
  const int col = pick(a.cols.rhs, ci);
  const slot_t* __restrict__ ap = (const slot_t*)a.ap + (int64_t)col * a.cols.fstride;
  rslot_t* __restrict__ r = (rslot_t*)a.r + (int64_t)col * a.r_fstride;
  const float al = (float)pick(a.alpha, ci);
  const int64_t per_blk = (int64_t)a.ls * 64;
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < a.n5; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t hb = i / per_blk, rem = i - hb * per_blk;
    const int s = (int)(rem >> 6);
    const int64_t h = (hb << 6) + (rem & 63);
    cplx<float> av[12], rv[12];
    SpinorIO<ST>::load(ap, a.ls, s, h, av);
    SpinorIO<SR>::load(r, a.ls, s, h, rv);
#pragma unroll
    for (int k = 0; k < 12; ++k) rv[k] = cmk<float>(rv[k].re - al * av[k].re, rv[k].im - al * av[k].im);
    store_q<SR>(r, a.ls, s, h, rv);
#pragma unroll
    for (int k = 0; k < 12; ++k) v += (double)rv[k].re * rv[k].re + (double)rv[k].im * rv[k].im;
  }
  __shared__ double sh[kBlasBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kBlasBlock / 64; ++i) t += sh[i];
    a.part[(int64_t)ci * gridDim.x + blockIdx.x] = t;
  }
}

static int site_grid(int64_t n5)
{
  int64_t g = (n5 + kBlasBlock - 1) / kBlasBlock;
  if (g > kMaxBlasGrid) g = kMaxBlasGrid;
  return (int)g;
}

int launch_axpy_norm_site(const AxpyNormSiteArgs& a, int storage, hipStream_t st)
{
  const int g = site_grid(a.n5);
  const dim3 grid((unsigned)g, a.cols.n);
  if (storage == 2) hipLaunchKernelGGL((k_axpy_norm_site<Half16, Half16>), grid, dim3(kBlasBlock), 0, st, a);
  else if (storage == 1) hipLaunchKernelGGL((k_axpy_norm_site<Half16, float>), grid, dim3(kBlasBlock), 0, st, a);
  else hipLaunchKernelGGL((k_axpy_norm_site<float, float>), grid, dim3(kBlasBlock), 0, st, a);
  return g;
}

// x += alpha p over the 5D sites of one field (p in storage ST, x in SX)
template <class ST, class SX>
__global__ void __launch_bounds__(kBlasBlock) k_axpy_site(float alpha, const void* pp, void* xx, int ls, int64_t n5)
{
  typedef typename SpinorIO<ST>::slot_t slot_t;
  typedef typename SpinorIO<SX>::slot_t xslot_t;
  const slot_t* __restrict__ p = (const slot_t*)pp;
  xslot_t* __restrict__ x = (xslot_t*)xx;
  const int64_t per_blk = (int64_t)ls * 64;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n5; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t hb = i / per_blk, rem = i - hb * per_blk;
    const int s = (int)(rem >> 6);
    const int64_t h = (hb << 6) + (rem & 63);
    cplx<float> pv[12], xv[12];
    SpinorIO<ST>::load(p, ls, s, h, pv);
    SpinorIO<SX>::load(x, ls, s, h, xv);
#pragma unroll
    for (int k = 0; k < 12; ++k) xv[k] = cmk<float>(xv[k].re + alpha * pv[k].re, xv[k].im + alpha * pv[k].im);
    SpinorIO<SX>::store(x, ls, s, h, xv);
  }
}

void launch_axpy_site(double alpha, const void* p, void* x, int ls, int64_t n5, int storage, hipStream_t st)
{
  const dim3 grid((unsigned)site_grid(n5));
  const float a = (float)alpha;
  if (storage == 2) hipLaunchKernelGGL((k_axpy_site<Half16, Half16>), grid, dim3(kBlasBlock), 0, st, a, p, x, ls, n5);
  else if (storage == 1) hipLaunchKernelGGL((k_axpy_site<Half16, float>), grid, dim3(kBlasBlock), 0, st, a, p, x, ls, n5);
  else hipLaunchKernelGGL((k_axpy_site<float, float>), grid, dim3(kBlasBlock), 0, st, a, p, x, ls, n5);
}

// FP64 blocked field <-> FP16 (per-site scale) stack column, over 5D sites:
// out = scale * in (d2h); out64 (+)= scale * in (h2d, acc = 1 adds).
__global__ void __launch_bounds__(kBlasBlock) k_d2h(const double2* __restrict__ in, uint4* __restrict__ out, int ls,
                                                    int64_t n5, double scale)
{
  const int64_t per_blk = (int64_t)ls * 64;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n5; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t hb = i / per_blk, rem = i - hb * per_blk;
    const int s = (int)(rem >> 6);
    const int64_t h = (hb << 6) + (rem & 63);
    cplx<double> d[12];
    SpinorIO<double>::load(in, ls, s, h, d);
    cplx<float> v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = cmk<float>((float)(scale * d[k].re), (float)(scale * d[k].im));
    SpinorIO<Half16>::store(out, ls, s, h, v);
  }
}
__global__ void __launch_bounds__(kBlasBlock) k_h2d(const uint4* __restrict__ in, double2* __restrict__ out, int ls,
                                                    int64_t n5, int acc, double scale)
{
  const int64_t per_blk = (int64_t)ls * 64;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n5; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t hb = i / per_blk, rem = i - hb * per_blk;
    const int s = (int)(rem >> 6);
    const int64_t h = (hb << 6) + (rem & 63);
    cplx<float> v[12];
    SpinorIO<Half16>::load(in, ls, s, h, v);
    cplx<double> d[12];
    if (acc) SpinorIO<double>::load(out, ls, s, h, d);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const double re = scale * v[k].re, im = scale * v[k].im;
      d[k] = acc ? cmk<double>(d[k].re + re, d[k].im + im) : cmk<double>(re, im);
    }
    SpinorIO<double>::store(out, ls, s, h, d);
  }
}
void launch_d2h(const double2* in, uint4* out, int ls, int64_t n5, double scale, hipStream_t st)
{
  hipLaunchKernelGGL(k_d2h, dim3((unsigned)site_grid(n5)), dim3(kBlasBlock), 0, st, in, out, ls, n5, scale);
}
void launch_h2d(const uint4* in, double2* out, int ls, int64_t n5, int acc, double scale, hipStream_t st)
{
  hipLaunchKernelGGL(k_h2d, dim3((unsigned)site_grid(n5)), dim3(kBlasBlock), 0, st, in, out, ls, n5, acc, scale);
}

// ============================================================ force
// Link derivative of S = phi^+ (Mpc^+ Mpc)^-1 phi.  With full-lattice
// fields A (Y on odd, w on even) and B (a on odd, c on even) -- see the host
// code for their definition -- dS = sum_{y,mu} Re Tr[G_mu(y) dU_mu(y)] with
//   G_mu(y) = - sum_s [ (1 - g_mu) B(y+mu) (x) A(y)^+  +  A(y+mu) (x) ((1 + g_mu) B(y))^+ ]
// (outer products in colour, summed over spin and s).  The kernel writes
// F_mu(y) = T_A(U_mu(y) G_mu(y)), T_A(M) = (M - M^+)/2 - Tr(M - M^+)/6, the
// su(3) projection of the reference's gauge force (matrix-hmc.h:202-213,
// hmc.cpp:64-75), and optionally G^+ (the gradient w.r.t. the complex link
// entries in the conjugate-Wirtinger convention).
// (1 + sg g_mu) psi = (w, sg B^+ w), w = u + sg B l  (dwf_common.h)
template <int MU, int SG>
__device__ __forceinline__ void full_proj(const cplx<double> psi[12], cplx<double> out[12])
{
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    cplx<double> w0, w1;
    proj<MU, SG>(psi[c], psi[3 + c], psi[6 + c], psi[9 + c], w0, w1);
    cplx<double> l0 = cmk<double>(0, 0), l1 = cmk<double>(0, 0);
    recon_lower<MU, SG>(w0, w1, l0, l1);
    out[c] = w0;
    out[3 + c] = w1;
    out[6 + c] = l0;
    out[9 + c] = l1;
  }
}

// G += sign * sum_spin a (x) b^*  (G[c][c'] += a[spin][c] conj(b[spin][c']))
__device__ __forceinline__ void outer_acc(cplx<double> G[9], const cplx<double> a[12], const cplx<double> b[12])
{
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int cp = 0; cp < 3; ++cp) {
      double re = 0, im = 0;
#pragma unroll
      for (int sp = 0; sp < 4; ++sp) {
        const cplx<double> x = a[sp * 3 + c], y = b[sp * 3 + cp];
        re += x.re * y.re + x.im * y.im;
        im += x.im * y.re - x.re * y.im;
      }
      G[c * 3 + cp].re += re;
      G[c * 3 + cp].im += im;
    }
}

template <int MU>
__device__ __forceinline__ void force_mu(const ForceArgs& a, int p, int64_t h, const int x[4], double2* __restrict__ fout,
                                         double2* __restrict__ gout)
{
  int xp[4] = {x[0], x[1], x[2], x[3]};
  xp[MU] = wrap_inc(x[MU], a.geo.n[MU]);
  const int64_t hp = half_index(a.geo, xp);
  const int q = 1 - p;  // parity slot of y + mu
  cplx<double> G[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) G[e] = cmk<double>(0, 0);
  for (int s = 0; s < a.ls; ++s) {
    cplx<double> v[12], w[12], pv[12];
    // (1 - g) B(y+mu) (x) A(y)^+
    SpinorIO<double>::load((const double2*)a.B[q], a.ls, s, hp, v);
    full_proj<MU, -1>(v, pv);
    SpinorIO<double>::load((const double2*)a.A[p], a.ls, s, h, w);
    outer_acc(G, pv, w);
    // A(y+mu) (x) ((1 + g) B(y))^+
    SpinorIO<double>::load((const double2*)a.B[p], a.ls, s, h, v);
    full_proj<MU, 1>(v, pv);
    SpinorIO<double>::load((const double2*)a.A[q], a.ls, s, hp, w);
    outer_acc(G, w, pv);
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) G[e] = cneg(G[e]);
  cplx<double> U[9];
  LinkIO<double>::load((const double2*)a.links, a.geo.stride, p, MU, h, U);
  // M = U G
  cplx<double> M[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cplx<double> t = cmul(U[i * 3 + 0], G[0 * 3 + j]);
      t = cfma(t, U[i * 3 + 1], G[1 * 3 + j]);
      t = cfma(t, U[i * 3 + 2], G[2 * 3 + j]);
      M[i * 3 + j] = t;
    }
  // T_A(M): anti-hermitian part, then remove the trace
  double F[18];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      F[2 * (i * 3 + j)] = 0.5 * (M[i * 3 + j].re - M[j * 3 + i].re);
      F[2 * (i * 3 + j) + 1] = 0.5 * (M[i * 3 + j].im + M[j * 3 + i].im);
    }
  const double tr = (F[1] + F[9] + F[17]) / 3.0;
  F[1] -= tr;
  F[9] -= tr;
  F[17] -= tr;
  const int64_t site = full_index(a.geo, x);
#pragma unroll
  for (int e = 0; e < 9; ++e) fout[(site * 4 + MU) * 9 + e] = make_double2(F[2 * e], F[2 * e + 1]);
  if (gout) {
    // G^+ : gradient w.r.t. the complex entries of U (conjugate Wirtinger)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) gout[(site * 4 + MU) * 9 + i * 3 + j] = make_double2(G[j * 3 + i].re, -G[j * 3 + i].im);
  }
}

// Block = 64 sites x the 4 directions (one wave per mu): the four waves
// read the sites' own X and Mpc X at every s together, so those come from
// HBM once and are shared through the caches, not re-read per direction.
__global__ void __launch_bounds__(256) k_force(ForceArgs a)
{
  const int64_t h = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const int p = blockIdx.y;  // parity slot of y: 0 even, 1 odd
  const int mu = threadIdx.y;  // wave-uniform
  if (h >= a.geo.vh) return;
  int x[4];
  half_coord(a.geo, h, p == 0 ? 2 : 1, x);
  double2* fout = (double2*)a.force;
  double2* gout = (double2*)a.grad;
  if (mu == 0) force_mu<0>(a, p, h, x, fout, gout);
  else if (mu == 1) force_mu<1>(a, p, h, x, fout, gout);
  else if (mu == 2) force_mu<2>(a, p, h, x, fout, gout);
  else force_mu<3>(a, p, h, x, fout, gout);
}

void launch_force(const ForceArgs& a, hipStream_t st)
{
  hipLaunchKernelGGL(k_force, dim3((unsigned)((a.geo.vh + 63) / 64), 2), dim3(64, 4), 0, st, a);
}

// ================================================================== M5
// out = alpha * (Mat_chi in) + beta * x, where Mat_chi is a dense Ls x Ls
// complex matrix per chirality (chi = 0: spins 0,1; chi = 1: spins 2,3).
// Every fifth-dimension operator of the path is of this form: M_ee, M_ee^+,
// M_ee^-1, M_ee^-+ (dslash.cpp:591-857), the Moebius pre-factor b + cS of
// M_eo (dslash.cpp:961-1003) and its adjoint (dslash.cpp:1005-1060), and
// their products.  Thread = (half-site, 16-byte slot); the slot index is
// blockIdx.y so chirality and matrix reads are wave-uniform.
template <class T, int LSMAX, bool XPAY, bool REAL, bool DOT>
__global__ void __launch_bounds__(kM5Block) k_m5(M5Args<T> a)
{
  typedef typename SpinorIO<T>::slot_t slot_t;
  const int64_t h = (int64_t)blockIdx.x * kM5Block + threadIdx.x;
  const int slot = blockIdx.y;
  double dot = 0.0;
  if (h < a.geo.vh) {
  constexpr int NS = SpinorIO<T>::kSlots;
  const int chi = (slot < NS / 2) ? 0 : 1;
  const int ls = a.ls;
  const cplx<T>* __restrict__ M = a.mat + (int64_t)chi * ls * ls;
  const int64_t coff = (int64_t)pick(a.cols.rhs, (int)blockIdx.z) * a.cols.fstride;
  const slot_t* __restrict__ in = (const slot_t*)a.in + coff;
  slot_t* __restrict__ out = (slot_t*)a.out + coff;
  // a slot holds one (FP64) or two (FP32) complex numbers
  constexpr int NC = (NS == 12) ? 1 : 2;
  cplx<T> v[LSMAX][NC];
#pragma unroll
  for (int t = 0; t < LSMAX; ++t) {
    if (t < ls) {
      const slot_t q = in[sidx<NS>(ls, t, slot, h)];
      if constexpr (NC == 1) {
        v[t][0] = cmk<T>(q.x, q.y);
      } else {
        v[t][0] = cmk<T>(q.x, q.y);
        v[t][1] = cmk<T>(q.z, q.w);
      }
    }
  }
  for (int s = 0; s < ls; ++s) {
    cplx<T> r[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) r[c] = cmk<T>(0, 0);
#pragma unroll
    for (int t = 0; t < LSMAX; ++t) {
      if (t < ls) {
        const cplx<T> m = M[s * ls + t];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (REAL) {
            r[c].re += m.re * v[t][c].re;
            r[c].im += m.re * v[t][c].im;
          } else {
            r[c] = cfma(r[c], m, v[t][c]);
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) r[c] = cmk<T>(a.alpha * r[c].re, a.alpha * r[c].im);
    const int64_t off = sidx<NS>(ls, s, slot, h);
    if (XPAY) {
      const slot_t q = ((const slot_t*)a.x)[coff + off];
      if constexpr (NC == 1) {
        r[0].re += a.beta * q.x;
        r[0].im += a.beta * q.y;
      } else {
        r[0].re += a.beta * q.x;
        r[0].im += a.beta * q.y;
        r[1].re += a.beta * q.z;
        r[1].im += a.beta * q.w;
      }
    }
    slot_t o;
    if constexpr (NC == 1) {
      o.x = r[0].re;
      o.y = r[0].im;
    } else {
      o.x = r[0].re;
      o.y = r[0].im;
      o.z = r[1].re;
      o.w = r[1].im;
    }
    out[off] = o;
    if (DOT) {
      const slot_t q = ((const slot_t*)a.dotp)[coff + off];
      if constexpr (NC == 1) {
        dot += (double)q.x * o.x + (double)q.y * o.y;
      } else {
        dot += (double)q.x * o.x + (double)q.y * o.y + (double)q.z * o.z + (double)q.w * o.w;
      }
    }
  }
  }
  if (DOT) {
    double v[1] = {dot};
    // block_reduce_store writes part[blockIdx.x]; offset rows by blockIdx.y
    __shared__ double sh[kM5Block / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[0] += __shfl_down(v[0], o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v[0];
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < kM5Block / 64; ++i) t += sh[i];
      a.part[((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = t;
    }
  }
}

// CG direction update fused with the first s-matrix of the next operator
// application:  x += alpha p (if UPDX);  p = r + beta p;  t = Mat p.
// Thread = (half-site, slot) looping over s, like k_m5.
template <class T, int LSMAX, bool UPDX, bool REAL>
__global__ void __launch_bounds__(kM5Block) k_cg_step(CgStepArgs<T> a)
{
  typedef typename SpinorIO<T>::slot_t slot_t;
  const int64_t h = (int64_t)blockIdx.x * kM5Block + threadIdx.x;
  if (h >= a.geo.vh) return;
  const int slot = blockIdx.y;
  constexpr int NS = SpinorIO<T>::kSlots;
  constexpr int NC = (NS == 12) ? 1 : 2;
  const int chi = (slot < NS / 2) ? 0 : 1;
  const int ls = a.ls;
  const cplx<T>* __restrict__ M = a.mat + (int64_t)chi * ls * ls;
  const int ci = blockIdx.z;
  const int64_t coff = (int64_t)pick(a.cols.rhs, ci) * a.cols.fstride;
  slot_t* __restrict__ x = (slot_t*)a.x + coff;
  slot_t* __restrict__ p = (slot_t*)a.p + coff;
  const slot_t* __restrict__ r = (const slot_t*)a.r + coff;
  const T al = (T)pick(a.alpha, ci), be = (T)pick(a.beta, ci);
  cplx<T> v[LSMAX][NC];
#pragma unroll
  for (int t = 0; t < LSMAX; ++t) {
    if (t < ls) {
      const int64_t o = sidx<NS>(ls, t, slot, h);
      const slot_t pv = p[o];
      const slot_t rv = r[o];
      if (UPDX) {
        slot_t xv = x[o];
        xv.x += al * pv.x;
        xv.y += al * pv.y;
        if constexpr (NC == 2) {
          xv.z += al * pv.z;
          xv.w += al * pv.w;
        }
        x[o] = xv;
      }
      slot_t pn;
      pn.x = be * pv.x + rv.x;
      pn.y = be * pv.y + rv.y;
      if constexpr (NC == 2) {
        pn.z = be * pv.z + rv.z;
        pn.w = be * pv.w + rv.w;
      }
      p[o] = pn;
      v[t][0] = cmk<T>(pn.x, pn.y);
      if constexpr (NC == 2) v[t][NC - 1] = cmk<T>(pn.z, pn.w);
    }
  }
  slot_t* __restrict__ out = (slot_t*)a.t + coff;
  for (int s = 0; s < ls; ++s) {
    cplx<T> res[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) res[c] = cmk<T>(0, 0);
#pragma unroll
    for (int t = 0; t < LSMAX; ++t) {
      if (t < ls) {
        const cplx<T> m = M[s * ls + t];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (REAL) {
            res[c].re += m.re * v[t][c].re;
            res[c].im += m.re * v[t][c].im;
          } else {
            res[c] = cfma(res[c], m, v[t][c]);
          }
        }
      }
    }
    slot_t o;
    o.x = res[0].re;
    o.y = res[0].im;
    if constexpr (NC == 2) {
      o.z = res[NC - 1].re;
      o.w = res[NC - 1].im;
    }
    out[sidx<NS>(ls, s, slot, h)] = o;
  }
}

template <class T, int LSMAX>
static void launch_cg_step_ls(const CgStepArgs<T>& a, hipStream_t st)
{
  const dim3 grid((unsigned)((a.geo.vh + kM5Block - 1) / kM5Block), SpinorIO<T>::kSlots, a.cols.n);
  const bool updx = a.x != nullptr;
  if (a.real_mat) {
    if (updx) hipLaunchKernelGGL((k_cg_step<T, LSMAX, true, true>), grid, dim3(kM5Block), 0, st, a);
    else hipLaunchKernelGGL((k_cg_step<T, LSMAX, false, true>), grid, dim3(kM5Block), 0, st, a);
  } else {
    if (updx) hipLaunchKernelGGL((k_cg_step<T, LSMAX, true, false>), grid, dim3(kM5Block), 0, st, a);
    else hipLaunchKernelGGL((k_cg_step<T, LSMAX, false, false>), grid, dim3(kM5Block), 0, st, a);
  }
}

template <class T>
void launch_cg_step(const CgStepArgs<T>& a, hipStream_t st)
{
  if (a.ls <= 8) launch_cg_step_ls<T, 8>(a, st);
  else if (a.ls <= 12) launch_cg_step_ls<T, 12>(a, st);
  else if (a.ls <= 16) launch_cg_step_ls<T, 16>(a, st);
  else if (a.ls <= 24) launch_cg_step_ls<T, 24>(a, st);
  else launch_cg_step_ls<T, 32>(a, st);
}

template void launch_cg_step<double>(const CgStepArgs<double>&, hipStream_t);
template void launch_cg_step<float>(const CgStepArgs<float>&, hipStream_t);

template <class T, int LSMAX>
static void launch_m5_ls(const M5Args<T>& a, hipStream_t st)
{
  const dim3 grid((unsigned)((a.geo.vh + kM5Block - 1) / kM5Block), SpinorIO<T>::kSlots, a.cols.n);
  if (a.dotp) {  // x-pay form with the CG dot product
    if (a.real_mat) hipLaunchKernelGGL((k_m5<T, LSMAX, true, true, true>), grid, dim3(kM5Block), 0, st, a);
    else hipLaunchKernelGGL((k_m5<T, LSMAX, true, false, true>), grid, dim3(kM5Block), 0, st, a);
  } else if (a.real_mat) {
    if (a.x) hipLaunchKernelGGL((k_m5<T, LSMAX, true, true, false>), grid, dim3(kM5Block), 0, st, a);
    else hipLaunchKernelGGL((k_m5<T, LSMAX, false, true, false>), grid, dim3(kM5Block), 0, st, a);
  } else {
    if (a.x) hipLaunchKernelGGL((k_m5<T, LSMAX, true, false, false>), grid, dim3(kM5Block), 0, st, a);
    else hipLaunchKernelGGL((k_m5<T, LSMAX, false, false, false>), grid, dim3(kM5Block), 0, st, a);
  }
}

template <class T>
void launch_m5(const M5Args<T>& a, hipStream_t st)
{
  if (a.ls <= 8) launch_m5_ls<T, 8>(a, st);
  else if (a.ls <= 12) launch_m5_ls<T, 12>(a, st);
  else if (a.ls <= 16) launch_m5_ls<T, 16>(a, st);
  else if (a.ls <= 24) launch_m5_ls<T, 24>(a, st);
  else launch_m5_ls<T, 32>(a, st);
}

template void launch_m5<double>(const M5Args<double>&, hipStream_t);
template void launch_m5<float>(const M5Args<float>&, hipStream_t);

// ================================================================ BLAS-1
// All vectors are flat arrays of 16-byte slots (n slots); pads are zero.

__device__ __forceinline__ double warp_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block reduction of NV doubles; thread 0 writes them to part[blockIdx.x*NV+i].
template <int NV>
__device__ __forceinline__ void block_reduce_store(double v[NV], double* part)
{
  __shared__ double sh[kBlasBlock / 64][NV];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = warp_sum(v[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) sh[wid][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double t = 0;
      for (int w = 0; w < kBlasBlock / 64; ++w) t += sh[w][i];
      part[(int64_t)blockIdx.x * NV + i] = t;
    }
  }
}

template <class S>
__device__ __forceinline__ double slot_norm2(const S& a);
template <>
__device__ __forceinline__ double slot_norm2<double2>(const double2& a)
{
  return a.x * a.x + a.y * a.y;
}
template <>
__device__ __forceinline__ double slot_norm2<float4>(const float4& a)
{
  return (double)a.x * a.x + (double)a.y * a.y + (double)a.z * a.z + (double)a.w * a.w;
}
// conj(a) b: (re, im)
__device__ __forceinline__ void slot_dot(const double2& a, const double2& b, double& re, double& im)
{
  re += a.x * b.x + a.y * b.y;
  im += a.x * b.y - a.y * b.x;
}
__device__ __forceinline__ void slot_dot(const float4& a, const float4& b, double& re, double& im)
{
  re += (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
  im += (double)a.x * b.y - (double)a.y * b.x + (double)a.z * b.w - (double)a.w * b.z;
}

__device__ __forceinline__ double2 slot_axpy(double a, double2 x, double2 y)
{
  return make_double2(y.x + a * x.x, y.y + a * x.y);
}
__device__ __forceinline__ float4 slot_axpy(double a, float4 x, float4 y)
{
  const float f = (float)a;
  return make_float4(y.x + f * x.x, y.y + f * x.y, y.z + f * x.z, y.w + f * x.w);
}
// x + a*y
__device__ __forceinline__ double2 slot_xpay(double2 x, double a, double2 y)
{
  return make_double2(a * y.x + x.x, a * y.y + x.y);
}
__device__ __forceinline__ float4 slot_xpay(float4 x, double a, float4 y)
{
  const float f = (float)a;
  return make_float4(f * y.x + x.x, f * y.y + x.y, f * y.z + x.z, f * y.w + x.w);
}

template <class S>
__global__ void __launch_bounds__(kBlasBlock) k_norm2(const S* __restrict__ a, int64_t n, double* part)
{
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock)
    v[0] += slot_norm2(a[i]);
  block_reduce_store<1>(v, part);
}

template <class S>
__global__ void __launch_bounds__(kBlasBlock) k_dot(const S* __restrict__ a, const S* __restrict__ b, int64_t n,
                                                    double* part)
{
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock)
    slot_dot(a[i], b[i], v[0], v[1]);
  block_reduce_store<2>(v, part);
}

// Compensated sums (Ogita, Rump and Oishi, "Accurate sum and dot product",
// SIAM J. Sci. Comput. 26 (2005), algorithms TwoSum / TwoProduct / Dot2):
// the pair (s, c) carries the sum to about twice FP64 precision.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
  s = a + b;
  const double z = s - a;
  e = (a - (s - z)) + (b - z);
}
__device__ __forceinline__ void dot2_add(double& s, double& c, double a, double b)
{
  const double p = a * b;
  const double q = fma(a, b, -p);  // exact: a*b = p + q
  double r;
  two_sum(s, p, s, r);
  c += q + r;
}

// part[blk] = (s, c) of sum_i 2 Re conj(phi_i) x_i - |y_i|^2
__global__ void __launch_bounds__(kBlasBlock) k_action_dd(const double2* __restrict__ phi,
                                                          const double2* __restrict__ x,
                                                          const double2* __restrict__ y, int64_t n, double* part)
{
  double s = 0.0, c = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock) {
    const double2 f = phi[i], a = x[i], b = y[i];
    dot2_add(s, c, 2.0 * f.x, a.x);
    dot2_add(s, c, 2.0 * f.y, a.y);
    dot2_add(s, c, -b.x, b.x);
    dot2_add(s, c, -b.y, b.y);
  }
  __shared__ double sh[kBlasBlock];
  __shared__ double shc[kBlasBlock];
  sh[threadIdx.x] = s;
  shc[threadIdx.x] = c;
  __syncthreads();
  for (int h = kBlasBlock / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      double r;
      two_sum(sh[threadIdx.x], sh[threadIdx.x + h], sh[threadIdx.x], r);
      shc[threadIdx.x] += shc[threadIdx.x + h] + r;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0];
    part[2 * blockIdx.x + 1] = shc[0];
  }
}

// out[0] + out[1] = compensated sum of nblk (s, c) pairs, out[0] the sum
// rounded to FP64 (one 256-thread block)
__global__ void __launch_bounds__(256) k_reduce_dd(const double* __restrict__ part, int nblk, double* out)
{
  double s = 0.0, c = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    double r;
    two_sum(s, part[2 * b], s, r);
    c += part[2 * b + 1] + r;
  }
  __shared__ double sh[256];
  __shared__ double shc[256];
  sh[threadIdx.x] = s;
  shc[threadIdx.x] = c;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      double r;
      two_sum(sh[threadIdx.x], sh[threadIdx.x + h], sh[threadIdx.x], r);
      shc[threadIdx.x] += shc[threadIdx.x + h] + r;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double hi, lo;
    two_sum(sh[0], shc[0], hi, lo);
    out[0] = hi;
    out[1] = lo;
  }
}

// y = a*x + y
template <class S>
__global__ void __launch_bounds__(kBlasBlock) k_axpy(double a, const S* __restrict__ x, S* __restrict__ y, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock)
    y[i] = slot_axpy(a, x[i], y[i]);
}

// y = x + a*y
template <class S>
__global__ void __launch_bounds__(kBlasBlock) k_xpay(const S* __restrict__ x, double a, S* __restrict__ y, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock)
    y[i] = slot_xpay(x[i], a, y[i]);
}

// x += a p ; r -= a ap ; partial |r|^2
template <class S>
__global__ void __launch_bounds__(kBlasBlock) k_cg_update(double a, const S* __restrict__ p, const S* __restrict__ ap,
                                                          S* __restrict__ x, S* __restrict__ r, int64_t n,
                                                          double* part)
{
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock) {
    x[i] = slot_axpy(a, p[i], x[i]);
    const S rn = slot_axpy(-a, ap[i], r[i]);
    r[i] = rn;
    v[0] += slot_norm2(rn);
  }
  block_reduce_store<1>(v, part);
}

// r -= a ap ; partial |r|^2
template <class S>
__global__ void __launch_bounds__(kBlasBlock) k_axpy_norm(double a, const S* __restrict__ ap, S* __restrict__ r,
                                                          int64_t n, double* part)
{
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock) {
    const S rn = slot_axpy(-a, ap[i], r[i]);
    r[i] = rn;
    v[0] += slot_norm2(rn);
  }
  block_reduce_store<1>(v, part);
}

template <class S>
__global__ void __launch_bounds__(kBlasBlock) k_axpy_norm_multi(AxpyNormArgs<S> a)
{
  const int ci = blockIdx.y;
  const int64_t coff = (int64_t)pick(a.cols.rhs, ci) * a.cols.fstride;
  const S* __restrict__ ap = a.ap + coff;
  S* __restrict__ r = a.r + coff;
  const double al = pick(a.alpha, ci);
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlasBlock) {
    const S rn = slot_axpy(-al, ap[i], r[i]);
    r[i] = rn;
    v += slot_norm2(rn);
  }
  __shared__ double sh[kBlasBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kBlasBlock / 64; ++i) t += sh[i];
    a.part[(int64_t)ci * gridDim.x + blockIdx.x] = t;
  }
}

// out[row] = sum of part[row*nblk + 0 .. nblk)
// One 1024-thread block per row (4 independent partial sums per thread, so
// several loads are in flight); the summation order depends only on nblk.
__global__ void __launch_bounds__(1024) k_reduce_rows(const double* __restrict__ part, int nblk, double* out)
{
  __shared__ double sh[16];
  const double* p = part + (int64_t)blockIdx.x * nblk;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int b = threadIdx.x;
  for (; b + 3 * 1024 < nblk; b += 4 * 1024) {
    v0 += p[b];
    v1 += p[b + 1024];
    v2 += p[b + 2 * 1024];
    v3 += p[b + 3 * 1024];
  }
  for (; b < nblk; b += 1024) v0 += p[b];
  double v = (v0 + v1) + (v2 + v3);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < 16; ++w) t += sh[w];
    out[blockIdx.x] = t;
  }
}

// Sum nblk partial rows of nv values (nv <= 2) with one 256-thread block;
// the summation order depends only on nblk, so results are reproducible.
__global__ void __launch_bounds__(256) k_reduce_final(const double* __restrict__ part, int nblk, int nv, double* out)
{
  __shared__ double sh[4][2];
  double v[2] = {0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += 256)
    for (int i = 0; i < nv; ++i) v[i] += part[(int64_t)b * nv + i];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int i = 0; i < 2; ++i) {
    double t = v[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
    if (lane == 0) sh[wid][i] = t;
  }
  __syncthreads();
  if (threadIdx.x < nv) out[threadIdx.x] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

// FP64 blocked field (12 double2 slots per 5D site) <-> FP32 blocked field
// (6 float4 slots per 5D site); i enumerates FP32 slots.
__global__ void __launch_bounds__(kBlasBlock) k_d2f(const double2* __restrict__ in, float4* __restrict__ out,
                                                    int64_t nslot32, double scale)
{
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < nslot32; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t lane = i & 63, q = i >> 6;  // q = (hb*ls + s)*6 + kp
    const int64_t hs = q / 6, kp = q - hs * 6;
    const int64_t o = ((hs * 12 + 2 * kp) << 6) + lane;
    const double2 a = in[o];
    const double2 b = in[o + 64];
    out[i] = make_float4((float)(scale * a.x), (float)(scale * a.y), (float)(scale * b.x), (float)(scale * b.y));
  }
}

// out64 (+)= scale * in32 : acc=0 overwrite, acc=1 add
__global__ void __launch_bounds__(kBlasBlock) k_f2d(const float4* __restrict__ in, double2* __restrict__ out,
                                                    int64_t nslot32, int acc, double scale)
{
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < nslot32; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t lane = i & 63, q = i >> 6;
    const int64_t hs = q / 6, kp = q - hs * 6;
    const int64_t o = ((hs * 12 + 2 * kp) << 6) + lane;
    const float4 f = in[i];
    const double2 a0 = make_double2(scale * f.x, scale * f.y), a1 = make_double2(scale * f.z, scale * f.w);
    double2* o0 = &out[o];
    double2* o1 = &out[o + 64];
    if (acc) {
      const double2 p0 = *o0, p1 = *o1;
      *o0 = make_double2(p0.x + a0.x, p0.y + a0.y);
      *o1 = make_double2(p1.x + a1.x, p1.y + a1.y);
    } else {
      *o0 = a0;
      *o1 = a1;
    }
  }
}

// Plain copy (reads n slots, writes n slots): the reference point for
// converting HBM traffic counters into bytes.
template <class S>
__global__ void __launch_bounds__(kBlasBlock) k_copy(const S* __restrict__ x, S* __restrict__ y, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock)
    y[i] = x[i];
}

static int blas_grid(int64_t n)
{
  int64_t g = (n + kBlasBlock - 1) / kBlasBlock;
  if (g > kMaxBlasGrid) g = kMaxBlasGrid;
  if (g < 1) g = 1;
  return (int)g;
}

template <class S>
void launch_norm2(const S* a, int64_t n, double* part, double* out, hipStream_t st)
{
  const int g = blas_grid(n);
  hipLaunchKernelGGL(k_norm2<S>, dim3(g), dim3(kBlasBlock), 0, st, a, n, part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, part, g, 1, out);
}
template <class S>
void launch_dot(const S* a, const S* b, int64_t n, double* part, double* out, hipStream_t st)
{
  const int g = blas_grid(n);
  hipLaunchKernelGGL(k_dot<S>, dim3(g), dim3(kBlasBlock), 0, st, a, b, n, part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, part, g, 2, out);
}
void launch_action_dd(const double2* phi, const double2* x, const double2* y, int64_t n, double* part, double* out,
                      hipStream_t st)
{
  const int g = blas_grid(n);
  hipLaunchKernelGGL(k_action_dd, dim3(g), dim3(kBlasBlock), 0, st, phi, x, y, n, part);
  hipLaunchKernelGGL(k_reduce_dd, dim3(1), dim3(256), 0, st, part, g, out);
}
template <class S>
void launch_axpy(double a, const S* x, S* y, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_axpy<S>, dim3(blas_grid(n)), dim3(kBlasBlock), 0, st, a, x, y, n);
}
template <class S>
void launch_xpay(const S* x, double a, S* y, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_xpay<S>, dim3(blas_grid(n)), dim3(kBlasBlock), 0, st, x, a, y, n);
}
template <class S>
void launch_cg_update(double a, const S* p, const S* ap, S* x, S* r, int64_t n, double* part, double* out,
                      hipStream_t st)
{
  const int g = blas_grid(n);
  hipLaunchKernelGGL(k_cg_update<S>, dim3(g), dim3(kBlasBlock), 0, st, a, p, ap, x, r, n, part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, part, g, 1, out);
}

void launch_copy(const float4* x, float4* y, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_copy<float4>, dim3(blas_grid(n)), dim3(kBlasBlock), 0, st, x, y, n);
}

template <class S>
void launch_axpy_norm(double a, const S* ap, S* r, int64_t n, double* part, double* out, hipStream_t st)
{
  const int g = blas_grid(n);
  hipLaunchKernelGGL(k_axpy_norm<S>, dim3(g), dim3(kBlasBlock), 0, st, a, ap, r, n, part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, part, g, 1, out);
}
void launch_reduce_rows(const double* part, int nblk, int nrows, double* out, hipStream_t st)
{
  hipLaunchKernelGGL(k_reduce_rows, dim3(nrows), dim3(1024), 0, st, part, nblk, out);
}
template <class S>
int launch_axpy_norm_multi(const AxpyNormArgs<S>& a, double* out, hipStream_t st)
{
  const int g = blas_grid(a.n);
  hipLaunchKernelGGL(k_axpy_norm_multi<S>, dim3(g, a.cols.n), dim3(kBlasBlock), 0, st, a);
  hipLaunchKernelGGL(k_reduce_rows, dim3(a.cols.n), dim3(1024), 0, st, a.part, g, out);
  return g;
}
template int launch_axpy_norm_multi<double2>(const AxpyNormArgs<double2>&, double*, hipStream_t);
template int launch_axpy_norm_multi<float4>(const AxpyNormArgs<float4>&, double*, hipStream_t);

void launch_reduce(const double* part, int nblk, int nv, double* out, hipStream_t st)
{
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, part, nblk, nv, out);
}
template void launch_axpy_norm<double2>(double, const double2*, double2*, int64_t, double*, double*, hipStream_t);
template void launch_axpy_norm<float4>(double, const float4*, float4*, int64_t, double*, double*, hipStream_t);

template void launch_norm2<double2>(const double2*, int64_t, double*, double*, hipStream_t);
template void launch_norm2<float4>(const float4*, int64_t, double*, double*, hipStream_t);
template void launch_dot<double2>(const double2*, const double2*, int64_t, double*, double*, hipStream_t);
template void launch_dot<float4>(const float4*, const float4*, int64_t, double*, double*, hipStream_t);
template void launch_axpy<double2>(double, const double2*, double2*, int64_t, hipStream_t);
template void launch_axpy<float4>(double, const float4*, float4*, int64_t, hipStream_t);
template void launch_xpay<double2>(const double2*, double, double2*, int64_t, hipStream_t);
template void launch_xpay<float4>(const float4*, double, float4*, int64_t, hipStream_t);
template void launch_cg_update<double2>(double, const double2*, const double2*, double2*, double2*, int64_t,
                                        double*, double*, hipStream_t);
template void launch_cg_update<float4>(double, const float4*, const float4*, float4*, float4*, int64_t, double*,
                                       double*, hipStream_t);

void launch_d2f(const double2* in, float4* out, int ls, int64_t st, hipStream_t s, double scale)
{
  const int64_t n = (int64_t)ls * 6 * st;
  hipLaunchKernelGGL(k_d2f, dim3(blas_grid(n)), dim3(kBlasBlock), 0, s, in, out, n, scale);
}
void launch_f2d(const float4* in, double2* out, int ls, int64_t st, int acc, hipStream_t s, double scale)
{
  const int64_t n = (int64_t)ls * 6 * st;
  hipLaunchKernelGGL(k_f2d, dim3(blas_grid(n)), dim3(kBlasBlock), 0, s, in, out, n, acc, scale);
}

// ============================================================ layouts
// Reference 5D field AoS complex128 [site][s][12] (full, V sites) or
// [half][s][12] (half field).  `full` selects which: for a full field the
// half-site h of parity eo sits at full index of half_coord(h, eo).
__global__ void __launch_bounds__(kBlasBlock) k_pack5(const double2* __restrict__ aos, double2* __restrict__ soa,
                                                      LatGeo g, int ls, int eo, int full)
{
  const int64_t n = g.vh * ls * 12;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock) {
    // i enumerates (s, k, h) with h fastest
    const int64_t h = i % g.vh, q = i / g.vh;
    const int64_t s = q / 12, k = q - s * 12;
    // full field: reference site index; half field: reference half index
    int x[4];
    half_coord(g, h, eo, x);
    const int64_t site = full ? full_index(g, x) : full_index(g, x) / 2;
    soa[sidx<12>(ls, (int)s, (int)k, h)] = aos[(site * ls + s) * 12 + k];
  }
}

__global__ void __launch_bounds__(kBlasBlock) k_unpack5(const double2* __restrict__ soa, double2* __restrict__ aos,
                                                        LatGeo g, int ls, int eo, int full)
{
  const int64_t n = g.vh * ls * 12;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t h = i % g.vh, q = i / g.vh;
    const int64_t s = q / 12, k = q - s * 12;
    int x[4];
    half_coord(g, h, eo, x);
    const int64_t site = full ? full_index(g, x) : full_index(g, x) / 2;
    aos[(site * ls + s) * 12 + k] = soa[sidx<12>(ls, (int)s, (int)k, h)];
  }
}

void launch_pack5(const double2* aos, double2* soa, const LatGeo& g, int ls, int eo, int full, hipStream_t st)
{
  const int64_t n = g.vh * ls * 12;
  hipLaunchKernelGGL(k_pack5, dim3(blas_grid(n)), dim3(kBlasBlock), 0, st, aos, soa, g, ls, eo, full);
}
void launch_unpack5(const double2* soa, double2* aos, const LatGeo& g, int ls, int eo, int full, hipStream_t st)
{
  const int64_t n = g.vh * ls * 12;
  hipLaunchKernelGGL(k_unpack5, dim3(blas_grid(n)), dim3(kBlasBlock), 0, st, soa, aos, g, ls, eo, full);
}

// Propagator column j ([site][12 rows][12 cols] complex128) -> 5D half field
// of parity eo: the chiral embedding of fermion_field_5d_from_4d
// (qcd-prop.h:96-116) with upper = 0, lower = Ls-1 (qcd-prop.h:157):
// spins 0,1 go to s = 0, spins 2,3 to s = Ls-1, everything else is zero.
// With col < 0 the input is a 4D fermion field [site][12].
__global__ void __launch_bounds__(kBlasBlock) k_prop_col_to_5d(const double2* __restrict__ src, double2* __restrict__ soa,
                                                               LatGeo g, int ls, int eo, int col)
{
  const int64_t n = g.vh * ls * 12;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t h = i % g.vh, q = i / g.vh;
    const int s = (int)(q / 12), k = (int)(q - (int64_t)(q / 12) * 12);
    double2 v = make_double2(0.0, 0.0);
    const bool take = (k < 6) ? (s == 0) : (s == ls - 1);
    if (take) {
      int x[4];
      half_coord(g, h, eo, x);
      const int64_t site = full_index(g, x);
      v = (col >= 0) ? src[(site * 12 + k) * 12 + col] : src[site * 12 + k];
    }
    soa[sidx<12>(ls, s, k, h)] = v;
  }
}

// 5D half field of parity eo -> propagator column j (fermion_field_4d_from_5d,
// qcd-prop.h:118-138, upper = Ls-1, lower = 0, qcd-prop.h:164).
__global__ void __launch_bounds__(kBlasBlock) k_5d_to_prop_col(const double2* __restrict__ soa, double2* __restrict__ dst,
                                                               LatGeo g, int ls, int eo, int col)
{
  const int64_t n = g.vh * 12;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t h = i % g.vh;
    const int k = (int)(i / g.vh);
    const int s = (k < 6) ? ls - 1 : 0;
    int x[4];
    half_coord(g, h, eo, x);
    const int64_t site = full_index(g, x);
    const double2 v = soa[sidx<12>(ls, s, k, h)];
    if (col >= 0) dst[(site * 12 + k) * 12 + col] = v;
    else dst[site * 12 + k] = v;
  }
}

void launch_prop_col_to_5d(const double2* src, double2* soa, const LatGeo& g, int ls, int eo, int col, hipStream_t st)
{
  const int64_t n = g.vh * ls * 12;
  hipLaunchKernelGGL(k_prop_col_to_5d, dim3(blas_grid(n)), dim3(kBlasBlock), 0, st, src, soa, g, ls, eo, col);
}
void launch_5d_to_prop_col(const double2* soa, double2* dst, const LatGeo& g, int ls, int eo, int col, hipStream_t st)
{
  const int64_t n = g.vh * 12;
  hipLaunchKernelGGL(k_5d_to_prop_col, dim3(blas_grid(n)), dim3(kBlasBlock), 0, st, soa, dst, g, ls, eo, col);
}

// Gauge AoS [site][mu][3][3] complex128 -> per-output-parity 8-link layout
// (dwf_common.h), FP64 and FP32 copies.  Backward links are stored already
// adjointed: U_mu(x - mu)^+.
__global__ void __launch_bounds__(kBlasBlock) k_pack_gauge(const double2* __restrict__ gauge, double2* __restrict__ l64,
                                                           float2* __restrict__ l32, unsigned* __restrict__ l16,
                                                           LatGeo g)
{
  const int64_t n = g.vh * 2 * 8;
  for (int64_t i = (int64_t)blockIdx.x * kBlasBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlasBlock) {
    const int64_t h = i % g.vh, q = i / g.vh;
    const int p = (int)(q / 8), d = (int)(q % 8);
    const int eo = (p == 0) ? 2 : 1;
    int x[4];
    half_coord(g, h, eo, x);
    const int mu = d & 3;
    if (d >= 4) x[mu] = wrap_dec(x[mu], g.n[mu]);
    const int64_t site = full_index(g, x);
    const double2* U = &gauge[(site * 4 + mu) * 9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double2 v = (d < 4) ? U[r * 3 + c] : make_double2(U[c * 3 + r].x, -U[c * 3 + r].y);
        const int64_t off = lidx(g.stride, p, d * 9 + r * 3 + c, h);
        l64[off] = v;
        l32[off] = make_float2((float)v.x, (float)v.y);
        if (l16) l16[off] = pack_h2((float)v.x, (float)v.y);
      }
  }
}

void launch_pack_gauge(const double2* gauge, double2* l64, float2* l32, unsigned* l16, const LatGeo& g,
                       hipStream_t st)
{
  const int64_t n = g.vh * 16;
  hipLaunchKernelGGL(k_pack_gauge, dim3(blas_grid(n)), dim3(kBlasBlock), 0, st, gauge, l64, l32, l16, g);
}


// ------------------------------------------------------------ RngState fill
// SHA-256 (FIPS 180-4): the round constants and one compression of a
// 64-byte block onto a state.
__constant__ uint32_t kSha256K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

__device__ __forceinline__ uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

__device__ void sha256_compress(uint32_t h[8], uint32_t w[16])
{
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
  for (int i = 0; i < 64; ++i) {
    uint32_t wi;
    if (i < 16) {
      wi = w[i];
    } else {  // message schedule kept in a rolling window of 16 words
      const uint32_t w15 = w[(i - 15) & 15], w2 = w[(i - 2) & 15];
      const uint32_t s0 = rotr32(w15, 7) ^ rotr32(w15, 18) ^ (w15 >> 3);
      const uint32_t s1 = rotr32(w2, 17) ^ rotr32(w2, 19) ^ (w2 >> 10);
      wi = w[i & 15] + s0 + w[(i - 7) & 15] + s1;
      w[i & 15] = wi;
    }
    const uint32_t S1 = rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25);
    const uint32_t ch = (e & f) ^ (~e & g);
    const uint32_t t1 = k + S1 + ch + kSha256K[i] + wi;
    const uint32_t S0 = rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22);
    const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
    const uint32_t t2 = S0 + mj;
    k = g;
    g = f;
    f = e;
    e = d + t1;
    d = c;
    c = b;
    b = a;
    a = t1 + t2;
  }
  h[0] += a;
  h[1] += b;
  h[2] += c;
  h[3] += d;
  h[4] += e;
  h[5] += f;
  h[6] += g;
  h[7] += k;
}

// decimal digits of v into buf (at most 20), returns the count
__device__ int u64_digits(char* buf, uint64_t v)
{
  char tmp[20];
  int n = 0;
  do {
    tmp[n++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  for (int i = 0; i < n; ++i) buf[i] = tmp[n - 1 - i];
  return n;
}

// One thread per (site g, hash k): the draws at indices index0 + 4k + 1 ..
// + 4 of the state newtype(g) hash "[g,index0 + 4k + 1]" once (rand_gen,
// rng-state.cpp:186-212: digest words 6-7 first, then the cached 4-5, 2-3,
// 0-1).  The final block holds the text, 0x80, zeros and the message length
// in bits (the text is at most 43 bytes, so one block always suffices).
__global__ void k_rng_fill_u(RngFillArgs a)
{
#pragma clang fp contract(off)
  const int64_t hps = (a.n_per_site + 3) / 4;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_sites * hps) return;
  const int64_t g = t / hps, kh = t - g * hps;
  char text[48];
  int len = 0;
  text[len++] = '[';
  len += u64_digits(text + len, (uint64_t)g);
  text[len++] = ',';
  len += u64_digits(text + len, a.index0 + 1 + 4 * (uint64_t)kh);
  text[len++] = ']';
  uint32_t w[16];
  for (int i = 0; i < 16; ++i) {
    uint32_t v = 0;
    for (int j = 0; j < 4; ++j) {
      const int b = 4 * i + j;
      const uint32_t byte = b < len ? (uint32_t)(uint8_t)text[b] : (b == len ? 0x80u : 0u);
      v = (v << 8) | byte;
    }
    w[i] = v;
  }
  const uint64_t bits = (a.num_bytes + (uint64_t)len) * 8;
  w[14] = (uint32_t)(bits >> 32);
  w[15] = (uint32_t)bits;
  uint32_t h[8];
  for (int i = 0; i < 8; ++i) h[i] = a.mid[i];
  sha256_compress(h, w);
  const double fac = 1.0 / 18446744073709551616.0;  // 2^-64
  double* out = a.out + g * a.n_per_site;
  for (int j = 0; j < 4; ++j) {
    const int64_t m = 4 * kh + j;
    if (m >= a.n_per_site) break;
    const int hi = 6 - 2 * j;
    const uint64_t u = ((uint64_t)h[hi] << 32) | h[hi + 1];
    out[m] = (double)u * fac * a.scale + a.lower;
  }
}

void launch_rng_fill_u(const RngFillArgs& a, hipStream_t st)
{
  const int64_t tot = a.n_sites * ((a.n_per_site + 3) / 4);
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_rng_fill_u, dim3((unsigned)((tot + kBlasBlock - 1) / kBlasBlock)), dim3(kBlasBlock), 0, st,
                     a);
}

// The RngState of one site as the reference keeps it (rng-state.h:31-40):
// index, three cached 64-bit words and a spare Gaussian; draws as rand_gen,
// u_rand_gen and g_rand_gen (rng-state.cpp:186-258).
struct SiteRng {
  const uint32_t* mid;
  uint64_t num_bytes, type, index;
  uint64_t cache[3];
  int avail;
  double gauss;
  bool gavail;
};

__device__ uint64_t site_rand(SiteRng& r)
{
  r.index += 1;
  if (r.avail > 0) {
    r.avail -= 1;
    return r.cache[r.avail];
  }
  char text[48];
  int len = 0;
  text[len++] = '[';
  len += u64_digits(text + len, r.type);
  text[len++] = ',';
  len += u64_digits(text + len, r.index);
  text[len++] = ']';
  uint32_t w[16];
  for (int i = 0; i < 16; ++i) {
    uint32_t v = 0;
    for (int j = 0; j < 4; ++j) {
      const int b = 4 * i + j;
      const uint32_t byte = b < len ? (uint32_t)(uint8_t)text[b] : (b == len ? 0x80u : 0u);
      v = (v << 8) | byte;
    }
    w[i] = v;
  }
  const uint64_t bits = (r.num_bytes + (uint64_t)len) * 8;
  w[14] = (uint32_t)(bits >> 32);
  w[15] = (uint32_t)bits;
  uint32_t h[8];
  for (int i = 0; i < 8; ++i) h[i] = r.mid[i];
  sha256_compress(h, w);
  r.cache[0] = ((uint64_t)h[0] << 32) | h[1];
  r.cache[1] = ((uint64_t)h[2] << 32) | h[3];
  r.cache[2] = ((uint64_t)h[4] << 32) | h[5];
  r.avail = 3;
  return ((uint64_t)h[6] << 32) | h[7];
}

__device__ double site_u_rand(SiteRng& r, double upper, double lower)
{
#pragma clang fp contract(off)
  const double fac = 1.0 / 18446744073709551616.0;  // 2^-64
  return (double)site_rand(r) * fac * (upper - lower) + lower;
}

__device__ double site_g_rand(SiteRng& r, double center, double sigma)
{
#pragma clang fp contract(off)
  r.index += 1;
  if (r.gavail) {
    r.gavail = false;
    return r.gauss * sigma + center;
  }
  int num_try = 1;
  double v1, v2, rsq;
  do {  // at most 9999 tries, as in the reference
    v1 = site_u_rand(r, 1.0, -1.0);
    v2 = site_u_rand(r, 1.0, -1.0);
    rsq = v1 * v1 + v2 * v2;
    num_try++;
  } while ((num_try < 10000) && (rsq >= 1.0 || rsq == 0));
  if (num_try > 9999) return 1e10;
  const double fac = sqrt(-2.0 * log(rsq) / rsq);
  r.gauss = v1 * fac;
  r.gavail = true;
  return v2 * fac * sigma + center;
}

__global__ void k_rng_fill_g(RngFillArgs a)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.n_sites) return;
  SiteRng r;
  r.mid = a.mid;
  r.num_bytes = a.num_bytes;
  r.type = (uint64_t)g;
  r.index = a.index0;
  r.cache[0] = r.cache[1] = r.cache[2] = 0;
  r.avail = 0;
  r.gauss = 0.0;
  r.gavail = false;
  double* out = a.out + g * a.n_per_site;
  for (int m = 0; m < a.n_per_site; ++m) out[m] = site_g_rand(r, a.lower, a.scale);
}

void launch_rng_fill_g(const RngFillArgs& a, hipStream_t st)
{
  if (a.n_sites <= 0) return;
  hipLaunchKernelGGL(k_rng_fill_g, dim3((unsigned)((a.n_sites + kBlasBlock - 1) / kBlasBlock)), dim3(kBlasBlock), 0,
                     st, a);
}

}  // namespace qdwf
