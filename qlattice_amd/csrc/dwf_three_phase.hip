// gfx950 three-phase kernels of the domain-wall library: the multi-RHS
// hop + s-matrix (k_hop_m5s) for FP32 and FP16 (per-site scale) spinor
// storage, and the CG direction update that feeds it (k_cg_step_s), with
// their s-matrix phases (matrix cores or vector units).  Split from
// dwf_kernels.hip so the two translation units compile in parallel.
#include "dwf_internal.h"

namespace qdwf {

// ============================================== three-phase hop + M5
// Variant of k_hop_m5 for any spinor storage ST (FP32, or FP16 with a
// per-site scale).  Phase 1 as in k_hop_m5: thread (site, s) computes
// (H in)(site, s) and parks it in LDS.  Phase 2: out(s) = sum_t Mat[s][t]
// (H in)(site, t) for all 12 components, on the matrix cores for blocks of
// 16 or 8 sites, or row by row (thread (site, s) forms its own row from
// LDS) for blocks of 32 sites.  Phase 3 adds beta * x, stores the whole 5D
// site (so a per-site scale sees all 24 reals) and accumulates the CG dot
// products from the stored values.  SMAT = false skips the LDS round trip
// (plain hop, out = alpha H in + beta x; selected by a null mats[0]).

// A block's share of a staging copy global -> LDS: element f = tid + i *
// blockDim for i < N (f < total).  The N loads are all issued before the
// first LDS store (a loop with a run-time trip count would wait for each
// load before issuing the next).  load(f) returns the value, store(f, v)
// writes it; the caller synchronises.
template <int N, class T, class Load, class Store>
__device__ __forceinline__ void stage_copy(int total, Load&& load, Store&& store)
{
  T v[N];
  const int bd = blockDim.x, t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int f = t + i * bd;
    if (f < total) v[i] = load(f);
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int f = t + i * bd;
    if (f < total) store(f, v[i]);
  }
}

// The block's s-matrices (both chiralities, [2][ls][ls] on the device) into
// LDS; every thread of the block takes part, the caller synchronises.
template <int LSMAX>
__device__ __forceinline__ void stage_km(cplx<float> (*km)[LSMAX][LSMAX], const cplx<float>* __restrict__ M, int ls)
{
  // compile-time divisors: (chi, r, t) of entry i of the LDS copy
  for (int i = threadIdx.x; i < 2 * LSMAX * LSMAX; i += blockDim.x) {
    const int chi = i / (LSMAX * LSMAX), r = (i / LSMAX) % LSMAX, t = i % LSMAX;
    if (r < ls && t < ls) km[chi][r][t] = M[(chi * ls + r) * ls + t];
  }
}

// Row s of the s-matrix product for one site: v = sum_t K[chi][s][t] h(t),
// h(t) = hr[t][.][site] (6 float4 = 12 complex; chirality 0 for complex
// 0..5, 1 for 6..11).  Thread (site, s) forms its own output row.
template <int SB, int LSMAX, bool REAL>
__device__ __forceinline__ void smat_row(const float4 (*hr)[6][SB], const cplx<float> (*km)[LSMAX][LSMAX], int ls,
                                         int site, int s, cplx<float> v[12])
{
  float4 o[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) o[k] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int t = 0; t < LSMAX; ++t) {
    if (t < ls) {
      const cplx<float> k0 = km[0][s][t], k1 = km[1][s][t];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const cplx<float> m = k < 3 ? k0 : k1;
        const float4 q = hr[t][k][site];
        o[k].x += m.re * q.x;
        o[k].y += m.re * q.y;
        o[k].z += m.re * q.z;
        o[k].w += m.re * q.w;
        if (!REAL) {
          o[k].x -= m.im * q.y;
          o[k].y += m.im * q.x;
          o[k].z -= m.im * q.w;
          o[k].w += m.im * q.z;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    v[2 * k] = cmk<float>(o[k].x, o[k].y);
    v[2 * k + 1] = cmk<float>(o[k].z, o[k].w);
  }
}

// The s-matrix phase of a block of 16 sites x ls rows (ls a multiple of 4,
// ls/4 waves) on the matrix cores: out(s) = sum_t K_chi[s][t] h(t) is, per
// chirality and real component r of the 12 of that chirality, the product
// of the matrix (rows s, in tiles of 16; columns t) with the (t x 16 sites)
// block of that component, done as ls/4 steps of v_mfma_f32_16x16x4_f32
// (full FP32 products and sums).  Lane l of a wave
// supplies A[s = l%16][t = 4 kb + l/16] (the matrix) and B[t = 4 kb + l/16]
// [site = l%16] (one real of hr), and receives D[s = 4 (l/16) + i][site =
// l%16], i < 4.  A float4 of hr (reals 4k..4k+3: complex 2k, 2k+1, one
// chirality) feeds four products at once; wave w takes the float4 columns
// k = w, w + ls/4, ... < 6.  The results are written back over hr (after
// every wave has read it), and thread (site, row) then reads its own row.
// km(chi, s, t): the block's matrices, staged in LDS (complex, or their
// real parts for real matrices).
// Complex (zMoebius) matrices take a second product per real component:
// im K times the partner real of the same complex number (-im K against
// im h for the real part, +im K against re h for the imaginary part).
typedef float mf4 __attribute__((ext_vector_type(4)));

template <int LSMAX, bool REAL, class KM>
__device__ __forceinline__ void smat_mfma16(float4 (*hr)[6][16], KM km, int ls, int site,
                                            int row, bool act, cplx<float> v[12])
{
  constexpr int MT = (LSMAX + 15) / 16;  // 16-row tiles of s
  constexpr int KB = LSMAX / 4;          // 4-column steps of t
  // float4 columns per wave: the dispatch gives this instance ls = 8
  // (2 waves) for LSMAX 8, else ls > 8 (at least 3 waves)
  constexpr int UMAX = LSMAX <= 8 ? 3 : 2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, lq = lane >> 4;
  const int nw = ls >> 2, nkb = ls >> 2;
  __syncthreads();  // hr (and the matrices) complete
  // matrix operands of this lane: K_chi[mt*16 + li][4 kb + lq]
  float are[2][MT][KB], aim[2][MT][KB];
#pragma unroll
  for (int chi = 0; chi < 2; ++chi)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        const int s = mt * 16 + li, t = 4 * kb + lq;
        cplx<float> m = cmk<float>(0.f, 0.f);
        if (kb < nkb && s < ls) m = km(chi, s, t);
        are[chi][mt][kb] = m.re;
        aim[chi][mt][kb] = m.im;
      }
  mf4 d[UMAX][MT][4];
#pragma unroll
  for (int u = 0; u < UMAX; ++u) {
    const int k = wv + u * nw;
    if (k < 6) {  // wave-uniform
      const int chi = k / 3;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int j = 0; j < 4; ++j) d[u][mt][j] = (mf4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        if (kb < nkb) {
          const float4 b = hr[4 * kb + lq][k][li];
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) {
            const float ar = chi ? are[1][mt][kb] : are[0][mt][kb];
            d[u][mt][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b.x, d[u][mt][0], 0, 0, 0);
            d[u][mt][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b.y, d[u][mt][1], 0, 0, 0);
            d[u][mt][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b.z, d[u][mt][2], 0, 0, 0);
            d[u][mt][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b.w, d[u][mt][3], 0, 0, 0);
            if (!REAL) {
              const float ai = chi ? aim[1][mt][kb] : aim[0][mt][kb];
              d[u][mt][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, b.y, d[u][mt][0], 0, 0, 0);
              d[u][mt][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b.x, d[u][mt][1], 0, 0, 0);
              d[u][mt][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, b.w, d[u][mt][2], 0, 0, 0);
              d[u][mt][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b.z, d[u][mt][3], 0, 0, 0);
            }
          }
        }
      }
    }
  }
  __syncthreads();  // every wave has read hr
#pragma unroll
  for (int u = 0; u < UMAX; ++u) {
    const int k = wv + u * nw;
    if (k < 6) {
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int s = mt * 16 + 4 * lq + i;
          if (s < ls) hr[s][k][li] = make_float4(d[u][mt][0][i], d[u][mt][1][i], d[u][mt][2][i], d[u][mt][3][i]);
        }
    }
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float4 q = hr[row][k][site];
      v[2 * k] = cmk<float>(q.x, q.y);
      v[2 * k + 1] = cmk<float>(q.z, q.w);
    }
  }
}

// The same for a block of 8 sites x ls rows (ls = 24: 3 waves): a product
// covers 8 sites x 2 reals, column n = l%16 of lane l being site n%8 and
// real 2j + n/8 of the float4 column k (j = 0, 1: two products per float4).
// For complex matrices the second product takes im K against the partner
// real of the same complex number, its sign folded into the operand (-im h
// for the real part, +re h for the imaginary part).
__device__ __forceinline__ float f4c(const float4& q, int c)
{
  return c == 0 ? q.x : c == 1 ? q.y : c == 2 ? q.z : q.w;
}
template <int LSMAX, bool REAL, class KM>
__device__ __forceinline__ void smat_mfma8(float4 (*hr)[6][8], KM km, int ls, int site,
                                           int row, bool act, cplx<float> v[12])
{
  constexpr int MT = (LSMAX + 15) / 16;
  constexpr int KB = LSMAX / 4;
  // float4 columns per wave: 6 over ls/8 waves (2 waves at Ls 16, 3 at 24)
  constexpr int UMAX = LSMAX <= 16 ? 3 : 2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, lq = lane >> 4;
  const int ns = li & 7, rr = li >> 3;  // this lane's site and real within the pair
  const int nw = ls >> 3, nkb = ls >> 2;
  __syncthreads();  // hr (and the matrices) complete
  float are[2][MT][KB], aim[2][MT][KB];
#pragma unroll
  for (int chi = 0; chi < 2; ++chi)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        const int s = mt * 16 + li, t = 4 * kb + lq;
        cplx<float> m = cmk<float>(0.f, 0.f);
        if (kb < nkb && s < ls) m = km(chi, s, t);
        are[chi][mt][kb] = m.re;
        aim[chi][mt][kb] = m.im;
      }
  mf4 d[UMAX][MT][2];
#pragma unroll
  for (int u = 0; u < UMAX; ++u) {
    const int k = wv + u * nw;
    if (k < 6) {  // wave-uniform
      const int chi = k / 3;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int j = 0; j < 2; ++j) d[u][mt][j] = (mf4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        if (kb < nkb) {
          const float4 q = hr[4 * kb + lq][k][ns];
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const float b = f4c(q, 2 * j + rr);
            const float bx = rr ? f4c(q, 2 * j) : -f4c(q, 2 * j + 1);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
              const float ar = chi ? are[1][mt][kb] : are[0][mt][kb];
              d[u][mt][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b, d[u][mt][j], 0, 0, 0);
              if (!REAL) {
                const float ai = chi ? aim[1][mt][kb] : aim[0][mt][kb];
                d[u][mt][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, bx, d[u][mt][j], 0, 0, 0);
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();  // every wave has read hr
#pragma unroll
  for (int u = 0; u < UMAX; ++u) {
    const int k = wv + u * nw;
    if (k < 6) {
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int s = mt * 16 + 4 * lq + i;
          if (s < ls) {
            float* p = reinterpret_cast<float*>(&hr[s][k][ns]);
            p[rr] = d[u][mt][0][i];
            p[2 + rr] = d[u][mt][1][i];
          }
        }
    }
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float4 q = hr[row][k][site];
      v[2 * k] = cmk<float>(q.x, q.y);
      v[2 * k + 1] = cmk<float>(q.z, q.w);
    }
  }
}

// streaming store that also returns the stored values
template <class ST>
__device__ __forceinline__ void store_qs(typename SpinorIO<ST>::slot_t* __restrict__ f, int ls, int s, int64_t h,
                                         cplx<float> v[12])
{
  if constexpr (std::is_same<ST, Half16>::value) SpinorIO<Half16>::put<true>(f, ls, s, h, v);
  else SpinorIO<ST>::store_stream(f, ls, s, h, v);
}

// HM: 0 FP32 arithmetic in the hop phase (8 neighbour spinors in flight per
// thread, a register budget of 3 waves per SIMD); 1 FP16 arithmetic
// (hop_term_h2) with 4 neighbour spinors in flight and 4 waves per SIMD.
// Measured at 24^3x64 Ls12 over 4 columns (b = K H t): 1.177 ms and
// 1.063 ms; FP16 arithmetic with 8 in flight needs 129 registers (3
// waves per SIMD, 1.28 ms), or spills when held to 128.
template <int HM>
struct HopMode {
  static constexpr bool h2 = HM != 0;
  static constexpr int nb = HM ? 4 : 8;
  static constexpr int wpe = HM ? 4 : 3;
};

template <class ST, int SB, int LSMAX, bool DAG, bool XPAY, bool REAL, bool DOT, bool SMAT, int HM>
__global__ void __launch_bounds__(SB * LSMAX) __attribute__((amdgpu_waves_per_eu(HopMode<HM>::wpe)))
k_hop_m5s(HopM5Args a)
{
  typedef typename SpinorIO<ST>::slot_t slot_t;
  constexpr bool H2 = HopMode<HM>::h2;
  static_assert(!H2 || std::is_same<ST, Half16>::value, "FP16 arithmetic needs FP16 storage");
  // links of the block's sites: FP32 complex, or (H2) the FP16 forms
  // la = (re, -im), lb = (im, re) of hop_term_h2
  __shared__ float2 sl[H2 ? 1 : 72][H2 ? 1 : SB];
  // With FP16 arithmetic and an s-matrix phase the links and the hop
  // results (hr below) share one LDS area; the hop phase ends at a barrier
  // before the results go in.  With the real matrices' real parts staged
  // alone (kmr below) a 16-site block at Ls 12 takes 20 KB instead of 30 KB:
  // 8 blocks per CU instead of 5
  constexpr bool ALIAS = H2 && SMAT;
  constexpr unsigned kLinkBytes = 2u * 72u * SB * 4u;
  constexpr unsigned kHrBytes = (unsigned)LSMAX * 6u * SB * 16u;
  __shared__ __attribute__((aligned(16))) unsigned char pool[ALIAS ? (kLinkBytes > kHrBytes ? kLinkBytes : kHrBytes) : 16];
  constexpr int LO = H2 && !ALIAS ? 72 : 1, LS_ = H2 && !ALIAS ? SB : 1;
  __shared__ h16x2 sla_own[LO][LS_], slb_own[LO][LS_];
  h16x2(*sla)[SB] = ALIAS ? (h16x2(*)[SB])pool : (h16x2(*)[SB])sla_own;
  h16x2(*slb)[SB] = ALIAS ? (h16x2(*)[SB])(pool + 72u * SB * 4u) : (h16x2(*)[SB])slb_own;
  // block numbers are below 2^32: 32-bit divisions
  const unsigned lbc = (unsigned)xcd_logical_block(blockIdx.x, (int64_t)gridDim.x);
  const int ncol = a.cols.n;
  const unsigned nblk = gridDim.x / (unsigned)ncol;
  // column-fastest (the columns of a site block back to back: link reuse)
  // or site-fastest (one column's sites back to back: neighbour reuse)
  const unsigned lb = a.site_major ? lbc % nblk : lbc / (unsigned)ncol;
  const int ci = a.site_major ? (int)(lbc / nblk) : (int)(lbc - lb * (unsigned)ncol);
  if (a.cg && a.cg->mode[ci] != 0) return;  // frozen column (the whole block)
  const int64_t h0 = (int64_t)lb * SB;
  const int pslot = (a.eo_out == 2) ? 0 : 1;
  // 72 SB link elements over SB ls >= 6 SB threads: at most 12 each
  if constexpr (H2) {
    const unsigned* __restrict__ l16 = (const unsigned*)a.links16 + lidx(a.geo.stride, pslot, 0, h0);
    stage_copy<12, unsigned>(
        72 * SB, [&](int i) { return l16[(i / SB) * 64 + (i % SB)]; },
        [&](int i, unsigned w) {
          const h16x2 u = __builtin_bit_cast(h16x2, w);
          sla[i / SB][i % SB] = (h16x2){u.x, -u.y};
          slb[i / SB][i % SB] = (h16x2){u.y, u.x};
        });
  } else {
    const float2* __restrict__ links = (const float2*)a.links + lidx(a.geo.stride, pslot, 0, h0);
    stage_copy<12, float2>(
        72 * SB, [&](int i) { return links[(i / SB) * 64 + (i % SB)]; },
        [&](int i, float2 w) { sl[i / SB][i % SB] = w; });
  }
  // the block's s-matrices (both chiralities) in LDS, loaded while the
  // links and neighbour offsets are staged
  // (real matrices on the matrix-core paths: their real parts only, 1.2 KB
  // less per block at Ls 12, which lets 8 blocks of 16 sites share a CU)
  constexpr bool KMR = SMAT && REAL && (SB == 16 || SB == 8);
  constexpr int KC = SMAT && !KMR ? LSMAX : 1, KRN = KMR ? LSMAX : 1;
  __shared__ cplx<float> km[SMAT && !KMR ? 2 : 1][KC][KC];
  __shared__ float kmr[KMR ? 2 : 1][KRN][KRN];
  if constexpr (KMR) {
    const cplx<float>* __restrict__ M = pick(a.mats, ci);
    for (int i = threadIdx.x; i < 2 * LSMAX * LSMAX; i += blockDim.x) {
      const int chi = i / (LSMAX * LSMAX), r = (i / LSMAX) % LSMAX, t = i % LSMAX;
      if (r < a.ls && t < a.ls) kmr[chi][r][t] = M[(chi * a.ls + r) * a.ls + t].re;
    }
  } else if constexpr (SMAT) {
    stage_km<LSMAX>(km, pick(a.mats, ci), a.ls);
  }
  // neighbour indices once per site, shared by the Ls rows (thread (site,
  // d) of the first 8 SB forms neighbour d of its site)
  __shared__ unsigned nbs[8][SB];
  for (int i = threadIdx.x; i < 8 * SB; i += blockDim.x) {
    const int st = i % SB, d = i / SB;
    if (h0 + st < a.geo.vh) nbs[d][st] = (unsigned)neighbour_of(a.geo, h0 + st, a.eo_out, d);
  }
  __syncthreads();
  const int ls = a.ls;
  const int site = threadIdx.x % SB, row = threadIdx.x / SB;
  const int64_t h = h0 + site;
  const bool active = h < a.geo.vh;
  const int64_t coff = (int64_t)pick(a.cols.rhs, ci) * a.cols.fstride;
  cplx<float> acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = cmk<float>(0, 0);
  if (active && row < ls) {
    unsigned nb[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) nb[d] = nbs[d][site];
    const slot_t* __restrict__ in = (const slot_t*)a.in + coff;
    constexpr int SF = DAG ? 1 : -1;
    constexpr int SB_ = -SF;
    // neighbour spinors are loaded in batches of NB directions before any
    // of them is used, so that each wave keeps several gathers in flight
    constexpr int NB = std::is_same<ST, Half16>::value ? HopMode<HM>::nb : 4;
    typename SpinorIO<ST>::Raw raw[NB];
#define QDWF_LD(J, D) SpinorIO<ST>::load_raw(in, ls, row, nb[D], raw[J]);
#define QDWF_CP(J, D, MU, SG)                                       \
  if constexpr (H2) {                                               \
    h16x2 la[9], lb[9], hp[12];                                     \
    _Pragma("unroll") for (int e = 0; e < 9; ++e)                   \
    {                                                               \
      la[e] = sla[(D) * 9 + e][site];                               \
      lb[e] = slb[(D) * 9 + e][site];                               \
    }                                                               \
    _Pragma("unroll") for (int j = 0; j < 3; ++j)                   \
    {                                                               \
      hp[4 * j + 0] = __builtin_bit_cast(h16x2, raw[J].q[j].x);     \
      hp[4 * j + 1] = __builtin_bit_cast(h16x2, raw[J].q[j].y);     \
      hp[4 * j + 2] = __builtin_bit_cast(h16x2, raw[J].q[j].z);     \
      hp[4 * j + 3] = __builtin_bit_cast(h16x2, raw[J].q[j].w);     \
    }                                                               \
    hop_term_h2<MU, SG>(la, lb, hp, raw[J].sc, acc);                \
  } else {                                                          \
    cplx<float> W[9];                                               \
    _Pragma("unroll") for (int e = 0; e < 9; ++e)                   \
    {                                                               \
      const float2 q = sl[(D) * 9 + e][site];                       \
      W[e] = cmk<float>(q.x, q.y);                                  \
    }                                                               \
    cplx<float> psi[12];                                            \
    SpinorIO<ST>::decode(raw[J], psi);                              \
    hop_term<MU, SG>(W, psi, acc);                                  \
  }
    if constexpr (NB == 8) {
      QDWF_LD(0, 0) QDWF_LD(1, 1) QDWF_LD(2, 2) QDWF_LD(3, 3)
      QDWF_LD(4, 4) QDWF_LD(5, 5) QDWF_LD(6, 6) QDWF_LD(7 % NB, 7)
      QDWF_CP(0, 0, 0, SF) QDWF_CP(1, 1, 1, SF) QDWF_CP(2, 2, 2, SF) QDWF_CP(3, 3, 3, SF)
      QDWF_CP(4 % NB, 4, 0, SB_) QDWF_CP(5 % NB, 5, 1, SB_) QDWF_CP(6 % NB, 6, 2, SB_) QDWF_CP(7 % NB, 7, 3, SB_)
    } else {
      QDWF_LD(0, 0) QDWF_LD(1, 1) QDWF_LD(2, 2) QDWF_LD(3, 3)
      QDWF_CP(0, 0, 0, SF) QDWF_CP(1, 1, 1, SF) QDWF_CP(2, 2, 2, SF) QDWF_CP(3, 3, 3, SF)
      QDWF_LD(0, 4) QDWF_LD(1, 5) QDWF_LD(2, 6) QDWF_LD(3, 7)
      QDWF_CP(0, 4, 0, SB_) QDWF_CP(1, 5, 1, SB_) QDWF_CP(2, 6, 2, SB_) QDWF_CP(3, 7, 3, SB_)
    }
#undef QDWF_LD
#undef QDWF_CP
  }
  const float sc = -0.5f * a.alpha;
  cplx<float> v[12];
  if constexpr (SMAT) {
    __shared__ float4 hr_own[ALIAS ? 1 : LSMAX][ALIAS ? 1 : 6][ALIAS ? 1 : SB];
    float4(*hr)[6][SB] = ALIAS ? (float4(*)[6][SB])pool : (float4(*)[6][SB])hr_own;
    if constexpr (ALIAS) __syncthreads();  // every wave is done with the links
    if (active && row < ls) {
#pragma unroll
      for (int k = 0; k < 6; ++k)
        hr[row][k][site] = make_float4(sc * acc[2 * k].re, sc * acc[2 * k].im, sc * acc[2 * k + 1].re,
                                       sc * acc[2 * k + 1].im);
    }
    if constexpr (SB == 16) {
      // on the matrix cores (ls is a multiple of 4 here: whole waves)
      if constexpr (KMR)
        smat_mfma16<LSMAX, REAL>(
            hr, [&](int c, int r, int t) { return cmk<float>(kmr[c][r][t], 0.f); }, ls, site, row,
            active && row < ls, v);
      else
        smat_mfma16<LSMAX, REAL>(
            hr, [&](int c, int r, int t) { return km[c][r][t]; }, ls, site, row, active && row < ls, v);
    } else if constexpr (SB == 8) {
      if constexpr (KMR)
        smat_mfma8<LSMAX, REAL>(
            hr, [&](int c, int r, int t) { return cmk<float>(kmr[c][r][t], 0.f); }, ls, site, row,
            active && row < ls, v);
      else
        smat_mfma8<LSMAX, REAL>(
            hr, [&](int c, int r, int t) { return km[c][r][t]; }, ls, site, row, active && row < ls, v);
    } else {
      // thread (site, s) forms its own output row s for all 12 components
      __syncthreads();
      if (active && row < ls) smat_row<SB, LSMAX, REAL>(hr, km, ls, site, row, v);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = cmk<float>(sc * acc[k].re, sc * acc[k].im);
  }
  double dot = 0.0, dotr = 0.0, dota = 0.0;
  if (active && row < ls) {
    if (XPAY) {  // (the streamed inputs are read with non-temporal loads)
      cplx<float> xv[12];
      SpinorIO<ST>::load_stream((const slot_t*)a.x + coff, ls, row, h, xv);
#pragma unroll
      for (int k = 0; k < 12; ++k) v[k] = cmk<float>(v[k].re + a.beta * xv[k].re, v[k].im + a.beta * xv[k].im);
    }
    store_qs<ST>((slot_t*)a.out + coff, ls, row, h, v);
    if (DOT) {
      cplx<float> pv[12];
      SpinorIO<ST>::load_stream((const slot_t*)a.dotp + coff, ls, row, h, pv);
#pragma unroll
      for (int k = 0; k < 12; ++k) dot += (double)pv[k].re * v[k].re + (double)pv[k].im * v[k].im;
      if (a.dotr) {
        cplx<float> rv[12];
        SpinorIO<float>::load_stream((const float4*)a.dotr + (int64_t)pick(a.cols.rhs, ci) * a.dotr_fstride, ls,
                                     row, h, rv);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          dotr += (double)rv[k].re * v[k].re + (double)rv[k].im * v[k].im;
          dota += (double)v[k].re * v[k].re + (double)v[k].im * v[k].im;
        }
      }
    }
  }
  if (DOT) {
    __shared__ double red[3][(SB * LSMAX + 63) / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nv = a.dotr ? 3 : 1;
    double vals[3] = {dot, dotr, dota};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (q < nv) {
        double t = vals[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
        if (lane == 0) red[q][wv] = t;
      }
    }
    __syncthreads();
    if (threadIdx.x < nv) {
      double t = 0.0;
      for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) t += red[threadIdx.x][i];
      a.part[((int64_t)threadIdx.x * ncol + ci) * nblk + lb] = t;
    }
  }
}

// plain hop: out = alpha H in + beta x (no s-matrix phase)
template <class ST, int SB, int LSMAX, int H2>
static void launch_hop_plain_ls(const HopM5Args& a, bool dag, hipStream_t st)
{
  const dim3 grid((unsigned)((a.geo.vh + SB - 1) / SB * a.cols.n));
  const dim3 blk(SB * a.ls);
  const bool xpay = a.x != nullptr;
  if (!dag && !xpay) hipLaunchKernelGGL((k_hop_m5s<ST, SB, LSMAX, false, false, true, false, false, H2>), grid, blk, 0, st, a);
  if (!dag && xpay) hipLaunchKernelGGL((k_hop_m5s<ST, SB, LSMAX, false, true, true, false, false, H2>), grid, blk, 0, st, a);
  if (dag && !xpay) hipLaunchKernelGGL((k_hop_m5s<ST, SB, LSMAX, true, false, true, false, false, H2>), grid, blk, 0, st, a);
  if (dag && xpay) hipLaunchKernelGGL((k_hop_m5s<ST, SB, LSMAX, true, true, true, false, false, H2>), grid, blk, 0, st, a);
}

template <class ST, int SB, int LSMAX, int H2>
static void launch_hop_m5s_ls(const HopM5Args& a, bool dag, hipStream_t st)
{
  const dim3 grid((unsigned)((a.geo.vh + SB - 1) / SB * a.cols.n));
  const dim3 blk(SB * a.ls);
  const bool xpay = a.x != nullptr;
#define QDWF_L(D, X, R, DT) hipLaunchKernelGGL((k_hop_m5s<ST, SB, LSMAX, D, X, R, DT, true, H2>), grid, blk, 0, st, a)
  if (!a.mats[0]) {
    launch_hop_plain_ls<ST, SB, LSMAX, H2>(a, dag, st);
    return;
  }
  if (a.dotp) {
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

// phase 2 needs >= 6 rows; whole waves (sb * ls multiple of 64) for the
// block reductions; above Ls 16 (the 24-row instances) only 16 sites per
// block, which keeps a block's LDS (rows x 12 complex x sites) below 40 KB
bool three_phase_ok(int ls, int sb)
{
  return ls >= 6 && ls <= 24 && (sb == 16 || (sb == 32 && ls <= 16) || (sb == 8 && (ls == 16 || ls == 24))) &&
         (sb * ls) % 64 == 0;
}
int64_t hop_m5s_blocks(const LatGeo& g, int sb) { return (g.vh + sb - 1) / sb; }

template <class ST, int SB, int H2>
static void launch_hop_m5s_sb(const HopM5Args& a, bool dag, hipStream_t st)
{
  if constexpr (SB == 8) {  // Ls 16 and 24 only
    if (a.ls == 16) launch_hop_m5s_ls<ST, SB, 16, H2>(a, dag, st);
    else launch_hop_m5s_ls<ST, SB, 24, H2>(a, dag, st);
    return;
  }
  if (a.ls <= 8) launch_hop_m5s_ls<ST, SB, 8, H2>(a, dag, st);
  else if (a.ls <= 12) launch_hop_m5s_ls<ST, SB, 12, H2>(a, dag, st);
  else if (a.ls <= 16) launch_hop_m5s_ls<ST, SB, 16, H2>(a, dag, st);
  else if constexpr (SB == 16) launch_hop_m5s_ls<ST, SB, 24, H2>(a, dag, st);
}

bool launch_hop_m5s(const HopM5Args& a, bool dag, bool half, hipStream_t st)
{
  if (!three_phase_ok(a.ls, a.sb)) return false;
  const bool h2 = half && a.links16 && a.hop_mode != 0;
  if (h2 && a.plain64 && !a.mats[0] && !a.dotp && a.ls <= 16) {
    // The plain hop (no s-matrix phase, no barrier after the hop) on blocks
    // of 64 sites: a wave then holds 64 sites of one row, so that a
    // neighbour gather that is not line-aligned touches 9 cache lines per
    // 16-byte unit of the wave's sites (1 KB) instead of 12 (blocks of 16
    // sites x 4 rows: 4 spans of 256 bytes, 3 lines each).  At
    // 24^3x64 Ls12 over 4 columns: 77.7M L1-to-L2 requests against 98.8M,
    // 8.5% faster.  The s-matrix kernels stay on 16-site blocks: with 32 or
    // 64 sites their block-wide barriers leave fewer requests in flight,
    // and they ran 10% and 40% slower.
    if (a.ls <= 8) launch_hop_plain_ls<Half16, 64, 8, 1>(a, dag, st);
    else if (a.ls <= 12) launch_hop_plain_ls<Half16, 64, 12, 1>(a, dag, st);
    else launch_hop_plain_ls<Half16, 64, 16, 1>(a, dag, st);
    return true;
  }
  if (h2) {
    if (a.sb == 8) launch_hop_m5s_sb<Half16, 8, 1>(a, dag, st);
    else if (a.sb == 16) launch_hop_m5s_sb<Half16, 16, 1>(a, dag, st);
    else launch_hop_m5s_sb<Half16, 32, 1>(a, dag, st);
  } else if (half) {
    if (a.sb == 8) launch_hop_m5s_sb<Half16, 8, 0>(a, dag, st);
    else if (a.sb == 16) launch_hop_m5s_sb<Half16, 16, 0>(a, dag, st);
    else launch_hop_m5s_sb<Half16, 32, 0>(a, dag, st);
  } else {
    if (a.sb == 8) launch_hop_m5s_sb<float, 8, 0>(a, dag, st);
    else if (a.sb == 16) launch_hop_m5s_sb<float, 16, 0>(a, dag, st);
    else launch_hop_m5s_sb<float, 32, 0>(a, dag, st);
  }
  return true;
}

// CG direction update for p, t in storage ST, r in SR and x in SX:
// [r -= alpha_r ap (merged form)]; x += alpha p (UPDX); p = r + beta p;
// t = Mat p, in three phases like
// k_hop_m5s (block = SB sites x Ls, blockIdx.y = column).  p and t are
// computed from the stored (rounded) p.
template <class ST, class SR, class SX, int SB, int LSMAX, bool UPDX, bool REAL>
__global__ void __launch_bounds__(SB * LSMAX) k_cg_step_s(CgStepArgs<float> a)
{
  typedef typename SpinorIO<ST>::slot_t slot_t;
  typedef typename SpinorIO<SR>::slot_t rslot_t;
  typedef typename SpinorIO<SX>::slot_t xslot_t;
  __shared__ float4 hr[LSMAX][6][SB];
  const int ci = blockIdx.y;
  const int ls = a.ls;
  const int site = threadIdx.x % SB, row = threadIdx.x / SB;
  const int64_t h = (int64_t)blockIdx.x * SB + site;
  const bool active = h < a.geo.vh;
  const int col = pick(a.cols.rhs, ci);
  const int64_t coff = (int64_t)col * a.cols.fstride;
  const int64_t roff = (int64_t)col * a.xr_fstride, xoff = (int64_t)col * a.x_fstride;
  // scalars from the kernel arguments, or from the device CG state
  // (block-uniform: the whole block returns together)
  float al, be, ar, aq = 0.f;
  if (a.cg) {
    const int mode = a.cg->mode[ci];
    if (mode == 1) return;
    al = (float)a.cg->alpha_x[ci];
    aq = a.p_out ? (float)a.cg->alpha_q[ci] : 0.f;
    if (mode == 2) {  // reliable-update point: only x += alpha_q q + alpha p
      if (UPDX && active && row < ls) {
        cplx<float> pv[12], xv[12];
        SpinorIO<ST>::load((const slot_t*)a.p + coff, ls, row, h, pv);
        xslot_t* x = (xslot_t*)a.x + xoff;
        SpinorIO<SX>::load(x, ls, row, h, xv);
        if (aq != 0.f) {  // block-uniform
          cplx<float> qv[12];
          SpinorIO<ST>::load((const slot_t*)a.p_out + coff, ls, row, h, qv);
#pragma unroll
          for (int k = 0; k < 12; ++k) xv[k] = cmk<float>(xv[k].re + aq * qv[k].re, xv[k].im + aq * qv[k].im);
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) xv[k] = cmk<float>(xv[k].re + al * pv[k].re, xv[k].im + al * pv[k].im);
        SpinorIO<SX>::store(x, ls, row, h, xv);
      }
      return;
    }
    be = (float)a.cg->beta[ci];
    ar = (float)a.cg->alpha_r[ci];
  } else {
    al = (float)pick(a.alpha, ci);
    be = (float)pick(a.beta, ci);
    ar = (float)pick(a.alpha_r, ci);
    aq = (float)pick(a.alpha_prev, ci);
  }
  slot_t* pout = (slot_t*)(a.p_out ? a.p_out : a.p) + coff;  // may be a.p
  if (active && row < ls) {
    // every field is streamed once (non-temporal loads and stores)
    cplx<float> pv[12], rv[12];
    SpinorIO<ST>::load_stream((const slot_t*)a.p + coff, ls, row, h, pv);
    SpinorIO<SR>::load_stream((const rslot_t*)a.r + roff, ls, row, h, rv);
    if (a.ap) {  // merged form: r -= alpha_r ap
      cplx<float> av[12];
      SpinorIO<ST>::load_stream((const slot_t*)a.ap + coff, ls, row, h, av);
#pragma unroll
      for (int k = 0; k < 12; ++k) rv[k] = cmk<float>(rv[k].re - ar * av[k].re, rv[k].im - ar * av[k].im);
      store_qs<SR>((rslot_t*)a.r + roff, ls, row, h, rv);
    }
    if (UPDX && (al != 0.f || aq != 0.f)) {  // block-uniform
      cplx<float> xv[12];
      xslot_t* x = (xslot_t*)a.x + xoff;
      SpinorIO<SX>::load_stream(x, ls, row, h, xv);
      if (aq != 0.f) {  // the previous iteration's deferred alpha q (q in p_out, not yet replaced)
        cplx<float> qv[12];
        SpinorIO<ST>::load_stream(pout, ls, row, h, qv);
#pragma unroll
        for (int k = 0; k < 12; ++k) xv[k] = cmk<float>(xv[k].re + aq * qv[k].re, xv[k].im + aq * qv[k].im);
      }
#pragma unroll
      for (int k = 0; k < 12; ++k) xv[k] = cmk<float>(xv[k].re + al * pv[k].re, xv[k].im + al * pv[k].im);
      store_qs<SX>(x, ls, row, h, xv);
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) pv[k] = cmk<float>(be * pv[k].re + rv[k].re, be * pv[k].im + rv[k].im);
    store_qs<ST>(pout, ls, row, h, pv);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      hr[row][k][site] = make_float4(pv[2 * k].re, pv[2 * k].im, pv[2 * k + 1].re, pv[2 * k + 1].im);
  }
  __shared__ cplx<float> km[2][LSMAX][LSMAX];
  stage_km<LSMAX>(km, pick(a.mats, ci), ls);
  if constexpr (SB == 16) {
    // t = Mat p on the matrix cores (16-site blocks: whole waves)
    cplx<float> v[12];
    smat_mfma16<LSMAX, REAL>(
        hr, [&](int c, int r, int t) { return km[c][r][t]; }, ls, site, row, active && row < ls, v);
    if (active && row < ls) store_qs<ST>((slot_t*)a.t + coff, ls, row, h, v);
  } else {
    __syncthreads();
    // t = Mat p: thread (site, s) forms its own row s
    if (active && row < ls) {
      cplx<float> v[12];
      smat_row<SB, LSMAX, REAL>(hr, km, ls, site, row, v);
      store_qs<ST>((slot_t*)a.t + coff, ls, row, h, v);
    }
  }
}

template <class ST, class SR, class SX, int SB, int LSMAX>
static void launch_cg_step_s_ls(const CgStepArgs<float>& a, hipStream_t st)
{
  const dim3 grid((unsigned)((a.geo.vh + SB - 1) / SB), a.cols.n);
  const dim3 blk(SB * a.ls);
  const bool updx = a.x != nullptr;
  if (a.real_mat) {
    if (updx) hipLaunchKernelGGL((k_cg_step_s<ST, SR, SX, SB, LSMAX, true, true>), grid, blk, 0, st, a);
    else hipLaunchKernelGGL((k_cg_step_s<ST, SR, SX, SB, LSMAX, false, true>), grid, blk, 0, st, a);
  } else {
    if (updx) hipLaunchKernelGGL((k_cg_step_s<ST, SR, SX, SB, LSMAX, true, false>), grid, blk, 0, st, a);
    else hipLaunchKernelGGL((k_cg_step_s<ST, SR, SX, SB, LSMAX, false, false>), grid, blk, 0, st, a);
  }
}

template <class ST, class SR, class SX, int SB>
static void launch_cg_step_s_sb(const CgStepArgs<float>& a, hipStream_t st)
{
  if (a.ls <= 8) launch_cg_step_s_ls<ST, SR, SX, SB, 8>(a, st);
  else if (a.ls <= 12) launch_cg_step_s_ls<ST, SR, SX, SB, 12>(a, st);
  else if (a.ls <= 16) launch_cg_step_s_ls<ST, SR, SX, SB, 16>(a, st);
  else if constexpr (SB == 16) launch_cg_step_s_ls<ST, SR, SX, SB, 24>(a, st);
}

// 16 sites per block when asked for (or needed: Ls above 16) and Ls allows
// whole waves, else 32.  (8-site blocks at Ls 24 measured slower for the
// step, 6.5 against 6.0 ms at 24^3x64 over 4 columns, unlike the hop.)
template <class ST, class SR, class SX>
static void launch_cg_step_s_st(const CgStepArgs<float>& a, hipStream_t st)
{
  if ((a.sb == 16 || !three_phase_ok(a.ls, 32)) && three_phase_ok(a.ls, 16)) launch_cg_step_s_sb<ST, SR, SX, 16>(a, st);
  else launch_cg_step_s_sb<ST, SR, SX, 32>(a, st);
}

// storage: 0 all FP32; 1 p, t FP16, r and x FP32; 2 all FP16
bool launch_cg_step_s(const CgStepArgs<float>& a, int storage, hipStream_t st)
{
  if (!three_phase_ok(a.ls, 32) && !three_phase_ok(a.ls, 16)) return false;
  if (storage == 2) launch_cg_step_s_st<Half16, Half16, Half16>(a, st);
  else if (storage == 1) launch_cg_step_s_st<Half16, float, float>(a, st);
  else launch_cg_step_s_st<float, float, float>(a, st);
  return true;
}

}  // namespace qdwf
