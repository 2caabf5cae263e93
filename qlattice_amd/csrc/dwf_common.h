// Shared device-side definitions of the gfx950 domain-wall library.
//
// Device layouts (blocked structure of arrays: lanes = 64 consecutive
// device half-sites; hb = h / 64, lane = h % 64):
//   half spinor field, FP64: double2  f[((hb*Ls + s)*12 + k)*64 + lane]   k = spin*3+colour
//   half spinor field, FP32: float4   f[((hb*Ls + s)*6 + kp)*64 + lane]   complex 2kp, 2kp+1
//   half spinor field, FP16 with a per-5D-site scale: 16-byte units
//                            u[(hb*Ls + s)*208 + j*64 + lane]  j < 3: 8 halves = complex 4j..4j+3
//                            scale: float at unit (hb*Ls + s)*208 + 192, element lane
//     value = half * scale, scale = max |re|,|im| over the site's 24 reals
//   links, FP64 / FP32:      double2 / float2
//                            g[((p*nhb + hb)*72 + d*9 + e)*64 + lane]
//     p = output parity slot (0 even, 1 odd), d = 0..3 forward U_mu(x),
//     d = 4..7 backward U_mu(x - mu)^dagger, e = row*3+col, nhb = stride/64.
// All Ls values of a 64-site block are contiguous (72 KB in FP32), and so
// are a block's 8 links (37 KB), which keeps the number of pages a wave
// touches small.  `stride` is the half volume rounded up to a multiple of
// 256; pad entries are zero and never written.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace qdwf {

constexpr int kStrideAlign = 256;

template <class T>
struct cplx {
  T re, im;
};

template <class T>
__device__ __forceinline__ cplx<T> cmk(T r, T i)
{
  cplx<T> z;
  z.re = r;
  z.im = i;
  return z;
}
template <class T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b)
{
  return cmk<T>(a.re + b.re, a.im + b.im);
}
template <class T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b)
{
  return cmk<T>(a.re - b.re, a.im - b.im);
}
template <class T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
  return cmk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// a + b*c
template <class T>
__device__ __forceinline__ cplx<T> cfma(cplx<T> a, cplx<T> b, cplx<T> c)
{
  return cmk<T>(a.re + b.re * c.re - b.im * c.im, a.im + b.re * c.im + b.im * c.re);
}
// i*a
template <class T>
__device__ __forceinline__ cplx<T> cmuli(cplx<T> a)
{
  return cmk<T>(-a.im, a.re);
}
template <class T>
__device__ __forceinline__ cplx<T> cneg(cplx<T> a)
{
  return cmk<T>(-a.re, -a.im);
}

// ---------------------------------------------------------------- geometry
// Lattice extents and the device site order.  The reference orders sites
// lexicographically with x fastest and a half-site index is full_index/2
// (qlat core.h:232-261).  On the device the half sites of one parity are
// ordered as
//     d = xh + X/2 * (yi + by * (zi + bz * (t + T * (zt + Z/bz * yt)))),
//     y = yt*by + yi,  z = zt*bz + zi,
// i.e. the lattice is cut into tiles of `by` rows in y and `bz` planes in z
// and each tile is swept in t.  A t neighbour is then X/2*by*bz half sites
// away instead of X/2*Y*Z (shorter reuse distance in the L2); the y and z
// neighbours across a tile face are read from far away.  by = Y, bz = Z is
// the reference order.  Layout conversion at the boundary applies the
// permutation (pack/unpack kernels).
// Unsigned division n / d for n < 2^31 by a divisor fixed at set-up, as a
// multiply and shift (Granlund and Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994): with 2^(s-1) < d <= 2^s,
// m = floor(2^(31+s) / d) + 1 < 2^32 and n / d = (n * m) >> (31 + s); the
// error term n (m - 2^(31+s)/d) / 2^(31+s) < 2^-s <= 1/d keeps the floor
// exact.
struct FastDiv {
  unsigned d, m, sh;
};
inline FastDiv make_fastdiv(unsigned d)
{
  unsigned s = 0;
  while ((1ull << s) < d) ++s;
  const unsigned long long m = ((1ull << (31 + s)) / d) + 1;
  return FastDiv{d, (unsigned)m, 31 + s};
}
__device__ __forceinline__ unsigned fdiv(unsigned n, const FastDiv& f)
{
  return (unsigned)(((unsigned long long)n * f.m) >> f.sh);
}

struct LatGeo {
  int n[4];       // X Y Z T
  int xh;         // X/2
  int by;         // tile height in y (divides Y)
  int bz;         // tile depth in z (divides Z)
  int64_t vh;     // V/2
  int64_t stride; // vh rounded up to kStrideAlign
  FastDiv dxh, dby, dbz, dnt, dnzt;  // divisions by xh, by, bz, T, Z/bz (set with the extents)
};

// Full coordinate of the device half-site d on parity eo (2 even, 1 odd).
// 32-bit arithmetic: a half volume is < 2^31 for any lattice that fits in
// device memory.
__device__ __forceinline__ void half_coord(const LatGeo& g, int64_t d64, int eo, int x[4])
{
  const unsigned d = (unsigned)d64;
  const unsigned xh = (unsigned)g.xh, by = (unsigned)g.by, bz = (unsigned)g.bz, nt = (unsigned)g.n[3];
  const unsigned nzt = (unsigned)(g.n[2] / g.bz);
  unsigned r = fdiv(d, g.dxh);
  const unsigned xr = d - r * xh;
  unsigned q = fdiv(r, g.dby);
  const unsigned yi = r - q * by;
  r = fdiv(q, g.dbz);
  const unsigned zi = q - r * bz;
  q = fdiv(r, g.dnt);
  x[3] = (int)(r - q * nt);
  r = fdiv(q, g.dnzt);
  x[2] = (int)((q - r * nzt) * bz + zi);
  x[1] = (int)(r * by + yi);
  const int odd = (eo == 1) ? 1 : 0;
  x[0] = 2 * (int)xr + ((x[1] + x[2] + x[3] + odd) & 1);
}

// Device half-site index of full coordinate x (0 <= x[mu] < n[mu]).
__device__ __forceinline__ int64_t half_index(const LatGeo& g, const int x[4])
{
  const unsigned by = (unsigned)g.by, bz = (unsigned)g.bz;
  const unsigned nzt = (unsigned)(g.n[2] / g.bz);
  const unsigned y = (unsigned)x[1], yt = fdiv(y, g.dby), yi = y - yt * by;
  const unsigned z = (unsigned)x[2], zt = fdiv(z, g.dbz), zi = z - zt * bz;
  return (int64_t)((unsigned)(x[0] >> 1) +
                   (unsigned)g.xh * (yi + by * (zi + bz * ((unsigned)x[3] + (unsigned)g.n[3] * (zt + nzt * yt)))));
}

// Slot index of (site h, fifth coordinate s, 16-byte slot) in a blocked
// spinor field with NS slots per 5D site.
// 32-bit: the inverter requires every field to have fewer than 2^32 slots.
template <int NS>
__device__ __forceinline__ unsigned sidx(int ls, int s, int slot, int64_t h)
{
  const unsigned hh = (unsigned)h;
  return ((((hh >> 6) * (unsigned)ls + (unsigned)s) * NS + (unsigned)slot) << 6) + (hh & 63u);
}

// Element index of link (p, d, e) of site h in the blocked link layout.
__device__ __forceinline__ int64_t lidx(int64_t stride, int p, int de, int64_t h)
{
  return ((((int64_t)p * (stride >> 6) + (h >> 6)) * 72 + de) << 6) + (h & 63);
}

// Reference (lexicographic) full site index of coordinate x.
__device__ __forceinline__ int64_t full_index(const LatGeo& g, const int x[4])
{
  return (int64_t)x[0] + (int64_t)g.n[0] * ((int64_t)x[1] + (int64_t)g.n[1] * ((int64_t)x[2] + (int64_t)g.n[2] * x[3]));
}

__device__ __forceinline__ int wrap_inc(int v, int n) { return (v + 1 == n) ? 0 : v + 1; }
__device__ __forceinline__ int wrap_dec(int v, int n) { return (v == 0) ? n - 1 : v - 1; }

// Half indices of the 8 neighbours x+mu (d = mu) and x-mu (d = 4+mu).
__device__ __forceinline__ void neighbours(const LatGeo& g, const int x[4], int64_t nb[8])
{
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    int xp[4] = {x[0], x[1], x[2], x[3]};
    int xm[4] = {x[0], x[1], x[2], x[3]};
    xp[mu] = wrap_inc(x[mu], g.n[mu]);
    xm[mu] = wrap_dec(x[mu], g.n[mu]);
    nb[mu] = half_index(g, xp);
    nb[4 + mu] = half_index(g, xm);
  }
}

// Half index of neighbour d (x+mu for d = mu, x-mu for d = 4+mu) of the
// device half-site h on parity eo.
__device__ __forceinline__ int64_t neighbour_of(const LatGeo& g, int64_t h, int eo, int d)
{
  int x[4];
  half_coord(g, h, eo, x);
  const int mu = d & 3;
  const bool fwd = d < 4;
#pragma unroll
  for (int c = 0; c < 4; ++c)
    if (c == mu) x[c] = fwd ? wrap_inc(x[c], g.n[c]) : wrap_dec(x[c], g.n[c]);
  return half_index(g, x);
}

// Workgroup id remapped so that consecutive logical blocks land on the same
// XCD (hardware dispatches workgroup b to XCD b % 8).
__device__ __forceinline__ int64_t xcd_logical_block(int64_t b, int64_t nblk)
{
  const int64_t q = nblk / 8, r = nblk % 8;
  const int64_t xcd = b % 8, local = b / 8;
  return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// Element i of a small by-value array (kernel arguments) without dynamic
// indexing, which would copy the array to scratch memory; i is wave-uniform.
template <class V, int N>
__device__ __forceinline__ V pick(const V (&arr)[N], int i)
{
  V r = arr[0];
#pragma unroll
  for (int j = 1; j < N; ++j)
    if (j == i) r = arr[j];
  return r;
}

// Non-temporal (streaming) stores: outputs that the next kernel reads in a
// different order should not evict this kernel's reusable inputs from L2.
__device__ __forceinline__ void store_nt(float4* p, const float4& v)
{
  typedef float v4f __attribute__((ext_vector_type(4)));
  v4f t = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(t, (v4f*)p);
}
__device__ __forceinline__ void store_nt(double2* p, const double2& v)
{
  typedef double v2d __attribute__((ext_vector_type(2)));
  v2d t = {v.x, v.y};
  __builtin_nontemporal_store(t, (v2d*)p);
}
__device__ __forceinline__ float2 load_nt(const float2* p)
{
  typedef float v2f __attribute__((ext_vector_type(2)));
  const v2f t = __builtin_nontemporal_load((const v2f*)p);
  return make_float2(t.x, t.y);
}
__device__ __forceinline__ double2 load_nt(const double2* p)
{
  typedef double v2d __attribute__((ext_vector_type(2)));
  const v2d t = __builtin_nontemporal_load((const v2d*)p);
  return make_double2(t.x, t.y);
}

// ---------------------------------------------------------------- spinors
template <class T>
struct SpinorIO;

template <>
struct SpinorIO<double> {
  typedef double2 slot_t;
  typedef double real;
  static constexpr int kSlots = 12;
  __device__ __forceinline__ static void load(const double2* __restrict__ f, int ls, int s, int64_t h,
                                              cplx<double> v[12])
  {
    const double2* p = f + sidx<12>(ls, s, 0, h);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const double2 a = p[k * 64];
      v[k] = cmk<double>(a.x, a.y);
    }
  }
  __device__ __forceinline__ static void store(double2* __restrict__ f, int ls, int s, int64_t h,
                                               const cplx<double> v[12])
  {
    double2* p = f + sidx<12>(ls, s, 0, h);
#pragma unroll
    for (int k = 0; k < 12; ++k) p[k * 64] = make_double2(v[k].re, v[k].im);
  }
  __device__ __forceinline__ static void store_stream(double2* __restrict__ f, int ls, int s, int64_t h,
                                                      const cplx<double> v[12])
  {
    double2* p = f + sidx<12>(ls, s, 0, h);
#pragma unroll
    for (int k = 0; k < 12; ++k) store_nt(&p[k * 64], make_double2(v[k].re, v[k].im));
  }
};

template <>
struct SpinorIO<float> {
  typedef float4 slot_t;
  typedef float real;
  static constexpr int kSlots = 6;
  __device__ __forceinline__ static void load(const float4* __restrict__ f, int ls, int s, int64_t h,
                                              cplx<float> v[12])
  {
    const float4* p = f + sidx<6>(ls, s, 0, h);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float4 a = p[k * 64];
      v[2 * k] = cmk<float>(a.x, a.y);
      v[2 * k + 1] = cmk<float>(a.z, a.w);
    }
  }
  __device__ __forceinline__ static void store(float4* __restrict__ f, int ls, int s, int64_t h,
                                               const cplx<float> v[12])
  {
    float4* p = f + sidx<6>(ls, s, 0, h);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      p[k * 64] = make_float4(v[2 * k].re, v[2 * k].im, v[2 * k + 1].re, v[2 * k + 1].im);
  }
  __device__ __forceinline__ static void store_stream(float4* __restrict__ f, int ls, int s, int64_t h,
                                                      const cplx<float> v[12])
  {
    float4* p = f + sidx<6>(ls, s, 0, h);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      store_nt(&p[k * 64], make_float4(v[2 * k].re, v[2 * k].im, v[2 * k + 1].re, v[2 * k + 1].im));
  }
  __device__ __forceinline__ static void load_stream(const float4* __restrict__ f, int ls, int s, int64_t h,
                                                     cplx<float> v[12])
  {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const float4* p = f + sidx<6>(ls, s, 0, h);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const v4f a = __builtin_nontemporal_load((const v4f*)&p[k * 64]);
      v[2 * k] = cmk<float>(a.x, a.y);
      v[2 * k + 1] = cmk<float>(a.z, a.w);
    }
  }
  // raw loads (issued together, decoded later: more loads in flight per wave)
  struct Raw {
    float4 q[6];
  };
  __device__ __forceinline__ static void load_raw(const float4* __restrict__ f, int ls, int s, int64_t h, Raw& r)
  {
    const float4* p = f + sidx<6>(ls, s, 0, h);
#pragma unroll
    for (int k = 0; k < 6; ++k) r.q[k] = p[k * 64];
  }
  __device__ __forceinline__ static void decode(const Raw& r, cplx<float> v[12])
  {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      v[2 * k] = cmk<float>(r.q[k].x, r.q[k].y);
      v[2 * k + 1] = cmk<float>(r.q[k].z, r.q[k].w);
    }
  }
};

// FP16 storage with a per-5D-site FP32 scale (arithmetic in FP32).  Stores
// round to nearest, and store a site whose largest component is below
// 2^-100 as zero (the solvers keep their Krylov vectors normalised, so such
// sites are far below the storage precision); store_q also returns the
// values as stored, so that a
// kernel's other uses of them (dot products, s-matrix inputs) see exactly
// what the next kernel will read.
struct Half16 {};

typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned pack_h2(float a, float b)
{
  h16x2 v;
  v.x = (_Float16)a;
  v.y = (_Float16)b;
  return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ cplx<float> unpack_h2(unsigned w, float sc)
{
  const h16x2 v = __builtin_bit_cast(h16x2, w);
  return cmk<float>((float)v.x * sc, (float)v.y * sc);
}

template <>
struct SpinorIO<Half16> {
  typedef uint4 slot_t;
  typedef float real;
  // A 64-site block holds Ls x 208 sixteen-byte units: row s of spinors at
  // units 192 s + 64 j + lane (j < 3), then the block's Ls x 64 FP32
  // scales, grouped by 16 sites: the scale of (s, lane) at byte
  // 64 ((lane / 16) Ls + s) + 4 (lane % 16) of that area, so that four
  // consecutive rows of 16 sites (one wave of a 16-site block) share two
  // 128-byte lines.
  static constexpr int kUnits = 208;  // 16-byte units per (64-site block, s)
  __device__ __forceinline__ static unsigned base(int ls, int s, int64_t h)
  {
    return ((unsigned)h >> 6) * (unsigned)ls * (unsigned)kUnits + (unsigned)s * 192u;
  }
  __device__ __forceinline__ static unsigned scale_byte(int ls, int s, int64_t h)
  {
    const unsigned lane = (unsigned)h & 63u;
    return (((unsigned)h >> 6) * (unsigned)ls * (unsigned)kUnits + (unsigned)ls * 192u) * 16u +
           (((lane >> 4) * (unsigned)ls + (unsigned)s) << 6) + ((lane & 15u) << 2);
  }
  struct Raw {
    uint4 q[3];
    float sc;
  };
  // Byte offsets within one field are 32-bit (a field is below 4 GB), so
  // that a wave-uniform field pointer plus a 32-bit offset addresses it.
  __device__ __forceinline__ static const uint4* slot_at(const uint4* f, unsigned unit)
  {
    return (const uint4*)((const char*)f + unit * 16u);
  }
  __device__ __forceinline__ static uint4* slot_at(uint4* f, unsigned unit)
  {
    return (uint4*)((char*)f + unit * 16u);
  }
  __device__ __forceinline__ static const float* scale_at(const uint4* f, unsigned byte)
  {
    return (const float*)((const char*)f + byte);
  }
  __device__ __forceinline__ static float* scale_at(uint4* f, unsigned byte)
  {
    return (float*)((char*)f + byte);
  }
  __device__ __forceinline__ static void load_raw(const uint4* __restrict__ f, int ls, int s, int64_t h, Raw& r)
  {
    const unsigned b = base(ls, s, h), lane = (unsigned)h & 63u;
    r.sc = *scale_at(f, scale_byte(ls, s, h));
#pragma unroll
    for (int j = 0; j < 3; ++j) r.q[j] = *slot_at(f, b + j * 64u + lane);
  }
  __device__ __forceinline__ static void decode(const Raw& r, cplx<float> v[12])
  {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      v[4 * j + 0] = unpack_h2(r.q[j].x, r.sc);
      v[4 * j + 1] = unpack_h2(r.q[j].y, r.sc);
      v[4 * j + 2] = unpack_h2(r.q[j].z, r.sc);
      v[4 * j + 3] = unpack_h2(r.q[j].w, r.sc);
    }
  }
  __device__ __forceinline__ static void load(const uint4* __restrict__ f, int ls, int s, int64_t h,
                                              cplx<float> v[12])
  {
    const unsigned b = base(ls, s, h), lane = (unsigned)h & 63u;
    const float sc = *scale_at(f, scale_byte(ls, s, h));
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const uint4 q = *slot_at(f, b + j * 64u + lane);
      v[4 * j + 0] = unpack_h2(q.x, sc);
      v[4 * j + 1] = unpack_h2(q.y, sc);
      v[4 * j + 2] = unpack_h2(q.z, sc);
      v[4 * j + 3] = unpack_h2(q.w, sc);
    }
  }
  // the same as load with non-temporal loads (streaming kernels)
  __device__ __forceinline__ static void load_stream(const uint4* __restrict__ f, int ls, int s, int64_t h,
                                                     cplx<float> v[12])
  {
    const unsigned b = base(ls, s, h), lane = (unsigned)h & 63u;
    const float sc = __builtin_nontemporal_load(scale_at(f, scale_byte(ls, s, h)));
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const u32x4 q = __builtin_nontemporal_load((const u32x4*)slot_at(f, b + j * 64u + lane));
      v[4 * j + 0] = unpack_h2(q.x, sc);
      v[4 * j + 1] = unpack_h2(q.y, sc);
      v[4 * j + 2] = unpack_h2(q.z, sc);
      v[4 * j + 3] = unpack_h2(q.w, sc);
    }
  }
  template <bool NT>
  __device__ __forceinline__ static void put(uint4* __restrict__ f, int ls, int s, int64_t h, cplx<float> v[12])
  {
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 12; ++k) m = fmaxf(m, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
    // sites below 2^-100 are stored as zero: 1/m must stay finite
    if (m < 0x1p-100f) m = 0.f;
    const float inv = m > 0.f ? 1.f / m : 0.f;
    const unsigned b = base(ls, s, h), lane = (unsigned)h & 63u;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      uint4 q;
      q.x = pack_h2(v[4 * j + 0].re * inv, v[4 * j + 0].im * inv);
      q.y = pack_h2(v[4 * j + 1].re * inv, v[4 * j + 1].im * inv);
      q.z = pack_h2(v[4 * j + 2].re * inv, v[4 * j + 2].im * inv);
      q.w = pack_h2(v[4 * j + 3].re * inv, v[4 * j + 3].im * inv);
      uint4* dst = slot_at(f, b + j * 64u + lane);
      if (NT) {
        const u32x4 t = {q.x, q.y, q.z, q.w};
        __builtin_nontemporal_store(t, (u32x4*)dst);
      } else {
        *dst = q;
      }
      v[4 * j + 0] = unpack_h2(q.x, m);
      v[4 * j + 1] = unpack_h2(q.y, m);
      v[4 * j + 2] = unpack_h2(q.z, m);
      v[4 * j + 3] = unpack_h2(q.w, m);
    }
    *scale_at(f, scale_byte(ls, s, h)) = m;
  }
  __device__ __forceinline__ static void store_q(uint4* __restrict__ f, int ls, int s, int64_t h, cplx<float> v[12])
  {
    put<false>(f, ls, s, h, v);
  }
  __device__ __forceinline__ static void store(uint4* __restrict__ f, int ls, int s, int64_t h,
                                               const cplx<float> v0[12])
  {
    cplx<float> v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = v0[k];
    put<false>(f, ls, s, h, v);
  }
  __device__ __forceinline__ static void store_stream(uint4* __restrict__ f, int ls, int s, int64_t h,
                                                      const cplx<float> v0[12])
  {
    cplx<float> v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = v0[k];
    put<true>(f, ls, s, h, v);
  }
};

// store_q for FP32/FP64 storage: values are stored exactly
template <class ST>
__device__ __forceinline__ void store_q(typename SpinorIO<ST>::slot_t* __restrict__ f, int ls, int s, int64_t h,
                                        cplx<typename SpinorIO<ST>::real> v[12])
{
  if constexpr (std::is_same<ST, Half16>::value) SpinorIO<Half16>::store_q(f, ls, s, h, v);
  else SpinorIO<ST>::store(f, ls, s, h, v);
}

template <class T>
struct LinkIO;

template <>
struct LinkIO<double> {
  typedef double2 elem_t;
  __device__ __forceinline__ static void load(const double2* __restrict__ g, int64_t stride, int p, int d,
                                              int64_t h, cplx<double> u[9])
  {
    const double2* q = g + lidx(stride, p, d * 9, h);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const double2 a = q[e * 64];
      u[e] = cmk<double>(a.x, a.y);
    }
  }
};

template <>
struct LinkIO<float> {
  typedef float2 elem_t;
  __device__ __forceinline__ static void load(const float2* __restrict__ g, int64_t stride, int p, int d,
                                              int64_t h, cplx<float> u[9])
  {
    const float2* q = g + lidx(stride, p, d * 9, h);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const float2 a = q[e * 64];
      u[e] = cmk<float>(a.x, a.y);
    }
  }
};

// ------------------------------------------------------- spin projection
// CPS gamma matrices in the chiral basis are gamma_mu = [[0, B],[B^+, 0]]
// (2x2 spin blocks) with
//   B_x = [[0, i],[i, 0]],  B_y = [[0,-1],[1, 0]],
//   B_z = [[i, 0],[0,-i]],  B_t = [[1, 0],[0, 1]]
// (the reference's cps_gammas, matrix.h:426-452).  For psi = (u, l),
//   (1 + sg*gamma_mu) psi = (w, sg*B^+ w),  w = u + sg*B l,
// so a hop term needs a 2-spin half spinor, one SU(3) product per spin and
// a cheap reconstruction.

// w = u + sg*B l for one colour: inputs u0,u1,l0,l1 (spins 0..3).
template <int MU, int SG, class T>
__device__ __forceinline__ void proj(cplx<T> u0, cplx<T> u1, cplx<T> l0, cplx<T> l1,
                                     cplx<T>& w0, cplx<T>& w1)
{
  if (MU == 0) {  // B l = (i l1, i l0)
    const cplx<T> a = cmuli(l1), b = cmuli(l0);
    w0 = SG > 0 ? cadd(u0, a) : csub(u0, a);
    w1 = SG > 0 ? cadd(u1, b) : csub(u1, b);
  } else if (MU == 1) {  // B l = (-l1, l0)
    w0 = SG > 0 ? csub(u0, l1) : cadd(u0, l1);
    w1 = SG > 0 ? cadd(u1, l0) : csub(u1, l0);
  } else if (MU == 2) {  // B l = (i l0, -i l1)
    const cplx<T> a = cmuli(l0), b = cmuli(l1);
    w0 = SG > 0 ? cadd(u0, a) : csub(u0, a);
    w1 = SG > 0 ? csub(u1, b) : cadd(u1, b);
  } else {  // B l = l
    w0 = SG > 0 ? cadd(u0, l0) : csub(u0, l0);
    w1 = SG > 0 ? cadd(u1, l1) : csub(u1, l1);
  }
}

// acc_l += sg * B^+ y for one colour (y = U w for spins 0,1).
template <int MU, int SG, class T>
__device__ __forceinline__ void recon_lower(cplx<T> y0, cplx<T> y1, cplx<T>& a2, cplx<T>& a3)
{
  if (MU == 0) {  // B^+ y = (-i y1, -i y0)
    const cplx<T> p = cmuli(y1), q = cmuli(y0);
    if (SG > 0) {
      a2 = csub(a2, p);
      a3 = csub(a3, q);
    } else {
      a2 = cadd(a2, p);
      a3 = cadd(a3, q);
    }
  } else if (MU == 1) {  // B^+ y = (y1, -y0)
    if (SG > 0) {
      a2 = cadd(a2, y1);
      a3 = csub(a3, y0);
    } else {
      a2 = csub(a2, y1);
      a3 = cadd(a3, y0);
    }
  } else if (MU == 2) {  // B^+ y = (-i y0, i y1)
    const cplx<T> p = cmuli(y0), q = cmuli(y1);
    if (SG > 0) {
      a2 = csub(a2, p);
      a3 = cadd(a3, q);
    } else {
      a2 = cadd(a2, p);
      a3 = csub(a3, q);
    }
  } else {  // B^+ y = y
    if (SG > 0) {
      a2 = cadd(a2, y0);
      a3 = cadd(a3, y1);
    } else {
      a2 = csub(a2, y0);
      a3 = csub(a3, y1);
    }
  }
}

// acc += (1 + sg*gamma_mu) contribution of link matrix W applied to psi:
//   acc_u += W w,  acc_l += sg B^+ (W w),  w = u + sg B l.
template <int MU, int SG, class T>
__device__ __forceinline__ void hop_term(const cplx<T> W[9], const cplx<T> psi[12], cplx<T> acc[12])
{
  cplx<T> w0[3], w1[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) proj<MU, SG>(psi[c], psi[3 + c], psi[6 + c], psi[9 + c], w0[c], w1[c]);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    cplx<T> y0 = cmul(W[r * 3 + 0], w0[0]);
    cplx<T> y1 = cmul(W[r * 3 + 0], w1[0]);
    y0 = cfma(y0, W[r * 3 + 1], w0[1]);
    y1 = cfma(y1, W[r * 3 + 1], w1[1]);
    y0 = cfma(y0, W[r * 3 + 2], w0[2]);
    y1 = cfma(y1, W[r * 3 + 2], w1[2]);
    acc[r] = cadd(acc[r], y0);
    acc[3 + r] = cadd(acc[3 + r], y1);
    recon_lower<MU, SG>(y0, y1, acc[6 + r], acc[9 + r]);
  }
}

// ------------------------------------------- FP16-arithmetic hop term
// The same hop term with the neighbour spinor as stored (FP16 values of a
// site, times the site's scale sc) and FP16 links: the projection is one
// packed FP16 add per colour and spin (the swizzle and signs of B fold into
// the add's operand modifiers), each SU(3) row is three 2-term FP16 dot
// products accumulated in FP32 (v_dot2_f32_f16: the products are exact in
// FP32), and the scale is applied when the FP32 result is accumulated.  A
// link element enters as la = (re, -im) and lb = (im, re), so that
// Re(U w) = la . w and Im(U w) = lb . w.

// w = u + sg*B l for one colour, packed (x = re, y = im)
template <int MU, int SG>
__device__ __forceinline__ void proj_h2(h16x2 u0, h16x2 u1, h16x2 l0, h16x2 l1, h16x2& w0, h16x2& w1)
{
  const _Float16 s = SG > 0 ? (_Float16)1 : (_Float16)-1;
  if (MU == 0) {  // B l = (i l1, i l0)
    w0 = u0 + s * (h16x2){-l1.y, l1.x};
    w1 = u1 + s * (h16x2){-l0.y, l0.x};
  } else if (MU == 1) {  // B l = (-l1, l0)
    w0 = u0 - s * l1;
    w1 = u1 + s * l0;
  } else if (MU == 2) {  // B l = (i l0, -i l1)
    w0 = u0 + s * (h16x2){-l0.y, l0.x};
    w1 = u1 - s * (h16x2){-l1.y, l1.x};
  } else {  // B l = l
    w0 = u0 + s * l0;
    w1 = u1 + s * l1;
  }
}

// (t2, t3) = B^+ (y0, y1)
template <int MU>
__device__ __forceinline__ void bdag(cplx<float> y0, cplx<float> y1, cplx<float>& t2, cplx<float>& t3)
{
  if (MU == 0) {  // (-i y1, -i y0)
    t2 = cmk<float>(y1.im, -y1.re);
    t3 = cmk<float>(y0.im, -y0.re);
  } else if (MU == 1) {  // (y1, -y0)
    t2 = y1;
    t3 = cneg(y0);
  } else if (MU == 2) {  // (-i y0, i y1)
    t2 = cmk<float>(y0.im, -y0.re);
    t3 = cmk<float>(-y1.im, y1.re);
  } else {  // (y0, y1)
    t2 = y0;
    t3 = y1;
  }
}

__device__ __forceinline__ void cfma_s(cplx<float>& acc, float s, cplx<float> t)
{
  acc.re = fmaf(s, t.re, acc.re);
  acc.im = fmaf(s, t.im, acc.im);
}

// a.x b.x + a.y b.y + 0 in FP32 from FP16 pairs as one v_dot2_f32_f16 with
// an inline-constant accumulator: the builtin with a zero accumulator is
// compiled to a move of 0 plus the two-operand form, one instruction more
// per dot-product chain.  An ALU instruction only (no memory operand).
__device__ __forceinline__ float fdot2_z(h16x2 a, h16x2 b)
{
  float r;
  asm("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// acc += sc * (1 + sg gamma_mu) U psi with psi, U in FP16 (la, lb as above)
template <int MU, int SG>
__device__ __forceinline__ void hop_term_h2(const h16x2 la[9], const h16x2 lb[9], const h16x2 psi[12], float sc,
                                            cplx<float> acc[12])
{
  h16x2 w0[3], w1[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) proj_h2<MU, SG>(psi[c], psi[3 + c], psi[6 + c], psi[9 + c], w0[c], w1[c]);
  const float ssg = SG > 0 ? sc : -sc;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float y0r = fdot2_z(la[r * 3], w0[0]), y0i = fdot2_z(lb[r * 3], w0[0]);
    float y1r = fdot2_z(la[r * 3], w1[0]), y1i = fdot2_z(lb[r * 3], w1[0]);
#pragma unroll
    for (int c = 1; c < 3; ++c) {
      y0r = __builtin_amdgcn_fdot2(la[r * 3 + c], w0[c], y0r, false);
      y0i = __builtin_amdgcn_fdot2(lb[r * 3 + c], w0[c], y0i, false);
      y1r = __builtin_amdgcn_fdot2(la[r * 3 + c], w1[c], y1r, false);
      y1i = __builtin_amdgcn_fdot2(lb[r * 3 + c], w1[c], y1i, false);
    }
    const cplx<float> y0 = cmk<float>(y0r, y0i), y1 = cmk<float>(y1r, y1i);
    cfma_s(acc[r], sc, y0);
    cfma_s(acc[3 + r], sc, y1);
    cplx<float> t2, t3;
    bdag<MU>(y0, y1, t2, t3);
    cfma_s(acc[6 + r], ssg, t2);
    cfma_s(acc[9 + r], ssg, t3);
  }
}

}  // namespace qdwf
