// Shared device-side definitions of the gfx950 domain-wall library.
//
// Device layouts (structure of arrays, lanes = consecutive half-sites):
//   half spinor field, FP64: double2  f[(s*12 + k)*stride + h]   k = spin*3+colour
//   half spinor field, FP32: float4   f[(s*6 + kp)*stride + h]   complex 2kp,2kp+1
//   links, FP64 / FP32:      double2 / float2
//                            g[((p*8 + d)*9 + e)*stride + h]
//     p = output parity slot (0 even, 1 odd), d = 0..3 forward U_mu(x),
//     d = 4..7 backward U_mu(x - mu)^dagger, e = row*3+col.
// `stride` is the half volume rounded up to a multiple of 256; pad entries
// are zero and never written.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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
// Lattice extents and parity bookkeeping.  Site order is the reference's
// (x fastest); a half-site index is full_index/2 (qlat core.h:232-261).
struct LatGeo {
  int n[4];       // X Y Z T
  int xh;         // X/2
  int64_t vh;     // V/2
  int64_t stride; // vh rounded up to kStrideAlign
};

// Full coordinate of half-site h on parity eo (2 even, 1 odd).
__device__ __forceinline__ void half_coord(const LatGeo& g, int64_t h, int eo,
                                           int x[4])
{
  int64_t r = h / g.xh;
  const int xh = (int)(h - r * g.xh);
  x[1] = (int)(r % g.n[1]);
  r /= g.n[1];
  x[2] = (int)(r % g.n[2]);
  x[3] = (int)(r / g.n[2]);
  const int odd = (eo == 1) ? 1 : 0;
  x[0] = 2 * xh + ((x[1] + x[2] + x[3] + odd) & 1);
}

__device__ __forceinline__ int64_t half_index(const LatGeo& g, const int x[4])
{
  return (int64_t)(x[0] >> 1) +
         (int64_t)g.xh * ((int64_t)x[1] + (int64_t)g.n[1] * ((int64_t)x[2] + (int64_t)g.n[2] * x[3]));
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

// Workgroup id remapped so that consecutive logical blocks land on the same
// XCD (hardware dispatches workgroup b to XCD b % 8).
__device__ __forceinline__ int64_t xcd_logical_block(int64_t b, int64_t nblk)
{
  const int64_t q = nblk / 8, r = nblk % 8;
  const int64_t xcd = b % 8, local = b / 8;
  return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// ---------------------------------------------------------------- spinors
template <class T>
struct SpinorIO;

template <>
struct SpinorIO<double> {
  typedef double2 slot_t;
  static constexpr int kSlots = 12;
  __device__ __forceinline__ static void load(const double2* __restrict__ f, int64_t stride,
                                              int s, int64_t h, cplx<double> v[12])
  {
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const double2 a = f[((int64_t)s * 12 + k) * stride + h];
      v[k] = cmk<double>(a.x, a.y);
    }
  }
  __device__ __forceinline__ static void store(double2* __restrict__ f, int64_t stride, int s,
                                               int64_t h, const cplx<double> v[12])
  {
#pragma unroll
    for (int k = 0; k < 12; ++k) f[((int64_t)s * 12 + k) * stride + h] = make_double2(v[k].re, v[k].im);
  }
};

template <>
struct SpinorIO<float> {
  typedef float4 slot_t;
  static constexpr int kSlots = 6;
  __device__ __forceinline__ static void load(const float4* __restrict__ f, int64_t stride,
                                              int s, int64_t h, cplx<float> v[12])
  {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float4 a = f[((int64_t)s * 6 + k) * stride + h];
      v[2 * k] = cmk<float>(a.x, a.y);
      v[2 * k + 1] = cmk<float>(a.z, a.w);
    }
  }
  __device__ __forceinline__ static void store(float4* __restrict__ f, int64_t stride, int s,
                                               int64_t h, const cplx<float> v[12])
  {
#pragma unroll
    for (int k = 0; k < 6; ++k)
      f[((int64_t)s * 6 + k) * stride + h] =
          make_float4(v[2 * k].re, v[2 * k].im, v[2 * k + 1].re, v[2 * k + 1].im);
  }
};

template <class T>
struct LinkIO;

template <>
struct LinkIO<double> {
  typedef double2 elem_t;
  __device__ __forceinline__ static void load(const double2* __restrict__ g, int64_t stride,
                                              int pd, int64_t h, cplx<double> u[9])
  {
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const double2 a = g[((int64_t)pd * 9 + e) * stride + h];
      u[e] = cmk<double>(a.x, a.y);
    }
  }
};

template <>
struct LinkIO<float> {
  typedef float2 elem_t;
  __device__ __forceinline__ static void load(const float2* __restrict__ g, int64_t stride,
                                              int pd, int64_t h, cplx<float> u[9])
  {
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const float2 a = g[((int64_t)pd * 9 + e) * stride + h];
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

}  // namespace qdwf
