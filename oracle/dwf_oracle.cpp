// TEST INFRASTRUCTURE ONLY -- the CPU FP64 oracle for the domain-wall path.
//
// This file restates, in plain C++17 + OpenMP, the algorithm of Qlattice's
// even-odd Moebius domain-wall operator and its CG inverter so that the HIP
// product path can be checked against it.  It is never linked into the
// product library: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it (through ctypes, see oracle/oracle.py).
//
// Layout follows the reference exactly (array-of-structures, complex128):
//   WilsonVector  = 12 complex, index spin*3+colour
//                   (qlat-utils/.../matrix.h:736-771, mat-vec.h:286)
//   ColorMatrix   = 3x3 row-major (mat-vec.h:52-74)
//   5D full field = [site][s][12]          (core.h:1203, multiplicity = Ls)
//   5D half field = [site/2][s][12]        (core.h:232-261, half index = full/2)
//   gauge field   = [site][mu][3][3]       (core.h:1187)
//   site index    = x + X*(y + Y*(z + Z*t)) (coordinate.h:169-188)
//   parity eo     = 2 - (x+y+z+t)%2 : 2 even, 1 odd (coordinate.h:190-193)
//
// Every function cites the reference lines it follows.  The loop structure
// (per-site full-spinor algebra, SpinMatrix*WilsonVector with zero skipping,
// per-call temporaries, serial dot_product) is kept on purpose so that the
// timing of this file is a fair stand-in for the reference CPU path.

#include <omp.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

typedef std::complex<double> cd;

struct WV {  // WilsonVector (matrix.h:736-771)
  cd p[12];
};
struct SM {  // SpinMatrix, row-major 4x4
  cd p[16];
};
struct CM {  // ColorMatrix, row-major 3x3
  cd p[9];
};

inline void set_zero(WV& v)
{
  for (int k = 0; k < 12; ++k) v.p[k] = 0.0;
}
inline WV operator+(const WV& a, const WV& b)
{
  WV r;
  for (int k = 0; k < 12; ++k) r.p[k] = a.p[k] + b.p[k];
  return r;
}
inline WV& operator+=(WV& a, const WV& b)
{
  for (int k = 0; k < 12; ++k) a.p[k] += b.p[k];
  return a;
}
inline WV& operator-=(WV& a, const WV& b)
{
  for (int k = 0; k < 12; ++k) a.p[k] -= b.p[k];
  return a;
}
inline WV operator*(const cd& c, const WV& a)
{
  WV r;
  for (int k = 0; k < 12; ++k) r.p[k] = c * a.p[k];
  return r;
}
inline WV& operator*=(WV& a, const cd& c)
{
  for (int k = 0; k < 12; ++k) a.p[k] *= c;
  return a;
}

// ColorMatrix * WilsonVector (matrix.h:736-750)
inline WV operator*(const CM& cm, const WV& m)
{
  WV ret;
  set_zero(ret);
  for (int s1 = 0; s1 < 4; ++s1)
    for (int c1 = 0; c1 < 3; ++c1)
      for (int c2 = 0; c2 < 3; ++c2)
        ret.p[s1 * 3 + c1] += cm.p[c1 * 3 + c2] * m.p[s1 * 3 + c2];
  return ret;
}

// SpinMatrix * WilsonVector with zero skipping (matrix.h:753-771)
inline WV operator*(const SM& sm, const WV& m)
{
  WV ret;
  set_zero(ret);
  for (int s1 = 0; s1 < 4; ++s1)
    for (int s2 = 0; s2 < 4; ++s2) {
      const cd& e = sm.p[s1 * 4 + s2];
      if (e == 0.0) continue;
      for (int c1 = 0; c1 < 3; ++c1) ret.p[s1 * 3 + c1] += e * m.p[s2 * 3 + c1];
    }
  return ret;
}

inline CM adjoint(const CM& a)
{
  CM r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.p[i * 3 + j] = std::conj(a.p[j * 3 + i]);
  return r;
}

inline SM sm_from(const double re[16], const double im[16])
{
  SM r;
  for (int k = 0; k < 16; ++k) r.p[k] = cd(re[k], im[k]);
  return r;
}
inline SM sm_lin(const cd a, const SM& x, const cd b, const SM& y)
{
  SM r;
  for (int k = 0; k < 16; ++k) r.p[k] = a * x.p[k] + b * y.p[k];
  return r;
}

// Spin constants (matrix.h:426-452).  The "standard" gammas are written as
// the integer patterns of the reference, gamma_{x,y,z} multiplied by -i, and
// CPS gammas = (-g_x, g_y, -g_z, g_t); gamma5 = diag(1,1,-1,-1).
struct SpinConsts {
  SM unit, gamma5, cps[4], pmu_p[4], pmu_m[4], p_p, p_m;
  SpinConsts()
  {
    const cd ii(0.0, 1.0);
    const cd z = 0.0;
    SM g[4];
    const cd gx[16] = {z, z, z, 1.0, z, z, 1.0, z, z, -1.0, z, z, -1.0, z, z, z};
    const cd gy[16] = {z, z, z, -ii, z, z, ii, z, z, ii, z, z, -ii, z, z, z};
    const cd gz[16] = {z, z, 1.0, z, z, z, z, -1.0, -1.0, z, z, z, z, 1.0, z, z};
    const cd gt[16] = {z, z, 1.0, z, z, z, z, 1.0, 1.0, z, z, z, z, 1.0, z, z};
    for (int k = 0; k < 16; ++k) {
      g[0].p[k] = gx[k] * (-ii);
      g[1].p[k] = gy[k] * (-ii);
      g[2].p[k] = gz[k] * (-ii);
      g[3].p[k] = gt[k];
      unit.p[k] = (k % 5 == 0) ? 1.0 : 0.0;
      gamma5.p[k] = 0.0;
    }
    gamma5.p[0] = 1.0;
    gamma5.p[5] = 1.0;
    gamma5.p[10] = -1.0;
    gamma5.p[15] = -1.0;
    for (int k = 0; k < 16; ++k) {
      cps[0].p[k] = -g[0].p[k];
      cps[1].p[k] = g[1].p[k];
      cps[2].p[k] = -g[2].p[k];
      cps[3].p[k] = g[3].p[k];
    }
    // dslash.cpp:345-354
    p_p = sm_lin(0.5, unit, 0.5, gamma5);
    p_m = sm_lin(0.5, unit, -0.5, gamma5);
    for (int mu = 0; mu < 4; ++mu) {
      pmu_p[mu] = sm_lin(0.5, unit, 0.5, cps[mu]);
      pmu_m[mu] = sm_lin(0.5, unit, -0.5, cps[mu]);
    }
  }
};
const SpinConsts& sc()
{
  static SpinConsts s;
  return s;
}

// ---------------------------------------------------------------- geometry
struct Geo {
  int n[4];
  long V;
};

inline void coord_from_index(int x[4], long index, const int n[4])
// coordinate.h:169-180
{
  x[0] = index % n[0];
  index /= n[0];
  x[1] = index % n[1];
  index /= n[1];
  x[2] = index % n[2];
  index /= n[2];
  x[3] = index % n[3];
}
inline int pmod(int a, int n) { return ((a % n) + n) % n; }
inline long index_from_coord(const int xin[4], const int n[4])
// coordinate.h:182-188 after Geometry::mirror (core.h:179-189): a single
// node wraps every direction periodically.
{
  int x[4];
  for (int i = 0; i < 4; ++i) x[i] = pmod(xin[i], n[i]);
  return (((long)x[3] * n[2] + x[2]) * n[1] + x[1]) * n[0] + x[0];
}
inline int eo_from_coord(const int x[4])
// coordinate.h:190-193
{
  return 2 - (x[0] + x[1] + x[2] + x[3] + 16 * 1024 * 1024) % 2;
}
inline void coord_from_half_index(int x[4], long index, int eo, const int n[4])
// core.h:246-261
{
  coord_from_index(x, index * 2, n);
  if (eo_from_coord(x) != eo) coord_from_index(x, index * 2 + 1, n);
}
inline long half_index_from_coord(const int x[4], const int n[4])
// core.h:232-244
{
  return index_from_coord(x, n) / 2;
}

// --------------------------------------------------------------- inverter
struct OInv {
  Geo geo;
  int ls;
  double mass, m5;
  double mobius_scale;
  bool is_multiplying_dminus;
  std::vector<cd> bs, cs;
  std::vector<CM> gf;  // [V][4]
  double stop_rsd;     // InverterParams defaults (dslash.h:114-133)
  long max_num_iter;
  long max_mixed_precision_cycle;
  long last_total_iter;
  int last_cycles;
  long hermop_calls;
};

typedef std::vector<WV> Field;  // [sites][ls]

// ---------------------------------------------------------------- M5 family
void multiply_m_e_e(Field& out, const Field& in, const OInv& fa, long nh)
// dslash.cpp:591-639 (aliasing handled by copying the input)
{
  const SpinConsts& S = sc();
  const int ls = fa.ls;
  const Field in_copy = in;
  out.resize(in.size());
  std::vector<cd> bee(ls), cee(ls);
  for (int m = 0; m < ls; ++m) {
    bee[m] = 1.0 + fa.bs[m] * (4.0 - fa.m5);
    cee[m] = 1.0 - fa.cs[m] * (4.0 - fa.m5);
  }
#pragma omp parallel for
  for (long index = 0; index < nh; ++index) {
    const WV* iv = &in_copy[index * ls];
    WV* v = &out[index * ls];
    for (int m = 0; m < ls; ++m) {
      v[m] = bee[m] * iv[m];
      const WV tmp =
          (S.p_m * (m < ls - 1 ? iv[m + 1] : (WV)(cd(-fa.mass) * iv[0]))) +
          (S.p_p * (m > 0 ? iv[m - 1] : (WV)(cd(-fa.mass) * iv[ls - 1])));
      v[m] -= cee[m] * tmp;
    }
  }
}

void multiply_mdag_e_e(Field& out, const Field& in, const OInv& fa, long nh)
// dslash.cpp:641-692
{
  const SpinConsts& S = sc();
  const int ls = fa.ls;
  const Field in_copy = in;
  out.resize(in.size());
  std::vector<cd> bee(ls), cee(ls);
  for (int m = 0; m < ls; ++m) {
    bee[m] = std::conj(1.0 + fa.bs[m] * (4.0 - fa.m5));
    cee[m] = std::conj(1.0 - fa.cs[m] * (4.0 - fa.m5));
  }
  const cd mc = -std::conj(cd(fa.mass));
#pragma omp parallel for
  for (long index = 0; index < nh; ++index) {
    const WV* iv = &in_copy[index * ls];
    WV* v = &out[index * ls];
    for (int m = 0; m < ls; ++m) {
      v[m] = bee[m] * iv[m];
      const WV tmp =
          (S.p_p * (m < ls - 1 ? (WV)(cee[m + 1] * iv[m + 1])
                               : (WV)((mc * cee[0]) * iv[0]))) +
          (S.p_m * (m > 0 ? (WV)(cee[m - 1] * iv[m - 1])
                          : (WV)((mc * cee[ls - 1]) * iv[ls - 1])));
      v[m] -= tmp;
    }
  }
}

struct LDU {
  std::vector<cd> bee, cee, lee, leem, dee, uee, ueem;
};
LDU make_ldu(const OInv& fa)
// dslash.cpp:712-736 (identical in :789-813)
{
  const int ls = fa.ls;
  LDU c;
  c.bee.resize(ls);
  c.cee.resize(ls);
  for (int m = 0; m < ls; ++m) {
    c.bee[m] = 1.0 + fa.bs[m] * (4.0 - fa.m5);
    c.cee[m] = 1.0 - fa.cs[m] * (4.0 - fa.m5);
  }
  c.lee.resize(ls - 1);
  c.leem.resize(ls - 1);
  for (int m = 0; m < ls - 1; ++m) {
    c.lee[m] = -c.cee[m + 1] / c.bee[m];
    c.leem[m] = m == 0 ? fa.mass * c.cee[ls - 1] / c.bee[0]
                       : c.leem[m - 1] * c.cee[m - 1] / c.bee[m];
  }
  c.dee.assign(ls, 0.0);
  c.dee[ls - 1] = fa.mass * c.cee[ls - 1];
  for (int m = 0; m < ls - 1; ++m) c.dee[ls - 1] *= c.cee[m] / c.bee[m];
  for (int m = 0; m < ls; ++m) c.dee[m] += c.bee[m];
  c.uee.resize(ls - 1);
  c.ueem.resize(ls - 1);
  for (int m = 0; m < ls - 1; ++m) {
    c.uee[m] = -c.cee[m] / c.bee[m];
    c.ueem[m] =
        m == 0 ? fa.mass * c.cee[0] / c.bee[0] : c.ueem[m - 1] * c.cee[m] / c.bee[m];
  }
  return c;
}

void multiply_m_e_e_inv(Field& out, const Field& in, const OInv& fa, long nh)
// dslash.cpp:694-769
{
  const SpinConsts& S = sc();
  const int ls = fa.ls;
  const LDU c = make_ldu(fa);
  if (&out != &in) out = in;
#pragma omp parallel for
  for (long index = 0; index < nh; ++index) {
    WV* v = &out[index * ls];
    WV tmp;
    set_zero(tmp);
    for (int m = 0; m < ls - 1; ++m) tmp += (-c.leem[m]) * v[m];
    v[ls - 1] += S.p_m * tmp;
    for (int m = 1; m < ls; ++m) v[m] += (-c.lee[m - 1]) * (S.p_p * v[m - 1]);
    for (int m = 0; m < ls; ++m) v[m] *= (1.0 / c.dee[m]);
    for (int m = ls - 2; m >= 0; --m) v[m] += (-c.uee[m]) * (S.p_m * v[m + 1]);
    for (int m = 0; m < ls - 1; ++m) v[m] += (-c.ueem[m]) * (S.p_p * v[ls - 1]);
  }
}

void multiply_mdag_e_e_inv(Field& out, const Field& in, const OInv& fa, long nh)
// dslash.cpp:771-857
{
  const SpinConsts& S = sc();
  const int ls = fa.ls;
  LDU c = make_ldu(fa);
  for (int m = 0; m < ls; ++m) {
    c.bee[m] = std::conj(c.bee[m]);
    c.cee[m] = std::conj(c.cee[m]);
    c.dee[m] = std::conj(c.dee[m]);
  }
  for (int m = 0; m < ls - 1; ++m) {
    c.lee[m] = std::conj(c.lee[m]);
    c.leem[m] = std::conj(c.leem[m]);
    c.uee[m] = std::conj(c.uee[m]);
    c.ueem[m] = std::conj(c.ueem[m]);
  }
  if (&out != &in) out = in;
#pragma omp parallel for
  for (long index = 0; index < nh; ++index) {
    WV* v = &out[index * ls];
    WV tmp;
    set_zero(tmp);
    for (int m = 0; m < ls - 1; ++m) tmp += (-c.ueem[m]) * v[m];
    v[ls - 1] += S.p_p * tmp;
    for (int m = 1; m < ls; ++m) v[m] += (-c.uee[m - 1]) * (S.p_m * v[m - 1]);
    for (int m = 0; m < ls; ++m) v[m] *= (1.0 / c.dee[m]);
    for (int m = ls - 2; m >= 0; --m) v[m] += (-c.lee[m]) * (S.p_p * v[m + 1]);
    for (int m = 0; m < ls - 1; ++m) v[m] += (-c.leem[m]) * (S.p_m * v[ls - 1]);
  }
}

// ---------------------------------------------------------------- hopping
void wilson_hop_e_o(Field& out, const Field& in, const OInv& inv, int eo_in,
                    bool dag)
// multiply_wilson_d_e_o_no_comm (dslash.cpp:859-908) and, with dag,
// multiply_wilson_ddag_e_o_no_comm (dslash.cpp:910-959).  Output parity is
// 3 - eo_in.
{
  const SpinConsts& S = sc();
  const int ls = inv.ls;
  const int* n = inv.geo.n;
  const long nh = inv.geo.V / 2;
  const int eo_out = 3 - eo_in;
  out.assign(nh * ls, WV());
#pragma omp parallel for
  for (long index = 0; index < nh; ++index) {
    int xl[4];
    coord_from_half_index(xl, index, eo_out, n);
    WV* v = &out[index * ls];
    for (int m = 0; m < ls; ++m) set_zero(v[m]);
    for (int mu = 0; mu < 4; ++mu) {
      int xp[4] = {xl[0], xl[1], xl[2], xl[3]};
      int xm[4] = {xl[0], xl[1], xl[2], xl[3]};
      xp[mu] += 1;
      xm[mu] -= 1;
      const CM& u_p = inv.gf[index_from_coord(xl, n) * 4 + mu];
      const CM u_m = adjoint(inv.gf[index_from_coord(xm, n) * 4 + mu]);
      const WV* iv_p = &in[half_index_from_coord(xp, n) * ls];
      const WV* iv_m = &in[half_index_from_coord(xm, n) * ls];
      const SM& pf = dag ? S.pmu_p[mu] : S.pmu_m[mu];
      const SM& pb = dag ? S.pmu_m[mu] : S.pmu_p[mu];
      for (int m = 0; m < ls; ++m) {
        v[m] -= u_p * (pf * iv_p[m]);
        v[m] -= u_m * (pb * iv_m[m]);
      }
    }
  }
}

void multiply_m_e_o(Field& out, const Field& in, const OInv& fa, int eo_in)
// dslash.cpp:961-1003
{
  const SpinConsts& S = sc();
  const int ls = fa.ls;
  const long nh = fa.geo.V / 2;
  Field in1(in.size());
  std::vector<cd> beo(ls), ceo(ls);
  for (int m = 0; m < ls; ++m) {
    beo[m] = fa.bs[m];
    ceo[m] = -fa.cs[m];
  }
#pragma omp parallel for
  for (long index = 0; index < nh; ++index) {
    const WV* iv = &in[index * ls];
    WV* v = &in1[index * ls];
    for (int m = 0; m < ls; ++m) {
      v[m] = beo[m] * iv[m];
      const WV tmp =
          (S.p_m * (m < ls - 1 ? iv[m + 1] : (WV)(cd(-fa.mass) * iv[0]))) +
          (S.p_p * (m > 0 ? iv[m - 1] : (WV)(cd(-fa.mass) * iv[ls - 1])));
      v[m] -= ceo[m] * tmp;
    }
  }
  wilson_hop_e_o(out, in1, fa, eo_in, false);
}

void multiply_mdag_e_o(Field& out, const Field& in, const OInv& fa, int eo_in)
// dslash.cpp:1005-1060
{
  const SpinConsts& S = sc();
  const int ls = fa.ls;
  const long nh = fa.geo.V / 2;
  std::vector<cd> beo(ls), ceo(ls);
  for (int m = 0; m < ls; ++m) {
    beo[m] = std::conj(fa.bs[m]);
    ceo[m] = std::conj(-fa.cs[m]);
  }
  Field out1;
  wilson_hop_e_o(out1, in, fa, eo_in, true);
  out.resize(out1.size());
  const cd mc = -std::conj(cd(fa.mass));
#pragma omp parallel for
  for (long index = 0; index < nh; ++index) {
    const WV* iv = &out1[index * ls];
    WV* v = &out[index * ls];
    for (int m = 0; m < ls; ++m) {
      v[m] = beo[m] * iv[m];
      const WV tmp =
          (S.p_p * (m < ls - 1 ? (WV)(ceo[m + 1] * iv[m + 1])
                               : (WV)((mc * ceo[0]) * iv[0]))) +
          (S.p_m * (m > 0 ? (WV)(ceo[m - 1] * iv[m - 1])
                          : (WV)((mc * ceo[ls - 1]) * iv[ls - 1])));
      v[m] -= tmp;
    }
  }
}

// ------------------------------------------------------- full-lattice ops
void get_half_fermion(Field& half, const Field& ff, int eo, const OInv& inv)
// dslash.cpp:542-561
{
  const int ls = inv.ls;
  const long nh = inv.geo.V / 2;
  half.resize(nh * ls);
#pragma omp parallel for
  for (long index = 0; index < nh; ++index) {
    int x[4];
    coord_from_half_index(x, index, eo, inv.geo.n);
    const long full = index_from_coord(x, inv.geo.n);
    for (int m = 0; m < ls; ++m) half[index * ls + m] = ff[full * ls + m];
  }
}

void set_half_fermion(Field& ff, const Field& half, int eo, const OInv& inv)
// dslash.cpp:563-580
{
  const int ls = inv.ls;
  const long nh = inv.geo.V / 2;
  ff.resize(inv.geo.V * ls);
#pragma omp parallel for
  for (long index = 0; index < nh; ++index) {
    int x[4];
    coord_from_half_index(x, index, eo, inv.geo.n);
    const long full = index_from_coord(x, inv.geo.n);
    for (int m = 0; m < ls; ++m) ff[full * ls + m] = half[index * ls + m];
  }
}

void wilson_d_full(Field& out, const Field& in, const OInv& inv, double mass)
// multiply_wilson_d_no_comm (dslash.cpp:400-446)
{
  const SpinConsts& S = sc();
  const int ls = inv.ls;
  const int* n = inv.geo.n;
  const long V = inv.geo.V;
  out.assign(V * ls, WV());
#pragma omp parallel for
  for (long index = 0; index < V; ++index) {
    int xl[4];
    coord_from_index(xl, index, n);
    WV* v = &out[index * ls];
    const WV* iv = &in[index * ls];
    for (int m = 0; m < ls; ++m) v[m] = cd(4.0 + mass) * iv[m];
    for (int mu = 0; mu < 4; ++mu) {
      int xp[4] = {xl[0], xl[1], xl[2], xl[3]};
      int xm[4] = {xl[0], xl[1], xl[2], xl[3]};
      xp[mu] += 1;
      xm[mu] -= 1;
      const CM& u_p = inv.gf[index * 4 + mu];
      const CM u_m = adjoint(inv.gf[index_from_coord(xm, n) * 4 + mu]);
      const WV* iv_p = &in[index_from_coord(xp, n) * ls];
      const WV* iv_m = &in[index_from_coord(xm, n) * ls];
      for (int m = 0; m < ls; ++m) {
        v[m] -= u_p * (S.pmu_m[mu] * iv_p[m]);
        v[m] -= u_m * (S.pmu_p[mu] * iv_m[m]);
      }
    }
  }
}

void multiply_d_minus(Field& out, const Field& in, const OInv& fa)
// dslash.cpp:448-480
{
  const int ls = fa.ls;
  const long V = fa.geo.V;
  Field in1(V * ls), out1(V * ls);
#pragma omp parallel for
  for (long index = 0; index < V; ++index)
    for (int m = 0; m < ls; ++m) {
      in1[index * ls + m] = (-fa.cs[m]) * in[index * ls + m];
      out1[index * ls + m] = in[index * ls + m];
    }
  wilson_d_full(out, in1, fa, -fa.m5);
  for (long k = 0; k < V * ls; ++k) out[k] += out1[k];
}

void multiply_m_full(Field& out, const Field& in, const OInv& fa)
// dslash.cpp:490-540
{
  const SpinConsts& S = sc();
  const int ls = fa.ls;
  const long V = fa.geo.V;
  Field in1(V * ls), fftmp(V * ls);
#pragma omp parallel for
  for (long index = 0; index < V; ++index) {
    const WV* iv = &in[index * ls];
    WV* v = &fftmp[index * ls];
    WV* v1 = &in1[index * ls];
    for (int m = 0; m < ls; ++m) {
      v1[m] = fa.bs[m] * iv[m];
      v[m] = iv[m];
      const WV tmp =
          (S.p_m * (m < ls - 1 ? iv[m + 1] : (WV)(cd(-fa.mass) * iv[0]))) +
          (S.p_p * (m > 0 ? iv[m - 1] : (WV)(cd(-fa.mass) * iv[ls - 1])));
      v1[m] += fa.cs[m] * tmp;
      v[m] -= tmp;
    }
  }
  wilson_d_full(out, in1, fa, -fa.m5);
  for (long k = 0; k < V * ls; ++k) out[k] += fftmp[k];
}

void multiply_m_dwf(Field& out, const Field& in, const OInv& fa)
// multiply_m_dwf_no_comm (dslash.cpp:329-386); Shamir only (:342)
{
  const SpinConsts& S = sc();
  const int ls = fa.ls;
  const int* n = fa.geo.n;
  const long V = fa.geo.V;
  out.assign(V * ls, WV());
#pragma omp parallel for
  for (long index = 0; index < V; ++index) {
    int xl[4];
    coord_from_index(xl, index, n);
    WV* v = &out[index * ls];
    const WV* iv = &in[index * ls];
    for (int m = 0; m < ls; ++m) {
      v[m] = cd(5.0 - fa.m5) * iv[m];
      v[m] -= S.p_m * (m < ls - 1 ? iv[m + 1] : (WV)(cd(-fa.mass) * iv[0]));
      v[m] -= S.p_p * (m > 0 ? iv[m - 1] : (WV)(cd(-fa.mass) * iv[ls - 1]));
    }
    for (int mu = 0; mu < 4; ++mu) {
      int xp[4] = {xl[0], xl[1], xl[2], xl[3]};
      int xm[4] = {xl[0], xl[1], xl[2], xl[3]};
      xp[mu] += 1;
      xm[mu] -= 1;
      const CM& u_p = fa.gf[index * 4 + mu];
      const CM u_m = adjoint(fa.gf[index_from_coord(xm, n) * 4 + mu]);
      const WV* iv_p = &in[index_from_coord(xp, n) * ls];
      const WV* iv_m = &in[index_from_coord(xm, n) * ls];
      for (int m = 0; m < ls; ++m) {
        v[m] -= u_p * (S.pmu_m[mu] * iv_p[m]);
        v[m] -= u_m * (S.pmu_p[mu] * iv_m[m]);
      }
    }
  }
}

void add_to(Field& a, const Field& b)
{
  for (size_t k = 0; k < a.size(); ++k) a[k] += b[k];
}
void sub_from(Field& a, const Field& b)
{
  for (size_t k = 0; k < a.size(); ++k) a[k] -= b[k];
}

void multiply_m(Field& out, const Field& in, const OInv& fa, bool dag)
// multiply_m / multiply_mdag (dslash.cpp:1062-1100)
{
  const long nh = fa.geo.V / 2;
  Field in_e, in_o, tmp_e, tmp_o, out_e, out_o;
  get_half_fermion(in_e, in, 2, fa);
  get_half_fermion(in_o, in, 1, fa);
  if (!dag) {
    multiply_m_e_o(tmp_e, in_o, fa, 1);
    multiply_m_e_e(out_e, in_e, fa, nh);
    multiply_m_e_e(tmp_o, in_o, fa, nh);
    multiply_m_e_o(out_o, in_e, fa, 2);
  } else {
    multiply_mdag_e_o(tmp_e, in_o, fa, 1);
    multiply_mdag_e_e(out_e, in_e, fa, nh);
    multiply_mdag_e_e(tmp_o, in_o, fa, nh);
    multiply_mdag_e_o(out_o, in_e, fa, 2);
  }
  add_to(out_e, tmp_e);
  add_to(out_o, tmp_o);
  set_half_fermion(out, out_e, 2, fa);
  set_half_fermion(out, out_o, 1, fa);
}

// -------------------------------------------------- Schur-complement ops
void multiply_mpc_sym2(Field& out, const Field& in, const OInv& fa, int eo)
// dslash.cpp:1102-1119
{
  const long nh = fa.geo.V / 2;
  Field tmp;
  multiply_m_e_e_inv(tmp, in, fa, nh);
  Field tmp2;
  multiply_m_e_o(tmp2, tmp, fa, eo);
  multiply_m_e_e_inv(tmp2, tmp2, fa, nh);
  multiply_m_e_o(tmp, tmp2, fa, 3 - eo);
  Field res = in;
  sub_from(res, tmp);
  out.swap(res);
}

void multiply_mpcdag_sym2(Field& out, const Field& in, const OInv& fa, int eo)
// dslash.cpp:1121-1138
{
  const long nh = fa.geo.V / 2;
  Field tmp;
  multiply_mdag_e_o(tmp, in, fa, eo);
  multiply_mdag_e_e_inv(tmp, tmp, fa, nh);
  Field tmp2;
  multiply_mdag_e_o(tmp2, tmp, fa, 3 - eo);
  multiply_mdag_e_e_inv(tmp2, tmp2, fa, nh);
  Field res = in;
  sub_from(res, tmp2);
  out.swap(res);
}

void multiply_hermop_sym2(Field& out, const Field& in, const OInv& fa, int eo)
// dslash.cpp:1140-1149
{
  multiply_mpc_sym2(out, in, fa, eo);
  multiply_mpcdag_sym2(out, out, fa, eo);
}

void multiply_m_with_prec_sym2(Field& out, const Field& in, const OInv& inv)
// dslash.cpp:1271-1296
{
  const long nh = inv.geo.V / 2;
  Field in_e, in_o, out_e, out_o;
  get_half_fermion(in_e, in, 2, inv);
  get_half_fermion(in_o, in, 1, inv);
  multiply_m_e_o(out_e, in_o, inv, 1);
  multiply_m_e_e_inv(out_e, out_e, inv, nh);
  add_to(out_e, in_e);
  multiply_m_e_e(out_o, in_o, inv, nh);
  multiply_m_e_e(out_e, out_e, inv, nh);
  multiply_mpc_sym2(out_o, out_o, inv, 1);
  Field tmp, tmp2;
  multiply_m_e_e_inv(tmp, out_e, inv, nh);
  multiply_m_e_o(tmp2, tmp, inv, 2);
  add_to(out_o, tmp2);
  set_half_fermion(out, out_e, 2, inv);
  set_half_fermion(out, out_o, 1, inv);
}

// ------------------------------------------------------------- BLAS-1
cd dot_product(const Field& a, const Field& b)
// dslash.cpp:1298-1321: a^dagger b, serial loop
{
  cd sum = 0.0;
  for (size_t k = 0; k < a.size(); ++k)
    for (int c = 0; c < 12; ++c) sum += std::conj(a[k].p[c]) * b[k].p[c];
  return sum;
}

double qnorm(const Field& a)
// field.h:170-195 (per-thread partial sums combined in thread order)
{
  double sum = 0.0;
  const long n = a.size();
#pragma omp parallel
  {
    double psum = 0.0;
#pragma omp for nowait schedule(static)
    for (long k = 0; k < n; ++k)
      for (int c = 0; c < 12; ++c) psum += std::norm(a[k].p[c]);
    for (int i = 0; i < omp_get_num_threads(); ++i) {
#pragma omp barrier
      if (omp_get_thread_num() == i) sum += psum;
    }
  }
  return sum;
}

// ------------------------------------------------------------- solver
long cg_with_f(Field& out, const Field& in, const OInv& inv, int eo,
               double stop_rsd, long max_num_iter, bool verbose)
// cg_with_f (dslash.h:331-399) with f = multiply_hermop_sym2
// (cg_with_herm_sym_2, dslash.cpp:1323-1333).  `out` is the initial guess.
{
  if (max_num_iter == 0) return 0;
  const size_t n = in.size();
  if (out.size() != n) out.assign(n, WV());
  Field r = in, p, tmp, ap;
  multiply_hermop_sym2(tmp, out, inv, eo);
  const_cast<OInv&>(inv).hermop_calls += 1;
  sub_from(r, tmp);
  p = r;
  const double qnorm_in = qnorm(in);
  double qnorm_r = qnorm(r);
  for (long iter = 1; iter <= max_num_iter; ++iter) {
    multiply_hermop_sym2(ap, p, inv, eo);
    const_cast<OInv&>(inv).hermop_calls += 1;
    const double alpha = qnorm_r / dot_product(p, ap).real();
    tmp = p;
    for (auto& w : tmp) w *= alpha;
    add_to(out, tmp);
    tmp = ap;
    for (auto& w : tmp) w *= alpha;
    sub_from(r, tmp);
    const double new_qnorm_r = qnorm(r);
    if (verbose)
      std::printf("oracle cg iter=%4ld sqrt(qnorm_r/qnorm_in)=%.3E\n", iter,
                  std::sqrt(new_qnorm_r / qnorm_in));
    if (new_qnorm_r <= qnorm_in * stop_rsd * stop_rsd) return iter;
    const double beta = new_qnorm_r / qnorm_r;
    for (size_t k = 0; k < n; ++k) p[k] = beta * p[k] + r[k];
    qnorm_r = new_qnorm_r;
  }
  return max_num_iter + 1;
}

void set_odd_prec_field_sym2(Field& in_o_p, Field& out_e_p, const Field& in,
                             const OInv& inv)
// dslash.h:401-417
{
  const long nh = inv.geo.V / 2;
  Field in_e, tmp;
  get_half_fermion(in_e, in, 2, inv);
  get_half_fermion(in_o_p, in, 1, inv);
  multiply_m_e_e_inv(out_e_p, in_e, inv, nh);
  multiply_m_e_o(tmp, out_e_p, inv, 2);
  sub_from(in_o_p, tmp);
  multiply_mpcdag_sym2(in_o_p, in_o_p, inv, 1);
}

void restore_field_from_odd_prec_sym2(Field& out, const Field& out_o_p,
                                      const Field& out_e_p, const OInv& inv)
// dslash.h:419-438
{
  const long nh = inv.geo.V / 2;
  Field out_e = out_e_p, out_o = out_o_p, tmp;
  multiply_m_e_e_inv(out_o, out_o, inv, nh);
  multiply_m_e_o(tmp, out_o, inv, 1);
  multiply_m_e_e_inv(tmp, tmp, inv, nh);
  sub_from(out_e, tmp);
  out.resize(inv.geo.V * inv.ls);
  set_half_fermion(out, out_e, 2, inv);
  set_half_fermion(out, out_o, 1, inv);
}

long invert_with_cg(Field& out, const Field& in, OInv& inv, bool verbose, double stop_rsd = -1,
                    bool dminus_multiplied_already = false)
// invert_with_cg (dslash.h:440-530) with cg = cg_with_herm_sym_2 and the
// default cg_diagonal_mee == 2; invert(5d) (dslash.cpp:1335-1340).
{
  if (stop_rsd < 0) stop_rsd = inv.stop_rsd;
  const long max_num_iter = inv.max_num_iter;
  const long max_cycle = inv.max_mixed_precision_cycle;
  Field dm_in;
  if (!dminus_multiplied_already && inv.is_multiplying_dminus) {
    multiply_d_minus(dm_in, in, inv);
  } else {
    dm_in = in;
  }
  const double dm_in_qnorm = qnorm(dm_in);
  if (dm_in_qnorm == 0.0) {
    out.assign(inv.geo.V * inv.ls, WV());
    inv.last_total_iter = 0;
    inv.last_cycles = 0;
    return 0;
  }
  long total_iter = 0;
  Field in_o_p, out_e_p, out_o_p;
  set_odd_prec_field_sym2(in_o_p, out_e_p, dm_in, inv);
  out_o_p.assign(in_o_p.size(), WV());
  const double qnorm_in_o_p = qnorm(in_o_p);
  Field tmp, itmp = in_o_p;
  double qnorm_itmp = qnorm_in_o_p;
  int cycle;
  for (cycle = 1; cycle <= max_cycle; ++cycle) {
    tmp.assign(in_o_p.size(), WV());
    const long iter = cg_with_f(tmp, itmp, inv, 1,
                                stop_rsd * std::sqrt(qnorm_in_o_p / qnorm_itmp),
                                max_num_iter, verbose);
    total_iter += iter;
    add_to(out_o_p, tmp);
    if (iter <= max_num_iter) break;
    multiply_hermop_sym2(itmp, out_o_p, inv, 1);
    inv.hermop_calls += 1;
    for (auto& w : itmp) w *= -1.0;
    add_to(itmp, in_o_p);
    qnorm_itmp = qnorm(itmp);
  }
  inv.last_total_iter = total_iter;
  inv.last_cycles = cycle;
  restore_field_from_odd_prec_sym2(out, out_o_p, out_e_p, inv);
  return total_iter;
}

long invert_with_cg_with_guess(Field& out, const Field& in, OInv& inv, bool verbose)
// invert_with_cg_with_guess (dslash.h:532-573) with cg = cg_with_herm_sym_2
// and the inverter's parameters: `out` holds the guess on entry.
{
  Field dm_in;
  if (inv.is_multiplying_dminus) {
    multiply_d_minus(dm_in, in, inv);
  } else {
    dm_in = in;
  }
  const double qnorm_dm_in = qnorm(dm_in);
  Field tmp;
  multiply_m(tmp, out, inv, false);
  sub_from(dm_in, tmp);
  const double qnorm_dm_in_sub = qnorm(dm_in);
  const double ratio = std::sqrt(qnorm_dm_in / qnorm_dm_in_sub);
  const long total_iter = invert_with_cg(tmp, dm_in, inv, verbose, inv.stop_rsd * ratio, true);
  add_to(out, tmp);
  return total_iter;
}

void ff5d_from_4d(Field& ff5d, const WV* ff4d, int upper, int lower,
                  const OInv& inv)
// fermion_field_5d_from_4d (qcd-prop.h:96-116): upper (right-handed) half
// of the WilsonVector goes to s=upper, the lower half to s=lower.
{
  const int ls = inv.ls;
  const long V = inv.geo.V;
  ff5d.assign(V * ls, WV());
  for (long index = 0; index < V; ++index) {
    for (int k = 0; k < 6; ++k) ff5d[index * ls + upper].p[k] = ff4d[index].p[k];
    for (int k = 6; k < 12; ++k) ff5d[index * ls + lower].p[k] = ff4d[index].p[k];
  }
}

void ff4d_from_5d(WV* ff4d, const Field& ff5d, int upper, int lower,
                  const OInv& inv)
// fermion_field_4d_from_5d (qcd-prop.h:118-138)
{
  const int ls = inv.ls;
  const long V = inv.geo.V;
  for (long index = 0; index < V; ++index) {
    for (int k = 0; k < 6; ++k) ff4d[index].p[k] = ff5d[index * ls + upper].p[k];
    for (int k = 6; k < 12; ++k) ff4d[index].p[k] = ff5d[index * ls + lower].p[k];
  }
}

Field field_from(const double* p, size_t nwv)
{
  Field f(nwv);
  std::memcpy((void*)f.data(), p, nwv * sizeof(WV));
  return f;
}
void field_to(double* p, const Field& f)
{
  std::memcpy(p, (const void*)f.data(), f.size() * sizeof(WV));
}

}  // namespace

// =================================================================== C ABI
extern "C" {

void* oracle_inv_create(const int* total_site, const double* gauge, int ls,
                        double mass, double m5, double mobius_scale,
                        const double* bs, const double* cs,
                        int is_multiplying_dminus)
// InverterDomainWall::setup(gf, fa) (dslash.h:154-162) + FermionAction
// (fermion-action.h:36-52).  bs/cs are complex (re,im) arrays of length ls;
// for Moebius they are 0.5*scale+0.5 and 0.5*scale-0.5, for zMoebius
// bs = 0.5*(1/omega+1), cs = bs-1 (cqlat/fermion-action.cpp:17-37).
{
  OInv* inv = new OInv;
  long V = 1;
  for (int i = 0; i < 4; ++i) {
    inv->geo.n[i] = total_site[i];
    V *= total_site[i];
  }
  inv->geo.V = V;
  inv->ls = ls;
  inv->mass = mass;
  inv->m5 = m5;
  inv->mobius_scale = mobius_scale;
  inv->is_multiplying_dminus = is_multiplying_dminus != 0;
  inv->bs.resize(ls);
  inv->cs.resize(ls);
  for (int m = 0; m < ls; ++m) {
    inv->bs[m] = cd(bs[2 * m], bs[2 * m + 1]);
    inv->cs[m] = cd(cs[2 * m], cs[2 * m + 1]);
  }
  inv->gf.resize(V * 4);
  std::memcpy((void*)inv->gf.data(), gauge, V * 4 * sizeof(CM));
  inv->stop_rsd = 1.0e-8;
  inv->max_num_iter = 200;
  inv->max_mixed_precision_cycle = 300;
  inv->last_total_iter = 0;
  inv->last_cycles = 0;
  inv->hermop_calls = 0;
  return inv;
}

void oracle_inv_destroy(void* p) { delete (OInv*)p; }

void oracle_inv_set_params(void* p, double stop_rsd, long max_num_iter,
                           long max_mixed_precision_cycle)
{
  OInv* inv = (OInv*)p;
  inv->stop_rsd = stop_rsd;
  inv->max_num_iter = max_num_iter;
  inv->max_mixed_precision_cycle = max_mixed_precision_cycle;
}

void oracle_inv_stats(void* p, long* total_iter, int* cycles, long* hermop_calls)
{
  OInv* inv = (OInv*)p;
  *total_iter = inv->last_total_iter;
  *cycles = inv->last_cycles;
  *hermop_calls = inv->hermop_calls;
}

int oracle_num_threads(void) { return omp_get_max_threads(); }

// op codes for oracle_apply_half (half fields [V/2][ls][12] complex)
//  0 m_e_e, 1 mdag_e_e, 2 m_e_e_inv, 3 mdag_e_e_inv, 4 m_e_o, 5 mdag_e_o,
//  6 mpc_sym2, 7 mpcdag_sym2, 8 hermop_sym2, 9 wilson hop H, 10 hop H^dagger
int oracle_apply_half(void* p, int op, double* out, const double* in, int eo_in)
{
  OInv* inv = (OInv*)p;
  const long nh = inv->geo.V / 2;
  Field fin = field_from(in, nh * inv->ls), fout;
  switch (op) {
    case 0: multiply_m_e_e(fout, fin, *inv, nh); break;
    case 1: multiply_mdag_e_e(fout, fin, *inv, nh); break;
    case 2: multiply_m_e_e_inv(fout, fin, *inv, nh); break;
    case 3: multiply_mdag_e_e_inv(fout, fin, *inv, nh); break;
    case 4: multiply_m_e_o(fout, fin, *inv, eo_in); break;
    case 5: multiply_mdag_e_o(fout, fin, *inv, eo_in); break;
    case 6: multiply_mpc_sym2(fout, fin, *inv, eo_in); break;
    case 7: multiply_mpcdag_sym2(fout, fin, *inv, eo_in); break;
    case 8: multiply_hermop_sym2(fout, fin, *inv, eo_in); break;
    case 9: wilson_hop_e_o(fout, fin, *inv, eo_in, false); break;
    case 10: wilson_hop_e_o(fout, fin, *inv, eo_in, true); break;
    default: return -1;
  }
  field_to(out, fout);
  return 0;
}

// op codes for oracle_apply_full (full fields [V][ls][12] complex)
//  0 multiply_m, 1 multiply_mdag, 2 multiply_d_minus, 3 multiply_m_full,
//  4 multiply_m_dwf (Shamir), 5 multiply_m_with_prec_sym2,
//  6 multiply_wilson_d (mass = -m5)
int oracle_apply_full(void* p, int op, double* out, const double* in)
{
  OInv* inv = (OInv*)p;
  const long V = inv->geo.V;
  Field fin = field_from(in, V * inv->ls), fout;
  switch (op) {
    case 0: multiply_m(fout, fin, *inv, false); break;
    case 1: multiply_m(fout, fin, *inv, true); break;
    case 2: multiply_d_minus(fout, fin, *inv); break;
    case 3: multiply_m_full(fout, fin, *inv); break;
    case 4:
      if (inv->mobius_scale != 1.0) return -2;  // dslash.cpp:342
      multiply_m_dwf(fout, fin, *inv);
      break;
    case 5: multiply_m_with_prec_sym2(fout, fin, *inv); break;
    case 6: wilson_d_full(fout, fin, *inv, -inv->m5); break;
    default: return -1;
  }
  field_to(out, fout);
  return 0;
}

long oracle_cg_with_herm_sym_2(void* p, double* sol, const double* src,
                               double stop_rsd, long max_num_iter, int verbose)
// cg_with_herm_sym_2 (dslash.cpp:1323-1333) on odd half fields; `sol` holds
// the initial guess on entry.
{
  OInv* inv = (OInv*)p;
  const long n = inv->geo.V / 2 * inv->ls;
  Field fsrc = field_from(src, n), fsol = field_from(sol, n);
  const long it = cg_with_f(fsol, fsrc, *inv, 1, stop_rsd, max_num_iter, verbose);
  field_to(sol, fsol);
  return it;
}

long oracle_invert_5d(void* p, double* out, const double* in, int verbose)
// invert(FermionField5d&, ...) (dslash.cpp:1335-1340)
{
  OInv* inv = (OInv*)p;
  const long n = inv->geo.V * inv->ls;
  Field fin = field_from(in, n), fout;
  const long it = invert_with_cg(fout, fin, *inv, verbose);
  field_to(out, fout);
  return it;
}

long oracle_invert_5d_with_guess(void* p, double* out, const double* in, int verbose)
// invert_with_cg_with_guess (dslash.h:532-573): `out` holds the guess
{
  OInv* inv = (OInv*)p;
  const long n = inv->geo.V * inv->ls;
  Field fin = field_from(in, n), fout = field_from(out, n);
  const long it = invert_with_cg_with_guess(fout, fin, *inv, verbose);
  field_to(out, fout);
  return it;
}

long oracle_invert_4d(void* p, double* out, const double* in, int verbose)
// invert(FermionField4d&, ...) (dslash.cpp:1342-1347) -> invert_dwf
// (qcd-prop.h:146-167)
{
  OInv* inv = (OInv*)p;
  const int ls = inv->ls;
  Field src5d, sol5d;
  ff5d_from_4d(src5d, (const WV*)in, 0, ls - 1, *inv);
  const long it = invert_with_cg(sol5d, src5d, *inv, verbose);
  ff4d_from_5d((WV*)out, sol5d, ls - 1, 0, *inv);
  return it;
}

long oracle_invert_prop(void* p, double* sol, const double* src, int verbose)
// invert(Propagator4d&, ...) (qcd-prop.h:169-186): 12 columns solved one
// after the other.  Prop layout [site][12 rows][12 cols] complex.
{
  OInv* inv = (OInv*)p;
  const long V = inv->geo.V;
  std::vector<WV> ff_src(V), ff_sol(V);
  long total = 0;
  for (int j = 0; j < 12; ++j) {
    for (long x = 0; x < V; ++x)
      for (int r = 0; r < 12; ++r)
        ff_src[x].p[r] = cd(src[((x * 12 + r) * 12 + j) * 2],
                            src[((x * 12 + r) * 12 + j) * 2 + 1]);
    total += oracle_invert_4d(p, (double*)ff_sol.data(),
                              (const double*)ff_src.data(), verbose);
    for (long x = 0; x < V; ++x)
      for (int r = 0; r < 12; ++r) {
        sol[((x * 12 + r) * 12 + j) * 2] = ff_sol[x].p[r].real();
        sol[((x * 12 + r) * 12 + j) * 2 + 1] = ff_sol[x].p[r].imag();
      }
  }
  return total;
}

double oracle_find_max_eigen_value_hermop_sym2(void* p, const double* start,
                                               long max_iter)
// find_max_eigen_value_hermop_sym2 (dslash.cpp:1349-1369) from a caller-
// supplied start vector (the reference draws it from its own RNG).
{
  OInv* inv = (OInv*)p;
  const long n = inv->geo.V / 2 * inv->ls;
  Field ff = field_from(start, n);
  const double s0 = 1.0 / std::sqrt(qnorm(ff));
  for (auto& w : ff) w *= s0;
  double ratio = 1.0;
  for (long i = 0; i < max_iter; ++i) {
    multiply_hermop_sym2(ff, ff, *inv, 1);
    ratio = std::sqrt(qnorm(ff));
    for (auto& w : ff) w *= 1.0 / ratio;
  }
  return ratio;
}

void oracle_dot_product(const double* a, const double* b, long nwv, double* re,
                        double* im)
// dslash.cpp:1298-1321
{
  Field fa = field_from(a, nwv), fb = field_from(b, nwv);
  const cd d = dot_product(fa, fb);
  *re = d.real();
  *im = d.imag();
}

double oracle_qnorm(const double* a, long nwv)
// field.h:170-195
{
  Field fa = field_from(a, nwv);
  return qnorm(fa);
}

}  // extern "C"
