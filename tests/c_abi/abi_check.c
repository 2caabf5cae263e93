/*
 * A C program against include/qlat_dwf.h alone: the boundary as a C caller
 * (a cgo / Fortran / plain-C host) sees it.  Built by the tests with
 *   gcc -std=c99 -Wall -Wextra -Werror -pedantic -I include ...
 *
 *   abi_check cpu   error paths and process-wide flags only (no GPU call)
 *   abi_check gpu   create an inverter on the unit gauge field of a 4^3x8
 *                   lattice (Shamir, Ls 16, M5 1.0, mass 0.05 -- the
 *                   reference's examples-py/free-invert.py setup), check
 *                   that Mpc^+ Mpc is Hermitian and positive on two seeded
 *                   vectors, solve the 12-column point-source propagator
 *                   with invert_prop, compare its qnorm with the
 *                   reference's golden 9.736817290497456 (eps 1e-10,
 *                   free-invert.log.json), read the solve statistics, and
 *                   destroy the inverter.
 * Exit status 0 on success; every failure prints one line.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qlat_dwf.h"

static int fail(const char* what, int rc)
{
  printf("FAIL %s (status %d): %s\n", what, rc, qlat_dwf_last_error());
  return 1;
}

/* a deterministic uniform stream in [-1, 1) (64-bit LCG, Knuth's MMIX constants) */
static double next_uniform(uint64_t* s)
{
  *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(*s >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static int cpu_checks(void)
{
  int rc;
  double s = 0;
  long n = 0, c = 0;
  if (qlat_dwf_abi_version() != 4) {
    printf("FAIL abi version %d\n", qlat_dwf_abi_version());
    return 1;
  }
  rc = qlat_dwf_inverter_set_params(NULL, 1e-8, 10, 10);
  if (rc != QLAT_DWF_EINVAL || strstr(qlat_dwf_last_error(), "null") == NULL) return fail("null handle", rc);
  rc = qlat_dwf_inverter_get_params(NULL, &s, &n, &c);
  if (rc != QLAT_DWF_EINVAL) return fail("null handle (get)", rc);
  if (qlat_dwf_is_checking_invert() != 1) return fail("is_checking_invert default", 0);
  if ((rc = qlat_dwf_set_checking_invert(0)) != QLAT_DWF_OK || qlat_dwf_is_checking_invert() != 0)
    return fail("set_checking_invert(0)", rc);
  if ((rc = qlat_dwf_set_checking_invert(1)) != QLAT_DWF_OK || qlat_dwf_is_checking_invert() != 1)
    return fail("set_checking_invert(1)", rc);
  printf("OK cpu\n");
  return 0;
}

static double cdot_re(const double* a, const double* b, size_t n, double* im)
{
  double re = 0, ii = 0;
  size_t k;
  for (k = 0; k < n; k += 2) {
    re += a[k] * b[k] + a[k + 1] * b[k + 1];
    ii += a[k] * b[k + 1] - a[k + 1] * b[k];
  }
  *im = ii;
  return re;
}

static int gpu_checks(void)
{
  const int site[4] = {4, 4, 4, 8};
  const size_t V = 4 * 4 * 4 * 8, ls = 16;
  const size_t nh = V / 2 * ls * 12 * 2; /* reals of a half field */
  const size_t np = V * 144 * 2;         /* reals of a propagator */
  qlat_dwf_fermion_action fa;
  qlat_dwf_inverter* inv = NULL;
  qlat_dwf_solve_stats st;
  double *gauge, *x, *y, *ax, *ay, *src, *sol;
  double q = 0, ip, ipi, ip2, ip2i, xax, xaxi;
  const double golden = 9.736817290497456;
  uint64_t seed = 2024;
  size_t i, k;
  long iter = 0;
  int rc, bad = 0;

  gauge = calloc(V * 4 * 9 * 2, sizeof(double));
  x = malloc(nh * sizeof(double));
  y = malloc(nh * sizeof(double));
  ax = malloc(nh * sizeof(double));
  ay = malloc(nh * sizeof(double));
  src = calloc(np, sizeof(double));
  sol = malloc(np * sizeof(double));
  if (!gauge || !x || !y || !ax || !ay || !src || !sol) {
    printf("FAIL host allocation\n");
    return 1;
  }
  for (i = 0; i < V * 4; ++i)
    for (k = 0; k < 3; ++k) gauge[(i * 9 + k * 4) * 2] = 1.0; /* unit links */
  memset(&fa, 0, sizeof fa);
  fa.mass = 0.05;
  fa.ls = (int)ls;
  fa.m5 = 1.0;
  fa.mobius_scale = 1.0;
  fa.is_multiplying_dminus = 1;
  fa.cg_diagonal_mee = 2;
  if ((rc = qlat_dwf_inverter_create(&inv, site, gauge, &fa)) != QLAT_DWF_OK) return fail("inverter_create", rc);

  for (k = 0; k < nh; ++k) x[k] = next_uniform(&seed);
  for (k = 0; k < nh; ++k) y[k] = next_uniform(&seed);
  if ((rc = qlat_dwf_multiply_hermop_sym2(inv, ax, x, 1)) != QLAT_DWF_OK) bad |= fail("hermop x", rc);
  if ((rc = qlat_dwf_multiply_hermop_sym2(inv, ay, y, 1)) != QLAT_DWF_OK) bad |= fail("hermop y", rc);
  ip = cdot_re(x, ay, nh, &ipi);   /* <x, A y> */
  ip2 = cdot_re(ax, y, nh, &ip2i); /* <A x, y> */
  xax = cdot_re(x, ax, nh, &xaxi);
  if (fabs(ip - ip2) > 1e-12 * fabs(ip) || fabs(ipi - ip2i) > 1e-12 * fabs(ip)) {
    printf("FAIL hermop not Hermitian: <x,Ay> = %.17g%+.17gi, <Ax,y> = %.17g%+.17gi\n", ip, ipi, ip2, ip2i);
    bad = 1;
  }
  if (!(xax > 0) || fabs(xaxi) > 1e-12 * xax) {
    printf("FAIL hermop not positive: <x,Ax> = %.17g%+.17gi\n", xax, xaxi);
    bad = 1;
  }

  for (k = 0; k < 12; ++k) src[(k * 12 + k) * 2] = 1.0; /* point source at the origin */
  if ((rc = qlat_dwf_invert_prop(inv, sol, src, &iter)) != QLAT_DWF_OK) bad |= fail("invert_prop", rc);
  for (k = 0; k < np; ++k) q += sol[k] * sol[k];
  if (2 * fabs(q - golden) / (fabs(q) + fabs(golden)) > 1e-10) {
    printf("FAIL propagator qnorm %.16g, reference %.16g\n", q, golden);
    bad = 1;
  }
  if ((rc = qlat_dwf_inverter_get_stats(inv, &st)) != QLAT_DWF_OK) bad |= fail("get_stats", rc);
  if (st.total_iter != iter || !(st.true_residual <= 1e-8) || !(st.true_residual_5d > 0) ||
      !(st.true_residual_5d < 1e-6)) {
    printf("FAIL stats: total_iter %ld (iter %ld), residual %g, 5D residual %g\n", st.total_iter, iter,
           st.true_residual, st.true_residual_5d);
    bad = 1;
  }
  if ((rc = qlat_dwf_inverter_destroy(inv)) != QLAT_DWF_OK) bad |= fail("inverter_destroy", rc);
  free(gauge);
  free(x);
  free(y);
  free(ax);
  free(ay);
  free(src);
  free(sol);
  if (!bad)
    printf("OK gpu: <x,Ax> %.6g, propagator qnorm %.16g, %ld iterations, 5D residual %.2e\n", xax, q, iter,
           st.true_residual_5d);
  return bad;
}

int main(int argc, char** argv)
{
  if (argc != 2 || (strcmp(argv[1], "cpu") != 0 && strcmp(argv[1], "gpu") != 0)) {
    printf("usage: %s cpu|gpu\n", argv[0]);
    return 2;
  }
  if (cpu_checks()) return 1;
  return strcmp(argv[1], "gpu") == 0 ? gpu_checks() : 0;
}
