/* Drives the C ABI's argument validation and error paths on the host, for the
 * AddressSanitizer / UBSan build of the library (liblri_hip_san.so).  Needs no
 * GPU: every call either rejects its arguments before touching a device or
 * fails cleanly when no device is present.  Exits non-zero on an unexpected
 * status; the sanitizers abort on a memory error or undefined behaviour. */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "lri.h"

static int failures = 0;

static void expect(int cond, const char* what) {
  if (!cond) {
    fprintf(stderr, "unexpected: %s\n", what);
    ++failures;
  }
}

int main(void) {
  expect(lri_version() >= 100, "lri_version");
  for (int s = -8; s <= 2; ++s) expect(lri_status_string(s) != NULL, "lri_status_string");
  expect(lri_last_error(NULL) != NULL, "lri_last_error(NULL)");

  /* lri_create: argument validation before any device call */
  lri_ctx* ctx = NULL;
  expect(lri_create(NULL, 0, 8, 16, 4) == LRI_EINVAL, "create: null out");
  expect(lri_create(&ctx, 0, 0, 16, 4) == LRI_EINVAL, "create: n = 0");
  expect(lri_create(&ctx, 0, 16, 16, 65) == LRI_EINVAL, "create: rcap > 64");
  expect(lri_create(&ctx, 0, 8, 16, 9) == LRI_EINVAL, "create: rcap > n");
  expect(ctx == NULL, "create: no context on failure");
  expect(strlen(lri_last_error(NULL)) > 0, "create: message");

  /* a valid request: a context when a device is present, a clean failure otherwise */
  int st = lri_create(&ctx, 0, 64, 48, 4);
  if (st == LRI_OK) {
    expect(ctx != NULL, "create: context");
    expect(lri_set_precision(ctx, 99) == LRI_EINVAL, "set_precision: bad mode");
    expect(lri_step_ksl(ctx, 7, 0, 1) != LRI_OK, "step_ksl: bad order");
    expect(lri_snapshots_reserve(ctx, -1, LRI_F64) == LRI_EINVAL, "reserve: negative count");
    expect(lri_destroy(ctx) == LRI_OK, "destroy");
  } else {
    expect(ctx == NULL, "create: no context after a device failure");
  }

  /* null contexts are rejected by every entry point that takes one */
  double x[16] = {0};
  int64_t r = 0;
  expect(lri_destroy(NULL) == LRI_OK, "destroy(NULL)");
  expect(lri_synchronize(NULL) == LRI_EINVAL, "synchronize(NULL)");
  expect(lri_set_precision(NULL, LRI_PREC_F64) == LRI_EINVAL, "set_precision(NULL)");
  expect(lri_set_state(NULL, 1, x, 4, x, 1, x, 4) == LRI_EINVAL, "set_state(NULL)");
  expect(lri_get_state(NULL, &r, x, 4, x, 1, x, 4) == LRI_EINVAL, "get_state(NULL)");
  expect(lri_step_ksl(NULL, LRI_PRIMAL, 0, 1) == LRI_EINVAL, "step_ksl(NULL)");
  expect(lri_step_ksl_strang(NULL, 0, 1, 2) == LRI_EINVAL, "step_ksl_strang(NULL)");
  expect(lri_step_bug(NULL, 0, 1) == LRI_EINVAL, "step_bug(NULL)");
  expect(lri_step_bug_ra(NULL, 0, 1, 1e-8, 8, &r) == LRI_EINVAL, "step_bug_ra(NULL)");
  expect(lri_step_bug_adjoint(NULL, 0, 1, x, 4, x, 1, x, 4) == LRI_EINVAL, "step_bug_adjoint(NULL)");
  expect(lri_step_ksl_adjoint(NULL, 0, 1, x, 4, x, 1, x, 4) == LRI_EINVAL, "step_ksl_adjoint(NULL)");
  expect(lri_step_ksl_dual_adjoint(NULL, 0, 1, x, 4, x, 1, x, 4) == LRI_EINVAL, "step_ksl_dual_adjoint(NULL)");
  expect(lri_trajectory_get_range(NULL, 0, 1, &r, x) == LRI_EINVAL, "trajectory_get_range(NULL)");
  expect(lri_step_de_bug_ra_adjoint(NULL, 1e-3, 1, 1e-8, 4, 1, x, 4, x, 1, x, 4, NULL) == LRI_EINVAL,
         "step_de_bug_ra_adjoint(NULL)");
  expect(lri_step_ksl_strang_adjoint(NULL, 0, 1, 2, x, 4, x, 1, x, 4) == LRI_EINVAL,
         "step_ksl_strang_adjoint(NULL)");
  expect(lri_step_greedy(NULL, 1) == LRI_EINVAL, "step_greedy(NULL)");
  expect(lri_step_de_bug(NULL, 1e-3, 4) == LRI_EINVAL, "step_de_bug(NULL)");
  expect(lri_step_de_bug_ra(NULL, 1e-3, 4, 1e-8, 8, &r) == LRI_EINVAL, "step_de_bug_ra(NULL)");
  expect(lri_trajectory_reserve(NULL, 4) == LRI_EINVAL, "trajectory_reserve(NULL)");
  expect(lri_kernel_timing(NULL, 1) == LRI_EINVAL, "kernel_timing(NULL)");
  expect(lri_autotune_passes(NULL, 0, 1, 3, x) == LRI_EINVAL, "autotune_passes(NULL)");

  /* batches */
  lri_batch* b = NULL;
  int64_t nn[2] = {32, 40}, mm[2] = {32, 40}, rc[2] = {4, 70};
  expect(lri_batch_create(NULL, 0, 2, nn, mm, rc) == LRI_EINVAL, "batch_create: null out");
  expect(lri_batch_create(&b, 0, 0, nn, mm, rc) == LRI_EINVAL, "batch_create: count 0");
  expect(lri_batch_create(&b, 0, 2, nn, mm, rc) == LRI_EINVAL, "batch_create: rcap 70");
  expect(b == NULL, "batch_create: no batch on failure");
  expect(lri_batch_destroy(NULL) == LRI_OK, "batch_destroy(NULL)");
  expect(lri_batch_synchronize(NULL) == LRI_EINVAL, "batch_synchronize(NULL)");
  expect(lri_batch_last_error(NULL) != NULL, "batch_last_error(NULL)");

  /* the allocation-free product rejects bad shapes before any device call */
  double out[4];
  expect(lri_product_noalloc(0, LRI_F64, LRI_PREC_F64, 0, 16, x, x, 16, x, 16, 4, out, 16, NULL) == LRI_EINVAL,
         "product_noalloc: n = 0");
  expect(lri_product_noalloc(2, LRI_F64, LRI_PREC_F64, 16, 16, x, x, 16, x, 16, 4, out, 16, NULL) == LRI_EINVAL,
         "product_noalloc: which = 2");
  expect(lri_product_noalloc(0, LRI_F64, LRI_PREC_F64, 65536, 65536, x, x, 65536, x, 65536, 64, out, 65536, NULL) ==
             LRI_EINVAL,
         "product_noalloc: beyond the scratch");
  expect(strlen(lri_noalloc_last_error()) > 0, "noalloc_last_error");

  if (failures) {
    fprintf(stderr, "%d unexpected results\n", failures);
    return 1;
  }
  printf("abi host checks passed\n");
  return 0;
}
