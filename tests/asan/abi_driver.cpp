// Host-side AddressSanitizer driver of the C ABI (no GPU needed).
//
// Runs every entry point's argument-checking and error paths -- null and
// invalid arguments, the error-message buffer, plan creation failing on a
// machine without a HIP device and its clean-up -- in a process built with
// -fsanitize=address, so heap misuse in the host code (the plan and
// ensemble bookkeeping, the thread-local error string) is reported.  Exits 0
// when every call returned the status it should.
#include <stdio.h>
#include <string.h>

#include "qj_api.h"

static int failures = 0;

static void expect(const char* what, int rc, int want) {
  if (rc != want) {
    printf("FAIL %s: rc %d, want %d (%s)\n", what, rc, want, qj_last_error());
    ++failures;
  }
}

static void expect_neg(const char* what, int rc) {
  if (rc >= 0) {
    printf("FAIL %s: rc %d, want an error\n", what, rc);
    ++failures;
  } else if (strlen(qj_last_error()) == 0) {
    printf("FAIL %s: empty error message\n", what);
    ++failures;
  }
}

int main() {
  qj_plan* p = NULL;
  expect("version", qj_version() > 0, 1);
  expect_neg("plan_create null out", qj_plan_create(NULL, 16, 16, 16, 1, 1, 1, QJ_C64));
  expect("plan_create bad size", qj_plan_create(&p, 12, 16, 16, 1, 1, 1, QJ_C64), QJ_EUNSUPPORTED);
  expect("plan_create size 1024", qj_plan_create(&p, 1024, 16, 16, 1, 1, 1, QJ_C64),
         QJ_EUNSUPPORTED);
  expect("plan_create bad box", qj_plan_create(&p, 16, 16, 16, 1, -1, 1, QJ_C64), QJ_EARG);
  expect("plan_create bad dtype", qj_plan_create(&p, 16, 16, 16, 1, 1, 1, 7), QJ_EARG);
  expect("plan untouched on error", p == NULL, 1);
  // with no HIP device the plan's device allocations fail: the partially
  // built plan must be released and an error returned
  int rc = qj_plan_create(&p, 16, 16, 16, 1, 1, 1, QJ_C64);
  if (rc == QJ_OK) {
    qj_plan_destroy(p);  // a device was present after all
  } else {
    expect_neg("plan_create without a device", rc);
    expect("plan null after failure", p == NULL, 1);
  }
  expect("plan_destroy null", qj_plan_destroy(NULL), QJ_OK);
  expect("workspace_size null", (int)qj_workspace_size(NULL), 0);
  expect("star_scratch_size null", (int)qj_star_scratch_size(NULL, 100), 0);
  expect("adjoint_scratch_size null", (int)qj_adjoint_scratch_size(NULL), 0);
  int v = 0;
  expect_neg("set_option null", qj_plan_set_option(NULL, 0, 1));
  expect_neg("get_option null", qj_plan_get_option(NULL, 0, &v));
  expect_neg("timing_enable null", qj_timing_enable(NULL, 1));
  expect_neg("timing_read null", qj_timing_read(NULL, NULL, NULL, 0));
  const int ns = qj_stage_count();
  expect("stage_count", ns > 0, 1);
  for (int k = -2; k < ns + 2; ++k) {
    const char* n = qj_stage_name(k);
    expect("stage_name non-null", n != NULL, 1);
    if (k >= 0 && k < ns) expect("stage_name non-empty", strlen(n) > 0, 1);
  }
  qj_state st;
  memset(&st, 0, sizeof st);
  qj_params pr;
  memset(&pr, 0, sizeof pr);
  expect_neg("get_potential null plan", qj_get_potential(NULL, NULL, 1, 1, 0, 0, NULL, NULL, NULL));
  expect_neg("acceleration_grid null plan",
             qj_acceleration_grid(NULL, NULL, 1, 1, 0, NULL, NULL, NULL, NULL, NULL));
  expect_neg("kick null plan", qj_kick(NULL, NULL, NULL, 0.1, NULL));
  expect_neg("drift null plan", qj_drift(NULL, NULL, 0.1, 1.0, NULL));
  expect_neg("cic null plan", qj_cic_indices(NULL, NULL, 0, NULL, NULL, NULL, NULL, NULL));
  expect_neg("bin null plan", qj_bin_particles(NULL, NULL, NULL, 1, 0, NULL, 0, NULL));
  expect_neg("interp null plan", qj_interp(NULL, NULL, 0, NULL, NULL, NULL));
  expect_neg("hydro null plan",
             qj_solve_hydro(NULL, NULL, NULL, NULL, NULL, 1, 1, 0, NULL, NULL, NULL, NULL, NULL));
  expect_neg("kdk_steps null plan", qj_kdk_steps(NULL, &st, &pr, 1, NULL, NULL));
  expect_neg("kdk_steps_noalloc null plan", qj_kdk_steps_noalloc(NULL, &st, &pr, 1, NULL));
  expect_neg("kdk_finalize null plan", qj_kdk_finalize(NULL, &st, &pr, NULL));
  expect_neg("kdk_record null plan", qj_kdk_record(NULL, &st, &pr, 1, NULL, NULL, NULL, NULL));
  expect_neg("kdk_adjoint null plan",
             qj_kdk_adjoint(NULL, &st, &pr, 1, NULL, NULL, NULL, NULL, NULL));
  expect_neg("bh_accretion null plan",
             qj_bh_accretion(NULL, NULL, NULL, NULL, NULL, 1, 1, 1, 1, 1, 1, NULL));
  qj_ensemble* e = NULL;
  expect_neg("ensemble_create no members", qj_ensemble_create(&e, 0, NULL, NULL, NULL, NULL));
  expect_neg("ensemble_create null out", qj_ensemble_create(NULL, 1, NULL, NULL, NULL, NULL));
  expect("ensemble_destroy null", qj_ensemble_destroy(NULL), QJ_OK);
  expect_neg("ensemble_steps null", qj_ensemble_steps(NULL, 1, NULL));
  expect("ensemble_group_count null", qj_ensemble_group_count(NULL) <= 0, 1);
  // a long error message is replaced, not appended, call after call
  for (int k = 0; k < 1000; ++k) (void)qj_plan_create(&p, 3, 3, 3, 1, 1, 1, QJ_C64);
  expect("error message bounded", strlen(qj_last_error()) < 1024, 1);
  printf("%s: %d failures\n", failures ? "FAILED" : "ok", failures);
  return failures ? 1 : 0;
}
