/* Links the static archive libhp_mi355x.a into a plain C program (the way the
 * reference's Rust glue links its bridge, backend/src/ascend910b3/ffi.rs:34
 * kind = "static") and calls entry points whose argument checks come before
 * any device use, so it runs without a GPU. */
#include <stdio.h>

#include "../../include/hp_mi355x.h"

int main(void) {
  HpMi355xStatus s = hp_mi355x_v1_context_create(0, NULL);
  if (s.kind != HP_MI355X_STATUS_INVALID_PARAMETERS || s.native_code != 0) return 1;
  s = hp_mi355x_v1_context_destroy(NULL);
  if (s.kind != HP_MI355X_STATUS_INVALID_PARAMETERS) return 2;
  HpMi355xDeviceAddress out = (HpMi355xDeviceAddress)1;
  s = hp_mi355x_v1_mle_upload(NULL, 1, NULL, 2, &out);
  if (s.kind != HP_MI355X_STATUS_INVALID_PARAMETERS || out != NULL) return 3;
  printf("STATIC_LINK PASS\n");
  return 0;
}
