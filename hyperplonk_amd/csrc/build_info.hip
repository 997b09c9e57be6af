// Build provenance: the SHA-256 of the sources this library was compiled from
// (every .hip and .hpp file of this folder in name order, then
// include/hp_mi355x.h, concatenated), computed by the Makefile at build time.
// A test recomputes it from the files in the tree, so a library that does not
// match its sources is caught wherever it is loaded.
#include <cstring>

#include "../../include/hp_mi355x.h"

#ifndef HP_SOURCE_DIGEST
#define HP_SOURCE_DIGEST "unknown"
#endif

extern "C" HpMi355xStatus hp_mi355x_v1_build_info(char* out, size_t capacity) {
  static const char kInfo[] = "sources-sha256=" HP_SOURCE_DIGEST " arch=gfx950";
  if (out == nullptr || capacity < sizeof kInfo) return {HP_MI355X_STATUS_INVALID_PARAMETERS, 0};
  std::memcpy(out, kInfo, sizeof kInfo);
  return {HP_MI355X_STATUS_SUCCESS, 0};
}
