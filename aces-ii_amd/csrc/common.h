// Shared helpers for the gfx950 CCSD amplitude library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/aces2_amd.h"  // status codes ACES_OK, ACES_ERR_*

namespace aces {

void set_error(const std::string& msg);
const char* last_error();

}  // namespace aces

#define ACES_HIP(call)                                                            \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess) {                                                       \
      char _b[512];                                                               \
      snprintf(_b, sizeof(_b), "%s:%d %s -> %s", __FILE__, __LINE__, #call,       \
               hipGetErrorString(_e));                                            \
      aces::set_error(_b);                                                        \
      return ACES_ERR_HIP;                                                  \
    }                                                                             \
  } while (0)

#define ACES_CHECK(expr)                 \
  do {                                   \
    int _s = (expr);                     \
    if (_s != ACES_OK) return _s;  \
  } while (0)

#define ACES_REQUIRE(cond, msg)            \
  do {                                     \
    if (!(cond)) {                         \
      aces::set_error(msg);                \
      return ACES_ERR_ARG;           \
    }                                      \
  } while (0)
