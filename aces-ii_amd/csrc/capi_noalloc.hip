// Grouped MFMA GEMM launch that allocates nothing: the split-K partials of
// aces_gemm_grouped live in a file-scope device array instead of a
// caller-provided workspace.
//
// For callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops): the caller passes operands and outputs only.  Every GPU
// the library runs on holds its own copy of `scratch` (32 MiB) until the
// process ends.  All calls on one GPU share that copy, so calls that overlap in
// time on one GPU (concurrent streams or threads) race on it; calls on one
// stream are ordered and safe.  The partials are written in full by the
// product kernels before the reduction reads them, so nothing needs clearing
// between calls.
#include <mutex>
#include <string>
#include "common.h"
#include "gemm.h"

using namespace aces;

static constexpr long kScratchElems = 1L << 22;  // 4 Mi doubles = 32 MiB

static __device__ __align__(256) double scratch[kScratchElems];

namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` per device.  In host code the array's name is a
// host-side stand-in, never a device address: only these pointers are used.
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_for_current_device(double** out) {
  int dev = -1;
  ACES_HIP(hipGetDevice(&dev));
  ACES_REQUIRE(dev >= 0 && dev < kMaxDevices, "aces_gemm_grouped_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    ACES_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ACES_REQUIRE(p != nullptr, "aces_gemm_grouped_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return ACES_OK;
}

}  // namespace

extern "C" {

int aces_gemm_grouped_noalloc(int prec, int ngroups, const int* M, const int* N, const int* K, const void* const* A,
                              const void* const* A_lo, const long* lda, int a_kc, const void* const* B,
                              const void* const* B_lo, const long* ldb, int b_kc, double* const* C,
                              const long* ldc, double alpha, double beta, long ws_elems, void* stream) {
  if (ngroups < 1 || ngroups > 8) {
    set_error("aces_gemm_grouped_noalloc: 1..8 groups");
    return ACES_ERR_ARG;
  }
  if (ws_elems < 0 || ws_elems > kScratchElems) {
    set_error("aces_gemm_grouped_noalloc: ws_elems = " + std::to_string(ws_elems) +
              " exceeds the library's static scratch of " + std::to_string(kScratchElems) + " doubles");
    return ACES_ERR_RANGE;
  }
  double* ws = nullptr;
  if (ws_elems > 0) ACES_CHECK(scratch_for_current_device(&ws));
  GemmDesc d[8];
  for (int g = 0; g < ngroups; ++g) {
    d[g].M = M[g]; d[g].N = N[g]; d[g].K = K[g];
    d[g].A = A[g]; d[g].A_lo = A_lo ? A_lo[g] : nullptr; d[g].lda = lda[g]; d[g].a_kc = a_kc;
    d[g].B = B[g]; d[g].B_lo = B_lo ? B_lo[g] : nullptr; d[g].ldb = ldb[g]; d[g].b_kc = b_kc;
    d[g].C = C[g]; d[g].ldc = ldc[g]; d[g].alpha = alpha; d[g].beta = beta;
    d[g].ws = ws; d[g].ws_elems = ws_elems;
  }
  return gemm_grouped(d, ngroups, (GemmPrec)prec, (hipStream_t)stream);
}

}  // extern "C"
