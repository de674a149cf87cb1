// MFMA GEMM engine for gfx950 (CDNA4).
//
//   C[m, n] = alpha * sum_k A(m, k) * B(n, k) + beta * C[m, n]      (C is fp64)
//
// A(m, k) is A[m*lda + k] when a_kc (k contiguous) else A[k*lda + m];
// B(n, k) is B[n*ldb + k] when b_kc (k contiguous) else B[k*ldb + n].
// Batched with element strides sA, sB, sC.
//
// Precision modes:
//   GP_F64    fp64 operands, v_mfma_f64_16x16x4_f64, fp64 accumulation
//   GP_F32    fp32 operands, v_mfma_f32_16x16x4_f32, fp32 accumulation
//   GP_BF16X3 operands split a = hi + lo (bf16 each), three products
//             hi*hi + hi*lo + lo*hi per k on v_mfma_f32_16x16x32_bf16 with
//             fp32 accumulation (Henry, Tang, Heinecke, ARITH 2019).
//             k-contiguous operands only, lda/ldb multiples of 8, K multiple of 8.
#pragma once
#include <hip/hip_runtime.h>

namespace aces {

enum GemmPrec : int { GP_F64 = 0, GP_F32 = 1, GP_BF16X3 = 2 };

struct GemmDesc {
  int M = 0, N = 0, K = 0;
  const void* A = nullptr;
  const void* A_lo = nullptr;  // bf16x3 only
  long lda = 0;
  int a_kc = 1;
  const void* B = nullptr;
  const void* B_lo = nullptr;  // bf16x3 only
  long ldb = 0;
  int b_kc = 1;
  double* C = nullptr;
  long ldc = 0;
  double alpha = 1.0, beta = 0.0;
  int batch = 1;
  long sA = 0, sB = 0, sC = 0;
  // Optional split-K workspace (fp64 elements).  When the output has too few
  // tiles to fill the chip and K is long, K is split over blockIdx.z; partial
  // products go to the workspace and a deterministic reduction adds them in
  // split order.
  double* ws = nullptr;
  long ws_elems = 0;
  // Optional bf16 split output: when C_hi is set, the result goes to the hi/lo
  // rows C_hi/C_lo (row r at r * ldc_lp, columns N.. left as they are) and C
  // is only read (beta != 0) or used as scratch: its contents afterwards are
  // unspecified.  batch 1.
  unsigned short* C_hi = nullptr;
  unsigned short* C_lo = nullptr;
  long ldc_lp = 0;
  // bf16x3 block shape request: 0 = chosen from the problem size; 8 = the
  // 256 x 256 direct-to-LDS kernel (K % 32 == 0) whenever M, N >= 256 and the
  // launch (tiles x batch) has at least 256 blocks.
  int tile_hint = 0;
};

int gemm(const GemmDesc& d, GemmPrec prec, hipStream_t stream);
// fp64 products that run as operand streams (grouped too): 1 = one output
// side at most 64 long, the other at least 4096, K at most 64 / ceil(short
// side / 16); 2 = output sides at most 32 and 64, K at least 4096; 0 =
// neither; 3 = an empty problem (decides nothing).
int gemm_skinny_kind(long M, long N, long K);
// Grouped launch: up to 8 independent problems (same operand orders and
// precision, batch 1) in one kernel launch, blockIdx.y selecting the problem.
// Split-K partials of all groups share d[0].ws.
int gemm_grouped(const GemmDesc* d, int ngroups, GemmPrec prec, hipStream_t stream);

}  // namespace aces
