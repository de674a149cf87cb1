// MFMA GEMM kernels for gfx950: fp64, fp32 and bf16x3 (split) operands.
//
// Block tile 128 x 128 computed by 4 waves (2 x 2, each 64 x 64 as a 4 x 4
// grid of 16 x 16 MFMA tiles).  Operand tiles are staged through LDS in
// [k][m] order (fp64/fp32) or [m][k] order (bf16), double buffered with a
// register prefetch of the next k tile.  Block -> tile mapping is XCD aware:
// the 8 XCDs each receive a contiguous range of logical tiles, and tiles are
// walked in groups of 8 row tiles so that co-resident blocks share A and B
// tiles in the XCD's L2.
#include "common.h"
#include "gemm.h"

namespace aces {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef short s8_t __attribute__((ext_vector_type(8)));

struct GemmArgs {
  int M, N, K;
  const void* A;
  const void* Alo;
  long lda;
  const void* B;
  const void* Blo;
  long ldb;
  double* C;
  long ldc;
  double alpha, beta;
  long sA, sB, sC;
  int tiles_m, tiles_n;
};

static constexpr int kBM = 128;
static constexpr int kBN = 128;
static constexpr int kGroupM = 8;
static constexpr int kXcd = 8;

// Logical tile for this block: XCD-contiguous ranges, then grouped order.
__device__ inline void tile_coords(int tiles_m, int tiles_n, int& tm, int& tn) {
  const int nb = tiles_m * tiles_n;
  const int bid = blockIdx.x;
  const int x = bid % kXcd;
  const int local = bid / kXcd;
  const int per = nb / kXcd;
  const int rem = nb % kXcd;
  const int logical = x * per + (x < rem ? x : rem) + local;
  const int group = kGroupM * tiles_n;
  const int gid = logical / group;
  const int first_m = gid * kGroupM;
  int gsz = tiles_m - first_m;
  if (gsz > kGroupM) gsz = kGroupM;
  const int in = logical % group;
  tm = first_m + in % gsz;
  tn = in / gsz;
}

template <typename T>
struct Mfma16;
// Output element r of lane l: f64 16x16x4 holds D[l/16 + 4r][l%16];
// f32 16x16x4 and bf16 16x16x32 hold D[4(l/16) + r][l%16].
template <>
struct Mfma16<double> {
  typedef d4_t acc_t;
  static __device__ inline int row(int lk, int r) { return lk + 4 * r; }
  static __device__ inline acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma16<float> {
  typedef f4_t acc_t;
  static __device__ inline int row(int lk, int r) { return 4 * lk + r; }
  static __device__ inline acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// ---------------------------------------------------------------------------
// fp64 / fp32 kernel (T = double | float), BK = 16.
// ---------------------------------------------------------------------------
template <typename T, bool AKC, bool BKC>
__global__ __launch_bounds__(256) void gemm16_kernel(GemmArgs p) {
  constexpr int BK = 16;
  constexpr int PAD = 16;
  constexpr int LD = kBM + PAD;
  __shared__ T As[2][BK][LD];
  __shared__ T Bs[2][BK][LD];
  typedef typename Mfma16<T>::acc_t acc_t;

  int tm, tn;
  tile_coords(p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * kBM;
  const int n0 = tn * kBN;
  const int bz = blockIdx.y;
  const T* __restrict__ A = reinterpret_cast<const T*>(p.A) + (long)bz * p.sA;
  const T* __restrict__ B = reinterpret_cast<const T*>(p.B) + (long)bz * p.sB;
  double* __restrict__ C = p.C + (long)bz * p.sC;
  const int M = p.M, N = p.N, K = p.K;

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave >> 1;
  const int wn = wave & 1;

  // Loader mapping.  k-contiguous: row = t & 127, 8 consecutive k from
  // (t >> 7) * 8.  m-contiguous: k row = t >> 4, 8 consecutive m from (t & 15) * 8.
  T ra[8], rb[8];
  auto load_a = [&](int k0) {
    if (AKC) {
      const int r = t & 127, kb = (t >> 7) * 8;
      const int m = m0 + r;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + kb + j;
        ra[j] = (m < M && k < K) ? A[(long)m * p.lda + k] : T(0);
      }
    } else {
      const int kr = t >> 4, mb = (t & 15) * 8;
      const int k = k0 + kr;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int m = m0 + mb + j;
        ra[j] = (m < M && k < K) ? A[(long)k * p.lda + m] : T(0);
      }
    }
  };
  auto load_b = [&](int k0) {
    if (BKC) {
      const int r = t & 127, kb = (t >> 7) * 8;
      const int n = n0 + r;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + kb + j;
        rb[j] = (n < N && k < K) ? B[(long)n * p.ldb + k] : T(0);
      }
    } else {
      const int kr = t >> 4, nb = (t & 15) * 8;
      const int k = k0 + kr;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int n = n0 + nb + j;
        rb[j] = (n < N && k < K) ? B[(long)k * p.ldb + n] : T(0);
      }
    }
  };
  auto store_tiles = [&](int buf) {
    if (AKC) {
      const int r = t & 127, kb = (t >> 7) * 8;
#pragma unroll
      for (int j = 0; j < 8; ++j) As[buf][kb + j][r] = ra[j];
    } else {
      const int kr = t >> 4, mb = (t & 15) * 8;
#pragma unroll
      for (int j = 0; j < 8; ++j) As[buf][kr][mb + j] = ra[j];
    }
    if (BKC) {
      const int r = t & 127, kb = (t >> 7) * 8;
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[buf][kb + j][r] = rb[j];
    } else {
      const int kr = t >> 4, nb = (t & 15) * 8;
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[buf][kr][nb + j] = rb[j];
    }
  };

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int nk = (K + BK - 1) / BK;
  load_a(0);
  load_b(0);
  store_tiles(0);
  __syncthreads();
  const int lr = lane & 15;
  const int lk = lane >> 4;
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) {
      load_a((kt + 1) * BK);
      load_b((kt + 1) * BK);
    }
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[buf][kk * 4 + lk][wm * 64 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[buf][kk * 4 + lk][wn * 64 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Mfma16<T>::op(a[i], b[j], acc[i][j]);
    }
    if (kt + 1 < nk) {
      store_tiles(buf ^ 1);
    }
    __syncthreads();
  }

  // Epilogue: acc[i][j][r] -> row wm*64 + i*16 + row(lk, r), col wn*64 + j*16 + lr.
  const double alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn * 64 + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + Mfma16<T>::row(lk, r);
        if (row < M && col < N) {
          double* c = C + (long)row * p.ldc + col;
          const double v = alpha * (double)acc[i][j][r];
          *c = (beta == 0.0) ? v : v + beta * (*c);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// bf16x3 kernel: A = Ahi + Alo, B = Bhi + Blo (bf16, k contiguous), BK = 32.
// acc += Ahi*Bhi + Ahi*Blo + Alo*Bhi with v_mfma_f32_16x16x32_bf16.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gemm_bf16x3_kernel(GemmArgs p) {
  constexpr int BK = 32;
  constexpr int LDK = BK + 8;  // bf16 elements per LDS row (80 bytes)
  __shared__ __attribute__((aligned(16))) unsigned short Ah[2][kBM * LDK];
  __shared__ __attribute__((aligned(16))) unsigned short Al[2][kBM * LDK];
  __shared__ __attribute__((aligned(16))) unsigned short Bh[2][kBN * LDK];
  __shared__ __attribute__((aligned(16))) unsigned short Bl[2][kBN * LDK];

  int tm, tn;
  tile_coords(p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * kBM;
  const int n0 = tn * kBN;
  const int bz = blockIdx.y;
  const unsigned short* __restrict__ Ahg = reinterpret_cast<const unsigned short*>(p.A) + (long)bz * p.sA;
  const unsigned short* __restrict__ Alg = reinterpret_cast<const unsigned short*>(p.Alo) + (long)bz * p.sA;
  const unsigned short* __restrict__ Bhg = reinterpret_cast<const unsigned short*>(p.B) + (long)bz * p.sB;
  const unsigned short* __restrict__ Blg = reinterpret_cast<const unsigned short*>(p.Blo) + (long)bz * p.sB;
  double* __restrict__ C = p.C + (long)bz * p.sC;
  const int M = p.M, N = p.N, K = p.K;

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave >> 1;
  const int wn = wave & 1;

  // Loader: row = t >> 1 (0..127), 16 consecutive k from (t & 1) * 16, as two
  // 16-byte chunks.
  const int lrow = t >> 1;
  const int lkb = (t & 1) * 16;
  uint4 rah[2], ral[2], rbh[2], rbl[2];
  auto load = [&](int k0) {
    const int m = m0 + lrow;
    const int n = n0 + lrow;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int k = k0 + lkb + c * 8;
      const bool ka = (k < K);
      if (m < M && ka) {
        rah[c] = *reinterpret_cast<const uint4*>(Ahg + (long)m * p.lda + k);
        ral[c] = *reinterpret_cast<const uint4*>(Alg + (long)m * p.lda + k);
      } else {
        rah[c] = make_uint4(0, 0, 0, 0);
        ral[c] = make_uint4(0, 0, 0, 0);
      }
      if (n < N && ka) {
        rbh[c] = *reinterpret_cast<const uint4*>(Bhg + (long)n * p.ldb + k);
        rbl[c] = *reinterpret_cast<const uint4*>(Blg + (long)n * p.ldb + k);
      } else {
        rbh[c] = make_uint4(0, 0, 0, 0);
        rbl[c] = make_uint4(0, 0, 0, 0);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int off = lrow * LDK + lkb + c * 8;
      *reinterpret_cast<uint4*>(&Ah[buf][off]) = rah[c];
      *reinterpret_cast<uint4*>(&Al[buf][off]) = ral[c];
      *reinterpret_cast<uint4*>(&Bh[buf][off]) = rbh[c];
      *reinterpret_cast<uint4*>(&Bl[buf][off]) = rbl[c];
    }
  };

  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0, 0, 0, 0};

  const int nk = (K + BK - 1) / BK;
  load(0);
  store(0);
  __syncthreads();
  const int lr = lane & 15;
  const int lk = lane >> 4;
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
    s8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int off = (wm * 64 + i * 16 + lr) * LDK + lk * 8;
      ah[i] = *reinterpret_cast<const s8_t*>(&Ah[buf][off]);
      al[i] = *reinterpret_cast<const s8_t*>(&Al[buf][off]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int off = (wn * 64 + j * 16 + lr) * LDK + lk * 8;
      bh[j] = *reinterpret_cast<const s8_t*>(&Bh[buf][off]);
      bl[j] = *reinterpret_cast<const s8_t*>(&Bl[buf][off]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf8_t, al[i]), __builtin_bit_cast(bf8_t, bh[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf8_t, ah[i]), __builtin_bit_cast(bf8_t, bl[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf8_t, ah[i]), __builtin_bit_cast(bf8_t, bh[j]), acc[i][j], 0, 0, 0);
      }
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }

  const double alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn * 64 + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + 4 * lk + r;
        if (row < M && col < N) {
          double* c = C + (long)row * p.ldc + col;
          const double v = alpha * (double)acc[i][j][r];
          *c = (beta == 0.0) ? v : v + beta * (*c);
        }
      }
    }
  }
}

template <typename T>
static void launch16(const GemmArgs& a, bool akc, bool bkc, dim3 grid, hipStream_t s) {
  if (akc && bkc)
    hipLaunchKernelGGL((gemm16_kernel<T, true, true>), grid, dim3(256), 0, s, a);
  else if (akc && !bkc)
    hipLaunchKernelGGL((gemm16_kernel<T, true, false>), grid, dim3(256), 0, s, a);
  else if (!akc && bkc)
    hipLaunchKernelGGL((gemm16_kernel<T, false, true>), grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((gemm16_kernel<T, false, false>), grid, dim3(256), 0, s, a);
}

int gemm(const GemmDesc& d, GemmPrec prec, hipStream_t stream) {
  ACES_REQUIRE(d.M >= 0 && d.N >= 0 && d.K >= 0 && d.batch >= 1, "gemm: bad dimensions");
  if (d.M == 0 || d.N == 0) return ACES_OK;
  ACES_REQUIRE(d.C != nullptr && d.ldc >= d.N, "gemm: bad C");
  if (d.K > 0) {
    ACES_REQUIRE(d.A != nullptr && d.B != nullptr, "gemm: null operand");
    ACES_REQUIRE(d.lda >= (d.a_kc ? d.K : d.M), "gemm: lda too small");
    ACES_REQUIRE(d.ldb >= (d.b_kc ? d.K : d.N), "gemm: ldb too small");
  }
  GemmArgs a;
  a.M = d.M; a.N = d.N; a.K = d.K;
  a.A = d.A; a.Alo = d.A_lo; a.lda = d.lda;
  a.B = d.B; a.Blo = d.B_lo; a.ldb = d.ldb;
  a.C = d.C; a.ldc = d.ldc;
  a.alpha = d.alpha; a.beta = d.beta;
  a.sA = d.sA; a.sB = d.sB; a.sC = d.sC;
  a.tiles_m = (d.M + kBM - 1) / kBM;
  a.tiles_n = (d.N + kBN - 1) / kBN;
  const long nb = (long)a.tiles_m * a.tiles_n;
  ACES_REQUIRE(nb < (1L << 31) && d.batch < 65536, "gemm: grid too large");
  dim3 grid((unsigned)nb, (unsigned)d.batch);
  if (d.K == 0) {
    // C = beta * C handled by a K = 0 launch of the fp64 kernel (acc stays 0).
    a.K = 0;
    launch16<double>(a, true, true, grid, stream);
    ACES_HIP(hipGetLastError());
    return ACES_OK;
  }
  switch (prec) {
    case GP_F64:
      launch16<double>(a, d.a_kc != 0, d.b_kc != 0, grid, stream);
      break;
    case GP_F32:
      launch16<float>(a, d.a_kc != 0, d.b_kc != 0, grid, stream);
      break;
    case GP_BF16X3:
      ACES_REQUIRE(d.a_kc && d.b_kc, "gemm bf16x3: operands must be k-contiguous");
      ACES_REQUIRE(d.A_lo != nullptr && d.B_lo != nullptr, "gemm bf16x3: missing lo parts");
      ACES_REQUIRE(d.lda % 8 == 0 && d.ldb % 8 == 0 && d.K % 8 == 0 && d.sA % 8 == 0 && d.sB % 8 == 0,
                   "gemm bf16x3: lda, ldb, K and strides must be multiples of 8");
      ACES_REQUIRE(((uintptr_t)d.A % 16) == 0 && ((uintptr_t)d.B % 16) == 0 &&
                       ((uintptr_t)d.A_lo % 16) == 0 && ((uintptr_t)d.B_lo % 16) == 0,
                   "gemm bf16x3: operands must be 16-byte aligned");
      hipLaunchKernelGGL(gemm_bf16x3_kernel, grid, dim3(256), 0, stream, a);
      break;
    default:
      ACES_REQUIRE(false, "gemm: unknown precision");
  }
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

}  // namespace aces
