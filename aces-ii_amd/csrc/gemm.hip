// MFMA GEMM kernels for gfx950: fp64, fp32 and bf16x3 (split) operands.
//
// Block tile 128 x 128 computed by 4 waves (2 x 2, each 64 x 64 as a 4 x 4
// grid of 16 x 16 MFMA tiles).  Operand tiles are staged through LDS in
// [k][m] order (fp64/fp32) or [m][k] order (bf16), double buffered with a
// register prefetch of the next k tile.  Block -> tile mapping is XCD aware:
// the 8 XCDs each receive a contiguous range of logical tiles, and tiles are
// walked in groups of 8 row tiles so that co-resident blocks share A and B
// tiles in the XCD's L2.
#include <algorithm>
#include <cstring>
#include "common.h"
#include "gemm.h"
#include "bf16_util.h"
#include "tensor_ops.h"

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
  // split-K: blockIdx.z selects k range [z*kchunk, min(K, (z+1)*kchunk)); the
  // operand pointers advance by z*kchunk*zkA / zkB and C by z*zC.
  long kchunk;
  long zkA, zkB, zC;
  int vecA, vecB;  // 16-byte aligned operand rows (vector load path allowed)
  int splits;      // number of k splits of this problem (grid.z may be larger)
  unsigned short* Chi;  // bf16 split output instead of C (direct-to-LDS bf16x3 kernel, no split-K)
  unsigned short* Clo;
  long ldcl;
};

// One launch computes up to kMaxGroups independent problems (the irrep blocks
// of one term); blockIdx.y selects the group when n > 1, the batch entry when
// n == 1.  Passed by value in the kernel arguments.
static constexpr int kMaxGroups = 8;
struct GemmGroups {
  int n;
  GemmArgs g[kMaxGroups];
};

// Uniform selection with constant indices (keeps the argument block in the
// kernarg segment).
__device__ inline GemmArgs select_group(const GemmGroups& gg, int gi) {
  GemmArgs p = gg.g[0];
#pragma unroll
  for (int k = 1; k < kMaxGroups; ++k)
    if (k == gi) p = gg.g[k];
  return p;
}

static constexpr int kGroupM = 8;
static constexpr int kNarrowN = 24;
static constexpr int kXcd = 8;

// Logical tile for this block: XCD-contiguous ranges, then row-major order
// (narrow outputs) or groups of kGroupM row tiles walked column by column.
__device__ inline void tile_coords(int tiles_m, int tiles_n, int& tm, int& tn) {
  const int nb = tiles_m * tiles_n;
  const int bid = blockIdx.x;
  const int x = bid % kXcd;
  const int local = bid / kXcd;
  const int per = nb / kXcd;
  const int rem = nb % kXcd;
  const int logical = x * per + (x < rem ? x : rem) + local;
  if (tiles_n <= kNarrowN) {
    // narrow output: all column tiles of a row tile run together, so each A
    // row panel is pulled through the XCD's L2 once
    tm = logical / tiles_n;
    tn = logical % tiles_n;
    return;
  }
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
// fp64 / fp32 kernel.  Block = WM x WN waves, each wave a 64 x 64 tile of
// 4 x 4 MFMA 16x16 blocks; block tile (64 WM) x (64 WN) x BK.
// LDS tiles are [m][k] with row pitch LDK = BK + 2 (fp64) or BK + 4 (fp32):
// LDK = 2*odd (fp64) / 4*odd (fp32) maps the 16 rows x (64/16) k values of one
// MFMA operand fetch onto 64 distinct banks.  Operands move in 16-byte chunks
// (2 fp64 / 4 fp32); interior tiles take a branch-free vector path.
// ---------------------------------------------------------------------------
template <typename T>
struct Chunk;
template <>
struct Chunk<double> {
  static constexpr int E = 2;
  static constexpr int PADK = 2;
};
template <>
struct Chunk<float> {
  static constexpr int E = 4;
  static constexpr int PADK = 4;
};

template <typename T, bool KC, int ROWS, int BK, int NT>
struct TileLoader {
  // ROWS x BK tile of a k-contiguous (KC) or row-contiguous (!KC) operand.
  static constexpr int E = Chunk<T>::E;
  static constexpr int LDK = BK + Chunk<T>::PADK;
  static constexpr int NCH = ROWS * BK / E;  // chunks per tile
  static constexpr int PER = NCH / NT;       // chunks per thread
  static_assert(NCH % NT == 0, "tile/threads mismatch");
  uint4 r[PER];

  __device__ inline void load(const T* __restrict__ G, long ld, int row0, int nrows, int k0, int K, bool vec,
                              int t) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int q = t + i * NT;
      int row, kk;
      if (KC) {
        row = q / (BK / E);
        kk = (q % (BK / E)) * E;
      } else {
        kk = q / (ROWS / E);
        row = (q % (ROWS / E)) * E;
      }
      const int gr = row0 + row, gk = k0 + kk;
      if (vec) {
        const T* src = KC ? G + (long)gr * ld + gk : G + (long)gk * ld + gr;
        r[i] = *reinterpret_cast<const uint4*>(src);
      } else {
        T v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int rr = KC ? gr : gr + e;
          const int kx = KC ? gk + e : gk;
          v[e] = (rr < nrows && kx < K) ? (KC ? G[(long)rr * ld + kx] : G[(long)kx * ld + rr]) : T(0);
        }
        r[i] = *reinterpret_cast<const uint4*>(v);
      }
    }
  }
  __device__ inline void store(T* __restrict__ S, int t) const {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int q = t + i * NT;
      if (KC) {
        const int row = q / (BK / E);
        const int kk = (q % (BK / E)) * E;
        *reinterpret_cast<uint4*>(S + row * LDK + kk) = r[i];
      } else {
        const int kk = q / (ROWS / E);
        const int row = (q % (ROWS / E)) * E;
        const T* v = reinterpret_cast<const T*>(&r[i]);
#pragma unroll
        for (int e = 0; e < E; ++e) S[(row + e) * LDK + kk] = v[e];
      }
    }
  }
};

template <typename T, bool AKC, bool BKC, int BK, int WM, int WN, int FM, int FN>
__global__ __launch_bounds__(64 * WM * WN) void gemm16_kernel(GemmGroups gg) {
  constexpr int NT = 64 * WM * WN;
  constexpr int BM = 16 * FM * WM;
  constexpr int BN = 16 * FN * WN;
  typedef TileLoader<T, AKC, BM, BK, NT> LA;
  typedef TileLoader<T, BKC, BN, BK, NT> LB;
  constexpr int LDK = LA::LDK;
  __shared__ __attribute__((aligned(16))) T As[2][BM * LDK];
  __shared__ __attribute__((aligned(16))) T Bs[2][BN * LDK];
  typedef typename Mfma16<T>::acc_t acc_t;

  const int gi = (gg.n > 1) ? (int)blockIdx.y : 0;
  const GemmArgs p = select_group(gg, gi);
  if ((int)blockIdx.x >= p.tiles_m * p.tiles_n || (int)blockIdx.z >= p.splits) return;
  int tm, tn;
  tile_coords(p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const int bz = (gg.n > 1) ? 0 : (int)blockIdx.y;
  const long kz = (long)blockIdx.z * p.kchunk;
  const T* __restrict__ A = reinterpret_cast<const T*>(p.A) + (long)bz * p.sA + kz * p.zkA;
  const T* __restrict__ B = reinterpret_cast<const T*>(p.B) + (long)bz * p.sB + kz * p.zkB;
  double* __restrict__ C = p.C + (long)bz * p.sC + (long)blockIdx.z * p.zC;
  const int M = p.M, N = p.N;
  const int K = (int)((p.K - kz) < p.kchunk ? (p.K - kz) : p.kchunk);

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave / WN;
  const int wn = wave % WN;

  // vector path: whole tile inside and 16-byte aligned rows
  const bool alA = p.vecA != 0;
  const bool alB = p.vecB != 0;
  const bool inM = (m0 + BM <= M), inN = (n0 + BN <= N);

  LA la;
  LB lb;
  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int nk = (K + BK - 1) / BK;
  {
    const bool kin = (BK <= K);
    la.load(A, p.lda, m0, M, 0, K, alA && inM && kin, t);
    lb.load(B, p.ldb, n0, N, 0, K, alB && inN && kin, t);
    la.store(As[0], t);
    lb.store(Bs[0], t);
  }
  __syncthreads();
  const int lr = lane & 15;
  const int lk = lane >> 4;
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) {
      const int k0 = (kt + 1) * BK;
      const bool kin = (k0 + BK <= K);
      la.load(A, p.lda, m0, M, k0, K, alA && inM && kin, t);
      lb.load(B, p.ldb, n0, N, k0, K, alB && inN && kin, t);
    }
    const T* __restrict__ as = As[buf];
    const T* __restrict__ bs = Bs[buf];
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      T a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = as[(wm * 16 * FM + i * 16 + lr) * LDK + kk * 4 + lk];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = bs[(wn * 16 * FN + j * 16 + lr) * LDK + kk * 4 + lk];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = Mfma16<T>::op(a[i], b[j], acc[i][j]);
    }
    if (kt + 1 < nk) {
      la.store(As[buf ^ 1], t);
      lb.store(Bs[buf ^ 1], t);
    }
    __syncthreads();
  }

  // Epilogue: acc[i][j][r] -> row wm*16FM + i*16 + row(lk, r), col wn*16FN + j*16 + lr.
  const double alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn * 16 * FN + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 16 * FM + i * 16 + Mfma16<T>::row(lk, r);
        if (row < M && col < N) {
          double* c = C + (long)row * p.ldc + col;
          const double v = alpha * (double)acc[i][j][r];
          *c = (beta == 0.0) ? v : v + beta * (*c);
        }
      }
    }
  }
}

// Position of 16-byte chunk 0 of a 64-byte LDS row r (chunk q sits at q ^ sw):
// the 16 rows of one MFMA fragment read (one chunk each) then cover all banks
// both when 8 lanes (32 banks) and when 16 lanes (64 banks) share a cycle:
// rows with equal 16r mod 128 bytes (r ^ 2, 4, 6) and with equal 16r mod 256
// bytes (r ^ 4, 8, 12) get four different positions.
__device__ inline int lds_swizzle(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

// ---------------------------------------------------------------------------
// bf16x3 kernel: A = Ahi + Alo, B = Bhi + Blo (bf16, k contiguous), BK = 32.
// acc += Alo*Bhi + Ahi*Blo + Ahi*Bhi with v_mfma_f32_16x16x32_bf16.
// Block tile (64 WM) x 128 on 2 WM waves (each 64 x 64); WM = 4 halves the
// operand bytes per flop that the block pulls through L2 (the bound of this
// kernel) and is used when M is large.
// ---------------------------------------------------------------------------
template <int WM>
__global__ __launch_bounds__(128 * WM) void gemm_bf16x3_kernel(GemmGroups gg) {
  constexpr int BK = 32;
  constexpr int BM = 64 * WM, BN = 128, NT = 128 * WM;
  // LDS rows of 32 bf16 (64 bytes, no padding); the 16-byte chunk q of row r
  // sits at position q ^ sw(r), sw(r) = ((r >> 1) ^ (r >> 2)) & 3, which keeps both the row-wise
  // stores and the MFMA fragment reads free of bank conflicts.
  constexpr int LDK = BK;
  constexpr int LA = BM * 4 / NT, LB = BN * 4 / NT;  // row loads per thread and array
  __shared__ __attribute__((aligned(16))) unsigned short Ah[2][BM * LDK];
  __shared__ __attribute__((aligned(16))) unsigned short Al[2][BM * LDK];
  __shared__ __attribute__((aligned(16))) unsigned short Bh[2][BN * LDK];
  __shared__ __attribute__((aligned(16))) unsigned short Bl[2][BN * LDK];

  const int gi = (gg.n > 1) ? (int)blockIdx.y : 0;
  const GemmArgs p = select_group(gg, gi);
  if ((int)blockIdx.x >= p.tiles_m * p.tiles_n || (int)blockIdx.z >= p.splits) return;
  int tm, tn;
  tile_coords(p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const int bz = (gg.n > 1) ? 0 : (int)blockIdx.y;
  // split-K: k range [kz, kz + kchunk), kchunk a multiple of 32 (keeps the
  // 16-byte row loads aligned)
  const long kz = (long)blockIdx.z * p.kchunk;
  const unsigned short* __restrict__ Ahg = reinterpret_cast<const unsigned short*>(p.A) + (long)bz * p.sA + kz;
  const unsigned short* __restrict__ Alg = reinterpret_cast<const unsigned short*>(p.Alo) + (long)bz * p.sA + kz;
  const unsigned short* __restrict__ Bhg = reinterpret_cast<const unsigned short*>(p.B) + (long)bz * p.sB + kz;
  const unsigned short* __restrict__ Blg = reinterpret_cast<const unsigned short*>(p.Blo) + (long)bz * p.sB + kz;
  double* __restrict__ C = p.C + (long)bz * p.sC + (long)blockIdx.z * p.zC;
  const int M = p.M, N = p.N;
  const int K = (int)((p.K - kz) < p.kchunk ? (p.K - kz) : p.kchunk);

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave >> 1;
  const int wn = wave & 1;

  // Loader: four consecutive lanes read one row's 32 k values (64 contiguous
  // bytes, two whole 32-byte sectors) per array; rows (t >> 2) + c NT/4.
  const int lrow = t >> 2;
  const int lkb = (t & 3) * 8;
  uint4 rah[LA], ral[LA], rbh[LB], rbl[LB];
  auto load = [&](int k0) {
    const int k = k0 + lkb;
    const bool ka = (k < K);
#pragma unroll
    for (int c = 0; c < LA; ++c) {
      const int m = m0 + lrow + (NT / 4) * c;
      if (m < M && ka) {
        rah[c] = *reinterpret_cast<const uint4*>(Ahg + (long)m * p.lda + k);
        ral[c] = *reinterpret_cast<const uint4*>(Alg + (long)m * p.lda + k);
      } else {
        rah[c] = make_uint4(0, 0, 0, 0);
        ral[c] = make_uint4(0, 0, 0, 0);
      }
    }
#pragma unroll
    for (int c = 0; c < LB; ++c) {
      const int n = n0 + lrow + (NT / 4) * c;
      if (n < N && ka) {
        rbh[c] = *reinterpret_cast<const uint4*>(Bhg + (long)n * p.ldb + k);
        rbl[c] = *reinterpret_cast<const uint4*>(Blg + (long)n * p.ldb + k);
      } else {
        rbh[c] = make_uint4(0, 0, 0, 0);
        rbl[c] = make_uint4(0, 0, 0, 0);
      }
    }
  };
  // (NT / 4) * c is a multiple of 16, so the swizzle of row lrow + (NT/4) c is that of lrow
  const int soff = lrow * LDK + (((t & 3) ^ lds_swizzle(lrow)) << 3);
  auto store = [&](int buf) {
#pragma unroll
    for (int c = 0; c < LA; ++c) {
      const int off = soff + (NT / 4) * c * LDK;
      *reinterpret_cast<uint4*>(&Ah[buf][off]) = rah[c];
      *reinterpret_cast<uint4*>(&Al[buf][off]) = ral[c];
    }
#pragma unroll
    for (int c = 0; c < LB; ++c) {
      const int off = soff + (NT / 4) * c * LDK;
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
  const int rsw = (lk ^ lds_swizzle(lr)) << 3;  // swizzled chunk of k values lk*8..lk*8+7
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
    s8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int off = (wm * 64 + i * 16 + lr) * LDK + rsw;
      ah[i] = *reinterpret_cast<const s8_t*>(&Ah[buf][off]);
      al[i] = *reinterpret_cast<const s8_t*>(&Al[buf][off]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int off = (wn * 64 + j * 16 + lr) * LDK + rsw;
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

// ---------------------------------------------------------------------------
// bf16x3 kernel with direct-to-LDS operand loads (global_load_lds_dwordx4)
// and an S-stage LDS ring: the loads of k tile kt + S - 1 are in flight while
// tile kt is multiplied, and no VGPR holds operand data on its way to LDS.
// Block (64 WM) x (32 FN) on 2 WM waves, each a 64 x 16 FN tile; same MFMA
// sequence and LDS row swizzle as gemm_bf16x3_kernel.  FN = 8 (256 x 256
// blocks) halves the operand bytes per flop of the 256 x 128 shape.
// Requires K % 32 == 0 (no k tail: every loaded chunk lies inside the
// operand rows).
//
// Loads: one wave instruction moves 16 rows x 64 bytes of one array into 1 KB
// of LDS (lane L -> row L / 4, LDS chunk position L % 4); the lane fetches the
// global chunk (L % 4) ^ sw(row), which puts chunk q of row r at position
// q ^ sw(r).  Rows past M (N) are clamped to the last row:
// they only feed accumulator rows (columns) that are never stored.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int vmcnt_imm(int n) {
  return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14);
}

template <int WM, int FN, int S>
__global__ __launch_bounds__(128 * WM) void gemm_bf16x3_dl_kernel(GemmGroups gg) {
  constexpr int BK = 32;
  constexpr int BM = 64 * WM, BN = 32 * FN, NT = 128 * WM, NW = NT / 64;  // wave tile 64 x 16 FN
  constexpr int GA = BM / 16, GB = BN / 16, NG = 2 * GA + 2 * GB;  // 16-row groups per stage
  constexpr int IPW = NG / NW;                                      // load instructions per wave and stage
  constexpr int SA = BM * BK, SB = BN * BK;                         // elements per array and stage
  constexpr int STAGE = 2 * SA + 2 * SB;
  static_assert(NG % NW == 0, "groups per wave");
  __shared__ __attribute__((aligned(16))) unsigned short lds[S * STAGE];

  const int gi = (gg.n > 1) ? (int)blockIdx.y : 0;
  const GemmArgs p = select_group(gg, gi);
  if ((int)blockIdx.x >= p.tiles_m * p.tiles_n || (int)blockIdx.z >= p.splits) return;
  int tm, tn;
  tile_coords(p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const int bz = (gg.n > 1) ? 0 : (int)blockIdx.y;
  const long kz = (long)blockIdx.z * p.kchunk;
  const unsigned short* __restrict__ Ahg = reinterpret_cast<const unsigned short*>(p.A) + (long)bz * p.sA + kz;
  const unsigned short* __restrict__ Alg = reinterpret_cast<const unsigned short*>(p.Alo) + (long)bz * p.sA + kz;
  const unsigned short* __restrict__ Bhg = reinterpret_cast<const unsigned short*>(p.B) + (long)bz * p.sB + kz;
  const unsigned short* __restrict__ Blg = reinterpret_cast<const unsigned short*>(p.Blo) + (long)bz * p.sB + kz;
  double* __restrict__ C = p.C + (long)bz * p.sC + (long)blockIdx.z * p.zC;
  const int M = p.M, N = p.N;
  const int K = (int)((p.K - kz) < p.kchunk ? (p.K - kz) : p.kchunk);

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave >> 1;
  const int wn = wave & 1;

  // per-instruction global row pointer (this lane) and LDS group offset (wave-uniform)
  const int lrow = lane >> 2;
  const int lchunk = (lane & 3) ^ lds_swizzle(lrow);
  const unsigned short* gp[IPW];
  int goff[IPW];
#pragma unroll
  for (int j = 0; j < IPW; ++j) {
    const int G = wave * IPW + j;
    const unsigned short* base;
    long ld;
    int row, nrows, off;
    if (G < GA) {
      base = Ahg; ld = p.lda; row = m0 + G * 16 + lrow; nrows = M; off = G * 16 * BK;
    } else if (G < 2 * GA) {
      base = Alg; ld = p.lda; row = m0 + (G - GA) * 16 + lrow; nrows = M; off = SA + (G - GA) * 16 * BK;
    } else if (G < 2 * GA + GB) {
      base = Bhg; ld = p.ldb; row = n0 + (G - 2 * GA) * 16 + lrow; nrows = N; off = 2 * SA + (G - 2 * GA) * 16 * BK;
    } else {
      base = Blg; ld = p.ldb; row = n0 + (G - 2 * GA - GB) * 16 + lrow; nrows = N;
      off = 2 * SA + SB + (G - 2 * GA - GB) * 16 * BK;
    }
    row = row < nrows ? row : nrows - 1;
    gp[j] = base + (long)row * ld + lchunk * 8;
    goff[j] = off;
  }
  auto issue = [&](int kt) {
#if __HIP_DEVICE_COMPILE__
    // (the host pass parses kernel bodies too but has no gfx950 target features)
    unsigned short* st = lds + (kt % S) * STAGE;
#pragma unroll
    for (int j = 0; j < IPW; ++j)
      __builtin_amdgcn_global_load_lds(gp[j] + kt * BK, (__attribute__((address_space(3))) void*)(st + goff[j]),
                                       16, 0, 0);
#endif
  };

  f4_t acc[4][FN];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f4_t{0, 0, 0, 0};

  const int nk = K / BK;
  const int lr = lane & 15;
  const int lk = lane >> 4;
  const int rsw = (lk ^ lds_swizzle(lr)) << 3;
#pragma unroll
  for (int s = 0; s < S - 1; ++s)
    if (s < nk) issue(s);
  for (int kt = 0; kt < nk; ++kt) {
    // this wave's loads of tile kt are done when at most min(S - 2, nk - 1 - kt) later tiles remain
    const int later = (nk - 1 - kt) < (S - 2) ? (nk - 1 - kt) : (S - 2);
    if (S >= 4 && later >= 2)
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(2 * IPW));
    else if (S >= 3 && later >= 1)
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(IPW));
    else
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
    // plain s_barrier: a __syncthreads() fence would wait for every load in
    // flight (the later stages too); the wait above covers this tile
    __builtin_amdgcn_s_barrier();
    const unsigned short* st = lds + (kt % S) * STAGE;
    s8_t ah[4], al[4], bh[FN], bl[FN];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int off = (wm * 64 + i * 16 + lr) * BK + rsw;
      ah[i] = *reinterpret_cast<const s8_t*>(st + off);
      al[i] = *reinterpret_cast<const s8_t*>(st + SA + off);
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int off = (wn * 16 * FN + j * 16 + lr) * BK + rsw;
      bh[j] = *reinterpret_cast<const s8_t*>(st + 2 * SA + off);
      bl[j] = *reinterpret_cast<const s8_t*>(st + 2 * SA + SB + off);
    }
    // the next tile's loads go out after this tile's fragment reads, so that
    // their LDS writes do not compete with the reads that start the products
    if (kt + S - 1 < nk) issue(kt + S - 1);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf8_t, al[i]), __builtin_bit_cast(bf8_t, bh[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf8_t, ah[i]), __builtin_bit_cast(bf8_t, bl[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf8_t, ah[i]), __builtin_bit_cast(bf8_t, bh[j]), acc[i][j], 0, 0, 0);
      }
    }
  }

  const double alpha = p.alpha, beta = p.beta;
  unsigned short* __restrict__ Chi = p.Chi;
  unsigned short* __restrict__ Clo = p.Clo;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn * 16 * FN + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + 4 * lk + r;
        if (row < M && col < N) {
          double* c = C + (long)row * p.ldc + col;
          double v = alpha * (double)acc[i][j][r];
          if (beta != 0.0) v += beta * (*c);
          if (Chi) {  // the result as the bf16 hi/lo operand of a later product
            const unsigned short h = bf16_rn((float)v);
            Chi[(long)row * p.ldcl + col] = h;
            Clo[(long)row * p.ldcl + col] = bf16_rn((float)(v - (double)bf16_to_f(h)));
          } else {
            *c = v;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// fp64 kernel with direct-to-LDS operand loads (global_load_lds_dwordx4) into
// an S-stage LDS ring; K % 16 == 0.  Block (16 FM WM) x (16 FN WN) on WM x WN
// waves, each a (16 FM) x (16 FN) tile of v_mfma_f64_16x16x4_f64.
// LDS layout per operand: k-contiguous (KC) rows hold 16 doubles (128 bytes,
// eight 16-byte chunks) with chunk q of row r at position q ^ (r & 7); a
// row-contiguous operand keeps its own [k][row] order (one k row of the tile
// is 8 ROWS bytes).  Either way one wave instruction moves 1 KB (64 lanes x
// 16 bytes) and the 16 rows of one fragment read hit distinct banks (KC: at
// most two rows per bank group).  Rows past M (N) are clamped to the last row
// (KC) or the last row pair (row-contiguous, even row counts): they only feed
// accumulators that are never stored.
// ---------------------------------------------------------------------------
template <bool KC, int ROWS>
struct DlOperand {
  static_assert(KC || (ROWS <= 128 && 128 % ROWS == 0), "row-contiguous tiles: one k row per 1 KB or less");
  static constexpr int BK = 16;
  static constexpr int NI = ROWS / 8;  // 1 KB wave instructions per stage
  // global pointer (this lane) and LDS offset (wave-uniform) of instruction G
  __device__ static inline void map(const double* base, long ld, int row0, int nrows, int G, int lane,
                                    const double*& gp, int& off) {
    if (KC) {
      const int r = lane >> 3;
      int row = row0 + G * 8 + r;
      row = row < nrows ? row : nrows - 1;
      gp = base + (long)row * ld + (((lane & 7) ^ r) << 1);
      off = G * 8 * BK;
    } else {
      constexpr int HALF = ROWS / 2;  // lanes per k row
      const int kr = G * (128 / ROWS) + lane / HALF;
      int m = row0 + 2 * (lane % HALF);
      m = m <= nrows - 2 ? m : nrows - 2;
      gp = base + (long)kr * ld + m;
      off = G * 128;
    }
  }
  __device__ static inline long kstep(long ld) { return KC ? BK : BK * ld; }
  __device__ static inline double read(const double* st, int r, int k) {
    return KC ? st[r * BK + (((k >> 1) ^ (r & 7)) << 1) + (k & 1)] : st[k * ROWS + r];
  }
};

template <bool AKC, bool BKC, int FM, int FN, int S, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void gemm_f64_dl_kernel(GemmGroups gg) {
  constexpr int BK = 16, NW = WM * WN;
  constexpr int BM = 16 * FM * WM, BN = 16 * FN * WN;
  typedef DlOperand<AKC, BM> OA;
  typedef DlOperand<BKC, BN> OB;
  constexpr int NG = OA::NI + OB::NI, IPW = NG / NW;
  constexpr int SA = BM * BK, STAGE = (BM + BN) * BK;  // doubles
  static_assert(NG % NW == 0, "groups per wave");
  __shared__ __attribute__((aligned(16))) double lds[S * STAGE];

  const int gi = (gg.n > 1) ? (int)blockIdx.y : 0;
  const GemmArgs p = select_group(gg, gi);
  if ((int)blockIdx.x >= p.tiles_m * p.tiles_n || (int)blockIdx.z >= p.splits) return;
  int tm, tn;
  tile_coords(p.tiles_m, p.tiles_n, tm, tn);
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const int bz = (gg.n > 1) ? 0 : (int)blockIdx.y;
  const long kz = (long)blockIdx.z * p.kchunk;
  const double* __restrict__ Ag = reinterpret_cast<const double*>(p.A) + (long)bz * p.sA + kz * p.zkA;
  const double* __restrict__ Bg = reinterpret_cast<const double*>(p.B) + (long)bz * p.sB + kz * p.zkB;
  double* __restrict__ C = p.C + (long)bz * p.sC + (long)blockIdx.z * p.zC;
  const int M = p.M, N = p.N;
  const int K = (int)((p.K - kz) < p.kchunk ? (p.K - kz) : p.kchunk);

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave / WN;
  const int wn = wave % WN;

  const double* gp[IPW];
  int goff[IPW];
  long gstep[IPW];
#pragma unroll
  for (int j = 0; j < IPW; ++j) {
    const int G = wave * IPW + j;
    if (G < OA::NI) {
      OA::map(Ag, p.lda, m0, M, G, lane, gp[j], goff[j]);
      gstep[j] = OA::kstep(p.lda);
    } else {
      OB::map(Bg, p.ldb, n0, N, G - OA::NI, lane, gp[j], goff[j]);
      goff[j] += SA;
      gstep[j] = OB::kstep(p.ldb);
    }
  }
  auto issue = [&](int kt) {
#if __HIP_DEVICE_COMPILE__
    double* st = lds + (kt % S) * STAGE;
#pragma unroll
    for (int j = 0; j < IPW; ++j)
      __builtin_amdgcn_global_load_lds(gp[j] + kt * gstep[j], (__attribute__((address_space(3))) void*)(st + goff[j]),
                                       16, 0, 0);
#endif
  };

  d4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = d4_t{0, 0, 0, 0};

  const int nk = K / BK;
  const int lr = lane & 15;
  const int lk = lane >> 4;
#pragma unroll
  for (int s = 0; s < S - 1; ++s)
    if (s < nk) issue(s);
  for (int kt = 0; kt < nk; ++kt) {
    const int later = (nk - 1 - kt) < (S - 2) ? (nk - 1 - kt) : (S - 2);
    if (S >= 4 && later >= 2)
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(2 * IPW));
    else if (S >= 3 && later >= 1)
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(IPW));
    else
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
    __builtin_amdgcn_s_barrier();
    if (kt + S - 1 < nk) issue(kt + S - 1);
    const double* st = lds + (kt % S) * STAGE;
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int k = kk * 4 + lk;
      double a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = OA::read(st, wm * 16 * FM + i * 16 + lr, k);
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = OB::read(st + SA, wn * 16 * FN + j * 16 + lr, k);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = Mfma16<double>::op(a[i], b[j], acc[i][j]);
    }
  }

  const double alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn * 16 * FN + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 16 * FM + i * 16 + Mfma16<double>::row(lk, r);
        if (row < M && col < N) {
          double* c = C + (long)row * p.ldc + col;
          const double v = alpha * acc[i][j][r];
          *c = (beta == 0.0) ? v : v + beta * (*c);
        }
      }
    }
  }
}

// C[z-sum] reduction for split-K: C = alpha-scaled partial sums (already in ws) + beta*C.
__global__ void splitk_reduce_kernel(int M, int N, int splits, int batch, const double* __restrict__ ws, long zC,
                                     double* __restrict__ C, long ldc, long sC, double beta) {
  const long per = (long)M * N;
  const long total = per * batch;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long bz = idx / per, r = idx - bz * per;
    const long row = r / N, col = r - row * N;
    double acc = 0.0;
    for (int z = 0; z < splits; ++z) acc += ws[z * zC + bz * per + r];
    double* c = C + bz * sC + row * ldc + col;
    *c = (beta == 0.0) ? acc : acc + beta * (*c);
  }
}

// Many splits over a small output: one wave per output element, lane l
// summing splits l, l + 64, ... and the lanes combined by a fixed shuffle
// tree (deterministic), so the partial sums are read 64 at a time instead of
// one after another.
__device__ inline void splitk_reduce_wave(long idx, long total, int N, int splits, const double* __restrict__ ws,
                                          long zC, double* __restrict__ C, long ldc, double beta) {
  const int lane = threadIdx.x & 63;
  if (idx >= total) return;
  double acc = 0.0;
  for (int z = lane; z < splits; z += 64) acc += ws[z * zC + idx];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) {
    const long row = idx / N, col = idx - row * N;
    double* c = C + row * ldc + col;
    *c = (beta == 0.0) ? acc : acc + beta * (*c);
  }
}

__global__ __launch_bounds__(256) void splitk_reduce_wide_kernel(int M, int N, int splits, int batch,
                                                                 const double* __restrict__ ws, long zC,
                                                                 double* __restrict__ C, long ldc, long sC,
                                                                 double beta) {
  const long per = (long)M * N;
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= per * batch) return;
  const long bz = w / per, r = w - bz * per;
  splitk_reduce_wave(r, per, N, splits, ws + bz * per, zC, C + bz * sC, ldc, beta);
}

constexpr int kWideSplits = 64;           // splits from which a wave reduces one element
constexpr long kWideMaxOutputs = 1L << 16;  // outputs up to which that pays

// The split-K reductions of every group of one grouped launch in one launch
// (blockIdx.y = group; groups without splits have splits = 0).
struct ReduceGroups {
  int n;
  struct G {
    int M, N, splits;
    const double* ws;
    long zC;
    double* C;
    long ldc;
    double beta;
  } g[kMaxGroups];
};

__global__ void splitk_reduce_groups_kernel(ReduceGroups rg, int wide) {
  ReduceGroups::G q = rg.g[0];
#pragma unroll
  for (int k = 1; k < kMaxGroups; ++k)
    if (k == (int)blockIdx.y) q = rg.g[k];
  if (q.splits <= 1) return;
  if (wide) {
    const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    splitk_reduce_wave(w, (long)q.M * q.N, q.N, q.splits, q.ws, q.zC, q.C, q.ldc, q.beta);
    return;
  }
  const long total = (long)q.M * q.N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long row = idx / q.N, col = idx - row * q.N;
    double acc = 0.0;
    for (int z = 0; z < q.splits; ++z) acc += q.ws[z * q.zC + idx];
    double* c = q.C + row * q.ldc + col;
    *c = (q.beta == 0.0) ? acc : acc + q.beta * (*c);
  }
}

// Tile configuration of the fp64/fp32 kernel: BK, waves along M/N, MFMA
// 16x16 blocks per wave along M/N.
// Measured on MI355X (o=48, v=320 contraction shapes): fp64 runs best with
// 64 x 64 blocks of four 32 x 32 waves, fp32 with 64 x 128 blocks of four
// 32 x 64 waves.
template <int BK_, int WM_, int WN_, int FM_, int FN_>
struct Cfg {
  static constexpr int BK = BK_, WM = WM_, WN = WN_, FM = FM_, FN = FN_;
  static constexpr int BM = 16 * FM * WM, BN = 16 * FN * WN, NT = 64 * WM * WN;
};
template <typename T>
struct TileCfg;
template <>
struct TileCfg<double> {
  typedef Cfg<16, 2, 2, 2, 2> type;
};
template <>
struct TileCfg<float> {
  typedef Cfg<16, 2, 2, 2, 4> type;
};

template <typename T>
static void launch16(const GemmGroups& a, bool akc, bool bkc, dim3 grid, hipStream_t s) {
  typedef typename TileCfg<T>::type C64;
  const dim3 blk(C64::NT);
#define L16(AK, BK2) \
  hipLaunchKernelGGL((gemm16_kernel<T, AK, BK2, C64::BK, C64::WM, C64::WN, C64::FM, C64::FN>), grid, blk, 0, s, a)
  if (akc && bkc)
    L16(true, true);
  else if (akc && !bkc)
    L16(true, false);
  else if (!akc && bkc)
    L16(false, true);
  else
    L16(false, false);
#undef L16
}

// LDS ring depth of the direct-to-LDS bf16x3 kernel per block shape
static constexpr int kDlStages2 = 2, kDlStages4 = 3, kDlStagesWide = 2;

// The direct-to-LDS bf16x3 kernel needs whole k tiles in every problem.
static bool bf16_direct(const GemmDesc* d, int ng) {
  for (int g = 0; g < ng; ++g)
    if (d[g].M > 0 && d[g].N > 0 && d[g].K % 32 != 0) return false;
  return true;
}

// bf16x3 block shape: 2 = 128 x 128 (some problem has M < 1024); 8 = 256 x 256
// (direct-to-LDS kernel; every problem at least 1024 x 1024 and with at least
// 256 such tiles or a split-K workspace to fill the chip); else 4 = 256 x 128.
static int bf16_shape(const GemmDesc* d, int ng) {
  if (ng == 1 && d[0].tile_hint == 8 && bf16_direct(d, 1) && d[0].M >= 256 && d[0].N >= 256 &&
      (long)((d[0].M + 255) / 256) * ((d[0].N + 255) / 256) * d[0].batch >= 256)
    return 8;
  bool any = false, wide = bf16_direct(d, ng);
  for (int g = 0; g < ng; ++g)
    if (d[g].M > 0 && d[g].N > 0) {
      if (d[g].M < 1024) return 2;
      const long tiles = (long)((d[g].M + 255) / 256) * ((d[g].N + 255) / 256);
      wide = wide && d[g].N >= 1024 && (tiles >= 256 || (d[0].ws != nullptr && d[g].K >= 1024));
      any = true;
    }
  return !any ? 2 : (wide ? 8 : 4);
}
// fp64 problems with 16-byte aligned rows, K % 16 == 0, M >= 256 and K >= 256
// (or K >= 32 with N >= 256: 16 % faster on 737280 x 320 x 48, slower on
// narrow outputs) run on the direct-to-LDS kernel: 128 x 128 blocks of 2 x 4 waves
// (64 x 32 each), two LDS stages (64 KB, two blocks per CU).  Measured against
// the 64 x 64 register-staged kernel and other direct-to-LDS shapes (128 x 64
// on 4 waves with 3 stages, 128 x 128 on 4 or 16 waves, 3 or 4 stages): 70 TF
// on 15360^3 and 62.5 TF on the 66720 x 1176 x 51360 ladder against 55 and 50.
static constexpr int kShapeF64Dl = 16;
static constexpr int kF64DlFM = 4, kF64DlFN = 2, kF64DlStages = 2, kF64DlWM = 2, kF64DlWN = 4;
static bool f64_direct(const GemmDesc* d, int ng) {
  bool any = false;
  for (int g = 0; g < ng; ++g)
    if (d[g].M > 0 && d[g].N > 0) {
      const GemmDesc& e = d[g];
      // (short K too when the output is wide: the write-bound rank-o updates)
      if ((e.K < 256 && !(e.K >= 32 && e.N >= 256)) || e.K % 16 != 0 || e.M < 256) return false;
      if ((uintptr_t)e.A % 16 || (uintptr_t)e.B % 16 || e.lda % 2 || e.ldb % 2 || e.sA % 2 || e.sB % 2) return false;
      // row-contiguous operands are loaded in row pairs
      if ((!e.a_kc && e.M % 2) || (!e.b_kc && e.N % 2)) return false;
      any = true;
    }
  return any;
}
static int gemm_shape(const GemmDesc* d, int ng, GemmPrec prec) {
  if (prec == GP_BF16X3) return bf16_shape(d, ng);
  if (prec == GP_F64 && f64_direct(d, ng)) return kShapeF64Dl;
  return 0;
}
static int bf16_bm(int shape) { return shape == 2 ? 128 : 256; }
static int bf16_bn(int shape) { return shape == 8 ? 256 : 128; }

// Fill the kernel arguments of one problem; split-K when the output cannot
// fill the chip and K is long, with partial products
// in ws[ws_off ...].  Returns the number of splits.
// Output block of each kernel: rows x columns.
static int block_rows(GemmPrec prec, int shape) {
  if (prec == GP_BF16X3) return bf16_bm(shape);
  if (prec == GP_F64 && shape == kShapeF64Dl) return 16 * kF64DlFM * kF64DlWM;
  return (prec == GP_F32) ? TileCfg<float>::type::BM : TileCfg<double>::type::BM;
}
static int block_cols(GemmPrec prec, int shape) {
  if (prec == GP_BF16X3) return bf16_bn(shape);
  if (prec == GP_F64 && shape == kShapeF64Dl) return 16 * kF64DlFN * kF64DlWN;
  return (prec == GP_F32) ? TileCfg<float>::type::BN : TileCfg<double>::type::BN;
}

static int prepare(const GemmDesc& d, GemmPrec prec, GemmArgs& a, double* ws, long ws_avail, long ws_off,
                   long tiles_total, int bf_wm) {
  a.M = d.M; a.N = d.N; a.K = d.K;
  a.A = d.A; a.Alo = d.A_lo; a.lda = d.lda;
  a.B = d.B; a.Blo = d.B_lo; a.ldb = d.ldb;
  a.C = d.C; a.ldc = d.ldc;
  a.alpha = d.alpha; a.beta = d.beta;
  a.sA = d.sA; a.sB = d.sB; a.sC = d.sC;
  a.Chi = nullptr; a.Clo = nullptr; a.ldcl = 0;  // set by gemm() where the kernel writes the split
  const int bm = block_rows(prec, bf_wm), bn = block_cols(prec, bf_wm);
  a.tiles_m = (d.M + bm - 1) / bm;
  a.tiles_n = (d.N + bn - 1) / bn;
  {
    const int esz = (prec == GP_F32) ? 4 : 8;
    auto aligned = [&](const void* ptr, long ld, long sb) {
      return ((uintptr_t)ptr % 16 == 0) && ((ld * esz) % 16 == 0) && ((sb * esz) % 16 == 0);
    };
    a.vecA = aligned(d.A, d.lda, d.sA) ? 1 : 0;
    a.vecB = aligned(d.B, d.ldb, d.sB) ? 1 : 0;
  }
  a.kchunk = d.K > 0 ? d.K : 1;
  a.zkA = d.a_kc ? 1 : d.lda;
  a.zkB = d.b_kc ? 1 : d.ldb;
  a.zC = 0;
  a.splits = 1;
  int splits = 1;
  if (ws != nullptr && d.K >= 1024 && tiles_total > 0 && tiles_total < 512) {
    long sp = 1024 / tiles_total;
    const long kmax = d.K / 512;
    if (sp > kmax) sp = kmax;
    const long per = (long)d.M * d.N * d.batch;
    if (sp * per > ws_avail) sp = ws_avail / per;
    if (sp > 1024) sp = 1024;
    if (sp >= 2) splits = (int)sp;
  }
  if (splits > 1) {
    long kc = (d.K + splits - 1) / splits;
    kc = (kc + 31) / 32 * 32;
    splits = (int)((d.K + kc - 1) / kc);
    a.kchunk = kc;
    a.C = ws + ws_off;
    a.ldc = d.N;
    a.sC = (long)d.M * d.N;
    a.zC = (long)d.M * d.N * d.batch;
    a.beta = 0.0;
    a.splits = splits;
  }
  return splits;
}

static int check_desc(const GemmDesc& d, GemmPrec prec) {
  ACES_REQUIRE(d.M >= 0 && d.N >= 0 && d.K >= 0 && d.batch >= 1, "gemm: bad dimensions");
  if (d.M == 0 || d.N == 0) return ACES_OK;
  ACES_REQUIRE(d.C != nullptr && d.ldc >= d.N, "gemm: bad C");
  ACES_REQUIRE(d.C_hi == nullptr || (d.C_lo != nullptr && d.ldc_lp >= d.N && d.batch == 1),
               "gemm: bad split output");
  if (d.K > 0) {
    ACES_REQUIRE(d.A != nullptr && d.B != nullptr, "gemm: null operand");
    ACES_REQUIRE(d.lda >= (d.a_kc ? d.K : d.M), "gemm: lda too small");
    ACES_REQUIRE(d.ldb >= (d.b_kc ? d.K : d.N), "gemm: ldb too small");
  }
  if (prec == GP_BF16X3 && d.K > 0) {
    ACES_REQUIRE(d.a_kc && d.b_kc, "gemm bf16x3: operands must be k-contiguous");
    ACES_REQUIRE(d.A_lo != nullptr && d.B_lo != nullptr, "gemm bf16x3: missing lo parts");
    ACES_REQUIRE(d.lda % 8 == 0 && d.ldb % 8 == 0 && d.K % 8 == 0 && d.sA % 8 == 0 && d.sB % 8 == 0,
                 "gemm bf16x3: lda, ldb, K and strides must be multiples of 8");
    ACES_REQUIRE(((uintptr_t)d.A % 16) == 0 && ((uintptr_t)d.B % 16) == 0 && ((uintptr_t)d.A_lo % 16) == 0 &&
                     ((uintptr_t)d.B_lo % 16) == 0,
                 "gemm bf16x3: operands must be 16-byte aligned");
  }
  return ACES_OK;
}

static int launch_groups(const GemmGroups& gg, GemmPrec prec, bool akc, bool bkc, bool kzero, dim3 grid, int bf_wm,
                         bool bf_dl, hipStream_t stream) {
  if (kzero) {
    // C = beta * C via a K = 0 launch of the fp64 kernel (acc stays 0).
    launch16<double>(gg, true, true, grid, stream);
  } else {
    switch (prec) {
      case GP_F64:
        if (bf_wm == kShapeF64Dl) {
#define F64DL(AK, BK2)                                                                                        \
  hipLaunchKernelGGL((gemm_f64_dl_kernel<AK, BK2, kF64DlFM, kF64DlFN, kF64DlStages, kF64DlWM, kF64DlWN>), grid,  \
                     dim3(64 * kF64DlWM * kF64DlWN), 0, stream, gg)
          if (akc && bkc)
            F64DL(true, true);
          else if (akc)
            F64DL(true, false);
          else if (bkc)
            F64DL(false, true);
          else
            F64DL(false, false);
#undef F64DL
        } else {
          launch16<double>(gg, akc, bkc, grid, stream);
        }
        break;
      case GP_F32:
        launch16<float>(gg, akc, bkc, grid, stream);
        break;
      case GP_BF16X3:
        if (bf_dl && bf_wm == 8)
          hipLaunchKernelGGL((gemm_bf16x3_dl_kernel<4, 8, kDlStagesWide>), grid, dim3(512), 0, stream, gg);
        else if (bf_dl && bf_wm == 4)
          hipLaunchKernelGGL((gemm_bf16x3_dl_kernel<4, 4, kDlStages4>), grid, dim3(512), 0, stream, gg);
        else if (bf_dl)
          hipLaunchKernelGGL((gemm_bf16x3_dl_kernel<2, 4, kDlStages2>), grid, dim3(256), 0, stream, gg);
        else if (bf_wm == 4)
          hipLaunchKernelGGL(gemm_bf16x3_kernel<4>, grid, dim3(512), 0, stream, gg);
        else
          hipLaunchKernelGGL(gemm_bf16x3_kernel<2>, grid, dim3(256), 0, stream, gg);
        break;
      default:
        ACES_REQUIRE(false, "gemm: unknown precision");
    }
  }
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

static int reduce_splits(const GemmDesc& d, const GemmArgs& a, hipStream_t stream) {
  if (a.splits <= 1) return ACES_OK;
  const long total = (long)d.M * d.N * d.batch;
  if (a.splits >= kWideSplits && total <= kWideMaxOutputs) {
    hipLaunchKernelGGL(splitk_reduce_wide_kernel, dim3((unsigned)((total + 3) / 4)), dim3(256), 0, stream, d.M, d.N,
                       a.splits, d.batch, a.C, a.zC, d.C, d.ldc, d.sC, d.beta);
    ACES_HIP(hipGetLastError());
    return ACES_OK;
  }
  long g = (total + 255) / 256;
  if (g > (1L << 20)) g = 1L << 20;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)g), dim3(256), 0, stream, d.M, d.N, a.splits, d.batch,
                     a.C, a.zC, d.C, d.ldc, d.sC, d.beta);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// Skinny fp64 products: one side of the output at most 64 long (the short
// side), the other (long) side taken 16 at a time.  Such products read far
// more operand bytes than they compute on, so they run as HBM streams rather
// than on 64 x 64 or 128 x 128 tiles that would idle most of their MFMA lanes:
//   short K (K <= 64 / TM): each wave keeps the short operand's fragments in
//     registers and walks 16-wide chunks of the long side over the whole K;
//   long K (both output sides <= 64): each wave takes one 16-wide chunk and
//     one K range; partial sums go to the split-K workspace and the grouped
//     reduction adds them in split order (one split writes C itself).
// S(s, k) = S[s*ss + k*sk], L(l, k) = L[l*ls + k*lk]; output (s, l) at
// C[s*cs + l*cl].  The side whose index is contiguous in C goes on the MFMA
// columns (swap = 1: long side on the rows), so that stores run along lanes.
struct SkinnyArgs {
  const double* S;
  long ss, sk;
  const double* L;
  long ls, lk;
  double* C;
  long cs, cl;
  int ns, nl, K, nchunks;
  double alpha, beta;
  int splits, swap;
  long kchunk;
  long zW;  // partial sums of split z at C + z * zW (long K)
};
struct SkinnyGroups {
  int n;
  SkinnyArgs g[kMaxGroups];
};
__device__ inline SkinnyArgs select_skinny(const SkinnyGroups& sg, int gi) {
  SkinnyArgs p = sg.g[0];
#pragma unroll
  for (int k = 1; k < kMaxGroups; ++k)
    if (k == gi) p = sg.g[k];
  return p;
}

constexpr int kSkinnyShort = 64;    // longest short side
constexpr int kSkinnyLong = 4096;   // shortest long side (short K) / K (long K)

// 16 x 16 output block of one MFMA chain: short-side tile t against the
// 16-wide long-side chunk; element r of lane (lr, lg) is (s, l) below.
__device__ inline d4_t skinny_mfma(double sv, double lv, d4_t acc, int swap) {
  return swap ? __builtin_amdgcn_mfma_f64_16x16x4f64(lv, sv, acc, 0, 0, 0)
              : __builtin_amdgcn_mfma_f64_16x16x4f64(sv, lv, acc, 0, 0, 0);
}
__device__ inline int skinny_s(int t, int lr, int lg, int r, int swap) { return 16 * t + (swap ? lr : lg + 4 * r); }
__device__ inline int skinny_l(int l0, int lr, int lg, int r, int swap) { return l0 + (swap ? lg + 4 * r : lr); }

template <int TM>
__global__ __launch_bounds__(256) void skinny_f64_shortk_kernel(SkinnyGroups sg) {
  constexpr int KSM = 16 / TM;  // MFMA k steps held in registers
  const SkinnyArgs p = select_skinny(sg, (int)blockIdx.y);
  if (p.nchunks == 0) return;
  const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const int ks = (p.K + 3) >> 2;
  double sf[TM][KSM];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int j = 0; j < KSM; ++j) {
      const int s = 16 * t + lr, k = 4 * j + lg;
      sf[t][j] = (s < p.ns && k < p.K) ? p.S[(long)s * p.ss + (long)k * p.sk] : 0.0;
    }
  const int nw = (int)gridDim.x * 4;
  for (int c = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6); c < p.nchunks; c += nw) {
    const int l = c * 16 + lr;
    double lf[KSM];
#pragma unroll
    for (int j = 0; j < KSM; ++j) {
      const int k = 4 * j + lg;
      lf[j] = (l < p.nl && k < p.K) ? p.L[(long)l * p.ls + (long)k * p.lk] : 0.0;
    }
    d4_t acc[TM];
#pragma unroll
    for (int t = 0; t < TM; ++t) acc[t] = d4_t{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < KSM; ++j)
      if (j < ks)
#pragma unroll
        for (int t = 0; t < TM; ++t) acc[t] = skinny_mfma(sf[t][j], lf[j], acc[t], p.swap);
#pragma unroll
    for (int t = 0; t < TM; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int so = skinny_s(t, lr, lg, r, p.swap), lo = skinny_l(c * 16, lr, lg, r, p.swap);
        if (so < p.ns && lo < p.nl) {
          double* o = p.C + (long)so * p.cs + (long)lo * p.cl;
          const double v = p.alpha * acc[t][r];
          *o = (p.beta == 0.0) ? v : v + p.beta * (*o);
        }
      }
  }
}

// KC: both operands k-contiguous (sk = lk = 1), the k steps at fixed offsets.
template <int TM, bool KC>
__global__ __launch_bounds__(256) void skinny_f64_longk_kernel(SkinnyGroups sg) {
  constexpr int KSTEP = (TM <= 2) ? 32 : 16;  // k per iteration
  const SkinnyArgs p = select_skinny(sg, (int)blockIdx.y);
  const long item = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (item >= (long)p.nchunks * p.splits) return;
  const int c = (int)(item % p.nchunks), z = (int)(item / p.nchunks);
  const long k0 = (long)z * p.kchunk;
  const long k1 = min((long)p.K, k0 + p.kchunk);
  const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const int l = c * 16 + lr;
  const bool lin = l < p.nl;
  bool sok[TM];
#pragma unroll
  for (int t = 0; t < TM; ++t) sok[t] = 16 * t + lr < p.ns;
  const long lk = KC ? 1 : p.lk, sk = KC ? 1 : p.sk;
  // lane (lr, lg) reads k = kb + 4u + lg of its rows
  const double* Lp = p.L + (long)(lin ? l : 0) * p.ls + (k0 + lg) * lk;
  const double* Sp[TM];
#pragma unroll
  for (int t = 0; t < TM; ++t) Sp[t] = p.S + (long)(sok[t] ? 16 * t + lr : 0) * p.ss + (k0 + lg) * sk;
  d4_t acc[TM];
#pragma unroll
  for (int t = 0; t < TM; ++t) acc[t] = d4_t{0, 0, 0, 0};
  const long kend = k1 - k0 - lg;  // this lane's k offsets run below kend
  for (long kb = 0; kb < k1 - k0; kb += KSTEP) {
    double lv[KSTEP / 4], sv[TM][KSTEP / 4];
#pragma unroll
    for (int u = 0; u < KSTEP / 4; ++u) {
      const bool ok = kb + 4 * u < kend;
      lv[u] = (ok && lin) ? Lp[(kb + 4 * u) * lk] : 0.0;
#pragma unroll
      for (int t = 0; t < TM; ++t) sv[t][u] = (ok && sok[t]) ? Sp[t][(kb + 4 * u) * sk] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < KSTEP / 4; ++u)
#pragma unroll
      for (int t = 0; t < TM; ++t) acc[t] = skinny_mfma(sv[t][u], lv[u], acc[t], p.swap);
  }
  // several splits: partial sums (no beta) at C + z * zW; one split: C itself
  double* W = p.C + (long)z * p.zW;
  const bool fin = p.splits == 1 && p.beta != 0.0;
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int so = skinny_s(t, lr, lg, r, p.swap), lo = skinny_l(c * 16, lr, lg, r, p.swap);
      if (so < p.ns && lo < p.nl) {
        double* o = W + (long)so * p.cs + (long)lo * p.cl;
        const double v = p.alpha * acc[t][r];
        *o = fin ? v + p.beta * (*o) : v;
      }
    }
}

int gemm_skinny_kind(long M, long N, long K) {
  if (M <= 0 || N <= 0 || K <= 0) return 3;  // empty blocks do not decide
  const long sh = std::min(M, N), lo = std::max(M, N);
  if (sh > kSkinnyShort) return 0;
  const int tm = (int)((sh + 15) / 16);
  if (K <= 4 * (16 / tm) && lo >= kSkinnyLong) return 1;
  // long K with the short side at most 32: wider short sides re-read it for
  // every 16-wide chunk of the other, where one 64 x 64 split-K tile reads both once
  return (sh <= 32 && lo <= kSkinnyShort && K >= kSkinnyLong) ? 2 : 0;
}

// Skinny path for fp64 groups (batch 1, fp64 output).  *ran = 1 when it ran,
// 0 when some nonempty group is not skinny (the tiled kernels take it).
static int gemm_skinny(const GemmDesc* d, int ng, hipStream_t stream, int* ran) {
  *ran = 0;
  bool any = false, longk = false, shortk = false;
  int tm = 1;
  for (int g = 0; g < ng; ++g) {
    const GemmDesc& e = d[g];
    if (e.M == 0 || e.N == 0) continue;
    const int kind = gemm_skinny_kind(e.M, e.N, e.K);
    if (e.batch != 1 || e.C_hi != nullptr || kind == 0 || kind == 3) return ACES_OK;
    const int sh = std::min(e.M, e.N);
    tm = std::max(tm, (sh + 15) / 16);
    const bool lk = kind == 2;
    longk = longk || lk;
    shortk = shortk || !lk;
    any = true;
  }
  if (!any || (longk && shortk)) return ACES_OK;
  if (shortk) {
    for (int g = 0; g < ng; ++g)
      if (d[g].M > 0 && d[g].N > 0 && d[g].K > 4 * (16 / tm)) return ACES_OK;
  }
  if (longk) {
    // partial sums need the workspace: at least one full output per group
    long need = 0;
    for (int g = 0; g < ng; ++g) need += (long)d[g].M * d[g].N;
    if (d[0].ws == nullptr || d[0].ws_elems < need) return ACES_OK;
  }
  SkinnyGroups sg;
  sg.n = ng;
  long maxitems = 0, ws_off = 0;
  ReduceGroups rg;
  rg.n = ng;
  long most = 0;
  int minsplit = 1 << 30;
  bool kc_all = true;
  for (int g = 0; g < kMaxGroups; ++g) {
    SkinnyArgs& a = sg.g[g];
    std::memset(&a, 0, sizeof(a));
    rg.g[g].splits = 0;
    if (g >= ng || d[g].M == 0 || d[g].N == 0) continue;
    const GemmDesc& e = d[g];
    const bool tr = e.N < e.M;  // short side on N: S = B, L = A
    const long Ask = e.a_kc ? 1 : e.lda, Asm = e.a_kc ? e.lda : 1;
    const long Bsk = e.b_kc ? 1 : e.ldb, Bsn = e.b_kc ? e.ldb : 1;
    a.S = static_cast<const double*>(tr ? e.B : e.A);
    a.ss = tr ? Bsn : Asm;
    a.sk = tr ? Bsk : Ask;
    a.L = static_cast<const double*>(tr ? e.A : e.B);
    a.ls = tr ? Asm : Bsn;
    a.lk = tr ? Ask : Bsk;
    a.ns = tr ? e.N : e.M;
    a.nl = tr ? e.M : e.N;
    a.K = e.K;
    a.nchunks = (a.nl + 15) / 16;
    a.alpha = e.alpha;
    a.beta = e.beta;
    a.splits = 1;
    a.kchunk = e.K;
    a.C = e.C;
    a.cs = tr ? 1 : e.ldc;
    a.cl = tr ? e.ldc : 1;
    a.swap = tr ? 1 : 0;
    kc_all = kc_all && a.sk == 1 && a.lk == 1;
    if (longk) {
      // about 1024 waves per group, whole k steps per split; the groups after
      // this one keep one output's worth of the workspace each
      const long per = (long)e.M * e.N;
      long later = 0;
      for (int h = g + 1; h < ng; ++h) later += (long)d[h].M * d[h].N;
      long sp = std::max<long>(1, 1024 / a.nchunks);
      sp = std::min<long>(sp, (e.K + 31) / 32);
      const long avail = d[0].ws_elems - ws_off - later;
      if (sp * per > avail) sp = std::max<long>(1, avail / per);
      long kc = (e.K + sp - 1) / sp;
      kc = (kc + 31) / 32 * 32;
      sp = (e.K + kc - 1) / kc;
      a.splits = (int)sp;
      a.kchunk = kc;
      maxitems = std::max(maxitems, (long)a.nchunks * sp);
      if (sp == 1) continue;  // the one split writes C itself
      a.C = d[0].ws + ws_off;
      a.zW = per;
      a.cs = tr ? 1 : e.N;
      a.cl = tr ? e.N : 1;
      ws_off += sp * per;
      ReduceGroups::G& q = rg.g[g];
      q.M = e.M;
      q.N = e.N;
      q.splits = (int)sp;
      q.ws = a.C;
      q.zC = per;
      q.C = e.C;
      q.ldc = e.ldc;
      q.beta = e.beta;
      most = std::max(most, per);
      minsplit = std::min(minsplit, (int)sp);
    } else {
      maxitems = std::max(maxitems, (long)a.nchunks);
    }
  }
  for (int g = ng; g < kMaxGroups; ++g) sg.g[g] = sg.g[0];
  if (longk) {
    const dim3 grid((unsigned)((maxitems + 3) / 4), (unsigned)ng);
#define SKL(T)                                                                          \
  if (kc_all)                                                                            \
    hipLaunchKernelGGL((skinny_f64_longk_kernel<T, true>), grid, dim3(256), 0, stream, sg); \
  else                                                                                   \
    hipLaunchKernelGGL((skinny_f64_longk_kernel<T, false>), grid, dim3(256), 0, stream, sg)
    if (tm == 1) {
      SKL(1);
    } else if (tm == 2) {
      SKL(2);
    } else if (tm == 3) {
      SKL(3);
    } else {
      SKL(4);
    }
#undef SKL
    ACES_HIP(hipGetLastError());
    for (int g = ng; g < kMaxGroups; ++g) rg.g[g].splits = 0;
    if (most > 0) {
      const int wide = (minsplit >= kWideSplits && most <= kWideMaxOutputs) ? 1 : 0;
      long gx = wide ? (most + 3) / 4 : (most + 255) / 256;
      if (!wide && gx > 65536) gx = 65536;
      hipLaunchKernelGGL(splitk_reduce_groups_kernel, dim3((unsigned)gx, (unsigned)kMaxGroups), dim3(256), 0,
                         stream, rg, wide);
    }
  } else {
    // enough waves to fill the chip a few times over, each walking chunks
    const long blocks = std::min<long>((maxitems + 3) / 4, std::max<long>(1, 4096 / ng));
    const dim3 grid((unsigned)blocks, (unsigned)ng);
#define SKS(T) hipLaunchKernelGGL(skinny_f64_shortk_kernel<T>, grid, dim3(256), 0, stream, sg)
    if (tm == 1) SKS(1); else if (tm == 2) SKS(2); else if (tm == 3) SKS(3); else SKS(4);
#undef SKS
  }
  ACES_HIP(hipGetLastError());
  *ran = 1;
  return ACES_OK;
}

static int gemm_one(const GemmDesc& d, GemmPrec prec, hipStream_t stream);

// Compute units of the current device (one value per process: one GPU).
static int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      cus = n;
    else
      cus = 256;
  }
  return cus;
}

// Wave quantization: an output of nb equal tiles runs in ceil(nb / CUs) rounds
// of one tile time each, so a last round that is under half full leaves most
// of the chip idle for a whole (long-K) tile.  Such a product runs as two:
// the rows that fill the full rounds, then the remaining rows, whose few
// tiles split K over the workspace (deterministic split-K) and finish in a
// fraction of a tile time.
int gemm(const GemmDesc& d, GemmPrec prec, hipStream_t stream) {
  ACES_CHECK(check_desc(d, prec));
  if (d.M == 0 || d.N == 0) return ACES_OK;
  if (prec == GP_F64) {
    int ran = 0;
    ACES_CHECK(gemm_skinny(&d, 1, stream, &ran));
    if (ran) return ACES_OK;
  }
  if (d.batch == 1 && d.ws != nullptr && d.K >= 1024) {
    const int wm = gemm_shape(&d, 1, prec);
    GemmArgs probe;
    prepare(d, prec, probe, nullptr, 0, 0, 1L << 30, wm);
    const long nb = (long)probe.tiles_m * probe.tiles_n;
    const long cus = device_cus();
    const long full = nb / cus, tail = nb - full * cus;
    if (full >= 1 && full <= 32 && tail > 0 && 2 * tail <= cus) {
      const long main_tm = full * cus / probe.tiles_n;
      const long rows = main_tm * block_rows(prec, wm);
      if (main_tm >= 1 && rows < d.M) {
        const long esz = (prec == GP_BF16X3) ? 2 : (prec == GP_F32) ? 4 : 8;
        const long aoff = (d.a_kc ? rows * d.lda : rows) * esz;
        GemmDesc top = d, bot = d;
        top.M = (int)rows;
        bot.M = d.M - (int)rows;
        bot.A = (const char*)d.A + aoff;
        if (d.A_lo) bot.A_lo = (const char*)d.A_lo + aoff;
        bot.C = d.C + rows * d.ldc;
        if (d.C_hi) {
          bot.C_hi = d.C_hi + rows * d.ldc_lp;
          bot.C_lo = d.C_lo + rows * d.ldc_lp;
        }
        // worth it only where the tail's split-K fits in much less than a round
        const int wb = gemm_shape(&bot, 1, prec);
        GemmArgs pb;
        prepare(bot, prec, pb, nullptr, 0, 0, 1L << 30, wb);
        const long nbot = (long)pb.tiles_m * pb.tiles_n;
        const int sp = prepare(bot, prec, pb, d.ws, d.ws_elems, 0, nbot, wb);
        const long ntop = main_tm * probe.tiles_n;
        // rounds in units of one full-K tile of the original block size
        const double unit = (double)block_rows(prec, wm) * block_cols(prec, wm) * d.K;
        const double bot_round = (double)block_rows(prec, wb) * block_cols(prec, wb) * pb.kchunk / unit;
        const double rounds = (double)((ntop + cus - 1) / cus) + (double)((nbot * sp + cus - 1) / cus) * bot_round;
        if (rounds <= (double)full + 0.75) {
          ACES_CHECK(gemm_one(top, prec, stream));
          return gemm_one(bot, prec, stream);
        }
      }
    }
  }
  return gemm_one(d, prec, stream);
}

static int gemm_one(const GemmDesc& d, GemmPrec prec, hipStream_t stream) {
  GemmGroups gg;
  gg.n = 1;
  GemmArgs& a = gg.g[0];
  // tiles for the split decision use this precision's block size
  const int wm = gemm_shape(&d, 1, prec);
  GemmArgs probe;
  prepare(d, prec, probe, nullptr, 0, 0, 1L << 30, wm);
  const long nb = (long)probe.tiles_m * probe.tiles_n;
  ACES_REQUIRE(nb < (1L << 31) && d.batch < 65536, "gemm: grid too large");
  const int splits = prepare(d, prec, a, d.ws, d.ws_elems, 0, nb * d.batch, wm);
  // split output straight from the direct-to-LDS bf16x3 epilogue; any other
  // path writes C and splits it afterwards
  const bool dl = prec == GP_BF16X3 && bf16_direct(&d, 1);
  if (d.C_hi && dl && splits == 1) {
    a.Chi = d.C_hi;
    a.Clo = d.C_lo;
    a.ldcl = d.ldc_lp;
  }
  for (int k = 1; k < kMaxGroups; ++k) gg.g[k] = a;
  dim3 grid((unsigned)nb, (unsigned)d.batch, (unsigned)splits);
  if (d.K == 0) a.K = 0;
  ACES_CHECK(launch_groups(gg, prec, d.a_kc != 0, d.b_kc != 0, d.K == 0, grid, wm, dl, stream));
  ACES_CHECK(reduce_splits(d, a, stream));
  if (d.C_hi && a.Chi == nullptr)
    ACES_CHECK(split_bf16(d.M, d.N, d.C, d.ldc, d.C_hi, d.C_lo, d.ldc_lp, stream));
  return ACES_OK;
}

int gemm_grouped(const GemmDesc* d, int ng, GemmPrec prec, hipStream_t stream) {
  ACES_REQUIRE(ng >= 1 && ng <= kMaxGroups, "gemm_grouped: 1..8 groups");
  if (ng == 1) return gemm(d[0], prec, stream);
  GemmGroups gg;
  gg.n = ng;
  long maxnb = 0, ws_off = 0;
  int maxsplit = 1;
  bool any = false;
  const int wm = gemm_shape(d, ng, prec);
  for (int g = 0; g < ng; ++g) {
    ACES_CHECK(check_desc(d[g], prec));
    ACES_REQUIRE(d[g].batch == 1, "gemm_grouped: batched groups are not supported");
    ACES_REQUIRE(d[g].C_hi == nullptr, "gemm_grouped: split output is not supported");
    ACES_REQUIRE(d[g].a_kc == d[0].a_kc && d[g].b_kc == d[0].b_kc, "gemm_grouped: operand orders must agree");
    ACES_REQUIRE((d[g].K == 0) == (d[0].K == 0) || d[g].M == 0 || d[g].N == 0,
                 "gemm_grouped: mixed K = 0 and K > 0 groups");
    GemmArgs probe;
    prepare(d[g], prec, probe, nullptr, 0, 0, 1L << 30, wm);
    const long nb = (d[g].M == 0 || d[g].N == 0) ? 0 : (long)probe.tiles_m * probe.tiles_n;
    const long avail = d[0].ws ? d[0].ws_elems - ws_off : 0;
    const int sp = prepare(d[g], prec, gg.g[g], d[0].ws, avail, ws_off, nb, wm);
    if (sp > 1) ws_off += (long)sp * d[g].M * d[g].N;
    if (nb == 0) gg.g[g].tiles_m = 0;  // empty block: every workgroup exits
    if (d[g].K == 0) gg.g[g].K = 0;
    maxnb = std::max(maxnb, nb);
    maxsplit = std::max(maxsplit, sp);
    any = any || nb > 0;
  }
  for (int g = ng; g < kMaxGroups; ++g) gg.g[g] = gg.g[0];
  if (!any) return ACES_OK;
  if (prec == GP_F64) {
    int ran = 0;
    ACES_CHECK(gemm_skinny(d, ng, stream, &ran));
    if (ran) return ACES_OK;
  }
  ACES_REQUIRE(maxnb < (1L << 31), "gemm_grouped: grid too large");
  int kz = 0;
  for (int g = 0; g < ng; ++g)
    if (d[g].M > 0 && d[g].N > 0) kz = (d[g].K == 0);
  dim3 grid((unsigned)maxnb, (unsigned)ng, (unsigned)maxsplit);
  ACES_CHECK(launch_groups(gg, prec, d[0].a_kc != 0, d[0].b_kc != 0, kz != 0, grid, wm, bf16_direct(d, ng), stream));
  // one launch for the split-K sums of all groups (same order of additions
  // as the single-problem reduction)
  ReduceGroups rg;
  rg.n = ng;
  long most = 0;
  for (int g = 0; g < kMaxGroups; ++g) {
    ReduceGroups::G& q = rg.g[g];
    q.splits = 0;
    if (g >= ng || d[g].M == 0 || d[g].N == 0 || gg.g[g].splits <= 1) continue;
    q.M = d[g].M;
    q.N = d[g].N;
    q.splits = gg.g[g].splits;
    q.ws = gg.g[g].C;
    q.zC = gg.g[g].zC;
    q.C = d[g].C;
    q.ldc = d[g].ldc;
    q.beta = d[g].beta;
    most = std::max(most, (long)q.M * q.N);
  }
  if (most > 0) {
    int minsplit = 1 << 30;
    for (int g = 0; g < kMaxGroups; ++g)
      if (rg.g[g].splits > 1) minsplit = std::min(minsplit, rg.g[g].splits);
    const int wide = (minsplit >= kWideSplits && most <= kWideMaxOutputs) ? 1 : 0;
    long gx = wide ? (most + 3) / 4 : (most + 255) / 256;
    if (!wide && gx > 65536) gx = 65536;
    hipLaunchKernelGGL(splitk_reduce_groups_kernel, dim3((unsigned)gx, (unsigned)kMaxGroups), dim3(256), 0, stream,
                       rg, wide);
    ACES_HIP(hipGetLastError());
  }
  return ACES_OK;
}

}  // namespace aces
