// CCSD-specific HBM-bound kernels: ladder packing/unpacking, Jacobi update,
// energy, DIIS vector algebra.  All fp64.
//
// Pair orderings follow the reference's DPD packing (acescore/aces_com_symloc.F,
// libr/symexp4.f:36, libr/assym.f):
//   'PCK2' (p <= q):  index = q*(q+1)/2 + p     (0-based)
//   'PACK' (p <  q):  index = q*(q-1)/2 + p
#include "common.h"
#include "ccsd_kernels.h"
#include "bf16_util.h"
#include "pairs.h"

namespace aces {

// ---------------------------------------------------------------------------
// Ladder right operand from tau[i][j][c][d]:
//   Sp[(i<=j)][(c<=d)] = (tau_ijcd + tau_jicd)/2,  Dp[(i<j)][(c<d)] = (tau_ijcd - tau_jicd)/2
// One block per (pair i<=j, 64 values of d).  The (c, d) pair index runs over
// c inside d, so a row segment of Sp/Dp is a column strip of the tau_ij and
// tau_ji blocks: strips of 32 c x 64 d are read along d (coalesced) into LDS
// and written along c.  Outputs go to fp64 rows and/or bf16 hi/lo rows.
constexpr int kPtC = 32, kPtD = 64;

__global__ __launch_bounds__(256) void pack_tau_kernel(int o, int v, const double* __restrict__ tau, PackTauOut out) {
  __shared__ double xs[kPtC][kPtD + 1], ys[kPtC][kPtD + 1];
  const int ndt = (v + kPtD - 1) / kPtD;
  const long p = blockIdx.x / ndt;
  const int d0 = (int)(blockIdx.x % ndt) * kPtD;
  int i, j;
  unpack_le(p, i, j);
  const long vv = (long)v * v;
  const double* X = tau + ((long)i * o + j) * vv;
  const double* Y = tau + ((long)j * o + i) * vv;
  const long pm = (i < j) ? pk_lt(i, j) : -1;
  const int t = threadIdx.x;
  const int dmax = (d0 + kPtD < v) ? d0 + kPtD : v;  // d in [d0, dmax)
  for (int c0 = 0; c0 < dmax; c0 += kPtC) {
    {
      const int dl = t & 63, cl = t >> 6;
#pragma unroll
      for (int k = 0; k < kPtC / 4; ++k) {
        const int c = c0 + cl + 4 * k, d = d0 + dl;
        const bool in = c < v && d < v;
        xs[cl + 4 * k][dl] = in ? X[(long)c * v + d] : 0.0;
        ys[cl + 4 * k][dl] = in ? Y[(long)c * v + d] : 0.0;
      }
    }
    __syncthreads();
    {
      const int cl = t & 31, dl0 = t >> 5;
      const int c = c0 + cl;
#pragma unroll
      for (int k = 0; k < kPtD / 8; ++k) {
        const int dl = dl0 + 8 * k, d = d0 + dl;
        if (d >= dmax || c > d) continue;
        const double x = xs[cl][dl], y = ys[cl][dl];
        const double sv = 0.5 * (x + y);
        const long ps = (long)d * (d + 1) / 2 + c;
        if (out.S) out.S[p * out.lds + ps] = sv;
        if (out.Shi) {
          const unsigned short h = bf16_rn((float)sv);
          out.Shi[p * out.ldsl + ps] = h;
          out.Slo[p * out.ldsl + ps] = bf16_rn((float)(sv - (double)bf16_to_f(h)));
        }
        if (pm >= 0 && c < d) {
          const double dv = 0.5 * (x - y);
          const long pd = (long)d * (d - 1) / 2 + c;
          if (out.D) out.D[pm * out.ldd + pd] = dv;
          if (out.Dhi) {
            const unsigned short h = bf16_rn((float)dv);
            out.Dhi[pm * out.lddl + pd] = h;
            out.Dlo[pm * out.lddl + pd] = bf16_rn((float)(dv - (double)bf16_to_f(h)));
          }
        }
      }
    }
    __syncthreads();
  }
}

int pack_tau(int o, int v, const double* tau, const PackTauOut& out, hipStream_t s) {
  const long nop = (long)o * (o + 1) / 2;
  const long nblk = nop * ((v + kPtD - 1) / kPtD);
  if (nblk == 0) return ACES_OK;
  ACES_REQUIRE(nblk < (1L << 31), "pack_tau: too many tiles");
  ACES_REQUIRE((out.Shi == nullptr) == (out.Slo == nullptr) && (out.Dhi == nullptr) == (out.Dlo == nullptr),
               "pack_tau: bf16 output needs hi and lo");
  hipLaunchKernelGGL(pack_tau_kernel, dim3((unsigned)nblk), dim3(256), 0, s, o, v, tau, out);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// Ladder left operand rows.
//   Vp[(a<=b)][(c<=d)] = <ab|cd> + <ab|dc>  (c<d),  <ab|cc> (c==d)
//   Vm[(a<b)][(c<d)]   = <ab|cd> - <ab|dc>
// from the reference's compressed RHF list 233 (rows A<=B 'PCK2', columns Cd
// with C fastest, column major): element (ab, cd) at L[(c + v*d)*nvp + pk_le(a,b)],
// value <ab|cd> = (ac|bd)  (intprc/intproc.F:250-257).
__global__ void build_vpm_vvvv_kernel(int v, const double* __restrict__ L233, double* __restrict__ Vp, long ldp,
                                      double* __restrict__ Vm, long ldm) {
  const long nvp = (long)v * (v + 1) / 2, nvm = (long)v * (v - 1) / 2;
  const long totP = nvp * nvp, totM = nvm * nvm;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < totP + totM;
       idx += (long)gridDim.x * blockDim.x) {
    int a, b, c, d;
    if (idx < totP) {
      const long rab = idx / nvp, pcd = idx - rab * nvp;
      unpack_le(pcd, c, d);
      const double x = L233[((long)c + (long)v * d) * nvp + rab];
      const double y = (c == d) ? 0.0 : L233[((long)d + (long)v * c) * nvp + rab];
      Vp[rab * ldp + pcd] = x + y;
    } else {
      const long r = idx - totP;
      const long rab = r / nvm, pcd = r - rab * nvm;
      unpack_lt(rab, a, b);
      unpack_lt(pcd, c, d);
      const long row = pk_le(a, b);
      const double x = L233[((long)c + (long)v * d) * nvp + row];
      const double y = L233[((long)d + (long)v * c) * nvp + row];
      Vm[rab * ldm + pcd] = x - y;
    }
  }
}

// G rows (k,b) appended to the ladder operand (T1 part of the dressed <ab||cd>):
//   Gp[(k,b)][(c<=d)] = (kc|bd) + (kd|bc) (c<d), (kc|bc) (c==d);  Gm = (kc|bd) - (kd|bc)
// from ovvv[k][c][b][d] = (kc|bd).
__global__ void build_g_ovvv_kernel(int o, int v, const double* __restrict__ ovvv, double* __restrict__ Gp,
                                    long ldp, double* __restrict__ Gm, long ldm) {
  const long nvp = (long)v * (v + 1) / 2, nvm = (long)v * (v - 1) / 2;
  const long rows = (long)o * v;
  const long totP = rows * nvp, totM = rows * nvm;
  const long v2 = (long)v * v, v3 = v2 * v;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < totP + totM;
       idx += (long)gridDim.x * blockDim.x) {
    int c, d;
    long row, pcd;
    if (idx < totP) {
      row = idx / nvp;
      pcd = idx - row * nvp;
      unpack_le(pcd, c, d);
    } else {
      const long r = idx - totP;
      row = r / nvm;
      pcd = r - row * nvm;
      unpack_lt(pcd, c, d);
    }
    const long k = row / v, b = row - (row / v) * v;
    const double x = ovvv[k * v3 + (long)c * v2 + b * v + d];  // (kc|bd)
    const double y = ovvv[k * v3 + (long)d * v2 + b * v + c];  // (kd|bc)
    if (idx < totP)
      Gp[row * ldp + pcd] = (c == d) ? x : x + y;
    else
      Gm[row * ldm + pcd] = x - y;
  }
}

int build_ladder_operands(int o, int v, const double* L233, const double* ovvv, double* Vp, long ldp, double* Vm,
                          long ldm, hipStream_t s) {
  const long nvp = (long)v * (v + 1) / 2, nvm = (long)v * (v - 1) / 2;
  long n = nvp * nvp + nvm * nvm;
  hipLaunchKernelGGL(build_vpm_vvvv_kernel, dim3(grid1(n)), dim3(256), 0, s, v, L233, Vp, ldp, Vm, ldm);
  ACES_HIP(hipGetLastError());
  n = (long)o * v * (nvp + nvm);
  hipLaunchKernelGGL(build_g_ovvv_kernel, dim3(grid1(n)), dim3(256), 0, s, o, v, ovvv, Vp + nvp * ldp, ldp,
                     Vm + nvm * ldm, ldm);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// Ladder result unpack.  Zp rows: [(a<=b) | (k,b)] x cols (i<=j); Zm rows [(a<b) | (k,b)] x (i<j).
//   inc2[i][j][a][b] += Zp[ab][ij] + s Zm[ab][ij],  s = sign(b-a) * sign(j-i)
//   Yb[b][i][j][k]   = Zp[kb][ij] + sign(j-i) Zm[kb][ij]
//
// inc2 part: one block per (32 consecutive (i<=j) columns, a, 64 values of b).
// The Z rows of the pairs (a, b) are read along the pair columns (coalesced),
// turned round in LDS, and added into inc2[i][j][a][b0..] and its mirror
// inc2[j][i][a][b0..] along b.
constexpr int kUpC = 32, kUpB = 64;

__global__ __launch_bounds__(256) void unpack_ladder_t2_kernel(int o, int v, const double* __restrict__ Zp,
                                                               long ldzp, const double* __restrict__ Zm, long ldzm,
                                                               double* __restrict__ inc2) {
  __shared__ double zp[kUpB][kUpC + 1];
  __shared__ double zm[kUpB][kUpC + 1];
  const long nop = (long)o * (o + 1) / 2;
  const int nbt = (v + kUpB - 1) / kUpB;
  long blk = blockIdx.x;
  const int bt = (int)(blk % nbt);
  blk /= nbt;
  const int a = (int)(blk % v);
  const long c0 = (blk / v) * kUpC;
  const int b0 = bt * kUpB;
  const int t = threadIdx.x;
  // read: column c0 + cj, rows b0 + rb + 8k
  {
    const int cj = t & 31, rb = t >> 5;
    const long p = c0 + cj;
    int i = 0, j = 0;
    if (p < nop) unpack_le(p, i, j);
    // branch-free loads (outside the ranges: element 0, dropped), all in
    // flight together
    // (Zm has elements only for o > 1 and v > 1)
    const bool hm = o > 1 && v > 1;
    double xp[kUpB / 8], xm[kUpB / 8];
#pragma unroll
    for (int k = 0; k < kUpB / 8; ++k) {
      const int b = b0 + rb + 8 * k;
      const bool okp = p < nop && b < v, okm = okp && i < j && a != b;
      const int lo = min(a, b), hi = max(a, b);
      xp[k] = Zp[okp ? pk_le(lo, hi) * ldzp + p : 0];
      xm[k] = hm ? Zm[okm ? pk_lt(lo, hi) * ldzm + pk_lt(i, j) : 0] : 0.0;
      if (!okp) xp[k] = 0.0;
      if (!okm) xm[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < kUpB / 8; ++k) {
      zp[rb + 8 * k][cj] = xp[k];
      zm[rb + 8 * k][cj] = xm[k];
    }
  }
  __syncthreads();
  // write: lanes along b, waves over the pair columns
  const int lb = t & 63, w = t >> 6;
  const int b = b0 + lb;
  if (b >= v) return;
  const double sab = (b > a) ? 1.0 : (b < a ? -1.0 : 0.0);
  const long vv = (long)v * v;
  for (int cj = w; cj < kUpC; cj += 4) {
    const long p = c0 + cj;
    if (p >= nop) break;
    int i, j;
    unpack_le(p, i, j);
    const double xp = zp[lb][cj], xm = zm[lb][cj];
    inc2[((long)i * o + j) * vv + (long)a * v + b] += xp + sab * xm;
    if (i < j) inc2[((long)j * o + i) * vv + (long)a * v + b] += xp - sab * xm;
  }
}

//   Yb[b][i][j][k] = Zp[kb][ij] + sign(j-i) Zm[kb][ij]   (v o^3 elements)
__global__ void unpack_ladder_yb_kernel(int o, int v, const double* __restrict__ Zp, long ldzp,
                                        const double* __restrict__ Zm, long ldzm, double* __restrict__ Yb) {
  const long nvp = (long)v * (v + 1) / 2, nvm = (long)v * (v - 1) / 2;
  const long tot = (long)v * o * o * o;
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < tot; r += (long)gridDim.x * blockDim.x) {
    // r = ((b*o + i)*o + j)*o + k
    const int k = (int)(r % o);
    long q = r / o;
    const int j = (int)(q % o);
    q /= o;
    const int i = (int)(q % o);
    const int b = (int)(q / o);
    const long row = (long)k * v + b;
    const int i0 = min(i, j), j0 = max(i, j);
    double y = Zp[(nvp + row) * ldzp + pk_le(i0, j0)];
    if (i != j) y += ((i < j) ? 1.0 : -1.0) * Zm[(nvm + row) * ldzm + pk_lt(i0, j0)];
    Yb[r] = y;
  }
}

int unpack_ladder(int o, int v, const double* Zp, long ldzp, const double* Zm, long ldzm, double* inc2, double* Yb,
                  hipStream_t s) {
  const long nop = (long)o * (o + 1) / 2;
  const long nblk = (nop + kUpC - 1) / kUpC * v * ((v + kUpB - 1) / kUpB);
  ACES_REQUIRE(nblk < (1L << 31), "unpack_ladder: too many tiles");
  if (nblk > 0)
    hipLaunchKernelGGL(unpack_ladder_t2_kernel, dim3((unsigned)nblk), dim3(256), 0, s, o, v, Zp, ldzp, Zm, ldzm,
                       inc2);
  ACES_HIP(hipGetLastError());
  const long n = (long)v * o * o * o;
  hipLaunchKernelGGL(unpack_ladder_yb_kernel, dim3(grid1(n)), dim3(256), 0, s, o, v, Zp, ldzp, Zm, ldzm, Yb);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// Ring operands from t2[i][j][a][b] and t1[i][a], rows (i,a), columns (j,b):
//   U = 2 t_ijab - t_ijba,  X = t_ijba,  Ut = U - 2 t1_ja t1_ib,  Xt = X + 2 t1_ja t1_ib
// One block per (i, j, 32 values of a, 64 values of b): t_ijab is read along b,
// t_ijba along a and turned round through LDS.  Each output goes to fp64
// (ld = ov) and/or to bf16 hi/lo rows (ld = ldl) where the pointer is set;
// tau_ijab = t_ijab + t1_ia t1_jb and the bf16 split of t_ijab keep t2's order;
// the bf16 splits of tau go wherever their strides put them.

__global__ __launch_bounds__(256) void ring_operands_kernel(int o, int v, const double* __restrict__ t1,
                                                            const double* __restrict__ t2, RingOperands r) {
  __shared__ double tr[64][33];
  const int nbt = (v + 63) / 64, nat = (v + 31) / 32;
  long blk = blockIdx.x;
  const int bt = (int)(blk % nbt);
  blk /= nbt;
  const int at = (int)(blk % nat);
  blk /= nat;
  const int j = (int)(blk % o);
  const int i = (int)(blk / o);
  const int a0 = at * 32, b0 = bt * 64;
  const long vv = (long)v * v, ov = (long)o * v;
  const double* T = t2 + ((long)i * o + j) * vv;  // T[a][b] = t_ijab
  const int tid = threadIdx.x;
  // Loads are branch-free (an element outside the block reads T[a0][b0],
  // which exists, and is dropped), so that a thread's loads are in flight
  // together.
  {
    // t_ijba for a in [a0, a0+32), b in [b0, b0+64): read along a (T[b][a])
    const int ca = tid & 31, rb = tid >> 5;
    double w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int b = b0 + rb + 8 * k, a = a0 + ca;
      const bool ok = a < v && b < v;
      w[k] = T[ok ? (long)b * v + a : (long)b0 * v + a0];
      if (!ok) w[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) tr[rb + 8 * k][ca] = w[k];
  }
  __syncthreads();
  const int cb = tid & 63, ra = tid >> 6;
  const int b = b0 + cb;
  if (b >= v) return;
  const double tib = t1[(long)i * v + b];
  const double tjb = t1[(long)j * v + b];
  const long trow = ((long)i * o + j) * v;  // row of t2 element (i, j, a = 0)
  double dv[8], tia[8], tja[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int a = min(a0 + ra + 4 * k, v - 1);
    dv[k] = T[(long)a * v + b];
    tia[k] = t1[(long)i * v + a];
    tja[k] = t1[(long)j * v + a];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int a = a0 + ra + 4 * k;
    if (a >= v) break;
    const double d = dv[k];
    const double tv = d + tia[k] * tjb;
    if (r.tau) r.tau[(trow + a) * v + b] = tv;
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if (r.tauhi[q]) {
        const long off = i * r.tsi[q] + j * r.tsj[q] + a * r.tsa[q] + b;
        const unsigned short h = bf16_rn((float)tv);
        r.tauhi[q][off] = h;
        r.taulo[q][off] = bf16_rn((float)(tv - (double)bf16_to_f(h)));
      }
    if (r.t2hi) {
      const unsigned short h = bf16_rn((float)d);
      r.t2hi[(trow + a) * r.ldt + b] = h;
      r.t2lo[(trow + a) * r.ldt + b] = bf16_rn((float)(d - (double)bf16_to_f(h)));
    }
    const double x = tr[cb][ra + 4 * k];
    const double u = 2.0 * d - x;
    const double p = 2.0 * tja[k] * tib;
    const double val[4] = {u, x, u - p, x + p};
    const long row = (long)i * v + a, col = (long)j * v + b;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (r.f64[q]) r.f64[q][row * ov + col] = val[q];
      if (r.hi[q]) {
        const unsigned short h = bf16_rn((float)val[q]);
        r.hi[q][row * r.ldl + col] = h;
        r.lo[q][row * r.ldl + col] = bf16_rn((float)(val[q] - (double)bf16_to_f(h)));
      }
    }
  }
}

int ring_operands(int o, int v, const double* t1, const double* t2, const RingOperands& r, hipStream_t s) {
  const long nblk = (long)o * o * ((v + 31) / 32) * ((v + 63) / 64);
  if (nblk == 0) return ACES_OK;
  ACES_REQUIRE(nblk < (1L << 31), "ring_operands: too many tiles");
  for (int q = 0; q < 4; ++q)
    ACES_REQUIRE((r.hi[q] == nullptr) == (r.lo[q] == nullptr) && (r.hi[q] == nullptr || r.ldl >= (long)o * v),
                 "ring_operands: bad bf16 output");
  ACES_REQUIRE((r.t2hi == nullptr) == (r.t2lo == nullptr) && (r.t2hi == nullptr || r.ldt >= v),
               "ring_operands: bad t2 split output");
  for (int q = 0; q < 3; ++q)
    ACES_REQUIRE((r.tauhi[q] == nullptr) == (r.taulo[q] == nullptr), "ring_operands: bad tau split output");
  hipLaunchKernelGGL(ring_operands_kernel, dim3((unsigned)nblk), dim3(256), 0, s, o, v, t1, t2, r);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// Jacobi step (NEWT2 / E4S): t1n = inc1 / (e_i - e_a), t2n = inc2 / (e_i + e_j - e_a - e_b).
__global__ void jacobi_kernel(int o, int v, const double* __restrict__ eo, const double* __restrict__ ev,
                              const double* __restrict__ inc1, const double* __restrict__ inc2,
                              double* __restrict__ t1n, double* __restrict__ t2n) {
  const long n1 = (long)o * v, n2 = n1 * n1;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n1 + n2;
       idx += (long)gridDim.x * blockDim.x) {
    if (idx < n1) {
      const int i = (int)(idx / v), a = (int)(idx % v);
      t1n[idx] = inc1[idx] / (eo[i] - ev[a]);
    } else {
      const long r = idx - n1;
      const int b = (int)(r % v);
      const int a = (int)((r / v) % v);
      const int j = (int)((r / ((long)v * v)) % o);
      const int i = (int)(r / ((long)v * v * o));
      t2n[r] = inc2[r] / (eo[i] + eo[j] - ev[a] - ev[b]);
    }
  }
}

int jacobi(int o, int v, const double* eo, const double* ev, const double* inc1, const double* inc2, double* t1n,
           double* t2n, hipStream_t s) {
  const long n = (long)o * v * (1 + (long)o * v);
  hipLaunchKernelGGL(jacobi_kernel, dim3(grid1(n)), dim3(256), 0, s, o, v, eo, ev, inc1, inc2, t1n, t2n);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// In-place SYMETRIZE_T2 on t2[i][j][a][b]: t <- (t_ijab + t_jiba)/2.
// One block per (ij, a-tile, b-tile) with i < j, or i == j and a-tile <= b-tile:
// it owns the 64x64 tile of block ij and its transposed partner in block ji,
// reads both coalesced and writes both back through an LDS transpose.
constexpr int kSymT = 64;

__global__ __launch_bounds__(256) void symmetrize_kernel(int o, int v, double* __restrict__ t2) {
  __shared__ double tp[kSymT][kSymT + 1];
  const int nt = (v + kSymT - 1) / kSymT;
  long b = blockIdx.x;
  const int tb = (int)(b % nt);
  b /= nt;
  const int ta = (int)(b % nt);
  b /= nt;
  const int j = (int)(b % o);
  const int i = (int)(b / o);
  if (i > j || (i == j && ta > tb)) return;  // whole block: the partner block owns it
  const long vv = (long)v * v;
  double* A = t2 + ((long)i * o + j) * vv;  // A[a][b]
  double* B = t2 + ((long)j * o + i) * vv;  // B[b][a]
  const int a0 = ta * kSymT, b0 = tb * kSymT;
  const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
  // B tile: rows b0.., columns a0.. (coalesced along a) into tp[b][a]
  double areg[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int bb = b0 + r + 4 * k, aa = a0 + c;
    tp[r + 4 * k][c] = (bb < v && aa < v) ? B[(long)bb * v + aa] : 0.0;
  }
  // A tile: rows a0.., columns b0.. (coalesced along b) into registers
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int aa = a0 + r + 4 * k, bb = b0 + c;
    areg[k] = (aa < v && bb < v) ? A[(long)aa * v + bb] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) areg[k] = 0.5 * (areg[k] + tp[c][r + 4 * k]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int aa = a0 + r + 4 * k, bb = b0 + c;
    tp[c][r + 4 * k] = areg[k];
    if (aa < v && bb < v) A[(long)aa * v + bb] = areg[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int bb = b0 + r + 4 * k, aa = a0 + c;
    if (bb < v && aa < v) B[(long)bb * v + aa] = tp[r + 4 * k][c];
  }
}

int symmetrize_t2(int o, int v, double* t2, hipStream_t s) {
  const long nt = (v + kSymT - 1) / kSymT;
  const long nblk = (long)o * o * nt * nt;
  if (nblk == 0) return ACES_OK;
  ACES_REQUIRE(nblk < (1L << 31), "symmetrize_t2: too many tiles");
  hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)nblk), dim3(256), 0, s, o, v, t2);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// Block reductions (fixed grid; partials then a single-block finish).
static constexpr int kRedBlocks = 1024;

__device__ inline double block_sum(double x, double* sh) {
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__global__ void finish_sum_kernel(const double* __restrict__ part, int n, double* __restrict__ out);

// ---------------------------------------------------------------------------
// The end of a CC cycle in one pass over T2: the Jacobi update
// t2n = inc2 / (e_i + e_j - e_a - e_b), the energy of the update
// (E = 2 sum f_ia t1n_ia + sum L_ijab (t2n_ijab + t1n_ia t1n_jb), taken before
// symmetrization as the reference does) and SYMETRIZE_T2 of t2n.  Blocks
// pair the tile (ij, a-tile, b-tile) with its transposed partner in block ji,
// as symmetrize_kernel does; each block leaves its energy partial in
// part[blockIdx.x] (0 for blocks the partner owns).  t1n comes from
// jacobi_t1_energy_kernel, launched first.
__global__ __launch_bounds__(256) void jacobi_t2_energy_sym_kernel(int o, int v, const double* __restrict__ eo,
                                                                   const double* __restrict__ ev,
                                                                   const double* __restrict__ t1n,
                                                                   const double* __restrict__ inc2,
                                                                   const double* __restrict__ L,
                                                                   double* __restrict__ t2n,
                                                                   double* __restrict__ part) {
  __shared__ double tp[kSymT][kSymT + 1];
  __shared__ double sh[256];
  const int nt = (v + kSymT - 1) / kSymT;
  long blk = blockIdx.x;
  const int tb = (int)(blk % nt);
  blk /= nt;
  const int ta = (int)(blk % nt);
  blk /= nt;
  const int j = (int)(blk % o);
  const int i = (int)(blk / o);
  if (i > j || (i == j && ta > tb)) {  // whole block: the partner block owns it
    if (threadIdx.x == 0) part[blockIdx.x] = 0.0;
    return;
  }
  const bool self = (i == j && ta == tb);  // the tile is its own partner
  const long vv = (long)v * v;
  const long oA = ((long)i * o + j) * vv, oB = ((long)j * o + i) * vv;  // A[a][b], B[b][a]
  const int a0 = ta * kSymT, b0 = tb * kSymT;
  const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
  double e = 0.0;
  // B tile (block ji, rows b, columns a): update, energy, into tp[b][a]
#pragma unroll 4
  for (int k = 0; k < 16; ++k) {
    const int bb = b0 + r + 4 * k, aa = a0 + c;
    double t = 0.0;
    if (bb < v && aa < v) {
      const long q = oB + (long)bb * v + aa;
      t = inc2[q] / (eo[j] + eo[i] - ev[bb] - ev[aa]);
      if (!self) e += L[q] * (t + t1n[(long)j * v + bb] * t1n[(long)i * v + aa]);
    }
    tp[r + 4 * k][c] = t;
  }
  // A tile (block ij, rows a, columns b) in registers
  double areg[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int aa = a0 + r + 4 * k, bb = b0 + c;
    double t = 0.0;
    if (aa < v && bb < v) {
      const long q = oA + (long)aa * v + bb;
      t = inc2[q] / (eo[i] + eo[j] - ev[aa] - ev[bb]);
      e += L[q] * (t + t1n[(long)i * v + aa] * t1n[(long)j * v + bb]);
    }
    areg[k] = t;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) areg[k] = 0.5 * (areg[k] + tp[c][r + 4 * k]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int aa = a0 + r + 4 * k, bb = b0 + c;
    tp[c][r + 4 * k] = areg[k];
    if (aa < v && bb < v) t2n[oA + (long)aa * v + bb] = areg[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int bb = b0 + r + 4 * k, aa = a0 + c;
    if (bb < v && aa < v) t2n[oB + (long)bb * v + aa] = tp[r + 4 * k][c];
  }
  const double sum = block_sum(e, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = sum;
}

// t1n = inc1 / (e_i - e_a) and 2 sum f_ia t1n_ia (one block, deterministic).
__global__ __launch_bounds__(256) void jacobi_t1_energy_kernel(int o, int v, const double* __restrict__ eo,
                                                               const double* __restrict__ ev,
                                                               const double* __restrict__ inc1,
                                                               const double* __restrict__ fov, long ldf,
                                                               double* __restrict__ t1n, double* __restrict__ out) {
  __shared__ double sh[256];
  const long n1 = (long)o * v;
  double e = 0.0;
  for (long idx = threadIdx.x; idx < n1; idx += blockDim.x) {
    const int i = (int)(idx / v), a = (int)(idx % v);
    const double t = inc1[idx] / (eo[i] - ev[a]);
    t1n[idx] = t;
    e += 2.0 * fov[(long)i * ldf + a] * t;
  }
  const double sum = block_sum(e, sh);
  if (threadIdx.x == 0) *out = sum;
}

long jacobi_energy_scratch(int o, int v) {
  const long nt = (v + kSymT - 1) / kSymT;
  return (long)o * o * nt * nt + 1;
}

int jacobi_energy_sym(int o, int v, const double* eo, const double* ev, const double* inc1, const double* inc2,
                      const double* fov, long ldf, const double* Lijab, double* t1n, double* t2n, double* part,
                      double* out, hipStream_t s) {
  const long nt = (v + kSymT - 1) / kSymT;
  const long nblk = (long)o * o * nt * nt;
  ACES_REQUIRE(nblk < (1L << 31), "jacobi_energy_sym: too many tiles");
  hipLaunchKernelGGL(jacobi_t1_energy_kernel, dim3(1), dim3(256), 0, s, o, v, eo, ev, inc1, fov, ldf, t1n,
                     part + nblk);
  if (nblk > 0)
    hipLaunchKernelGGL(jacobi_t2_energy_sym_kernel, dim3((unsigned)nblk), dim3(256), 0, s, o, v, eo, ev, t1n,
                       inc2, Lijab, t2n, part);
  hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(256), 0, s, part, (int)(nblk + 1), out);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// Energy: E = 2 sum f_ia t_ia + sum_ijab L[i][j][a][b] (t2_ijab + t1_ia t1_jb).
__global__ void energy_partial_kernel(int o, int v, const double* __restrict__ fov, long ldf,
                                      const double* __restrict__ Lijab, const double* __restrict__ t1,
                                      const double* __restrict__ t2, double* __restrict__ part) {
  __shared__ double sh[256];
  const long n1 = (long)o * v, n2 = n1 * n1, vv = (long)v * v;
  double acc = 0.0;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n1 + n2;
       idx += (long)gridDim.x * blockDim.x) {
    if (idx < n1) {
      const int i = (int)(idx / v), a = (int)(idx % v);
      acc += 2.0 * fov[(long)i * ldf + a] * t1[idx];
    } else {
      const long r = idx - n1;
      const long ij = r / vv, ab = r - ij * vv;
      const int i = (int)(ij / o), j = (int)(ij % o), a = (int)(ab / v), b = (int)(ab % v);
      acc += Lijab[r] * (t2[r] + t1[(long)i * v + a] * t1[(long)j * v + b]);
    }
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void finish_sum_kernel(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) *out = s;
}

int energy(int o, int v, const double* fov, long ldf, const double* Lijab, const double* t1, const double* t2,
           double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL(energy_partial_kernel, dim3(kRedBlocks), dim3(256), 0, s, o, v, fov, ldf, Lijab, t1, t2, part);
  hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(256), 0, s, part, kRedBlocks, out);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// DIIS inner products of vectors x = [t1 | t2(ijab)] including the implicit
// same-spin block (t_ijab - t_jiab for i<j, a<b) that DIISLST appends
// (vcc/diislst.F:20-29).  Computes m dots <x_k, y> at once.
__global__ void diis_dot_partial_kernel(int o, int v, int m, const double* const* __restrict__ xs,
                                        const double* __restrict__ y, double* __restrict__ part) {
  __shared__ double sh[256];
  const long n1 = (long)o * v, vv = (long)v * v, n2 = n1 * n1;
  double acc[8];
  for (int k = 0; k < 8; ++k) acc[k] = 0.0;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n1 + n2;
       idx += (long)gridDim.x * blockDim.x) {
    const double yv = y[idx];
    double yaa = 0.0;
    long tw = -1;
    if (idx >= n1) {
      const long r = idx - n1;
      const long ij = r / vv, ab = r - ij * vv;
      const int i = (int)(ij / o), j = (int)(ij % o), a = (int)(ab / v), b = (int)(ab % v);
      if (i < j && a < b) {
        tw = n1 + ((long)j * o + i) * vv + ab;
        yaa = yv - y[tw];
      }
    }
    for (int k = 0; k < m; ++k) {
      const double* x = xs[k];
      double c = x[idx] * yv;
      if (tw >= 0) c += (x[idx] - x[tw]) * yaa;
      acc[k] += c;
    }
  }
  for (int k = 0; k < m; ++k) {
    const double s = block_sum(acc[k], sh);
    if (threadIdx.x == 0) part[k * kRedBlocks + blockIdx.x] = s;
  }
}

__global__ void finish_multi_kernel(const double* __restrict__ part, int m, double* __restrict__ out) {
  __shared__ double sh[256];
  for (int k = 0; k < m; ++k) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += blockDim.x) acc += part[k * kRedBlocks + i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[k] = s;
  }
}

int diis_dots(int o, int v, int m, const double* const* d_xs, const double* y, double* part, double* out,
              hipStream_t s) {
  ACES_REQUIRE(m >= 1 && m <= 8, "diis_dots: 1..8 vectors per call");
  hipLaunchKernelGGL(diis_dot_partial_kernel, dim3(kRedBlocks), dim3(256), 0, s, o, v, m, d_xs, y, part);
  hipLaunchKernelGGL(finish_multi_kernel, dim3(1), dim3(256), 0, s, part, m, out);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// Signed value of the first element of largest magnitude of the DIIS vector in
// the reference's order: [T1 (AI) | T2(Ab,Ij) list order | T2(A<B,I<J) packed]
// (ISAMAX semantics, vcc/dodiis0.F:96-101).  Partial: (|x|, position, value).
struct AbsMax {
  double a;
  long pos;
  double val;
};

__device__ inline void am_update(AbsMax& m, double x, long pos) {
  const double a = fabs(x);
  if (a > m.a || (a == m.a && pos < m.pos)) {
    m.a = a;
    m.pos = pos;
    m.val = x;
  }
}

__global__ void absmax_partial_kernel(int o, int v, const double* __restrict__ e, AbsMax* __restrict__ part) {
  __shared__ AbsMax sh[256];
  const long n1 = (long)o * v, vv = (long)v * v, n2 = n1 * n1;
  const long nom = (long)o * (o - 1) / 2, nvm = (long)v * (v - 1) / 2;
  AbsMax m{-1.0, 0x7fffffffffffffffL, 0.0};
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n1; idx += (long)gridDim.x * blockDim.x)
    am_update(m, e[idx], idx);
  // t2 rows (i,j) over the blocks, (a,b) over the threads: 32-bit index
  // arithmetic per element (the candidates and the tie rule are those of a
  // flat scan, so the result does not depend on the order)
  const int nij = o * o, nab = v * v;
  for (int ij = blockIdx.x; ij < nij; ij += gridDim.x) {
    const int i = ij / o, j = ij - i * o;
    const double* __restrict__ row = e + n1 + (long)ij * vv;
    const double* __restrict__ twin = e + n1 + ((long)j * o + i) * vv;
    // list 46 position of element (A=a,b; I=i,j): a + v*b + v*v*(i + o*j)
    const long base46 = n1 + vv * ((long)i + (long)o * j);
    const long base44 = n1 + n2 + nvm * pk_lt(i, j);
    // four elements per thread in flight: the loads go out before the updates
    for (int ab0 = threadIdx.x; ab0 < nab; ab0 += 4 * blockDim.x) {
      double x[4], t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int ab = ab0 + u * blockDim.x;
        x[u] = ab < nab ? row[ab] : 0.0;
        t[u] = (ab < nab && i < j) ? twin[ab] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int ab = ab0 + u * blockDim.x;
        if (ab >= nab) break;
        const int a = ab / v, b = ab - a * v;
        am_update(m, x[u], base46 + a + (long)v * b);
        if (i < j && a < b) am_update(m, x[u] - t[u], base44 + pk_lt(a, b));
      }
    }
  }
  (void)nom;
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      AbsMax o2 = sh[threadIdx.x + w];
      AbsMax& mine = sh[threadIdx.x];
      if (o2.a > mine.a || (o2.a == mine.a && o2.pos < mine.pos)) mine = o2;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void absmax_finish_kernel(const AbsMax* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ AbsMax sh[256];
  AbsMax m{-1.0, 0x7fffffffffffffffL, 0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const AbsMax o2 = part[i];
    if (o2.a > m.a || (o2.a == m.a && o2.pos < m.pos)) m = o2;
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      AbsMax o2 = sh[threadIdx.x + w];
      AbsMax& mine = sh[threadIdx.x];
      if (o2.a > mine.a || (o2.a == mine.a && o2.pos < mine.pos)) mine = o2;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0].val;
}

// The first half of a DODIIS0 step in one pass: e = xn - x is written to
// the error slot and x to the solution slot, with the ISAMAX of e (as
// absmax_partial_kernel) and the m dots <x_k, e> (as diis_dot_partial_kernel;
// x_self is e itself).  The same-spin twin e[tw] is recomputed from xn and x.
__global__ void diis_error_partial_kernel(int o, int v, int m, int self, const double* const* __restrict__ xs,
                                          const double* __restrict__ x, const double* __restrict__ xn,
                                          double* __restrict__ e_out, double* __restrict__ sol_out,
                                          double* __restrict__ dpart, AbsMax* __restrict__ apart) {
  __shared__ double shd[256];
  __shared__ AbsMax sh[256];
  const long n1 = (long)o * v, vv = (long)v * v, n2 = n1 * n1;
  const long nvm = (long)v * (v - 1) / 2;
  AbsMax am{-1.0, 0x7fffffffffffffffL, 0.0};
  double acc[8];
  for (int k = 0; k < 8; ++k) acc[k] = 0.0;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n1 + n2;
       idx += (long)gridDim.x * blockDim.x) {
    const double xv = x[idx];
    const double ev = xn[idx] - xv;
    e_out[idx] = ev;
    sol_out[idx] = xv;
    long tw = -1;
    double etw = 0.0, eaa = 0.0;
    if (idx < n1) {
      am_update(am, ev, idx);
    } else {
      const long r = idx - n1;
      const long ij = r / vv, ab = r - ij * vv;
      const int i = (int)(ij / o), j = (int)(ij % o), a = (int)(ab / v), b = (int)(ab % v);
      am_update(am, ev, n1 + (long)a + (long)v * b + vv * ((long)i + (long)o * j));
      if (i < j && a < b) {
        tw = n1 + ((long)j * o + i) * vv + ab;
        etw = xn[tw] - x[tw];
        eaa = ev - etw;
        am_update(am, eaa, n1 + n2 + pk_lt(a, b) + nvm * pk_lt(i, j));
      }
    }
    for (int k = 0; k < m; ++k) {
      const double xk = (k == self) ? ev : xs[k][idx];
      double c = xk * ev;
      if (tw >= 0) c += (xk - ((k == self) ? etw : xs[k][tw])) * eaa;
      acc[k] += c;
    }
  }
  for (int k = 0; k < m; ++k) {
    const double sum = block_sum(acc[k], shd);
    if (threadIdx.x == 0) dpart[k * kRedBlocks + blockIdx.x] = sum;
  }
  sh[threadIdx.x] = am;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      AbsMax o2 = sh[threadIdx.x + w];
      AbsMax& mine = sh[threadIdx.x];
      if (o2.a > mine.a || (o2.a == mine.a && o2.pos < mine.pos)) mine = o2;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) apart[blockIdx.x] = sh[0];
}

int diis_error(int o, int v, const double* x, const double* xn, int m, const double* const* d_xs, int self,
               double* e_out, double* sol_out, void* part, double* out_dots, double* out_absmax, hipStream_t s) {
  ACES_REQUIRE(m >= 1 && m <= 8 && self >= 0 && self < m, "diis_error: 1..8 vectors, self among them");
  double* dpart = reinterpret_cast<double*>(part);
  AbsMax* apart = reinterpret_cast<AbsMax*>(dpart + 8L * kRedBlocks);
  hipLaunchKernelGGL(diis_error_partial_kernel, dim3(kRedBlocks), dim3(256), 0, s, o, v, m, self, d_xs, x, xn, e_out,
                     sol_out, dpart, apart);
  hipLaunchKernelGGL(finish_multi_kernel, dim3(1), dim3(256), 0, s, dpart, m, out_dots);
  hipLaunchKernelGGL(absmax_finish_kernel, dim3(1), dim3(256), 0, s, apart, kRedBlocks, out_absmax);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

int diis_absmax(int o, int v, const double* e, void* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL(absmax_partial_kernel, dim3(kRedBlocks), dim3(256), 0, s, o, v, e,
                     reinterpret_cast<AbsMax*>(part));
  hipLaunchKernelGGL(absmax_finish_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<const AbsMax*>(part),
                     kRedBlocks, out);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

long reduction_scratch_bytes() { return (long)kRedBlocks * 8 * (long)sizeof(double) + kRedBlocks * sizeof(AbsMax); }

// out[idx] = sum_k c[k] * x_k[idx] evaluated in the given order (k = 0..m-1).
__global__ void lincomb_kernel(long n, int m, const double* const* __restrict__ xs, const double* __restrict__ c,
                               double* __restrict__ out) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
    double acc = c[0] * xs[0][idx];
    for (int k = 1; k < m; ++k) acc += c[k] * xs[k][idx];
    out[idx] = acc;
  }
}

int lincomb(long n, int m, const double* const* d_xs, const double* d_c, double* out, hipStream_t s) {
  ACES_REQUIRE(m >= 1, "lincomb: empty");
  hipLaunchKernelGGL(lincomb_kernel, dim3(grid1(n)), dim3(256), 0, s, n, m, d_xs, d_c, out);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// y = x - z (error vector)
__global__ void sub_kernel(long n, const double* __restrict__ x, const double* __restrict__ z,
                           double* __restrict__ y) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x)
    y[idx] = x[idx] - z[idx];
}

int sub(long n, const double* x, const double* z, double* y, hipStream_t s) {
  hipLaunchKernelGGL(sub_kernel, dim3(grid1(n)), dim3(256), 0, s, n, x, z, y);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// Remove the diagonal: X[p][p] -= d[p]  (square n x n, leading dimension ld).
__global__ void sub_diag_kernel(int n, double* __restrict__ X, long ld, const double* __restrict__ d) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) X[(long)p * ld + p] -= d[p];
}

int sub_diag(int n, double* X, long ld, const double* d, hipStream_t s) {
  hipLaunchKernelGGL(sub_diag_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, X, ld, d);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

}  // namespace aces
