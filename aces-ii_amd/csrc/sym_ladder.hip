// Particle-particle ladder kernels on irrep-blocked storage (see sym_ladder.h).
// Every kernel writes each output element from one thread (gathers only).
#include "common.h"
#include "sym_ladder.h"

namespace aces {
namespace dpd {

LadderDesc make_ladder(const Space& O, const Space& V) {
  LadderDesc d;
  d.nirr = O.nirr;
  d.O = O;
  d.V = V;
  d.VP = make_pairs(V, false);
  d.VM = make_pairs(V, true);
  d.OP = make_pairs(O, false);
  d.OM = make_pairs(O, true);
  Layout ovL = make_layout(4, 2, &O, &V, &O, &V);
  long a = 0, b = 0, c = 0, e = 0, f = 0, g = 0;
  for (int G = 0; G < d.nirr; ++G) {
    d.nkb[G] = ovL.nrow[G];
    for (int h = 0; h < d.nirr; ++h) d.kbsub[G][h] = ovL.subr[G][h];
    d.vp_off[G] = a;
    d.vm_off[G] = b;
    d.sp_off[G] = c;
    d.dp_off[G] = e;
    d.zp_off[G] = f;
    d.zm_off[G] = g;
    a += (d.VP.n[G] + d.nkb[G]) * d.VP.n[G];
    b += (d.VM.n[G] + d.nkb[G]) * d.VM.n[G];
    c += d.OP.n[G] * d.VP.n[G];
    e += d.OM.n[G] * d.VM.n[G];
    f += (d.VP.n[G] + d.nkb[G]) * d.OP.n[G];
    g += (d.VM.n[G] + d.nkb[G]) * d.OM.n[G];
  }
  d.vp_off[d.nirr] = a;
  d.vm_off[d.nirr] = b;
  d.sp_off[d.nirr] = c;
  d.dp_off[d.nirr] = e;
  d.zp_off[d.nirr] = f;
  d.zm_off[d.nirr] = g;
  return d;
}

__device__ inline int block_of(const long* off, int nirr, long idx) {
  int G = 0;
  while (G + 1 < nirr && idx >= off[G + 1]) ++G;
  return G;
}

__device__ inline void kb_decode(const LadderDesc& d, int G, long q, int& gk, int& lk, int& gb, int& lb) {
  gk = gb = lk = lb = 0;
  for (int h = 0; h < d.nirr; ++h) {
    const int hb = G ^ h;
    const long sz = (long)d.O.cnt[h] * d.V.cnt[hb];
    const long b0 = d.kbsub[G][h];
    if (q >= b0 && q < b0 + sz) {
      const long t = q - b0;
      gk = h;
      gb = hb;
      lk = (int)(t / d.V.cnt[hb]);
      lb = (int)(t - (long)lk * d.V.cnt[hb]);
      return;
    }
  }
}

__device__ inline long kb_row(const LadderDesc& d, int G, int gk, int lk, int gb, int lb) {
  return d.kbsub[G][gk] + (long)lk * d.V.cnt[gb] + lb;
}

// Sort two orbitals of one space into global order.
__device__ inline void order2(const Space& S, int& g0, int& l0, int& g1, int& l1) {
  const int a = S.off[g0] + l0, b = S.off[g1] + l1;
  if (a > b) {
    int t = g0; g0 = g1; g1 = t;
    t = l0; l0 = l1; l1 = t;
  }
}

// ---------------------------------------------------------------------------
__global__ void build_ladder_sym_kernel(const LadderDesc* __restrict__ dp, const double* __restrict__ L233,
                                        const double* __restrict__ ovvv, double* __restrict__ Vp,
                                        double* __restrict__ Vm) {
  const LadderDesc& d = *dp;
  const long v = d.V.n;
  const long nvpg = v * (v + 1) / 2;
  const long totP = d.vp_off[d.nirr], tot = totP + d.vm_off[d.nirr];
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < tot; idx += (long)gridDim.x * blockDim.x) {
    const bool plus = idx < totP;
    const long e = plus ? idx : idx - totP;
    const int G = block_of(plus ? d.vp_off : d.vm_off, d.nirr, e);
    const PairSpace& P = plus ? d.VP : d.VM;
    const long nc = P.n[G];
    const long r = (e - (plus ? d.vp_off[G] : d.vm_off[G])) / nc;
    const long cc = (e - (plus ? d.vp_off[G] : d.vm_off[G])) - r * nc;
    int gc, lc, gd, ld;
    pair_decode(P, d.V, G, cc, gc, lc, gd, ld);
    const long c = d.V.off[gc] + lc, dd = d.V.off[gd] + ld;
    double val;
    if (r < P.n[G]) {
      int ga, la, gb, lb;
      pair_decode(P, d.V, G, r, ga, la, gb, lb);
      const long a = d.V.off[ga] + la, b = d.V.off[gb] + lb;
      const long pab = b * (b + 1) / 2 + a;
      const double x = L233[(c + v * dd) * nvpg + pab];
      const double y = (c == dd) ? 0.0 : L233[(dd + v * c) * nvpg + pab];
      val = plus ? x + y : x - y;
    } else {
      int gk, lk, gb, lb;
      kb_decode(d, G, r - P.n[G], gk, lk, gb, lb);
      const long k = d.O.off[gk] + lk, b = d.V.off[gb] + lb;
      const double x = ovvv[((k * v + c) * v + b) * v + dd];                    // (kc|bd)
      const double y = (c == dd) ? 0.0 : ovvv[((k * v + dd) * v + b) * v + c];  // (kd|bc)
      val = plus ? x + y : x - y;
    }
    (plus ? Vp : Vm)[e] = val;
  }
}

int build_ladder_sym(const LadderDesc* d_desc, const LadderDesc& h, const double* L233, const double* ovvv,
                     double* Vp, double* Vm, hipStream_t s) {
  const long n = h.vp_off[h.nirr] + h.vm_off[h.nirr];
  if (n == 0) return ACES_OK;
  hipLaunchKernelGGL(build_ladder_sym_kernel, dim3((unsigned)std::min<long>((n + 255) / 256, 1L << 20)),
                     dim3(256), 0, s, d_desc, L233, ovvv, Vp, Vm);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
__global__ void pack_tau_sym_kernel(const LadderDesc* __restrict__ dp, const Layout* __restrict__ tl,
                                    const double* __restrict__ tau, double* __restrict__ Sp,
                                    double* __restrict__ Dp) {
  const LadderDesc& d = *dp;
  const long totS = d.sp_off[d.nirr], tot = totS + d.dp_off[d.nirr];
  // kPackBatch elements per thread and pass, their gathers in flight together
  // (past the end: element 0, not written)
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i0 = blockIdx.x * (long)blockDim.x + threadIdx.x; i0 < tot; i0 += kPackBatch * stride) {
    long o1[kPackBatch], o2[kPackBatch];
#pragma unroll
    for (int u = 0; u < kPackBatch; ++u) {
      const long idx0 = i0 + u * stride;
      const long idx = idx0 < tot ? idx0 : 0;
      const bool sym = idx < totS;
      const long e = sym ? idx : idx - totS;
      const int G = block_of(sym ? d.sp_off : d.dp_off, d.nirr, e);
      const PairSpace& PO = sym ? d.OP : d.OM;
      const PairSpace& PV = sym ? d.VP : d.VM;
      const long nc = PV.n[G];
      const long loc = e - (sym ? d.sp_off[G] : d.dp_off[G]);
      const long r = loc / nc, cc = loc - r * nc;
      int gi, li, gj, lj, gc, lc, gd, ld;
      pair_decode(PO, d.O, G, r, gi, li, gj, lj);
      pair_decode(PV, d.V, G, cc, gc, lc, gd, ld);
      const int g1[4] = {gi, gj, gc, gd}, l1[4] = {li, lj, lc, ld};
      const int g2[4] = {gj, gi, gc, gd}, l2[4] = {lj, li, lc, ld};
      o1[u] = tl->offset(g1, l1);
      o2[u] = tl->offset(g2, l2);
    }
    double x[kPackBatch], y[kPackBatch];
#pragma unroll
    for (int u = 0; u < kPackBatch; ++u) {
      x[u] = tau[o1[u]];
      y[u] = tau[o2[u]];
    }
#pragma unroll
    for (int u = 0; u < kPackBatch; ++u) {
      const long idx = i0 + u * stride;
      if (idx < tot) {
        const bool sym = idx < totS;
        (sym ? Sp : Dp)[sym ? idx : idx - totS] = sym ? 0.5 * (x[u] + y[u]) : 0.5 * (x[u] - y[u]);
      }
    }
  }
}

int pack_tau_sym(const LadderDesc* d_desc, const LadderDesc& h, const Layout* d_tauL, const double* tau, double* Sp,
                 double* Dp, hipStream_t s) {
  const long n = h.sp_off[h.nirr] + h.dp_off[h.nirr];
  if (n == 0) return ACES_OK;
  hipLaunchKernelGGL(pack_tau_sym_kernel,
                     dim3((unsigned)std::min<long>((n + 256 * kPackBatch - 1) / (256 * kPackBatch), 1L << 20)),
                     dim3(256), 0, s, d_desc, d_tauL, tau, Sp, Dp);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// mode 0: walk inc2 (i,j,a,b) and add Z; mode 1: walk Yb (b,i,j,k) and set the (k,b) rows.
__global__ __launch_bounds__(256) void unpack_sym_kernel(const LadderDesc* __restrict__ dp, const Box* boxes, int nbox,
                                                         int mode, const double* __restrict__ Zp,
                                                         const double* __restrict__ Zm, double* __restrict__ out) {
  __shared__ int sb;
  if (threadIdx.x == 0) sb = box_of_chunk(boxes, nbox, blockIdx.x);
  __syncthreads();
  const LadderDesc& d = *dp;
  const Box& bx = boxes[sb];
  const long elems = (long)bx.d[0] * bx.d[1] * bx.d[2] * bx.d[3];
  const long e0 = (blockIdx.x - bx.first) * kChunk;
  const long e1 = min(elems, e0 + kChunk);
  for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int l3 = (int)(e % bx.d[3]);
    const long r2 = e / bx.d[3];
    const int l2 = (int)(r2 % bx.d[2]);
    const long r1 = r2 / bx.d[2];
    const int l1 = (int)(r1 % bx.d[1]);
    const int l0 = (int)(r1 / bx.d[1]);
    double* o = out + bx.ob + l0 * bx.os[0] + l1 * bx.os[1] + l2 * bx.os[2] + l3 * bx.os[3];
    if (mode == 0) {
      int gi = bx.g[0], li = l0, gj = bx.g[1], lj = l1, ga = bx.g[2], la = l2, gb = bx.g[3], lb = l3;
      const int G = ga ^ gb;
      const long I = d.O.off[gi] + li, J = d.O.off[gj] + lj, A = d.V.off[ga] + la, B = d.V.off[gb] + lb;
      order2(d.O, gi, li, gj, lj);
      order2(d.V, ga, la, gb, lb);
      const long pab = pair_index(d.VP, d.V, ga, la, gb, lb);
      const long pij = pair_index(d.OP, d.O, gi, li, gj, lj);
      double z = Zp[d.zp_off[G] + pab * d.OP.n[G] + pij];
      if (I != J && A != B) {
        const double sg = ((I < J) == (A < B)) ? 1.0 : -1.0;
        const long mab = pair_index(d.VM, d.V, ga, la, gb, lb);
        const long mij = pair_index(d.OM, d.O, gi, li, gj, lj);
        z += sg * Zm[d.zm_off[G] + mab * d.OM.n[G] + mij];
      }
      *o += z;
    } else {
      const int gb = bx.g[0], lb = l0, gk = bx.g[3], lk = l3;
      int gi = bx.g[1], li = l1, gj = bx.g[2], lj = l2;
      const int G = gk ^ gb;
      const long I = d.O.off[gi] + li, J = d.O.off[gj] + lj;
      order2(d.O, gi, li, gj, lj);
      const long kb = kb_row(d, G, gk, lk, gb, lb);
      double y = Zp[d.zp_off[G] + (d.VP.n[G] + kb) * d.OP.n[G] + pair_index(d.OP, d.O, gi, li, gj, lj)];
      if (I != J)
        y += ((I < J) ? 1.0 : -1.0) *
             Zm[d.zm_off[G] + (d.VM.n[G] + kb) * d.OM.n[G] + pair_index(d.OM, d.O, gi, li, gj, lj)];
      *o = y;
    }
  }
}

// Occupied pairs of an inc2 box (irreps gi, gj) in Z's column order: gi < gj
// the rectangle (li, lj), li slowest; gi > gj the rectangle (lj, li), lj
// slowest; gi == gj the triangle lp <= lq, lq slowest, each pair standing for
// the outputs (lp, lq) and (lq, lp).
__host__ __device__ inline int box_pairs(int gi, int gj, int oi, int oj) {
  return gi == gj ? oi * (oi + 1) / 2 : oi * oj;
}

__device__ inline void tri_decode(int q, int& lp, int& lq) {
  int r = (int)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > q) --r;
  while ((r + 1) * (r + 2) / 2 <= q) ++r;
  lq = r;
  lp = q - r * (r + 1) / 2;
}

void unpack_tiles(const Plan& p, std::vector<UnpackTile>& tiles) {
  tiles.clear();
  for (int k = 0; k < p.nbox; ++k) {
    const Box& b = p.host[k];
    const int nq = box_pairs(b.g[0], b.g[1], b.d[0], b.d[1]);
    const int vb = b.d[3];
    const int nbc = (vb + 63) / 64, tb = (vb + nbc - 1) / nbc;
    const int qmax = std::max(1, kUnpackLds / tb);
    const int nqc = (nq + qmax - 1) / qmax, tq = (nq + nqc - 1) / nqc;
    for (int la = 0; la < b.d[2]; ++la)
      for (int lb0 = 0; lb0 < vb; lb0 += tb)
        for (int q0 = 0; q0 < nq; q0 += tq) {
          UnpackTile t;
          t.box = k;
          t.la = (short)la;
          t.lb0 = (short)lb0;
          t.tb = (short)tb;
          t.tq = (short)tq;
          t.q0 = q0;
          tiles.push_back(t);
        }
  }
}

__global__ __launch_bounds__(256) void unpack_sym_tiled_kernel(const LadderDesc* __restrict__ dp, const Box* boxes,
                                                               const UnpackTile* __restrict__ tiles,
                                                               const double* __restrict__ Zp,
                                                               const double* __restrict__ Zm,
                                                               double* __restrict__ out) {
  __shared__ double tp[kUnpackLds], tm[kUnpackLds];
  const UnpackTile tl = tiles[blockIdx.x];
  const LadderDesc& d = *dp;
  const Box& bx = boxes[tl.box];
  const int gi = bx.g[0], gj = bx.g[1], ga = bx.g[2], gb = bx.g[3];
  const int oi = bx.d[0], oj = bx.d[1];
  const int G = ga ^ gb;
  const int la = tl.la, lb0 = tl.lb0, q0 = tl.q0;
  const int nq = box_pairs(gi, gj, oi, oj);
  const int eq = min((int)tl.tq, nq - q0), eb = min((int)tl.tb, bx.d[3] - lb0);
  const long nOP = d.OP.n[G], nOM = d.OM.n[G];
  const int glo = gi < gj ? gi : gj;
  const long pij0 = d.OP.sub[G][glo] + q0, mij0 = d.OM.sub[G][glo] + q0;
  const long A = d.V.off[ga] + la;
  // Z rows of this tile (one per b) read along the box's pairs into LDS.
  // Loads are branch-free in batches of kUnpackBatch (an element outside the
  // tile or without a Z- term reads element 0 of Zp / Zm and is dropped), so
  // that a thread's loads are in flight together.
  const int nrd = eb * eq;
  const bool zm_any = d.zm_off[d.nirr] > 0;
  for (int base = threadIdx.x; base < nrd; base += kUnpackBatch * blockDim.x) {
    double pv[kUnpackBatch], mv[kUnpackBatch];
#pragma unroll
    for (int u = 0; u < kUnpackBatch; ++u) {
      const int idx = base + u * blockDim.x;
      const bool ok = idx < nrd;
      const int id = ok ? idx : 0;
      const int row = id / eq, qq = id - row * eq;
      const int lb = lb0 + row;
      int g0 = ga, l0 = la, g1 = gb, l1 = lb;
      order2(d.V, g0, l0, g1, l1);
      const long pab = pair_index(d.VP, d.V, g0, l0, g1, l1);
      long mij = -1;
      if (gi != gj) {
        mij = mij0 + qq;
      } else {
        int lp, lq;
        tri_decode(q0 + qq, lp, lq);
        if (lp != lq) mij = d.OM.sub[0][gi] + (long)lq * (lq - 1) / 2 + lp;
      }
      const bool okm = ok && zm_any && A != d.V.off[gb] + lb && mij >= 0;
      pv[u] = Zp[d.zp_off[G] + pab * nOP + pij0 + qq];
      mv[u] = zm_any ? Zm[okm ? d.zm_off[G] + pair_index(d.VM, d.V, g0, l0, g1, l1) * nOM + mij : 0] : 0.0;
      if (!okm) mv[u] = 0.0;
    }
#pragma unroll
    for (int u = 0; u < kUnpackBatch; ++u) {
      const int idx = base + u * blockDim.x;
      if (idx < nrd) {
        tp[idx] = pv[u];
        tm[idx] = mv[u];
      }
    }
  }
  __syncthreads();
  // inc2 += Zp + sign Zm, written along b (a pair of the triangle writes twice)
  const int nout = (gi == gj) ? 2 * eq : eq;
  const long o2 = la * bx.os[2];
  // (the old values of a batch are loaded together, then updated; the
  // duplicate of a diagonal pair and elements past the tile read element
  // (0, 0, la, lb0) of the box and write nothing)
  const int nwr = nout * eb;
  for (int base = threadIdx.x; base < nwr; base += kUnpackBatch * blockDim.x) {
    long off[kUnpackBatch];
    double z[kUnpackBatch], old[kUnpackBatch];
    bool wr[kUnpackBatch];
#pragma unroll
    for (int u = 0; u < kUnpackBatch; ++u) {
      const int idx = base + u * blockDim.x;
      const bool ok = idx < nwr;
      const int id = ok ? idx : 0;
      const int oq = id / eb, row = id - oq * eb;
      int qq, li, lj;
      bool dup = false;
      if (gi == gj) {
        qq = oq >> 1;
        int lp, lq;
        tri_decode(q0 + qq, lp, lq);
        dup = (oq & 1) && lp == lq;
        li = (oq & 1) ? lq : lp;
        lj = (oq & 1) ? lp : lq;
      } else {
        qq = oq;
        const int q = q0 + qq;
        if (gi < gj) {
          li = q / oj;
          lj = q - li * oj;
        } else {
          lj = q / oi;
          li = q - lj * oi;
        }
      }
      const int lb = lb0 + row;
      const long I = d.O.off[gi] + li, J = d.O.off[gj] + lj, B = d.V.off[gb] + lb;
      double zz = tp[row * eq + qq];
      if (I != J && A != B) zz += (((I < J) == (A < B)) ? 1.0 : -1.0) * tm[row * eq + qq];
      wr[u] = ok && !dup;
      off[u] = wr[u] ? bx.ob + li * bx.os[0] + lj * bx.os[1] + o2 + lb * bx.os[3] : bx.ob + o2 + lb0 * bx.os[3];
      z[u] = zz;
      old[u] = out[off[u]];
    }
#pragma unroll
    for (int u = 0; u < kUnpackBatch; ++u)
      if (wr[u]) out[off[u]] = old[u] + z[u];
  }
}

int unpack_ladder_sym(const LadderDesc* d_desc, const Plan& inc2_plan, const UnpackTile* tiles, long ntiles,
                      const Plan& yb_plan, const double* Zp, const double* Zm, double* inc2, double* Yb,
                      hipStream_t s) {
  ACES_REQUIRE(ntiles < (1L << 31), "unpack: too many tiles");
  if (ntiles > 0)
    hipLaunchKernelGGL(unpack_sym_tiled_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, d_desc, inc2_plan.dev,
                       tiles, Zp, Zm, inc2);
  if (yb_plan.total > 0)
    hipLaunchKernelGGL(unpack_sym_kernel, dim3((unsigned)yb_plan.total), dim3(256), 0, s, d_desc, yb_plan.dev,
                       yb_plan.nbox, 1, Zp, Zm, Yb);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
__global__ void unpack_sym_adjoint_kernel(const LadderDesc* __restrict__ dp, const Layout* __restrict__ L2,
                                          const Layout* __restrict__ LY, const double* __restrict__ g2,
                                          const double* __restrict__ gYb, double* __restrict__ gZp,
                                          double* __restrict__ gZm) {
  const LadderDesc& d = *dp;
  const long totP = d.zp_off[d.nirr], tot = totP + d.zm_off[d.nirr];
  auto G2 = [&](int gi, int li, int gj, int lj, int ga, int la, int gb, int lb) {
    const int g[4] = {gi, gj, ga, gb}, l[4] = {li, lj, la, lb};
    return g2[L2->offset(g, l)];
  };
  auto GY = [&](int gb, int lb, int gi, int li, int gj, int lj, int gk, int lk) {
    const int g[4] = {gb, gi, gj, gk}, l[4] = {lb, li, lj, lk};
    return gYb[LY->offset(g, l)];
  };
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < tot; idx += (long)gridDim.x * blockDim.x) {
    const bool plus = idx < totP;
    const long e = plus ? idx : idx - totP;
    const int G = block_of(plus ? d.zp_off : d.zm_off, d.nirr, e);
    const PairSpace& PO = plus ? d.OP : d.OM;
    const PairSpace& PV = plus ? d.VP : d.VM;
    const long nc = PO.n[G];
    const long loc = e - (plus ? d.zp_off[G] : d.zm_off[G]);
    const long r = loc / nc, cc = loc - r * nc;
    int gi, li, gj, lj;
    pair_decode(PO, d.O, G, cc, gi, li, gj, lj);
    const bool dij = (gi != gj) || (li != lj);
    double z;
    if (r < PV.n[G]) {
      int ga, la, gb, lb;
      pair_decode(PV, d.V, G, r, ga, la, gb, lb);
      const bool dab = (ga != gb) || (la != lb);
      if (plus) {
        z = G2(gi, li, gj, lj, ga, la, gb, lb);
        if (dab) z += G2(gi, li, gj, lj, gb, lb, ga, la);
        if (dij) {
          z += G2(gj, lj, gi, li, ga, la, gb, lb);
          if (dab) z += G2(gj, lj, gi, li, gb, lb, ga, la);
        }
      } else {
        z = G2(gi, li, gj, lj, ga, la, gb, lb) - G2(gj, lj, gi, li, ga, la, gb, lb) -
            G2(gi, li, gj, lj, gb, lb, ga, la) + G2(gj, lj, gi, li, gb, lb, ga, la);
      }
    } else {
      int gk, lk, gb, lb;
      kb_decode(d, G, r - PV.n[G], gk, lk, gb, lb);
      if (plus) {
        z = GY(gb, lb, gi, li, gj, lj, gk, lk);
        if (dij) z += GY(gb, lb, gj, lj, gi, li, gk, lk);
      } else {
        z = GY(gb, lb, gi, li, gj, lj, gk, lk) - GY(gb, lb, gj, lj, gi, li, gk, lk);
      }
    }
    (plus ? gZp : gZm)[e] = z;
  }
}

int unpack_ladder_sym_adjoint(const LadderDesc* d_desc, const LadderDesc& h, const Layout* d_inc2L,
                              const Layout* d_ybL, const double* g2, const double* gYb, double* gZp, double* gZm,
                              hipStream_t s) {
  const long n = h.zp_off[h.nirr] + h.zm_off[h.nirr];
  if (n == 0) return ACES_OK;
  hipLaunchKernelGGL(unpack_sym_adjoint_kernel, dim3((unsigned)std::min<long>((n + 255) / 256, 1L << 20)),
                     dim3(256), 0, s, d_desc, d_inc2L, d_ybL, g2, gYb, gZp, gZm);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pack_tau_sym_adjoint_kernel(const LadderDesc* __restrict__ dp,
                                                                   const Box* boxes, int nbox,
                                                                   const double* __restrict__ gSp,
                                                                   const double* __restrict__ gDp,
                                                                   double* __restrict__ gtau) {
  __shared__ int sb;
  if (threadIdx.x == 0) sb = box_of_chunk(boxes, nbox, blockIdx.x);
  __syncthreads();
  const LadderDesc& d = *dp;
  const Box& bx = boxes[sb];
  const long elems = (long)bx.d[0] * bx.d[1] * bx.d[2] * bx.d[3];
  const long e0 = (blockIdx.x - bx.first) * kChunk;
  const long e1 = min(elems, e0 + kChunk);
  for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int l3 = (int)(e % bx.d[3]);
    const long r2 = e / bx.d[3];
    const int l2 = (int)(r2 % bx.d[2]);
    const long r1 = r2 / bx.d[2];
    const int l1 = (int)(r1 % bx.d[1]);
    const int l0 = (int)(r1 / bx.d[1]);
    int gi = bx.g[0], li = l0, gj = bx.g[1], lj = l1;
    const int gc = bx.g[2], lc = l2, gd = bx.g[3], ld = l3;
    const long C = d.V.off[gc] + lc, D = d.V.off[gd] + ld;
    if (C > D) continue;
    const long I = d.O.off[gi] + li, J = d.O.off[gj] + lj;
    const int G = gc ^ gd;
    order2(d.O, gi, li, gj, lj);
    const long pcd = pair_index(d.VP, d.V, gc, lc, gd, ld);
    double g = gSp[d.sp_off[G] + pair_index(d.OP, d.O, gi, li, gj, lj) * d.VP.n[G] + pcd] * (I == J ? 1.0 : 0.5);
    if (C < D && I != J)
      g += (I < J ? 0.5 : -0.5) *
           gDp[d.dp_off[G] + pair_index(d.OM, d.O, gi, li, gj, lj) * d.VM.n[G] + pair_index(d.VM, d.V, gc, lc, gd, ld)];
    gtau[bx.ob + l0 * bx.os[0] + l1 * bx.os[1] + l2 * bx.os[2] + l3 * bx.os[3]] += g;
  }
}

int pack_tau_sym_adjoint(const LadderDesc* d_desc, const Plan& tau_plan, const double* gSp, const double* gDp,
                         double* gtau, hipStream_t s) {
  if (tau_plan.total == 0) return ACES_OK;
  hipLaunchKernelGGL(pack_tau_sym_adjoint_kernel, dim3((unsigned)tau_plan.total), dim3(256), 0, s, d_desc,
                     tau_plan.dev, tau_plan.nbox, gSp, gDp, gtau);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

}  // namespace dpd
}  // namespace aces
