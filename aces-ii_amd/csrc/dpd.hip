// Irrep-blocked tensor layouts, box plans and the HBM-bound kernels that walk them.
#include <algorithm>
#include <cstring>
#include "common.h"
#include "dpd.h"
#include "bf16_util.h"

namespace aces {
namespace dpd {

Space Space::make(int nirr, const int* counts) {
  Space s;
  s.nirr = nirr;
  int o = 0;
  for (int g = 0; g < nirr; ++g) {
    s.cnt[g] = counts[g];
    s.off[g] = o;
    o += counts[g];
  }
  s.n = o;
  return s;
}

static long compound_count(const Space* sp, int k, int G, int nirr, long* sub) {
  // compound of k indices (spaces sp[0..k-1]) with total irrep G; sub[t] for
  // the tuple t of the leading k-1 irreps (lexicographic).
  long tot = 0;
  if (k == 1) {
    sub[0] = 0;
    return sp[0].cnt[G];
  }
  if (k == 2) {
    for (int g0 = 0; g0 < nirr; ++g0) {
      sub[g0] = tot;
      tot += (long)sp[0].cnt[g0] * sp[1].cnt[G ^ g0];
    }
    return tot;
  }
  for (int g0 = 0; g0 < nirr; ++g0)
    for (int g1 = 0; g1 < nirr; ++g1) {
      sub[g0 * nirr + g1] = tot;
      tot += (long)sp[0].cnt[g0] * sp[1].cnt[g1] * sp[2].cnt[G ^ g0 ^ g1];
    }
  return tot;
}

Layout make_layout(int rank, int split, const Space* s0, const Space* s1, const Space* s2, const Space* s3) {
  Layout L;
  L.rank = rank;
  L.split = split;
  const Space* ss[4] = {s0, s1, s2, s3};
  for (int k = 0; k < rank; ++k) L.sp[k] = *ss[k];
  L.nirr = s0->nirr;
  long off = 0;
  for (int G = 0; G < L.nirr; ++G) {
    L.nrow[G] = compound_count(L.sp, split, G, L.nirr, L.subr[G]);
    L.ncol[G] = compound_count(L.sp + split, rank - split, G, L.nirr, L.subc[G]);
    L.blk_off[G] = off;
    off += L.nrow[G] * L.ncol[G];
  }
  L.blk_off[L.nirr] = off;
  return L;
}

Layout make_dense(int rank, const Space* const* sp, const long* strides) {
  Layout L;
  L.rank = rank;
  L.split = rank / 2;
  L.dense = 1;
  L.nirr = sp[0]->nirr;
  long n = 1;
  for (int k = 0; k < rank; ++k) {
    L.sp[k] = *sp[k];
    L.dstride[k] = strides[k];
    n *= sp[k]->n;
  }
  L.blk_off[L.nirr] = n;
  return L;
}

bool Layout::same(const Layout& o) const {
  if (rank != o.rank || split != o.split || nirr != o.nirr) return false;
  for (int k = 0; k < rank; ++k)
    for (int g = 0; g < nirr; ++g)
      if (sp[k].cnt[g] != o.sp[k].cnt[g]) return false;
  return true;
}

PairSpace make_pairs(const Space& s, bool strict) {
  PairSpace P;
  P.strict = strict;
  P.nirr = s.nirr;
  for (int G = 0; G < s.nirr; ++G) {
    long tot = 0;
    for (int gp = 0; gp < s.nirr; ++gp) {
      const int gq = G ^ gp;
      P.sub[G][gp] = tot;
      if (G == 0) {
        const long c = s.cnt[gp];
        tot += strict ? c * (c - 1) / 2 : c * (c + 1) / 2;
      } else if (gp < gq) {
        tot += (long)s.cnt[gp] * s.cnt[gq];
      }
    }
    P.n[G] = tot;
  }
  return P;
}

// ---------------------------------------------------------------------------

// Choose the tiled walk for a rank-4 plan from its largest box: a source is
// read through LDS when its stride along output index 3 is large and it has
// unit stride along an index q of 0..2; all such sources must agree on q
// (with none, q is the largest of indices 0..2).
// (The tiled kernel uses every box's own strides, so this choice only decides
// speed.)  Sets p.tq, p.tr, p.ttotal and Box::tfirst.
int source_fast_index(const Plan& p, int rank, int s) {
  if (rank != 4 || p.nbox == 0 || s < 0 || s >= p.nsrc) return -1;
  const Box* big = &p.host[0];
  long best = 0;
  for (const Box& b : p.host) {
    const long e = (long)b.d[0] * b.d[1] * b.d[2] * b.d[3];
    if (e > best) {
      best = e;
      big = &b;
    }
  }
  if (big->d[3] < kTileMinFast) return -1;
  const long s3 = big->ss[s][3] < 0 ? -big->ss[s][3] : big->ss[s][3];
  if (s3 < 8) return 3;
  for (int kk = 2; kk >= 0; --kk)
    if (big->d[kk] >= 8 && big->ss[s][kk] == 1) return kk;
  return 3;
}

// Reciprocal for the tiled kernel's divisions: floor(e / d) = umulhi(e, m)
// with m = ceil(2^32 / d), exact for e * d < 2^32 (here e, d <= 4096).
static unsigned div_magic(int d) {
  return d <= 1 ? 0u : (unsigned)(((1ull << 32) + (unsigned long long)d - 1) / (unsigned long long)d);
}

// Tile shape of one box.  The kernel walks a tile as rows along index 3:
// 256 / te[0] rows at a time, at most kTileRows rows per thread, so a tile
// holds te[1] * te[2] <= kTileRows * (256 / te[0]) rows.  Indices 3 and q
// are cut into balanced pieces of at most kTileSpan (a box of 48 x 9 is one
// 48 x 9 piece, not a 64 x 32 tile that it fills to a fifth), then as many
// values of index rb join as the row budget and the LDS image (pitch
// te[1] | 1) allow.
static void box_tiles(Box& b, int q, int rb) {
  const int d3 = b.d[3], dq = b.d[q], dr = b.d[rb];
  const int n3 = (d3 + kTileSpan - 1) / kTileSpan, t3 = (d3 + n3 - 1) / n3;
  const int rows = kTileRows * (256 / t3);
  const int qmax = std::min(kTileSpan, rows);
  const int nq = (dq + qmax - 1) / qmax, tq = (dq + nq - 1) / nq;
  const int pitch = tq | 1;
  int tr = std::max(1, std::min(dr, rows / tq));
  while (tr > 1 && (long)tr * t3 * pitch > kTileLds) --tr;
  const int nr = (dr + tr - 1) / tr;
  tr = (dr + nr - 1) / nr;
  b.te[0] = t3;
  b.te[1] = tq;
  b.te[2] = tr;
  b.tn[0] = n3;
  b.tn[1] = nq;
  b.tn[2] = nr;
  b.mg[0] = div_magic(t3);
  b.mg[1] = div_magic(tq);
}

static void plan_tiles(Plan& p, int rank) {
  p.tq = -1;
  p.tr = 0;
  p.ttotal = 0;
  if (rank != 4 || p.nbox == 0) return;
  const Box* big = &p.host[0];
  long best = 0;
  for (const Box& b : p.host) {
    const long e = (long)b.d[0] * b.d[1] * b.d[2] * b.d[3];
    if (e > best) {
      best = e;
      big = &b;
    }
  }
  if (big->d[3] < kTileMinFast) return;
  int q = -1, mask = 0;
  for (int s = 0; s < p.nsrc; ++s) {
    const long s3 = big->ss[s][3] < 0 ? -big->ss[s][3] : big->ss[s][3];
    if (s3 < 8) continue;
    int k = -1;
    for (int kk = 2; kk >= 0; --kk)
      if (big->d[kk] >= 8 && big->ss[s][kk] == 1) {
        k = kk;
        break;
      }
    if (k < 0) continue;
    if (q >= 0 && q != k) return;
    q = k;
    mask |= 1 << s;
  }
  if (q < 0) {
    // every source reads in the output's order: tile over the largest of
    // indices 0..2 (no LDS turn), which drops the per-element index division
    q = 2;
    for (int kk = 1; kk >= 0; --kk)
      if (big->d[kk] > big->d[q]) q = kk;
  }
  const int ra = (q == 0) ? 1 : 0, rb = (q == 2) ? 1 : 2;
  long t = 0;
  for (Box& b : p.host) {
    b.tfirst = t;
    box_tiles(b, q, rb);
    t += (long)b.tn[0] * b.tn[1] * b.tn[2] * b.d[ra];
  }
  p.tq = q;
  p.tr = mask;
  p.ttotal = t;
  // 32-bit in-tile offsets when every stride walk of a tile stays below 2^31
  p.wide = false;
  auto span = [&](const long* st, const Box& b) {
    return (long)(b.te[0] - 1) * std::labs(st[3]) + (long)(b.te[1] - 1) * std::labs(st[q]) +
           (long)(b.te[2] - 1) * std::labs(st[rb]);
  };
  for (const Box& b : p.host) {
    long m = span(b.os, b);
    for (int s = 0; s < p.nsrc; ++s) m = std::max(m, span(b.ss[s], b));
    if (p.prod) m = std::max(m, std::max(span(b.xs, b), span(b.ys, b)));
    if (p.low) m = std::max(m, span(b.los, b));
    if (m >= (1L << 31)) p.wide = true;
  }
}

int build_boxes(Plan& p, const Layout& out, const std::vector<Src>& srcs, const Src* X, const Src* Y) {
  ACES_REQUIRE(out.rank == 2 || out.rank == 4, "dpd plan: output rank must be 2 or 4");
  ACES_REQUIRE(srcs.size() <= 3, "dpd plan: at most three sources");
  const int nirr = out.nirr;
  p.host.clear();
  p.nsrc = (int)srcs.size();
  p.prod = (X != nullptr);
  for (const Src& s : srcs) ACES_REQUIRE(s.L->rank == out.rank, "dpd plan: source rank differs from output");
  long first = 0;
  // split image of a blocked rank-4 output: block offsets with padded rows
  const bool low = out.rank == 4 && !out.dense && out.split == 2;
  long loff[kMaxIrr + 1] = {0};
  for (int G = 0; G < nirr; ++G) loff[G + 1] = loff[G] + out.nrow[G] * ((std::max<long>(out.ncol[G], 1) + 31) / 32 * 32);
  p.low = low;
  auto add_box = [&](const int* g) {
    Box b;
    std::memset(&b, 0, sizeof(b));
    long elems = 1;
    for (int k = 0; k < 4; ++k) {
      b.d[k] = (k < out.rank) ? out.sp[k].cnt[g[k]] : 1;
      b.g[k] = (k < out.rank) ? g[k] : 0;
      elems *= b.d[k];
    }
    if (elems == 0) return;
    long st[4] = {0, 0, 0, 0};
    out.affine(g, b.ob, st);
    for (int k = 0; k < 4; ++k) b.os[k] = st[k];
    if (low) {
      int G = 0;
      for (int k = 0; k < out.split; ++k) G ^= g[k];
      const long nc = out.ncol[G], ld = (std::max<long>(nc, 1) + 31) / 32 * 32;
      const long rel = b.ob - out.blk_off[G];
      b.lob = loff[G] + rel / nc * ld + rel % nc;
      for (int k = 0; k < 4; ++k) b.los[k] = (k < out.split) ? st[k] / nc * ld : st[k];
    }
    for (size_t s = 0; s < srcs.size(); ++s) {
      const Src& sr = srcs[s];
      int gs[4];
      for (int m = 0; m < sr.L->rank; ++m) gs[m] = g[sr.map[m]];
      long sst[4] = {0, 0, 0, 0};
      sr.L->affine(gs, b.sb[s], sst);
      for (int m = 0; m < sr.L->rank; ++m) b.ss[s][sr.map[m]] += sst[m];
    }
    if (X) {
      const int gx0 = g[X->map[0]], gx1 = g[X->map[1]];
      const int gy0 = g[Y->map[0]], gy1 = g[Y->map[1]];
      if (gx0 != gx1 || gy0 != gy1) {
        b.xb = -1;
      } else {
        int gx[2] = {gx0, gx1}, gy[2] = {gy0, gy1};
        long xs[4] = {0, 0, 0, 0}, ys[4] = {0, 0, 0, 0};
        X->L->affine(gx, b.xb, xs);
        Y->L->affine(gy, b.yb, ys);
        for (int m = 0; m < 2; ++m) {
          b.xs[X->map[m]] += xs[m];
          b.ys[Y->map[m]] += ys[m];
        }
      }
    }
    b.first = first;
    first += (elems + kChunk - 1) / kChunk;
    p.host.push_back(b);
  };
  if (out.rank == 4) {
    for (int g0 = 0; g0 < nirr; ++g0)
      for (int g1 = 0; g1 < nirr; ++g1)
        for (int g2 = 0; g2 < nirr; ++g2) {
          int g[4] = {g0, g1, g2, g0 ^ g1 ^ g2};
          add_box(g);
        }
  } else {
    for (int g0 = 0; g0 < nirr; ++g0) {
      int g[4] = {g0, g0, 0, 0};
      add_box(g);
    }
  }
  p.nbox = (int)p.host.size();
  p.total = first;  // chunks
  plan_tiles(p, out.rank);
  return ACES_OK;
}

int build_plan(Plan& p, const Layout& out, const std::vector<Src>& srcs, const Src* X, const Src* Y) {
  ACES_CHECK(build_boxes(p, out, srcs, X, Y));
  free_plan(p);
  if (p.nbox > 0) {
    const long nt = (p.tq >= 0) ? p.ttotal : 0;
    ACES_REQUIRE(nt < (1L << 31), "dpd plan: too many tiles");
    ACES_HIP(hipMalloc((void**)&p.dev, sizeof(Box) * p.nbox + sizeof(int) * nt));
    ACES_HIP(hipMemcpy(p.dev, p.host.data(), sizeof(Box) * p.nbox, hipMemcpyHostToDevice));
    if (nt > 0) {
      std::vector<int> tb((size_t)nt);
      for (int k = 0; k < p.nbox; ++k) {
        const long end = (k + 1 < p.nbox) ? p.host[k + 1].tfirst : nt;
        for (long t = p.host[k].tfirst; t < end; ++t) tb[(size_t)t] = k;
      }
      int* dt = reinterpret_cast<int*>(p.dev + p.nbox);
      ACES_HIP(hipMemcpy(dt, tb.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
      p.tbox = dt;
    }
  }
  return ACES_OK;
}

void free_plan(Plan& p) {
  if (p.dev) (void)hipFree(p.dev);
  p.dev = nullptr;
  p.tbox = nullptr;
}

struct ApplyArgs {
  const Box* boxes;
  int nbox;
  double* out;
  double beta;
  const double* src[3];
  double alpha[3];
  int nsrc;
  const double* X;
  const double* Y;
  double gamma;
  Denom den;
  unsigned short* lhi;  // split image of the output (tiled walk only), or null
  unsigned short* llo;
};

__global__ __launch_bounds__(256) void apply_plan_kernel(ApplyArgs a) {
  __shared__ int sb;
  if (threadIdx.x == 0) sb = box_of_chunk(a.boxes, a.nbox, blockIdx.x);
  __syncthreads();
  const Box& b = a.boxes[sb];
  const long elems = (long)b.d[0] * b.d[1] * b.d[2] * b.d[3];
  const long e0 = (blockIdx.x - b.first) * kChunk;
  const long e1 = min(elems, e0 + kChunk);
  const int d1 = b.d[1], d2 = b.d[2], d3 = b.d[3];
  for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int l3 = (int)(e % d3);
    const long r2 = e / d3;
    const int l2 = (int)(r2 % d2);
    const long r1 = r2 / d2;
    const int l1 = (int)(r1 % d1);
    const int l0 = (int)(r1 / d1);
    double acc = 0.0;
    for (int s = 0; s < a.nsrc; ++s)
      acc += a.alpha[s] * a.src[s][b.sb[s] + l0 * b.ss[s][0] + l1 * b.ss[s][1] + l2 * b.ss[s][2] + l3 * b.ss[s][3]];
    if (a.den.pow != 0) {
      const int l[4] = {l0, l1, l2, l3};
      double D = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (a.den.sign[k] != 0) D += a.den.sign[k] * a.den.e[k][a.den.off[k][b.g[k]] + l[k]];
      acc = (a.den.pow > 0) ? acc * D : acc / D;
    }
    if (a.X != nullptr && b.xb >= 0)
      acc += a.gamma * a.X[b.xb + l0 * b.xs[0] + l1 * b.xs[1] + l2 * b.xs[2] + l3 * b.xs[3]] *
             a.Y[b.yb + l0 * b.ys[0] + l1 * b.ys[1] + l2 * b.ys[2] + l3 * b.ys[3]];
    double* o = a.out + b.ob + l0 * b.os[0] + l1 * b.os[1] + l2 * b.os[2] + l3 * b.os[3];
    *o = (a.beta == 0.0) ? acc : a.beta * (*o) + acc;
  }
}

// Tiled walk: one workgroup per tile of a box (Box::te values of output
// indices 3, tq and rb, one value of index ra).  Sources in the mask tr are
// read along tq (their unit-stride index, q fastest, then 3, then rb) into an
// LDS image of the tile and read back in the output's order; the others, the
// denominator and the product term follow the output order (3 fastest).
__device__ inline int div_by(int e, int d, unsigned m) { return d == 1 ? e : (int)__umulhi((unsigned)e, m); }

template <typename Off>
__global__ __launch_bounds__(256) void apply_plan_tiled_kernel(ApplyArgs a, int tq, int tr, const int* tbox) {
  __shared__ double tile[kTileLds];
  const Box& b = a.boxes[tbox[blockIdx.x]];
  const int ra = (tq == 0) ? 1 : 0, rb = (tq == 2) ? 1 : 2;
  const int T3 = b.te[0], TQ = b.te[1], TR = b.te[2];
  const unsigned m3 = b.mg[0], mq = b.mg[1];
  int t = (int)(blockIdx.x - b.tfirst);
  const int t3 = t % b.tn[0];
  t /= b.tn[0];
  const int tqi = t % b.tn[1];
  t /= b.tn[1];
  const int tri = t % b.tn[2];
  const int ia = t / b.tn[2];
  const int i3b = t3 * T3, iqb = tqi * TQ, irb = tri * TR;
  const int e3 = min(T3, b.d[3] - i3b), eq = min(TQ, b.d[tq] - iqb), er = min(TR, b.d[rb] - irb);
  const int P = TQ | 1;
  const int tid = threadIdx.x;
  // output walk: thread (c3, j0) takes rows j0, j0 + RP, ... (row j = q + TQ r)
  const int RP = 256 / T3;
  const int j0 = div_by(tid, T3, m3);
  const int c3 = tid - j0 * T3;
  const bool act = j0 < RP && c3 < e3;
  const int R = TQ * TR;
  double acc[kTileRows];
#pragma unroll
  for (int k = 0; k < kTileRows; ++k) acc[k] = 0.0;
  for (int s = 0; s < a.nsrc; ++s) {
    const long* st = b.ss[s];
    const double al = a.alpha[s];
    const Off s3 = (Off)st[3], sq = (Off)st[tq], sr = (Off)st[rb];
    const double* base = a.src[s] + b.sb[s] + ia * st[ra] + (long)i3b * s3 + (long)iqb * sq + (long)irb * sr;
    // Loads are branch-free (an element outside the tile reads the tile's
    // first element, which exists, and is then dropped), so that a thread's
    // loads are all in flight together.
    if ((tr >> s) & 1) {
      // read along q: thread (q, jq) takes rows jq, jq + RPq, ... (row = c3 + T3 r)
      const int RPq = 256 / TQ;
      const int jq = div_by(tid, TQ, mq);
      const int qq = tid - jq * TQ;
      const int n = T3 * TR;
      __syncthreads();  // the previous reader of the tile is done
      if (jq < RPq && qq < eq) {
        for (int jb = jq; jb < n; jb += kReadBatch * RPq) {
          double v[kReadBatch];
#pragma unroll
          for (int u = 0; u < kReadBatch; ++u) {
            const int j = jb + u * RPq;
            const int r = div_by(j, T3, m3);
            const int cc = j - r * T3;
            const bool ok = j < n && cc < e3 && r < er;
            v[u] = base[ok ? qq * sq + cc * s3 + r * sr : 0];
            if (!ok) v[u] = 0.0;
          }
#pragma unroll
          for (int u = 0; u < kReadBatch; ++u) {
            const int j = jb + u * RPq;
            if (j < n) tile[j * P + qq] = v[u];
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kTileRows; ++k) {
        const int j = j0 + RP * k;
        if (act && j < R) {
          const int r = div_by(j, TQ, mq);
          const int q = j - r * TQ;
          acc[k] += al * tile[(r * T3 + c3) * P + q];
        }
      }
    } else {
      double v[kTileRows];
#pragma unroll
      for (int k = 0; k < kTileRows; ++k) {
        const int j = j0 + RP * k;
        const int r = div_by(j, TQ, mq);
        const int q = j - r * TQ;
        const bool ok = act && j < R && q < eq && r < er;
        v[k] = base[ok ? c3 * s3 + q * sq + r * sr : 0];
        if (!ok) v[k] = 0.0;
      }
#pragma unroll
      for (int k = 0; k < kTileRows; ++k) acc[k] += al * v[k];
    }
  }
  const Off os3 = (Off)b.os[3], osq = (Off)b.os[tq], osr = (Off)b.os[rb];
  double* obase = a.out + b.ob + ia * b.os[ra] + (long)i3b * os3 + (long)iqb * osq + (long)irb * osr;
  const bool prod = a.X != nullptr && b.xb >= 0;
  const bool lsplit = a.lhi != nullptr;
  unsigned short* lhi = a.lhi;
  unsigned short* llo = a.llo;
  Off ls3 = 0, lsq = 0, lsr = 0;
  if (lsplit) {
    const long lb = b.lob + ia * b.los[ra] + (long)i3b * b.los[3] + (long)iqb * b.los[tq] + (long)irb * b.los[rb];
    lhi += lb;
    llo += lb;
    ls3 = (Off)b.los[3], lsq = (Off)b.los[tq], lsr = (Off)b.los[rb];
  }
  const double* xbase = a.X;
  const double* ybase = a.Y;
  Off xs3 = 0, xsq = 0, xsr = 0, ys3 = 0, ysq = 0, ysr = 0;
  if (prod) {
    xbase += b.xb + ia * b.xs[ra] + i3b * b.xs[3] + iqb * b.xs[tq] + irb * b.xs[rb];
    ybase += b.yb + ia * b.ys[ra] + i3b * b.ys[3] + iqb * b.ys[tq] + irb * b.ys[rb];
    xs3 = (Off)b.xs[3], xsq = (Off)b.xs[tq], xsr = (Off)b.xs[rb];
    ys3 = (Off)b.ys[3], ysq = (Off)b.ys[tq], ysr = (Off)b.ys[rb];
  }
  // denominator: the parts of indices 3 and ra are fixed for this thread
  double Dfix = 0.0;
  if (a.den.pow != 0 && act) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
      if (a.den.sign[m] != 0 && (m == 3 || m == ra))
        Dfix += a.den.sign[m] * a.den.e[m][a.den.off[m][b.g[m]] + (m == 3 ? i3b + c3 : ia)];
  }
  // the old output values and the product terms, loaded together
  double prev[kTileRows], xy[kTileRows];
#pragma unroll
  for (int k = 0; k < kTileRows; ++k) {
    const int j = j0 + RP * k;
    const int r = div_by(j, TQ, mq);
    const int q = j - r * TQ;
    const bool ok = act && j < R && q < eq && r < er;
    prev[k] = 0.0;
    xy[k] = 0.0;
    if (a.beta != 0.0) prev[k] = obase[ok ? c3 * os3 + q * osq + r * osr : 0];
    if (prod) xy[k] = xbase[ok ? c3 * xs3 + q * xsq + r * xsr : 0] * ybase[ok ? c3 * ys3 + q * ysq + r * ysr : 0];
  }
#pragma unroll
  for (int k = 0; k < kTileRows; ++k) {
    const int j = j0 + RP * k;
    if (!act || j >= R) continue;
    const int r = div_by(j, TQ, mq);
    const int q = j - r * TQ;
    if (q >= eq || r >= er) continue;
    double v = acc[k];
    if (a.den.pow != 0) {
      double D = Dfix;
#pragma unroll
      for (int m = 0; m < 3; ++m)
        if (a.den.sign[m] != 0 && m != ra)
          D += a.den.sign[m] * a.den.e[m][a.den.off[m][b.g[m]] + (m == tq ? iqb + q : irb + r)];
      v = (a.den.pow > 0) ? v * D : v / D;
    }
    if (prod) v += a.gamma * xy[k];
    double* o = obase + c3 * os3 + q * osq + r * osr;
    if (a.beta != 0.0) v += a.beta * prev[k];
    *o = v;
    if (lsplit) {
      const Off lo = c3 * ls3 + q * lsq + r * lsr;
      const unsigned short h = bf16_rn((float)v);
      lhi[lo] = h;
      llo[lo] = bf16_rn((float)(v - (double)bf16_to_f(h)));
    }
  }
}

int apply_plan(const Plan& p, double* out, double beta, const double* const* src, const double* alpha,
               const double* X, const double* Y, double gamma, const Denom* den, hipStream_t s) {
  return apply_plan_split(p, out, beta, src, alpha, X, Y, gamma, den, nullptr, nullptr, s);
}

int apply_plan_split(const Plan& p, double* out, double beta, const double* const* src, const double* alpha,
                     const double* X, const double* Y, double gamma, const Denom* den, unsigned short* hi,
                     unsigned short* lo, hipStream_t s) {
  ACES_REQUIRE(hi == nullptr || (lo != nullptr && p.low && p.tq >= 0), "dpd apply: split image needs a tiled plan");
  if (p.nbox == 0 || p.total == 0) return ACES_OK;
  ApplyArgs a;
  std::memset(&a, 0, sizeof(a));
  a.lhi = hi;
  a.llo = lo;
  a.boxes = p.dev;
  a.nbox = p.nbox;
  a.out = out;
  a.beta = beta;
  a.nsrc = p.nsrc;
  for (int k = 0; k < p.nsrc; ++k) {
    a.src[k] = src[k];
    a.alpha[k] = alpha[k];
  }
  a.X = p.prod ? X : nullptr;
  a.Y = p.prod ? Y : nullptr;
  a.gamma = gamma;
  if (den) a.den = *den;
  if (p.tq >= 0) {
    ACES_REQUIRE(p.ttotal < (1L << 31) && p.tbox != nullptr, "dpd apply: tile table missing or too many tiles");
    if (p.wide)
      hipLaunchKernelGGL(apply_plan_tiled_kernel<long>, dim3((unsigned)p.ttotal), dim3(256), 0, s, a, p.tq, p.tr,
                         p.tbox);
    else
      hipLaunchKernelGGL(apply_plan_tiled_kernel<int>, dim3((unsigned)p.ttotal), dim3(256), 0, s, a, p.tq, p.tr,
                         p.tbox);
    ACES_HIP(hipGetLastError());
    return ACES_OK;
  }
  ACES_REQUIRE(p.total < (1L << 31), "dpd apply: too many chunks");
  hipLaunchKernelGGL(apply_plan_kernel, dim3((unsigned)p.total), dim3(256), 0, s, a);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
struct BlockInfo {
  long blk_off[kMaxIrr];
  long nrow[kMaxIrr], ncol[kMaxIrr];
  long rstart[kMaxIrr + 1];
  int nirr;
};

static BlockInfo block_info(const Layout& L) {
  BlockInfo b;
  b.nirr = L.nirr;
  long r = 0;
  for (int g = 0; g < L.nirr; ++g) {
    b.blk_off[g] = L.blk_off[g];
    b.nrow[g] = L.nrow[g];
    b.ncol[g] = L.ncol[g];
    b.rstart[g] = r;
    r += L.nrow[g];
  }
  b.rstart[L.nirr] = r;
  return b;
}

__global__ void block_rowsum_kernel(BlockInfo b, const double* __restrict__ x, double* __restrict__ out,
                                    double alpha, double beta) {
  const int lane = threadIdx.x & 63;
  const long wid = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long nw = (long)gridDim.x * (blockDim.x >> 6);
  for (long r = wid; r < b.rstart[b.nirr]; r += nw) {
    int g = 0;
    while (r >= b.rstart[g + 1]) ++g;
    const long lr = r - b.rstart[g];
    const double* row = x + b.blk_off[g] + lr * b.ncol[g];
    double acc = 0.0;
    for (long c = lane; c < b.ncol[g]; c += 64) acc += row[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) out[r] = alpha * acc + (beta == 0.0 ? 0.0 : beta * out[r]);
  }
}

// Long rows: block (row, chunk) sums a contiguous piece of its row into
// part[row * nchunk + chunk]; the pieces are added in chunk order.
__global__ __launch_bounds__(256) void block_rowsum_part_kernel(BlockInfo b, const double* __restrict__ x,
                                                                double* __restrict__ part, int nchunk) {
  __shared__ double sh[4];
  const long r = blockIdx.x / nchunk;
  const int ch = (int)(blockIdx.x - r * nchunk);
  int g = 0;
  while (r >= b.rstart[g + 1]) ++g;
  const long nc = b.ncol[g];
  const double* row = x + b.blk_off[g] + (r - b.rstart[g]) * nc;
  const long len = (nc + nchunk - 1) / nchunk, c0 = ch * len, c1 = min(nc, c0 + len);
  double acc = 0.0;
  for (long c = c0 + threadIdx.x; c < c1; c += 4 * blockDim.x) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long cc = c + u * (long)blockDim.x;
      v[u] = row[cc < c1 ? cc : c0];
      if (cc >= c1) v[u] = 0.0;
    }
    acc += (v[0] + v[1]) + (v[2] + v[3]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void block_rowsum_finish_kernel(long rows, int nchunk, const double* __restrict__ part,
                                           double* __restrict__ out, double alpha, double beta) {
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  double acc = 0.0;
  for (int c = 0; c < nchunk; ++c) acc += part[r * nchunk + c];
  out[r] = alpha * acc + (beta == 0.0 ? 0.0 : beta * out[r]);
}

int block_rowsum(const Layout& L, const double* x, double* out, double alpha, double beta, hipStream_t s,
                 double* ws, long ws_elems) {
  BlockInfo b = block_info(L);
  const long rows = b.rstart[b.nirr];
  if (rows == 0) return ACES_OK;
  long widest = 0;
  for (int g = 0; g < b.nirr; ++g) widest = std::max(widest, b.ncol[g]);
  // rows of at least 16384 elements: pieces over about 4096 blocks
  const long nchunk = std::min<long>(std::max<long>(1, 4096 / rows), std::max<long>(1, widest / 4096));
  if (ws != nullptr && widest >= 16384 && nchunk > 1 && rows * nchunk <= ws_elems) {
    hipLaunchKernelGGL(block_rowsum_part_kernel, dim3((unsigned)(rows * nchunk)), dim3(256), 0, s, b, x, ws,
                       (int)nchunk);
    ACES_HIP(hipGetLastError());
    hipLaunchKernelGGL(block_rowsum_finish_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, rows,
                       (int)nchunk, ws, out, alpha, beta);
    ACES_HIP(hipGetLastError());
    return ACES_OK;
  }
  const unsigned grid = (unsigned)std::min<long>((rows + 3) / 4, 65536);
  hipLaunchKernelGGL(block_rowsum_kernel, dim3(grid), dim3(256), 0, s, b, x, out, alpha, beta);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// mode 0: x_ll += alpha e_l ; mode 1: e_l += alpha x_ll
__global__ void diag_kernel(BlockInfo b, double* __restrict__ x, double* __restrict__ e, double alpha, int mode) {
  const long n = b.rstart[b.nirr];
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
    int g = 0;
    while (r >= b.rstart[g + 1]) ++g;
    const long l = r - b.rstart[g];
    double* xd = x + b.blk_off[g] + l * b.ncol[g] + l;
    if (mode == 0)
      *xd += alpha * e[r];
    else
      e[r] += alpha * (*xd);
  }
}

static int diag_launch(const Layout& L, double* x, double* e, double alpha, int mode, hipStream_t s) {
  ACES_REQUIRE(L.rank == 2, "dpd diag: rank-2 layout required");
  BlockInfo b = block_info(L);
  const long n = b.rstart[b.nirr];
  if (n == 0) return ACES_OK;
  hipLaunchKernelGGL(diag_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, x, e, alpha, mode);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

int sub_diag_blocks(const Layout& L, double* x, const double* e, double alpha, hipStream_t s) {
  return diag_launch(L, x, const_cast<double*>(e), alpha, 0, s);
}

int diag_to_vec(const Layout& L, const double* x, double* e, double alpha, hipStream_t s) {
  return diag_launch(L, const_cast<double*>(x), e, alpha, 1, s);
}

}  // namespace dpd
}  // namespace aces
