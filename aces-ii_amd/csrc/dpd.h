// Irrep-blocked (DPD) storage of totally symmetric tensors over abelian point
// groups (D2h and subgroups: irrep product = XOR of 0-based labels).
//
// Orbitals are sorted by irrep inside the occupied and the virtual space; a
// Space records the per-irrep counts.  A tensor of rank 2 or 4 is stored as
// one dense block per row irrep G: rows are the compound of the first `split`
// indices with total irrep G, columns the compound of the rest with irrep G
// (the reference's list layout idea, acescore/aces_com_symloc.F:11-106,
// libr/inipck.F:3-33, but with its own row-major block order).  Inside a
// block, compounds are ordered by the irreps of their leading indices
// (lexicographically), then by the local orbital numbers (row major).
//
// With all four irreps fixed, every element offset is affine in the local
// orbital numbers; a "box" is such a fixed-irrep sub-block, and the permute
// kernels below walk boxes.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <vector>

namespace aces {
namespace dpd {

constexpr int kMaxIrr = 8;

struct Space {
  int nirr = 1;
  int cnt[kMaxIrr] = {0};
  int off[kMaxIrr] = {0};  // first orbital (0-based within the space) of each irrep
  int n = 0;
  static Space make(int nirr, const int* counts);
};

struct Layout {
  int rank = 0;   // 2 or 4
  int split = 1;  // number of row indices
  // dense: a plain row-major array over the full index ranges (orbitals in
  // irrep-sorted order) with element strides dstride (no blocks)
  int dense = 0;
  long dstride[4] = {0, 0, 0, 0};
  Space sp[4];
  int nirr = 1;
  long blk_off[kMaxIrr + 1] = {0};
  long nrow[kMaxIrr] = {0}, ncol[kMaxIrr] = {0};
  // compound sub-block offsets: row compound of `split` indices, column
  // compound of rank-split indices; tuple index over the leading irreps.
  long subr[kMaxIrr][kMaxIrr * kMaxIrr] = {{0}};
  long subc[kMaxIrr][kMaxIrr * kMaxIrr] = {{0}};
  long size() const { return blk_off[nirr]; }
  // Affine map for fixed index irreps g[0..rank-1] (total irrep 0):
  // offset = base + sum_k loc_k * st[k].
  __host__ __device__ void affine(const int* g, long& base, long* st) const {
    if (dense) {
      base = 0;
      for (int k = 0; k < rank; ++k) {
        base += (long)sp[k].off[g[k]] * dstride[k];
        st[k] = dstride[k];
      }
      return;
    }
    int G = 0;
    for (int k = 0; k < split; ++k) G ^= g[k];
    const long nc = ncol[G];
    long rsub = 0, csub = 0;
    if (split == 1) {
      st[0] = nc;
    } else if (split == 2) {
      rsub = subr[G][g[0]];
      st[0] = (long)sp[1].cnt[g[1]] * nc;
      st[1] = nc;
    } else {
      rsub = subr[G][g[0] * nirr + g[1]];
      st[0] = (long)sp[1].cnt[g[1]] * sp[2].cnt[g[2]] * nc;
      st[1] = (long)sp[2].cnt[g[2]] * nc;
      st[2] = nc;
    }
    const int c = rank - split;
    const int s = split;
    if (c == 1) {
      st[s] = 1;
    } else if (c == 2) {
      csub = subc[G][g[s]];
      st[s] = sp[s + 1].cnt[g[s + 1]];
      st[s + 1] = 1;
    } else {
      csub = subc[G][g[s] * nirr + g[s + 1]];
      st[s] = (long)sp[s + 1].cnt[g[s + 1]] * sp[s + 2].cnt[g[s + 2]];
      st[s + 1] = sp[s + 2].cnt[g[s + 2]];
      st[s + 2] = 1;
    }
    base = blk_off[G] + rsub * nc + csub;
  }
  // Offset of the element with irreps g and local indices l.
  __host__ __device__ long offset(const int* g, const int* l) const {
    long base, st[4] = {0, 0, 0, 0};
    affine(g, base, st);
    long o = base;
    for (int k = 0; k < rank; ++k) o += (long)l[k] * st[k];
    return o;
  }
  bool same(const Layout& o) const;
};

Layout make_layout(int rank, int split, const Space* s0, const Space* s1, const Space* s2 = nullptr,
                   const Space* s3 = nullptr);
// Dense row-major array (leading dimensions given as element strides).
Layout make_dense(int rank, const Space* const* sp, const long* strides);

// Packed pair spaces for the particle-particle ladder: (p<=q) or (p<q) over
// one Space, blocked by pair irrep.  G = 0: per irrep g a triangle of pairs
// inside g (q_loc*(q_loc+1)/2 + p_loc, resp. q_loc*(q_loc-1)/2 + p_loc);
// G != 0: per irrep pair (gp < gq, gp^gq = G) a rectangle p_loc*cnt[gq] + q_loc.
struct PairSpace {
  bool strict = false;  // p < q
  int nirr = 1;
  long n[kMaxIrr] = {0};               // pairs per irrep G
  long sub[kMaxIrr][kMaxIrr] = {{0}};  // offset of sub-block (G, gp) inside block G
};
PairSpace make_pairs(const Space& s, bool strict);

// Index of the pair (p, q), p before q in the global orbital order, inside
// its block G = gp ^ gq.
__host__ __device__ inline long pair_index(const PairSpace& P, const Space& S, int gp, int lp, int gq, int lq) {
  if (gp == gq) {
    const long t = P.strict ? (long)lq * (lq - 1) / 2 + lp : (long)lq * (lq + 1) / 2 + lp;
    return P.sub[0][gp] + t;
  }
  return P.sub[gp ^ gq][gp] + (long)lp * S.cnt[gq] + lq;
}

// Inverse of pair_index inside block G.
__host__ __device__ inline void pair_decode(const PairSpace& P, const Space& S, int G, long idx, int& gp, int& lp,
                                            int& gq, int& lq) {
  gp = gq = lp = lq = 0;
  for (int g = 0; g < P.nirr; ++g) {
    const int h = g ^ G;
    long sz;
    if (G == 0) {
      const long c = S.cnt[g];
      sz = P.strict ? c * (c - 1) / 2 : c * (c + 1) / 2;
    } else {
      sz = (g < h) ? (long)S.cnt[g] * S.cnt[h] : 0;
    }
    const long b = P.sub[G][g];
    if (idx >= b && idx < b + sz) {
      const long t = idx - b;
      gp = g;
      gq = h;
      if (G == 0) {
        int q;
        if (P.strict) {
          q = (int)((sqrt(8.0 * (double)t + 1.0) + 1.0) * 0.5);
          while ((long)q * (q - 1) / 2 > t) --q;
          while ((long)(q + 1) * q / 2 <= t) ++q;
          lp = (int)(t - (long)q * (q - 1) / 2);
        } else {
          q = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
          while ((long)q * (q + 1) / 2 > t) --q;
          while ((long)(q + 1) * (q + 2) / 2 <= t) ++q;
          lp = (int)(t - (long)q * (q + 1) / 2);
        }
        lq = q;
      } else {
        lp = (int)(t / S.cnt[h]);
        lq = (int)(t - (long)lp * S.cnt[h]);
      }
      return;
    }
  }
}

// ---------------------------------------------------------------------------
// Box plans: out[...] = beta*out + scale(D) * sum_s alpha_s * src_s[...] + gamma * X[...] * Y[...]
struct Box {
  int d[4];
  int g[4];
  long first;   // first element chunk (prefix over boxes)
  long tfirst;  // first tile of the tiled walk (prefix over boxes)
  // Tiled walk: this box's tile extents along output indices 3, tq and rb
  // (te), the number of tiles along each (tn), and the reciprocals that
  // divide an element number by te[0] and te[1] (umulhi magic, 0 for 1).
  int te[3], tn[3];
  unsigned mg[2];
  long ob, os[4];
  // the same output elements in the bf16 split image of the output (rows of
  // each irrep block padded to a multiple of 32 columns), Plan::low only
  long lob, los[4];
  long sb[3], ss[3][4];
  long xb, xs[4], yb, ys[4];  // xb < 0: product term vanishes in this box
};

constexpr long kChunk = 4096;  // elements per workgroup in box walks
// Tiled walk: a tile spans at most kTileSpan along index 3 and along tq, a
// thread of its 256 takes at most kTileRows output elements, and the tile's
// LDS image (pitch tq-extent | 1) fits kTileLds doubles.
constexpr int kTileSpan = 64, kTileRows = 8, kTileLds = 2304;
// tiled walk when the largest box spans at least this along the output's
// fastest index (shorter rows take the element walk)
constexpr int kTileMinFast = 4;
// turned sources: LDS-image loads per thread issued together
constexpr int kReadBatch = 8;

// Box of workgroup `chunk` (binary search over Box::first).
__device__ inline int box_of_chunk(const Box* boxes, int nbox, long chunk) {
  int lo = 0, hi = nbox - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (boxes[mid].first <= chunk)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

struct Plan {
  std::vector<Box> host;
  Box* dev = nullptr;
  int nbox = 0;
  long total = 0;
  int nsrc = 0;
  bool prod = false;
  // Tiled walk (rank 4): tq = the output index (0..2) along which the sources
  // in the mask tr have unit stride; they are read along it and turned round
  // in LDS.  tq < 0: element walk.  ttotal: tiles over all boxes.
  int tq = -1;
  int tr = 0;
  long ttotal = 0;
  // box of every tile (device, stored after the boxes), and whether a tile's
  // element offsets relative to its first element need 64 bits
  const int* tbox = nullptr;
  bool wide = false;
  // Box::lob/los hold the split-image map (blocked rank-4 outputs)
  bool low = false;
};

// Source description for plan building: layout and, for each source
// dimension m, the output dimension that supplies its index.
struct Src {
  const Layout* L = nullptr;
  int map[4] = {0, 1, 2, 3};
};

// Build a plan for output layout `out` with up to three sources and an
// optional product X*Y (rank-2 operands).  Sources with total irrep rules
// violated in a box are simply zero there (never happens for totally
// symmetric operands that map one to one).
int build_plan(Plan& p, const Layout& out, const std::vector<Src>& srcs, const Src* X, const Src* Y);
void free_plan(Plan& p);
// Host part of build_plan (boxes only, nothing uploaded).
int build_boxes(Plan& p, const Layout& out, const std::vector<Src>& srcs, const Src* X, const Src* Y);
// For a rank-4 plan whose largest box spans at least kTileMinFast along the output's
// fastest index: the index (0..2) along which source s has unit stride and
// that a tile can span, 3 when the source reads well in the output's order,
// -1 when the plan cannot be tiled at all.
int source_fast_index(const Plan& p, int rank, int s);

// Denominator scaling: D = sum_k dsign[k] * e(space of dim k)[global index],
// e = eo for occupied dims, ev for virtual dims.
struct Denom {
  int pow = 0;  // 0: no scaling, 1: times D, -1: divided by D
  int sign[4] = {0, 0, 0, 0};
  const double* e[4] = {nullptr, nullptr, nullptr, nullptr};
  int off[4][kMaxIrr] = {{0}};  // orbital offset of each irrep in the space of dim k
};

int apply_plan(const Plan& p, double* out, double beta, const double* const* src, const double* alpha,
               const double* X, const double* Y, double gamma, const Denom* den, hipStream_t s);
// The same, also writing every output value's bf16 split (hi = bf16(v), lo =
// bf16(v - hi)) to its place in the split image hi/lo: each irrep block G at
// element offset sum_{g<G} nrow[g] * pad32(ncol[g]), rows padded to
// pad32(ncol[G]) columns (the pad columns are left as they are).  Tiled plans
// with Plan::low only.
int apply_plan_split(const Plan& p, double* out, double beta, const double* const* src, const double* alpha,
                     const double* X, const double* Y, double gamma, const Denom* den, unsigned short* hi,
                     unsigned short* lo, hipStream_t s);

// Per-row sums of every block of a layout: out[r] for the rows in block
// order (for rank-2 (O|O) this is the orbital order).
// ws (ws_elems doubles, optional): partial sums of long rows, split over
// many workgroups and added in a fixed order.
int block_rowsum(const Layout& L, const double* x, double* out, double alpha, double beta, hipStream_t s,
                 double* ws = nullptr, long ws_elems = 0);
// X -= diag(e) on a rank-2 layout with both indices in the same space.
int sub_diag_blocks(const Layout& L, double* x, const double* e, double alpha, hipStream_t s);
// e += alpha * diag(X) (the adjoint of sub_diag_blocks).
int diag_to_vec(const Layout& L, const double* x, double* e, double alpha, hipStream_t s);

}  // namespace dpd
}  // namespace aces
