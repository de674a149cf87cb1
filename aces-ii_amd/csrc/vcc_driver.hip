// xvcc driver entry points over the device list store (the drop-in boundary).
//
// ACES II's CC loop (vcc/vcc.F:811-1085) calls, in order, INITIN, GENINT,
// FEACONT, FMICONT, FMECONT, T1INT2A, T1INT2B, T1INT1, DRE3EN, E4S, NEWT2,
// CMPENG, DODIIS0, DRMOVE, RNABIJ; each reads lists and accumulates into
// increment lists.  Here the same entry points (Fortran linkage, INTEGER*8
// arguments by reference, trailing underscore) act on the device list store:
//
//   GENINT   builds the engine from the integral lists on first use, loads T
//            from lists 46 (T2 AB) and 90 group 1 (T1) and evaluates every
//            CCSD increment term in one pass (the W/F intermediates of GENINT
//            and the contributions FEACONT..DRE3EN and E4S add are fused).
//   FEACONT, FMICONT, FMECONT, T1INT2A, T1INT2B, T1INT1, INITIN, INITSN, RNABIJ
//            no separate work (their terms are part of the fused increment).
//   DRE3EN   writes the T2 increment to list 63 (AB) and 61 (AA).
//   E4S      writes T1_new = T1 increment / D1 to list 90 group 3.
//   NEWT2    list 63 <- list 63 / D (libr2/newt2.F:233-244), list 61 <- ABTOAA(63).
//   CMPENG   E(AA), E(AB), the total and the T2-only (linear) energy from lists
//            NLIST2+1..+3 and T1 group NLIST1+1 (libr/cmpeng.F:99-180).
//   DODIIS0  DIIS on (lists 44/46/90 g1) vs (61/63/90 g3); extrapolated vector
//            written to the increment lists as in vcc/dodiis0.F.  Only xvcc's
//            list arguments (44, 61, 90, 0, 90, 2, 70) are implemented; any
//            other set ends the process through the fatal path.
//   DRMOVE   lists 61/63 -> 44+NLIST/46+NLIST and, with SING1, list 90 group 3
//            -> list 90+NLIST group 1 (libr/drmove.f).
//
// Integral lists read by GENINT (RHF, C1; types from intprc/clmoio.F):
//   13 <Mn|Ij> (14,14), 16 <Ab|Ij> (13,14), 10 <Ij|Ka> (14,18),
//   30 <Ab|Ci> (13,11), 233 <Ab|Cd> (5,13), 23 <AJ||BI> ordered (AI,BJ) (9,9);
// orbital energies from record SCFEVALA, the MO Fock matrix from record FOCKMOA
// when present (non-HF), otherwise diag(SCFEVALA).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#include "aces_files.h"
#include "ccsd_engine.h"
#include "ccsd_kernels.h"
#include "common.h"
#include "dpd_lists.h"
#include "intprc_df.h"
#include "liststore.h"
#include "sym_engine.h"
#include "tensor_ops.h"

namespace aces {

// Dense amplitude-space buffers of the irrep-blocked path (nirrep > 1): the
// symmetry-packed amplitude lists are moved through C1 order, in which the
// irrep-blocked engine takes and returns its amplitudes.
struct SymBuffers {
  double* base = nullptr;
  double *t1 = nullptr, *t2 = nullptr, *t1n = nullptr, *t2n = nullptr, *inc1 = nullptr, *inc2 = nullptr;
  double *x1 = nullptr, *x2 = nullptr, *aa = nullptr, *l16 = nullptr, *eo = nullptr, *ev = nullptr;
  ~SymBuffers() {
    if (base) (void)hipFree(base);
  }
};

struct AcesContext {
  hipStream_t stream = nullptr;
  std::unique_ptr<ListStore> lists;
  RecordStore recs;
  std::unique_ptr<CcsdEngine> eng;   // C1 lists (nirrep = 1)
  std::unique_ptr<SymEngine> seng;   // symmetry-packed lists (nirrep > 1)
  SymBuffers sb;
  int o = 0, v = 0;
  int prec = GP_F64;
  int last_cycle = 0;
  double* fov = nullptr;    // f(i,a) of the MO Fock matrix (zero for HF), [i][a]
  double* epart = nullptr;  // CMPENG partial sums: o^2 blocks x 4, then the 5 totals
  bool ready() const { return eng != nullptr || seng != nullptr; }
  ~AcesContext() {
    eng.reset();
    seng.reset();
    if (fov) (void)hipFree(fov);
    if (epart) (void)hipFree(epart);
  }
};

static AcesContext* g_ctx = nullptr;

static int ctx(AcesContext** c) {
  if (!g_ctx) {
    set_error("aces_init has not been called");
    return ACES_ERR_STATE;
  }
  *c = g_ctx;
  return ACES_OK;
}

// AA (A<B, I<J) list 44/61 from the AB amplitudes t2[i][j][a][b]: AA = t_ijab - t_jiab.
__global__ void abtoaa_kernel(int o, int v, const double* __restrict__ t2, double* __restrict__ aa) {
  const long nom = (long)o * (o - 1) / 2, nvm = (long)v * (v - 1) / 2;
  const long vv = (long)v * v;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < nom * nvm;
       idx += (long)gridDim.x * blockDim.x) {
    const long cij = idx / nvm, rab = idx - cij * nvm;
    int j = (int)((sqrt(8.0 * (double)cij + 1.0) + 1.0) * 0.5);
    while ((long)j * (j - 1) / 2 > cij) --j;
    while ((long)(j + 1) * j / 2 <= cij) ++j;
    const int i = (int)(cij - (long)j * (j - 1) / 2);
    int b = (int)((sqrt(8.0 * (double)rab + 1.0) + 1.0) * 0.5);
    while ((long)b * (b - 1) / 2 > rab) --b;
    while ((long)(b + 1) * b / 2 <= rab) ++b;
    const int a = (int)(rab - (long)b * (b - 1) / 2);
    aa[idx] = t2[((long)i * o + j) * vv + (long)a * v + b] - t2[((long)j * o + i) * vv + (long)a * v + b];
  }
}

static int abtoaa(int o, int v, const double* t2, double* aa, hipStream_t s) {
  const long n = (long)o * (o - 1) / 2 * ((long)v * (v - 1) / 2);
  if (n == 0) return ACES_OK;
  long g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(abtoaa_kernel, dim3((unsigned)g), dim3(256), 0, s, o, v, t2, aa);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// CMPENG pieces (libr/cmpeng.F:99-180 with libr/tener.F, libr/ftau.f) on the
// closed-shell lists, one block per (p, q) pair of occupied indices.  At memory
// index m = ((p o + q) v + r) v + s the T2 list (46 or 63) holds t_pq^rs and
// list 16 <rs|pq> = (pr|qs); the same-spin amplitudes are t_pq^rs - t_qp^rs
// (ABTOAA) and <RS||PQ> = <rs|pq> - <sr|pq>.  part[4 (p o + q) + k]:
//   k = 0  E(AB) = sum tau(Ab,Ij) <Ab|Ij>,  tau = t2 + t1 t1   1  its T2-only part
//   k = 2  E(AA) = sum_{A<B, I<J} tau(AB,IJ) <AB||IJ> (a quarter of the unrestricted
//          sum),  tau = tAA + t1(A,I) t1(B,J) - t1(A,J) t1(B,I)  3  its T2-only part
__global__ void cmpeng_partial_kernel(int o, int v, const double* __restrict__ t1, const double* __restrict__ t2,
                                      const double* __restrict__ w, double* __restrict__ part) {
  const int p = blockIdx.x / o, q = blockIdx.x - p * o;
  const long vv = (long)v * v;
  const double* tpq = t2 + ((long)p * o + q) * vv;
  const double* tqp = t2 + ((long)q * o + p) * vv;
  const double* wpq = w + ((long)p * o + q) * vv;
  const double* t1p = t1 + (long)p * v;
  const double* t1q = t1 + (long)q * v;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (long m = threadIdx.x; m < vv; m += blockDim.x) {
    const int r = (int)(m / v), s = (int)(m - (long)r * v);
    const double t = tpq[m], wab = wpq[m], wba = wpq[(long)s * v + r];
    const double taa = t - tqp[m];
    const double ss = t1p[r] * t1q[s];
    const double waa = wab - wba;
    s0 += (t + ss) * wab;
    s1 += t * wab;
    s2 += (taa + ss - t1q[r] * t1p[s]) * waa;
    s3 += taa * waa;
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  red[2][threadIdx.x] = s2;
  red[3][threadIdx.x] = s3;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4L * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Sums the block partials in block order (deterministic) and the non-HF singles
// term sum t1(i,a) f(i,a): out[0..3] as the partials (AA ones quartered), out[4]
// the singles term.
__global__ void cmpeng_final_kernel(int nblk, int ov, const double* __restrict__ part,
                                    const double* __restrict__ t1, const double* __restrict__ fov,
                                    double* __restrict__ out) {
  __shared__ double red[5][256];
  double acc[5] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
    for (int k = 0; k < 4; ++k) acc[k] += part[4L * b + k];
  for (int x = threadIdx.x; x < ov; x += blockDim.x) acc[4] += t1[x] * fov[x];
  for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = red[0][0];
    out[1] = red[1][0];
    out[2] = 0.25 * red[2][0];
    out[3] = 0.25 * red[3][0];
    out[4] = red[4][0];
  }
}

static int require_list(AcesContext* c, int list, long rows, long cols, const char* what) {
  long r = 0, k = 0;
  if (c->lists->rows(1, list, &r) != ACES_OK || c->lists->cols(1, list, &k) != ACES_OK || r != rows ||
      k != cols) {
    set_error(std::string("list ") + std::to_string(list) + " (" + what + ") missing or of wrong shape: have " +
              std::to_string(r) + "x" + std::to_string(k) + ", need " + std::to_string(rows) + "x" +
              std::to_string(cols));
    return ACES_ERR_RANGE;
  }
  return ACES_OK;
}

// Dense integral blocks (the engines' set-up inputs) from the integral lists in
// C1 order (device pointers; see the list contract at the top of this file).
struct DenseBlocks {
  double *f = nullptr, *ovov = nullptr, *oovv = nullptr, *ovoo = nullptr, *oooo = nullptr, *ovvv = nullptr;
  ~DenseBlocks() {
    for (double* p : {f, ovov, oovv, ovoo, oooo, ovvv})
      if (p) (void)hipFree(p);
  }
};

static int dense_blocks(int o, int v, const double* L16, const double* L13, const double* L10, const double* L30,
                        const double* L23, hipStream_t s, DenseBlocks& b) {
  const long oo = (long)o * o, vv = (long)v * v, ov = (long)o * v;
  ACES_HIP(hipMalloc(&b.ovov, sizeof(double) * ov * ov));
  ACES_HIP(hipMalloc(&b.oovv, sizeof(double) * ov * ov));
  ACES_HIP(hipMalloc(&b.ovoo, sizeof(double) * ov * oo));
  ACES_HIP(hipMalloc(&b.oooo, sizeof(double) * oo * oo));
  ACES_HIP(hipMalloc(&b.ovvv, sizeof(double) * ov * vv));
  auto perm = [&](double* out, long d0, long d1, long d2, long d3, const double* a, double aa, const long* sa,
                  const double* bb, double ab, const long* sbb) {
    Permute4 q;
    q.d[0] = d0; q.d[1] = d1; q.d[2] = d2; q.d[3] = d3;
    q.out = out;
    q.in[0] = a; q.alpha[0] = aa;
    for (int k = 0; k < 4; ++k) q.st[0][k] = sa[k];
    if (bb) {
      q.in[1] = bb; q.alpha[1] = ab;
      for (int k = 0; k < 4; ++k) q.st[1][k] = sbb[k];
    }
    return permute4(q, s);
  };
  // ovov[i][a][j][b] = <ab|ij>, list 16 element a + v b + v^2 (i + o j)
  { const long st[4] = {vv, 1, vv * o, v}; ACES_CHECK(perm(b.ovov, o, v, o, v, L16, 1.0, st, nullptr, 0, nullptr)); }
  // oooo[m][i][n][j] = <mn|ij>, list 13 element m + o n + o^2 (i + o j)
  { const long st[4] = {1, oo, o, oo * o}; ACES_CHECK(perm(b.oooo, o, o, o, o, L13, 1.0, st, nullptr, 0, nullptr)); }
  // ovoo[j][a][i][k] = <ij|ka>, list 10 element i + o j + o^2 (k + o a)
  { const long st[4] = {o, oo * o, 1, oo}; ACES_CHECK(perm(b.ovoo, o, v, o, o, L10, 1.0, st, nullptr, 0, nullptr)); }
  // ovvv[i][b][a][c] = <ab|ci>, list 30 element a + v b + v^2 (c + v i)
  { const long st[4] = {vv * v, v, 1, vv}; ACES_CHECK(perm(b.ovvv, o, v, v, v, L30, 1.0, st, nullptr, 0, nullptr)); }
  // oovv[i][j][a][b] = (ij|ab) = <AJ||BI>(AI,BJ) + (ai|bj); list 23 element a + v i + ov (b + v j)
  {
    const long st[4] = {v, ov * v, 1, ov};
    const long so[4] = {ov * v, v, ov, 1};  // ovov[i][a][j][b] over output (i, j, a, b)
    ACES_CHECK(perm(b.oovv, o, o, v, v, L23, 1.0, st, b.ovov, 1.0, so));
  }
  return ACES_OK;
}

// The MO Fock matrix (host, n x n, orbitals in list order: occupied then
// virtual, irrep-sorted in each) from record FOCKMOA, else diag(SCFEVALA); and
// f(i,a) plus the CMPENG scratch on the device.
static int fock_from_records(AcesContext* c, int o, int v, std::vector<double>& f) {
  const int n = o + v;
  f.assign((size_t)n * n, 0.0);
  long len = (long)n * n;
  if (c->recs.has("FOCKMOA")) {
    ACES_CHECK(c->recs.getrec(1, "FOCKMOA", &len, f.data()));
  } else {
    std::vector<double> e(n);
    long ne = n;
    ACES_CHECK(c->recs.getrec(1, "SCFEVALA", &ne, e.data()));
    for (int p = 0; p < n; ++p) f[(size_t)p * n + p] = e[p];
  }
  const long ov = (long)o * v, oo = (long)o * o;
  std::vector<double> fo((size_t)ov);
  for (int i = 0; i < o; ++i)
    for (int a = 0; a < v; ++a) fo[(size_t)i * v + a] = f[(size_t)i * n + o + a];
  if (c->fov) (void)hipFree(c->fov);
  if (c->epart) (void)hipFree(c->epart);
  c->fov = nullptr;
  c->epart = nullptr;
  ACES_HIP(hipMalloc(&c->fov, sizeof(double) * ov));
  ACES_HIP(hipMalloc(&c->epart, sizeof(double) * (4 * oo + 8)));
  ACES_HIP(hipMemcpyAsync(c->fov, fo.data(), sizeof(double) * ov, hipMemcpyHostToDevice, c->stream));
  ACES_HIP(hipStreamSynchronize(c->stream));
  return ACES_OK;
}

// Build the engine from the integral lists (C1 RHF).
static int setup_engine_c1(AcesContext* c) {
  const DpdSym& sy = c->lists->sym();
  const int o = sy.pop[0][0], v = sy.vrt[0][0], n = o + v;
  ACES_REQUIRE(o >= 1 && v >= 1, "need occupied and virtual orbitals");
  const long oo = (long)o * o, vv = (long)v * v, ov = (long)o * v;
  ACES_CHECK(require_list(c, 13, oo, oo, "<Mn|Ij>"));
  ACES_CHECK(require_list(c, 16, vv, oo, "<Ab|Ij>"));
  ACES_CHECK(require_list(c, 10, oo, ov, "<Ij|Ka>"));
  ACES_CHECK(require_list(c, 30, vv, ov, "<Ab|Ci>"));
  ACES_CHECK(require_list(c, 233, (long)v * (v + 1) / 2, vv, "<Ab|Cd>"));
  ACES_CHECK(require_list(c, 23, ov, ov, "<AJ||BI>"));
  hipStream_t s = c->stream;
  std::vector<double> f;
  ACES_CHECK(fock_from_records(c, o, v, f));
  DenseBlocks b;
  ACES_HIP(hipMalloc(&b.f, sizeof(double) * n * n));
  ACES_HIP(hipMemcpyAsync(b.f, f.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, s));
  ACES_CHECK(dense_blocks(o, v, c->lists->device_ptr(1, 16), c->lists->device_ptr(1, 13), c->lists->device_ptr(1, 10),
                          c->lists->device_ptr(1, 30), c->lists->device_ptr(1, 23), s, b));
  c->eng.reset(new CcsdEngine(o, v, s));
  const int rc = c->eng->setup(b.f, b.ovov, b.oovv, b.ovoo, b.oooo, b.ovvv, c->lists->device_ptr(1, 233));
  if (rc != ACES_OK) {
    c->eng.reset();
    return rc;
  }
  c->o = o;
  c->v = v;
  return ACES_OK;
}

// Build the irrep-blocked engine from symmetry-packed integral lists (nirrep
// 2, 4 or 8; pair types as intprc/clmoio.F writes them for RHF): each list goes
// through C1 order into the same dense blocks as the C1 path.
static int setup_engine_sym(AcesContext* c) {
  const DpdSym& sy = c->lists->sym();
  int occ[8], vir[8], o = 0, v = 0;
  for (int h = 0; h < sy.nirrep; ++h) {
    occ[h] = sy.pop[0][h];
    vir[h] = sy.vrt[0][h];
    ACES_REQUIRE(sy.pop[1][h] == occ[h] && sy.vrt[1][h] == vir[h], "the list-fed CCSD engine takes RHF lists");
    o += occ[h];
    v += vir[h];
  }
  ACES_REQUIRE(o >= 1 && v >= 1, "need occupied and virtual orbitals");
  const int n = o + v;
  const long ov = (long)o * v;
  hipStream_t s = c->stream;
  struct In { int list, tl, tr; const char* what; };
  const In ins[6] = {{13, 14, 14, "<Mn|Ij>"}, {16, 13, 14, "<Ab|Ij>"}, {10, 14, 18, "<Ij|Ka>"},
                     {30, 13, 11, "<Ab|Ci>"}, {23, 9, 9, "<AJ||BI>"}, {233, 5, 13, "<Ab|Cd>"}};
  double* c1[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  struct Free {
    double** p;
    ~Free() {
      for (int k = 0; k < 6; ++k)
        if (p[k]) (void)hipFree(p[k]);
    }
  } release{c1};
  for (int k = 0; k < 6; ++k) {
    int tl = 0, tr = 0;
    ACES_REQUIRE(c->lists->isytyp(ins[k].list, &tl, &tr) == ACES_OK && tl == ins[k].tl && tr == ins[k].tr,
                 std::string("list ") + std::to_string(ins[k].list) + " (" + ins[k].what + ") missing or not of pair types (" +
                     std::to_string(ins[k].tl) + "," + std::to_string(ins[k].tr) + ")");
    const long n1 = c1_pairs(sy, tl) * c1_pairs(sy, tr);
    ACES_HIP(hipMalloc(&c1[k], sizeof(double) * n1));
    ACES_CHECK(dpd_to_c1(*c->lists, ins[k].list, tl, tr, 1, c1[k], s));
  }
  std::vector<double> f;
  ACES_CHECK(fock_from_records(c, o, v, f));
  DenseBlocks b;
  ACES_HIP(hipMalloc(&b.f, sizeof(double) * n * n));
  ACES_HIP(hipMemcpyAsync(b.f, f.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, s));
  ACES_CHECK(dense_blocks(o, v, c1[1], c1[0], c1[2], c1[3], c1[4], s, b));
  // amplitude-space buffers; list 16 kept in C1 order for CMPENG
  SymBuffers& B = c->sb;
  const long n2 = ov * ov, naa = (long)o * (o - 1) / 2 * ((long)v * (v - 1) / 2);
  const long total = 5 * n2 + 4 * ov + naa + o + v;
  if (B.base) (void)hipFree(B.base);
  B.base = nullptr;
  ACES_HIP(hipMalloc(&B.base, sizeof(double) * total));
  double* p = B.base;
  for (double** q : {&B.t2, &B.t2n, &B.inc2, &B.x2, &B.l16}) {
    *q = p;
    p += n2;
  }
  for (double** q : {&B.t1, &B.t1n, &B.inc1, &B.x1}) {
    *q = p;
    p += ov;
  }
  B.aa = p;
  p += naa;
  B.eo = p;
  p += o;
  B.ev = p;
  ACES_HIP(hipMemcpyAsync(B.l16, c1[1], sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
  {
    std::vector<double> e(n);
    for (int q = 0; q < n; ++q) e[q] = f[(size_t)q * n + q];
    ACES_HIP(hipMemcpyAsync(B.eo, e.data(), sizeof(double) * o, hipMemcpyHostToDevice, s));
    ACES_HIP(hipMemcpyAsync(B.ev, e.data() + o, sizeof(double) * v, hipMemcpyHostToDevice, s));
  }
  c->seng.reset(new SymEngine(sy.nirrep, occ, vir, s));
  const int rc = c->seng->setup(b.f, b.ovov, b.oovv, b.ovoo, b.oooo, b.ovvv, c1[5]);
  ACES_HIP(hipStreamSynchronize(s));
  if (rc != ACES_OK) {
    c->seng.reset();
    return rc;
  }
  c->o = o;
  c->v = v;
  return ACES_OK;
}

static int setup_engine(AcesContext* c) {
  return c->lists->sym().nirrep == 1 ? setup_engine_c1(c) : setup_engine_sym(c);
}

static int ensure_amp_lists(AcesContext* c) {
  // T1 vector: the (a,i) pairs of irrep 1 (pair type 9), o v elements for C1
  const long nt1 = c->lists->sym().irpdpd(0, 9);
  long r;
  // list 46 / 63: (13,14) AB; 44 / 61: (1,3) AA; 90 groups 1..4: T1 and increments
  if (c->lists->rows(1, 46, &r) != ACES_OK) ACES_CHECK(c->lists->inipck(1, 13, 14, 46));
  if (c->lists->rows(1, 63, &r) != ACES_OK) ACES_CHECK(c->lists->inipck(1, 13, 14, 63));
  if (c->lists->rows(1, 44, &r) != ACES_OK) ACES_CHECK(c->lists->inipck(1, 1, 3, 44));
  if (c->lists->rows(1, 61, &r) != ACES_OK) ACES_CHECK(c->lists->inipck(1, 1, 3, 61));
  for (int g = 1; g <= 4; ++g)
    if (c->lists->rows(g, 90, &r) != ACES_OK) ACES_CHECK(c->lists->updmoi(1, nt1, g, 90));
  return ACES_OK;
}

// Irrep-blocked path: amplitude lists <-> the dense buffers.
// T2 AB list (13,14) in C1 order is t2[j][i][b][a] = t2[i][j][a][b] (pair symmetry).
static int sym_get_t(AcesContext* c, int g90, int l2, double* t1, double* t2) {
  if (t1) ACES_CHECK(t1_to_dense(*c->lists, g90, 90, t1, c->stream));
  if (t2) ACES_CHECK(dpd_to_c1(*c->lists, l2, 13, 14, 1, t2, c->stream));
  return ACES_OK;
}

// T2 (dense AB) -> AB list l2 and its same-spin list laa = ABTOAA (l2 < 0: AA only)
static int sym_put_t2(AcesContext* c, int l2, int laa, const double* t2) {
  if (l2 > 0) ACES_CHECK(c1_to_dpd(*c->lists, l2, 13, 14, 1, t2, c->stream));
  if (laa > 0) {
    ACES_CHECK(abtoaa(c->o, c->v, t2, c->sb.aa, c->stream));
    ACES_CHECK(c1_to_dpd(*c->lists, laa, 1, 3, 1, c->sb.aa, c->stream));
  }
  return ACES_OK;
}

}  // namespace aces

using namespace aces;

extern "C" {

// ---- runtime (aces_init / aces_fin, acescore/aces_init.F:26) ----
int aces_init(void* stream) {
  if (g_ctx) return ACES_OK;
  g_ctx = new AcesContext();
  g_ctx->stream = (hipStream_t)stream;
  g_ctx->lists.reset(new ListStore(g_ctx->stream));
  return ACES_OK;
}

int aces_fin(void) {
  delete g_ctx;
  g_ctx = nullptr;
  return ACES_OK;
}

int aces_set_symmetry(int nirrep, const int* popa, const int* popb, const int* vrta, const int* vrtb) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->lists->set_symmetry(nirrep, popa, popb, vrta, vrtb);
}

long aces_irpdpd(int irrep, int type) {
  if (!g_ctx) return -1;
  return g_ctx->lists->sym().irpdpd(irrep - 1, type);
}

long aces_isymoff(int irrep_q, int irrep_pq, int type) {
  if (!g_ctx) return -1;
  const long off = g_ctx->lists->sym().isymoff(irrep_q - 1, irrep_pq - 1, type);
  return off < 0 ? -1 : off + 1;
}

// ---- list interface (C names) ----
int aces_inipck(int irrepx, int typel, int typer, int list) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->lists->inipck(irrepx, typel, typer, list);
}
int aces_updmoi(long ncols, long nrows, int group, int list) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->lists->updmoi(ncols, nrows, group, list);
}
int aces_list_shape(int group, int list, long* nrows, long* ncols) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_CHECK(c->lists->rows(group, list, nrows));
  return c->lists->cols(group, list, ncols);
}
int aces_list_isytyp(int list, int* typel, int* typer) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->lists->isytyp(list, typel, typer);
}
int aces_getlst(double* dest, long ifirst, long ncols, int group, int list, int dest_on_device) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->lists->getlst(dest, ifirst, ncols, group, list, dest_on_device != 0);
}
int aces_putlst(const double* src, long ifirst, long ncols, int group, int list, int src_on_device) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->lists->putlst(src, ifirst, ncols, group, list, src_on_device != 0);
}
int aces_getall(double* dest, long length, int irrepx, int list, int on_device) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->lists->getall(dest, length, irrepx, list, on_device != 0);
}
int aces_putall(const double* src, long length, int irrepx, int list, int on_device) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->lists->putall(src, length, irrepx, list, on_device != 0);
}
int aces_sumsym2(const double* add, int group, int list, int on_device) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->lists->sumsym2(add, group, list, on_device != 0);
}
int aces_list_c1(int list, int typel, int typer, int irrepx, double* d_c1, int to_c1) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(d_c1 != nullptr, "aces_list_c1: null buffer");
  const int rc = to_c1 ? dpd_to_c1(*c->lists, list, typel, typer, irrepx, d_c1, c->stream)
                       : c1_to_dpd(*c->lists, list, typel, typer, irrepx, d_c1, c->stream);
  ACES_CHECK(rc);
  ACES_HIP(hipStreamSynchronize(c->stream));
  return ACES_OK;
}

double* aces_list_device_ptr(int group, int list) { return g_ctx ? g_ctx->lists->device_ptr(group, list) : nullptr; }

// ---- record interface (C names) ----
// ---- the job archive and list files on disk (ACES_JA_INIT / ACES_JA_FIN,
// ACES_IO_INIT / ACES_IO_FIN: acescore/aces_ja_init.F, aces_ja_fin.F,
// aces_io_init.F, aces_io_fin.F) ----
int aces_jobarc_read(const char* dir) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return jobarc_read(c->recs, dir ? dir : "");
}

int aces_jobarc_write(const char* dir) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return jobarc_write(c->recs, dir ? dir : "");
}

int aces_lists_read(const char* dir) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(!c->ready(), "aces_lists_read: lists are read before the engine is built (GENINT)");
  return lists_read(*c->lists, c->recs, dir ? dir : "");
}

int aces_lists_write(const char* dir) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return lists_write(*c->lists, c->recs, dir ? dir : "");
}

int aces_intprc_df(const double* d_B, int naux, const double* eps) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(!c->ready(), "aces_intprc_df: the lists are written before the engine is built (GENINT)");
  return intprc_df(*c->lists, c->recs, d_B, naux, eps, c->stream);
}

int aces_putrec(const char* name, long nwords, const void* src) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->recs.putrec(name, nwords, src);
}
int aces_getrec(int iflag, const char* name, long* nwords, void* dest) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  return c->recs.getrec(iflag, name, nwords, dest);
}

int aces_vcc_set_precision(int prec) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(prec >= 0 && prec <= 3,
               "precision must be 0 (fp64), 1 (fp32), 2 (bf16x3) or 3 (bf16x3 refinement of an fp64 anchor)");
  if (prec == kPrecRefine && c->prec != kPrecRefine) {  // anchor at the switch
    if (c->eng) c->eng->reset_anchor();
    if (c->seng) c->seng->reset_anchor();
  }
  c->prec = prec;
  return ACES_OK;
}

// ---- Fortran linkage (INTEGER*8 by reference, hidden CHARACTER lengths) ----
typedef int64_t i8;
static void fatal(int rc, const char* where) {
  if (rc != ACES_OK) {
    fprintf(stderr, "@%s: %s\n", where, last_error());
    fflush(stderr);
    // The reference's fatal paths end the process through aces_exit(1).
    exit(1);
  }
}

void getlst_(double* dest, i8* ifirst, i8* ncols, i8* /*xcache*/, i8* ileft, i8* iright) {
  fatal(aces_getlst(dest, *ifirst, *ncols, (int)*ileft, (int)*iright, 0), "GETLST");
}
void putlst_(double* src, i8* ifirst, i8* ncols, i8* /*xcache*/, i8* ileft, i8* iright) {
  fatal(aces_putlst(src, *ifirst, *ncols, (int)*ileft, (int)*iright, 0), "PUTLST");
}
void getall_(double* dest, i8* length, i8* irrepx, i8* list) {
  fatal(aces_getall(dest, *length, (int)*irrepx, (int)*list, 0), "GETALL");
}
void putall_(double* src, i8* length, i8* irrepx, i8* list) {
  fatal(aces_putall(src, *length, (int)*irrepx, (int)*list, 0), "PUTALL");
}
void sumsym2_(double* add, double* /*scr*/, i8* /*dimscr*/, i8* /*junk*/, i8* ileft, i8* iright) {
  fatal(aces_sumsym2(add, (int)*ileft, (int)*iright, 0), "SUMSYM2");
}
void inipck_(i8* irrepx, i8* typel, i8* typer, i8* list, i8* /*iarg1*/, i8* /*iarg2*/, i8* /*iset*/) {
  fatal(aces_inipck((int)*irrepx, (int)*typel, (int)*typer, (int)*list), "INIPCK");
}
void updmoi_(i8* ncols, i8* nrows, i8* group, i8* list, i8* /*ienter*/, i8* /*ioff*/) {
  fatal(aces_updmoi(*ncols, *nrows, (int)*group, (int)*list), "UPDMOI");
}
void getrec_(i8* iflag, const char* /*archive*/, const char* name, i8* len, void* dest, size_t /*la*/,
             size_t ln) {
  std::string nm(name, ln);
  long l = *len;
  fatal(aces_getrec((int)*iflag, nm.c_str(), &l, dest), "GETREC");
  if (*iflag == 0) *len = l;
}
void putrec_(i8* /*iflag*/, const char* /*archive*/, const char* name, i8* len, const void* src, size_t /*la*/,
             size_t ln) {
  std::string nm(name, ln);
  fatal(aces_putrec(nm.c_str(), *len, src), "PUTREC");
}

// ---- xvcc driver entry points ----
// C1 lists run the C1 engine on the lists' own memory; symmetry-packed lists
// run the irrep-blocked engine, the amplitude lists moving through C1 order.
int aces_genint(void) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  if (!c->ready()) ACES_CHECK(setup_engine(c));
  ACES_CHECK(ensure_amp_lists(c));
  if (c->seng) {
    SymBuffers& B = c->sb;
    ACES_CHECK(sym_get_t(c, 1, 46, B.t1, B.t2));
    ACES_CHECK(c->seng->set_amplitudes(B.t1, B.t2));
    return c->seng->increments((GemmPrec)c->prec, B.inc1, B.inc2);
  }
  CcsdEngine& e = *c->eng;
  ACES_CHECK(e.set_amplitudes(c->lists->device_ptr(1, 90), c->lists->device_ptr(1, 46)));
  return e.increments((GemmPrec)c->prec, nullptr, nullptr);
}

int aces_dre3en(void) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(c->ready(), "DRE3EN before GENINT");
  if (c->seng) {
    ACES_CHECK(sym_put_t2(c, 63, 61, c->sb.inc2));
    ACES_HIP(hipStreamSynchronize(c->stream));
    return ACES_OK;
  }
  CcsdEngine& e = *c->eng;
  const long ov = (long)e.o * e.v;
  ACES_HIP(hipMemcpyAsync(c->lists->device_ptr(1, 63), e.Inc() + ov, sizeof(double) * ov * ov,
                          hipMemcpyDeviceToDevice, c->stream));
  ACES_CHECK(abtoaa(e.o, e.v, c->lists->device_ptr(1, 63), c->lists->device_ptr(1, 61), c->stream));
  ACES_HIP(hipStreamSynchronize(c->stream));
  return ACES_OK;
}

int aces_e4s(void) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(c->ready(), "E4S before GENINT");
  if (c->seng) {
    SymBuffers& B = c->sb;
    ACES_CHECK(jacobi(c->o, c->v, B.eo, B.ev, B.inc1, B.inc2, B.t1n, B.t2n, c->stream));
    ACES_CHECK(dense_to_t1(*c->lists, 3, 90, B.t1n, c->stream));
    ACES_HIP(hipStreamSynchronize(c->stream));
    return ACES_OK;
  }
  CcsdEngine& e = *c->eng;
  ACES_CHECK(e.jacobi_step());
  const long ov = (long)e.o * e.v;
  ACES_HIP(hipMemcpyAsync(c->lists->device_ptr(3, 90), e.Tnew(), sizeof(double) * ov, hipMemcpyDeviceToDevice,
                          c->stream));
  ACES_HIP(hipStreamSynchronize(c->stream));
  return ACES_OK;
}

int aces_newt2(void) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(c->ready(), "NEWT2 before GENINT");
  if (c->seng) {
    SymBuffers& B = c->sb;
    // the increment on list 63 (possibly modified by the caller) divided by D
    ACES_CHECK(dpd_to_c1(*c->lists, 63, 13, 14, 1, B.inc2, c->stream));
    ACES_CHECK(jacobi(c->o, c->v, B.eo, B.ev, B.inc1, B.inc2, B.t1n, B.t2n, c->stream));
    ACES_CHECK(sym_put_t2(c, 63, 61, B.t2n));
    ACES_HIP(hipStreamSynchronize(c->stream));
    return ACES_OK;
  }
  CcsdEngine& e = *c->eng;
  const long ov = (long)e.o * e.v;
  // the increment on list 63 (possibly modified by the caller) divided by D
  ACES_HIP(hipMemcpyAsync(e.Inc() + ov, c->lists->device_ptr(1, 63), sizeof(double) * ov * ov,
                          hipMemcpyDeviceToDevice, c->stream));
  ACES_CHECK(e.jacobi_step());
  ACES_HIP(hipMemcpyAsync(c->lists->device_ptr(1, 63), e.Tnew() + ov, sizeof(double) * ov * ov,
                          hipMemcpyDeviceToDevice, c->stream));
  ACES_CHECK(abtoaa(e.o, e.v, c->lists->device_ptr(1, 63), c->lists->device_ptr(1, 61), c->stream));
  ACES_HIP(hipStreamSynchronize(c->stream));
  return ACES_OK;
}

// CMPENG (libr/cmpeng.F:60-188): nlist2 = 43 (T lists 44/46) or 60 (increment
// lists 61/63), T1 from list 90 group nlist1 + 1.  Closed shell (IUHF = 0):
//   ecorr3[0] = E(AA) (+ the non-HF singles term sum t1 f), ecorr3[2] = E(AB),
//   ecorr3[1] untouched (the reference does not set it for RHF),
//   *etot = 2 ecorr3[0] + ecorr3[2], *etott2 the same with T2 in place of tau.
// The AA pieces are evaluated from the AB lists (list 44 = ABTOAA(46) for RHF);
// symmetry-packed lists are read through C1 order.
int aces_cmpeng(int nlist2, int nlist1, double* ecorr3, double* etot, double* etott2) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(c->ready(), "CMPENG before GENINT");
  ACES_REQUIRE(nlist2 == 43 || nlist2 == 60, "CMPENG: NLIST2 must be 43 (lists 44/46) or 60 (lists 61/63)");
  ACES_REQUIRE(nlist1 == 0 || nlist1 == 2, "CMPENG: NLIST1 must be 0 (list 90 group 1) or 2 (group 3)");
  const int o = c->o, v = c->v;
  const double *t1, *t2, *w;
  if (c->seng) {
    SymBuffers& B = c->sb;
    ACES_CHECK(sym_get_t(c, nlist1 + 1, nlist2 + 3, B.x1, B.x2));
    t1 = B.x1;
    t2 = B.x2;
    w = B.l16;
  } else {
    t1 = c->lists->device_ptr(nlist1 + 1, 90);
    t2 = c->lists->device_ptr(1, nlist2 + 3);
    w = c->lists->device_ptr(1, 16);
  }
  ACES_REQUIRE(t1 && t2 && w, "CMPENG: lists 90, 16 and the T2 list must exist");
  hipStream_t s = c->stream;
  double* out = c->epart + 4L * o * o;
  hipLaunchKernelGGL(cmpeng_partial_kernel, dim3(o * o), dim3(256), 0, s, o, v, t1, t2, w, c->epart);
  ACES_HIP(hipGetLastError());
  hipLaunchKernelGGL(cmpeng_final_kernel, dim3(1), dim3(256), 0, s, o * o, o * v, c->epart, t1, c->fov, out);
  ACES_HIP(hipGetLastError());
  double h[5];
  ACES_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
  ACES_HIP(hipStreamSynchronize(s));
  const double eaa = h[2] + h[4];
  if (etot) *etot = 2.0 * eaa + h[0];
  if (etott2) *etott2 = 2.0 * (h[3] + h[4]) + h[1];
  if (ecorr3) {
    ecorr3[0] = eaa;
    ecorr3[2] = h[0];
  }
  return ACES_OK;
}

// DODIIS0 on T (44/46, 90 g1) and T_new (61/63, 90 g3).  stats as aces_ccsd_cycle.
int aces_dodiis0(int icycle, double tol, int* iconvg, double* stats) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(c->ready(), "DODIIS0 before GENINT");
  hipStream_t s = c->stream;
  CcsdStats st;
  if (c->seng) {
    SymBuffers& B = c->sb;
    ACES_CHECK(sym_get_t(c, 1, 46, B.t1, B.t2));
    ACES_CHECK(sym_get_t(c, 3, 63, B.t1n, B.t2n));
    ACES_CHECK(c->seng->diis_dense(icycle, tol, B.t1, B.t2, B.t1n, B.t2n, st));
    ACES_CHECK(c->seng->get_amplitudes(B.x1, B.x2));
    if (!st.converged) {
      // extrapolated (or Jacobi) vector -> increment lists
      ACES_CHECK(dense_to_t1(*c->lists, 3, 90, B.x1, s));
      ACES_CHECK(sym_put_t2(c, 63, 61, B.x2));
    } else {
      // T stays (symmetrised in place); the extrapolant still goes over the
      // increment lists (vcc/dodiis0.F:181-233), the Lambda start DRMOVE(100) moves
      ACES_CHECK(sym_put_t2(c, 46, -1, B.x2));
      ACES_CHECK(c->seng->get_increment_lists(B.x1, B.x2));
      ACES_CHECK(dense_to_t1(*c->lists, 3, 90, B.x1, s));
      ACES_CHECK(sym_put_t2(c, 63, 61, B.x2));
    }
  } else {
    CcsdEngine& e = *c->eng;
    const long ov = (long)e.o * e.v;
    ACES_HIP(hipMemcpyAsync(e.T(), c->lists->device_ptr(1, 90), sizeof(double) * ov, hipMemcpyDeviceToDevice, s));
    ACES_HIP(hipMemcpyAsync(e.T() + ov, c->lists->device_ptr(1, 46), sizeof(double) * ov * ov,
                            hipMemcpyDeviceToDevice, s));
    ACES_HIP(hipMemcpyAsync(e.Tnew(), c->lists->device_ptr(3, 90), sizeof(double) * ov, hipMemcpyDeviceToDevice,
                            s));
    ACES_HIP(hipMemcpyAsync(e.Tnew() + ov, c->lists->device_ptr(1, 63), sizeof(double) * ov * ov,
                            hipMemcpyDeviceToDevice, s));
    ACES_CHECK(e.energy_dev(e.Tnew(), e.Tnew() + ov, nullptr));
    ACES_CHECK(e.diis_step(icycle, tol, st));
    // extrapolated (or Jacobi) vector -> increment lists (vcc/dodiis0.F "WRITE IT
    // OVER CURRENT INCREMENT LISTS", in the converged cycle too: there it is the
    // Lambda start that DRMOVE(...,100) moves, and the engine keeps it in Tnew)
    const double* nx = st.converged ? e.Tnew_ro() : e.T_ro();
    ACES_HIP(hipMemcpyAsync(c->lists->device_ptr(3, 90), nx, sizeof(double) * ov, hipMemcpyDeviceToDevice, s));
    ACES_HIP(hipMemcpyAsync(c->lists->device_ptr(1, 63), nx + ov, sizeof(double) * ov * ov,
                            hipMemcpyDeviceToDevice, s));
    ACES_CHECK(abtoaa(e.o, e.v, c->lists->device_ptr(1, 63), c->lists->device_ptr(1, 61), s));
    if (st.converged) {
      // DODIIS0 symmetrised list 46 in place (libr2/symetrize_t2.F)
      ACES_HIP(hipMemcpyAsync(c->lists->device_ptr(1, 46), e.T_ro() + ov, sizeof(double) * ov * ov,
                              hipMemcpyDeviceToDevice, s));
    }
  }
  *iconvg = st.converged ? 0 : 1;
  if (stats) {
    stats[0] = st.energy_new;
    stats[1] = st.resmiter;
    stats[2] = st.resmdiis;
    stats[3] = st.converged;
    stats[4] = st.energy_final;
  }
  ACES_HIP(hipStreamSynchronize(s));
  return ACES_OK;
}

// DRMOVE (libr/drmove.f): lists 61 -> 44 + nlist and 63 -> 46 + nlist (MVSYM,
// every irrep), with sing1 list 90 group 3 -> list 90 + nlist group 1 (MVLIST).
// nlist = 0 moves the increments into the T lists; nlist = 100 fills the Lambda
// start lists 144/146/190 and leaves the T lists unchanged.
int aces_drmove_lists(int nlist, int sing1) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(nlist >= 0 && nlist + 90 <= 500, "DRMOVE: NLIST out of range");
  ListStore& L = *c->lists;
  const int moves[2][2] = {{61, 44 + nlist}, {63, 46 + nlist}};
  for (const auto& mv : moves) {
    const std::vector<int> g = L.groups(mv[0]);
    ACES_REQUIRE(!g.empty(), "DRMOVE: increment list " + std::to_string(mv[0]) + " does not exist");
    for (int grp : g) ACES_CHECK(L.move(grp, mv[0], grp, mv[1]));
  }
  if (sing1) ACES_CHECK(L.move(3, 90, 1, 90 + nlist));
  return ACES_OK;
}

int aces_drmove(void) { return aces_drmove_lists(0, 1); }

// MP1 start written by xintprc (intprc/dfrmt2.F) into lists 46/44 and T1 = 0.
int aces_vcc_mp1(void) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  if (!c->ready()) ACES_CHECK(setup_engine(c));
  ACES_CHECK(ensure_amp_lists(c));
  if (c->seng) {
    SymBuffers& B = c->sb;
    ACES_CHECK(c->seng->start_mp1());
    ACES_CHECK(c->seng->get_amplitudes(B.t1, B.t2));
    ACES_CHECK(dense_to_t1(*c->lists, 1, 90, B.t1, c->stream));
    ACES_CHECK(sym_put_t2(c, 46, 44, B.t2));
    ACES_HIP(hipStreamSynchronize(c->stream));
    return ACES_OK;
  }
  CcsdEngine& e = *c->eng;
  const long ov = (long)e.o * e.v;
  ACES_CHECK(e.start_mp1());
  ACES_HIP(hipMemcpyAsync(c->lists->device_ptr(1, 90), e.T(), sizeof(double) * ov, hipMemcpyDeviceToDevice,
                          c->stream));
  ACES_HIP(hipMemcpyAsync(c->lists->device_ptr(1, 46), e.T() + ov, sizeof(double) * ov * ov,
                          hipMemcpyDeviceToDevice, c->stream));
  ACES_CHECK(abtoaa(e.o, e.v, c->lists->device_ptr(1, 46), c->lists->device_ptr(1, 44), c->stream));
  ACES_HIP(hipStreamSynchronize(c->stream));
  return ACES_OK;
}

int aces_vcc_set_diis_order(int order) {
  AcesContext* c;
  ACES_CHECK(ctx(&c));
  ACES_REQUIRE(c->ready(), "DIISLST before GENINT/MP1");
  if (c->seng) {
    ACES_CHECK(c->seng->set_diis_order(order));
    return c->seng->reset_diis();
  }
  ACES_CHECK(c->eng->set_diis_order(order));
  return c->eng->reset_diis();
}

// Fortran spellings of the driver entry points (signatures of vcc/vcc.F's calls).
void genint_(i8*, i8*, i8*, i8*, i8*) { fatal(aces_genint(), "GENINT"); }
void feacont_(i8*, i8*, i8*) {}
void fmicont_(i8*, i8*, i8*) {}
void fmecont_(i8*, i8*, i8*, i8*) {}
void t1int2a_(i8*, i8*, i8*) {}
void t1int2b_(i8*, i8*, i8*) {}
void t1int1_(i8*, i8*, i8*, i8*) {}
void initin_(i8*, i8*, i8*) {}
void initsn_(i8*, i8*, i8*) {}
// RNABIJ (libr/rnabij.f) writes re-sorted copies of <Ab|Ij> ('W'), T2 ('T'), the
// T2 increment ('R') or L2 ('L') in the ring orders (lists 17-22, 34-39); the
// engine forms those orders inside its ring products, so these four are no-ops
// here.  'C' (iterative MBPT from another list set) is not implemented.
void rnabij_(i8*, i8*, i8*, const char* type, size_t len) {
  const char t = (type && len > 0) ? type[0] : ' ';
  if (t != 'W' && t != 'T' && t != 'R' && t != 'L') {
    set_error(std::string("RNABIJ: re-sort type '") + t + "' is not implemented");
    fatal(ACES_ERR_ARG, "RNABIJ");
  }
}
void dre3en_(i8*, i8*, i8*, i8*) { fatal(aces_dre3en(), "DRE3EN"); }
// SE4 is the fourth-order singles energy of finite-order MBPT (METHOD 3/4,
// vcc/e4seng.F); for the CC methods the reference returns 0 as well.
void e4s_(i8*, i8*, i8*, double* se4) {
  fatal(aces_e4s(), "E4S");
  if (se4) *se4 = 0.0;
}
void newt2_(i8*, i8*, i8*) { fatal(aces_newt2(), "NEWT2"); }
void cmpeng_(i8*, i8*, i8* nlist2, i8* nlist1, double* ecorr, double* etot, double* etott2, i8* iuhf,
             i8* iprint) {
  if (iuhf && *iuhf != 0) {
    set_error("CMPENG: only closed-shell (IUHF = 0) lists are implemented");
    fatal(ACES_ERR_ARG, "CMPENG");
  }
  fatal(aces_cmpeng((int)*nlist2, (int)*nlist1, ecorr, etot, etott2), "CMPENG");
  // output lines of libr/cmpeng.F (formats 80 and 81)
  if (iprint && *iprint) {
    printf("   The AA contribution to the correlation energy is: %12.7f a.u.\n", ecorr[0]);
    printf("   The AB contribution to the correlation energy is: %12.7f a.u.\n", ecorr[2]);
    printf("   The total correlation energy is %15.12f a.u.\n", *etot);
  }
  fflush(stdout);
}
// DODIIS0 with the argument list of vcc/dodiis0.F:1-4.  xvcc's call is
// DODIIS0(...,1,ICYCLE,ICONVG,ICONTL,SING1,44,61,90,0,90,2,70,'     ',DAMP)
// (vcc/vcc.F:1073-1075); that is the set implemented.
void dodiis0_(double*, i8*, i8* iuhf, i8* irrepx, i8* icycle, i8* iconvg, i8* icontl, i8*, i8* listt20,
              i8* listt21, i8* listt10, i8* ioff0, i8* listt11, i8* ioff1, i8* lsterr, const char*,
              double* damp, size_t) {
  const bool ok = *iuhf == 0 && *irrepx == 1 && *listt20 == 44 && *listt21 == 61 && *listt10 == 90 &&
                  *ioff0 == 0 && *listt11 == 90 && *ioff1 == 2 && *lsterr == 70;
  if (!ok) {
    set_error("DODIIS0: only the closed-shell amplitude lists (44, 61, 90, 0, 90, 2, 70) are implemented");
    fatal(ACES_ERR_ARG, "DODIIS0");
  }
  if (*damp != 1.0) {
    set_error("DODIIS0: DAMP_PARAMETER other than 1 is not implemented");
    fatal(ACES_ERR_ARG, "DODIIS0");
  }
  int conv = 1;
  double st[8];
  fatal(aces_dodiis0((int)*icycle, std::pow(10.0, -(double)*icontl), &conv, st), "DODIIS0");
  *iconvg = conv;
  // output lines of vcc/dodiis0.F (formats 1001-1004)
  printf("  Convergence information after %5ld iterations: \n", (long)*icycle);
  printf("  Largest element of residual vector : %15.8E.\n", st[1]);
  printf("  Largest element of DIIS residual   : %15.8E.\n", st[2]);
  if (conv == 0) printf("  Amplitude equations converged in %5ld iterations.\n", (long)*icycle);
  fflush(stdout);
}
void drmove_(i8*, i8*, i8* iuhf, i8* nlist, i8* sing1) {
  if (*iuhf != 0) {
    set_error("DRMOVE: only closed-shell (IUHF = 0) lists are implemented");
    fatal(ACES_ERR_ARG, "DRMOVE");
  }
  fatal(aces_drmove_lists((int)*nlist, *sing1 != 0), "DRMOVE");
}

}  // extern "C"
