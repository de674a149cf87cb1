// Device-resident DPD list store and record store (see liststore.h).
#include <cstring>
#include "common.h"
#include "liststore.h"

namespace aces {

namespace {
// Pair types 1..25: (left space, right space, left spin, right spin, packing).
// The table is the ACES II DPD pair-type convention (acescore/aces_com_symloc.F:11-34):
// 1-4 same-spin p<q (VV a, VV b, OO a, OO b), 5-8 the same with p<=q,
// 9-12 VO (a a, b b, a b, b a), 13 Ab, 14 Ij, 15 Ab, 16 IA, 17 ia, 18 Ia,
// 19 AB full, 20 ab full, 21 IJ full, 22 ij full, 23 aB, 24 iJ, 25 iA.
enum Pack { FULL = 0, PACK = 1, PCK2 = 2 };
struct PairType {
  char l, r;
  int sl, sr;
  Pack pk;
};
const PairType kTypes[26] = {
    {'-', '-', 0, 0, FULL},                                                               // unused 0
    {'V', 'V', 0, 0, PACK}, {'V', 'V', 1, 1, PACK}, {'O', 'O', 0, 0, PACK}, {'O', 'O', 1, 1, PACK},
    {'V', 'V', 0, 0, PCK2}, {'V', 'V', 1, 1, PCK2}, {'O', 'O', 0, 0, PCK2}, {'O', 'O', 1, 1, PCK2},
    {'V', 'O', 0, 0, FULL}, {'V', 'O', 1, 1, FULL}, {'V', 'O', 0, 1, FULL}, {'V', 'O', 1, 0, FULL},
    {'V', 'V', 0, 1, FULL}, {'O', 'O', 0, 1, FULL}, {'V', 'V', 0, 1, FULL}, {'O', 'V', 0, 0, FULL},
    {'O', 'V', 1, 1, FULL}, {'O', 'V', 0, 1, FULL}, {'V', 'V', 0, 0, FULL}, {'V', 'V', 1, 1, FULL},
    {'O', 'O', 0, 0, FULL}, {'O', 'O', 1, 1, FULL}, {'V', 'V', 1, 0, FULL}, {'O', 'O', 1, 0, FULL},
    {'O', 'V', 1, 0, FULL}};
}  // namespace

bool pair_type(int type, int* space_p, int* space_q, int* pack) {
  if (type < 1 || type > 25) return false;
  const PairType& t = kTypes[type];
  *space_p = 2 * t.sl + (t.l == 'V' ? 1 : 0);
  *space_q = 2 * t.sr + (t.r == 'V' ? 1 : 0);
  *pack = (t.pk == FULL) ? 0 : (t.pk == PACK ? 1 : 2);
  return true;
}

static inline int space_count(const DpdSym& s, char sp, int spin, int irrep) {
  return sp == 'O' ? s.pop[spin][irrep] : s.vrt[spin][irrep];
}

long DpdSym::irpdpd(int irp_pq, int type) const {
  if (type < 1 || type > 25 || irp_pq < 0 || irp_pq >= nirrep) return -1;
  const PairType& t = kTypes[type];
  long n = 0;
  for (int irp_q = 0; irp_q < nirrep; ++irp_q) {
    const int irp_p = dirprd(irp_q, irp_pq);
    const long p = space_count(*this, t.l, t.sl, irp_p), q = space_count(*this, t.r, t.sr, irp_q);
    if (t.pk == FULL) {
      n += p * q;
    } else if (irp_p < irp_q) {
      n += p * q;
    } else if (irp_p == irp_q) {
      n += (t.pk == PACK) ? p * (p - 1) / 2 : p * (p + 1) / 2;
    }
  }
  return n;
}

long DpdSym::isymoff(int irp_q_target, int irp_pq, int type) const {
  if (type < 1 || type > 25) return -1;
  const PairType& t = kTypes[type];
  long n = 0;
  for (int irp_q = 0; irp_q < nirrep; ++irp_q) {
    const int irp_p = dirprd(irp_q, irp_pq);
    const long p = space_count(*this, t.l, t.sl, irp_p), q = space_count(*this, t.r, t.sr, irp_q);
    if (irp_q == irp_q_target) return n;
    if (t.pk == FULL || irp_p < irp_q)
      n += p * q;
    else if (irp_p == irp_q)
      n += (t.pk == PACK) ? p * (p - 1) / 2 : p * (p + 1) / 2;
  }
  return -1;
}

ListStore::~ListStore() {
  for (auto& l : lists_)
    for (auto& g : l.second.groups)
      if (g.second.d) (void)hipFree(g.second.d);
}

int ListStore::set_symmetry(int nirrep, const int* popa, const int* popb, const int* vrta, const int* vrtb) {
  ACES_REQUIRE(nirrep == 1 || nirrep == 2 || nirrep == 4 || nirrep == 8, "symmetry: nirrep must be 1, 2, 4 or 8");
  sym_.nirrep = nirrep;
  for (int h = 0; h < 8; ++h) {
    sym_.pop[0][h] = h < nirrep ? popa[h] : 0;
    sym_.pop[1][h] = h < nirrep ? (popb ? popb[h] : popa[h]) : 0;
    sym_.vrt[0][h] = h < nirrep ? vrta[h] : 0;
    sym_.vrt[1][h] = h < nirrep ? (vrtb ? vrtb[h] : vrta[h]) : 0;
    ACES_REQUIRE(sym_.pop[0][h] >= 0 && sym_.vrt[0][h] >= 0 && sym_.pop[1][h] >= 0 && sym_.vrt[1][h] >= 0,
                 "symmetry: negative orbital count");
  }
  return ACES_OK;
}

int ListStore::updmoi(long ncols, long nrows, int group, int list) {
  ACES_REQUIRE(list >= 1 && list <= 500 && group >= 1 && group <= 255, "UPDMOI: list 1..500, group 1..255");
  ACES_REQUIRE(ncols >= 0 && nrows >= 0, "UPDMOI: negative size");
  Block& b = lists_[list].groups[group];
  if (b.d && b.nrows * b.ncols == nrows * ncols) {
    b.nrows = nrows;
    b.ncols = ncols;
  } else {
    if (b.d) {
      ACES_HIP(hipFree(b.d));
      bytes_ -= b.nrows * b.ncols * 8;
      b.d = nullptr;
    }
    b.nrows = nrows;
    b.ncols = ncols;
    if (nrows * ncols > 0) {
      hipError_t e = hipMalloc(&b.d, sizeof(double) * nrows * ncols);
      if (e != hipSuccess) {
        set_error("UPDMOI: device allocation failed for list " + std::to_string(list));
        return ACES_ERR_NOMEM;
      }
      bytes_ += nrows * ncols * 8;
    }
  }
  if (b.d) ACES_HIP(hipMemsetAsync(b.d, 0, sizeof(double) * nrows * ncols, s_));
  return ACES_OK;
}

int ListStore::inipck(int irrepx, int typel, int typer, int list) {
  ACES_REQUIRE(typel >= 1 && typel <= 25 && typer >= 1 && typer <= 25, "INIPCK: pair types must be 1..25");
  ACES_REQUIRE(irrepx >= 1 && irrepx <= sym_.nirrep, "INIPCK: irrepx out of range");
  for (int r = 0; r < sym_.nirrep; ++r) {
    const int l = DpdSym::dirprd(r, irrepx - 1);
    ACES_CHECK(updmoi(sym_.irpdpd(r, typer), sym_.irpdpd(l, typel), r + 1, list));
  }
  lists_[list].tl = typel;
  lists_[list].tr = typer;
  return ACES_OK;
}

int ListStore::isytyp(int list, int* tl, int* tr) const {
  auto it = lists_.find(list);
  if (it == lists_.end()) {
    *tl = *tr = 0;
    return ACES_ERR_RANGE;
  }
  *tl = it->second.tl;
  *tr = it->second.tr;
  return ACES_OK;
}

int ListStore::block(int group, int list, Block** b) {
  auto it = lists_.find(list);
  if (it == lists_.end() || it->second.groups.count(group) == 0) {
    set_error("list (" + std::to_string(group) + "," + std::to_string(list) + ") does not exist");
    return ACES_ERR_RANGE;
  }
  *b = &it->second.groups[group];
  return ACES_OK;
}

int ListStore::rows(int group, int list, long* n) const {
  auto it = lists_.find(list);
  if (it == lists_.end() || !it->second.groups.count(group)) {
    *n = 0;
    return ACES_ERR_RANGE;
  }
  *n = it->second.groups.at(group).nrows;
  return ACES_OK;
}

int ListStore::cols(int group, int list, long* n) const {
  auto it = lists_.find(list);
  if (it == lists_.end() || !it->second.groups.count(group)) {
    *n = 0;
    return ACES_ERR_RANGE;
  }
  *n = it->second.groups.at(group).ncols;
  return ACES_OK;
}

double* ListStore::device_ptr(int group, int list) const {
  auto it = lists_.find(list);
  if (it == lists_.end() || !it->second.groups.count(group)) return nullptr;
  return it->second.groups.at(group).d;
}

int ListStore::getlst(double* dest, long ifirst, long ncols, int group, int list, bool dev) {
  Block* b;
  ACES_CHECK(block(group, list, &b));
  // range check as getlst.F:100-108
  if (ifirst < 1 || ncols < 0 || ifirst - 1 + ncols > b->ncols) {
    set_error("GETLST: columns " + std::to_string(ifirst) + ".." + std::to_string(ifirst + ncols - 1) +
              " outside list (" + std::to_string(group) + "," + std::to_string(list) + ") with " +
              std::to_string(b->ncols) + " columns");
    return ACES_ERR_RANGE;
  }
  const long n = ncols * b->nrows;
  if (n == 0) return ACES_OK;
  ACES_HIP(hipMemcpyAsync(dest, b->d + (ifirst - 1) * b->nrows, sizeof(double) * n,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

int ListStore::putlst(const double* src, long ifirst, long ncols, int group, int list, bool dev) {
  Block* b;
  ACES_CHECK(block(group, list, &b));
  if (ifirst < 1 || ncols < 0 || ifirst - 1 + ncols > b->ncols) {
    set_error("PUTLST: columns outside list (" + std::to_string(group) + "," + std::to_string(list) + ")");
    return ACES_ERR_RANGE;
  }
  const long n = ncols * b->nrows;
  if (n == 0) return ACES_OK;
  ACES_HIP(hipMemcpyAsync(b->d + (ifirst - 1) * b->nrows, src, sizeof(double) * n,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

int ListStore::getall(double* dest, long length, int irrepx, int list, bool dev) {
  long off = 0;
  for (int r = 1; r <= sym_.nirrep; ++r) {
    long nr, nc;
    if (rows(r, list, &nr) != ACES_OK || cols(r, list, &nc) != ACES_OK) continue;
    if (nr * nc == 0) continue;
    ACES_REQUIRE(off + nr * nc <= length, "GETALL: destination too short");
    ACES_CHECK(getlst(dest + off, 1, nc, r, list, dev));
    off += nr * nc;
  }
  (void)irrepx;
  return ACES_OK;
}

int ListStore::putall(const double* src, long length, int irrepx, int list, bool dev) {
  long off = 0;
  for (int r = 1; r <= sym_.nirrep; ++r) {
    long nr, nc;
    if (rows(r, list, &nr) != ACES_OK || cols(r, list, &nc) != ACES_OK) continue;
    if (nr * nc == 0) continue;
    ACES_REQUIRE(off + nr * nc <= length, "PUTALL: source too short");
    ACES_CHECK(putlst(src + off, 1, nc, r, list, dev));
    off += nr * nc;
  }
  (void)irrepx;
  return ACES_OK;
}

__global__ void axpy_kernel(long n, const double* __restrict__ x, double* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] += x[i];
}

int ListStore::sumsym2(const double* add, int group, int list, bool dev) {
  Block* b;
  ACES_CHECK(block(group, list, &b));
  const long n = b->nrows * b->ncols;
  if (n == 0) return ACES_OK;
  const double* src = add;
  double* tmp = nullptr;
  if (!dev) {
    ACES_HIP(hipMalloc(&tmp, sizeof(double) * n));
    ACES_HIP(hipMemcpyAsync(tmp, add, sizeof(double) * n, hipMemcpyHostToDevice, s_));
    src = tmp;
  }
  long g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)g), dim3(256), 0, s_, n, src, b->d);
  ACES_HIP(hipGetLastError());
  ACES_HIP(hipStreamSynchronize(s_));
  if (tmp) ACES_HIP(hipFree(tmp));
  return ACES_OK;
}

int ListStore::zero(int group, int list) {
  Block* b;
  ACES_CHECK(block(group, list, &b));
  if (b->nrows * b->ncols) ACES_HIP(hipMemsetAsync(b->d, 0, sizeof(double) * b->nrows * b->ncols, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

int ListStore::move(int src_group, int src_list, int dst_group, int dst_list) {
  Block* sb;
  ACES_CHECK(block(src_group, src_list, &sb));
  if (src_group == dst_group && src_list == dst_list) return ACES_OK;
  const long nr = sb->nrows, nc = sb->ncols;
  long r = -1, k = -1;
  if (rows(dst_group, dst_list, &r) != ACES_OK || cols(dst_group, dst_list, &k) != ACES_OK || r != nr || k != nc)
    ACES_CHECK(updmoi(nc, nr, dst_group, dst_list));
  List& dl = lists_[dst_list];
  if (dl.tl == 0 && dl.tr == 0) {
    dl.tl = lists_[src_list].tl;
    dl.tr = lists_[src_list].tr;
  }
  Block* db;
  ACES_CHECK(block(dst_group, dst_list, &db));
  if (nr * nc) ACES_HIP(hipMemcpyAsync(db->d, sb->d, sizeof(double) * nr * nc, hipMemcpyDeviceToDevice, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

std::vector<int> ListStore::groups(int list) const {
  std::vector<int> g;
  auto it = lists_.find(list);
  if (it != lists_.end())
    for (const auto& kv : it->second.groups) g.push_back(kv.first);
  return g;
}

// ---------------------------------------------------------------------------
std::string RecordStore::key(const std::string& name) {
  std::string k = name.substr(0, 8);
  while (!k.empty() && k.back() == ' ') k.pop_back();
  return k;
}

int RecordStore::getrec(int iflag, const std::string& name, long* len, void* dest) {
  auto it = index_.find(key(name));
  if (iflag == 0) {
    *len = (it == index_.end()) ? -1 : (long)slots_[it->second].data.size();
    return ACES_OK;
  }
  if (it == index_.end()) {
    if (iflag > 0) {
      set_error("GETREC: record " + key(name) + " not found");
      return ACES_ERR_RANGE;
    }
    if (*len > 0) std::memset(dest, 0, sizeof(long) * (*len));
    return ACES_OK;
  }
  const std::vector<long>& d = slots_[it->second].data;
  if (*len > (long)d.size()) {
    set_error("GETREC: record " + key(name) + " is shorter than requested");
    return ACES_ERR_RANGE;
  }
  if (*len > 0) std::memcpy(dest, d.data(), sizeof(long) * (*len));
  return ACES_OK;
}

int RecordStore::putrec(const std::string& name, long len, const void* src) {
  ACES_REQUIRE(len >= 0, "PUTREC: negative length");
  const std::string k = key(name);
  ACES_REQUIRE(!k.empty() && k != "OPENSLOT" && k != ". LOST .", "PUTREC: invalid record name '" + k + "'");
  auto it = index_.find(k);
  if (it != index_.end() && len > (long)slots_[it->second].data.size()) {
    slots_[it->second].name = ". LOST .";
    index_.erase(it);
    it = index_.end();
  }
  if (it == index_.end()) {
    ACES_REQUIRE(slots_.size() < 1000, "PUTREC: maximum number of JOBARC records exceeded");
    slots_.push_back(Slot{k, std::vector<long>((size_t)len, 0)});
    it = index_.emplace(k, slots_.size() - 1).first;
  }
  std::vector<long>& d = slots_[it->second].data;
  if (len) std::memcpy(d.data(), src, sizeof(long) * len);
  return ACES_OK;
}

void RecordStore::assign(std::vector<Slot> slots) {
  slots_ = std::move(slots);
  index_.clear();
  for (size_t i = 0; i < slots_.size(); ++i)
    if (slots_[i].name != ". LOST ." && !index_.count(slots_[i].name)) index_.emplace(slots_[i].name, i);
}

}  // namespace aces
