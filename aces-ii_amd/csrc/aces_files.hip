// The ACES II job archive (JAINDX + JOBARC) and list files (MOINTS, GAMLAM,
// MOABCD, DERINT, DERGAM) on disk <-> the record and list stores (see
// aces_files.h).  Host code: the files are streamed in column blocks through
// a host buffer into the device-resident lists.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "aces_files.h"
#include "common.h"

namespace aces {

namespace {

constexpr int kJaRecs = 1000;      // logical records (include/jobarc.h _MAX_JA_RECS)
constexpr long kJaRecWd = 128;     // JOBARC physical record, integers (_JA_RECWD)
constexpr int kGrps = 10, kFams = 500, kFamsPerFile = 100, kLuns = 5, kFirstLun = 50;  // include/lists.h
constexpr long kDefaultRecWd = 8192;  // list-file physical record when no IFLAGS record exists (64 KB)
const char* const kListFiles[kLuns] = {"MOINTS", "GAMLAM", "MOABCD", "DERINT", "DERGAM"};  // szfiles.dat

std::string path_of(const std::string& dir, const char* name) {
  return dir.empty() ? std::string(name) : dir + "/" + name;
}

struct File {
  FILE* f = nullptr;
  File() = default;
  File(const std::string& p, const char* mode) { f = std::fopen(p.c_str(), mode); }
  File(const File&) = delete;
  File& operator=(const File&) = delete;
  ~File() {
    if (f) std::fclose(f);
  }
};

int read_at(FILE* f, long byte_off, void* dst, size_t bytes, const std::string& what) {
  if (bytes == 0) return ACES_OK;
  if (std::fseek(f, byte_off, SEEK_SET) != 0 || std::fread(dst, 1, bytes, f) != bytes) {
    set_error(what + ": short read at byte " + std::to_string(byte_off));
    return ACES_ERR_RANGE;
  }
  return ACES_OK;
}

int write_at(FILE* f, long byte_off, const void* src, size_t bytes, const std::string& what) {
  if (bytes == 0) return ACES_OK;
  if (std::fseek(f, byte_off, SEEK_SET) != 0 || std::fwrite(src, 1, bytes, f) != bytes) {
    set_error(what + ": write failed at byte " + std::to_string(byte_off));
    return ACES_ERR_RANGE;
  }
  return ACES_OK;
}

// Extend a file with zeros to `bytes` (direct-access records that exist but
// hold no list yet read as zeros).
int extend_to(FILE* f, long bytes, const std::string& what) {
  if (std::fseek(f, 0, SEEK_END) != 0) return ACES_ERR_RANGE;
  long have = std::ftell(f);
  std::vector<char> z(1 << 16, 0);
  while (have < bytes) {
    const size_t n = (size_t)std::min<long>((long)z.size(), bytes - have);
    if (std::fwrite(z.data(), 1, n, f) != n) {
      set_error(what + ": cannot extend the file");
      return ACES_ERR_RANGE;
    }
    have += (long)n;
  }
  return ACES_OK;
}

std::string pad8(const std::string& s) {
  std::string r = s.substr(0, 8);
  r.resize(8, ' ');
  return r;
}

// Location (1-based integer index) of each logical record, as PUTREC assigns
// them: after the previous slot, moved to the next physical record when a
// record of 2..128 integers would straddle one.
std::vector<long> record_locations(const std::vector<RecordStore::Slot>& slots) {
  std::vector<long> loc(slots.size());
  for (size_t i = 0; i < slots.size(); ++i) {
    long r = (i == 0) ? 1 : loc[i - 1] + (long)slots[i - 1].data.size();
    const long len = (long)slots[i].data.size();
    if (len >= 2 && len <= kJaRecWd) {
      const long rec0 = 1 + (r - 1) / kJaRecWd, rec1 = 1 + (r + len - 2) / kJaRecWd;
      if (rec0 != rec1) r = 1 + (rec1 - 1) * kJaRecWd;
    }
    loc[i] = r;
  }
  return loc;
}

int get_array(RecordStore& recs, const char* name, long n, std::vector<long>& out, bool required) {
  out.assign((size_t)n, 0);
  long len = 0;
  recs.getrec(0, name, &len, nullptr);
  if (len < 0) {
    if (required) {
      set_error(std::string("list files: record ") + name + " not found in the job archive");
      return ACES_ERR_RANGE;
    }
    return ACES_OK;
  }
  long m = std::min(len, n);
  return recs.getrec(1, name, &m, out.data());
}

}  // namespace

int jobarc_read(RecordStore& recs, const std::string& dir) {
  File ix(path_of(dir, "JAINDX"), "rb");
  ACES_REQUIRE(ix.f != nullptr, "JOBARC: cannot open " + path_of(dir, "JAINDX"));
  int32_t head = 0, tail = 0;
  const long body = 8L * kJaRecs + 8L * kJaRecs + 8L * kJaRecs + 8;
  std::vector<char> marker(8 * kJaRecs);
  std::vector<int64_t> rloc(kJaRecs), rsize(kJaRecs);
  int64_t nrecs = 0;
  ACES_CHECK(read_at(ix.f, 0, &head, 4, "JAINDX"));
  ACES_REQUIRE(head == body, "JAINDX: record length " + std::to_string(head) + ", expected " + std::to_string(body) +
                                 " (8-byte integers, 4-byte record markers)");
  ACES_CHECK(read_at(ix.f, 4, marker.data(), marker.size(), "JAINDX"));
  ACES_CHECK(read_at(ix.f, 4 + 8L * kJaRecs, rloc.data(), 8L * kJaRecs, "JAINDX"));
  ACES_CHECK(read_at(ix.f, 4 + 16L * kJaRecs, rsize.data(), 8L * kJaRecs, "JAINDX"));
  ACES_CHECK(read_at(ix.f, 4 + 24L * kJaRecs, &nrecs, 8, "JAINDX"));
  ACES_CHECK(read_at(ix.f, 4 + body, &tail, 4, "JAINDX"));
  ACES_REQUIRE(tail == head, "JAINDX: trailing record marker does not match");
  File ja(path_of(dir, "JOBARC"), "rb");
  ACES_REQUIRE(ja.f != nullptr, "JOBARC: cannot open " + path_of(dir, "JOBARC"));
  std::vector<RecordStore::Slot> slots;
  for (int i = 0; i < kJaRecs; ++i) {
    const std::string name(&marker[8 * (size_t)i], 8);
    if (name == "OPENSLOT") continue;
    ACES_REQUIRE(rloc[i] >= 1 && rsize[i] >= 0, "JAINDX: bad location of record '" + name + "'");
    ACES_REQUIRE(rsize[i] == 0 || (rloc[i] + rsize[i] - 1 + kJaRecWd - 1) / kJaRecWd <= nrecs,
                 "JAINDX: record '" + name + "' lies past the last physical record");
    RecordStore::Slot s;
    s.name = RecordStore::key(name);
    s.data.resize((size_t)rsize[i]);
    ACES_CHECK(read_at(ja.f, 8L * (rloc[i] - 1), s.data.data(), 8 * (size_t)rsize[i], "JOBARC"));
    slots.push_back(std::move(s));
  }
  recs.assign(std::move(slots));
  return ACES_OK;
}

int jobarc_write(const RecordStore& recs, const std::string& dir) {
  const std::vector<RecordStore::Slot>& slots = recs.slots();
  ACES_REQUIRE((int)slots.size() <= kJaRecs, "JOBARC: more than 1000 records");
  const std::vector<long> loc = record_locations(slots);
  long nrecs = 1;  // the archive starts with one all-zero physical record
  for (size_t i = 0; i < slots.size(); ++i)
    if (!slots[i].data.empty())
      nrecs = std::max(nrecs, (loc[i] + (long)slots[i].data.size() - 1 + kJaRecWd - 1) / kJaRecWd);
  {
    File ja(path_of(dir, "JOBARC"), "wb");
    ACES_REQUIRE(ja.f != nullptr, "JOBARC: cannot create " + path_of(dir, "JOBARC"));
    ACES_CHECK(extend_to(ja.f, nrecs * kJaRecWd * 8, "JOBARC"));
    for (size_t i = 0; i < slots.size(); ++i)
      ACES_CHECK(write_at(ja.f, 8L * (loc[i] - 1), slots[i].data.data(), 8 * slots[i].data.size(), "JOBARC"));
  }
  std::vector<char> marker(8 * kJaRecs);
  std::vector<int64_t> rloc(kJaRecs, 0), rsize(kJaRecs, 0);
  for (int i = 0; i < kJaRecs; ++i) {
    const std::string m = (i < (int)slots.size()) ? pad8(slots[(size_t)i].name) : std::string("OPENSLOT");
    std::memcpy(&marker[8 * (size_t)i], m.data(), 8);
    if (i < (int)slots.size()) {
      rloc[i] = loc[(size_t)i];
      rsize[i] = (int64_t)slots[(size_t)i].data.size();
    }
  }
  if (slots.empty()) rloc[0] = 1;
  const int64_t nr = nrecs;
  const int32_t len = (int32_t)(24L * kJaRecs + 8);
  File ix(path_of(dir, "JAINDX"), "wb");
  ACES_REQUIRE(ix.f != nullptr, "JOBARC: cannot create " + path_of(dir, "JAINDX"));
  ACES_CHECK(write_at(ix.f, 0, &len, 4, "JAINDX"));
  ACES_CHECK(write_at(ix.f, 4, marker.data(), marker.size(), "JAINDX"));
  ACES_CHECK(write_at(ix.f, 4 + 8L * kJaRecs, rloc.data(), 8L * kJaRecs, "JAINDX"));
  ACES_CHECK(write_at(ix.f, 4 + 16L * kJaRecs, rsize.data(), 8L * kJaRecs, "JAINDX"));
  ACES_CHECK(write_at(ix.f, 4 + 24L * kJaRecs, &nr, 8, "JAINDX"));
  ACES_CHECK(write_at(ix.f, 4 + 24L * kJaRecs + 8, &len, 4, "JAINDX"));
  return ACES_OK;
}

int lists_read(ListStore& lists, RecordStore& recs, const std::string& dir) {
  const long n = (long)kGrps * kFams;
  std::vector<long> moio, moiowd, moiosz, moiods, moiofl, isytyp, iflags;
  ACES_CHECK(get_array(recs, "MOIOVEC", n, moio, true));
  ACES_CHECK(get_array(recs, "MOIOWRD", n, moiowd, true));
  ACES_CHECK(get_array(recs, "MOIOSIZ", n, moiosz, true));
  ACES_CHECK(get_array(recs, "MOIODIS", n, moiods, true));
  ACES_CHECK(get_array(recs, "MOIOFIL", n, moiofl, true));
  ACES_CHECK(get_array(recs, "ISYMTYP", 2L * kFams, isytyp, false));
  ACES_CHECK(get_array(recs, "IFLAGS", 100, iflags, true));
  const long recwd = iflags[36];
  ACES_REQUIRE(recwd > 0, "list files: IFLAGS(37), the physical record length, is not set");
  // symmetry of the lists
  long nirr = 0;
  recs.getrec(0, "COMPNIRR", &nirr, nullptr);
  if (nirr > 0) {
    long ni = 1;
    ACES_CHECK(recs.getrec(1, "COMPNIRR", &ni, &nirr));
    ACES_REQUIRE(nirr >= 1 && nirr <= 8, "list files: COMPNIRR out of range");
    std::vector<long> p[4];
    const char* names[4] = {"SYMPOPOA", "SYMPOPOB", "SYMPOPVA", "SYMPOPVB"};
    for (int k = 0; k < 4; ++k) ACES_CHECK(get_array(recs, names[k], nirr, p[k], k == 0 || k == 2));
    int pa[8] = {0}, pb[8] = {0}, va[8] = {0}, vb[8] = {0};
    for (int h = 0; h < nirr; ++h) {
      pa[h] = (int)p[0][h];
      va[h] = (int)p[2][h];
      pb[h] = recs.has("SYMPOPOB") ? (int)p[1][h] : pa[h];
      vb[h] = recs.has("SYMPOPVB") ? (int)p[3][h] : va[h];
    }
    ACES_CHECK(lists.set_symmetry((int)nirr, pa, pb, va, vb));
  }
  File files[kLuns];
  std::vector<double> buf;
  for (int f = 1; f <= kFams; ++f) {
    for (int g = 1; g <= kGrps; ++g) {
      const long i = (long)(f - 1) * kGrps + (g - 1);
      if (moio[(size_t)i] == 0) continue;
      const long nr = moiosz[(size_t)i], nc = moiods[(size_t)i];
      const int h = (int)moiofl[(size_t)i] - kFirstLun;
      ACES_REQUIRE(h >= 0 && h < kLuns && nr >= 0 && nc >= 0 && moiowd[(size_t)i] >= 1,
                   "list files: bad MOIO entry for list (" + std::to_string(g) + "," + std::to_string(f) + ")");
      if (!files[h].f) {
        files[h].f = std::fopen(path_of(dir, kListFiles[h]).c_str(), "rb");
        ACES_REQUIRE(files[h].f != nullptr, std::string("list files: cannot open ") + path_of(dir, kListFiles[h]));
      }
      ACES_CHECK(lists.updmoi(nc, nr, g, f));
      const long word0 = (moio[(size_t)i] - 1) * recwd + (moiowd[(size_t)i] - 1);
      // whole columns through a host buffer of at most 32 Mi doubles
      const long cols_per = std::max<long>(1, (32L << 20) / std::max<long>(nr, 1));
      for (long c0 = 0; c0 < nc; c0 += cols_per) {
        const long m = std::min(cols_per, nc - c0);
        buf.resize((size_t)(m * nr));
        ACES_CHECK(read_at(files[h].f, 8 * (word0 + c0 * nr), buf.data(), 8 * (size_t)(m * nr), kListFiles[h]));
        if (m * nr > 0) ACES_CHECK(lists.putlst(buf.data(), c0 + 1, m, g, f, false));
      }
    }
    const int tl = (int)isytyp[2 * (size_t)(f - 1)], tr = (int)isytyp[2 * (size_t)(f - 1) + 1];
    if (tl > 0 || tr > 0) lists.set_types(f, tl, tr);
  }
  return ACES_OK;
}

int lists_write(ListStore& lists, RecordStore& recs, const std::string& dir) {
  std::vector<long> iflags;
  ACES_CHECK(get_array(recs, "IFLAGS", 100, iflags, false));
  if (!recs.has("IFLAGS") || iflags[36] <= 0) {
    iflags[36] = kDefaultRecWd;
    ACES_CHECK(recs.putrec("IFLAGS", 100, iflags.data()));
  }
  const long recwd = iflags[36];
  const long n = (long)kGrps * kFams;
  std::vector<long> moio(n, 0), moiowd(n, 0), moiosz(n, 0), moiods(n, 0), moiofl(n, 0), moiomx(n, 0);
  std::vector<long> isytyp(2L * kFams, 0), prec(kLuns, 1), ioff(kLuns, 0);
  File files[kLuns];
  std::vector<double> buf;
  for (int f = 1; f <= kFams; ++f) {
    const std::vector<int> groups = lists.groups(f);
    if (groups.empty()) continue;
    const int h = (f - 1) / kFamsPerFile;
    if (!files[h].f) {
      files[h].f = std::fopen(path_of(dir, kListFiles[h]).c_str(), "wb");
      ACES_REQUIRE(files[h].f != nullptr, std::string("list files: cannot create ") + path_of(dir, kListFiles[h]));
    }
    for (int g : groups) {
      ACES_REQUIRE(g >= 1 && g <= kGrps, "list files: group " + std::to_string(g) + " of list " + std::to_string(f) +
                                             " beyond the 10 groups of the MOIO records");
      long nr = 0, nc = 0;
      ACES_CHECK(lists.rows(g, f, &nr));
      ACES_CHECK(lists.cols(g, f, &nc));
      const long i = (long)(f - 1) * kGrps + (g - 1);
      // placement as aces_list_touch
      moio[(size_t)i] = prec[(size_t)h];
      moiowd[(size_t)i] = 1 + ioff[(size_t)h];
      moiosz[(size_t)i] = nr;
      moiods[(size_t)i] = nc;
      moiofl[(size_t)i] = kFirstLun + h;
      moiomx[(size_t)i] = nr * nc;
      const long word0 = (prec[(size_t)h] - 1) * recwd + ioff[(size_t)h];
      const long words = nr * nc + ioff[(size_t)h];
      prec[(size_t)h] += words / recwd;
      ioff[(size_t)h] = words % recwd;
      const long cols_per = std::max<long>(1, (32L << 20) / std::max<long>(nr, 1));
      for (long c0 = 0; c0 < nc; c0 += cols_per) {
        const long m = std::min(cols_per, nc - c0);
        buf.resize((size_t)(m * nr));
        if (m * nr > 0) ACES_CHECK(lists.getlst(buf.data(), c0 + 1, m, g, f, false));
        ACES_CHECK(write_at(files[h].f, 8 * (word0 + c0 * nr), buf.data(), 8 * (size_t)(m * nr), kListFiles[h]));
      }
    }
    int tl = 0, tr = 0;
    if (lists.isytyp(f, &tl, &tr) == ACES_OK) {
      isytyp[2 * (size_t)(f - 1)] = tl;
      isytyp[2 * (size_t)(f - 1) + 1] = tr;
    }
  }
  // the record holding the free space exists (aces_list_touch writes it)
  for (int h = 0; h < kLuns; ++h)
    if (files[h].f) ACES_CHECK(extend_to(files[h].f, prec[(size_t)h] * recwd * 8, kListFiles[h]));
  ACES_CHECK(recs.putrec("MOIOVEC", n, moio.data()));
  ACES_CHECK(recs.putrec("MOIOWRD", n, moiowd.data()));
  ACES_CHECK(recs.putrec("MOIOSIZ", n, moiosz.data()));
  ACES_CHECK(recs.putrec("MOIODIS", n, moiods.data()));
  ACES_CHECK(recs.putrec("MOIOFIL", n, moiofl.data()));
  ACES_CHECK(recs.putrec("MOIOMXSZ", n, moiomx.data()));
  ACES_CHECK(recs.putrec("ISYMTYP", 2L * kFams, isytyp.data()));
  ACES_CHECK(recs.putrec("TOTRECMO", kLuns, prec.data()));
  ACES_CHECK(recs.putrec("TOTWRDMO", kLuns, ioff.data()));
  // the lists' symmetry
  const DpdSym& sy = lists.sym();
  const long nirr = sy.nirrep;
  ACES_CHECK(recs.putrec("COMPNIRR", 1, &nirr));
  long pops[4][8];
  for (int h2 = 0; h2 < 8; ++h2) {
    pops[0][h2] = sy.pop[0][h2];
    pops[1][h2] = sy.pop[1][h2];
    pops[2][h2] = sy.vrt[0][h2];
    pops[3][h2] = sy.vrt[1][h2];
  }
  ACES_CHECK(recs.putrec("SYMPOPOA", nirr, pops[0]));
  ACES_CHECK(recs.putrec("SYMPOPOB", nirr, pops[1]));
  ACES_CHECK(recs.putrec("SYMPOPVA", nirr, pops[2]));
  ACES_CHECK(recs.putrec("SYMPOPVB", nirr, pops[3]));
  return ACES_OK;
}

}  // namespace aces
