// Device-resident DPD list store and JOBARC-style record store.
//
// The list interface of ACES II (acescore/getlst.F:19, putlst.F:19,
// acesutil/getall.F, acesutil/sumsym2.F, libr/inipck.F:3-33,
// libr/updmoi.F:11) with the lists held in HBM instead of the MOINTS files:
// a list is addressed by (group, family) = (irrep or group 1..ngroups, list
// 1..500); each group is a column-major block of nrows x ncols doubles.
// Pair-type sizes follow acescore/aces_com_symloc.F:11-106 (25 types, PACK
// p<q, PCK2 p<=q, FULL) with DIRPRD = XOR of 0-based irreps.
#pragma once
#include <hip/hip_runtime.h>
#include <map>
#include <string>
#include <vector>

namespace aces {

// Pair type `type` (1..25): orbital spaces of its left (p) and right (q) member,
// each 0 = occupied alpha, 1 = virtual alpha, 2 = occupied beta, 3 = virtual
// beta, and its packing (0 full, 1 p < q, 2 p <= q).  False for an unknown type.
bool pair_type(int type, int* space_p, int* space_q, int* pack);

struct DpdSym {
  int nirrep = 1;
  int pop[2][8] = {{0}};  // occupied per irrep, spin 0 = alpha, 1 = beta
  int vrt[2][8] = {{0}};  // virtual per irrep
  // size of pair type `type` (1..25) in total irrep irp_pq (0-based)
  long irpdpd(int irp_pq, int type) const;
  // offset (0-based) of the sub-block with right-index irrep irp_q inside irrep irp_pq
  long isymoff(int irp_q, int irp_pq, int type) const;
  static int dirprd(int a, int b) { return a ^ b; }
};

class ListStore {
 public:
  explicit ListStore(hipStream_t s) : s_(s) {}
  ~ListStore();
  int set_symmetry(int nirrep, const int* popa, const int* popb, const int* vrta, const int* vrtb);
  const DpdSym& sym() const { return sym_; }

  // UPDMOI(ncols, nrows, group, list): (re)create one group block, zero filled.
  int updmoi(long ncols, long nrows, int group, int list);
  // INIPCK(irrepx, typel, typer, list): one block per irrep (right index irrep r,
  // left irrep r xor irrepx), records ISYTYP(list) = (typel, typer).
  int inipck(int irrepx, int typel, int typer, int list);
  int isytyp(int list, int* tl, int* tr) const;
  // Record ISYTYP(list) without creating groups (lists read from MOINTS files).
  void set_types(int list, int tl, int tr) {
    lists_[list].tl = tl;
    lists_[list].tr = tr;
  }

  int rows(int group, int list, long* n) const;
  int cols(int group, int list, long* n) const;
  double* device_ptr(int group, int list) const;

  // GETLST/PUTLST: ncols whole columns from column ifirst (1-based); dest/src on
  // the host (dev = false) or on the device (dev = true).
  int getlst(double* dest, long ifirst, long ncols, int group, int list, bool dev);
  int putlst(const double* src, long ifirst, long ncols, int group, int list, bool dev);
  // GETALL/PUTALL: all irreps of a list with overall symmetry irrepx (1-based)
  int getall(double* dest, long length, int irrepx, int list, bool dev);
  int putall(const double* src, long length, int irrepx, int list, bool dev);
  // SUMSYM2: list(group) += add (the whole block)
  int sumsym2(const double* add, int group, int list, bool dev);
  int zero(int group, int list);
  // Copy the block (src_group, src_list) into (dst_group, dst_list), creating or
  // reshaping the destination to the source's shape; a destination without pair
  // types takes the source's (the list moves of libr/drmove.f: MVSYM, MVLIST).
  int move(int src_group, int src_list, int dst_group, int dst_list);
  // Groups that exist for a list (ascending).
  std::vector<int> groups(int list) const;
  long device_bytes() const { return bytes_; }

 private:
  struct Block {
    double* d = nullptr;
    long nrows = 0, ncols = 0;
  };
  struct List {
    int tl = 0, tr = 0;
    std::map<int, Block> groups;
  };
  int block(int group, int list, Block** b);
  hipStream_t s_;
  DpdSym sym_;
  std::map<int, List> lists_;
  long bytes_ = 0;
};

// The job archive (JOBARC records) as ACES II keeps it: logical records in
// slot order, each a name of at most 8 characters and a length in 8-byte words.
class RecordStore {
 public:
  // GETREC semantics (acescore/getrec.F:5-20): iflag > 0 record must exist;
  // iflag == 0 returns the length (-1 if absent) in *len; iflag < 0 zero fills
  // a missing record.  A shorter read than the stored length is allowed.
  int getrec(int iflag, const std::string& name, long* len, void* dest);
  // PUTREC semantics (acescore/putrec.F): a new name takes the next slot; an
  // existing record keeps its slot and stored length when the new length is
  // not larger (the first len words are replaced), and moves to a new slot
  // when it is (the old slot stays as the lost record '. LOST .').
  int putrec(const std::string& name, long len, const void* src);
  bool has(const std::string& name) const { return index_.count(key(name)) != 0; }
  static std::string key(const std::string& name);

  struct Slot {
    std::string name;  // key form (no trailing blanks); ". LOST ." for a lost record
    std::vector<long> data;
  };
  const std::vector<Slot>& slots() const { return slots_; }
  // Replace the whole archive (a JOBARC read).
  void assign(std::vector<Slot> slots);

 private:
  std::vector<Slot> slots_;
  std::map<std::string, size_t> index_;
};

}  // namespace aces
