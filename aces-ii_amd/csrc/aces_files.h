// The ACES II job archive and list files on disk (ILP64 builds: 8-byte
// integers, one double per integer word), read into and written from the
// record and list stores.
//
//   JAINDX   one Fortran unformatted sequential record (4-byte length markers
//            before and after): marker(1000) CHARACTER*8, rloc(1000),
//            rsize(1000), nrecs (acescore/aces_ja_fin.F, include/jobarc.h)
//   JOBARC   direct access, physical records of 128 integers; logical record i
//            occupies integers rloc(i) .. rloc(i) + rsize(i) - 1, a record of
//            2..128 integers never straddles a physical record
//            (acescore/putrec.F, getrec.F)
//   MOINTS, GAMLAM, MOABCD, DERINT, DERGAM
//            direct access, physical records of IFLAGS(37) integers; list
//            (group g, family f) lives in file 1 + (f - 1) / 100 at physical
//            record MOIOVEC(g,f), integer MOIOWRD(g,f), MOIOSIZ(g,f) rows by
//            MOIODIS(g,f) columns, column major, contiguous across records;
//            lists are placed one after the other in each file
//            (acescore/aces_list_touch.F, getlst.F, aces_io_fin.F)
// The MOIO* arrays (10 x 500), ISYMTYP (2 x 500), TOTRECMO and TOTWRDMO (5)
// are JOBARC records; COMPNIRR and SYMPOPOA/OB/VA/VB give the symmetry.
#pragma once
#include <string>
#include "liststore.h"

namespace aces {

int jobarc_read(RecordStore& recs, const std::string& dir);
int jobarc_write(const RecordStore& recs, const std::string& dir);
// Lists named by the MOIO records (read the job archive first); sets the list
// store's symmetry from COMPNIRR/SYMPOP* when present.
int lists_read(ListStore& lists, RecordStore& recs, const std::string& dir);
// Every list of the store into the list files, and the MOIO/ISYMTYP/TOTRECMO/
// TOTWRDMO records into the archive (write the archive afterwards).  The
// physical record length is IFLAGS(37) when an IFLAGS record exists, else
// 8192 integers (64 KB), recorded in a new IFLAGS record.
int lists_write(ListStore& lists, RecordStore& recs, const std::string& dir);

}  // namespace aces
