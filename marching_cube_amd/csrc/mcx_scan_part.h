// mcx_scan_part.h -- how the scan kernel splits the list of non-empty tiles
// over its workgroups.
//
// Plain constexpr functions (host and device in HIP, plain C++ elsewhere), so
// that the kernel and the host check tests/cpp/scan_part_check.cpp run the
// same arithmetic.
//
// The list holds nz entries; nblocks workgroups of 256 threads take
// contiguous ranges of L = ceil(nz / nblocks) entries (L <= 1024 because
// nz <= tiles <= 1024 * nblocks), and thread tid of a workgroup takes c =
// ceil(L / 256) <= 4 consecutive slots j0 + k (j0 = lb + tid * c, k < c).
// The slots of the last threads can run past the workgroup's range (256 c
// may exceed L by up to 255): a slot counts only when it lies in [lb, le).
// Anything keyed on a slot index alone (such as "the slot holding entry
// nz - 1") can therefore match in two workgroups; the list's totals are
// published by the workgroup with the last ticket instead.
#pragma once

#include <stdint.h>

namespace mcx {

struct ListPart {
  int64_t lb, le;  // this workgroup's entries [lb, le)
  int c;           // slots per thread
};

constexpr int64_t min_i64(int64_t a, int64_t b) { return a < b ? a : b; }

constexpr ListPart list_part(int64_t nz, int64_t nblocks, int64_t blk) {
  const int64_t L = nz > 0 ? (nz + nblocks - 1) / nblocks : 0;
  const int c = (int)((L + 255) / 256);
  const int64_t lb = min_i64(blk * L, nz);
  return ListPart{lb, min_i64(lb + L, nz), c};
}

// Slot k of thread tid: its list entry, or -1 when it is not in the range.
constexpr int64_t list_slot(const ListPart& p, int tid, int k) {
  const int64_t j = p.lb + (int64_t)tid * p.c + k;
  return (k < p.c && j < p.le) ? j : -1;
}

}  // namespace mcx
