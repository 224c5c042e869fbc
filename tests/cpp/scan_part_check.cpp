// Host check of the scan kernel's split of the non-empty-tile list over its
// workgroups (mcx_scan_part.h, the same functions the kernel calls).
//
// For tile counts that give 1..600 scan workgroups and list lengths nz from 1
// to the tile count (every nz below 600, then about 400 more per count): every list entry lies in exactly one in-range slot,
// slots are ascending within a workgroup, and no thread has more than four.
// It also counts the (nz, workgroups) pairs in which a thread's slot indices
// (ignoring the range) reach entry nz - 1 in a second workgroup: a rule keyed
// on the slot index alone picks two writers there, which is why the list's
// totals come from the workgroup with the last ticket.
// Prints "ok <pairs> <ambiguous pairs>" and exits 0, or names the first failure.
#include <cstdio>
#include <vector>

#include "mcx_scan_part.h"

using mcx::ListPart;

int main() {
  long pairs = 0, ambiguous = 0;
  std::vector<int> hits;
  for (int nblocks = 1; nblocks <= 600; nblocks += (nblocks < 24 ? 1 : 53)) {
    const long tiles = 1024L * nblocks;
    const long step = tiles > 400 ? tiles / 400 + 1 : 1;
    for (long nz = 1; nz <= tiles; nz += (nz < 600 ? 1 : step)) {
      hits.assign(nz, 0);
      int writers_by_index = 0;
      for (int blk = 0; blk < nblocks; ++blk) {
        const ListPart p = mcx::list_part(nz, nblocks, blk);
        if (p.c > 4) {
          std::printf("FAIL c=%d > 4 (nz %ld, blocks %d)\n", p.c, nz, nblocks);
          return 1;
        }
        long prev = -1;
        bool reaches_last = false;
        for (int tid = 0; tid < 256; ++tid) {
          for (int k = 0; k < 4; ++k) {
            const long j = mcx::list_slot(p, tid, k);
            if (k < p.c && p.lb + (long)tid * p.c + k == nz - 1) reaches_last = true;
            if (j < 0) continue;
            if (j >= nz || j <= prev) {
              std::printf("FAIL slot %ld after %ld (nz %ld, blocks %d, blk %d)\n", j, prev, nz,
                          nblocks, blk);
              return 1;
            }
            prev = j;
            ++hits[j];
          }
        }
        if (reaches_last) ++writers_by_index;
      }
      for (long j = 0; j < nz; ++j) {
        if (hits[j] != 1) {
          std::printf("FAIL entry %ld taken %d times (nz %ld, blocks %d)\n", j, hits[j], nz,
                      nblocks);
          return 1;
        }
      }
      ++pairs;
      if (writers_by_index > 1) ++ambiguous;
    }
  }
  std::printf("ok %ld %ld\n", pairs, ambiguous);
  return 0;
}
