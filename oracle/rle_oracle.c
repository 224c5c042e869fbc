/*
 * rle_oracle.c -- CPU oracle for the run-length volume codec of the `.vol`
 * format (TEST INFRASTRUCTURE; never linked by the product).
 *
 * Follows the reference's (disabled) codec for the live voxel type
 * cu_voxel_t {short tsdf; ushort w;} (structure.h:14-18, 4 bytes):
 *   cu_voxel_run_length_encode_slice   cu_kinect_fusion_v1.cu:1367-1413
 *       runs of equal tsdf (the weight is not compared), at most 0x3fff
 *       voxels each; record = u16 {n : 14 (low bits), v : 2 (high bits)},
 *       v = 1 for a run of +32766, v = 2 for -32766 (no payload), v = 3 for
 *       any other tsdf followed by the run's first voxel (4 bytes);
 *   cu_voxel_run_length_decode_slice   :1415-1447
 *       v = 1 -> (+32766, w 0), v = 2 -> (-32766, w 0), else a 4-byte voxel
 *       (v = 0 decodes as a literal too); each record repeated n times.
 *   header cu_voxel_run_length_write_header :1303-1310: 16 bytes
 *       "VOLUME_TSDF_RGB\0", int size[3], float min_t, max_t, grid_s;
 *   dump :1345-1351: per z slice, uint32 byte count then the slice bytes.
 * The reference decoder trusts its input; this one reports malformed
 * slices (a run of length 0, runs past the slice, a truncated record,
 * bytes left after the slice is full) with return value 1.
 */
#include <stdint.h>
#include <string.h>

#include "mc_oracle.h"

#define RLE_MAX_RUN 0x3fff
#define RLE_MAX_TSDF 32766

static int16_t vox_tsdf(uint32_t v) { return (int16_t)(v & 0xffffu); }

static uint8_t* put_record(uint8_t* out, uint32_t n, uint32_t voxel) {
  const int16_t t = vox_tsdf(voxel);
  const uint32_t v = t == RLE_MAX_TSDF ? 1u : t == -RLE_MAX_TSDF ? 2u : 3u;
  const uint16_t hdr = (uint16_t)(n | (v << 14));
  memcpy(out, &hdr, 2);
  out += 2;
  if (v == 3u) {
    memcpy(out, &voxel, 4);
    out += 4;
  }
  return out;
}

int64_t orc_rle_encode_slice(const uint32_t* vox, int64_t n, uint8_t* out) {
  if (n <= 0) return 0;
  uint8_t* p = out;
  uint32_t count = 0;
  uint32_t old = vox[0];
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t cur = vox[i];
    if (vox_tsdf(cur) == vox_tsdf(old)) {
      if (++count <= RLE_MAX_RUN) continue;
      --count; /* a full run: write it and start a new one at cur */
    }
    p = put_record(p, count, old);
    count = 1;
    old = cur;
  }
  p = put_record(p, count, old);
  return (int64_t)(p - out);
}

int orc_rle_decode_slice(const uint8_t* in, int64_t nbytes, uint32_t* vox, int64_t n) {
  int64_t pos = 0, filled = 0;
  while (filled < n) {
    if (pos + 2 > nbytes) return 1;
    uint16_t hdr;
    memcpy(&hdr, in + pos, 2);
    pos += 2;
    const uint32_t cnt = hdr & RLE_MAX_RUN, v = hdr >> 14;
    uint32_t val;
    if (v == 1u) {
      val = (uint32_t)(uint16_t)RLE_MAX_TSDF;
    } else if (v == 2u) {
      val = (uint32_t)(uint16_t)(int16_t)-RLE_MAX_TSDF;
    } else {
      if (pos + 4 > nbytes) return 1;
      memcpy(&val, in + pos, 4);
      pos += 4;
    }
    if (cnt == 0 || filled + cnt > n) return 1;
    for (uint32_t k = 0; k < cnt; ++k) vox[filled + k] = val;
    filled += cnt;
  }
  return pos == nbytes ? 0 : 1;
}

/* Dump of slices z0 .. z0+nz-1 of a [Z][Y][X] volume: per slice a uint32
 * byte count then the slice's records.  Returns the dump's size. */
int64_t orc_rle_encode_dump(const uint32_t* vol, int X, int Y, int z0, int nz, uint8_t* out) {
  const int64_t XY = (int64_t)X * Y;
  int64_t off = 0;
  for (int z = z0; z < z0 + nz; ++z) {
    const int64_t b = orc_rle_encode_slice(vol + (int64_t)z * XY, XY, out + off + 4);
    const uint32_t b32 = (uint32_t)b;
    memcpy(out + off, &b32, 4);
    off += 4 + b;
  }
  return off;
}

/* Decodes slices z0 .. z0+nz-1 of a dump holding at least z0+nz slices into
 * out[nz][Y][X].  Returns 0, or 1 for a malformed dump. */
int orc_rle_decode_dump(const uint8_t* dump, int64_t nbytes, int X, int Y, int z0, int nz,
                        uint32_t* out) {
  const int64_t XY = (int64_t)X * Y;
  int64_t off = 0;
  for (int z = 0; z < z0 + nz; ++z) {
    if (off + 4 > nbytes) return 1;
    uint32_t b;
    memcpy(&b, dump + off, 4);
    off += 4;
    if ((int64_t)b > nbytes - off) return 1;
    if (z >= z0 && orc_rle_decode_slice(dump + off, b, out + (int64_t)(z - z0) * XY, XY))
      return 1;
    off += b;
  }
  return 0;
}
