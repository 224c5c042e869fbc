/*
 * mc_oracle.c -- CPU oracle for marching-cubes extraction (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this file; the product never links it.  Each function restates the
 * reference's algorithm and cites the lines it follows in
 * cu_marching_cubes_v1.cu / cu_kinect_fusion_v1.cu.
 *
 * Pinning (tests/test_oracle.py): the tables against the fixture's sha256 and
 * the case invariants; cube indices by brute force over all 256 sign patterns
 * and every weight-0 placement; the single-corner case against the closed
 * form t = f0/(f0-f1) on edges 0, 8, 3 with reversed winding; planar fields
 * lie on the plane; a sphere gives a closed 2-manifold with Euler
 * characteristic 2 and O(h^2) distance error; the fp32 path agrees with the
 * FP64 path; the FP64 gradient agrees with central differences.
 */
#include "mc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "mc_tables_ref.inc"

/* CU_VOXEL_MAX_TSDF = 0x7ffe (cu_kinect_fusion_v1.h:44); the unpack divides
 * by it (d_unpack_tsdf, cu_kinect_fusion_v1.h:253-256). */
#define ORC_TSDF_DIV 32766.0f

/* Corner offsets in the reference's corner order (computeCubeIndex,
 * cu_marching_cubes_v1.cu:77-84 and getNodeCoo calls :279-286). */
static const int kCx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int kCy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int kCz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
/* Edge e interpolates from corner kEa[e] to corner kEb[e] (:291-302). */
static const int kEa[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
static const int kEb[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};

const int* orc_tables_tri(void) { return &orc_tri_table[0][0]; }
const int* orc_tables_num_verts(void) { return orc_num_verts_table; }
const int* orc_tables_edge(void) { return orc_edge_table; }

static inline int64_t lin(int X, int Y, int x, int y, int z) {
  return ((int64_t)z * Y + y) * X + x;
}

/* readTsdf (cu_marching_cubes_v1.cu:59-68): f = tsdf / 32766, weight raw. */
static inline void read_voxel(const void* vol, int elem, int64_t i, float* f,
                              int* w) {
  if (elem == ORC_ELEM_SHORT2) {
    const int16_t* p = (const int16_t*)vol + 2 * i;
    *f = (float)p[0] / ORC_TSDF_DIV;
    *w = (int)(uint16_t)p[1];
  } else if (elem == ORC_ELEM_F64) {
    *f = (float)((const double*)vol)[i];
    *w = 1;
  } else {
    *f = ((const float*)vol)[i];
    *w = 1;
  }
}

/* computeCubeIndex (cu_marching_cubes_v1.cu:70-94): any weight 0 -> 0; bit k
 * set when corner k's value is below the iso value 0 (:57). */
int orc_cube_index(const void* vol, int elem, int X, int Y, int Z, int x, int y,
                   int z, float f[8]) {
  (void)Z;
  int ci = 0;
  for (int k = 0; k < 8; ++k) {
    int w;
    read_voxel(vol, elem, lin(X, Y, x + kCx[k], y + kCy[k], z + kCz[k]), &f[k],
               &w);
    if (w == 0) return 0;
  }
  for (int k = 0; k < 8; ++k) ci += (f[k] < 0.f) << k;
  return ci;
}

/* numVerts rule of OccupiedVoxels (:127-131). */
static inline int num_verts(int ci) {
  return (ci == 0 || ci == 255) ? 0 : orc_num_verts_table[ci];
}

int64_t orc_classify(const void* vol, int elem, int X, int Y, int Z, uint8_t* ci,
                     uint8_t* nv) {
  int64_t n = (int64_t)X * Y * Z, active = 0;
  if (ci) memset(ci, 0, (size_t)n);
  if (nv) memset(nv, 0, (size_t)n);
  for (int z = 0; z + 1 < Z; ++z)
    for (int y = 0; y + 1 < Y; ++y)
      for (int x = 0; x + 1 < X; ++x) {
        float f[8];
        int c = orc_cube_index(vol, elem, X, Y, Z, x, y, z, f);
        int v = num_verts(c);
        int64_t i = lin(X, Y, x, y, z);
        if (ci) ci[i] = (uint8_t)c;
        if (nv) nv[i] = (uint8_t)v;
        active += v > 0;
      }
  return active;
}

/* getNodeCoo (:243-254): cell-centre convention (i + 0.5) * h. */
static inline float node_coord(int i, float h) { return ((float)i + 0.5f) * h; }

/* vertex_interp (:256-264).  nvcc contracts p0 + t*(p1-p0) into one FMA; the
 * oracle fixes that choice with fmaf. */
static inline void vertex_interp(const float p0[3], const float p1[3], float f0,
                                 float f1, float out[3]) {
  float t = (0.0f - f0) / (f1 - f0 + 1e-15f);
  for (int d = 0; d < 3; ++d) out[d] = fmaf(t, p1[d] - p0[d], p0[d]);
}

typedef struct {
  const void* vol;
  int elem, X, Y, Z;
  float h;
  orc_vertex_voxel* vv;
  float* pos;
  int32_t* v2vox;
  int16_t* ver_voxel;
  int64_t cap;
} emit_ctx;

/* store_point (:336-367): record position, vox_pos = floor(p/h - 0.5),
 * tsdf/w of that voxel, int3 map and the ver_voxel scatter.  Where vox_pos
 * falls outside the volume (possible at index 0 when t == 0 rounds below the
 * node) the reference reads out of bounds; here tsdf = w = 0 and ver_voxel is
 * not written. */
static void store_point(const emit_ctx* c, int64_t index, const float p[3]) {
  if (index >= c->cap) return;
  float coo[3];
  int v[3];
  for (int d = 0; d < 3; ++d) {
    coo[d] = p[d] / c->h;
    coo[d] = coo[d] - 0.5f;
    v[d] = (int)floorf(coo[d]);
  }
  int16_t tsdf = 0, w = 0;
  int inb = v[0] >= 0 && v[1] >= 0 && v[2] >= 0 && v[0] < c->X && v[1] < c->Y &&
            v[2] < c->Z;
  int64_t li = inb ? lin(c->X, c->Y, v[0], v[1], v[2]) : -1;
  if (inb && c->elem == ORC_ELEM_SHORT2) {
    const int16_t* q = (const int16_t*)c->vol + 2 * li;
    tsdf = q[0];
    w = q[1];
  }
  if (c->vv) {
    orc_vertex_voxel* r = &c->vv[index];
    r->x = p[0];
    r->y = p[1];
    r->z = p[2];
    r->vx = v[0];
    r->vy = v[1];
    r->vz = v[2];
    r->tsdf = tsdf;
    r->w = w;
  }
  if (c->pos) {
    c->pos[3 * index + 0] = p[0];
    c->pos[3 * index + 1] = p[1];
    c->pos[3 * index + 2] = p[2];
  }
  if (c->v2vox) {
    c->v2vox[3 * index + 0] = v[0];
    c->v2vox[3 * index + 1] = v[1];
    c->v2vox[3 * index + 2] = v[2];
  }
  if (c->ver_voxel && inb) __atomic_store_n(&c->ver_voxel[li], tsdf, __ATOMIC_RELAXED);
}

/* TrianglesGenerator::operator() (:266-333) for one cube at offset off. */
static void emit_cube(const emit_ctx* c, int x, int y, int z, int64_t off) {
  float f[8];
  int ci = orc_cube_index(c->vol, c->elem, c->X, c->Y, c->Z, x, y, z, f);
  int nv = num_verts(ci);
  if (nv == 0) return;
  float v[8][3], vert[12][3];
  for (int k = 0; k < 8; ++k) {
    v[k][0] = node_coord(x + kCx[k], c->h);
    v[k][1] = node_coord(y + kCy[k], c->h);
    v[k][2] = node_coord(z + kCz[k], c->h);
  }
  for (int e = 0; e < 12; ++e)
    vertex_interp(v[kEa[e]], v[kEb[e]], f[kEa[e]], f[kEb[e]], vert[e]);
  for (int i = 0; i < nv; i += 3) {
    int e1 = orc_tri_table[ci][i + 0];
    int e2 = orc_tri_table[ci][i + 1];
    int e3 = orc_tri_table[ci][i + 2];
    store_point(c, off + i + 2, vert[e1]);
    store_point(c, off + i + 1, vert[e2]);
    store_point(c, off + i + 0, vert[e3]);
  }
}

static int64_t plane_count(const void* vol, int elem, int X, int Y, int Z, int z) {
  int64_t s = 0;
  for (int y = 0; y + 1 < Y; ++y)
    for (int x = 0; x + 1 < X; ++x) {
      float f[8];
      s += num_verts(orc_cube_index(vol, elem, X, Y, Z, x, y, z, f));
    }
  return s;
}

static void plane_emit(const emit_ctx* c, int z, int64_t off) {
  for (int y = 0; y + 1 < c->Y; ++y)
    for (int x = 0; x + 1 < c->X; ++x) {
      float f[8];
      int nv = num_verts(orc_cube_index(c->vol, c->elem, c->X, c->Y, c->Z, x, y, z, f));
      if (nv) {
        emit_cube(c, x, y, z, off);
        off += nv;
      }
    }
}

int64_t orc_extract(const void* vol, int elem, int X, int Y, int Z, float h,
                    orc_vertex_voxel* vv, float* pos, int32_t* v2vox,
                    int16_t* ver_voxel, int64_t cap, int nthreads) {
  if (X < 1 || Y < 1 || Z < 1) return 0;
  int64_t n = (int64_t)X * Y * Z;
  if (ver_voxel) memset(ver_voxel, 0, (size_t)n * sizeof(int16_t));
  if (Z < 2 || Y < 2 || X < 2) return 0;
  int64_t* plane_off = (int64_t*)calloc((size_t)Z + 1, sizeof(int64_t));
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int z = 0; z < Z - 1; ++z) plane_off[z + 1] = plane_count(vol, elem, X, Y, Z, z);
  for (int z = 0; z < Z - 1; ++z) plane_off[z + 1] += plane_off[z];
  int64_t total = plane_off[Z - 1];
  emit_ctx c = {vol, elem, X, Y, Z, h, vv, pos, v2vox, ver_voxel, cap};
  if (vv || pos || v2vox || ver_voxel) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int z = 0; z < Z - 1; ++z)
      if (plane_off[z + 1] > plane_off[z] && plane_off[z] < cap)
        plane_emit(&c, z, plane_off[z]);
  }
  free(plane_off);
  return total;
}

/* ---- FP64 path: same topology/order, double arithmetic ------------------ */

static inline double read_f64(const void* vol, int elem, int64_t i) {
  if (elem == ORC_ELEM_SHORT2) return (double)((const int16_t*)vol)[2 * i] / 32766.0;
  if (elem == ORC_ELEM_F64) return ((const double*)vol)[i];
  return (double)((const float*)vol)[i];
}

int64_t orc_extract_f64(const void* vol, int elem, int X, int Y, int Z, double h,
                        double* pos, int64_t* cube, int8_t* edge, int64_t cap) {
  int64_t off = 0;
  if (X < 2 || Y < 2 || Z < 2) return 0;
  for (int z = 0; z + 1 < Z; ++z)
    for (int y = 0; y + 1 < Y; ++y)
      for (int x = 0; x + 1 < X; ++x) {
        float ff[8];
        int ci = orc_cube_index(vol, elem, X, Y, Z, x, y, z, ff);
        int nv = num_verts(ci);
        if (!nv) continue;
        double f[8], v[8][3];
        for (int k = 0; k < 8; ++k) {
          f[k] = read_f64(vol, elem, lin(X, Y, x + kCx[k], y + kCy[k], z + kCz[k]));
          v[k][0] = ((double)(x + kCx[k]) + 0.5) * h;
          v[k][1] = ((double)(y + kCy[k]) + 0.5) * h;
          v[k][2] = ((double)(z + kCz[k]) + 0.5) * h;
        }
        for (int i = 0; i < nv; i += 3) {
          for (int m = 0; m < 3; ++m) {
            int e = orc_tri_table[ci][i + m];
            int64_t idx = off + i + 2 - m;
            if (idx >= cap) continue;
            int a = kEa[e], b = kEb[e];
            double t = (0.0 - f[a]) / (f[b] - f[a] + 1e-15);
            if (pos)
              for (int d = 0; d < 3; ++d) pos[3 * idx + d] = v[a][d] + t * (v[b][d] - v[a][d]);
            if (cube) cube[idx] = lin(X, Y, x, y, z);
            if (edge) edge[idx] = (int8_t)e;
          }
        }
        off += nv;
      }
  return off;
}

void orc_extract_backward_f64(const float* vol, int X, int Y, int Z, double h,
                              const double* gpos, int64_t nverts, double* grad) {
  int64_t n = (int64_t)X * Y * Z, off = 0;
  for (int64_t i = 0; i < n; ++i) grad[i] = 0.0;
  if (X < 2 || Y < 2 || Z < 2) return;
  for (int z = 0; z + 1 < Z; ++z)
    for (int y = 0; y + 1 < Y; ++y)
      for (int x = 0; x + 1 < X; ++x) {
        float ff[8];
        int ci = orc_cube_index(vol, ORC_ELEM_F32, X, Y, Z, x, y, z, ff);
        int nv = num_verts(ci);
        if (!nv) continue;
        int64_t li[8];
        double f[8];
        for (int k = 0; k < 8; ++k) {
          li[k] = lin(X, Y, x + kCx[k], y + kCy[k], z + kCz[k]);
          f[k] = (double)vol[li[k]];
        }
        for (int i = 0; i < nv; i += 3)
          for (int m = 0; m < 3; ++m) {
            int e = orc_tri_table[ci][i + m];
            int64_t idx = off + i + 2 - m;
            if (idx >= nverts) continue;
            int a = kEa[e], b = kEb[e];
            double dp[3] = {(double)(kCx[b] - kCx[a]) * h, (double)(kCy[b] - kCy[a]) * h,
                            (double)(kCz[b] - kCz[a]) * h};
            double s = gpos[3 * idx] * dp[0] + gpos[3 * idx + 1] * dp[1] +
                       gpos[3 * idx + 2] * dp[2];
            double d = f[b] - f[a] + 1e-15;
            grad[li[a]] += s * (-(f[b] + 1e-15) / (d * d));
            grad[li[b]] += s * (f[a] / (d * d));
          }
        off += nv;
      }
}

/* ---- normals: _d_voxel_extract_surface_normal (cu_kinect_fusion_v1.cu:3146-3235)
 * p_f = ver_pos / grid_s (:3241, without the +0.5 shift, as the reference);
 * p_i = (int)p_f; zero normal when any of the 8 tsdf values is +-32766; the
 * bilinear x/y/z differences of :3201-3218; float3_unit divides only when the
 * norm exceeds FLT_EPSILON (cu_vector_functions.h:258-264).  Where p_i+1 leaves
 * the volume (the reference reads out of bounds) the normal is zero. */
static inline float bilinear(const float v[4], float x, float y) {
  float a = 1.0f - x, b = 1.0f - y;
  return v[0] * a * b + v[1] * x * b + v[2] * a * y + v[3] * x * y;
}

void orc_normals(const void* vol, int X, int Y, int Z, float h, const float* pos,
                 int64_t stride_floats, int64_t n, float* normals) {
  const int16_t* s = (const int16_t*)vol;
  for (int64_t i = 0; i < n; ++i) {
    float pf[3], t[3];
    int pi[3];
    for (int d = 0; d < 3; ++d) {
      pf[d] = pos[i * stride_floats + d] / h;
      pi[d] = (int)pf[d];
      t[d] = pf[d] - (float)pi[d];
    }
    float* out = &normals[3 * i];
    out[0] = out[1] = out[2] = 0.0f;
    if (pf[0] < 0.f || pf[1] < 0.f || pf[2] < 0.f) continue;
    if (pi[0] + 1 >= X || pi[1] + 1 >= Y || pi[2] + 1 >= Z) continue;
    /* sdfs[] slots: 0 V000, 1 V100, 6 V110, 2 V010, 3 V001, 4 V101, 7 V111, 5 V011 */
    static const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    static const int oy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    static const int oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    static const int slot[8] = {0, 1, 6, 2, 3, 4, 7, 5};
    float sd[8];
    int bad = 0;
    for (int k = 0; k < 8 && !bad; ++k) {
      int16_t v = s[2 * lin(X, Y, pi[0] + ox[k], pi[1] + oy[k], pi[2] + oz[k])];
      if (v == 32766 || v == -32766) bad = 1;
      sd[slot[k]] = (float)v / ORC_TSDF_DIV;
    }
    if (bad) continue;
    float vxy[4], s1, s2, nf[3];
    vxy[0] = sd[0]; vxy[1] = sd[2]; vxy[2] = sd[3]; vxy[3] = sd[5];
    s1 = bilinear(vxy, t[1], t[2]);
    vxy[0] = sd[1]; vxy[1] = sd[6]; vxy[2] = sd[4]; vxy[3] = sd[7];
    s2 = bilinear(vxy, t[1], t[2]);
    nf[0] = s2 - s1;
    vxy[0] = sd[0]; vxy[1] = sd[1]; vxy[2] = sd[3]; vxy[3] = sd[4];
    s1 = bilinear(vxy, t[0], t[2]);
    vxy[0] = sd[2]; vxy[1] = sd[6]; vxy[2] = sd[5]; vxy[3] = sd[7];
    s2 = bilinear(vxy, t[0], t[2]);
    nf[1] = s2 - s1;
    vxy[0] = sd[0]; vxy[1] = sd[1]; vxy[2] = sd[2]; vxy[3] = sd[6];
    s1 = bilinear(vxy, t[0], t[1]);
    vxy[0] = sd[3]; vxy[1] = sd[4]; vxy[2] = sd[5]; vxy[3] = sd[7];
    s2 = bilinear(vxy, t[0], t[1]);
    nf[2] = s2 - s1;
    float nn = sqrtf(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
    if (nn > FLT_EPSILON || nn < -FLT_EPSILON) {
      nf[0] /= nn;
      nf[1] /= nn;
      nf[2] /= nn;
    }
    out[0] = nf[0];
    out[1] = nf[1];
    out[2] = nf[2];
  }
}
