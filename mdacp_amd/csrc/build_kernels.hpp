// build_kernels.hpp -- pair-list construction on gfx950: periodic wrap,
// counting-sort cell binning (atomics + prefix scan, then each cell's members
// put back in input order, so the result does not depend on atomic ordering)
// and the ghost images of the single-domain halo.
//
// Reference semantics (kaityo256/mdacp): AdjustPeriodicBoundary
// (variables.cc:59-79), MakeMesh (meshlist.cc:153-205), FindBorderParticles
// and MakeBufferForBorderParticles (mdunit.cc:125-176, mdrect.cc:24-47),
// SearchMesh (meshlist.cc:220-271).  The distance test is evaluated without
// fused multiply-adds (the library is compiled with -ffp-contract=off), so the
// pair set is bit-exact with the reference's default build.
#pragma once
#include "device_common.hpp"

namespace mdgpu {

struct Grid {
  double s[3];      // domain start (MDRect::s)
  double im[3];     // 1 / mesh size
  int m[3];         // cells per axis incl. one ghost layer per side
};

__device__ __forceinline__ int cell_of(double x, double y, double z, const Grid &g) {
  int ix = (int)((x - g.s[0]) * g.im[0]) + 1;
  int iy = (int)((y - g.s[1]) * g.im[1]) + 1;
  int iz = (int)((z - g.s[2]) * g.im[2]) + 1;
  if (ix < 0) ix = g.m[0] - 1; else if (ix >= g.m[0]) ix = 0;
  if (iy < 0) iy = g.m[1] - 1; else if (iy >= g.m[1]) iy = 0;
  if (iz < 0) iz = g.m[2] - 1; else if (iz >= g.m[2]) iz = 0;
  return g.m[0] * g.m[1] * iz + g.m[0] * iy + ix;
}

// Wrap owned particles into the box (one +-L per component), flag positions
// still outside, bin, and take a slot in the cell by atomic increment.
__global__ void k_wrap_bin(double4 *__restrict__ pos, int n, double Lx, double Ly, double Lz,
                           Grid g, int *__restrict__ cell, int *__restrict__ slot,
                           int *__restrict__ ccount, int *__restrict__ err, int wrap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 q = pos[i];
  if (wrap) {
    if (q.x < 0.0) q.x += Lx; else if (q.x > Lx) q.x -= Lx;
    if (q.y < 0.0) q.y += Ly; else if (q.y > Ly) q.y -= Ly;
    if (q.z < 0.0) q.z += Lz; else if (q.z > Lz) q.z -= Lz;
    pos[i] = q;
    // the reference lets NaN through this test; a NaN is refused here
    if (!(q.x >= 0.0 && q.x <= Lx && q.y >= 0.0 && q.y <= Ly && q.z >= 0.0 && q.z <= Lz))
      atomicOr(err, kErrInvalidPosition);
  }
  const int c = cell_of(q.x, q.y, q.z, g);
  cell[i] = c;
  slot[i] = atomicAdd(&ccount[c], 1);
}

__global__ void k_scatter_perm(const int *__restrict__ cell, const int *__restrict__ slot,
                               const int *__restrict__ cstart, int n, int *__restrict__ perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  perm[cstart[cell[i]] + slot[i]] = i;
}

// Order each cell's members by their index in the input order (insertion
// sort of the cell's contiguous segment of perm; cells hold tens of
// particles).  With the slots of k_wrap_bin taken by atomics, this makes the
// order deterministic: a stable counting sort.
__global__ void k_sort_cells(int *__restrict__ perm, const int *__restrict__ cstart,
                             const int *__restrict__ ccount, int ncell) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int *a = perm + cstart[c];
  const int n = ccount[c];
  for (int t = 1; t < n; t++) {
    const int v = a[t];
    int u = t - 1;
    while (u >= 0 && a[u] > v) {
      a[u + 1] = a[u];
      u--;
    }
    a[u + 1] = v;
  }
}

__global__ void k_gather_owned(const int *__restrict__ perm, int n, const double4 *__restrict__ pos_in,
                               const double4 *__restrict__ mom_in, const int *__restrict__ id_in,
                               const int *__restrict__ cell_in, double4 *__restrict__ pos_out,
                               double4 *__restrict__ mom_out, int *__restrict__ id_out,
                               int *__restrict__ cell_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
  pos_out[k] = pos_in[i];
  mom_out[k] = mom_in[i];
  id_out[k] = id_in[i];
  cell_out[k] = cell_in[i];
}

// Halo of one domain (MDManager::MakePairList's six sequential
// FindBorderParticles/SendBorderParticlesSub passes, mdmanager.cc:269-275):
// along each axis a particle strictly within SL of the low face gets an image
// at +L, strictly within SL of the high face an image at -L; images of images
// make the edges and corners.  Equivalently, one image per non-zero shift
// vector whose every non-zero component passes its face test.
struct Edges {
  double lo_hi[3];   // s + SL
  double hi_lo[3];   // e - SL
  double s[3], e[3];
};

__device__ __forceinline__ void face_options(const double4 q, const Edges &E, int lo[3], int hi[3]) {
  const double v[3] = {q.x, q.y, q.z};
  for (int d = 0; d < 3; d++) {
    lo[d] = (v[d] < E.lo_hi[d] && v[d] > E.s[d]);
    hi[d] = (v[d] > E.hi_lo[d] && v[d] < E.e[d]);
  }
}

__global__ void k_count_images(const double4 *__restrict__ pos, int n, Edges E, int *__restrict__ nimg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo[3], hi[3];
  face_options(pos[i], E, lo, hi);
  nimg[i] = (1 + lo[0] + hi[0]) * (1 + lo[1] + hi[1]) * (1 + lo[2] + hi[2]) - 1;
}

__global__ void k_emit_images(const double4 *__restrict__ pos, int n, Edges E, double Lx, double Ly,
                              double Lz, Grid g, const int *__restrict__ goff, int ng,
                              int *__restrict__ gsrc, int *__restrict__ gcode,
                              double4 *__restrict__ gpos, int *__restrict__ gcell,
                              int *__restrict__ gslot, int *__restrict__ gcount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 q = pos[i];
  int lo[3], hi[3];
  face_options(q, E, lo, hi);
  int k = goff[i];
  const int sxs[3] = {0, 1, -1};
  for (int a = 0; a < 3; a++) {
    const int sx = sxs[a];
    if ((sx == 1 && !lo[0]) || (sx == -1 && !hi[0])) continue;
    for (int b = 0; b < 3; b++) {
      const int sy = sxs[b];
      if ((sy == 1 && !lo[1]) || (sy == -1 && !hi[1])) continue;
      for (int c = 0; c < 3; c++) {
        const int sz = sxs[c];
        if ((sz == 1 && !lo[2]) || (sz == -1 && !hi[2])) continue;
        if (!sx && !sy && !sz) continue;
        if (k >= ng) return;  // cannot happen: goff is the scan of k_count_images
        const int code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
        const double3 y = image_of(q, code, Lx, Ly, Lz);
        gsrc[k] = i;
        gcode[k] = code;
        gpos[k] = make_double4(y.x, y.y, y.z, 0.0);
        const int cl = cell_of(y.x, y.y, y.z, g);
        gcell[k] = cl;
        gslot[k] = atomicAdd(&gcount[cl], 1);
        k++;
      }
    }
  }
}

__global__ void k_gather_ghosts(const int *__restrict__ perm, int ng, int pn,
                                const int *__restrict__ src_in, const int *__restrict__ code_in,
                                const double4 *__restrict__ gpos_in, const int *__restrict__ gcell_in,
                                int *__restrict__ gsrc, int *__restrict__ gcode,
                                double4 *__restrict__ pos, int *__restrict__ cell) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ng) return;
  const int g = perm[k];
  gsrc[k] = src_in[g];
  gcode[k] = code_in[g];
  pos[pn + k] = gpos_in[g];
  cell[pn + k] = gcell_in[g];
}

__global__ void k_set_sentinel(double4 *a, double4 *b, float4 *ra, float4 *rb, int idx) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double far = 1.0e200;
    a[idx] = make_double4(far, far, far, 0.0);
    b[idx] = make_double4(far, far, far, 0.0);
    if (ra) ra[idx] = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
    if (rb) rb[idx] = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
  }
}

}  // namespace mdgpu
