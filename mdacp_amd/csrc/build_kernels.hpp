// build_kernels.hpp -- pair-list construction on gfx950: periodic wrap,
// counting-sort cell binning (atomics + prefix scan, then a per-cell order by
// canonical key so the result does not depend on atomic ordering), ghost
// images of the single-domain halo, and the full neighbour list in sliced-ELL
// layout (64 rows per slice, one wavefront).
//
// Reference semantics (kaityo256/mdacp): AdjustPeriodicBoundary
// (variables.cc:172-193), MakeMesh (meshlist.cc:217-270), FindBorderParticles
// and MakeBufferForBorderParticles (mdunit.cc:202-254, mdrect.cc:356-380),
// SearchMesh (meshlist.cc:284-336).  The distance test is evaluated without
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

// Order each cell's members by a canonical 64-bit key (insertion sort; cells
// hold tens of particles).
__global__ void k_sort_cells(int *__restrict__ perm, const int *__restrict__ cstart,
                             const int *__restrict__ ccount, const long long *__restrict__ key,
                             int ncell) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int *a = perm + cstart[c];
  const int n = ccount[c];
  for (int t = 1; t < n; t++) {
    const int v = a[t];
    const long long kv = key[v];
    int u = t - 1;
    while (u >= 0 && key[a[u]] > kv) {
      a[u + 1] = a[u];
      u--;
    }
    a[u + 1] = v;
  }
}

__global__ void k_owned_key(const int *__restrict__ id, int n, long long *__restrict__ key) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = (long long)id[i];
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
                              int *__restrict__ gslot, int *__restrict__ gcount,
                              long long *__restrict__ gkey, const int *__restrict__ id) {
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
        gkey[k] = (long long)id[i] * 32 + code;
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

// ---- neighbour search: one lane per owned row, 27 cells of the extended
// grid (no wrap), owned members then ghost members of each cell in order.
struct CellTables {
  const int *ostart, *ocount, *gstart, *gcount;
};

// Tiles: kWave * W consecutive owned rows (cell order) per workgroup.  The
// partners of a tile's rows lie in 9 contiguous cell ranges (one per (dy,dz)
// neighbour row of the cell grid), i.e. in 9 owned and 9 ghost index ranges;
// the force kernels stage those 18 ranges into LDS and the list stores each
// partner as a 16-bit index into the staged array.
constexpr int kTileRanges = 18;

__device__ __forceinline__ void range_of(int k, int c0, int c1, Grid g, CellTables T, int pn, int &start,
                                         int &count) {
  const int kk = k % 9;
  const int dz = kk / 3 - 1, dy = kk % 3 - 1;
  const int ncell = g.m[0] * g.m[1] * g.m[2];
  const int o = dy * g.m[0] + dz * g.m[0] * g.m[1];
  const int lo = max(0, c0 - 1 + o), hi = min(ncell - 1, c1 + 1 + o);
  if (lo > hi) {
    start = 0;
    count = 0;
    return;
  }
  if (k < 9) {
    start = T.ostart[lo];
    count = T.ostart[hi] + T.ocount[hi] - start;
  } else {
    start = pn + T.gstart[lo];
    count = T.gstart[hi] + T.gcount[hi] - T.gstart[lo];
  }
}

// tstart[t][18] = first particle index of each range, tbase[t][19] = offset of
// each range in the staged array (tbase[t][18] = staged count = the pad slot)
__global__ void k_tile_ranges(const int *__restrict__ cell, int pn, int rows_per_tile, int nt, Grid g,
                              CellTables T, int *__restrict__ tstart, int *__restrict__ tbase,
                              double *__restrict__ maxpart) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double total = 0.0;
  if (t < nt) {
    const int r0 = t * rows_per_tile, r1 = min(pn, r0 + rows_per_tile) - 1;
    const int c0 = cell[r0], c1 = cell[r1];
    int base = 0;
    for (int k = 0; k < kTileRanges; k++) {
      int st, n;
      range_of(k, c0, c1, g, T, pn, st, n);
      tstart[t * kTileRanges + k] = st;
      tbase[t * (kTileRanges + 1) + k] = base;
      base += n;
    }
    tbase[t * (kTileRanges + 1) + kTileRanges] = base;
    total = (double)base;
  }
  block_max_to_partial(total, maxpart);
}

template <bool kFill>
__device__ __forceinline__ int search_row(int i, const double4 *__restrict__ pos, const int *__restrict__ cell,
                                          int pn, Grid g, CellTables T, double S2, uint16_t *__restrict__ out,
                                          uint8_t *__restrict__ code_out, long long stride, int cap,
                                          const int *__restrict__ ts, const int *__restrict__ tb,
                                          int *__restrict__ err) {
  const double4 qi = pos[i];
  const int c = cell[i];
  const int mx = g.m[0], my = g.m[1], mz = g.m[2];
  const int ix = c % mx, iy = (c / mx) % my, iz = c / (mx * my);
  int cnt = 0;
  for (int dz = -1; dz <= 1; dz++) {
    const int jz = iz + dz;
    if (jz < 0 || jz >= mz) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int jy = iy + dy;
      if (jy < 0 || jy >= my) continue;
      const int k = (dz + 1) * 3 + (dy + 1);
      for (int dx = -1; dx <= 1; dx++) {
        const int jx = ix + dx;
        if (jx < 0 || jx >= mx) continue;
        const int c2 = mx * my * jz + mx * jy + jx;
        const uint8_t code = (uint8_t)((dx + 1) | ((dy + 1) << 2) | ((dz + 1) << 4));
        for (int pass = 0; pass < 2; pass++) {
          const int b = pass ? pn + T.gstart[c2] : T.ostart[c2];
          const int e = b + (pass ? T.gcount[c2] : T.ocount[c2]);
          const int kk = pass ? 9 + k : k;
          for (int j = b; j < e; j++) {
            if (j == i) continue;
            const double4 qj = pos[j];
            const double ex = qi.x - qj.x, ey = qi.y - qj.y, ez = qi.z - qj.z;
            const double r2 = (ex * ex + ey * ey + ez * ez);
            if (r2 > S2) continue;
            if (kFill) {
              const int off = j - ts[kk];
              const int n = tb[kk + 1] - tb[kk];
              if (off < 0 || off >= n) {
                atomicOr(err, kErrCapacity);  // partner outside its tile range: cannot happen
              } else if (cnt < cap) {
                out[(long long)cnt * stride] = (uint16_t)(tb[kk] + off);
                code_out[(long long)cnt * stride] = code;
              }
            }
            cnt++;
          }
        }
      }
    }
  }
  return cnt;
}

__global__ void k_count_rows(const double4 *__restrict__ pos, const int *__restrict__ cell, int pn, Grid g,
                             CellTables T, double S2, int *__restrict__ row_len) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= pn) return;
  row_len[i] = search_row<false>(i, pos, cell, pn, g, T, S2, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr);
}

// slice length = longest row of its 64 rows, rounded up to a multiple of
// kSliceRound so the force loop runs whole batches; slots = 64 * length
__global__ void k_slice_len(const int *__restrict__ row_len, int pn, int nslice, int *__restrict__ slen,
                            long long *__restrict__ slots) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslice) return;
  int m = 0;
  const int b = s * kWave, e = min(pn, b + kWave);
  for (int i = b; i < e; i++) m = max(m, row_len[i]);
  m = (m + kSliceRound - 1) / kSliceRound * kSliceRound;
  slen[s] = m;
  slots[s] = (long long)m * kWave;
}

__global__ void k_fill_rows(const double4 *__restrict__ pos, const int *__restrict__ cell, int pn, Grid g,
                            CellTables T, double S2, const long long *__restrict__ sbase,
                            const int *__restrict__ slen, int rows_per_tile, const int *__restrict__ tstart,
                            const int *__restrict__ tbase, uint16_t *__restrict__ nbr, uint8_t *__restrict__ ncode,
                            int *__restrict__ err) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= pn) return;
  const int s = i / kWave, lane = i % kWave, t = i / rows_per_tile;
  const int *ts = tstart + (long long)t * kTileRanges;
  const int *tb = tbase + (long long)t * (kTileRanges + 1);
  const int len = slen[s];
  uint16_t *row = nbr + sbase[s] + lane;
  uint8_t *crow = ncode + sbase[s] + lane;
  const int cnt = search_row<true>(i, pos, cell, pn, g, T, S2, row, crow, kWave, len, ts, tb, err);
  if (cnt > len) atomicOr(err, kErrCapacity);
  const uint16_t pad = (uint16_t)tb[kTileRanges];
  for (int k = cnt; k < len; k++) {
    row[(long long)k * kWave] = pad;
    crow[(long long)k * kWave] = (uint8_t)kZeroOffset;
  }
}

// rows past pn inside the last slice are all padding
__global__ void k_fill_tail(int pn, int nslice, int rows_per_tile, const long long *__restrict__ sbase,
                            const int *__restrict__ slen, const int *__restrict__ tbase, uint16_t *__restrict__ nbr,
                            uint8_t *__restrict__ ncode) {
  const int lane = threadIdx.x;
  const int s = nslice - 1;
  const int i = s * kWave + lane;
  if (lane >= kWave || i < pn) return;
  const int t = (s * kWave) / rows_per_tile;
  const uint16_t pad = (uint16_t)tbase[(long long)t * (kTileRanges + 1) + kTileRanges];
  const int len = slen[s];
  for (int k = 0; k < len; k++) {
    nbr[sbase[s] + (long long)k * kWave + lane] = pad;
    ncode[sbase[s] + (long long)k * kWave + lane] = (uint8_t)kZeroOffset;
  }
}

}  // namespace mdgpu
