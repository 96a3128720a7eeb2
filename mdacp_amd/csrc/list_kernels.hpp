// list_kernels.hpp -- neighbour list on gfx950, organised by cell bricks.
//
// The extended cell grid (meshlist.cc:95-120) is cut into bricks of
// bd[0] x bd[1] x bd[2] cells.  Cells are ranked brick by brick, and owned
// particles are stored in that order, so the rows of one brick are contiguous
// and one workgroup of the force kernels takes one brick.  The partners of a
// brick's rows lie in its neighbourhood: the brick grown by one cell on every
// side.  Each neighbourhood cell holds two contiguous ranges: its owned
// particles and its ghost images (materialised by the halo refresh, binned in
// the extended grid exactly as MeshList bins them).  The force kernels stage
// those ranges into LDS and the list stores each partner as a 16-bit index
// into the staged array plus the partner's cell offset from the row's cell.
// A partner is staged in the very cell where the search found it, so every
// listed partner is covered whatever the rounding of its image position.
//
// Which pairs are listed is decided exactly as in the reference
// (MeshList::SearchMesh, meshlist.cc:284-336): owned rows against owned and
// materialised ghost particles binned in the extended grid, r^2 = dx*dx +
// dy*dy + dz*dz evaluated without fused multiply-adds, kept when !(r2 > S2).
#pragma once
#include "build_kernels.hpp"

namespace mdgpu {

struct CellTables {
  const int *ostart, *ocount, *gstart, *gcount;
};

struct Bricks {
  int bd[3];   // cells per brick along x, y, z
  int nb[3];   // bricks along x, y, z
  int m[3];    // cells along x, y, z (extended grid)
  int rmax;    // neighbourhood cells of a full brick: prod(bd + 2)
};

// staged ranges of a brick: two per neighbourhood cell (owned, then ghosts)
constexpr int kRangesPerCell = 2;

struct BrickBox {
  int lo[3];  // first neighbourhood cell (may be -1)
  int n[3];   // neighbourhood cells along x, y, z
};

__device__ __forceinline__ BrickBox brick_box(int b, const Bricks &B) {
  BrickBox X;
  const int bx = b % B.nb[0];
  const int by = (b / B.nb[0]) % B.nb[1];
  const int bz = b / (B.nb[0] * B.nb[1]);
  const int bi[3] = {bx, by, bz};
  for (int d = 0; d < 3; d++) {
    const int c0 = bi[d] * B.bd[d];
    const int c1 = min(B.m[d], c0 + B.bd[d]);
    X.lo[d] = c0 - 1;
    X.n[d] = c1 - c0 + 2;
  }
  return X;
}

// ---- rank order of cells -------------------------------------------------
__global__ void k_rank_counts(const int *__restrict__ count, const int *__restrict__ rank, int ncell,
                              int *__restrict__ count_r) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncell) count_r[rank[c]] = count[c];
}

__global__ void k_unrank_starts(const int *__restrict__ start_r, const int *__restrict__ rank, int ncell,
                                int *__restrict__ start) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncell) start[c] = start_r[rank[c]];
}

// rows of each brick: [row0, row0 + nrow), in 64-row slices
__global__ void k_brick_rows(const int *__restrict__ brick_rank0, const int *__restrict__ start_r, int nbrick,
                             int ncell, int pn, int *__restrict__ row0, int *__restrict__ nrow,
                             int *__restrict__ nsl) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbrick) return;
  const int r0 = brick_rank0[b], r1 = brick_rank0[b + 1];
  const int s0 = r0 < ncell ? start_r[r0] : pn;
  const int s1 = r1 < ncell ? start_r[r1] : pn;
  row0[b] = s0;
  nrow[b] = s1 - s0;
  nsl[b] = (s1 - s0 + kWave - 1) / kWave;
}

__global__ void k_slice_brick(const int *__restrict__ vslice0, const int *__restrict__ nsl, int nbrick,
                              int *__restrict__ slice_brick) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbrick) return;
  for (int k = 0; k < nsl[b]; k++) slice_brick[vslice0[b] + k] = b;
}

// Staged ranges of each brick, range 2r = owned particles of neighbourhood
// cell r, range 2r+1 = its ghosts (cells outside the extended grid are
// empty).  tstart = first index in the owned+ghost arrays (ghosts from pn on),
// tbase = offset in the staged array (tbase[R] = staged count = pad slot).
// The largest staged count goes to maxpart (per workgroup).
__global__ void k_brick_ranges(Bricks B, CellTables T, int nbrick, int pn, int *__restrict__ tstart,
                               int *__restrict__ tbase, int *__restrict__ nrange, double *__restrict__ maxpart) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  double total = 0.0;
  if (b < nbrick) {
    const BrickBox X = brick_box(b, B);
    const int nc = X.n[0] * X.n[1] * X.n[2];
    const int stride = kRangesPerCell * B.rmax;
    int *ts = tstart + (long long)b * stride;
    int *tb = tbase + (long long)b * (stride + 1);
    int base = 0;
    for (int r = 0; r < nc; r++) {
      const int x = X.lo[0] + r % X.n[0];
      const int y = X.lo[1] + (r / X.n[0]) % X.n[1];
      const int z = X.lo[2] + r / (X.n[0] * X.n[1]);
      int os = 0, oc = 0, gs = pn, gc = 0;
      if (x >= 0 && x < B.m[0] && y >= 0 && y < B.m[1] && z >= 0 && z < B.m[2]) {
        const int c = (z * B.m[1] + y) * B.m[0] + x;
        os = T.ostart[c];
        oc = T.ocount[c];
        gs = pn + T.gstart[c];
        gc = T.gcount[c];
      }
      ts[2 * r] = os;
      tb[2 * r] = base;
      base += oc;
      ts[2 * r + 1] = gs;
      tb[2 * r + 1] = base;
      base += gc;
    }
    tb[kRangesPerCell * nc] = base;
    nrange[b] = kRangesPerCell * nc;
    total = (double)base;
  }
  block_max_to_partial(total, maxpart);
}

// ---- search --------------------------------------------------------------
struct RowOut {
  uint16_t *nbr;
  uint8_t *code;
  long long stride;
  int cap;
  // brick of the row and its staged ranges
  BrickBox box;
  const int *ts, *tb;
};

// cell offset code of a partner cell relative to the row's cell (each -1..1)
__device__ __forceinline__ uint8_t offset_code(int ox, int oy, int oz) {
  return (uint8_t)((ox + 1) | ((oy + 1) << 2) | ((oz + 1) << 4));
}

template <bool kFill>
__device__ __forceinline__ int search_row(int i, const double4 *__restrict__ pos, const int *__restrict__ cell,
                                          int pn, Grid g, CellTables T, double S2, RowOut *o,
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
      for (int dx = -1; dx <= 1; dx++) {
        const int jx = ix + dx;
        if (jx < 0 || jx >= mx) continue;
        const int c2 = mx * my * jz + mx * jy + jx;
        for (int pass = 0; pass < 2; pass++) {
          const int b = pass ? pn + T.gstart[c2] : T.ostart[c2];
          const int e = b + (pass ? T.gcount[c2] : T.ocount[c2]);
          for (int j = b; j < e; j++) {
            if (j == i) continue;
            const double4 qj = pos[j];
            const double ex = qi.x - qj.x, ey = qi.y - qj.y, ez = qi.z - qj.z;
            const double r2 = (ex * ex + ey * ey + ez * ez);
            if (r2 > S2) continue;
            if (kFill) {
              // staged index: range of (partner cell, owned or ghost) + offset
              const int lx = jx - o->box.lo[0], ly = jy - o->box.lo[1], lz = jz - o->box.lo[2];
              const int r = kRangesPerCell * ((lz * o->box.n[1] + ly) * o->box.n[0] + lx) + pass;
              const int off = j - o->ts[r];
              // always true by construction (the box covers every cell next to a row's cell)
              const bool good = lx >= 0 && lx < o->box.n[0] && ly >= 0 && ly < o->box.n[1] && lz >= 0 &&
                                lz < o->box.n[2] && off >= 0 && off < o->tb[r + 1] - o->tb[r];
              if (!good) {
                atomicOr(err, kErrLayout);
              } else if (cnt < o->cap) {
                o->nbr[(long long)cnt * o->stride] = (uint16_t)(o->tb[r] + off);
                o->code[(long long)cnt * o->stride] = offset_code(dx, dy, dz);
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
  row_len[i] = search_row<false>(i, pos, cell, pn, g, T, S2, nullptr, nullptr);
}

// slice length = longest row of its 64 rows, rounded up to a multiple of
// kSliceRound so the force loop runs whole batches; slots = 64 * length
__global__ void k_slice_len(const int *__restrict__ row_len, const int *__restrict__ slice_brick,
                            const int *__restrict__ vslice0, const int *__restrict__ row0,
                            const int *__restrict__ nrow, int nslice, int *__restrict__ slen,
                            long long *__restrict__ slots) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslice) return;
  const int b = slice_brick[s];
  const int v0 = (s - vslice0[b]) * kWave;
  const int v1 = min(nrow[b], v0 + kWave);
  int m = 0;
  for (int v = v0; v < v1; v++) m = max(m, row_len[row0[b] + v]);
  m = (m + kSliceRound - 1) / kSliceRound * kSliceRound;
  slen[s] = m;
  slots[s] = (long long)m * kWave;
}

struct BrickRows {
  const int *brick_of_cell, *row0, *vslice0;
  const int *tstart, *tbase;
};

__global__ void k_fill_rows(const double4 *__restrict__ pos, const int *__restrict__ cell, int pn, Grid g,
                            CellTables T, double S2, Bricks B, BrickRows BR, const long long *__restrict__ sbase,
                            const int *__restrict__ slen, uint16_t *__restrict__ nbr, uint8_t *__restrict__ ncode,
                            int *__restrict__ err) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= pn) return;
  const int c = cell[i];
  const int b = BR.brick_of_cell[c];
  const int v = i - BR.row0[b];
  const int s = BR.vslice0[b] + v / kWave, lane = v % kWave;
  RowOut o;
  o.box = brick_box(b, B);
  o.ts = BR.tstart + (long long)b * kRangesPerCell * B.rmax;
  o.tb = BR.tbase + (long long)b * (kRangesPerCell * B.rmax + 1);
  const int len = slen[s];
  o.nbr = nbr + sbase[s] + lane;
  o.code = ncode + sbase[s] + lane;
  o.stride = kWave;
  o.cap = len;
  const int cnt = search_row<true>(i, pos, cell, pn, g, T, S2, &o, err);
  if (cnt > len) atomicOr(err, kErrCapacity);
  const int R = kRangesPerCell * o.box.n[0] * o.box.n[1] * o.box.n[2];
  const uint16_t pad = (uint16_t)o.tb[R];
  for (int k = cnt; k < len; k++) {
    o.nbr[(long long)k * kWave] = pad;
    o.code[(long long)k * kWave] = (uint8_t)kZeroOffset;
  }
}

// lanes of a brick's last slice past its rows are all padding
__global__ void k_fill_tail(int nslice, Bricks B, const int *__restrict__ slice_brick,
                            const int *__restrict__ vslice0, const int *__restrict__ nrow,
                            const int *__restrict__ nrange, const int *__restrict__ tbase,
                            const long long *__restrict__ sbase, const int *__restrict__ slen,
                            uint16_t *__restrict__ nbr, uint8_t *__restrict__ ncode) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int s = (int)(gid / kWave), lane = (int)(gid % kWave);
  if (s >= nslice) return;
  const int b = slice_brick[s];
  const int v = (s - vslice0[b]) * kWave + lane;
  if (v < nrow[b]) return;
  const uint16_t pad = (uint16_t)tbase[(long long)b * (kRangesPerCell * B.rmax + 1) + nrange[b]];
  const int len = slen[s];
  for (int k = 0; k < len; k++) {
    nbr[sbase[s] + (long long)k * kWave + lane] = pad;
    ncode[sbase[s] + (long long)k * kWave + lane] = (uint8_t)kZeroOffset;
  }
}

}  // namespace mdgpu
