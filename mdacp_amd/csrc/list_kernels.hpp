// list_kernels.hpp -- neighbour list on gfx950, organised by cell bricks.
//
// The extended cell grid (meshlist.cc:31-55) of each replica, m cells per
// axis, is cut into bricks of bd[0] x bd[1] x bd[2] cells (Replica::bd) from
// cell 1 to cell m - 1: the owned cells and the upper ghost layer, which holds
// an owned particle that sits exactly on the upper face (q == L bins there).
// The lower ghost layer (cell 0) only ever holds ghost images and belongs to
// no brick.  Cells are ranked brick by brick, and owned
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
// (MeshList::SearchMesh, meshlist.cc:220-271): owned rows against owned and
// materialised ghost particles binned in the extended grid, r^2 = dx*dx +
// dy*dy + dz*dz evaluated without fused multiply-adds, kept when !(r2 > S2).
#pragma once
#include "build_kernels.hpp"

namespace mdgpu {

struct CellTables {
  const int *ostart, *ocount, *gstart, *gcount;
};

// Bricks of a domain: each brick's replica, the replica table, and the
// stride of the per-brick range tables.
struct Bricks {
  const int *brep;         // replica of each brick
  const Replica *reps;
  int rmax;                // largest neighbourhood of a brick: max over replicas of prod(bd + 2)
};

// staged ranges of a brick: two per neighbourhood cell (owned, then ghosts)
constexpr int kRangesPerCell = 2;
// cells of a brick's neighbourhood, (bd[0] + 2)(bd[1] + 2)(bd[2] + 2), at most,
// and the longest brick edge in cells
constexpr int kMaxBrickCells = 216;
constexpr int kMaxBrickEdge = 8;

struct BrickBox {
  int lo[3];  // first neighbourhood cell (may be -1)
  int n[3];   // neighbourhood cells along x, y, z
};

// neighbourhood box of global brick b of replica R, in the replica's cell
// coordinates (the last box along an axis reaches one cell past the grid;
// that cell is empty)
__device__ __forceinline__ BrickBox brick_box(int b, const Replica &R) {
  BrickBox X;
  const int l = b - R.brick0;
  const int bx = l % R.nb[0];
  const int by = (l / R.nb[0]) % R.nb[1];
  const int bz = l / (R.nb[0] * R.nb[1]);
  const int bi[3] = {bx, by, bz};
  for (int d = 0; d < 3; d++) {
    const int c0 = 1 + bi[d] * R.bd[d];
    const int c1 = min(R.m[d], c0 + R.bd[d]);
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
// One wavefront per brick, a lane per neighbourhood cell, the offsets from a
// wavefront prefix sum; each brick's staged count also goes to total_d.
__global__ __launch_bounds__(kBlock) void k_brick_ranges(Bricks B, CellTables T, int nbrick, int pn,
                                                         int *__restrict__ tstart, int *__restrict__ tbase,
                                                         int *__restrict__ nrange, int *__restrict__ btotal,
                                                         double *__restrict__ total_d) {
  const int b = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (b >= nbrick) return;  // the whole wavefront
  const Replica &Rp = B.reps[B.brep[b]];
  const BrickBox X = brick_box(b, Rp);
  const int nc = X.n[0] * X.n[1] * X.n[2];
  const int stride = kRangesPerCell * B.rmax;
  int *ts = tstart + (long long)b * stride;
  int *tb = tbase + (long long)b * (stride + 1);
  int base = 0;
  for (int r0 = 0; r0 < nc; r0 += kWave) {
    const int r = r0 + lane;
    int os = 0, oc = 0, gs = pn, gc = 0;
    if (r < nc) {
      const int x = X.lo[0] + r % X.n[0];
      const int y = X.lo[1] + (r / X.n[0]) % X.n[1];
      const int z = X.lo[2] + r / (X.n[0] * X.n[1]);
      if (x >= 0 && x < Rp.m[0] && y >= 0 && y < Rp.m[1] && z >= 0 && z < Rp.m[2]) {
        const int c = Rp.cell0 + (z * Rp.m[1] + y) * Rp.m[0] + x;
        os = T.ostart[c];
        oc = T.ocount[c];
        gs = pn + T.gstart[c];
        gc = T.gcount[c];
      }
    }
    const int cnt = oc + gc;
    int inc = cnt;  // inclusive prefix sum over the wavefront
    for (int o = 1; o < kWave; o <<= 1) {
      const int v = __shfl_up(inc, o, kWave);
      if (lane >= o) inc += v;
    }
    const int ex = base + inc - cnt;
    if (r < nc) {
      ts[2 * r] = os;
      tb[2 * r] = ex;
      ts[2 * r + 1] = gs;
      tb[2 * r + 1] = ex + oc;
    }
    base += __shfl(inc, kWave - 1, kWave);
  }
  if (lane == 0) {
    tb[kRangesPerCell * nc] = base;
    nrange[b] = kRangesPerCell * nc;
    btotal[b] = base;
    total_d[b] = (double)base;
  }
}

// Gather list of each brick: glist[gbase[b] + u] = index (owned, or pn + ghost)
// of staged entry u, so the force kernels stage with one indexed load per
// entry.  One workgroup per brick.
__global__ void k_brick_glist(int rstride, const int *__restrict__ tstart, const int *__restrict__ tbase,
                              const int *__restrict__ nrange, const long long *__restrict__ gbase,
                              int *__restrict__ glist) {
  __shared__ int tb[kRangesPerCell * kMaxBrickCells + 1];
  const int b = blockIdx.x;
  const int R = nrange[b];
  const int *tbg = tbase + (long long)b * (rstride + 1);
  const int *ts = tstart + (long long)b * rstride;
  for (int r = threadIdx.x; r <= R; r += blockDim.x) tb[r] = tbg[r];
  __syncthreads();
  const int total = tb[R];
  int *out = glist + gbase[b];
  for (int u = threadIdx.x; u < total; u += blockDim.x) {
    int lo = 0, hi = R;  // largest r with tb[r] <= u
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (tb[mid] <= u) lo = mid;
      else hi = mid;
    }
    out[u] = ts[lo] + (u - tb[lo]);
  }
}

// ---- search --------------------------------------------------------------
constexpr int kMaxRanges = kRangesPerCell * kMaxBrickCells;
constexpr int kErrSearchCap = 16;  // a row has more candidates than its bitmask holds

// cell offset code of a partner cell relative to the row's cell (each -1..1)
__device__ __forceinline__ uint8_t offset_code(int ox, int oy, int oz) {
  return (uint8_t)((ox + 1) | ((oy + 1) << 2) | ((oz + 1) << 4));
}

constexpr int kStageBatch = 4;  // staged entries per thread with loads in flight together

// Position of gather-list entry j: an owned particle (j < pn) from src, a
// ghost image from `ghost`.  The pair kernels of a fused step read the owned
// particles' half drift from one array and the images, which the halo refresh
// writes once, from the current positions.
__device__ __forceinline__ double4 gather_pos(const double4 *__restrict__ src, const double4 *__restrict__ ghost,
                                              int pn, int j) {
  return (j < pn ? src : ghost)[j];
}

// Stage positions (3 doubles per entry) of a brick's gather list into LDS,
// plus a far sentinel at index `total` (the pad entry of the list); entries
// from pn on come from `ghost`.
template <int T>
__device__ __forceinline__ void stage_fp64(const int *__restrict__ gl, int total, const double4 *__restrict__ src,
                                           double *__restrict__ lds, int pn, const double4 *__restrict__ ghost) {
  for (int u0 = threadIdx.x; u0 < total; u0 += T * kStageBatch) {
    int j[kStageBatch];
#pragma unroll
    for (int k = 0; k < kStageBatch; k++) {
      const int u = u0 + k * T;
      j[k] = u < total ? gl[u] : 0;
    }
    double4 q[kStageBatch];
#pragma unroll
    for (int k = 0; k < kStageBatch; k++) q[k] = gather_pos(src, ghost, pn, j[k]);
#pragma unroll
    for (int k = 0; k < kStageBatch; k++) {
      const int u = u0 + k * T;
      if (u < total) {
        lds[3 * u] = q[k].x;
        lds[3 * u + 1] = q[k].y;
        lds[3 * u + 2] = q[k].z;
      }
    }
  }
  if (threadIdx.x == 0) {
    lds[3 * total] = 1.0e200;
    lds[3 * total + 1] = 1.0e200;
    lds[3 * total + 2] = 1.0e200;
  }
}

// Per-brick tables the search and emit kernels read.
struct SearchTables {
  const int *tstart, *tbase, *nrange;  // staged ranges (list_kernels: k_brick_ranges)
  const int *glist;                    // staged entry -> particle index
  const long long *gbase;              // [brick + 1]
  const int *row0, *nrow, *vslice0;
  int rstride;                         // 2 rmax
  int lds_cap;                         // staged entries that fit the LDS budget (incl. pad)
};

template <int T>
__device__ __forceinline__ int load_tb(const SearchTables &st, int b, int *tb) {
  const int R = st.nrange[b];
  const int *tbg = st.tbase + (long long)b * (st.rstride + 1);
  for (int r = threadIdx.x; r <= R; r += T) tb[r] = tbg[r];
  return R;
}

// Box coordinates of a row's cell and the staged index of the row itself.
struct RowCell {
  int lx, ly, lz, self;
};

__device__ __forceinline__ RowCell row_cell(int i, int c, const BrickBox &X, const Replica &R, const int *tb,
                                            const int *__restrict__ ts) {
  RowCell rc;
  c -= R.cell0;
  rc.lx = c % R.m[0] - X.lo[0];
  rc.ly = (c / R.m[0]) % R.m[1] - X.lo[1];
  rc.lz = c / (R.m[0] * R.m[1]) - X.lo[2];
  const int rs = kRangesPerCell * ((rc.lz * X.n[1] + rc.ly) * X.n[0] + rc.lx);
  rc.self = tb[rs] + (i - ts[rs]);
  return rc;
}

// MeshList::SearchMesh over a brick (meshlist.cc:220-271): each owned row of
// the brick tests the owned and ghost particles of the 27 cells around its
// cell, staged in LDS.  For a row the candidates come as 9 contiguous staged
// runs (x-neighbouring cells are adjacent in the staged order); bit k of the
// row's mask marks candidate k, in run order, as listed: r^2 = dx*dx + dy*dy +
// dz*dz without fused multiply-adds, kept when !(r2 > S2), the row itself
// skipped.  row_len receives the count.
//
// Threads take consecutive rows (rows of one cell share their candidates, so
// most staged reads are broadcasts); each thread tests kSearchBatch
// candidates at a time and collects the bits in a 64-bit accumulator that is
// flushed to the mask a word at a time.
constexpr int kSearchBatch = 4;

template <int W>
__global__ __launch_bounds__(kWave *W) void k_search_brick(const double4 *__restrict__ pos,
                                                          const int *__restrict__ cell, Bricks B,
                                                          SearchTables st, double S2, int mcap,
                                                          int *__restrict__ row_len, uint32_t *__restrict__ mask,
                                                          int *__restrict__ err) {
  extern __shared__ __align__(16) double lds64[];
  __shared__ int tb[kMaxRanges + 1];
  constexpr int T = kWave * W;
  const int b = blockIdx.x;
  if (st.nrow[b] == 0) return;  // a brick of ghost-layer cells only
  const int R = load_tb<T>(st, b, tb);
  __syncthreads();
  const int total = tb[R];
  const int *gl = st.glist + st.gbase[b];
  const bool staged = total < st.lds_cap;
  if (staged) stage_fp64<T>(gl, total, pos, lds64, 0, pos);
  __syncthreads();
  const Replica &Rp = B.reps[B.brep[b]];
  const BrickBox X = brick_box(b, Rp);
  const int *ts = st.tstart + (long long)b * st.rstride;
  const int nrow = st.nrow[b], row0 = st.row0[b];
  const int nwords = mcap;
  for (int v = threadIdx.x; v < nrow; v += T) {
    const int i = row0 + v;
    const double4 qi = pos[i];
    const RowCell rc = row_cell(i, cell[i], X, Rp, tb, ts);
    uint32_t *mk = mask + (long long)i * mcap;
    unsigned long long acc = 0;  // pending bits, oldest in bit 0
    int accbits = 0, word = 0, cnt = 0;
    for (int dz = -1; dz <= 1; dz++) {
      for (int dy = -1; dy <= 1; dy++) {
        const int r0 = ((rc.lz + dz) * X.n[1] + (rc.ly + dy)) * X.n[0] + rc.lx - 1;
        const int u0 = tb[kRangesPerCell * r0], u1 = tb[kRangesPerCell * (r0 + 3)];
        for (int u = u0; u < u1; u += kSearchBatch) {
          const int nc = min(kSearchBatch, u1 - u);
          unsigned h = 0;
#pragma unroll
          for (int t = 0; t < kSearchBatch; t++) {
            const int uu = min(u + t, u1 - 1);  // in range; surplus lanes of the batch are masked below
            double qx, qy, qz;
            if (staged) {
              qx = lds64[3 * uu];
              qy = lds64[3 * uu + 1];
              qz = lds64[3 * uu + 2];
            } else {
              const double4 q = pos[gl[uu]];
              qx = q.x;
              qy = q.y;
              qz = q.z;
            }
            const double ex = qi.x - qx, ey = qi.y - qy, ez = qi.z - qz;
            const double r2 = (ex * ex + ey * ey + ez * ez);
            const bool hit = t < nc && !(r2 > S2) && uu != rc.self;
            h |= (hit ? 1u : 0u) << t;
          }
          cnt += __popc(h);
          acc |= (unsigned long long)h << accbits;
          accbits += nc;
          if (accbits >= 32) {
            if (word < nwords) mk[word] = (uint32_t)acc;
            word++;
            acc >>= 32;
            accbits -= 32;
          }
        }
      }
    }
    if (accbits > 0) {
      if (word < nwords) mk[word] = (uint32_t)acc;
      word++;
    }
    if (word > nwords) atomicOr(err, kErrSearchCap);
    row_len[i] = cnt;
  }
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Rows ur0 .. ur0 + nr - 1 (staged indices) of a cell against this lane's
// candidate (cx, cy, cz): r^2 without fused multiply-adds, kept when
// !(r2 > S2).  Lane r keeps row r's ballot in (lo, hi).  Rows come from
// LDS (kStaged) or from lane r's registers (ax, ay, az), four at a time.
template <bool kStaged>
__device__ __forceinline__ void row_position(int r, int ur0, const double *__restrict__ lds, double ax, double ay,
                                             double az, double &qx, double &qy, double &qz) {
  if (kStaged) {
    qx = lds[3 * (ur0 + r)];
    qy = lds[3 * (ur0 + r) + 1];
    qz = lds[3 * (ur0 + r) + 2];
  } else {
    qx = readlane_f64(ax, r);
    qy = readlane_f64(ay, r);
    qz = readlane_f64(az, r);
  }
}

// v with lane r (uniform) replaced by the uniform word s: one v_writelane_b32
// (a compare and a select per word otherwise, plus a copy of s to a VGPR).  The
// lane select is read from M0, as the instruction reads one SGPR besides it;
// the nop covers a write of s to VCC, or of M0, just before.
__device__ __forceinline__ uint32_t put_lane(uint32_t v, uint32_t s, int r) {
  asm("s_nop 1\n\tv_writelane_b32 %0, %1, m0" : "+v"(v) : "s"(s), "{m0}"(r));
  return v;
}

__device__ __forceinline__ void test_row(int r, double qx, double qy, double qz, double cx, double cy,
                                         double cz, double S2, uint32_t &lo, uint32_t &hi) {
  const double ex = qx - cx, ey = qy - cy, ez = qz - cz;
  const double r2 = (ex * ex + ey * ey + ez * ez);
  const unsigned long long m = __ballot(!(r2 > S2));
  lo = put_lane(lo, (uint32_t)m, r);  // lane r takes the ballot
  hi = put_lane(hi, (uint32_t)(m >> 32), r);
}

template <bool kStaged>
__device__ __forceinline__ void test_rows(int nr, int ur0, const double *__restrict__ lds, double ax, double ay,
                                          double az, double cx, double cy, double cz, double S2, uint32_t &lo,
                                          uint32_t &hi) {
  int r = 0;
  for (; r + 4 <= nr; r += 4) {
    double q[4][3];
#pragma unroll
    for (int j = 0; j < 4; j++) row_position<kStaged>(r + j, ur0, lds, ax, ay, az, q[j][0], q[j][1], q[j][2]);
#pragma unroll
    for (int j = 0; j < 4; j++) test_row(r + j, q[j][0], q[j][1], q[j][2], cx, cy, cz, S2, lo, hi);
  }
  for (; r < nr; r++) {
    double qx, qy, qz;
    row_position<kStaged>(r, ur0, lds, ax, ay, az, qx, qy, qz);
    test_row(r, qx, qy, qz, cx, cy, cz, S2, lo, hi);
  }
}

// The next item of a workgroup's shared counter, for a whole wavefront.
__device__ __forceinline__ int take_next(int *counter, int lane) {
  int t = 0;
  if (lane == 0) t = atomicAdd(counter, 1);
  return __shfl(t, 0, kWave);
}

// The same search and the same masks, one wavefront per cell of the brick:
// the cell's candidates (its 9 runs, concatenated) are taken 64 at a time, one
// per lane, and each row of the cell is tested against all 64 at once; the
// wavefront's ballot of the test is bits [64c, 64c + 64) of that row's mask.
// Rows are read at one address for the whole wavefront (an LDS broadcast) and
// every lane does the same work.  Lane r collects the mask words of row r of
// each block of up to 64 rows of the cell, clears the row's own bit (the
// row sits in the middle run, at its offset in its cell's owned range) and
// counts the rest.
template <int W>
__global__ __launch_bounds__(kWave *W) void k_search_cells(const double4 *__restrict__ pos, Bricks B,
                                                          SearchTables st, double S2, int mcap,
                                                          int *__restrict__ row_len, uint32_t *__restrict__ mask,
                                                          int *__restrict__ err) {
  extern __shared__ __align__(16) double lds64[];
  __shared__ int tb[kMaxRanges + 1];
  __shared__ int cell_next;  // the next cell a wavefront takes
  constexpr int T = kWave * W;
  const int b = blockIdx.x;
  if (st.nrow[b] == 0) return;  // a brick without rows
  const int R = load_tb<T>(st, b, tb);
  if (threadIdx.x == 0) cell_next = W;
  __syncthreads();
  const int total = tb[R];
  const int *gl = st.glist + st.gbase[b];
  const bool staged = total < st.lds_cap;
  if (staged) stage_fp64<T>(gl, total, pos, lds64, 0, pos);
  __syncthreads();
  const Replica &Rp = B.reps[B.brep[b]];
  const BrickBox X = brick_box(b, Rp);
  const int *ts = st.tstart + (long long)b * st.rstride;
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  // the brick's own cells are the neighbourhood less its outer layer
  const int ox = X.n[0] - 2, oy = X.n[1] - 2, oz = X.n[2] - 2;
  // cells differ in rows and candidates: each wavefront takes the next one as it finishes
  for (int c = wv; c < ox * oy * oz; c = take_next(&cell_next, lane)) {
    const int lx = 1 + c % ox, ly = 1 + (c / ox) % oy, lz = 1 + c / (ox * oy);
    const int rs = kRangesPerCell * ((lz * X.n[1] + ly) * X.n[0] + lx);
    const int urow0 = __builtin_amdgcn_readfirstlane(tb[rs]);  // staged index of the cell's first row
    const int nrows = __builtin_amdgcn_readfirstlane(tb[rs + 1]) - urow0;
    if (nrows == 0) continue;
    const int irow0 = __builtin_amdgcn_readfirstlane(ts[rs]);  // particle index of that row
    // runs of the candidate sequence: staged start ru[q], first candidate
    // rp[q], and du[q] = ru[q] - rp[q] (candidate k of run q is staged at k + du[q])
    int ru[9], rp[10], du[9];
    rp[0] = 0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const int r0 = ((lz + q / 3 - 1) * X.n[1] + (ly + q % 3 - 1)) * X.n[0] + lx - 1;
      ru[q] = __builtin_amdgcn_readfirstlane(tb[kRangesPerCell * r0]);
      rp[q + 1] = rp[q] + __builtin_amdgcn_readfirstlane(tb[kRangesPerCell * (r0 + 3)]) - ru[q];
      du[q] = ru[q] - rp[q];
    }
    const int ncand = rp[9];
    if (ncand > 32 * mcap) {  // the host retries with longer masks
      if (lane == 0) atomicOr(err, kErrSearchCap);
      continue;
    }
    const int nw = (ncand + 31) / 32;
    for (int rb = 0; rb < nrows; rb += kWave) {
      const int nr = min(kWave, nrows - rb);
      int cnt = 0;
      double ax = 0.0, ay = 0.0, az = 0.0;  // without staging, lane r holds row rb + r
      if (!staged && lane < nr) {
        const double4 q = pos[irow0 + rb + lane];
        ax = q.x;
        ay = q.y;
        az = q.z;
      }
      for (int c0 = 0; c0 < ncand; c0 += kWave) {
        const int k = c0 + lane;
        // the last run that starts at or before k holds it (an empty run starts
        // where the next one does): one compare and one select per run
        int dk = du[0];
#pragma unroll
        for (int q = 1; q < 9; q++) dk = k >= rp[q] ? du[q] : dk;
        const int u = k + dk;
        double cx = 1.0e200, cy = 1.0e200, cz = 1.0e200;  // lanes past the last candidate
        if (k < ncand) {
          if (staged) {
            cx = lds64[3 * u];
            cy = lds64[3 * u + 1];
            cz = lds64[3 * u + 2];
          } else {
            const double4 q = pos[gl[u]];
            cx = q.x;
            cy = q.y;
            cz = q.z;
          }
        }
        uint32_t lo = 0, hi = 0;
        if (staged)
          test_rows<true>(nr, urow0 + rb, lds64, 0.0, 0.0, 0.0, cx, cy, cz, S2, lo, hi);
        else
          test_rows<false>(nr, urow0 + rb, nullptr, ax, ay, az, cx, cy, cz, S2, lo, hi);
        if (lane < nr) {
          // the row itself is candidate kself (the middle run holds the row's cell)
          const int kself = rp[4] + (urow0 + rb + lane - ru[4]) - c0;
          if (kself >= 0 && kself < 32) lo &= ~(1u << kself);
          else if (kself >= 32 && kself < 64) hi &= ~(1u << (kself - 32));
          cnt += __popc(lo) + __popc(hi);
          uint32_t *mk = mask + (long long)(irow0 + rb + lane) * mcap + c0 / 32;
          mk[0] = lo;
          if (c0 / 32 + 1 < nw) mk[1] = hi;
        }
      }
      if (lane < nr) row_len[irow0 + rb + lane] = cnt;
    }
  }
}

// ---- row order inside a brick ------------------------------------------
constexpr int kSortBins = kBlock;  // lengths >= kSortBins - 1 share the last bin

// Order the rows of each brick by decreasing list length counted in steps of
// `sort` (ceil(length / sort)), ties by index, i.e. in cell order (a stable
// counting sort, so the order is deterministic): rowsrc[new] = old,
// rowinv[old] = new.  Slices of 64 consecutive rows then hold rows of nearly
// equal length and carry little padding.  Steps of kSliceRound, to which
// slice lengths are rounded anyway, cost no more padding than exact lengths
// and keep many more rows of one cell together in a slice: their partners lie
// close together in the staged array, and the pair kernels' LDS gathers
// conflict less (the FP64 kick of 4M particles at rho = 0.5 takes 6% less
// time than with exact lengths).  With sort == 0 the rows keep their cell
// order (rowsrc and rowinv are the identity).  One workgroup of kBlock
// threads per brick.
__global__ __launch_bounds__(kBlock) void k_brick_row_order(const int *__restrict__ row_len,
                                                            const int *__restrict__ row0,
                                                            const int *__restrict__ nrow,
                                                            int *__restrict__ rowsrc, int *__restrict__ rowinv,
                                                            int sort) {
  __shared__ int start[kSortBins];
  __shared__ int scan[kSortBins];
  __shared__ int bins[kBlock];
  const int b = blockIdx.x;
  const int r0 = row0[b], n = nrow[b];
  const int t = threadIdx.x;
  if (!sort) {
    for (int v = t; v < n; v += kBlock) {
      rowsrc[r0 + v] = r0 + v;
      rowinv[r0 + v] = r0 + v;
    }
    return;
  }
  start[t] = 0;
  __syncthreads();
  // bin of a row: longest rows first, lengths taken in steps of `sort`
  for (int v = t; v < n; v += kBlock)
    atomicAdd(&start[kSortBins - 1 - min((row_len[r0 + v] + sort - 1) / sort, kSortBins - 1)], 1);
  __syncthreads();
  // exclusive scan of the bin counts (Hillis-Steele over kSortBins == kBlock)
  int x = start[t];
  scan[t] = x;
  __syncthreads();
  for (int o = 1; o < kSortBins; o <<= 1) {
    const int y = t >= o ? scan[t - o] : 0;
    __syncthreads();
    scan[t] += y;
    __syncthreads();
  }
  start[t] = scan[t] - x;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += kBlock) {
    const int v = c0 + t;
    const int bin = v < n ? kSortBins - 1 - min((row_len[r0 + v] + sort - 1) / sort, kSortBins - 1) : -1;
    bins[t] = bin;
    __syncthreads();
    if (v < n) {
      int rank = 0;
      for (int k = 0; k < t; k++) rank += bins[k] == bin ? 1 : 0;
      const int p = r0 + start[bin] + rank;
      rowsrc[p] = r0 + v;
      rowinv[r0 + v] = p;
    }
    __syncthreads();
    if (v < n) atomicAdd(&start[bin], 1);
    __syncthreads();
  }
}

__global__ void k_gather_int(const int *__restrict__ perm, int n, const int *__restrict__ in, int *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = in[perm[k]];
}

// new indices of owned particles (entries >= pn are ghosts and stay)
__global__ void k_remap_owned(int *__restrict__ idx, long long n, int pn, const int *__restrict__ rowinv) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int j = idx[k];
  if (j < pn) idx[k] = rowinv[j];
}

// slice length = longest row of its 64 rows, rounded up to a multiple of
// kSliceRound so the force loop runs whole batches; slots = 64 * length.
// The largest slice length goes to partial (per workgroup).
__global__ void k_slice_len(const int *__restrict__ row_len, const int *__restrict__ slice_brick,
                            const int *__restrict__ vslice0, const int *__restrict__ row0,
                            const int *__restrict__ nrow, int nslice, int *__restrict__ slen,
                            long long *__restrict__ slots, double *__restrict__ partial) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  int m = 0;
  if (s < nslice) {
    const int b = slice_brick[s];
    const int v0 = (s - vslice0[b]) * kWave;
    const int v1 = min(nrow[b], v0 + kWave);
    for (int v = v0; v < v1; v++) m = max(m, row_len[row0[b] + v]);
    m = (m + kSliceRound - 1) / kSliceRound * kSliceRound;
    slen[s] = m;
    slots[s] = (long long)m * kWave;
  }
  block_max_to_partial((double)m, partial, blockIdx.x);
}

constexpr int kEmitWaves = 4;  // wavefronts per brick workgroup of k_emit_brick

// A row's walk through its candidate mask during the emit: the current
// 64-bit mask word (bits already taken cleared) and the next one, loaded ahead;
// the current run q of the candidate sequence with its first candidate kb, its
// end ke, its first staged entry u0 and the starts t1, t2 of its second and
// third cells (x-neighbouring cells are adjacent in the staged order).
struct MaskCursor {
  const unsigned long long *mk;
  unsigned long long w, wn;
  int wi, nw, ncand;
  int q, kb, ke, u0, t1, t2;
};

__device__ __forceinline__ unsigned long long mask_word(const MaskCursor &c, int wi) {
  if (wi >= c.nw) return 0ull;
  const unsigned long long v = c.mk[wi];
  const int valid = c.ncand - 64 * wi;  // the search writes whole 32-bit words up to the last candidate
  return valid >= 64 ? v : v & ((1ull << valid) - 1ull);
}

__device__ __forceinline__ void cursor_run(MaskCursor &c, const RowCell &rc, const BrickBox &X, const int *tb) {
  const int r0 = ((rc.lz + c.q / 3 - 1) * X.n[1] + (rc.ly + c.q % 3 - 1)) * X.n[0] + rc.lx - 1;
  c.u0 = tb[kRangesPerCell * r0];
  c.t1 = tb[kRangesPerCell * (r0 + 1)];
  c.t2 = tb[kRangesPerCell * (r0 + 2)];
  c.kb = c.ke;
  c.ke = c.kb + tb[kRangesPerCell * (r0 + 3)] - c.u0;
}

// MeshList::MakeList (meshlist.cc:58-87) in sliced-ELL form: one wavefront per
// 64-row slice (rows in the order of k_brick_row_order, rowsrc[row] = the
// row's index in the search) turns the rows' candidate masks into staged
// indices and, with kCodes, cell offset codes.  All lanes produce entry j of
// group g together, each from its own row's mask (or the brick's pad entry
// once the row is exhausted), and store group g as one 16-byte piece of
// indices (and one 8-byte piece of codes) per lane: piece g * 64 + lane of the
// slice, the layout of slot_of().  Without codes, a lane holds its row's nine
// candidate runs (first candidate, staged offset) in registers and maps a
// candidate to its staged index with eight compare-selects; with codes, it
// walks the runs through the brick's range table, which also gives the
// partner's cell offset.
template <int W, bool kCodes>
__global__ __launch_bounds__(kWave *W) void k_emit_brick(const int *__restrict__ cell, Bricks B, SearchTables st,
                                                        const int *__restrict__ rowsrc,
                                                        const uint32_t *__restrict__ mask, int mcap,
                                                        const long long *__restrict__ sbase,
                                                        const int *__restrict__ slen, uint16_t *__restrict__ nbr,
                                                        uint8_t *__restrict__ ncode, int *__restrict__ err) {
  __shared__ int tb[kMaxRanges + 1];
  constexpr int T = kWave * W;
  const int b = blockIdx.x;
  if (st.nrow[b] == 0) return;  // a brick without rows
  const int R = load_tb<T>(st, b, tb);
  __syncthreads();
  const unsigned total = (unsigned)tb[R];
  const Replica &Rp = B.reps[B.brep[b]];
  const BrickBox X = brick_box(b, Rp);
  const int *ts = st.tstart + (long long)b * st.rstride;
  const int nrow = st.nrow[b], row0 = st.row0[b], vs0 = st.vslice0[b];
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int nsl = (nrow + kWave - 1) / kWave;
  for (int sl = wv; sl < nsl; sl += W) {
    const int s = vs0 + sl;
    const int len = slen[s];  // a multiple of kGroup
    const int v = sl * kWave + lane;
    MaskCursor c;
    RowCell rc = {0, 0, 0, 0};
    c.ncand = 0;
    c.nw = 0;
    c.mk = nullptr;
    int kb[9], du[9];  // without codes: run q's first candidate, and staged index - candidate in it
#pragma unroll
    for (int q = 0; q < 9; q++) kb[q] = du[q] = 0;
    if (v < nrow) {
      const int i = row0 + v;
      rc = row_cell(i, cell[i], X, Rp, tb, ts);
      // masks are indexed by the row order of the search; mcap is even
      c.mk = reinterpret_cast<const unsigned long long *>(mask + (long long)rowsrc[i] * mcap);
      int n = 0;
#pragma unroll
      for (int q = 0; q < 9; q++) {
        const int r0 = ((rc.lz + q / 3 - 1) * X.n[1] + (rc.ly + q % 3 - 1)) * X.n[0] + rc.lx - 1;
        const int u0 = tb[kRangesPerCell * r0];
        kb[q] = n;
        du[q] = u0 - n;
        n += tb[kRangesPerCell * (r0 + 3)] - u0;
      }
      c.ncand = n;
      c.nw = (n + 63) / 64;
    }
    c.wi = 0;
    c.w = mask_word(c, 0);
    c.wn = mask_word(c, 1);
    c.q = 0;
    c.ke = 0;
    if (kCodes && c.nw > 0) cursor_run(c, rc, X, tb);
    uint4 *gn = reinterpret_cast<uint4 *>(nbr + sbase[s]);  // sbase: a multiple of 512
    uint2 *gc = kCodes ? reinterpret_cast<uint2 *>(ncode + sbase[s]) : nullptr;
    for (int g = 0; g < len / kGroup; g++) {
      unsigned e[kGroup], cd[kGroup];
#pragma unroll
      for (int j = 0; j < kGroup; j++) {
        while (c.w == 0ull && c.wi + 1 < c.nw) {
          c.wi++;
          c.w = c.wn;
          c.wn = mask_word(c, c.wi + 1);
        }
        e[j] = total;
        cd[j] = kZeroOffset;
        if (c.w != 0ull) {
          const int p = __ffsll((long long)c.w) - 1;
          c.w &= c.w - 1ull;
          const int kc = 64 * c.wi + p;
          if (kCodes) {
            while (kc >= c.ke) {
              c.q++;
              cursor_run(c, rc, X, tb);
            }
            const int u = c.u0 + (kc - c.kb);
            const int dx = u < c.t1 ? -1 : (u < c.t2 ? 0 : 1);
            e[j] = (unsigned)u;
            cd[j] = offset_code(dx, c.q % 3 - 1, c.q / 3 - 1);
          } else {
            // the last run that starts at or before kc holds it (an empty run
            // starts where the next one does)
            int d = du[0];
#pragma unroll
            for (int q = 1; q < 9; q++) d = kc >= kb[q] ? du[q] : d;
            e[j] = (unsigned)(kc + d);
          }
        }
      }
      gn[g * kWave + lane] = make_uint4(e[0] | e[1] << 16, e[2] | e[3] << 16, e[4] | e[5] << 16, e[6] | e[7] << 16);
      if (kCodes)
        gc[g * kWave + lane] = make_uint2(cd[0] | cd[1] << 8 | cd[2] << 16 | cd[3] << 24,
                                          cd[4] | cd[5] << 8 | cd[6] << 16 | cd[7] << 24);
    }
    // entries left over: the slice is shorter than the row
    if (c.w != 0ull || c.wn != 0ull || c.wi + 2 < c.nw) {
      bool more = c.w != 0ull || c.wn != 0ull;
      for (int k = c.wi + 2; k < c.nw && !more; k++) more = mask_word(c, k) != 0ull;
      if (more) atomicOr(err, kErrCapacity);
    }
  }
}

}  // namespace mdgpu
