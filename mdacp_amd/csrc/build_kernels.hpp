// build_kernels.hpp -- pair-list construction on gfx950: periodic wrap,
// counting-sort cell binning (atomics + prefix scan, then each cell's members
// put back in input order, so the result does not depend on atomic ordering)
// and the ghost images of the single-domain halo, for every replica of a
// domain at once (each particle carries its replica index).
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

// MeshList's cell of a position (meshlist.cc:164-182): int((q - s) * (1/w)) + 1
// per axis with s the box origin (0 for a periodic box, where q - 0.0 is q), as
// a global cell index.
__device__ __forceinline__ int cell_of(double x, double y, double z, const Replica &R) {
  int ix = (int)((x - R.o[0]) * R.im[0]) + 1;
  int iy = (int)((y - R.o[1]) * R.im[1]) + 1;
  int iz = (int)((z - R.o[2]) * R.im[2]) + 1;
  if (ix < 0) ix = R.m[0] - 1; else if (ix >= R.m[0]) ix = 0;
  if (iy < 0) iy = R.m[1] - 1; else if (iy >= R.m[1]) iy = 0;
  if (iz < 0) iz = R.m[2] - 1; else if (iz >= R.m[2]) iz = 0;
  return R.cell0 + R.m[0] * R.m[1] * iz + R.m[0] * iy + ix;
}

// A slot in cell c's counter for each active lane of the wavefront: one
// atomic add per distinct cell among the active lanes, whose lanes take
// consecutive slots in lane order.  The lanes are first grouped by cell with
// ballots (no memory traffic); then the first lane of every group adds its
// group's size in the same instruction, so the wavefront waits for one atomic
// round trip however many cells it meets (rows sorted by list length put many
// cells in one wavefront).  Per-particle atomics on one counter would
// serialise.  The order within a cell is settled later by particle id.
__device__ __forceinline__ int cell_slot(int *__restrict__ counts, int c) {
  unsigned long long todo = __ballot(1);
  const int lane = __lane_id();
  int leader_of = lane, rank = 0, size = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int cl = __shfl(c, leader, kWave);
    const unsigned long long same = __ballot(c == cl) & todo;
    if (c == cl) {
      leader_of = leader;
      rank = __popcll(same & ((1ull << lane) - 1ull));
    }
    if (lane == leader) size = __popcll(same);
    todo &= ~same;
  }
  int base = 0;
  if (lane == leader_of) base = atomicAdd(&counts[c], size);
  return __shfl(base, leader_of, kWave) + rank;
}

// Wrap owned particles into their replica's box (one +-L per component), flag
// positions still outside, bin, and take a slot in the cell (cell_slot).
__global__ void k_wrap_bin(double4 *__restrict__ pos, int n, const int *__restrict__ rep,
                           const Replica *__restrict__ reps, int *__restrict__ cell, int *__restrict__ slot,
                           int *__restrict__ ccount, int *__restrict__ err, int wrap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Replica &R = reps[rep[i]];
  double4 q = pos[i];
  if (wrap) {
    const double Lx = R.L[0], Ly = R.L[1], Lz = R.L[2];
    if (q.x < 0.0) q.x += Lx; else if (q.x > Lx) q.x -= Lx;
    if (q.y < 0.0) q.y += Ly; else if (q.y > Ly) q.y -= Ly;
    if (q.z < 0.0) q.z += Lz; else if (q.z > Lz) q.z -= Lz;
    pos[i] = q;
    // the reference lets NaN through this test; a NaN is refused here
    if (!(q.x >= 0.0 && q.x <= Lx && q.y >= 0.0 && q.y <= Ly && q.z >= 0.0 && q.z <= Lz))
      atomicOr(err, kErrInvalidPosition);
  }
  const int c = cell_of(q.x, q.y, q.z, R);
  // 0 <= q <= L bins into cells 1 .. m - 1, which bricks cover; the lower
  // ghost layer is refused like a position outside the box
  const int l = c - R.cell0;
  if (l % R.m[0] == 0 || (l / R.m[0]) % R.m[1] == 0 || l / (R.m[0] * R.m[1]) == 0)
    atomicOr(err, kErrInvalidPosition);
  cell[i] = c;
  slot[i] = cell_slot(ccount, c);
}

__global__ void k_scatter_perm(const int *__restrict__ cell, const int *__restrict__ slot,
                               const int *__restrict__ cstart, int n, int *__restrict__ perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  perm[cstart[cell[i]] + slot[i]] = i;
}

// Order each cell's members by (key, input index) when `key` is given, else
// by input index: perm_in holds the items grouped by cell (k_scatter_perm, in
// atomic order); each thread takes one item, counts the members of its cell
// that sort before it and writes the item at that rank into perm_out.  Keys
// are unique, so this is a permutation, and with the slots of k_wrap_bin
// taken by atomics it makes the order deterministic.  Owned particles are
// keyed by particle id, so the order after a construction depends only on
// the particles and their ids, not on the order an earlier one left.
__device__ __forceinline__ unsigned long long sort_key(const int *__restrict__ key, int i) {
  const unsigned long long k = key ? (unsigned long long)((unsigned)key[i] ^ 0x80000000u) : 0ull;
  return (k << 32) | (unsigned)i;
}

__global__ void k_rank_in_cell(const int *__restrict__ perm_in, const int *__restrict__ cellv,
                               const int *__restrict__ cstart, const int *__restrict__ ccount,
                               const int *__restrict__ key, int n, int *__restrict__ perm_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int item = perm_in[p];
  const int c = cellv[item];
  const int s0 = cstart[c], m = ccount[c];
  const unsigned long long mine = sort_key(key, item);
  int rank = 0;
  for (int q = 0; q < m; q++) rank += sort_key(key, perm_in[s0 + q]) < mine ? 1 : 0;
  perm_out[s0 + rank] = item;
}

__global__ void k_gather_owned(const int *__restrict__ perm, int n, const double4 *__restrict__ pos_in,
                               const double4 *__restrict__ mom_in, const int *__restrict__ id_in,
                               const int *__restrict__ cell_in, const int *__restrict__ rep_in,
                               double4 *__restrict__ pos_out, double4 *__restrict__ mom_out,
                               int *__restrict__ id_out, int *__restrict__ cell_out, int *__restrict__ rep_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
  pos_out[k] = pos_in[i];
  mom_out[k] = mom_in[i];
  id_out[k] = id_in[i];
  cell_out[k] = cell_in[i];
  rep_out[k] = rep_in[i];
}

// Halo of one domain (MDManager::MakePairList's six sequential
// FindBorderParticles/SendBorderParticlesSub passes, mdmanager.cc:269-275):
// along each axis a particle strictly within SL of the low face gets an image
// at +L, strictly within SL of the high face an image at -L; images of images
// make the edges and corners.  Equivalently, one image per non-zero shift
// vector whose every non-zero component passes its face test.  The box is
// [s, e) = [0, L) (MDRect::IsInsideEdge, mdrect.cc:24-47: s < q < s + SL and
// e - SL < q < e).
__device__ __forceinline__ void face_options(const double4 q, const Replica &R, double SL, int lo[3], int hi[3]) {
  const double v[3] = {q.x, q.y, q.z};
  for (int d = 0; d < 3; d++) {
    lo[d] = (v[d] < 0.0 + SL && v[d] > 0.0);
    hi[d] = (v[d] > R.L[d] - SL && v[d] < R.L[d]);
  }
}

__global__ void k_count_images(const double4 *__restrict__ pos, int n, const int *__restrict__ rep,
                               const Replica *__restrict__ reps, double SL, int *__restrict__ nimg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo[3], hi[3];
  face_options(pos[i], reps[rep[i]], SL, lo, hi);
  nimg[i] = (1 + lo[0] + hi[0]) * (1 + lo[1] + hi[1]) * (1 + lo[2] + hi[2]) - 1;
}

__global__ void k_emit_images(const double4 *__restrict__ pos, int n, const int *__restrict__ rep,
                              const Replica *__restrict__ reps, double SL, const int *__restrict__ goff, int ng,
                              int *__restrict__ gsrc, int *__restrict__ gcode,
                              double4 *__restrict__ gpos, int *__restrict__ gcell,
                              int *__restrict__ gslot, int *__restrict__ gcount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = rep[i];
  const Replica &R = reps[r];
  const double4 q = pos[i];
  int lo[3], hi[3];
  face_options(q, R, SL, lo, hi);
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
        const double3 y = image_of(q, code, R.L);
        gsrc[k] = i;
        gcode[k] = ghost_word(code, r);
        gpos[k] = make_double4(y.x, y.y, y.z, 0.0);
        const int cl = cell_of(y.x, y.y, y.z, R);
        gcell[k] = cl;
        gslot[k] = cell_slot(gcount, cl);
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

// ---- units of a process grid: MDManager with several MDUnits in one process
// (mdmanager.cc:194-348), each unit a box [s, e) of the grid on the same device.

// Variables::AdjustPeriodicBoundary (variables.cc:59-79) with the whole box's
// edges: one +-L per component, then a position still outside [0, L] (or NaN)
// is flagged.
__global__ void k_unit_wrap(double4 *__restrict__ pos, int n, double Lx, double Ly, double Lz,
                            int *__restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 q = pos[i];
  if (q.x < 0.0) q.x += Lx; else if (q.x > Lx) q.x -= Lx;
  if (q.y < 0.0) q.y += Ly; else if (q.y > Ly) q.y -= Ly;
  if (q.z < 0.0) q.z += Lz; else if (q.z > Lz) q.z -= Lz;
  pos[i] = q;
  if (!(q.x >= 0.0 && q.x <= Lx && q.y >= 0.0 && q.y <= Ly && q.z >= 0.0 && q.z <= Lz))
    atomicOr(err, kErrInvalidPosition);
}

// MDRect::IsOverBoundary (mdrect.cc:4-21) of direction dir (D_LEFT, D_RIGHT,
// D_BACK, D_FORWARD, D_DOWN, D_UP = 0..5: axis dir / 2, the low face for even
// dir) for a unit whose box spans [s, e) along that axis.
__global__ void k_leave_flags(const double4 *__restrict__ pos, int n, int dir, double s, double e,
                              int *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 q = pos[i];
  const int a = dir >> 1;
  const double v = a == 0 ? q.x : (a == 1 ? q.y : q.z);
  flag[i] = (dir & 1) ? (v >= e) : (v < s);
}

// MDUnit::MakeBufferForSendingParticle (mdunit.cc:78-101): the particles that
// leave go to the send buffer in their order (off = exclusive scan of the
// flags), the others are compacted in theirs.
__global__ void k_partition(const int *__restrict__ flag, const int *__restrict__ off, int n,
                            const double4 *__restrict__ pos, const double4 *__restrict__ mom,
                            const int *__restrict__ id, double4 *__restrict__ pos_keep,
                            double4 *__restrict__ mom_keep, int *__restrict__ id_keep,
                            double4 *__restrict__ pos_send, double4 *__restrict__ mom_send,
                            int *__restrict__ id_send) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = off[i];
  if (flag[i]) {
    pos_send[o] = pos[i];
    mom_send[o] = mom[i];
    id_send[o] = id[i];
  } else {
    pos_keep[i - o] = pos[i];
    mom_keep[i - o] = mom[i];
    id_keep[i - o] = id[i];
  }
}

// One unit of the group as a source of another unit's ghosts.  Along each
// axis the source is in the receiver's slab (same grid coordinate), or its
// left or right neighbour, or several of these at once (a grid of one or two
// units along the axis).  What the reference relays through six sequential
// FindBorderParticles / SendBorderParticlesSub passes (mdmanager.cc:269-275,
// mdunit.cc:125-191) is then, per axis and independently: the position
// itself from the same slab; from the left neighbour, a particle strictly
// within SL of that neighbour's high face (MDRect::IsInsideEdge D_RIGHT,
// mdrect.cc:30-31), shifted by -L when the neighbour's D_RIGHT crosses the
// periodic boundary (ParaInfo::IsOverBoundary, parainfo.cc:256-275); from the
// right neighbour, strictly within SL of its low face (D_LEFT), shifted by +L
// when that crosses.  The ghosts are every combination of one option per axis
// except the particle itself; a ghost relayed through other units carries the
// same position, since each relay adds 0 or +-L to one component only.
struct SrcUnit {
  const double4 *pos;     // the source's owned particles
  int n, base;            // their number and first entry in the concatenation
  int same[3];            // same grid coordinate as the receiver along the axis
  int left[3], right[3];  // the source is the receiver's left / right neighbour
  int lsh[3], rsh[3];     // image shift of what comes from the left / right: -1, 0, +1
  double s[3], e[3];      // the source's box, for its border tests
};
constexpr int kMaxUnits = 64;  // units in one group

// The options of one axis in the order k_emit_images takes its shifts (0,
// +1, -1 for a unit that is its own neighbour), so that a grid of one unit
// places its ghosts exactly as a periodic domain does.
__device__ __forceinline__ int unit_axis_options(const SrcUnit &S, int a, double v, double SL, int sh[3]) {
  int k = 0;
  if (S.same[a]) sh[k++] = 0;
  if (S.right[a] && v < S.s[a] + SL && v > S.s[a]) sh[k++] = S.rsh[a];
  if (S.left[a] && v > S.e[a] - SL && v < S.e[a]) sh[k++] = S.lsh[a];
  return k;
}

__device__ __forceinline__ int unit_source(const SrcUnit *__restrict__ tab, int nsrc, int t) {
  int w = nsrc - 1;
  while (w > 0 && t < tab[w].base) w--;
  return w;
}

__global__ void k_unit_count(const SrcUnit *__restrict__ tab, int nsrc, int total, int self, double SL,
                             int *__restrict__ cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int w = unit_source(tab, nsrc, t);
  const SrcUnit &S = tab[w];
  const double4 q = S.pos[t - S.base];
  int sh[3];
  const int nx = unit_axis_options(S, 0, q.x, SL, sh);
  const int ny = unit_axis_options(S, 1, q.y, SL, sh);
  const int nz = unit_axis_options(S, 2, q.z, SL, sh);
  cnt[t] = nx * ny * nz - (w == self ? 1 : 0);
}

// The images of k_unit_count's count, at off[t]: source index and a word
// holding the shift code and the source unit (ghost_word), the position, the
// receiver's cell and a slot in it.  The shift is added as image_of adds it.
__global__ void k_unit_emit(const SrcUnit *__restrict__ tab, int nsrc, int total, int self, double SL,
                            const int *__restrict__ off, int ng, double Lx, double Ly, double Lz,
                            const Replica *__restrict__ reps, int *__restrict__ gsrc, int *__restrict__ gword,
                            double4 *__restrict__ gpos, int *__restrict__ gcell, int *__restrict__ gslot,
                            int *__restrict__ gcount) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int w = unit_source(tab, nsrc, t);
  const SrcUnit &S = tab[w];
  const int i = t - S.base;
  const double4 q = S.pos[i];
  int ox[3], oy[3], oz[3];
  const int nx = unit_axis_options(S, 0, q.x, SL, ox);
  const int ny = unit_axis_options(S, 1, q.y, SL, oy);
  const int nz = unit_axis_options(S, 2, q.z, SL, oz);
  const Replica &R = reps[0];
  const double L[3] = {Lx, Ly, Lz};
  int k = off[t];
  for (int a = 0; a < nx; a++)
    for (int b = 0; b < ny; b++)
      for (int c = 0; c < nz; c++) {
        const int sx = ox[a], sy = oy[b], sz = oz[c];
        if (w == self && !sx && !sy && !sz) continue;  // the particle itself
        if (k >= ng) return;  // cannot happen: off is the scan of k_unit_count
        const int code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
        const double3 y = image_of(q, code, L);
        gsrc[k] = i;
        gword[k] = ghost_word(code, w);
        gpos[k] = make_double4(y.x, y.y, y.z, 0.0);
        const int cl = cell_of(y.x, y.y, y.z, R);
        gcell[k] = cl;
        gslot[k] = cell_slot(gcount, cl);
        k++;
      }
}

// k_gather_ghosts for a unit: the shift code goes to gcode (replica 0), the
// source unit to gunit.
__global__ void k_gather_unit_ghosts(const int *__restrict__ perm, int ng, int pn, const int *__restrict__ src_in,
                                     const int *__restrict__ word_in, const double4 *__restrict__ gpos_in,
                                     const int *__restrict__ gcell_in, int *__restrict__ gsrc,
                                     int *__restrict__ gcode, int *__restrict__ gunit, double4 *__restrict__ pos,
                                     int *__restrict__ cell) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ng) return;
  const int g = perm[k];
  gsrc[k] = src_in[g];
  gcode[k] = ghost_shift(word_in[g]);
  gunit[k] = ghost_rep(word_in[g]);
  pos[pn + k] = gpos_in[g];
  cell[pn + k] = gcell_in[g];
}

// Owned positions of every unit of a group, by unit slot.
struct UnitPos {
  const double4 *pos[kMaxUnits];
};

// The row order each unit of a group gave its owned particles at its last
// construction (old index -> new), by unit slot.
struct UnitRowInv {
  const int *inv[kMaxUnits];
};

// Every unit of a group in one launch: the flat index t of a launch over all
// units' items (owned particles, ghosts or blocks) belongs to the last unit u
// with base[u] <= t (units without items share their successor's base).
__device__ __forceinline__ int group_unit_of(const int *base, int n, int t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The owned particles of every unit (UpdatePositionHalf, fcalculator.cc:22-32,
// with k_drift's operations).
struct GroupDrift {
  double4 *pos[kMaxUnits];
  const double4 *mom[kMaxUnits];
  int base[kMaxUnits + 1];
  int n;
};

__global__ void k_group_drift(GroupDrift G, double dt2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.base[G.n]) return;
  const int u = group_unit_of(G.base, G.n, t);
  const int i = t - G.base[u];
  double4 q = G.pos[u][i];
  const double4 p = G.mom[u][i];
  q.x += p.x * dt2;
  q.y += p.y * dt2;
  q.z += p.z * dt2;
  G.pos[u][i] = q;
}

// The ghosts of every unit from their sources' current positions (k_unit_refresh
// for the whole group).
struct GroupHalo {
  double4 *pos[kMaxUnits];
  const int *gsrc[kMaxUnits];
  const int *gcode[kMaxUnits];
  const int *gunit[kMaxUnits];
  int pn[kMaxUnits];
  int base[kMaxUnits + 1];
  int n;
};

__global__ void k_group_refresh(GroupHalo G, double Lx, double Ly, double Lz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.base[G.n]) return;
  const int u = group_unit_of(G.base, G.n, t);
  const int k = t - G.base[u];
  const double L[3] = {Lx, Ly, Lz};
  const double3 y = image_of(G.pos[G.gunit[u][k]][G.gsrc[u][k]], G.gcode[u][k], L);
  G.pos[u][G.pn[u] + k] = make_double4(y.x, y.y, y.z, 0.0);
}

// max |p|^2 of every unit: blocks aligned to units (bbase: first block of each
// unit), a partial maximum per block, then one workgroup per unit.
struct GroupMomenta {
  const double4 *mom[kMaxUnits];
  int pn[kMaxUnits];
  int bbase[kMaxUnits + 1];
  int n;
};

__global__ void k_group_max_p2(GroupMomenta G, double *__restrict__ partial) {
  const int b = blockIdx.x;
  const int u = group_unit_of(G.bbase, G.n, b);
  const int i = (b - G.bbase[u]) * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < G.pn[u]) {
    const double4 p = G.mom[u][i];
    v = p.x * p.x + p.y * p.y + p.z * p.z;
  }
  block_max_to_partial(v, partial, b);
}

__global__ void k_group_reduce_max(GroupMomenta G, const double *__restrict__ partial,
                                   unsigned long long *__restrict__ out) {
  __shared__ double red[kBlock];
  const int u = blockIdx.x;
  double m = 0.0;
  for (int k = G.bbase[u] + threadIdx.x; k < G.bbase[u + 1]; k += blockDim.x) m = fmax(m, partial[k]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[u] = (unsigned long long)__double_as_longlong(red[0]);
}

// A unit's ghost sources follow their source units' row orders: the indices
// were taken before the sources ordered their rows.
__global__ void k_remap_unit_ghosts(int *__restrict__ gsrc, const int *__restrict__ gunit, int ng, UnitRowInv R) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ng) return;
  gsrc[k] = R.inv[gunit[k]][gsrc[k]];
}

// MDManager::SendBorderParticles for a unit (mdmanager.cc:283-348): every
// ghost from its source unit's current position plus its shift.
__global__ void k_unit_refresh(double4 *__restrict__ pos, const int *__restrict__ gsrc,
                               const int *__restrict__ gcode, const int *__restrict__ gunit, int pn, int ng,
                               UnitPos P, double Lx, double Ly, double Lz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ng) return;
  const double L[3] = {Lx, Ly, Lz};
  const double3 y = image_of(P.pos[gunit[k]][gsrc[k]], gcode[k], L);
  pos[pn + k] = make_double4(y.x, y.y, y.z, 0.0);
}

}  // namespace mdgpu
