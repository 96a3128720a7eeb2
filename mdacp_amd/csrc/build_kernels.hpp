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

// MeshList's cell of a position (meshlist.cc:164-182): int((q - s)/w) + 1 per
// axis with s = 0, as a global cell index.
__device__ __forceinline__ int cell_of(double x, double y, double z, const Replica &R) {
  int ix = (int)(x * R.im[0]) + 1;
  int iy = (int)(y * R.im[1]) + 1;
  int iz = (int)(z * R.im[2]) + 1;
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

}  // namespace mdgpu
