// force_kernels.hpp -- the pair loops on gfx950: the smoothed-LJ force kick
// (FP64 and mixed FP32) and the pair observers, one workgroup per cell brick.
//
// Reference semantics (kaityo256/mdacp):
//   CalculateForceNext  fcalculator.cc:229-290: for each listed pair within rc,
//                       p_i += df*(q_j - q_i), df = ((24 r^6 - 48)/(r^6 r^6 r^2) + 8 C2) dt
//   potential           4(r^-12 - r^-6 + C2 r^2 + C0), C0/C2 from variables.cc:16-28
//   observers           observer.cc:18-79 (PE and virial, ghost pairs halved)
//
// A workgroup first copies its brick's neighbourhood (list_kernels.hpp) into
// LDS: the owned particles and ghost images of the cells around the brick,
// gathered through the brick's gather list with several loads in flight per
// thread.  Its wavefronts then walk the brick's 64-row slices of the
// sliced-ELL list (16-bit indices into the staged array) and read partners
// from LDS.  A brick whose staged set does not fit the LDS budget reads the
// same partners from global memory (same arithmetic, slower).
//
// Positions of owned particles and ghosts live in one array (`src`, ghosts at
// [pn, pn + ng)), as the halo refresh left them.
//
// Full neighbour list: each owned row lists every partner within SL, so no
// momentum scatter and no atomics: results are bit-reproducible run to run.
#pragma once
#include "device_common.hpp"
#include "list_kernels.hpp"
#include "step_kernels.hpp"

namespace mdgpu {

enum Epilogue {
  kKickOnly = 0,      // p += kick                       (CalculateForce)
  kKickDrift = 1,     // p += kick; q_out = q + p*dt2    (CalculateForce + UpdatePositionHalf)
  kKickLangevin = 2,  // p += kick; Langevin update      (CalculateLangevin)
};

struct BrickList {
  const uint16_t *nbr;      // staged index per (slice, slot, lane)
  const uint8_t *code;      // cell offset of the partner: (dx+1) | (dy+1)<<2 | (dz+1)<<4
  const long long *sbase;   // first slot of each 64-row slice
  const int *slen;          // slots per slice (multiple of kSliceRound)
  const int *glist;         // staged entry -> particle index (owned, or pn + ghost)
  const long long *gbase;   // [brick + 1] first gather-list entry of each brick
  const int *row0, *nrow, *vslice0;
  const int *brep;          // replica of each brick
  const Replica *reps;
  const double4 *ghost;     // FP64 positions of the ghost images (gather entries from pn on)
  int pn;                   // owned particles
  int lds_cap;              // staged entries that fit the LDS budget (incl. the pad entry)
};

// Time step and Langevin step counter of a pair-force launch: the counter is
// read from device memory when `clock` is set (a launch replayed from a graph
// reads the current step), else `step` is used.
struct StepClock {
  double dt;
  uint64_t step;
  const unsigned long long *clock;
  const RunGate *gate;  // a step of a run: nothing to do once the run has halted
};

__device__ __forceinline__ uint64_t step_of(const StepClock &sc) { return sc.clock ? sc.clock[0] : sc.step; }

// 1/x for the pair force: the hardware reciprocal (relative error below
// 5e-8 on gfx950) refined by one Newton step, which squares the error: below
// 2.3e-15 relative, about 10 ulp (measured over 1e-3 <= x <= 1e3).  A second
// step would give the correctly rounded quotient at two more FMAs per pair.
__device__ __forceinline__ double rcp64(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}

// 1/x with a second Newton step: within about an ulp of the quotient, for
// sums that must agree with divisions closely (the observers' PE and virial
// sum r^-12 terms, which would scale the one-step error by 6 with one sign).
__device__ __forceinline__ double rcp64_exact(double x) {
  const double r = rcp64(x);
  return fma(r, fma(-x, r, 1.0), r);
}

// The 8 staged indices of one group of a lane's list slots (one 16-byte load).
__device__ __forceinline__ void unpack_group(const uint4 g, int e[kGroup]) {
  e[0] = g.x & 0xffff; e[1] = g.x >> 16; e[2] = g.y & 0xffff; e[3] = g.y >> 16;
  e[4] = g.z & 0xffff; e[5] = g.z >> 16; e[6] = g.w & 0xffff; e[7] = g.w >> 16;
}

__device__ __forceinline__ void unpack_codes(const uint2 g, int c[kGroup]) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    c[u] = (g.x >> (8 * u)) & 0xff;
    c[u + 4] = (g.y >> (8 * u)) & 0xff;
  }
}

// The list groups of slice s, group g of lane l at [g * kWave + l]: the base
// made wavefront-uniform, so the loads address it from scalar registers with
// the lane's offset as the one vector operand (no 64-bit vector address to
// compute per load, whose registers would collide with the buffers in flight).
__device__ __forceinline__ const uint4 *slice_groups(const uint16_t *nbr, long long first) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)first);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)first >> 32));
  return reinterpret_cast<const uint4 *>(nbr + (long long)(((uint64_t)hi << 32) | lo));
}

// Keeps a prefetched list group at its load: an empty asm that takes the
// group as an operand together with values computed from the current group
// (the row's running sums), so the compiler cannot move the load down to the
// iteration that uses it, where its latency would stall the wavefront.
template <class A>
__device__ __forceinline__ void hold_prefetch(uint4 &g, const A &a, const A &b, const A &c) {
  asm volatile("" : "+v"(g.x), "+v"(g.y), "+v"(g.z), "+v"(g.w) : "v"(a), "v"(b), "v"(c));
}
template <class A>
__device__ __forceinline__ void hold_prefetch(uint2 &g, const A &a, const A &b, const A &c) {
  asm volatile("" : "+v"(g.x), "+v"(g.y) : "v"(a), "v"(b), "v"(c));
}

// A group of 8 list indices as one 128-bit register tuple: as an asm operand
// it stays in four consecutive registers, where a load writes it.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <class A>
__device__ __forceinline__ void hold_prefetch(u32x4 &g, const A &a, const A &b, const A &c) {
  asm volatile("" : "+v"(g) : "v"(a), "v"(b), "v"(c));
}

__device__ __forceinline__ void unpack_group(const u32x4 g, int e[kGroup]) {
  e[0] = g.x & 0xffff; e[1] = g.x >> 16; e[2] = g.y & 0xffff; e[3] = g.y >> 16;
  e[4] = g.z & 0xffff; e[5] = g.z >> 16; e[6] = g.w & 0xffff; e[7] = g.w >> 16;
}

// One group of kGroup list entries of a row: partners from LDS (kStaged) or
// through the gather list, the smoothed-LJ kick added to a.
template <bool kStaged>
__device__ __forceinline__ void group_kick_fp64(const int e[kGroup], const double4 qi,
                                                const double *__restrict__ lds, const double4 *__restrict__ src,
                                                const int *__restrict__ gl, int total, const ForceConsts &fc, int pn,
                                                const double4 *__restrict__ ghost, double &ax, double &ay,
                                                double &az) {
  double qx[kGroup], qy[kGroup], qz[kGroup];
#pragma unroll
  for (int u = 0; u < kGroup; u++) {
    if (kStaged) {
      const double *q = lds + 3 * e[u];
      qx[u] = q[0];
      qy[u] = q[1];
      qz[u] = q[2];
    } else if (e[u] < total) {
      const double4 q = gather_pos(src, ghost, pn, gl[e[u]]);
      qx[u] = q.x;
      qy[u] = q.y;
      qz[u] = q.z;
    } else {
      qx[u] = qy[u] = qz[u] = 1.0e200;
    }
  }
#pragma unroll
  for (int u = 0; u < kGroup; u++) {
    const double dx = qx[u] - qi.x, dy = qy[u] - qi.y, dz = qz[u] - qi.z;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double ir2 = rcp64(r2);
    const double ir6 = ir2 * ir2 * ir2;
    double df = fma(ir2 * ir6, fma(-fc.B, ir6, fc.A), fc.C);
    df = (r2 <= fc.CL2) ? df : 0.0;
    ax = fma(df, dx, ax);
    ay = fma(df, dy, ay);
    az = fma(df, dz, az);
  }
}

// The FP64 kick of one row (one lane of a 64-row slice) over its list
// entries, one group of kGroup per batch: group g at sg[g * kWave + lane]
// (slice_groups).  P buffers hold the indices of the next P
// groups: the loop is unrolled by P so that each buffer is refilled in place
// right after it is unpacked, and waited for only after the P - 1 groups
// computed in between (rotating the buffers instead would copy registers
// whose loads are still in flight, and wait for them at once).  Prefetches
// past the end of the slice re-read its last group, which is in the cache,
// rather than the next slice's first groups, which would add list traffic
// (P <= kMaxAhead).  The cutoff is a select, not a branch (padding
// entries point at a far sentinel).
template <bool kStaged, int P>
__device__ __forceinline__ double3 row_kick_fp64(const uint4 *__restrict__ sg, int lane, int len, const double4 qi,
                                                 const double *__restrict__ lds, const double4 *__restrict__ src,
                                                 const int *__restrict__ gl, int total, const ForceConsts &fc,
                                                 int pn, const double4 *__restrict__ ghost) {
  double ax = 0.0, ay = 0.0, az = 0.0;
  const u32x4 *vg = reinterpret_cast<const u32x4 *>(sg);
  u32x4 ahead[P];
  const unsigned ul = (unsigned)lane;
  const int last = max(len / kGroup - 1, 0);  // loads past the slice re-read its last group
#pragma unroll
  for (int j = 0; j < P; j++) ahead[j] = (vg + min(j, last) * kWave)[ul];
  int k = 0;
  for (; k + P * kGroup <= len; k += P * kGroup) {
#pragma unroll
    for (int j = 0; j < P; j++) {
      int e[kGroup];
      unpack_group(ahead[j], e);
      ahead[j] = (vg + min(k / kGroup + j + P, last) * kWave)[ul];
      __builtin_amdgcn_sched_barrier(0);  // the prefetch is issued before this group's work
      group_kick_fp64<kStaged>(e, qi, lds, src, gl, total, fc, pn, ghost, ax, ay, az);
      // the buffer unpacked next was loaded P - 1 groups of work ago
      hold_prefetch(ahead[(j + 1) % P], ax, ay, az);
    }
  }
  // the last groups of a slice whose length is not a multiple of P groups
  // (len is the slice's: uniform across the wavefront)
#pragma unroll
  for (int j = 0; j + 1 < P; j++) {
    if (k + j * kGroup < len) {
      int e[kGroup];
      unpack_group(ahead[j], e);
      group_kick_fp64<kStaged>(e, qi, lds, src, gl, total, fc, pn, ghost, ax, ay, az);
    }
  }
  return make_double3(ax, ay, az);
}

// How the mixed kernel sums a row's FP32 pair kicks.
enum MixedSum {
  kSumPlain = 0,    // one FP32 running sum
  kSumKahan = 1,    // compensated (Kahan) FP32 running sum
  kSumGroup64 = 2,  // FP32 sum of each group of kGroup kicks, the group sums added in FP64
};

// Mixed-precision kick of one row: FP32 cell-relative partner coordinates,
// FP32 pair forces, summed as kSum says; same groups and loads as
// row_kick_fp64 (`cg`: the lane's groups of cell-offset codes).
template <bool kStaged, int kSum, int P>
__device__ __forceinline__ double3 row_kick_mixed(const uint4 *__restrict__ rg, const uint2 *__restrict__ cg,
                                                  int len, const float4 ri, const float4 *__restrict__ lds,
                                                  const float4 *__restrict__ offs, const float4 *__restrict__ rel,
                                                  const int *__restrict__ gl, int total, const ForceConstsF &fc) {
  float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  uint4 ahead[P];
  uint2 cahead[P];
#pragma unroll
  for (int j = 0; j < P; j++) {
    ahead[j] = rg[j * kWave];
    cahead[j] = cg[j * kWave];
  }
  for (int k = 0; k < len; k += kGroup) {
    int e[kGroup], cc[kGroup];
    unpack_group(ahead[0], e);
    unpack_codes(cahead[0], cc);
#pragma unroll
    for (int j = 0; j + 1 < P; j++) {
      ahead[j] = ahead[j + 1];
      cahead[j] = cahead[j + 1];
    }
    ahead[P - 1] = rg[(k / kGroup + P) * kWave];
    cahead[P - 1] = cg[(k / kGroup + P) * kWave];
    __builtin_amdgcn_sched_barrier(0);  // the prefetch is issued before this group's work
    float4 rj[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; u++) {
      if (kStaged) rj[u] = lds[e[u]];
      else rj[u] = e[u] < total ? rel[gl[e[u]]] : make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
    }
    if (kSum == kSumGroup64) ax = ay = az = 0.f;
#pragma unroll
    for (int u = 0; u < kGroup; u++) {
      const float4 o = offs[cc[u]];
      const float dx = (rj[u].x - ri.x) + o.x, dy = (rj[u].y - ri.y) + o.y, dz = (rj[u].z - ri.z) + o.z;
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      const float ir2 = __builtin_amdgcn_rcpf(r2);
      const float ir6 = ir2 * ir2 * ir2;
      float df = fmaf(ir2 * ir6, fmaf(-fc.B, ir6, fc.A), fc.C);
      df = (r2 <= fc.CL2) ? df : 0.0f;
      if (kSum == kSumKahan) {
        float y, q;
        y = fmaf(df, dx, -cx); q = ax + y; cx = (q - ax) - y; ax = q;
        y = fmaf(df, dy, -cy); q = ay + y; cy = (q - ay) - y; ay = q;
        y = fmaf(df, dz, -cz); q = az + y; cz = (q - az) - y; az = q;
      } else {
        ax = fmaf(df, dx, ax);
        ay = fmaf(df, dy, ay);
        az = fmaf(df, dz, az);
      }
    }
    if (kSum == kSumGroup64) {
      sx += (double)ax;
      sy += (double)ay;
      sz += (double)az;
    }
    hold_prefetch(ahead[P - 1], ax, ay, az);
    hold_prefetch(cahead[P - 1], ax, ay, az);
  }
  if (kSum == kSumGroup64) return make_double3(sx, sy, sz);
  return make_double3((double)ax - (double)cx, (double)ay - (double)cy, (double)az - (double)cz);
}

// One group of a row's list entries in the brick-centred FP32 frame: FP32
// pair kicks summed as kSum says into a (and its compensation cx, or the FP64
// group sums s).
template <bool kStaged, int kSum>
__device__ __forceinline__ void group_kick_mixed_brick(const int e[kGroup], const float4 ri,
                                                       const float4 *__restrict__ lds,
                                                       const double4 *__restrict__ src,
                                                       const double4 *__restrict__ ghost, int pn,
                                                       const int *__restrict__ gl, int total,
                                                       const ForceConstsF &fc, const double3 c, float &ax, float &ay,
                                                       float &az, float &cx, float &cy, float &cz, double &sx,
                                                       double &sy, double &sz) {
  float4 rj[kGroup];
#pragma unroll
  for (int u = 0; u < kGroup; u++) {
    if (kStaged) {
      rj[u] = lds[e[u]];
    } else if (e[u] < total) {
      const double4 q = gather_pos(src, ghost, pn, gl[e[u]]);
      rj[u] = make_float4((float)(q.x - c.x), (float)(q.y - c.y), (float)(q.z - c.z), 0.0f);
    } else {
      rj[u] = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
    }
  }
  if (kSum == kSumGroup64) ax = ay = az = 0.f;
#pragma unroll
  for (int u = 0; u < kGroup; u++) {
    const float dx = rj[u].x - ri.x, dy = rj[u].y - ri.y, dz = rj[u].z - ri.z;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float ir2 = __builtin_amdgcn_rcpf(r2);
    const float ir6 = ir2 * ir2 * ir2;
    float df = fmaf(ir2 * ir6, fmaf(-fc.B, ir6, fc.A), fc.C);
    df = (r2 <= fc.CL2) ? df : 0.0f;
    if (kSum == kSumKahan) {
      float y, q;
      y = fmaf(df, dx, -cx); q = ax + y; cx = (q - ax) - y; ax = q;
      y = fmaf(df, dy, -cy); q = ay + y; cy = (q - ay) - y; ay = q;
      y = fmaf(df, dz, -cz); q = az + y; cz = (q - az) - y; az = q;
    } else {
      ax = fmaf(df, dx, ax);
      ay = fmaf(df, dy, ay);
      az = fmaf(df, dz, az);
    }
  }
  if (kSum == kSumGroup64) {
    sx += (double)ax;
    sy += (double)ay;
    sz += (double)az;
  }
}

// Mixed-precision kick of one row in a brick-centred FP32 frame: partners and
// the row are (float)(q - c) with c the centre of the row's brick, so dx =
// rj - ri needs no cell offset.  Groups, loads and buffers as row_kick_fp64.
template <bool kStaged, int kSum, int P>
__device__ __forceinline__ double3 row_kick_mixed_brick(const uint4 *__restrict__ sg, int lane, int len,
                                                        const float4 ri, const float4 *__restrict__ lds,
                                                        const double4 *__restrict__ src,
                                                        const double4 *__restrict__ ghost, int pn,
                                                        const int *__restrict__ gl, int total,
                                                        const ForceConstsF &fc, const double3 c) {
  float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  const u32x4 *vg = reinterpret_cast<const u32x4 *>(sg);
  u32x4 ahead[P];
  const unsigned ul = (unsigned)lane;
  const int last = max(len / kGroup - 1, 0);  // loads past the slice re-read its last group
#pragma unroll
  for (int j = 0; j < P; j++) ahead[j] = (vg + min(j, last) * kWave)[ul];
  int k = 0;
  for (; k + P * kGroup <= len; k += P * kGroup) {
#pragma unroll
    for (int j = 0; j < P; j++) {
      int e[kGroup];
      unpack_group(ahead[j], e);
      ahead[j] = (vg + min(k / kGroup + j + P, last) * kWave)[ul];
      __builtin_amdgcn_sched_barrier(0);  // the prefetch is issued before this group's work
      group_kick_mixed_brick<kStaged, kSum>(e, ri, lds, src, ghost, pn, gl, total, fc, c, ax, ay, az, cx, cy, cz,
                                            sx, sy, sz);
      hold_prefetch(ahead[(j + 1) % P], ax, ay, az);
    }
  }
#pragma unroll
  for (int j = 0; j + 1 < P; j++) {
    if (k + j * kGroup < len) {
      int e[kGroup];
      unpack_group(ahead[j], e);
      group_kick_mixed_brick<kStaged, kSum>(e, ri, lds, src, ghost, pn, gl, total, fc, c, ax, ay, az, cx, cy, cz,
                                            sx, sy, sz);
    }
  }
  if (kSum == kSumGroup64) return make_double3(sx, sy, sz);
  return make_double3((double)ax - (double)cx, (double)ay - (double)cy, (double)az - (double)cz);
}

// Where the epilogue also writes the next step's first half drift (q + p dt/2
// with the final q and p of this step, the same roundings as the next step's
// UpdatePositionHalf): pos (owned rows) and, for mixed precision, its
// cell-relative FP32 copy.  pos == nullptr: not written.
struct NextDrift {
  double4 *pos;
  float4 *rel;
  const int *cell;
  const double *org;  // column origins of the cell-relative coordinates
};

// The row's id and cell, loaded with its position before the pair loop so
// that their latency overlaps it (only what the epilogue needs).
struct RowExtra {
  int id, cell;
};

__device__ __forceinline__ RowExtra row_extra(int i, int epilogue, const int *__restrict__ id, const NextDrift &nd) {
  RowExtra x;
  x.id = epilogue == kKickLangevin ? id[i] : 0;
  x.cell = nd.rel != nullptr ? nd.cell[i] : 0;
  return x;
}

__device__ __forceinline__ double finish_row(int i, double4 p, const double4 qi, double4 *__restrict__ mom,
                                             double4 *__restrict__ pos_out, int epilogue, double dt2,
                                             const LangevinConsts &lc, const RowExtra &x, const NextDrift &nd,
                                             const Replica &R) {
  if (epilogue == kKickLangevin) langevin_update(p, lc, (uint32_t)x.id);
  mom[i] = p;
  double4 q = qi;
  if (epilogue == kKickDrift) {
    q.x += p.x * dt2;
    q.y += p.y * dt2;
    q.z += p.z * dt2;
    if (pos_out != nullptr) pos_out[i] = q;
  } else if (epilogue == kKickLangevin) {
    if (pos_out != nullptr) pos_out[i] = qi;  // one half drift per Langevin step (mdmanager.cc:186-192)
  }
  if (nd.pos != nullptr && epilogue != kKickOnly) {
    q.x += p.x * dt2;
    q.y += p.y * dt2;
    q.z += p.z * dt2;
    nd.pos[i] = q;
    if (nd.rel != nullptr) nd.rel[i] = to_rel(q, x.cell, nd.org, R);
  }
  return p.x * p.x + p.y * p.y + p.z * p.z;
}

// FP64 kick: one workgroup per brick, wavefronts over the brick's slices;
// U groups of list indices in flight ahead of the group computed.
// The FP64 kick of brick b by the calling workgroup (the body of
// k_force_brick_fp64 and of the batched k_force_group_fp64).
template <int W, int U>
__device__ __forceinline__ void brick_kick_fp64(const int b, const double4 *__restrict__ src,
                                                double4 *__restrict__ mom, double4 *__restrict__ pos_out,
                                                const BrickList &bl, const ForceConsts &fc, int epilogue,
                                                const StepClock &sc, const int *__restrict__ id,
                                                double *__restrict__ vpart, const NextDrift &nd) {
  extern __shared__ __align__(16) double lds64[];
  __shared__ int slice_next;  // the next slice a wavefront takes
  constexpr int T = kWave * W;
  const Replica &Rp = bl.reps[bl.brep[b]];
  const LangevinConsts lc = langevin_consts(Rp, sc.dt, step_of(sc));
  const int nrow = bl.nrow[b], row0 = bl.row0[b], vs0 = bl.vslice0[b];
  if (nrow == 0) {  // a brick without rows: nothing to stage
    if (vpart) block_max_to_partial(0.0, vpart, b);
    return;
  }
  const long long g0 = bl.gbase[b];
  const int total = (int)(bl.gbase[b + 1] - g0);
  const int *gl = bl.glist + g0;
  const bool staged = total < bl.lds_cap;  // uniform across the workgroup
  if (threadIdx.x == 0) slice_next = W;
  if (staged) stage_fp64<T>(gl, total, src, lds64, bl.pn, bl.ghost);
  __syncthreads();
  const int nsl = (nrow + kWave - 1) / kWave;
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  double vmax = 0.0;
  // slices come longest first: each wavefront takes the next one as it finishes
  for (int sl = wv; sl < nsl; sl = take_next(&slice_next, lane)) {
    const int s = vs0 + sl;
    const int v = sl * kWave + lane;
    const bool ok = v < nrow;
    const int i = row0 + (ok ? v : nrow - 1);
    const double4 qi = src[i];
    double4 p = mom[i];  // loaded before the pair loop, which hides its latency
    const RowExtra x = row_extra(i, epilogue, id, nd);
    const int len = bl.slen[s];
    const uint4 *sg = slice_groups(bl.nbr, bl.sbase[s]);
    const double3 a = staged ? row_kick_fp64<true, U>(sg, lane, len, qi, lds64, src, gl, total, fc, bl.pn, bl.ghost)
                             : row_kick_fp64<false, U>(sg, lane, len, qi, lds64, src, gl, total, fc, bl.pn, bl.ghost);
    if (ok) {
      p.x += a.x;
      p.y += a.y;
      p.z += a.z;
      vmax = fmax(vmax, finish_row(i, p, qi, mom, pos_out, epilogue, fc.dt2, lc, x, nd, Rp));
    }
  }
  if (vpart) block_max_to_partial(vmax, vpart, b);
}

template <int W, int U>
__global__ __launch_bounds__(kWave *W) void k_force_brick_fp64(const double4 *__restrict__ src,
                                                              double4 *__restrict__ mom,
                                                              double4 *__restrict__ pos_out, BrickList bl,
                                                              ForceConsts fc, int epilogue, StepClock sc,
                                                              const int *__restrict__ id,
                                                              double *__restrict__ vpart, NextDrift nd) {
  if (run_halted(sc.gate)) return;
  brick_kick_fp64<W, U>(blockIdx.x, src, mom, pos_out, bl, fc, epilogue, sc, id, vpart, nd);
}

// The units of a process grid, kicked in one launch: unit u's bricks are the
// launch's workgroups [brick0 of u, brick0 of u + 1), each doing what the
// unit's own kick does (CalculateForce in place, no epilogue), on the unit's
// own arrays.  One launch of all the units' bricks fills the CUs where one
// launch per unit ends each unit with a partial round of workgroups.  The
// table travels in the kernel arguments (1 KB for kMaxKickUnits units).
struct UnitKick {
  const double4 *src;  // positions, owned then ghosts (kicked in place: no position output)
  double4 *mom;
  BrickList bl;
  int brick0;          // the unit's first workgroup in the launch
};
constexpr int kMaxKickUnits = 8;
struct UnitKicks {
  UnitKick u[kMaxKickUnits];
  int n;
};

// The unit of workgroup b (units in brick0 order; uniform across the workgroup).
__device__ __forceinline__ int group_kick_unit(const UnitKicks &uk, int b) {
  int u = 0;
  while (u + 1 < uk.n && b >= uk.u[u + 1].brick0) u++;
  return u;
}

template <int W, int U>
__global__ __launch_bounds__(kWave *W) void k_force_group_fp64(UnitKicks uk, ForceConsts fc, StepClock sc) {
  const int b = blockIdx.x;
  const UnitKick &k = uk.u[group_kick_unit(uk, b)];
  brick_kick_fp64<W, U>(b - k.brick0, k.src, k.mom, nullptr, k.bl, fc, kKickOnly, sc, nullptr, nullptr,
                        NextDrift{nullptr, nullptr, nullptr, nullptr});
}

// ---------------------------------------------------------------------------
// Pipelined FP64 kick: the same per-brick work as k_force_brick_fp64, with the
// staging of the next brick overlapped with the pair loop of the current one.
//
// The grid is persistent: workgroup g serves bricks g, g + G, g + 2G, ... (G =
// gridDim.x, one workgroup per CU).  Each workgroup has W compute wavefronts
// and P producer wavefronts and two LDS buffers.  While the compute wavefronts
// walk the slices of brick b from one buffer, the producers copy the
// neighbourhood of brick b + G into the other with LDS-direct loads
// (global_load_lds_dword{,x4}: lane l of one load writes at base + l * size,
// for size 4 or 16 bytes).  A buffer keeps x, y of entry e at xy[e] (16 bytes,
// one x4 load per 64 entries) and z at z[e] (8 bytes, two lanes per entry, one
// dword load per 32 entries): 24 bytes per entry, as the one-workgroup-per-brick
// kernel stages them.  These loads hold no registers, and a producer waits only
// for its own loads, so the compute wavefronts never wait for the staging
// (their memory waits count their own loads only).  One barrier per brick hands
// the buffers over.  Every workgroup runs the same number of iterations, so
// every wavefront reaches every barrier.
// ---------------------------------------------------------------------------
constexpr int kStageChunk = 64;  // entries one producer x4 load moves

struct PipeBuf {
  double2 *xy;  // [cap] x, y
  double *z;    // [cap] z
};

// Copy the positions of gather-list entries [0, total) into buf with LDS-direct
// loads, chunks of kStageChunk entries spread over the P producer wavefronts
// (this one is producer pw), plus the far sentinel at entry `total` (the pad
// entry of the list).  Returns after this wavefront's loads have landed.
__device__ __forceinline__ void stage_fp64_dma(const int *__restrict__ gl, int total, const double4 *__restrict__ src,
                                               const double4 *__restrict__ ghost, int pn, PipeBuf buf, int pw,
                                               int P, int lane) {
  constexpr int kBatch = 8;  // chunks whose indices are loaded together
  const int nchunk = (total + kStageChunk - 1) / kStageChunk;
  for (int c0 = pw; c0 < nchunk; c0 += P * kBatch) {
    int j[kBatch];
#pragma unroll
    for (int k = 0; k < kBatch; k++) {
      const int u = (c0 + k * P) * kStageChunk + lane;
      j[k] = u < total ? gl[u] : -1;
    }
#pragma unroll
    for (int k = 0; k < kBatch; k++) {
      const int c = c0 + k * P;
      if (c >= nchunk) break;  // uniform
      const double4 *rec = j[k] >= 0 ? (j[k] < pn ? src : ghost) + j[k] : nullptr;
      if (rec != nullptr)
        __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)rec,
                                         (void __attribute__((address_space(3))) *)(buf.xy + kStageChunk * c),
                                         16, 0, 0);
      // z: lane l moves dword l & 1 of entry (l >> 1) of each half of the chunk
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int jz = __shfl(j[k], 32 * h + (lane >> 1), kWave);
        if (jz >= 0) {
          const char *g = reinterpret_cast<const char *>((jz < pn ? src : ghost) + jz) + 16 + 4 * (lane & 1);
          __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)g,
                                           (void __attribute__((address_space(3))) *)(buf.z + kStageChunk * c + 32 * h),
                                           4, 0, 0);
        }
      }
    }
  }
  if (pw == 0 && lane == 0) {
    buf.xy[total] = make_double2(1.0e200, 1.0e200);
    buf.z[total] = 1.0e200;
  }
  __builtin_amdgcn_s_waitcnt(0);  // this wavefront's loads have landed in LDS
}

// row_kick_fp64 over a pipelined buffer (x, y and z in two arrays).
template <int P>
__device__ __forceinline__ double3 row_kick_fp64_buf(const uint4 *__restrict__ rg, int len, const double4 qi,
                                                     const PipeBuf buf, const ForceConsts &fc) {
  double ax = 0.0, ay = 0.0, az = 0.0;
  uint4 ahead[P];
#pragma unroll
  for (int j = 0; j < P; j++) ahead[j] = rg[j * kWave];
  for (int k = 0; k < len; k += kGroup) {
    int e[kGroup];
    unpack_group(ahead[0], e);
#pragma unroll
    for (int j = 0; j + 1 < P; j++) ahead[j] = ahead[j + 1];
    ahead[P - 1] = rg[(k / kGroup + P) * kWave];
    __builtin_amdgcn_sched_barrier(0);  // the prefetch is issued before this group's work
    double qx[kGroup], qy[kGroup], qz[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; u++) {
      const double2 q = buf.xy[e[u]];
      qx[u] = q.x;
      qy[u] = q.y;
      qz[u] = buf.z[e[u]];
    }
#pragma unroll
    for (int u = 0; u < kGroup; u++) {
      const double dx = qx[u] - qi.x, dy = qy[u] - qi.y, dz = qz[u] - qi.z;
      const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
      const double ir2 = rcp64(r2);
      const double ir6 = ir2 * ir2 * ir2;
      double df = fma(ir2 * ir6, fma(-fc.B, ir6, fc.A), fc.C);
      df = (r2 <= fc.CL2) ? df : 0.0;
      ax = fma(df, dx, ax);
      ay = fma(df, dy, ay);
      az = fma(df, dz, az);
    }
    hold_prefetch(ahead[P - 1], ax, ay, az);
  }
  return make_double3(ax, ay, az);
}

template <int W, int U, int P>
__global__ __launch_bounds__(kWave *(W + P)) void k_force_pipe_fp64(const double4 *__restrict__ src,
                                                                   double4 *__restrict__ mom,
                                                                   double4 *__restrict__ pos_out, BrickList bl,
                                                                   ForceConsts fc, int epilogue, StepClock sc,
                                                                   const int *__restrict__ id,
                                                                   double *__restrict__ vpart, NextDrift nd,
                                                                   int nbrick) {
  extern __shared__ __align__(16) double lds64[];
  __shared__ int slice_next[2];
  __shared__ double wmax[2][W];
  if (run_halted(sc.gate)) return;
  const int cap = bl.lds_cap;  // entries per buffer, pad entry included
  const PipeBuf buf0{reinterpret_cast<double2 *>(lds64), lds64 + 2 * (size_t)cap};
  const PipeBuf buf1{reinterpret_cast<double2 *>(lds64 + 3 * (size_t)cap), lds64 + 5 * (size_t)cap};
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const bool producer = wv >= W;
  const int G = gridDim.x;
  int b = blockIdx.x;
  // brick b's staging: skipped for a brick without rows or one too large for a buffer
  auto stage = [&](int bb, const PipeBuf &buf) {
    if (bb >= nbrick || bl.nrow[bb] == 0) return;
    const long long g0 = bl.gbase[bb];
    const int total = (int)(bl.gbase[bb + 1] - g0);
    if (total >= cap) return;
    stage_fp64_dma(bl.glist + g0, total, src, bl.ghost, bl.pn, buf, wv - W, P, lane);
  };
  if (producer) stage(b, buf0);
  if (threadIdx.x == 0) slice_next[0] = W;
  __syncthreads();
  for (int it = 0; b < nbrick; it++, b += G) {
    const int cur = it & 1;
    const PipeBuf lbuf = cur ? buf1 : buf0;
    if (producer) {
      stage(b + G, cur ? buf0 : buf1);
      if (threadIdx.x == W * kWave) slice_next[cur ^ 1] = W;
    } else {
      double vmax = 0.0;
      const int nrow = bl.nrow[b];
      if (nrow > 0) {
        const Replica &Rp = bl.reps[bl.brep[b]];
        const LangevinConsts lc = langevin_consts(Rp, sc.dt, step_of(sc));
        const int row0 = bl.row0[b], vs0 = bl.vslice0[b];
        const long long g0 = bl.gbase[b];
        const int total = (int)(bl.gbase[b + 1] - g0);
        const int *gl = bl.glist + g0;
        const bool staged = total < cap;  // uniform across the workgroup
        const int nsl = (nrow + kWave - 1) / kWave;
        for (int sl = wv; sl < nsl; sl = take_next(&slice_next[cur], lane)) {
          const int s = vs0 + sl;
          const int v = sl * kWave + lane;
          const bool ok = v < nrow;
          const int i = row0 + (ok ? v : nrow - 1);
          const double4 qi = src[i];
          double4 p = mom[i];
          const RowExtra x = row_extra(i, epilogue, id, nd);
          const int len = bl.slen[s];
          const uint4 *rg = reinterpret_cast<const uint4 *>(bl.nbr + bl.sbase[s] + lane * kGroup);
          const double3 a = staged ? row_kick_fp64_buf<U>(rg, len, qi, lbuf, fc)
                                   : row_kick_fp64<false, U>(slice_groups(bl.nbr, bl.sbase[s]), lane, len, qi,
                                                             nullptr, src, gl, total, fc, bl.pn, bl.ghost);
          if (ok) {
            p.x += a.x;
            p.y += a.y;
            p.z += a.z;
            vmax = fmax(vmax, finish_row(i, p, qi, mom, pos_out, epilogue, fc.dt2, lc, x, nd, Rp));
          }
        }
      }
      vmax = wave_max(vmax);
      if (lane == 0) wmax[cur][wv] = vmax;
    }
    __syncthreads();  // brick b done; brick b + G staged in the other buffer
    if (vpart && threadIdx.x == 0) {
      double m = wmax[cur][0];
      for (int k = 1; k < W; k++) m = fmax(m, wmax[cur][k]);
      vpart[b] = m;
    }
  }
}

// Mixed precision kick: FP32 cell-relative coordinates (staged as float4),
// FP32 pair forces, compensated (Kahan) FP32 sums; the sum is added to the
// FP64 momentum and the drift is FP64.  The partner's cell offset comes from
// the code array, so dx = (rel_j - rel_i) + off*w stays small.  Ghost images
// carry coordinates relative to the extended-grid cell they are binned in,
// like owned particles.
template <int W, int U, int kSum>
__global__ __launch_bounds__(kWave *W) void k_force_brick_mixed(
    const float4 *__restrict__ rel, const double4 *__restrict__ pos, double4 *__restrict__ mom,
    double4 *__restrict__ pos_out, BrickList bl, ForceConstsF fc, double dt2, int epilogue, StepClock sc,
    const int *__restrict__ id, double *__restrict__ vpart, NextDrift nd) {
  extern __shared__ float4 lds32[];
  __shared__ float4 offs[64];
  __shared__ int slice_next;  // the next slice a wavefront takes
  constexpr int T = kWave * W;
  if (run_halted(sc.gate)) return;
  const int b = blockIdx.x;
  const Replica &Rp = bl.reps[bl.brep[b]];
  const LangevinConsts lc = langevin_consts(Rp, sc.dt, step_of(sc));
  if (threadIdx.x < 64) {
    const int c = threadIdx.x;
    const int ox = (c & 3) - 1, oy = ((c >> 2) & 3) - 1, oz = ((c >> 4) & 3) - 1;
    const float w0 = (float)Rp.w[0], w1 = (float)Rp.w[1], w2 = (float)Rp.w[2];
    offs[c] = make_float4(ox * w0, oy * w1, oz * w2, 0.0f);
  }
  const int nrow = bl.nrow[b], row0 = bl.row0[b], vs0 = bl.vslice0[b];
  if (nrow == 0) {  // a brick without rows: nothing to stage
    if (vpart) block_max_to_partial(0.0, vpart, b);
    return;
  }
  const long long g0 = bl.gbase[b];
  const int total = (int)(bl.gbase[b + 1] - g0);
  const int *gl = bl.glist + g0;
  const bool staged = total < bl.lds_cap;
  if (staged) {
    for (int u0 = threadIdx.x; u0 < total; u0 += T * kStageBatch) {
      int j[kStageBatch];
#pragma unroll
      for (int k = 0; k < kStageBatch; k++) {
        const int u = u0 + k * T;
        j[k] = u < total ? gl[u] : 0;
      }
      float4 q[kStageBatch];
#pragma unroll
      for (int k = 0; k < kStageBatch; k++) q[k] = rel[j[k]];
#pragma unroll
      for (int k = 0; k < kStageBatch; k++)
        if (u0 + k * T < total) lds32[u0 + k * T] = q[k];
    }
    if (threadIdx.x == 0) lds32[total] = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
  }
  if (threadIdx.x == 0) slice_next = W;
  __syncthreads();
  const int nsl = (nrow + kWave - 1) / kWave;
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  double vmax = 0.0;
  // slices come longest first: each wavefront takes the next one as it finishes
  for (int sl = wv; sl < nsl; sl = take_next(&slice_next, lane)) {
    const int s = vs0 + sl;
    const int v = sl * kWave + lane;
    const bool ok = v < nrow;
    const int i = row0 + (ok ? v : nrow - 1);
    const float4 ri = rel[i];
    const int len = bl.slen[s];
    const uint4 *rg = reinterpret_cast<const uint4 *>(bl.nbr + bl.sbase[s] + lane * kGroup);
    const uint2 *cg = reinterpret_cast<const uint2 *>(bl.code + bl.sbase[s] + lane * kGroup);
    const double3 a = staged ? row_kick_mixed<true, kSum, U>(rg, cg, len, ri, lds32, offs, rel, gl, total, fc)
                             : row_kick_mixed<false, kSum, U>(rg, cg, len, ri, lds32, offs, rel, gl, total, fc);
    if (ok) {
      // loaded after the pair loop: held across it, they would cost this
      // kernel a workgroup per CU (registers)
      double4 p = mom[i];
      const double4 qi = pos[i];
      const RowExtra x = row_extra(i, epilogue, id, nd);
      p.x += a.x;
      p.y += a.y;
      p.z += a.z;
      vmax = fmax(vmax, finish_row(i, p, qi, mom, pos_out, epilogue, dt2, lc, x, nd, Rp));
    }
  }
  if (vpart) block_max_to_partial(vmax, vpart, b);
}

// Mixed-precision kick in a brick-centred FP32 frame: the workgroup stages its
// brick's neighbourhood from the FP64 positions as (float)(q - c), c the
// centre of the brick's cells, 16 bytes per entry; rows use the same frame, so
// a partner separation is one FP32 subtraction per axis, with no cell offset,
// and the list's cell-offset codes are not read.  The coordinates stay within
// about half a neighbourhood (9 at 4x4x4-cell bricks of 3-wide cells), so
// each carries a rounding of at most 2^-21 to 2^-20; the cell-relative frame
// (k_force_brick_mixed) rounds 4 times finer at the cost of one offset lookup
// per pair.  Sums, epilogue and FP64 momenta as in k_force_brick_mixed.
template <int W, int U, int kSum>
__device__ __forceinline__ void brick_kick_mixed_brick(const int b, const double4 *__restrict__ src,
                                                       double4 *__restrict__ mom, double4 *__restrict__ pos_out,
                                                       const BrickList &bl, const ForceConstsF &fc, double dt2,
                                                       int epilogue, const StepClock &sc,
                                                       const int *__restrict__ id, double *__restrict__ vpart,
                                                       const NextDrift &nd) {
  extern __shared__ float4 lds32[];
  __shared__ int slice_next;  // the next slice a wavefront takes
  constexpr int T = kWave * W;
  const Replica &Rp = bl.reps[bl.brep[b]];
  const LangevinConsts lc = langevin_consts(Rp, sc.dt, step_of(sc));
  const int nrow = bl.nrow[b], row0 = bl.row0[b], vs0 = bl.vslice0[b];
  if (nrow == 0) {  // a brick without rows: nothing to stage
    if (vpart) block_max_to_partial(0.0, vpart, b);
    return;
  }
  const BrickBox X = brick_box(b, Rp);
  const double3 c = make_double3(Rp.o[0] + (X.lo[0] + 0.5 * (X.n[0] - 2)) * Rp.w[0],
                                 Rp.o[1] + (X.lo[1] + 0.5 * (X.n[1] - 2)) * Rp.w[1],
                                 Rp.o[2] + (X.lo[2] + 0.5 * (X.n[2] - 2)) * Rp.w[2]);
  const long long g0 = bl.gbase[b];
  const int total = (int)(bl.gbase[b + 1] - g0);
  const int *gl = bl.glist + g0;
  const bool staged = total < bl.lds_cap;
  if (staged) {
    for (int u0 = threadIdx.x; u0 < total; u0 += T * kStageBatch) {
      int j[kStageBatch];
#pragma unroll
      for (int k = 0; k < kStageBatch; k++) {
        const int u = u0 + k * T;
        j[k] = u < total ? gl[u] : 0;
      }
      double4 q[kStageBatch];
#pragma unroll
      for (int k = 0; k < kStageBatch; k++) q[k] = gather_pos(src, bl.ghost, bl.pn, j[k]);
#pragma unroll
      for (int k = 0; k < kStageBatch; k++)
        if (u0 + k * T < total)
          lds32[u0 + k * T] =
              make_float4((float)(q[k].x - c.x), (float)(q[k].y - c.y), (float)(q[k].z - c.z), 0.0f);
    }
    if (threadIdx.x == 0) lds32[total] = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
  }
  if (threadIdx.x == 0) slice_next = W;
  __syncthreads();
  const int nsl = (nrow + kWave - 1) / kWave;
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  double vmax = 0.0;
  // slices come longest first: each wavefront takes the next one as it finishes
  for (int sl = wv; sl < nsl; sl = take_next(&slice_next, lane)) {
    const int s = vs0 + sl;
    const int v = sl * kWave + lane;
    const bool ok = v < nrow;
    const int i = row0 + (ok ? v : nrow - 1);
    float4 ri;
    {
      const double4 q = src[i];
      ri = make_float4((float)(q.x - c.x), (float)(q.y - c.y), (float)(q.z - c.z), 0.0f);
    }
    const int len = bl.slen[s];
    const uint4 *sg = slice_groups(bl.nbr, bl.sbase[s]);
    const double3 a =
        staged ? row_kick_mixed_brick<true, kSum, U>(sg, lane, len, ri, lds32, src, bl.ghost, bl.pn, gl, total, fc, c)
               : row_kick_mixed_brick<false, kSum, U>(sg, lane, len, ri, lds32, src, bl.ghost, bl.pn, gl, total, fc,
                                                      c);
    if (ok) {
      // loaded after the pair loop (registers)
      double4 p = mom[i];
      const double4 qi = src[i];
      const RowExtra x = row_extra(i, epilogue, id, nd);
      p.x += a.x;
      p.y += a.y;
      p.z += a.z;
      vmax = fmax(vmax, finish_row(i, p, qi, mom, pos_out, epilogue, dt2, lc, x, nd, Rp));
    }
  }
  if (vpart) block_max_to_partial(vmax, vpart, b);
}

template <int W, int U, int kSum>
__global__ __launch_bounds__(kWave *W) void k_force_brick_mixed_brick(
    const double4 *__restrict__ src, double4 *__restrict__ mom, double4 *__restrict__ pos_out, BrickList bl,
    ForceConstsF fc, double dt2, int epilogue, StepClock sc, const int *__restrict__ id, double *__restrict__ vpart,
    NextDrift nd) {
  if (run_halted(sc.gate)) return;
  brick_kick_mixed_brick<W, U, kSum>(blockIdx.x, src, mom, pos_out, bl, fc, dt2, epilogue, sc, id, vpart, nd);
}

// The units of a process grid kicked in one launch (UnitKick), mixed
// precision in the brick-centred frame.
template <int W, int U, int kSum>
__global__ __launch_bounds__(kWave *W) void k_force_group_mixed_brick(UnitKicks uk, ForceConstsF fc, double dt2,
                                                                     StepClock sc) {
  const int b = blockIdx.x;
  const UnitKick &k = uk.u[group_kick_unit(uk, b)];
  brick_kick_mixed_brick<W, U, kSum>(b - k.brick0, k.src, k.mom, nullptr, k.bl, fc, dt2, kKickOnly, sc, nullptr,
                                     nullptr, NextDrift{nullptr, nullptr, nullptr, nullptr});
}

// Dynamic pruning of the pair list (no counterpart in the reference: it
// changes which list entries the fused step's kick visits, not which pairs
// interact).  The entries of each row whose partner lies within rin2 of the
// row at the current half-drifted positions are copied, in list order, into
// the inner list, which has the slice layout of the outer list; slen_in[s]
// receives the slice's inner length (a multiple of kGroup) and the rest of
// each row up to it is the brick's pad entry.  The host keeps the inner list
// only while every pair outside it stays beyond rc (the PairList budget
// argument with rin - rc in place of the margin), so the kick sums the same
// pairs in the same order.
__device__ __forceinline__ uint32_t put16(uint32_t w, int half, uint32_t v) {
  return half ? ((w & 0xffffu) | (v << 16)) : ((w & 0xffff0000u) | v);
}

__device__ __forceinline__ uint32_t put8(uint32_t w, int byte, uint32_t v) {
  const int sh = 8 * byte;
  return (w & ~(0xffu << sh)) | (v << sh);
}

__device__ __forceinline__ int wave_max_int(int v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

template <int W>
__global__ __launch_bounds__(kWave *W) void k_prune_brick(const double4 *__restrict__ src, BrickList bl, double rin2,
                                                         uint16_t *__restrict__ nbr_in, uint8_t *__restrict__ code_in,
                                                         int *__restrict__ slen_in) {
  extern __shared__ __align__(16) double lds64[];
  constexpr int T = kWave * W;
  const int b = blockIdx.x;
  const int nrow = bl.nrow[b], row0 = bl.row0[b], vs0 = bl.vslice0[b];
  if (nrow == 0) return;
  const long long g0 = bl.gbase[b];
  const int total = (int)(bl.gbase[b + 1] - g0);
  const int *gl = bl.glist + g0;
  const bool staged = total < bl.lds_cap;
  if (staged) stage_fp64<T>(gl, total, src, lds64, bl.pn, bl.ghost);
  __syncthreads();
  const uint32_t pad16 = (uint32_t)total | ((uint32_t)total << 16);
  const uint32_t pad8 = kZeroOffset * 0x01010101u;
  const int nsl = (nrow + kWave - 1) / kWave;
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int sl = wv; sl < nsl; sl += W) {
    const int s = vs0 + sl;
    const int v = sl * kWave + lane;
    const bool ok = v < nrow;
    const int i = row0 + (ok ? v : nrow - 1);
    const double4 qi = src[i];
    const int len = bl.slen[s];
    const long long base = bl.sbase[s] + lane * kGroup;
    const uint4 *rg = reinterpret_cast<const uint4 *>(bl.nbr + base);
    // a list without cell offset codes (FP64) gets an inner list without them
    const bool codes = bl.code != nullptr;
    const uint2 *cg = codes ? reinterpret_cast<const uint2 *>(bl.code + base) : nullptr;
    uint4 *og = reinterpret_cast<uint4 *>(nbr_in + base);
    uint2 *oc = codes ? reinterpret_cast<uint2 *>(code_in + base) : nullptr;
    uint32_t w0 = pad16, w1 = pad16, w2 = pad16, w3 = pad16, c0 = pad8, c1 = pad8;
    int kept = 0;
    for (int k = 0; k < len; k += kGroup) {
      int e[kGroup], cc[kGroup];
      unpack_group(rg[(k / kGroup) * kWave], e);
      if (codes) {
        unpack_codes(cg[(k / kGroup) * kWave], cc);
      } else {
#pragma unroll
        for (int u = 0; u < kGroup; u++) cc[u] = kZeroOffset;
      }
#pragma unroll
      for (int u = 0; u < kGroup; u++) {
        double qx, qy, qz;
        if (staged) {
          qx = lds64[3 * e[u]];
          qy = lds64[3 * e[u] + 1];
          qz = lds64[3 * e[u] + 2];
        } else if (e[u] < total) {
          const double4 q = gather_pos(src, bl.ghost, bl.pn, gl[e[u]]);
          qx = q.x;
          qy = q.y;
          qz = q.z;
        } else {
          qx = qy = qz = 1.0e200;
        }
        const double dx = qx - qi.x, dy = qy - qi.y, dz = qz - qi.z;
        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
        if (ok && e[u] < total && r2 <= rin2) {
          const int j = kept % kGroup, wd = j >> 1, hf = j & 1;
          const uint32_t ev = (uint32_t)e[u], cv = (uint32_t)cc[u];
          w0 = wd == 0 ? put16(w0, hf, ev) : w0;
          w1 = wd == 1 ? put16(w1, hf, ev) : w1;
          w2 = wd == 2 ? put16(w2, hf, ev) : w2;
          w3 = wd == 3 ? put16(w3, hf, ev) : w3;
          c0 = j < 4 ? put8(c0, j, cv) : c0;
          c1 = j >= 4 ? put8(c1, j - 4, cv) : c1;
          kept++;
          if (kept % kGroup == 0) {
            og[(kept / kGroup - 1) * kWave] = make_uint4(w0, w1, w2, w3);
            if (codes) oc[(kept / kGroup - 1) * kWave] = make_uint2(c0, c1);
            w0 = w1 = w2 = w3 = pad16;
            c0 = c1 = pad8;
          }
        }
      }
    }
    const int mine = (kept + kGroup - 1) / kGroup;  // groups this row fills
    const int groups = wave_max_int(mine);          // groups of the slice's inner list
    if (kept % kGroup) {
      og[(kept / kGroup) * kWave] = make_uint4(w0, w1, w2, w3);
      if (codes) oc[(kept / kGroup) * kWave] = make_uint2(c0, c1);
    }
    for (int g = mine; g < groups; g++) {
      og[g * kWave] = make_uint4(pad16, pad16, pad16, pad16);
      if (codes) oc[g * kWave] = make_uint2(pad8, pad8);
    }
    if (lane == 0) slen_in[s] = groups * kGroup;
  }
}

// Observers over the list: kinetic energy, PE and virial sums per brick
// (partial[brick*3 + {0,1,2}]), reduced in fixed order by k_reduce_partials
// or, per replica, k_replica_sum3,
// so the results are deterministic.  Each list entry carries weight 1/2:
// owned-owned pairs appear in two rows, owned-ghost pairs in one row and are
// halved by the reference (observer.cc:39-41,70-72).  Owned partners are at
// their current positions, ghosts where the last halo refresh put them, as
// the reference's are.  The pair terms come from one reciprocal per pair,
// ir2 = 1/r^2 (rcp64_exact) and ir6 = ir2^3: half of PE 4 (ir6^2 - ir6 + C2 r^2
// + C0) and half of the virial ir6 (24 - 48 ir6) + 8 C2 r^2, which are the
// reference's 4 (1/r^12 - 1/r^6 + C2 r^2 + C0) and
// ((24 r^6 - 48)/(r^6 r^6 r^2) + 8 C2) r^2 to a few roundings.  A row sums
// ir6^2 - ir6 + C2 r^2 + C0 and 3 ir6 - 6 ir6^2 + C2 r^2 and scales the sums
// by 2 and 4 (exact) at the end.  r^2 is summed without fused multiply-adds,
// so the cutoff keeps the reference's pairs.  Indices come 8 per 16-byte load,
// the next group in flight while a group is computed, as in the kick.
template <int W>
__global__ __launch_bounds__(kWave *W) void k_observe_brick(const double4 *__restrict__ pos,
                                                          const double4 *__restrict__ mom, BrickList bl,
                                                          double CL2, double C2, double C0,
                                                          double *__restrict__ partial) {
  extern __shared__ __align__(16) double lds64[];
  __shared__ double red[3][W];
  constexpr int T = kWave * W;
  const int b = blockIdx.x;
  const int nrow = bl.nrow[b], row0 = bl.row0[b], vs0 = bl.vslice0[b];
  if (nrow == 0) {
    if (threadIdx.x < 3) partial[(long long)b * 3 + threadIdx.x] = 0.0;
    return;
  }
  const long long g0 = bl.gbase[b];
  const int total = (int)(bl.gbase[b + 1] - g0);
  const int *gl = bl.glist + g0;
  const bool staged = total < bl.lds_cap;
  if (staged) stage_fp64<T>(gl, total, pos, lds64, 0, pos);
  __syncthreads();
  const int nsl = (nrow + kWave - 1) / kWave;
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  double ke = 0.0, pe = 0.0, vir = 0.0;
  for (int sl = wv; sl < nsl; sl += W) {
    const int v = sl * kWave + lane;
    if (v >= nrow) continue;
    const int s = vs0 + sl;
    const int i = row0 + v;
    const double4 p = mom[i];
    ke += 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z + 0.5 * p.w * p.w;
    const double4 qi = pos[i];
    const int len = bl.slen[s];
    const uint4 *rg = reinterpret_cast<const uint4 *>(bl.nbr + bl.sbase[s]) + lane;  // group g at rg[g * kWave]
    double spe = 0.0, svir = 0.0;
    uint4 ahead = rg[0];  // one group past the slice stays inside the list arrays (kListTail)
    for (int k = 0; k < len; k += kGroup) {
      int e[kGroup];
      unpack_group(ahead, e);
      ahead = rg[(k / kGroup + 1) * kWave];
      __builtin_amdgcn_sched_barrier(0);  // the prefetch is issued before this group's work
      // two halves of four partners, each loaded then summed in list order:
      // all eight loaded at once held the kernel above 128 registers, one
      // workgroup per CU instead of two
#pragma unroll
      for (int h = 0; h < kGroup; h += kGroup / 2) {
        double qx[kGroup / 2], qy[kGroup / 2], qz[kGroup / 2];
#pragma unroll
        for (int u = 0; u < kGroup / 2; u++) {
          const int eu = e[h + u];
          if (staged) {  // padding reads the staged far sentinel
            qx[u] = lds64[3 * eu];
            qy[u] = lds64[3 * eu + 1];
            qz[u] = lds64[3 * eu + 2];
          } else if (eu < total) {
            const double4 q = pos[gl[eu]];
            qx[u] = q.x;
            qy[u] = q.y;
            qz[u] = q.z;
          } else {
            qx[u] = qy[u] = qz[u] = 1.0e200;
          }
        }
#pragma unroll
        for (int u = 0; u < kGroup / 2; u++) {
          const double dx = qi.x - qx[u], dy = qi.y - qy[u], dz = qi.z - qz[u];
          const double r2 = (dx * dx + dy * dy + dz * dz);
          const double ir2 = rcp64_exact(r2);  // padding: r2 = inf, NaN terms, dropped by the select
          const double ir6 = ir2 * ir2 * ir2;
          const double c2r2 = C2 * r2;
          const bool in = !(r2 > CL2);
          const double tpe = fma(ir6, ir6, -ir6) + (c2r2 + C0);
          const double tvir = fma(ir6, fma(-6.0, ir6, 3.0), c2r2);
          spe += in ? tpe : 0.0;
          svir += in ? tvir : 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);  // keeps the halves apart
      }
      hold_prefetch(ahead, spe, svir, ke);
    }
    pe += 2.0 * spe;
    vir += 4.0 * svir;
  }
  ke = wave_sum(ke);
  pe = wave_sum(pe);
  vir = wave_sum(vir);
  if (lane == 0) {
    red[0][wv] = ke;
    red[1][wv] = pe;
    red[2][wv] = vir;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double acc = 0.0;
    for (int k = 0; k < W; k++) acc += red[threadIdx.x][k];
    partial[(long long)b * 3 + threadIdx.x] = acc;
  }
}

}  // namespace mdgpu
