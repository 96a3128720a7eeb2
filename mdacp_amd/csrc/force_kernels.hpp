// force_kernels.hpp -- the pair loops on gfx950: the smoothed-LJ force kick
// (FP64 and mixed FP32) and the pair observers, over LDS-staged cell tiles.
//
// Reference semantics (kaityo256/mdacp):
//   CalculateForceNext  fcalculator.cc:228-290: for each listed pair within rc,
//                       p_i += df*(q_j - q_i), df = ((24 r^6 - 48)/(r^6 r^6 r^2) + 8 C2) dt
//   potential           4(r^-12 - r^-6 + C2 r^2 + C0), C0/C2 from variables.cc:130-142
//   observers           observer.cc:17-79 (PE and virial, ghost pairs halved)
//
// A workgroup takes one tile of kWave*W consecutive owned rows (cell order).
// It first copies the 18 contiguous particle ranges its rows' partners can
// lie in (9 owned + 9 ghost, see k_tile_ranges) into LDS with coalesced
// loads, then each lane walks its row of the sliced-ELL list (16-bit indices
// into the staged array) and gathers partners from LDS instead of through the
// texture path.  Tiles whose staged set does not fit the LDS budget read the
// same partners from global memory instead (same arithmetic, slower).
//
// Full neighbour list: each owned row lists every partner within SL, so no
// momentum scatter and no atomics: results are bit-reproducible run to run.
#pragma once
#include "build_kernels.hpp"
#include "device_common.hpp"
#include "step_kernels.hpp"

namespace mdgpu {

enum Epilogue {
  kKickOnly = 0,      // p += kick                       (CalculateForce)
  kKickDrift = 1,     // p += kick; q_out = q + p*dt2    (CalculateForce + UpdatePositionHalf)
  kKickLangevin = 2,  // p += kick; Langevin update      (CalculateLangevin)
};

struct TileList {
  const uint16_t *nbr;      // staged index per (slice, slot, lane)
  const uint8_t *code;      // cell offset of the partner: (dx+1) | (dy+1)<<2 | (dz+1)<<4
  const long long *sbase;   // first slot of each 64-row slice
  const int *slen;          // slots per slice (multiple of kSliceRound)
  const int *tstart;        // [tile][18] first particle of each staged range
  const int *tbase;         // [tile][19] offset of each range in the staged array
  int nslice;
  int lds_cap;              // staged entries that fit the LDS budget
};

// staged index -> particle index, for tiles read from global memory
__device__ __forceinline__ int tile_global(int e, const int *__restrict__ ts, const int *__restrict__ tb, int total,
                                           int sentinel) {
  if (e >= total) return sentinel;
  int k = 0;
#pragma unroll
  for (int r = 1; r < kTileRanges; r++) k += (e >= tb[r]) ? 1 : 0;
  return ts[k] + (e - tb[k]);
}

// 1/x to double precision: hardware reciprocal estimate + two Newton steps.
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

__device__ __forceinline__ void finish_particle(int i, int lane_ok, double4 p, const double4 qi,
                                                double4 *__restrict__ mom, double4 *__restrict__ pos_out,
                                                int epilogue, double dt2, const LangevinConsts &lc,
                                                const int *__restrict__ id, double *__restrict__ vpart) {
  double v2 = 0.0;
  if (lane_ok) {
    if (epilogue == kKickLangevin) langevin_update(p, lc, (uint32_t)id[i]);
    mom[i] = p;
    if (epilogue == kKickDrift) {
      double4 q = qi;
      q.x += p.x * dt2;
      q.y += p.y * dt2;
      q.z += p.z * dt2;
      pos_out[i] = q;
    } else if (epilogue == kKickLangevin) {
      pos_out[i] = qi;  // one half drift per Langevin step (mdmanager.cc:185-192)
    }
    v2 = p.x * p.x + p.y * p.y + p.z * p.z;
  }
  if (vpart) block_max_to_partial(v2, vpart);
}

// Stage the tile's partner ranges as SoA doubles (24 bytes per particle).
template <int T>
__device__ __forceinline__ void stage_fp64(const double4 *__restrict__ src, const int *__restrict__ ts,
                                           const int *__restrict__ tb, int total, double *lx, double *ly,
                                           double *lz) {
  for (int k = 0; k < kTileRanges; k++) {
    const int b = tb[k], n = tb[k + 1] - b, st = ts[k];
    for (int u = threadIdx.x; u < n; u += T) {
      const double4 q = src[st + u];
      lx[b + u] = q.x;
      ly[b + u] = q.y;
      lz[b + u] = q.z;
    }
  }
  if (threadIdx.x == 0) {
    lx[total] = 1.0e200;
    ly[total] = 1.0e200;
    lz[total] = 1.0e200;
  }
}

// FP64 kick.  Entries are consumed in batches of U: the next batch's indices
// are loaded before the current batch is computed.  The cutoff is a select,
// not a branch (padding entries point at a far sentinel: r2 = inf).
template <int W, int U>
__global__ __launch_bounds__(kWave *W) void k_force_tile_fp64(
    const double4 *__restrict__ src, double4 *__restrict__ mom, double4 *__restrict__ pos_out, TileList tl,
    int pn, int tn, ForceConsts fc, int epilogue, LangevinConsts lc, const int *__restrict__ id,
    double *__restrict__ vpart) {
  extern __shared__ double lds64[];
  constexpr int T = kWave * W;
  const int t = xcd_block(blockIdx.x, gridDim.x);
  const int i = t * T + threadIdx.x;
  const int s = i / kWave, lane = threadIdx.x % kWave;
  const int ok = i < pn;
  const int ii = ok ? i : pn - 1;
  const int *ts = tl.tstart + (long long)t * kTileRanges;
  const int *tb = tl.tbase + (long long)t * (kTileRanges + 1);
  const int total = tb[kTileRanges];
  const bool staged = total < tl.lds_cap;  // uniform across the workgroup
  double *lx = lds64, *ly = lds64 + tl.lds_cap, *lz = lds64 + 2 * tl.lds_cap;
  if (staged) stage_fp64<T>(src, ts, tb, total, lx, ly, lz);
  __syncthreads();
  const double4 qi = src[ii];
  double ax = 0.0, ay = 0.0, az = 0.0;
  const bool live = s < tl.nslice;  // uniform across the wavefront
  const int len = live ? tl.slen[s] : 0;
  const uint16_t *row = tl.nbr + (live ? tl.sbase[s] : 0) + lane;
  int e[U];
  if (len > 0) {
#pragma unroll
    for (int u = 0; u < U; u++) e[u] = row[(long long)u * kWave];
  }
  for (int k = 0; k < len; k += U) {
    double qx[U], qy[U], qz[U];
    if (staged) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        qx[u] = lx[e[u]];
        qy[u] = ly[e[u]];
        qz[u] = lz[e[u]];
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const double4 q = src[tile_global(e[u], ts, tb, total, tn)];
        qx[u] = q.x;
        qy[u] = q.y;
        qz[u] = q.z;
      }
    }
    if (k + U < len) {
#pragma unroll
      for (int u = 0; u < U; u++) e[u] = row[(long long)(k + U + u) * kWave];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
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
  double4 p = mom[ii];
  p.x += ax;
  p.y += ay;
  p.z += az;
  finish_particle(i, ok, p, qi, mom, pos_out, epilogue, fc.dt2, lc, id, vpart);
}

// Mixed precision kick: FP32 cell-relative coordinates (staged as float4),
// FP32 pair forces, compensated (Kahan) FP32 sums; the sum is added to the
// FP64 momentum and the drift is FP64.  The partner's cell offset comes from
// the code array, so dx = (rel_j - rel_i) + off*w stays small.
template <int W, int U, bool kKahan>
__global__ __launch_bounds__(kWave *W) void k_force_tile_mixed(
    const float4 *__restrict__ rel, const double4 *__restrict__ pos, double4 *__restrict__ mom,
    double4 *__restrict__ pos_out, TileList tl, int pn, int tn, ForceConstsF fc, double dt2, int epilogue,
    LangevinConsts lc, const int *__restrict__ id, double *__restrict__ vpart) {
  extern __shared__ float4 lds32[];
  __shared__ float4 offs[64];
  constexpr int T = kWave * W;
  if (threadIdx.x < 64) {
    const int c = threadIdx.x;
    const int ox = (c & 3) - 1, oy = ((c >> 2) & 3) - 1, oz = ((c >> 4) & 3) - 1;
    offs[c] = make_float4(ox * fc.w[0], oy * fc.w[1], oz * fc.w[2], 0.0f);
  }
  const int t = xcd_block(blockIdx.x, gridDim.x);
  const int i = t * T + threadIdx.x;
  const int s = i / kWave, lane = threadIdx.x % kWave;
  const int ok = i < pn;
  const int ii = ok ? i : pn - 1;
  const int *ts = tl.tstart + (long long)t * kTileRanges;
  const int *tb = tl.tbase + (long long)t * (kTileRanges + 1);
  const int total = tb[kTileRanges];
  const bool staged = total < tl.lds_cap;
  if (staged) {
    for (int k = 0; k < kTileRanges; k++) {
      const int b = tb[k], n = tb[k + 1] - b, st = ts[k];
      for (int u = threadIdx.x; u < n; u += T) lds32[b + u] = rel[st + u];
    }
    if (threadIdx.x == 0) lds32[total] = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
  }
  __syncthreads();
  const float4 ri = rel[ii];
  float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
  const bool live = s < tl.nslice;
  const int len = live ? tl.slen[s] : 0;
  const long long base = (live ? tl.sbase[s] : 0) + lane;
  const uint16_t *row = tl.nbr + base;
  const uint8_t *crow = tl.code + base;
  int e[U], c[U];
  if (len > 0) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      e[u] = row[(long long)u * kWave];
      c[u] = crow[(long long)u * kWave];
    }
  }
  for (int k = 0; k < len; k += U) {
    float4 rj[U];
    int cc[U];
    if (staged) {
#pragma unroll
      for (int u = 0; u < U; u++) rj[u] = lds32[e[u]];
    } else {
#pragma unroll
      for (int u = 0; u < U; u++) rj[u] = rel[tile_global(e[u], ts, tb, total, tn)];
    }
#pragma unroll
    for (int u = 0; u < U; u++) cc[u] = c[u];
    if (k + U < len) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        e[u] = row[(long long)(k + U + u) * kWave];
        c[u] = crow[(long long)(k + U + u) * kWave];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const float4 o = offs[cc[u]];
      const float dx = (rj[u].x - ri.x) + o.x, dy = (rj[u].y - ri.y) + o.y, dz = (rj[u].z - ri.z) + o.z;
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      const float ir2 = __builtin_amdgcn_rcpf(r2);
      const float ir6 = ir2 * ir2 * ir2;
      float df = fmaf(ir2 * ir6, fmaf(-fc.B, ir6, fc.A), fc.C);
      df = (r2 <= fc.CL2) ? df : 0.0f;
      if (kKahan) {
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
  }
  double4 p = mom[ii];
  p.x += (double)ax - (double)cx;
  p.y += (double)ay - (double)cy;
  p.z += (double)az - (double)cz;
  const double4 qi = pos[ii];
  finish_particle(i, ok, p, qi, mom, pos_out, epilogue, dt2, lc, id, vpart);
}

// Observers over the list: kinetic energy, PE and virial sums per workgroup
// (partial[block*3 + {0,1,2}]), reduced in fixed order by k_reduce_partials,
// so the results are deterministic.  Each list entry carries weight 1/2:
// owned-owned pairs appear in two rows, owned-ghost pairs in one row and are
// halved by the reference (observer.cc:39-41,70-72).  Ghost positions are the
// ones of the last halo refresh, as in the reference.
template <int W>
__global__ __launch_bounds__(kWave *W) void k_observe_tile(const double4 *__restrict__ pos,
                                                         const double4 *__restrict__ mom, TileList tl, int pn,
                                                         int tn, double CL2, double C2, double C0,
                                                         double *__restrict__ partial) {
  extern __shared__ double lds64[];
  __shared__ double red[3][W];
  constexpr int T = kWave * W;
  const int t = xcd_block(blockIdx.x, gridDim.x);
  const int i = t * T + threadIdx.x;
  const int s = i / kWave, lane = threadIdx.x % kWave;
  const int *ts = tl.tstart + (long long)t * kTileRanges;
  const int *tb = tl.tbase + (long long)t * (kTileRanges + 1);
  const int total = tb[kTileRanges];
  const bool staged = total < tl.lds_cap;
  double *lx = lds64, *ly = lds64 + tl.lds_cap, *lz = lds64 + 2 * tl.lds_cap;
  if (staged) stage_fp64<T>(pos, ts, tb, total, lx, ly, lz);
  __syncthreads();
  double ke = 0.0, pe = 0.0, vir = 0.0;
  if (i < pn) {
    const double4 p = mom[i];
    ke = 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z + 0.5 * p.w * p.w;
    const double4 qi = pos[i];
    const int len = tl.slen[s];
    const uint16_t *row = tl.nbr + tl.sbase[s] + lane;
    for (int k = 0; k < len; k++) {
      const int e = row[(long long)k * kWave];
      double xj, yj, zj;
      if (staged) {
        xj = lx[e];
        yj = ly[e];
        zj = lz[e];
      } else {
        const double4 q = pos[tile_global(e, ts, tb, total, tn)];
        xj = q.x;
        yj = q.y;
        zj = q.z;
      }
      const double dx = qi.x - xj, dy = qi.y - yj, dz = qi.z - zj;
      const double r2 = (dx * dx + dy * dy + dz * dz);
      if (r2 > CL2) continue;
      const double r6 = r2 * r2 * r2;
      pe += 0.5 * (4.0 * (1.0 / (r2 * r2 * r2 * r2 * r2 * r2) - 1.0 / r6 + C2 * r2 + C0));
      vir += 0.5 * (((24.0 * r6 - 48.0) / (r6 * r6 * r2) + C2 * 8.0) * r2);
    }
  }
  ke = wave_sum(ke);
  pe = wave_sum(pe);
  vir = wave_sum(vir);
  const int w = threadIdx.x / kWave;
  if (lane == 0) {
    red[0][w] = ke;
    red[1][w] = pe;
    red[2][w] = vir;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double acc = 0.0;
    for (int k = 0; k < W; k++) acc += red[threadIdx.x][k];
    partial[(long long)blockIdx.x * 3 + threadIdx.x] = acc;
  }
}

}  // namespace mdgpu
