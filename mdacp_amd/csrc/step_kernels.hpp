// step_kernels.hpp -- per-step kernels on gfx950: half drift, halo refresh,
// the smoothed-LJ pair-force kick (FP64 and mixed FP32), Langevin noise,
// Nose-Hoover scaling, observers and the Verlet-list expiry test.
//
// Reference semantics (kaityo256/mdacp):
//   UpdatePositionHalf   fcalculator.cc:21-32   q += p*(dt*0.5)
//   SendBorderParticles  mdmanager.cc:282-348, mdunit.cc:214-269
//   CalculateForceNext   fcalculator.cc:228-290 (kick p += F*dt, F from the
//                        potential 4(r^-12 - r^-6 + C2 r^2 + C0), variables.cc:130-142)
//   Langevin             fcalculator.cc:764-778
//   HeatbathMomenta      fcalculator.cc:751-762
//   observers            observer.cc:4-79
//   expiry               pairlist.cc:58-154
//
// The force kernels read a full neighbour list (each owned row lists every
// owned or ghost partner within SL), so no momentum scatter and no atomics:
// results are bit-reproducible run to run.  Where the reference applies
// Newton's third law inside a half list, this computes both sides.
#pragma once
#include "device_common.hpp"

namespace mdgpu {

struct ForceConsts {
  double CL2;      // rc^2
  double A, B, C;  // 24 dt, 48 dt, 8 C2 dt: df = ir2*ir6*(A - B*ir6) + C
  double dt2;      // dt*0.5
};

struct ForceConstsF {
  float CL2, A, B, C;
  float w[3];       // cell widths (cell-relative coordinates)
};

struct LangevinConsts {
  double gamma, hb_D, dt;
  uint64_t seed, step;
};

// ---------------------------------------------------------------- drift
__global__ void k_drift(double4 *__restrict__ pos, const double4 *__restrict__ mom, int pn, double dt2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pn) return;
  double4 q = pos[i];
  const double4 p = mom[i];
  q.x += p.x * dt2;
  q.y += p.y * dt2;
  q.z += p.z * dt2;
  pos[i] = q;
}

__global__ void k_refresh(double4 *__restrict__ pos, const int *__restrict__ gsrc,
                          const int *__restrict__ gcode, int pn, int ng, double Lx, double Ly, double Lz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ng) return;
  const double3 y = image_of(pos[gsrc[k]], gcode[k], Lx, Ly, Lz);
  pos[pn + k] = make_double4(y.x, y.y, y.z, 0.0);
}

// Cell-relative FP32 coordinates: q - origin(cell), origin = s + (ix-1)*w.
__device__ __forceinline__ float4 to_rel(double4 q, int c, const double *org_x, const double *org_y,
                                         const double *org_z, int mx, int my) {
  const int ix = c % mx, iy = (c / mx) % my, iz = c / (mx * my);
  return make_float4((float)(q.x - org_x[ix]), (float)(q.y - org_y[iy]), (float)(q.z - org_z[iz]), 0.0f);
}

struct Origins {
  const double *x, *y, *z;
  int mx, my;
};

__global__ void k_to_rel(const double4 *__restrict__ pos, const int *__restrict__ cell, int tn, Origins O,
                         float4 *__restrict__ rel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tn) return;
  rel[i] = to_rel(pos[i], cell[i], O.x, O.y, O.z, O.mx, O.my);
}

// Fused first half drift + halo refresh into a second buffer:
//   owned:  tmp[i] = q_i + p_i*dt2
//   ghost:  tmp[g] = pos[g] = (q_src + p_src*dt2) + shift*L
// (the same two roundings as UpdatePositionHalf then SendBorderParticles).
// `pos` keeps the owned positions until the force kernel writes them.
template <bool kRel>
__global__ void k_drift_refresh(double4 *__restrict__ pos, const double4 *__restrict__ mom,
                                double4 *__restrict__ tmp, const int *__restrict__ gsrc,
                                const int *__restrict__ gcode, int pn, int ng, double dt2, double Lx,
                                double Ly, double Lz, const int *__restrict__ cell, Origins O,
                                float4 *__restrict__ rel, unsigned long long *__restrict__ vmax_bits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *vmax_bits = 0ull;
  if (i < pn) {
    double4 q = pos[i];
    const double4 p = mom[i];
    q.x += p.x * dt2;
    q.y += p.y * dt2;
    q.z += p.z * dt2;
    tmp[i] = q;
    if (kRel) rel[i] = to_rel(q, cell[i], O.x, O.y, O.z, O.mx, O.my);
  } else if (i < pn + ng) {
    const int k = i - pn;
    const int src = gsrc[k];
    double4 q = pos[src];
    const double4 p = mom[src];
    q.x += p.x * dt2;
    q.y += p.y * dt2;
    q.z += p.z * dt2;
    const double3 y = image_of(q, gcode[k], Lx, Ly, Lz);
    const double4 g = make_double4(y.x, y.y, y.z, 0.0);
    tmp[i] = g;
    pos[i] = g;
    if (kRel) rel[i] = to_rel(g, cell[i], O.x, O.y, O.z, O.mx, O.my);
  }
}

// --------------------------------------------------------------- force
enum Epilogue {
  kKickOnly = 0,      // p += kick                       (CalculateForce)
  kKickDrift = 1,     // p += kick; q_out = q + p*dt2    (CalculateForce + UpdatePositionHalf)
  kKickLangevin = 2,  // p += kick; Langevin update      (CalculateLangevin)
};

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint64_t m0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t m1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(m1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = (uint32_t)m1;
    const uint32_t n2 = (uint32_t)(m0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = (uint32_t)m0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Three standard normals per (seed, particle id, step): one Philox4x32-10
// block (Salmon et al., SC'11), four uniforms in (0,1), two Box-Muller pairs.
__device__ __forceinline__ void philox_normal3(uint64_t seed, uint32_t id, uint64_t step, double n[3]) {
  uint32_t c[4] = {id, (uint32_t)step, (uint32_t)(step >> 32), 0x4C414E47u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double k = 2.3283064365386963e-10;
  const double u0 = ((double)c[0] + 0.5) * k, u1 = ((double)c[1] + 0.5) * k;
  const double u2 = ((double)c[2] + 0.5) * k, u3 = ((double)c[3] + 0.5) * k;
  const double twopi = 6.283185307179586;
  const double a = sqrt(-2.0 * log(u0)), b = sqrt(-2.0 * log(u2));
  n[0] = a * cos(twopi * u1);
  n[1] = a * sin(twopi * u1);
  n[2] = b * cos(twopi * u3);
}

__device__ __forceinline__ void langevin_update(double4 &p, const LangevinConsts &lc, uint32_t id) {
  double n[3];
  philox_normal3(lc.seed, id, lc.step, n);
  const double g = lc.gamma;
  p.x += (-g * p.x + (n[0] * lc.hb_D + 0.0)) * lc.dt;
  p.y += (-g * p.y + (n[1] * lc.hb_D + 0.0)) * lc.dt;
  p.z += (-g * p.z + (n[2] * lc.hb_D + 0.0)) * lc.dt;
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
                                                const int *__restrict__ id,
                                                unsigned long long *__restrict__ vmax_bits) {
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
  if (vmax_bits) {
    v2 = wave_max(v2);
    if ((threadIdx.x % kWave) == 0) atomic_max_nonneg(vmax_bits, v2);
  }
}

// FP64 kick.  One lane per owned row, rows in cell order, one wavefront per
// 64-row slice of the neighbour list (coalesced 256-byte loads per entry).
__global__ __launch_bounds__(kBlock) void k_force_fp64(
    const double4 *__restrict__ pos, double4 *__restrict__ mom, double4 *__restrict__ pos_out,
    const uint32_t *__restrict__ nbr, const long long *__restrict__ sbase, const int *__restrict__ slen,
    int pn, ForceConsts fc, int epilogue, LangevinConsts lc, const int *__restrict__ id,
    unsigned long long *__restrict__ vmax_bits) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int s = i / kWave, lane = i % kWave;
  const int ok = i < pn;
  const int ii = ok ? i : pn - 1;
  const double4 qi = pos[ii];
  double ax = 0.0, ay = 0.0, az = 0.0;
  const int len = ok || (s * kWave < pn) ? slen[s] : 0;
  const uint32_t *row = nbr + sbase[s < (pn + kWave - 1) / kWave ? s : 0] + lane;
  for (int k = 0; k < len; k++) {
    const uint32_t e = row[(long long)k * kWave];
    const double4 qj = pos[e & kIndexMask];
    const double dx = qj.x - qi.x, dy = qj.y - qi.y, dz = qj.z - qi.z;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    if (r2 <= fc.CL2) {
      const double ir2 = rcp64(r2);
      const double ir6 = ir2 * ir2 * ir2;
      const double df = fma(ir2 * ir6, fma(-fc.B, ir6, fc.A), fc.C);
      ax = fma(df, dx, ax);
      ay = fma(df, dy, ay);
      az = fma(df, dz, az);
    }
  }
  double4 p = mom[ii];
  p.x += ax;
  p.y += ay;
  p.z += az;
  finish_particle(i, ok, p, qi, mom, pos_out, epilogue, fc.dt2, lc, id, vmax_bits);
}

// Mixed precision kick: FP32 cell-relative coordinates, FP32 pair forces,
// compensated (Kahan) FP32 sums; the sum is added to the FP64 momentum and
// the drift is FP64.  The neighbour entry carries the cell offset of the
// partner, so dx = (rel_j - rel_i) + off*w stays small.
template <bool kKahan>
__global__ __launch_bounds__(kBlock) void k_force_mixed(
    const float4 *__restrict__ rel, const double4 *__restrict__ pos, double4 *__restrict__ mom,
    double4 *__restrict__ pos_out, const uint32_t *__restrict__ nbr, const long long *__restrict__ sbase,
    const int *__restrict__ slen, int pn, ForceConstsF fc, double dt2, int epilogue, LangevinConsts lc,
    const int *__restrict__ id, unsigned long long *__restrict__ vmax_bits) {
  __shared__ float4 offs[64];
  if (threadIdx.x < 64) {
    const int t = threadIdx.x;
    const int ox = (t & 3) - 1, oy = ((t >> 2) & 3) - 1, oz = ((t >> 4) & 3) - 1;
    offs[t] = make_float4(ox * fc.w[0], oy * fc.w[1], oz * fc.w[2], 0.0f);
  }
  __syncthreads();
  const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int s = i / kWave, lane = i % kWave;
  const int ok = i < pn;
  const int ii = ok ? i : pn - 1;
  const float4 ri = rel[ii];
  float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
  const int len = ok || (s * kWave < pn) ? slen[s] : 0;
  const uint32_t *row = nbr + sbase[s < (pn + kWave - 1) / kWave ? s : 0] + lane;
  for (int k = 0; k < len; k++) {
    const uint32_t e = row[(long long)k * kWave];
    const float4 rj = rel[e & kIndexMask];
    const float4 o = offs[e >> kIndexBits];
    const float dx = (rj.x - ri.x) + o.x, dy = (rj.y - ri.y) + o.y, dz = (rj.z - ri.z) + o.z;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    if (r2 <= fc.CL2) {
      const float ir2 = __builtin_amdgcn_rcpf(r2);
      const float ir6 = ir2 * ir2 * ir2;
      const float df = fmaf(ir2 * ir6, fmaf(-fc.B, ir6, fc.A), fc.C);
      if (kKahan) {
        float y, t;
        y = fmaf(df, dx, -cx); t = ax + y; cx = (t - ax) - y; ax = t;
        y = fmaf(df, dy, -cy); t = ay + y; cy = (t - ay) - y; ay = t;
        y = fmaf(df, dz, -cz); t = az + y; cz = (t - az) - y; az = t;
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
  finish_particle(i, ok, p, qi, mom, pos_out, epilogue, dt2, lc, id, vmax_bits);
}

// ----------------------------------------------------------- thermostats
__global__ void k_langevin(double4 *__restrict__ mom, const int *__restrict__ id, int pn, LangevinConsts lc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pn) return;
  double4 p = mom[i];
  langevin_update(p, lc, (uint32_t)id[i]);
  mom[i] = p;
}

__global__ void k_scale_momenta(double4 *__restrict__ mom, int pn, double f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pn) return;
  double4 p = mom[i];
  p.x *= f; p.y *= f; p.z *= f; p.w *= f;
  mom[i] = p;
}

// ------------------------------------------------------------- observers
// Per-block partial sums, reduced in fixed order by k_reduce_partials, so the
// results are deterministic.  Each list entry carries weight 1/2: owned-owned
// pairs appear in two rows, owned-ghost pairs in one row and are halved by
// the reference (observer.cc:39-41,70-72).
__global__ __launch_bounds__(kBlock) void k_observe(const double4 *__restrict__ pos,
                                                    const double4 *__restrict__ mom,
                                                    const uint32_t *__restrict__ nbr,
                                                    const long long *__restrict__ sbase,
                                                    const int *__restrict__ slen, int pn, double CL2,
                                                    double C2, double C0, double *__restrict__ partial) {
  __shared__ double red[3][kBlock / kWave];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = i / kWave, lane = i % kWave;
  double ke = 0.0, pe = 0.0, vir = 0.0;
  if (i < pn) {
    const double4 p = mom[i];
    ke = 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z + 0.5 * p.w * p.w;
    const double4 qi = pos[i];
    const int len = slen[s];
    const uint32_t *row = nbr + sbase[s] + lane;
    for (int k = 0; k < len; k++) {
      const double4 qj = pos[row[(long long)k * kWave] & kIndexMask];
      const double dx = qi.x - qj.x, dy = qi.y - qj.y, dz = qi.z - qj.z;
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
  if ((threadIdx.x % kWave) == 0) {
    red[0][w] = ke;
    red[1][w] = pe;
    red[2][w] = vir;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int k = 0; k < kBlock / kWave; k++) t += red[threadIdx.x][k];
    partial[(long long)blockIdx.x * 3 + threadIdx.x] = t;
  }
}

// One workgroup reduces nblk triples of partial sums in a fixed order.
__global__ void k_reduce_partials(const double *__restrict__ partial, int nblk, int ncomp,
                                  double *__restrict__ out) {
  __shared__ double red[kBlock];
  for (int c = 0; c < ncomp; c++) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) t += partial[(long long)b * ncomp + c];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
    __syncthreads();
  }
}

__global__ void k_kinetic(const double4 *__restrict__ mom, int pn, double *__restrict__ partial) {
  __shared__ double red[kBlock / kWave];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ke = 0.0;
  if (i < pn) {
    const double4 p = mom[i];
    ke = 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z + 0.5 * p.w * p.w;
  }
  ke = wave_sum(ke);
  if ((threadIdx.x % kWave) == 0) red[threadIdx.x / kWave] = ke;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kBlock / kWave; k++) t += red[k];
    partial[blockIdx.x] = t;
  }
}

// ---------------------------------------------------------------- expiry
__global__ void k_max_p2(const double4 *__restrict__ mom, int pn, unsigned long long *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < pn) {
    const double4 p = mom[i];
    v = p.x * p.x + p.y * p.y + p.z * p.z;
  }
  v = wave_max(v);
  if ((threadIdx.x % kWave) == 0) atomic_max_nonneg(out, v);
}

// PairList::CheckByDisplacement, first pass: squared displacement since the
// build for owned and ghost entries, and its maximum.
__global__ void k_displacement(const double4 *__restrict__ pos, const double4 *__restrict__ old, int tn,
                               double *__restrict__ disp, unsigned long long *__restrict__ dmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < tn) {
    const double4 a = old[i], b = pos[i];
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    v = dx * dx + dy * dy + dz * dz;
    disp[i] = v;
  }
  v = wave_max(v);
  if ((threadIdx.x % kWave) == 0) atomic_max_nonneg(dmax, v);
}

// second pass: particles displaced by more than (margin - max displacement);
// up to `cap` of them are collected (the reference stops at CheckListLength).
__global__ void k_candidates(const double *__restrict__ disp, int tn, double c_d2, int *__restrict__ count,
                             int *__restrict__ list, int cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tn) return;
  if (disp[i] > c_d2) {
    const int k = atomicAdd(count, 1);
    if (k < cap) list[k] = i;
  }
}

// third pass: a candidate pair that was beyond SL at the build and is now
// inside rc expires the list.
__global__ void k_candidate_pairs(const double4 *__restrict__ pos, const double4 *__restrict__ old,
                                  const int *__restrict__ list, int l, double SL2, double CL2,
                                  int *__restrict__ flag) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)l * l;
  if (t >= total) return;
  const int a = (int)(t / l), b = (int)(t % l);
  if (b <= a) return;
  const int i1 = list[a], i2 = list[b];
  const double4 o1 = old[i1], o2 = old[i2];
  double dx = o1.x - o2.x, dy = o1.y - o2.y, dz = o1.z - o2.z;
  double dr2 = dx * dx + dy * dy + dz * dz;
  if (dr2 < SL2) return;
  const double4 q1 = pos[i1], q2 = pos[i2];
  dx = q1.x - q2.x;
  dy = q1.y - q2.y;
  dz = q1.z - q2.z;
  dr2 = dx * dx + dy * dy + dz * dz;
  if (dr2 < CL2) atomicOr(flag, 1);
}

__global__ void k_copy_pos(const double4 *__restrict__ a, double4 *__restrict__ b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_scale_pos(double4 *__restrict__ pos, int pn, double alpha) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pn) return;
  double4 q = pos[i];
  q.x *= alpha;
  q.y *= alpha;
  q.z *= alpha;
  pos[i] = q;
}

__global__ void k_fill_int(int *__restrict__ a, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

}  // namespace mdgpu
