// step_kernels.hpp -- per-step kernels on gfx950 other than the pair loops:
// half drift, halo refresh, cell-relative FP32 coordinates, Langevin noise,
// Nose-Hoover scaling, kinetic energy and the Verlet-list expiry test.
//
// Reference semantics (kaityo256/mdacp):
//   UpdatePositionHalf   fcalculator.cc:22-32   q += p*(dt*0.5)
//   SendBorderParticles  mdmanager.cc:283-348, mdunit.cc:137-191
//   Langevin             fcalculator.cc:765-778
//   HeatbathMomenta      fcalculator.cc:752-762
//   KineticEnergyObserver observer.cc:5-15
//   expiry               pairlist.cc:26-121
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

// SendBorderParticles: ghosts [pn, pn + ng) from the current owned positions.
__global__ void k_refresh(double4 *__restrict__ pos, const int *__restrict__ gsrc, const int *__restrict__ gcode,
                          int pn, int ng, double Lx, double Ly, double Lz) {
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

// Fused first half drift + halo refresh:
//   owned:  tmp[i] = q_i + p_i*dt2                    (and its cell-relative FP32 copy)
//   ghost:  pos[g] = tmp[g] = (q_src + p_src*dt2) + shift*L   (and its FP32 copy)
// (the same two roundings as UpdatePositionHalf then SendBorderParticles).
// `tmp` then holds every position the force kernel reads; `pos` keeps the
// owned positions until the force kernel writes the drifted ones, and the
// ghosts until the next refresh.
template <bool kRel>
__global__ void k_drift_refresh(double4 *__restrict__ pos, const double4 *__restrict__ mom,
                                double4 *__restrict__ tmp, const int *__restrict__ gsrc,
                                const int *__restrict__ gcode, int pn, int ng, double dt2, double Lx,
                                double Ly, double Lz, const int *__restrict__ cell, Origins O,
                                float4 *__restrict__ rel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
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
    pos[i] = g;
    tmp[i] = g;
    if (kRel) rel[i] = to_rel(g, cell[i], O.x, O.y, O.z, O.mx, O.my);
  }
}

// Halo refresh when `tmp` already holds the half-drifted owned positions
// (written by the previous fused step's epilogue): ghosts only,
// pos[g] = tmp[g] = tmp[src] + shift*L (and the FP32 copy) -- the values
// k_drift_refresh would write.
template <bool kRel>
__global__ void k_refresh_from_drifted(double4 *__restrict__ pos, double4 *__restrict__ tmp,
                                       const int *__restrict__ gsrc, const int *__restrict__ gcode, int pn,
                                       int ng, double Lx, double Ly, double Lz, const int *__restrict__ cell,
                                       Origins O, float4 *__restrict__ rel) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ng) return;
  const int i = pn + k;
  const double3 y = image_of(tmp[gsrc[k]], gcode[k], Lx, Ly, Lz);
  const double4 g = make_double4(y.x, y.y, y.z, 0.0);
  pos[i] = g;
  tmp[i] = g;
  if (kRel) rel[i] = to_rel(g, cell[i], O.x, O.y, O.z, O.mx, O.my);
}

// ----------------------------------------------------------- Langevin RNG
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

// ------------------------------------------------------------- reductions
// One workgroup reduces nblk tuples of partial sums in a fixed order.
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

// KineticEnergyObserver: sum of p^2/2 over the four lanes (padding lane 0)
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
__global__ void k_max_p2(const double4 *__restrict__ mom, int pn, double *__restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < pn) {
    const double4 p = mom[i];
    v = p.x * p.x + p.y * p.y + p.z * p.z;
  }
  block_max_to_partial(v, partial);
}

// PairList::CheckByDisplacement, first pass: squared displacement since the
// build for owned and ghost entries, and its maximum.
__global__ void k_displacement(const double4 *__restrict__ pos, const double4 *__restrict__ old, int tn,
                               double *__restrict__ disp, double *__restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < tn) {
    const double4 a = old[i], b = pos[i];
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    v = dx * dx + dy * dy + dz * dz;
    disp[i] = v;
  }
  block_max_to_partial(v, partial);
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

}  // namespace mdgpu
