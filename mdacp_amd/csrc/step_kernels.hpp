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
};

struct LangevinConsts {
  double gamma, hb_D, dt;
  uint64_t seed, step;
};

// The Langevin constants of a replica at a step.
__device__ __forceinline__ LangevinConsts langevin_consts(const Replica &R, double dt, uint64_t step) {
  LangevinConsts lc;
  lc.gamma = R.gamma;
  lc.hb_D = R.hb_D;
  lc.dt = dt;
  lc.seed = R.seed;
  lc.step = step;
  return lc;
}

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
                          int pn, int ng, const Replica *__restrict__ reps) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ng) return;
  const int g = gcode[k];
  const double3 y = image_of(pos[gsrc[k]], ghost_shift(g), reps[ghost_rep(g)].L);
  pos[pn + k] = make_double4(y.x, y.y, y.z, 0.0);
}

// Cell-relative FP32 coordinates: q - origin(cell), origin = s + (ix-1)*w per
// axis, read from the table of column origins (org[R.col0 ...]: the replica's
// x columns, then its y, then its z columns).
__device__ __forceinline__ float4 to_rel(double4 q, int c, const double *__restrict__ org, const Replica &R) {
  const int l = c - R.cell0;
  const int ix = l % R.m[0], iy = (l / R.m[0]) % R.m[1], iz = l / (R.m[0] * R.m[1]);
  const double *o = org + R.col0;
  return make_float4((float)(q.x - o[ix]), (float)(q.y - o[R.m[0] + iy]), (float)(q.z - o[R.m[0] + R.m[1] + iz]),
                     0.0f);
}

// replica of entry i of the owned + ghost arrays
__device__ __forceinline__ int replica_of(int i, int pn, const int *__restrict__ rep, const int *__restrict__ gcode) {
  return i < pn ? rep[i] : ghost_rep(gcode[i - pn]);
}

__global__ void k_to_rel(const double4 *__restrict__ pos, const int *__restrict__ cell, int pn, int tn,
                         const int *__restrict__ rep, const int *__restrict__ gcode, const Replica *__restrict__ reps,
                         const double *__restrict__ org, float4 *__restrict__ rel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tn) return;
  rel[i] = to_rel(pos[i], cell[i], org, reps[replica_of(i, pn, rep, gcode)]);
}

// Fused first half drift + halo refresh:
//   owned:  tmp[i] = q_i + p_i*dt2                    (and its cell-relative FP32 copy)
//   ghost:  pos[g] = (q_src + p_src*dt2) + shift*L   (and its FP32 copy)
// (the same two roundings as UpdatePositionHalf then SendBorderParticles).
// The force kernel reads the owned particles from `tmp` and the ghosts from
// `pos` (gather_pos); `pos` keeps the owned positions until the force kernel
// writes the drifted ones, and the ghosts until the next refresh.
template <bool kRel>
__global__ void k_drift_refresh(double4 *__restrict__ pos, const double4 *__restrict__ mom,
                                double4 *__restrict__ tmp, const int *__restrict__ gsrc,
                                const int *__restrict__ gcode, int pn, int ng, double dt2,
                                const int *__restrict__ rep, const Replica *__restrict__ reps,
                                const int *__restrict__ cell, const double *__restrict__ org,
                                float4 *__restrict__ rel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < pn) {
    double4 q = pos[i];
    const double4 p = mom[i];
    q.x += p.x * dt2;
    q.y += p.y * dt2;
    q.z += p.z * dt2;
    tmp[i] = q;
    if (kRel) rel[i] = to_rel(q, cell[i], org, reps[rep[i]]);
  } else if (i < pn + ng) {
    const int k = i - pn;
    const int src = gsrc[k];
    const int w = gcode[k];
    const Replica &R = reps[ghost_rep(w)];
    double4 q = pos[src];
    const double4 p = mom[src];
    q.x += p.x * dt2;
    q.y += p.y * dt2;
    q.z += p.z * dt2;
    const double3 y = image_of(q, ghost_shift(w), R.L);
    const double4 g = make_double4(y.x, y.y, y.z, 0.0);
    pos[i] = g;
    if (kRel) rel[i] = to_rel(g, cell[i], org, R);
  }
}

// The owned particles' positions at the end of a step, from that step's first
// half drift qh and final momenta: qh + p dt/2 (drift) or qh (Langevin), the
// operations and roundings of the kick's epilogue (finish_row).  A run's steps
// leave them out, since no step of the run reads them; the host writes them
// once after the run, from the last completed step.
__global__ void k_final_positions(double4 *__restrict__ pos, const double4 *__restrict__ qh,
                                  const double4 *__restrict__ mom, int pn, double dt2, int drift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pn) return;
  double4 q = qh[i];
  if (drift) {
    const double4 p = mom[i];
    q.x += p.x * dt2;
    q.y += p.y * dt2;
    q.z += p.z * dt2;
  }
  pos[i] = q;
}

// Halo refresh when `tmp` already holds the half-drifted owned positions
// (written by the previous fused step's epilogue): ghosts only,
// pos[g] = tmp[src] + shift*L (and the FP32 copy) -- the values
// k_drift_refresh would write.
template <bool kRel>
__global__ void k_refresh_from_drifted(double4 *__restrict__ pos, const double4 *__restrict__ tmp,
                                       const int *__restrict__ gsrc, const int *__restrict__ gcode, int pn,
                                       int ng, const Replica *__restrict__ reps, const int *__restrict__ cell,
                                       const double *__restrict__ org, float4 *__restrict__ rel,
                                       const RunGate *__restrict__ gate) {
  if (run_halted(gate)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ng) return;
  const int i = pn + k;
  const int w = gcode[k];
  const Replica &R = reps[ghost_rep(w)];
  const double3 y = image_of(tmp[gsrc[k]], ghost_shift(w), R.L);
  const double4 g = make_double4(y.x, y.y, y.z, 0.0);
  pos[i] = g;
  if (kRel) rel[i] = to_rel(g, cell[i], org, R);
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
__global__ void k_langevin(double4 *__restrict__ mom, const int *__restrict__ id, int pn, const int *__restrict__ rep,
                           const Replica *__restrict__ reps, double dt, uint64_t step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pn) return;
  double4 p = mom[i];
  langevin_update(p, langevin_consts(reps[rep[i]], dt, step), (uint32_t)id[i]);
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
// x, y, z of n 32-byte entries to and from 3 doubles per entry (host transfers)
__global__ void k_pack3(const double4 *__restrict__ src, int n, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 v = src[i];
  double *o = out + 3 * (size_t)i;
  o[0] = v.x;
  o[1] = v.y;
  o[2] = v.z;
}

__global__ void k_unpack3(const double *__restrict__ in, int n, double4 *__restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *v = in + 3 * (size_t)i;
  dst[i] = make_double4(v[0], v[1], v[2], 0.0);
}

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
  block_max_to_partial(v, partial, blockIdx.x);
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
  block_max_to_partial(v, partial, blockIdx.x);
}

// second pass: particles displaced by more than (margin - max displacement);
// up to `cap` of them are collected (the reference stops at CheckListLength).
// The threshold comes from the maximum of the first pass (the bits of a
// double in dmax), c = margin - sqrt(dmax), compared squared.
__global__ void k_candidates(const double *__restrict__ disp, int tn, const unsigned long long *__restrict__ dmax,
                             double margin, int *__restrict__ count, int *__restrict__ list, int cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tn) return;
  const double c_distance = margin - sqrt(__longlong_as_double((long long)dmax[0]));
  const double c_d2 = c_distance * c_distance;
  if (disp[i] > c_d2) {
    const int k = atomicAdd(count, 1);
    if (k < cap) list[k] = i;
  }
}

// third pass, for 1 < l < Nl candidates (l = count[0]; Nl or more expire the
// list without it): a candidate pair that was beyond SL at the build and is
// now inside rc expires the list.  Launched for Nl * Nl pairs.
__global__ void k_candidate_pairs(const double4 *__restrict__ pos, const double4 *__restrict__ old,
                                  const int *__restrict__ list, const int *__restrict__ count, int Nl, double SL2,
                                  double CL2, int *__restrict__ flag) {
  const int l = count[0];
  if (l <= 1 || l >= Nl) return;
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

// ------------------------------------------------------------- replicas
// Reductions and the expiry test of an ensemble, one workgroup per replica
// (kBlock threads, fixed order: repeated calls agree bitwise).

template <int T = kBlock>
__device__ __forceinline__ double block_max_all(double v) {
  __shared__ double wm[T / kWave];
  __shared__ double res;
  v = wave_max(v);
  if ((threadIdx.x % kWave) == 0) wm[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wm[0];
    for (int k = 1; k < T / kWave; k++) m = fmax(m, wm[k]);
    res = m;
  }
  __syncthreads();
  return res;
}

__device__ __forceinline__ double block_sum0(double v) {
  __shared__ double ws[kBlock / kWave];
  v = wave_sum(v);
  if ((threadIdx.x % kWave) == 0) ws[threadIdx.x / kWave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kBlock / kWave; k++) t += ws[k];
  __syncthreads();
  return t;  // valid in thread 0
}

// Largest per-brick value of each replica: out[r] = max of part over the
// replica's bricks (the force kernel's maxima of |p|^2).  clock (optional) is
// the step counter of the fused step; it advances by one here.
__global__ __launch_bounds__(kBlock) void k_replica_max(const double *__restrict__ part,
                                                        const Replica *__restrict__ reps, double *__restrict__ out,
                                                        unsigned long long *__restrict__ clock,
                                                        RunGate *__restrict__ gate) {
  if (run_halted(gate)) return;
  const Replica &R = reps[blockIdx.x];
  const int b0 = R.brick0, nb = replica_bricks(R);
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;  // four loads in flight per thread
  int k = threadIdx.x;
  for (; k + 3 * kBlock < nb; k += 4 * kBlock) {
    const double a = part[b0 + k], b = part[b0 + k + kBlock];
    const double c = part[b0 + k + 2 * kBlock], e = part[b0 + k + 3 * kBlock];
    m0 = fmax(m0, a);
    m1 = fmax(m1, b);
    m2 = fmax(m2, c);
    m3 = fmax(m3, e);
  }
  for (; k < nb; k += kBlock) m0 = fmax(m0, part[b0 + k]);
  double m = fmax(fmax(m0, m1), fmax(m2, m3));
  m = block_max_all(m);
  if (threadIdx.x == 0) {
    out[blockIdx.x] = m;
    if (clock != nullptr && blockIdx.x == 0) clock[0] = clock[0] + 1ull;
    if (gate != nullptr && blockIdx.x == 0) gate->done = gate->done + 1;
  }
}

// PairList::IsPairListExpired with UpdatePairListValidity (pairlist.cc:26-65)
// for the next step of a run, every replica at once: the lifetime counts up
// and a budget still positive shrinks by 2 max|p| dt, from the replica's
// max |p|^2 after the last step (v2).  When some replica's particles are too
// fast or its budget ends below zero, nothing is committed and the run halts:
// the host takes that step (and its displacement check) from the same values.
// One workgroup.
__device__ __forceinline__ bool gate_budget(double v2, double dt, double margin, double &b) {
  if (b > 0.0) {
    const double vmax = sqrt(v2);
    if (vmax * 2.0 * dt > margin) return true;
    b = b - vmax * 2.0 * dt;
  }
  return b < 0.0;
}

__global__ __launch_bounds__(kBlock) void k_gate(RunGate *__restrict__ gate, double *__restrict__ budget,
                                                 int *__restrict__ lifetime, const double *__restrict__ v2,
                                                 int nrep, double dt, double margin) {
  __shared__ int stop;
  if (gate->halted) return;
  if (threadIdx.x == 0) stop = 0;
  __syncthreads();
  for (int r = threadIdx.x; r < nrep; r += kBlock) {
    double b = budget[r];
    if (gate_budget(v2[r], dt, margin, b)) stop = 1;
  }
  __syncthreads();
  if (stop) {
    if (threadIdx.x == 0) gate->halted = 1;
    return;
  }
  for (int r = threadIdx.x; r < nrep; r += kBlock) {
    double b = budget[r];
    gate_budget(v2[r], dt, margin, b);
    budget[r] = b;
    lifetime[r] = lifetime[r] + 1;
  }
}

// max |p|^2 over each replica's owned particles (PairList::UpdatePairListValidity,
// pairlist.cc:48-65, before the square root)
__global__ __launch_bounds__(kBlock) void k_replica_max_p2(const double4 *__restrict__ mom,
                                                           const Replica *__restrict__ reps,
                                                           double *__restrict__ out) {
  const Replica &R = reps[blockIdx.x];
  double m = 0.0;
  for (int k = threadIdx.x; k < R.pn; k += kBlock) {
    const double4 p = mom[R.p0 + k];
    m = fmax(m, p.x * p.x + p.y * p.y + p.z * p.z);
  }
  m = block_max_all(m);
  if (threadIdx.x == 0) out[blockIdx.x] = m;
}

// ForceCalculator::HeatbathMomenta (fcalculator.cc:752-762) of every replica
// with its own factor exp(-dt/2 zeta_r).
__global__ void k_scale_momenta_replicas(double4 *__restrict__ mom, const int *__restrict__ rep, int n,
                                         const double *__restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = f[rep[i]];
  double4 p = mom[i];
  p.x *= s; p.y *= s; p.z *= s; p.w *= s;
  mom[i] = p;
}

// Per-replica sums of the observers' per-brick partials (3 per brick):
// out[3r + c] = sum over the replica's bricks of part[3b + c].
__global__ __launch_bounds__(kBlock) void k_replica_sum3(const double *__restrict__ part,
                                                         const Replica *__restrict__ reps, double *__restrict__ out) {
  const Replica &R = reps[blockIdx.x];
  const int b0 = R.brick0, nb = replica_bricks(R);
  for (int c = 0; c < 3; c++) {
    double t = 0.0;
    for (int k = threadIdx.x; k < nb; k += kBlock) t += part[3 * (long long)(b0 + k) + c];
    t = block_sum0(t);
    if (threadIdx.x == 0) out[3 * blockIdx.x + c] = t;
  }
}

// KineticEnergyObserver (observer.cc:5-15) of each replica.
__global__ __launch_bounds__(kBlock) void k_replica_kinetic(const double4 *__restrict__ mom,
                                                            const Replica *__restrict__ reps,
                                                            double *__restrict__ out) {
  const Replica &R = reps[blockIdx.x];
  double t = 0.0;
  for (int k = threadIdx.x; k < R.pn; k += kBlock) {
    const double4 p = mom[R.p0 + k];
    t += 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z + 0.5 * p.w * p.w;
  }
  t = block_sum0(t);
  if (threadIdx.x == 0) out[blockIdx.x] = 0.0 + t;
}

// PairList::CheckByDisplacement (pairlist.cc:68-121) of the replicas listed
// in `which`, over each replica's owned entries [p0, p0 + pn) and its ghosts
// [pn_all + g0, pn_all + g0 + ng) (grng[r] = {g0, ng}), in three launches:
//  1. k_check_dmax, kCheckBlocks workgroups per listed replica: each entry's
//     squared displacement since the build into disp, and the largest into
//     dmax[k] as the bits of a double (non-negative doubles order like their
//     bits);
//  2. k_check_cand, as many workgroups: the entries displaced by more than
//     margin - max displacement (read back from disp), up to Nl = min(cap, n)
//     of them stored, all of them counted in cnt[k];
//  3. k_check_pairs, kPairBlocks workgroups per listed replica sharing its
//     candidate pairs: expired once Nl candidates were found, else when a
//     candidate pair that was beyond SL at the build is now inside rc.
// dmax, cnt and expired start at zero (k_zero_check); expired[k] receives the
// answer for replica which[k].  The decision does not depend on the order in which
// candidates are stored.
constexpr int kCheckCap = 1024;   // largest CheckListLength an ensemble accepts (cap² pair tests per replica)
constexpr int kCheckBlocks = 32;  // workgroups per listed replica in the first two launches
constexpr int kPairBlocks = 8;    // workgroups per listed replica in the candidate pair test

__device__ __forceinline__ int check_entry(int k, const Replica &R, int no, int g0, int pn_all) {
  return k < no ? R.p0 + k : pn_all + g0 + (k - no);
}

__global__ __launch_bounds__(kBlock) void k_check_dmax(const double4 *__restrict__ pos,
                                                       const double4 *__restrict__ old,
                                                       const int *__restrict__ which, const Replica *__restrict__ reps,
                                                       const int2 *__restrict__ grng, int pn_all,
                                                       double *__restrict__ disp,
                                                       unsigned long long *__restrict__ dmax) {
  const int kr = blockIdx.x / kCheckBlocks, part = blockIdx.x % kCheckBlocks;
  const int r = which[kr];
  const Replica &R = reps[r];
  const int no = R.pn, g0 = grng[r].x, n = no + grng[r].y;
  double m = 0.0;
  for (int k = part * kBlock + threadIdx.x; k < n; k += kCheckBlocks * kBlock) {
    const int i = check_entry(k, R, no, g0, pn_all);
    const double4 a = old[i], b = pos[i];
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    disp[i] = d2;
    m = fmax(m, d2);
  }
  m = block_max_all(m);
  if (threadIdx.x == 0) atomicMax(&dmax[kr], (unsigned long long)__double_as_longlong(m));
}

__global__ __launch_bounds__(kBlock) void k_check_cand(const double *__restrict__ disp,
                                                       const int *__restrict__ which, const Replica *__restrict__ reps,
                                                       const int2 *__restrict__ grng, int pn_all, double margin,
                                                       int cap, const unsigned long long *__restrict__ dmax,
                                                       int *__restrict__ cnt, int *__restrict__ cand) {
  const int kr = blockIdx.x / kCheckBlocks, part = blockIdx.x % kCheckBlocks;
  const int r = which[kr];
  const Replica &R = reps[r];
  const int no = R.pn, g0 = grng[r].x, n = no + grng[r].y;
  const int Nl = n > cap ? cap : n;
  const double c_distance = margin - sqrt(__longlong_as_double((long long)dmax[kr]));
  const double c_d2 = c_distance * c_distance;
  for (int k = part * kBlock + threadIdx.x; k < n; k += kCheckBlocks * kBlock) {
    const int i = check_entry(k, R, no, g0, pn_all);
    if (disp[i] > c_d2) {
      const int c = atomicAdd(&cnt[kr], 1);
      if (c < Nl) cand[(long long)kr * cap + c] = i;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_check_pairs(const double4 *__restrict__ pos,
                                                        const double4 *__restrict__ old,
                                                        const int *__restrict__ which,
                                                        const Replica *__restrict__ reps,
                                                        const int2 *__restrict__ grng, int cap, double SL2,
                                                        double CL2, const int *__restrict__ cnt,
                                                        const int *__restrict__ cand, int *__restrict__ expired) {
  __shared__ int flag;
  const int kr = blockIdx.x;
  const int r = which[kr];
  const int n = reps[r].pn + grng[r].y;
  const int Nl = n > cap ? cap : n;
  const int l = cnt[kr];
  if (l > 0 && l >= Nl) {
    if (threadIdx.x == 0 && blockIdx.y == 0) expired[kr] = 1;
    return;
  }
  if (threadIdx.x == 0) flag = 0;
  __syncthreads();
  const int *c = cand + (long long)kr * cap;
  // gridDim.y workgroups share the replica's pairs; any that finds one sets
  // the flag (zeroed by k_zero_check), so the outcome is an OR over all pairs
  for (int t = blockIdx.y * kBlock + threadIdx.x; t < l * l; t += kBlock * gridDim.y) {
    const int a = t / l, b = t % l;
    if (b <= a) continue;
    const double4 o1 = old[c[a]], o2 = old[c[b]];
    double dx = o1.x - o2.x, dy = o1.y - o2.y, dz = o1.z - o2.z;
    if (dx * dx + dy * dy + dz * dz < SL2) continue;
    const double4 q1 = pos[c[a]], q2 = pos[c[b]];
    dx = q1.x - q2.x;
    dy = q1.y - q2.y;
    dz = q1.z - q2.z;
    if (dx * dx + dy * dy + dz * dz < CL2) flag = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0 && flag) expired[kr] = 1;
}

__global__ void k_zero_check(unsigned long long *__restrict__ dmax, int *__restrict__ cnt,
                             int *__restrict__ expired, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    dmax[k] = 0ull;
    cnt[k] = 0;
    expired[k] = 0;
  }
}

// Ghost ranges of each replica after a construction: replica r's images are
// [goff[p0], goff[p0 + pn]) (goff = exclusive scan of the image counts over
// the replica-major owned order; total = the number of images).
__global__ void k_replica_ghosts(const int *__restrict__ goff, const Replica *__restrict__ reps, int nrep,
                                 int total, int2 *__restrict__ grng) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  const Replica &R = reps[r];
  const int a = R.pn > 0 ? goff[R.p0] : 0;
  int e = total;
  for (int s = r + 1; s < nrep; s++)
    if (reps[s].pn > 0) {
      e = goff[reps[s].p0];
      break;
    }
  const int g0 = R.pn > 0 ? a : e;
  grng[r] = make_int2(g0, e - g0);
}

__global__ void k_set_u64(unsigned long long *__restrict__ p, unsigned long long v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) p[0] = v;
}

}  // namespace mdgpu
