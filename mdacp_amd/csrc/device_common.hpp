// device_common.hpp -- shared device helpers for the gfx950 LJ kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdgpu {

constexpr int kWave = 64;      // CDNA wavefront width
constexpr int kBlock = 256;    // 4 waves per workgroup
constexpr int kMaxParticles = (1 << 26) - 2;        // owned + ghosts + sentinel per domain
constexpr uint32_t kZeroOffset = 0x15u;             // cell offset (0,0,0): 01 01 01
constexpr int kGroup = 8;                           // list slots a lane holds contiguously
constexpr int kSliceRound = kGroup;                 // slice lengths are multiples of this
constexpr int kMaxAhead = 4;                        // list-index groups the force kernels read ahead
constexpr int kListTail = kMaxAhead * kGroup * kWave;  // readable padding after the list arrays

// Position of slot k of lane `lane` inside a 64-row slice of the list.  Slots
// come in groups of kGroup: a lane's kGroup consecutive slots are contiguous
// (16 bytes of uint16 indices, 8 bytes of codes) and the 64 lanes' groups
// follow one another, so group g of a slice is one contiguous run that a
// wavefront loads with one 16-byte (8-byte) load per lane.
__host__ __device__ __forceinline__ long long slot_of(int k, int lane) {
  return (long long)(k / kGroup) * (kGroup * kWave) + lane * kGroup + (k % kGroup);
}

// Error bits raised by kernels; read back at the existing sync points.
constexpr int kErrInvalidPosition = 1;
constexpr int kErrInvalidIndex = 2;
constexpr int kErrCapacity = 4;
constexpr int kErrLayout = 8;  // a listed partner outside its brick's staged ranges

// Brick kernels serve brick blockIdx.x.  The hardware hands consecutive
// workgroups to the 8 XCDs in turn, so every XCD (each with its own L2) gets
// an even share of every region of a domain and of every replica of an
// ensemble.  Giving each XCD one contiguous run of bricks instead, for L2
// locality, was slower (the 4M-particle kick by 3.5%, the 256-replica
// ensemble's by 1.7%): the last-level cache shared by the XCDs holds most of
// the positions anyway, and the XCDs' shares of the work then differ.

__device__ __forceinline__ double wave_max(double v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Maximum over the workgroup, written to partial[slot] by thread 0 (no
// same-address atomics: those serialise in L2 across ~10^4 workgroups).
// Every thread of the workgroup must call it.
__device__ __forceinline__ void block_max_to_partial(double v, double *__restrict__ partial, int slot) {
  __shared__ double wm[16];  // up to 1024 threads
  v = wave_max(v);
  if ((threadIdx.x % kWave) == 0) wm[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wm[0];
    for (int k = 1; k < (int)(blockDim.x / kWave); k++) m = fmax(m, wm[k]);
    partial[slot] = m;
  }
}

// One workgroup: max of n partials, stored as the bits of a double.
__global__ void k_reduce_max(const double *__restrict__ partial, int n, unsigned long long *__restrict__ out) {
  __shared__ double red[kBlock];
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;  // four loads in flight per thread
  int k = threadIdx.x;
  for (; k + 3 * (int)blockDim.x < n; k += 4 * blockDim.x) {
    const double a = partial[k], b = partial[k + blockDim.x];
    const double c = partial[k + 2 * blockDim.x], e = partial[k + 3 * blockDim.x];
    m0 = fmax(m0, a);
    m1 = fmax(m1, b);
    m2 = fmax(m2, c);
    m3 = fmax(m3, e);
  }
  for (; k < n; k += blockDim.x) m0 = fmax(m0, partial[k]);
  red[threadIdx.x] = fmax(fmax(m0, m1), fmax(m2, m3));
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = (unsigned long long)__double_as_longlong(red[0]);
}

// One replica's periodic box and cell grid.  A domain advances one or more
// independent replicas (an ensemble); their particles, cells and bricks are
// concatenated in the domain's arrays, each replica's contiguous, and every
// kernel that needs geometry finds its replica here.  A single-domain run has
// one replica.  The box of a replica is [0, L) (MDRect of a 1x1x1 grid,
// mdunit.cc:17-24), or [o, o + L) for a unit of a larger grid; its extended
// cell grid is MeshList's (meshlist.cc:31-55).
struct Replica {
  double L[3];     // box edges
  double o[3];     // box origin (MDRect's start): cells count from it; 0 for a periodic box
  double im[3];    // 1 / cell width
  double w[3];     // cell width
  double hb_D;     // Langevin noise amplitude sqrt(2 gamma T / dt) (fcalculator.cc:766)
  double gamma;    // Langevin friction
  uint64_t seed;   // Langevin key
  int m[3];        // cells per axis, one ghost layer per side included
  int nb[3];       // bricks per axis
  int bd[3];       // cells per brick along each axis
  int cell0;       // first global cell
  int brick0;      // first global brick
  int col0;        // first column origin (x columns, then y, then z) in the origin table
  int p0, pn;      // owned particles [p0, p0 + pn) once binned
};

__device__ __forceinline__ int replica_bricks(const Replica &R) { return R.nb[0] * R.nb[1] * R.nb[2]; }

// A run of fused steps queued without waiting for the host.  The gate kernel
// at the head of each step halts the run when that step's validity test needs
// the host (a budget spent: the displacement check or a construction is due;
// or a particle too fast); the step's other kernels then return at once, and
// so does every step queued after it, which leaves the state of the last
// completed step for the host to take the halted step from.
struct RunGate {
  int halted;  // 0 while the run goes on
  int done;    // steps completed in this run
};

__device__ __forceinline__ bool run_halted(const RunGate *g) { return g != nullptr && g->halted != 0; }

// A ghost's code word: image shift code (sx+1)+3(sy+1)+9(sz+1) in the low
// five bits, the replica above them.
constexpr int kShiftBits = 5;
__host__ __device__ __forceinline__ int ghost_word(int code, int rep) { return code | (rep << kShiftBits); }
__host__ __device__ __forceinline__ int ghost_shift(int word) { return word & ((1 << kShiftBits) - 1); }
__host__ __device__ __forceinline__ int ghost_rep(int word) { return word >> kShiftBits; }

// Image displacement of a shift code (sx+1)+3(sy+1)+9(sz+1): one +-L or 0.0
// per component, added as the reference's halo adds it (mdunit.cc:142-161).
__device__ __forceinline__ double3 image_of(double4 q, int code, const double *L) {
  const int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
  const double dx = sx > 0 ? L[0] : (sx < 0 ? -L[0] : 0.0);
  const double dy = sy > 0 ? L[1] : (sy < 0 ? -L[1] : 0.0);
  const double dz = sz > 0 ? L[2] : (sz < 0 ? -L[2] : 0.0);
  return make_double3(q.x + dx, q.y + dy, q.z + dz);
}

}  // namespace mdgpu
