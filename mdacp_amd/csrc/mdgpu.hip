// mdgpu.hip -- host side of the gfx950 LJ domain: device state, the pair-list
// rebuild pipeline, the step sequence of MDManager::Calculate and the C ABI
// declared in include/mdgpu.h.
//
// Build (see mdacp_amd/build.py):
//   hipcc --offload-arch=gfx950 -O3 -ffp-contract=off -fPIC -shared
//         -Xclang -target-feature -Xclang -packed-fp32-ops
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mdgpu.h"
#include "build_kernels.hpp"
#include "force_kernels.hpp"
#include "list_kernels.hpp"
#include "scan.hpp"
#include "step_kernels.hpp"

using namespace mdgpu;

// Device scratch of mdgpu_observe_noalloc: the per-brick partial sums of the
// observer kernel (3 doubles per brick, up to kNoAllocBricks bricks) and the
// reduced totals after them (3 doubles).  Each GPU holds its own copy, at its
// own address, from the library's load until the process ends.  Host code
// never uses this name as an address: create_domain resolves the device copy
// once per device with hipGetSymbolAddress and every launch takes that pointer.
// Every call on one GPU shares the array, so calls that overlap on one GPU
// (several domains or threads at once) race on it.
constexpr int kNoAllocBricks = 131072;
constexpr size_t kScratchDoubles = 3 * (size_t)kNoAllocBricks + 4;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
// Parameters and PairList state of one replica (pairlist.h:10-31).
struct ReplicaState {
  double T = 1.0, gamma = 0.1;  // Langevin target temperature and friction
  uint64_t seed = 0;            // Langevin key
  double buffer_length = 0.0;   // validity budget
  int lifetime = 0;
  double zeta = 0.0;            // Nose-Hoover friction (Variables::Zeta)
};

// A captured fused step (ghost refresh, pair force, per-replica maxima and
// their copy to the host), replayed while the pointers it was captured with
// stay valid.
struct StepGraph {
  hipGraphExec_t exec = nullptr;
  const void *key = nullptr;  // the half-drifted positions it reads
  long epoch = -1;
};

// The units of one process grid (MDManager's MDUnits of one process,
// mdmanager.cc:34-45) on one device: they share one stream, so that work
// each unit queues after MDManager's phase barriers runs after what every
// unit queued before them, and each unit's ghosts come straight from the
// owned particles of the units they image.
struct UnitGroup {
  std::mutex mu;  // the slot table, and the host-side pointers of units read by other units
  hipStream_t stream = nullptr;
  int device = 0;
  double L[3] = {0.0, 0.0, 0.0};  // the whole box
  int grid[3] = {1, 1, 1};        // units per axis
  std::vector<mdgpu_domain *> units;  // by slot; nullptr once destroyed
  int live = 0;
  long epoch = 0;  // bumped when a unit is destroyed: ghosts placed before then may image it
  // the whole-group expiry test: per-block maxima, then each unit's max |p|^2 (bits) and its pinned copy
  double *part = nullptr;
  size_t part_n = 0;
  unsigned long long *vmax2 = nullptr, *h_vmax2 = nullptr;
};
}  // namespace

struct mdgpu_domain {
  mdgpu_params prm{};
  hipStream_t stream = nullptr;
  int device = 0;
  std::string err;
  int precision = MDGPU_FP64;

  // tuning
  int kahan = 2;          // mixed kernel's row sums: 0 plain FP32, 1 Kahan FP32, 2 FP32 groups summed in FP64
  int unroll = 0;         // list entries whose indices are loaded ahead in the force kernels (8, 16 or 32;
                          // 0: unroll_of's default for the precision)
  int waves = 8;          // wavefronts per brick workgroup: 2, 4, 8 or 16
  int bd[3] = {4, 4, 4};  // cells per brick
  bool brick_auto = true; // a single domain picks bd from its density at the first construction
  bool brick_user = false; // the caller set bd (MDGPU_OPT_BRICK): kept through scale changes
  int brick_kb = 0;         // LDS budget the automatic bricks are sized for, KiB (0: the default)
  int lds_kb = 0;         // LDS budget for one brick's staged partners, KiB (0: the bricks' budget)
  int lds_max_kb = 60;    // device limit per workgroup less the kernels' static LDS
  int num_cu = 256;       // compute units (the pipelined kick's persistent grid)
  int disp_check = 1;     // 1 = PairList::CheckByDisplacement once the budget is spent (pairlist.cc:26-45),
                          // 0 = the list expires as soon as the budget is spent
  int mixed_frame = 1;    // mixed kick: 0 = cell-relative FP32 copies with cell-offset codes,
                          // 1 = brick-centred FP32 frame staged from the FP64 positions (default)
  int pipeline = 0;       // FP64 kick: 1 = persistent workgroups staging the next brick while
                          // computing the current one (k_force_pipe_fp64), 0 = a workgroup per brick
  int search_waves = 16;  // wavefronts per brick workgroup of the neighbour search: 4, 8 or 16
  int search_stage = 1;   // neighbour search reads candidates staged in LDS (1) or from global memory (0)
  int row_sort = kSliceRound;  // rows of a brick by list length in steps of this (0: cell order)
  int search_cells = 1;   // neighbour search by a wavefront per cell (1) or a thread per row (0)

  // SimulationInfo / MDRect / MeshList geometry
  double L[3], rs[3], re[3];
  double dt, rc, margin, SL, S2, CL2, C0, C2;
  double msize[3];
  int m[3], ncell = 0;
  int check_list_length = 500;
  // bricks
  int nb[3], nbrick = 0, rmax = 0;
  std::vector<int> h_brick_of_cell;

  // replicas: one for a single domain, several for an ensemble
  int nrep = 1;
  std::vector<Replica> rh;         // host copy of the replica table
  std::vector<ReplicaState> rst;   // per-replica parameters and list budget
  Replica *reps = nullptr;         // device replica table
  int *brep = nullptr;             // replica of each brick
  int *rep = nullptr, *rep_alt = nullptr;  // replica of each owned particle
  int2 *grng = nullptr;            // ghost range {first, count} of each replica
  double *rep_v2 = nullptr;        // per replica: max |p|^2 after the last fused step
  double *rep_out = nullptr;       // per replica observer sums (3 each)
  int *which = nullptr, *rep_exp = nullptr;  // replicas to check, their answers
  double *h_rep_v2 = nullptr, *h_rep_out = nullptr;
  double *rep_factor = nullptr, *h_rep_factor = nullptr;  // per replica: Nose-Hoover momentum factor
  int *h_which = nullptr, *h_rep_exp = nullptr;
  unsigned long long *clock = nullptr;  // step counter read by replayed force launches
  unsigned long long *chk_dmax = nullptr;  // displacement check of listed replicas: max |dq|^2 bits,
  unsigned long long *h_chk_dmax = nullptr;  // and its host copy
  int *chk_cnt = nullptr, *chk_cand = nullptr;  // candidate counts, candidates (cap per replica)
  bool clock_dirty = true;              // clock differs from `step`

  // unit of a process grid (mdgpu_create_unit); none for a periodic domain or an ensemble
  UnitGroup *group = nullptr;
  int unit_slot = -1;          // index in the group: a ghost's source unit
  int upos[3] = {0, 0, 0};     // position in the grid
  bool own_stream = true;      // false: the group's stream
  bool ghosts_ready = false;   // ghosts placed for the current owned particles and their order
  long ghost_epoch = -1;       // the group's epoch when they were placed
  bool ghost_remap = false;    // ghost sources still index the sources' rows as they were before their row order
  int *gunit = nullptr;        // source unit of each ghost
  double4 *st_pos = nullptr, *st_mom = nullptr;  // send buffer of the last MakeBufferForSendingParticle
  int *st_id = nullptr;
  int st_n = 0, st_cap = 0;
  SrcUnit *src_tab = nullptr, *h_src_tab = nullptr;  // ghost sources (device; pinned mirror)
  int *ucnt = nullptr, *uoff = nullptr;  // per source particle: images, first image
  long long ucap = 0;

  // captured fused steps, one per parity of the position buffers
  bool use_graph = false;   // fused steps replayed from a captured graph (launched directly when false)
  long epoch = 0;  // bumped whenever a captured step may no longer match
  StepGraph graphs[2];
  // runs of steps queued without waiting for the host (k_gate at the head of
  // each step): at most run_ahead steps per run (0: every step waits)
  int run_ahead = 64;
  StepGraph gated[2];
  RunGate *gate = nullptr, *h_gate = nullptr;  // device state and its pinned mirror
  double *gbudget = nullptr, *h_gbudget = nullptr;  // per replica: validity budget
  int *glife = nullptr, *h_glife = nullptr;         // per replica: list lifetime

  // thermostat
  int thermostat = MDGPU_NVE;
  double T_aim = 1.0, gamma = 0.1, tau = 0.1, zeta = 0.0;
  uint64_t seed = 0, step = 0;
  double sim_time = 0.0;

  // PairList state (pairlist.h:10-31)
  bool initialized = false;
  double buffer_length = 0.0;
  int lifetime = 0;
  bool vmax_valid = false;

  // counts and capacities
  int pn = 0, ng = 0, tn = 0;
  int cap = 0, gcap = 0;
  int nslice = 0, slice_cap = 0, max_staged = 0;
  long long nbr_slots = 0, nbr_entries = 0, nbr_cap = 0;

  // device state: positions as double4 (x, y, z, 0); owned [0, pn), ghosts [pn, tn)
  double4 *pos = nullptr;      // current owned positions; ghosts of the last halo refresh
  double4 *pos_tmp = nullptr;  // half-drifted owned positions inside the fused step; build scratch
  double4 *pos_next = nullptr; // the next step's half-drifted owned positions (fused epilogue)
  double4 *pos_old = nullptr;  // PairList::Init snapshot
  double4 *mom = nullptr, *mom_alt = nullptr;
  float4 *rel = nullptr, *rel_tmp = nullptr;  // cell-relative FP32 copies (owned + ghosts)
  float4 *rel_next = nullptr;
  bool mid_valid = false;  // pos_tmp (rel_tmp) already holds this step's first half drift
  int *id = nullptr, *id_alt = nullptr, *cell = nullptr, *cell_alt = nullptr;
  int *gsrc = nullptr, *gcode = nullptr;
  int *c_ocount = nullptr, *c_ostart = nullptr, *c_gcount = nullptr, *c_gstart = nullptr;
  int *count_r = nullptr, *start_r = nullptr;
  double *org = nullptr;  // origins of the cell columns, x then y then z
  // bricks
  int *rank = nullptr, *brick_of_cell = nullptr, *brick_rank0 = nullptr;
  int *row0 = nullptr, *nrow = nullptr, *nsl = nullptr, *vslice0 = nullptr, *slice_brick = nullptr;
  int *tstart = nullptr, *tbase = nullptr, *nrange = nullptr, *btotal = nullptr;
  long long *gbase = nullptr;  // [nbrick + 1] first gather-list entry of each brick
  int *glist = nullptr;        // staged entry -> particle index, brick by brick
  long long glist_cap = 0;
  // list
  uint16_t *nbr = nullptr;   // staged partner index per list slot
  uint8_t *ncode = nullptr;  // partner cell offset per list slot (cell-frame mixed kick only)
  long long code_cap = 0;    // slots ncode and ncode_in hold
  bool list_coded = false;   // the current list (and inner list) carry the codes
  // inner (pruned) list of the fused step: same slots, per-slice inner length
  uint16_t *nbr_in = nullptr;
  uint8_t *ncode_in = nullptr;
  int *slen_in = nullptr;
  double prune_delta = 0.0;    // inner radius rc + prune_delta; 0: the fused step reads the full list
  bool inner_valid = false;    // the inner list holds every pair that can come within rc
  double inner_budget = 0.0;   // what is left of prune_delta for particle motion
  double step_vmax = 0.0;      // max |p| at this step's expiry test (all replicas)
  int prunes = 0;
  long long *sbase = nullptr, *slots64 = nullptr;
  int *slen = nullptr, *row_len = nullptr;
  // build scratch
  int *rowsrc = nullptr, *rowinv = nullptr;  // row order inside bricks: new -> old, old -> new
  long long gtotal = 0;                       // gather-list entries
  uint32_t *mask = nullptr;  // candidate bitmask per row, mcap words
  size_t mask_cap = 0;
  int mcap = 32;
  double *spart = nullptr;  // per-workgroup maxima of slice lengths
  int *slot = nullptr, *perm = nullptr, *nimg = nullptr, *goff = nullptr;
  int *perm_in = nullptr;  // binned items grouped by cell, before the in-cell order
  int *t_gsrc = nullptr, *t_gcode = nullptr, *t_gcell = nullptr;
  long long *key = nullptr, *scan_scratch = nullptr;
  long long scan_scratch_n = 0;
  double *xfer = nullptr;  // 3 doubles per particle: host buffers in the boundary's layout
  size_t xfer_cap = 0;
  double *disp = nullptr, *partial = nullptr, *maxpart = nullptr;
  double *bpart = nullptr;  // one maximum per brick workgroup
  int partial_n = 0;
  int *cand = nullptr;
  // small device scalars: u64 [0] vmax2 [1] dmax2 [2] max staged; i32 [0] err [1] count [2] flag
  unsigned long long *u64 = nullptr;
  long long *i64 = nullptr;
  int *i32 = nullptr;
  double *d3 = nullptr;
  // pinned host mirrors
  unsigned long long *h_u64 = nullptr;
  long long *h_i64 = nullptr;
  int *h_i32 = nullptr;
  double *h_d3 = nullptr;
  double *scratch_dev = nullptr;  // this device's copy of `scratch` (mdgpu_observe_noalloc)

  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;
  int force_launches = 0, rebuild_count = 0;
  double force_ms = 0.0, total_ms = 0.0, rebuild_ms = 0.0;
};

namespace {

#define HIPCHK(d, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) {                                                               \
      (d)->err = std::string(#call) + ": " + hipGetErrorString(e_);                        \
      return MDGPU_ERR_HIP;                                                               \
    }                                                                                     \
  } while (0)

#define CHK(call)                  \
  do {                             \
    int r_ = (call);               \
    if (r_ != MDGPU_OK) return r_; \
  } while (0)

// The cell-relative FP32 copies (rel, rel_tmp, rel_next) are kept only for
// the mixed kick in the cell-relative frame.
// List entries loaded ahead by the force kernels: the caller's choice, else
// two groups for the FP64 kick (measured 1-2% faster than one at 4M
// particles) and one for the mixed kick (whose two-group form runs 3% slower).
inline int unroll_of(const mdgpu_domain *d) {
  return d->unroll ? d->unroll : (d->precision == MDGPU_MIXED ? 8 : 16);
}

inline bool rel_frame(const mdgpu_domain *d) { return d->precision == MDGPU_MIXED && d->mixed_frame == 0; }
// Only the cell-frame mixed kick (rel_frame) reads each partner's cell offset
// code: a construction writes them for that frame alone, and a switch to it
// writes them for the current list (recode_list).
inline bool list_codes(const mdgpu_domain *d) { return d->list_coded; }

inline unsigned nblk(long long n, int b = kBlock) { return (unsigned)std::max<long long>(1, (n + b - 1) / b); }

template <class T>
int dalloc(mdgpu_domain *d, T **p, size_t n) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (n == 0) n = 1;
  HIPCHK(d, hipMalloc((void **)p, n * sizeof(T)));
  return MDGPU_OK;
}

template <class T>
int dgrow_copy(mdgpu_domain *d, T **p, size_t keep, size_t new_n) {
  T *np = nullptr;
  HIPCHK(d, hipMalloc((void **)&np, std::max<size_t>(1, new_n) * sizeof(T)));
  if (*p && keep) HIPCHK(d, hipMemcpyAsync(np, *p, keep * sizeof(T), hipMemcpyDeviceToDevice, d->stream));
  if (*p) {
    HIPCHK(d, hipStreamSynchronize(d->stream));
    (void)hipFree(*p);
  }
  *p = np;
  return MDGPU_OK;
}

int sync(mdgpu_domain *d) {
  HIPCHK(d, hipStreamSynchronize(d->stream));
  return MDGPU_OK;
}

int check_launch(mdgpu_domain *d, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    d->err = std::string(what) + ": " + hipGetErrorString(e);
    return MDGPU_ERR_HIP;
  }
  return MDGPU_OK;
}

// Invalidate the captured steps and the half drift the next fused step would
// reuse (state that the step reads has changed).
void invalidate(mdgpu_domain *d) {
  d->mid_valid = false;
  d->inner_valid = false;
  d->epoch++;
}

// The replica table on the device: Langevin constants from each replica's
// parameters (hb_D = sqrt(2 gamma T / dt), fcalculator.cc:766) and, for a
// single domain, its owned count.
int upload_replicas(mdgpu_domain *d) {
  if (d->nrep == 1) {
    d->rst[0].T = d->T_aim;
    d->rst[0].gamma = d->gamma;
    d->rst[0].seed = d->seed;
    d->rh[0].p0 = 0;
    d->rh[0].pn = d->pn;
  }
  for (int r = 0; r < d->nrep; r++) {
    Replica &R = d->rh[r];
    const ReplicaState &S = d->rst[r];
    R.gamma = S.gamma;
    R.hb_D = std::sqrt(2.0 * S.gamma * S.T / d->dt);
    R.seed = S.seed;
  }
  CHK(sync(d));
  HIPCHK(d, hipMemcpy(d->reps, d->rh.data(), d->nrep * sizeof(Replica), hipMemcpyHostToDevice));
  return MDGPU_OK;
}

// MeshList::ChangeScale (meshlist.cc:31-55) of every replica, the column
// origins of the cell-relative FP32 coordinates, and the brick order of the
// cells: cells, bricks and origins of the replicas are concatenated, each
// replica's cells ranked brick by brick.
int set_geometry(mdgpu_domain *d) {
  CHK(sync(d));
  if (d->nrep == 1)
    for (int k = 0; k < 3; k++) {
      d->rh[0].L[k] = d->re[k] - d->rs[k];
      d->rh[0].o[k] = d->rs[k];
      d->rh[0].bd[k] = d->bd[k];
    }
  long long ncell = 0, nbrick = 0;
  int rmax = 0;
  std::vector<double> org;
  for (int r = 0; r < d->nrep; r++) {
    Replica &R = d->rh[r];
    for (int k = 0; k < 3; k++) {
      const double w = R.L[k];
      const int ms = (int)(w / d->SL);
      if (ms < 2) {
        d->err = "box edge shorter than two search lengths (rc + margin)";
        return MDGPU_ERR_ARGUMENT;
      }
      R.w[k] = w / (double)ms;
      R.im[k] = 1.0 / R.w[k];
      R.m[k] = (int)(w / R.w[k]) + 2;
      R.nb[k] = (R.m[k] - 1 + R.bd[k] - 1) / R.bd[k];  // bricks tile cells 1 .. m - 1
    }
    R.cell0 = (int)ncell;
    R.brick0 = (int)nbrick;
    R.col0 = (int)org.size();
    // Origins of the FP32 cell-relative coordinates: multiples of the cell width
    // rounded to FP32, exact in FP64 (24-bit width times an integer below 2^12).
    // The mixed kernels add (cell_j - cell_i) * (float)w for a partner one cell
    // over, which is then exactly the difference of the two origins: a partner
    // separation carries no per-cell bias, only the rounding of the two
    // coordinates.  The cells themselves (binning) keep the FP64 width.
    // A unit's box starts at R.o: its origins are R.o + (c - 1) * w32, and a
    // partner separation then carries at most the rounding of those sums.
    for (int k = 0; k < 3; k++) {
      const double w32 = (double)(float)R.w[k];
      for (int c = 0; c < R.m[k]; c++) org.push_back(R.o[k] + (double)(c - 1) * w32);
    }
    ncell += (long long)R.m[0] * R.m[1] * R.m[2];
    nbrick += (long long)R.nb[0] * R.nb[1] * R.nb[2];
    rmax = std::max(rmax, (R.bd[0] + 2) * (R.bd[1] + 2) * (R.bd[2] + 2));
    if (ncell > (1ll << 30)) {
      d->err = "too many cells";
      return MDGPU_ERR_CAPACITY;
    }
  }
  d->ncell = (int)ncell;
  d->nbrick = (int)nbrick;
  d->rmax = rmax;
  if (d->nrep == 1)
    for (int k = 0; k < 3; k++) {
      d->msize[k] = d->rh[0].w[k];
      d->m[k] = d->rh[0].m[k];
      d->nb[k] = d->rh[0].nb[k];
    }
  CHK(dalloc(d, &d->c_ocount, d->ncell));
  CHK(dalloc(d, &d->c_ostart, d->ncell));
  CHK(dalloc(d, &d->c_gcount, d->ncell));
  CHK(dalloc(d, &d->c_gstart, d->ncell));
  CHK(dalloc(d, &d->count_r, d->ncell + 1));
  CHK(dalloc(d, &d->start_r, d->ncell + 1));
  CHK(dalloc(d, &d->org, org.size()));
  HIPCHK(d, hipMemcpy(d->org, org.data(), org.size() * sizeof(double), hipMemcpyHostToDevice));
  // bricks z-major inside each replica, cells z, y, x inside each brick; then
  // the replica's lower ghost-layer cells, which belong to no brick (no owned
  // particle bins there: k_wrap_bin refuses one)
  std::vector<int> rank(d->ncell), boc(d->ncell, -1), rank0(d->nbrick + 1), brep(d->nbrick);
  int rk = 0;
  for (int r = 0; r < d->nrep; r++) {
    const Replica &R = d->rh[r];
    for (int bz = 0; bz < R.nb[2]; bz++)
      for (int by = 0; by < R.nb[1]; by++)
        for (int bx = 0; bx < R.nb[0]; bx++) {
          const int b = R.brick0 + (bz * R.nb[1] + by) * R.nb[0] + bx;
          rank0[b] = rk;
          brep[b] = r;
          for (int z = 1 + bz * R.bd[2]; z < std::min(R.m[2], 1 + (bz + 1) * R.bd[2]); z++)
            for (int y = 1 + by * R.bd[1]; y < std::min(R.m[1], 1 + (by + 1) * R.bd[1]); y++)
              for (int x = 1 + bx * R.bd[0]; x < std::min(R.m[0], 1 + (bx + 1) * R.bd[0]); x++) {
                const int c = R.cell0 + (z * R.m[1] + y) * R.m[0] + x;
                rank[c] = rk++;
                boc[c] = b;
              }
        }
    for (int z = 0; z < R.m[2]; z++)
      for (int y = 0; y < R.m[1]; y++)
        for (int x = 0; x < R.m[0]; x++)
          if (x == 0 || y == 0 || z == 0) rank[R.cell0 + (z * R.m[1] + y) * R.m[0] + x] = rk++;
  }
  rank0[d->nbrick] = rk;
  d->h_brick_of_cell = boc;
  CHK(dalloc(d, &d->rank, d->ncell));
  CHK(dalloc(d, &d->brick_of_cell, d->ncell));
  CHK(dalloc(d, &d->brick_rank0, d->nbrick + 1));
  CHK(dalloc(d, &d->brep, d->nbrick));
  HIPCHK(d, hipMemcpy(d->rank, rank.data(), rank.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(d, hipMemcpy(d->brick_of_cell, boc.data(), boc.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(d, hipMemcpy(d->brick_rank0, rank0.data(), rank0.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(d, hipMemcpy(d->brep, brep.data(), brep.size() * sizeof(int), hipMemcpyHostToDevice));
  CHK(dalloc(d, &d->row0, d->nbrick));
  CHK(dalloc(d, &d->nrow, d->nbrick));
  CHK(dalloc(d, &d->nsl, d->nbrick));
  CHK(dalloc(d, &d->vslice0, d->nbrick));
  CHK(dalloc(d, &d->nrange, d->nbrick));
  CHK(dalloc(d, &d->bpart, d->nbrick));
  CHK(dalloc(d, &d->btotal, d->nbrick));
  CHK(dalloc(d, &d->gbase, d->nbrick + 1));
  CHK(dalloc(d, &d->tstart, (size_t)d->nbrick * kRangesPerCell * d->rmax));
  CHK(dalloc(d, &d->tbase, (size_t)d->nbrick * (kRangesPerCell * d->rmax + 1)));
  CHK(upload_replicas(d));
  d->initialized = false;
  invalidate(d);
  return MDGPU_OK;
}

Bricks bricks_of(const mdgpu_domain *d) {
  Bricks B;
  B.brep = d->brep;
  B.reps = d->reps;
  B.rmax = d->rmax;
  return B;
}

// capacity for owned particles (arrays indexed by owned index)
int ensure_owned(mdgpu_domain *d, int need) {
  if (need <= d->cap) return MDGPU_OK;
  if (need > kMaxParticles) {
    d->err = "particle capacity exceeded";
    return MDGPU_ERR_CAPACITY;
  }
  int nc = std::max(need, std::max(1024, d->cap + d->cap / 4));
  nc = std::min(nc, kMaxParticles);
  const size_t tc_new = (size_t)nc + d->gcap + 1;
  CHK(dgrow_copy(d, &d->pos, d->pn, tc_new));
  CHK(dgrow_copy(d, &d->mom, d->pn, nc));
  CHK(dgrow_copy(d, &d->id, d->pn, nc));
  CHK(dgrow_copy(d, &d->rep, d->pn, nc));
  CHK(dalloc(d, &d->rep_alt, nc));
  CHK(dalloc(d, &d->pos_tmp, tc_new));
  CHK(dalloc(d, &d->pos_next, tc_new));
  CHK(dalloc(d, &d->pos_old, tc_new));
  CHK(dalloc(d, &d->rel, tc_new));
  CHK(dalloc(d, &d->rel_tmp, tc_new));
  CHK(dalloc(d, &d->rel_next, tc_new));
  CHK(dalloc(d, &d->cell, tc_new));
  CHK(dalloc(d, &d->cell_alt, tc_new));
  CHK(dalloc(d, &d->disp, tc_new));
  CHK(dalloc(d, &d->maxpart, nblk((long long)tc_new, kWave)));  // one per workgroup of >= 64 threads
  CHK(dalloc(d, &d->mom_alt, nc));
  CHK(dalloc(d, &d->id_alt, nc));
  CHK(dalloc(d, &d->slot, tc_new));
  CHK(dalloc(d, &d->perm, tc_new));
  CHK(dalloc(d, &d->perm_in, tc_new));
  CHK(dalloc(d, &d->key, tc_new));
  CHK(dalloc(d, &d->nimg, nc));
  CHK(dalloc(d, &d->goff, nc));
  CHK(dalloc(d, &d->row_len, nc));
  CHK(dalloc(d, &d->rowsrc, nc));
  CHK(dalloc(d, &d->rowinv, nc));
  d->cap = nc;
  invalidate(d);
  return MDGPU_OK;
}

// capacity for ghost images (arrays indexed by owned+ghost index)
int ensure_ghosts(mdgpu_domain *d, int need) {
  if (need <= d->gcap) return MDGPU_OK;
  if ((long long)d->cap + need + 1 > kMaxParticles) {
    d->err = "particle capacity exceeded (owned + ghosts)";
    return MDGPU_ERR_CAPACITY;
  }
  // room for the ghost count to fluctuate between constructions: growing
  // reallocates and copies every per-entry array
  const long long want = std::max<long long>((long long)need + need / 8 + 1024, (long long)d->gcap + d->gcap / 4);
  const int ngc = (int)std::min<long long>(want, kMaxParticles - 1 - (long long)d->cap);
  const size_t tc_new = (size_t)d->cap + ngc + 1;
  CHK(dgrow_copy(d, &d->pos, d->pn, tc_new));
  CHK(dgrow_copy(d, &d->cell, d->pn, tc_new));
  CHK(dalloc(d, &d->pos_tmp, tc_new));
  CHK(dalloc(d, &d->pos_next, tc_new));
  CHK(dalloc(d, &d->pos_old, tc_new));
  CHK(dalloc(d, &d->rel, tc_new));
  CHK(dalloc(d, &d->rel_tmp, tc_new));
  CHK(dalloc(d, &d->rel_next, tc_new));
  CHK(dalloc(d, &d->cell_alt, tc_new));
  CHK(dalloc(d, &d->disp, tc_new));
  CHK(dalloc(d, &d->maxpart, nblk((long long)tc_new, kWave)));
  CHK(dalloc(d, &d->slot, tc_new));
  CHK(dalloc(d, &d->perm, tc_new));
  CHK(dalloc(d, &d->perm_in, tc_new));
  CHK(dalloc(d, &d->key, tc_new));
  CHK(dalloc(d, &d->gsrc, ngc));
  CHK(dalloc(d, &d->gcode, ngc));
  CHK(dalloc(d, &d->gunit, ngc));
  CHK(dalloc(d, &d->t_gsrc, ngc));
  CHK(dalloc(d, &d->t_gcode, ngc));
  CHK(dalloc(d, &d->t_gcell, ngc));
  d->gcap = ngc;
  invalidate(d);
  return MDGPU_OK;
}

int ensure_scan(mdgpu_domain *d, long long n) {
  const long long need = scan_scratch_size(n);
  if (d->scan_scratch && need <= d->scan_scratch_n) return MDGPU_OK;
  d->scan_scratch_n = 0;  // dalloc frees the old buffer first; a failed allocation leaves none
  CHK(dalloc(d, &d->scan_scratch, (size_t)need));
  d->scan_scratch_n = need;
  return MDGPU_OK;
}

int ensure_partial(mdgpu_domain *d, int nb) {
  if (d->partial && nb * 3 <= d->partial_n) return MDGPU_OK;
  d->partial_n = 0;
  CHK(dalloc(d, &d->partial, (size_t)nb * 3));
  d->partial_n = nb * 3;
  return MDGPU_OK;
}

int read_err_flags(mdgpu_domain *d) {
  HIPCHK(d, hipMemcpyAsync(d->h_i32, d->i32, sizeof(int) * 4, hipMemcpyDeviceToHost, d->stream));
  CHK(sync(d));
  const int e = d->h_i32[0];
  if (e & kErrInvalidPosition) {
    d->err = "Invalid position";
    return MDGPU_ERR_INVALID_POSITION;
  }
  if (e & kErrInvalidIndex) {
    d->err = "Invalid index";
    return MDGPU_ERR_INVALID_INDEX;
  }
  if (e & kErrLayout) {
    d->err = "internal: a listed partner lies outside its brick's staged ranges";
    return MDGPU_ERR_HIP;
  }
  if (e & (kErrCapacity | kErrSearchCap)) {
    d->err = "pair list capacity";
    return MDGPU_ERR_CAPACITY;
  }
  return MDGPU_OK;
}

int read_total(mdgpu_domain *d, long long *v) {
  HIPCHK(d, hipMemcpyAsync(d->h_i64, d->i64, sizeof(long long), hipMemcpyDeviceToHost, d->stream));
  CHK(sync(d));
  *v = d->h_i64[0];
  return MDGPU_OK;
}

// a double stored as bits in u64[k]
int read_max(mdgpu_domain *d, int k, double *v) {
  HIPCHK(d, hipMemcpyAsync(d->h_u64 + k, d->u64 + k, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           d->stream));
  CHK(sync(d));
  std::memcpy(v, d->h_u64 + k, sizeof(double));
  return MDGPU_OK;
}

// Order n binned items (cellv[], slot[] from atomics, counts[] per cell):
// starts[] over cells (in brick rank order when `ranked`), perm[] = new order,
// members of a cell by (key, input index), or input index when key is null
// (k_rank_in_cell).
int cell_sort(mdgpu_domain *d, const int *cellv, int n, int *counts, int *starts, bool ranked,
              const int *key) {
  hipStream_t st = d->stream;
  CHK(ensure_scan(d, d->ncell));
  if (ranked) {
    hipLaunchKernelGGL(k_rank_counts, dim3(nblk(d->ncell)), dim3(kBlock), 0, st, counts, d->rank, d->ncell,
                       d->count_r);
    exclusive_scan<int, int>(d->count_r, d->start_r, d->ncell, d->scan_scratch, nullptr, st);
    hipLaunchKernelGGL(k_unrank_starts, dim3(nblk(d->ncell)), dim3(kBlock), 0, st, d->start_r, d->rank, d->ncell,
                       starts);
  } else {
    exclusive_scan<int, int>(counts, starts, d->ncell, d->scan_scratch, nullptr, st);
  }
  hipLaunchKernelGGL(k_scatter_perm, dim3(nblk(n)), dim3(kBlock), 0, st, cellv, d->slot, starts, n, d->perm_in);
  if (n > 0)
    hipLaunchKernelGGL(k_rank_in_cell, dim3(nblk(n)), dim3(kBlock), 0, st, d->perm_in, cellv, starts, counts, key, n,
                       d->perm);
  return check_launch(d, "cell sort");
}

template <class K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// staged entries of bytes_per_entry that fit the LDS budget (incl. the pad entry)
int lds_cap_for(const mdgpu_domain *d, int bytes_per_entry, int kb) {
  const int budget = std::min(kb, d->lds_max_kb) * 1024 / bytes_per_entry;
  return std::max(1, std::min(d->max_staged + 1, budget));
}

// LDS budget of the neighbour search's staging (independent of the force
// kernels' budget: a construction runs one search per brick, without the
// per-step trade-off between staged bricks and workgroups per CU).
constexpr int kSearchLdsKb = 96;

SearchTables search_tables(const mdgpu_domain *d) {
  SearchTables t;
  t.tstart = d->tstart;
  t.tbase = d->tbase;
  t.nrange = d->nrange;
  t.glist = d->glist;
  t.gbase = d->gbase;
  t.row0 = d->row0;
  t.nrow = d->nrow;
  t.vslice0 = d->vslice0;
  t.rstride = kRangesPerCell * d->rmax;
  t.lds_cap = lds_cap_for(d, 3 * sizeof(double), kSearchLdsKb);
  return t;
}

template <int W>
void launch_search_w(mdgpu_domain *d, const Bricks &B, SearchTables t) {
  if (!d->search_stage) t.lds_cap = 0;  // every brick reads its candidates from global memory
  const size_t shm = (size_t)t.lds_cap * 3 * sizeof(double);
  if (d->search_cells) {
    allow_lds(k_search_cells<W>, shm);
    hipLaunchKernelGGL(k_search_cells<W>, dim3(d->nbrick), dim3(kWave * W), shm, d->stream, d->pos, B, t, d->S2,
                       d->mcap, d->row_len, d->mask, d->i32);
  } else {
    allow_lds(k_search_brick<W>, shm);
    hipLaunchKernelGGL(k_search_brick<W>, dim3(d->nbrick), dim3(kWave * W), shm, d->stream, d->pos, d->cell, B, t,
                       d->S2, d->mcap, d->row_len, d->mask, d->i32);
  }
}

// One workgroup per brick, a wavefront per cell (or a thread per row);
// candidates from the brick's neighbourhood staged in LDS like the force
// kernel, or read from global memory (L2).
int launch_search(mdgpu_domain *d, const Bricks &B, const SearchTables &t) {
  switch (d->search_waves) {
    case 4: launch_search_w<4>(d, B, t); break;
    case 8: launch_search_w<8>(d, B, t); break;
    default: launch_search_w<16>(d, B, t); break;
  }
  return check_launch(d, "neighbour search");
}

// Room for a code per slot of the list and of the inner list.
int ensure_codes(mdgpu_domain *d) {
  if (d->ncode && d->code_cap >= d->nbr_cap) return MDGPU_OK;
  d->code_cap = 0;
  CHK(dalloc(d, &d->ncode, (size_t)(d->nbr_cap + kListTail)));
  CHK(dalloc(d, &d->ncode_in, (size_t)(d->nbr_cap + kListTail)));
  d->code_cap = d->nbr_cap;
  return MDGPU_OK;
}

// one wavefront per 64-row slice, four per brick workgroup; with codes for
// the cell-frame mixed kick only (recode_list gives them to a built list)
int launch_emit(mdgpu_domain *d, const Bricks &B, const SearchTables &t) {
  const bool codes = rel_frame(d);
  if (codes) {
    CHK(ensure_codes(d));
    hipLaunchKernelGGL((k_emit_brick<kEmitWaves, true>), dim3(d->nbrick), dim3(kWave * kEmitWaves), 0, d->stream,
                       d->cell, B, t, d->rowsrc, d->mask, d->mcap, d->sbase, d->slen, d->nbr, d->ncode, d->i32);
  } else {
    hipLaunchKernelGGL((k_emit_brick<kEmitWaves, false>), dim3(d->nbrick), dim3(kWave * kEmitWaves), 0, d->stream,
                       d->cell, B, t, d->rowsrc, d->mask, d->mcap, d->sbase, d->slen, d->nbr, nullptr, d->i32);
  }
  d->list_coded = codes;
  return check_launch(d, "neighbour list emit");
}

// A switch to the cell frame gives the current list its codes: the emit runs
// again from the construction's candidate masks, which stay in place until
// the next construction, and writes the same indices with the codes.  The
// inner list is pruned again at the next step (the option change invalidates
// it).
int recode_list(mdgpu_domain *d) {
  if (!rel_frame(d) || d->list_coded || !d->initialized || d->pn == 0 || d->nbr_slots == 0) return MDGPU_OK;
  return launch_emit(d, bricks_of(d), search_tables(d));
}

// Rows of each brick by decreasing list length: permute the owned particles
// (positions, momenta, ids, cells, row lengths) and remap the indices that
// point at them (gather lists, ghost sources).  Ghosts keep their places.
// With row_sort 0 the rows stay in cell order (rowsrc is the identity) and
// nothing moves: a unit's owned particles keep the indices other units'
// ghosts refer to.
int order_rows(mdgpu_domain *d) {
  hipStream_t st = d->stream;
  const int pn = d->pn, ng = d->ng;
  hipLaunchKernelGGL(k_brick_row_order, dim3(d->nbrick), dim3(kBlock), 0, st, d->row_len, d->row0, d->nrow,
                     d->rowsrc, d->rowinv, d->row_sort);
  if (d->row_sort == 0) return check_launch(d, "row order");
  hipLaunchKernelGGL(k_gather_owned, dim3(nblk(pn)), dim3(kBlock), 0, st, d->rowsrc, pn, d->pos, d->mom, d->id,
                     d->cell, d->rep, d->pos_tmp, d->mom_alt, d->id_alt, d->cell_alt, d->rep_alt);
  hipLaunchKernelGGL(k_gather_int, dim3(nblk(pn)), dim3(kBlock), 0, st, d->rowsrc, pn, d->row_len, d->nimg);
  if (ng > 0) {
    HIPCHK(d, hipMemcpyAsync(d->pos_tmp + pn, d->pos + pn, ng * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIPCHK(d, hipMemcpyAsync(d->cell_alt + pn, d->cell + pn, ng * sizeof(int), hipMemcpyDeviceToDevice, st));
    // a unit's ghost sources index other units' rows: remapped once every
    // unit of the group has ordered its own (remap_unit_ghosts)
    if (!d->group)
      hipLaunchKernelGGL(k_remap_owned, dim3(nblk(ng)), dim3(kBlock), 0, st, d->gsrc, (long long)ng, pn, d->rowinv);
  }
  if (d->gtotal > 0)
    hipLaunchKernelGGL(k_remap_owned, dim3(nblk(d->gtotal)), dim3(kBlock), 0, st, d->glist, d->gtotal, pn,
                       d->rowinv);
  std::swap(d->pos, d->pos_tmp);
  std::swap(d->mom, d->mom_alt);
  std::swap(d->id, d->id_alt);
  std::swap(d->cell, d->cell_alt);
  std::swap(d->rep, d->rep_alt);
  std::swap(d->row_len, d->nimg);
  return check_launch(d, "row order");
}

// Cells per brick for a box of density rho (cubic bricks): the largest edge
// whose staged neighbourhood (bd + 2 cells per axis, FP64 entries of 24
// bytes) is expected to fill at most 90% of kb KiB of LDS.
// Cells of the neighbourhoods of all bricks along one axis, counting only
// cells inside the grid: bricks of bd cells tile cells 1 .. ms + 1 of a grid
// of ms + 2 (ms owned cells and the two ghost layers).
long long axis_staged_cells(int ms, int bd) {
  long long t = 0;
  for (int c0 = 1; c0 <= ms + 1; c0 += bd) {
    const int c1 = std::min(ms + 2, c0 + bd);
    t += std::min(c1, ms + 1) - (c0 - 1) + 1;
  }
  return t;
}

// Brick dimensions for a box of edges L at density rho: among the bricks
// whose neighbourhood fits kMaxBrickCells cells and, at the mean density,
// the fraction `fill` of kb KiB of LDS (positions, 24 B per staged entry),
// the ones that stage the fewest cells in all (ties: fewer bricks).  Staging
// the neighbourhood is the per-brick overhead of the search and force kernels.
void auto_brick(const mdgpu_domain *d, double rho, const double L[3], int kb, int bd[3], double fill) {
  const double budget = fill * std::min(kb, d->lds_max_kb) * 1024.0;
  int ms[3];
  double vol = 1.0;
  for (int k = 0; k < 3; k++) {
    ms[k] = std::max(1, (int)(L[k] / d->SL));
    vol *= L[k] / ms[k];
  }
  long long best = -1, best_n = 0;
  for (int k = 0; k < 3; k++) bd[k] = 1;
  for (int bz = 1; bz <= kMaxBrickEdge; bz++)
    for (int by = 1; by <= kMaxBrickEdge; by++)
      for (int bx = 1; bx <= kMaxBrickEdge; bx++) {
        const int e[3] = {bx, by, bz};
        const int cells = (bx + 2) * (by + 2) * (bz + 2);
        if (cells > kMaxBrickCells || rho * vol * cells * 24.0 > budget) continue;
        long long staged = 1, n = 1;
        for (int k = 0; k < 3; k++) {
          staged *= axis_staged_cells(ms[k], e[k]);
          n *= (ms[k] + e[k]) / e[k];
        }
        if (best < 0 || staged < best || (staged == best && n < best_n)) {
          best = staged;
          best_n = n;
          for (int k = 0; k < 3; k++) bd[k] = e[k];
        }
      }
}

// Ensemble replicas: bricks for two workgroups per CU (160 KiB of LDS),
// their mean neighbourhood within 90% of the budget (small boxes: the
// neighbourhoods' counts spread more).
constexpr int kReplicaBrickKb = 76;
constexpr double kReplicaFill = 0.9;
// A single domain, unless the caller chose its bricks: bricks for two
// workgroups per CU as well, chosen at the first construction (and after a
// scale change) from the density.  The mean neighbourhood may take 96% of
// the budget: in a uniform fluid of millions of particles the fullest
// neighbourhood holds about 3% more than the mean (4M particles at rho =
// 0.5: 3,099 entries against 3,003 for 4x4x4-cell bricks).  At rho = 0.6
// this picks 4x4x3 cells, whose kick (16.4M particles) takes 1.45 ms against
// 1.84 ms for 4x4x4, which holds one workgroup per CU.
constexpr int kDomainBrickKb = 78;
constexpr double kDomainFill = 0.96;

int brick_budget(const mdgpu_domain *d) {
  return d->brick_kb ? d->brick_kb : (d->nrep == 1 ? kDomainBrickKb : kReplicaBrickKb);
}

// The force kernels' LDS budget for a single domain whose bricks are sized
// automatically: more than the bricks' 78 KiB, so that a state whose fullest
// neighbourhoods outgrow that sizing (density fluctuations) runs one workgroup
// per CU instead of reading those bricks' partners from memory, several
// times slower.
constexpr int kDomainForceKb = 96;

// The force kernels' LDS budget: unless set, the one the bricks are sized for
// (an ensemble's replicas: two workgroups per CU at 76 KiB), or
// kDomainForceKb for a single domain with automatic bricks.
int force_lds_kb(const mdgpu_domain *d) {
  if (d->lds_kb) return d->lds_kb;
  if (d->nrep == 1 && !d->brick_kb) return kDomainForceKb;
  return brick_budget(d);
}

bool is_unit(const mdgpu_domain *d) { return d->group != nullptr; }

// A single domain (or unit) picks its bricks from its density at its first
// construction, unless the caller chose them.
int auto_bricks(mdgpu_domain *d) {
  if (d->nrep == 1 && d->brick_auto && d->pn > 0) {
    d->brick_auto = false;
    const double rho = d->pn / (d->rh[0].L[0] * d->rh[0].L[1] * d->rh[0].L[2]);
    int b[3];
    auto_brick(d, rho, d->rh[0].L, brick_budget(d), b, kDomainFill);
    if (b[0] != d->bd[0] || b[1] != d->bd[1] || b[2] != d->bd[2]) {
      for (int k = 0; k < 3; k++) d->bd[k] = b[k];
      CHK(set_geometry(d));
    }
  }
  return MDGPU_OK;
}

// Bin the owned particles (after the periodic wrap of each replica's box when
// `wrap`) and put them in brick order, each cell's members by particle id.
int bin_owned(mdgpu_domain *d, int wrap) {
  hipStream_t st = d->stream;
  invalidate(d);  // particles are reordered and wrapped
  const int pn = d->pn;
  d->ng = 0;
  d->tn = pn;
  d->initialized = false;
  d->ghosts_ready = false;
  HIPCHK(d, hipMemsetAsync(d->c_ocount, 0, sizeof(int) * d->ncell, st));
  if (pn > 0) {
    hipLaunchKernelGGL(k_wrap_bin, dim3(nblk(pn)), dim3(kBlock), 0, st, d->pos, pn, d->rep, d->reps, d->cell,
                       d->slot, d->c_ocount, d->i32, wrap);
  }
  CHK(read_err_flags(d));
  CHK(cell_sort(d, d->cell, pn, d->c_ocount, d->c_ostart, true, d->id));
  if (pn > 0) {
    hipLaunchKernelGGL(k_gather_owned, dim3(nblk(pn)), dim3(kBlock), 0, st, d->perm, pn, d->pos, d->mom, d->id,
                       d->cell, d->rep, d->pos_tmp, d->mom_alt, d->id_alt, d->cell_alt, d->rep_alt);
    std::swap(d->pos, d->pos_tmp);
    std::swap(d->mom, d->mom_alt);
    std::swap(d->id, d->id_alt);
    std::swap(d->cell, d->cell_alt);
    std::swap(d->rep, d->rep_alt);
  }
  return MDGPU_OK;
}

// The ghost images of a periodic domain or of every replica of an ensemble
// (FindBorderParticles / SendBorderParticlesSub x 6), cell-sorted after the
// owned particles.
int place_images(mdgpu_domain *d) {
  hipStream_t st = d->stream;
  const int pn = d->pn;
  CHK(ensure_scan(d, std::max(pn, d->ncell)));
  long long ngl = 0;
  if (pn > 0) {
    hipLaunchKernelGGL(k_count_images, dim3(nblk(pn)), dim3(kBlock), 0, st, d->pos, pn, d->rep, d->reps, d->SL,
                       d->nimg);
    exclusive_scan<int, int>(d->nimg, d->goff, pn, d->scan_scratch, d->i64, st);
    CHK(read_total(d, &ngl));
  }
  if (ngl + pn + 1 > kMaxParticles) {
    d->err = "particle capacity exceeded (owned + ghosts)";
    return MDGPU_ERR_CAPACITY;
  }
  const int ng = (int)ngl;
  CHK(ensure_ghosts(d, ng));
  HIPCHK(d, hipMemsetAsync(d->c_gcount, 0, sizeof(int) * d->ncell, st));
  if (ng > 0) {
    // pos_tmp[0..ng) holds the unsorted images
    hipLaunchKernelGGL(k_emit_images, dim3(nblk(pn)), dim3(kBlock), 0, st, d->pos, pn, d->rep, d->reps, d->SL,
                       d->goff, ng, d->t_gsrc, d->t_gcode, d->pos_tmp, d->t_gcell, d->slot, d->c_gcount);
    CHK(check_launch(d, "emit images"));
    CHK(cell_sort(d, d->t_gcell, ng, d->c_gcount, d->c_gstart, false, nullptr));
    hipLaunchKernelGGL(k_gather_ghosts, dim3(nblk(ng)), dim3(kBlock), 0, st, d->perm, ng, pn, d->t_gsrc,
                       d->t_gcode, d->pos_tmp, d->t_gcell, d->gsrc, d->gcode, d->pos, d->cell);
  } else {
    exclusive_scan<int, int>(d->c_gcount, d->c_gstart, d->ncell, d->scan_scratch, nullptr, st);
  }
  d->ng = ng;
  d->tn = pn + ng;
  if (d->nrep > 1 && pn > 0)
    hipLaunchKernelGGL(k_replica_ghosts, dim3(nblk(d->nrep)), dim3(kBlock), 0, st, d->goff, d->reps, d->nrep, ng,
                       d->grng);
  d->ghosts_ready = true;
  return check_launch(d, "ghost images");
}

// Bricks, the neighbour list and the PairList::Init snapshot for the owned
// particles and ghosts in place (bin_owned, then place_images or a unit's
// ghost placement).
int build_lists(mdgpu_domain *d) {
  hipStream_t st = d->stream;
  const int pn = d->pn;
  HIPCHK(d, hipMemsetAsync(d->i32, 0, sizeof(int) * 4, st));
  // 4. bricks: rows, slices, staged ranges
  const Bricks B = bricks_of(d);
  const CellTables T{d->c_ostart, d->c_ocount, d->c_gstart, d->c_gcount};
  hipLaunchKernelGGL(k_brick_rows, dim3(nblk(d->nbrick)), dim3(kBlock), 0, st, d->brick_rank0, d->start_r,
                     d->nbrick, d->ncell, pn, d->row0, d->nrow, d->nsl);
  CHK(ensure_scan(d, d->nbrick));
  exclusive_scan<int, int>(d->nsl, d->vslice0, d->nbrick, d->scan_scratch, d->i64, st);
  long long nsl_total = 0;
  CHK(read_total(d, &nsl_total));
  d->nslice = (int)nsl_total;
  if (d->nslice + 1 > d->slice_cap) {
    const int sc = d->nslice + d->nslice / 4 + 64;
    CHK(dalloc(d, &d->slen, sc));
    CHK(dalloc(d, &d->slen_in, sc));
    CHK(dalloc(d, &d->sbase, sc));
    CHK(dalloc(d, &d->slots64, sc));
    CHK(dalloc(d, &d->slice_brick, sc));
    CHK(dalloc(d, &d->spart, nblk(sc)));
    d->slice_cap = sc;
  }
  hipLaunchKernelGGL(k_slice_brick, dim3(nblk(d->nbrick)), dim3(kBlock), 0, st, d->vslice0, d->nsl, d->nbrick,
                     d->slice_brick);
  hipLaunchKernelGGL(k_brick_ranges, dim3(nblk((long long)d->nbrick * kWave)), dim3(kBlock), 0, st, B, T,
                     d->nbrick, pn, d->tstart, d->tbase, d->nrange, d->btotal, d->bpart);
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(kBlock), 0, st, d->bpart, d->nbrick, d->u64 + 2);
  // gather lists: gbase = exclusive scan of the staged counts, gbase[nbrick] = total;
  // the largest neighbourhood and the total are read after one wait
  exclusive_scan<int, long long>(d->btotal, d->gbase, d->nbrick, d->scan_scratch, d->i64, st);
  HIPCHK(d, hipMemcpyAsync(d->gbase + d->nbrick, d->i64, sizeof(long long), hipMemcpyDeviceToDevice, st));
  HIPCHK(d, hipMemcpyAsync(d->h_u64 + 2, d->u64 + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  long long gtotal = 0;
  CHK(read_total(d, &gtotal));
  double mstaged = 0.0;
  std::memcpy(&mstaged, d->h_u64 + 2, sizeof(double));
  d->max_staged = (int)mstaged;
  if (d->max_staged > 65534) {
    d->err = "a brick's neighbourhood exceeds 16-bit indexing (extreme local density): use smaller bricks";
    return MDGPU_ERR_CAPACITY;
  }
  d->gtotal = gtotal;
  if (gtotal > d->glist_cap) {
    const long long gc = gtotal + gtotal / 8 + 1024;
    CHK(dalloc(d, &d->glist, (size_t)gc));
    d->glist_cap = gc;
  }
  hipLaunchKernelGGL(k_brick_glist, dim3(d->nbrick), dim3(kBlock), 0, st, kRangesPerCell * d->rmax, d->tstart,
                     d->tbase, d->nrange, d->gbase, d->glist);
  // 5. MeshList::MakeList: full neighbour list in sliced-ELL layout
  long long slots = 0;
  int scap = 0;
  const SearchTables STb = search_tables(d);
  if (pn > 0) {
    // candidate masks: retried with room for more candidates if a row overflows
    for (;;) {
      const size_t need = (size_t)pn * d->mcap;
      if (!d->mask || need > d->mask_cap) {
        d->mask_cap = 0;
        CHK(dalloc(d, &d->mask, need + need / 8));
        d->mask_cap = need + need / 8;
      }
      HIPCHK(d, hipMemsetAsync(d->i32, 0, sizeof(int), st));
      CHK(launch_search(d, B, STb));
      HIPCHK(d, hipMemcpyAsync(d->h_i32, d->i32, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
      CHK(sync(d));
      if (!(d->h_i32[0] & kErrSearchCap)) break;
      if (d->mcap >= (1 << 16)) {
        d->err = "a row has too many neighbour candidates";
        return MDGPU_ERR_CAPACITY;
      }
      d->mcap *= 2;
    }
    CHK(order_rows(d));
    // scratch for both scans before either is queued (a reallocation frees
    // the buffer a queued scan would use)
    CHK(ensure_scan(d, std::max<long long>(pn, d->nslice)));
    // number of listed (row, partner) entries, for accounting: copied now,
    // read after the next wait
    exclusive_scan<int, long long>(d->row_len, d->key, pn, d->scan_scratch, d->i64, st);
    HIPCHK(d, hipMemcpyAsync(d->h_i64 + 1, d->i64, sizeof(long long), hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(k_slice_len, dim3(nblk(d->nslice)), dim3(kBlock), 0, st, d->row_len, d->slice_brick,
                       d->vslice0, d->row0, d->nrow, d->nslice, d->slen, d->slots64, d->spart);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(kBlock), 0, st, d->spart, (int)nblk(d->nslice), d->u64 + 3);
    exclusive_scan<long long, long long>(d->slots64, d->sbase, d->nslice, d->scan_scratch, d->i64, st);
    // the slots and the longest slice after one wait
    HIPCHK(d, hipMemcpyAsync(d->h_u64 + 3, d->u64 + 3, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    CHK(read_total(d, &slots));
    d->nbr_entries = d->h_i64[1];
    double smax = 0.0;
    std::memcpy(&smax, d->h_u64 + 3, sizeof(double));
    scap = (int)smax;
  } else {
    d->nbr_entries = 0;
  }
  if (slots > d->nbr_cap) {
    // kListTail readable entries past the last slice (the force kernels read ahead)
    const long long nc = slots + slots / 8 + 1024;
    CHK(dalloc(d, &d->nbr, (size_t)(nc + kListTail)));
    CHK(dalloc(d, &d->nbr_in, (size_t)(nc + kListTail)));
    d->nbr_cap = nc;
  }
  d->nbr_slots = slots;
  if (pn > 0 && scap > 0) CHK(launch_emit(d, B, STb));
  else d->list_coded = false;  // no slots to code
  CHK(check_launch(d, "neighbour list"));
  // PairList::Init: snapshot of owned + ghost positions (pairlist.cc:11-23)
  if (d->tn > 0)
    hipLaunchKernelGGL(k_copy_pos, dim3(nblk(d->tn)), dim3(kBlock), 0, st, d->pos, d->pos_old, d->tn);
  CHK(read_err_flags(d));
  d->initialized = true;
  d->buffer_length = d->margin;
  d->lifetime = 0;
  for (ReplicaState &S : d->rst) {
    S.buffer_length = d->margin;
    S.lifetime = 0;
  }
  d->vmax_valid = false;
  invalidate(d);
  return MDGPU_OK;
}

// MDManager::MakePairList for a periodic domain or an ensemble
// (mdmanager.cc:262-280).  The reference also permutes particles that cross
// a face (MakeBufferForSendingParticle, mdunit.cc:78-101); that only renames
// indices, and particles are put in brick order here anyway.  A unit builds
// its lists from the ghosts its group placed (mdgpu_unit_find_border_particles).
int make_pair_list(mdgpu_domain *d) {
  if (is_unit(d)) {
    if (!d->ghosts_ready) {
      d->err = "a unit's pair list needs its group's migration and ghosts first "
               "(mdgpu_group_make_pair_list, or the mdgpu_unit_* sequence)";
      return MDGPU_ERR_ARGUMENT;
    }
    return build_lists(d);
  }
  CHK(auto_bricks(d));
  HIPCHK(d, hipMemsetAsync(d->i32, 0, sizeof(int) * 4, d->stream));
  CHK(bin_owned(d, 1));
  CHK(place_images(d));
  return build_lists(d);
}

// max |p|^2 of a single domain; after a fused step it is the one the step's
// force launch reduced and copied to the host
int max_p2(mdgpu_domain *d, double *v2) {
  if (d->vmax_valid) {
    CHK(sync(d));
    *v2 = d->h_rep_v2[0];
    return MDGPU_OK;
  }
  const unsigned nb = nblk(d->pn);
  hipLaunchKernelGGL(k_max_p2, dim3(nb), dim3(kBlock), 0, d->stream, d->mom, d->pn, d->maxpart);
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(kBlock), 0, d->stream, d->maxpart, (int)nb, d->u64);
  return read_max(d, 0, v2);
}

// PairList::CheckByDisplacement (pairlist.cc:68-121): the passes queued (the
// outcome copied to h_i32), then read once the stream has run them.
int enqueue_displacement_check(mdgpu_domain *d) {
  hipStream_t st = d->stream;
  const int tn = d->tn;
  hipLaunchKernelGGL(k_displacement, dim3(nblk(tn)), dim3(kBlock), 0, st, d->pos, d->pos_old, tn, d->disp,
                     d->maxpart);
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(kBlock), 0, st, d->maxpart, (int)nblk(tn), d->u64 + 1);
  // the three passes run back to back, each taking the last one's result from
  // device memory, and the host reads the outcome once
  const int Nl = (tn > d->check_list_length) ? d->check_list_length : tn;
  HIPCHK(d, hipMemsetAsync(d->i32 + 1, 0, sizeof(int) * 2, st));
  if (tn > 0)
    hipLaunchKernelGGL(k_candidates, dim3(nblk(tn)), dim3(kBlock), 0, st, d->disp, tn, d->u64 + 1, d->margin,
                       d->i32 + 1, d->cand, d->check_list_length);
  if (Nl > 1) {
    const long long total = (long long)Nl * Nl;
    hipLaunchKernelGGL(k_candidate_pairs, dim3(nblk(total)), dim3(kBlock), 0, st, d->pos, d->pos_old, d->cand,
                       d->i32 + 1, Nl, d->SL * d->SL, d->CL2, d->i32 + 2);
  }
  CHK(check_launch(d, "displacement check"));
  HIPCHK(d, hipMemcpyAsync(d->h_i32, d->i32, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(d, hipMemcpyAsync(d->h_u64 + 1, d->u64 + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  return MDGPU_OK;
}

// A list that passed the displacement check takes a budget again from the
// largest displacement dmax the check measured.  No particle has moved more
// than dmax since the construction, so a pair outside the list stays beyond
// rc while 2 (dmax + the motion since) stays below the margin, which the
// budget counts down as before.  While that budget lasts the check could not
// expire the list (its candidates have moved more than margin - dmax, and
// none has while 2 dmax < margin), so the list expires on the step it would
// with a check on every step, after fewer checks and with the steps in
// between queued ahead of the host.  `dmax2` is max |dq|^2.
double budget_after_check(double margin, double dmax2) { return margin - 2.0 * std::sqrt(dmax2); }

// the check's max |dq|^2, read with its outcome
void rearm_after_check(mdgpu_domain *d) {
  double d2;
  std::memcpy(&d2, d->h_u64 + 1, sizeof d2);
  const double left = budget_after_check(d->margin, d2);
  if (left > 0.0) d->buffer_length = left;
}

int displacement_check_result(const mdgpu_domain *d, int *expired) {
  const int tn = d->tn;
  const int Nl = (tn > d->check_list_length) ? d->check_list_length : tn;
  const int count = d->h_i32[1];
  if (count > 0 && count >= Nl) {
    *expired = 1;
    return MDGPU_OK;
  }
  *expired = count > 1 && d->h_i32[2] != 0;
  return MDGPU_OK;
}

int check_by_displacement(mdgpu_domain *d, int *expired) {
  *expired = 0;
  CHK(enqueue_displacement_check(d));
  CHK(sync(d));
  return displacement_check_result(d, expired);
}

// PairList::IsPairListExpired (pairlist.cc:26-45) with
// UpdatePairListValidity (pairlist.cc:48-65)
int is_expired(mdgpu_domain *d, int *expired) {
  if (!d->initialized) {
    *expired = 1;
    return MDGPU_OK;
  }
  d->lifetime++;
  if (d->buffer_length <= 0.0 && d->prune_delta > 0.0 && d->inner_valid) {
    // the budget is spent, so the reference skips max|p|; the inner list's budget still needs it
    double v2 = 0.0;
    CHK(max_p2(d, &v2));
    d->step_vmax = std::sqrt(v2);
  }
  if (d->buffer_length > 0.0) {
    double v2 = 0.0;
    CHK(max_p2(d, &v2));
    const double vmax = std::sqrt(v2);
    d->step_vmax = vmax;
    if (vmax * 2.0 * d->dt > d->margin) {
      char b[160];
      snprintf(b, sizeof b, "Too fast particles exists. Try smaller time step (max_velocity = %f)", vmax);
      d->err = b;
      return MDGPU_ERR_TOO_FAST;
    }
    d->buffer_length = d->buffer_length - vmax * 2.0 * d->dt;
  }
  *expired = 0;
  if (d->buffer_length < 0.0) {
    if (d->disp_check) {
      CHK(check_by_displacement(d, expired));
      if (!*expired) rearm_after_check(d);
    } else {
      *expired = 1;  // rebuild cadence without the exact check: the budget alone
    }
  }
  return MDGPU_OK;
}

// The same test for each replica of an ensemble; *expired = 1 when the list
// of at least one replica has expired (the ensemble then rebuilds every
// replica's list).  The maxima of |p|^2 come from the last fused step; the
// displacement check runs, one workgroup per replica, for the replicas whose
// budget ran out.
int is_expired_ensemble(mdgpu_domain *d, int *expired) {
  if (!d->initialized) {
    *expired = 1;
    return MDGPU_OK;
  }
  hipStream_t st = d->stream;
  const int R = d->nrep;
  if (!d->vmax_valid) {
    hipLaunchKernelGGL(k_replica_max_p2, dim3(R), dim3(kBlock), 0, st, d->mom, d->reps, d->rep_v2);
    HIPCHK(d, hipMemcpyAsync(d->h_rep_v2, d->rep_v2, R * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  CHK(sync(d));
  double v2max = 0.0;
  for (int r = 0; r < R; r++) v2max = std::max(v2max, d->h_rep_v2[r]);
  d->step_vmax = std::sqrt(v2max);
  int nflag = 0;
  for (int r = 0; r < R; r++) {
    ReplicaState &S = d->rst[r];
    S.lifetime++;
    if (S.buffer_length > 0.0) {
      const double vmax = std::sqrt(d->h_rep_v2[r]);
      if (vmax * 2.0 * d->dt > d->margin) {
        char b[200];
        snprintf(b, sizeof b, "Too fast particles exists. Try smaller time step (replica %d, max_velocity = %f)", r,
                 vmax);
        d->err = b;
        return MDGPU_ERR_TOO_FAST;
      }
      S.buffer_length = S.buffer_length - vmax * 2.0 * d->dt;
    }
    if (S.buffer_length < 0.0) d->h_which[nflag++] = r;
  }
  *expired = 0;
  if (nflag == 0) return MDGPU_OK;
  if (!d->disp_check) {  // the budget alone: a spent budget expires the list
    *expired = 1;
    return MDGPU_OK;
  }
  HIPCHK(d, hipMemcpyAsync(d->which, d->h_which, nflag * sizeof(int), hipMemcpyHostToDevice, st));
  const int cap = d->check_list_length;
  hipLaunchKernelGGL(k_zero_check, dim3(nblk(nflag)), dim3(kBlock), 0, st, d->chk_dmax, d->chk_cnt, d->rep_exp,
                     nflag);
  hipLaunchKernelGGL(k_check_dmax, dim3(nflag * kCheckBlocks), dim3(kBlock), 0, st, d->pos, d->pos_old, d->which,
                     d->reps, d->grng, d->pn, d->disp, d->chk_dmax);
  hipLaunchKernelGGL(k_check_cand, dim3(nflag * kCheckBlocks), dim3(kBlock), 0, st, d->disp, d->which, d->reps,
                     d->grng, d->pn, d->margin, cap, d->chk_dmax, d->chk_cnt, d->chk_cand);
  hipLaunchKernelGGL(k_check_pairs, dim3(nflag, kPairBlocks), dim3(kBlock), 0, st, d->pos, d->pos_old, d->which,
                     d->reps, d->grng, cap, d->SL * d->SL, d->CL2, d->chk_cnt, d->chk_cand, d->rep_exp);
  CHK(check_launch(d, "displacement check"));
  HIPCHK(d, hipMemcpyAsync(d->h_rep_exp, d->rep_exp, nflag * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(d, hipMemcpyAsync(d->h_chk_dmax, d->chk_dmax, nflag * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           st));
  CHK(sync(d));
  for (int k = 0; k < nflag; k++)
    if (d->h_rep_exp[k]) *expired = 1;
  if (*expired) return MDGPU_OK;
  // each replica whose list passed takes a budget again (budget_after_check)
  for (int k = 0; k < nflag; k++) {
    double d2;
    std::memcpy(&d2, &d->h_chk_dmax[k], sizeof d2);
    const double left = budget_after_check(d->margin, d2);
    if (left > 0.0) d->rst[d->h_which[k]].buffer_length = left;
  }
  return MDGPU_OK;
}

int check_expired(mdgpu_domain *d, int *expired) {
  return d->nrep == 1 ? is_expired(d, expired) : is_expired_ensemble(d, expired);
}

ForceConsts fconsts(const mdgpu_domain *d) {
  ForceConsts fc;
  fc.CL2 = d->CL2;
  fc.A = 24.0 * d->dt;
  fc.B = 48.0 * d->dt;
  fc.C = 8.0 * d->C2 * d->dt;
  fc.dt2 = d->dt * 0.5;
  return fc;
}

// FP32 constants of the mixed kick.  C is chosen so that the kick evaluated in
// FP32 with the rounded A, B at r^2 = rc^2 (rounded) is zero to one rounding:
// the smoothed force vanishes at the cutoff in the arithmetic the kernel uses,
// not only in exact arithmetic (8 C2 dt rounded to FP32 would leave a step in
// the force where pairs cross rc).
ForceConstsF fconsts32(const mdgpu_domain *d) {
  ForceConstsF f;
  f.CL2 = (float)d->CL2;
  f.A = (float)(24.0 * d->dt);
  f.B = (float)(48.0 * d->dt);
  const double ir2 = 1.0 / (double)f.CL2;
  const double ir6 = ir2 * ir2 * ir2;
  f.C = (float)(-(ir2 * ir6) * ((double)f.A - (double)f.B * ir6));
  return f;
}

// The list as the brick kernels read it; lds_cap = staged entries per brick
// that fit the LDS budget (bricks with more read partners from global memory).
BrickList brick_list(const mdgpu_domain *d, int bytes_per_entry, bool inner = false) {
  BrickList bl;
  bl.nbr = inner ? d->nbr_in : d->nbr;
  bl.code = !list_codes(d) ? nullptr : inner ? d->ncode_in : d->ncode;
  bl.sbase = d->sbase;
  bl.slen = inner ? d->slen_in : d->slen;
  bl.glist = d->glist;
  bl.gbase = d->gbase;
  bl.row0 = d->row0;
  bl.nrow = d->nrow;
  bl.vslice0 = d->vslice0;
  bl.brep = d->brep;
  bl.reps = d->reps;
  bl.ghost = d->pos;  // the halo refresh writes the images there only
  bl.pn = d->pn;
  bl.lds_cap = lds_cap_for(d, bytes_per_entry, force_lds_kb(d));
  return bl;
}

template <int W, int U>
void launch_fp64(mdgpu_domain *d, const double4 *src, double4 *pos_out, int epi, const ForceConsts &fc,
                 const StepClock &sc, double *vb, const NextDrift &nd, bool inner) {
  const BrickList bl = brick_list(d, 3 * sizeof(double), inner);
  const size_t shm = (size_t)bl.lds_cap * 3 * sizeof(double);
  allow_lds(k_force_brick_fp64<W, U>, shm);
  hipLaunchKernelGGL((k_force_brick_fp64<W, U>), dim3(d->nbrick), dim3(kWave * W), shm, d->stream, src, d->mom,
                     pos_out, bl, fc, epi, sc, d->id, vb, nd);
}

// Staged entries per buffer of the pipelined kick (two buffers in the LDS of one
// workgroup), pad entry included.
int pipe_cap(const mdgpu_domain *d) {
  const int kb = d->lds_kb ? std::min(d->lds_kb, d->lds_max_kb) : d->lds_max_kb;
  const int budget = (kb * 1024 / (2 * 3 * (int)sizeof(double))) & ~1;
  // even: the second buffer's 16-byte xy entries start 16-byte aligned
  return std::max(2, std::min((d->max_staged + 2) & ~1, budget));
}

template <int W, int U>
void launch_fp64_pipe(mdgpu_domain *d, const double4 *src, double4 *pos_out, int epi, const ForceConsts &fc,
                      const StepClock &sc, double *vb, const NextDrift &nd, bool inner) {
  constexpr int P = 2;  // producer wavefronts
  static_assert(W + P <= 16, "at most 1024 threads per workgroup");
  BrickList bl = brick_list(d, 3 * sizeof(double), inner);
  bl.lds_cap = pipe_cap(d);
  const size_t shm = 2 * (size_t)bl.lds_cap * 3 * sizeof(double);
  allow_lds(k_force_pipe_fp64<W, U, P>, shm);
  const int grid = std::max(1, std::min(d->nbrick, d->num_cu));
  hipLaunchKernelGGL((k_force_pipe_fp64<W, U, P>), dim3(grid), dim3(kWave * (W + P)), shm, d->stream, src,
                     d->mom, pos_out, bl, fc, epi, sc, d->id, vb, nd, d->nbrick);
}

template <int W, int U, int K>
void launch_mixed(mdgpu_domain *d, const double4 *src, const float4 *rel, double4 *pos_out, int epi,
                  const ForceConstsF &f32, double dt2, const StepClock &sc, double *vb, const NextDrift &nd,
                  bool inner) {
  const BrickList bl = brick_list(d, sizeof(float4), inner);
  const size_t shm = (size_t)bl.lds_cap * sizeof(float4);
  allow_lds(k_force_brick_mixed<W, U, K>, shm);
  hipLaunchKernelGGL((k_force_brick_mixed<W, U, K>), dim3(d->nbrick), dim3(kWave * W), shm, d->stream, rel, src,
                     d->mom, pos_out, bl, f32, dt2, epi, sc, d->id, vb, nd);
}

// the mixed kick in the brick-centred frame (no FP32 copies, no cell codes)
template <int W, int U, int K>
void launch_mixed_brick(mdgpu_domain *d, const double4 *src, double4 *pos_out, int epi, const ForceConstsF &f32,
                        double dt2, const StepClock &sc, double *vb, const NextDrift &nd, bool inner) {
  const BrickList bl = brick_list(d, sizeof(float4), inner);
  const size_t shm = (size_t)bl.lds_cap * sizeof(float4);
  const NextDrift nd64{nd.pos, nullptr, nd.cell, nd.org};
  allow_lds(k_force_brick_mixed_brick<W, U, K>, shm);
  hipLaunchKernelGGL((k_force_brick_mixed_brick<W, U, K>), dim3(d->nbrick), dim3(kWave * W), shm, d->stream, src,
                     d->mom, pos_out, bl, f32, dt2, epi, sc, d->id, vb, nd64);
}

template <int W, int K>
void launch_mixed_sum(mdgpu_domain *d, const double4 *src, const float4 *rel, double4 *pos_out, int epi,
                      const ForceConstsF &f32, double dt2, const StepClock &sc, double *vb, const NextDrift &nd,
                      bool inner) {
  if (d->mixed_frame == 1) {
    if (unroll_of(d) == 32) launch_mixed_brick<W, 4, K>(d, src, pos_out, epi, f32, dt2, sc, vb, nd, inner);
    else if (unroll_of(d) == 16) launch_mixed_brick<W, 2, K>(d, src, pos_out, epi, f32, dt2, sc, vb, nd, inner);
    else launch_mixed_brick<W, 1, K>(d, src, pos_out, epi, f32, dt2, sc, vb, nd, inner);
    return;
  }
  if (unroll_of(d) == 32) launch_mixed<W, 4, K>(d, src, rel, pos_out, epi, f32, dt2, sc, vb, nd, inner);
  else if (unroll_of(d) == 16) launch_mixed<W, 2, K>(d, src, rel, pos_out, epi, f32, dt2, sc, vb, nd, inner);
  else launch_mixed<W, 1, K>(d, src, rel, pos_out, epi, f32, dt2, sc, vb, nd, inner);
}

template <int W>
void launch_force_w(mdgpu_domain *d, const double4 *src, const float4 *rel, double4 *pos_out, int epi,
                    const ForceConsts &fc, const ForceConstsF &f32, const StepClock &sc, double *vb,
                    const NextDrift &nd, bool in) {
  // groups of kGroup list indices in flight: unroll / kGroup (1, 2 or 4)
  if (d->precision == MDGPU_FP64 && d->pipeline && W >= 4) {
    // compute wavefronts: 14 with 16 requested (two producers make 16), else as requested
    constexpr int WP = W > 14 ? 14 : W;
    if (unroll_of(d) == 16) launch_fp64_pipe<WP, 2>(d, src, pos_out, epi, fc, sc, vb, nd, in);
    else launch_fp64_pipe<WP, 1>(d, src, pos_out, epi, fc, sc, vb, nd, in);
  } else if (d->precision == MDGPU_FP64) {
    if (unroll_of(d) == 32) launch_fp64<W, 4>(d, src, pos_out, epi, fc, sc, vb, nd, in);
    else if (unroll_of(d) == 16) launch_fp64<W, 2>(d, src, pos_out, epi, fc, sc, vb, nd, in);
    else launch_fp64<W, 1>(d, src, pos_out, epi, fc, sc, vb, nd, in);
  } else {
    switch (d->kahan) {
      case kSumPlain: launch_mixed_sum<W, kSumPlain>(d, src, rel, pos_out, epi, f32, fc.dt2, sc, vb, nd, in); break;
      case kSumKahan: launch_mixed_sum<W, kSumKahan>(d, src, rel, pos_out, epi, f32, fc.dt2, sc, vb, nd, in); break;
      default: launch_mixed_sum<W, kSumGroup64>(d, src, rel, pos_out, epi, f32, fc.dt2, sc, vb, nd, in); break;
    }
  }
}

// One pair-force launch.  Rows and partners (owned, and ghosts at
// [pn, pn + ng)) are read from `src` (`rel` in FP32); `pos_out` receives the
// drifted (or copied) owned positions; `nd` (optional) the next step's first
// half drift.  With want_vmax, each replica's max |p|^2 goes to rep_v2 and its
// host mirror.  `clock` set: the Langevin step is read from it and advanced
// (a launch that is captured and replayed).
int launch_force(mdgpu_domain *d, const double4 *src, const float4 *rel, double4 *pos_out, int epilogue,
                 bool want_vmax, const NextDrift &nd, unsigned long long *clock = nullptr, bool inner = false,
                 RunGate *gate = nullptr) {
  if (d->pn == 0) return MDGPU_OK;
  hipStream_t st = d->stream;
  const ForceConsts fc = fconsts(d);
  const ForceConstsF f32 = fconsts32(d);
  const StepClock sc{d->dt, d->step, clock, gate};
  double *vb = want_vmax ? d->bpart : nullptr;
  const int ev0 = d->force_launches * 2;
  const bool ev = d->timing && clock == nullptr && ev0 + 1 < (int)d->ev.size();
  if (ev) HIPCHK(d, hipEventRecord(d->ev[ev0], st));
  switch (d->waves) {
    case 2: launch_force_w<2>(d, src, rel, pos_out, epilogue, fc, f32, sc, vb, nd, inner); break;
    case 8: launch_force_w<8>(d, src, rel, pos_out, epilogue, fc, f32, sc, vb, nd, inner); break;
    case 16: launch_force_w<16>(d, src, rel, pos_out, epilogue, fc, f32, sc, vb, nd, inner); break;
    default: launch_force_w<4>(d, src, rel, pos_out, epilogue, fc, f32, sc, vb, nd, inner); break;
  }
  if (ev) {
    HIPCHK(d, hipEventRecord(d->ev[ev0 + 1], st));
    d->force_launches++;
  }
  if (want_vmax) {
    hipLaunchKernelGGL(k_replica_max, dim3(d->nrep), dim3(kBlock), 0, st, d->bpart, d->reps, d->rep_v2, clock,
                       gate);
    // a run's steps keep the maxima on the device for the next step's gate
    if (!gate) HIPCHK(d, hipMemcpyAsync(d->h_rep_v2, d->rep_v2, d->nrep * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  return check_launch(d, "force");
}

int kinetic(mdgpu_domain *d, double *ke) {
  const unsigned nb = nblk(d->pn);
  CHK(ensure_partial(d, (int)nb));
  if (d->pn > 0) {
    hipLaunchKernelGGL(k_kinetic, dim3(nb), dim3(kBlock), 0, d->stream, d->mom, d->pn, d->partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, d->stream, d->partial, (int)nb, 1, d->d3);
    HIPCHK(d, hipMemcpyAsync(d->h_d3, d->d3, sizeof(double), hipMemcpyDeviceToHost, d->stream));
    CHK(sync(d));
    *ke = 0.0 + d->h_d3[0];
  } else {
    *ke = 0.0;
  }
  return MDGPU_OK;
}

int temperature(mdgpu_domain *d, double *t) {
  double ke;
  CHK(kinetic(d, &ke));
  *t = (ke / (double)d->pn) / 1.5;
  return MDGPU_OK;
}

int heatbath_zeta(mdgpu_domain *d, double t) {
  const double dt2 = d->dt * 0.5;
  const double t1 = (t - d->T_aim) / (d->tau * d->tau);
  d->zeta += t1 * dt2;
  return MDGPU_OK;
}

int heatbath_momenta(mdgpu_domain *d) {
  const double e1 = std::exp(-(d->dt * 0.5) * d->zeta);
  if (d->pn > 0)
    hipLaunchKernelGGL(k_scale_momenta, dim3(nblk(d->pn)), dim3(kBlock), 0, d->stream, d->mom, d->pn, e1);
  d->vmax_valid = false;
  return check_launch(d, "heatbath momenta");
}

// Nose-Hoover for every replica of an ensemble, each with its own temperature,
// target and zeta (MDManager::CalculateNoseHoover, mdmanager.cc:160-175, per system).
int replica_temperatures(mdgpu_domain *d, std::vector<double> &t) {
  const int R = d->nrep;
  hipLaunchKernelGGL(k_replica_kinetic, dim3(R), dim3(kBlock), 0, d->stream, d->mom, d->reps, d->rep_out);
  CHK(check_launch(d, "replica kinetic energy"));
  HIPCHK(d, hipMemcpyAsync(d->h_rep_out, d->rep_out, R * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  CHK(sync(d));
  t.resize(R);
  for (int r = 0; r < R; r++) t[r] = (d->h_rep_out[r] / (double)d->rh[r].pn) / 1.5;
  return MDGPU_OK;
}

void replica_heatbath_zeta(mdgpu_domain *d, const std::vector<double> &t) {
  const double dt2 = d->dt * 0.5;
  for (int r = 0; r < d->nrep; r++) {
    ReplicaState &S = d->rst[r];
    const double t1 = (t[r] - S.T) / (d->tau * d->tau);
    S.zeta += t1 * dt2;
  }
}

int replica_heatbath_momenta(mdgpu_domain *d) {
  for (int r = 0; r < d->nrep; r++) d->h_rep_factor[r] = std::exp(-(d->dt * 0.5) * d->rst[r].zeta);
  HIPCHK(d, hipMemcpyAsync(d->rep_factor, d->h_rep_factor, d->nrep * sizeof(double), hipMemcpyHostToDevice,
                           d->stream));
  if (d->pn > 0)
    hipLaunchKernelGGL(k_scale_momenta_replicas, dim3(nblk(d->pn)), dim3(kBlock), 0, d->stream, d->mom, d->rep,
                       d->pn, d->rep_factor);
  d->vmax_valid = false;
  CHK(check_launch(d, "replica heat bath"));
  // the pinned factors are read by the copy above before the next write: the
  // next call follows a temperature read, which waits for the stream
  return MDGPU_OK;
}

int drift(mdgpu_domain *d) {
  if (d->pn > 0)
    hipLaunchKernelGGL(k_drift, dim3(nblk(d->pn)), dim3(kBlock), 0, d->stream, d->pos, d->mom, d->pn,
                       d->dt * 0.5);
  return check_launch(d, "drift");
}

int refresh(mdgpu_domain *d) {
  if (d->ng > 0)
    hipLaunchKernelGGL(k_refresh, dim3(nblk(d->ng)), dim3(kBlock), 0, d->stream, d->pos, d->gsrc, d->gcode, d->pn,
                       d->ng, d->reps);
  return check_launch(d, "refresh");
}

// CalculateForce on the current positions (owned and the ghosts of the last
// halo refresh, all in `pos`)
int force_in_place(mdgpu_domain *d) {
  if (rel_frame(d) && d->tn > 0)
    hipLaunchKernelGGL(k_to_rel, dim3(nblk(d->tn)), dim3(kBlock), 0, d->stream, d->pos, d->cell, d->pn, d->tn,
                       d->rep, d->gcode, d->reps, d->org, d->rel);
  invalidate(d);
  CHK(launch_force(d, d->pos, d->rel, d->pos, kKickOnly, false, NextDrift{nullptr, nullptr, nullptr, nullptr}));
  d->vmax_valid = false;
  return MDGPU_OK;
}

// The launches of one fused step once pos_tmp (rel_tmp) holds the owned
// particles' first half drift: ghost refresh from those positions, the kick
// whose epilogue writes the final positions to pos and the next step's first
// half drift to pos_next (rel_next), and each replica's max |p|^2, copied to
// the host.  With `clock` the launches read and advance the device step
// counter (they are being captured into a graph).
template <int W>
void launch_prune_w(mdgpu_domain *d) {
  const BrickList bl = brick_list(d, 3 * sizeof(double));
  const size_t shm = (size_t)bl.lds_cap * 3 * sizeof(double);
  const double rin = d->rc + d->prune_delta;
  allow_lds(k_prune_brick<W>, shm);
  hipLaunchKernelGGL(k_prune_brick<W>, dim3(d->nbrick), dim3(kWave * W), shm, d->stream, d->pos_tmp, bl, rin * rin,
                     d->nbr_in, d->ncode_in, d->slen_in);
}

// The inner list from the full list at the half-drifted positions (pos_tmp,
// ghosts refreshed): every listed pair within rc + prune_delta.
int launch_prune(mdgpu_domain *d) {
  if (d->pn == 0) return MDGPU_OK;
  switch (d->waves) {
    case 2: launch_prune_w<2>(d); break;
    case 8: launch_prune_w<8>(d); break;
    case 16: launch_prune_w<16>(d); break;
    default: launch_prune_w<4>(d); break;
  }
  d->prunes++;
  return check_launch(d, "prune");
}

// The fused step's launches.  gated: a step of a run, headed by the validity
// gate (k_gate), whose launches do nothing once the run has halted.
int enqueue_fused(mdgpu_domain *d, unsigned long long *clock, bool prune = false, bool gated = false) {
  hipStream_t st = d->stream;
  const bool mixed = rel_frame(d);
  RunGate *gate = gated ? d->gate : nullptr;
  if (gated)
    hipLaunchKernelGGL(k_gate, dim3(1), dim3(kBlock), 0, st, d->gate, d->gbudget, d->glife, d->rep_v2, d->nrep, d->dt,
                       d->margin);
  if (d->ng > 0) {
    if (mixed)
      hipLaunchKernelGGL((k_refresh_from_drifted<true>), dim3(nblk(d->ng)), dim3(kBlock), 0, st, d->pos, d->pos_tmp,
                         d->gsrc, d->gcode, d->pn, d->ng, d->reps, d->cell, d->org, d->rel_tmp, gate);
    else
      hipLaunchKernelGGL((k_refresh_from_drifted<false>), dim3(nblk(d->ng)), dim3(kBlock), 0, st, d->pos,
                         d->pos_tmp, d->gsrc, d->gcode, d->pn, d->ng, d->reps, d->cell, d->org, d->rel_tmp, gate);
  }
  CHK(check_launch(d, "refresh"));
  if (prune) CHK(launch_prune(d));
  const int epi = d->thermostat == MDGPU_LANGEVIN ? kKickLangevin : kKickDrift;
  const NextDrift nd{d->pos_next, mixed ? d->rel_next : nullptr, d->cell, d->org};
  // a run's steps leave the final owned positions out (run_gated writes them
  // after the run); a timed step does too and writes them after its timed kick,
  // so that the events time the kernel the runs launch
  const bool timed = d->timing && clock == nullptr && !gated;
  CHK(launch_force(d, d->pos_tmp, d->rel_tmp, gated || timed ? nullptr : d->pos, epi, true, nd, clock,
                   d->prune_delta > 0.0, gate));
  if (timed && d->pn > 0) {
    hipLaunchKernelGGL(k_final_positions, dim3(nblk(d->pn)), dim3(kBlock), 0, st, d->pos, d->pos_tmp, d->mom, d->pn,
                       d->dt * 0.5, d->thermostat == MDGPU_LANGEVIN ? 0 : 1);
    CHK(check_launch(d, "final positions"));
  }
  return MDGPU_OK;
}

void drop_graph(StepGraph &g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
  g.key = nullptr;
  g.epoch = -1;
}

// enqueue_fused replayed from a graph.  A graph is captured for the current
// buffers: one per parity of the swapped position buffers, both recaptured
// after anything that changes what the launches read (a construction,
// reallocation, option or thermostat change).
int launch_fused_graph(mdgpu_domain *d, bool gated = false) {
  StepGraph *gs = gated ? d->gated : d->graphs;
  StepGraph *g = nullptr;
  for (int k = 0; k < 2; k++)
    if (gs[k].exec && gs[k].epoch == d->epoch && gs[k].key == (const void *)d->pos_tmp) g = &gs[k];
  if (!g) {
    g = &gs[0];
    for (int k = 0; k < 2; k++)
      if (!gs[k].exec || gs[k].epoch != d->epoch) {
        g = &gs[k];
        break;
      }
    drop_graph(*g);
    hipGraph_t graph = nullptr;
    HIPCHK(d, hipStreamBeginCapture(d->stream, hipStreamCaptureModeRelaxed));
    const int r = enqueue_fused(d, d->clock, false, gated);
    const hipError_t e = hipStreamEndCapture(d->stream, &graph);
    if (r != MDGPU_OK) {
      if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return r;
    }
    if (e != hipSuccess) {
      d->err = std::string("step capture: ") + hipGetErrorString(e);
      return MDGPU_ERR_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      g->exec = nullptr;
      d->err = std::string("step graph: ") + hipGetErrorString(ei);
      return MDGPU_ERR_HIP;
    }
    g->key = d->pos_tmp;
    g->epoch = d->epoch;
  }
  HIPCHK(d, hipGraphLaunch(g->exec, d->stream));
  return MDGPU_OK;
}

int rebuild_timed(mdgpu_domain *d) {
  hipStream_t st = d->stream;
  if (d->timing) {
    hipEvent_t a, b;
    HIPCHK(d, hipEventCreate(&a));
    HIPCHK(d, hipEventCreate(&b));
    HIPCHK(d, hipEventRecord(a, st));
    CHK(make_pair_list(d));
    HIPCHK(d, hipEventRecord(b, st));
    HIPCHK(d, hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(d, hipEventElapsedTime(&ms, a, b));
    d->rebuild_ms += ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  } else {
    CHK(make_pair_list(d));
  }
  d->rebuild_count++;
  return MDGPU_OK;
}

// MDManager::Calculate (mdmanager.cc:116-148): one step of every replica
int calculate_step(mdgpu_domain *d) {
  hipStream_t st = d->stream;
  int expired = 0;
  CHK(check_expired(d, &expired));
  if (expired) CHK(rebuild_timed(d));
  if (d->thermostat == MDGPU_NOSEHOOVER && d->nrep > 1) {
    // CalculateNoseHoover (mdmanager.cc:160-175) in every replica
    invalidate(d);
    CHK(drift(d));
    CHK(refresh(d));
    std::vector<double> t;
    CHK(replica_temperatures(d, t));
    replica_heatbath_zeta(d, t);
    CHK(replica_heatbath_momenta(d));
    CHK(force_in_place(d));
    CHK(replica_heatbath_momenta(d));
    CHK(replica_temperatures(d, t));
    replica_heatbath_zeta(d, t);
    CHK(drift(d));
  } else if (d->thermostat == MDGPU_NOSEHOOVER) {
    // CalculateNoseHoover (mdmanager.cc:160-175), single domain
    invalidate(d);
    CHK(drift(d));
    CHK(refresh(d));
    double t;
    CHK(temperature(d, &t));
    CHK(heatbath_zeta(d, t));
    CHK(heatbath_momenta(d));
    CHK(force_in_place(d));
    CHK(heatbath_momenta(d));
    CHK(temperature(d, &t));
    CHK(heatbath_zeta(d, t));
    CHK(drift(d));
  } else {
    // fused: [drift into pos_tmp +] halo refresh, then kick (+ drift or Langevin)
    // whose epilogue also writes the next step's first half drift into pos_next
    const bool mixed = rel_frame(d);
    if (!d->mid_valid) {
      const int tot = d->pn + d->ng;
      if (tot > 0) {
        if (mixed)
          hipLaunchKernelGGL((k_drift_refresh<true>), dim3(nblk(tot)), dim3(kBlock), 0, st, d->pos, d->mom,
                             d->pos_tmp, d->gsrc, d->gcode, d->pn, d->ng, d->dt * 0.5, d->rep, d->reps, d->cell,
                             d->org, d->rel_tmp);
        else
          hipLaunchKernelGGL((k_drift_refresh<false>), dim3(nblk(tot)), dim3(kBlock), 0, st, d->pos, d->mom,
                             d->pos_tmp, d->gsrc, d->gcode, d->pn, d->ng, d->dt * 0.5, d->rep, d->reps, d->cell,
                             d->org, d->rel_tmp);
      }
      CHK(check_launch(d, "drift+refresh"));
    }
    // the inner list: pruned again once the motion since the last pruning
    // could have brought a pair from beyond rc + delta to within rc
    bool prune = false;
    if (d->prune_delta > 0.0) {
      if (d->inner_valid) d->inner_budget -= d->step_vmax * 2.0 * d->dt;
      if (!d->inner_valid || !(d->inner_budget > 0.0)) {
        prune = true;
        d->inner_valid = true;
        d->inner_budget = d->prune_delta;
      }
    }
    if (d->use_graph && !d->timing && !prune) {
      if (d->clock_dirty) {
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, d->clock, (unsigned long long)d->step);
        d->clock_dirty = false;
      }
      CHK(launch_fused_graph(d));
    } else {
      CHK(enqueue_fused(d, nullptr, prune));
      d->clock_dirty = true;
    }
    std::swap(d->pos_tmp, d->pos_next);
    if (mixed) std::swap(d->rel_tmp, d->rel_next);
    d->mid_valid = true;
    d->vmax_valid = true;
  }
  d->sim_time += d->dt;
  d->step++;
  return MDGPU_OK;
}

// Whether the next steps may be queued without waiting for the host: fused
// NVE or Langevin steps (replayed from graphs or launched), no inner list, and the last step
// a fused one (its maxima of |p|^2 and the next first half drift are on the
// device).
bool can_run_ahead(const mdgpu_domain *d) {
  return d->run_ahead > 0 && !d->timing && d->thermostat != MDGPU_NOSEHOOVER &&
         !(d->prune_delta > 0.0) && d->initialized && d->mid_valid && d->vmax_valid && d->pn > 0;
}

double &budget_of(mdgpu_domain *d, int r) { return d->nrep == 1 ? d->buffer_length : d->rst[r].buffer_length; }
int &lifetime_of(mdgpu_domain *d, int r) { return d->nrep == 1 ? d->lifetime : d->rst[r].lifetime; }

// Steps every validity budget lasts at the last maxima of |p| (0: the next
// step needs the host).  The maxima change from step to step, so a run may
// still halt earlier, or end before the budgets do.
int run_length(mdgpu_domain *d, int limit, int *k) {
  *k = 0;
  for (int r = 0; r < d->nrep; r++)
    if (!(budget_of(d, r) > 0.0)) return MDGPU_OK;  // a displacement check is due: no wait for the maxima
  CHK(sync(d));  // the last step's maxima
  double steps = (double)limit;
  for (int r = 0; r < d->nrep; r++) {
    const double b = budget_of(d, r);
    const double v = std::sqrt(d->h_rep_v2[r]);
    if (v > 0.0) steps = std::min(steps, b / (v * 2.0 * d->dt));
  }
  if (steps >= 1.0) *k = (int)steps;
  return MDGPU_OK;
}

// A run of k fused steps queued at once, each headed by k_gate; *done
// receives the steps completed, fewer than k when a step's validity test
// needs the host.  The budgets and lifetimes go to the device before the run
// and come back after it, as do the last maxima of |p|^2.
int run_gated(mdgpu_domain *d, int k, int *done) {
  hipStream_t st = d->stream;
  const int R = d->nrep;
  const bool mixed = rel_frame(d);
  for (int r = 0; r < R; r++) {
    d->h_gbudget[r] = budget_of(d, r);
    d->h_glife[r] = lifetime_of(d, r);
  }
  d->h_gate->halted = 0;
  d->h_gate->done = 0;
  HIPCHK(d, hipMemcpyAsync(d->gbudget, d->h_gbudget, R * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(d, hipMemcpyAsync(d->glife, d->h_glife, R * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(d, hipMemcpyAsync(d->gate, d->h_gate, sizeof(RunGate), hipMemcpyHostToDevice, st));
  if (d->clock_dirty) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, d->clock, (unsigned long long)d->step);
    d->clock_dirty = false;
  }
  for (int j = 0; j < k; j++) {
    if (d->use_graph) CHK(launch_fused_graph(d, true));
    else CHK(enqueue_fused(d, d->clock, false, true));  // the same launches, each queued directly
    std::swap(d->pos_tmp, d->pos_next);
    if (mixed) std::swap(d->rel_tmp, d->rel_next);
  }
  HIPCHK(d, hipMemcpyAsync(d->h_gate, d->gate, sizeof(RunGate), hipMemcpyDeviceToHost, st));
  HIPCHK(d, hipMemcpyAsync(d->h_gbudget, d->gbudget, R * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(d, hipMemcpyAsync(d->h_glife, d->glife, R * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(d, hipMemcpyAsync(d->h_rep_v2, d->rep_v2, R * sizeof(double), hipMemcpyDeviceToHost, st));
  CHK(sync(d));
  const int n = d->h_gate->done;
  if (n < 0 || n > k) {
    d->err = "step run: inconsistent count of completed steps";
    return MDGPU_ERR_HIP;
  }
  if ((k - n) % 2) {  // the halted steps' buffer swaps did not happen on the device
    std::swap(d->pos_tmp, d->pos_next);
    if (mixed) std::swap(d->rel_tmp, d->rel_next);
  }
  for (int r = 0; r < R; r++) {
    budget_of(d, r) = d->h_gbudget[r];
    lifetime_of(d, r) = d->h_glife[r];
  }
  if (n > 0 && d->pn > 0) {
    // the final owned positions of the last completed step, whose first half
    // drift is now in pos_next (the next step's is in pos_tmp)
    hipLaunchKernelGGL(k_final_positions, dim3(nblk(d->pn)), dim3(kBlock), 0, st, d->pos, d->pos_next, d->mom,
                       d->pn, d->dt * 0.5, d->thermostat == MDGPU_LANGEVIN ? 0 : 1);
    CHK(check_launch(d, "final positions"));
  }
  for (int j = 0; j < n; j++) {
    d->sim_time += d->dt;
    d->step++;
  }
  *done = n;
  return MDGPU_OK;
}

template <int W>
void launch_observe_w(mdgpu_domain *d, const BrickList &bl, size_t shm, double *partial) {
  allow_lds(k_observe_brick<W>, shm);
  hipLaunchKernelGGL(k_observe_brick<W>, dim3(d->nbrick), dim3(kWave * W), shm, d->stream, d->pos,
                     d->mom, bl, d->CL2, d->C2, d->C0, partial);
}

// The observer kernel over every brick: KE, PE and virial per brick into
// partial[0, 3 nbrick).
void launch_observe(mdgpu_domain *d, double *partial) {
  const BrickList bl = brick_list(d, 3 * sizeof(double));
  const size_t shm = (size_t)bl.lds_cap * 3 * sizeof(double);
  switch (d->waves) {
    case 2: launch_observe_w<2>(d, bl, shm, partial); break;
    case 8: launch_observe_w<8>(d, bl, shm, partial); break;
    case 16: launch_observe_w<16>(d, bl, shm, partial); break;
    default: launch_observe_w<4>(d, bl, shm, partial); break;
  }
}

// KE, PE and virial sums of a single domain: per-brick partial sums into
// partial[0, 3 nbrick), reduced in fixed order into out_dev[0, 3), copied to
// the host.  Allocates nothing.
int observe_into(mdgpu_domain *d, double *partial, double *out_dev, double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
  if (d->pn == 0) return MDGPU_OK;
  const int nb = d->nbrick;
  launch_observe(d, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, d->stream, partial, nb, 3, out_dev);
  CHK(check_launch(d, "observe"));
  HIPCHK(d, hipMemcpyAsync(d->h_d3, out_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  CHK(sync(d));
  std::memcpy(out, d->h_d3, 3 * sizeof(double));
  return MDGPU_OK;
}

}  // namespace

// ======================================================================= C ABI
namespace {

// The address of `scratch` on the current device (`device`), looked up once per
// device and kept: every GPU holds its own copy of the array.
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxScratchDevices] = {};

hipError_t scratch_address(int device, double **out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = static_cast<double *>(ptr);
  }
  *out = g_scratch_dev[device];
  return hipSuccess;
}

// A domain of nrep replicas with the common parameters of prm (dt, rc,
// margin, check list length, precision, device); rl[r] gives replica r's box
// and Langevin parameters (nullptr: one replica with box prm->L).
// A unit (`unit`) has the box [unit->s, unit->e) of the whole box prm->L and
// queues its work on the group's stream `shared`.
int create_domain(const mdgpu_params *prm, int nrep, const mdgpu_replica *rl, const int *counts,
                  mdgpu_domain **out, const mdgpu_unit *unit = nullptr, hipStream_t shared = nullptr) {
  *out = nullptr;
  mdgpu_domain *d = new mdgpu_domain();
  d->prm = *prm;
  d->device = prm->device;
  if (hipSetDevice(prm->device) != hipSuccess) {
    delete d;
    return MDGPU_ERR_HIP;
  }
  if (shared) {
    d->stream = shared;
    d->own_stream = false;
  } else if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    delete d;
    return MDGPU_ERR_HIP;
  }
  {
    int shm = 0;
    if (hipDeviceGetAttribute(&shm, hipDeviceAttributeMaxSharedMemoryPerBlock, prm->device) == hipSuccess &&
        shm > 8 * 1024)
      d->lds_max_kb = shm / 1024 - 4;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, prm->device) == hipSuccess && ncu > 0)
      d->num_cu = ncu;
  }
  for (int k = 0; k < 3; k++) {
    d->L[k] = rl ? rl[0].L[k] : prm->L[k];
    if (unit) {
      d->rs[k] = unit->s[k];
      d->re[k] = unit->e[k];
      d->upos[k] = unit->pos[k];
    } else {
      // single domain of a 1x1x1 grid: s = L*0, e = s + L (mdunit.cc:17-24)
      d->rs[k] = d->L[k] * 0.0;
      d->re[k] = d->rs[k] + d->L[k];
    }
  }
  d->dt = prm->dt;
  d->rc = prm->rc;
  d->margin = prm->margin;
  d->SL = prm->rc + prm->margin;
  d->S2 = d->SL * d->SL;
  d->CL2 = prm->rc * prm->rc;
  d->check_list_length = prm->check_list_length > 0 ? prm->check_list_length : 500;
  d->precision = prm->precision;
  // Variables::Variables (variables.cc:16-28)
  const double s2 = 1.0 / (prm->rc * prm->rc);
  const double s6 = s2 * s2 * s2;
  const double s8 = s6 * s2;
  const double s12 = s6 * s6;
  const double s14 = s12 * s2;
  d->C2 = 6.0 * s14 - 3.0 * s8;
  d->C0 = -s12 + s6 - d->C2 / s2;
  d->nrep = nrep;
  d->rh.assign(nrep, Replica{});
  d->rst.assign(nrep, ReplicaState{});
  long long p0 = 0;
  for (int r = 0; r < nrep; r++) {
    Replica &R = d->rh[r];
    if (rl) {
      for (int k = 0; k < 3; k++) R.L[k] = rl[r].L[k];
      d->rst[r].T = rl[r].T;
      d->rst[r].gamma = rl[r].gamma;
      d->rst[r].seed = rl[r].seed;
      const double rho = counts[r] / (R.L[0] * R.L[1] * R.L[2]);
      auto_brick(d, rho, R.L, brick_budget(d), R.bd, kReplicaFill);
      R.p0 = (int)p0;
      R.pn = counts[r];
      p0 += counts[r];
    } else {
      for (int k = 0; k < 3; k++) R.bd[k] = d->bd[k];
    }
  }
  int r;
  if ((r = dalloc(d, &d->reps, nrep)) || (r = dalloc(d, &d->rep_v2, nrep)) || (r = dalloc(d, &d->rep_out, 3 * nrep)) ||
      (r = dalloc(d, &d->which, nrep)) || (r = dalloc(d, &d->rep_exp, nrep)) || (r = dalloc(d, &d->grng, nrep)) ||
      (r = dalloc(d, &d->clock, 1)) || (r = dalloc(d, &d->chk_dmax, nrep)) || (r = dalloc(d, &d->chk_cnt, nrep)) ||
      (r = dalloc(d, &d->gate, 1)) || (r = dalloc(d, &d->gbudget, nrep)) || (r = dalloc(d, &d->glife, nrep)) ||
      (r = dalloc(d, &d->chk_cand, (size_t)nrep * d->check_list_length)) ||
      (r = dalloc(d, &d->rep_factor, nrep))) {
    mdgpu_destroy(d);
    return r;
  }
  if (hipHostMalloc((void **)&d->h_rep_v2, nrep * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_rep_out, 3 * nrep * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_which, nrep * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_rep_exp, nrep * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_gate, sizeof(RunGate)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_gbudget, nrep * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_glife, nrep * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_rep_factor, nrep * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_chk_dmax, nrep * sizeof(unsigned long long)) != hipSuccess) {
    mdgpu_destroy(d);
    return MDGPU_ERR_HIP;
  }
  const int cap0 = rl ? (int)std::min<long long>(p0, kMaxParticles) : prm->capacity;
  if ((r = set_geometry(d)) != MDGPU_OK || (r = ensure_owned(d, std::max(cap0, 1024))) != MDGPU_OK ||
      (r = ensure_ghosts(d, std::max(1024, cap0 / 4))) != MDGPU_OK) {
    mdgpu_destroy(d);
    return r;
  }
  if ((r = dalloc(d, &d->u64, 4)) || (r = dalloc(d, &d->i64, 4)) || (r = dalloc(d, &d->i32, 4)) ||
      (r = dalloc(d, &d->d3, 4)) || (r = dalloc(d, &d->cand, (size_t)d->check_list_length + 1))) {
    mdgpu_destroy(d);
    return r;
  }
  if (unit && ((r = dalloc(d, &d->src_tab, kMaxUnits)) != MDGPU_OK ||
               hipHostMalloc((void **)&d->h_src_tab, kMaxUnits * sizeof(SrcUnit)) != hipSuccess)) {
    mdgpu_destroy(d);
    return r ? r : MDGPU_ERR_HIP;
  }
  if (hipHostMalloc((void **)&d->h_u64, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_i64, 4 * sizeof(long long)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_i32, 4 * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void **)&d->h_d3, 4 * sizeof(double)) != hipSuccess) {
    mdgpu_destroy(d);
    return MDGPU_ERR_HIP;
  }
  (void)hipMemset(d->i32, 0, 4 * sizeof(int));
  (void)hipMemset(d->u64, 0, 4 * sizeof(unsigned long long));
  (void)hipMemset(d->clock, 0, sizeof(unsigned long long));
  {
    // set-up of mdgpu_observe_noalloc on this device (the device is current)
    const hipError_t e = scratch_address(prm->device, &d->scratch_dev);
    if (e != hipSuccess) {
      // the domain works without it; mdgpu_observe_noalloc reports the failure
      d->scratch_dev = nullptr;
      d->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    }
  }
  *out = d;
  return MDGPU_OK;
}

// ---- units of a process grid ------------------------------------------------

bool same_group(const mdgpu_domain *a, const mdgpu_domain *b) { return a->group && a->group == b->group; }

// MDUnit::MakeBufferForSendingParticle (mdunit.cc:78-101): the owned
// particles over the face of direction dir (MDRect::IsOverBoundary) move to
// the unit's send buffer, the others are compacted in their order.
int unit_send(mdgpu_domain *d, int dir, int *nsend) {
  hipStream_t st = d->stream;
  invalidate(d);
  const int pn = d->pn;
  d->ng = 0;
  d->tn = pn;
  d->initialized = false;
  d->ghosts_ready = false;
  d->st_n = 0;
  if (pn > 0) {
    const int a = dir >> 1;
    CHK(ensure_scan(d, pn));
    hipLaunchKernelGGL(k_leave_flags, dim3(nblk(pn)), dim3(kBlock), 0, st, d->pos, pn, dir, d->rs[a], d->re[a],
                       d->nimg);
    exclusive_scan<int, int>(d->nimg, d->goff, pn, d->scan_scratch, d->i64, st);
    long long ns = 0;
    CHK(read_total(d, &ns));
    if (ns > d->st_cap) {
      const int c = (int)std::min<long long>(ns + ns / 4 + 256, kMaxParticles);
      d->st_cap = 0;
      CHK(dalloc(d, &d->st_pos, c));
      CHK(dalloc(d, &d->st_mom, c));
      CHK(dalloc(d, &d->st_id, c));
      d->st_cap = c;
    }
    if (ns > 0) {
      hipLaunchKernelGGL(k_partition, dim3(nblk(pn)), dim3(kBlock), 0, st, d->nimg, d->goff, pn, d->pos, d->mom,
                         d->id, d->pos_tmp, d->mom_alt, d->id_alt, d->st_pos, d->st_mom, d->st_id);
      std::swap(d->pos, d->pos_tmp);
      std::swap(d->mom, d->mom_alt);
      std::swap(d->id, d->id_alt);
      d->pn = pn - (int)ns;
      d->tn = d->pn;
    }
    d->st_n = (int)ns;
  }
  *nsend = d->st_n;
  d->vmax_valid = false;
  CHK(check_launch(d, "send particles"));
  return upload_replicas(d);
}

// MDUnit::ReceiveParticles (mdunit.cc:104-122): src's send buffer appended to
// the unit's owned particles, device to device.
int unit_receive(mdgpu_domain *d, mdgpu_domain *src) {
  const int n = src->st_n;
  if (n == 0) return MDGPU_OK;
  hipStream_t st = d->stream;
  invalidate(d);
  CHK(ensure_owned(d, d->pn + n));
  HIPCHK(d, hipMemcpyAsync(d->pos + d->pn, src->st_pos, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
  HIPCHK(d, hipMemcpyAsync(d->mom + d->pn, src->st_mom, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
  HIPCHK(d, hipMemcpyAsync(d->id + d->pn, src->st_id, n * sizeof(int), hipMemcpyDeviceToDevice, st));
  HIPCHK(d, hipMemsetAsync(d->rep + d->pn, 0, n * sizeof(int), st));
  d->pn += n;
  d->tn = d->pn;
  d->ng = 0;
  d->initialized = false;
  d->ghosts_ready = false;
  d->vmax_valid = false;
  return upload_replicas(d);
}

// MDUnit::AdjustPeriodicBoundary (mdunit.h:71 -> variables.cc:59-79, the whole
// box's edges), then the owned particles binned in the unit's cells
// (MeshList::MakeMesh, meshlist.cc:153-205) and put in brick order, so that
// the indices other units' ghosts take from this unit are final.
int unit_adjust(mdgpu_domain *d) {
  hipStream_t st = d->stream;
  CHK(auto_bricks(d));
  HIPCHK(d, hipMemsetAsync(d->i32, 0, sizeof(int) * 4, st));
  if (d->pn > 0)
    hipLaunchKernelGGL(k_unit_wrap, dim3(nblk(d->pn)), dim3(kBlock), 0, st, d->pos, d->pn, d->L[0], d->L[1],
                       d->L[2], d->i32);
  return bin_owned(d, 0);
}

// The unit's ghosts: what MDManager's six FindBorderParticles /
// SendBorderParticlesSub passes bring it (mdmanager.cc:269-275), taken from
// the owned particles of every unit of the group (SrcUnit), cell-sorted after
// its own.  Every unit of the group must have been through unit_adjust.  The
// group's lock keeps the units' positions where this reads them: another
// unit placing its ghosts may reallocate its arrays.
int unit_ghosts(mdgpu_domain *d) {
  UnitGroup *g = d->group;
  std::lock_guard<std::mutex> lock(g->mu);
  hipStream_t st = d->stream;
  const int pn = d->pn;
  const int nsrc = (int)g->units.size();
  long long total = 0;
  for (int w = 0; w < nsrc; w++) {
    const mdgpu_domain *W = g->units[w];
    SrcUnit &S = d->h_src_tab[w];
    S = SrcUnit{};
    S.base = (int)total;
    if (!W) continue;
    S.pos = W->pos;
    S.n = W->pn;
    for (int a = 0; a < 3; a++) {
      const int G = g->grid[a], u = d->upos[a], v = W->upos[a];
      S.same[a] = v == u;
      S.left[a] = v == (u - 1 + G) % G;   // W sends to u through its D_RIGHT-type face
      S.right[a] = v == (u + 1) % G;      // ... through its D_LEFT-type face
      S.lsh[a] = v + 1 >= G ? -1 : 0;     // ParaInfo::IsOverBoundary(W, high face)
      S.rsh[a] = v - 1 < 0 ? 1 : 0;       // ParaInfo::IsOverBoundary(W, low face)
      S.s[a] = W->rs[a];
      S.e[a] = W->re[a];
    }
    total += W->pn;
  }
  if (total > (long long)INT32_MAX - 1) {
    d->err = "too many particles in the group";
    return MDGPU_ERR_CAPACITY;
  }
  if (total + 1 > d->ucap) {
    const long long c = total + total / 8 + 1024;
    d->ucap = 0;
    CHK(dalloc(d, &d->ucnt, (size_t)c));
    CHK(dalloc(d, &d->uoff, (size_t)c));
    d->ucap = c;
  }
  HIPCHK(d, hipMemcpyAsync(d->src_tab, d->h_src_tab, nsrc * sizeof(SrcUnit), hipMemcpyHostToDevice, st));
  CHK(ensure_scan(d, std::max<long long>(total, d->ncell)));
  long long ngl = 0;
  if (total > 0) {
    hipLaunchKernelGGL(k_unit_count, dim3(nblk(total)), dim3(kBlock), 0, st, d->src_tab, nsrc, (int)total,
                       d->unit_slot, d->SL, d->ucnt);
    exclusive_scan<int, int>(d->ucnt, d->uoff, (int)total, d->scan_scratch, d->i64, st);
    CHK(read_total(d, &ngl));
  }
  if (ngl + pn + 1 > kMaxParticles) {
    d->err = "particle capacity exceeded (owned + ghosts)";
    return MDGPU_ERR_CAPACITY;
  }
  const int ng = (int)ngl;
  const double4 *own = d->pos;
  CHK(ensure_ghosts(d, ng));
  if (d->pos != own) {
    // the unit's own positions moved: the table's entry for it follows
    d->h_src_tab[d->unit_slot].pos = d->pos;
    HIPCHK(d, hipMemcpyAsync(d->src_tab, d->h_src_tab, nsrc * sizeof(SrcUnit), hipMemcpyHostToDevice, st));
  }
  HIPCHK(d, hipMemsetAsync(d->c_gcount, 0, sizeof(int) * d->ncell, st));
  if (ng > 0) {
    // pos_tmp[0..ng) holds the unsorted images
    hipLaunchKernelGGL(k_unit_emit, dim3(nblk(total)), dim3(kBlock), 0, st, d->src_tab, nsrc, (int)total,
                       d->unit_slot, d->SL, d->uoff, ng, g->L[0], g->L[1], g->L[2], d->reps, d->t_gsrc, d->t_gcode,
                       d->pos_tmp, d->t_gcell, d->slot, d->c_gcount);
    CHK(check_launch(d, "unit ghosts"));
    CHK(cell_sort(d, d->t_gcell, ng, d->c_gcount, d->c_gstart, false, nullptr));
    hipLaunchKernelGGL(k_gather_unit_ghosts, dim3(nblk(ng)), dim3(kBlock), 0, st, d->perm, ng, pn, d->t_gsrc,
                       d->t_gcode, d->pos_tmp, d->t_gcell, d->gsrc, d->gcode, d->gunit, d->pos, d->cell);
  } else {
    exclusive_scan<int, int>(d->c_gcount, d->c_gstart, d->ncell, d->scan_scratch, nullptr, st);
  }
  d->ng = ng;
  d->tn = pn + ng;
  d->ghosts_ready = true;
  d->ghost_epoch = g->epoch;
  d->ghost_remap = true;
  // the host table is read by the copies above before it is written again:
  // the next write follows this unit's next read_total
  return check_launch(d, "unit ghosts");
}

// The unit's ghost sources after every source unit has built its list and
// ordered its rows (order_rows permutes a unit's owned particles): each
// source index mapped through its unit's row order, once per construction.
// The group's lock is held.
int remap_unit_ghosts_locked(mdgpu_domain *d) {
  UnitGroup *g = d->group;
  if (!d->ghost_remap) return MDGPU_OK;
  if (d->ghost_epoch != g->epoch) {
    d->err = "a unit of the group was destroyed after this unit's ghosts were placed: rebuild the group's lists";
    return MDGPU_ERR_ARGUMENT;
  }
  UnitRowInv R{};
  for (size_t w = 0; w < g->units.size(); w++) {
    const mdgpu_domain *W = g->units[w];
    if (W && !W->initialized) {
      d->err = "a unit's halo before every unit of its group built its pair list";
      return MDGPU_ERR_ARGUMENT;
    }
    R.inv[w] = W ? W->rowinv : nullptr;
  }
  if (d->ng > 0)
    hipLaunchKernelGGL(k_remap_unit_ghosts, dim3(nblk(d->ng)), dim3(kBlock), 0, d->stream, d->gsrc, d->gunit, d->ng, R);
  d->ghost_remap = false;
  return check_launch(d, "unit ghost sources");
}

// MDManager::SendBorderParticles for a unit (mdmanager.cc:283-348): each ghost
// from its source unit's current owned position.
int unit_refresh(mdgpu_domain *d) {
  if (d->ng == 0) return MDGPU_OK;
  UnitGroup *g = d->group;
  UnitPos P{};
  {
    std::lock_guard<std::mutex> lock(g->mu);
    if (d->ghost_epoch != g->epoch) {
      d->err = "a unit of the group was destroyed after this unit's ghosts were placed: rebuild the group's lists";
      return MDGPU_ERR_ARGUMENT;
    }
    CHK(remap_unit_ghosts_locked(d));
    for (size_t w = 0; w < g->units.size(); w++) P.pos[w] = g->units[w] ? g->units[w]->pos : nullptr;
  }
  hipLaunchKernelGGL(k_unit_refresh, dim3(nblk(d->ng)), dim3(kBlock), 0, d->stream, d->pos, d->gsrc, d->gcode,
                     d->gunit, d->pn, d->ng, P, g->L[0], g->L[1], g->L[2]);
  return check_launch(d, "unit halo");
}

// The unit of the group at grid position p (nullptr if there is none).
mdgpu_domain *unit_at(UnitGroup *g, const int p[3]) {
  for (mdgpu_domain *u : g->units)
    if (u && u->upos[0] == p[0] && u->upos[1] == p[1] && u->upos[2] == p[2]) return u;
  return nullptr;
}

// The neighbour of unit u in direction dir (ParaInfo::GetNeighborID, parainfo.cc:191-224).
mdgpu_domain *unit_neighbor(mdgpu_domain *u, int dir) {
  UnitGroup *g = u->group;
  int p[3] = {u->upos[0], u->upos[1], u->upos[2]};
  const int a = dir >> 1;
  p[a] += (dir & 1) ? 1 : -1;
  p[a] = (p[a] + g->grid[a]) % g->grid[a];
  return unit_at(g, p);
}

// Ghost source ids: the id of the owned particle each ghost images (the
// unit's own, or the source unit's for a unit).
int ghost_source_ids(mdgpu_domain *d, std::vector<int> &out) {
  const int ng = d->ng;
  out.assign(ng, 0);
  if (ng == 0) return MDGPU_OK;
  if (is_unit(d)) {
    std::lock_guard<std::mutex> lock(d->group->mu);
    CHK(remap_unit_ghosts_locked(d));
  }
  std::vector<int> gs(ng), gu(ng, 0);
  HIPCHK(d, hipMemcpyAsync(gs.data(), d->gsrc, ng * sizeof(int), hipMemcpyDeviceToHost, d->stream));
  if (is_unit(d)) HIPCHK(d, hipMemcpyAsync(gu.data(), d->gunit, ng * sizeof(int), hipMemcpyDeviceToHost, d->stream));
  CHK(sync(d));
  std::vector<std::vector<int>> ids;
  if (!is_unit(d)) {
    ids.resize(1);
    ids[0].resize(d->pn);
    if (d->pn) HIPCHK(d, hipMemcpy(ids[0].data(), d->id, d->pn * sizeof(int), hipMemcpyDeviceToHost));
  } else {
    UnitGroup *g = d->group;
    std::lock_guard<std::mutex> lock(g->mu);
    if (d->ghost_epoch != g->epoch) {
      d->err = "a unit of the group was destroyed after this unit's ghosts were placed: rebuild the group's lists";
      return MDGPU_ERR_ARGUMENT;
    }
    ids.resize(g->units.size());
    for (size_t w = 0; w < g->units.size(); w++) {
      const mdgpu_domain *W = g->units[w];
      if (!W || W->pn == 0) continue;
      ids[w].resize(W->pn);
      HIPCHK(d, hipMemcpy(ids[w].data(), W->id, W->pn * sizeof(int), hipMemcpyDeviceToHost));
    }
  }
  for (int k = 0; k < ng; k++) {
    if (gu[k] < 0 || gu[k] >= (int)ids.size() || gs[k] < 0 || gs[k] >= (int)ids[gu[k]].size()) {
      d->err = "ghost source out of range: a source unit changed since this unit's construction";
      return MDGPU_ERR_ARGUMENT;
    }
    out[k] = ids[gu[k]][gs[k]];
  }
  return MDGPU_OK;
}

// The live units of d's group, in slot order.
std::vector<mdgpu_domain *> group_units(mdgpu_domain *d) {
  std::vector<mdgpu_domain *> us;
  std::lock_guard<std::mutex> lock(d->group->mu);
  for (mdgpu_domain *u : d->group->units)
    if (u) us.push_back(u);
  return us;
}

// A unit's error, reported on the unit the caller called.
int group_fail(mdgpu_domain *d, mdgpu_domain *u, int r) {
  if (u != d) d->err = u->err;
  return r;
}

// MDManager::MakePairList (mdmanager.cc:262-280) for every unit of the group.
int group_make_pair_list(mdgpu_domain *d, const std::vector<mdgpu_domain *> &us) {
  int r;
  // MDManager::SendParticles (mdmanager.cc:194-260): every unit fills its
  // buffer for a direction, then every unit takes its opposite neighbour's
  for (int dir = 0; dir < 6; dir++) {
    for (mdgpu_domain *u : us) {
      int n = 0;
      if ((r = unit_send(u, dir, &n)) != MDGPU_OK) return group_fail(d, u, r);
    }
    for (mdgpu_domain *u : us) {
      mdgpu_domain *src = unit_neighbor(u, dir ^ 1);
      if (!src) {
        d->err = "the group does not fill its grid";
        return MDGPU_ERR_ARGUMENT;
      }
      if ((r = unit_receive(u, src)) != MDGPU_OK) return group_fail(d, u, r);
    }
  }
  for (mdgpu_domain *u : us)
    if ((r = unit_adjust(u)) != MDGPU_OK) return group_fail(d, u, r);
  for (mdgpu_domain *u : us)
    if ((r = unit_ghosts(u)) != MDGPU_OK) return group_fail(d, u, r);
  for (mdgpu_domain *u : us)
    if ((r = make_pair_list(u)) != MDGPU_OK) return group_fail(d, u, r);
  std::lock_guard<std::mutex> lock(d->group->mu);
  for (mdgpu_domain *u : us)
    if ((r = remap_unit_ghosts_locked(u)) != MDGPU_OK) return group_fail(d, u, r);
  return MDGPU_OK;
}

// MDManager::IsPairListExpired (mdmanager.cc:369-377): every unit's
// PairList::IsPairListExpired (pairlist.cc:26-65, is_expired), OR'ed.  The
// units share a stream, so their maxima of |p| are reduced together and read
// after one wait, and so are the displacement checks of the units whose budget
// ran out.
int group_expired(mdgpu_domain *d, const std::vector<mdgpu_domain *> &us, bool *any) {
  *any = false;
  std::vector<char> vmax_of(us.size(), 0), check_of(us.size(), 0);
  bool queued = false;
  for (size_t k = 0; k < us.size(); k++) {
    mdgpu_domain *u = us[k];
    if (!u->initialized) {
      *any = true;
      continue;
    }
    u->lifetime++;
    if (u->buffer_length > 0.0) {
      vmax_of[k] = 1;
      queued = true;
    }
  }
  if (queued) {
    // every unit's max |p|^2 in two launches (the maximum does not depend on
    // the order it is taken in), read after one wait
    UnitGroup *g = d->group;
    const int ns = (int)g->units.size();
    GroupMomenta G{};
    int nb = 0;
    for (int w = 0; w < ns; w++) {
      const mdgpu_domain *W = g->units[w];
      G.mom[w] = W ? W->mom : nullptr;
      G.pn[w] = W ? W->pn : 0;
      G.bbase[w] = nb;
      nb += (int)nblk(G.pn[w]);
    }
    G.bbase[ns] = nb;
    G.n = ns;
    if ((size_t)nb > g->part_n) {
      if (g->part) (void)hipFree(g->part);
      g->part = nullptr;
      g->part_n = 0;
      HIPCHK(d, hipMalloc((void **)&g->part, (size_t)nb * 2 * sizeof(double)));
      g->part_n = (size_t)nb * 2;
    }
    if (!g->vmax2) {
      HIPCHK(d, hipMalloc((void **)&g->vmax2, kMaxUnits * sizeof(unsigned long long)));
      HIPCHK(d, hipHostMalloc((void **)&g->h_vmax2, kMaxUnits * sizeof(unsigned long long)));
    }
    hipLaunchKernelGGL(k_group_max_p2, dim3(nb), dim3(kBlock), 0, d->stream, G, g->part);
    hipLaunchKernelGGL(k_group_reduce_max, dim3(ns), dim3(kBlock), 0, d->stream, G, g->part, g->vmax2);
    CHK(check_launch(d, "group max |p|"));
    HIPCHK(d, hipMemcpyAsync(g->h_vmax2, g->vmax2, ns * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             d->stream));
    CHK(sync(d));
  }
  for (size_t k = 0; k < us.size(); k++) {
    if (!vmax_of[k]) continue;
    mdgpu_domain *u = us[k];
    double v2 = 0.0;
    std::memcpy(&v2, d->group->h_vmax2 + u->unit_slot, sizeof(double));
    const double vmax = std::sqrt(v2);
    u->step_vmax = vmax;
    if (vmax * 2.0 * u->dt > u->margin) {
      char b[160];
      snprintf(b, sizeof b, "Too fast particles exists. Try smaller time step (max_velocity = %f)", vmax);
      u->err = b;
      return group_fail(d, u, MDGPU_ERR_TOO_FAST);
    }
    u->buffer_length = u->buffer_length - vmax * 2.0 * u->dt;
  }
  queued = false;
  for (size_t k = 0; k < us.size(); k++) {
    mdgpu_domain *u = us[k];
    if (!u->initialized || !(u->buffer_length < 0.0)) continue;
    if (!u->disp_check) {
      *any = true;
      continue;
    }
    const int r = enqueue_displacement_check(u);
    if (r != MDGPU_OK) return group_fail(d, u, r);
    check_of[k] = 1;
    queued = true;
  }
  if (queued) CHK(sync(d));
  for (size_t k = 0; k < us.size(); k++) {
    if (!check_of[k]) continue;
    int e = 0;
    CHK(displacement_check_result(us[k], &e));
    if (e) *any = true;
  }
  // without an expiry, the units whose list passed take a budget again
  if (!*any)
    for (size_t k = 0; k < us.size(); k++)
      if (check_of[k]) rearm_after_check(us[k]);
  return MDGPU_OK;
}

// UpdatePositionHalf of every unit in one launch.
int group_drift(mdgpu_domain *d, const std::vector<mdgpu_domain *> &us) {
  UnitGroup *g = d->group;
  const int ns = (int)g->units.size();
  GroupDrift G{};
  int t = 0;
  for (int w = 0; w < ns; w++) {
    mdgpu_domain *W = g->units[w];
    G.pos[w] = W ? W->pos : nullptr;
    G.mom[w] = W ? W->mom : nullptr;
    G.base[w] = t;
    t += W ? W->pn : 0;
  }
  G.base[ns] = t;
  G.n = ns;
  for (mdgpu_domain *u : us) invalidate(u);
  if (t > 0) hipLaunchKernelGGL(k_group_drift, dim3(nblk(t)), dim3(kBlock), 0, d->stream, G, d->dt * 0.5);
  return check_launch(d, "group drift");
}

// MDManager::SendBorderParticles for every unit in one launch (each unit's
// ghost sources remapped first where its construction left that pending).
int group_refresh(mdgpu_domain *d) {
  UnitGroup *g = d->group;
  std::lock_guard<std::mutex> lock(g->mu);
  const int ns = (int)g->units.size();
  GroupHalo G{};
  int t = 0;
  for (int w = 0; w < ns; w++) {
    mdgpu_domain *W = g->units[w];
    G.pos[w] = W ? W->pos : nullptr;
    G.base[w] = t;
    if (!W || W->ng == 0) continue;
    if (W->ghost_epoch != g->epoch) {
      d->err = "a unit of the group was destroyed after the ghosts were placed: rebuild the group's lists";
      return MDGPU_ERR_ARGUMENT;
    }
    const int r = remap_unit_ghosts_locked(W);
    if (r != MDGPU_OK) return group_fail(d, W, r);
    G.gsrc[w] = W->gsrc;
    G.gcode[w] = W->gcode;
    G.gunit[w] = W->gunit;
    G.pn[w] = W->pn;
    t += W->ng;
  }
  G.base[ns] = t;
  G.n = ns;
  if (t > 0)
    hipLaunchKernelGGL(k_group_refresh, dim3(nblk(t)), dim3(kBlock), 0, d->stream, G, g->L[0], g->L[1], g->L[2]);
  return check_launch(d, "group halo");
}

// Temperature over the whole group (MDManager::Temperature, mdmanager.cc:390-397).
int group_temperature(mdgpu_domain *d, const std::vector<mdgpu_domain *> &us, double *t) {
  double ke = 0.0;
  long long n = 0;
  for (mdgpu_domain *u : us) {
    double k = 0.0;
    const int r = kinetic(u, &k);
    if (r != MDGPU_OK) return group_fail(d, u, r);
    ke += k;
    n += u->pn;
  }
  *t = (ke / (double)n) / 1.5;
  return MDGPU_OK;
}

// MDManager::Calculate (mdmanager.cc:116-192) for every unit of the group,
// nsteps times: the expiry test of every unit, the group's construction if any
// expired, the first half drift of every unit, every unit's halo, then the
// force and the second half drift (NVE), the force and the Langevin update
// (Langevin), or the Nose-Hoover sequence with the group's temperature.  The
// units' thermostat is the one d has.  With timing on (mdgpu_enable_timing on
// d), every unit records its force launches and d the wall time of the steps.
// CalculateForce of every unit of the group (force_in_place on each), in one
// launch of all their bricks when the units run the same kick (FP64, or mixed
// precision in the brick frame) and there are at most kMaxKickUnits of them,
// with the launch shape of the unit stepping the group; else one launch per
// unit.  The units' results are those of their own launches bit for bit: each
// workgroup does what the unit's own kick does to that brick.
template <int W, int U>
void launch_group_kick(mdgpu_domain *d, const UnitKicks &uk, int nb, size_t shm) {
  const StepClock sc{d->dt, d->step, nullptr, nullptr};
  if (d->precision == MDGPU_FP64) {
    allow_lds(k_force_group_fp64<W, U>, shm);
    hipLaunchKernelGGL((k_force_group_fp64<W, U>), dim3(nb), dim3(kWave * W), shm, d->stream, uk, fconsts(d), sc);
    return;
  }
  const ForceConstsF f32 = fconsts32(d);
  const double dt2 = d->dt * 0.5;
  switch (d->kahan) {
    case kSumPlain:
      allow_lds(k_force_group_mixed_brick<W, U, kSumPlain>, shm);
      hipLaunchKernelGGL((k_force_group_mixed_brick<W, U, kSumPlain>), dim3(nb), dim3(kWave * W), shm, d->stream, uk,
                         f32, dt2, sc);
      break;
    case kSumKahan:
      allow_lds(k_force_group_mixed_brick<W, U, kSumKahan>, shm);
      hipLaunchKernelGGL((k_force_group_mixed_brick<W, U, kSumKahan>), dim3(nb), dim3(kWave * W), shm, d->stream, uk,
                         f32, dt2, sc);
      break;
    default:
      allow_lds(k_force_group_mixed_brick<W, U, kSumGroup64>, shm);
      hipLaunchKernelGGL((k_force_group_mixed_brick<W, U, kSumGroup64>), dim3(nb), dim3(kWave * W), shm, d->stream,
                         uk, f32, dt2, sc);
      break;
  }
}

// The group kick's indices ahead: unless set, 8 for both precisions.  The
// FP64 group kernel with 16 takes 136 registers, which fits one 512-thread
// workgroup per CU instead of two (94 registers with 8): 0.381 against
// 0.310 ms per launch for 2x2x2 units of 4M particles.  The sums are the same
// at any depth.
int group_unroll(const mdgpu_domain *d) { return d->unroll ? d->unroll : 8; }

template <int W>
void launch_group_kick_w(mdgpu_domain *d, const UnitKicks &uk, int nb, size_t shm) {
  if (group_unroll(d) == 32) launch_group_kick<W, 4>(d, uk, nb, shm);
  else if (group_unroll(d) == 16) launch_group_kick<W, 2>(d, uk, nb, shm);
  else launch_group_kick<W, 1>(d, uk, nb, shm);
}

int group_force(mdgpu_domain *d, const std::vector<mdgpu_domain *> &us) {
  bool batch = (d->precision == MDGPU_FP64 || (d->precision == MDGPU_MIXED && d->mixed_frame == 1)) &&
               !d->pipeline && us.size() <= (size_t)kMaxKickUnits;
  for (mdgpu_domain *u : us)
    batch = batch && u->precision == d->precision && u->mixed_frame == d->mixed_frame && u->kahan == d->kahan &&
            u->dt == d->dt && u->CL2 == d->CL2 && u->C2 == d->C2 && u->step == d->step;
  if (!batch) {
    for (mdgpu_domain *u : us) CHK(force_in_place(u));
    return MDGPU_OK;
  }
  const int bytes = d->precision == MDGPU_FP64 ? 3 * (int)sizeof(double) : (int)sizeof(float4);
  UnitKicks uk{};
  uk.n = (int)us.size();
  int nb = 0;
  size_t shm = 0;
  for (size_t k = 0; k < us.size(); k++) {
    mdgpu_domain *u = us[k];
    invalidate(u);
    u->vmax_valid = false;
    UnitKick &e = uk.u[k];
    e.src = u->pos;
    e.mom = u->mom;
    e.bl = brick_list(u, bytes);
    e.brick0 = nb;
    if (u->pn > 0) {
      nb += u->nbrick;
      shm = std::max(shm, (size_t)e.bl.lds_cap * bytes);
    }
  }
  if (nb == 0) return MDGPU_OK;
  const int ev0 = d->force_launches * 2;
  const bool ev = d->timing && ev0 + 1 < (int)d->ev.size();
  if (ev) HIPCHK(d, hipEventRecord(d->ev[ev0], d->stream));
  switch (d->waves) {
    case 2: launch_group_kick_w<2>(d, uk, nb, shm); break;
    case 8: launch_group_kick_w<8>(d, uk, nb, shm); break;
    case 16: launch_group_kick_w<16>(d, uk, nb, shm); break;
    default: launch_group_kick_w<4>(d, uk, nb, shm); break;
  }
  if (ev) {
    HIPCHK(d, hipEventRecord(d->ev[ev0 + 1], d->stream));
    d->force_launches++;
  }
  return check_launch(d, "group force");
}

int group_calculate(mdgpu_domain *d, int nsteps, int *rebuilds) {
  const std::vector<mdgpu_domain *> us = group_units(d);
  for (mdgpu_domain *u : us)
    if (u->thermostat != d->thermostat) {
      d->err = "the units of a group step with one thermostat";
      return MDGPU_ERR_ARGUMENT;
    }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  for (mdgpu_domain *u : us) {
    u->force_launches = 0;
    u->force_ms = u->total_ms = u->rebuild_ms = 0.0;
    u->rebuild_count = 0;
    u->timing = d->timing;
    if (d->timing) {
      const size_t need = (size_t)nsteps * 2 + 2;
      while (u->ev.size() < need) {
        hipEvent_t e;
        HIPCHK(d, hipEventCreate(&e));
        u->ev.push_back(e);
      }
    }
  }
  if (d->timing) {
    HIPCHK(d, hipEventCreate(&t0));
    HIPCHK(d, hipEventCreate(&t1));
    HIPCHK(d, hipEventRecord(t0, d->stream));
  }
  int count = 0;
  int r = MDGPU_OK;
  for (int s = 0; s < nsteps && r == MDGPU_OK; s++) {
    bool any = false;
    if ((r = group_expired(d, us, &any)) != MDGPU_OK) break;
    if (any) {
      if ((r = group_make_pair_list(d, us)) != MDGPU_OK) break;
      count++;
    }
    if ((r = group_drift(d, us)) != MDGPU_OK || (r = group_refresh(d)) != MDGPU_OK) break;
    if (d->thermostat == MDGPU_NOSEHOOVER) {
      double t = 0.0;
      if ((r = group_temperature(d, us, &t)) != MDGPU_OK) break;
      // each unit's sequence (scale, kick, scale) touches only its own particles
      for (mdgpu_domain *u : us)
        if ((r = heatbath_zeta(u, t)) != MDGPU_OK || (r = heatbath_momenta(u)) != MDGPU_OK) break;
      if (r != MDGPU_OK || (r = group_force(d, us)) != MDGPU_OK) break;
      for (mdgpu_domain *u : us)
        if ((r = heatbath_momenta(u)) != MDGPU_OK) break;
      if (r != MDGPU_OK) break;
      if ((r = group_temperature(d, us, &t)) != MDGPU_OK) break;
      for (mdgpu_domain *u : us)
        if ((r = heatbath_zeta(u, t)) != MDGPU_OK) break;
      if (r != MDGPU_OK || (r = group_drift(d, us)) != MDGPU_OK) break;
    } else if (d->thermostat == MDGPU_LANGEVIN) {
      if ((r = group_force(d, us)) != MDGPU_OK) break;
      for (mdgpu_domain *u : us)
        if ((r = mdgpu_langevin(u)) != MDGPU_OK) break;
    } else {
      if ((r = group_force(d, us)) != MDGPU_OK || (r = group_drift(d, us)) != MDGPU_OK) break;
    }
    if (r != MDGPU_OK) break;
    for (mdgpu_domain *u : us) {
      u->sim_time += u->dt;
      if (d->thermostat != MDGPU_LANGEVIN) u->step++;  // mdgpu_langevin counts its own calls
    }
  }
  if (r == MDGPU_OK) r = sync(d);
  if (d->timing && t0) {
    HIPCHK(d, hipEventRecord(t1, d->stream));
    HIPCHK(d, hipEventSynchronize(t1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    d->total_ms = ms;
    for (mdgpu_domain *u : us)
      for (int k = 0; k < u->force_launches; k++) {
        (void)hipEventElapsedTime(&ms, u->ev[2 * k], u->ev[2 * k + 1]);
        u->force_ms += ms;
      }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
  }
  d->rebuild_count = count;
  if (rebuilds) *rebuilds = count;
  return r;
}

}  // namespace

extern "C" {

int mdgpu_create(const mdgpu_params *prm, mdgpu_domain **out) {
  if (!prm || !out) return MDGPU_ERR_ARGUMENT;
  return create_domain(prm, 1, nullptr, nullptr, out);
}

int mdgpu_create_unit(const mdgpu_params *prm, const mdgpu_unit *u, mdgpu_domain *join, mdgpu_domain **out) {
  if (!prm || !u || !out) return MDGPU_ERR_ARGUMENT;
  *out = nullptr;
  for (int a = 0; a < 3; a++)
    if (u->grid[a] < 1 || u->pos[a] < 0 || u->pos[a] >= u->grid[a] || !(u->s[a] < u->e[a]) || !(prm->L[a] > 0.0))
      return MDGPU_ERR_ARGUMENT;
  UnitGroup *g = nullptr;
  bool fresh = false;
  if (join) {
    g = join->group;
    if (!g) {
      join->err = "mdgpu_create_unit: the domain to join is not a unit";
      return MDGPU_ERR_ARGUMENT;
    }
  } else {
    g = new UnitGroup();
    fresh = true;
    g->device = prm->device;
    for (int a = 0; a < 3; a++) {
      g->L[a] = prm->L[a];
      g->grid[a] = u->grid[a];
    }
    if (hipSetDevice(prm->device) != hipSuccess ||
        hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
      delete g;
      return MDGPU_ERR_HIP;
    }
  }
  mdgpu_domain *d = nullptr;
  const int r = create_domain(prm, 1, nullptr, nullptr, &d, u, g->stream);
  if (r != MDGPU_OK) {
    if (fresh) {
      (void)hipStreamDestroy(g->stream);
      delete g;
    }
    return r;
  }
  std::string why;
  {
    std::lock_guard<std::mutex> lock(g->mu);
    for (int a = 0; a < 3; a++)
      if (g->grid[a] != u->grid[a] || g->L[a] != prm->L[a]) why = "the units of a group share the grid and the box";
    if (g->device != prm->device) why = "the units of a group share the device";
    if ((int)g->units.size() >= kMaxUnits) why = "too many units in one group";
    if (unit_at(g, u->pos)) why = "a unit of the group already has this grid position";
    if (why.empty()) {
      d->group = g;
      d->unit_slot = (int)g->units.size();
      g->units.push_back(d);
      g->live++;
    }
  }
  if (!why.empty()) {
    mdgpu_destroy(d);  // not in the group: leaves the group's stream alone
    if (join) join->err = "mdgpu_create_unit: " + why;
    return MDGPU_ERR_ARGUMENT;
  }
  *out = d;
  return MDGPU_OK;
}

int mdgpu_unit_send_particles(mdgpu_domain *d, int dir, int *nsend) {
  if (!d || dir < 0 || dir > 5 || !nsend) return MDGPU_ERR_ARGUMENT;
  if (!is_unit(d)) {
    d->err = "not a unit (mdgpu_create_unit)";
    return MDGPU_ERR_ARGUMENT;
  }
  return unit_send(d, dir, nsend);
}

int mdgpu_unit_receive_particles(mdgpu_domain *d, mdgpu_domain *src) {
  if (!d || !src) return MDGPU_ERR_ARGUMENT;
  if (!same_group(d, src)) {
    d->err = "particles are received from a unit of the same group";
    return MDGPU_ERR_ARGUMENT;
  }
  return unit_receive(d, src);
}

int mdgpu_unit_adjust_periodic_boundary(mdgpu_domain *d) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (!is_unit(d)) {
    d->err = "not a unit (mdgpu_create_unit)";
    return MDGPU_ERR_ARGUMENT;
  }
  return unit_adjust(d);
}

int mdgpu_unit_find_border_particles(mdgpu_domain *d) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (!is_unit(d)) {
    d->err = "not a unit (mdgpu_create_unit)";
    return MDGPU_ERR_ARGUMENT;
  }
  return unit_ghosts(d);
}

int mdgpu_group_make_pair_list(mdgpu_domain *d) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (!is_unit(d)) {
    d->err = "not a unit (mdgpu_create_unit)";
    return MDGPU_ERR_ARGUMENT;
  }
  return group_make_pair_list(d, group_units(d));
}

int mdgpu_group_calculate(mdgpu_domain *d, int nsteps, int *rebuilds) {
  if (!d || nsteps < 0) return MDGPU_ERR_ARGUMENT;
  if (!is_unit(d)) {
    d->err = "not a unit (mdgpu_create_unit)";
    return MDGPU_ERR_ARGUMENT;
  }
  return group_calculate(d, nsteps, rebuilds);
}

int mdgpu_create_ensemble(const mdgpu_params *prm, int thermostat, int nrep, const mdgpu_replica *reps,
                          const int *counts, const double *q, const double *p, const int *type, mdgpu_domain **out) {
  if (!prm || !out || nrep < 1 || !reps || !counts || !q) return MDGPU_ERR_ARGUMENT;
  *out = nullptr;
  if (thermostat != MDGPU_NVE && thermostat != MDGPU_LANGEVIN && thermostat != MDGPU_NOSEHOOVER)
    return MDGPU_ERR_ARGUMENT;
  if (prm->check_list_length > kCheckCap) return MDGPU_ERR_ARGUMENT;
  long long total = 0;
  for (int r = 0; r < nrep; r++) {
    if (counts[r] < 1) return MDGPU_ERR_ARGUMENT;
    total += counts[r];
  }
  if (total + 1 > kMaxParticles) return MDGPU_ERR_CAPACITY;
  // every particle inside its replica's box (MDRect::IsInside, mdrect.h:35-38)
  {
    long long i = 0;
    for (int r = 0; r < nrep; r++)
      for (int k = 0; k < counts[r]; k++, i++)
        for (int a = 0; a < 3; a++) {
          const double x = q[3 * i + a];
          if (!(0.0 <= x && x < reps[r].L[a])) return MDGPU_ERR_ARGUMENT;
        }
  }
  mdgpu_domain *d = nullptr;
  int rc = create_domain(prm, nrep, reps, counts, &d);
  if (rc != MDGPU_OK) return rc;
  d->thermostat = thermostat;
  const int n = (int)total;
  std::vector<double4> hq(n), hp(n);
  std::vector<int> hid(n), hrep(n);
  {
    int i = 0;
    for (int r = 0; r < nrep; r++)
      for (int k = 0; k < counts[r]; k++, i++) {
        hq[i] = make_double4(q[3 * (size_t)i], q[3 * (size_t)i + 1], q[3 * (size_t)i + 2], 0.0);
        hp[i] = p ? make_double4(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2], 0.0)
                  : make_double4(0.0, 0.0, 0.0, 0.0);
        hid[i] = type ? type[i] : k;
        hrep[i] = r;
      }
  }
  if (hipMemcpy(d->pos, hq.data(), n * sizeof(double4), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d->mom, hp.data(), n * sizeof(double4), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d->id, hid.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d->rep, hrep.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    mdgpu_destroy(d);
    return MDGPU_ERR_HIP;
  }
  d->pn = d->tn = n;
  if ((rc = upload_replicas(d)) != MDGPU_OK || (rc = make_pair_list(d)) != MDGPU_OK) {
    mdgpu_destroy(d);
    return rc;
  }
  *out = d;
  return MDGPU_OK;
}

int mdgpu_destroy(mdgpu_domain *d) {
  if (!d) return MDGPU_OK;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  void *bufs[] = {d->pos,     d->pos_tmp,     d->pos_next,    d->rel_next,  d->pos_old,     d->mom,       d->mom_alt,  d->rel,
                  d->rel_tmp, d->id,          d->id_alt,      d->cell,      d->cell_alt, d->gsrc,
                  d->gcode,   d->c_ocount,    d->c_ostart,    d->c_gcount,  d->c_gstart, d->count_r,
                  d->start_r, d->org,         d->rank,        d->brick_of_cell, d->brick_rank0, d->row0,
                  d->nrow,    d->nsl,         d->vslice0,     d->slice_brick, d->tstart,  d->tbase,
                  d->nrange,  d->btotal,      d->gbase,       d->glist,     d->mask,     d->spart,
                  d->nbr,     d->ncode,       d->sbase,       d->slots64,   d->slen,     d->row_len,
                  d->rowsrc,  d->rowinv,      d->slot,        d->perm,      d->nimg,     d->goff,
                  d->t_gsrc,  d->t_gcode,     d->t_gcell,     d->key,       d->scan_scratch, d->disp,
                  d->partial, d->maxpart,     d->bpart,       d->cand,      d->u64,      d->i64,
                  d->i32,     d->d3,          d->reps,        d->brep,      d->rep,      d->rep_alt,
                  d->grng,    d->rep_v2,      d->rep_out,     d->which,     d->rep_exp,  d->clock,
                  d->perm_in, d->nbr_in,      d->ncode_in,    d->slen_in,   d->chk_dmax, d->chk_cnt,
                  d->chk_cand, d->gate,       d->gbudget,     d->glife,     d->xfer,     d->gunit,
                  d->st_pos,  d->st_mom,      d->st_id,       d->src_tab,   d->ucnt,     d->uoff,
                  d->rep_factor};
  for (StepGraph &g : d->graphs) drop_graph(g);
  for (StepGraph &g : d->gated) drop_graph(g);
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (void *h : {(void *)d->h_rep_v2, (void *)d->h_rep_out, (void *)d->h_which, (void *)d->h_rep_exp,
                  (void *)d->h_gate, (void *)d->h_gbudget, (void *)d->h_glife, (void *)d->h_rep_factor,
                  (void *)d->h_chk_dmax})
    if (h) (void)hipHostFree(h);
  if (d->h_u64) (void)hipHostFree(d->h_u64);
  if (d->h_i64) (void)hipHostFree(d->h_i64);
  if (d->h_i32) (void)hipHostFree(d->h_i32);
  if (d->h_d3) (void)hipHostFree(d->h_d3);
  if (d->h_src_tab) (void)hipHostFree(d->h_src_tab);
  for (hipEvent_t e : d->ev) (void)hipEventDestroy(e);
  if (d->stream && d->own_stream) (void)hipStreamDestroy(d->stream);
  if (d->group) {
    // the last unit of a group takes the group and its stream with it
    UnitGroup *g = d->group;
    bool last = false;
    {
      std::lock_guard<std::mutex> lock(g->mu);
      if (d->unit_slot >= 0 && d->unit_slot < (int)g->units.size()) g->units[d->unit_slot] = nullptr;
      g->epoch++;
      last = --g->live == 0;
    }
    if (last) {
      (void)hipStreamDestroy(g->stream);
      if (g->part) (void)hipFree(g->part);
      if (g->vmax2) (void)hipFree(g->vmax2);
      if (g->h_vmax2) (void)hipHostFree(g->h_vmax2);
      delete g;
    }
  }
  delete d;
  return MDGPU_OK;
}

const char *mdgpu_last_error(const mdgpu_domain *d) { return d ? d->err.c_str() : "null domain"; }

int mdgpu_set_precision(mdgpu_domain *d, int precision) {
  if (!d || (precision != MDGPU_FP64 && precision != MDGPU_MIXED)) return MDGPU_ERR_ARGUMENT;
  invalidate(d);  // the next fused step recomputes its first half drift
  d->precision = precision;
  return recode_list(d);
}

int mdgpu_add_particles(mdgpu_domain *d, int n, const double *q, const double *p, const int *type) {
  if (!d || n < 0 || (n > 0 && !q)) return MDGPU_ERR_ARGUMENT;
  if (d->nrep > 1) {
    d->err = "an ensemble's particles are set when it is created";
    return MDGPU_ERR_ARGUMENT;
  }
  invalidate(d);  // the next fused step recomputes its first half drift
  std::vector<double4> hq, hp;
  std::vector<int> hid;
  hq.reserve(n);
  for (int i = 0; i < n; i++) {
    const double *x = q + 3 * (size_t)i;
    // MDRect::IsInside (mdrect.h:35-38)
    if (!(d->rs[0] <= x[0] && x[0] < d->re[0] && d->rs[1] <= x[1] && x[1] < d->re[1] && d->rs[2] <= x[2] &&
          x[2] < d->re[2]))
      continue;
    hq.push_back(make_double4(x[0], x[1], x[2], 0.0));
    hp.push_back(p ? make_double4(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2], 0.0)
                   : make_double4(0.0, 0.0, 0.0, 0.0));
    hid.push_back(type ? type[i] : 0);
  }
  const int add = (int)hq.size();
  if (add == 0) return MDGPU_OK;
  CHK(ensure_owned(d, d->pn + add));
  // ghosts of an earlier build are dropped (Variables::AddParticle sets total = owned)
  HIPCHK(d, hipMemcpyAsync(d->pos + d->pn, hq.data(), add * sizeof(double4), hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipMemcpyAsync(d->mom + d->pn, hp.data(), add * sizeof(double4), hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipMemcpyAsync(d->id + d->pn, hid.data(), add * sizeof(int), hipMemcpyHostToDevice, d->stream));
  HIPCHK(d, hipMemsetAsync(d->rep + d->pn, 0, add * sizeof(int), d->stream));
  CHK(sync(d));
  d->pn += add;
  d->tn = d->pn;
  d->ng = 0;
  d->ghosts_ready = false;
  CHK(upload_replicas(d));
  d->vmax_valid = false;
  // indices of the old list are stale: the next step rebuilds
  d->initialized = false;
  return MDGPU_OK;
}

int mdgpu_particle_number(const mdgpu_domain *d, int *pn, int *tn) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (pn) *pn = d->pn;
  if (tn) *tn = d->tn;
  return MDGPU_OK;
}

// The owned particles' x, y, z (of q or p) between the device's 32-byte
// entries and a caller's array of 3 doubles per particle: packed or unpacked
// on the device in `xfer`, so that one copy moves exactly the caller's bytes
// and no host loop touches them.  The stream orders the reuse of `xfer`.
int ensure_xfer(mdgpu_domain *d, size_t count) {
  if (d->xfer && count <= d->xfer_cap) return MDGPU_OK;
  d->xfer_cap = 0;  // dalloc frees the old buffer first; a failed allocation leaves none
  CHK(dalloc(d, &d->xfer, count));
  d->xfer_cap = count;
  return MDGPU_OK;
}

int copy_out3(mdgpu_domain *d, const double4 *src, int n, double *dst) {
  CHK(ensure_xfer(d, 3 * (size_t)n));
  hipLaunchKernelGGL(k_pack3, dim3(nblk(n)), dim3(kBlock), 0, d->stream, src, n, d->xfer);
  HIPCHK(d, hipMemcpyAsync(dst, d->xfer, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  return MDGPU_OK;
}

int copy_in3(mdgpu_domain *d, const double *src, int n, double4 *dst) {
  CHK(ensure_xfer(d, 3 * (size_t)n));
  HIPCHK(d, hipMemcpyAsync(d->xfer, src, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, d->stream));
  hipLaunchKernelGGL(k_unpack3, dim3(nblk(n)), dim3(kBlock), 0, d->stream, d->xfer, n, dst);
  return MDGPU_OK;
}

int mdgpu_get_particles(mdgpu_domain *d, double *q, double *p, int *type) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  const int n = d->pn;
  if (n > 0) {
    if (q) CHK(copy_out3(d, d->pos, n, q));
    if (p) CHK(copy_out3(d, d->mom, n, p));
    if (type) HIPCHK(d, hipMemcpyAsync(type, d->id, n * sizeof(int), hipMemcpyDeviceToHost, d->stream));
  }
  CHK(check_launch(d, "copy to host"));
  return sync(d);
}

int mdgpu_set_momenta(mdgpu_domain *d, const double *p) {
  if (!d || !p) return MDGPU_ERR_ARGUMENT;
  invalidate(d);  // the next fused step recomputes its first half drift
  const int n = d->pn;
  if (n > 0) CHK(copy_in3(d, p, n, d->mom));
  CHK(check_launch(d, "copy from host"));
  CHK(sync(d));
  d->vmax_valid = false;
  return MDGPU_OK;
}

int mdgpu_set_positions(mdgpu_domain *d, const double *q) {
  if (!d || !q) return MDGPU_ERR_ARGUMENT;
  invalidate(d);  // the next fused step recomputes its first half drift
  const int n = d->pn;
  if (n > 0) CHK(copy_in3(d, q, n, d->pos));
  CHK(check_launch(d, "copy from host"));
  CHK(sync(d));
  return MDGPU_OK;
}

int mdgpu_clear_particles(mdgpu_domain *d) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (d->nrep > 1) {
    d->err = "an ensemble's particles are set when it is created";
    return MDGPU_ERR_ARGUMENT;
  }
  invalidate(d);  // the next fused step recomputes its first half drift
  CHK(sync(d));
  d->pn = d->ng = d->tn = 0;
  d->initialized = false;
  d->ghosts_ready = false;
  d->vmax_valid = false;
  return upload_replicas(d);
}

int mdgpu_get_ghosts(mdgpu_domain *d, int *src_type, int *shift_code, double *q) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  const int ng = d->ng, pn = d->pn;
  std::vector<int> gc(ng), gid;
  std::vector<double4> gq(ng);
  if (ng > 0) {
    HIPCHK(d, hipMemcpyAsync(gc.data(), d->gcode, ng * sizeof(int), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(d, hipMemcpyAsync(gq.data(), d->pos + pn, ng * sizeof(double4), hipMemcpyDeviceToHost, d->stream));
  }
  CHK(ghost_source_ids(d, gid));  // synchronises
  for (int k = 0; k < ng; k++) {
    if (src_type) src_type[k] = gid[k];
    if (shift_code) shift_code[k] = ghost_shift(gc[k]);
    if (q) {
      q[3 * (size_t)k] = gq[k].x;
      q[3 * (size_t)k + 1] = gq[k].y;
      q[3 * (size_t)k + 2] = gq[k].z;
    }
  }
  return MDGPU_OK;
}

int mdgpu_set_thermostat(mdgpu_domain *d, int kind, double T, double gamma, double tau) {
  if (!d || kind < MDGPU_NVE || kind > MDGPU_LANGEVIN) return MDGPU_ERR_ARGUMENT;
  invalidate(d);  // the next fused step recomputes its first half drift
  d->thermostat = kind;
  d->T_aim = T;
  d->gamma = gamma;
  d->tau = tau;
  if (d->nrep > 1)
    for (ReplicaState &S : d->rst) {
      S.T = T;
      S.gamma = gamma;
    }
  return upload_replicas(d);
}

int mdgpu_set_seed(mdgpu_domain *d, uint64_t seed) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (d->nrep > 1) {
    d->err = "seed the replicas of an ensemble with mdgpu_set_replica_thermostat";
    return MDGPU_ERR_ARGUMENT;
  }
  d->seed = seed;
  return upload_replicas(d);
}

int mdgpu_get_zeta(const mdgpu_domain *d, double *zeta) {
  if (!d || !zeta) return MDGPU_ERR_ARGUMENT;
  *zeta = d->zeta;
  return MDGPU_OK;
}

int mdgpu_replica_zetas(const mdgpu_domain *d, double *zeta) {
  if (!d || !zeta) return MDGPU_ERR_ARGUMENT;
  if (d->nrep == 1) {
    zeta[0] = d->zeta;
    return MDGPU_OK;
  }
  for (int r = 0; r < d->nrep; r++) zeta[r] = d->rst[r].zeta;
  return MDGPU_OK;
}

int mdgpu_get_time(const mdgpu_domain *d, double *sim_time, uint64_t *step) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (sim_time) *sim_time = d->sim_time;
  if (step) *step = d->step;
  return MDGPU_OK;
}

int mdgpu_make_pair_list(mdgpu_domain *d) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  return make_pair_list(d);
}

int mdgpu_is_pair_list_expired(mdgpu_domain *d, int *expired) {
  if (!d || !expired) return MDGPU_ERR_ARGUMENT;
  return check_expired(d, expired);
}

int mdgpu_update_position_half(mdgpu_domain *d) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  invalidate(d);  // the next fused step recomputes its first half drift
  return drift(d);
}

int mdgpu_send_border_particles(mdgpu_domain *d) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  invalidate(d);  // the next fused step recomputes its first half drift
  return is_unit(d) ? unit_refresh(d) : refresh(d);
}

int mdgpu_calculate_force(mdgpu_domain *d) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (!d->initialized) {
    d->err = "pair list not built";
    return MDGPU_ERR_ARGUMENT;
  }
  return force_in_place(d);
}

int mdgpu_langevin(mdgpu_domain *d) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  invalidate(d);  // the next fused step recomputes its first half drift
  if (d->pn > 0)
    hipLaunchKernelGGL(k_langevin, dim3(nblk(d->pn)), dim3(kBlock), 0, d->stream, d->mom, d->id, d->pn, d->rep,
                       d->reps, d->dt, (uint64_t)d->step);
  d->vmax_valid = false;
  d->step++;  // the next call draws the next numbers of each particle's stream
  d->clock_dirty = true;
  return check_launch(d, "langevin");
}

int mdgpu_set_zeta(mdgpu_domain *d, double zeta) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  d->zeta = zeta;
  return MDGPU_OK;
}

int mdgpu_heatbath_zeta(mdgpu_domain *d, double t) {
  if (!d || d->nrep > 1) return MDGPU_ERR_ARGUMENT;
  return heatbath_zeta(d, t);
}

int mdgpu_heatbath_momenta(mdgpu_domain *d) {
  if (!d || d->nrep > 1) return MDGPU_ERR_ARGUMENT;
  invalidate(d);  // the next fused step recomputes its first half drift
  return heatbath_momenta(d);
}

int mdgpu_calculate(mdgpu_domain *d, int nsteps, int *rebuilds) {
  if (!d || nsteps < 0) return MDGPU_ERR_ARGUMENT;
  if (is_unit(d)) {
    d->err = "mdgpu_calculate: the units of a group step through MDManager's sequence (the mdgpu_unit_* calls)";
    return MDGPU_ERR_ARGUMENT;
  }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  d->force_launches = 0;
  d->force_ms = d->total_ms = d->rebuild_ms = 0.0;
  d->rebuild_count = 0;
  if (d->timing) {
    const size_t need = (size_t)nsteps * 2 + 2;
    while (d->ev.size() < need) {
      hipEvent_t e;
      HIPCHK(d, hipEventCreate(&e));
      d->ev.push_back(e);
    }
    HIPCHK(d, hipEventCreate(&t0));
    HIPCHK(d, hipEventCreate(&t1));
    HIPCHK(d, hipEventRecord(t0, d->stream));
  }
  int r = MDGPU_OK;
  int s = 0;
  while (s < nsteps && r == MDGPU_OK) {
    // steps whose validity the budgets settle run ahead of the host; a step
    // that needs the host (the displacement check, a construction, an error)
    // is taken here, one at a time
    if (can_run_ahead(d) && nsteps - s > 1) {
      int k = 0;
      r = run_length(d, std::min(nsteps - s, d->run_ahead), &k);
      if (r != MDGPU_OK) break;
      if (k > 0) {
        int done = 0;
        r = run_gated(d, k, &done);
        s += done;
        if (r != MDGPU_OK || done == k) continue;
      }
    }
    if (s < nsteps) {
      r = calculate_step(d);
      s++;
    }
  }
  if (r == MDGPU_OK) r = sync(d);
  if (d->timing && t0) {
    HIPCHK(d, hipEventRecord(t1, d->stream));
    HIPCHK(d, hipEventSynchronize(t1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    d->total_ms = ms;
    for (int k = 0; k < d->force_launches; k++) {
      (void)hipEventElapsedTime(&ms, d->ev[2 * k], d->ev[2 * k + 1]);
      d->force_ms += ms;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
  }
  if (rebuilds) *rebuilds = d->rebuild_count;
  return r;
}

int mdgpu_observe(mdgpu_domain *d, double *ke, double *pe, double *virial) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (d->nrep > 1) {
    d->err = "observe the replicas of an ensemble with mdgpu_observe_replicas";
    return MDGPU_ERR_ARGUMENT;
  }
  if (!d->initialized) {
    d->err = "pair list not built";
    return MDGPU_ERR_ARGUMENT;
  }
  CHK(ensure_partial(d, std::max(1, d->nbrick)));
  double out[3];
  CHK(observe_into(d, d->partial, d->d3, out));
  if (ke) *ke = out[0];
  if (pe) *pe = out[1];
  if (virial) *virial = out[2] / 3.0;
  return MDGPU_OK;
}

int mdgpu_observe_noalloc(mdgpu_domain *d, double *ke, double *pe, double *virial) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (d->nrep > 1) {
    d->err = "observe the replicas of an ensemble with mdgpu_observe_replicas";
    return MDGPU_ERR_ARGUMENT;
  }
  if (!d->initialized) {
    d->err = "pair list not built";
    return MDGPU_ERR_ARGUMENT;
  }
  if (!d->scratch_dev) {
    d->err = "mdgpu_observe_noalloc: the device scratch array was not resolved when the domain was created";
    return MDGPU_ERR_HIP;
  }
  if (d->nbrick > kNoAllocBricks) {
    d->err = "mdgpu_observe_noalloc: " + std::to_string(d->nbrick) + " bricks exceed the scratch limit of " +
             std::to_string(kNoAllocBricks) + " (use mdgpu_observe)";
    return MDGPU_ERR_CAPACITY;
  }
  // every partial the kernel reads is written by it first: nothing to clear
  double out[3];
  CHK(observe_into(d, d->scratch_dev, d->scratch_dev + 3 * (size_t)kNoAllocBricks, out));
  if (ke) *ke = out[0];
  if (pe) *pe = out[1];
  if (virial) *virial = out[2] / 3.0;
  return MDGPU_OK;
}

int mdgpu_observe_kinetic(mdgpu_domain *d, double *ke) {
  if (!d || !ke || d->nrep > 1) return MDGPU_ERR_ARGUMENT;
  return kinetic(d, ke);
}

int mdgpu_change_scale(mdgpu_domain *d, double alpha) {
  if (!d || !(alpha > 0.0)) return MDGPU_ERR_ARGUMENT;
  if (is_unit(d)) {
    d->err = "mdgpu_change_scale: a unit of a process grid does not rescale";
    return MDGPU_ERR_ARGUMENT;
  }
  invalidate(d);  // the next fused step recomputes its first half drift
  CHK(make_pair_list(d));
  if (d->nrep == 1) {
    for (int k = 0; k < 3; k++) {
      d->L[k] = d->L[k] * alpha;
      d->rs[k] *= alpha;
      d->re[k] *= alpha;
    }
    // bricks picked from the density are picked again for the new one (at the
    // construction below), as an ensemble's replicas' are
    if (!d->brick_user) d->brick_auto = true;
  } else {
    // every replica, as MDManager::ChangeScale does to its one system; each
    // replica's bricks are chosen again for its new density, as at creation
    for (int r = 0; r < d->nrep; r++) {
      Replica &R = d->rh[r];
      for (int k = 0; k < 3; k++) R.L[k] = R.L[k] * alpha;
      const double rho = R.pn / (R.L[0] * R.L[1] * R.L[2]);
      auto_brick(d, rho, R.L, brick_budget(d), R.bd, kReplicaFill);
    }
    for (int k = 0; k < 3; k++) d->L[k] = d->rh[0].L[k];
  }
  if (d->pn > 0)
    hipLaunchKernelGGL(k_scale_pos, dim3(nblk(d->pn)), dim3(kBlock), 0, d->stream, d->pos, d->pn, alpha);
  CHK(check_launch(d, "scale"));
  CHK(set_geometry(d));
  return make_pair_list(d);
}

}  // extern "C"

namespace {

// Canonical pairs of the rows of replica rsel (all rows when rsel < 0).
int collect_pairs(mdgpu_domain *d, int rsel, int64_t *npairs, int32_t *out3) {
  if (!d->initialized) {
    d->err = "pair list not built";
    return MDGPU_ERR_ARGUMENT;
  }
  const int pn = d->pn, ng = d->ng, nbk = d->nbrick;
  std::vector<uint16_t> nbr((size_t)d->nbr_slots);
  std::vector<long long> sb(d->nslice), gb(nbk + 1);
  std::vector<int> sl(d->nslice), ids(pn), cl(pn), rp(pn), r0(nbk), vs0(nbk), gc(ng), gid;
  if (d->nbr_slots) HIPCHK(d, hipMemcpyAsync(nbr.data(), d->nbr, nbr.size() * 2, hipMemcpyDeviceToHost, d->stream));
  if (d->nslice) {
    HIPCHK(d, hipMemcpyAsync(sb.data(), d->sbase, d->nslice * 8, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(d, hipMemcpyAsync(sl.data(), d->slen, d->nslice * 4, hipMemcpyDeviceToHost, d->stream));
  }
  if (pn) {
    HIPCHK(d, hipMemcpyAsync(ids.data(), d->id, pn * 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(d, hipMemcpyAsync(cl.data(), d->cell, pn * 4, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(d, hipMemcpyAsync(rp.data(), d->rep, pn * 4, hipMemcpyDeviceToHost, d->stream));
  }
  if (ng) HIPCHK(d, hipMemcpyAsync(gc.data(), d->gcode, ng * 4, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipMemcpyAsync(r0.data(), d->row0, nbk * 4, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipMemcpyAsync(vs0.data(), d->vslice0, nbk * 4, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipMemcpyAsync(gb.data(), d->gbase, (nbk + 1) * 8, hipMemcpyDeviceToHost, d->stream));
  CHK(ghost_source_ids(d, gid));  // synchronises
  std::vector<int> gl((size_t)gb[nbk]);
  if (!gl.empty()) HIPCHK(d, hipMemcpy(gl.data(), d->glist, gl.size() * 4, hipMemcpyDeviceToHost));
  int64_t cnt = 0;
  for (int i = 0; i < pn; i++) {
    if (rsel >= 0 && rp[i] != rsel) continue;
    const int b = d->h_brick_of_cell[cl[i]];
    const int v = i - r0[b];
    const int s = vs0[b] + v / kWave, lane = v % kWave;
    const long long total = gb[b + 1] - gb[b];
    for (int k = 0; k < sl[s]; k++) {
      const int e = nbr[sb[s] + slot_of(k, lane)];
      if (e >= total) continue;  // padding
      const int j = gl[gb[b] + e];
      const int a = ids[i];
      int bb, code;
      if (j < pn) {
        bb = ids[j];
        code = 13;
        if (bb < a) continue;  // owned-owned pairs appear in both rows
      } else {
        bb = gid[j - pn];
        code = ghost_shift(gc[j - pn]);
      }
      if (out3) {
        out3[3 * cnt] = a;
        out3[3 * cnt + 1] = bb;
        out3[3 * cnt + 2] = code;
      }
      cnt++;
    }
  }
  *npairs = cnt;
  return MDGPU_OK;
}

}  // namespace

extern "C" {

int mdgpu_get_pairs(mdgpu_domain *d, int64_t *npairs, int32_t *out3) {
  if (!d || !npairs) return MDGPU_ERR_ARGUMENT;
  if (d->nrep > 1) {
    d->err = "pairs of an ensemble are read per replica (mdgpu_get_replica_pairs)";
    return MDGPU_ERR_ARGUMENT;
  }
  return collect_pairs(d, -1, npairs, out3);
}

int mdgpu_get_replica_pairs(mdgpu_domain *d, int replica, int64_t *npairs, int32_t *out3) {
  if (!d || !npairs || replica < 0 || replica >= d->nrep) return MDGPU_ERR_ARGUMENT;
  return collect_pairs(d, replica, npairs, out3);
}

int mdgpu_get_pair_indices(mdgpu_domain *d, int64_t *npairs, int32_t *key, int32_t *partner) {
  if (!d || !npairs) return MDGPU_ERR_ARGUMENT;
  if (!d->initialized) {
    d->err = "pair list not built";
    return MDGPU_ERR_ARGUMENT;
  }
  const int pn = d->pn, nbk = d->nbrick;
  std::vector<uint16_t> nbr((size_t)d->nbr_slots);
  std::vector<long long> sb(d->nslice), gb(nbk + 1);
  std::vector<int> sl(d->nslice), cl(pn), r0(nbk), vs0(nbk);
  if (d->nbr_slots) HIPCHK(d, hipMemcpyAsync(nbr.data(), d->nbr, nbr.size() * 2, hipMemcpyDeviceToHost, d->stream));
  if (d->nslice) {
    HIPCHK(d, hipMemcpyAsync(sb.data(), d->sbase, d->nslice * 8, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(d, hipMemcpyAsync(sl.data(), d->slen, d->nslice * 4, hipMemcpyDeviceToHost, d->stream));
  }
  if (pn) HIPCHK(d, hipMemcpyAsync(cl.data(), d->cell, pn * 4, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipMemcpyAsync(r0.data(), d->row0, nbk * 4, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipMemcpyAsync(vs0.data(), d->vslice0, nbk * 4, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipMemcpyAsync(gb.data(), d->gbase, (nbk + 1) * 8, hipMemcpyDeviceToHost, d->stream));
  CHK(sync(d));
  std::vector<int> gl((size_t)gb[nbk]);
  if (!gl.empty()) HIPCHK(d, hipMemcpy(gl.data(), d->glist, gl.size() * 4, hipMemcpyDeviceToHost));
  int64_t cnt = 0;
  for (int i = 0; i < pn; i++) {
    const int b = d->h_brick_of_cell[cl[i]];
    const int v = i - r0[b];
    const int s = vs0[b] + v / kWave, lane = v % kWave;
    const long long total = gb[b + 1] - gb[b];
    for (int k = 0; k < sl[s]; k++) {
      const int e = nbr[sb[s] + slot_of(k, lane)];
      if (e >= total) continue;  // padding
      const int j = gl[gb[b] + e];
      if (j < pn && j < i) continue;  // owned-owned pairs once, as (smaller, larger)
      if (key) key[cnt] = i;
      if (partner) partner[cnt] = j;
      cnt++;
    }
  }
  *npairs = cnt;
  return MDGPU_OK;
}

int mdgpu_compute_kick(mdgpu_domain *d, double *dp) {
  if (!d || !dp) return MDGPU_ERR_ARGUMENT;
  if (!d->initialized) {
    d->err = "pair list not built";
    return MDGPU_ERR_ARGUMENT;
  }
  const int n = d->pn;
  if (n == 0) return MDGPU_OK;
  HIPCHK(d, hipMemcpyAsync(d->mom_alt, d->mom, n * sizeof(double4), hipMemcpyDeviceToDevice, d->stream));
  HIPCHK(d, hipMemsetAsync(d->mom, 0, n * sizeof(double4), d->stream));
  CHK(force_in_place(d));
  std::vector<double4> hp(n);
  HIPCHK(d, hipMemcpyAsync(hp.data(), d->mom, n * sizeof(double4), hipMemcpyDeviceToHost, d->stream));
  HIPCHK(d, hipMemcpyAsync(d->mom, d->mom_alt, n * sizeof(double4), hipMemcpyDeviceToDevice, d->stream));
  CHK(sync(d));
  for (int i = 0; i < n; i++) {
    dp[3 * (size_t)i] = hp[i].x;
    dp[3 * (size_t)i + 1] = hp[i].y;
    dp[3 * (size_t)i + 2] = hp[i].z;
  }
  return MDGPU_OK;
}

int mdgpu_enable_timing(mdgpu_domain *d, int on) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  d->timing = on != 0;
  return MDGPU_OK;
}

int mdgpu_get_timing(const mdgpu_domain *d, int *force_launches, double *force_ms, double *total_ms,
                     int *rebuilds, double *rebuild_ms) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (force_launches) *force_launches = d->force_launches;
  if (force_ms) *force_ms = d->force_ms;
  if (total_ms) *total_ms = d->total_ms;
  if (rebuilds) *rebuilds = d->rebuild_count;
  if (rebuild_ms) *rebuild_ms = d->rebuild_ms;
  return MDGPU_OK;
}

int mdgpu_set_option(mdgpu_domain *d, int option, int value) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  invalidate(d);  // the next fused step recomputes its first half drift
  switch (option) {
    case MDGPU_OPT_UNROLL:
      if (value != 8 && value != 16 && value != 32) return MDGPU_ERR_ARGUMENT;
      d->unroll = value;
      return MDGPU_OK;
    case MDGPU_OPT_KAHAN:
      if (value < kSumPlain || value > kSumGroup64) return MDGPU_ERR_ARGUMENT;
      d->kahan = value;
      return MDGPU_OK;
    case MDGPU_OPT_WAVES:
      if (value != 2 && value != 4 && value != 8 && value != 16) return MDGPU_ERR_ARGUMENT;
      d->waves = value;
      return MDGPU_OK;
    case MDGPU_OPT_PIPELINE:
      if (value != 0 && value != 1) return MDGPU_ERR_ARGUMENT;
      d->pipeline = value;
      return MDGPU_OK;
    case MDGPU_OPT_MIXED_FRAME:
      if (value != 0 && value != 1) return MDGPU_ERR_ARGUMENT;
      d->mixed_frame = value;
      return recode_list(d);
    case MDGPU_OPT_DISPLACEMENT_CHECK:
      if (value != 0 && value != 1) return MDGPU_ERR_ARGUMENT;
      d->disp_check = value;
      return MDGPU_OK;
    case MDGPU_OPT_LDS_KB:
      if (value < 8) return MDGPU_ERR_ARGUMENT;
      d->lds_kb = value;
      return MDGPU_OK;
    case MDGPU_OPT_BRICK: {
      const int bx = value & 0xff, by = (value >> 8) & 0xff, bz = (value >> 16) & 0xff;
      if (bx < 1 || bx > kMaxBrickEdge || by < 1 || by > kMaxBrickEdge || bz < 1 || bz > kMaxBrickEdge ||
          (bx + 2) * (by + 2) * (bz + 2) > kMaxBrickCells)
        return MDGPU_ERR_ARGUMENT;
      d->brick_auto = false;
      d->brick_user = true;
      d->bd[0] = bx;
      d->bd[1] = by;
      d->bd[2] = bz;
      for (Replica &R : d->rh) {
        R.bd[0] = bx;
        R.bd[1] = by;
        R.bd[2] = bz;
      }
      return set_geometry(d);  // the next step rebuilds the list
    }
    case MDGPU_OPT_GRAPH:
      d->use_graph = value != 0;
      return MDGPU_OK;
    case MDGPU_OPT_SEARCH_WAVES:
      if (value != 4 && value != 8 && value != 16) return MDGPU_ERR_ARGUMENT;
      d->search_waves = value;
      return MDGPU_OK;
    case MDGPU_OPT_SEARCH_STAGE:
      d->search_stage = value != 0;
      return MDGPU_OK;
    case MDGPU_OPT_PRUNE:
      if (value < 0 || value > 1000) return MDGPU_ERR_ARGUMENT;
      d->prune_delta = value * 1.0e-3;
      return MDGPU_OK;
    case MDGPU_OPT_ROW_SORT:
      if (value < 0 || value > 64) return MDGPU_ERR_ARGUMENT;
      d->row_sort = value;
      d->initialized = false;  // the next step rebuilds the list
      return MDGPU_OK;
    case MDGPU_OPT_SEARCH_CELLS:
      d->search_cells = value != 0;
      return MDGPU_OK;
    case MDGPU_OPT_RUN_AHEAD:
      if (value < 0) return MDGPU_ERR_ARGUMENT;
      d->run_ahead = value;
      return MDGPU_OK;
    case MDGPU_OPT_BRICK_KB:
      if (value < 8 || value > 160) return MDGPU_ERR_ARGUMENT;
      d->brick_kb = value;
      if (d->nrep == 1) {
        d->brick_user = false;
        d->brick_auto = true;  // chosen at the next construction, which the next step runs
        d->initialized = false;
        return MDGPU_OK;
      }
      for (Replica &R : d->rh) auto_brick(d, R.pn / (R.L[0] * R.L[1] * R.L[2]), R.L, value, R.bd, kReplicaFill);
      return set_geometry(d);  // the next step rebuilds the list
  }
  return MDGPU_ERR_ARGUMENT;
}

int mdgpu_get_option(const mdgpu_domain *d, int option, int *value) {
  if (!d || !value) return MDGPU_ERR_ARGUMENT;
  switch (option) {
    case MDGPU_OPT_UNROLL: *value = unroll_of(d); return MDGPU_OK;
    case MDGPU_OPT_KAHAN: *value = d->kahan; return MDGPU_OK;
    case MDGPU_OPT_WAVES: *value = d->waves; return MDGPU_OK;
    case MDGPU_OPT_LDS_KB: *value = std::min(force_lds_kb(d), d->lds_max_kb); return MDGPU_OK;
    case MDGPU_OPT_BRICK: *value = d->bd[0] | (d->bd[1] << 8) | (d->bd[2] << 16); return MDGPU_OK;
    case MDGPU_OPT_GRAPH: *value = d->use_graph ? 1 : 0; return MDGPU_OK;
    case MDGPU_OPT_SEARCH_WAVES: *value = d->search_waves; return MDGPU_OK;
    case MDGPU_OPT_SEARCH_STAGE: *value = d->search_stage; return MDGPU_OK;
    case MDGPU_OPT_PRUNE: *value = (int)std::lround(d->prune_delta * 1.0e3); return MDGPU_OK;
    case MDGPU_OPT_ROW_SORT: *value = d->row_sort; return MDGPU_OK;
    case MDGPU_OPT_BRICK_KB: *value = brick_budget(d); return MDGPU_OK;
    case MDGPU_OPT_SEARCH_CELLS: *value = d->search_cells; return MDGPU_OK;
    case MDGPU_OPT_RUN_AHEAD: *value = d->run_ahead; return MDGPU_OK;
    case MDGPU_OPT_PIPELINE: *value = d->pipeline; return MDGPU_OK;
    case MDGPU_OPT_MIXED_FRAME: *value = d->mixed_frame; return MDGPU_OK;
    case MDGPU_OPT_DISPLACEMENT_CHECK: *value = d->disp_check; return MDGPU_OK;
  }
  return MDGPU_ERR_ARGUMENT;
}

int mdgpu_get_list_stats(const mdgpu_domain *d, int64_t *list_entries, int64_t *list_slots, int *ncells,
                         int *ghosts) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (list_entries) *list_entries = d->nbr_entries;
  if (list_slots) *list_slots = d->nbr_slots;
  if (ncells) *ncells = d->ncell;
  if (ghosts) *ghosts = d->ng;
  return MDGPU_OK;
}

int mdgpu_replica_number(const mdgpu_domain *d, int *nrep) {
  if (!d || !nrep) return MDGPU_ERR_ARGUMENT;
  *nrep = d->nrep;
  return MDGPU_OK;
}

int mdgpu_replica_counts(const mdgpu_domain *d, int *counts) {
  if (!d || !counts) return MDGPU_ERR_ARGUMENT;
  if (d->nrep == 1) counts[0] = d->pn;
  else
    for (int r = 0; r < d->nrep; r++) counts[r] = d->rh[r].pn;
  return MDGPU_OK;
}

int mdgpu_set_replica_thermostat(mdgpu_domain *d, int replica, double T, double gamma, uint64_t seed) {
  if (!d || replica < 0 || replica >= d->nrep) return MDGPU_ERR_ARGUMENT;
  if (d->nrep == 1) {
    d->T_aim = T;
    d->gamma = gamma;
    d->seed = seed;
  } else {
    d->rst[replica].T = T;
    d->rst[replica].gamma = gamma;
    d->rst[replica].seed = seed;
  }
  return upload_replicas(d);
}

int mdgpu_observe_replicas(mdgpu_domain *d, double *ke, double *pe, double *virial) {
  if (!d) return MDGPU_ERR_ARGUMENT;
  if (!d->initialized) {
    d->err = "pair list not built";
    return MDGPU_ERR_ARGUMENT;
  }
  const int nb = d->nbrick, R = d->nrep;
  CHK(ensure_partial(d, std::max(1, nb)));
  if (d->pn > 0) {
    launch_observe(d, d->partial);
    hipLaunchKernelGGL(k_replica_sum3, dim3(R), dim3(kBlock), 0, d->stream, d->partial, d->reps, d->rep_out);
    CHK(check_launch(d, "observe"));
    HIPCHK(d, hipMemcpyAsync(d->h_rep_out, d->rep_out, 3 * R * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    CHK(sync(d));
  } else {
    for (int k = 0; k < 3 * R; k++) d->h_rep_out[k] = 0.0;
  }
  for (int r = 0; r < R; r++) {
    if (ke) ke[r] = d->h_rep_out[3 * r];
    if (pe) pe[r] = d->h_rep_out[3 * r + 1];
    if (virial) virial[r] = d->h_rep_out[3 * r + 2] / 3.0;
  }
  return MDGPU_OK;
}

}  // extern "C"
