/*
 * mdgpu.h -- C ABI of the MI355X (gfx950) implementation of MDACP's
 * short-range Lennard-Jones step: cell binning, Verlet pair list with a search
 * margin, smoothed-LJ force kick, position-Verlet / Langevin / Nose-Hoover
 * updates with periodic images, and the observers.
 *
 * One `mdgpu_domain` is one MDACP domain (one MDUnit: its Variables, MeshList,
 * PairList and the ghost images of its box) resident in HBM, with its own HIP
 * stream.  Plain pointers and sizes only; host arrays are xyz triplets
 * (n*3 doubles).  Every function returns MDGPU_OK or an error code; the text
 * of the last error of a domain is available from mdgpu_last_error().
 *
 * The entry points replace these reference interfaces (kaityo256/mdacp):
 *   mdgpu_create            Variables::Variables (variables.cc:16-28),
 *                           SimulationInfo (simulationinfo.cc:5-51),
 *                           MeshList::MeshList/ChangeScale (meshlist.cc:8-55)
 *   mdgpu_add_particles     MDUnit::AddParticle (mdunit.cc:36-49)
 *   mdgpu_set_momenta       writes to Variables::p (variables.h:18)
 *   mdgpu_get_particles     reads Variables::q/p/type (variables.h:16-18)
 *   mdgpu_make_pair_list    MDManager::MakePairList (mdmanager.cc:262-280) on one
 *                           domain: SendParticles wrap, FindBorderParticles,
 *                           MDUnit::MakePairList -> PairList::Init + MeshList::MakeList
 *   mdgpu_is_pair_list_expired  PairList::IsPairListExpired (pairlist.cc:26-121)
 *   mdgpu_update_position_half  ForceCalculator::UpdatePositionHalf (fcalculator.cc:22-32)
 *   mdgpu_send_border_particles MDManager::SendBorderParticles (mdmanager.cc:283-348)
 *   mdgpu_calculate_force   ForceCalculator::CalculateForce (fcalculator.cc:35-50,
 *                           semantics of CalculateForceNext :228-290)
 *   mdgpu_langevin          ForceCalculator::Langevin (fcalculator.cc:765-778); the noise is
 *                           Philox4x32-10 keyed by (seed, particle id, call count), not
 *                           the unit's std::mt19937 stream
 *   mdgpu_heatbath_zeta     ForceCalculator::HeatbathZeta (fcalculator.cc:744-749)
 *   mdgpu_heatbath_momenta  ForceCalculator::HeatbathMomenta (fcalculator.cc:752-762)
 *   mdgpu_calculate         MDManager::Calculate (mdmanager.cc:116-192), n steps
 *   mdgpu_observe           KineticEnergyObserver, PotentialEnergyObserver,
 *                           VirialObserver (observer.cc:5-79)
 *   mdgpu_change_scale      MDManager::ChangeScale (mdmanager.cc:432-443)
 *   mdgpu_get_pairs         MeshList::GetKeyParticles/GetPartnerParticles
 *                           (meshlist.h:46-48), as canonical pairs
 *
 * Ensembles (no counterpart in the reference, which runs one system per
 * process): mdgpu_create_ensemble makes one domain that advances many
 * independent replicas -- each its own periodic box, particle count,
 * temperature and Langevin seed, every one with the single-domain semantics
 * above -- with one pair-force launch per step for all of them.  Each replica
 * keeps its own PairList budget and displacement check (pairlist.cc:26-121);
 * when the list of any replica expires, the lists of all replicas are rebuilt.
 * NVE, Langevin or Nose-Hoover run in every replica with its own parameters;
 * mdgpu_change_scale rescales every replica.
 *
 * Units of a process grid (MDManager with several MDUnits in one process,
 * mdmanager.cc:34-45, e.g. OMP_NUM_THREADS > 1): mdgpu_create_unit makes one
 * domain per unit, each the box [s, e) of the grid, joined in a group that
 * shares one device and one stream.  Particles migrate and ghosts are placed
 * device to device between the units of the group:
 *   mdgpu_unit_send_particles     MDUnit::MakeBufferForSendingParticle (mdunit.cc:78-101)
 *   mdgpu_unit_receive_particles  MDUnit::ReceiveParticles (mdunit.cc:104-122) from the
 *                                 sending unit's buffer
 *   mdgpu_unit_adjust_periodic_boundary  MDUnit::AdjustPeriodicBoundary
 *                                 (variables.cc:59-79), then the unit's owned particles
 *                                 binned (MeshList::MakeMesh, meshlist.cc:153-205)
 *   mdgpu_unit_find_border_particles  the six FindBorderParticles /
 *                                 MakeBufferForBorderParticles / ReceiveBorderParticles
 *                                 passes of MDManager::MakePairList (mdmanager.cc:269-275,
 *                                 mdunit.cc:125-191): every ghost the relay would bring,
 *                                 straight from the unit that owns the particle
 *   mdgpu_make_pair_list          MDUnit::MakePairList (mdunit.cc:194-198) of the unit
 *   mdgpu_send_border_particles   MDManager::SendBorderParticles (mdmanager.cc:283-348)
 *                                 for the unit: its ghosts from their owners' positions
 *   mdgpu_group_make_pair_list    MDManager::MakePairList (mdmanager.cc:262-280) for every
 *                                 unit of the group, in the reference's order
 *   mdgpu_group_calculate         MDManager::Calculate (mdmanager.cc:116-192) for every unit
 *                                 of the group, n steps (the units' kicks in one launch)
 * MDManager calls each unit's operations between its phase barriers
 * (omp parallel for): a unit's work is queued on the group's stream after the
 * work of every unit's previous phase.  The other step operations
 * (mdgpu_update_position_half, mdgpu_calculate_force, the thermostats,
 * mdgpu_observe) act on one unit as on a single domain; the fused
 * mdgpu_calculate and mdgpu_change_scale are refused for a unit.
 */
#ifndef MDGPU_H
#define MDGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdgpu_domain mdgpu_domain;

/* status codes */
#define MDGPU_OK 0
#define MDGPU_ERR_TOO_FAST 1         /* "Too fast particles exists" (pairlist.cc:58-62) */
#define MDGPU_ERR_INVALID_POSITION 2 /* "Invalid position" (variables.cc:70-78) */
#define MDGPU_ERR_INVALID_INDEX 3    /* "Invalid index" (meshlist.cc:178-182) */
#define MDGPU_ERR_CAPACITY 4         /* particle or pair capacity exceeded */
#define MDGPU_ERR_HIP 5              /* a HIP runtime call failed */
#define MDGPU_ERR_ARGUMENT 6         /* invalid argument */

/* force precision */
#define MDGPU_FP64 0   /* FP64 throughout */
#define MDGPU_MIXED 1  /* FP32 pair forces from FP32 coordinates relative to the row's
                          cell brick (or its cell, MDGPU_OPT_MIXED_FRAME), FP32 sums of
                          each 8 list entries added in FP64 (MDGPU_OPT_KAHAN), FP64
                          positions, momenta, integration and reductions */

/* thermostat (SimulationInfo::ControlTemperature / HeatbathType) */
#define MDGPU_NVE 0
#define MDGPU_NOSEHOOVER 1
#define MDGPU_LANGEVIN 2

typedef struct {
  double L[3];            /* box edges (SimulationInfo::L) */
  double dt;              /* TimeStep */
  double rc;              /* cutoff (CUTOFF_LENGTH, mdconfig.h:15) */
  double margin;          /* BufferLength (simulationinfo.cc:6) */
  int check_list_length;  /* CheckListLength (simulationinfo.cc:29), 500 */
  int precision;          /* MDGPU_FP64 or MDGPU_MIXED */
  int capacity;           /* initial owned-particle capacity (grows on demand) */
  int device;             /* HIP device ordinal */
} mdgpu_params;

/* one replica of an ensemble: box, Langevin temperature, friction and key */
typedef struct {
  double L[3];
  double T;
  double gamma;
  uint64_t seed;
} mdgpu_replica;

/* lifetime */
int mdgpu_create(const mdgpu_params *prm, mdgpu_domain **out);
/* An ensemble of nrep replicas sharing prm's dt, rc, margin, check list length,
   precision and device (prm->L, prm->capacity unused); thermostat MDGPU_NVE,
   MDGPU_LANGEVIN or MDGPU_NOSEHOOVER (each replica with its own temperature,
   target T and zeta; tau from mdgpu_set_thermostat, 0.1 by default).  Replica r owns counts[r] particles, given replica after
   replica in q, p (n*3 doubles; p may be NULL) and type (NULL: 0..counts[r]-1),
   each inside [0, L) of its replica.  The pair lists are built before return. */
int mdgpu_create_ensemble(const mdgpu_params *prm, int thermostat, int nrep, const mdgpu_replica *reps,
                          const int *counts, const double *q, const double *p, const int *type,
                          mdgpu_domain **out);
/* One unit of a process grid (ParaInfo's grid, parainfo.cc): units per axis,
   the unit's grid position and its box [s, e) (MDUnit::MDUnit, mdunit.cc:8-27). */
typedef struct {
  int grid[3];
  int pos[3];
  double s[3];
  double e[3];
} mdgpu_unit;

/* A unit of the grid, with prm's parameters (prm->L the whole box).  join:
   NULL for the first unit of a group, else any unit of the group (same grid,
   box and device, another grid position; at most 64 units).  Each edge of the
   unit's box must hold two search lengths (rc + margin). */
int mdgpu_create_unit(const mdgpu_params *prm, const mdgpu_unit *u, mdgpu_domain *join, mdgpu_domain **out);
/* dir: D_LEFT, D_RIGHT, D_BACK, D_FORWARD, D_DOWN, D_UP = 0..5 (mdconfig.h:25);
   nsend: particles now in the unit's send buffer */
int mdgpu_unit_send_particles(mdgpu_domain *d, int dir, int *nsend);
int mdgpu_unit_receive_particles(mdgpu_domain *d, mdgpu_domain *src);
int mdgpu_unit_adjust_periodic_boundary(mdgpu_domain *d);
/* needs every unit of the group through mdgpu_unit_adjust_periodic_boundary */
int mdgpu_unit_find_border_particles(mdgpu_domain *d);
int mdgpu_group_make_pair_list(mdgpu_domain *d);
/* nsteps of MDManager::Calculate (mdmanager.cc:116-192) for every unit of d's
   group: each unit's expiry test (OR'ed), the group's construction when any
   expired, then the half drifts, halo and force of every unit, with the
   thermostat of d (NVE, Langevin, or Nose-Hoover on the group's temperature).
   With mdgpu_enable_timing on d, every unit records its force launches
   (mdgpu_get_timing of each unit) and d the wall time of the steps. */
int mdgpu_group_calculate(mdgpu_domain *d, int nsteps, int *rebuilds);

/* replicas of a domain (1 for mdgpu_create) and owned particles of each;
   mdgpu_get_particles returns them replica after replica */
int mdgpu_replica_number(const mdgpu_domain *d, int *nrep);
int mdgpu_replica_counts(const mdgpu_domain *d, int *counts);
/* Langevin temperature, friction and key of one replica (its Nose-Hoover target too) */
int mdgpu_set_replica_thermostat(mdgpu_domain *d, int replica, double T, double gamma, uint64_t seed);
/* the Nose-Hoover zeta of every replica (Variables::Zeta of each system) */
int mdgpu_replica_zetas(const mdgpu_domain *d, double *zeta);
int mdgpu_destroy(mdgpu_domain *d);
const char *mdgpu_last_error(const mdgpu_domain *d);
int mdgpu_set_precision(mdgpu_domain *d, int precision);

/* state (host arrays, n*3 doubles; type = canonical particle id) */
int mdgpu_add_particles(mdgpu_domain *d, int n, const double *q, const double *p, const int *type);
int mdgpu_particle_number(const mdgpu_domain *d, int *pn, int *tn);
int mdgpu_get_particles(mdgpu_domain *d, double *q, double *p, int *type);
int mdgpu_set_momenta(mdgpu_domain *d, const double *p); /* in mdgpu_get_particles order */
/* overwrite owned positions in mdgpu_get_particles order (writes to Variables::q,
   variables.h:17, by host code between steps); ghosts follow at the next refresh */
int mdgpu_set_positions(mdgpu_domain *d, const double *q);
/* drop all particles (Variables::SetParticleNumber(0), variables.h:30) */
int mdgpu_clear_particles(mdgpu_domain *d);
/* ghost images: owner id, shift code, position (the ghost part of Variables::q,
   indices [pn, tn) in the reference) */
int mdgpu_get_ghosts(mdgpu_domain *d, int *src_type, int *shift_code, double *q);

/* thermostat and run parameters */
int mdgpu_set_thermostat(mdgpu_domain *d, int kind, double T, double gamma, double tau);
int mdgpu_set_seed(mdgpu_domain *d, uint64_t seed);
int mdgpu_get_zeta(const mdgpu_domain *d, double *zeta);
/* Variables::Zeta (variables.h:19), used by mdgpu_heatbath_momenta */
int mdgpu_set_zeta(mdgpu_domain *d, double zeta);
int mdgpu_get_time(const mdgpu_domain *d, double *sim_time, uint64_t *step);

/* the step, piece by piece (same order as MDManager::Calculate) */
int mdgpu_make_pair_list(mdgpu_domain *d);
int mdgpu_is_pair_list_expired(mdgpu_domain *d, int *expired);
int mdgpu_update_position_half(mdgpu_domain *d);
int mdgpu_send_border_particles(mdgpu_domain *d);
int mdgpu_calculate_force(mdgpu_domain *d);
int mdgpu_langevin(mdgpu_domain *d);
int mdgpu_heatbath_zeta(mdgpu_domain *d, double current_temperature);
int mdgpu_heatbath_momenta(mdgpu_domain *d);

/* n fused steps of MDManager::Calculate; *rebuilds (nullable) receives the
   number of pair-list constructions done inside */
int mdgpu_calculate(mdgpu_domain *d, int nsteps, int *rebuilds);

/* observers: raw sums as returned by the reference observers */
int mdgpu_observe(mdgpu_domain *d, double *ke, double *pe, double *virial);
/* KineticEnergyObserver alone (observer.cc:5-15); needs no pair list */
int mdgpu_observe_kinetic(mdgpu_domain *d, double *ke);
/* mdgpu_observe without any device allocation, for callers that cannot
   allocate (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  No counterpart in the reference.  Its per-brick partial
   sums live in a fixed device array of the library (3 MiB per GPU, held until
   the process ends) instead of the domain's growable buffer: a domain of more
   than 131072 bricks gets MDGPU_ERR_CAPACITY (use mdgpu_observe).  Every call
   on one GPU shares that array, so calls that overlap on one GPU race.  Same
   kernels and summation order as mdgpu_observe: the same bits. */
int mdgpu_observe_noalloc(mdgpu_domain *d, double *ke, double *pe, double *virial);
/* the three observers of every replica (arrays of mdgpu_replica_number entries) */
int mdgpu_observe_replicas(mdgpu_domain *d, double *ke, double *pe, double *virial);

/* MDManager::ChangeScale: rebuild, scale box and owned positions by alpha, rebuild */
int mdgpu_change_scale(mdgpu_domain *d, double alpha);

/* parity export: canonical pairs (owned id, partner owner id, partner image
   shift code (sx+1)+3(sy+1)+9(sz+1)); owned-owned pairs as (min id, max id, 13).
   Call with out3 == NULL to get the count. */
int mdgpu_get_pairs(mdgpu_domain *d, int64_t *npairs, int32_t *out3);
/* the same for the rows of one replica (ids are those given for that replica) */
int mdgpu_get_replica_pairs(mdgpu_domain *d, int replica, int64_t *npairs, int32_t *out3);
/* the same pairs by index, as MeshList::GetKeyParticles/GetPartnerParticles
   hold them (meshlist.h:46-48): key = owned index, partner = owned index
   (key < partner) or pn + ghost index, in mdgpu_get_particles / mdgpu_get_ghosts
   order.  Call with key == partner == NULL to get the count. */
int mdgpu_get_pair_indices(mdgpu_domain *d, int64_t *npairs, int32_t *key, int32_t *partner);
/* kick of CalculateForce from zero momenta, per owned particle in
   mdgpu_get_particles order (dp = F*dt) */
int mdgpu_compute_kick(mdgpu_domain *d, double *dp);

/* timing of the last mdgpu_calculate call, from HIP events recorded on the
   domain's stream: count and total milliseconds of the pair-force launches and
   of the whole step sequence.  Enabled by mdgpu_enable_timing(d, 1). */
int mdgpu_enable_timing(mdgpu_domain *d, int on);
int mdgpu_get_timing(const mdgpu_domain *d, int *force_launches, double *force_ms, double *total_ms,
                     int *rebuilds, double *rebuild_ms);

/* tuning knobs (no counterpart in the reference; results do not depend on
   them beyond floating-point summation order) */
#define MDGPU_OPT_UNROLL 1   /* list entries whose indices the force kernels load ahead of
                                the 8 being computed: 8, 16 or 32 (default 16 for FP64,
                                8 for mixed precision and for a group's one kick of all
                                its units) */
#define MDGPU_OPT_KAHAN 2    /* mixed precision row sums: 2 = FP32 sums of each 8 list
                                entries added in FP64 (default), 1 = compensated FP32
                                (Kahan), 0 = plain FP32 */
#define MDGPU_OPT_WAVES 3    /* wavefronts per cell-brick workgroup: 2, 4, 8 or 16 */
#define MDGPU_OPT_LDS_KB 4   /* LDS budget for one brick's staged partners in the force and
                                observer kernels, KiB (>= 8; default: MDGPU_OPT_BRICK_KB's
                                value); bricks with more staged partners read them from
                                global memory */
#define MDGPU_OPT_BRICK 5    /* cells per brick: bx | by << 8 | bz << 16, each 1..8 with
                                (bx+2)(by+2)(bz+2) <= 216 (the next step rebuilds the
                                list); by default a domain picks, at its first
                                construction, and an ensemble, per replica from its
                                density, the bricks that stage the fewest cells within
                                the LDS budget */
#define MDGPU_OPT_GRAPH 6    /* 1 = fused steps replayed from a captured HIP graph (not
                                while timing is enabled), 0 = launched (default: about
                                1.5% faster per step on the 4M workload) */
#define MDGPU_OPT_SEARCH_WAVES 7  /* wavefronts per brick in the neighbour search: 4, 8 or 16 */
#define MDGPU_OPT_SEARCH_STAGE 8  /* neighbour search candidates staged in LDS (1) or read
                                     from global memory (0) */
#define MDGPU_OPT_PRUNE 9    /* dynamic pruning: the fused step's kick reads an inner list of
                                the pairs within rc + value/1000, pruned from the full list
                                whenever the particles could have moved value/1000 since
                                (0 = off, the default; the same pairs in the same order) */
#define MDGPU_OPT_ROW_SORT 10  /* rows of a brick ordered by list length counted in steps of
                                  this value, 1..64 (default 8, the slices' rounding: no
                                  more padding than exact lengths, and rows of one cell stay
                                  together), or kept in cell order (0); the next step
                                  rebuilds the list */
#define MDGPU_OPT_BRICK_KB 11  /* LDS budget, KiB (8..160), that the automatic bricks are sized
                                  for (default 78 for a domain, 76 for an ensemble); setting it
                                  chooses the bricks again and the next step rebuilds the list */
#define MDGPU_OPT_SEARCH_CELLS 12  /* neighbour search with a wavefront per cell, lanes over 64
                                      candidates at a time (1, the default), or a thread per
                                      row (0); the same masks either way */
#define MDGPU_OPT_RUN_AHEAD 13  /* mdgpu_calculate queues up to this many fused steps without
                                   waiting for the host (default 64; 0 = every step waits): a
                                   device-side gate does each step's validity budget and stops
                                   the run at the step whose displacement check or construction
                                   needs the host, which then takes that step; the same
                                   trajectory either way */
#define MDGPU_OPT_PIPELINE 14  /* FP64 kick with 4, 8 or 16 wavefronts: 1 = persistent workgroups (one
                                  per CU) whose producer wavefronts stage the next brick's
                                  neighbourhood into a second LDS buffer with LDS-direct loads while
                                  the others compute the current brick; 0 = one workgroup per brick
                                  (default); the same sums in the same order */
#define MDGPU_OPT_MIXED_FRAME 15  /* mixed precision's FP32 coordinates: 1 = relative to the centre of
                                     the row's cell brick, staged from the FP64 positions (default: no
                                     offsets to read; coordinates up to about half a brick
                                     neighbourhood, so rounding about 2x coarser); 0 = relative to each
                                     particle's cell, with the partner's cell offset per list entry
                                     (written by constructions in this frame only; setting 0 writes
                                     them for the current list) */
#define MDGPU_OPT_DISPLACEMENT_CHECK 16  /* rebuild cadence: 1 = once the Verlet budget is spent, the
                                            exact displacement check each step (pairlist.cc:26-121;
                                            the default and the reference's schedule); 0 = the list
                                            expires as soon as the budget is spent (more constructions,
                                            no checks).  The pairs within rc are the same either way;
                                            only the construction steps, and with them the summation
                                            order of the kicks, differ */
int mdgpu_set_option(mdgpu_domain *d, int option, int value);
int mdgpu_get_option(const mdgpu_domain *d, int option, int *value);

/* sizes of the device-resident structures (for roofline accounting) */
int mdgpu_get_list_stats(const mdgpu_domain *d, int64_t *list_entries, int64_t *list_slots,
                         int *ncells, int *ghosts);

#ifdef __cplusplus
}
#endif
#endif
