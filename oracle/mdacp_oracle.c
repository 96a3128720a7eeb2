/*
 * mdacp_oracle.c -- CPU restatement of MDACP's single-domain LJ step.
 *
 * TEST INFRASTRUCTURE ONLY: the HIP path's results are tested against it (see mdacp_oracle.h).
 * Build with -O2 -ffp-contract=off and no -march flag so that no fused
 * multiply-add is emitted: the reference's default build has none either.
 *
 * Reference files restated here (kaityo256/mdacp):
 *   simulationinfo.cc:5-51, variables.cc:16-28,59-108,134-140,
 *   mdrect.cc:5-60, meshlist.cc:31-87,153-271,288-310, pairlist.cc:11-121,
 *   fcalculator.cc:22-32,229-290,744-778, observer.cc:5-79,
 *   mdunit.cc:78-205, mdmanager.cc:116-192,251-348,369-443, confmaker.cc:8-44.
 * Third-party arithmetic restated from its published definition:
 *   std::mt19937 (Matsumoto & Nishimura 1998), libstdc++-11
 *   generate_canonical<double,53>, uniform_real_distribution and the Marsaglia
 *   polar normal_distribution; Philox4x32-10 (Salmon et al., SC'11).
 */
#include "mdacp_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { DLEFT, DRIGHT, DBACK, DFORWARD, DDOWN, DUP, NDIR };

/* ------------------------------------------------------------------ RNG */
typedef struct {
  uint32_t mt[624];
  int idx;
} mt19937;

static void mt_seed(mt19937 *g, uint32_t s) {
  g->mt[0] = s;
  for (int i = 1; i < 624; i++)
    g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->idx = 624;
}

static uint32_t mt_next(mt19937 *g) {
  if (g->idx >= 624) {
    for (int i = 0; i < 624; i++) {
      uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
      g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    g->idx = 0;
  }
  uint32_t y = g->mt[g->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* generate_canonical<double,53> over a 32-bit engine: two draws, the second
   scaled by 2^32, summed in double, divided by 2^64; 1.0 maps to the
   largest double below 1. */
static double mt_canonical(mt19937 *g) {
  double sum = 0.0, tmp = 1.0;
  for (int k = 0; k < 2; k++) {
    sum += (double)mt_next(g) * tmp;
    tmp *= 4294967296.0;
  }
  double r = sum / tmp;
  if (r >= 1.0) r = nextafter(1.0, 0.0);
  return r;
}

typedef struct {
  int saved_available;
  double saved;
} normal_state;

/* Marsaglia polar method as in libstdc++ normal_distribution<double> */
static double mt_normal(mt19937 *g, normal_state *ns, double mean, double sd) {
  double ret;
  if (ns->saved_available) {
    ns->saved_available = 0;
    ret = ns->saved;
  } else {
    double x, y, r2;
    do {
      x = 2.0 * mt_canonical(g) - 1.0;
      y = 2.0 * mt_canonical(g) - 1.0;
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double mult = sqrt(-2.0 * log(r2) / r2);
    ns->saved = x * mult;
    ns->saved_available = 1;
    ret = y * mult;
  }
  return ret * sd + mean;
}

void ora_mt_raw(uint32_t seed, int n, uint32_t *out) {
  mt19937 g;
  mt_seed(&g, seed);
  for (int i = 0; i < n; i++) out[i] = mt_next(&g);
}

void ora_mt_uniform(uint32_t seed, int n, double *out) {
  mt19937 g;
  mt_seed(&g, seed);
  for (int i = 0; i < n; i++) out[i] = mt_canonical(&g) * (1.0 - 0.0) + 0.0;
}

void ora_mt_normal(uint32_t seed, double sd, int n, double *out) {
  mt19937 g;
  normal_state ns = {0, 0.0};
  mt_seed(&g, seed);
  for (int i = 0; i < n; i++) out[i] = mt_normal(&g, &ns, 0.0, sd);
}

/* Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
   easy as 1, 2, 3", SC'11: ten rounds of two 32x32->64 multiplies with a
   Weyl-sequence key schedule. */
void ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    const uint64_t m0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t m1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(m1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)m1;
    const uint32_t n2 = (uint32_t)(m0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)m0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Three standard normals for (seed, particle id, step): one Philox block,
   four 32-bit uniforms in (0,1), two Box-Muller transforms. */
void ora_philox_normal3(uint64_t seed, uint32_t id, uint64_t step, double out[3]) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const uint32_t ctr[4] = {id, (uint32_t)step, (uint32_t)(step >> 32), 0x4C414E47u};
  uint32_t r[4];
  ora_philox4x32(ctr, key, r);
  const double k = 2.3283064365386963e-10; /* 2^-32 */
  double u0 = ((double)r[0] + 0.5) * k, u1 = ((double)r[1] + 0.5) * k;
  double u2 = ((double)r[2] + 0.5) * k, u3 = ((double)r[3] + 0.5) * k;
  const double twopi = 6.283185307179586;
  double a = sqrt(-2.0 * log(u0)), b = sqrt(-2.0 * log(u2));
  out[0] = a * cos(twopi * u1);
  out[1] = a * sin(twopi * u1);
  out[2] = b * cos(twopi * u3);
}

/* --------------------------------------------------------------- system */
struct ora_sys {
  /* SimulationInfo (simulationinfo.cc:5-51) */
  double L[3], dt, rc, margin, SL, CL2, S2;
  double T_aim, gamma, tau;
  int thermostat, check_list_length;
  int noise_kind;
  uint64_t philox_seed;
  uint64_t step_count;
  /* MDRect of the single domain: s = 0, e = L (mdunit.cc:8-27) */
  double rs[3], re[3];
  /* Variables (variables.h:9-39) */
  int pn, tn, cap;
  double *q, *p; /* [cap][4] */
  int *type;
  double C0, C2, zeta, sim_time;
  /* per-ghost bookkeeping for canonical export (not used by the step) */
  int *g_src, *g_shift;
  /* MDUnit border lists (mdunit.h:33) */
  int *border[NDIR];
  int nborder[NDIR], border_cap[NDIR];
  /* MeshList (meshlist.h:10-62) */
  int mx, my, mz, nmesh;
  double msize[3];
  int *mesh_count, *mesh_index, *mesh_index2, *sortbuf, *ppos;
  long npairs, pair_cap;
  int *key, *partner, *npartners, *key_pointer, *key_pointer2, *sorted_list;
  int nconstructions;
  /* PairList (pairlist.h:10-31) */
  int initialized, lifetime, is_fresh;
  double buffer_length;
  double *qb_old, *disp;
  /* MDUnit::mt (mdunit.h:35) */
  mt19937 mt;
  int rebuilt;
};

static void *xrealloc(void *p, size_t n) {
  void *r = realloc(p, n ? n : 1);
  if (!r) abort();
  return r;
}

static void ensure_cap(ora_sys *s, int n) {
  if (n <= s->cap) return;
  int c = s->cap ? s->cap : 1024;
  while (c < n) c *= 2;
  s->q = (double *)xrealloc(s->q, sizeof(double) * 4 * (size_t)c);
  s->p = (double *)xrealloc(s->p, sizeof(double) * 4 * (size_t)c);
  memset(s->q + 4 * (size_t)s->cap, 0, sizeof(double) * 4 * (size_t)(c - s->cap));
  memset(s->p + 4 * (size_t)s->cap, 0, sizeof(double) * 4 * (size_t)(c - s->cap));
  s->type = (int *)xrealloc(s->type, sizeof(int) * (size_t)c);
  s->g_src = (int *)xrealloc(s->g_src, sizeof(int) * (size_t)c);
  s->g_shift = (int *)xrealloc(s->g_shift, sizeof(int) * (size_t)c);
  s->sortbuf = (int *)xrealloc(s->sortbuf, sizeof(int) * (size_t)c);
  s->ppos = (int *)xrealloc(s->ppos, sizeof(int) * (size_t)c);
  s->npartners = (int *)xrealloc(s->npartners, sizeof(int) * (size_t)c);
  s->key_pointer = (int *)xrealloc(s->key_pointer, sizeof(int) * (size_t)c);
  s->key_pointer2 = (int *)xrealloc(s->key_pointer2, sizeof(int) * (size_t)c);
  s->qb_old = (double *)xrealloc(s->qb_old, sizeof(double) * 4 * (size_t)c);
  s->disp = (double *)xrealloc(s->disp, sizeof(double) * (size_t)c);
  s->cap = c;
}

/* MeshList::ChangeScale (meshlist.cc:31-55) */
static void mesh_change_scale(ora_sys *s) {
  double w[3];
  int ms[3], m[3];
  for (int d = 0; d < 3; d++) {
    w[d] = s->re[d] - s->rs[d];
    ms[d] = (int)(w[d] / s->SL);
    s->msize[d] = w[d] / (double)ms[d];
    m[d] = (int)(w[d] / s->msize[d]) + 2;
  }
  s->mx = m[0]; s->my = m[1]; s->mz = m[2];
  s->nmesh = s->mx * s->my * s->mz;
  s->mesh_count = (int *)xrealloc(s->mesh_count, sizeof(int) * (size_t)s->nmesh);
  s->mesh_index = (int *)xrealloc(s->mesh_index, sizeof(int) * (size_t)s->nmesh);
  s->mesh_index2 = (int *)xrealloc(s->mesh_index2, sizeof(int) * (size_t)s->nmesh);
}

ora_sys *ora_create(double Lx, double Ly, double Lz, double dt, double rc, double margin) {
  ora_sys *s = (ora_sys *)calloc(1, sizeof(ora_sys));
  s->L[0] = Lx; s->L[1] = Ly; s->L[2] = Lz;
  s->dt = dt;
  s->rc = rc;
  s->margin = margin;
  s->SL = rc + margin;          /* simulationinfo.cc:6-7 */
  s->CL2 = rc * rc;             /* fcalculator.cc:231 */
  s->S2 = s->SL * s->SL;        /* meshlist.cc:247 */
  s->T_aim = 1.0; s->gamma = 0.1; s->tau = 0.1;
  s->check_list_length = 500;   /* simulationinfo.cc:29 */
  s->thermostat = ORA_NVE;
  /* Variables::Variables (variables.cc:16-28) */
  const double s2 = 1.0 / (rc * rc);
  const double s6 = s2 * s2 * s2;
  const double s8 = s6 * s2;
  const double s12 = s6 * s6;
  const double s14 = s12 * s2;
  s->C2 = 6.0 * s14 - 3.0 * s8;
  s->C0 = -s12 + s6 - s->C2 / s2;
  /* single domain: grid 1x1x1, s = L*0, e = s + L (mdunit.cc:17-24) */
  for (int d = 0; d < 3; d++) {
    s->rs[d] = s->L[d] * 0.0;
    s->re[d] = s->rs[d] + s->L[d];
  }
  mesh_change_scale(s);
  mt_seed(&s->mt, 0u); /* mdunit.cc:26, id = 0 */
  ensure_cap(s, 1024);
  return s;
}

void ora_destroy(ora_sys *s) {
  if (!s) return;
  free(s->q); free(s->p); free(s->type); free(s->g_src); free(s->g_shift);
  for (int d = 0; d < NDIR; d++) free(s->border[d]);
  free(s->mesh_count); free(s->mesh_index); free(s->mesh_index2);
  free(s->sortbuf); free(s->ppos);
  free(s->key); free(s->partner); free(s->npartners);
  free(s->key_pointer); free(s->key_pointer2); free(s->sorted_list);
  free(s->qb_old); free(s->disp);
  free(s);
}

void ora_set_thermostat(ora_sys *s, int kind, double T, double gamma, double tau) {
  s->thermostat = kind;
  s->T_aim = T;
  s->gamma = gamma;
  s->tau = tau;
}

void ora_set_noise(ora_sys *s, int kind, uint64_t seed) {
  s->noise_kind = kind;
  s->philox_seed = seed;
}

void ora_set_check_list_length(ora_sys *s, int n) { s->check_list_length = n; }

/* MDUnit::AddParticle -> MDRect::IsInside -> Variables::AddParticle
   (mdunit.cc:36-49, mdrect.h:35-38, variables.cc:31-42) */
int ora_add_particles(ora_sys *s, int n, const double *x, const double *v, const int *type) {
  int added = 0;
  for (int i = 0; i < n; i++) {
    const double *xi = x + 3 * (size_t)i;
    if (!(s->rs[0] <= xi[0] && xi[0] < s->re[0] && s->rs[1] <= xi[1] && xi[1] < s->re[1] &&
          s->rs[2] <= xi[2] && xi[2] < s->re[2]))
      continue;
    ensure_cap(s, s->pn + 1);
    double *q = s->q + 4 * (size_t)s->pn, *p = s->p + 4 * (size_t)s->pn;
    q[0] = xi[0]; q[1] = xi[1]; q[2] = xi[2];
    p[0] = v ? v[3 * (size_t)i] : 0.0;
    p[1] = v ? v[3 * (size_t)i + 1] : 0.0;
    p[2] = v ? v[3 * (size_t)i + 2] : 0.0;
    s->type[s->pn] = type ? type[i] : 0;
    s->pn++;
    s->tn = s->pn;
    added++;
  }
  return added;
}

/* Variables::SetInitialVelocity (variables.cc:82-108) */
void ora_set_initial_velocity(ora_sys *s, double V0, int seed) {
  mt19937 g;
  mt_seed(&g, (uint32_t)seed);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < s->pn; i++) {
    double z = mt_canonical(&g) * 2.0 - 1.0;
    double phi = mt_canonical(&g) * 2.0 * pi;
    double *p = s->p + 4 * (size_t)i;
    p[0] = V0 * sqrt(1 - z * z) * cos(phi);
    p[1] = V0 * sqrt(1 - z * z) * sin(phi);
    p[2] = V0 * z;
  }
  double av[3] = {0, 0, 0};
  for (int i = 0; i < s->pn; i++)
    for (int d = 0; d < 3; d++) av[d] += s->p[4 * (size_t)i + d];
  for (int d = 0; d < 3; d++) av[d] /= (double)s->pn;
  for (int i = 0; i < s->pn; i++)
    for (int d = 0; d < 3; d++) s->p[4 * (size_t)i + d] -= av[d];
}

/* MDUnit::SetSeed (mdunit.h:58), id = 0 */
void ora_set_seed(ora_sys *s, int seed) { mt_seed(&s->mt, (uint32_t)(seed + 0)); }

/* ----------------------------------------------------- migration + wrap */
/* MDRect::IsOverBoundary (mdrect.cc:5-21) */
static int over_boundary(const ora_sys *s, int dir, const double *q) {
  switch (dir) {
    case DLEFT: return q[0] < s->rs[0];
    case DRIGHT: return q[0] >= s->re[0];
    case DBACK: return q[1] < s->rs[1];
    case DFORWARD: return q[1] >= s->re[1];
    case DDOWN: return q[2] < s->rs[2];
    case DUP: return q[2] >= s->re[2];
  }
  return 0;
}

/* MDUnit::MakeBufferForSendingParticle + ReceiveParticles on the same unit
   (mdunit.cc:78-122, mdmanager.cc:195-248): stable compaction of the staying
   particles, crossing ones appended in order. */
static void send_particles_sub(ora_sys *s, int dir) {
  int nsend = 0;
  for (int i = 0; i < s->pn; i++)
    if (over_boundary(s, dir, s->q + 4 * (size_t)i)) nsend++;
  if (!nsend) return;
  double *bq = (double *)malloc(sizeof(double) * 3 * (size_t)nsend);
  double *bp = (double *)malloc(sizeof(double) * 3 * (size_t)nsend);
  int *bt = (int *)malloc(sizeof(int) * (size_t)nsend);
  int k = 0, index = 0;
  for (int i = 0; i < s->pn; i++) {
    double *q = s->q + 4 * (size_t)i, *p = s->p + 4 * (size_t)i;
    if (over_boundary(s, dir, q)) {
      bq[3 * k] = q[0]; bq[3 * k + 1] = q[1]; bq[3 * k + 2] = q[2];
      bp[3 * k] = p[0]; bp[3 * k + 1] = p[1]; bp[3 * k + 2] = p[2];
      bt[k] = s->type[i];
      k++;
    } else {
      double *qd = s->q + 4 * (size_t)index, *pd = s->p + 4 * (size_t)index;
      qd[0] = q[0]; qd[1] = q[1]; qd[2] = q[2];
      pd[0] = p[0]; pd[1] = p[1]; pd[2] = p[2];
      s->type[index] = s->type[i];
      index++;
    }
  }
  for (int j = 0; j < nsend; j++) {
    double *qd = s->q + 4 * (size_t)index, *pd = s->p + 4 * (size_t)index;
    qd[0] = bq[3 * j]; qd[1] = bq[3 * j + 1]; qd[2] = bq[3 * j + 2];
    pd[0] = bp[3 * j]; pd[1] = bp[3 * j + 1]; pd[2] = bp[3 * j + 2];
    s->type[index] = bt[j];
    index++;
  }
  free(bq); free(bp); free(bt);
}

/* Variables::AdjustPeriodicBoundary (variables.cc:59-79) */
static int adjust_periodic(ora_sys *s) {
  for (int i = 0; i < s->pn; i++) {
    double *q = s->q + 4 * (size_t)i;
    for (int d = 0; d < 3; d++) {
      if (q[d] < 0.0) q[d] += s->L[d];
      else if (q[d] > s->L[d]) q[d] -= s->L[d];
    }
  }
  for (int i = 0; i < s->pn; i++) {
    double *q = s->q + 4 * (size_t)i;
    for (int d = 0; d < 3; d++)
      if (q[d] < 0.0 || q[d] > s->L[d]) return ORA_ERR_INVALID_POSITION;
  }
  return ORA_OK;
}

/* ---------------------------------------------------------------- ghosts */
/* MDRect::IsInsideEdge (mdrect.cc:24-47) */
static int inside_edge(const ora_sys *s, int dir, const double *q) {
  const double SL = s->SL;
  switch (dir) {
    case DLEFT: return q[0] < s->rs[0] + SL && q[0] > s->rs[0];
    case DRIGHT: return q[0] > s->re[0] - SL && q[0] < s->re[0];
    case DBACK: return q[1] < s->rs[1] + SL && q[1] > s->rs[1];
    case DFORWARD: return q[1] > s->re[1] - SL && q[1] < s->re[1];
    case DDOWN: return q[2] < s->rs[2] + SL && q[2] > s->rs[2];
    case DUP: return q[2] > s->re[2] - SL && q[2] < s->re[2];
  }
  return 0;
}

/* image displacement of MDUnit::MakeBufferForBorderParticles (mdunit.cc:142-161):
   a particle leaving through the left face reappears at +L, and so on. */
static void dir_diff(const ora_sys *s, int dir, double diff[3]) {
  diff[0] = diff[1] = diff[2] = 0.0;
  switch (dir) {
    case DLEFT: diff[0] = s->L[0]; break;
    case DRIGHT: diff[0] -= s->L[0]; break;
    case DBACK: diff[1] = s->L[1]; break;
    case DFORWARD: diff[1] = -s->L[1]; break;
    case DDOWN: diff[2] = s->L[2]; break;
    case DUP: diff[2] = -s->L[2]; break;
  }
}

static const int dir_axis[NDIR] = {0, 0, 1, 1, 2, 2};
static const int dir_sign[NDIR] = {+1, -1, +1, -1, +1, -1};

/* MakeBufferForBorderParticles + ReceiveBorderParticles on the same unit */
static void send_border_sub(ora_sys *s, int dir) {
  double diff[3];
  dir_diff(s, dir, diff);
  const int nb = s->nborder[dir];
  ensure_cap(s, s->tn + nb);
  double *buf = (double *)malloc(sizeof(double) * 3 * (size_t)(nb ? nb : 1));
  for (int k = 0; k < nb; k++) {
    const double *q = s->q + 4 * (size_t)s->border[dir][k];
    buf[3 * k] = q[0] + diff[0];
    buf[3 * k + 1] = q[1] + diff[1];
    buf[3 * k + 2] = q[2] + diff[2];
  }
  for (int k = 0; k < nb; k++) {
    double *q = s->q + 4 * (size_t)(s->tn + k);
    q[0] = buf[3 * k]; q[1] = buf[3 * k + 1]; q[2] = buf[3 * k + 2];
  }
  free(buf);
  s->tn += nb;
}

/* MDUnit::FindBorderParticles (mdunit.cc:125-134) */
static void find_border(ora_sys *s, int dir) {
  s->nborder[dir] = 0;
  for (int i = 0; i < s->tn; i++) {
    if (inside_edge(s, dir, s->q + 4 * (size_t)i)) {
      if (s->nborder[dir] == s->border_cap[dir]) {
        s->border_cap[dir] = s->border_cap[dir] ? 2 * s->border_cap[dir] : 1024;
        s->border[dir] = (int *)xrealloc(s->border[dir], sizeof(int) * (size_t)s->border_cap[dir]);
      }
      s->border[dir][s->nborder[dir]++] = i;
    }
  }
}

/* MDManager::SendBorderParticles (mdmanager.cc:283-292) */
void ora_send_border_particles(ora_sys *s) {
  s->tn = s->pn;
  for (int dir = 0; dir < NDIR; dir++) send_border_sub(s, dir);
}

/* ------------------------------------------------------------ mesh list */
static void register_pair(ora_sys *s, int a, int b) {
  int i1 = a < b ? a : b, i2 = a < b ? b : a;
  if (s->npairs == s->pair_cap) {
    s->pair_cap = s->pair_cap ? 2 * s->pair_cap : (1 << 16);
    s->key = (int *)xrealloc(s->key, sizeof(int) * (size_t)s->pair_cap);
    s->partner = (int *)xrealloc(s->partner, sizeof(int) * (size_t)s->pair_cap);
    s->sorted_list = (int *)xrealloc(s->sorted_list, sizeof(int) * (size_t)s->pair_cap);
  }
  s->key[s->npairs] = i1;
  s->partner[s->npairs] = i2;
  s->npartners[i1]++;
  s->npairs++;
}

/* MeshList::MakeMesh (meshlist.cc:153-205): stable counting sort by cell */
static int make_mesh(ora_sys *s) {
  const int pn = s->tn;
  const double imx = 1.0 / s->msize[0], imy = 1.0 / s->msize[1], imz = 1.0 / s->msize[2];
  memset(s->mesh_count, 0, sizeof(int) * (size_t)s->nmesh);
  for (int i = 0; i < pn; i++) {
    const double *q = s->q + 4 * (size_t)i;
    int ix = (int)((q[0] - s->rs[0]) * imx) + 1;
    int iy = (int)((q[1] - s->rs[1]) * imy) + 1;
    int iz = (int)((q[2] - s->rs[2]) * imz) + 1;
    if (ix < 0) ix = s->mx - 1; else if (ix >= s->mx) ix = 0;
    if (iy < 0) iy = s->my - 1; else if (iy >= s->my) iy = 0;
    if (iz < 0) iz = s->mz - 1; else if (iz >= s->mz) iz = 0;
    int index = s->mx * s->my * iz + s->mx * iy + ix;
    if (index >= s->nmesh || index < 0) return ORA_ERR_INVALID_INDEX;
    s->ppos[i] = index;
    s->mesh_count[index]++;
  }
  s->mesh_index[0] = 0;
  int sum = 0;
  for (int i = 0; i < s->nmesh - 1; i++) {
    sum += s->mesh_count[i];
    s->mesh_index[i + 1] = sum;
  }
  memset(s->mesh_index2, 0, sizeof(int) * (size_t)s->nmesh);
  for (int i = 0; i < pn; i++) {
    int index = s->ppos[i];
    s->sortbuf[s->mesh_index[index] + s->mesh_index2[index]] = i;
    s->mesh_index2[index]++;
  }
  return ORA_OK;
}

/* MeshList::SearchMesh (meshlist.cc:220-271): self cell plus the 13-cell
   half shell, pairs (home particle, later particle), no ghost-ghost pairs */
static void search_mesh(ora_sys *s, int index, int *v, int vcap) {
  (void)vcap;
  int ix = index % s->mx, rest = index / s->mx;
  int iy = rest % s->my, iz = rest / s->my;
  static const int off[14][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 1, 0},  {0, 1, 0},  {1, 1, 0},
                                 {-1, 0, 1},  {0, 0, 1},   {1, 0, 1},   {-1, -1, 1}, {0, -1, 1},
                                 {1, -1, 1},  {-1, 1, 1},  {0, 1, 1},   {1, 1, 1}};
  int ln = 0;
  for (int k = 0; k < 14; k++) {
    int jx = ix + off[k][0], jy = iy + off[k][1], jz = iz + off[k][2];
    if (jx < 0 || jx >= s->mx || jy < 0 || jy >= s->my || jz < 0 || jz >= s->mz) continue;
    int c = s->mx * s->my * jz + s->mx * jy + jx;
    int in = s->mesh_count[c], mi = s->mesh_index[c];
    for (int t = 0; t < in; t++) v[ln++] = s->sortbuf[mi + t];
  }
  const int pn = s->pn;
  const int in = s->mesh_count[index];
  for (int i = 0; i < in; i++) {
    const int i1 = v[i];
    const double x1 = s->q[4 * (size_t)i1], y1 = s->q[4 * (size_t)i1 + 1], z1 = s->q[4 * (size_t)i1 + 2];
    for (int j = i + 1; j < ln; j++) {
      const int i2 = v[j];
      if (i1 >= pn && i2 >= pn) continue;
      const double dx = x1 - s->q[4 * (size_t)i2];
      const double dy = y1 - s->q[4 * (size_t)i2 + 1];
      const double dz = z1 - s->q[4 * (size_t)i2 + 2];
      const double r2 = (dx * dx + dy * dy + dz * dz);
      if (r2 > s->S2) continue;
      register_pair(s, i1, i2);
    }
  }
}

/* MeshList::MakeList (meshlist.cc:117-150): COO via the mesh, then CSR */
static int make_list(ora_sys *s) {
  s->npairs = 0;
  const int pn = s->tn;
  for (int i = 0; i < pn; i++) s->npartners[i] = 0;
  int err = make_mesh(s);
  if (err) return err;
  int maxc = 0;
  for (int c = 0; c < s->nmesh; c++)
    if (s->mesh_count[c] > maxc) maxc = s->mesh_count[c];
  int *v = (int *)malloc(sizeof(int) * (size_t)(14 * maxc + 1));
  for (int c = 0; c < s->nmesh; c++) search_mesh(s, c, v, 14 * maxc + 1);
  free(v);
  int pos = 0;
  s->key_pointer[0] = 0;
  for (int i = 0; i < pn - 1; i++) {
    pos += s->npartners[i];
    s->key_pointer[i + 1] = pos;
  }
  for (int i = 0; i < pn; i++) s->key_pointer2[i] = 0;
  for (long k = 0; k < s->npairs; k++) {
    int i = s->key[k], j = s->partner[k];
    s->sorted_list[s->key_pointer[i] + s->key_pointer2[i]] = j;
    s->key_pointer2[i]++;
  }
  s->nconstructions++;
  return ORA_OK;
}

/* PairList::Init (pairlist.cc:11-23) */
static void pairlist_init(ora_sys *s) {
  s->initialized = 1;
  s->is_fresh = 1;
  s->buffer_length = s->margin;
  s->lifetime = 0;
  for (int i = 0; i < s->tn; i++)
    for (int d = 0; d < 3; d++) s->qb_old[4 * (size_t)i + d] = s->q[4 * (size_t)i + d];
}

/* MDManager::MakePairList (mdmanager.cc:262-280) for one unit */
int ora_make_pair_list(ora_sys *s) {
  for (int dir = 0; dir < NDIR; dir++) send_particles_sub(s, dir);
  int err = adjust_periodic(s);
  if (err) return err;
  s->tn = s->pn;
  for (int i = 0; i < s->pn; i++) {
    s->g_src[i] = i;
    s->g_shift[i] = 13;
  }
  for (int dir = 0; dir < NDIR; dir++) {
    find_border(s, dir);
    /* bookkeeping of (owned source, composed shift) for every new ghost */
    ensure_cap(s, s->tn + s->nborder[dir]);
    for (int k = 0; k < s->nborder[dir]; k++) {
      int j = s->border[dir][k];
      int g = s->tn + k;
      s->g_src[g] = s->g_src[j];
      int code = s->g_shift[j];
      int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
      int sh[3] = {sx, sy, sz};
      sh[dir_axis[dir]] += dir_sign[dir];
      s->g_shift[g] = (sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1);
    }
    send_border_sub(s, dir);
  }
  /* MDUnit::MakePairList (mdunit.cc:194-198); MeshList::Sort is off */
  pairlist_init(s);
  err = make_list(s);
  s->rebuilt = 1;
  return err;
}

/* ---------------------------------------------------------------- force */
/* ForceCalculator::UpdatePositionHalf (fcalculator.cc:22-32) */
void ora_update_position_half(ora_sys *s) {
  const double dt2 = s->dt * 0.5;
  for (int i = 0; i < s->pn; i++) {
    double *q = s->q + 4 * (size_t)i, *p = s->p + 4 * (size_t)i;
    q[0] += p[0] * dt2;
    q[1] += p[1] * dt2;
    q[2] += p[2] * dt2;
  }
}

/* ForceCalculator::CalculateForceNext (fcalculator.cc:229-290): half-list
   Newton-3 kick.  The reference pipelines the next partner one iteration
   ahead; the arithmetic below is that loop's, pair for pair: each in-cutoff
   partner's p_j is decremented in row order, and p_i receives the running
   sum of all but the last term plus the last term. */
void ora_calculate_force(ora_sys *s) {
  const double CL2 = s->CL2;
  const double C2 = s->C2 * 8.0;
  const double dt = s->dt;
  double *q = s->q, *p = s->p;
  for (int i = 0; i < s->pn; i++) {
    const double qx = q[4 * (size_t)i], qy = q[4 * (size_t)i + 1], qz = q[4 * (size_t)i + 2];
    double pfx = 0, pfy = 0, pfz = 0;
    double df = 0.0, dxb = 0.0, dyb = 0.0, dzb = 0.0;
    int jb = 0;
    const int kp = s->key_pointer[i], np = s->npartners[i];
    for (int k = kp; k < kp + np; k++) {
      const int j = s->sorted_list[k];
      const double dx = q[4 * (size_t)j] - qx;
      const double dy = q[4 * (size_t)j + 1] - qy;
      const double dz = q[4 * (size_t)j + 2] - qz;
      const double r2 = (dx * dx + dy * dy + dz * dz);
      if (r2 > CL2) continue;
      pfx += df * dxb;
      pfy += df * dyb;
      pfz += df * dzb;
      p[4 * (size_t)jb] -= df * dxb;
      p[4 * (size_t)jb + 1] -= df * dyb;
      p[4 * (size_t)jb + 2] -= df * dzb;
      const double r6 = r2 * r2 * r2;
      df = ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + C2) * dt;
      jb = j;
      dxb = dx; dyb = dy; dzb = dz;
    }
    p[4 * (size_t)jb] -= df * dxb;
    p[4 * (size_t)jb + 1] -= df * dyb;
    p[4 * (size_t)jb + 2] -= df * dzb;
    p[4 * (size_t)i] += pfx + df * dxb;
    p[4 * (size_t)i + 1] += pfy + df * dyb;
    p[4 * (size_t)i + 2] += pfz + df * dzb;
  }
}

void ora_kick_only(ora_sys *s, double *dp_out) {
  double *save = (double *)malloc(sizeof(double) * 4 * (size_t)s->tn);
  memcpy(save, s->p, sizeof(double) * 4 * (size_t)s->tn);
  memset(s->p, 0, sizeof(double) * 4 * (size_t)s->tn);
  ora_calculate_force(s);
  for (int i = 0; i < s->pn; i++)
    for (int d = 0; d < 3; d++) dp_out[3 * (size_t)i + d] = s->p[4 * (size_t)i + d];
  memcpy(s->p, save, sizeof(double) * 4 * (size_t)s->tn);
  free(save);
}

/* ForceCalculator::Langevin (fcalculator.cc:765-778) */
static void langevin(ora_sys *s) {
  const double dt = s->dt;
  const double g = s->gamma;
  const double hb_D = sqrt(2.0 * g * s->T_aim / dt);
  if (s->noise_kind == ORA_NOISE_MT19937) {
    normal_state ns = {0, 0.0};
    for (int i = 0; i < s->pn; i++) {
      double *p = s->p + 4 * (size_t)i;
      p[0] += (-g * p[0] + mt_normal(&s->mt, &ns, 0.0, hb_D)) * dt;
      p[1] += (-g * p[1] + mt_normal(&s->mt, &ns, 0.0, hb_D)) * dt;
      p[2] += (-g * p[2] + mt_normal(&s->mt, &ns, 0.0, hb_D)) * dt;
    }
  } else {
    for (int i = 0; i < s->pn; i++) {
      double n[3];
      ora_philox_normal3(s->philox_seed, (uint32_t)s->type[i], s->step_count, n);
      double *p = s->p + 4 * (size_t)i;
      p[0] += (-g * p[0] + (n[0] * hb_D + 0.0)) * dt;
      p[1] += (-g * p[1] + (n[1] * hb_D + 0.0)) * dt;
      p[2] += (-g * p[2] + (n[2] * hb_D + 0.0)) * dt;
    }
  }
}

/* ForceCalculator::HeatbathZeta / HeatbathMomenta (fcalculator.cc:744-762) */
static void heatbath_zeta(ora_sys *s, double t) {
  const double dt2 = s->dt * 0.5;
  double t1 = (t - s->T_aim) / (s->tau * s->tau);
  s->zeta += t1 * dt2;
}

static void heatbath_momenta(ora_sys *s) {
  const double dt2 = s->dt * 0.5;
  const double e1 = exp(-dt2 * s->zeta);
  for (int i = 0; i < s->pn; i++)
    for (int d = 0; d < 4; d++) s->p[4 * (size_t)i + d] *= e1;
}

/* ------------------------------------------------------------ observers */
/* KineticEnergyObserver (observer.cc:5-15): four lanes incl. padding */
double ora_kinetic_energy_raw(ora_sys *s) {
  double e = 0;
  for (int i = 0; i < s->pn; i++)
    for (int d = 0; d < 4; d++) e += 0.5 * s->p[4 * (size_t)i + d] * s->p[4 * (size_t)i + d];
  return e;
}

/* PotentialEnergyObserver (observer.cc:18-48) over the COO order */
double ora_potential_energy_raw(ora_sys *s) {
  const double CL2 = s->CL2, C2 = s->C2, C0 = s->C0;
  double energy = 0.0;
  for (long k = 0; k < s->npairs; k++) {
    int i = s->key[k], j = s->partner[k];
    double dx = s->q[4 * (size_t)i] - s->q[4 * (size_t)j];
    double dy = s->q[4 * (size_t)i + 1] - s->q[4 * (size_t)j + 1];
    double dz = s->q[4 * (size_t)i + 2] - s->q[4 * (size_t)j + 2];
    const double r2 = (dx * dx + dy * dy + dz * dz);
    if (r2 > CL2) continue;
    double e = 4.0 * (1.0 / (r2 * r2 * r2 * r2 * r2 * r2) - 1.0 / (r2 * r2 * r2) + C2 * r2 + C0);
    if (i >= s->pn || j >= s->pn) e *= 0.5;
    energy += e;
  }
  return energy;
}

/* VirialObserver (observer.cc:51-79) */
double ora_virial_raw(ora_sys *s) {
  double phi = 0.0;
  const double C_2 = s->C2, CL2 = s->CL2;
  for (long k = 0; k < s->npairs; k++) {
    int i = s->key[k], j = s->partner[k];
    double dx = s->q[4 * (size_t)j] - s->q[4 * (size_t)i];
    double dy = s->q[4 * (size_t)j + 1] - s->q[4 * (size_t)i + 1];
    double dz = s->q[4 * (size_t)j + 2] - s->q[4 * (size_t)i + 2];
    const double r2 = (dx * dx + dy * dy + dz * dz);
    if (r2 > CL2) continue;
    const double r6 = r2 * r2 * r2;
    double df = ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + C_2 * 8.0) * r2;
    if (i >= s->pn || j >= s->pn) df *= 0.5;
    phi += df;
  }
  return phi / 3.0;
}

/* MDManager observables (mdmanager.cc:390-429); ObserveDouble sums one unit
   onto 0.0 (omp reduction) and through a one-rank MPI_Allreduce */
static double observe_sum(double unit_value) { return 0.0 + unit_value; }

double ora_temperature(ora_sys *s) {
  const double pn = (double)s->pn;
  return (observe_sum(ora_kinetic_energy_raw(s)) / pn) / 1.5;
}

double ora_pressure(ora_sys *s) {
  const double pn = (double)s->pn;
  const double phi = observe_sum(ora_virial_raw(s)) / pn;
  const double T = ora_temperature(s);
  const double V = s->L[0] * s->L[1] * s->L[2];
  return (T - phi) * pn / V;
}

double ora_total_energy(ora_sys *s) {
  const double pn = (double)s->pn;
  double ke = observe_sum(ora_kinetic_energy_raw(s)) / pn;
  double pe = observe_sum(ora_potential_energy_raw(s)) / pn;
  return ke + pe;
}

/* -------------------------------------------------------------- expiry */
/* PairList::UpdatePairListValidity (pairlist.cc:48-65) */
static int update_validity(ora_sys *s) {
  const double *p = s->p;
  double vmax = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
  for (int i = 1; i < s->pn; i++) {
    double v = p[4 * (size_t)i] * p[4 * (size_t)i] + p[4 * (size_t)i + 1] * p[4 * (size_t)i + 1] +
               p[4 * (size_t)i + 2] * p[4 * (size_t)i + 2];
    if (v > vmax) vmax = v;
  }
  vmax = sqrt(vmax);
  if (vmax * 2.0 * s->dt > s->margin) return ORA_ERR_TOO_FAST;
  s->buffer_length = s->buffer_length - vmax * 2.0 * s->dt;
  return ORA_OK;
}

/* PairList::CheckByDisplacement (pairlist.cc:68-121) */
static int check_by_displacement(ora_sys *s) {
  const int pn = s->tn;
  double dr2_max = 0.0;
  for (int i = 0; i < pn; i++) {
    double dx = s->qb_old[4 * (size_t)i] - s->q[4 * (size_t)i];
    double dy = s->qb_old[4 * (size_t)i + 1] - s->q[4 * (size_t)i + 1];
    double dz = s->qb_old[4 * (size_t)i + 2] - s->q[4 * (size_t)i + 2];
    double dr2 = dx * dx + dy * dy + dz * dz;
    s->disp[i] = dr2;
    if (dr2_max < dr2) dr2_max = dr2;
  }
  const int Nl = (pn > s->check_list_length) ? s->check_list_length : pn;
  int *v = (int *)malloc(sizeof(int) * (size_t)(Nl + 1));
  int l = 0;
  const double c_distance = s->margin - sqrt(dr2_max);
  const double c_d2 = c_distance * c_distance;
  for (int i = 0; i < pn; i++) {
    if (s->disp[i] > c_d2) {
      v[l++] = i;
      if (l >= Nl) {
        free(v);
        return 1;
      }
    }
  }
  const double SL2 = s->SL * s->SL;
  const double CL2 = s->CL2;
  for (int i = 0; i < l - 1; i++) {
    for (int j = i + 1; j < l; j++) {
      int i1 = v[i], i2 = v[j];
      double dx = s->qb_old[4 * (size_t)i1] - s->qb_old[4 * (size_t)i2];
      double dy = s->qb_old[4 * (size_t)i1 + 1] - s->qb_old[4 * (size_t)i2 + 1];
      double dz = s->qb_old[4 * (size_t)i1 + 2] - s->qb_old[4 * (size_t)i2 + 2];
      double dr2 = dx * dx + dy * dy + dz * dz;
      if (dr2 < SL2) continue;
      dx = s->q[4 * (size_t)i1] - s->q[4 * (size_t)i2];
      dy = s->q[4 * (size_t)i1 + 1] - s->q[4 * (size_t)i2 + 1];
      dz = s->q[4 * (size_t)i1 + 2] - s->q[4 * (size_t)i2 + 2];
      dr2 = dx * dx + dy * dy + dz * dz;
      if (dr2 < CL2) {
        free(v);
        return 1;
      }
    }
  }
  free(v);
  return 0;
}

/* PairList::IsPairListExpired (pairlist.cc:26-45) */
static int is_expired(ora_sys *s, int *expired) {
  if (!s->initialized) {
    *expired = 1;
    return ORA_OK;
  }
  s->is_fresh = 0;
  s->lifetime++;
  if (s->buffer_length > 0.0) {
    int err = update_validity(s);
    if (err) return err;
  }
  *expired = 0;
  if (s->buffer_length < 0.0) *expired = check_by_displacement(s);
  return ORA_OK;
}

/* ----------------------------------------------------------------- step */
/* MDManager::Calculate (mdmanager.cc:116-148) with CalculateForce (:150-157),
   CalculateNoseHoover (:159-175) and CalculateLangevin (:185-192) */
int ora_calculate(ora_sys *s) {
  int expired = 0, err;
  s->rebuilt = 0;
  err = is_expired(s, &expired);
  if (err) return err;
  if (expired) {
    err = ora_make_pair_list(s);
    if (err) return err;
  }
  ora_update_position_half(s);
  ora_send_border_particles(s);
  if (s->thermostat == ORA_NOSEHOOVER) {
    double t = ora_temperature(s);
    heatbath_zeta(s, t);
    heatbath_momenta(s);
    ora_calculate_force(s);
    heatbath_momenta(s);
    t = ora_temperature(s);
    heatbath_zeta(s, t);
    ora_update_position_half(s);
  } else if (s->thermostat == ORA_LANGEVIN) {
    ora_calculate_force(s);
    langevin(s);
  } else {
    ora_calculate_force(s);
    ora_update_position_half(s);
  }
  s->sim_time += s->dt;
  s->step_count++;
  return ORA_OK;
}

int ora_last_step_rebuilt(const ora_sys *s) { return s->rebuilt; }

/* MDManager::ChangeScale (mdmanager.cc:432-443) with MDUnit::ChangeScale
   (mdunit.cc:201-205), MDRect::ChangeScale (mdrect.cc:55-60) and
   Variables::ChangeScale (variables.cc:134-140) */
void ora_change_scale(ora_sys *s, double alpha) {
  ora_make_pair_list(s);
  for (int d = 0; d < 3; d++) s->L[d] = s->L[d] * alpha;
  for (int d = 0; d < 3; d++) {
    s->rs[d] *= alpha;
    s->re[d] *= alpha;
  }
  for (int i = 0; i < s->pn; i++) {
    s->q[4 * (size_t)i] *= alpha;
    s->q[4 * (size_t)i + 1] *= alpha;
    s->q[4 * (size_t)i + 2] *= alpha;
  }
  mesh_change_scale(s);
  ora_make_pair_list(s);
}

/* -------------------------------------------------------------- export */
double ora_sim_time(const ora_sys *s) { return s->sim_time; }
double ora_zeta(const ora_sys *s) { return s->zeta; }
int ora_particle_number(const ora_sys *s) { return s->pn; }
int ora_total_particle_number(const ora_sys *s) { return s->tn; }
long ora_pair_number(const ora_sys *s) { return s->npairs; }
double ora_buffer_length(const ora_sys *s) { return s->buffer_length; }
int ora_lifetime(const ora_sys *s) { return s->lifetime; }

int ora_mesh_dims(const ora_sys *s, int *m3) {
  m3[0] = s->mx; m3[1] = s->my; m3[2] = s->mz;
  return s->nmesh;
}

void ora_get_state(const ora_sys *s, double *q, double *p, int *type) {
  for (int i = 0; i < s->pn; i++) {
    for (int d = 0; d < 3; d++) {
      if (q) q[3 * (size_t)i + d] = s->q[4 * (size_t)i + d];
      if (p) p[3 * (size_t)i + d] = s->p[4 * (size_t)i + d];
    }
    if (type) type[i] = s->type[i];
  }
}

void ora_set_momenta(ora_sys *s, const double *p) {
  for (int i = 0; i < s->pn; i++)
    for (int d = 0; d < 3; d++) s->p[4 * (size_t)i + d] = p[3 * (size_t)i + d];
}

void ora_get_ghosts(const ora_sys *s, int *src_type, int *shift_code, double *q) {
  for (int g = s->pn; g < s->tn; g++) {
    int k = g - s->pn;
    if (src_type) src_type[k] = s->type[s->g_src[g]];
    if (shift_code) shift_code[k] = s->g_shift[g];
    if (q)
      for (int d = 0; d < 3; d++) q[3 * (size_t)k + d] = s->q[4 * (size_t)g + d];
  }
}

/* canonical pair: (id of the owned key, id of the partner's owner, shift of
   the partner's image); owned-owned pairs are ordered (min id, max id) */
long ora_get_pairs(const ora_sys *s, int *out3) {
  for (long k = 0; k < s->npairs; k++) {
    int i = s->key[k], j = s->partner[k];
    int a = s->type[s->g_src[i]], b = s->type[s->g_src[j]];
    int sh = s->g_shift[j];
    if (sh == 13 && s->g_shift[i] == 13 && b < a) {
      int t = a; a = b; b = t;
    }
    out3[3 * k] = a;
    out3[3 * k + 1] = b;
    out3[3 * k + 2] = sh;
  }
  return s->npairs;
}

void ora_get_cells(const ora_sys *s, int *cell_of_owned) {
  for (int i = 0; i < s->pn; i++) cell_of_owned[i] = s->ppos[i];
}

/* ConfigurationMaker::Execute (confmaker.cc:8-44) */
long ora_fcc(double Lx, double Ly, double Lz, double density, long cap, double *x) {
  const double s = 1.0 / pow(density * 0.25, 1.0 / 3.0);
  const double hs = s * 0.5;
  const int sx = (int)(Lx / s), sy = (int)(Ly / s), sz = (int)(Lz / s);
  const double e = 0.0000001;
  long n = 0;
  for (int iz = 0; iz < sz; iz++)
    for (int iy = 0; iy < sy; iy++)
      for (int ix = 0; ix < sx; ix++) {
        const double b[4][3] = {{ix * s + e, iy * s + e, iz * s + e},
                                {ix * s + e, iy * s + hs + e, iz * s + hs + e},
                                {ix * s + hs + e, iy * s + e, iz * s + hs + e},
                                {ix * s + hs + e, iy * s + hs + e, iz * s + e}};
        for (int k = 0; k < 4; k++) {
          if (!(0.0 <= b[k][0] && b[k][0] < Lx && 0.0 <= b[k][1] && b[k][1] < Ly && 0.0 <= b[k][2] &&
                b[k][2] < Lz))
            continue;
          if (n < cap) {
            x[3 * n] = b[k][0];
            x[3 * n + 1] = b[k][1];
            x[3 * n + 2] = b[k][2];
          }
          n++;
        }
      }
  return n;
}
