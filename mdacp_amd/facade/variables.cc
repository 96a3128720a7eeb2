//----------------------------------------------------------------------
// variables.cc -- host side of the Variables facade: host mirrors of the
// device-resident particle state and the host-only routines of
// kaityo256/mdacp variables.cc (initial velocities, file formats), which
// operate on the mirrors.
//----------------------------------------------------------------------
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <iostream>
#include <mutex>
#include <random>
#include <vector>
#include "mpistream.h"
#include "variables.h"
//----------------------------------------------------------------------
namespace {
// The caller's type of every particle id, by the tag of the unit that named
// it (ids of a plain counter: tag 0).  Units of one process add particles
// from their own threads (MDManager::ExecuteAll) and read each other's
// entries once particles migrate, so the tables share one lock.
const int kIdBits = 24;
const int kMaxTags = 128;
std::vector<int> g_types[kMaxTags];
std::mutex g_types_mu;

int type_of_id(int id, bool tagged) {
  const int tag = tagged ? (id >> kIdBits) : 0;
  const int local = tagged ? (id & ((1 << kIdBits) - 1)) : id;
  return g_types[tag][local];
}

template <class T>
T *grow_aligned(T *old, size_t keep, size_t n) {
  void *p = NULL;
  if (posix_memalign(&p, 64, std::max<size_t>(1, n) * sizeof(T)) != 0) {
    mout << "# Error: host mirror allocation failed" << std::endl;
    exit(1);
  }
  memset(p, 0, std::max<size_t>(1, n) * sizeof(T));
  if (old) {
    memcpy(p, old, keep * sizeof(T));
    free(old);
  }
  return static_cast<T *>(p);
}
}  // namespace
//----------------------------------------------------------------------
// Variables::Variables (variables.cc:16-28): C0 and C2 of the smoothed
// potential for CUTOFF_LENGTH, in the reference's order of operations.
Variables::Variables(void) {
  particle_number = 0;
  total_particle_number = 0;
  SimulationTime = 0.0;
  Zeta = 0.0;
  const double s2 = 1.0 / (CUTOFF_LENGTH * CUTOFF_LENGTH);
  const double s6 = s2 * s2 * s2;
  const double s8 = s6 * s2;
  const double s12 = s6 * s6;
  const double s14 = s12 * s2;
  C2 = 6.0 * s14 - 3.0 * s8;
  C0 = -s12 + s6 - C2 / s2;
  capacity = 0;
  q = NULL;
  p = NULL;
  type = NULL;
  ids = NULL;
  unit_tag = -1;
  dev = NULL;
  dev_count = 0;
  host_stale = false;
  host_dirty = false;
  pair_sums_valid = false;
  Reserve(1024);
}
//----------------------------------------------------------------------
Variables::~Variables(void) {
  free(q);
  free(p);
  free(type);
  free(ids);
}
//----------------------------------------------------------------------
void
Variables::Reserve(int n) {
  if (n <= capacity) return;
  const int nc = std::max(n, capacity + capacity / 2);
  const size_t keep = static_cast<size_t>(capacity);
  q = reinterpret_cast<double(*)[D]>(grow_aligned(reinterpret_cast<double *>(q), keep * D, (size_t)nc * D));
  p = reinterpret_cast<double(*)[D]>(grow_aligned(reinterpret_cast<double *>(p), keep * D, (size_t)nc * D));
  type = grow_aligned(type, keep, nc);
  ids = grow_aligned(ids, keep, nc);
  capacity = nc;
}
//----------------------------------------------------------------------
void
Variables::Check(int status, const char *what) {
  if (status == MDGPU_OK) return;
  show_error(what);
  mout << mdgpu_last_error(dev) << std::endl;
  exit(1);
}
//----------------------------------------------------------------------
int
Variables::NewId(int t) {
  std::lock_guard<std::mutex> lock(g_types_mu);
  std::vector<int> &v = g_types[unit_tag < 0 ? 0 : unit_tag];
  const size_t local = v.size();
  if (unit_tag >= 0 && local >= (size_t(1) << kIdBits)) {
    show_error("a unit of a process grid names at most 2^24 particles");
    exit(1);
  }
  v.push_back(t);
  return unit_tag < 0 ? static_cast<int>(local) : ((unit_tag << kIdBits) | static_cast<int>(local));
}
//----------------------------------------------------------------------
int
Variables::DeviceParticleNumber(void) {
  if (dev == NULL || !host_stale) return particle_number;
  int pn = 0, tn = 0;
  Check(mdgpu_particle_number(dev, &pn, &tn), "particle number");
  return pn;
}
//----------------------------------------------------------------------
// Variables::AddParticle (variables.cc:31-42): appended on the host; pushed to
// the device before the next device operation.
void
Variables::AddParticle(double x[D], double v[D], int t) {
  ToHost();
  Reserve(particle_number + 1);
  const int i = particle_number;
  q[i][X] = x[X];
  q[i][Y] = x[Y];
  q[i][Z] = x[Z];
  p[i][X] = v[X];
  p[i][Y] = v[Y];
  p[i][Z] = v[Z];
  type[i] = t;
  ids[i] = NewId(t);
  particle_number++;
  total_particle_number = particle_number;
}
//----------------------------------------------------------------------
void
Variables::SetParticleNumber(int pn) {
  ToHost();
  if (pn < particle_number) {
    // drop the tail: the device keeps the first pn particles of its order
    particle_number = pn;
    total_particle_number = pn;
    Check(mdgpu_clear_particles(dev), "clear particles");
    dev_count = 0;
    host_dirty = false;
  } else {
    particle_number = pn;
  }
}
//----------------------------------------------------------------------
// Fill the mirrors from the device: owned particles in the device's current
// order, then the ghost images (type and momentum of their owner).
void
Variables::ToHost(void) {
  if (!host_stale || dev == NULL) return;
  int pn = 0, tn = 0;
  Check(mdgpu_particle_number(dev, &pn, &tn), "particle number");
  Reserve(tn);
  std::vector<double> hq((size_t)pn * 3), hp((size_t)pn * 3);
  std::vector<int> hid(pn);
  Check(mdgpu_get_particles(dev, hq.data(), hp.data(), hid.data()), "download particles");
  const int ng = tn - pn;
  std::vector<int> gsrc(ng), gcode(ng);
  std::vector<double> gq((size_t)ng * 3);
  if (ng > 0) Check(mdgpu_get_ghosts(dev, gsrc.data(), gcode.data(), gq.data()), "download ghosts");
  const bool tagged = unit_tag >= 0;
  std::lock_guard<std::mutex> lock(g_types_mu);
  for (int i = 0; i < pn; i++) {
    for (int d = 0; d < 3; d++) {
      q[i][d] = hq[3 * (size_t)i + d];
      p[i][d] = hp[3 * (size_t)i + d];
    }
    ids[i] = hid[i];
    type[i] = type_of_id(hid[i], tagged);
  }
  // ghosts: the owner's type; the owner's momentum when the owner is here
  // (a ghost of another unit of a process grid gets 0: the reference's ghost
  // momentum rows hold nothing either)
  std::vector<int> row_of_id;
  if (!tagged && ng > 0) {
    row_of_id.assign(g_types[0].size(), -1);
    for (int i = 0; i < pn; i++) row_of_id[hid[i]] = i;
  }
  for (int k = 0; k < ng; k++) {
    const int i = pn + k;
    const int src = tagged ? -1 : row_of_id[gsrc[k]];
    for (int d = 0; d < 3; d++) {
      q[i][d] = gq[3 * (size_t)k + d];
      p[i][d] = src >= 0 ? p[src][d] : 0.0;
    }
    ids[i] = gsrc[k];
    type[i] = type_of_id(gsrc[k], tagged);
  }
  particle_number = pn;
  total_particle_number = tn;
  dev_count = pn;
  host_stale = false;
}
//----------------------------------------------------------------------
// Push to the device what the host holds that the device lacks: positions
// and momenta changed on the host (in the device's order) and particles added
// since the last push.
void
Variables::ToDevice(void) {
  if (dev == NULL) return;
  if (host_dirty || particle_number > dev_count) pair_sums_valid = false;
  if (host_dirty && dev_count > 0) {
    std::vector<double> hq((size_t)dev_count * 3), hp((size_t)dev_count * 3);
    for (int i = 0; i < dev_count; i++) {
      for (int d = 0; d < 3; d++) {
        hq[3 * (size_t)i + d] = q[i][d];
        hp[3 * (size_t)i + d] = p[i][d];
      }
    }
    Check(mdgpu_set_positions(dev, hq.data()), "upload positions");
    Check(mdgpu_set_momenta(dev, hp.data()), "upload momenta");
  }
  if (particle_number > dev_count) {
    const int n = particle_number - dev_count;
    std::vector<double> hq((size_t)n * 3), hp((size_t)n * 3);
    std::vector<int> hid(n);
    for (int k = 0; k < n; k++) {
      const int i = dev_count + k;
      for (int d = 0; d < 3; d++) {
        hq[3 * (size_t)k + d] = q[i][d];
        hp[3 * (size_t)k + d] = p[i][d];
      }
      hid[k] = ids[i];
    }
    Check(mdgpu_add_particles(dev, n, hq.data(), hp.data(), hid.data()), "upload particles");
    int pn = 0, tn = 0;
    Check(mdgpu_particle_number(dev, &pn, &tn), "particle number");
    if (pn != particle_number) {
      // a particle outside the box: the device does not take it
      show_error("Invalid position");
      exit(1);
    }
    dev_count = pn;
    // ghosts of an earlier construction are gone until the next one
    host_stale = true;
  }
  host_dirty = false;
}
//----------------------------------------------------------------------
// Variables::SaveConfiguration (variables.cc:45-56): owned and ghost rows.
void
Variables::SaveConfiguration(std::ostream &fs) {
  ToHost();
  for (int i = 0; i < total_particle_number; i++) {
    fs << q[i][X];
    fs << " " << q[i][Y];
    fs << " " << q[i][Z];
    fs << " " << p[i][X];
    fs << " " << p[i][Y];
    fs << " " << p[i][Z];
    fs << "\n";
  }
}
//----------------------------------------------------------------------
// Variables::AdjustPeriodicBoundary (variables.cc:59-79) on the mirrors; the
// device applies the same wrap at every pair-list construction.
void
Variables::AdjustPeriodicBoundary(SimulationInfo *sinfo) {
  ToHost();
  for (int i = 0; i < particle_number; i++) {
    for (int d = 0; d < 3; d++) {
      if (q[i][d] < 0.0) {
        q[i][d] += sinfo->L[d];
      } else if (q[i][d] > sinfo->L[d]) {
        q[i][d] -= sinfo->L[d];
      }
    }
  }
  for (int i = 0; i < particle_number; i++) {
    for (int d = 0; d < 3; d++) {
      if (q[i][d] < 0.0 || q[i][d] > sinfo->L[d]) {
        show_error("Invalid position");
        printf("%f: %d %d %f\n", SimulationTime, i, d, q[i][d]);
        exit(1);
      }
    }
  }
  host_dirty = true;
}
//----------------------------------------------------------------------
// Variables::SetInitialVelocity (variables.cc:82-108): |p| = V0 in a random
// direction from std::mt19937(seed) and uniform_real_distribution(0, 1),
// then the mean removed -- the same library calls in the same order, so the
// momenta are those of the reference for the same particle order.
void
Variables::SetInitialVelocity(const double V0, const int seed) {
  ToHost();
  std::mt19937 mt(seed);
  std::uniform_real_distribution<double> ud(0.0, 1.0);
  for (int i = 0; i < particle_number; i++) {
    double z = ud(mt) * 2.0 - 1.0;
    double phi = ud(mt) * 2.0 * M_PI;
    p[i][X] = V0 * sqrt(1 - z * z) * cos(phi);
    p[i][Y] = V0 * sqrt(1 - z * z) * sin(phi);
    p[i][Z] = V0 * z;
  }
  double avx = 0;
  double avy = 0;
  double avz = 0;
  for (int i = 0; i < particle_number; i++) {
    avx += p[i][X];
    avy += p[i][Y];
    avz += p[i][Z];
  }
  avx /= static_cast<double>(particle_number);
  avy /= static_cast<double>(particle_number);
  avz /= static_cast<double>(particle_number);
  for (int i = 0; i < particle_number; i++) {
    p[i][X] -= avx;
    p[i][Y] -= avy;
    p[i][Z] -= avz;
  }
  host_dirty = true;
}
//----------------------------------------------------------------------
// Variables::SaveToStream / LoadFromStream (variables.cc:111-131): the same
// binary layout (q and p interleaved per component, D components).
void
Variables::SaveToStream(std::ostream &fs) {
  ToHost();
  for (int i = 0; i < particle_number; i++) {
    for (int d = 0; d < D; d++) {
      fs.write((const char *)&q[i][d], sizeof(double));
      fs.write((const char *)&p[i][d], sizeof(double));
    }
  }
}
//----------------------------------------------------------------------
void
Variables::LoadFromStream(std::istream &fs) {
  int pn = 0;
  fs.read((char *)&pn, sizeof(pn));
  fs.read((char *)&SimulationTime, sizeof(SimulationTime));
  fs.read((char *)&Zeta, sizeof(Zeta));
  if (dev != NULL) Check(mdgpu_clear_particles(dev), "clear particles");
  dev_count = 0;
  host_stale = false;
  particle_number = 0;
  total_particle_number = 0;
  Reserve(pn);
  for (int i = 0; i < pn; i++) {
    for (int d = 0; d < D; d++) {
      fs.read((char *)&q[i][d], sizeof(double));
      fs.read((char *)&p[i][d], sizeof(double));
    }
    type[i] = 0;
    ids[i] = NewId(0);
  }
  particle_number = pn;
  total_particle_number = pn;
}
//----------------------------------------------------------------------
// Variables::ChangeScale (variables.cc:134-140) on the mirrors.  MDUnit's
// ChangeScale scales the device state itself.
void
Variables::ChangeScale(double alpha) {
  ToHost();
  for (int i = 0; i < particle_number; i++) {
    q[i][X] *= alpha;
    q[i][Y] *= alpha;
    q[i][Z] *= alpha;
  }
  host_dirty = true;
}
//----------------------------------------------------------------------
// Variables::Shuffle (variables.cc:145-172, debugging): a random permutation
// of the owned particles.
void
Variables::Shuffle(void) {
  ToHost();
  const int pn = particle_number;
  std::vector<int> a(pn);
  for (int i = 0; i < pn; i++) a[i] = i;
  std::shuffle(a.begin(), a.end(), std::mt19937(static_cast<unsigned>(rand())));
  std::vector<double> qb((size_t)pn * D), pb((size_t)pn * D);
  std::vector<int> tb(pn), ib(pn);
  for (int i = 0; i < pn; i++) {
    for (int d = 0; d < D; d++) {
      qb[(size_t)i * D + d] = q[i][d];
      pb[(size_t)i * D + d] = p[i][d];
    }
    tb[i] = type[i];
    ib[i] = ids[i];
  }
  for (int i = 0; i < pn; i++) {
    for (int d = 0; d < D; d++) {
      q[i][d] = qb[(size_t)a[i] * D + d];
      p[i][d] = pb[(size_t)a[i] * D + d];
    }
    type[i] = tb[a[i]];
    ids[i] = ib[a[i]];
  }
  // the device order no longer matches: give it the shuffled set
  if (dev != NULL) Check(mdgpu_clear_particles(dev), "clear particles");
  dev_count = 0;
  host_dirty = false;
}
//----------------------------------------------------------------------
