//----------------------------------------------------------------------
// variables.h -- MDACP's Variables (kaityo256/mdacp variables.h:9-39) on the
// MI355X: the same public members and methods, with the particle state
// resident in HBM in an mdgpu domain.  q, p and type are host mirrors of
// that state: they are filled from the device when host code asks for them
// (MDUnit::GetVariables, executors, observers that are not device
// observers) and pushed back before the next device operation.
//
// Differences from the reference class, none visible to its callers:
//  - q, p, type are pointers (double (*)[D], int *) into mirrors that grow
//    with the particle count, so `double (*q)[D] = vars->q;` and q[i][X]
//    read as before and there is no compile-time limit N;
//  - each particle also carries a unique id used by the device to order
//    particles inside a cell deterministically; `type` is the caller's value.
//    The unit that adds a particle names it: its own counter, or, for the
//    units of a process grid, unit id * 2^24 + counter, so that ids stay
//    unique and reproducible when particles migrate between units.
//----------------------------------------------------------------------
#ifndef variables_h
#define variables_h
//----------------------------------------------------------------------
#include <iostream>
#include <vector>
#include "mdconfig.h"
#include "simulationinfo.h"
#include "mdgpu.h"
//----------------------------------------------------------------------
class Variables {
private:
  int particle_number;
  int total_particle_number;
  double C0, C2;
  int capacity;                 // rows of the host mirrors (owned + ghosts)
  int *ids;                     // particle id per mirror row
  int unit_tag;                 // -1: ids are a plain counter; else the unit id of a process grid
  void Reserve(int n);
  int NewId(int t);             // the next id of this unit, with its caller's type

public:
  Variables(void);
  ~Variables(void);
  int *type;
  double (*q)[D];
  double (*p)[D];
  double Zeta;
  double SimulationTime;
  double GetC0(void) {return C0;}
  double GetC2(void) {return C2;}
  int GetParticleNumber(void) {return particle_number;}
  int GetTotalParticleNumber(void) {return total_particle_number;}

  void AddParticle(double x[D], double v[D], int t = 0);
  void SaveToStream(std::ostream &fs);
  void LoadFromStream(std::istream &fs);
  void SaveConfiguration(std::ostream &fs);
  void SetParticleNumber(int pn);
  void SetTotalParticleNumber(int pn) {total_particle_number = pn;}
  void AdjustPeriodicBoundary(SimulationInfo *sinfo);
  void SetInitialVelocity(double V0, const int id);

  void ChangeScale(double alpha);

  // For Debug
  void Shuffle(void);

  // ---- device side (no counterpart in the reference) ----
  mdgpu_domain *dev;  // the domain holding the particle state
  int dev_count;      // owned particles the device holds
  bool host_stale;    // the device changed since the mirrors were filled
  bool host_dirty;    // the mirrors may hold changes the device lacks
  void AttachDevice(mdgpu_domain *d) {dev = d;}
  void SetUnitTag(int tag) {unit_tag = tag;}  // before the first AddParticle
  int DeviceParticleNumber(void);  // owned particles, current with the device
  void ToHost(void);    // refresh the mirrors (owned and ghosts) if stale
  void ToDevice(void);  // push particles added on the host and host changes
  void DeviceChanged(void) {host_stale = true; pair_sums_valid = false;}
  // the last mdgpu_observe pair sums of the resident state (kinetic energy,
  // potential energy, virial), kept until the state changes so that the
  // potential-energy and virial observers of one observation share one pass
  bool pair_sums_valid;
  double pair_sums[3];
  void Check(int status, const char *what);  // report a device error and exit(1)
};
//----------------------------------------------------------------------
#endif
//----------------------------------------------------------------------
