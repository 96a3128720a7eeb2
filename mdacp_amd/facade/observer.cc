//----------------------------------------------------------------------
// observer.cc -- the device observers (interface of kaityo256/mdacp
// observer.h; sums of observer.cc:5-79 reduced on the GPU by mdgpu_observe).
//----------------------------------------------------------------------
#include "observer.h"
//----------------------------------------------------------------------
namespace {
// pe and virial over the pair list of the resident state: one pass per state,
// shared by the observers that ask for it (MDManager::Pressure and
// TotalEnergy each ask at every observation)
void observe(Variables *vars, double *ke, double *pe, double *vir) {
  vars->ToDevice();
  if (!vars->pair_sums_valid) {
    double *s = vars->pair_sums;
    vars->Check(mdgpu_observe(vars->dev, &s[0], &s[1], &s[2]), "observer");
    vars->pair_sums_valid = true;
  }
  *ke = vars->pair_sums[0];
  *pe = vars->pair_sums[1];
  *vir = vars->pair_sums[2];
}
}  // namespace
//----------------------------------------------------------------------
// KineticEnergyObserver (observer.cc:5-15): needs no pair list.
double
KineticEnergyObserver::Observe(Variables *vars, MeshList *mesh) {
  vars->ToDevice();
  double ke = 0.0;
  vars->Check(mdgpu_observe_kinetic(vars->dev, &ke), "observer");
  return ke;
}
//----------------------------------------------------------------------
double
PotentialEnergyObserver::Observe(Variables *vars, MeshList *mesh) {
  double ke, pe, vir;
  observe(vars, &ke, &pe, &vir);
  return pe;
}
//----------------------------------------------------------------------
double
VirialObserver::Observe(Variables *vars, MeshList *mesh) {
  double ke, pe, vir;
  observe(vars, &ke, &pe, &vir);
  return vir;
}
//----------------------------------------------------------------------
