//----------------------------------------------------------------------
// observer.cc -- the device observers (interface of kaityo256/mdacp
// observer.h; sums of observer.cc:5-79 reduced on the GPU by mdgpu_observe).
//----------------------------------------------------------------------
#include "observer.h"
//----------------------------------------------------------------------
namespace {
// pe and virial over the pair list of the resident state
void observe(Variables *vars, double *ke, double *pe, double *vir) {
  vars->ToDevice();
  vars->Check(mdgpu_observe(vars->dev, ke, pe, vir), "observer");
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
