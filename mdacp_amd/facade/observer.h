//----------------------------------------------------------------------
// observer.h -- MDACP's observers (kaityo256/mdacp observer.h:8-40) on the
// MI355X.  The kinetic energy, potential energy and virial observers are
// device observers: their sums are reduced on the GPU from the resident
// state and pair list (in a fixed order, so repeated calls agree bitwise).
// Observers written against the reference interface (subclasses of
// DoubleObserver / IntegerObserver that read vars->q, vars->p or the mesh
// lists) still work: MDUnit fills the host mirrors before calling them.
//----------------------------------------------------------------------
#ifndef observe_h
#define observe_h
#include "variables.h"
#include "meshlist.h"
#include "simulationinfo.h"
//----------------------------------------------------------------------
class IntegerObserver {
public:
  virtual int Observe(Variables * vars, MeshList *mesh) = 0;
};
//----------------------------------------------------------------------
class DoubleObserver {
public:
  virtual double Observe(Variables * vars, MeshList *mesh) = 0;
};
//----------------------------------------------------------------------
// Marks an observer whose Observe runs on the device (no counterpart in the
// reference): MDUnit does not fill the host mirrors for it.
class DeviceObserver {
public:
  virtual ~DeviceObserver(void) {}
};
//----------------------------------------------------------------------
class ParticleNumberObserver : public IntegerObserver, public DeviceObserver {
public:
  int Observe(Variables *vars, MeshList *mesh = NULL) {return vars->GetParticleNumber();}
};
//----------------------------------------------------------------------
class KineticEnergyObserver : public DoubleObserver, public DeviceObserver {
public:
  double Observe(Variables * vars, MeshList *mesh);
};
//----------------------------------------------------------------------
class PotentialEnergyObserver : public DoubleObserver, public DeviceObserver {
private:
  SimulationInfo *sinfo;
public:
  PotentialEnergyObserver(SimulationInfo *sinfo_) {sinfo = sinfo_;}
  double Observe(Variables * vars, MeshList *mesh);
};
//----------------------------------------------------------------------
class VirialObserver : public DoubleObserver, public DeviceObserver {
public:
  double Observe(Variables * vars, MeshList *mesh);
};
//----------------------------------------------------------------------
#endif
//----------------------------------------------------------------------
