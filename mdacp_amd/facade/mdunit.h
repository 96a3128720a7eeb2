//----------------------------------------------------------------------
// mdunit.h -- MDACP's MDUnit and Executor (kaityo256/mdacp mdunit.h:20-75)
// on the MI355X.  One MDUnit is one mdgpu domain: its particles, ghost
// images, cell grid and Verlet list live in HBM and every step operation
// MDManager calls runs on the GPU through the mdgpu C ABI (include/mdgpu.h).
//
// The public interface is the reference's, so MDManager, ProjectManager and
// the projects compile and run unchanged on top of it.
//
// One unit per process (OMP_NUM_THREADS=1): the particle migration and halo
// exchange MDManager drives through send_buffer / Receive* reduce, for a
// domain that is its own neighbour on all six faces, to the periodic wrap and
// ghost images the device makes itself (at MakePairList, and at the halo
// refresh MakeBufferForBorderParticles(D_LEFT) triggers).
//
// Several units in one process (OMP_NUM_THREADS > 1, one MPI process): each
// unit is a device unit of one group (mdgpu_create_unit) on the process's
// GPU, and MDManager's calls move particles and ghosts device to device:
// MakeBufferForSendingParticle fills the unit's device send buffer,
// ReceiveParticles(sender's send_buffer) appends the sender's device buffer
// (the host vectors stay empty; the sender is found by the address of its
// send_buffer), AdjustPeriodicBoundary wraps and bins on the device, and
// FindBorderParticles(D_LEFT) places every ghost the six passes would bring,
// from the units that own the particles.
//
// Environment (read when the unit is created):
//   MDACP_GPU_PRECISION  "fp64" (default) or "mixed"
//   MDACP_GPU_DEVICE     HIP device ordinal (default 0)
//----------------------------------------------------------------------
#ifndef mdunit_h
#define mdunit_h
#include <stdio.h>
#include <vector>
#include <random>
#include "mdconfig.h"
#include "parainfo.h"
#include "simulationinfo.h"
#include "mdrect.h"
#include "parameter.h"
#include "variables.h"
#include "meshlist.h"
#include "pairlist.h"
#include "observer.h"
#include "fcalculator.h"
//----------------------------------------------------------------------
class MDUnit;
//----------------------------------------------------------------------
class Executor {
public:
  virtual void Execute(MDUnit *) {};
};
//----------------------------------------------------------------------
class MDUnit {
private:
  Variables *vars;
  SimulationInfo *sinfo;
  ParaInfo *pinfo;
  MeshList *mesh;
  PairList *plist;
  const int id;
  MDRect myrect;
  std::mt19937 mt;
  bool list_built;  // the device holds a list for its current particles
  bool unit_mode;   // one of several units of the process (a device unit)
  static MDUnit *Sender(const std::vector<ParticleInfo> &buffer);
public:
  MDUnit (int id_, SimulationInfo *si, ParaInfo *pi);
  ~MDUnit(void);
  std::vector<ParticleInfo> send_buffer;
  int GetID(void) {return id;}
  MDRect * GetRect(void) {return &myrect;}
  Variables *GetVariables(void);
  void SaveConfiguration(void);
  void SaveAsCdview(std::ofstream &ofs);
  void AddParticle(double x[D], double v[D], int type = 0);
  void AddParticle(double x[D], int type = 0);
  double * GetSystemSize(void) {return sinfo->L;}
  double GetTimeStep(void) {return sinfo->TimeStep;}
  void SetInitialVelocity(double v0);
  int GetParticleNumber(void) {return vars->DeviceParticleNumber();}
  int GetTotalParticleNumber(void) {return vars->GetTotalParticleNumber();}
  void SetTotalParticleNumber(int n) {vars->SetTotalParticleNumber(n);}

  void CalculateForce(void) {ForceCalculator::CalculateForce(vars, mesh, sinfo);}
  void UpdatePositionHalf(void) {ForceCalculator::UpdatePositionHalf(vars, sinfo);}
  void HeatbathZeta(double t) {ForceCalculator::HeatbathZeta(vars, t, sinfo);}
  void HeatbathMomenta(void) {ForceCalculator::HeatbathMomenta(vars, sinfo);}
  void SetSeed(const int seed);
  void Langevin(void) {ForceCalculator::Langevin(vars, sinfo, mt);}

  void MakeBufferForSendingParticle(const int dir);
  void FindBorderParticles(const int dir);
  void MakeBufferForBorderParticles(const int dir);
  void ReceiveParticles(std::vector<ParticleInfo> &recv_buffer);
  void AdjustPeriodicBoundary(void);
  void ReceiveBorderParticles(std::vector<ParticleInfo> &recv_buffer);
  double ObserveDouble(DoubleObserver *obs);
  int IntegerDouble(IntegerObserver *obs);
  void Execute(Executor *ex);
  void MakePairList(void);
  void ShowPairs(void) {vars->ToHost(); mesh->ShowPairs();}
  bool IsPairListExpired(void) {return plist->IsPairListExpired(vars, sinfo);}
  //
  void ChangeScale(double alpha);
};
//----------------------------------------------------------------------
#endif
//----------------------------------------------------------------------
