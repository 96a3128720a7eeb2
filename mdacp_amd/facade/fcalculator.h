//----------------------------------------------------------------------
// fcalculator.h -- MDACP's ForceCalculator (kaityo256/mdacp
// fcalculator.h:12-28) on the MI355X.  Every force routine the reference
// declares runs the device's pair kick over the Verlet list built by
// MDUnit::MakePairList (FP64 or mixed precision, as the domain is set up);
// the integrator and thermostat routines run on the device too.
//----------------------------------------------------------------------
#ifndef fcalculator_h
#define fcalculator_h
#include <random>
#include "mdconfig.h"
#include "variables.h"
#include "meshlist.h"
#include "simulationinfo.h"
//----------------------------------------------------------------------
namespace ForceCalculator {
void CalculateForceBruteforce(Variables *vars, SimulationInfo *sinfo);
void CalculateForceUnroll(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
void CalculateForceSorted(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
void CalculateForceNext(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
void CalculateForcePair(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
void CalculateForceReactless(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
void CalculateForceReactlessSIMD(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
void CalculateForceReactlessSIMD_errsafe(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
#ifdef AVX2
void CalculateForceAVX2(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
#endif
void UpdatePositionHalf(Variables *vars, SimulationInfo *sinfo);
void CalculateForce(Variables *vars, MeshList *mesh, SimulationInfo *sinfo);
void HeatbathZeta(Variables *vars, double ct, SimulationInfo *sinfo);
void HeatbathMomenta(Variables *vars, SimulationInfo *sinfo);
// The noise comes from the device's Philox4x32-10 streams keyed by the unit's
// seed (MDUnit::SetSeed), particle id and call count; `mt` is not drawn from.
void Langevin(Variables *vars, SimulationInfo *sinfo, std::mt19937 &mt);
};
//----------------------------------------------------------------------
#endif
//----------------------------------------------------------------------
