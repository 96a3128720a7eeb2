//----------------------------------------------------------------------
// fcalculator.cc -- ForceCalculator facade (interface of kaityo256/mdacp
// fcalculator.h): the pair kick, the half drift and the thermostats on the
// device.
//----------------------------------------------------------------------
#include <math.h>
#include "fcalculator.h"
//----------------------------------------------------------------------
namespace {
void device_kick(Variables *vars) {
  vars->ToDevice();
  vars->Check(mdgpu_calculate_force(vars->dev), "force");
  vars->DeviceChanged();
}
}  // namespace
//----------------------------------------------------------------------
// ForceCalculator::CalculateForce (fcalculator.cc:35-50) with the semantics of
// CalculateForceNext (fcalculator.cc:229-290): p_i += df (q_j - q_i) over the
// listed partners within the cutoff.  The reference's other variants compute
// the same kick.
void
ForceCalculator::CalculateForce(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  device_kick(vars);
}
void
ForceCalculator::CalculateForceBruteforce(Variables *vars, SimulationInfo *sinfo) {
  device_kick(vars);
}
void
ForceCalculator::CalculateForceUnroll(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  device_kick(vars);
}
void
ForceCalculator::CalculateForceSorted(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  device_kick(vars);
}
void
ForceCalculator::CalculateForceNext(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  device_kick(vars);
}
void
ForceCalculator::CalculateForcePair(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  device_kick(vars);
}
void
ForceCalculator::CalculateForceReactless(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  device_kick(vars);
}
void
ForceCalculator::CalculateForceReactlessSIMD(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  device_kick(vars);
}
void
ForceCalculator::CalculateForceReactlessSIMD_errsafe(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  device_kick(vars);
}
#ifdef AVX2
void
ForceCalculator::CalculateForceAVX2(Variables *vars, MeshList *mesh, SimulationInfo *sinfo) {
  device_kick(vars);
}
#endif
//----------------------------------------------------------------------
// ForceCalculator::UpdatePositionHalf (fcalculator.cc:22-32)
void
ForceCalculator::UpdatePositionHalf(Variables *vars, SimulationInfo *sinfo) {
  vars->ToDevice();
  vars->Check(mdgpu_update_position_half(vars->dev), "half drift");
  vars->DeviceChanged();
}
//----------------------------------------------------------------------
// ForceCalculator::HeatbathZeta (fcalculator.cc:744-749): a scalar update of
// Variables::Zeta, done on the host in the reference's order of operations.
void
ForceCalculator::HeatbathZeta(Variables *vars, double current_temperature, SimulationInfo *sinfo) {
  const double dt2 = sinfo->TimeStep * 0.5;
  const double tau = sinfo->HeatbathTau;
  double t1 = (current_temperature - sinfo->AimedTemperature) / (tau * tau);
  vars->Zeta += t1 * dt2;
}
//----------------------------------------------------------------------
// ForceCalculator::HeatbathMomenta (fcalculator.cc:752-762): p *= exp(-dt/2 Zeta)
// on the device, with the host's Zeta.
void
ForceCalculator::HeatbathMomenta(Variables *vars, SimulationInfo *sinfo) {
  vars->ToDevice();
  vars->Check(mdgpu_set_thermostat(vars->dev, MDGPU_NOSEHOOVER, sinfo->AimedTemperature, sinfo->HeatbathGamma,
                                   sinfo->HeatbathTau),
              "thermostat");
  vars->Check(mdgpu_set_zeta(vars->dev, vars->Zeta), "thermostat");
  vars->Check(mdgpu_heatbath_momenta(vars->dev), "heat bath");
  vars->DeviceChanged();
}
//----------------------------------------------------------------------
// ForceCalculator::Langevin (fcalculator.cc:765-778): p += (-gamma p + N(0, hb_D)) dt
// per component on the device; see fcalculator.h for the noise source.
void
ForceCalculator::Langevin(Variables *vars, SimulationInfo *sinfo, std::mt19937 &mt) {
  vars->ToDevice();
  vars->Check(mdgpu_set_thermostat(vars->dev, MDGPU_LANGEVIN, sinfo->AimedTemperature, sinfo->HeatbathGamma,
                                   sinfo->HeatbathTau),
              "thermostat");
  vars->Check(mdgpu_langevin(vars->dev), "Langevin");
  vars->DeviceChanged();
}
//----------------------------------------------------------------------
