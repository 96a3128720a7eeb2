//----------------------------------------------------------------------
// meshlist.h -- MDACP's MeshList (kaityo256/mdacp meshlist.h:10-62) on the
// MI355X.  The cell grid and the Verlet list are built on the device
// (MDUnit::MakePairList); this class keeps the reference's public interface.
// The list accessors (GetKeyParticles, GetPartnerParticles, GetPairNumber,
// GetSortedList, GetKeyPointer, ...) return host copies of the device list
// in the reference's COO / CSR form, with indices into the host mirrors of
// Variables (owned particles, then ghosts), made when first asked for after
// each construction.
//----------------------------------------------------------------------
#ifndef meshlist_h
#define meshlist_h
#include <vector>
#include "mdconfig.h"
#include "variables.h"
#include "mdrect.h"
//----------------------------------------------------------------------
class MeshList {
private:
  int number_of_pairs;
  int number_of_constructions;
  long built;         // constructions done on the device
  long materialized;  // construction the host copies below belong to
  std::vector<int> key_particles;
  std::vector<int> partner_particles;
  std::vector<int> number_of_partners;
  std::vector<int> key_pointer;
  std::vector<int> key_pointer2;
  std::vector<int> sorted_list;
  Variables *vars;
  void Materialize(void);

public:
  MeshList(SimulationInfo *sinfo, MDRect &r);
  ~MeshList(void);

  void ChangeScale(SimulationInfo *sinfo, MDRect &myrect);

  int *GetKeyParticles(void) {Materialize(); return key_particles.data();}
  int *GetPartnerParticles(void) {Materialize(); return partner_particles.data();}
  int GetPairNumber(void) {Materialize(); return number_of_pairs;}
  int GetPartnerNumber(int i) {Materialize(); return number_of_partners[i];}
  int *GetSortedList(void) {Materialize(); return sorted_list.data();}
  int GetKeyPointer(int i) {Materialize(); return key_pointer[i];}
  int *GetKeyPointerP(void) {Materialize(); return key_pointer.data();}
  int *GetNumberOfPartners(void) {Materialize(); return number_of_partners.data();}
  void Sort(Variables *vars, SimulationInfo *sinfo, MDRect &myrect);
  int GetNumberOfConstructions(void) {return number_of_constructions;}
  void ClearNumberOfConstructions(void) {number_of_constructions = 0;}

  void MakeList(Variables *vars, SimulationInfo *sinfo, MDRect &myrect);
  void MakeListBruteforce(Variables *vars, SimulationInfo *sinfo, MDRect &myrect);
  void ShowPairs(void);
  void ShowSortedList(Variables *vars);

  // ---- device side (no counterpart in the reference) ----
  void Attach(Variables *v) {vars = v;}
  void Built(void) {built++; number_of_constructions++;}
};
//----------------------------------------------------------------------
#endif
//----------------------------------------------------------------------
