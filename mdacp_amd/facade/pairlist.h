//----------------------------------------------------------------------
// pairlist.h -- MDACP's PairList (kaityo256/mdacp pairlist.h:10-31) on the
// MI355X: the validity budget and the displacement check run on the device
// (mdgpu_is_pair_list_expired); this class keeps the reference's interface
// and its lifetime / freshness bookkeeping.
//----------------------------------------------------------------------
#ifndef pairlist_h
#define pairlist_h
#include "mdconfig.h"
#include "meshlist.h"
#include "variables.h"
//----------------------------------------------------------------------
class PairList {
private:
  int lifetime;
  bool initialized;
  bool is_fresh;

public:
  PairList(void) {initialized = false; is_fresh = false; lifetime = 0;}
  void Init(Variables *vars, SimulationInfo *sinfo);
  bool IsFresh(void) {return is_fresh;}
  void SetFresh(bool b) {is_fresh = b;}
  bool IsPairListExpired(Variables *vars, SimulationInfo *sinfo);
  int GetLifeTime(void) {return lifetime;}
};
//----------------------------------------------------------------------
#endif
//----------------------------------------------------------------------
