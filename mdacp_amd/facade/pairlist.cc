//----------------------------------------------------------------------
// pairlist.cc -- PairList facade (interface of kaityo256/mdacp pairlist.h).
// The snapshot, the validity budget and the displacement check
// (pairlist.cc:11-121) run on the device; this class keeps the lifetime and
// freshness the reference reports.
//----------------------------------------------------------------------
#include "pairlist.h"
//----------------------------------------------------------------------
// PairList::Init (pairlist.cc:11-23): the device took its snapshot during the
// construction that precedes this call.
void
PairList::Init(Variables *vars, SimulationInfo *sinfo) {
  initialized = true;
  is_fresh = true;
  lifetime = 0;
}
//----------------------------------------------------------------------
// PairList::IsPairListExpired (pairlist.cc:26-45)
bool
PairList::IsPairListExpired(Variables *vars, SimulationInfo *sinfo) {
  if (!initialized) {
    return true;
  }
  is_fresh = false;
  lifetime++;
  vars->ToDevice();
  int expired = 0;
  vars->Check(mdgpu_is_pair_list_expired(vars->dev, &expired), "pair-list expiry check");
  return expired != 0;
}
//----------------------------------------------------------------------
