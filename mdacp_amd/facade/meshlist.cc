//----------------------------------------------------------------------
// meshlist.cc -- MeshList facade (interface of kaityo256/mdacp meshlist.h).
// Construction happens on the device; the COO and CSR lists of the reference
// (key_particles / partner_particles, number_of_partners / key_pointer /
// sorted_list, meshlist.cc:58-87) are made on the host from the device list
// when first asked for after a construction.
//----------------------------------------------------------------------
#include <stdio.h>
#include <stdint.h>
#include <vector>
#include "mpistream.h"
#include "meshlist.h"
//----------------------------------------------------------------------
MeshList::MeshList(SimulationInfo *sinfo, MDRect &r) {
  number_of_pairs = 0;
  number_of_constructions = 0;
  built = 0;
  materialized = -1;
  vars = NULL;
}
//----------------------------------------------------------------------
MeshList::~MeshList(void) {
}
//----------------------------------------------------------------------
// The cell grid follows the box on the device (MDUnit::ChangeScale).
void
MeshList::ChangeScale(SimulationInfo *sinfo, MDRect &myrect) {
}
//----------------------------------------------------------------------
// MeshList::MakeList (meshlist.cc:58-87) and the MDUnit::MakePairList around
// it: on the device.
void
MeshList::MakeList(Variables *v, SimulationInfo *sinfo, MDRect &myrect) {
  v->ToDevice();
  v->Check(mdgpu_make_pair_list(v->dev), "pair-list construction");
  v->DeviceChanged();
  Built();
}
//----------------------------------------------------------------------
// The reference's brute-force construction lists the same pairs.
void
MeshList::MakeListBruteforce(Variables *v, SimulationInfo *sinfo, MDRect &myrect) {
  MakeList(v, sinfo, myrect);
}
//----------------------------------------------------------------------
// Particles are always kept in cell order on the device.
void
MeshList::Sort(Variables *v, SimulationInfo *sinfo, MDRect &myrect) {
}
//----------------------------------------------------------------------
void
MeshList::Materialize(void) {
  if (materialized == built || vars == NULL) return;
  vars->ToHost();
  int64_t n = 0;
  vars->Check(mdgpu_get_pair_indices(vars->dev, &n, NULL, NULL), "pair list export");
  key_particles.assign((size_t)n, 0);
  partner_particles.assign((size_t)n, 0);
  vars->Check(mdgpu_get_pair_indices(vars->dev, &n, key_particles.data(), partner_particles.data()),
              "pair list export");
  number_of_pairs = static_cast<int>(n);
  // CSR by key particle, partners in COO order (meshlist.cc:58-87)
  const int tn = vars->GetTotalParticleNumber();
  number_of_partners.assign(tn + 1, 0);
  key_pointer.assign(tn + 1, 0);
  key_pointer2.assign(tn + 1, 0);
  for (int k = 0; k < number_of_pairs; k++) number_of_partners[key_particles[k]]++;
  int sum = 0;
  for (int i = 0; i < tn; i++) {
    key_pointer[i] = sum;
    sum += number_of_partners[i];
  }
  key_pointer[tn] = sum;
  sorted_list.assign((size_t)number_of_pairs, 0);
  for (int i = 0; i < tn; i++) key_pointer2[i] = key_pointer[i];
  for (int k = 0; k < number_of_pairs; k++) {
    const int i = key_particles[k];
    sorted_list[key_pointer2[i]] = partner_particles[k];
    key_pointer2[i]++;
  }
  materialized = built;
}
//----------------------------------------------------------------------
// MeshList::ShowPairs / ShowSortedList (meshlist.cc:313-330): the same format.
void
MeshList::ShowPairs(void) {
  Materialize();
  for (int i = 0; i < number_of_pairs; i++) {
    printf("(%05d,%05d)\n", key_particles[i], partner_particles[i]);
  }
}
//----------------------------------------------------------------------
void
MeshList::ShowSortedList(Variables *v) {
  Materialize();
  const int pn = v->GetTotalParticleNumber();
  for (int i = 0; i < pn; i++) {
    const int np = GetPartnerNumber(i);
    const int kp = GetKeyPointer(i);
    for (int k = 0; k < np; k++) {
      const int j = sorted_list[kp + k];
      printf("(%05d,%05d)\n", i, j);
    }
  }
}
//----------------------------------------------------------------------
