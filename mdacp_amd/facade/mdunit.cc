//----------------------------------------------------------------------
// mdunit.cc -- MDUnit on the MI355X (interface of kaityo256/mdacp
// mdunit.h:25-75).  Each unit owns one mdgpu domain; the operations MDManager
// calls map onto the C ABI of include/mdgpu.h.
//----------------------------------------------------------------------
#include <iostream>
#include <fstream>
#include <string>
#include <stdlib.h>
#include <string.h>
#include "mpistream.h"
#include "mdunit.h"
//----------------------------------------------------------------------
namespace {
int env_int(const char *name, int def) {
  const char *v = getenv(name);
  return v ? atoi(v) : def;
}
int env_precision(void) {
  const char *v = getenv("MDACP_GPU_PRECISION");
  if (v == NULL || std::string(v) == "fp64") return MDGPU_FP64;
  if (std::string(v) == "mixed") return MDGPU_MIXED;
  mout << "# Error: MDACP_GPU_PRECISION must be fp64 or mixed" << std::endl;
  exit(1);
}
}  // namespace
//----------------------------------------------------------------------
// MDUnit::MDUnit (mdunit.cc:8-27): the unit's box from the process grid, then
// a device domain for it (cutoff CUTOFF_LENGTH, margin BufferLength, the
// CheckListLength of the displacement check).
MDUnit::MDUnit(int id_, SimulationInfo *si, ParaInfo *pi):
  id(id_) {
  vars = new Variables();
  plist = new PairList();
  sinfo = si;
  pinfo = pi;
  int grid_size[D];
  int grid_position[D];
  pinfo->GetGridSize(grid_size);
  pinfo->GetGridPosition(id, grid_position);
  if (grid_size[X] != 1 || grid_size[Y] != 1 || grid_size[Z] != 1) {
    show_error("the GPU build runs one domain per process: use mpirun -np 1 and OMP_NUM_THREADS=1");
    exit(1);
  }
  double s[D], e[D];
  for (int d = 0; d < D; d++) {
    double ul = sinfo->L[d] / static_cast<double>(grid_size[d]);
    s[d] = ul * static_cast<double>(grid_position[d]);
    e[d] = s[d] + ul;
  }
  myrect = MDRect(s, e);
  mesh = new MeshList(sinfo, myrect);
  mt.seed(GetID());
  list_built = false;

  mdgpu_params prm;
  memset(&prm, 0, sizeof(prm));
  for (int d = 0; d < 3; d++) prm.L[d] = sinfo->L[d];
  prm.dt = sinfo->TimeStep;
  prm.rc = CUTOFF_LENGTH;
  prm.margin = sinfo->BufferLength;
  prm.check_list_length = sinfo->CheckListLength;
  prm.precision = env_precision();
  prm.capacity = 0;
  prm.device = env_int("MDACP_GPU_DEVICE", 0);
  mdgpu_domain *d = NULL;
  if (mdgpu_create(&prm, &d) != MDGPU_OK) {
    show_error("could not create the GPU domain (no HIP device, or out of memory)");
    exit(1);
  }
  vars->AttachDevice(d);
  mesh->Attach(vars);
  vars->Check(mdgpu_set_seed(d, static_cast<uint64_t>(GetID())), "seed");
}
//----------------------------------------------------------------------
MDUnit::~MDUnit(void) {
  mdgpu_domain *d = vars->dev;
  delete vars;
  delete mesh;
  delete plist;
  mdgpu_destroy(d);
}
//----------------------------------------------------------------------
// Host code gets the mirrors, current with the device; they are pushed back
// before the next device operation in case it changed them.
Variables *
MDUnit::GetVariables(void) {
  vars->ToHost();
  vars->host_dirty = true;
  return vars;
}
//----------------------------------------------------------------------
void
MDUnit::AddParticle(double x[D], int type) {
  double v[D];
  v[X] = 0.0;
  v[Y] = 0.0;
  v[Z] = 0.0;
  AddParticle(x, v, type);
}
//----------------------------------------------------------------------
// MDUnit::AddParticle (mdunit.cc:36-42): only particles inside the unit's box.
void
MDUnit::AddParticle(double x[D], double v[D], int type) {
  if (myrect.IsInside(x)) {
    vars->AddParticle(x, v, type);
    list_built = false;
  }
}
//----------------------------------------------------------------------
void
MDUnit::SetInitialVelocity(double v0) {
  vars->SetInitialVelocity(v0, id);
}
//----------------------------------------------------------------------
// MDUnit::SaveAsCdview (mdunit.cc:52-67): the same line format.
void
MDUnit::SaveAsCdview(std::ofstream &ofs) {
  vars->ToHost();
  double (*q)[D] = vars->q;
  double (*p)[D] = vars->p;
  int *type = vars->type;
  const int pn = vars->GetParticleNumber();
  for (int i = 0; i < pn; i++) {
    ofs << "0 ";
    ofs << type[i] << " ";
    ofs << q[i][X] << " ";
    ofs << q[i][Y] << " ";
    ofs << q[i][Z] << " ";
    ofs << p[i][X] << " ";
    ofs << p[i][Y] << " ";
    ofs << p[i][Z] << "\n";
  }
}
//----------------------------------------------------------------------
// MDUnit::SaveConfiguration (mdunit.cc:70-75): conf<id>.cnf.
void
MDUnit::SaveConfiguration(void) {
  char filename[256];
  sprintf(filename, "conf%03d.cnf", id);
  std::ofstream fs(filename);
  vars->SaveConfiguration(fs);
}
//----------------------------------------------------------------------
void
MDUnit::SetSeed(const int seed) {
  mt.seed(seed + GetID());
  vars->Check(mdgpu_set_seed(vars->dev, static_cast<uint64_t>(seed + GetID())), "seed");
}
//----------------------------------------------------------------------
// Migration (MakeBufferForSendingParticle / ReceiveParticles, mdunit.cc:78-122):
// a single domain is its own neighbour on every face, so particles that leave
// through one face come back through the opposite one -- the periodic wrap,
// which the device applies at every pair-list construction.  Nothing is sent.
void
MDUnit::MakeBufferForSendingParticle(const int dir) {
  send_buffer.clear();
}
//----------------------------------------------------------------------
void
MDUnit::ReceiveParticles(std::vector<ParticleInfo> &recv_buffer) {
  for (size_t k = 0; k < recv_buffer.size(); k++) {
    ParticleInfo &pi = recv_buffer[k];
    AddParticle(pi.q, pi.p, pi.type);
  }
}
//----------------------------------------------------------------------
// Ghosts (FindBorderParticles / MakeBufferForBorderParticles /
// ReceiveBorderParticles, mdunit.cc:125-191): the device finds the border
// particles and places their images at each construction, and refreshes the
// images from the owners' current positions in one pass over all six
// directions, done here for the first direction of MDManager's sequence.
void
MDUnit::FindBorderParticles(const int dir) {
}
//----------------------------------------------------------------------
void
MDUnit::MakeBufferForBorderParticles(const int dir) {
  send_buffer.clear();
  if (dir == D_LEFT && list_built) {
    vars->ToDevice();
    vars->Check(mdgpu_send_border_particles(vars->dev), "halo refresh");
    vars->DeviceChanged();
  }
}
//----------------------------------------------------------------------
void
MDUnit::ReceiveBorderParticles(std::vector<ParticleInfo> &recv_buffer) {
  if (!recv_buffer.empty()) {
    show_error("the GPU build runs one domain per process: ghost images from another unit");
    exit(1);
  }
}
//----------------------------------------------------------------------
// Observers: device observers reduce on the GPU; any other observer gets the
// host mirrors (owned particles, ghosts) and the host copy of the pair list.
double
MDUnit::ObserveDouble(DoubleObserver *obs) {
  if (dynamic_cast<DeviceObserver *>(obs) == NULL) vars->ToHost();
  return obs->Observe(vars, mesh);
}
//----------------------------------------------------------------------
int
MDUnit::IntegerDouble(IntegerObserver *obs) {
  if (dynamic_cast<DeviceObserver *>(obs) == NULL) vars->ToHost();
  return obs->Observe(vars, mesh);
}
//----------------------------------------------------------------------
void
MDUnit::Execute(Executor *ex) {
  vars->ToHost();
  vars->host_dirty = true;
  const int before = vars->GetParticleNumber();
  ex->Execute(this);
  if (vars->GetParticleNumber() != before) list_built = false;
}
//----------------------------------------------------------------------
// MDUnit::MakePairList (mdunit.cc:194-198): PairList::Init and MeshList::MakeList
// on the device (wrap, ghost images, cell binning, Verlet list, snapshot).
void
MDUnit::MakePairList(void) {
  mesh->MakeList(vars, sinfo, myrect);
  plist->Init(vars, sinfo);
  list_built = true;
}
//----------------------------------------------------------------------
// MDUnit::ChangeScale (mdunit.cc:201-205): positions, the box and the cell
// grid scaled by alpha on the device (which rebuilds the list around it).
void
MDUnit::ChangeScale(double alpha) {
  vars->ToDevice();
  myrect.ChangeScale(alpha);
  vars->Check(mdgpu_change_scale(vars->dev, alpha), "change scale");
  vars->DeviceChanged();
  mesh->Built();
  plist->Init(vars, sinfo);
  list_built = true;
}
//----------------------------------------------------------------------
