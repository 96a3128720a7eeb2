//----------------------------------------------------------------------
// mdunit.cc -- MDUnit on the MI355X (interface of kaityo256/mdacp
// mdunit.h:25-75).  Each unit owns one mdgpu domain; the operations MDManager
// calls map onto the C ABI of include/mdgpu.h.
//----------------------------------------------------------------------
#include <iostream>
#include <fstream>
#include <mutex>
#include <string>
#include <vector>
#include <stdlib.h>
#include <string.h>
#include <mpi.h>
#include "mpistream.h"
#include "mdunit.h"
//----------------------------------------------------------------------
namespace {
// The units of this process that are device units of one group, in the order
// they were made (MDManager makes them from its OpenMP threads).
std::mutex g_units_mu;
std::vector<MDUnit *> g_units;
std::vector<mdgpu_domain *> g_unit_devs;

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
  unit_mode = grid_size[X] != 1 || grid_size[Y] != 1 || grid_size[Z] != 1;
  if (unit_mode) {
    int procs = 1;
    MPI_Comm_size(MPI_COMM_WORLD, &procs);
    if (procs != 1) {
      show_error("the GPU build runs the units of one process on its GPU: use mpirun -np 1");
      exit(1);
    }
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
  if (!unit_mode) {
    if (mdgpu_create(&prm, &d) != MDGPU_OK) {
      show_error("could not create the GPU domain (no HIP device, or out of memory)");
      exit(1);
    }
  } else {
    // a device unit of the process's group: the first unit makes the group
    mdgpu_unit u;
    for (int k = 0; k < 3; k++) {
      u.grid[k] = grid_size[k];
      u.pos[k] = grid_position[k];
      u.s[k] = s[k];
      u.e[k] = e[k];
    }
    std::lock_guard<std::mutex> lock(g_units_mu);
    mdgpu_domain *join = g_unit_devs.empty() ? NULL : g_unit_devs[0];
    if (mdgpu_create_unit(&prm, &u, join, &d) != MDGPU_OK) {
      show_error("could not create the GPU unit (no HIP device, out of memory, or a unit box under two search lengths)");
      if (join) mout << mdgpu_last_error(join) << std::endl;
      exit(1);
    }
    g_units.push_back(this);
    g_unit_devs.push_back(d);
    vars->SetUnitTag(GetID());
  }
  vars->AttachDevice(d);
  mesh->Attach(vars);
  vars->Check(mdgpu_set_seed(d, static_cast<uint64_t>(GetID())), "seed");
}
//----------------------------------------------------------------------
MDUnit::~MDUnit(void) {
  mdgpu_domain *d = vars->dev;
  if (unit_mode) {
    std::lock_guard<std::mutex> lock(g_units_mu);
    for (size_t k = 0; k < g_units.size(); k++)
      if (g_units[k] == this) {
        g_units.erase(g_units.begin() + k);
        g_unit_devs.erase(g_unit_devs.begin() + k);
        break;
      }
  }
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
// The unit of this process whose send_buffer this is (NULL: another buffer).
MDUnit *
MDUnit::Sender(const std::vector<ParticleInfo> &buffer) {
  std::lock_guard<std::mutex> lock(g_units_mu);
  for (size_t k = 0; k < g_units.size(); k++)
    if (&g_units[k]->send_buffer == &buffer) return g_units[k];
  return NULL;
}
//----------------------------------------------------------------------
// Migration (MakeBufferForSendingParticle / ReceiveParticles, mdunit.cc:78-122).
// A single domain is its own neighbour on every face, so particles that leave
// through one face come back through the opposite one -- the periodic wrap,
// which the device applies at every pair-list construction: nothing is sent.
// A device unit moves the particles over the face of `dir` into its device
// send buffer; the unit MDManager hands that send_buffer to takes them from
// there, device to device.
void
MDUnit::MakeBufferForSendingParticle(const int dir) {
  send_buffer.clear();
  if (!unit_mode) return;
  vars->ToDevice();
  int n = 0;
  vars->Check(mdgpu_unit_send_particles(vars->dev, dir, &n), "send particles");
  vars->DeviceChanged();
  list_built = false;
}
//----------------------------------------------------------------------
void
MDUnit::ReceiveParticles(std::vector<ParticleInfo> &recv_buffer) {
  if (unit_mode) {
    MDUnit *src = Sender(recv_buffer);
    if (src != NULL) {
      vars->ToDevice();
      vars->Check(mdgpu_unit_receive_particles(vars->dev, src->vars->dev), "receive particles");
      vars->DeviceChanged();
      list_built = false;
      return;
    }
    if (!recv_buffer.empty()) {
      show_error("the GPU build runs the units of one process on its GPU: particles from another process");
      exit(1);
    }
    return;
  }
  for (size_t k = 0; k < recv_buffer.size(); k++) {
    ParticleInfo &pi = recv_buffer[k];
    AddParticle(pi.q, pi.p, pi.type);
  }
}
//----------------------------------------------------------------------
// MDUnit::AdjustPeriodicBoundary (mdunit.h:71): a device unit wraps its
// particles into the box and bins them in its cells, so that the indices the
// other units' ghosts take from it are final; a single domain wraps at its
// pair-list construction.
void
MDUnit::AdjustPeriodicBoundary(void) {
  if (!unit_mode) return;
  vars->ToDevice();
  vars->Check(mdgpu_unit_adjust_periodic_boundary(vars->dev), "periodic boundary");
  vars->DeviceChanged();
  list_built = false;
}
//----------------------------------------------------------------------
// Ghosts (FindBorderParticles / MakeBufferForBorderParticles /
// ReceiveBorderParticles, mdunit.cc:125-191): the device finds the border
// particles and places their images at each construction, and refreshes the
// images from the owners' current positions in one pass over all six
// directions, done here for the first direction of MDManager's sequence.  A
// device unit places, at its first direction, every ghost the six passes
// bring it, from the units that own them (all units are through
// AdjustPeriodicBoundary by then: MDManager's phases are separated).
void
MDUnit::FindBorderParticles(const int dir) {
  if (!unit_mode || dir != D_LEFT) return;
  vars->Check(mdgpu_unit_find_border_particles(vars->dev), "border particles");
  vars->DeviceChanged();
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
    show_error("the GPU build runs the units of one process on its GPU: ghost images from another process");
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
  if (unit_mode) {
    show_error("ChangeScale with several units in a process is not supported by the GPU build");
    exit(1);
  }
  vars->ToDevice();
  myrect.ChangeScale(alpha);
  vars->Check(mdgpu_change_scale(vars->dev, alpha), "change scale");
  vars->DeviceChanged();
  mesh->Built();
  plist->Init(vars, sinfo);
  list_built = true;
}
//----------------------------------------------------------------------
