// ref_driver.cc -- drives the reference's own MDManager (kaityo256/mdacp,
// compiled unmodified from its sources) on given inputs and dumps what the
// tests compare against: owned state, canonical pair list, kicks and
// observables.  TEST INFRASTRUCTURE ONLY: it runs in the build container to
// pin the oracle restatement and to write golden fixtures; the product never
// loads it.
//
// usage: mdacp_ref <spec.cfg> <particles.bin> <out.bin>
//   spec.cfg holds the reference's own keys (SystemSizeX/Y/Z, TimeStep,
//   ControlTemperature, HeatbathType, HeatbathGamma, AimedTemperature,
//   HeatbathTau) plus driver keys read here: Steps, DumpEvery, InitV0, Seed,
//   ChangeScaleAlpha and ChangeScaleStep (MDManager::ChangeScale(alpha) called
//   after that step, mdmanager.cc:432-443, followed by a record flagged in
//   header[6]), Kick (0: the kick section of each record is left zero),
//   DumpFileStep (MDManager::SaveAsDumpFile("conf.dump") after that step, 0:
//   before the first; mdmanager.cc:474-504, the binary q, p dump).
//
// The same source also links against the MI355X facade (mdacp_amd/facade,
// `make -C oracle facade_driver` -> oracle/_ref/mdacp_gpu_driver), where it
// reads the facade's list accessors through the same calls.
//   particles.bin: int32 n, then n*(x,y,z,vx,vy,vz) doubles, then n int32 ids.
//   out.bin: a sequence of records (see write_record).
#include <omp.h>

#include <cstdint>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "mdmanager.h"
#include "observer.h"

namespace {

struct Spec {
  int steps = 0, dump_every = 1, seed = 1, scale_step = -1, kick = 1, dump_file_step = -1;
  double init_v0 = -1.0, scale_alpha = 1.0;
};

Spec read_spec(const char *path) {
  Spec s;
  std::ifstream is(path);
  std::string line;
  while (std::getline(is, line)) {
    size_t k = line.find('=');
    if (k == std::string::npos) continue;
    std::string key = line.substr(0, k), val = line.substr(k + 1);
    if (key == "Steps") s.steps = std::stoi(val);
    if (key == "DumpEvery") s.dump_every = std::stoi(val);
    if (key == "Seed") s.seed = std::stoi(val);
    if (key == "InitV0") s.init_v0 = std::stod(val);
    if (key == "ChangeScaleAlpha") s.scale_alpha = std::stod(val);
    if (key == "ChangeScaleStep") s.scale_step = std::stoi(val);
    if (key == "Kick") s.kick = std::stoi(val);
    if (key == "DumpFileStep") s.dump_file_step = std::stoi(val);
  }
  return s;
}

struct Loader : public Executor {
  std::vector<double> xv;
  std::vector<int> ids;
  void Execute(MDUnit *u) override {
    for (size_t i = 0; i < ids.size(); i++) {
      double x[D] = {xv[6 * i], xv[6 * i + 1], xv[6 * i + 2], 0.0};
      double v[D] = {xv[6 * i + 3], xv[6 * i + 4], xv[6 * i + 5], 0.0};
      u->AddParticle(x, v, ids[i]);
    }
  }
};

struct MeshGrab : public DoubleObserver {
  MeshList *mesh = nullptr;
  double Observe(Variables *, MeshList *m) override {
    mesh = m;
    return 0.0;
  }
};

struct Key3 {
  uint64_t a, b, c;
  bool operator==(const Key3 &o) const { return a == o.a && b == o.b && c == o.c; }
};
struct Key3Hash {
  size_t operator()(const Key3 &k) const { return k.a * 1000003u ^ k.b * 10007u ^ k.c; }
};
uint64_t bits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}

// Owner and image shift of every ghost.  Ghosts are refreshed after the first
// half drift and owned particles drift once more, so a ghost sits within
// dt*|p|/2 of (owner + shift*L); owners are found by a grid search within 0.05.
void map_ghosts(Variables *vars, const double *L, std::vector<int> &src, std::vector<int> &code) {
  const int pn = vars->GetParticleNumber(), tn = vars->GetTotalParticleNumber();
  const double h = 0.5;
  auto cell = [&](double x) { return (long)std::floor(x / h); };
  std::unordered_map<Key3, std::vector<int>, Key3Hash> grid;
  for (int i = 0; i < pn; i++)
    grid[Key3{(uint64_t)cell(vars->q[i][0]), (uint64_t)cell(vars->q[i][1]),
              (uint64_t)cell(vars->q[i][2])}].push_back(i);
  src.assign(tn - pn, -1);
  code.assign(tn - pn, -1);
  for (int g = pn; g < tn; g++) {
    int found = 0;
    for (int c = 0; c < 27; c++) {
      if (c == 13) continue;
      int sh[3] = {c % 3 - 1, (c / 3) % 3 - 1, c / 9 - 1};
      double y[3];
      for (int d = 0; d < 3; d++) y[d] = vars->q[g][d] - sh[d] * L[d];
      for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
          for (int e = -1; e <= 1; e++) {
            auto it = grid.find(Key3{(uint64_t)(cell(y[0]) + a), (uint64_t)(cell(y[1]) + b),
                                     (uint64_t)(cell(y[2]) + e)});
            if (it == grid.end()) continue;
            for (int i : it->second) {
              if (std::fabs(vars->q[i][0] - y[0]) < 0.05 && std::fabs(vars->q[i][1] - y[1]) < 0.05 &&
                  std::fabs(vars->q[i][2] - y[2]) < 0.05) {
                src[g - pn] = vars->type[i];
                code[g - pn] = c;
                found++;
              }
            }
          }
    }
    if (found != 1) {
      fprintf(stderr, "ref_driver: ghost %d matches %d owners\n", g, found);
      exit(3);
    }
  }
}

template <class T>
void put(std::ofstream &o, const T *p, size_t n) {
  o.write(reinterpret_cast<const char *>(p), sizeof(T) * n);
}

// record: int64 header[8] = {step, pn, tn, npairs, rebuilt, nconstructions, scaled, 0}
//         (scaled = 1: the record written right after MDManager::ChangeScale)
//         double obs[8] = {ke_raw, pe_raw, vir_raw, T, P, E, sim_time, 0}
//         q[pn*3], p[pn*3], type[pn] (int32), ghost_src[tn-pn], ghost_code[tn-pn],
//         ghost_q[(tn-pn)*3], pairs[npairs*3] (int32), pairs_csr[npairs*3] (int32),
//         kick[pn*3]
// pairs come from the COO accessors (GetKeyParticles/GetPartnerParticles,
// meshlist.h:46-48), pairs_csr from the CSR ones (GetPartnerNumber,
// GetKeyPointer, GetSortedList, meshlist.h:49-53), each as canonical triples.
void write_record(std::ofstream &o, MDManager *mdm, int step, int rebuilt, double SL, bool kick_on,
                  int scaled = 0) {
  MDUnit *u = mdm->GetMDUnit(0);
  Variables *vars = u->GetVariables();
  MeshGrab grab;
  u->ObserveDouble(&grab);
  MeshList *mesh = grab.mesh;
  const int pn = vars->GetParticleNumber(), tn = vars->GetTotalParticleNumber();
  const int np = mesh->GetPairNumber();
  std::vector<int> gsrc, gcode;
  map_ghosts(vars, mdm->GetSystemSize(), gsrc, gcode);
  KineticEnergyObserver ke;
  PotentialEnergyObserver pe(nullptr);
  VirialObserver vir;
  int64_t hdr[8] = {step, pn, tn, np, rebuilt, mesh->GetNumberOfConstructions(), scaled, 0};
  double obs[8] = {u->ObserveDouble(&ke), u->ObserveDouble(&pe), u->ObserveDouble(&vir),
                   mdm->Temperature(), mdm->Pressure(), mdm->TotalEnergy(),
                   mdm->GetSimulationTime(), 0.0};
  put(o, hdr, 8);
  put(o, obs, 8);
  std::vector<double> q(3 * pn), p(3 * pn);
  for (int i = 0; i < pn; i++)
    for (int d = 0; d < 3; d++) {
      q[3 * i + d] = vars->q[i][d];
      p[3 * i + d] = vars->p[i][d];
    }
  put(o, q.data(), q.size());
  put(o, p.data(), p.size());
  put(o, vars->type, pn);
  put(o, gsrc.data(), gsrc.size());
  put(o, gcode.data(), gcode.size());
  std::vector<double> gq(3 * (tn - pn));
  for (int g = pn; g < tn; g++)
    for (int d = 0; d < 3; d++) gq[3 * (g - pn) + d] = vars->q[g][d];
  put(o, gq.data(), gq.size());
  // canonical pairs: (owned key id, partner owner id, partner image shift)
  auto canon = [&](int i, int j, int32_t *out) {
    int a = vars->type[i], b, sh = 13;
    if (j < pn) {
      b = vars->type[j];
      if (b < a) std::swap(a, b);
    } else {
      b = gsrc[j - pn];
      sh = gcode[j - pn];
    }
    out[0] = a;
    out[1] = b;
    out[2] = sh;
  };
  std::vector<int32_t> pr(3 * (size_t)np);
  int *key = mesh->GetKeyParticles(), *par = mesh->GetPartnerParticles();
  for (int k = 0; k < np; k++) canon(key[k], par[k], &pr[3 * (size_t)k]);
  put(o, pr.data(), pr.size());
  // the same list through the CSR accessors, row by row
  std::vector<int32_t> pc(3 * (size_t)np, -1);
  int *sorted = mesh->GetSortedList();
  size_t m = 0;
  for (int i = 0; i < pn; i++) {
    const int kp = mesh->GetKeyPointer(i), nn = mesh->GetPartnerNumber(i);
    for (int k = 0; k < nn && m < (size_t)np; k++, m++) canon(i, sorted[kp + k], &pc[3 * m]);
  }
  put(o, pc.data(), pc.size());
  // kick from zero momenta through the reference's own force routine
  std::vector<double> kick(3 * pn, 0.0);
  if (kick_on) {
    std::vector<double> save(4 * (size_t)tn);
    std::memcpy(save.data(), vars->p, sizeof(double) * 4 * tn);
    std::memset(vars->p, 0, sizeof(double) * 4 * tn);
    u->CalculateForce();
    for (int i = 0; i < pn; i++)
      for (int d = 0; d < 3; d++) kick[3 * i + d] = vars->p[i][d];
    std::memcpy(vars->p, save.data(), sizeof(double) * 4 * tn);
  }
  put(o, kick.data(), kick.size());
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s spec.cfg particles.bin out.bin\n", argv[0]);
    return 2;
  }
  omp_set_num_threads(1);
  Spec spec = read_spec(argv[1]);
  Loader loader;
  {
    std::ifstream is(argv[2], std::ios::binary);
    int32_t n = 0;
    is.read(reinterpret_cast<char *>(&n), 4);
    loader.xv.resize(6 * (size_t)n);
    loader.ids.resize(n);
    is.read(reinterpret_cast<char *>(loader.xv.data()), sizeof(double) * 6 * n);
    is.read(reinterpret_cast<char *>(loader.ids.data()), sizeof(int32_t) * n);
  }
  int margc = 2;
  char *margv_store[3] = {argv[0], argv[1], nullptr};
  char **margv = margv_store;
  MDManager *mdm = new MDManager(margc, margv);
  if (!mdm->IsValid()) return 4;
  const double SL = CUTOFF_LENGTH + 0.3;  // SimulationInfo::SearchLength
  mdm->ExecuteAll(&loader);
  if (spec.init_v0 >= 0.0) mdm->SetInitialVelocity(spec.init_v0);
  mdm->SetSeed(spec.seed);
  std::ofstream out(argv[3], std::ios::binary);
  mdm->MakePairList();
  const bool kick_on = spec.kick != 0;
  write_record(out, mdm, 0, 1, SL, kick_on);
  if (spec.dump_file_step == 0) mdm->SaveAsDumpFile("conf.dump");
  MeshGrab grab;
  for (int step = 1; step <= spec.steps; step++) {
    if (step - 1 == spec.scale_step) {
      mdm->ChangeScale(spec.scale_alpha);
      write_record(out, mdm, step - 1, 1, SL, kick_on, 1);
    }
    mdm->GetMDUnit(0)->ObserveDouble(&grab);
    const int before = grab.mesh->GetNumberOfConstructions();
    mdm->Calculate();
    const int rebuilt = grab.mesh->GetNumberOfConstructions() != before;
    if (step == spec.dump_file_step) mdm->SaveAsDumpFile("conf.dump");
    if (step % spec.dump_every == 0) write_record(out, mdm, step, rebuilt, SL, kick_on);
    else {
      int64_t hdr[8] = {step, -1, 0, 0, rebuilt, 0, 0, 0};
      put(out, hdr, 8);
    }
  }
  out.close();
  delete mdm;
  return 0;
}
