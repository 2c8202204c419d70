// charles_ig_gpu: the ideal-gas driver of the MI355X path.
//
// Mirrors the reference's charles_ig.cpp main (src/charles/charles_ig.cpp:292-352):
// read the input deck, build the solver, init(), run(). Decks use the reference's
// keys for physics and time stepping (GAMMA, P_REF, ..., SGS_MODEL, TIMESTEP, NSTEPS,
// CHECK_INTERVAL); the mesh comes from this driver's MESH key (periodic box meshes):
//   MESH LATTICE <n> [L]
//   MESH JITTERED <n> <jitter> <seed> [L]
// or, as the reference, RESTART <mesh.mles> [<data.sles>] (StaticSolver.hpp:1168-1230;
// periodic mesh files). INIT TGV [Ma] sets the Taylor-Green vortex at the cell
// centroids, as an initialHook would (charles_ig.cpp:209-222); INIT K4 <k0> <Mt>
// <seed> the seeded k^4-spectrum decaying-turbulence field (cg_ic_k4). WRITE_SLES
// <interval> or WRITE_SLES INTERVAL <interval> [NAME <prefix>] writes data files
// (StaticSolver.hpp:3378-3455).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "IdealGasSolver.hpp"

namespace {

struct Deck {
  std::string text;
  std::string mesh_kind = "LATTICE";
  int n = 16;
  double L = 2.0 * M_PI, jitter = 0.3;
  unsigned long long seed = 1;
  bool tgv = false;
  double tgv_ma = 0.1;
  bool k4 = false;
  double k0 = 4.0, mt = 0.3;
  unsigned long long k4_seed = 1;
  int nsteps = -1;
  std::string restart_mles, restart_sles;
  int write_sles_interval = 0;
  std::string write_sles_prefix = "result";
};

Deck read_deck(const char* path) {
  std::ifstream f(path);
  if (!f) {
    std::fprintf(stderr, "ERROR: cannot open %s\n", path);
    throw(0);
  }
  Deck d;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key) || key[0] == '#') {
      d.text += line + "\n";
      continue;
    }
    if (key == "MESH") {
      ls >> d.mesh_kind >> d.n;
      if (d.mesh_kind == "JITTERED") ls >> d.jitter >> d.seed;
      double L;
      if (ls >> L) d.L = L;
      continue;
    }
    if (key == "INIT") {
      std::string what;
      ls >> what;
      if (what == "TGV") {
        d.tgv = true;
        double ma;
        if (ls >> ma) d.tgv_ma = ma;
      } else if (what == "K4") {
        d.k4 = true;
        ls >> d.k0 >> d.mt >> d.k4_seed;
      } else {
        std::fprintf(stderr, "ERROR: INIT %s: expected TGV [Ma] or K4 <k0> <Mt> <seed>\n", what.c_str());
        throw(0);
      }
      continue;
    }
    if (key == "RESTART") {
      ls >> d.restart_mles;
      ls >> d.restart_sles;
      continue;
    }
    if (key == "WRITE_SLES") {
      std::string tok;
      while (ls >> tok) {
        if (tok == "INTERVAL") ls >> d.write_sles_interval;
        else if (tok == "NAME") ls >> d.write_sles_prefix;
        else d.write_sles_interval = std::atoi(tok.c_str());
      }
      if (d.write_sles_interval <= 0) {
        std::fprintf(stderr, "ERROR: WRITE_SLES: expected <interval> or INTERVAL <interval> [NAME <prefix>]\n");
        throw(0);
      }
      continue;
    }
    if (key == "NSTEPS") ls >> d.nsteps;
    d.text += line + "\n";
  }
  return d;
}

class MyIdealGasSolver : public charles::IdealGasSolver {
 public:
  bool tgv = false;
  double ma = 0.1;
  bool k4 = false;
  double k0 = 4.0, mt = 0.3;
  unsigned long long k4_seed = 1;
  void initialHook() override {
    if (k4) {
      setK4(k0, mt, k4_seed);
      return;
    }
    if (!tgv) return;
    const double gamma = params().gamma;
    const double c0 = 1.0 / ma, p0 = c0 * c0 / gamma, R = params().p_ref / params().rho_ref / params().T_ref;
    for (int i = 0; i < ncv; ++i) {
      const double x = x_cv[3 * i], y = x_cv[3 * i + 1], z = x_cv[3 * i + 2];
      u[3 * i + 0] = std::sin(x) * std::cos(y) * std::cos(z);
      u[3 * i + 1] = -std::cos(x) * std::sin(y) * std::cos(z);
      u[3 * i + 2] = 0.0;
      const double pp = p0 + (1.0 / 16.0) * (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2.0);
      rho[i] = pp / (R * params().T_ref);
      const double usq = u[3 * i] * u[3 * i] + u[3 * i + 1] * u[3 * i + 1];
      rhoE[i] = pp / (gamma - 1.0) + 0.5 * rho[i] * usq;
    }
  }
};

}  // namespace

int main(int argc, char** argv) {
  const char* input = "charles.in";
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], "-i")) input = argv[i + 1];
  try {
    Deck d = read_deck(input);
    MyIdealGasSolver solver;
    solver.parseParams(d.text);
    solver.tgv = d.tgv;
    solver.ma = d.tgv_ma;
    solver.k4 = d.k4;
    solver.k0 = d.k0;
    solver.mt = d.mt;
    solver.k4_seed = d.k4_seed;
    solver.nsteps = d.nsteps;
    solver.restart_sles = d.restart_sles;
    solver.write_sles_interval = d.write_sles_interval;
    solver.write_sles_prefix = d.write_sles_prefix;
    if (!d.restart_mles.empty()) solver.initMles(d.restart_mles);
    else if (d.mesh_kind == "JITTERED") solver.initJittered(d.n, d.L, d.jitter, d.seed);
    else solver.initLattice(d.n, d.L);
    solver.init();
    std::printf(" > mesh: %d cells\n", solver.ncv);
    solver.run();
    cg_diag r = solver.report(true);
    return r.nan_flag ? 1 : 0;
  } catch (int) {
    return 1;
  }
}
