// lmp.cpp — a minimal LAMMPS-shaped host harness around the device path.
//
// It provides what LAMMPS provides to the reference's styles on this path:
// the box, ghost images (Comm::borders / forward comm, done on the device),
// the re-neighbour trigger (neigh_modify every k) and velocity-Verlet NVE.
// By default the pair style runs only the path (lists, bond orders, charges)
// and NVE moves atoms with the velocities they have (zero force); positions
// still change every step, and every step recomputes the lists, bond orders
// and charges.  The harness command `forces yes` (not a LAMMPS command:
// LAMMPS always computes forces) adds the energy terms built here (bond,
// lone-pair, over-/under-coordination, valence angle, penalty, coalition,
// torsion, four-body conjugation, hydrogen bonds, van der Waals, Coulomb,
// polarization: the full ReaxFF energy of the reference) and integrates with
// their forces, as FixNVESunway does; `mass T m` sets atom->mass.
//
// The style classes keep the reference's command surface and call order:
//   PairReaxC  = pair_style reax/c/sunway  (PairReaxCSunway, pair_reaxc_sunway.h:26)
//   FixQEqReax = fix qeq/reax/sunway       (FixQEqReaxSunway, fix_qeq_reax_sunway.h:26)
// The un-suffixed names reax/c and qeq/reax resolve to them, as they do in
// LAMMPS under `-sf sunway` / `suffix sunway` (run.sh:2); reax/c/amd and
// qeq/reax/amd are accepted as aliases.
// One step: initial_integrate, neighbor (pbc + borders + list build every k
// steps, else forward comm), FixQEqReax::pre_force, PairReaxC::compute.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "swr_internal.h"

namespace {

// A style name as LAMMPS resolves it: the base name, its /sunway variant
// (the kept surface), or the /amd alias.
bool is_style(const std::string &name, const std::string &base) {
  return name == base || name == base + "/sunway" || name == base + "/amd";
}

std::vector<std::string> split(const std::string &s) {
  std::istringstream is(s);
  std::vector<std::string> t;
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}

// pair_style reax/c/sunway <NULL> [keywords]  (PairReaxCSunway::settings,
// pair_reaxc_sunway.cpp:202-290) and pair_coeff * * ffield el... (:294-362).
struct PairReaxC {
  swr_control_t ctl;
  bool have_style = false, have_coeff = false;
  int qeqflag = 1;
  std::string ffield;
  std::vector<std::string> elements;

  PairReaxC() { swr_control_default(&ctl); }

  std::string settings(const std::vector<std::string> &a) {
    if (a.empty()) return "Illegal pair_style command";
    if (a[0] != "NULL") return "control files are not supported; use NULL with keywords";
    for (size_t k = 1; k < a.size(); k += 2) {
      if (k + 1 >= a.size()) return "Illegal pair_style reax/c command";
      const std::string &key = a[k], &v = a[k + 1];
      if (key == "checkqeq") {
        if (v == "yes") qeqflag = 1; else if (v == "no") qeqflag = 0; else return "Illegal pair_style reax/c command";
      } else if (key == "enobonds" || key == "lgvdw") {
        if (v != "yes" && v != "no") return "Illegal pair_style reax/c command";
        if (key == "lgvdw" && v == "yes") return "lgvdw yes is not supported on this path";
        if (key == "enobonds") ctl.enobonds = v == "yes" ? 1 : 0;
      } else if (key == "safezone") {
        ctl.safezone = std::atof(v.c_str());
        if (ctl.safezone < 0.0) return "Illegal pair_style reax/c safezone command";
        ctl.saferzone = ctl.safezone * 1.2 + 0.2;
      } else if (key == "mincap") {
        ctl.mincap = std::atoi(v.c_str());
        if (ctl.mincap < 0) return "Illegal pair_style reax/c mincap command";
      } else if (key == "maxfar") {
        ctl.maxfar = std::atoi(v.c_str());
        if (ctl.maxfar < 0) return "Illegal pair_style reax/c maxfar command";
      } else if (key == "bond_cut") {
        ctl.bond_cut = std::atof(v.c_str());
      } else if (key == "hbond_cut") {
        ctl.hbond_cut = std::atof(v.c_str());
      } else {
        return "Illegal pair_style reax/c command";
      }
    }
    have_style = true;
    return "";
  }

  std::string coeff(const std::vector<std::string> &a) {
    if (a.size() < 4 || a[0] != "*" || a[1] != "*") return "Incorrect args for pair coefficients";
    ffield = a[2];
    elements.assign(a.begin() + 3, a.end());
    have_coeff = true;
    return "";
  }
};

// fix ID group qeq/reax/sunway nevery swa swb tol reax/c [mixed]
// (FixQEqReaxSunway ctor, fix_qeq_reax_sunway.cpp:76-99).
struct FixQEqReax {
  bool present = false;
  int nevery = 1;
  double swa = 0.0, swb = 10.0, tol = 1e-6;
  int mixed = 0;
  std::string parse(const std::vector<std::string> &a) {
    // a = [ID, group, style, nevery, swa, swb, tol, reax/c, (mixed)]
    if (a.size() < 8 || a.size() > 9) return "Illegal fix qeq/reax command";
    nevery = std::atoi(a[3].c_str());
    if (nevery <= 0) return "Illegal fix qeq/reax command";
    swa = std::atof(a[4].c_str());
    swb = std::atof(a[5].c_str());
    tol = std::atof(a[6].c_str());
    if (a[7] != "reax/c") return "only the reax/c parameter source is supported";
    if (a.size() == 9) {
      if (a[8] == "mixed") mixed = 1;
      else if (a[8] == "dual") return "Dual keyword only supported with fix qeq/reax/omp";
      else return "Illegal fix qeq/reax command";
    }
    if (swb < 0) return "Fix qeq/reax has negative upper Taper radius cutoff";
    present = true;
    return "";
  }
};

enum Stage { ST_INTEGRATE = 0, ST_NEIGH, ST_COMM, ST_QEQ_H, ST_QEQ_CG, ST_QEQ_Q, ST_PACK_RESET, ST_LISTS,
             ST_VALIDATE, ST_BO, ST_COUNT };
const char *kStageNames[ST_COUNT] = {"integrate", "neighbor", "comm",     "qeq_H", "qeq_cg",
                                     "qeq_q",     "pack_reset", "lists", "validate", "bo"};

}  // namespace

struct swr_lmp {
  PairReaxC pair;
  FixQEqReax qeq;
  double skin = 2.0;
  int every = 10;
  double dt = 0.25;
  bool forces = false;
  std::vector<double> mass_set;   // [type] from `mass` commands, 0 = force-field default
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int periodic = 1;
  int n = 0;
  std::vector<double> x0, v0;
  std::vector<double> s_hist0, t_hist0;  // optional warm start for setup (n x 5 each)
  std::vector<int> type0, tag0;
  swr_system *sys = nullptr;
  bool is_setup = false;
  long long step = 0;
  long long last_rebuild = 0;
  int iters_s = 0, iters_t = 0;
  long long total_iters_s = 0, total_iters_t = 0;
  long long cg_work_launches = 0;  // iteration launches that did work: max(iters_s, iters_t) - 1 per solve
  int profile = 0;
  double stage_ms[ST_COUNT] = {0};
  double wall_ms = 0;
  long long steps_run = 0;
  std::vector<hipEvent_t> ev;

  ~swr_lmp() {
    for (auto e : ev) (void)hipEventDestroy(e);
    if (sys) swr_destroy(sys);
  }

  void mark(int k) {
    if (profile && (size_t)k < ev.size()) (void)hipEventRecord(ev[k], sys->stream);
  }

  int neighbor_build() {
    int rc;
    const double cutmax = std::max(std::max(sys->ff.nonb_cut, pair.ctl.hbond_cut), 2 * pair.ctl.bond_cut);
    const double cutghost = cutmax + skin;
    if (periodic) {
      if ((rc = swr_nbr_pbc(sys, lo, hi))) return rc;
      if ((rc = swr_nbr_borders(sys, lo, hi, cutghost))) return rc;
      if ((rc = swr_nbr_set_grid(sys, lo, hi, cutghost, cutghost))) return rc;
    } else {
      sys->N = sys->n;
      if ((rc = swr_nbr_set_grid(sys, lo, hi, 0.0, cutghost))) return rc;
    }
    if ((rc = swr_nbr_build(sys, 0))) return rc;
    last_rebuild = step;
    return SWR_OK;
  }

  // FixQEqReaxSunway::pre_force then PairReaxCSunway::compute
  int force_stages() {
    int rc;
    if (qeq.present && step % qeq.nevery == 0) {
      mark(3);
      if ((rc = swr_qeq_init_matvec(sys))) return rc;
      mark(4);
      if ((rc = swr_qeq_cg(sys, &iters_s, &iters_t))) return rc;
      total_iters_s += iters_s;
      total_iters_t += iters_t;
      cg_work_launches += std::max(0, std::max(iters_s, iters_t) - 1);
      mark(5);
      if ((rc = swr_calculate_q(sys))) return rc;
    } else {
      mark(3); mark(4); mark(5);
    }
    mark(6);
    if ((rc = swr_write_reax_atoms_and_pack(sys))) return rc;
    if ((rc = swr_reset(sys))) return rc;
    mark(7);
    if ((rc = swr_write_reax_lists_init_forces(sys))) return rc;
    mark(8);
    if ((rc = swr_validate_lists(sys)) || (rc = swr_sort_bonds(sys))) return rc;
    mark(9);
    if ((rc = swr_bo(sys))) return rc;
    mark(10);
    if (forces) {
      // Compute_Bonded_Forces, Compute_NonBonded_Forces, Compute_Total_Force
      // (reaxc_forces_sunway.cpp), then the copy into atom->f + reverse comm
      if ((rc = swr_atom_energy(sys)) || (rc = swr_valence_angles(sys)) || (rc = swr_torsion_angles(sys)) ||
          (rc = swr_hydrogen_bonds(sys)) ||
          (rc = swr_vdw_coulomb_energy(sys)) ||
          (rc = swr_add_dbond_forces(sys)) || (rc = swr_compute_forces(sys)))
        return rc;
    }
    return SWR_OK;
  }

  // force->ftm2v in real units (LAMMPS force.cpp / update.cpp)
  static constexpr double FTM2V = 1.0 / 48.88821291 / 48.88821291;

  // potential energy of the terms computed (eng_vdwl + eng_coul)
  int potential_energy(double *pe) {
    int rc;
    if ((rc = swr_sync(sys))) return rc;
    double nb[8], be[5], va[3], to[2], hb = 0.0;
    if (!forces || swr_get(sys, "nonbonded_energy", nb) != (int64_t)sizeof(nb) ||
        swr_get(sys, "bonded_energy", be) != (int64_t)sizeof(be) || swr_get(sys, "hbond_energy", &hb) != 8 ||
        swr_get(sys, "valence_energy", va) != (int64_t)sizeof(va) ||
        swr_get(sys, "torsion_energy", to) != (int64_t)sizeof(to)) {
      swr_set_error_msg("the potential energy needs `forces yes`");
      return SWR_ERR_STATE;
    }
    *pe = nb[0] + nb[1] + be[0] + be[1] + be[2] + be[3] + be[4] + hb + va[0] + va[1] + va[2] + to[0] + to[1];
    return SWR_OK;
  }

  int setup() {
    if (!pair.have_style || !pair.have_coeff) { swr_set_error_msg("pair_style/pair_coeff missing"); return SWR_ERR_STATE; }
    if (pair.qeqflag && !qeq.present) { swr_set_error_msg("Pair reax/c requires use of fix qeq/reax"); return SWR_ERR_STATE; }
    if (n <= 0) { swr_set_error_msg("no atoms"); return SWR_ERR_STATE; }
    swr_control_t ctl = pair.ctl;
    ctl.swa = qeq.swa; ctl.swb = qeq.swb; ctl.tol = qeq.tol; ctl.qeq_mixed = qeq.mixed; ctl.skin = skin;
    std::vector<const char *> el;
    for (auto &e : pair.elements) el.push_back(e.c_str());
    int rc = swr_create(pair.ffield.c_str(), (int)el.size(), el.data(), &ctl, &sys);
    if (rc) return rc;
    if (!mass_set.empty()) {
      std::vector<double> m(sys->hp.type_mass, sys->hp.type_mass + sys->hp.ntypes + 1);
      for (size_t t = 1; t < mass_set.size() && t < m.size(); t++)
        if (mass_set[t] > 0.0) m[t] = mass_set[t];
      if ((rc = swr_set_masses(sys, sys->hp.ntypes, m.data()))) return rc;
    }
    for (int t : type0)
      if (t < 1 || t > (int)el.size()) { swr_set_error_msg("atom type without pair_coeff element"); return SWR_ERR_ARG; }
    if ((rc = swr_set_atoms(sys, n, n, tag0.data(), type0.data(), x0.data(), nullptr, nullptr, nullptr))) return rc;
    HIP_CHECK(hipMemcpyAsync(sys->v, v0.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, sys->stream));
    HIP_CHECK(hipStreamSynchronize(sys->stream));
    if (!s_hist0.empty()) {
      if ((rc = swr_put(sys, "s_hist", s_hist0.data(), (int64_t)(sizeof(double) * s_hist0.size()))) ||
          (rc = swr_put(sys, "t_hist", t_hist0.data(), (int64_t)(sizeof(double) * t_hist0.size()))))
        return rc;
    }
    ev.resize(16);
    for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
    step = 0;
    if ((rc = neighbor_build())) return rc;
    if ((rc = force_stages())) return rc;
    if ((rc = swr_sync(sys))) return rc;
    is_setup = true;
    return SWR_OK;
  }

  int one_step() {
    int rc;
    step++;
    mark(0);
    if (forces) {
      if ((rc = swr_fix_nve_initial_integrate(sys, dt, 0.5 * dt * FTM2V))) return rc;
    } else if ((rc = swr_nbr_nve(sys, dt))) {
      return rc;  // initial_integrate (+ final: F = 0)
    }
    mark(1);
    if (step - last_rebuild >= every) {
      if ((rc = neighbor_build())) return rc;
      mark(2);
    } else {
      mark(2);
      if (periodic && (rc = swr_nbr_forward_x(sys, lo, hi))) return rc;
    }
    if ((rc = force_stages())) return rc;
    if (forces && (rc = swr_fix_nve_final_integrate(sys, dt, 0.5 * dt * FTM2V))) return rc;
    return SWR_OK;
  }

  int run(int nsteps) {
    if (!is_setup) {
      int rc = setup();
      if (rc) return rc;
    }
    HIP_CHECK(hipStreamSynchronize(sys->stream));
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nsteps; s++) {
      int rc = one_step();
      if (rc) return rc;
      if (profile) {
        HIP_CHECK(hipEventSynchronize(ev[10]));
        float ms;
        for (int k = 0; k < ST_COUNT; k++) {
          if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) == hipSuccess) stage_ms[k] += ms;
        }
      }
    }
    int rc = swr_sync(sys);
    auto t1 = std::chrono::steady_clock::now();
    wall_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    steps_run += nsteps;
    return rc;
  }
};

extern "C" {

int swr_lmp_open(swr_lmp **out) {
  if (!out) return SWR_ERR_ARG;
  *out = new swr_lmp();
  return SWR_OK;
}

int swr_lmp_close(swr_lmp *L) {
  delete L;
  return SWR_OK;
}

int swr_lmp_command(swr_lmp *L, const char *cmd) {
  if (!L || !cmd) return SWR_ERR_ARG;
  std::vector<std::string> t = split(cmd);
  if (t.empty() || t[0][0] == '#') return SWR_OK;
  std::string err;
  const std::string &c = t[0];
  std::vector<std::string> a(t.begin() + 1, t.end());
  if (c == "pair_style") {
    if (a.empty() || !is_style(a[0], "reax/c")) err = "Unknown pair style";
    else err = L->pair.settings(std::vector<std::string>(a.begin() + 1, a.end()));
  } else if (c == "pair_coeff") {
    err = L->pair.coeff(a);
  } else if (c == "fix") {
    if (a.size() >= 3 && is_style(a[2], "qeq/reax")) err = L->qeq.parse(a);
    else if (a.size() >= 3 && is_style(a[2], "nve")) err = "";
    else err = "Unknown fix style";
  } else if (c == "neighbor") {
    if (a.size() != 2 || a[1] != "bin") err = "Illegal neighbor command";
    else L->skin = std::atof(a[0].c_str());
  } else if (c == "neigh_modify") {
    for (size_t k = 0; k + 1 < a.size(); k += 2)
      if (a[k] == "every") L->every = std::max(1, std::atoi(a[k + 1].c_str()));
  } else if (c == "timestep") {
    if (a.size() != 1) err = "Illegal timestep command";
    else L->dt = std::atof(a[0].c_str());
  } else if (c == "forces") {
    if (a.size() != 1 || (a[0] != "yes" && a[0] != "no")) err = "Illegal forces command";
    else if (L->is_setup) err = "forces must be set before the run";
    else L->forces = a[0] == "yes";
  } else if (c == "mass") {
    const int t = a.size() == 2 ? std::atoi(a[0].c_str()) : 0;
    const double m = a.size() == 2 ? std::atof(a[1].c_str()) : 0.0;
    if (t < 1 || t > SWR_MAX_TYPES) err = "Illegal mass command";
    else if (!(m > 0.0)) err = "Invalid mass value";
    else if (L->is_setup) err = "mass must be set before the run";
    else {
      if ((int)L->mass_set.size() <= t) L->mass_set.resize(t + 1, 0.0);
      L->mass_set[t] = m;
    }
  } else if (c == "suffix") {
    // `suffix sunway` (the -sf sunway of run.sh:2): the un-suffixed style
    // names already resolve to the sunway styles here
    if (a.size() != 1) err = "Illegal suffix command";
  } else if (c == "profile") {
    L->profile = (a.size() == 1 && a[0] == "yes") ? 1 : 0;
  } else if (c == "units" || c == "atom_style") {
    if (a.size() != 1 || (c == "units" && a[0] != "real") || (c == "atom_style" && a[0] != "charge"))
      err = "only units real / atom_style charge";
  } else {
    err = "Unknown command: " + c;
  }
  if (!err.empty()) {
    swr_set_error_msg(err);
    return SWR_ERR_ARG;
  }
  return SWR_OK;
}

int swr_lmp_create_atoms(swr_lmp *L, int n, const double *lo, const double *hi, int periodic, const double *x,
                         const double *v, const int *type, const int *tag) {
  if (!L || n <= 0 || !lo || !hi || !x || !type || !tag) return SWR_ERR_ARG;
  L->n = n;
  for (int d = 0; d < 3; d++) { L->lo[d] = lo[d]; L->hi[d] = hi[d]; }
  L->periodic = periodic;
  L->x0.assign(x, x + 3 * (size_t)n);
  L->v0.assign(3 * (size_t)n, 0.0);
  if (v) L->v0.assign(v, v + 3 * (size_t)n);
  L->type0.assign(type, type + n);
  L->tag0.assign(tag, tag + n);
  return SWR_OK;
}

int swr_lmp_set_history(swr_lmp *L, int n, const double *s_hist, const double *t_hist) {
  if (!L || !s_hist || !t_hist || n != L->n || L->is_setup) {
    swr_set_error_msg("swr_lmp_set_history: needs the created atoms' count, before setup");
    return SWR_ERR_ARG;
  }
  L->s_hist0.assign(s_hist, s_hist + 5 * (size_t)n);
  L->t_hist0.assign(t_hist, t_hist + 5 * (size_t)n);
  return SWR_OK;
}

int swr_lmp_setup(swr_lmp *L) { return L ? L->setup() : SWR_ERR_ARG; }

int swr_lmp_run(swr_lmp *L, int nsteps) { return L ? L->run(nsteps) : SWR_ERR_ARG; }

swr_system *swr_lmp_system(swr_lmp *L) { return L ? L->sys : nullptr; }

double swr_lmp_stat(swr_lmp *L, const char *name) {
  if (!L || !name) return NAN;
  std::string k = name;
  if (k == "step") return (double)L->step;
  if (k == "pe") {
    double pe = NAN;
    if (!L->sys || L->potential_energy(&pe)) return NAN;
    return pe;
  }
  if (k == "wall_ms") return L->wall_ms;
  if (k == "steps_run") return (double)L->steps_run;
  if (k == "iters_s") return L->iters_s;
  if (k == "iters_t") return L->iters_t;
  if (k == "total_iters_s") return (double)L->total_iters_s;
  if (k == "total_iters_t") return (double)L->total_iters_t;
  if (k == "cg_work_launches") return (double)L->cg_work_launches;
  if (k == "n") return L->n;
  if (k == "N") return L->sys ? L->sys->N : 0;
  for (int s = 0; s < ST_COUNT; s++)
    if (k == std::string("ms_") + kStageNames[s]) return L->stage_ms[s];
  return NAN;
}

}  // extern "C"
