// reaxff_oracle.cpp — CPU FP64 oracle for the nbr + BO + QEq path.
//
// TEST INFRASTRUCTURE ONLY (see reaxff_oracle.h).  Every stage below restates
// the reference's live algorithm; the citation on each function names the
// reference lines it follows.  Arithmetic is plain IEEE double with libm
// exp/pow/log/sqrt, built with -ffp-contract=off, and expressions keep the
// reference's operand order so that results can be compared tightly.
//
// Deliberate departures from the reference, each decided explicitly:
//  * Lists are stored CSR-exact (index[i] = prefix sum of counts) instead of
//    the reference's i*maxfar stride; consumers only read index/end_index.
//  * The Verlet list is the mathematically complete set {j != i : r^2 <= cut^2}
//    over all N atoms; the reference's `ibin+stencil[k] > 0` guard
//    (npair_full_bin_ghost_sw5.c:151) would drop bin 0 — not reproduced.
//  * Bond-order exponentials follow the host form (libm, no clamping), as in
//    BOp_single (reaxc_forces_sunway.cpp:695-717), not the CPE's p_expd clamp
//    (reaxc_forces_sw64.c:1160-1193).
//  * BO section 2 follows the host branch (reaxc_bond_orders_sw64.c:203-219),
//    which resets C1dbo/C2dbo in the no-correction branch; section 3 follows
//    the host branch (:451-481), which writes nlp_temp.
//  * Capacity overflows return an error (the reference writes past them).

#include "reaxff_oracle.h"
#include "../include/swreax_layout.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double EV_TO_KCAL_PER_MOL = 14.4;   // fix_qeq_reax_sw64.c:8
constexpr int    MIN_BONDS  = 25;             // reaxc_defs_sunway.h:97
constexpr int    MIN_HBONDS = 25;             // reaxc_defs_sunway.h:98
constexpr double SAFE_ZONE  = 1.2;            // reaxc_defs_sunway.h:102
constexpr double SAFER_ZONE = 1.4;            // reaxc_defs_sunway.h:103
constexpr int    MIN_CAP    = 50;             // reaxc_defs_sunway.h:93

inline double SQR(double x) { return x * x; }

struct SBP {
  std::string name;
  double r_s, valency, mass, r_vdw, epsilon, gamma, r_pi, valency_e, nlp_opt;
  double alpha, gamma_w, valency_boc, p_ovun5, chi, eta;
  int p_hbond;
  double r_pi_pi, p_lp2, b_o_131, b_o_132, b_o_133;
  double p_ovun2, p_val3, valency_val, p_val5, rcore2, ecore2, acore2;
};

struct HBP {
  double r0_hb, p_hb1, p_hb2, p_hb3;
};

// three_body_parameters (reaxc_ctypes_sunway.h); REAX_MAX_3BODY_PARAM = 5
struct THBP {
  double theta_00, p_val1, p_val2, p_coa1, p_val7, p_pen1, p_val4;
};
constexpr int MAX_3BODY_PARAM = 5;

// four_body_header cnt + four_body_parameters prm[0]
struct FBP {
  int cnt;
  double V1, V2, V3, p_tor1, p_cot1;
};

struct TBP {
  double p_bo1, p_bo2, p_bo3, p_bo4, p_bo5, p_bo6;
  double r_s, r_p, r_pp;
  double p_boc3, p_boc4, p_boc5;
  double p_be1, p_be2, De_s, De_p, De_pp, p_ovun1;
  double D, alpha, r_vdW, gamma_w, rcore, ecore, acore, lgcij, lgre, gamma;
  double v13cor, ovc;
};

// Tokenize with the reference's separators (reaxc_tool_box_sunway.cpp:42-57).
std::vector<std::string> tokenize(const std::string &line) {
  std::vector<std::string> out;
  const char *sep = "\t \n\r\f!=";
  std::string cur;
  for (char c : line) {
    if (std::strchr(sep, c) && c != '\0') {
      if (!cur.empty()) { out.push_back(cur); cur.clear(); }
    } else {
      cur.push_back(c);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

double tokd(const std::vector<std::string> &t, size_t k) {
  return k < t.size() ? std::atof(t[k].c_str()) : 0.0;
}

}  // namespace

struct orc_sys {
  // ---- force field + control
  std::vector<double> gp;
  std::vector<SBP> sbp;
  std::vector<std::vector<TBP>> tbp;
  std::vector<HBP> hbp;          // [nt][nt][nt], r0_hb < 0: no hydrogen bond
  std::vector<int> thb_cnt;      // [nt][nt][nt]
  std::vector<THBP> thbp;        // [nt][nt][nt][MAX_3BODY_PARAM], zero-filled
  double thb_cut = 0.001, thb_cutsq = 0.00001;   // pair_reaxc_sunway.cpp:218-219
  std::vector<FBP> fbp;          // [nt][nt][nt][nt]
  double bo_cut = 0, nonb_low = 0, nonb_cut = 0;
  int vdw_type = 0;
  double bond_cut = 5.0, hbond_cut = 7.5;
  double swa = 0, swb = 10, tol = 1e-6;
  double safezone = SAFE_ZONE, saferzone = SAFER_ZONE;
  int mincap = MIN_CAP;
  int nelem = 0;                 // LAMMPS types 1..nelem
  std::vector<int> map;          // map[t] = element index
  std::vector<double> chi, eta, gamma;  // per LAMMPS type (pair->extract)
  std::vector<double> shld;      // (nelem+1)^2
  double Tap[8] = {0};

  // ---- atoms
  int n = 0, N = 0;
  std::vector<int> tag, ltype, owner;
  std::vector<double> x, q;
  std::vector<swr_atom_pack_t> packed;

  // ---- Verlet list (a1)
  std::vector<int64_t> voff;
  std::vector<int> vlist;

  // ---- far list (a3)
  std::vector<int64_t> far_index, far_end;   // int64 offsets (oracle-internal)
  std::vector<swr_far_neighbor_t> far;

  // ---- capacities + bonds/hbonds (a4-a7)
  std::vector<int> num_bonds, num_hbonds, Hindex;
  int numH = 0;
  bool estimated = false;
  std::vector<int> bindex, bend, hindex, hend;
  std::vector<swr_bond_t> bonds;
  std::vector<swr_bond_order_t> bo_data;
  std::vector<double> BO, Cdbo, Cdbopi, Cdbopi2;
  std::vector<double> BOpi;   // [slot][2]
  std::vector<swr_hbond_t> hbonds;
  std::vector<int64_t> hbond_far;   // far-list slot each hbond points to
  int num_bonds_total = 0, num_hbonds_total = 0;

  // ---- workspace (a8-a10)
  std::vector<double> bo_dboc;       // [N][2]
  std::vector<double> dDeltap_self;  // [N][3]
  std::vector<int> bond_mark;
  std::vector<double> Deltap, Deltap_boc, Delta, Delta_e, Delta_val, vlpex, nlp,
      nlp_temp, Clp, Delta_lp, dDelta_lp, Delta_lp_temp, dDelta_lp_temp;

  // ---- non-bonded (f2): workspace fCdDelta [N][4] (gradient xyz + CdDelta)
  std::vector<double> fCdDelta;

  // ---- QEq (a11-a16)
  std::vector<int> Hfirst, Hnum, Hj;
  std::vector<double> Hval;
  std::vector<double> Hdia_inv, b_s, b_t, s, t, s_hist, t_hist;  // hist [n][5]
  std::vector<double> p, qv, r, d, u, w, m, v, z, ss;
  int matvecs_s = 0, matvecs_t = 0;
  int cg_imax = 200;   // CG_v2's imax (fix_qeq_reax_sunway.cpp:998)
  bool qeq_storage_init = false;

  std::string err;
};

namespace {

// ------------------------------------------------------------------------
// Force field: Read_Force_Field, reaxc_ffield_sunway.cpp:35-386 (the part
// this path consumes: globals, atoms, bonds, off-diagonal overrides, and the
// hydrogen-bond block :654-686 after skipping the angle and torsion blocks).
// ------------------------------------------------------------------------
bool read_ffield(orc_sys *S, const char *path) {
  FILE *fp = std::fopen(path, "r");
  if (!fp) { S->err = std::string("cannot open ") + path; return false; }
  std::vector<std::string> lines;
  char buf[4096];
  while (std::fgets(buf, sizeof(buf), fp)) lines.push_back(buf);
  std::fclose(fp);
  size_t li = 0;
  auto next = [&]() -> std::vector<std::string> {
    if (li >= lines.size()) return {};
    return tokenize(lines[li++]);
  };
  next();                                   // header comment
  auto t = next();
  int ng = t.empty() ? 0 : std::atoi(t[0].c_str());
  if (ng < 1) { S->err = "no globals"; return false; }
  S->gp.resize(ng);
  for (int i = 0; i < ng; i++) { t = next(); S->gp[i] = tokd(t, 0); }
  S->bo_cut = 0.01 * S->gp[29];             // :86
  S->nonb_low = S->gp[11];                  // :87
  S->nonb_cut = S->gp[12];                  // :88
  t = next();
  int nt = std::atoi(t[0].c_str());
  next(); next(); next();
  S->sbp.assign(nt, SBP());
  S->tbp.assign(nt, std::vector<TBP>(nt));
  for (auto &row : S->tbp) for (auto &e : row) std::memset(&e, 0, sizeof(TBP));
  for (int i = 0; i < nt; i++) {
    SBP &a = S->sbp[i];
    t = next();                               // :191-205
    a.name = t[0];
    for (auto &c : a.name) c = (char)std::toupper((unsigned char)c);
    a.r_s = tokd(t, 1); a.valency = tokd(t, 2); a.mass = tokd(t, 3);
    a.r_vdw = tokd(t, 4); a.epsilon = tokd(t, 5); a.gamma = tokd(t, 6);
    a.r_pi = tokd(t, 7); a.valency_e = tokd(t, 8);
    a.nlp_opt = 0.5 * (a.valency_e - a.valency);
    t = next();                               // :207-218
    a.alpha = tokd(t, 0); a.gamma_w = tokd(t, 1); a.valency_boc = tokd(t, 2);
    a.p_ovun5 = tokd(t, 3); a.chi = tokd(t, 5); a.eta = 2.0 * tokd(t, 6);
    a.p_hbond = (int)tokd(t, 7);
    t = next();                               // :220-231
    a.r_pi_pi = tokd(t, 0); a.p_lp2 = tokd(t, 1); a.b_o_131 = tokd(t, 3);
    a.b_o_132 = tokd(t, 4); a.b_o_133 = tokd(t, 5);
    t = next();                               // :233-251
    a.p_ovun2 = tokd(t, 0); a.p_val3 = tokd(t, 1); a.valency_val = tokd(t, 3);
    a.p_val5 = tokd(t, 4); a.rcore2 = tokd(t, 5); a.ecore2 = tokd(t, 6);
    a.acore2 = tokd(t, 7);
  }
  // vdWaals method (:241-292): 1 shielding only, 2 inner wall only, 3 both;
  // an element disagreeing with an earlier one keeps the earlier setting,
  // one with neither is an error
  S->vdw_type = 0;
  for (int i = 0; i < nt; i++) {
    const SBP &a = S->sbp[i];
    const bool wall = a.rcore2 > 0.01 && a.acore2 > 0.01, shield = a.gamma_w > 0.5;
    if (!wall && !shield) { S->err = "No shielding or inner-wall set for element " + a.name; return false; }
    const int want = wall ? (shield ? 3 : 2) : 1;
    if (S->vdw_type == 0) S->vdw_type = want;
  }
  for (int i = 0; i < nt; i++)              // :327-335
    if (S->sbp[i].mass < 21 && S->sbp[i].valency_val != S->sbp[i].valency_boc)
      S->sbp[i].valency_val = S->sbp[i].valency_boc;
  t = next();                               // :338-342
  int nb = std::atoi(t[0].c_str());
  next();
  for (int k = 0; k < nb; k++) {            // :344-398
    t = next();
    int j = std::atoi(t[0].c_str()) - 1, kk = std::atoi(t[1].c_str()) - 1;
    auto t2 = next();
    if (j < nt && kk < nt) {
      for (int sw = 0; sw < 2; sw++) {
        TBP &b = sw ? S->tbp[kk][j] : S->tbp[j][kk];
        b.De_s = tokd(t, 2); b.De_p = tokd(t, 3); b.De_pp = tokd(t, 4);
        b.p_be1 = tokd(t, 5); b.p_bo5 = tokd(t, 6); b.v13cor = tokd(t, 7);
        b.p_bo6 = tokd(t, 8); b.p_ovun1 = tokd(t, 9);
        b.p_be2 = tokd(t2, 0); b.p_bo3 = tokd(t2, 1); b.p_bo4 = tokd(t2, 2);
        b.p_bo1 = tokd(t2, 4); b.p_bo2 = tokd(t2, 5); b.ovc = tokd(t2, 6);
      }
    }
  }
  for (int i = 0; i < nt; i++)              // :402-484
    for (int j = i; j < nt; j++) {
      const SBP &a = S->sbp[i], &c = S->sbp[j];
      for (int sw = 0; sw < 2; sw++) {
        TBP &b = sw ? S->tbp[j][i] : S->tbp[i][j];
        const SBP &p1 = sw ? c : a, &p2 = sw ? a : c;
        b.r_s = 0.5 * (p1.r_s + p2.r_s);
        b.r_p = 0.5 * (p1.r_pi + p2.r_pi);
        b.r_pp = 0.5 * (p1.r_pi_pi + p2.r_pi_pi);
        b.p_boc3 = std::sqrt(p1.b_o_132 * p2.b_o_132);
        b.p_boc4 = std::sqrt(p1.b_o_131 * p2.b_o_131);
        b.p_boc5 = std::sqrt(p1.b_o_133 * p2.b_o_133);
        b.D = std::sqrt(p1.epsilon * p2.epsilon);
        b.alpha = std::sqrt(p1.alpha * p2.alpha);
        b.r_vdW = 2.0 * std::sqrt(p1.r_vdw * p2.r_vdw);
        b.gamma_w = std::sqrt(p1.gamma_w * p2.gamma_w);
        b.gamma = std::pow(p1.gamma * p2.gamma, -1.5);
        b.rcore = std::sqrt(p1.rcore2 * p2.rcore2);   // :445-452
        b.ecore = std::sqrt(p1.ecore2 * p2.ecore2);
        b.acore = std::sqrt(p1.acore2 * p2.acore2);
      }
    }
  t = next();                               // :488-545
  int no = t.empty() ? 0 : std::atoi(t[0].c_str());
  for (int k = 0; k < no; k++) {
    t = next();
    int j = std::atoi(t[0].c_str()) - 1, kk = std::atoi(t[1].c_str()) - 1;
    if (j < nt && kk < nt) {
      double val;
      if ((val = tokd(t, 2)) > 0.0) S->tbp[j][kk].D = S->tbp[kk][j].D = val;
      if ((val = tokd(t, 3)) > 0.0) S->tbp[j][kk].r_vdW = S->tbp[kk][j].r_vdW = 2 * val;
      if ((val = tokd(t, 4)) > 0.0) S->tbp[j][kk].alpha = S->tbp[kk][j].alpha = val;
      if ((val = tokd(t, 5)) > 0.0) S->tbp[j][kk].r_s = S->tbp[kk][j].r_s = val;
      if ((val = tokd(t, 6)) > 0.0) S->tbp[j][kk].r_p = S->tbp[kk][j].r_p = val;
      if ((val = tokd(t, 7)) > 0.0) S->tbp[j][kk].r_pp = S->tbp[kk][j].r_pp = val;
    }
  }
  // valence-angle block (:520-572): line "j k m" writes prm[cnt(j,k,m)] of
  // both (j,k,m) and (m,k,j) and counts both up; then the torsion block,
  // skipped by its count; then the hydrogen-bond block: r0_hb = -1 unless
  // listed, then j k m r0 p1 p2 p3
  S->hbp.assign((size_t)nt * nt * nt, HBP{-1.0, 0.0, 0.0, 0.0});
  S->thb_cnt.assign((size_t)nt * nt * nt, 0);
  S->thbp.assign((size_t)nt * nt * nt * MAX_3BODY_PARAM, THBP{0, 0, 0, 0, 0, 0, 0});
  {
    t = next();
    const int l = t.empty() ? 0 : std::atoi(t[0].c_str());
    for (int k = 0; k < l; k++) {
      t = next();
      if (t.size() < 10) continue;
      const int a = std::atoi(t[0].c_str()) - 1, b = std::atoi(t[1].c_str()) - 1, c = std::atoi(t[2].c_str()) - 1;
      if (a < 0 || b < 0 || c < 0 || a >= nt || b >= nt || c >= nt) continue;
      const size_t abc = ((size_t)a * nt + b) * nt + c, cba = ((size_t)c * nt + b) * nt + a;
      const int cnt = S->thb_cnt[abc];
      if (cnt >= MAX_3BODY_PARAM || S->thb_cnt[cba] >= MAX_3BODY_PARAM) continue;
      S->thb_cnt[abc]++;
      S->thb_cnt[cba]++;
      const THBP v{tokd(t, 3), tokd(t, 4), tokd(t, 5), tokd(t, 6), tokd(t, 7), tokd(t, 8), tokd(t, 9)};
      S->thbp[abc * MAX_3BODY_PARAM + cnt] = v;
      S->thbp[cba * MAX_3BODY_PARAM + cnt] = v;
    }
    // torsion block (:574-652): explicit quadruples set and flag both
    // orientations; "0 j k 0" sets cnt of every (p,j,k,o), (o,k,j,p) and the
    // values where not flagged
    S->fbp.assign((size_t)nt * nt * nt * nt, FBP{0, 0, 0, 0, 0, 0});
    std::vector<char> tor_flag((size_t)nt * nt * nt * nt, 0);
    auto Q = [&](int a, int b, int c, int d) { return (((size_t)a * nt + b) * nt + c) * nt + d; };
    t = next();
    const int l2 = t.empty() ? 0 : std::atoi(t[0].c_str());
    for (int k = 0; k < l2; k++) {
      t = next();
      if (t.size() < 9) continue;
      const int j = std::atoi(t[0].c_str()) - 1, kk = std::atoi(t[1].c_str()) - 1, m = std::atoi(t[2].c_str()) - 1,
                nn = std::atoi(t[3].c_str()) - 1;
      const FBP v{1, tokd(t, 4), tokd(t, 5), tokd(t, 6), tokd(t, 7), tokd(t, 8)};
      if (j >= 0 && nn >= 0) {
        if (j < nt && kk < nt && m < nt && nn < nt) {
          tor_flag[Q(j, kk, m, nn)] = tor_flag[Q(nn, m, kk, j)] = 1;
          S->fbp[Q(j, kk, m, nn)] = v;
          S->fbp[Q(nn, m, kk, j)] = v;
        }
      } else if (kk < nt && m < nt && kk >= 0 && m >= 0) {
        for (int p = 0; p < nt; p++)
          for (int o = 0; o < nt; o++) {
            S->fbp[Q(p, kk, m, o)].cnt = 1;
            S->fbp[Q(o, m, kk, p)].cnt = 1;
            if (tor_flag[Q(p, kk, m, o)] == 0) S->fbp[Q(p, kk, m, o)] = v;
            if (tor_flag[Q(o, m, kk, p)] == 0) S->fbp[Q(o, m, kk, p)] = v;
          }
      }
    }
  }
  t = next();
  const int nh = t.empty() ? 0 : std::atoi(t[0].c_str());
  for (int k = 0; k < nh; k++) {
    t = next();
    if (t.size() < 7) break;
    const int a = std::atoi(t[0].c_str()) - 1, b = std::atoi(t[1].c_str()) - 1, c = std::atoi(t[2].c_str()) - 1;
    if (a < nt && c < nt && a >= 0 && b >= 0 && b < nt && c >= 0)
      S->hbp[((size_t)a * nt + b) * nt + c] = HBP{tokd(t, 3), tokd(t, 4), tokd(t, 5), tokd(t, 6)};
  }
  return true;
}

// init_taper, fix_qeq_reax_sunway.cpp:458-484.
void init_taper(orc_sys *S) {
  double swa = S->swa, swb = S->swb;
  double d7 = std::pow(swb - swa, 7);
  double swa2 = SQR(swa), swa3 = swa * swa * swa;
  double swb2 = SQR(swb), swb3 = swb * swb * swb;
  S->Tap[7] = 20.0 / d7;
  S->Tap[6] = -70.0 * (swa + swb) / d7;
  S->Tap[5] = 84.0 * (swa2 + 3.0 * swa * swb + swb2) / d7;
  S->Tap[4] = -35.0 * (swa3 + 9.0 * swa2 * swb + 9.0 * swa * swb2 + swb3) / d7;
  S->Tap[3] = 140.0 * (swa3 * swb + 3.0 * swa2 * swb2 + swa * swb3) / d7;
  S->Tap[2] = -210.0 * (swa3 * swb2 + swa2 * swb3) / d7;
  S->Tap[1] = 140.0 * swa3 * swb3 / d7;
  S->Tap[0] = (-35.0 * swa3 * swb2 * swb2 + 21.0 * swa2 * swb3 * swb2 +
               7.0 * swa * swb3 * swb3 + swb3 * swb3 * swb) / d7;
}

// Forward comm of a per-atom vector: ghost g takes its owner's value
// (pack/unpack_forward_comm, fix_qeq_reax_sunway.cpp:1759-1813).
void forward(orc_sys *S, std::vector<double> &vec) {
  for (int g = S->n; g < S->N; g++) vec[g] = vec[S->owner[g]];
}

template <class T>
int64_t cp(const std::vector<T> &v, void *out) {
  if (out && !v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(T));
  return (int64_t)(v.size() * sizeof(T));
}

}  // namespace

extern "C" {

static int g_threads = 1;
void orc_set_threads(int nthreads) {
  g_threads = nthreads < 1 ? 1 : nthreads;
#ifdef _OPENMP
  omp_set_num_threads(g_threads);
#endif
}

orc_sys *orc_create(const char *ffield_path, int nelem, const char *const *elements,
                    double bond_cut, double hbond_cut, double swa, double swb, double tol,
                    char *errbuf, int errlen) {
  orc_sys *S = new orc_sys();
  auto fail = [&](const std::string &m) -> orc_sys * {
    if (errbuf && errlen > 0) std::snprintf(errbuf, errlen, "%s", m.c_str());
    delete S;
    return nullptr;
  };
  if (!read_ffield(S, ffield_path)) return fail(S->err);
  S->bond_cut = bond_cut;
  S->hbond_cut = hbond_cut;
  S->swa = swa; S->swb = swb; S->tol = tol;
  S->nelem = nelem;
  S->map.assign(nelem + 1, -1);
  for (int t = 1; t <= nelem; t++) {       // pair_coeff map, pair_reaxc_sunway.cpp:322-345
    std::string e = elements[t - 1];
    for (auto &c : e) c = (char)std::toupper((unsigned char)c);
    for (size_t k = 0; k < S->sbp.size(); k++)
      if (S->sbp[k].name == e) S->map[t] = (int)k;
    if (S->map[t] < 0) return fail("Non-existent ReaxFF type " + e);
  }
  // pair->extract("chi"/"eta"/"gamma"), pair_reaxc_sunway.cpp:1106-1128
  S->chi.assign(nelem + 1, 0.0); S->eta.assign(nelem + 1, 0.0); S->gamma.assign(nelem + 1, 0.0);
  for (int t = 1; t <= nelem; t++) {
    S->chi[t] = S->sbp[S->map[t]].chi;
    S->eta[t] = S->sbp[S->map[t]].eta;
    S->gamma[t] = S->sbp[S->map[t]].gamma;
  }
  // init_shielding, fix_qeq_reax_sunway.cpp:440-454
  S->shld.assign((nelem + 1) * (nelem + 1), 0.0);
  for (int i = 1; i <= nelem; i++)
    for (int j = 1; j <= nelem; j++)
      S->shld[i * (nelem + 1) + j] = std::pow(S->gamma[i] * S->gamma[j], -1.5);
  init_taper(S);
  return S;
}

void orc_destroy(orc_sys *s) { delete s; }

double orc_gp(orc_sys *s, int k) { return (k >= 0 && k < (int)s->gp.size()) ? s->gp[k] : NAN; }

// Field order: r_s valency mass gamma r_pi valency_e nlp_opt valency_boc chi eta
//              p_hbond r_pi_pi b_o_131 b_o_132 b_o_133 valency_val
int orc_sbp(orc_sys *s, int e, double *o) {
  if (e < 0 || e >= (int)s->sbp.size()) return -1;
  const SBP &a = s->sbp[e];
  double v[16] = {a.r_s, a.valency, a.mass, a.gamma, a.r_pi, a.valency_e, a.nlp_opt,
                  a.valency_boc, a.chi, a.eta, (double)a.p_hbond, a.r_pi_pi,
                  a.b_o_131, a.b_o_132, a.b_o_133, a.valency_val};
  std::memcpy(o, v, sizeof(v));
  return 0;
}

// Field order: p_bo1..6 r_s r_p r_pp p_boc3 p_boc4 p_boc5 v13cor ovc
int orc_tbp(orc_sys *s, int i, int j, double *o) {
  int nt = (int)s->sbp.size();
  if (i < 0 || j < 0 || i >= nt || j >= nt) return -1;
  const TBP &b = s->tbp[i][j];
  double v[14] = {b.p_bo1, b.p_bo2, b.p_bo3, b.p_bo4, b.p_bo5, b.p_bo6, b.r_s, b.r_p,
                  b.r_pp, b.p_boc3, b.p_boc4, b.p_boc5, b.v13cor, b.ovc};
  std::memcpy(o, v, sizeof(v));
  return 0;
}

double orc_control(orc_sys *s, const char *key) {
  std::string k = key;
  if (k == "bo_cut") return s->bo_cut;
  if (k == "nonb_cut") return s->nonb_cut;
  if (k == "nonb_low") return s->nonb_low;
  if (k == "bond_cut") return s->bond_cut;
  if (k == "hbond_cut") return s->hbond_cut;
  if (k == "swa") return s->swa;
  if (k == "swb") return s->swb;
  if (k == "tol") return s->tol;
  return NAN;
}

void orc_taper(orc_sys *s, double *tap8) { std::memcpy(tap8, s->Tap, sizeof(s->Tap)); }

int orc_set_atoms(orc_sys *S, int n, int N, const int *tag, const int *type,
                  const int *owner, const double *x, const double *q) {
  bool resize = (S->N != N) || (S->n != n);
  S->n = n; S->N = N;
  S->tag.assign(tag, tag + N);
  S->ltype.assign(type, type + N);
  S->owner.assign(owner, owner + N);
  S->x.assign(x, x + 3 * (size_t)N);
  S->q.assign(q, q + N);
  for (int i = 0; i < N; i++)
    if (S->ltype[i] < 1 || S->ltype[i] > S->nelem) { S->err = "bad type"; return -1; }
  if (resize) {
    if ((int)S->num_bonds.size() < N) {
      S->num_bonds.resize(N, 0);
      S->num_hbonds.resize(N, 0);
    }
    if ((int)S->s_hist.size() != 5 * n) {
      S->s_hist.assign(5 * (size_t)n, 0.0);
      S->t_hist.assign(5 * (size_t)n, 0.0);
    }
    S->qeq_storage_init = false;
  }
  return 0;
}

int orc_set_q(orc_sys *S, const double *q) {
  S->q.assign(q, q + S->N);
  return 0;
}

// ------------------------------------------------------------------------
// a1: full neighbour list over local + ghost atoms.  Restates the candidate
// test of npair_full_bin_ghost_sunway_build_packed_para
// (npair_full_bin_ghost_sw5.c:183-199): j != i, del = x_i - x_j,
// rsq = delx^2+dely^2+delz^2 <= cutneighsq, over the mathematically complete
// set (no bin-0 exclusion).  Cell order: 27 cells of edge >= cut, ascending j
// inside a cell.  brute_force=1 enumerates all j in index order.
// ------------------------------------------------------------------------
int64_t orc_build_verlet(orc_sys *S, double cut, int brute) {
  const int N = S->N;
  const double *x = S->x.data();
  const double cutsq = cut * cut;
  std::vector<std::vector<int>> per(N);
  if (brute) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; i++) {
      for (int j = 0; j < N; j++) {
        if (j == i) continue;
        double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
        double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq <= cutsq) per[i].push_back(j);
      }
    }
  } else {
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++) { lo[d] = 1e300; hi[d] = -1e300; }
    for (int i = 0; i < N; i++)
      for (int d = 0; d < 3; d++) { lo[d] = std::min(lo[d], x[3 * i + d]); hi[d] = std::max(hi[d], x[3 * i + d]); }
    int nb[3];
    for (int d = 0; d < 3; d++) nb[d] = std::max(1, (int)std::floor((hi[d] - lo[d]) / cut));
    auto cell = [&](int i, int d) {
      int c = (int)((x[3 * i + d] - lo[d]) / (hi[d] - lo[d] + 1e-9) * nb[d]);
      return std::min(std::max(c, 0), nb[d] - 1);
    };
    std::vector<std::vector<int>> cells((size_t)nb[0] * nb[1] * nb[2]);
    std::vector<int> ci(3 * (size_t)N);
    for (int i = 0; i < N; i++) {
      for (int d = 0; d < 3; d++) ci[3 * i + d] = cell(i, d);
      cells[((size_t)ci[3 * i + 2] * nb[1] + ci[3 * i + 1]) * nb[0] + ci[3 * i]].push_back(i);
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; i++) {
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int cx = ci[3 * i] + dx, cy = ci[3 * i + 1] + dy, cz = ci[3 * i + 2] + dz;
            if (cx < 0 || cy < 0 || cz < 0 || cx >= nb[0] || cy >= nb[1] || cz >= nb[2]) continue;
            for (int j : cells[((size_t)cz * nb[1] + cy) * nb[0] + cx]) {
              if (j == i) continue;
              double ddx = x[3 * i] - x[3 * j], ddy = x[3 * i + 1] - x[3 * j + 1], ddz = x[3 * i + 2] - x[3 * j + 2];
              double rsq = ddx * ddx + ddy * ddy + ddz * ddz;
              if (rsq <= cutsq) per[i].push_back(j);
            }
          }
    }
  }
  S->voff.assign(N + 1, 0);
  for (int i = 0; i < N; i++) S->voff[i + 1] = S->voff[i] + (int64_t)per[i].size();
  S->vlist.resize(S->voff[N]);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) std::copy(per[i].begin(), per[i].end(), S->vlist.begin() + S->voff[i]);
  return S->voff[N];
}

int orc_set_verlet(orc_sys *S, const int64_t *off, const int *list) {
  S->voff.assign(off, off + S->N + 1);
  S->vlist.assign(list, list + off[S->N]);
  return 0;
}

// a2: write_reax_atoms_and_pack_para, pair_reaxc_sw64.c:114-190.
int orc_write_reax_atoms(orc_sys *S) {
  S->packed.resize(S->N);
  for (int i = 0; i < S->N; i++) {
    swr_atom_pack_t &a = S->packed[i];
    a.orig_id = S->tag[i];
    a.type = S->map[S->ltype[i]];
    a.x[0] = S->x[3 * i]; a.x[1] = S->x[3 * i + 1]; a.x[2] = S->x[3 * i + 2];
    a.q = S->q[i];
  }
  return 0;
}

// a3: write_reax_lists_c_para, pair_reaxc_sw64.c:193-341 (host copy
// pair_reaxc_sunway.cpp:913-937): for every atom and every Verlet neighbour,
// dvec = x_j - x_i, d^2 = dx^2+dy^2+dz^2; keep d^2 <= nonb_cut^2 with
// d = sqrt(d^2) (the CPE's d^2*inv_sqrt(d^2) is replaced by IEEE sqrt).
int64_t orc_write_reax_lists(orc_sys *S) {
  const int N = S->N;
  const double nonb_cutsq = S->nonb_cut * S->nonb_cut;
  const swr_atom_pack_t *P = S->packed.data();
  std::vector<int64_t> cnt(N, 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < N; i++) {
    int64_t c = 0;
    for (int64_t k = S->voff[i]; k < S->voff[i + 1]; k++) {
      int j = S->vlist[k];
      double dx = P[j].x[0] - P[i].x[0], dy = P[j].x[1] - P[i].x[1], dz = P[j].x[2] - P[i].x[2];
      double dsq = dx * dx + dy * dy + dz * dz;
      if (dsq <= nonb_cutsq) c++;
    }
    cnt[i] = c;
  }
  S->far_index.assign(N, 0);
  S->far_end.assign(N, 0);
  int64_t tot = 0;
  for (int i = 0; i < N; i++) { S->far_index[i] = tot; tot += cnt[i]; S->far_end[i] = tot; }
  S->far.resize(tot);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < N; i++) {
    int64_t o = S->far_index[i];
    for (int64_t k = S->voff[i]; k < S->voff[i + 1]; k++) {
      int j = S->vlist[k];
      double dx = P[j].x[0] - P[i].x[0], dy = P[j].x[1] - P[i].x[1], dz = P[j].x[2] - P[i].x[2];
      double dsq = dx * dx + dy * dy + dz * dz;
      if (dsq <= nonb_cutsq) {
        swr_far_neighbor_t &e = S->far[o++];
        e.nbr = j;
        e.type = P[j].type;
        e.orig_id = P[j].orig_id;
        e.rel_box[0] = e.rel_box[1] = e.rel_box[2] = 0;
        e.d = std::sqrt(dsq);
        e.q = P[j].q;
        e.dvec[0] = dx; e.dvec[1] = dy; e.dvec[2] = dz;
      }
    }
  }
  return tot;
}

namespace {

// Uncorrected bond order of one far entry: BOp_single,
// reaxc_forces_sunway.cpp:677-717 (host form of reaxc_forces_sw64.c:1154-1199).
struct BOp { double BO, BO_s, BO_pi, BO_pi2, C12, C34, C56; };

BOp bop_uncorrected(const orc_sys *S, int ei, int ej, double d) {
  const SBP &si = S->sbp[ei], &sj = S->sbp[ej];
  const TBP &tw = S->tbp[ei][ej];
  BOp b;
  if (si.r_s > 0.0 && sj.r_s > 0.0) {
    b.C12 = tw.p_bo1 * std::pow(d / tw.r_s, tw.p_bo2);
    b.BO_s = (1.0 + S->bo_cut) * std::exp(b.C12);
  } else b.BO_s = b.C12 = 0.0;
  if (si.r_pi > 0.0 && sj.r_pi > 0.0) {
    b.C34 = tw.p_bo3 * std::pow(d / tw.r_p, tw.p_bo4);
    b.BO_pi = std::exp(b.C34);
  } else b.BO_pi = b.C34 = 0.0;
  if (si.r_pi_pi > 0.0 && sj.r_pi_pi > 0.0) {
    b.C56 = tw.p_bo5 * std::pow(d / tw.r_pp, tw.p_bo6);
    b.BO_pi2 = std::exp(b.C56);
  } else b.BO_pi2 = b.C56 = 0.0;
  b.BO = b.BO_s + b.BO_pi + b.BO_pi2;
  return b;
}

}  // namespace

// a4 (setup): Estimate_Storages, reaxc_forces_sunway.cpp:1159-1294, and
// Init_Lists, reaxc_init_md_sunway.cpp:171-238.  Bond counts are taken for
// local atoms only (ghosts start at MIN_BONDS), hbond counts for local H.
int orc_estimate_storages(orc_sys *S) {
  const int n = S->n, N = S->N;
  std::vector<int> bond_top(N, 0), hb_top(n, 0);
  for (int i = 0; i < n; i++) {
    int ei = S->packed[i].type;
    const SBP &si = S->sbp[ei];
    int ihb = si.p_hbond;
    for (int64_t pj = S->far_index[i]; pj < S->far_end[i]; pj++) {
      const swr_far_neighbor_t &e = S->far[pj];
      if (e.d <= S->nonb_cut) {
        int ej = S->packed[e.nbr].type;
        if (S->hbond_cut > 0.1 && (ihb == 1 || ihb == 2) && e.d <= S->hbond_cut) {
          if (ihb == 1 && S->sbp[ej].p_hbond == 2) ++hb_top[i];
        }
        if (e.d <= S->bond_cut) {
          BOp b = bop_uncorrected(S, ei, ej, e.d);
          if (b.BO >= S->bo_cut) ++bond_top[i];
        }
      }
    }
  }
  if ((int)S->num_bonds.size() < N) { S->num_bonds.resize(N); S->num_hbonds.resize(N); }
  for (int i = 0; i < n; i++) S->num_hbonds[i] = (int)std::max(hb_top[i] * S->saferzone, (double)MIN_HBONDS);
  for (int i = n; i < N; i++) S->num_hbonds[i] = 0;
  for (int i = 0; i < N; i++) S->num_bonds[i] = std::max(bond_top[i] * 2, MIN_BONDS);
  S->estimated = true;
  return 0;
}

// a4: Reset, reaxc_reset_tools_sunway.cpp:36-57 (Reset_Atoms), :122-131
// (Reset_Workspace) and :134-197 (Reset_Neighbor_Lists).
int orc_reset(orc_sys *S) {
  const int n = S->n, N = S->N;
  S->Hindex.assign(N, -1);
  S->numH = 0;
  if (S->hbond_cut > 0)
    for (int i = 0; i < n; i++)
      S->Hindex[i] = (S->sbp[S->map[S->ltype[i]]].p_hbond == 1) ? S->numH++ : -1;
  S->bo_dboc.assign(2 * (size_t)N, 0.0);
  S->dDeltap_self.assign(3 * (size_t)N, 0.0);
  S->fCdDelta.assign(4 * (size_t)N, 0.0);
  S->bindex.assign(N, 0); S->bend.assign(N, 0);
  int tot = 0;
  for (int i = 0; i < N; i++) { S->bindex[i] = S->bend[i] = tot; tot += S->num_bonds[i]; }
  S->num_bonds_total = tot;
  S->hindex.assign(S->numH, 0); S->hend.assign(S->numH, 0);
  int htot = 0;
  for (int i = 0; i < n; i++)
    if (S->Hindex[i] > -1) {
      S->hindex[S->Hindex[i]] = S->hend[S->Hindex[i]] = htot;
      htot += S->num_hbonds[i];
    }
  S->num_hbonds_total = htot;
  size_t cap = (size_t)tot;
  S->bonds.assign(cap, swr_bond_t());
  S->bo_data.assign(cap, swr_bond_order_t());
  S->BO.assign(cap, 0.0); S->Cdbo.assign(cap, 0.0); S->Cdbopi.assign(cap, 0.0); S->Cdbopi2.assign(cap, 0.0);
  S->BOpi.assign(2 * cap, 0.0);
  S->hbonds.assign(htot, swr_hbond_t());
  S->hbond_far.assign(htot, -1);
  return 0;
}

// a5 + a6: Init_Forces_noQEq_Full_C_para, reaxc_forces_sw64.c:996-1318
// (bonds, uncorrected BO, dBOp, per-atom sort by nbr via my_bubble_sort
// :31-75), and Init_Forces_noQEq_HB_Full_C, :787-863 (hbonds of local H).
int orc_init_forces(orc_sys *S) {
  const int n = S->n, N = S->N;
  S->bond_mark.assign(N, 0);
  int overflow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : overflow)
  for (int i = 0; i < N; i++) {
    S->bond_mark[i] = (i < n) ? 0 : 1000;
    int ei = S->packed[i].type;
    if (ei < 0) continue;
    int start = S->bindex[i];
    int cap = S->num_bonds[i];
    int btop = S->bend[i];
    struct Rec { swr_bond_t b; swr_bond_order_t o; double BO, pi, pi2; };
    std::vector<Rec> recs;
    for (int64_t pj = S->far_index[i]; pj < S->far_end[i]; pj++) {
      const swr_far_neighbor_t &e = S->far[pj];
      if (!(e.d <= S->bond_cut)) continue;
      int ej = e.type;
      if (ej < 0) continue;
      BOp b = bop_uncorrected(S, ei, ej, e.d);
      if (b.BO >= S->bo_cut) {
        const TBP &tw = S->tbp[ei][ej];
        Rec rc;
        std::memset(&rc, 0, sizeof(rc));
        rc.b.nbr = e.nbr;
        rc.b.d = e.d;
        for (int k = 0; k < 3; k++) { rc.b.dvec[k] = e.dvec[k]; rc.b.rel_box[k] = e.rel_box[k]; }
        rc.b.dbond_index = btop;
        rc.b.sym_index = 0;
        rc.BO = b.BO;
        rc.o.BO_s = b.BO_s;
        rc.pi = b.BO_pi;
        rc.pi2 = b.BO_pi2;
        double r2 = SQR(e.d);
        double Cln_BOp_s = tw.p_bo2 * b.C12 / r2;
        double Cln_BOp_pi = tw.p_bo4 * b.C34 / r2;
        double Cln_BOp_pi2 = tw.p_bo6 * b.C56 / r2;
        for (int k = 0; k < 3; k++) {
          rc.o.dln_BOp_s[k] = -rc.o.BO_s * Cln_BOp_s * rc.b.dvec[k];
          rc.o.dln_BOp_pi[k] = -rc.pi * Cln_BOp_pi * rc.b.dvec[k];
          rc.o.dln_BOp_pi2[k] = -rc.pi2 * Cln_BOp_pi2 * rc.b.dvec[k];
        }
        double cdb = -(rc.o.BO_s * Cln_BOp_s + rc.pi * Cln_BOp_pi + rc.pi2 * Cln_BOp_pi2);
        for (int k = 0; k < 3; k++) {
          rc.o.dBOp[k] = cdb * rc.b.dvec[k];
          S->dDeltap_self[3 * (size_t)i + k] += rc.o.dBOp[k];
        }
        rc.o.BO_s -= S->bo_cut;
        rc.BO -= S->bo_cut;
        S->bo_dboc[2 * (size_t)i] += rc.BO;
        recs.push_back(rc);
        btop++;
      }
    }
    if ((int)recs.size() > cap) { overflow = 1; continue; }
    std::stable_sort(recs.begin(), recs.end(), [](const Rec &a, const Rec &b) { return a.b.nbr < b.b.nbr; });
    for (size_t k = 0; k < recs.size(); k++) {
      size_t p = (size_t)start + k;
      S->bonds[p] = recs[k].b;
      S->bo_data[p] = recs[k].o;
      S->BO[p] = recs[k].BO;
      S->BOpi[2 * p] = recs[k].pi;
      S->BOpi[2 * p + 1] = recs[k].pi2;
      S->Cdbo[p] = S->Cdbopi[p] = S->Cdbopi2[p] = 0.0;
    }
    S->bend[i] = start + (int)recs.size();
  }
  if (overflow) { S->err = "bond capacity overflow"; return 2; }
  // hbonds (local H only), reaxc_forces_sw64.c:810-862
  int hover = 0;
  if (S->hbond_cut > 0) {
#pragma omp parallel for schedule(dynamic, 64) reduction(| : hover)
    for (int i = 0; i < n; i++) {
      int ei = S->packed[i].type;
      if (ei < 0) continue;
      if (S->sbp[ei].p_hbond != 1) continue;
      int H = S->Hindex[i];
      int top = S->hend[H];
      int lim = S->hindex[H] + S->num_hbonds[i];
      for (int64_t pj = S->far_index[i]; pj < S->far_end[i]; pj++) {
        const swr_far_neighbor_t &e = S->far[pj];
        if (e.d <= S->hbond_cut) {
          int ej = S->packed[e.nbr].type;
          if (ej < 0) continue;
          if (S->sbp[ej].p_hbond == 2) {
            if (top >= lim) { hover = 1; break; }
            S->hbonds[top].nbr = e.nbr;
            S->hbonds[top].scl = 1;
            S->hbonds[top].ptr = &S->far[pj];
            S->hbond_far[top] = pj;
            ++top;
          }
        }
      }
      S->hend[H] = top;
    }
  }
  if (hover) { S->err = "hbond capacity overflow"; return 3; }
  return 0;
}

// a7: Validate_Lists_C, reaxc_forces_sw64.c:866-930 (capacity update), and
// sort_bonds_C, :932-971 (single in-order pass bond_mark[i] =
// min(bond_mark[i], bond_mark[j]+1)).  Ghost num_hbonds are left untouched
// (the reference's Hindex of ghosts is stale; only local H are meaningful).
int orc_validate_lists(orc_sys *S) {
  const int n = S->n, N = S->N;
  for (int i = 0; i < N; i++) {
    S->num_bonds[i] = std::max((S->bend[i] - S->bindex[i]) * 2, MIN_BONDS);
    int comp = (i < N - 1) ? S->bindex[i + 1] : S->num_bonds_total;
    if (S->bend[i] > comp) { S->err = "bondchk failed"; return 2; }
  }
  for (int i = 0; i < n; i++) {
    int H = S->Hindex[i];
    if (H > -1) {
      S->num_hbonds[i] = (int)std::max((S->hend[H] - S->hindex[H]) * S->saferzone, (double)MIN_HBONDS);
    }
  }
  for (int i = 0; i < N; i++)
    for (int pj = S->bindex[i]; pj < S->bend[i]; pj++) {
      int j = S->bonds[pj].nbr;
      if (S->bond_mark[i] > S->bond_mark[j] + 1) S->bond_mark[i] = S->bond_mark[j] + 1;
    }
  return 0;
}

// a8-a10: BO, reaxc_bond_orders_sw64.c:487-840 with the host branches of
// BO_Section1_C (:56-64), BO_Section2_C (:163-376) and BO_Section3_C (:451-481).
int orc_bo(orc_sys *S) {
  const int N = S->N;
  const double p_boc1 = S->gp[0], p_boc2 = S->gp[1], p_lp1 = S->gp[15];
  auto ensure = [&](std::vector<double> &v) { v.assign(N, 0.0); };
  ensure(S->Deltap); ensure(S->Deltap_boc);
  // Section 1
  for (int i = 0; i < N; i++) {
    int ei = S->packed[i].type;
    if (ei < 0) continue;
    const SBP &si = S->sbp[ei];
    S->Deltap[i] = S->bo_dboc[2 * (size_t)i] - si.valency;
    S->Deltap_boc[i] = S->bo_dboc[2 * (size_t)i] - si.valency_val;
    S->bo_dboc[2 * (size_t)i] = 0;
  }
  // Section 2
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < N; i++) {
    int ei = S->packed[i].type;
    if (ei < 0) continue;
    const SBP &si = S->sbp[ei];
    double val_i = si.valency, Deltap_i = S->Deltap[i], Deltap_boc_i = S->Deltap_boc[i];
    for (int pj = S->bindex[i]; pj < S->bend[i]; pj++) {
      int j = S->bonds[pj].nbr;
      int ej = S->packed[j].type;
      if (ej < 0) continue;
      swr_bond_order_t &bo = S->bo_data[pj];
      for (int pk = S->bindex[j]; pk < S->bend[j]; pk++)
        if (S->bonds[pk].nbr == i) S->bonds[pj].sym_index = pk;
      const TBP &tw = S->tbp[ei][ej];
      double *BOp = &S->BO[pj];
      double *pi = &S->BOpi[2 * (size_t)pj];
      if (tw.ovc < 0.001 && tw.v13cor < 0.001) {
        bo.C1dbo = 1.0; bo.C2dbo = 0.0; bo.C3dbo = 0.0;
        bo.C1dbopi = pi[0]; bo.C2dbopi = 0.0; bo.C3dbopi = 0.0; bo.C4dbopi = 0.0;
        bo.C1dbopi2 = pi[1]; bo.C2dbopi2 = 0.0; bo.C3dbopi2 = 0.0; bo.C4dbopi2 = 0.0;
      } else {
        double val_j = S->sbp[ej].valency;
        double Deltap_j = S->Deltap[j], Deltap_boc_j = S->Deltap_boc[j];
        double f1, Cf1_ij, Cf1_ji;
        if (tw.ovc >= 0.001) {
          double exp_p1i = std::exp(-p_boc1 * Deltap_i);
          double exp_p2i = std::exp(-p_boc2 * Deltap_i);
          double exp_p1j = std::exp(-p_boc1 * Deltap_j);
          double exp_p2j = std::exp(-p_boc2 * Deltap_j);
          double f2 = exp_p1i + exp_p1j;
          double f3 = -1.0 / p_boc2 * std::log(0.5 * (exp_p2i + exp_p2j));
          f1 = 0.5 * ((val_i + f2) / (val_i + f2 + f3) + (val_j + f2) / (val_j + f2 + f3));
          double temp = f2 + f3;
          double u1_ij = val_i + temp, u1_ji = val_j + temp;
          double Cf1A_ij = 0.5 * f3 * (1.0 / SQR(u1_ij) + 1.0 / SQR(u1_ji));
          double Cf1B_ij = -0.5 * ((u1_ij - f3) / SQR(u1_ij) + (u1_ji - f3) / SQR(u1_ji));
          Cf1_ij = 0.50 * (-p_boc1 * exp_p1i / u1_ij -
                           ((val_i + f2) / SQR(u1_ij)) * (-p_boc1 * exp_p1i + exp_p2i / (exp_p2i + exp_p2j)) +
                           -p_boc1 * exp_p1i / u1_ji -
                           ((val_j + f2) / SQR(u1_ji)) * (-p_boc1 * exp_p1i + exp_p2i / (exp_p2i + exp_p2j)));
          Cf1_ji = -Cf1A_ij * p_boc1 * exp_p1j + Cf1B_ij * exp_p2j / (exp_p2i + exp_p2j);
        } else {
          f1 = 1.0;
          Cf1_ij = Cf1_ji = 0.0;
        }
        double f4, f5, f4f5, Cf45_ij, Cf45_ji;
        if (tw.v13cor >= 0.001) {
          double exp_f4 = std::exp(-(tw.p_boc4 * SQR(*BOp) - Deltap_boc_i) * tw.p_boc3 + tw.p_boc5);
          double exp_f5 = std::exp(-(tw.p_boc4 * SQR(*BOp) - Deltap_boc_j) * tw.p_boc3 + tw.p_boc5);
          f4 = 1. / (1. + exp_f4);
          f5 = 1. / (1. + exp_f5);
          f4f5 = f4 * f5;
          Cf45_ij = -f4 * exp_f4;
          Cf45_ji = -f5 * exp_f5;
        } else {
          f4 = f5 = f4f5 = 1.0;
          Cf45_ij = Cf45_ji = 0.0;
        }
        double A0_ij = f1 * f4f5;
        double A1_ij = -2 * tw.p_boc3 * tw.p_boc4 * (*BOp) * (Cf45_ij + Cf45_ji);
        double A2_ij = Cf1_ij / f1 + tw.p_boc3 * Cf45_ij;
        double A2_ji = Cf1_ji / f1 + tw.p_boc3 * Cf45_ji;
        double A3_ij = A2_ij + Cf1_ij / f1;
        double A3_ji = A2_ji + Cf1_ji / f1;
        *BOp = *BOp * A0_ij;
        pi[0] = pi[0] * A0_ij * f1;
        pi[1] = pi[1] * A0_ij * f1;
        bo.BO_s = *BOp - (pi[0] + pi[1]);
        bo.C1dbo = A0_ij + *BOp * A1_ij;
        bo.C2dbo = *BOp * A2_ij;
        bo.C3dbo = *BOp * A2_ji;
        bo.C1dbopi = f1 * f1 * f4 * f5;
        bo.C2dbopi = pi[0] * A1_ij;
        bo.C3dbopi = pi[0] * A3_ij;
        bo.C4dbopi = pi[0] * A3_ji;
        bo.C1dbopi2 = f1 * f1 * f4 * f5;
        bo.C2dbopi2 = pi[1] * A1_ij;
        bo.C3dbopi2 = pi[1] * A3_ij;
        bo.C4dbopi2 = pi[1] * A3_ji;
      }
      if (*BOp < 1e-10) *BOp = 0.0;
      if (bo.BO_s < 1e-10) bo.BO_s = 0.0;
      if (pi[0] < 1e-10) pi[0] = 0.0;
      if (pi[1] < 1e-10) pi[1] = 0.0;
      S->bo_dboc[2 * (size_t)i] += *BOp;
    }
  }
  // Section 3
  ensure(S->Delta); ensure(S->Delta_e); ensure(S->Delta_val); ensure(S->vlpex); ensure(S->nlp);
  ensure(S->nlp_temp); ensure(S->Clp); ensure(S->Delta_lp); ensure(S->dDelta_lp);
  ensure(S->Delta_lp_temp); ensure(S->dDelta_lp_temp);
  for (int j = 0; j < N; j++) {
    int ej = S->packed[j].type;
    if (ej < 0) continue;
    const SBP &sj = S->sbp[ej];
    double tot = S->bo_dboc[2 * (size_t)j];
    S->Delta[j] = tot - sj.valency;
    S->Delta_e[j] = tot - sj.valency_e;
    S->bo_dboc[2 * (size_t)j + 1] = tot - sj.valency_boc;
    S->Delta_val[j] = tot - sj.valency_val;
    S->vlpex[j] = S->Delta_e[j] - 2.0 * (int)(S->Delta_e[j] / 2.0);
    double explp1 = std::exp(-p_lp1 * SQR(2.0 + S->vlpex[j]));
    S->nlp[j] = explp1 - (int)(S->Delta_e[j] / 2.0);
    S->Delta_lp[j] = sj.nlp_opt - S->nlp[j];
    S->Clp[j] = 2.0 * p_lp1 * explp1 * (2.0 + S->vlpex[j]);
    S->dDelta_lp[j] = S->Clp[j];
    if (sj.mass > 21.0) {
      S->nlp_temp[j] = 0.5 * (sj.valency_e - sj.valency);
      S->Delta_lp_temp[j] = sj.nlp_opt - S->nlp_temp[j];
      S->dDelta_lp_temp[j] = 0.;
    } else {
      S->nlp_temp[j] = S->nlp[j];
      S->Delta_lp_temp[j] = sj.nlp_opt - S->nlp_temp[j];
      S->dDelta_lp_temp[j] = S->Clp[j];
    }
  }
  return 0;
}

// ------------------------------------------------------------------------
// QEq
// ------------------------------------------------------------------------
static void qeq_init_storage(orc_sys *S) {
  // init_storage, fix_qeq_reax_sunway.cpp:518-535
  const int N = S->N;
  S->Hdia_inv.assign(N, 0.0); S->b_s.assign(N, 0.0); S->b_t.assign(N, 0.0);
  S->s.assign(N, 0.0); S->t.assign(N, 0.0);
  for (int i = 0; i < N; i++) {
    int ti = S->ltype[i];
    S->Hdia_inv[i] = 1. / S->eta[ti];
    S->b_s[i] = -S->chi[ti];
    S->b_t[i] = -1.0;
  }
  S->qeq_storage_init = true;
}

// a11-a13: init_matvec, fix_qeq_reax_sunway.cpp:618-723, with the H build of
// compute_H_Full_C_para, fix_qeq_reax_sw64.c:768-911 (full rows for local i,
// every Verlet neighbour with r^2 <= swb^2, value calculate_H :965-981).
int orc_qeq_init_matvec(orc_sys *S) {
  const int n = S->n, N = S->N;
  if (!S->qeq_storage_init || (int)S->s.size() != N) qeq_init_storage(S);
  const double swbsq = S->swb * S->swb;
  const double *x = S->x.data();
  const int ne = S->nelem + 1;
  std::vector<int> cnt(n, 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    int c = 0;
    for (int64_t k = S->voff[i]; k < S->voff[i + 1]; k++) {
      int j = S->vlist[k];
      double dx = x[3 * j] - x[3 * i], dy = x[3 * j + 1] - x[3 * i + 1], dz = x[3 * j + 2] - x[3 * i + 2];
      double r_sqr = SQR(dx) + SQR(dy) + SQR(dz);
      if (r_sqr <= swbsq) c++;
    }
    cnt[i] = c;
  }
  S->Hfirst.assign(n, 0); S->Hnum.assign(n, 0);
  int tot = 0;
  for (int i = 0; i < n; i++) { S->Hfirst[i] = tot; S->Hnum[i] = cnt[i]; tot += cnt[i]; }
  S->Hj.resize(tot); S->Hval.resize(tot);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    int o = S->Hfirst[i];
    int ti = S->ltype[i];
    for (int64_t k = S->voff[i]; k < S->voff[i + 1]; k++) {
      int j = S->vlist[k];
      double dx = x[3 * j] - x[3 * i], dy = x[3 * j + 1] - x[3 * i + 1], dz = x[3 * j + 2] - x[3 * i + 2];
      double r_sqr = SQR(dx) + SQR(dy) + SQR(dz);
      if (r_sqr <= swbsq) {
        double r = std::sqrt(r_sqr);
        double gamma = S->shld[ti * ne + S->ltype[j]];
        double Taper = S->Tap[7] * r + S->Tap[6];
        Taper = Taper * r + S->Tap[5];
        Taper = Taper * r + S->Tap[4];
        Taper = Taper * r + S->Tap[3];
        Taper = Taper * r + S->Tap[2];
        Taper = Taper * r + S->Tap[1];
        Taper = Taper * r + S->Tap[0];
        double denom = r * r * r + gamma;
        denom = std::pow(denom, 0.3333333333333);
        S->Hj[o] = j;
        S->Hval[o] = Taper * EV_TO_KCAL_PER_MOL / denom;
        o++;
      }
    }
  }
  // :700-722 preconditioner, right-hand sides, history extrapolation
  for (int i = 0; i < n; i++) {
    int ti = S->ltype[i];
    S->Hdia_inv[i] = 1. / S->eta[ti];
    S->b_s[i] = -S->chi[ti];
    S->b_t[i] = -1.0;
    const double *th = &S->t_hist[5 * (size_t)i], *sh = &S->s_hist[5 * (size_t)i];
    S->t[i] = th[2] + 3 * (th[0] - th[1]);
    S->s[i] = 4 * (sh[0] + sh[2]) - (6 * sh[1] + sh[3]);
  }
  forward(S, S->s);
  forward(S, S->t);
  return 0;
}

// a14: sparse_matvec, fix_qeq_reax_sunway.cpp:1563-1623 -> sparse_matvec_C_para,
// fix_qeq_reax_sw64.c:357-487: b_i = eta_i x_i + sum_k val_k x_{j_k}, i < n.
static void spmv(orc_sys *S, const double *xv, double *b) {
  const int n = S->n;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    double bi = S->eta[S->ltype[i]] * xv[i];
    for (int k = S->Hfirst[i]; k < S->Hfirst[i] + S->Hnum[i]; k++) bi += S->Hval[k] * xv[S->Hj[k]];
    b[i] = bi;
  }
}

int orc_sparse_matvec(orc_sys *S, const double *x_N, double *b_n) {
  spmv(S, x_N, b_n);
  return 0;
}

// a15: CG_v2, fix_qeq_reax_sunway.cpp:983-1167 — pipelined Jacobi PCG,
// imax = 200, stop when sqrt(u.r)/|b| <= tol.  Dots are summed sequentially
// in atom order (one rank's loop, then an identity allreduce).
static int cg_v2(orc_sys *S, std::vector<double> &bv, std::vector<double> &xv) {
  const int nn = S->n, N = S->N;
  const int imax = S->cg_imax;
  auto &r = S->r, &u = S->u, &w = S->w, &m = S->m, &v = S->v, &z = S->z, &ss = S->ss, &p = S->p,
       &d = S->d, &q = S->qv;
  for (auto *vec : {&r, &u, &w, &m, &v, &z, &ss, &p, &d, &q}) vec->assign(N, 0.0);
  const double *Hd = S->Hdia_inv.data();
  const double *b = bv.data();
  double *x = xv.data();
  const double tol = S->tol;

  spmv(S, x, q.data());
  for (int j = 0; j < nn; j++) { r[j] = b[j] - 1 * q[j]; u[j] = r[j] * Hd[j]; d[j] = u[j]; }
  forward(S, d);
  spmv(S, d.data(), q.data());
  for (int j = 0; j < nn; j++) { w[j] = q[j]; m[j] = q[j] * Hd[j]; d[j] = m[j]; }
  forward(S, d);
  spmv(S, d.data(), q.data());
  double sum0 = 0, sum1 = 0, sum2 = 0;
  for (int j = 0; j < nn; j++) {
    p[j] = u[j]; ss[j] = w[j]; v[j] = m[j];
    sum0 += SQR(b[j]); sum1 += u[j] * r[j]; sum2 += u[j] * w[j];
  }
  double b_norm = std::sqrt(sum0);
  double sig_old = sum1;
  double deta = sum2;
  double heta = deta;
  double alpha = sig_old / deta;
  double dg0 = sig_old, dg1 = 0;
  for (int j = 0; j < nn; j++) z[j] = q[j];
  int i;
  for (i = 1; i < imax && std::sqrt(dg0) / b_norm > tol; ++i) {
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < nn; jj++) {
      r[jj] -= alpha * ss[jj];
      u[jj] -= alpha * v[jj];
      w[jj] -= alpha * z[jj];
      d[jj] = w[jj] * Hd[jj];
    }
    double l0 = 0, l1 = 0;
    for (int jj = 0; jj < nn; jj++) { l0 += u[jj] * r[jj]; l1 += u[jj] * w[jj]; }
    forward(S, d);
    spmv(S, d.data(), q.data());
    dg0 = l0; dg1 = l1;
    double beta = dg0 / sig_old;
    heta = dg1 - beta * beta * heta;
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < nn; jj++) {
      x[jj] += alpha * p[jj];
      p[jj] = u[jj] + p[jj] * beta;
      ss[jj] = w[jj] + ss[jj] * beta;
      v[jj] = d[jj] + v[jj] * beta;
    }
    alpha = dg0 / heta;
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < nn; jj++) z[jj] = q[jj] + z[jj] * beta;
    sig_old = dg0;
  }
  return i;
}

void orc_set_cg_imax(orc_sys *S, int imax) { S->cg_imax = imax > 1 ? imax : 1; }

int orc_qeq_cg(orc_sys *S, int which) {
  if (which == 0) return S->matvecs_s = cg_v2(S, S->b_s, S->s);
  return S->matvecs_t = cg_v2(S, S->b_t, S->t);
}

// a16: calculate_Q, fix_qeq_reax_sunway.cpp:1697-1755.
int orc_qeq_calculate_q(orc_sys *S) {
  const int nn = S->n;
  double s_sum = 0, t_sum = 0;
  for (int i = 0; i < nn; i++) { s_sum += S->s[i]; t_sum += S->t[i]; }
  double uu = s_sum / t_sum;
  for (int i = 0; i < nn; i++) {
    S->q[i] = S->s[i] - uu * S->t[i];
    double *sh = &S->s_hist[5 * (size_t)i], *th = &S->t_hist[5 * (size_t)i];
    for (int k = 4; k > 0; --k) { sh[k] = sh[k - 1]; th[k] = th[k - 1]; }
    sh[0] = S->s[i];
    th[0] = S->t[i];
  }
  forward(S, S->q);
  return 0;
}

// pre_force, fix_qeq_reax_sunway.cpp:539-600.
int orc_qeq_pre_force(orc_sys *S, int *its, int *itt) {
  orc_qeq_init_matvec(S);
  int a = orc_qeq_cg(S, 0);
  int b = orc_qeq_cg(S, 1);
  if (its) *its = a;
  if (itt) *itt = b;
  return orc_qeq_calculate_q(S);
}

// One step of the path in the reference's order (call stack: Verlet step ->
// FixQEqReaxSunway::pre_force -> PairReaxCSunway::compute).
// f2: vdW_Coulomb_Energy_Full_C, reaxc_nonbonded_sw64.c:260-500, in its
// general host form :522-830 (the vdw_type 1 CPE kernel,
// reaxc_nonbonded_cpe.h:1-420, evaluates the same expressions), with libm
// pow/exp in place of p_powd/p_expd and lgvdw off (pair_reaxc_sunway.cpp:236).
// Over local atoms i and every far entry j (types >= 0): the taper of
// Init_Taper (reaxc_init_md_sunway.cpp:100-134, nonb_low..nonb_cut), the
// shielded Morse vdW (+ inner core for vdw_type 2/3) and the tapered
// shielded Coulomb; CEvd and CEclmb are (dE/dr)/r.  The gradient
// -(CEvd + CEclmb) dvec is added to fCdDelta[i][0..2]; eng_vdwl, eng_coul and
// the virial take half of every full-list pair.  out: eng_vdwl, eng_coul,
// virial[6].
int orc_vdw_coulomb(orc_sys *S, double *out) {
  const int n = S->n;
  const double C_ele = 332.06371;
  const double p_vdW1 = S->gp[28], p_vdW1i = 1.0 / p_vdW1;
  if ((int64_t)S->fCdDelta.size() < 4 * (int64_t)S->N) S->fCdDelta.assign(4 * (size_t)S->N, 0.0);
  double tap[8];
  {
    const double swa = S->nonb_low, swb = S->nonb_cut;
    const double d7 = std::pow(swb - swa, 7.0);
    const double swa2 = swa * swa, swa3 = swa * swa * swa, swb2 = swb * swb, swb3 = swb * swb * swb;
    tap[7] = 20.0 / d7;
    tap[6] = -70.0 * (swa + swb) / d7;
    tap[5] = 84.0 * (swa2 + 3.0 * swa * swb + swb2) / d7;
    tap[4] = -35.0 * (swa3 + 9.0 * swa2 * swb + 9.0 * swa * swb2 + swb3) / d7;
    tap[3] = 140.0 * (swa3 * swb + 3.0 * swa2 * swb2 + swa * swb3) / d7;
    tap[2] = -210.0 * (swa3 * swb2 + swa2 * swb3) / d7;
    tap[1] = 140.0 * swa3 * swb3 / d7;
    tap[0] = (-35.0 * swa3 * swb2 * swb2 + 21.0 * swa2 * swb3 * swb2 + 7.0 * swa * swb3 * swb3 + swb3 * swb3 * swb) / d7;
  }
  const int vt = S->vdw_type;
  std::vector<double> per(8 * (size_t)n, 0.0);   // per-atom energies + virial, summed in index order below
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    const swr_atom_pack_t &ai = S->packed[i];
    if (ai.type < 0) continue;
    double *acc = &per[8 * (size_t)i];
    double g[3] = {0, 0, 0};
    for (int64_t pj = S->far_index[i]; pj < S->far_end[i]; pj++) {
      const swr_far_neighbor_t &e = S->far[pj];
      if (e.type < 0) continue;
      const TBP &tw = S->tbp[ai.type][e.type];
      // per-pair factors as the host form precomputes them (:386-402)
      const double r_vdWinv = 1.0 / tw.r_vdW, alpha_div_rvdW = tw.alpha / tw.r_vdW;
      const double powgi = std::pow(tw.gamma_w, -p_vdW1);
      const double r = e.d, rinv = 1.0 / r;
      double Tap = tap[7] * r + tap[6];
      Tap = Tap * r + tap[5]; Tap = Tap * r + tap[4]; Tap = Tap * r + tap[3];
      Tap = Tap * r + tap[2]; Tap = Tap * r + tap[1]; Tap = Tap * r + tap[0];
      double dTap = 7 * tap[7] * r + 6 * tap[6];
      dTap = dTap * r + 5 * tap[5]; dTap = dTap * r + 4 * tap[4];
      dTap = dTap * r + 3 * tap[3]; dTap = dTap * r + 2 * tap[2];
      dTap += tap[1] * rinv;
      double e_vdW, CEvd, e_core = 0;
      if (vt == 1 || vt == 3) {
        const double powr = std::pow(r, p_vdW1);
        const double fn13 = std::pow(powr + powgi, p_vdW1i);
        const double exp2 = std::exp(0.5 * tw.alpha * (1.0 - fn13 * r_vdWinv));
        const double exp1 = exp2 * exp2;
        e_vdW = tw.D * (exp1 - 2.0 * exp2);
        const double dfn13 = std::pow(powr + powgi, p_vdW1i - 1.0) * std::pow(r, p_vdW1 - 2.0);
        CEvd = dTap * e_vdW - Tap * tw.D * alpha_div_rvdW * (exp1 - exp2) * dfn13;
      } else {
        const double exp2 = std::exp(0.5 * tw.alpha * (1.0 - r * r_vdWinv));
        const double exp1 = exp2 * exp2;
        e_vdW = tw.D * (exp1 - 2.0 * exp2);
        CEvd = dTap * e_vdW - Tap * tw.D * alpha_div_rvdW * (exp1 - exp2) * rinv;
      }
      if (vt == 2 || vt == 3) {
        const double rcoreinv = 1.0 / tw.rcore, a_div_rcore = tw.acore / tw.rcore;
        e_core = tw.ecore * std::exp(tw.acore * (1.0 - r * rcoreinv));
        const double de_core = -a_div_rcore * e_core;
        CEvd += dTap * e_core + Tap * de_core * rinv;
      }
      const double dr3gamij_1 = r * r * r + tw.gamma;
      const double dr3gamij_3 = std::pow(dr3gamij_1, 0.33333333333333);
      const double qq = C_ele * ai.q * e.q;
      const double e_ele = qq * Tap / dr3gamij_3;
      const double CEclmb = qq * (dTap - Tap * r / dr3gamij_1) / dr3gamij_3;
      const double fpair = -(CEvd + CEclmb);
      acc[0] += 0.5 * Tap * (e_vdW + e_core);
      acc[1] += 0.5 * e_ele;
      const double *del = e.dvec;
      acc[2] += 0.5 * del[0] * del[0] * fpair;
      acc[3] += 0.5 * del[1] * del[1] * fpair;
      acc[4] += 0.5 * del[2] * del[2] * fpair;
      acc[5] += 0.5 * del[0] * del[1] * fpair;
      acc[6] += 0.5 * del[0] * del[2] * fpair;
      acc[7] += 0.5 * del[1] * del[2] * fpair;
      for (int k = 0; k < 3; k++) g[k] += fpair * del[k];
    }
    for (int k = 0; k < 3; k++) S->fCdDelta[4 * (size_t)i + k] += g[k];
  }
  for (int k = 0; k < 8; k++) out[k] = 0.0;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 8; k++) out[k] += per[8 * (size_t)i + k];
  return 0;
}

// Whether local atom i owns bond (i, j) for the bond energy: the smaller
// orig_id, or for two images of one atom the one lower in z, then y, then x
// (Merge_Bonds_Atom_Energy_C_New, reaxc_multi_body_sw64.c:245-255).
static bool owns_bond(const orc_sys *S, int i, int j) {
  const swr_atom_pack_t &a = S->packed[i], &b = S->packed[j];
  if (a.orig_id > b.orig_id) return false;
  if (a.orig_id == b.orig_id) {
    if (b.x[2] < a.x[2]) return false;
    if (b.x[2] == a.x[2] && b.x[1] < a.x[1]) return false;
    if (b.x[2] == a.x[2] && b.x[1] == a.x[1] && b.x[0] < a.x[0]) return false;
  }
  return true;
}

// Atom_Energy (reaxc_multi_body_sunway.cpp:41-57; Bonds() returns at once,
// reaxc_bonds_sunway.cpp:49-52) = Merge_Bonds_Atom_Energy_C_New
// (reaxc_multi_body_sw64.c:21-337), the host loop over local atoms: the
// polarization energy, lone pair (+ the C-C lone-pair correction), over- and
// under-coordination, and for each bond the atom owns the bond energy and the
// terminal triple-bond stabilisation.  Accumulates Cdbo/Cdbopi/Cdbopi2 of
// the atom's bond slots and CdDelta (fCdDelta[.][3]) of the atom and its
// bonded neighbours, in the reference's order.  Departures: a neighbour of
// type < 0 is skipped in the last bond loop as in the first (the reference
// indexes tbp with it there); e_pol is the plain sum over atoms (the value
// the reference adds to eng_coul; its data->my_en.e_pol adds the running
// sum at every atom, :97-98).  out: e_bond (ebond + estriph), e_lp (e_lp +
// e_lph), e_ov, e_un, e_pol.
int orc_atom_energy(orc_sys *S, int enobonds, double *out) {
  const int n = S->n;
  const std::vector<double> &gp = S->gp;
  const double KCALpMOL_to_EV = 23.02;        // reaxc_defs_sunway.h:68
  const double p_lp3 = gp[5], p_ovun3 = gp[32], p_ovun4 = gp[31], p_ovun6 = gp[6], p_ovun7 = gp[8],
               p_ovun8 = gp[9];
  const double gp3 = gp[3], gp4 = gp[4], gp7 = gp[7], gp10 = gp[10];
  const int gp37 = (int)gp[37];
  if ((int64_t)S->fCdDelta.size() < 4 * (int64_t)S->N) S->fCdDelta.assign(4 * (size_t)S->N, 0.0);
  double *fc = S->fCdDelta.data();
  double e_lph = 0, e_lp = 0, e_ov = 0, e_un_sum = 0, ebond = 0, estriph = 0, e_pol = 0;
  for (int i = 0; i < n; i++) {
    const swr_atom_pack_t &ai = S->packed[i];
    const int ti = ai.type;
    if (ti < 0) continue;
    const SBP &si = S->sbp[ti];
    const double q = ai.q;
    e_pol += KCALpMOL_to_EV * (si.chi * q + (si.eta / 2.) * SQR(q));
    // lone pair
    const double p_lp2 = si.p_lp2;
    const double expvd2 = std::exp(-75 * S->Delta_lp[i]);
    const double inv_expvd2 = 1. / (1. + expvd2);
    int numbonds = 0;
    const double p_ovun2 = si.p_ovun2;
    double sum_ovun1 = 0, sum_ovun2 = 0;
    const double dfvl = si.mass > 21.0 ? 0.0 : 1.0;   // 1st-row elements only
    for (int pj = S->bindex[i]; pj < S->bend[i]; ++pj) {
      numbonds++;
      const int j = S->bonds[pj].nbr;
      const int tj = S->packed[j].type;
      if (tj < 0) continue;
      const TBP &tw = S->tbp[ti][tj];
      sum_ovun1 += tw.p_ovun1 * tw.De_s * S->BO[pj];
      sum_ovun2 += (S->Delta[j] - dfvl * S->Delta_lp_temp[j]) * (S->BOpi[2 * (size_t)pj] + S->BOpi[2 * (size_t)pj + 1]);
    }
    const double dElp = p_lp2 * inv_expvd2 + 75 * p_lp2 * S->Delta_lp[i] * expvd2 * SQR(inv_expvd2);
    const double CElp = dElp * S->dDelta_lp[i];
    if (p_lp3 > 0.001 && si.name == "C") {
      for (int pj = S->bindex[i]; pj < S->bend[i]; ++pj) {
        const int j = S->bonds[pj].nbr;
        const int tj = S->packed[j].type;
        if (tj < 0) continue;
        if (S->sbp[tj].name == "C") {
          const double Di = S->Delta[i];
          const double vov3 = S->BO[pj] - Di - 0.040 * std::pow(Di, 4.);
          if (vov3 > 3.) {
            e_lph += p_lp3 * SQR(vov3 - 3.0);
            const double deahu2dbo = 2. * p_lp3 * (vov3 - 3.);
            const double deahu2dsbo = 2. * p_lp3 * (vov3 - 3.) * (-1. - 0.16 * std::pow(Di, 3.));
            S->Cdbo[pj] += deahu2dbo;
            fc[4 * (size_t)i + 3] += deahu2dsbo;
          }
        }
      }
    }
    // over-coordination
    const double exp_ovun1 = p_ovun3 * std::exp(p_ovun4 * sum_ovun2);
    const double inv_exp_ovun1 = 1.0 / (1 + exp_ovun1);
    const double Delta_lpcorr = S->Delta[i] - (dfvl * S->Delta_lp_temp[i]) * inv_exp_ovun1;
    const double exp_ovun2 = std::exp(p_ovun2 * Delta_lpcorr);
    const double inv_exp_ovun2 = 1.0 / (1.0 + exp_ovun2);
    const double DlpVi = 1.0 / (Delta_lpcorr + si.valency + 1e-8);
    const double CEover1 = Delta_lpcorr * DlpVi * inv_exp_ovun2;
    e_ov += sum_ovun1 * CEover1;
    const double CEover2 =
        sum_ovun1 * DlpVi * inv_exp_ovun2 * (1.0 - Delta_lpcorr * (DlpVi + p_ovun2 * exp_ovun2 * inv_exp_ovun2));
    const double CEover3 = CEover2 * (1.0 - dfvl * S->dDelta_lp[i] * inv_exp_ovun1);
    const double CEover4 = CEover2 * (dfvl * S->Delta_lp_temp[i]) * p_ovun4 * exp_ovun1 * SQR(inv_exp_ovun1);
    // under-coordination
    const double p_ovun5 = si.p_ovun5;
    const double exp_ovun2n = 1.0 / exp_ovun2;
    const double exp_ovun6 = std::exp(p_ovun6 * Delta_lpcorr);
    const double exp_ovun8 = p_ovun7 * std::exp(p_ovun8 * sum_ovun2);
    const double inv_exp_ovun2n = 1.0 / (1.0 + exp_ovun2n);
    const double inv_exp_ovun8 = 1.0 / (1.0 + exp_ovun8);
    double e_un = 0.0;
    fc[4 * (size_t)i + 3] += CEover3;
    const bool under = numbonds > 0 || enobonds;
    if (under) e_un = -p_ovun5 * (1.0 - exp_ovun6) * inv_exp_ovun2n * inv_exp_ovun8;
    const double CEunder1 =
        inv_exp_ovun2n * (p_ovun5 * p_ovun6 * exp_ovun6 * inv_exp_ovun8 + p_ovun2 * e_un * exp_ovun2n);
    const double CEunder2 = -e_un * p_ovun8 * exp_ovun8 * inv_exp_ovun8;
    const double CEunder3 = CEunder1 * (1.0 - dfvl * S->dDelta_lp[i] * inv_exp_ovun1);
    const double CEunder4 =
        CEunder1 * (dfvl * S->Delta_lp_temp[i]) * p_ovun4 * exp_ovun1 * SQR(inv_exp_ovun1) + CEunder2;
    if (under) {
      e_un_sum += e_un;
      fc[4 * (size_t)i + 3] += CEunder3;
      fc[4 * (size_t)i + 3] += CElp;
      e_lp += p_lp2 * S->Delta_lp[i] * inv_expvd2;
    }
    for (int pj = S->bindex[i]; pj < S->bend[i]; ++pj) {
      const int j = S->bonds[pj].nbr;
      const int tj = S->packed[j].type;
      if (tj < 0) continue;
      const TBP &tw = S->tbp[ti][tj];
      const double pi_sum = S->BOpi[2 * (size_t)pj] + S->BOpi[2 * (size_t)pj + 1];
      S->Cdbo[pj] += CEover1 * tw.p_ovun1 * tw.De_s;
      const double ftmp2 = (1.0 - dfvl * S->dDelta_lp[j]) * pi_sum;
      fc[4 * (size_t)j + 3] += CEover4 * ftmp2;
      const double ftmp = CEover4 * (S->Delta[j] - dfvl * S->Delta_lp_temp[j]);
      S->Cdbopi[pj] += ftmp;
      S->Cdbopi2[pj] += ftmp;
      fc[4 * (size_t)j + 3] += CEunder4 * ftmp2;
      const double fu = CEunder4 * (S->Delta[j] - dfvl * S->Delta_lp_temp[j]);
      S->Cdbopi[pj] += fu;
      S->Cdbopi2[pj] += fu;
      if (!owns_bond(S, i, j)) continue;
      const SBP &sj = S->sbp[tj];
      const double BO_s = S->bo_data[pj].BO_s;
      const double pow_BOs_be2 = BO_s == 0.0 ? 0.0 : std::pow(BO_s, tw.p_be2);
      const double exp_be12 = std::exp(tw.p_be1 * (1.0 - pow_BOs_be2));
      const double CEbo = -tw.De_s * exp_be12 * (1.0 - tw.p_be1 * tw.p_be2 * pow_BOs_be2);
      ebond += -tw.De_s * BO_s * exp_be12 - tw.De_p * S->BOpi[2 * (size_t)pj] - tw.De_pp * S->BOpi[2 * (size_t)pj + 1];
      S->Cdbo[pj] += CEbo;
      S->Cdbopi[pj] -= (CEbo + tw.De_p);
      S->Cdbopi2[pj] -= (CEbo + tw.De_pp);
      const double BOij = S->BO[pj];
      if (BOij >= 1.00 && (gp37 == 2 || (si.mass == 12.0000 && sj.mass == 15.9990) ||
                           (sj.mass == 12.0000 && si.mass == 15.9990))) {
        const double exphu = std::exp(-gp7 * SQR(BOij - 2.50));
        const double exphua1 = std::exp(-gp3 * (S->bo_dboc[2 * (size_t)i] - BOij));
        const double exphub1 = std::exp(-gp3 * (S->bo_dboc[2 * (size_t)j] - BOij));
        const double exphuov = std::exp(gp4 * (S->Delta[i] + S->Delta[j]));
        const double hulpov = 1.0 / (1.0 + 25.0 * exphuov);
        estriph += gp10 * exphu * hulpov * (exphua1 + exphub1);
        const double decobdbo = gp10 * exphu * hulpov * (exphua1 + exphub1) * (gp3 - 2.0 * gp7 * (BOij - 2.50));
        const double decobdboua =
            -gp10 * exphu * hulpov * (gp3 * exphua1 + 25.0 * gp4 * exphuov * hulpov * (exphua1 + exphub1));
        const double decobdboub =
            -gp10 * exphu * hulpov * (gp3 * exphub1 + 25.0 * gp4 * exphuov * hulpov * (exphua1 + exphub1));
        S->Cdbo[pj] += decobdbo;
        fc[4 * (size_t)i + 3] += decobdboua;
        fc[4 * (size_t)j + 3] += decobdboub;
      }
    }
  }
  out[0] = ebond + estriph;
  out[1] = e_lp + e_lph;
  out[2] = e_ov;
  out[3] = e_un_sum;
  out[4] = e_pol;
  return 0;
}

// Hydrogen_Bonds (reaxc_forces_sunway.cpp:123-129) = Hydrogen_Bonds_C
// (reaxc_hydrogen_bonds_sw64.c:26-283), its host loop (the #else branch
// :145-253; the CPE form reaxc_hydrogen_bonds_cpe.h evaluates the same
// terms): for every local H atom j (p_hbond 1), every bond j-i to an
// acceptor (p_hbond 2) with BO >= HB_THRESHOLD = 0.01 (reaxc_defs_sunway.h:91)
// and every hydrogen-bond partner k of j with orig_id(k) != orig_id(i) and
// r0_hb(i, j, k) > 0: the energy, CEhb1 into Cdbo of j's slot, and the
// angle and distance terms into fCdDelta of i, j and k.  The angle through
// Calculate_Theta / Calculate_dCos_Theta (:336-372) with libm acos and pow
// where the reference calls xacos and p_powd.  Departure: an atom of type
// < 0 is skipped before its p_hbond is read.  Returns e_hb.
double orc_hydrogen_bonds(orc_sys *S) {
  const int n = S->n;
  const int nt = (int)S->sbp.size();
  const double HB_THRESHOLD = 1e-2;
  double *fc = S->fCdDelta.data();
  double e_hb_total = 0.0;
  std::vector<int> hblist;
  for (int j = 0; j < n; j++) {
    const int tj = S->packed[j].type;
    if (tj < 0 || S->sbp[tj].p_hbond != 1) continue;
    hblist.clear();
    for (int pi = S->bindex[j]; pi < S->bend[j]; ++pi) {
      const int i = S->bonds[pi].nbr;
      const int ti = S->packed[i].type;
      if (ti < 0) continue;
      if (S->sbp[ti].p_hbond == 2 && S->BO[pi] >= HB_THRESHOLD) hblist.push_back(pi);
    }
    const int H = S->Hindex[j];
    if (H < 0) continue;
    for (int pk = S->hindex[H]; pk < S->hend[H]; ++pk) {
      const int k = S->hbonds[pk].nbr;
      const int tk = S->packed[k].type;
      if (tk < 0) continue;
      const swr_far_neighbor_t &fe = S->far[S->hbond_far[pk]];
      const double r_jk = fe.d;
      double dvec_jk[3];
      for (int c = 0; c < 3; c++) dvec_jk[c] = S->hbonds[pk].scl * fe.dvec[c];
      for (const int pi : hblist) {
        const swr_bond_t &bij = S->bonds[pi];
        const int i = bij.nbr;
        if (S->packed[i].orig_id == S->packed[k].orig_id) continue;
        const int ti = S->packed[i].type;
        const HBP &hbp = S->hbp[((size_t)ti * nt + tj) * nt + tk];
        if (hbp.r0_hb <= 0.0) continue;
        // Calculate_Theta
        const double *dvec_ji = bij.dvec;
        const double d_ji = bij.d;
        const double dot = dvec_ji[0] * dvec_jk[0] + dvec_ji[1] * dvec_jk[1] + dvec_ji[2] * dvec_jk[2];
        double cos_theta = dot / (d_ji * r_jk);
        if (cos_theta > 1.) cos_theta = 1.0;
        if (cos_theta < -1.) cos_theta = -1.0;
        const double theta = std::acos(cos_theta);
        // Calculate_dCos_Theta
        const double sqr_d_ji = SQR(d_ji), sqr_d_jk = SQR(r_jk);
        const double inv_dists = 1.0 / (d_ji * r_jk);
        const double inv_dists3 = std::pow(inv_dists, 3.0);
        const double Cdot_inv3 = dot * inv_dists3;
        double di[3], dj[3], dk[3];
        for (int c = 0; c < 3; c++) {
          di[c] = dvec_jk[c] * inv_dists - Cdot_inv3 * sqr_d_jk * dvec_ji[c];
          dj[c] = -(dvec_jk[c] + dvec_ji[c]) * inv_dists + Cdot_inv3 * (sqr_d_jk * dvec_ji[c] + sqr_d_ji * dvec_jk[c]);
          dk[c] = dvec_ji[c] * inv_dists - Cdot_inv3 * sqr_d_ji * dvec_jk[c];
        }
        const double sin_theta2 = std::sin(theta / 2.0);
        double sin_xhz4 = SQR(sin_theta2);
        sin_xhz4 *= sin_xhz4;
        const double cos_xhz1 = (1.0 - cos_theta);
        const double exp_hb2 = std::exp(-hbp.p_hb2 * S->BO[pi]);
        const double exp_hb3 = std::exp(-hbp.p_hb3 * (hbp.r0_hb / r_jk + r_jk / hbp.r0_hb - 2.0));
        const double e_hb = hbp.p_hb1 * (1.0 - exp_hb2) * exp_hb3 * sin_xhz4;
        e_hb_total += e_hb;
        const double CEhb1 = hbp.p_hb1 * hbp.p_hb2 * exp_hb2 * exp_hb3 * sin_xhz4;
        const double CEhb2 = -hbp.p_hb1 / 2.0 * (1.0 - exp_hb2) * exp_hb3 * cos_xhz1;
        const double CEhb3 = -hbp.p_hb3 * (-hbp.r0_hb / SQR(r_jk) + 1.0 / hbp.r0_hb) * e_hb;
        S->Cdbo[pi] += CEhb1;
        for (int c = 0; c < 3; c++) fc[4 * (size_t)i + c] += CEhb2 * di[c];
        for (int c = 0; c < 3; c++) fc[4 * (size_t)j + c] += CEhb2 * dj[c];
        for (int c = 0; c < 3; c++) fc[4 * (size_t)k + c] += CEhb2 * dk[c];
        for (int c = 0; c < 3; c++) fc[4 * (size_t)j + c] += -CEhb3 / r_jk * dvec_jk[c];
        for (int c = 0; c < 3; c++) fc[4 * (size_t)k + c] += CEhb3 / r_jk * dvec_jk[c];
      }
    }
  }
  return e_hb_total;
}

// Calculate_Omega (reaxc_torsion_angles_sw64.c:21-104): the dihedral i-j-k-l
// from the bond vectors, and the derivatives of cos(omega).
double orc_omega(const double *dvec_ij, double r_ij, const double *dvec_jk, double r_jk, const double *dvec_kl,
                 double r_kl, const double *dvec_li, double r_li, double theta_ijk, const double *ijk_di,
                 const double *ijk_dj, const double *ijk_dk, double theta_jkl, const double *jkl_di,
                 const double *jkl_dj, const double *jkl_dk, double *d_i, double *d_j, double *d_k, double *d_l) {
  const double MIN_SINE = 1e-10;
  auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  double sin_ijk = std::sin(theta_ijk), cos_ijk = std::cos(theta_ijk);
  double sin_jkl = std::sin(theta_jkl), cos_jkl = std::cos(theta_jkl);
  const double unnorm_cos_omega = -dot(dvec_ij, dvec_jk) * dot(dvec_jk, dvec_kl) + SQR(r_jk) * dot(dvec_ij, dvec_kl);
  const double cross[3] = {dvec_jk[1] * dvec_kl[2] - dvec_jk[2] * dvec_kl[1],
                           dvec_jk[2] * dvec_kl[0] - dvec_jk[0] * dvec_kl[2],
                           dvec_jk[0] * dvec_kl[1] - dvec_jk[1] * dvec_kl[0]};
  const double unnorm_sin_omega = -r_jk * dot(dvec_ij, cross);
  const double omega = std::atan2(unnorm_sin_omega, unnorm_cos_omega);
  const double htra = r_ij + cos_ijk * (r_kl * cos_jkl - r_jk);
  const double htrb = r_jk - r_ij * cos_ijk - r_kl * cos_jkl;
  const double htrc = r_kl + cos_jkl * (r_ij * cos_ijk - r_jk);
  const double hthd = r_ij * sin_ijk * (r_jk - r_kl * cos_jkl);
  const double hthe = r_kl * sin_jkl * (r_jk - r_ij * cos_ijk);
  const double hnra = r_kl * sin_ijk * sin_jkl;
  const double hnrc = r_ij * sin_ijk * sin_jkl;
  const double hnhd = r_ij * r_kl * cos_ijk * sin_jkl;
  const double hnhe = r_ij * r_kl * sin_ijk * cos_jkl;
  double poem = 2.0 * r_ij * r_kl * sin_ijk * sin_jkl;
  if (poem < 1e-20) poem = 1e-20;
  const double tel = SQR(r_ij) + SQR(r_jk) + SQR(r_kl) - SQR(r_li) -
                     2.0 * (r_ij * r_jk * cos_ijk - r_ij * r_kl * cos_ijk * cos_jkl + r_jk * r_kl * cos_jkl);
  double arg = tel / poem;
  if (arg > 1.0) arg = 1.0;
  if (arg < -1.0) arg = -1.0;
  if (sin_ijk >= 0 && sin_ijk <= MIN_SINE) sin_ijk = MIN_SINE;
  else if (sin_ijk <= 0 && sin_ijk >= -MIN_SINE) sin_ijk = -MIN_SINE;
  if (sin_jkl >= 0 && sin_jkl <= MIN_SINE) sin_jkl = MIN_SINE;
  else if (sin_jkl <= 0 && sin_jkl >= -MIN_SINE) sin_jkl = -MIN_SINE;
  const double s = 2.0 / poem;
  for (int c = 0; c < 3; c++) {
    d_i[c] = s * ((htra - arg * hnra) / r_ij * dvec_ij[c] - dvec_li[c] - (hthd - arg * hnhd) / sin_ijk * ijk_dk[c]);
    d_j[c] = s * (-(htra - arg * hnra) / r_ij * dvec_ij[c] - htrb / r_jk * dvec_jk[c] -
                  (hthd - arg * hnhd) / sin_ijk * ijk_dj[c] - (hthe - arg * hnhe) / sin_jkl * jkl_di[c]);
    d_k[c] = s * (-(htrc - arg * hnrc) / r_kl * dvec_kl[c] + htrb / r_jk * dvec_jk[c] -
                  (hthd - arg * hnhd) / sin_ijk * ijk_di[c] - (hthe - arg * hnhe) / sin_jkl * jkl_dj[c]);
    d_l[c] = s * ((htrc - arg * hnrc) / r_kl * dvec_kl[c] + dvec_li[c] - (hthe - arg * hnhe) / sin_jkl * jkl_dk[c]);
  }
  return omega;
}

// Torsion part of Merge_Torsion_Valence_Angles, host form
// (reaxc_torsion_angles_sw64.c:812-1408): for every local j and bond j-k
// (pk) with BO > thb_cut (k with bonds), taken once per physical bond
// (orig_id(j) < orig_id(k), or equal with k after j by z, y, x), every other
// bond j-i (ph) with BO > thb_cut and every bond k-l (pw) other than k-j with
// i != l, a torsion term for (type i, j, k, l), BO_kl > thb_cut and
// BO_ij BO_jk BO_kl > thb_cut: the torsion and conjugation energies and their
// coefficients into Cdbopi of j-k, Cdbo of j-i, j-k and k-l, CdDelta of j
// and k, and the angle and dihedral derivatives into fCdDelta of i, j, k, l.
// theta by libm acos where the reference calls xacos; MIN_SINE 1e-10 (:12).
// Returns e_tor, e_con in out[0..1].
int orc_torsion_angles(orc_sys *S, double *out) {
  const int n = S->n;
  const int nt = (int)S->sbp.size();
  const double MIN_SINE = 1e-10;
  auto g = [&](size_t k) { return k < S->gp.size() ? S->gp[k] : 0.0; };
  const double p_tor2 = g(23), p_tor3 = g(24), p_tor4 = g(25), p_cot2 = g(27);
  const double thb_cut = S->thb_cut;
  double *fc = S->fCdDelta.data();
  double e_tor_t = 0.0, e_con_t = 0.0;
  auto theta_of = [](const swr_bond_t &a, const swr_bond_t &b, double *di, double *dj, double *dk) {
    const double dot = a.dvec[0] * b.dvec[0] + a.dvec[1] * b.dvec[1] + a.dvec[2] * b.dvec[2];
    double c = dot / (a.d * b.d);
    if (c > 1.) c = 1.0;
    if (c < -1.) c = -1.0;
    const double sa = SQR(a.d), sb = SQR(b.d);
    const double inv = 1.0 / (a.d * b.d);
    const double ci3 = dot * std::pow(inv, 3.0);
    for (int q = 0; q < 3; q++) {
      di[q] = b.dvec[q] * inv - ci3 * sb * a.dvec[q];
      dj[q] = -(b.dvec[q] + a.dvec[q]) * inv + ci3 * (sb * a.dvec[q] + sa * b.dvec[q]);
      dk[q] = a.dvec[q] * inv - ci3 * sa * b.dvec[q];
    }
    return std::acos(c);
  };
  auto add4 = [&](int a, double s, const double *v) {
    for (int q = 0; q < 3; q++) fc[4 * (size_t)a + q] += s * v[q];
  };
  for (int j = 0; j < n; ++j) {
    const int type_j = S->packed[j].type;
    const int start_j = S->bindex[j], end_j = S->bend[j];
    if (type_j < 0 || end_j <= start_j) continue;
    const double Delta_j = S->bo_dboc[2 * (size_t)j + 1];
    for (int pk = start_j; pk < end_j; ++pk) {
      const swr_bond_t &bjk = S->bonds[pk];
      const int k = bjk.nbr;
      const double BOA_jk = S->BO[pk] - thb_cut;
      if (!(BOA_jk > 0.0)) continue;
      const int start_k = S->bindex[k], end_k = S->bend[k];
      if (end_k <= start_k) continue;
      const swr_atom_pack_t &aj = S->packed[j], &ak = S->packed[k];
      if (aj.orig_id > ak.orig_id) continue;
      if (aj.orig_id == ak.orig_id) {
        if (ak.x[2] < aj.x[2]) continue;
        if (ak.x[2] == aj.x[2] && ak.x[1] < aj.x[1]) continue;
        if (ak.x[2] == aj.x[2] && ak.x[1] == aj.x[1] && ak.x[0] < aj.x[0]) continue;
      }
      const int pj = bjk.sym_index;
      const swr_bond_t &bkj = S->bonds[pj];
      const int type_k = ak.type;
      const double Delta_k = S->bo_dboc[2 * (size_t)k + 1];
      const double r_jk = bjk.d;
      const double exp_tor2_jk = std::exp(-p_tor2 * BOA_jk);
      const double exp_cot2_jk = std::exp(-p_cot2 * SQR(BOA_jk - 1.5));
      const double exp_tor3_DjDk = std::exp(-p_tor3 * (Delta_j + Delta_k));
      const double exp_tor4_DjDk = std::exp(p_tor4 * (Delta_j + Delta_k));
      const double exp_tor34_inv = 1.0 / (1.0 + exp_tor3_DjDk + exp_tor4_DjDk);
      const double f11_DjDk = (2.0 + exp_tor3_DjDk) * exp_tor34_inv;
      for (int ph = start_j; ph < end_j; ++ph) {
        if (ph == pk) continue;
        const swr_bond_t &bij = S->bonds[ph];
        if (!(S->BO[ph] > thb_cut)) continue;
        const int i = bij.nbr;
        const int type_i = S->packed[i].type;
        double ijk_di[3], ijk_dj[3], ijk_dk[3];   // of k, j, i
        const double theta_ijk = theta_of(bjk, bij, ijk_di, ijk_dj, ijk_dk);
        const double r_ij = bij.d;
        const double BOA_ij = S->BO[ph] - thb_cut;
        const double sin_ijk = std::sin(theta_ijk), cos_ijk = std::cos(theta_ijk);
        double tan_ijk_i;
        if (sin_ijk >= 0 && sin_ijk <= MIN_SINE) tan_ijk_i = cos_ijk / MIN_SINE;
        else if (sin_ijk <= 0 && sin_ijk >= -MIN_SINE) tan_ijk_i = cos_ijk / -MIN_SINE;
        else tan_ijk_i = cos_ijk / sin_ijk;
        const double exp_tor2_ij = std::exp(-p_tor2 * BOA_ij);
        const double exp_cot2_ij = std::exp(-p_cot2 * SQR(BOA_ij - 1.5));
        for (int pw = start_k; pw < end_k; ++pw) {
          if (pw == pj) continue;
          const swr_bond_t &bkl = S->bonds[pw];
          const int l = bkl.nbr;
          const int type_l = S->packed[l].type;
          double jkl_di[3], jkl_dj[3], jkl_dk[3];   // of j, k, l
          const double theta_jkl = theta_of(bkj, bkl, jkl_di, jkl_dj, jkl_dk);
          if (type_i < 0 || type_l < 0) continue;
          const FBP &fb = S->fbp[(((size_t)type_i * nt + type_j) * nt + type_k) * nt + type_l];
          if (!(i != l && fb.cnt && S->BO[pw] > thb_cut && S->BO[ph] * S->BO[pk] * S->BO[pw] > thb_cut)) continue;
          const double r_kl = bkl.d;
          const double BOA_kl = S->BO[pw] - thb_cut;
          const double sin_jkl = std::sin(theta_jkl), cos_jkl = std::cos(theta_jkl);
          double tan_jkl_i;
          if (sin_jkl >= 0 && sin_jkl <= MIN_SINE) tan_jkl_i = cos_jkl / MIN_SINE;
          else if (sin_jkl <= 0 && sin_jkl >= -MIN_SINE) tan_jkl_i = cos_jkl / -MIN_SINE;
          else tan_jkl_i = cos_jkl / sin_jkl;
          double dvec_li[3];
          for (int q = 0; q < 3; q++) dvec_li[q] = S->packed[i].x[q] - S->packed[l].x[q];
          const double r_li = std::sqrt(SQR(dvec_li[0]) + SQR(dvec_li[1]) + SQR(dvec_li[2]));
          double dw_i[3], dw_j[3], dw_k[3], dw_l[3];
          const double omega = orc_omega(bij.dvec, r_ij, bjk.dvec, r_jk, bkl.dvec, r_kl, dvec_li, r_li, theta_ijk,
                                         ijk_di, ijk_dj, ijk_dk, theta_jkl, jkl_di, jkl_dj, jkl_dk, dw_i, dw_j,
                                         dw_k, dw_l);
          const double cos_omega = std::cos(omega), cos2omega = std::cos(2. * omega),
                       cos3omega = std::cos(3. * omega);
          const double BOpi_jk = S->BOpi[2 * (size_t)pk];
          const double exp_tor1 = std::exp(fb.p_tor1 * SQR(2.0 - BOpi_jk - f11_DjDk));
          const double exp_tor2_kl = std::exp(-p_tor2 * BOA_kl);
          const double exp_cot2_kl = std::exp(-p_cot2 * SQR(BOA_kl - 1.5));
          const double fn10 = (1.0 - exp_tor2_ij) * (1.0 - exp_tor2_jk) * (1.0 - exp_tor2_kl);
          const double CV = 0.5 * (fb.V1 * (1.0 + cos_omega) + fb.V2 * exp_tor1 * (1.0 - cos2omega) +
                                   fb.V3 * (1.0 + cos3omega));
          const double e_tor = fn10 * sin_ijk * sin_jkl * CV;
          e_tor_t += e_tor;
          const double dfn11 = (-p_tor3 * exp_tor3_DjDk +
                                (p_tor3 * exp_tor3_DjDk - p_tor4 * exp_tor4_DjDk) * (2.0 + exp_tor3_DjDk) *
                                    exp_tor34_inv) *
                               exp_tor34_inv;
          const double CEtors1 = sin_ijk * sin_jkl * CV;
          const double CEtors2 = -fn10 * 2.0 * fb.p_tor1 * fb.V2 * exp_tor1 * (2.0 - BOpi_jk - f11_DjDk) *
                                 (1.0 - SQR(cos_omega)) * sin_ijk * sin_jkl;
          const double CEtors3 = CEtors2 * dfn11;
          const double CEtors4 = CEtors1 * p_tor2 * exp_tor2_ij * (1.0 - exp_tor2_jk) * (1.0 - exp_tor2_kl);
          const double CEtors5 = CEtors1 * p_tor2 * (1.0 - exp_tor2_ij) * exp_tor2_jk * (1.0 - exp_tor2_kl);
          const double CEtors6 = CEtors1 * p_tor2 * (1.0 - exp_tor2_ij) * (1.0 - exp_tor2_jk) * exp_tor2_kl;
          const double cmn = -fn10 * CV;
          const double CEtors7 = cmn * sin_jkl * tan_ijk_i;
          const double CEtors8 = cmn * sin_ijk * tan_jkl_i;
          const double CEtors9 = fn10 * sin_ijk * sin_jkl *
                                 (0.5 * fb.V1 - 2.0 * fb.V2 * exp_tor1 * cos_omega +
                                  1.5 * fb.V3 * (cos2omega + 2.0 * SQR(cos_omega)));
          const double fn12 = exp_cot2_ij * exp_cot2_jk * exp_cot2_kl;
          const double e_con = fb.p_cot1 * fn12 * (1.0 + (SQR(cos_omega) - 1.0) * sin_ijk * sin_jkl);
          e_con_t += e_con;
          const double Cconj = -2.0 * fn12 * fb.p_cot1 * p_cot2 * (1.0 + (SQR(cos_omega) - 1.0) * sin_ijk * sin_jkl);
          const double CEconj1 = Cconj * (BOA_ij - 1.5e0);
          const double CEconj2 = Cconj * (BOA_jk - 1.5e0);
          const double CEconj3 = Cconj * (BOA_kl - 1.5e0);
          const double CEconj4 = -fb.p_cot1 * fn12 * (SQR(cos_omega) - 1.0) * sin_jkl * tan_ijk_i;
          const double CEconj5 = -fb.p_cot1 * fn12 * (SQR(cos_omega) - 1.0) * sin_ijk * tan_jkl_i;
          const double CEconj6 = 2.0 * fb.p_cot1 * fn12 * cos_omega * sin_ijk * sin_jkl;
          S->Cdbopi[pk] += CEtors2;
          fc[4 * (size_t)j + 3] += CEtors3;
          fc[4 * (size_t)k + 3] += CEtors3;
          S->Cdbo[ph] += (CEtors4 + CEconj1);
          S->Cdbo[pk] += (CEtors5 + CEconj2);
          S->Cdbo[pw] += (CEtors6 + CEconj3);
          add4(i, CEtors7 + CEconj4, ijk_dk);
          add4(j, CEtors7 + CEconj4, ijk_dj);
          add4(k, CEtors7 + CEconj4, ijk_di);
          add4(j, CEtors8 + CEconj5, jkl_di);
          add4(k, CEtors8 + CEconj5, jkl_dj);
          add4(l, CEtors8 + CEconj5, jkl_dk);
          add4(i, CEtors9 + CEconj6, dw_i);
          add4(j, CEtors9 + CEconj6, dw_j);
          add4(k, CEtors9 + CEconj6, dw_k);
          add4(l, CEtors9 + CEconj6, dw_l);
        }
      }
    }
  }
  out[0] = e_tor_t;
  out[1] = e_con_t;
  return 0;
}

// Valence part of Merge_Torsion_Valence_Angles, host form
// (reaxc_torsion_angles_sw64.c:812-1117): for every atom j < n with bonds,
// SBO and its derivatives from j's bonds (:833-884), then for every bond
// pair (pk, ph > pk) with BO > thb_cut each and product > thb_cutsq and
// every parameter set of thbp[type k][type j][type h] with |p_val1| > 0.001
// (:957-1115): angle, penalty and coalition energies and their coefficients,
// added to Cdbo/Cdbopi/Cdbopi2, CdDelta and fCdDelta in the reference's
// order.  DEG2RAD with constPI = 3.14159265 (reaxc_defs_sunway.h:55,61);
// theta by libm acos where the reference calls xacos.  Returns e_ang,
// e_pen, e_coa in out[0..2].
int orc_valence_angles(orc_sys *S, double *out) {
  const int n = S->n;
  const int nt = (int)S->sbp.size();
  const double constPI = 3.14159265;
  auto DEG2RAD = [&](double a) { return a * constPI / 180.0; };
  const std::vector<double> &gp = S->gp;
  auto g = [&](size_t k) { return k < gp.size() ? gp[k] : 0.0; };
  const double p_val6 = g(14), p_val8 = g(33), p_val9 = g(16), p_val10 = g(17);
  const double p_pen2 = g(19), p_pen3 = g(20), p_pen4 = g(21);
  const double p_coa2 = g(2), p_coa3 = g(38), p_coa4 = g(30);
  const double thb_cut = S->thb_cut, thb_cutsq = S->thb_cutsq;
  double *fc = S->fCdDelta.data();
  double e_ang_t = 0.0, e_pen_t = 0.0, e_coa_t = 0.0;
  for (int j = 0; j < n; ++j) {
    const int type_j = S->packed[j].type;
    const int start_j = S->bindex[j], end_j = S->bend[j];
    if (type_j < 0 || end_j <= start_j) continue;
    const double Dboc_j = S->bo_dboc[2 * (size_t)j + 1];
    const double p_val3 = S->sbp[type_j].p_val3, p_val5 = S->sbp[type_j].p_val5;
    double SBOp = 0, prod_SBO = 1;
    for (int t = start_j; t < end_j; ++t) {
      SBOp += S->BOpi[2 * (size_t)t] + S->BOpi[2 * (size_t)t + 1];
      double temp = SQR(S->BO[t]);
      temp *= temp;
      temp *= temp;
      prod_SBO *= std::exp(-temp);
    }
    double vlpadj, dSBO2;
    if (S->vlpex[j] >= 0) {
      vlpadj = 0;
      dSBO2 = prod_SBO - 1;
    } else {
      vlpadj = S->nlp[j];
      dSBO2 = (prod_SBO - 1) * (1 - p_val8 * S->dDelta_lp[j]);
    }
    const double SBO = SBOp + (1 - prod_SBO) * (-Dboc_j - p_val8 * vlpadj);
    const double dSBO1 = -8 * prod_SBO * (Dboc_j + p_val8 * vlpadj);
    double SBO2, CSBO2;
    if (SBO <= 0) {
      SBO2 = 0; CSBO2 = 0;
    } else if (SBO > 0 && SBO <= 1) {
      SBO2 = std::pow(SBO, p_val9); CSBO2 = p_val9 * std::pow(SBO, p_val9 - 1);
    } else if (SBO > 1 && SBO < 2) {
      SBO2 = 2 - std::pow(2 - SBO, p_val9); CSBO2 = p_val9 * std::pow(2 - SBO, p_val9 - 1);
    } else {
      SBO2 = 2; CSBO2 = 0;
    }
    const double expval6 = std::exp(p_val6 * Dboc_j);
    for (int pk = start_j; pk < end_j; ++pk) {
      const swr_bond_t &bjk = S->bonds[pk];
      const int k = bjk.nbr;
      const double BOA_jk = S->BO[pk] - thb_cut;
      if (!(BOA_jk > 0.0)) continue;
      if (S->bend[k] <= S->bindex[k]) continue;
      const int type_k = S->packed[k].type;
      for (int ph = start_j; ph < end_j; ++ph) {
        if (ph == pk) continue;
        const swr_bond_t &bhj = S->bonds[ph];
        const double BOA_hj = S->BO[ph] - thb_cut;
        const int h = bhj.nbr;
        const int type_h = S->packed[h].type;
        // Calculate_Theta / Calculate_dCos_Theta with the j-k bond first
        const double dot = bjk.dvec[0] * bhj.dvec[0] + bjk.dvec[1] * bhj.dvec[1] + bjk.dvec[2] * bhj.dvec[2];
        double cos_theta = dot / (bjk.d * bhj.d);
        if (cos_theta > 1.) cos_theta = 1.0;
        if (cos_theta < -1.) cos_theta = -1.0;
        const double theta = std::acos(cos_theta);
        const double sqr_d_ji = SQR(bjk.d), sqr_d_jk = SQR(bhj.d);
        const double inv_dists = 1.0 / (bjk.d * bhj.d);
        const double Cdot_inv3 = dot * std::pow(inv_dists, 3.0);
        double dcos_di[3], dcos_dj[3], dcos_dk[3];
        for (int c = 0; c < 3; ++c) {
          dcos_di[c] = bhj.dvec[c] * inv_dists - Cdot_inv3 * sqr_d_jk * bjk.dvec[c];
          dcos_dj[c] = -(bhj.dvec[c] + bjk.dvec[c]) * inv_dists +
                       Cdot_inv3 * (sqr_d_jk * bjk.dvec[c] + sqr_d_ji * bhj.dvec[c]);
          dcos_dk[c] = bjk.dvec[c] * inv_dists - Cdot_inv3 * sqr_d_ji * bhj.dvec[c];
        }
        double sin_theta = std::sin(theta);
        if (sin_theta < 1.0e-5) sin_theta = 1.0e-5;
        if (!(ph > pk && S->BO[pk] > thb_cut && S->BO[ph] > thb_cut && S->BO[pk] * S->BO[ph] > thb_cutsq))
          continue;
        if (type_k < 0 || type_h < 0) continue;
        const size_t tri = ((size_t)type_k * nt + type_j) * nt + type_h;
        for (int cnt = 0; cnt < S->thb_cnt[tri] && cnt < MAX_3BODY_PARAM; ++cnt) {
          const THBP &q = S->thbp[tri * MAX_3BODY_PARAM + cnt];
          if (!(std::fabs(q.p_val1) > 0.001)) continue;
          const double p_val1 = q.p_val1, p_val2 = q.p_val2, p_val4 = q.p_val4, p_val7 = q.p_val7;
          const double theta_00 = q.theta_00;
          const double exp3jk = std::exp(-p_val3 * std::pow(BOA_jk, p_val4));
          const double f7_jk = 1.0 - exp3jk;
          const double Cf7jk = p_val3 * p_val4 * std::pow(BOA_jk, p_val4 - 1.0) * exp3jk;
          const double exp3hj = std::exp(-p_val3 * std::pow(BOA_hj, p_val4));
          const double f7_hj = 1.0 - exp3hj;
          const double Cf7hj = p_val3 * p_val4 * std::pow(BOA_hj, p_val4 - 1.0) * exp3hj;
          const double expval7 = std::exp(-p_val7 * Dboc_j);
          const double trm8 = 1.0 + expval6 + expval7;
          const double f8_Dj = p_val5 - ((p_val5 - 1.0) * (2.0 + expval6) / trm8);
          const double Cf8j = ((1.0 - p_val5) / SQR(trm8)) *
                              (p_val6 * expval6 * trm8 - (2.0 + expval6) * (p_val6 * expval6 - p_val7 * expval7));
          const double theta_0 = DEG2RAD(180.0 - theta_00 * (1.0 - std::exp(-p_val10 * (2.0 - SBO2))));
          const double expval2theta = std::exp(-p_val2 * SQR(theta_0 - theta));
          const double expval12theta = p_val1 >= 0 ? p_val1 * (1.0 - expval2theta) : p_val1 * -expval2theta;
          const double CEval1 = Cf7jk * f7_hj * f8_Dj * expval12theta;
          const double CEval2 = Cf7hj * f7_jk * f8_Dj * expval12theta;
          const double CEval3 = Cf8j * f7_hj * f7_jk * expval12theta;
          const double CEval4 = -2.0 * p_val1 * p_val2 * f7_jk * f7_hj * f8_Dj * expval2theta * (theta_0 - theta);
          const double Ctheta_0 = p_val10 * DEG2RAD(theta_00) * std::exp(-p_val10 * (2.0 - SBO2));
          const double CEval5 = -CEval4 * Ctheta_0 * CSBO2;
          const double CEval6 = CEval5 * dSBO1;
          const double CEval7 = CEval5 * dSBO2;
          const double CEval8 = -CEval4 / sin_theta;
          const double e_ang = f7_jk * f7_hj * f8_Dj * expval12theta;
          e_ang_t += e_ang;
          const double p_pen1 = q.p_pen1;
          const double exp_pen2jk = std::exp(-p_pen2 * SQR(BOA_jk - 2.0));
          const double exp_pen2hj = std::exp(-p_pen2 * SQR(BOA_hj - 2.0));
          const double exp_pen3 = std::exp(-p_pen3 * S->Delta[j]);
          const double exp_pen4 = std::exp(p_pen4 * S->Delta[j]);
          const double trm_pen34 = 1.0 + exp_pen3 + exp_pen4;
          const double f9_Dj = (2.0 + exp_pen3) / trm_pen34;
          const double Cf9j = (-p_pen3 * exp_pen3 * trm_pen34 - (2.0 + exp_pen3) * (-p_pen3 * exp_pen3 + p_pen4 * exp_pen4)) /
                              SQR(trm_pen34);
          const double e_pen = p_pen1 * f9_Dj * exp_pen2jk * exp_pen2hj;
          e_pen_t += e_pen;
          const double CEpen1 = e_pen * Cf9j / f9_Dj;
          const double temp = -2.0 * p_pen2 * e_pen;
          const double CEpen2 = temp * (BOA_jk - 2.0), CEpen3 = temp * (BOA_hj - 2.0);
          const double p_coa1 = q.p_coa1;
          const double exp_coa2 = std::exp(p_coa2 * S->Delta_val[j]);
          const double bk0 = S->bo_dboc[2 * (size_t)k], bh0 = S->bo_dboc[2 * (size_t)h];
          const double e_coa = p_coa1 / (1. + exp_coa2) * std::exp(-p_coa3 * SQR(bk0 - BOA_jk)) *
                               std::exp(-p_coa3 * SQR(bh0 - BOA_hj)) * std::exp(-p_coa4 * SQR(BOA_jk - 1.5)) *
                               std::exp(-p_coa4 * SQR(BOA_hj - 1.5));
          e_coa_t += e_coa;
          const double CEcoa1 = -2 * p_coa4 * (BOA_jk - 1.5) * e_coa;
          const double CEcoa2 = -2 * p_coa4 * (BOA_hj - 1.5) * e_coa;
          const double CEcoa3 = -p_coa2 * exp_coa2 * e_coa / (1 + exp_coa2);
          const double CEcoa4 = -2 * p_coa3 * (bk0 - BOA_jk) * e_coa;
          const double CEcoa5 = -2 * p_coa3 * (bh0 - BOA_hj) * e_coa;
          S->Cdbo[pk] += (CEval1 + CEpen2 + (CEcoa1 - CEcoa4));
          S->Cdbo[ph] += (CEval2 + CEpen3 + (CEcoa2 - CEcoa5));
          fc[4 * (size_t)j + 3] += ((CEval3 + CEval7) + CEpen1 + CEcoa3);
          fc[4 * (size_t)k + 3] += CEcoa4;
          fc[4 * (size_t)h + 3] += CEcoa5;
          for (int t = start_j; t < end_j; ++t) {
            const double bo = S->BO[t];
            const double c3 = bo * bo * bo;
            S->Cdbo[t] += CEval6 * (c3 * c3 * bo);
            S->Cdbopi[t] += CEval5;
            S->Cdbopi2[t] += CEval5;
          }
          for (int c = 0; c < 3; ++c) {
            fc[4 * (size_t)k + c] += CEval8 * dcos_di[c];
            fc[4 * (size_t)j + c] += CEval8 * dcos_dj[c];
            fc[4 * (size_t)h + c] += CEval8 * dcos_dk[c];
          }
        }
      }
    }
  }
  out[0] = e_ang_t;
  out[1] = e_pen_t;
  out[2] = e_coa_t;
  return 0;
}

// Add_All_dBond_to_Forces_C (reaxc_forces_sw64.c:417-589), its host loop
// (the #else branch :495-586; the CPE form :1370-1570 evaluates the same
// terms): over all N atoms and their bonds, the chain rule of the bond-order
// derivatives with the coefficients Cdbo/Cdbopi/Cdbopi2 of both halves of
// each bond and the CdDelta of both atoms; adds the atom's own terms to
// fCdDelta[i][0..2] and buf_c dBOp to each bonded neighbour.  fCdDelta[.][3]
// is only read.
int orc_dbond_forces(orc_sys *S) {
  const int N = S->N;
  double *fc = S->fCdDelta.data();
  std::vector<double> bufd;
  std::vector<int> bufj;
  for (int i = 0; i < N; ++i) {
    double buf_c = 0.0;
    bufd.clear();
    bufj.clear();
    const double *ds = &S->dDeltap_self[3 * (size_t)i];
    for (int pj = S->bindex[i]; pj < S->bend[i]; ++pj) {
      const swr_bond_t &b = S->bonds[pj];
      const int j = b.nbr, sym = b.sym_index;
      const swr_bond_order_t &bo = S->bo_data[pj];
      const double cdbo = S->Cdbo[pj] + S->Cdbo[sym];
      const double cdbopi = S->Cdbopi[pj] + S->Cdbopi[sym];
      const double cdbopi2 = S->Cdbopi2[pj] + S->Cdbopi2[sym];
      const double C1dbo = bo.C1dbo * cdbo, C2dbo = bo.C2dbo * cdbo;
      const double C1dbopi = bo.C1dbopi * cdbopi, C2dbopi = bo.C2dbopi * cdbopi, C3dbopi = bo.C3dbopi * cdbopi;
      const double C1dbopi2 = bo.C1dbopi2 * cdbopi2, C2dbopi2 = bo.C2dbopi2 * cdbopi2,
                   C3dbopi2 = bo.C3dbopi2 * cdbopi2;
      const double cdd = fc[4 * (size_t)i + 3] + fc[4 * (size_t)j + 3];
      const double C1dDelta = bo.C1dbo * cdd, C2dDelta = bo.C2dbo * cdd;
      double t[3];
      for (int k = 0; k < 3; k++) t[k] = C1dbo * bo.dBOp[k];
      for (int k = 0; k < 3; k++) t[k] += C2dbo * ds[k];
      for (int k = 0; k < 3; k++) t[k] += C1dDelta * bo.dBOp[k];
      for (int k = 0; k < 3; k++) t[k] += C2dDelta * ds[k];
      for (int k = 0; k < 3; k++) t[k] += C1dbopi * bo.dln_BOp_pi[k];
      for (int k = 0; k < 3; k++) t[k] += C2dbopi * bo.dBOp[k];
      for (int k = 0; k < 3; k++) t[k] += C3dbopi * ds[k];
      for (int k = 0; k < 3; k++) t[k] += C1dbopi2 * bo.dln_BOp_pi2[k];
      for (int k = 0; k < 3; k++) t[k] += C2dbopi2 * bo.dBOp[k];
      for (int k = 0; k < 3; k++) t[k] += C3dbopi2 * ds[k];
      for (int k = 0; k < 3; k++) fc[4 * (size_t)i + k] += t[k];
      buf_c += (-C2dbo - C2dDelta - C3dbopi - C3dbopi2);
      bufj.push_back(j);
      for (int k = 0; k < 3; k++) bufd.push_back(bo.dBOp[k]);
    }
    for (size_t m = 0; m < bufj.size(); m++)
      for (int k = 0; k < 3; k++) fc[4 * (size_t)bufj[m] + k] += buf_c * bufd[3 * m + k];
  }
  return 0;
}

int orc_step(orc_sys *S, int *its, int *itt) {
  int rc = orc_qeq_pre_force(S, its, itt);
  if (rc) return rc;
  orc_write_reax_atoms(S);
  if (!S->estimated) {
    orc_write_reax_lists(S);
    orc_estimate_storages(S);
  }
  if ((rc = orc_reset(S))) return rc;
  orc_write_reax_lists(S);
  if ((rc = orc_init_forces(S))) return rc;
  if ((rc = orc_validate_lists(S))) return rc;
  return orc_bo(S);
}

// ------------------------------------------------------------------------
// Named array access
// ------------------------------------------------------------------------
int64_t orc_get_array(orc_sys *S, const char *name, void *out) {
  std::string k = name;
#define A(nm, vec) if (k == nm) return cp(S->vec, out);
  A("verlet_offsets", voff) A("verlet_list", vlist) A("packed_atoms", packed)
  A("far_index", far_index) A("far_end", far_end) A("far", far)
  A("num_bonds", num_bonds) A("num_hbonds", num_hbonds) A("Hindex", Hindex)
  A("bond_index", bindex) A("bond_end", bend) A("bonds", bonds) A("bo_data", bo_data)
  A("BO", BO) A("BOpi", BOpi) A("Cdbo", Cdbo) A("Cdbopi", Cdbopi) A("Cdbopi2", Cdbopi2)
  A("hbond_index", hindex) A("hbond_end", hend) A("hbonds", hbonds) A("hbond_far", hbond_far)
  A("bo_dboc", bo_dboc) A("dDeltap_self", dDeltap_self) A("bond_mark", bond_mark)
  A("Deltap", Deltap) A("Deltap_boc", Deltap_boc) A("Delta", Delta) A("Delta_e", Delta_e)
  A("Delta_val", Delta_val) A("vlpex", vlpex) A("nlp", nlp) A("nlp_temp", nlp_temp)
  A("Clp", Clp) A("Delta_lp", Delta_lp) A("dDelta_lp", dDelta_lp) A("Delta_lp_temp", Delta_lp_temp)
  A("dDelta_lp_temp", dDelta_lp_temp)
  A("H_firstnbr", Hfirst) A("H_numnbrs", Hnum) A("H_jlist", Hj) A("H_val", Hval)
  A("Hdia_inv", Hdia_inv) A("b_s", b_s) A("b_t", b_t) A("s", s) A("t", t)
  A("s_hist", s_hist) A("t_hist", t_hist) A("q", q) A("x", x) A("fCdDelta", fCdDelta)
#undef A
  return -1;
}

int orc_set_array(orc_sys *S, const char *name, const void *in, int64_t bytes) {
  std::string k = name;
  auto set_i = [&](std::vector<int> &v) { v.resize(bytes / sizeof(int)); std::memcpy(v.data(), in, bytes); };
  auto set_d = [&](std::vector<double> &v) { v.resize(bytes / sizeof(double)); std::memcpy(v.data(), in, bytes); };
  if (k == "num_bonds") { set_i(S->num_bonds); S->estimated = true; return 0; }
  if (k == "num_hbonds") { set_i(S->num_hbonds); return 0; }
  if (k == "s_hist") { set_d(S->s_hist); return 0; }
  if (k == "t_hist") { set_d(S->t_hist); return 0; }
  if (k == "s") { set_d(S->s); return 0; }
  if (k == "t") { set_d(S->t); return 0; }
  return -1;
}

int64_t orc_get_scalar_i(orc_sys *S, const char *name) {
  std::string k = name;
  if (k == "n") return S->n;
  if (k == "N") return S->N;
  if (k == "numH") return S->numH;
  if (k == "matvecs_s") return S->matvecs_s;
  if (k == "matvecs_t") return S->matvecs_t;
  if (k == "num_far") return (int64_t)S->far.size();
  if (k == "num_bonds_total") return S->num_bonds_total;
  if (k == "num_hbonds_total") return S->num_hbonds_total;
  if (k == "H_nnz") return (int64_t)S->Hj.size();
  if (k == "vdw_type") return S->vdw_type;
  return -1;
}

}  // extern "C"
