// swr_ffield.h — force-field parameters as read from a ReaxFF `ffield` file.
#pragma once

#include <string>
#include <vector>

struct FfAtom {
  std::string name;
  double r_s = 0, valency = 0, mass = 0, r_vdw = 0, epsilon = 0, gamma = 0, r_pi = 0, valency_e = 0;
  double nlp_opt = 0;
  double alpha = 0, gamma_w = 0, valency_boc = 0, p_ovun5 = 0, chi = 0, eta = 0;
  int p_hbond = 0;
  double r_pi_pi = 0, p_lp2 = 0, b_o_131 = 0, b_o_132 = 0, b_o_133 = 0;
  double p_ovun2 = 0, p_val3 = 0, valency_val = 0, p_val5 = 0, rcore2 = 0, ecore2 = 0, acore2 = 0;
};

struct FfPair {
  double p_bo1 = 0, p_bo2 = 0, p_bo3 = 0, p_bo4 = 0, p_bo5 = 0, p_bo6 = 0;
  double r_s = 0, r_p = 0, r_pp = 0;
  double p_boc3 = 0, p_boc4 = 0, p_boc5 = 0;
  double De_s = 0, De_p = 0, De_pp = 0, p_be1 = 0, p_be2 = 0, p_ovun1 = 0;
  double D = 0, alpha = 0, r_vdW = 0, gamma_w = 0, gamma = 0;
  double rcore = 0, ecore = 0, acore = 0;   // inner wall (vdw_type 2, 3)
  double v13cor = 0, ovc = 0;
};

struct FfHbond {
  double r0_hb = -1.0, p_hb1 = 0, p_hb2 = 0, p_hb3 = 0;   // r0_hb <= 0: no hydrogen bond
};

// One valence-angle parameter set (reaxc_ffield_sunway.cpp:526-572).
struct FfAngle {
  double theta_00 = 0, p_val1 = 0, p_val2 = 0, p_coa1 = 0, p_val7 = 0, p_pen1 = 0, p_val4 = 0;
};
constexpr int SWR_FF_MAX_THB = 5;   // parameter sets per (i, j, k) triple, REAX_MAX_3BODY_PARAM

// Torsion parameters of one (i, j, k, l) quadruple (reaxc_ffield_sunway.cpp:574-652).
struct FfTorsion {
  int cnt = 0;   // 1: the quadruple has a torsion term
  double V1 = 0, V2 = 0, V3 = 0, p_tor1 = 0, p_cot1 = 0;
};

struct swr_ffield_t {
  std::vector<double> gp;           // general parameters l[0..]
  std::vector<FfAngle> thb;         // [ne*ne*ne][SWR_FF_MAX_THB] (end atom i, centre j, end atom k)
  std::vector<int> thb_cnt;         // [ne*ne*ne]
  std::vector<FfTorsion> fbp;       // [ne*ne*ne*ne]
  std::vector<FfHbond> hb;          // [ne*ne*ne] (acceptor i, hydrogen j, acceptor k)
  std::vector<FfAtom> atom;
  std::vector<FfPair> pair;         // [ne*ne]
  double bo_cut = 0, nonb_low = 0, nonb_cut = 0;
  int vdw_type = 0;                 // 1 shielding, 2 inner wall, 3 both
  int ne() const { return (int)atom.size(); }
  const FfPair &tb(int i, int j) const { return pair[(size_t)i * atom.size() + j]; }
  FfPair &tb(int i, int j) { return pair[(size_t)i * atom.size() + j]; }
  const FfHbond &hbp(int i, int j, int k) const { return hb[((size_t)i * atom.size() + j) * atom.size() + k]; }
  size_t tri(int i, int j, int k) const { return ((size_t)i * atom.size() + j) * atom.size() + k; }
  size_t quad(int i, int j, int k, int l) const { return tri(i, j, k) * atom.size() + l; }
};

// Returns an empty string on success, else the reason.
std::string swr_read_ffield(const char *path, swr_ffield_t *ff);
