// ffield.cpp — ReaxFF force-field reader for the device path.
//
// Reads the sections this path consumes — general parameters, atom blocks,
// bond blocks and off-diagonal overrides — with the reference's semantics
// (Read_Force_Field, reaxc_ffield_sunway.cpp:35-545):
//   * tokens split on "\t \n\r\f!=" (Tokenize, reaxc_tool_box_sunway.cpp:42-57);
//   * bo_cut = 0.01*l[29], nonb_low = l[11], nonb_cut = l[12]          (:86-88);
//   * eta is stored as twice the file value                             (:213);
//   * nlp_opt = (valency_e - valency)/2                                 (:205);
//   * first-row elements (mass < 21) get valency_val = valency_boc      (:327-335);
//   * pair radii are arithmetic means, p_boc3/4/5 geometric means,
//     gamma = (g_i g_j)^-1.5, then positive off-diagonal values override
//     D, r_vdW (x2), alpha, r_s, r_p, r_pp                              (:402-545).
// The valence-angle, torsion and hydrogen-bond blocks are read as well
// (:547-686).
#include "swr_ffield.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace {

class Lines {
 public:
  explicit Lines(std::istream &in) : in_(in) {}
  // Next line split into tokens; false at end of file.
  bool next(std::vector<std::string> &tok) {
    std::string line;
    if (!std::getline(in_, line)) return false;
    tok.clear();
    size_t i = 0;
    while (i < line.size()) {
      while (i < line.size() && is_sep(line[i])) i++;
      size_t j = i;
      while (j < line.size() && !is_sep(line[j])) j++;
      if (j > i) tok.emplace_back(line.substr(i, j - i));
      i = j;
    }
    return true;
  }

 private:
  static bool is_sep(char c) { return c == '\t' || c == ' ' || c == '\n' || c == '\r' || c == '\f' || c == '!' || c == '='; }
  std::istream &in_;
};

double num(const std::vector<std::string> &t, size_t k) { return k < t.size() ? std::atof(t[k].c_str()) : 0.0; }

}  // namespace

std::string swr_read_ffield(const char *path, swr_ffield_t *ff) {
  std::ifstream f(path);
  if (!f) return std::string("cannot open force field ") + path;
  Lines L(f);
  std::vector<std::string> t, t2;
  if (!L.next(t)) return "empty force field";
  if (!L.next(t) || t.empty()) return "missing global count";
  const int ngl = std::atoi(t[0].c_str());
  if (ngl < 1) return "number of globals in ffield file is 0";
  ff->gp.assign(ngl, 0.0);
  for (int k = 0; k < ngl; k++) {
    if (!L.next(t)) return "truncated globals";
    ff->gp[k] = num(t, 0);
  }
  if (ngl < 30) return "too few global parameters";
  ff->bo_cut = 0.01 * ff->gp[29];
  ff->nonb_low = ff->gp[11];
  ff->nonb_cut = ff->gp[12];

  if (!L.next(t) || t.empty()) return "missing atom count";
  const int ne = std::atoi(t[0].c_str());
  for (int k = 0; k < 3; k++) L.next(t);
  ff->atom.assign(ne, FfAtom());
  ff->pair.assign((size_t)ne * ne, FfPair());
  for (int e = 0; e < ne; e++) {
    FfAtom &a = ff->atom[e];
    if (!L.next(t) || t.size() < 9) return "bad atom line 1";
    a.name = t[0];
    for (char &c : a.name) c = (char)std::toupper((unsigned char)c);
    a.r_s = num(t, 1); a.valency = num(t, 2); a.mass = num(t, 3); a.r_vdw = num(t, 4);
    a.epsilon = num(t, 5); a.gamma = num(t, 6); a.r_pi = num(t, 7); a.valency_e = num(t, 8);
    a.nlp_opt = 0.5 * (a.valency_e - a.valency);
    if (!L.next(t)) return "bad atom line 2";
    a.alpha = num(t, 0); a.gamma_w = num(t, 1); a.valency_boc = num(t, 2); a.p_ovun5 = num(t, 3);
    a.chi = num(t, 5); a.eta = 2.0 * num(t, 6); a.p_hbond = (int)num(t, 7);
    if (!L.next(t)) return "bad atom line 3";
    a.r_pi_pi = num(t, 0); a.p_lp2 = num(t, 1);
    a.b_o_131 = num(t, 3); a.b_o_132 = num(t, 4); a.b_o_133 = num(t, 5);
    if (!L.next(t) || t.size() < 3) return "Inconsistent ffield file (atom line 4)";
    a.p_ovun2 = num(t, 0); a.p_val3 = num(t, 1); a.valency_val = num(t, 3); a.p_val5 = num(t, 4);
    a.rcore2 = num(t, 5); a.ecore2 = num(t, 6); a.acore2 = num(t, 7);
  }
  // vdWaals method (reaxc_ffield_sunway.cpp:241-292): the first element
  // decides, later disagreeing elements keep it; neither shielding nor an
  // inner wall is an error
  ff->vdw_type = 0;
  for (const FfAtom &a : ff->atom) {
    const bool wall = a.rcore2 > 0.01 && a.acore2 > 0.01, shield = a.gamma_w > 0.5;
    if (!wall && !shield) return "No shielding or inner-wall set for element " + a.name;
    if (ff->vdw_type == 0) ff->vdw_type = wall ? (shield ? 3 : 2) : 1;
  }
  for (FfAtom &a : ff->atom)
    if (a.mass < 21 && a.valency_val != a.valency_boc) a.valency_val = a.valency_boc;

  if (!L.next(t) || t.empty()) return "missing bond count";
  const int nbond = std::atoi(t[0].c_str());
  L.next(t);
  for (int k = 0; k < nbond; k++) {
    if (!L.next(t) || !L.next(t2)) return "truncated bond block";
    const int a = std::atoi(t[0].c_str()) - 1, b = std::atoi(t[1].c_str()) - 1;
    if (a >= ne || b >= ne || a < 0 || b < 0) continue;
    for (int pass = 0; pass < 2; pass++) {
      FfPair &p = pass ? ff->tb(b, a) : ff->tb(a, b);
      p.De_s = num(t, 2); p.De_p = num(t, 3); p.De_pp = num(t, 4); p.p_be1 = num(t, 5);
      p.p_bo5 = num(t, 6); p.v13cor = num(t, 7); p.p_bo6 = num(t, 8); p.p_ovun1 = num(t, 9);
      p.p_be2 = num(t2, 0); p.p_bo3 = num(t2, 1); p.p_bo4 = num(t2, 2);
      p.p_bo1 = num(t2, 4); p.p_bo2 = num(t2, 5); p.ovc = num(t2, 6);
    }
  }
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < ne; j++) {
      const FfAtom &a = ff->atom[i], &b = ff->atom[j];
      FfPair &p = ff->tb(i, j);
      p.r_s = 0.5 * (a.r_s + b.r_s);
      p.r_p = 0.5 * (a.r_pi + b.r_pi);
      p.r_pp = 0.5 * (a.r_pi_pi + b.r_pi_pi);
      p.p_boc3 = std::sqrt(a.b_o_132 * b.b_o_132);
      p.p_boc4 = std::sqrt(a.b_o_131 * b.b_o_131);
      p.p_boc5 = std::sqrt(a.b_o_133 * b.b_o_133);
      p.D = std::sqrt(a.epsilon * b.epsilon);
      p.alpha = std::sqrt(a.alpha * b.alpha);
      p.r_vdW = 2.0 * std::sqrt(a.r_vdw * b.r_vdw);
      p.gamma_w = std::sqrt(a.gamma_w * b.gamma_w);
      p.gamma = std::pow(a.gamma * b.gamma, -1.5);
      p.rcore = std::sqrt(a.rcore2 * b.rcore2);   // :445-452
      p.ecore = std::sqrt(a.ecore2 * b.ecore2);
      p.acore = std::sqrt(a.acore2 * b.acore2);
    }
  if (L.next(t) && !t.empty()) {
    const int noff = std::atoi(t[0].c_str());
    for (int k = 0; k < noff; k++) {
      if (!L.next(t)) return "truncated off-diagonal block";
      const int a = std::atoi(t[0].c_str()) - 1, b = std::atoi(t[1].c_str()) - 1;
      if (a >= ne || b >= ne || a < 0 || b < 0) continue;
      double v;
      for (int pass = 0; pass < 2; pass++) {
        FfPair &p = pass ? ff->tb(b, a) : ff->tb(a, b);
        if ((v = num(t, 2)) > 0.0) p.D = v;
        if ((v = num(t, 3)) > 0.0) p.r_vdW = 2 * v;
        if ((v = num(t, 4)) > 0.0) p.alpha = v;
        if ((v = num(t, 5)) > 0.0) p.r_s = v;
        if ((v = num(t, 6)) > 0.0) p.r_p = v;
        if ((v = num(t, 7)) > 0.0) p.r_pp = v;
      }
    }
  }
  // valence-angle block (:520-572): a line "a b c" adds one parameter set
  // to (a,b,c) and to (c,b,a) at the count (a,b,c) had, and counts both up
  // (so a == c counts twice and leaves a zero set, which p_val1 = 0 skips)
  ff->hb.assign((size_t)ne * ne * ne, FfHbond());
  ff->thb.assign((size_t)ne * ne * ne * SWR_FF_MAX_THB, FfAngle());
  ff->thb_cnt.assign((size_t)ne * ne * ne, 0);
  if (!L.next(t) || t.empty()) return "";
  const int nang = std::atoi(t[0].c_str());
  for (int k = 0; k < nang; k++) {
    if (!L.next(t)) return "truncated valence-angle block";
    if (t.size() < 10) return "bad valence-angle line";
    const int a = std::atoi(t[0].c_str()) - 1, b = std::atoi(t[1].c_str()) - 1, c = std::atoi(t[2].c_str()) - 1;
    if (a < 0 || b < 0 || c < 0 || a >= ne || b >= ne || c >= ne) continue;
    const int cnt = ff->thb_cnt[ff->tri(a, b, c)];
    if (cnt >= SWR_FF_MAX_THB || ff->thb_cnt[ff->tri(c, b, a)] >= SWR_FF_MAX_THB)
      return "too many valence-angle parameter sets for one triple";
    ff->thb_cnt[ff->tri(a, b, c)]++;
    ff->thb_cnt[ff->tri(c, b, a)]++;
    FfAngle g;
    g.theta_00 = num(t, 3); g.p_val1 = num(t, 4); g.p_val2 = num(t, 5); g.p_coa1 = num(t, 6);
    g.p_val7 = num(t, 7); g.p_pen1 = num(t, 8); g.p_val4 = num(t, 9);
    ff->thb[ff->tri(a, b, c) * SWR_FF_MAX_THB + cnt] = g;
    ff->thb[ff->tri(c, b, a) * SWR_FF_MAX_THB + cnt] = g;
  }
  // torsion block (:574-652): "i j k l" (all non-zero) sets (i,j,k,l) and
  // (l,k,j,i) and marks them; "0 j k 0" gives every (p,j,k,o) and (o,k,j,p)
  // a term, with these values where no explicit line set them
  ff->fbp.assign((size_t)ne * ne * ne * ne, FfTorsion());
  {
    std::vector<char> flag((size_t)ne * ne * ne * ne, 0);
    if (!L.next(t) || t.empty()) return "";
    const int l = std::atoi(t[0].c_str());
    for (int k = 0; k < l; k++) {
      if (!L.next(t)) return "truncated torsion block";
      if (t.size() < 9) return "bad torsion line";
      const int a = std::atoi(t[0].c_str()) - 1, b = std::atoi(t[1].c_str()) - 1, c = std::atoi(t[2].c_str()) - 1,
                d = std::atoi(t[3].c_str()) - 1;
      FfTorsion v;
      v.cnt = 1; v.V1 = num(t, 4); v.V2 = num(t, 5); v.V3 = num(t, 6); v.p_tor1 = num(t, 7); v.p_cot1 = num(t, 8);
      if (a >= 0 && d >= 0) {
        if (a < ne && b < ne && c < ne && d < ne && b >= 0 && c >= 0) {
          flag[ff->quad(a, b, c, d)] = flag[ff->quad(d, c, b, a)] = 1;
          ff->fbp[ff->quad(a, b, c, d)] = v;
          ff->fbp[ff->quad(d, c, b, a)] = v;
        }
      } else if (b >= 0 && c >= 0 && b < ne && c < ne) {
        for (int p = 0; p < ne; p++)
          for (int o = 0; o < ne; o++) {
            FfTorsion &f1 = ff->fbp[ff->quad(p, b, c, o)], &f2 = ff->fbp[ff->quad(o, c, b, p)];
            f1.cnt = f2.cnt = 1;
            if (!flag[ff->quad(p, b, c, o)]) f1 = v;
            if (!flag[ff->quad(o, c, b, p)]) f2 = v;
          }
      }
    }
  }
  if (!L.next(t) || t.empty()) return "";
  const int nh = std::atoi(t[0].c_str());
  for (int k = 0; k < nh; k++) {
    if (!L.next(t) || t.size() < 7) return "truncated hydrogen-bond block";
    const int a = std::atoi(t[0].c_str()) - 1, b = std::atoi(t[1].c_str()) - 1, c = std::atoi(t[2].c_str()) - 1;
    if (a < 0 || b < 0 || c < 0 || a >= ne || b >= ne || c >= ne) continue;
    FfHbond &h = ff->hb[((size_t)a * ne + b) * ne + c];
    h.r0_hb = num(t, 3); h.p_hb1 = num(t, 4); h.p_hb2 = num(t, 5); h.p_hb3 = num(t, 6);
  }
  return "";
}
