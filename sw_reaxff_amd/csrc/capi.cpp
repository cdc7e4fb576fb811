// capi.cpp — the C-ABI of include/swreax.h: system lifecycle, device
// allocation, the per-stage entry points, status and named mirrors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>

#include "swr_internal.h"

namespace {

constexpr int MARK_SWEEPS = 1024;

int copy_grow(void **p, size_t old_bytes, size_t new_bytes, hipStream_t s) {
  void *np = nullptr;
  hipError_t e = hipMalloc(&np, new_bytes ? new_bytes : 1);
  if (e != hipSuccess) {
    swr_set_error_msg(std::string("hipMalloc: ") + hipGetErrorString(e));
    return SWR_ERR_NOMEM;
  }
  if (*p && old_bytes) {
    e = hipMemcpyAsync(np, *p, std::min(old_bytes, new_bytes), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      swr_set_error_msg(std::string("hipMemcpy: ") + hipGetErrorString(e));
      return SWR_ERR_HIP;
    }
    (void)hipFree(*p);
  }
  *p = np;
  return SWR_OK;
}

void fill_params(swr_system *S, const std::vector<int> &map) {
  DevParams &h = S->hp;
  std::memset(&h, 0, sizeof(h));
  const swr_ffield_t &ff = S->ff;
  h.nelem = ff.ne();
  h.ntypes = (int)map.size() - 1;
  for (size_t t = 0; t < map.size(); t++) h.map[t] = map[t];
  // atom->mass defaults to the force field's element masses (the data files
  // of the reference list the same values); swr_set_masses overrides them
  for (size_t t = 1; t < map.size(); t++) h.type_mass[t] = map[t] >= 0 ? ff.atom[map[t]].mass : 0.0;
  for (int e = 0; e < ff.ne() && e < SWR_MAX_ELEM; e++) {
    const FfAtom &a = ff.atom[e];
    h.r_s[e] = a.r_s; h.r_pi[e] = a.r_pi; h.r_pi_pi[e] = a.r_pi_pi;
    h.valency[e] = a.valency; h.valency_e[e] = a.valency_e; h.valency_boc[e] = a.valency_boc;
    h.valency_val[e] = a.valency_val; h.nlp_opt[e] = a.nlp_opt; h.mass[e] = a.mass;
    h.p_hbond[e] = a.p_hbond;
    for (int f = 0; f < ff.ne() && f < SWR_MAX_ELEM; f++) {
      const FfPair &p = ff.tb(e, f);
      DevTbp &d = h.tbp[e * SWR_MAX_ELEM + f];
      d.p_bo1 = p.p_bo1; d.p_bo2 = p.p_bo2; d.p_bo3 = p.p_bo3; d.p_bo4 = p.p_bo4; d.p_bo5 = p.p_bo5;
      d.p_bo6 = p.p_bo6; d.r_s = p.r_s; d.r_p = p.r_p; d.r_pp = p.r_pp; d.p_boc3 = p.p_boc3;
      d.p_boc4 = p.p_boc4; d.p_boc5 = p.p_boc5; d.v13cor = p.v13cor; d.ovc = p.ovc;
    }
  }
  // bond-order radius per pair (host bisection, widened by 1e-9 relative)
  for (int e = 0; e < ff.ne() && e < SWR_MAX_ELEM; e++)
    for (int f = 0; f < ff.ne() && f < SWR_MAX_ELEM; f++) {
      const FfAtom &a = ff.atom[e], &b = ff.atom[f];
      const FfPair &p = ff.tb(e, f);
      const bool es = a.r_s > 0.0 && b.r_s > 0.0, ep = a.r_pi > 0.0 && b.r_pi > 0.0,
                 epp = a.r_pi_pi > 0.0 && b.r_pi_pi > 0.0;
      const bool mono = (!es || (p.p_bo1 < 0 && p.p_bo2 > 0 && p.r_s > 0)) &&
                        (!ep || (p.p_bo3 < 0 && p.p_bo4 > 0 && p.r_p > 0)) &&
                        (!epp || (p.p_bo5 < 0 && p.p_bo6 > 0 && p.r_pp > 0));
      auto bop = [&](double d) {
        double v = 0.0;
        if (es) v += (1.0 + ff.bo_cut) * std::exp(p.p_bo1 * std::pow(d / p.r_s, p.p_bo2));
        if (ep) v += std::exp(p.p_bo3 * std::pow(d / p.r_p, p.p_bo4));
        if (epp) v += std::exp(p.p_bo5 * std::pow(d / p.r_pp, p.p_bo6));
        return v;
      };
      double r;
      if (!es && !ep && !epp) r = -1.0;
      else if (!mono || bop(S->ctl.bond_cut) >= ff.bo_cut) r = S->ctl.bond_cut;
      else {
        double lo = 0.0, hi = S->ctl.bond_cut;
        for (int it = 0; it < 200; it++) {
          const double mid = 0.5 * (lo + hi);
          if (bop(mid) >= ff.bo_cut) lo = mid; else hi = mid;
        }
        r = std::min(S->ctl.bond_cut, hi * (1.0 + 1e-9) + 1e-12);
      }
      h.rbo[e * SWR_MAX_ELEM + f] = r;
    }
  h.p_boc1 = ff.gp[0];
  h.p_boc2 = ff.gp[1];
  h.p_lp1 = ff.gp[15];
  h.bo_cut = ff.bo_cut;
  h.bond_cut = S->ctl.bond_cut;
  h.hbond_cut = S->ctl.hbond_cut;
  h.nonb_cut = ff.nonb_cut;
  h.nonb_cutsq = ff.nonb_cut * ff.nonb_cut;
  // bond and atom energies
  for (int e = 0; e < ff.ne() && e < SWR_MAX_ELEM; e++) {
    const FfAtom &a = ff.atom[e];
    h.p_lp2[e] = a.p_lp2; h.p_ovun2[e] = a.p_ovun2; h.p_ovun5[e] = a.p_ovun5;
    h.chi_e[e] = a.chi; h.eta_e[e] = a.eta;
    h.is_carbon[e] = a.name == "C" ? 1 : 0;
    for (int f = 0; f < ff.ne() && f < SWR_MAX_ELEM; f++) {
      const FfPair &p = ff.tb(e, f);
      DevBe &d = h.be[e * SWR_MAX_ELEM + f];
      d.De_s = p.De_s; d.De_p = p.De_p; d.De_pp = p.De_pp; d.p_be1 = p.p_be1; d.p_be2 = p.p_be2;
      d.p_ovun1 = p.p_ovun1;
    }
  }
  {
    auto g = [&](size_t k) { return k < ff.gp.size() ? ff.gp[k] : 0.0; };
    h.p_lp3 = g(5); h.p_ovun3 = g(32); h.p_ovun4 = g(31); h.p_ovun6 = g(6); h.p_ovun7 = g(8);
    h.p_ovun8 = g(9); h.gp3 = g(3); h.gp4 = g(4); h.gp7 = g(7); h.gp10 = g(10);
    h.gp37 = (int)g(37);
    h.enobonds = S->ctl.enobonds;
    h.p_val6 = g(14); h.p_val8 = g(33); h.p_val9 = g(16); h.p_val10 = g(17);
    h.p_pen2 = g(19); h.p_pen3 = g(20); h.p_pen4 = g(21);
    h.p_coa2 = g(2); h.p_coa3 = g(38); h.p_coa4 = g(30);
    h.thb_cut = S->ctl.thb_cut; h.thb_cutsq = S->ctl.thb_cutsq;
    h.p_tor2 = g(23); h.p_tor3 = g(24); h.p_tor4 = g(25); h.p_cot2 = g(27);
  }
  for (int a = 0; a < ff.ne() && a < SWR_MAX_ELEM; a++)
    for (int b = 0; b < ff.ne() && b < SWR_MAX_ELEM; b++)
      for (int c = 0; c < ff.ne() && c < SWR_MAX_ELEM; c++)
        for (int d = 0; d < ff.ne() && d < SWR_MAX_ELEM; d++) {
          DevTor &o = h.tor[((a * SWR_MAX_ELEM + b) * SWR_MAX_ELEM + c) * SWR_MAX_ELEM + d];
          if (ff.fbp.empty()) { o.cnt = 0; continue; }
          const FfTorsion &f = ff.fbp[ff.quad(a, b, c, d)];
          o.cnt = f.cnt; o.V1 = f.V1; o.V2 = f.V2; o.V3 = f.V3; o.p_tor1 = f.p_tor1; o.p_cot1 = f.p_cot1;
        }
  for (int e = 0; e < ff.ne() && e < SWR_MAX_ELEM; e++) {
    h.p_val3[e] = ff.atom[e].p_val3;
    h.p_val5[e] = ff.atom[e].p_val5;
  }
  // valence-angle sets the energy uses (Merge_Torsion_Valence_Angles skips
  // |p_val1| <= 0.001, reaxc_torsion_angles_sw64.c:960)
  for (int a = 0; a < ff.ne() && a < SWR_MAX_ELEM; a++)
    for (int b = 0; b < ff.ne() && b < SWR_MAX_ELEM; b++)
      for (int c = 0; c < ff.ne() && c < SWR_MAX_ELEM; c++) {
        const int d = (a * SWR_MAX_ELEM + b) * SWR_MAX_ELEM + c;
        h.thb_n[d] = 0;
        h.thb_pc[d] = 0;
        if (ff.thb_cnt.empty()) continue;
        const size_t t = ff.tri(a, b, c);
        for (int k = 0; k < ff.thb_cnt[t] && k < SWR_FF_MAX_THB; k++) {
          const FfAngle &f = ff.thb[t * SWR_FF_MAX_THB + k];
          if (!(std::fabs(f.p_val1) > 0.001)) continue;
          DevThb &o = h.thb[d][h.thb_n[d]++];
          o.theta_00 = f.theta_00; o.p_val1 = f.p_val1; o.p_val2 = f.p_val2; o.p_coa1 = f.p_coa1;
          o.p_val7 = f.p_val7; o.p_pen1 = f.p_pen1; o.p_val4 = f.p_val4;
          if (f.p_pen1 != 0.0 || f.p_coa1 != 0.0) h.thb_pc[d] = 1;
        }
      }
  for (int a = 0; a < ff.ne() && a < SWR_MAX_ELEM; a++)
    for (int b = 0; b < ff.ne() && b < SWR_MAX_ELEM; b++)
      for (int c = 0; c < ff.ne() && c < SWR_MAX_ELEM; c++) {
        DevHb &d = h.hb[(a * SWR_MAX_ELEM + b) * SWR_MAX_ELEM + c];
        if (ff.hb.empty()) { d.r0_hb = -1.0; continue; }
        const FfHbond &f = ff.hbp(a, b, c);
        d.r0_hb = f.r0_hb; d.p_hb1 = f.p_hb1; d.p_hb2 = f.p_hb2; d.p_hb3 = f.p_hb3;
      }
  // non-bonded pair factors and taper
  h.vdw_type = ff.vdw_type;
  h.p_vdW1 = ff.gp.size() > 28 ? ff.gp[28] : 0.0;
  h.p_vdW1i = 1.0 / h.p_vdW1;
  for (int e = 0; e < ff.ne() && e < SWR_MAX_ELEM; e++)
    for (int f = 0; f < ff.ne() && f < SWR_MAX_ELEM; f++) {
      const FfPair &p = ff.tb(e, f);
      DevNb &d = h.nb[e * SWR_MAX_ELEM + f];
      d.D = p.D;
      d.alpha = p.alpha;
      d.r_vdWinv = 1.0 / p.r_vdW;
      d.alpha_div_rvdW = p.alpha / p.r_vdW;
      d.powgi_vdW1 = std::pow(p.gamma_w, -h.p_vdW1);
      d.gamma = p.gamma;
      d.ecore = p.ecore;
      d.acore = p.acore;
      d.rcoreinv = p.rcore > 0 ? 1.0 / p.rcore : 0.0;
      d.a_div_rcore = p.rcore > 0 ? p.acore / p.rcore : 0.0;
    }
  {
    const double swa = ff.nonb_low, swb = ff.nonb_cut;
    const double d7 = std::pow(swb - swa, 7.0);
    const double swa2 = swa * swa, swa3 = swa * swa * swa, swb2 = swb * swb, swb3 = swb * swb * swb;
    h.nb_tap[7] = 20.0 / d7;
    h.nb_tap[6] = -70.0 * (swa + swb) / d7;
    h.nb_tap[5] = 84.0 * (swa2 + 3.0 * swa * swb + swb2) / d7;
    h.nb_tap[4] = -35.0 * (swa3 + 9.0 * swa2 * swb + 9.0 * swa * swb2 + swb3) / d7;
    h.nb_tap[3] = 140.0 * (swa3 * swb + 3.0 * swa2 * swb2 + swa * swb3) / d7;
    h.nb_tap[2] = -210.0 * (swa3 * swb2 + swa2 * swb3) / d7;
    h.nb_tap[1] = 140.0 * swa3 * swb3 / d7;
    h.nb_tap[0] = (-35.0 * swa3 * swb2 * swb2 + 21.0 * swa2 * swb3 * swb2 + 7.0 * swa * swb3 * swb3 +
                   swb3 * swb3 * swb) / d7;
  }
  // pair->extract("chi"/"eta"/"gamma") per LAMMPS type (pair_reaxc_sunway.cpp:1106-1128)
  for (int t = 1; t <= h.ntypes; t++) {
    const FfAtom &a = ff.atom[map[t]];
    h.chi[t] = a.chi; h.eta[t] = a.eta; h.gamma[t] = a.gamma;
  }
  // init_shielding (fix_qeq_reax_sunway.cpp:440-454)
  const int nt1 = SWR_MAX_TYPES + 1;
  for (int i = 1; i <= h.ntypes; i++)
    for (int j = 1; j <= h.ntypes; j++) h.shld[i * nt1 + j] = std::pow(h.gamma[i] * h.gamma[j], -1.5);
  // init_taper (fix_qeq_reax_sunway.cpp:458-484)
  const double swa = S->ctl.swa, swb = S->ctl.swb;
  const double d7 = std::pow(swb - swa, 7);
  const double swa2 = swa * swa, swa3 = swa * swa * swa, swb2 = swb * swb, swb3 = swb * swb * swb;
  h.Tap[7] = 20.0 / d7;
  h.Tap[6] = -70.0 * (swa + swb) / d7;
  h.Tap[5] = 84.0 * (swa2 + 3.0 * swa * swb + swb2) / d7;
  h.Tap[4] = -35.0 * (swa3 + 9.0 * swa2 * swb + 9.0 * swa * swb2 + swb3) / d7;
  h.Tap[3] = 140.0 * (swa3 * swb + 3.0 * swa2 * swb2 + swa * swb3) / d7;
  h.Tap[2] = -210.0 * (swa3 * swb2 + swa2 * swb3) / d7;
  h.Tap[1] = 140.0 * swa3 * swb3 / d7;
  h.Tap[0] = (-35.0 * swa3 * swb2 * swb2 + 21.0 * swa2 * swb3 * swb2 + 7.0 * swa * swb3 * swb3 +
              swb3 * swb3 * swb) / d7;
  h.swa = swa;
  h.swb = swb;
  h.swbsq = swb * swb;
  h.tol = S->ctl.tol;
}

int status_to_rc(unsigned int f) {
  if (f & (SWR_F_BOND_CAP | SWR_F_BOND_TOTAL | SWR_F_BOND_LDS)) return SWR_ERR_BOND_CAP;
  if (f & (SWR_F_HBOND_CAP | SWR_F_HBOND_TOTAL)) return SWR_ERR_HBOND_CAP;
  if (f & SWR_F_BIN) return SWR_ERR_BIN;
  if (f & (SWR_F_NAN | SWR_F_FIXED_RANGE)) return SWR_ERR_NUMERIC;
  if (f & SWR_F_MARK_DEPTH) return SWR_ERR_STATE;
  return SWR_OK;
}

}  // namespace

int swr_status_to_rc(unsigned int f) { return status_to_rc(f); }

// Grow every per-atom array to hold Ncap atoms (contents preserved).
int swr_nbr_ensure_atoms(swr_system *S, int Ncap) {
  if (Ncap <= S->Ncap) return SWR_OK;
  const size_t o = S->Ncap, nn = Ncap;
  int rc;
#define GROW(ptr, T, per)                                                                              \
  if ((rc = copy_grow((void **)&S->ptr, o * (per) * sizeof(T), nn * (per) * sizeof(T), S->stream))) \
    return rc;
  GROW(x, double, 3) GROW(q, double, 1) GROW(type, int, 1) GROW(tag, int, 1) GROW(owner, int, 1)
  GROW(img, signed char, 3) GROW(num_bonds, int, 1) GROW(num_hbonds, int, 1)
#undef GROW
  // scratch-like arrays: contents need not survive
  if ((rc = swr_alloc(&S->packed, nn)) || (rc = swr_alloc(&S->Hindex, nn)) || (rc = swr_alloc(&S->bond_index, nn)) ||
      (rc = swr_alloc(&S->bond_end, nn)) || (rc = swr_alloc(&S->hbond_index, nn)) ||
      (rc = swr_alloc(&S->hbond_end, nn)) || (rc = swr_alloc(&S->bo_dboc, 2 * nn)) ||
      (rc = swr_alloc(&S->dDeltap_self, 3 * nn)) || (rc = swr_alloc(&S->bond_mark, nn)) ||
      (rc = swr_alloc(&S->bond_mark2, nn)) || (rc = swr_alloc(&S->counts, 2 * nn)) ||
      (rc = swr_alloc(&S->ws, WS_COUNT * nn)) || (rc = swr_alloc(&S->far_index, nn)) ||
      (rc = swr_alloc(&S->far_end, nn)) || (rc = swr_alloc(&S->fcd, 4 * nn)))
    return rc;
  HIP_CHECK(hipMemsetAsync(S->counts, 0, 2 * nn * sizeof(int), S->stream));
  S->Ncap = Ncap;
  return SWR_OK;
}

static int ensure_local(swr_system *S, int ncap) {
  if (ncap <= S->ncap) return SWR_OK;
  const size_t o = S->ncap, nn = ncap;
  int rc;
  if ((rc = copy_grow((void **)&S->v, o * 3 * sizeof(double), nn * 3 * sizeof(double), S->stream))) return rc;
  if ((rc = copy_grow((void **)&S->s_hist, o * 5 * sizeof(double), nn * 5 * sizeof(double), S->stream))) return rc;
  if ((rc = copy_grow((void **)&S->t_hist, o * 5 * sizeof(double), nn * 5 * sizeof(double), S->stream))) return rc;
  if ((rc = copy_grow((void **)&S->st, o * 2 * sizeof(double), nn * 2 * sizeof(double), S->stream))) return rc;
  if (o < nn) {
    HIP_CHECK(hipMemsetAsync(S->s_hist + o * 5, 0, (nn - o) * 5 * sizeof(double), S->stream));
    HIP_CHECK(hipMemsetAsync(S->t_hist + o * 5, 0, (nn - o) * 5 * sizeof(double), S->stream));
    HIP_CHECK(hipMemsetAsync(S->st + o * 2, 0, (nn - o) * 2 * sizeof(double), S->stream));
  }
  if ((rc = swr_alloc(&S->H_firstnbr, nn)) || (rc = swr_alloc(&S->H_numnbrs, nn)) ||
      (rc = swr_alloc(&S->Hdia_inv, nn)) || (rc = swr_alloc(&S->b_s, 2 * nn)) ||
      (rc = swr_alloc(&S->cgv, (size_t)CGV_COUNT * 2 * nn)))
    return rc;
  S->layout_dirty = true;
  S->b_t = nullptr;
  S->ncap = ncap;
  return SWR_OK;
}

extern "C" {

void swr_control_default(swr_control_t *c) {
  std::memset(c, 0, sizeof(*c));
  c->bond_cut = 5.0;    // pair_reaxc_sunway.cpp:216
  c->hbond_cut = 7.5;   // :217
  c->thb_cut = 0.001;   // :218
  c->thb_cutsq = 0.00001;   // :219
  c->safezone = 1.2;    // SAFE_ZONE
  c->saferzone = 1.4;   // SAFER_ZONE
  c->mincap = 50;       // MIN_CAP
  c->maxfar = 1024;     // :241
  c->swa = 0.0;
  c->swb = 10.0;
  c->tol = 1e-6;
  c->skin = 2.0;
  c->qeq_mixed = 0;
  c->qeq_imax = 200;    // fix_qeq_reax_sunway.cpp:998
  c->enobonds = 1;      // pair_reaxc_sunway.cpp:237
}

int swr_create(const char *ffield_path, int ntypes, const char *const *elements, const swr_control_t *ctl,
               swr_system **out) {
  if (!ffield_path || !elements || !out || ntypes < 1 || ntypes > SWR_MAX_TYPES) {
    swr_set_error_msg("swr_create: bad arguments");
    return SWR_ERR_ARG;
  }
  swr_system *S = new swr_system();
  if (ctl) S->ctl = *ctl; else swr_control_default(&S->ctl);
  std::string err = swr_read_ffield(ffield_path, &S->ff);
  if (!err.empty()) {
    swr_set_error_msg(err);
    delete S;
    return SWR_ERR_FFIELD;
  }
  if (S->ff.ne() > SWR_MAX_ELEM) {
    swr_set_error_msg("force field has more elements than supported");
    delete S;
    return SWR_ERR_FFIELD;
  }
  std::vector<int> map(ntypes + 1, -1);
  for (int t = 1; t <= ntypes; t++) {
    std::string e = elements[t - 1];
    for (char &c : e) c = (char)std::toupper((unsigned char)c);
    for (int k = 0; k < S->ff.ne(); k++)
      if (S->ff.atom[k].name == e) map[t] = k;
    if (map[t] < 0) {
      swr_set_error_msg("Non-existent ReaxFF type " + e);
      delete S;
      return SWR_ERR_FFIELD;
    }
  }
  S->mixed = S->ctl.qeq_mixed;
  S->imax = S->ctl.qeq_imax > 0 ? S->ctl.qeq_imax : 200;
  fill_params(S, map);
  hipError_t e;
  if ((e = hipGetDevice(&S->device)) != hipSuccess || (e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking)) != hipSuccess) {
    swr_set_error_msg(std::string("HIP init: ") + hipGetErrorString(e));
    delete S;
    return SWR_ERR_HIP;
  }
  int rc;
  if ((rc = swr_alloc(&S->dp, 1)) || (rc = swr_alloc(&S->status, 1)) || (rc = swr_alloc(&S->counters, 4))) {
    swr_destroy(S);
    return rc;
  }
  // the device address of the no-allocation entry point's scratch, once
  // per device (with this device current)
  if ((rc = swr_noalloc_init(S->device))) {
    swr_destroy(S);
    return rc;
  }
  if (hipHostMalloc((void **)&S->status_host, sizeof(DevStatus), 0) != hipSuccess) {
    swr_set_error_msg("hipHostMalloc failed");
    swr_destroy(S);
    return SWR_ERR_NOMEM;
  }
  std::memset(S->status_host, 0, sizeof(DevStatus));
  if (hipHostMalloc((void **)&S->pin_done, 64, 0) != hipSuccess ||
      hipHostMalloc(&S->pin_mix, SWR_PIN_MIX, 0) != hipSuccess) {
    swr_set_error_msg("hipHostMalloc failed");
    swr_destroy(S);
    return SWR_ERR_NOMEM;
  }
  if ((rc = swr_pin_iters_reserve(S, 64))) {
    swr_destroy(S);
    return rc;
  }
  // the system's stream does not wait for the null stream: everything goes on it
  if (hipMemcpyAsync(S->dp, &S->hp, sizeof(DevParams), hipMemcpyHostToDevice, S->stream) != hipSuccess ||
      hipMemsetAsync(S->status, 0, sizeof(DevStatus), S->stream) != hipSuccess ||
      hipMemsetAsync(S->counters, 0, 4 * sizeof(unsigned int), S->stream) != hipSuccess ||
      hipStreamSynchronize(S->stream) != hipSuccess) {
    swr_set_error_msg("HIP init: parameter upload failed");
    swr_destroy(S);
    return SWR_ERR_HIP;
  }
  *out = S;
  return SWR_OK;
}

int swr_destroy(swr_system *S) {
  if (!S) return SWR_OK;
  if (S->stream) (void)hipStreamSynchronize(S->stream);
  void *ptrs[] = {S->dp, S->x, S->v, S->q, S->type, S->tag, S->owner, S->img, S->mask, S->packed,
                  S->bin_of, S->bin_count, S->bin_start, S->bin_occ, S->rs_keys, S->rs_vals, S->rs_hist, S->bp_pos, S->bp_id, S->stencil_rows,
                  S->voff, S->vcount, S->vlist, S->far, S->far_index, S->far_end, S->num_bonds, S->num_hbonds,
                  S->Hindex, S->bond_index, S->bond_end, S->bonds, S->bond_nbr, S->bo_data, S->prow, S->pbeg, S->pcnt, S->pcol, S->pkey, S->jloc, S->sys_panel, S->BO, S->BOpi, S->Cdbo,
                  S->Cdbopi,
                  S->Cdbopi2, S->hbond_index, S->hbond_end, S->hbonds, S->bo_dboc, S->dDeltap_self, S->bond_mark,
                  S->fcd, S->nb_partials, S->nb_result, S->nb_eatom, S->be_part, S->be_ce, S->be_bufc, S->be_result, S->hb_part,
                  S->hb_result, S->hb_acc, S->va_part, S->va_result, S->va_acc, S->va_cnt, S->va_order, S->va_ctl, S->to_part, S->to_result, S->to_acc,
                  S->to_sacc, S->to_tc, S->to_task, S->to_ctl, S->to_row, S->f, S->gcsr_cnt, S->gcsr_start, S->gcsr_list,
                  S->bond_mark2, S->counts, S->ws, S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val, S->H_val32,
                  S->Hdia_inv, S->b_s, S->st, S->s_hist, S->t_hist, S->cgv, S->cgs, S->partials, S->counters,
                  S->status, S->scan_tmp, S->tmp_i, S->mixw, S->cand, S->cand_count, S->row_sys, S->tile_beg,
                  S->tile_sys, S->sys_tile, S->sblk_beg, S->sys_sblk, S->sys_iters, S->done_count, S->qsum,
                  S->mark_chg};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (S->status_host) (void)hipHostFree(S->status_host);
  if (S->pin_done) (void)hipHostFree(S->pin_done);
  if (S->pin_mix) (void)hipHostFree(S->pin_mix);
  if (S->pin_iters) (void)hipHostFree(S->pin_iters);
  if (S->mark_chg_host) (void)hipHostFree(S->mark_chg_host);
  if (S->stream) (void)hipStreamDestroy(S->stream);
  delete S;
  return SWR_OK;
}

int swr_set_atoms(swr_system *S, int n, int N, const int *tag, const int *type, const double *x, const double *q,
                  const int *owner, const signed char *img) {
  if (!S || n < 0 || N < n || !tag || !type || !x) {
    swr_set_error_msg("swr_set_atoms: bad arguments");
    return SWR_ERR_ARG;
  }
  if (int rc = swr_quiesce(S)) return rc;
  for (int i = 0; i < N; i++)
    if (type[i] < 1 || type[i] > S->hp.ntypes) {
      swr_set_error_msg("swr_set_atoms: atom type out of range");
      return SWR_ERR_ARG;
    }
  int rc;
  if ((rc = swr_nbr_ensure_atoms(S, std::max(N, 1))) || (rc = ensure_local(S, std::max(n, 1)))) return rc;
  if (n != S->n) {
    S->qeq_initialized = 0;
    S->sys_row.clear();  // a new atom set is one system until swr_set_batch
  }
  S->layout_dirty = true;
  S->next_totals_known = false;
  S->vtotal_local = -1;
  S->gcsr_valid = false;
  S->have_f = false;
  S->n = n;
  S->N = N;
  HIP_CHECK(hipMemcpyAsync(S->x, x, sizeof(double) * 3 * N, hipMemcpyHostToDevice, S->stream));
  HIP_CHECK(hipMemcpyAsync(S->tag, tag, sizeof(int) * N, hipMemcpyHostToDevice, S->stream));
  HIP_CHECK(hipMemcpyAsync(S->type, type, sizeof(int) * N, hipMemcpyHostToDevice, S->stream));
  if (q) HIP_CHECK(hipMemcpyAsync(S->q, q, sizeof(double) * N, hipMemcpyHostToDevice, S->stream));
  else HIP_CHECK(hipMemsetAsync(S->q, 0, sizeof(double) * N, S->stream));
  std::vector<int> own(N);
  std::vector<signed char> im(3 * (size_t)N, 0);
  for (int i = 0; i < N; i++) own[i] = owner ? owner[i] : i;
  if (img) std::memcpy(im.data(), img, 3 * (size_t)N);
  for (int i = n; i < N; i++)
    if (own[i] < 0 || own[i] >= n) {
      swr_set_error_msg("swr_set_atoms: ghost owner out of range");
      return SWR_ERR_ARG;
    }
  HIP_CHECK(hipMemcpyAsync(S->owner, own.data(), sizeof(int) * N, hipMemcpyHostToDevice, S->stream));
  HIP_CHECK(hipMemcpyAsync(S->img, im.data(), 3 * (size_t)N, hipMemcpyHostToDevice, S->stream));
  HIP_CHECK(hipStreamSynchronize(S->stream));
  return SWR_OK;
}

int swr_new_atoms(swr_system *S) {
  if (!S) return SWR_ERR_ARG;
  if (int rc = swr_quiesce(S)) return rc;
  if (S->ncap > 0) {
    HIP_CHECK(hipMemsetAsync(S->s_hist, 0, sizeof(double) * 5 * S->ncap, S->stream));
    HIP_CHECK(hipMemsetAsync(S->t_hist, 0, sizeof(double) * 5 * S->ncap, S->stream));
    HIP_CHECK(hipMemsetAsync(S->st, 0, sizeof(double) * 2 * S->ncap, S->stream));
  }
  S->capacities_set = false;
  S->next_totals_known = false;
  S->h_sys_iters.assign(S->h_sys_iters.size(), 0);
  return SWR_OK;
}

int swr_set_capacities(swr_system *S, const int *num_bonds, const int *num_hbonds) {
  if (!S || !num_bonds || !num_hbonds) return SWR_ERR_ARG;
  if (int rc = swr_quiesce(S)) return rc;
  HIP_CHECK(hipMemcpyAsync(S->num_bonds, num_bonds, sizeof(int) * S->N, hipMemcpyHostToDevice, S->stream));
  HIP_CHECK(hipMemcpyAsync(S->num_hbonds, num_hbonds, sizeof(int) * S->N, hipMemcpyHostToDevice, S->stream));
  HIP_CHECK(hipStreamSynchronize(S->stream));
  S->capacities_set = true;
  S->next_totals_known = false;
  return SWR_OK;
}

static int nbr_build_common(swr_system *S, double cutneigh, int local_only) {
  if (!S || S->N <= 0) return SWR_ERR_STATE;
  if (cutneigh <= 0) {
    double cutmax = std::max(std::max(S->ff.nonb_cut, S->ctl.hbond_cut), 2 * S->ctl.bond_cut);
    cutneigh = cutmax + S->ctl.skin;
  }
  // bin grid over the atoms' bounding box (reduced on the device)
  double lo[3], hi[3];
  int rc = swr_nbr_bbox(S, lo, hi);
  if (rc) return rc;
  rc = swr_nbr_set_grid(S, lo, hi, 0.0, cutneigh);
  if (rc) return rc;
  return swr_nbr_build(S, local_only);
}

int swr_npair_full_bin_ghost_build(swr_system *S, double cutneigh) { return nbr_build_common(S, cutneigh, 0); }

int swr_npair_full_bin_atomonly_build(swr_system *S, double cutneigh) { return nbr_build_common(S, cutneigh, 1); }

int swr_write_reax_atoms_and_pack(swr_system *S) {
  if (!S || S->N <= 0) return SWR_ERR_STATE;
  return swr_lists_pack(S);
}

static int ensure_bond_lists(swr_system *S, int need_bonds, int need_hbonds) {
  int rc;
  if (need_bonds > S->bond_cap) {
    int cap = std::max((int)(need_bonds * S->ctl.safezone), S->ctl.mincap * 25);
    if ((rc = swr_alloc(&S->bonds, cap)) || (rc = swr_alloc(&S->bond_nbr, cap)) ||
        (rc = swr_alloc(&S->bo_data, cap)) || (rc = swr_alloc(&S->BO, cap)) ||
        (rc = swr_alloc(&S->BOpi, 2 * (size_t)cap)) || (rc = swr_alloc(&S->Cdbo, cap)) ||
        (rc = swr_alloc(&S->Cdbopi, cap)) || (rc = swr_alloc(&S->Cdbopi2, cap)))
      return rc;
    S->bond_cap = cap;
  }
  if (need_hbonds > S->hbond_cap) {
    int cap = std::max((int)(need_hbonds * S->ctl.saferzone), S->ctl.mincap * 25);
    if ((rc = swr_alloc(&S->hbonds, cap))) return rc;
    S->hbond_cap = cap;
  }
  return SWR_OK;
}

// Reads the device status block (synchronising) and grows the bond lists
// when the capacity sum approaches the allocation (ReAllocate,
// reaxc_allocate_sunway.cpp:433-549, DANGER_ZONE 0.9).
static int refresh_status(swr_system *S) {
  HIP_CHECK(hipMemcpyAsync(S->status_host, S->status, sizeof(DevStatus), hipMemcpyDeviceToHost, S->stream));
  HIP_CHECK(hipStreamSynchronize(S->stream));
  if (swr_after_sync(S)) {  // a finished relaxation raised a flag after the copy
    HIP_CHECK(hipMemcpyAsync(S->status_host, S->status, sizeof(DevStatus), hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
  }
  return SWR_OK;
}

int swr_reset(swr_system *S) {
  if (!S || S->N <= 0 || !S->vlist) return SWR_ERR_STATE;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;  // Reset rewrites the bond ranges a pending relaxation reads
  if (!S->capacities_set) {
    // setup-time estimate (Estimate_Storages) from a counting pass
    if ((rc = swr_lists_pack(S)) || (rc = swr_lists_estimate(S))) return rc;
    S->capacities_set = true;
    S->next_totals_known = false;
  }
  // The previous step's Validate_Lists wrote the capacities and their totals
  // (read back at the host's next sync, normally QEq's); when they fit, Reset runs
  // without waiting for its own totals.
  const DevStatus &h = *S->status_host;
  if (S->next_totals_known && S->bonds && S->hbonds && h.next_total_bonds <= S->bond_cap * 0.9 &&
      h.next_total_hbonds <= S->hbond_cap * 0.9) {
    S->next_totals_known = false;
    return swr_lists_reset(S);
  }
  S->next_totals_known = false;
  // totals first, so the lists can be sized before they are written
  if ((rc = swr_lists_reset(S)) || (rc = refresh_status(S))) return rc;
  if (h.total_bonds > S->bond_cap * 0.9 || h.total_hbonds > S->hbond_cap * 0.9 || !S->bonds || !S->hbonds) {
    if ((rc = ensure_bond_lists(S, std::max(h.total_bonds, 1), std::max(h.total_hbonds, 1)))) return rc;
    HIP_CHECK(hipMemsetAsync(&S->status->flags, 0, sizeof(unsigned int), S->stream));
    if ((rc = swr_lists_reset(S))) return rc;
  }
  return SWR_OK;
}

int swr_write_reax_lists(swr_system *S) {
  if (!S || !S->vlist) return SWR_ERR_STATE;
  return swr_lists_run(S, 0);
}

int swr_init_forces_noqeq_full(swr_system *S) {
  if (!S || !S->bonds) return SWR_ERR_STATE;
  return swr_lists_run(S, 2);
}

int swr_write_reax_lists_init_forces(swr_system *S) {
  if (!S || !S->vlist || !S->bonds) return SWR_ERR_STATE;
  return swr_lists_run(S, 1);
}

int swr_validate_lists(swr_system *S) {
  if (!S || !S->bonds) return SWR_ERR_STATE;
  return swr_lists_validate(S);
}

int swr_sort_bonds(swr_system *S) {
  if (!S || !S->bonds) return SWR_ERR_STATE;
  return swr_lists_mark(S, MARK_SWEEPS);
}

int swr_bo(swr_system *S) {
  if (!S || !S->bonds) return SWR_ERR_STATE;
  return swr_bo_run(S);
}

int swr_qeq_init_matvec(swr_system *S) {
  if (!S || !S->vlist || S->n <= 0) return SWR_ERR_STATE;
  long long need = S->vtotal_local;
  if (need < 0) {
    HIP_CHECK(hipMemcpyAsync(&need, S->voff + S->n, sizeof(long long), hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
  }
  // H rows start at their Verlet offset + 4 i, rounded up to 4 entries
  // (qeq.hip): the last row ends by (voff[n] + 4 n + 3) & ~3
  need = (need + 4LL * S->n + 3) & ~3LL;
  int rc;
  if (need > S->Hcap) {
    long long cap = (long long)(need * 1.15) + 1024;
    if ((rc = swr_alloc(&S->H_jlist, cap)) || (rc = swr_alloc(&S->H_val, cap))) return rc;
    if (S->mixed && (rc = swr_alloc(&S->H_val32, cap))) return rc;
    S->Hcap = cap;
  }
  if (S->mixed && !S->H_val32 && (rc = swr_alloc(&S->H_val32, S->Hcap))) return rc;
  return swr_qeq_build_H(S);
}

int swr_sparse_matvec(swr_system *S, const double *x_dev, double *b_dev) {
  if (!S || !S->H_val || !x_dev || !b_dev) return SWR_ERR_STATE;
  return swr_qeq_spmv(S, x_dev, b_dev);
}

int swr_sparse_matvec_h(swr_system *S, const double *x_host, double *b_host) {
  if (!S || !S->H_val || !x_host || !b_host) return SWR_ERR_STATE;
  double *xd = nullptr, *bd = nullptr;
  int rc;
  if ((rc = swr_alloc(&xd, S->n)) || (rc = swr_alloc(&bd, S->n))) { swr_free(xd); return rc; }
  hipError_t e = hipMemcpyAsync(xd, x_host, sizeof(double) * S->n, hipMemcpyHostToDevice, S->stream);
  if (e == hipSuccess) rc = swr_qeq_spmv(S, xd, bd);
  if (e == hipSuccess && rc == SWR_OK)
    e = hipMemcpyAsync(b_host, bd, sizeof(double) * S->n, hipMemcpyDeviceToHost, S->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(S->stream);
  swr_free(xd);
  swr_free(bd);
  if (e != hipSuccess) { swr_set_error_msg(hipGetErrorString(e)); return SWR_ERR_HIP; }
  return rc;
}

int swr_qeq_cg(swr_system *S, int *iters_s, int *iters_t) {
  if (!S || !S->H_val) return SWR_ERR_STATE;
  int its0 = 0;
  for (size_t k = 0; k < S->h_sys_iters.size(); k += 2) its0 = std::max(its0, S->h_sys_iters[k]);
  int rc = S->mixed ? swr_qeq_mixed_solve(S) : swr_qeq_cg_run(S, its0 + 2);
  if (rc) return rc;
  if (S->mixed) {  // the FP64 solve's last done check already brought status and counts back
    S->h_sys_iters.resize(2 * (size_t)S->nsys);
    HIP_CHECK(hipMemcpyAsync(S->pin_iters, S->sys_iters, sizeof(int) * 2 * S->nsys, hipMemcpyDeviceToHost,
                             S->stream));
    if ((rc = refresh_status(S))) return rc;   // (synchronises)
    std::memcpy(S->h_sys_iters.data(), S->pin_iters, sizeof(int) * 2 * S->nsys);
  }
  int ms = 0, mt = 0;
  for (int k = 0; k < S->nsys; k++) {
    ms = std::max(ms, S->h_sys_iters[2 * k]);
    mt = std::max(mt, S->h_sys_iters[2 * k + 1]);
  }
  if (iters_s) *iters_s = ms;
  if (iters_t) *iters_t = mt;
  return status_to_rc(S->status_host->flags);
}

int swr_qeq_cg_noalloc(swr_system *S, const double *b, double *x, int *iters_s, int *iters_t) {
  if (!S) return SWR_ERR_ARG;
  return swr_qeq_cg_noalloc_run(S, b, x, iters_s, iters_t);
}

int swr_set_batch(swr_system *S, int nsys, const int *row_offsets) {
  if (!S || nsys < 1 || !row_offsets || row_offsets[0] != 0 || row_offsets[nsys] != S->n) {
    swr_set_error_msg("swr_set_batch: offsets must run from 0 to the local atom count");
    return SWR_ERR_ARG;
  }
  for (int k = 0; k < nsys; k++)
    if (row_offsets[k + 1] < row_offsets[k]) {
      swr_set_error_msg("swr_set_batch: offsets must not decrease");
      return SWR_ERR_ARG;
    }
  S->sys_row.assign(row_offsets, row_offsets + nsys + 1);
  S->layout_dirty = true;
  return swr_qeq_layout(S);
}

int swr_calculate_q(swr_system *S) {
  if (!S || S->n <= 0) return SWR_ERR_STATE;
  return swr_qeq_calc_q(S);
}

int swr_qeq_pre_force(swr_system *S, int *iters_s, int *iters_t) {
  int rc;
  if ((rc = swr_qeq_init_matvec(S))) return rc;
  if ((rc = swr_qeq_cg(S, iters_s, iters_t))) return rc;
  return swr_calculate_q(S);
}

int swr_pair_compute(swr_system *S) {
  int rc;
  if ((rc = swr_write_reax_atoms_and_pack(S)) || (rc = swr_reset(S)) || (rc = swr_write_reax_lists_init_forces(S)) ||
      (rc = swr_validate_lists(S)) || (rc = swr_sort_bonds(S)) || (rc = swr_bo(S)))
    return rc;
  return SWR_OK;
}

int swr_vdw_coulomb_energy(swr_system *S) {
  if (!S) return SWR_ERR_ARG;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  return swr_nonbonded_run(S);
}

int swr_atom_energy(swr_system *S) {
  if (!S) return SWR_ERR_ARG;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  return swr_atom_energy_run(S);
}

int swr_add_dbond_forces(swr_system *S) {
  if (!S) return SWR_ERR_ARG;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  return swr_dbond_forces_run(S);
}

int swr_hydrogen_bonds(swr_system *S) {
  if (!S) return SWR_ERR_ARG;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  return swr_hydrogen_bonds_run(S);
}

int swr_valence_angles(swr_system *S) {
  if (!S) return SWR_ERR_ARG;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  return swr_valence_angles_run(S);
}

int swr_torsion_angles(swr_system *S) {
  if (!S) return SWR_ERR_ARG;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  return swr_torsion_angles_run(S);
}

int swr_compute_forces(swr_system *S) {
  if (!S) return SWR_ERR_ARG;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  return swr_forces_run(S);
}

int swr_set_masses(swr_system *S, int ntypes, const double *mass) {
  if (!S || !mass || ntypes != S->hp.ntypes) {
    swr_set_error_msg("swr_set_masses: needs mass[1..ntypes] for every atom type");
    return SWR_ERR_ARG;
  }
  for (int t = 1; t <= ntypes; t++)
    if (!(mass[t] > 0.0)) {
      swr_set_error_msg("Invalid mass value");
      return SWR_ERR_ARG;
    }
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  for (int t = 1; t <= ntypes; t++) S->hp.type_mass[t] = mass[t];
  HIP_CHECK(hipMemcpyAsync(S->dp, &S->hp, sizeof(DevParams), hipMemcpyHostToDevice, S->stream));
  HIP_CHECK(hipStreamSynchronize(S->stream));  // S->hp may change again before the copy ran
  return SWR_OK;
}

int swr_fix_nve_initial_integrate(swr_system *S, double dtv, double dtf) {
  if (!S) return SWR_ERR_ARG;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  return swr_nve_run(S, 0, dtv, dtf);
}

int swr_fix_nve_final_integrate(swr_system *S, double dtv, double dtf) {
  if (!S) return SWR_ERR_ARG;
  int rc;
  if ((rc = swr_quiesce(S))) return rc;
  return swr_nve_run(S, 1, dtv, dtf);
}

int swr_sync(swr_system *S) {
  if (!S) return SWR_ERR_ARG;
  int rc = refresh_status(S);
  if (rc) return rc;
  unsigned int f = S->status_host->flags;
  if (f) {
    S->sticky_flags |= f;
    HIP_CHECK(hipMemsetAsync(&S->status->flags, 0, sizeof(unsigned int), S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
  }
  return status_to_rc(f);
}

int swr_device_synchronize(void) {
  HIP_CHECK(hipDeviceSynchronize());
  return SWR_OK;
}

// ---------------------------------------------------------------- mirrors
namespace {

struct View {
  const void *ptr;
  int64_t bytes;
  int stride_kind;  // 0 plain, 1 de-interleave .x of double2, 2 de-interleave .y
};

bool lookup(swr_system *S, const std::string &k, View &v) {
  const int64_t N = S->N, n = S->n;
  const size_t c = S->Ncap;
  auto set = [&](const void *p, int64_t b, int kind = 0) { v.ptr = p; v.bytes = b; v.stride_kind = kind; return true; };
  if (k == "x") return set(S->x, 24 * N);
  if (k == "q") return set(S->q, 8 * N);
  if (k == "tag") return set(S->tag, 4 * N);
  if (k == "type") return set(S->type, 4 * N);
  if (k == "owner") return set(S->owner, 4 * N);
  if (k == "img") return set(S->img, 3 * N);
  if (k == "v") return set(S->v, 24 * n);
  if (k == "packed_atoms") return set(S->packed, sizeof(swr_atom_pack_t) * N);
  if (k == "verlet_offsets") return set(S->voff, 8 * (N + 1));
  if (k == "verlet_list") return set(S->vlist, 4 * S->vtotal);
  if (k == "far") return set(S->far, sizeof(swr_far_neighbor_t) * S->vtotal);
  if (k == "far_index") return set(S->far_index, 4 * N);
  if (k == "far_end") return set(S->far_end, 4 * N);
  if (k == "num_bonds") return set(S->num_bonds, 4 * N);
  if (k == "num_hbonds") return set(S->num_hbonds, 4 * N);
  if (k == "Hindex") return set(S->Hindex, 4 * N);
  if (k == "bond_index") return set(S->bond_index, 4 * N);
  if (k == "bond_end") return set(S->bond_end, 4 * N);
  if (k == "bonds") return set(S->bonds, sizeof(swr_bond_t) * (int64_t)S->bond_cap);
  if (k == "bo_data") return set(S->bo_data, sizeof(swr_bond_order_t) * (int64_t)S->bond_cap);
  if (k == "BO") return set(S->BO, 8 * (int64_t)S->bond_cap);
  if (k == "BOpi") return set(S->BOpi, 16 * (int64_t)S->bond_cap);
  if (k == "Cdbo") return set(S->Cdbo, 8 * (int64_t)S->bond_cap);
  if (k == "Cdbopi") return set(S->Cdbopi, 8 * (int64_t)S->bond_cap);
  if (k == "Cdbopi2") return set(S->Cdbopi2, 8 * (int64_t)S->bond_cap);
  if (k == "hbond_index") return set(S->hbond_index, 4 * (int64_t)S->status_host->numH);
  if (k == "hbond_end") return set(S->hbond_end, 4 * (int64_t)S->status_host->numH);
  if (k == "hbonds") return set(S->hbonds, sizeof(swr_hbond_t) * (int64_t)S->hbond_cap);
  if (k == "bo_dboc") return set(S->bo_dboc, 16 * N);
  if (k == "dDeltap_self") return set(S->dDeltap_self, 24 * N);
  if (k == "bond_mark") return set(S->bond_mark, 4 * N);
  if (k == "fCdDelta") return set(S->fcd, 32 * N);
  if (k == "nonbonded_energy") return set(S->nb_result, S->nb_result ? 64 : 0);
  if (k == "nonbonded_eatom") return set(S->nb_eatom, S->nb_eatom ? 8 * n : 0);
  if (k == "bonded_energy") return set(S->be_result, S->be_result ? 40 : 0);
  if (k == "bonded_eatom") return set(S->be_part, S->be_part ? 40 * (int64_t)S->be_stride : 0);
  if (k == "hbond_energy") return set(S->hb_result, S->hb_result ? 8 : 0);
  if (k == "hbond_eatom") return set(S->hb_part, S->hb_part ? 8 * n : 0);
  if (k == "f") return set(S->f, S->have_f ? 24 * n : 0);
  if (k == "valence_energy") return set(S->va_result, S->va_result ? 24 : 0);
  if (k == "valence_eatom") return set(S->va_part, S->va_part ? 24 * (int64_t)S->va_stride : 0);
  if (k == "torsion_energy") return set(S->to_result, S->to_result ? 16 : 0);
  if (k == "torsion_eatom") return set(S->to_part, S->to_part ? 16 * (int64_t)S->to_stride : 0);
  static const char *wsn[WS_COUNT] = {"Deltap", "Deltap_boc", "Delta", "Delta_e", "Delta_val", "vlpex", "nlp",
                                      "nlp_temp", "Clp", "Delta_lp", "dDelta_lp", "Delta_lp_temp", "dDelta_lp_temp"};
  for (int w = 0; w < WS_COUNT; w++)
    if (k == wsn[w]) return set(S->ws + (size_t)w * c, 8 * N);
  if (k == "H_firstnbr") return set(S->H_firstnbr, 4 * n);
  if (k == "H_numnbrs") return set(S->H_numnbrs, 4 * n);
  if (k == "H_jlist") return set(S->H_jlist, 4 * S->Hcap);
  if (k == "H_val") return set(S->H_val, 8 * S->Hcap);
  if (k == "H_val32") return set(S->H_val32, S->H_val32 ? 4 * S->Hcap : 0);
  if (k == "Hdia_inv") return set(S->Hdia_inv, 8 * n);
  if (k == "panel_rows") return set(S->prow, S->prow ? 4 * (int64_t)S->prow_n : 0);
  if (k == "panel_begin") return set(S->pbeg, S->pbeg ? 4 * ((int64_t)S->npanel + 1) : 0);
  if (k == "panel_count") return set(S->pcnt, S->pcnt ? 4 * (int64_t)S->npanel : 0);
  if (k == "panel_columns") return set(S->pcol, S->pcol ? 4 * (int64_t)S->npanel * SWR_PANEL_MAXC : 0);
  if (k == "H_jloc") return set(S->jloc, S->jloc ? 2 * S->Hcap : 0);
  if (k == "b_s") return set(S->b_s, 8 * n, 1);
  if (k == "b_t") return set(S->b_s, 8 * n, 2);
  if (k == "s") return set(S->st, 8 * n, 1);
  if (k == "t") return set(S->st, 8 * n, 2);
  if (k == "s_hist") return set(S->s_hist, 40 * n);
  if (k == "bo_radius") return set(S->hp.rbo, sizeof(S->hp.rbo), 3);
  if (k == "dev_params") return set(S->dp, sizeof(DevParams));
  if (k == "sys_iters") return set(S->h_sys_iters.data(), 4 * (int64_t)S->h_sys_iters.size(), 3);
  if (k == "t_hist") return set(S->t_hist, 40 * n);
  return false;
}

}  // namespace

int64_t swr_get(swr_system *S, const char *name, void *out) {
  if (!S || !name) return -1;
  std::string k = name;
  if (k == "hbond_far") {
    // far-list slot of each hbond (ptr rebased onto the far array)
    int64_t cnt = S->hbond_cap;
    if (!out) return 8 * cnt;
    std::vector<swr_hbond_t> hb(cnt);
    if (cnt) {
      HIP_CHECK(hipMemcpyAsync(hb.data(), S->hbonds, sizeof(swr_hbond_t) * cnt, hipMemcpyDeviceToHost, S->stream));
      HIP_CHECK(hipStreamSynchronize(S->stream));
    }
    int64_t *o = (int64_t *)out;
    for (int64_t i = 0; i < cnt; i++) {
      intptr_t off = (intptr_t)hb[i].ptr - (intptr_t)S->far;
      o[i] = (hb[i].ptr && off >= 0) ? (int64_t)(off / (intptr_t)sizeof(swr_far_neighbor_t)) : -1;
    }
    return 8 * cnt;
  }
  View v;
  if (!lookup(S, k, v)) return -1;
  if (!out || v.bytes <= 0 || !v.ptr) return v.bytes < 0 ? 0 : v.bytes;
  (void)hipStreamSynchronize(S->stream);
  (void)swr_after_sync(S);  // a pending bond_mark relaxation finishes before anything is read
  if (v.stride_kind == 3) {  // host-side table
    std::memcpy(out, v.ptr, v.bytes);
  } else if (v.stride_kind == 0) {
    HIP_CHECK(hipMemcpyAsync(out, v.ptr, v.bytes, hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
  } else {
    int64_t cnt = v.bytes / 8;
    std::vector<double> tmp(2 * cnt);
    HIP_CHECK(hipMemcpyAsync(tmp.data(), v.ptr, 16 * cnt, hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
    double *o = (double *)out;
    for (int64_t i = 0; i < cnt; i++) o[i] = tmp[2 * i + (v.stride_kind - 1)];
  }
  return v.bytes;
}

int swr_put(swr_system *S, const char *name, const void *in, int64_t bytes) {
  if (!S || !name || !in) return SWR_ERR_ARG;
  std::string k = name;
  View v;
  if (!lookup(S, k, v) || bytes > v.bytes || v.stride_kind != 0) return SWR_ERR_ARG;
  if (int rc = swr_quiesce(S)) return rc;
  HIP_CHECK(hipStreamSynchronize(S->stream));
  HIP_CHECK(hipMemcpyAsync((void *)v.ptr, in, bytes, hipMemcpyHostToDevice, S->stream));
  HIP_CHECK(hipStreamSynchronize(S->stream));
  if (k == "num_bonds") S->capacities_set = true;
  S->next_totals_known = false;
  return SWR_OK;
}

void *swr_device_ptr(swr_system *S, const char *name) {
  if (!S || !name) return nullptr;
  View v;
  if (!lookup(S, name, v)) return nullptr;
  return (void *)v.ptr;
}

int64_t swr_get_scalar(swr_system *S, const char *name) {
  if (!S || !name) return -1;
  std::string k = name;
  if (k == "n") return S->n;
  if (k == "N") return S->N;
  if (k == "verlet_total") return S->vtotal;
  if (k == "bond_cap") return S->bond_cap;
  if (k == "hbond_cap") return S->hbond_cap;
  if (refresh_status(S)) return -1;
  const DevStatus &h = *S->status_host;
  if (k == "numH") return h.numH;
  if (k == "total_bonds") return h.total_bonds;
  if (k == "total_hbonds") return h.total_hbonds;
  if (k == "need_bond_cap") return h.need_bond_cap;
  if (k == "need_hbond_cap") return h.need_hbond_cap;
  if (k == "flags") return h.flags | S->sticky_flags;
  if (k == "matvecs_s" || k == "matvecs_t") {
    int m = 0;
    for (size_t j = (k == "matvecs_s" ? 0 : 1); j < S->h_sys_iters.size(); j += 2) m = std::max(m, S->h_sys_iters[j]);
    return m;
  }
  if (k == "nsys") return S->nsys;
  if (k == "mixed_outer") return S->mixed_outer;
  if (k == "bonded_stride") return (int64_t)S->be_stride;
  // first outer residual of the last mixed solve, sqrt(u.r)/|b| x 1e15
  if (k == "mixed_rho0_s_x1e15") return (int64_t)(S->mixed_rho0[0] * 1e15);
  if (k == "mixed_rho0_t_x1e15") return (int64_t)(S->mixed_rho0[1] * 1e15);
  return -1;
}

}  // extern "C"
