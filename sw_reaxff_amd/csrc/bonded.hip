// bonded.hip — bond and atom energies and the bond-order chain rule of the
// forces:
//
//   Atom_Energy (reaxc_multi_body_sunway.cpp:41-57) = Merge_Bonds_Atom_Energy_C_New
//   (reaxc_multi_body_sw64.c:21-337): per local atom the polarization, lone-pair
//   (+ C-C correction), over- and under-coordination energies, and for each
//   bond the atom owns the bond energy and the terminal triple-bond
//   stabilisation; the coefficients Cdbo/Cdbopi/Cdbopi2 of each bond slot and
//   CdDelta (fCdDelta[.][3]) of the atoms.  (Bonds() itself returns at once,
//   reaxc_bonds_sunway.cpp:49-52.)
//
//   Add_All_dBond_to_Forces_C (reaxc_forces_sw64.c:417-589): over all N atoms
//   and their bonds, the chain rule of the bond-order derivatives into
//   fCdDelta[.][0..2] (the energy gradient, as the non-bonded pass writes it).
//
// The reference loops over atoms and adds into the neighbours' entries as it
// goes.  Here every kernel writes only entries of its own atom, so no
// floating-point atomics are needed and results are the same on every run:
//
//   k_atom_energy     one thread per local atom: energies (per-atom rows,
//                     summed in a fixed order by k_sum_rows), its own bond
//                     slots' coefficients, its own CdDelta, and the two
//                     per-atom factors (CEover4, CEunder4) its neighbours need;
//   k_atom_energy_nbr one thread per atom (local and ghost): the CdDelta terms
//                     the reference adds to atom j from each bonded local atom
//                     i, read through the bond's reverse slot (sym_index), so
//                     the values are those i's own slot holds;
//   k_dbond_own       one thread per atom: its own chain-rule terms and the
//                     factor buf_c it passes to its neighbours;
//   k_dbond_nbr       one thread per atom: sum over its bonds of the
//                     neighbour's buf_c times the neighbour's dBOp toward it.
//
// Each atom's terms are added in the reference's order for that atom; only
// the interleaving of its own and its neighbours' terms differs, which moves
// results at the rounding level.
#include "swr_internal.h"
#include "swr_device.h"

#include <algorithm>

namespace {

constexpr int BT = 256;
constexpr double KCALpMOL_to_EV = 23.02;   // reaxc_defs_sunway.h:68

__device__ __forceinline__ double sqr(double x) { return x * x; }

// Local atom i owns bond (i, j) for the bond energy: the smaller orig_id, or
// for two images of one atom the lower in z, then y, then x (:245-255).
__device__ __forceinline__ bool owns_bond(const swr_atom_pack_t &a, const swr_atom_pack_t &b) {
  if (a.orig_id > b.orig_id) return false;
  if (a.orig_id == b.orig_id) {
    if (b.x[2] < a.x[2]) return false;
    if (b.x[2] == a.x[2] && b.x[1] < a.x[1]) return false;
    if (b.x[2] == a.x[2] && b.x[1] == a.x[1] && b.x[0] < a.x[0]) return false;
  }
  return true;
}

// Terminal triple-bond stabilisation of bond (i, j) owned by i (:283-308):
// returns whether it applies; dbo, dua, dub are the terms for Cdbo of i's
// slot, CdDelta of i and CdDelta of j; e the energy.
__device__ __forceinline__ bool triple_bond(const DevParams *__restrict__ P, int ti, int tj, double BOij,
                                            double dboc_i, double dboc_j, double Delta_i, double Delta_j, double &e,
                                            double &dbo, double &dua, double &dub) {
  const double mi = P->mass[ti], mj = P->mass[tj];
  if (!(BOij >= 1.00 && (P->gp37 == 2 || (mi == 12.0000 && mj == 15.9990) || (mj == 12.0000 && mi == 15.9990))))
    return false;
  const double gp3 = P->gp3, gp4 = P->gp4, gp7 = P->gp7, gp10 = P->gp10;
  const double exphu = exp(-gp7 * sqr(BOij - 2.50));
  const double exphua1 = exp(-gp3 * (dboc_i - BOij));
  const double exphub1 = exp(-gp3 * (dboc_j - BOij));
  const double exphuov = exp(gp4 * (Delta_i + Delta_j));
  const double hulpov = 1.0 / (1.0 + 25.0 * exphuov);
  e = gp10 * exphu * hulpov * (exphua1 + exphub1);
  dbo = gp10 * exphu * hulpov * (exphua1 + exphub1) * (gp3 - 2.0 * gp7 * (BOij - 2.50));
  dua = -gp10 * exphu * hulpov * (gp3 * exphua1 + 25.0 * gp4 * exphuov * hulpov * (exphua1 + exphub1));
  dub = -gp10 * exphu * hulpov * (gp3 * exphub1 + 25.0 * gp4 * exphuov * hulpov * (exphua1 + exphub1));
  return true;
}

struct BeArgs {
  const DevParams *P;
  const swr_atom_pack_t *atoms;
  const int *bidx, *bend;
  const swr_bond_t *bonds;
  const swr_bond_order_t *bo;
  const double *BO, *BOpi;       // BOpi: [slot][2]
  const double *Delta, *Delta_lp, *dDelta_lp, *Delta_lp_temp;
  const double *bo_dboc;         // [N][2]
  double *Cdbo, *Cdbopi, *Cdbopi2;
  double *fcd;                   // [N][4]
  double *ce;                    // [2][stride]: CEover4, CEunder4
  double *part;                  // [5][stride]
  size_t stride;
};

__global__ void __launch_bounds__(BT) k_atom_energy(int n, BeArgs A) {
  const int i = blockIdx.x * BT + threadIdx.x;
  if (i >= n) return;
  const DevParams *__restrict__ P = A.P;
  const swr_atom_pack_t ai = A.atoms[i];
  const int ti = ai.type;
  double e_bond = 0, e_lp = 0, e_ov = 0, e_un = 0, e_pol = 0;
  double ceo4 = 0, ceu4 = 0;
  if (ti >= 0) {
    const double q = ai.q;
    e_pol = KCALpMOL_to_EV * (P->chi_e[ti] * q + (P->eta_e[ti] / 2.) * sqr(q));
    const int b0 = A.bidx[i], b1 = A.bend[i];
    // lone pair
    const double p_lp2 = P->p_lp2[ti];
    const double Dlp = A.Delta_lp[i];
    const double expvd2 = exp(-75 * Dlp);
    const double inv_expvd2 = 1. / (1. + expvd2);
    const double p_ovun2 = P->p_ovun2[ti];
    const double dfvl = P->mass[ti] > 21.0 ? 0.0 : 1.0;   // 1st-row elements only
    int numbonds = 0;
    double sum_ovun1 = 0, sum_ovun2 = 0;
    for (int pj = b0; pj < b1; ++pj) {
      numbonds++;
      const int j = A.bonds[pj].nbr;
      const int tj = A.atoms[j].type;
      if (tj < 0) continue;
      const DevBe &tw = P->be[ti * SWR_MAX_ELEM + tj];
      sum_ovun1 += tw.p_ovun1 * tw.De_s * A.BO[pj];
      sum_ovun2 += (A.Delta[j] - dfvl * A.Delta_lp_temp[j]) * (A.BOpi[2 * (size_t)pj] + A.BOpi[2 * (size_t)pj + 1]);
    }
    const double dElp = p_lp2 * inv_expvd2 + 75 * p_lp2 * Dlp * expvd2 * sqr(inv_expvd2);
    const double CElp = dElp * A.dDelta_lp[i];
    double cdd = A.fcd[4 * (size_t)i + 3];
    const double Di = A.Delta[i];
    if (P->p_lp3 > 0.001 && P->is_carbon[ti]) {
      for (int pj = b0; pj < b1; ++pj) {
        const int j = A.bonds[pj].nbr;
        const int tj = A.atoms[j].type;
        if (tj < 0 || !P->is_carbon[tj]) continue;
        const double vov3 = A.BO[pj] - Di - 0.040 * pow(Di, 4.);
        if (vov3 > 3.) {
          e_lp += P->p_lp3 * sqr(vov3 - 3.0);
          const double deahu2dbo = 2. * P->p_lp3 * (vov3 - 3.);
          const double deahu2dsbo = 2. * P->p_lp3 * (vov3 - 3.) * (-1. - 0.16 * pow(Di, 3.));
          A.Cdbo[pj] += deahu2dbo;
          cdd += deahu2dsbo;
        }
      }
    }
    // over-coordination
    const double p_ovun4 = P->p_ovun4;
    const double exp_ovun1 = P->p_ovun3 * exp(p_ovun4 * sum_ovun2);
    const double inv_exp_ovun1 = 1.0 / (1 + exp_ovun1);
    const double Dlpt = A.Delta_lp_temp[i];
    const double Delta_lpcorr = Di - (dfvl * Dlpt) * inv_exp_ovun1;
    const double exp_ovun2 = exp(p_ovun2 * Delta_lpcorr);
    const double inv_exp_ovun2 = 1.0 / (1.0 + exp_ovun2);
    const double DlpVi = 1.0 / (Delta_lpcorr + P->valency[ti] + 1e-8);
    const double CEover1 = Delta_lpcorr * DlpVi * inv_exp_ovun2;
    e_ov = sum_ovun1 * CEover1;
    const double CEover2 =
        sum_ovun1 * DlpVi * inv_exp_ovun2 * (1.0 - Delta_lpcorr * (DlpVi + p_ovun2 * exp_ovun2 * inv_exp_ovun2));
    const double dDlp = A.dDelta_lp[i];
    const double CEover3 = CEover2 * (1.0 - dfvl * dDlp * inv_exp_ovun1);
    ceo4 = CEover2 * (dfvl * Dlpt) * p_ovun4 * exp_ovun1 * sqr(inv_exp_ovun1);
    // under-coordination
    const double p_ovun5 = P->p_ovun5[ti];
    const double exp_ovun2n = 1.0 / exp_ovun2;
    const double exp_ovun6 = exp(P->p_ovun6 * Delta_lpcorr);
    const double exp_ovun8 = P->p_ovun7 * exp(P->p_ovun8 * sum_ovun2);
    const double inv_exp_ovun2n = 1.0 / (1.0 + exp_ovun2n);
    const double inv_exp_ovun8 = 1.0 / (1.0 + exp_ovun8);
    cdd += CEover3;
    const bool under = numbonds > 0 || P->enobonds;
    double eu = 0.0;
    if (under) eu = -p_ovun5 * (1.0 - exp_ovun6) * inv_exp_ovun2n * inv_exp_ovun8;
    const double CEunder1 =
        inv_exp_ovun2n * (p_ovun5 * P->p_ovun6 * exp_ovun6 * inv_exp_ovun8 + p_ovun2 * eu * exp_ovun2n);
    const double CEunder2 = -eu * P->p_ovun8 * exp_ovun8 * inv_exp_ovun8;
    const double CEunder3 = CEunder1 * (1.0 - dfvl * dDlp * inv_exp_ovun1);
    ceu4 = CEunder1 * (dfvl * Dlpt) * p_ovun4 * exp_ovun1 * sqr(inv_exp_ovun1) + CEunder2;
    if (under) {
      e_un = eu;
      cdd += CEunder3;
      cdd += CElp;
      e_lp += p_lp2 * Dlp * inv_expvd2;
    }
    // per bond: over/under-coordination coefficients of this slot, bond
    // energy and triple-bond stabilisation of the bonds this atom owns
    for (int pj = b0; pj < b1; ++pj) {
      const int j = A.bonds[pj].nbr;
      const swr_atom_pack_t aj = A.atoms[j];
      const int tj = aj.type;
      if (tj < 0) continue;
      const DevBe &tw = P->be[ti * SWR_MAX_ELEM + tj];
      double cdbo = A.Cdbo[pj], cdbopi = A.Cdbopi[pj], cdbopi2 = A.Cdbopi2[pj];
      cdbo += CEover1 * tw.p_ovun1 * tw.De_s;
      const double dj = A.Delta[j] - dfvl * A.Delta_lp_temp[j];
      const double ftmp = ceo4 * dj;
      cdbopi += ftmp;
      cdbopi2 += ftmp;
      const double fu = ceu4 * dj;
      cdbopi += fu;
      cdbopi2 += fu;
      if (owns_bond(ai, aj)) {
        const double BO_s = A.bo[pj].BO_s;
        const double pow_BOs_be2 = BO_s == 0.0 ? 0.0 : pow(BO_s, tw.p_be2);
        const double exp_be12 = exp(tw.p_be1 * (1.0 - pow_BOs_be2));
        const double CEbo = -tw.De_s * exp_be12 * (1.0 - tw.p_be1 * tw.p_be2 * pow_BOs_be2);
        e_bond += -tw.De_s * BO_s * exp_be12 - tw.De_p * A.BOpi[2 * (size_t)pj] - tw.De_pp * A.BOpi[2 * (size_t)pj + 1];
        cdbo += CEbo;
        cdbopi -= (CEbo + tw.De_p);
        cdbopi2 -= (CEbo + tw.De_pp);
        double e3, dbo, dua, dub;
        if (triple_bond(P, ti, tj, A.BO[pj], A.bo_dboc[2 * (size_t)i], A.bo_dboc[2 * (size_t)j], Di, A.Delta[j], e3,
                        dbo, dua, dub)) {
          e_bond += e3;
          cdbo += dbo;
          cdd += dua;
        }
      }
      A.Cdbo[pj] = cdbo;
      A.Cdbopi[pj] = cdbopi;
      A.Cdbopi2[pj] = cdbopi2;
    }
    A.fcd[4 * (size_t)i + 3] = cdd;
  }
  A.ce[i] = ceo4;
  A.ce[A.stride + i] = ceu4;
  A.part[i] = e_bond;
  A.part[A.stride + i] = e_lp;
  A.part[2 * A.stride + i] = e_ov;
  A.part[3 * A.stride + i] = e_un;
  A.part[4 * A.stride + i] = e_pol;
}

// CdDelta terms of atom j from each bonded local atom i (:270-279 and the
// triple-bond term :305), read through i's slot toward j.
__global__ void __launch_bounds__(BT) k_atom_energy_nbr(int n, int N, BeArgs A) {
  const int j = blockIdx.x * BT + threadIdx.x;
  if (j >= N) return;
  const DevParams *__restrict__ P = A.P;
  const swr_atom_pack_t aj = A.atoms[j];
  const int tj = aj.type;
  if (tj < 0) return;
  const int b0 = A.bidx[j], b1 = A.bend[j];
  if (b0 >= b1) return;
  double cdd = A.fcd[4 * (size_t)j + 3];
  const double dDlp_j = A.dDelta_lp[j];
  for (int pk = b0; pk < b1; ++pk) {
    const swr_bond_t b = A.bonds[pk];
    const int i = b.nbr;
    if (i >= n) continue;
    const swr_atom_pack_t ai = A.atoms[i];
    const int ti = ai.type;
    if (ti < 0) continue;
    const int pj = b.sym_index;   // i's slot toward j
    const double dfvl = P->mass[ti] > 21.0 ? 0.0 : 1.0;
    const double ftmp2 = (1.0 - dfvl * dDlp_j) * (A.BOpi[2 * (size_t)pj] + A.BOpi[2 * (size_t)pj + 1]);
    cdd += A.ce[i] * ftmp2;
    cdd += A.ce[A.stride + i] * ftmp2;
    if (owns_bond(ai, aj)) {
      double e3, dbo, dua, dub;
      if (triple_bond(P, ti, tj, A.BO[pj], A.bo_dboc[2 * (size_t)i], A.bo_dboc[2 * (size_t)j], A.Delta[i],
                      A.Delta[j], e3, dbo, dua, dub))
        cdd += dub;
    }
  }
  A.fcd[4 * (size_t)j + 3] = cdd;
}

// Fixed-order sums of the per-atom energy rows: block r sums row r.
__global__ void __launch_bounds__(1024) k_sum_rows(int n, const double *__restrict__ part, size_t stride,
                                                   double *__restrict__ out) {
  __shared__ double s[1024];
  const double *row = part + blockIdx.x * stride;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) acc += row[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

struct DbArgs {
  const swr_bond_t *bonds;
  const swr_bond_order_t *bo;
  const int *bidx, *bend;
  const double *Cdbo, *Cdbopi, *Cdbopi2;
  const double *dDeltap_self;    // [N][3]
  double *fcd;                   // [N][4]
  double *bufc;                  // [N]
};

// The atom's own terms (the reference's forces on i, :532-548) and buf_c.
__global__ void __launch_bounds__(BT) k_dbond_own(int N, DbArgs A) {
  const int i = blockIdx.x * BT + threadIdx.x;
  if (i >= N) return;
  const int b0 = A.bidx[i], b1 = A.bend[i];
  double buf_c = 0.0;
  if (b0 < b1) {
    const double ds0 = A.dDeltap_self[3 * (size_t)i], ds1 = A.dDeltap_self[3 * (size_t)i + 1],
                 ds2 = A.dDeltap_self[3 * (size_t)i + 2];
    const double cd_i = A.fcd[4 * (size_t)i + 3];
    double f0 = A.fcd[4 * (size_t)i], f1 = A.fcd[4 * (size_t)i + 1], f2 = A.fcd[4 * (size_t)i + 2];
    for (int pj = b0; pj < b1; ++pj) {
      const swr_bond_t b = A.bonds[pj];
      const int j = b.nbr, sym = b.sym_index;
      const swr_bond_order_t &bo = A.bo[pj];
      const double cdbo = A.Cdbo[pj] + A.Cdbo[sym];
      const double cdbopi = A.Cdbopi[pj] + A.Cdbopi[sym];
      const double cdbopi2 = A.Cdbopi2[pj] + A.Cdbopi2[sym];
      const double C1dbo = bo.C1dbo * cdbo, C2dbo = bo.C2dbo * cdbo;
      const double C1dbopi = bo.C1dbopi * cdbopi, C2dbopi = bo.C2dbopi * cdbopi, C3dbopi = bo.C3dbopi * cdbopi;
      const double C1dbopi2 = bo.C1dbopi2 * cdbopi2, C2dbopi2 = bo.C2dbopi2 * cdbopi2,
                   C3dbopi2 = bo.C3dbopi2 * cdbopi2;
      const double cdd = cd_i + A.fcd[4 * (size_t)j + 3];
      const double C1dDelta = bo.C1dbo * cdd, C2dDelta = bo.C2dbo * cdd;
      const double dbx = bo.dBOp[0], dby = bo.dBOp[1], dbz = bo.dBOp[2];
      double t0 = C1dbo * dbx, t1 = C1dbo * dby, t2 = C1dbo * dbz;
      t0 += C2dbo * ds0; t1 += C2dbo * ds1; t2 += C2dbo * ds2;
      t0 += C1dDelta * dbx; t1 += C1dDelta * dby; t2 += C1dDelta * dbz;
      t0 += C2dDelta * ds0; t1 += C2dDelta * ds1; t2 += C2dDelta * ds2;
      t0 += C1dbopi * bo.dln_BOp_pi[0]; t1 += C1dbopi * bo.dln_BOp_pi[1]; t2 += C1dbopi * bo.dln_BOp_pi[2];
      t0 += C2dbopi * dbx; t1 += C2dbopi * dby; t2 += C2dbopi * dbz;
      t0 += C3dbopi * ds0; t1 += C3dbopi * ds1; t2 += C3dbopi * ds2;
      t0 += C1dbopi2 * bo.dln_BOp_pi2[0]; t1 += C1dbopi2 * bo.dln_BOp_pi2[1]; t2 += C1dbopi2 * bo.dln_BOp_pi2[2];
      t0 += C2dbopi2 * dbx; t1 += C2dbopi2 * dby; t2 += C2dbopi2 * dbz;
      t0 += C3dbopi2 * ds0; t1 += C3dbopi2 * ds1; t2 += C3dbopi2 * ds2;
      f0 += t0; f1 += t1; f2 += t2;
      buf_c += (-C2dbo - C2dDelta - C3dbopi - C3dbopi2);
    }
    A.fcd[4 * (size_t)i] = f0;
    A.fcd[4 * (size_t)i + 1] = f1;
    A.fcd[4 * (size_t)i + 2] = f2;
  }
  A.bufc[i] = buf_c;
}

// The terms the reference adds to k from each bonded atom i (forces on k,
// :580-584): buf_c of i times i's dBOp toward k.
__global__ void __launch_bounds__(BT) k_dbond_nbr(int N, DbArgs A) {
  const int k = blockIdx.x * BT + threadIdx.x;
  if (k >= N) return;
  const int b0 = A.bidx[k], b1 = A.bend[k];
  if (b0 >= b1) return;
  double f0 = A.fcd[4 * (size_t)k], f1 = A.fcd[4 * (size_t)k + 1], f2 = A.fcd[4 * (size_t)k + 2];
  for (int pk = b0; pk < b1; ++pk) {
    const swr_bond_t b = A.bonds[pk];
    const double c = A.bufc[b.nbr];
    const double *d = A.bo[b.sym_index].dBOp;
    f0 += c * d[0];
    f1 += c * d[1];
    f2 += c * d[2];
  }
  A.fcd[4 * (size_t)k] = f0;
  A.fcd[4 * (size_t)k + 1] = f1;
  A.fcd[4 * (size_t)k + 2] = f2;
}

// ---------------------------------------------------------------- hydrogen bonds
// Hydrogen_Bonds_C (reaxc_hydrogen_bonds_sw64.c:26-283, host loop :145-253):
// a triple is a local H atom j (p_hbond 1), a bond j-i to an acceptor
// (p_hbond 2) with BO >= 0.01, and a hydrogen-bond partner k of j with
// orig_id(k) != orig_id(i) and r0_hb(i, j, k) > 0.  Its energy and CEhb1
// (into Cdbo of j's slot) and the angle/distance terms into fCdDelta of i,
// j and k.  k_hb computes each triple once, a wave per H atom with its lanes
// over the atom's hydrogen-bond entries: the energy, CEhb1 and j's own
// terms are summed by a fixed butterfly; the terms of i (summed over the
// entries, one set per bond) and of k (summed over the bonds, one set per
// entry) go into fixed-point accumulators (fix_add: 2^-40 resolution for
// every term, sums up to 2^43) by integer atomics, whose sums do not depend
// on the order they arrive in; k_hb_fold adds them to fCdDelta and clears them.
// The results are the same on every run.
constexpr double HB_THRESHOLD = 1e-2;   // reaxc_defs_sunway.h:91
constexpr double HB_FIX = 1099511627776.0;   // 2^40

struct HbArgs {
  unsigned int *flags;   // DevStatus::flags
  const DevParams *P;
  const swr_atom_pack_t *atoms;
  const int *bidx, *bend;
  const swr_bond_t *bonds;
  const double *BO;
  const int *Hindex, *hidx, *hend;
  const swr_hbond_t *hbonds;
  double *Cdbo, *fcd, *part;
  unsigned long long *acc;   // [2][N][3] fixed-point terms of i and k (fine, then coarse words)
  size_t coff;               // 3N: offset of the coarse words
};

// One triple (Calculate_Theta, Calculate_dCos_Theta :336-372, energy and
// coefficients :224-235); the cube of 1/(d_ji d_jk) by two products where
// the reference calls p_powd.
struct HbTriple {
  double e, ce1, ce2, ce3;
  double di[3], dj[3], dk[3];
};

__device__ __forceinline__ void hb_triple(const DevHb &hbp, double BOij, const double *dvec_ji, double d_ji,
                                          const double *dvec_jk, double r_jk, HbTriple &t) {
  const double dot = dvec_ji[0] * dvec_jk[0] + dvec_ji[1] * dvec_jk[1] + dvec_ji[2] * dvec_jk[2];
  double cos_theta = dot / (d_ji * r_jk);
  if (cos_theta > 1.) cos_theta = 1.0;
  if (cos_theta < -1.) cos_theta = -1.0;
  const double sqr_d_ji = d_ji * d_ji, sqr_d_jk = r_jk * r_jk;
  const double inv_dists = 1.0 / (d_ji * r_jk);
  const double inv_dists3 = inv_dists * inv_dists * inv_dists;
  const double Cdot_inv3 = dot * inv_dists3;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    t.di[c] = dvec_jk[c] * inv_dists - Cdot_inv3 * sqr_d_jk * dvec_ji[c];
    t.dj[c] = -(dvec_jk[c] + dvec_ji[c]) * inv_dists + Cdot_inv3 * (sqr_d_jk * dvec_ji[c] + sqr_d_ji * dvec_jk[c]);
    t.dk[c] = dvec_ji[c] * inv_dists - Cdot_inv3 * sqr_d_ji * dvec_jk[c];
  }
  // sin^4(theta/2) = ((1 - cos theta)/2)^2: the reference goes through acos
  // and sin; the identity gives the same value to rounding without either
  const double cos_xhz1 = (1.0 - cos_theta);
  double sin_xhz4 = 0.5 * cos_xhz1;
  sin_xhz4 *= sin_xhz4;
  const double exp_hb2 = exp(-hbp.p_hb2 * BOij);
  const double exp_hb3 = exp(-hbp.p_hb3 * (hbp.r0_hb / r_jk + r_jk / hbp.r0_hb - 2.0));
  t.e = hbp.p_hb1 * (1.0 - exp_hb2) * exp_hb3 * sin_xhz4;
  t.ce1 = hbp.p_hb1 * hbp.p_hb2 * exp_hb2 * exp_hb3 * sin_xhz4;
  t.ce2 = -hbp.p_hb1 / 2.0 * (1.0 - exp_hb2) * exp_hb3 * cos_xhz1;
  t.ce3 = -hbp.p_hb3 * (-hbp.r0_hb / (r_jk * r_jk) + 1.0 / hbp.r0_hb) * t.e;
}

// One fixed-point accumulator is a pair of 64-bit words: a fine word in
// units of 2^-40 (range 2^23) and a coarse word in units of 2^-20 (range
// 2^43), coff elements after it.  A term below 2^10 in magnitude goes to the
// fine word whole; a larger one is split exactly into a multiple of 2^-20
// (coarse word) and a remainder of at most 2^-21 (fine word).  Every term
// thus keeps the 2^-40 resolution (at most 2^-41 rounding per add), the fine
// word holds at most 2^10 per add (8192 adds before it could wrap), and the
// pair holds sums up to 2^43.  Both words are integers, so the sums do not
// depend on the order the adds arrive in.  A non-finite term, a term beyond
// ~3.8e12 (2^62 coarse units), or an add that wraps a word (detected from the
// value atomicAdd returns) raises SWR_F_FIXED_RANGE, which the next swr_sync
// reports as SWR_ERR_NUMERIC.
constexpr double FIX_COARSE = 1048576.0;   // 2^20: coarse units
constexpr double FIX_SPLIT = 1024.0;       // terms at least this large are split

__device__ __forceinline__ void fix_word_add(unsigned int *flags, unsigned long long *a, long long d) {
  const long long old = (long long)atomicAdd(a, (unsigned long long)d);
  const long long sum = (long long)((unsigned long long)old + (unsigned long long)d);
  if (((old ^ sum) & (d ^ sum)) < 0) atomicOr(flags, SWR_F_FIXED_RANGE);   // signed wrap
}

__device__ __forceinline__ void fix_add(unsigned int *flags, unsigned long long *a, size_t coff, double v) {
  if (!(fabs(v) < FIX_SPLIT)) {
    const double hi = rint(v * FIX_COARSE);
    if (!(fabs(hi) < 4.0e18)) {   // not finite, or beyond ~3.8e12
      atomicOr(flags, SWR_F_FIXED_RANGE);
      return;
    }
    fix_word_add(flags, a + coff, (long long)hi);
    v -= hi * (1.0 / FIX_COARSE);   // exact: |v - hi 2^-20| <= 2^-21
  }
  const long long f = llrint(v * HB_FIX);
  if (f != 0) fix_word_add(flags, a, f);
}

// K fix_adds whose fine-word results are checked for a wrap only after all
// K were issued, so a lane waits for the returned words once per group
// instead of once per add.  Terms of FIX_SPLIT and more (rare) take fix_add.
template <int K>
__device__ __forceinline__ void fix_add_group(unsigned int *flags, unsigned long long *const (&a)[K],
                                              const size_t (&coff)[K], const double (&v)[K]) {
  unsigned long long old[K];
  long long d[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    old[k] = 0ull;
    d[k] = 0;
    if (!(fabs(v[k]) < FIX_SPLIT)) {
      fix_add(flags, a[k], coff[k], v[k]);
    } else {
      d[k] = llrint(v[k] * HB_FIX);
      if (d[k] != 0) old[k] = atomicAdd(a[k], (unsigned long long)d[k]);
    }
  }
  bool wrap = false;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const long long o = (long long)old[k];
    const long long sum = (long long)(old[k] + (unsigned long long)d[k]);
    wrap |= ((o ^ sum) & (d[k] ^ sum)) < 0;
  }
  if (wrap) atomicOr(flags, SWR_F_FIXED_RANGE);
}

// The value of the accumulator pair at a (coarse word coff elements on).
__device__ __forceinline__ double fix_value(const unsigned long long *a, size_t coff) {
  return (double)(long long)a[0] * (1.0 / HB_FIX) + (double)(long long)a[coff] * (1.0 / FIX_COARSE);
}

__global__ void __launch_bounds__(BT) k_hb(int n, HbArgs A) {
  const int j = blockIdx.x * (BT / SWR_WAVE) + wave_in_block();
  if (j >= n) return;
  const int lane = lane_id();
  const DevParams *__restrict__ P = A.P;
  const swr_atom_pack_t aj = A.atoms[j];
  const int tj = aj.type;
  const int H = (tj >= 0 && P->p_hbond[tj] == 1) ? A.Hindex[j] : -1;
  if (H < 0) {
    if (lane == 0) A.part[j] = 0.0;
    return;
  }
  const int h0 = A.hidx[H], h1 = A.hend[H];
  double e = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int pi = A.bidx[j]; pi < A.bend[j]; ++pi) {
    const swr_bond_t bij = A.bonds[pi];
    const int i = bij.nbr;
    const int ti = A.atoms[i].type;
    const double BOij = A.BO[pi];
    if (ti < 0 || P->p_hbond[ti] != 2 || !(BOij >= HB_THRESHOLD)) continue;
    const int oi = A.atoms[i].orig_id;
    double ce1 = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int pk = h0 + lane; pk < h1; pk += SWR_WAVE) {
      const swr_hbond_t hb = A.hbonds[pk];
      const int k = hb.nbr;
      const swr_atom_pack_t ak = A.atoms[k];
      if (ak.type < 0 || ak.orig_id == oi) continue;
      const DevHb &hbp = P->hb[(ti * SWR_MAX_ELEM + tj) * SWR_MAX_ELEM + ak.type];
      if (hbp.r0_hb <= 0.0) continue;
      const swr_far_neighbor_t *fe = hb.ptr;
      const double r_jk = fe->d;
      const double dvec_jk[3] = {hb.scl * fe->dvec[0], hb.scl * fe->dvec[1], hb.scl * fe->dvec[2]};
      HbTriple t;
      hb_triple(hbp, BOij, bij.dvec, bij.d, dvec_jk, r_jk, t);
      e += t.e;
      ce1 += t.ce1;
      f0 += t.ce2 * t.dj[0] - t.ce3 / r_jk * dvec_jk[0];
      f1 += t.ce2 * t.dj[1] - t.ce3 / r_jk * dvec_jk[1];
      f2 += t.ce2 * t.dj[2] - t.ce3 / r_jk * dvec_jk[2];
      g0 += t.ce2 * t.di[0];
      g1 += t.ce2 * t.di[1];
      g2 += t.ce2 * t.di[2];
      unsigned long long *ak3 = A.acc + 3 * (size_t)k;
      {
        unsigned long long *const ga[3] = {ak3, ak3 + 1, ak3 + 2};
        const size_t gc[3] = {A.coff, A.coff, A.coff};
        const double gv[3] = {t.ce2 * t.dk[0] + t.ce3 / r_jk * dvec_jk[0], t.ce2 * t.dk[1] + t.ce3 / r_jk * dvec_jk[1],
                              t.ce2 * t.dk[2] + t.ce3 / r_jk * dvec_jk[2]};
        fix_add_group<3>(A.flags, ga, gc, gv);
      }
    }
    ce1 = wave_sum(ce1);
    g0 = wave_sum(g0);
    g1 = wave_sum(g1);
    g2 = wave_sum(g2);
    if (lane == 0) {
      A.Cdbo[pi] += ce1;
      unsigned long long *ai3 = A.acc + 3 * (size_t)i;
      unsigned long long *const ga[3] = {ai3, ai3 + 1, ai3 + 2};
      const size_t gc[3] = {A.coff, A.coff, A.coff};
      const double gv[3] = {g0, g1, g2};
      fix_add_group<3>(A.flags, ga, gc, gv);
    }
  }
  e = wave_sum(e);
  f0 = wave_sum(f0);
  f1 = wave_sum(f1);
  f2 = wave_sum(f2);
  if (lane == 0) {
    A.part[j] = e;
    A.fcd[4 * (size_t)j] += f0;
    A.fcd[4 * (size_t)j + 1] += f1;
    A.fcd[4 * (size_t)j + 2] += f2;
  }
}

// fCdDelta += the fixed-point terms; the accumulators are left at zero.
__global__ void __launch_bounds__(BT) k_hb_fold(int N, unsigned long long *__restrict__ acc,
                                                double *__restrict__ fcd) {
  const int a = blockIdx.x * BT + threadIdx.x;
  if (a >= N) return;
  const size_t coff = 3 * (size_t)N;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    unsigned long long *w = acc + 3 * (size_t)a + c;
    if (w[0] != 0ull || w[coff] != 0ull) {
      fcd[4 * (size_t)a + c] += fix_value(w, coff);
      w[0] = 0ull;
      w[coff] = 0ull;
    }
  }
}

// ---------------------------------------------------------------- valence angles
// The valence part of Merge_Torsion_Valence_Angles (host form
// reaxc_torsion_angles_sw64.c:812-1117): for every local atom j and pair of
// its bonds j-k (pk) and j-h (ph > pk) with BO above thb_cut and their
// product above thb_cutsq, and every parameter set of (type k, type j,
// type h) with |p_val1| > 0.001: the angle, penalty and coalition energies;
// their coefficients into Cdbo of both bonds and of all of j's bonds
// (CEval6 BO^7), Cdbopi/Cdbopi2 of j's bonds (CEval5), CdDelta of j, k and
// h, and the angle derivative into fCdDelta[.][0..2] of k, j and h.  One
// thread per j writes j's own bond slots and row; the terms of k and h go
// into 64-bit fixed-point accumulators by integer atomics (order-independent
// sums), folded into fCdDelta by k_va_fold.  The centres with angles are
// listed by angle count (k_va_count, k_wl_scan, k_va_scatter: a centre with
// ten bonds has 45 pairs, one with two bonds one) and the threads take them
// largest first.
constexpr double DEG2RAD_PI = 3.14159265;   // constPI, reaxc_defs_sunway.h:61 (DEG2RAD :55)

// ---------------------------------------------------------------- work lists
// The valence and torsion passes give each lane one item (a centre atom, a
// central bond) and most items have little or no work while a few have
// dozens of angles: items are counted first, bucketed by count and handed
// to lanes largest first, so the lanes of a wave do about the same work.
// Counters: the list length, a histogram of the counts and each bucket's
// next free position.  Which lane takes which item changes no result: every
// item writes only its own entries and adds into the others through
// fixed-point integer atomics.
constexpr int WL_NB = 128;    // count buckets; counts of WL_NB - 1 and more share the last
enum { WL_NTASK = 0, WL_HIST = 1, WL_CUR = WL_HIST + WL_NB, WL_CTL_INTS = WL_CUR + WL_NB };
__device__ __forceinline__ int wl_bucket(int c) { return c < WL_NB - 1 ? c : WL_NB - 1; }

// Bucket starts, largest count first; the histogram is left at zero.
__global__ void __launch_bounds__(WL_NB) k_wl_scan(int *__restrict__ ctl) {
  __shared__ int h[WL_NB];
  const int b = threadIdx.x;
  h[b] = ctl[WL_HIST + b];
  ctl[WL_HIST + b] = 0;
  __syncthreads();
  int start = 0;
  for (int q = b + 1; q < WL_NB; ++q) start += h[q];
  ctl[WL_CUR + b] = start;
  if (b == 0) ctl[WL_NTASK] = start + h[0];
}

// Workgroup histogram of counts c (> 0 adds) into the global one.
__device__ __forceinline__ void wl_hist_flush(int *hist, int *__restrict__ ctl) {
  __syncthreads();
  for (int b = threadIdx.x; b < WL_NB; b += BT)
    if (hist[b]) atomicAdd(&ctl[WL_HIST + b], hist[b]);
}

struct VaArgs {
  unsigned int *flags;   // DevStatus::flags
  const DevParams *P;
  const swr_atom_pack_t *atoms;
  const int *bidx, *bend;
  const swr_bond_t *bonds;
  const double *BO, *BOpi, *bo_dboc;
  const double *Delta, *Delta_val, *vlpex, *nlp, *dDelta_lp;
  double *Cdbo, *Cdbopi, *Cdbopi2, *fcd, *part;
  size_t stride;
  unsigned long long *acc;   // [2][N][4] (fine, then coarse words)
  size_t coff;               // 4N
  int *cnt;                  // [n] angle count, then rank | bucket
  int *order;                // [n] centres with angles, largest count first
  int *ctl;                  // WL_CTL_INTS counters
};

// Angle evaluations of local centre j: the pairs of its bonds that pass the
// body's filters, times their parameter sets.
__device__ __forceinline__ int va_count(const VaArgs &A, int j) {
  const DevParams *__restrict__ P = A.P;
  const int tj = A.atoms[j].type;
  const int b0 = A.bidx[j], b1 = A.bend[j];
  if (tj < 0 || b1 <= b0) return 0;
  const double thb_cut = P->thb_cut, thb_cutsq = P->thb_cutsq;
  int c = 0;
  for (int pk = b0; pk < b1; ++pk) {
    const double BO_jk = A.BO[pk];
    if (!(BO_jk > thb_cut)) continue;
    const int k = A.bonds[pk].nbr;
    if (A.bend[k] <= A.bidx[k]) continue;
    const int tk = A.atoms[k].type;
    for (int ph = pk + 1; ph < b1; ++ph) {
      const double BO_hj = A.BO[ph];
      if (!(BO_hj > thb_cut) || !(BO_jk * BO_hj > thb_cutsq)) continue;
      const int th = A.atoms[A.bonds[ph].nbr].type;
      if (tk < 0 || th < 0) continue;
      c += P->thb_n[(tk * SWR_MAX_ELEM + tj) * SWR_MAX_ELEM + th];
    }
  }
  return c;
}

// One thread per local atom: its count and the histogram; a centre without
// angles gets its zero energy rows here (the valence kernel skips it).
__global__ void __launch_bounds__(BT) k_va_count(int n, VaArgs A) {
  __shared__ int hist[WL_NB];
  for (int b = threadIdx.x; b < WL_NB; b += BT) hist[b] = 0;
  __syncthreads();
  const int j = blockIdx.x * BT + threadIdx.x;
  if (j < n) {
    const int c = va_count(A, j);
    A.cnt[j] = c;
    if (c > 0) {
      atomicAdd(&hist[wl_bucket(c)], 1);
    } else {
      A.part[j] = 0.0;
      A.part[A.stride + j] = 0.0;
      A.part[2 * A.stride + j] = 0.0;
    }
  }
  wl_hist_flush(hist, A.ctl);
}

__global__ void __launch_bounds__(BT) k_va_scatter(int n, VaArgs A) {
  __shared__ int cnt[WL_NB], base[WL_NB];
  for (int b = threadIdx.x; b < WL_NB; b += BT) cnt[b] = 0;
  __syncthreads();
  const int j = blockIdx.x * BT + threadIdx.x;
  int b = -1, r = 0;
  if (j < n && A.cnt[j] > 0) {
    b = wl_bucket(A.cnt[j]);
    r = atomicAdd(&cnt[b], 1);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < WL_NB; q += BT) base[q] = cnt[q] ? atomicAdd(&A.ctl[WL_CUR + q], cnt[q]) : 0;
  __syncthreads();
  if (b >= 0) A.order[base[b] + r] = j;
}

// One thread per listed centre j (largest angle count first).
__global__ void __launch_bounds__(BT) k_valence(VaArgs A) {
  const int t = blockIdx.x * BT + threadIdx.x;
  if (t >= A.ctl[WL_NTASK]) return;
  const int j = A.order[t];
  const DevParams *__restrict__ P = A.P;
  const int tj = A.atoms[j].type;
  const int b0 = A.bidx[j], b1 = A.bend[j];
  double e_ang = 0.0, e_pen = 0.0, e_coa = 0.0;
  if (tj >= 0 && b1 > b0) {
    const double thb_cut = P->thb_cut, thb_cutsq = P->thb_cutsq;
    const double Dboc_j = A.bo_dboc[2 * (size_t)j + 1];
    const double p_val3 = P->p_val3[tj], p_val5 = P->p_val5[tj];
    const double p_val6 = P->p_val6, p_val8 = P->p_val8, p_val9 = P->p_val9, p_val10 = P->p_val10;
    double SBOp = 0.0, prod_SBO = 1.0;
    for (int t = b0; t < b1; ++t) {
      SBOp += A.BOpi[2 * (size_t)t] + A.BOpi[2 * (size_t)t + 1];
      double b2 = A.BO[t] * A.BO[t];
      b2 *= b2;
      b2 *= b2;
      prod_SBO *= exp(-b2);
    }
    double vlpadj, dSBO2;
    if (A.vlpex[j] >= 0) {
      vlpadj = 0.0;
      dSBO2 = prod_SBO - 1.0;
    } else {
      vlpadj = A.nlp[j];
      dSBO2 = (prod_SBO - 1.0) * (1.0 - p_val8 * A.dDelta_lp[j]);
    }
    const double SBO = SBOp + (1.0 - prod_SBO) * (-Dboc_j - p_val8 * vlpadj);
    const double dSBO1 = -8.0 * prod_SBO * (Dboc_j + p_val8 * vlpadj);
    double SBO2, CSBO2;
    if (SBO <= 0.0) {
      SBO2 = 0.0; CSBO2 = 0.0;
    } else if (SBO <= 1.0) {
      SBO2 = pow(SBO, p_val9); CSBO2 = p_val9 * pow(SBO, p_val9 - 1.0);
    } else if (SBO < 2.0) {
      SBO2 = 2.0 - pow(2.0 - SBO, p_val9); CSBO2 = p_val9 * pow(2.0 - SBO, p_val9 - 1.0);
    } else {
      SBO2 = 2.0; CSBO2 = 0.0;
    }
    const double expval6 = exp(p_val6 * Dboc_j);
    const double exp10 = exp(-p_val10 * (2.0 - SBO2));   // theta_0 and Ctheta_0 of every set
    const double Delta_j = A.Delta[j];
    const double exp_pen3 = exp(-P->p_pen3 * Delta_j), exp_pen4 = exp(P->p_pen4 * Delta_j);
    const double trm_pen34 = 1.0 + exp_pen3 + exp_pen4;
    const double f9_Dj = (2.0 + exp_pen3) / trm_pen34;
    const double Cf9j = (-P->p_pen3 * exp_pen3 * trm_pen34 - (2.0 + exp_pen3) * (-P->p_pen3 * exp_pen3 + P->p_pen4 * exp_pen4)) /
                        (trm_pen34 * trm_pen34);
    const double exp_coa2 = exp(P->p_coa2 * A.Delta_val[j]);
    double ce5 = 0.0, ce6 = 0.0, cdj = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int pk = b0; pk < b1; ++pk) {
      const double BO_jk = A.BO[pk];
      if (!(BO_jk > thb_cut)) continue;
      const swr_bond_t bjk = A.bonds[pk];
      const int k = bjk.nbr;
      if (A.bend[k] <= A.bidx[k]) continue;
      const int tk = A.atoms[k].type;
      const double BOA_jk = BO_jk - thb_cut;
      for (int ph = pk + 1; ph < b1; ++ph) {
        const double BO_hj = A.BO[ph];
        if (!(BO_hj > thb_cut) || !(BO_jk * BO_hj > thb_cutsq)) continue;
        const int th = A.atoms[A.bonds[ph].nbr].type;
        if (tk < 0 || th < 0) continue;
        const int tri = (tk * SWR_MAX_ELEM + tj) * SWR_MAX_ELEM + th;
        const int nset = P->thb_n[tri];
        if (nset == 0) continue;
        const swr_bond_t bhj = A.bonds[ph];
        const int h = bhj.nbr;
        const double BOA_hj = BO_hj - thb_cut;
        // Calculate_Theta / Calculate_dCos_Theta (:1510-1547) with the j-k
        // bond first: di belongs to k, dk to h
        const double dot = bjk.dvec[0] * bhj.dvec[0] + bjk.dvec[1] * bhj.dvec[1] + bjk.dvec[2] * bhj.dvec[2];
        double cos_t = dot / (bjk.d * bhj.d);
        if (cos_t > 1.0) cos_t = 1.0;
        if (cos_t < -1.0) cos_t = -1.0;
        const double theta = acos(cos_t);
        double sin_t = sin(theta);
        if (sin_t < 1.0e-5) sin_t = 1.0e-5;
        const double sq_jk = bjk.d * bjk.d, sq_hj = bhj.d * bhj.d;
        const double inv_d = 1.0 / (bjk.d * bhj.d);
        const double cinv3 = dot * (inv_d * inv_d * inv_d);
        const double bo_k0 = A.bo_dboc[2 * (size_t)k], bo_h0 = A.bo_dboc[2 * (size_t)h];
        // the penalty and coalition factors of the pair, the same for every
        // set; only evaluated when a set has p_pen1 or p_coa1 (a term with a
        // zero prefactor is zero either way)
        const bool need_pc = P->thb_pc[tri] != 0;
        const double pen2 = need_pc ? exp(-P->p_pen2 * (BOA_jk - 2.0) * (BOA_jk - 2.0)) *
                                          exp(-P->p_pen2 * (BOA_hj - 2.0) * (BOA_hj - 2.0))
                                    : 0.0;
        const double coa = need_pc ? 1.0 / (1. + exp_coa2) * exp(-P->p_coa3 * (bo_k0 - BOA_jk) * (bo_k0 - BOA_jk)) *
                                         exp(-P->p_coa3 * (bo_h0 - BOA_hj) * (bo_h0 - BOA_hj)) *
                                         exp(-P->p_coa4 * (BOA_jk - 1.5) * (BOA_jk - 1.5)) *
                                         exp(-P->p_coa4 * (BOA_hj - 1.5) * (BOA_hj - 1.5))
                                   : 0.0;
        double ck[4] = {0.0, 0.0, 0.0, 0.0}, chh[4] = {0.0, 0.0, 0.0, 0.0};
        double cdbo_jk = 0.0, cdbo_hj = 0.0;
        for (int c = 0; c < nset; ++c) {
          const DevThb &q = P->thb[tri][c];
          const double p_val1 = q.p_val1, p_val2 = q.p_val2, p_val4 = q.p_val4, p_val7 = q.p_val7;
          const double pjk = pow(BOA_jk, p_val4 - 1.0), phj = pow(BOA_hj, p_val4 - 1.0);
          const double exp3jk = exp(-p_val3 * (pjk * BOA_jk));
          const double f7_jk = 1.0 - exp3jk;
          const double Cf7jk = p_val3 * p_val4 * pjk * exp3jk;
          const double exp3hj = exp(-p_val3 * (phj * BOA_hj));
          const double f7_hj = 1.0 - exp3hj;
          const double Cf7hj = p_val3 * p_val4 * phj * exp3hj;
          const double expval7 = exp(-p_val7 * Dboc_j);
          const double trm8 = 1.0 + expval6 + expval7;
          const double f8_Dj = p_val5 - ((p_val5 - 1.0) * (2.0 + expval6) / trm8);
          const double Cf8j = ((1.0 - p_val5) / (trm8 * trm8)) *
                              (p_val6 * expval6 * trm8 - (2.0 + expval6) * (p_val6 * expval6 - p_val7 * expval7));
          const double theta_0 = (180.0 - q.theta_00 * (1.0 - exp10)) * DEG2RAD_PI / 180.0;
          const double expval2theta = exp(-p_val2 * (theta_0 - theta) * (theta_0 - theta));
          const double expval12theta = p_val1 >= 0 ? p_val1 * (1.0 - expval2theta) : p_val1 * -expval2theta;
          const double CEval1 = Cf7jk * f7_hj * f8_Dj * expval12theta;
          const double CEval2 = Cf7hj * f7_jk * f8_Dj * expval12theta;
          const double CEval3 = Cf8j * f7_hj * f7_jk * expval12theta;
          const double CEval4 = -2.0 * p_val1 * p_val2 * f7_jk * f7_hj * f8_Dj * expval2theta * (theta_0 - theta);
          const double Ctheta_0 = p_val10 * (q.theta_00 * DEG2RAD_PI / 180.0) * exp10;
          const double CEval5 = -CEval4 * Ctheta_0 * CSBO2;
          const double CEval8 = -CEval4 / sin_t;
          e_ang += f7_jk * f7_hj * f8_Dj * expval12theta;
          // penalty
          const double ep = q.p_pen1 * f9_Dj * pen2;
          e_pen += ep;
          const double CEpen1 = ep * Cf9j / f9_Dj;
          const double tp = -2.0 * P->p_pen2 * ep;
          const double CEpen2 = tp * (BOA_jk - 2.0), CEpen3 = tp * (BOA_hj - 2.0);
          // coalition
          const double ec = q.p_coa1 * coa;
          e_coa += ec;
          const double CEcoa1 = -2 * P->p_coa4 * (BOA_jk - 1.5) * ec;
          const double CEcoa2 = -2 * P->p_coa4 * (BOA_hj - 1.5) * ec;
          const double CEcoa3 = -P->p_coa2 * exp_coa2 * ec / (1 + exp_coa2);
          const double CEcoa4 = -2 * P->p_coa3 * (bo_k0 - BOA_jk) * ec;
          const double CEcoa5 = -2 * P->p_coa3 * (bo_h0 - BOA_hj) * ec;
          cdbo_jk += CEval1 + CEpen2 + (CEcoa1 - CEcoa4);
          cdbo_hj += CEval2 + CEpen3 + (CEcoa2 - CEcoa5);
          cdj += (CEval3 + CEval5 * dSBO2) + CEpen1 + CEcoa3;
          ck[3] += CEcoa4;
          chh[3] += CEcoa5;
          ce5 += CEval5;
          ce6 += CEval5 * dSBO1;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double di = bhj.dvec[d] * inv_d - cinv3 * sq_hj * bjk.dvec[d];
            const double dj = -(bhj.dvec[d] + bjk.dvec[d]) * inv_d + cinv3 * (sq_hj * bjk.dvec[d] + sq_jk * bhj.dvec[d]);
            const double dk = bjk.dvec[d] * inv_d - cinv3 * sq_jk * bhj.dvec[d];
            ck[d] += CEval8 * di;
            chh[d] += CEval8 * dk;
            if (d == 0) g0 += CEval8 * dj;
            if (d == 1) g1 += CEval8 * dj;
            if (d == 2) g2 += CEval8 * dj;
          }
        }
        A.Cdbo[pk] += cdbo_jk;
        A.Cdbo[ph] += cdbo_hj;
        unsigned long long *ak = A.acc + 4 * (size_t)k, *ah = A.acc + 4 * (size_t)h;
        {
          unsigned long long *const ga[8] = {ak, ak + 1, ak + 2, ak + 3, ah, ah + 1, ah + 2, ah + 3};
          const size_t gc[8] = {A.coff, A.coff, A.coff, A.coff, A.coff, A.coff, A.coff, A.coff};
          const double gv[8] = {ck[0], ck[1], ck[2], ck[3], chh[0], chh[1], chh[2], chh[3]};
          fix_add_group<8>(A.flags, ga, gc, gv);
        }
      }
    }
    if (ce5 != 0.0 || ce6 != 0.0)
      for (int t = b0; t < b1; ++t) {
        const double bo = A.BO[t];
        const double bo3 = bo * bo * bo;
        A.Cdbo[t] += ce6 * (bo3 * bo3 * bo);
        A.Cdbopi[t] += ce5;
        A.Cdbopi2[t] += ce5;
      }
    A.fcd[4 * (size_t)j] += g0;
    A.fcd[4 * (size_t)j + 1] += g1;
    A.fcd[4 * (size_t)j + 2] += g2;
    A.fcd[4 * (size_t)j + 3] += cdj;
  }
  A.part[j] = e_ang;
  A.part[A.stride + j] = e_pen;
  A.part[2 * A.stride + j] = e_coa;
}

// fCdDelta (all four columns) += the fixed-point terms; the accumulators are
// left at zero.
__global__ void __launch_bounds__(BT) k_va_fold(int N, unsigned long long *__restrict__ acc,
                                                double *__restrict__ fcd) {
  const int a = blockIdx.x * BT + threadIdx.x;
  if (a >= N) return;
  const size_t coff = 4 * (size_t)N;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    unsigned long long *w = acc + 4 * (size_t)a + c;
    if (w[0] != 0ull || w[coff] != 0ull) {
      fcd[4 * (size_t)a + c] += fix_value(w, coff);
      w[0] = 0ull;
      w[coff] = 0ull;
    }
  }
}

// ---------------------------------------------------------------- torsions
// The torsion part of Merge_Torsion_Valence_Angles (host form
// reaxc_torsion_angles_sw64.c:812-1408; Calculate_Omega :21-104).  The work
// is one task per central bond j-k of a local atom j (taken once per
// physical bond: orig_id order, then position for images of one atom): over
// the other bonds j-i of j and the bonds k-l of k, the torsion and
// conjugation energies.  Most bonds make no torsion and a few make dozens,
// so the tasks are listed with their exact quadruple counts and handed to
// lanes in decreasing count (a wave's lanes then do about the same work):
//
//   k_to_count    four lanes per local atom: per bond slot the number of
//                 quadruples the torsion body evaluates (the same filters),
//                 a histogram of the counts;
//   k_wl_scan     start of each count bucket in the list (largest first);
//   k_to_scatter  one thread per local atom: its central bonds with work
//                 into the list (the order inside a bucket is arbitrary and
//                 changes no result);
//   k_torsion     one lane per listed central bond: the slot's Cdbopi and
//                 Cdbo written directly, j's row terms and energies into
//                 the slot's row entry, Cdbo of j-i and k-l and the terms of
//                 i, k, l into 64-bit fixed-point accumulators by integer
//                 atomics (folded by k_va_fold / k_to_slot_fold);
//   k_to_row      one thread per local atom: its slots' row entries summed
//                 in slot order into its row and energies.
struct ToArgs {
  unsigned int *flags;   // DevStatus::flags
  const DevParams *P;
  const swr_atom_pack_t *atoms;
  const int *bidx, *bend;
  const swr_bond_t *bonds;
  const double *BO, *BOpi, *bo_dboc;
  double *Cdbo, *Cdbopi, *fcd, *part;
  size_t stride;
  unsigned long long *acc;    // [2][N][4] (fine, then coarse words)
  unsigned long long *sacc;   // [2][bond slots]
  size_t coff, scoff;         // 4N, bond slots
  const int *bnbr;            // bonds[].nbr, compact
  int *tc;                    // [bond slots] quadruple count, then rank | bucket
  int2 *task;                 // [bond slots] (j, slot) of the listed central bonds
  int *ctl;                   // WL_CTL_INTS counters
  double *row;                // [bond slots][6] j's row terms and energies per central bond
};


// Calculate_Theta + Calculate_dCos_Theta of bonds a (first) and b from one
// centre: di belongs to a's far atom, dj to the centre, dk to b's.
__device__ __forceinline__ double theta_dcos(const swr_bond_t &a, const swr_bond_t &b, double *di, double *dj,
                                             double *dk) {
  const double dot = a.dvec[0] * b.dvec[0] + a.dvec[1] * b.dvec[1] + a.dvec[2] * b.dvec[2];
  double c = dot / (a.d * b.d);
  if (c > 1.0) c = 1.0;
  if (c < -1.0) c = -1.0;
  const double sa = a.d * a.d, sb = b.d * b.d;
  const double inv = 1.0 / (a.d * b.d);
  const double ci3 = dot * (inv * inv * inv);
#pragma unroll
  for (int q = 0; q < 3; q++) {
    di[q] = b.dvec[q] * inv - ci3 * sb * a.dvec[q];
    dj[q] = -(b.dvec[q] + a.dvec[q]) * inv + ci3 * (sb * a.dvec[q] + sa * b.dvec[q]);
    dk[q] = a.dvec[q] * inv - ci3 * sa * b.dvec[q];
  }
  return acos(c);
}

__device__ __forceinline__ double tan_inv(double s, double c) {
  constexpr double MIN_SINE = 1e-10;   // reaxc_torsion_angles_sw64.c:12
  if (s >= 0 && s <= MIN_SINE) return c / MIN_SINE;
  if (s <= 0 && s >= -MIN_SINE) return c / -MIN_SINE;
  return c / s;
}

// Quadruples the torsion body evaluates for the bond in slot pk of local
// atom j (type tj >= 0, slots [b0, b1)): the body's filters in its order;
// 0 when the bond is no central bond (its BO, the orig_id order, a k without
// bonds or type) or makes no torsion.
__device__ __forceinline__ int to_count(const ToArgs &A, const swr_atom_pack_t &aj, int tj, int b0, int b1,
                                        int pk, double thb_cut) {
  const double BO_jk = A.BO[pk];
  if (!(BO_jk - thb_cut > 0.0)) return 0;
  const int k = A.bnbr[pk];
  const int k0 = A.bidx[k], k1 = A.bend[k];
  if (k1 <= k0) return 0;
  const swr_atom_pack_t ak = A.atoms[k];
  if (!owns_bond(aj, ak)) return 0;
  const int tk = ak.type;
  if (tk < 0) return 0;
  const int pj = A.bonds[pk].sym_index;
  int c = 0;
  for (int ph = b0; ph < b1; ++ph) {
    if (ph == pk) continue;
    const double BO_ij = A.BO[ph];
    if (!(BO_ij > thb_cut)) continue;
    const int i = A.bnbr[ph];
    const int ti = A.atoms[i].type;
    if (ti < 0) continue;
    for (int pw = k0; pw < k1; ++pw) {
      if (pw == pj) continue;
      const double BO_kl = A.BO[pw];
      if (!(BO_kl > thb_cut) || !(BO_ij * BO_jk * BO_kl > thb_cut)) continue;
      const int l = A.bnbr[pw];
      if (i == l) continue;
      const int tl = A.atoms[l].type;
      if (tl < 0) continue;
      if (!A.P->tor[((ti * SWR_MAX_ELEM + tj) * SWR_MAX_ELEM + tk) * SWR_MAX_ELEM + tl].cnt) continue;
      ++c;
    }
  }
  return c;
}

constexpr int TO_CNT_LANES = 4;   // lanes per atom in k_to_count, each over every 4th slot

__global__ void __launch_bounds__(BT) k_to_count(int n, ToArgs A) {
  __shared__ int hist[WL_NB];
  for (int b = threadIdx.x; b < WL_NB; b += BT) hist[b] = 0;
  __syncthreads();
  const int tid = blockIdx.x * BT + threadIdx.x;
  const int j = tid / TO_CNT_LANES;
  if (j < n) {
    const swr_atom_pack_t aj = A.atoms[j];
    const int b0 = A.bidx[j], b1 = A.bend[j];
    const double thb_cut = A.P->thb_cut;
    for (int pk = b0 + tid % TO_CNT_LANES; pk < b1; pk += TO_CNT_LANES) {
      const int c = aj.type >= 0 ? to_count(A, aj, aj.type, b0, b1, pk, thb_cut) : 0;
      A.tc[pk] = c;
      if (c > 0) {
        atomicAdd(&hist[wl_bucket(c)], 1);
      } else {
        double *rw = A.row + 6 * (size_t)pk;
#pragma unroll
        for (int q = 0; q < 6; q++) rw[q] = 0.0;
      }
    }
  }
  wl_hist_flush(hist, A.ctl);
}

__global__ void __launch_bounds__(BT) k_to_scatter(int n, ToArgs A) {
  __shared__ int cnt[WL_NB], base[WL_NB];
  for (int b = threadIdx.x; b < WL_NB; b += BT) cnt[b] = 0;
  __syncthreads();
  const int j = blockIdx.x * BT + threadIdx.x;
  const int b0 = j < n ? A.bidx[j] : 0, b1 = j < n ? A.bend[j] : 0;
  for (int pk = b0; pk < b1; ++pk) {   // rank in the workgroup's share of the bucket
    const int c = A.tc[pk];
    if (c > 0) {
      const int b = wl_bucket(c);
      A.tc[pk] = (atomicAdd(&cnt[b], 1) << 8) | b;
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < WL_NB; b += BT) base[b] = cnt[b] ? atomicAdd(&A.ctl[WL_CUR + b], cnt[b]) : 0;
  __syncthreads();
  for (int pk = b0; pk < b1; ++pk) {
    const int v = A.tc[pk];
    if (v > 0) A.task[base[v & 255] + (v >> 8)] = make_int2(j, pk);
  }
}

__global__ void __launch_bounds__(BT) k_torsion(ToArgs A) {
  const int task = blockIdx.x * BT + threadIdx.x;
  if (task >= A.ctl[WL_NTASK]) return;
  const int2 jp = A.task[task];
  const int j = jp.x, pk = jp.y;
  const DevParams *__restrict__ P = A.P;
  double e_tor = 0.0, e_con = 0.0;
  double fj[4] = {0.0, 0.0, 0.0, 0.0};
  const swr_atom_pack_t aj = A.atoms[j];
  const int tj = aj.type;
  const int b0 = A.bidx[j], b1 = A.bend[j];
  {
    constexpr double MIN_SINE = 1e-10;
    const double thb_cut = P->thb_cut;
    const double p_tor2 = P->p_tor2, p_tor3 = P->p_tor3, p_tor4 = P->p_tor4, p_cot2 = P->p_cot2;
    const double Delta_j = A.bo_dboc[2 * (size_t)j + 1];
    do {
      const double BO_jk = A.BO[pk];
      if (!(BO_jk - thb_cut > 0.0)) break;
      const swr_bond_t bjk = A.bonds[pk];
      const int k = bjk.nbr;
      const int k0 = A.bidx[k], k1 = A.bend[k];
      if (k1 <= k0) break;
      const swr_atom_pack_t ak = A.atoms[k];
      if (aj.orig_id > ak.orig_id) break;
      if (aj.orig_id == ak.orig_id) {
        if (ak.x[2] < aj.x[2]) break;
        if (ak.x[2] == aj.x[2] && ak.x[1] < aj.x[1]) break;
        if (ak.x[2] == aj.x[2] && ak.x[1] == aj.x[1] && ak.x[0] < aj.x[0]) break;
      }
      const int tk = ak.type;
      if (tk < 0) break;
      const int pj = bjk.sym_index;
      const swr_bond_t bkj = A.bonds[pj];
      const double BOA_jk = BO_jk - thb_cut;
      const double Delta_k = A.bo_dboc[2 * (size_t)k + 1];
      const double r_jk = bjk.d;
      const double exp_tor2_jk = exp(-p_tor2 * BOA_jk);
      const double exp_cot2_jk = exp(-p_cot2 * (BOA_jk - 1.5) * (BOA_jk - 1.5));
      const double exp_tor3_DjDk = exp(-p_tor3 * (Delta_j + Delta_k));
      const double exp_tor4_DjDk = exp(p_tor4 * (Delta_j + Delta_k));
      const double exp_tor34_inv = 1.0 / (1.0 + exp_tor3_DjDk + exp_tor4_DjDk);
      const double f11_DjDk = (2.0 + exp_tor3_DjDk) * exp_tor34_inv;
      const double dfn11 = (-p_tor3 * exp_tor3_DjDk +
                            (p_tor3 * exp_tor3_DjDk - p_tor4 * exp_tor4_DjDk) * (2.0 + exp_tor3_DjDk) * exp_tor34_inv) *
                           exp_tor34_inv;
      const double BOpi_jk = A.BOpi[2 * (size_t)pk];
      double cdbopi_jk = 0.0, cdbo_jk = 0.0, cd_k = 0.0;
      double fk[3] = {0.0, 0.0, 0.0};
      for (int ph = b0; ph < b1; ++ph) {
        if (ph == pk) continue;
        const double BO_ij = A.BO[ph];
        if (!(BO_ij > thb_cut)) continue;
        const swr_bond_t bij = A.bonds[ph];
        const int i = bij.nbr;
        const int ti = A.atoms[i].type;
        if (ti < 0) continue;
        // the i-j-k angle terms, evaluated at the first k-l bond that makes
        // a torsion with this j-i bond (most j-i bonds make none)
        double ijk_di[3], ijk_dj[3], ijk_dk[3];   // of k, j, i
        double theta_ijk = 0.0, sin_ijk = 0.0, cos_ijk = 0.0, tan_ijk_i = 0.0, exp_tor2_ij = 0.0, exp_cot2_ij = 0.0;
        bool have_ijk = false;
        const double r_ij = bij.d;
        const double BOA_ij = BO_ij - thb_cut;
        const double xi[3] = {A.atoms[i].x[0], A.atoms[i].x[1], A.atoms[i].x[2]};
        double cdbo_ij = 0.0;
        double fi[3] = {0.0, 0.0, 0.0};
        for (int pw = k0; pw < k1; ++pw) {
          if (pw == pj) continue;
          const double BO_kl = A.BO[pw];
          if (!(BO_kl > thb_cut) || !(BO_ij * BO_jk * BO_kl > thb_cut)) continue;
          const swr_bond_t bkl = A.bonds[pw];
          const int l = bkl.nbr;
          if (i == l) continue;
          const swr_atom_pack_t al = A.atoms[l];
          if (al.type < 0) continue;
          const DevTor &fb = P->tor[((ti * SWR_MAX_ELEM + tj) * SWR_MAX_ELEM + tk) * SWR_MAX_ELEM + al.type];
          if (!fb.cnt) continue;
          if (!have_ijk) {
            theta_ijk = theta_dcos(bjk, bij, ijk_di, ijk_dj, ijk_dk);
            sin_ijk = sin(theta_ijk);
            cos_ijk = cos(theta_ijk);
            tan_ijk_i = tan_inv(sin_ijk, cos_ijk);
            exp_tor2_ij = exp(-p_tor2 * BOA_ij);
            exp_cot2_ij = exp(-p_cot2 * (BOA_ij - 1.5) * (BOA_ij - 1.5));
            have_ijk = true;
          }
          double jkl_di[3], jkl_dj[3], jkl_dk[3];   // of j, k, l
          const double theta_jkl = theta_dcos(bkj, bkl, jkl_di, jkl_dj, jkl_dk);
          const double r_kl = bkl.d;
          const double BOA_kl = BO_kl - thb_cut;
          double sin_jkl = sin(theta_jkl);
          const double cos_jkl = cos(theta_jkl);
          const double tan_jkl_i = tan_inv(sin_jkl, cos_jkl);
          const double dvec_li[3] = {xi[0] - al.x[0], xi[1] - al.x[1], xi[2] - al.x[2]};
          const double r_li = sqrt(dvec_li[0] * dvec_li[0] + dvec_li[1] * dvec_li[1] + dvec_li[2] * dvec_li[2]);
          // Calculate_Omega
          const double *dij = bij.dvec, *djk = bjk.dvec, *dkl = bkl.dvec;
          const double d_ij_jk = dij[0] * djk[0] + dij[1] * djk[1] + dij[2] * djk[2];
          const double d_jk_kl = djk[0] * dkl[0] + djk[1] * dkl[1] + djk[2] * dkl[2];
          const double d_ij_kl = dij[0] * dkl[0] + dij[1] * dkl[1] + dij[2] * dkl[2];
          const double cr[3] = {djk[1] * dkl[2] - djk[2] * dkl[1], djk[2] * dkl[0] - djk[0] * dkl[2],
                                djk[0] * dkl[1] - djk[1] * dkl[0]};
          const double uc = -d_ij_jk * d_jk_kl + r_jk * r_jk * d_ij_kl;
          const double us = -r_jk * (dij[0] * cr[0] + dij[1] * cr[1] + dij[2] * cr[2]);
          const double omega = atan2(us, uc);
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
          const double tel = r_ij * r_ij + r_jk * r_jk + r_kl * r_kl - r_li * r_li -
                             2.0 * (r_ij * r_jk * cos_ijk - r_ij * r_kl * cos_ijk * cos_jkl + r_jk * r_kl * cos_jkl);
          double arg = tel / poem;
          if (arg > 1.0) arg = 1.0;
          if (arg < -1.0) arg = -1.0;
          double s_ijk = sin_ijk, s_jkl = sin_jkl;   // clamped copies for the divisions
          if (s_ijk >= 0 && s_ijk <= MIN_SINE) s_ijk = MIN_SINE;
          else if (s_ijk <= 0 && s_ijk >= -MIN_SINE) s_ijk = -MIN_SINE;
          if (s_jkl >= 0 && s_jkl <= MIN_SINE) s_jkl = MIN_SINE;
          else if (s_jkl <= 0 && s_jkl >= -MIN_SINE) s_jkl = -MIN_SINE;
          const double sc = 2.0 / poem;
          const double a_ij = (htra - arg * hnra) / r_ij, a_kl = (htrc - arg * hnrc) / r_kl, a_jk = htrb / r_jk;
          const double t_d = (hthd - arg * hnhd) / s_ijk, t_e = (hthe - arg * hnhe) / s_jkl;
          // torsion and conjugation energies
          const double cos_omega = cos(omega), cos2omega = cos(2. * omega), cos3omega = cos(3. * omega);
          const double exp_tor1 = exp(fb.p_tor1 * (2.0 - BOpi_jk - f11_DjDk) * (2.0 - BOpi_jk - f11_DjDk));
          const double exp_tor2_kl = exp(-p_tor2 * BOA_kl);
          const double exp_cot2_kl = exp(-p_cot2 * (BOA_kl - 1.5) * (BOA_kl - 1.5));
          const double fn10 = (1.0 - exp_tor2_ij) * (1.0 - exp_tor2_jk) * (1.0 - exp_tor2_kl);
          const double CV = 0.5 * (fb.V1 * (1.0 + cos_omega) + fb.V2 * exp_tor1 * (1.0 - cos2omega) +
                                   fb.V3 * (1.0 + cos3omega));
          e_tor += fn10 * sin_ijk * sin_jkl * CV;
          const double CEtors1 = sin_ijk * sin_jkl * CV;
          const double CEtors2 = -fn10 * 2.0 * fb.p_tor1 * fb.V2 * exp_tor1 * (2.0 - BOpi_jk - f11_DjDk) *
                                 (1.0 - cos_omega * cos_omega) * sin_ijk * sin_jkl;
          const double CEtors3 = CEtors2 * dfn11;
          const double CEtors4 = CEtors1 * p_tor2 * exp_tor2_ij * (1.0 - exp_tor2_jk) * (1.0 - exp_tor2_kl);
          const double CEtors5 = CEtors1 * p_tor2 * (1.0 - exp_tor2_ij) * exp_tor2_jk * (1.0 - exp_tor2_kl);
          const double CEtors6 = CEtors1 * p_tor2 * (1.0 - exp_tor2_ij) * (1.0 - exp_tor2_jk) * exp_tor2_kl;
          const double cmn = -fn10 * CV;
          const double CEtors7 = cmn * sin_jkl * tan_ijk_i;
          const double CEtors8 = cmn * sin_ijk * tan_jkl_i;
          const double CEtors9 = fn10 * sin_ijk * sin_jkl *
                                 (0.5 * fb.V1 - 2.0 * fb.V2 * exp_tor1 * cos_omega +
                                  1.5 * fb.V3 * (cos2omega + 2.0 * cos_omega * cos_omega));
          const double fn12 = exp_cot2_ij * exp_cot2_jk * exp_cot2_kl;
          const double conj = 1.0 + (cos_omega * cos_omega - 1.0) * sin_ijk * sin_jkl;
          e_con += fb.p_cot1 * fn12 * conj;
          const double Cconj = -2.0 * fn12 * fb.p_cot1 * p_cot2 * conj;
          const double CEconj4 = -fb.p_cot1 * fn12 * (cos_omega * cos_omega - 1.0) * sin_jkl * tan_ijk_i;
          const double CEconj5 = -fb.p_cot1 * fn12 * (cos_omega * cos_omega - 1.0) * sin_ijk * tan_jkl_i;
          const double CEconj6 = 2.0 * fb.p_cot1 * fn12 * cos_omega * sin_ijk * sin_jkl;
          cdbopi_jk += CEtors2;
          fj[3] += CEtors3;
          cd_k += CEtors3;
          cdbo_ij += CEtors4 + Cconj * (BOA_ij - 1.5e0);
          cdbo_jk += CEtors5 + Cconj * (BOA_jk - 1.5e0);
          const double cdbo_kl = CEtors6 + Cconj * (BOA_kl - 1.5e0);
          const double c7 = CEtors7 + CEconj4, c8 = CEtors8 + CEconj5, c9 = CEtors9 + CEconj6;
          double fl[3];
#pragma unroll
          for (int q = 0; q < 3; q++) {
            const double w_i = sc * (a_ij * dij[q] - dvec_li[q] - t_d * ijk_dk[q]);
            const double w_j = sc * (-a_ij * dij[q] - a_jk * djk[q] - t_d * ijk_dj[q] - t_e * jkl_di[q]);
            const double w_k = sc * (-a_kl * dkl[q] + a_jk * djk[q] - t_d * ijk_di[q] - t_e * jkl_dj[q]);
            const double w_l = sc * (a_kl * dkl[q] + dvec_li[q] - t_e * jkl_dk[q]);
            fi[q] += c7 * ijk_dk[q] + c9 * w_i;
            fj[q] += c7 * ijk_dj[q] + c8 * jkl_di[q] + c9 * w_j;
            fk[q] += c7 * ijk_di[q] + c8 * jkl_dj[q] + c9 * w_k;
            fl[q] = c8 * jkl_dk[q] + c9 * w_l;
          }
          unsigned long long *al4 = A.acc + 4 * (size_t)l;
          unsigned long long *const ga[4] = {A.sacc + pw, al4, al4 + 1, al4 + 2};
          const size_t gc[4] = {A.scoff, A.coff, A.coff, A.coff};
          const double gv[4] = {cdbo_kl, fl[0], fl[1], fl[2]};
          fix_add_group<4>(A.flags, ga, gc, gv);
        }
        if (cdbo_ij != 0.0 || fi[0] != 0.0 || fi[1] != 0.0 || fi[2] != 0.0) {
          unsigned long long *ai4 = A.acc + 4 * (size_t)i;
          unsigned long long *const ga[4] = {A.sacc + ph, ai4, ai4 + 1, ai4 + 2};
          const size_t gc[4] = {A.scoff, A.coff, A.coff, A.coff};
          const double gv[4] = {cdbo_ij, fi[0], fi[1], fi[2]};
          fix_add_group<4>(A.flags, ga, gc, gv);
        }
      }
      if (cdbopi_jk != 0.0) A.Cdbopi[pk] += cdbopi_jk;
      if (cdbo_jk != 0.0) A.Cdbo[pk] += cdbo_jk;
      if (cd_k != 0.0 || fk[0] != 0.0 || fk[1] != 0.0 || fk[2] != 0.0) {
        unsigned long long *ak4 = A.acc + 4 * (size_t)k;
        unsigned long long *const ga[4] = {ak4, ak4 + 1, ak4 + 2, ak4 + 3};
        const size_t gc[4] = {A.coff, A.coff, A.coff, A.coff};
        const double gv[4] = {fk[0], fk[1], fk[2], cd_k};
        fix_add_group<4>(A.flags, ga, gc, gv);
      }
    } while (false);
  }
  double *rw = A.row + 6 * (size_t)pk;
#pragma unroll
  for (int q = 0; q < 4; q++) rw[q] = fj[q];
  rw[4] = e_tor;
  rw[5] = e_con;
}

// j's row and torsion/conjugation energies: its slots' entries in slot order.
__global__ void __launch_bounds__(BT) k_to_row(int n, ToArgs A) {
  const int j = blockIdx.x * BT + threadIdx.x;
  if (j >= n) return;
  const int b0 = A.bidx[j], b1 = A.bend[j];
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int pk = b0; pk < b1; ++pk) {
    const double *rw = A.row + 6 * (size_t)pk;
#pragma unroll
    for (int q = 0; q < 6; q++) s[q] += rw[q];
  }
  if (A.atoms[j].type >= 0) {
#pragma unroll
    for (int q = 0; q < 4; q++) A.fcd[4 * (size_t)j + q] += s[q];
  }
  A.part[j] = s[4];
  A.part[A.stride + j] = s[5];
}

// Cdbo of the bond slots += their fixed-point terms; accumulators left at zero.
__global__ void __launch_bounds__(BT) k_to_slot_fold(int nslots, unsigned long long *__restrict__ sacc,
                                                     double *__restrict__ Cdbo) {
  const int s = blockIdx.x * BT + threadIdx.x;
  if (s >= nslots) return;
  unsigned long long *w = sacc + s;
  if (w[0] != 0ull || w[nslots] != 0ull) {
    Cdbo[s] += fix_value(w, (size_t)nslots);
    w[0] = 0ull;
    w[nslots] = 0ull;
  }
}

}  // namespace

int swr_atom_energy_run(swr_system *S) {
  const int n = S->n, N = S->N;
  if (!S->bonds || !S->bo_data || !S->fcd || N <= 0) {
    swr_set_error_msg("Atom_Energy needs the bonds and bond orders of the step (swr_bo first)");
    return SWR_ERR_STATE;
  }
  int rc;
  const size_t stride = (size_t)std::max(n, 1);   // rows as long as the local atoms: the getter copies 5n
  if ((rc = swr_reserve(&S->be_part, S->cap_be_part, 5 * stride)) ||
      (rc = swr_reserve(&S->be_ce, S->cap_be_ce, 2 * stride)) ||
      (rc = swr_reserve(&S->be_result, S->cap_be_result, 8)))
    return rc;
  S->be_stride = stride;
  BeArgs A;
  A.P = S->dp;
  A.atoms = S->packed;
  A.bidx = S->bond_index;
  A.bend = S->bond_end;
  A.bonds = S->bonds;
  A.bo = S->bo_data;
  A.BO = S->BO;
  A.BOpi = S->BOpi;
  const double *ws = S->ws;
  A.Delta = ws + (size_t)WS_DELTA * S->Ncap;
  A.Delta_lp = ws + (size_t)WS_DELTA_LP * S->Ncap;
  A.dDelta_lp = ws + (size_t)WS_DDELTA_LP * S->Ncap;
  A.Delta_lp_temp = ws + (size_t)WS_DELTA_LP_TEMP * S->Ncap;
  A.bo_dboc = S->bo_dboc;
  A.Cdbo = S->Cdbo;
  A.Cdbopi = S->Cdbopi;
  A.Cdbopi2 = S->Cdbopi2;
  A.fcd = S->fcd;
  A.ce = S->be_ce;
  A.part = S->be_part;
  A.stride = stride;
  swr_kt_begin(S, KT_ATOM_ENERGY);
  if (n > 0) hipLaunchKernelGGL(k_atom_energy, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, n, A);
  hipLaunchKernelGGL(k_atom_energy_nbr, dim3(div_up(N, BT)), dim3(BT), 0, S->stream, n, N, A);
  hipLaunchKernelGGL(k_sum_rows, dim3(5), dim3(1024), 0, S->stream, n, (const double *)S->be_part, stride,
                     S->be_result);
  swr_kt_end(S, KT_ATOM_ENERGY);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_dbond_forces_run(swr_system *S) {
  const int N = S->N;
  if (!S->bonds || !S->bo_data || !S->fcd || N <= 0) {
    swr_set_error_msg("Add_All_dBond_to_Forces needs the bonds and bond orders of the step");
    return SWR_ERR_STATE;
  }
  int rc;
  if ((rc = swr_reserve(&S->be_bufc, S->cap_be_bufc, (size_t)S->Ncap))) return rc;
  DbArgs A;
  A.bonds = S->bonds;
  A.bo = S->bo_data;
  A.bidx = S->bond_index;
  A.bend = S->bond_end;
  A.Cdbo = S->Cdbo;
  A.Cdbopi = S->Cdbopi;
  A.Cdbopi2 = S->Cdbopi2;
  A.dDeltap_self = S->dDeltap_self;
  A.fcd = S->fcd;
  A.bufc = S->be_bufc;
  swr_kt_begin(S, KT_DBOND);
  hipLaunchKernelGGL(k_dbond_own, dim3(div_up(N, BT)), dim3(BT), 0, S->stream, N, A);
  hipLaunchKernelGGL(k_dbond_nbr, dim3(div_up(N, BT)), dim3(BT), 0, S->stream, N, A);
  swr_kt_end(S, KT_DBOND);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_hydrogen_bonds_run(swr_system *S) {
  const int n = S->n, N = S->N;
  if (!S->bonds || !S->BO || !S->fcd || N <= 0 || !S->hbonds) {
    swr_set_error_msg("Hydrogen_Bonds needs the bonds, bond orders and hydrogen-bond lists of the step");
    return SWR_ERR_STATE;
  }
  int rc;
  const size_t old_acc = S->cap_hb_acc;
  if ((rc = swr_reserve(&S->hb_part, S->cap_hb_part, (size_t)std::max(n, 1))) ||
      (rc = swr_reserve(&S->hb_result, S->cap_hb_result, 1)) ||
      (rc = swr_reserve(&S->hb_acc, S->cap_hb_acc, 6 * (size_t)N)))
    return rc;
  // a new accumulator starts at zero; afterwards k_hb_fold leaves it so
  if (S->cap_hb_acc != old_acc)
    HIP_CHECK(hipMemsetAsync(S->hb_acc, 0, sizeof(unsigned long long) * S->cap_hb_acc, S->stream));
  HbArgs A;
  A.P = S->dp;
  A.atoms = S->packed;
  A.bidx = S->bond_index;
  A.bend = S->bond_end;
  A.bonds = S->bonds;
  A.BO = S->BO;
  A.Hindex = S->Hindex;
  A.hidx = S->hbond_index;
  A.hend = S->hbond_end;
  A.hbonds = S->hbonds;
  A.Cdbo = S->Cdbo;
  A.fcd = S->fcd;
  A.part = S->hb_part;
  A.acc = S->hb_acc;
  A.coff = 3 * (size_t)N;
  A.flags = &S->status->flags;
  constexpr int WPB = BT / SWR_WAVE;
  swr_kt_begin(S, KT_HBONDS);
  if (n > 0) hipLaunchKernelGGL(k_hb, dim3(div_up(n, WPB)), dim3(BT), 0, S->stream, n, A);
  hipLaunchKernelGGL(k_hb_fold, dim3(div_up(N, BT)), dim3(BT), 0, S->stream, N, S->hb_acc, S->fcd);
  hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(1024), 0, S->stream, n, (const double *)S->hb_part, (size_t)0,
                     S->hb_result);
  swr_kt_end(S, KT_HBONDS);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_valence_angles_run(swr_system *S) {
  const int n = S->n, N = S->N;
  if (!S->bonds || !S->BO || !S->fcd || N <= 0) {
    swr_set_error_msg("Valence_Angles needs the bonds and bond orders of the step (swr_bo first)");
    return SWR_ERR_STATE;
  }
  int rc;
  const size_t stride = (size_t)std::max(n, 1);
  const size_t old_acc = S->cap_va_acc;
  const size_t old_ctl = S->cap_va_ctl;
  if ((rc = swr_reserve(&S->va_part, S->cap_va_part, 3 * stride)) ||
      (rc = swr_reserve(&S->va_result, S->cap_va_result, 4)) ||
      (rc = swr_reserve(&S->va_acc, S->cap_va_acc, 8 * (size_t)N)) ||
      (rc = swr_reserve(&S->va_cnt, S->cap_va_cnt, stride)) ||
      (rc = swr_reserve(&S->va_order, S->cap_va_order, stride)) ||
      (rc = swr_reserve(&S->va_ctl, S->cap_va_ctl, (size_t)WL_CTL_INTS)))
    return rc;
  if (S->cap_va_acc != old_acc)
    HIP_CHECK(hipMemsetAsync(S->va_acc, 0, sizeof(unsigned long long) * S->cap_va_acc, S->stream));
  // the histogram starts at zero; k_wl_scan leaves it so
  if (S->cap_va_ctl != old_ctl) HIP_CHECK(hipMemsetAsync(S->va_ctl, 0, sizeof(int) * S->cap_va_ctl, S->stream));
  S->va_stride = stride;
  VaArgs A;
  A.P = S->dp;
  A.atoms = S->packed;
  A.bidx = S->bond_index;
  A.bend = S->bond_end;
  A.bonds = S->bonds;
  A.BO = S->BO;
  A.BOpi = S->BOpi;
  A.bo_dboc = S->bo_dboc;
  const double *ws = S->ws;
  A.Delta = ws + (size_t)WS_DELTA * S->Ncap;
  A.Delta_val = ws + (size_t)WS_DELTA_VAL * S->Ncap;
  A.vlpex = ws + (size_t)WS_VLPEX * S->Ncap;
  A.nlp = ws + (size_t)WS_NLP * S->Ncap;
  A.dDelta_lp = ws + (size_t)WS_DDELTA_LP * S->Ncap;
  A.Cdbo = S->Cdbo;
  A.Cdbopi = S->Cdbopi;
  A.Cdbopi2 = S->Cdbopi2;
  A.fcd = S->fcd;
  A.part = S->va_part;
  A.stride = stride;
  A.acc = S->va_acc;
  A.coff = 4 * (size_t)N;
  A.flags = &S->status->flags;
  A.cnt = S->va_cnt;
  A.order = S->va_order;
  A.ctl = S->va_ctl;
  swr_kt_begin(S, KT_VALENCE);
  if (n > 0) {
    hipLaunchKernelGGL(k_va_count, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, n, A);
    hipLaunchKernelGGL(k_wl_scan, dim3(1), dim3(WL_NB), 0, S->stream, S->va_ctl);
    hipLaunchKernelGGL(k_va_scatter, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, n, A);
    hipLaunchKernelGGL(k_valence, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, A);
  }
  hipLaunchKernelGGL(k_va_fold, dim3(div_up(N, BT)), dim3(BT), 0, S->stream, N, S->va_acc, S->fcd);
  hipLaunchKernelGGL(k_sum_rows, dim3(3), dim3(1024), 0, S->stream, n, (const double *)S->va_part, stride,
                     S->va_result);
  swr_kt_end(S, KT_VALENCE);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_torsion_angles_run(swr_system *S) {
  const int n = S->n, N = S->N;
  if (!S->bonds || !S->BO || !S->fcd || N <= 0) {
    swr_set_error_msg("Torsion_Angles needs the bonds and bond orders of the step (swr_bo first)");
    return SWR_ERR_STATE;
  }
  int rc;
  const size_t stride = (size_t)std::max(n, 1);
  const int nslots = S->bond_cap;
  const size_t old_acc = S->cap_to_acc, old_sacc = S->cap_to_sacc, old_ctl = S->cap_to_ctl;
  const size_t slots = (size_t)std::max(nslots, 1);
  if ((rc = swr_reserve(&S->to_part, S->cap_to_part, 2 * stride)) ||
      (rc = swr_reserve(&S->to_result, S->cap_to_result, 4)) ||
      (rc = swr_reserve(&S->to_acc, S->cap_to_acc, 8 * (size_t)N)) ||
      (rc = swr_reserve(&S->to_sacc, S->cap_to_sacc, 2 * slots)) ||
      (rc = swr_reserve(&S->to_tc, S->cap_to_tc, slots)) ||
      (rc = swr_reserve(&S->to_task, S->cap_to_task, 2 * slots)) ||
      (rc = swr_reserve(&S->to_row, S->cap_to_row, 6 * slots)) ||
      (rc = swr_reserve(&S->to_ctl, S->cap_to_ctl, (size_t)WL_CTL_INTS)))
    return rc;
  if (S->cap_to_acc != old_acc)
    HIP_CHECK(hipMemsetAsync(S->to_acc, 0, sizeof(unsigned long long) * S->cap_to_acc, S->stream));
  if (S->cap_to_sacc != old_sacc)
    HIP_CHECK(hipMemsetAsync(S->to_sacc, 0, sizeof(unsigned long long) * S->cap_to_sacc, S->stream));
  // the histogram starts at zero; k_wl_scan leaves it so
  if (S->cap_to_ctl != old_ctl) HIP_CHECK(hipMemsetAsync(S->to_ctl, 0, sizeof(int) * S->cap_to_ctl, S->stream));
  S->to_stride = stride;
  ToArgs A;
  A.P = S->dp;
  A.atoms = S->packed;
  A.bidx = S->bond_index;
  A.bend = S->bond_end;
  A.bonds = S->bonds;
  A.BO = S->BO;
  A.BOpi = S->BOpi;
  A.bo_dboc = S->bo_dboc;
  A.Cdbo = S->Cdbo;
  A.Cdbopi = S->Cdbopi;
  A.fcd = S->fcd;
  A.part = S->to_part;
  A.stride = stride;
  A.acc = S->to_acc;
  A.coff = 4 * (size_t)N;
  A.flags = &S->status->flags;
  A.sacc = S->to_sacc;
  A.scoff = slots;
  A.bnbr = S->bond_nbr;
  A.tc = S->to_tc;
  A.task = reinterpret_cast<int2 *>(S->to_task);
  A.ctl = S->to_ctl;
  A.row = S->to_row;
  swr_kt_begin(S, KT_TORSION);
  if (n > 0) {
    hipLaunchKernelGGL(k_to_count, dim3(div_up(n * TO_CNT_LANES, BT)), dim3(BT), 0, S->stream, n, A);
    hipLaunchKernelGGL(k_wl_scan, dim3(1), dim3(WL_NB), 0, S->stream, S->to_ctl);
    hipLaunchKernelGGL(k_to_scatter, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, n, A);
    // (at most one task per bond slot; lanes past the list's length exit)
    hipLaunchKernelGGL(k_torsion, dim3(div_up(nslots, BT)), dim3(BT), 0, S->stream, A);
    hipLaunchKernelGGL(k_to_row, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, n, A);
  }
  hipLaunchKernelGGL(k_va_fold, dim3(div_up(N, BT)), dim3(BT), 0, S->stream, N, S->to_acc, S->fcd);
  if (nslots > 0)
    hipLaunchKernelGGL(k_to_slot_fold, dim3(div_up(nslots, BT)), dim3(BT), 0, S->stream, nslots, S->to_sacc, S->Cdbo);
  hipLaunchKernelGGL(k_sum_rows, dim3(2), dim3(1024), 0, S->stream, n, (const double *)S->to_part, stride,
                     S->to_result);
  swr_kt_end(S, KT_TORSION);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}
