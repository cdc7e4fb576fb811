// bo.hip — corrected bond orders (a9) and the Delta / lone-pair terms (a10).
//
// Restates BO_Section2_C / BO_Section3_C (host branches,
// reaxc_bond_orders_sw64.c:163-376 and :451-481).  Section 1 (Deltap,
// Deltap_boc) is produced by the list pass, which owns each atom's BO' sum.
// Eight lanes per atom, one bond each (an atom's bonds are consecutive
// slots, so a group's loads and stores touch neighbouring records); an atom
// with more than eight bonds takes them eight at a time.  bo_dboc[i][0]
// sums the corrected bond orders in list (nbr-sorted) order, as the
// reference's loop does: the group's first lane adds the lanes' values in
// slot order.  sym_index is found by binary search in j's nbr-sorted list,
// over the compact bond_nbr array the list pass writes (the reference scans
// the list linearly and keeps the last match; pairs are unique, so both
// give the same slot).
#include "swr_internal.h"
#include "swr_device.h"

namespace {

constexpr int BO_LPA = 8;    // lanes per atom
constexpr int BO_T = 256;    // threads per block

__global__ void __launch_bounds__(BO_T)
k_bo(int N, const DevParams *__restrict__ P, const swr_atom_pack_t *__restrict__ atoms,
     const int *__restrict__ bond_index, const int *__restrict__ bond_end, const int *__restrict__ bond_nbr,
     swr_bond_t *__restrict__ bonds, swr_bond_order_t *__restrict__ bo_data, double *__restrict__ BO,
     double *__restrict__ BOpi, double *__restrict__ bo_dboc, const double *__restrict__ Deltap,
     const double *__restrict__ Deltap_boc, double *__restrict__ Delta, double *__restrict__ Delta_e,
     double *__restrict__ Delta_val, double *__restrict__ vlpex, double *__restrict__ nlp,
     double *__restrict__ nlp_temp, double *__restrict__ Clp, double *__restrict__ Delta_lp,
     double *__restrict__ dDelta_lp, double *__restrict__ Delta_lp_temp, double *__restrict__ dDelta_lp_temp) {
  const int gid = xcd_block(gridDim.x) * BO_T + threadIdx.x;
  const int i = gid / BO_LPA, sl = gid % BO_LPA;
  const int lead = lane_id() - sl;        // the group's first lane
  if (i >= N) return;                     // whole groups leave together
  const int ei = atoms[i].type;
  if (ei < 0) return;
  const double p_boc1 = P->p_boc1, p_boc2 = P->p_boc2;
  const double val_i = P->valency[ei];
  const double Deltap_i = Deltap[i], Deltap_boc_i = Deltap_boc[i];
  const int bs = bond_index[i], be = bond_end[i];
  double total = 0.0;
  for (int c0 = bs; c0 < be; c0 += BO_LPA) {
    const int pj = c0 + sl;
    double out = 0.0;
    const int j = pj < be ? bond_nbr[pj] : 0;
    const int ej = pj < be ? atoms[j].type : -1;
    if (ej >= 0) {
      // sym_index: slot of i in j's nbr-sorted list
      // (stored only when found, which it always is for pairs of atoms of
      // the system; the record is not read)
      int lo = bond_index[j], hi = bond_end[j] - 1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = bond_nbr[mid];
        if (v == i) { bonds[pj].sym_index = mid; break; }
        if (v < i) lo = mid + 1; else hi = mid - 1;
      }
      const DevTbp tw = P->tbp[ei * SWR_MAX_ELEM + ej];
      swr_bond_order_t &bo = bo_data[pj];
      double bo_ij = BO[pj];
      double pi = BOpi[2 * (size_t)pj], pi2 = BOpi[2 * (size_t)pj + 1];
      double BO_s = bo.BO_s;
      if (tw.ovc < 0.001 && tw.v13cor < 0.001) {
        bo.C1dbo = 1.0; bo.C2dbo = 0.0; bo.C3dbo = 0.0;
        bo.C1dbopi = pi; bo.C2dbopi = 0.0; bo.C3dbopi = 0.0; bo.C4dbopi = 0.0;
        bo.C1dbopi2 = pi2; bo.C2dbopi2 = 0.0; bo.C3dbopi2 = 0.0; bo.C4dbopi2 = 0.0;
      } else {
        const double val_j = P->valency[ej];
        const double Deltap_j = Deltap[j], Deltap_boc_j = Deltap_boc[j];
        double f1, Cf1_ij, Cf1_ji;
        if (tw.ovc >= 0.001) {
          const double exp_p1i = exp(-p_boc1 * Deltap_i);
          const double exp_p2i = exp(-p_boc2 * Deltap_i);
          const double exp_p1j = exp(-p_boc1 * Deltap_j);
          const double exp_p2j = exp(-p_boc2 * Deltap_j);
          const double f2 = exp_p1i + exp_p1j;
          const double f3 = -1.0 / p_boc2 * log(0.5 * (exp_p2i + exp_p2j));
          f1 = 0.5 * ((val_i + f2) / (val_i + f2 + f3) + (val_j + f2) / (val_j + f2 + f3));
          const double temp = f2 + f3;
          const double u1_ij = val_i + temp, u1_ji = val_j + temp;
          const double Cf1A_ij = 0.5 * f3 * (1.0 / (u1_ij * u1_ij) + 1.0 / (u1_ji * u1_ji));
          const double Cf1B_ij = -0.5 * ((u1_ij - f3) / (u1_ij * u1_ij) + (u1_ji - f3) / (u1_ji * u1_ji));
          Cf1_ij = 0.50 * (-p_boc1 * exp_p1i / u1_ij -
                           ((val_i + f2) / (u1_ij * u1_ij)) * (-p_boc1 * exp_p1i + exp_p2i / (exp_p2i + exp_p2j)) +
                           -p_boc1 * exp_p1i / u1_ji -
                           ((val_j + f2) / (u1_ji * u1_ji)) * (-p_boc1 * exp_p1i + exp_p2i / (exp_p2i + exp_p2j)));
          Cf1_ji = -Cf1A_ij * p_boc1 * exp_p1j + Cf1B_ij * exp_p2j / (exp_p2i + exp_p2j);
        } else {
          f1 = 1.0;
          Cf1_ij = Cf1_ji = 0.0;
        }
        double f4, f5, f4f5, Cf45_ij, Cf45_ji;
        if (tw.v13cor >= 0.001) {
          const double exp_f4 = exp(-(tw.p_boc4 * (bo_ij * bo_ij) - Deltap_boc_i) * tw.p_boc3 + tw.p_boc5);
          const double exp_f5 = exp(-(tw.p_boc4 * (bo_ij * bo_ij) - Deltap_boc_j) * tw.p_boc3 + tw.p_boc5);
          f4 = 1. / (1. + exp_f4);
          f5 = 1. / (1. + exp_f5);
          f4f5 = f4 * f5;
          Cf45_ij = -f4 * exp_f4;
          Cf45_ji = -f5 * exp_f5;
        } else {
          f4 = f5 = f4f5 = 1.0;
          Cf45_ij = Cf45_ji = 0.0;
        }
        const double A0_ij = f1 * f4f5;
        const double A1_ij = -2 * tw.p_boc3 * tw.p_boc4 * bo_ij * (Cf45_ij + Cf45_ji);
        const double A2_ij = Cf1_ij / f1 + tw.p_boc3 * Cf45_ij;
        const double A2_ji = Cf1_ji / f1 + tw.p_boc3 * Cf45_ji;
        const double A3_ij = A2_ij + Cf1_ij / f1;
        const double A3_ji = A2_ji + Cf1_ji / f1;
        bo_ij = bo_ij * A0_ij;
        pi = pi * A0_ij * f1;
        pi2 = pi2 * A0_ij * f1;
        BO_s = bo_ij - (pi + pi2);
        bo.C1dbo = A0_ij + bo_ij * A1_ij;
        bo.C2dbo = bo_ij * A2_ij;
        bo.C3dbo = bo_ij * A2_ji;
        bo.C1dbopi = f1 * f1 * f4 * f5;
        bo.C2dbopi = pi * A1_ij;
        bo.C3dbopi = pi * A3_ij;
        bo.C4dbopi = pi * A3_ji;
        bo.C1dbopi2 = f1 * f1 * f4 * f5;
        bo.C2dbopi2 = pi2 * A1_ij;
        bo.C3dbopi2 = pi2 * A3_ij;
        bo.C4dbopi2 = pi2 * A3_ji;
      }
      if (bo_ij < 1e-10) bo_ij = 0.0;
      if (BO_s < 1e-10) BO_s = 0.0;
      if (pi < 1e-10) pi = 0.0;
      if (pi2 < 1e-10) pi2 = 0.0;
      BO[pj] = bo_ij;
      BOpi[2 * (size_t)pj] = pi;
      BOpi[2 * (size_t)pj + 1] = pi2;
      bo.BO_s = BO_s;
      out = bo_ij;
    }
    // the group's sum in slot order (slots past the end add nothing)
#pragma unroll
    for (int k = 0; k < BO_LPA; k++) {
      const double v = __shfl(out, lead + k, SWR_WAVE);
      if (c0 + k < be) total += v;
    }
  }
  if (sl != 0) return;
  // Section 3 (reaxc_bond_orders_sw64.c:451-481)
  const double p_lp1 = P->p_lp1;
  bo_dboc[2 * (size_t)i] = total;
  const double De = total - P->valency_e[ei];
  Delta[i] = total - P->valency[ei];
  Delta_e[i] = De;
  bo_dboc[2 * (size_t)i + 1] = total - P->valency_boc[ei];
  Delta_val[i] = total - P->valency_val[ei];
  const double vx = De - 2.0 * (int)(De / 2.0);
  vlpex[i] = vx;
  const double explp1 = exp(-p_lp1 * ((2.0 + vx) * (2.0 + vx)));
  const double nl = explp1 - (int)(De / 2.0);
  nlp[i] = nl;
  const double nlp_opt = P->nlp_opt[ei];
  Delta_lp[i] = nlp_opt - nl;
  const double clp = 2.0 * p_lp1 * explp1 * (2.0 + vx);
  Clp[i] = clp;
  dDelta_lp[i] = clp;
  if (P->mass[ei] > 21.0) {
    const double nt = 0.5 * (P->valency_e[ei] - P->valency[ei]);
    nlp_temp[i] = nt;
    Delta_lp_temp[i] = nlp_opt - nt;
    dDelta_lp_temp[i] = 0.;
  } else {
    nlp_temp[i] = nl;
    Delta_lp_temp[i] = nlp_opt - nl;
    dDelta_lp_temp[i] = clp;
  }
}

}  // namespace

int swr_bo_run(swr_system *S) {
  const int N = S->N;
  const size_t c = S->Ncap;
  double *w = S->ws;
  swr_kt_begin(S, KT_BO);
  hipLaunchKernelGGL(k_bo, dim3(div_up((long long)N * BO_LPA, BO_T)), dim3(BO_T), 0, S->stream, N, S->dp, S->packed,
                     S->bond_index, S->bond_end, S->bond_nbr, S->bonds, S->bo_data, S->BO, S->BOpi, S->bo_dboc,
                     w + WS_DELTAP * c,
                     w + WS_DELTAP_BOC * c, w + WS_DELTA * c, w + WS_DELTA_E * c, w + WS_DELTA_VAL * c,
                     w + WS_VLPEX * c, w + WS_NLP * c, w + WS_NLP_TEMP * c, w + WS_CLP * c, w + WS_DELTA_LP * c,
                     w + WS_DDELTA_LP * c, w + WS_DELTA_LP_TEMP * c, w + WS_DDELTA_LP_TEMP * c);
  swr_kt_end(S, KT_BO);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}
