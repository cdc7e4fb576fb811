// nonbonded.hip — the non-bonded pass over the far list: shielded Morse vdW
// (+ inner wall) and tapered shielded Coulomb energies, their gradient and
// virial (vdW_Coulomb_Energy_Full_C, reaxc_nonbonded_sw64.c:260-500; host
// form :522-830; vdw_type 1 CPE kernel reaxc_nonbonded_cpe.h:1-420).
//
// One wave per local atom streams its far row 64 entries at a time (one
// 64-byte entry per lane, read as four 16-byte loads).  Every entry is
// evaluated with the reference's expressions; the powers are formed from
// exp/log (r^p = exp(p ln r), (r^p + g)^(1/p) = exp(ln(.)/p), and dfn13 =
// fn13 / (r^p + g) * r^p / r^2, which equals (r^p + g)^(1/p - 1) r^(p - 2))
// and the Coulomb cube root with cbrt, so a pair costs 4 exp, 2 log and one
// cbrt.  The gradient -(CEvd + CEclmb) dvec of atom i is summed over the
// wave and added to fCdDelta[i]; energies and virial take half of each
// full-list pair and are reduced in a fixed order (per-block partials, then
// one workgroup), so results are identical run to run; each atom's share
// (eflag_atom's ei) is written too, for per-system energies of a batch.
#include "swr_internal.h"
#include "swr_device.h"

namespace {

constexpr int NBT = 256;                // threads per block
constexpr int NBW = NBT / SWR_WAVE;     // atoms per block
constexpr double C_ELE = 332.06371;     // kcal/mol A e^-2

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = SWR_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, SWR_WAVE);
  return v;
}

// 4 waves per SIMD: the register budget that keeps the prefetched entry
// without spills (740 vs 817 us at 3 waves on the 96k-atom box)
__global__ void __launch_bounds__(NBT) __attribute__((amdgpu_waves_per_eu(4)))
k_vdw_coulomb(int n, const DevParams *__restrict__ P, const swr_atom_pack_t *__restrict__ atoms,
              const int *__restrict__ far_index, const int *__restrict__ far_end,
              const swr_far_neighbor_t *__restrict__ far, double *__restrict__ fcd, double *__restrict__ eatom,
              double *__restrict__ partials) {
  __shared__ double sacc[NBW][8];
  const int w = wave_in_block(), lane = lane_id();
  const int blk = blockIdx.x;   // rows and partial slot
  const int i = blk * NBW + w;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // eng_vdwl, eng_coul, virial[6]
  int ei = -1;
  swr_atom_pack_t ai = {};
  if (i < n) {
    ai = atoms[i];
    ei = ai.type;
  }
  if (ei >= 0) {
    const int vt = P->vdw_type;
    const double p = P->p_vdW1, pinv = P->p_vdW1i;
    const double *tp = P->nb_tap;
    const DevNb *nbrow = P->nb + ei * SWR_MAX_ELEM;
    const int beg = far_index[i], end = far_end[i];
    double g0 = 0, g1 = 0, g2 = 0;
    // the next entry is requested while this one is evaluated
    int k = beg + lane;
    int4 nh0 = make_int4(0, -1, 0, 0);
    double2 ne1 = make_double2(0, 0), ne2 = ne1, ne3 = ne1;
    if (k < end) {
      const double2 *e = reinterpret_cast<const double2 *>(far + k);
      nh0 = *reinterpret_cast<const int4 *>(far + k);
      ne1 = e[1]; ne2 = e[2]; ne3 = e[3];
    }
    for (; k < end; k += SWR_WAVE) {
      const int4 h0 = nh0;
      const double2 e1 = ne1, e2 = ne2, e3 = ne3;
      if (k + SWR_WAVE < end) {
        const double2 *e = reinterpret_cast<const double2 *>(far + k + SWR_WAVE);
        nh0 = *reinterpret_cast<const int4 *>(far + k + SWR_WAVE);
        ne1 = e[1]; ne2 = e[2]; ne3 = e[3];
      }
      const int tj = h0.y;
      if (tj < 0) continue;
      const double r = e1.y, qj = e2.x, dx = e2.y, dy = e3.x, dz = e3.y;
      const DevNb nb = nbrow[tj];
      const double rinv = 1.0 / r;
      double Tap = tp[7] * r + tp[6];
      Tap = Tap * r + tp[5]; Tap = Tap * r + tp[4]; Tap = Tap * r + tp[3];
      Tap = Tap * r + tp[2]; Tap = Tap * r + tp[1]; Tap = Tap * r + tp[0];
      double dTap = 7 * tp[7] * r + 6 * tp[6];
      dTap = dTap * r + 5 * tp[5]; dTap = dTap * r + 4 * tp[4];
      dTap = dTap * r + 3 * tp[3]; dTap = dTap * r + 2 * tp[2];
      dTap += tp[1] * rinv;
      double e_vdW, CEvd, e_core = 0.0;
      if (vt == 1 || vt == 3) {
        const double lnr = log(r);
        const double powr = exp(p * lnr);
        const double base = powr + nb.powgi_vdW1;
        const double fn13 = exp(log(base) * pinv);
        const double exp2 = exp(0.5 * nb.alpha * (1.0 - fn13 * nb.r_vdWinv));
        const double exp1 = exp2 * exp2;
        e_vdW = nb.D * (exp1 - 2.0 * exp2);
        const double dfn13 = fn13 / base * powr * (rinv * rinv);
        CEvd = dTap * e_vdW - Tap * nb.D * nb.alpha_div_rvdW * (exp1 - exp2) * dfn13;
      } else {
        const double exp2 = exp(0.5 * nb.alpha * (1.0 - r * nb.r_vdWinv));
        const double exp1 = exp2 * exp2;
        e_vdW = nb.D * (exp1 - 2.0 * exp2);
        CEvd = dTap * e_vdW - Tap * nb.D * nb.alpha_div_rvdW * (exp1 - exp2) * rinv;
      }
      if (vt == 2 || vt == 3) {
        e_core = nb.ecore * exp(nb.acore * (1.0 - r * nb.rcoreinv));
        const double de_core = -nb.a_div_rcore * e_core;
        CEvd += dTap * e_core + Tap * de_core * rinv;
      }
      const double dr3gamij_1 = r * r * r + nb.gamma;
      const double dr3gamij_3 = cbrt(dr3gamij_1);
      const double qq = C_ELE * ai.q * qj;
      const double e_ele = qq * Tap / dr3gamij_3;
      const double CEclmb = qq * (dTap - Tap * r / dr3gamij_1) / dr3gamij_3;
      const double fpair = -(CEvd + CEclmb);
      acc[0] += 0.5 * Tap * (e_vdW + e_core);
      acc[1] += 0.5 * e_ele;
      acc[2] += 0.5 * dx * dx * fpair;
      acc[3] += 0.5 * dy * dy * fpair;
      acc[4] += 0.5 * dz * dz * fpair;
      acc[5] += 0.5 * dx * dy * fpair;
      acc[6] += 0.5 * dx * dz * fpair;
      acc[7] += 0.5 * dy * dz * fpair;
      g0 += fpair * dx;
      g1 += fpair * dy;
      g2 += fpair * dz;
    }
    g0 = wave_sum_d(g0);
    g1 = wave_sum_d(g1);
    g2 = wave_sum_d(g2);
    const double ea = wave_sum_d(acc[0] + acc[1]);  // the atom's share (eflag_atom's ei)
    if (lane == 0) {
      fcd[4 * (size_t)i] += g0;
      fcd[4 * (size_t)i + 1] += g1;
      fcd[4 * (size_t)i + 2] += g2;
      eatom[i] = ea;
    }
  } else if (i < n && lane == 0) {
    eatom[i] = 0.0;
  }
#pragma unroll
  for (int c = 0; c < 8; c++) acc[c] = wave_sum_d(acc[c]);
  if (lane == 0)
    for (int c = 0; c < 8; c++) sacc[w][c] = acc[c];
  __syncthreads();
  if (threadIdx.x < 8) {
    double t = 0.0;
    for (int ww = 0; ww < NBW; ww++) t += sacc[ww][threadIdx.x];
    partials[(size_t)threadIdx.x * gridDim.x + blk] = t;  // component-major
  }
}

// Sum of the per-block partials in a fixed order: one workgroup per
// component (8), each thread a strided share, then a tree in LDS.
constexpr int SUMT = 1024;
__global__ void __launch_bounds__(SUMT) k_vdw_coulomb_sum(int nblk, const double *__restrict__ partials,
                                                          double *__restrict__ out) {
  __shared__ double s[SUMT];
  const double *p = partials + (size_t)blockIdx.x * nblk;
  double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
  int b = threadIdx.x;
  for (; b + 3 * SUMT < nblk; b += 4 * SUMT) {
    t0 += p[b]; t1 += p[b + SUMT]; t2 += p[b + 2 * SUMT]; t3 += p[b + 3 * SUMT];
  }
  for (; b < nblk; b += SUMT) t0 += p[b];
  s[threadIdx.x] = (t0 + t1) + (t2 + t3);
  __syncthreads();
  for (int stride = SUMT / 2; stride > 0; stride >>= 1) {
    if ((int)threadIdx.x < stride) s[threadIdx.x] += s[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

}  // namespace

int swr_nonbonded_run(swr_system *S) {
  const int n = S->n;
  if (n <= 0) return SWR_OK;
  if (!S->fcd || !S->far || !S->far_index) {
    swr_set_error_msg("swr_vdw_coulomb_energy: the far list is not built");
    return SWR_ERR_STATE;
  }
  const int nblk = div_up(n, NBW);
  if ((size_t)nblk * 8 > S->nb_partials_cap) {
    int rc = swr_alloc(&S->nb_partials, (size_t)nblk * 8);
    if (rc) return rc;
    S->nb_partials_cap = (size_t)nblk * 8;
  }
  if (!S->nb_result) {
    int rc = swr_alloc(&S->nb_result, 8);
    if (rc) return rc;
  }
  if ((size_t)n > S->nb_eatom_cap) {
    int rc = swr_alloc(&S->nb_eatom, (size_t)S->ncap);
    if (rc) return rc;
    S->nb_eatom_cap = (size_t)S->ncap;
  }
  swr_kt_begin(S, KT_NONBONDED);
  hipLaunchKernelGGL(k_vdw_coulomb, dim3(nblk), dim3(NBT), 0, S->stream, n, S->dp, S->packed, S->far_index,
                     S->far_end, S->far, S->fcd, S->nb_eatom, S->nb_partials);
  hipLaunchKernelGGL(k_vdw_coulomb_sum, dim3(8), dim3(SUMT), 0, S->stream, nblk, S->nb_partials, S->nb_result);
  swr_kt_end(S, KT_NONBONDED);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}
