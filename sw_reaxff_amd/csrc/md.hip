// md.hip — forces on the local atoms and the velocity-Verlet NVE update.
//
// Forces: PairReaxCSunway::compute copies the energy gradient of every atom,
// ghosts included, into atom->f (pair_reaxc_sunway.cpp:746-760:
// f[k] += -fCdDelta[k]), and LAMMPS' reverse communication then adds each
// ghost's force onto the local atom it images.  On one domain the owner of a
// ghost is owner[g]; k_forces gathers, for each local atom, its own row and
// the rows of its ghosts in ascending ghost order (a CSR built by
// k_ghost_count / scan / k_ghost_fill / k_ghost_sort whenever the ghost set
// changes), so every sum has a fixed order and no floating-point atomics.
//
// NVE: fix_nve_initial_integrate_para / fix_nve_final_integrate_para
// (fix_nve_sw64.c:25-173; dtv = dt, dtf = dt/2 * ftm2v at
// fix_nve_sunway.cpp:56-57): v += dtf/m f; x += dtv v, then after the force
// computation v += dtf/m f, with the per-type mass of atom->mass.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "swr_internal.h"

namespace {

constexpr int BT = 256;

__global__ void __launch_bounds__(BT) k_ghost_count(const int *__restrict__ owner, int n, int N,
                                                    int *__restrict__ cnt) {
  const int g = n + blockIdx.x * BT + threadIdx.x;
  if (g >= N) return;
  atomicAdd(&cnt[owner[g]], 1);
}

__global__ void __launch_bounds__(BT) k_ghost_fill(const int *__restrict__ owner, int n, int N,
                                                   const int *__restrict__ start, int *__restrict__ fill,
                                                   int *__restrict__ list) {
  const int g = n + blockIdx.x * BT + threadIdx.x;
  if (g >= N) return;
  const int o = owner[g];
  list[start[o] + atomicAdd(&fill[o], 1)] = g;
}

// The fill order above depends on arrival; sorting each owner's segment
// makes the summation order of k_forces the ascending ghost order.
__global__ void __launch_bounds__(BT) k_ghost_sort(int n, const int *__restrict__ start, int *__restrict__ list) {
  const int o = blockIdx.x * BT + threadIdx.x;
  if (o >= n) return;
  const int b = start[o], e = start[o + 1];
  for (int p = b + 1; p < e; ++p) {
    const int g = list[p];
    int q = p - 1;
    while (q >= b && list[q] > g) {
      list[q + 1] = list[q];
      --q;
    }
    list[q + 1] = g;
  }
}

__global__ void __launch_bounds__(BT) k_forces(int n, const double *__restrict__ fcd, const int *__restrict__ start,
                                               const int *__restrict__ list, double *__restrict__ f) {
  const int o = blockIdx.x * BT + threadIdx.x;
  if (o >= n) return;
  double g0 = fcd[4 * (size_t)o], g1 = fcd[4 * (size_t)o + 1], g2 = fcd[4 * (size_t)o + 2];
  for (int p = start[o]; p < start[o + 1]; ++p) {
    const size_t g = list[p];
    g0 += fcd[4 * g];
    g1 += fcd[4 * g + 1];
    g2 += fcd[4 * g + 2];
  }
  f[3 * (size_t)o] = -g0;
  f[3 * (size_t)o + 1] = -g1;
  f[3 * (size_t)o + 2] = -g2;
}

__global__ void __launch_bounds__(BT) k_nve_initial(int n, const int *__restrict__ type,
                                                    const DevParams *__restrict__ P, double dtv, double dtf,
                                                    const double *__restrict__ f, double *__restrict__ v,
                                                    double *__restrict__ x) {
  const int i = blockIdx.x * BT + threadIdx.x;
  if (i >= n) return;
  const double dtfm = dtf / P->type_mass[type[i]];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t k = 3 * (size_t)i + c;
    const double vk = v[k] + dtfm * f[k];
    v[k] = vk;
    x[k] += dtv * vk;
  }
}

__global__ void __launch_bounds__(BT) k_nve_final(int n, const int *__restrict__ type,
                                                  const DevParams *__restrict__ P, double dtf,
                                                  const double *__restrict__ f, double *__restrict__ v) {
  const int i = blockIdx.x * BT + threadIdx.x;
  if (i >= n) return;
  const double dtfm = dtf / P->type_mass[type[i]];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t k = 3 * (size_t)i + c;
    v[k] += dtfm * f[k];
  }
}

int ghost_csr(swr_system *S) {
  const int n = S->n, N = S->N;
  if (S->gcsr_valid && S->gcsr_n == n && S->gcsr_N == N) return SWR_OK;
  int rc;
  if ((rc = swr_reserve(&S->gcsr_cnt, S->cap_gcsr_cnt, 2 * (size_t)n + 2)) ||
      (rc = swr_reserve(&S->gcsr_start, S->cap_gcsr_start, (size_t)n + 1)) ||
      (rc = swr_reserve(&S->gcsr_list, S->cap_gcsr_list, (size_t)std::max(N - n, 1))))
    return rc;
  int *cnt = S->gcsr_cnt, *fill = S->gcsr_cnt + n + 1;
  HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * (2 * (size_t)n + 2), S->stream));
  const int ng = N - n;
  if (ng > 0)
    hipLaunchKernelGGL(k_ghost_count, dim3(div_up(ng, BT)), dim3(BT), 0, S->stream, S->owner, n, N, cnt);
  if ((rc = swr_scan_i32(S, cnt, S->gcsr_start, n, S->gcsr_start + n))) return rc;
  if (ng > 0) {
    hipLaunchKernelGGL(k_ghost_fill, dim3(div_up(ng, BT)), dim3(BT), 0, S->stream, S->owner, n, N,
                       (const int *)S->gcsr_start, fill, S->gcsr_list);
    hipLaunchKernelGGL(k_ghost_sort, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, n, (const int *)S->gcsr_start,
                       S->gcsr_list);
  }
  HIP_CHECK(hipGetLastError());
  S->gcsr_valid = true;
  S->gcsr_n = n;
  S->gcsr_N = N;
  return SWR_OK;
}

}  // namespace

int swr_forces_run(swr_system *S) {
  const int n = S->n;
  if (!S->fcd || n <= 0) {
    swr_set_error_msg("forces need the gradient (fCdDelta) of the step");
    return SWR_ERR_STATE;
  }
  int rc;
  if ((rc = ghost_csr(S))) return rc;
  if ((rc = swr_reserve(&S->f, S->cap_f, 3 * (size_t)n))) return rc;
  hipLaunchKernelGGL(k_forces, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, n, (const double *)S->fcd,
                     (const int *)S->gcsr_start, (const int *)S->gcsr_list, S->f);
  HIP_CHECK(hipGetLastError());
  S->have_f = true;
  return SWR_OK;
}

int swr_nve_run(swr_system *S, int stage, double dtv, double dtf) {
  const int n = S->n;
  if (n <= 0) return SWR_OK;
  if (!S->have_f) {
    swr_set_error_msg("fix nve needs the forces of the step (swr_compute_forces)");
    return SWR_ERR_STATE;
  }
  for (int t = 1; t <= S->hp.ntypes; t++)
    if (!(S->hp.type_mass[t] > 0.0)) {
      swr_set_error_msg("All masses are not set");
      return SWR_ERR_STATE;
    }
  if (stage == 0)
    hipLaunchKernelGGL(k_nve_initial, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, n, (const int *)S->type,
                       (const DevParams *)S->dp, dtv, dtf, (const double *)S->f, S->v, S->x);
  else
    hipLaunchKernelGGL(k_nve_final, dim3(div_up(n, BT)), dim3(BT), 0, S->stream, n, (const int *)S->type,
                       (const DevParams *)S->dp, dtf, (const double *)S->f, S->v);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}
