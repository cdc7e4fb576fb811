// qeq.hip — QEq: H matrix (a11-a13), SpMV (a14), the s/t pipelined PCG
// (a15) and the charge update (a16).
//
// H (compute_H_Full_C_para, fix_qeq_reax_sw64.c:768-911): full rows for the
// local atoms, every Verlet neighbour with r^2 <= swb^2,
//   val = Tap(r) * 14.4 / (r^3 + shld[ti][tj])^0.3333333333333
// with Tap by Horner (calculate_H, fix_qeq_reax_sunway.cpp:965-981).  Row i
// holds jlist/val[firstnbr[i] .. firstnbr[i]+numnbrs[i]); firstnbr[i] is the
// row's Verlet offset moved up by 4 i and rounded up to a multiple of 4
// entries (the reference uses a 600*i stride, fix_qeq_reax_sw64.c:828): rows
// start 16-byte aligned, are padded to a multiple of 4 entries (column i,
// value 0), and never overlap (a row holds at most its Verlet row's count),
// so the SpMV reads them with 16-byte loads.  Ghost columns are folded onto their owning
// local atom: on one domain the forward comm of x makes x[ghost] ==
// x[owner], so the SpMV result is unchanged and x needs only n entries.
//
// CG: CG_v2 (fix_qeq_reax_sunway.cpp:983-1167) for s and t at once; the
// kernels are templates in qeq_cg.h, instantiated here in FP64.
#include "qeq_cg.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

namespace {

using namespace swr_cg;
constexpr double EV_TO_KCAL_PER_MOL = 14.4;

// ---------------------------------------------------------------- H build
// First slot of H row i: its Verlet offset plus 4 i, rounded up to a
// multiple of 4.  Row i (at most voff[i+1] - voff[i] entries, padded to a
// multiple of 4) ends at or before row i+1 starts, since
// ceil(a/4) + ceil(c/4) <= ceil((a + c + 4)/4); the matrix needs
// h_row_start(voff[n], n) slots.
__device__ __host__ __forceinline__ long long h_row_start(long long voff_i, long long i) {
  return (voff_i + 4 * i + 3) & ~3LL;
}

// A Verlet candidate of the H build: its id, position, type and owner.
struct HCand {
  int j, t, o;
  double x, y, z;
};

__device__ __forceinline__ void h_fetch(HCand &c, int j, const double *__restrict__ x, const int *__restrict__ type,
                                        const int *__restrict__ owner) {
  c.j = j;
  c.x = x[3 * j];
  c.y = x[3 * j + 1];
  c.z = x[3 * j + 2];
  c.t = type[j];
  c.o = owner[j];
}

// One wave per local row over its Verlet candidates, 64 at a time.  The
// loop is software-pipelined as the far pass is: while chunk c is
// evaluated, chunk c+2's positions, types and owners and chunk c+6's ids
// are requested, into three candidate sets rotating with period 3 and six
// id registers rotating with period 6, so an id load has four chunks and a
// gather two to arrive (a row has ~10 chunks at 96k water atoms).
__global__ void __launch_bounds__(QT)
k_build_H(int n, const DevParams *__restrict__ P, const double *__restrict__ x, const int *__restrict__ type,
          const int *__restrict__ owner, const long long *__restrict__ voff, const int *__restrict__ vlist,
          int *__restrict__ firstnbr, int *__restrict__ numnbrs, int *__restrict__ jlist, double *__restrict__ val,
          float *__restrict__ val32, double *__restrict__ Hdia_inv, double2 *__restrict__ b2,
          double2 *__restrict__ st, const double *__restrict__ s_hist, const double *__restrict__ t_hist) {
  // the shielding table in LDS: a global load in the loop would make each
  // chunk wait for every prefetched load before it (the wait counter
  // retires loads in issue order)
  constexpr int nt1 = SWR_MAX_TYPES + 1;
  __shared__ double shl[nt1 * nt1];
  for (int k = threadIdx.x; k < nt1 * nt1; k += QT) shl[k] = P->shld[k];
  __syncthreads();
  const int i = xcd_block(gridDim.x) * QW + wave_in_block();
  if (i >= n) return;
  const int lane = lane_id();
  const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
  const int ti = type[i];
  const double swbsq = P->swbsq;
  double T[8];
#pragma unroll
  for (int k = 0; k < 8; k++) T[k] = P->Tap[k];
  const long long beg = voff[i], end = voff[i + 1];
  const long long dst = h_row_start(beg, i);
  const int nch = (int)((end - beg + SWR_WAVE - 1) / SWR_WAVE);
  const long long clast = end - 1;
  int cnt = 0;
  HCand A = {}, B = {}, C = {};
  int I0 = 0, I1 = 0, I2 = 0, I3 = 0, I4 = 0, I5 = 0;
  if (nch > 0) {
    I0 = vlist[min(beg + lane, clast)];
    I1 = vlist[min(beg + SWR_WAVE + lane, clast)];
    I2 = vlist[min(beg + 2 * SWR_WAVE + lane, clast)];
    I3 = vlist[min(beg + 3 * SWR_WAVE + lane, clast)];
    I4 = vlist[min(beg + 4 * SWR_WAVE + lane, clast)];
    I5 = vlist[min(beg + 5 * SWR_WAVE + lane, clast)];
    h_fetch(A, I0, x, type, owner);
    h_fetch(B, I1, x, type, owner);
  }
  // chunk c: evaluate cur, request chunk c+2's candidates (ids jaddr) into
  // far2 and then chunk c+6's ids into jfree (in this order: waiting for a
  // chunk's candidates then never waits for the ids requested after them)
  auto chunk = [&](int c, const HCand &cur, HCand &far2, const int &jaddr, int &jfree) {
    const long long k = beg + (long long)c * SWR_WAVE + lane;
    h_fetch(far2, jaddr, x, type, owner);
    jfree = vlist[min(k + 6 * SWR_WAVE, clast)];
    bool hit = false;
    double hv = 0.0;
    if (k < end) {
      const double dx = cur.x - xi, dy = cur.y - yi, dz = cur.z - zi;
      const double r_sqr = dx * dx + dy * dy + dz * dz;
      if (r_sqr <= swbsq) {
        hit = true;
        const double r = sqrt(r_sqr);
        const double gamma = shl[ti * nt1 + cur.t];
        double Taper = T[7] * r + T[6];
        Taper = Taper * r + T[5];
        Taper = Taper * r + T[4];
        Taper = Taper * r + T[3];
        Taper = Taper * r + T[2];
        Taper = Taper * r + T[1];
        Taper = Taper * r + T[0];
        // (r^3 + gamma)^0.3333333333333 as the reference writes it
        // (fix_qeq_reax_sunway.cpp:978), evaluated as cbrt(x) x^(a - 1/3):
        // with a - 1/3 = -3.3e-14 the second factor is 1 + (a - 1/3) ln x to
        // far below FP64 rounding, so a float log suffices.  Same value as
        // pow to ~1 ulp, with a third of pow's registers.
        const double denom = r * r * r + gamma;
        const double root = cbrt(denom) * (1.0 + (0.3333333333333 - 1.0 / 3.0) * (double)__logf((float)denom));
        hv = Taper * EV_TO_KCAL_PER_MOL / root;
      }
    }
    const unsigned long long m = __ballot(hit);
    if (hit) {
      const long long slot = dst + cnt + ballot_rank(m);
      jlist[slot] = cur.o;
      if (val) val[slot] = hv;
      if (val32) val32[slot] = (float)hv;
    }
    cnt += __popcll(m);
  };
  // six chunks per trip (the period of the id ring); a loop with one exit,
  // then the last nch % 6 chunks in straight-line code continuing the
  // rotation (early exits from the loop make the compiler's wait counts at
  // its header conservative)
  int c = 0;
  for (; c + 6 <= nch; c += 6) {
    chunk(c, A, C, I2, I0);
    chunk(c + 1, B, A, I3, I1);
    chunk(c + 2, C, B, I4, I2);
    chunk(c + 3, A, C, I5, I3);
    chunk(c + 4, B, A, I0, I4);
    chunk(c + 5, C, B, I1, I5);
  }
  if (c < nch) chunk(c, A, C, I2, I0);
  if (c + 1 < nch) chunk(c + 1, B, A, I3, I1);
  if (c + 2 < nch) chunk(c + 2, C, B, I4, I2);
  if (c + 3 < nch) chunk(c + 3, A, C, I5, I3);
  if (c + 4 < nch) chunk(c + 4, B, A, I0, I4);
  // padding to a multiple of 4 entries: column i with value 0
  if (lane < ((cnt + 3) & ~3) - cnt) {
    const long long slot = dst + cnt + lane;
    jlist[slot] = i;
    if (val) val[slot] = 0.0;
    if (val32) val32[slot] = 0.0f;
  }
  if (lane == 0) {
    firstnbr[i] = (int)dst;
    numnbrs[i] = cnt;
    // init_matvec vectors (fix_qeq_reax_sunway.cpp:700-716)
    Hdia_inv[i] = 1. / P->eta[ti];
    b2[i] = make_double2(-P->chi[ti], -1.0);
    const double *sh = s_hist + 5 * (size_t)i, *th = t_hist + 5 * (size_t)i;
    const double tv = th[2] + 3 * (th[0] - th[1]);
    const double sv = 4 * (sh[0] + sh[2]) - (6 * sh[1] + sh[3]);
    st[i] = make_double2(sv, tv);
  }
}

// ---------------------------------------------------------------- SpMV
// One wave per row, FP64 accumulation; VT is the stored value type (FP32
// when the system runs mixed-precision QEq).
template <typename VT>
__global__ void __launch_bounds__(QT)
k_spmv1(int n, const DevParams *__restrict__ P, const int *__restrict__ type, const int *__restrict__ firstnbr,
        const int *__restrict__ numnbrs, const int *__restrict__ jlist, const VT *__restrict__ val,
        const double *__restrict__ xv, double *__restrict__ b) {
  const int i = xcd_block(gridDim.x) * QW + wave_in_block();
  if (i >= n) return;
  const int lane = lane_id();
  const int beg = firstnbr[i], end = beg + numnbrs[i];
  double a0 = 0.0, a1 = 0.0;
  int k = beg + lane;
  for (; k + SWR_WAVE < end; k += 2 * SWR_WAVE) {
    const int j0 = jlist[k], j1 = jlist[k + SWR_WAVE];
    const double v0 = (double)val[k], v1 = (double)val[k + SWR_WAVE];
    a0 += v0 * xv[j0];
    a1 += v1 * xv[j1];
  }
  if (k < end) a0 += (double)val[k] * xv[jlist[k]];
  const double a = wave_sum(a0 + a1);
  if (lane == 0) b[i] = P->eta[type[i]] * xv[i] + a;
}

// ---------------------------------------------------------------- calculate_Q
// Tile partial sums of s and t over the local atoms (calculate_Q,
// fix_qeq_reax_sunway.cpp:1733-1734: the two parallel_vector_acc calls).
__global__ void __launch_bounds__(UT)
k_q_sums(const double2 *__restrict__ st, CgLayout L, double *__restrict__ partials, int pstride) {
  const int tile = blockIdx.x;
  const int i = L.tile_beg[tile] + threadIdx.x;
  double d[2] = {0.0, 0.0};
  if (i < L.tile_beg[tile + 1]) { const double2 v = st[i]; d[0] = v.x; d[1] = v.y; }
  block_partials<2, UT>(d, partials, pstride, tile);
}

// Per system: sum s, sum t and u = sum s / sum t (one workgroup each).
__global__ void __launch_bounds__(RT)
k_q_total(const double *__restrict__ partials, int pstride, CgLayout L, double *__restrict__ qsum) {
  const int sys = blockIdx.x;
  double tot[2];
  reduce_partials<2>(partials, pstride, L.sys_tile[sys], L.sys_tile[sys + 1], tot);
  if (threadIdx.x != 0) return;
  qsum[3 * sys] = tot[0];
  qsum[3 * sys + 1] = tot[1];
  qsum[3 * sys + 2] = tot[0] / tot[1];
}

// q = s - u t for local atoms (u of the atom's system), history shift, and
// the ghost q refresh (calculate_Q :1736-1754).
__global__ void k_q_apply(int n, int N, const double2 *__restrict__ st, const double *__restrict__ qsum,
                          const int *__restrict__ owner, const int *__restrict__ row_sys, double *__restrict__ q,
                          double *__restrict__ s_hist, double *__restrict__ t_hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int o = i < n ? i : owner[i];
  const double u = qsum[3 * row_sys[o] + 2];
  const double2 v = st[o];
  q[i] = v.x - u * v.y;
  if (i < n) {
    double *sh = s_hist + 5 * (size_t)i, *th = t_hist + 5 * (size_t)i;
    for (int k = 4; k > 0; --k) { sh[k] = sh[k - 1]; th[k] = th[k - 1]; }
    sh[0] = v.x;
    th[0] = v.y;
  }
}

CgVecsT<double> make_vecs(swr_system *S) {
  CgVecsT<double> V;
  double2 *base = (double2 *)S->cgv;
  const size_t c = S->ncap;
  V.r = base + CGV_R * c; V.u = base + CGV_U * c; V.w = base + CGV_W * c; V.z = base + CGV_Z * c;
  V.ss = base + CGV_SS * c; V.v = base + CGV_V * c; V.p = base + CGV_P * c; V.m = base + CGV_M * c;
  V.d[0] = base + CGV_D0 * c; V.d[1] = base + CGV_D1 * c; V.q = base + CGV_Q * c;
  V.x = (double2 *)S->st;
  V.b = (const double2 *)S->b_s;
  V.Hd = S->Hdia_inv;
  return V;
}

}  // namespace

// ---------------------------------------------------------------- host side
int swr_qeq_build_H(swr_system *S) {
  const int n = S->n;
  swr_kt_begin(S, KT_BUILD_H);
  hipLaunchKernelGGL(k_build_H, dim3(div_up(n, QW)), dim3(QT), 0, S->stream, n, S->dp, S->x, S->type, S->owner,
                     S->voff, S->vlist, S->H_firstnbr, S->H_numnbrs, S->H_jlist,
                     S->H_val, S->mixed ? S->H_val32 : (float *)nullptr, S->Hdia_inv, (double2 *)S->b_s, (double2 *)S->st,
                     S->s_hist, S->t_hist);
  swr_kt_end(S, KT_BUILD_H);
  HIP_CHECK(hipGetLastError());
  return swr_panel_build(S);
}

int swr_qeq_spmv(swr_system *S, const double *xv, double *b) {
  const int n = S->n;
  swr_kt_begin(S, KT_SPMV);
  if (S->mixed)
    hipLaunchKernelGGL(k_spmv1<float>, dim3(div_up(n, QW)), dim3(QT), 0, S->stream, n, S->dp, S->type,
                       S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val32, xv, b);
  else
    hipLaunchKernelGGL(k_spmv1<double>, dim3(div_up(n, QW)), dim3(QT), 0, S->stream, n, S->dp, S->type,
                       S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val, xv, b);
  swr_kt_end(S, KT_SPMV);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Row/tile layout of the (batch of) systems and the buffers sized by it.
// Rebuilt when the atoms or the batch partition changed.
int swr_pin_iters_reserve(swr_system *S, int nsys) {
  const int need = 2 * std::max(nsys, 1);
  if (need <= S->pin_iters_cap) return SWR_OK;
  const int cap = std::max(need, 2 * S->pin_iters_cap);
  if (S->pin_iters) {
    HIP_CHECK(hipStreamSynchronize(S->stream));
    HIP_CHECK(hipHostFree(S->pin_iters));
    S->pin_iters = nullptr;
    S->pin_iters_cap = 0;
  }
  if (hipHostMalloc((void **)&S->pin_iters, sizeof(int) * (size_t)cap, 0) != hipSuccess) {
    swr_set_error_msg("hipHostMalloc failed");
    return SWR_ERR_NOMEM;
  }
  S->pin_iters_cap = cap;
  return SWR_OK;
}

int swr_qeq_layout(swr_system *S) {
  if (!S->layout_dirty) return SWR_OK;
  const int n = S->n;
  if (S->sys_row.size() < 2 || S->sys_row.back() != n) S->sys_row = {0, n};
  const int nsys = (int)S->sys_row.size() - 1;
  if (int rc = swr_pin_iters_reserve(S, nsys)) return rc;
  std::vector<int> row_sys(std::max(n, 1)), tile_beg, tile_sys, sys_tile(nsys + 1), sblk_beg, sys_sblk(nsys + 1);
  for (int s = 0; s < nsys; s++) {
    sys_tile[s] = (int)tile_sys.size();
    for (int r = S->sys_row[s]; r < S->sys_row[s + 1]; r += UT) {
      tile_beg.push_back(r);
      tile_sys.push_back(s);
    }
    sys_sblk[s] = (int)sblk_beg.size();
    for (int r = S->sys_row[s]; r < S->sys_row[s + 1]; r += RPB) sblk_beg.push_back(r);
    for (int r = S->sys_row[s]; r < S->sys_row[s + 1]; r++) row_sys[r] = s;
  }
  sys_tile[nsys] = (int)tile_sys.size();
  tile_beg.push_back(n);
  const int nsblk = (int)sblk_beg.size();
  sys_sblk[nsys] = nsblk;
  sblk_beg.push_back(n);
  if (tile_sys.empty()) {  // no rows: one empty tile keeps every grid non-empty
    tile_beg = {0, 0};
    tile_sys = {0};
    sys_tile.assign(nsys + 1, 1);
    sys_tile[0] = 0;
  }
  const int ntiles = (int)tile_sys.size();
  int rc;
  if ((rc = swr_reserve(&S->row_sys, S->cap_row_sys, row_sys.size())) ||
      (rc = swr_reserve(&S->tile_beg, S->cap_tile_beg, tile_beg.size())) ||
      (rc = swr_reserve(&S->tile_sys, S->cap_tile_sys, tile_sys.size())) ||
      (rc = swr_reserve(&S->sys_tile, S->cap_sys_tile, sys_tile.size())) ||
      (rc = swr_reserve(&S->sblk_beg, S->cap_sblk_beg, sblk_beg.size())) ||
      (rc = swr_reserve(&S->sys_sblk, S->cap_sys_sblk, sys_sblk.size())) ||
      (rc = swr_reserve(&S->sys_iters, S->cap_sys_iters, 2 * (size_t)nsys)) ||
      (rc = swr_reserve(&S->cgs, S->cap_cgs, 4 * (size_t)nsys)) ||
      (rc = swr_reserve(&S->qsum, S->cap_qsum, 3 * (size_t)nsys)))
    return rc;
  HIP_CHECK(hipMemcpyAsync(S->row_sys, row_sys.data(), sizeof(int) * row_sys.size(), hipMemcpyHostToDevice, S->stream));
  HIP_CHECK(hipMemcpyAsync(S->tile_beg, tile_beg.data(), sizeof(int) * tile_beg.size(), hipMemcpyHostToDevice,
                           S->stream));
  HIP_CHECK(hipMemcpyAsync(S->tile_sys, tile_sys.data(), sizeof(int) * tile_sys.size(), hipMemcpyHostToDevice,
                           S->stream));
  HIP_CHECK(hipMemcpyAsync(S->sys_tile, sys_tile.data(), sizeof(int) * sys_tile.size(), hipMemcpyHostToDevice,
                           S->stream));
  HIP_CHECK(hipMemcpyAsync(S->sblk_beg, sblk_beg.data(), sizeof(int) * sblk_beg.size(), hipMemcpyHostToDevice,
                           S->stream));
  HIP_CHECK(hipMemcpyAsync(S->sys_sblk, sys_sblk.data(), sizeof(int) * sys_sblk.size(), hipMemcpyHostToDevice,
                           S->stream));
  HIP_CHECK(hipMemsetAsync(S->sys_iters, 0, sizeof(int) * 2 * nsys, S->stream));
  HIP_CHECK(hipMemsetAsync(S->cgs, 0, sizeof(CgScalars) * 4 * nsys, S->stream));
  const int ps = std::max(std::max(div_up(std::max(n, 1), RPB), nsblk), ntiles);
  if (ps > S->cg_pstride || !S->partials) {
    if ((rc = swr_alloc(&S->partials, (size_t)ps * 8))) return rc;
    S->cg_pstride = ps;
  }
  if (!S->done_count && (rc = swr_alloc(&S->done_count, 1))) return rc;
  HIP_CHECK(hipStreamSynchronize(S->stream));
  S->nsys = nsys;
  S->ntiles = ntiles;
  S->nsblk = nsblk;
  S->h_sys_iters.assign(2 * nsys, 0);
  S->layout_dirty = false;
  S->layout_gen++;
  return SWR_OK;
}

CgLayout swr_qeq_cg_layout(swr_system *S) {
  CgLayout L;
  L.row_sys = S->row_sys;
  L.tile_beg = S->tile_beg;
  L.tile_sys = S->tile_sys;
  L.sys_tile = S->sys_tile;
  L.nsys = S->nsys;
  L.ntiles = S->ntiles;
  L.done_count = S->done_count;
  return L;
}

// Autotuner hook: the average time of `reps` launches of the FP64 CG's SpMV
// in shape sh on the current H and vectors (the setup-phase instance, which
// multiplies d0 into q and always runs; both are scratch between solves).
int swr_qeq_time_spmv(swr_system *S, int sh, int reps, float *ms) {
  int rc = swr_qeq_layout(S);
  if (rc) return rc;
  if (!S->H_val || S->n <= 0) return SWR_ERR_STATE;
  CgVecsT<double> V = make_vecs(S);
  const int n = S->n;
  const CgScalars *cgs = S->cgs;
  const int *row_sys = S->row_sys;
  hipEvent_t e[2];
  HIP_CHECK(hipEventCreate(&e[0]));
  HIP_CHECK(hipEventCreate(&e[1]));
  spmv_dispatch(sh, [&](auto shc) {
    constexpr int SH = decltype(shc)::value;
    const dim3 gb(S->npanel), tb(SpmvShape<SH>::NT);
    for (int r = -1; r < reps; r++) {   // one untimed launch first
      if (r == 0) (void)hipEventRecord(e[0], S->stream);
      if (S->nsys > 1)
        hipLaunchKernelGGL((k_cg_spmv<PH_S2, true, double, double, SH>), gb, tb, 0, S->stream, n, S->dp, S->type,
                           S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val, V, 0, cgs, row_sys, panel_args(S));
      else
        hipLaunchKernelGGL((k_cg_spmv<PH_S2, false, double, double, SH>), gb, tb, 0, S->stream, n, S->dp, S->type,
                           S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val, V, 0, cgs, row_sys, panel_args(S));
    }
    (void)hipEventRecord(e[1], S->stream);
  });
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventSynchronize(e[1]));
  float t = 0.f;
  HIP_CHECK(hipEventElapsedTime(&t, e[0], e[1]));
  *ms = t / std::max(reps, 1);
  (void)hipEventDestroy(e[0]);
  (void)hipEventDestroy(e[1]);
  return SWR_OK;
}

// Runs the FP64 solves of every system to convergence (first batch of
// first_batch iteration launches, then batches of 4 between host checks).
int swr_qeq_cg_run(swr_system *S, int first_batch) {
  int rc = swr_qeq_layout(S);
  if (rc) return rc;
  CgRun<double, double> R{};
  R.V = make_vecs(S);
  R.val = S->H_val;
  R.L = swr_qeq_cg_layout(S);
  R.cgs = S->cgs;
  R.iters_out = S->sys_iters;
  S->h_sys_iters.resize(2 * (size_t)S->nsys);
  R.iters_host = S->h_sys_iters.data();
  R.pin_done = S->pin_done;
  R.pin_iters = S->pin_iters;
  R.tol = S->ctl.tol;
  R.imax = S->imax;
  R.kt_iter = KT_CG_ITER;
  R.kt_update = KT_CG_UPDATE;
  R.kt_setup = KT_CG_SETUP;
  R.kt_scalars = KT_CG_SCALARS;
  R.partials = S->partials;
  R.pstride = S->cg_pstride;
  return cg_solve(S, R, first_batch, 0);
}

int swr_qeq_calc_q(swr_system *S) {
  int rc = swr_qeq_layout(S);
  if (rc) return rc;
  const int n = S->n;
  const int ps = S->cg_pstride;
  const CgLayout L = swr_qeq_cg_layout(S);
  swr_kt_begin(S, KT_QEQ_Q);
  hipLaunchKernelGGL(k_q_sums, dim3(L.ntiles), dim3(UT), 0, S->stream, (const double2 *)S->st, L, S->partials, ps);
  hipLaunchKernelGGL(k_q_total, dim3(L.nsys), dim3(RT), 0, S->stream, S->partials, ps, L, S->qsum);
  hipLaunchKernelGGL(k_q_apply, dim3(div_up(S->N, 256)), dim3(256), 0, S->stream, n, S->N, (const double2 *)S->st,
                     S->qsum, S->owner, S->row_sys, S->q, S->s_hist, S->t_hist);
  swr_kt_end(S, KT_QEQ_Q);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// ---------------------------------------------------------------- warm-start store
struct swr_hist_store {
  double *dev = nullptr;   // [2][cap][5]: s rows, then t rows
  int cap = 0;
};

namespace {
// slots 1..4 of each history row = slot 0 (rows of s and t alike)
__global__ void k_hist_fill(int rows, double *__restrict__ h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  double *r = h + 5 * (size_t)i;
  const double v = r[0];
  r[1] = v; r[2] = v; r[3] = v; r[4] = v;
}
}  // namespace

extern "C" int swr_hist_save(swr_system *S, int row0, int n, int fill, swr_hist_store **store) {
  if (!S || !store || row0 < 0 || n < 0 || row0 + n > S->n) {
    swr_set_error_msg("swr_hist_save: rows out of range");
    return SWR_ERR_ARG;
  }
  swr_hist_store *h = *store;
  if (!h) *store = h = new swr_hist_store();
  if (h->cap < n) {
    if (h->dev) HIP_CHECK(hipFree(h->dev));
    h->dev = nullptr;
    h->cap = 0;
    if (hipMalloc(&h->dev, sizeof(double) * 10 * (size_t)n) != hipSuccess) {
      swr_set_error_msg("swr_hist_save: hipMalloc failed");
      return SWR_ERR_NOMEM;
    }
    h->cap = n;
  }
  if (n == 0) return SWR_OK;
  const size_t bytes = sizeof(double) * 5 * (size_t)n;
  HIP_CHECK(hipMemcpyAsync(h->dev, S->s_hist + 5 * (size_t)row0, bytes, hipMemcpyDeviceToDevice, S->stream));
  HIP_CHECK(hipMemcpyAsync(h->dev + 5 * (size_t)h->cap, S->t_hist + 5 * (size_t)row0, bytes,
                           hipMemcpyDeviceToDevice, S->stream));
  if (fill) {
    hipLaunchKernelGGL(k_hist_fill, dim3(div_up(n, 256)), dim3(256), 0, S->stream, n, h->dev);
    hipLaunchKernelGGL(k_hist_fill, dim3(div_up(n, 256)), dim3(256), 0, S->stream, n, h->dev + 5 * (size_t)h->cap);
    HIP_CHECK(hipGetLastError());
  }
  return SWR_OK;
}

extern "C" int swr_hist_load(swr_system *S, int row0, int n, const swr_hist_store *h) {
  if (!S || !h || row0 < 0 || n < 0 || row0 + n > S->n || n > h->cap) {
    swr_set_error_msg("swr_hist_load: rows out of range or store too small");
    return SWR_ERR_ARG;
  }
  if (n == 0) return SWR_OK;
  const size_t bytes = sizeof(double) * 5 * (size_t)n;
  HIP_CHECK(hipMemcpyAsync(S->s_hist + 5 * (size_t)row0, h->dev, bytes, hipMemcpyDeviceToDevice, S->stream));
  HIP_CHECK(hipMemcpyAsync(S->t_hist + 5 * (size_t)row0, h->dev + 5 * (size_t)h->cap, bytes,
                           hipMemcpyDeviceToDevice, S->stream));
  return SWR_OK;
}

extern "C" int swr_hist_free(swr_hist_store *h) {
  if (!h) return SWR_OK;
  if (h->dev) (void)hipFree(h->dev);
  delete h;
  return SWR_OK;
}

// ---------------------------------------------------------------- no-allocation CG
// swr_qeq_cg_noalloc runs the FP64 solve above for callers that cannot
// allocate device memory (code under graph capture, tight serving loops):
// the caller passes the right-hand sides and the solutions, and everything
// else the kernels need beyond sys's H -- the eleven CG vectors, the
// per-workgroup partial dots, the CG scalars, the iteration counts and the
// one-system tile layout -- lives in `scratch`, a file-scope device array of
// NA_BYTES per GPU, held until the process ends.  Its device address is
// taken once per device at set-up (swr_noalloc_init, called by swr_create)
// with hipGetSymbolAddress; host code uses only that address.  Every call
// on one GPU shares its array: calls that overlap on one GPU race.
namespace {

constexpr int NA_MAX_ROWS = 32768;                       // rows (local atoms) at most
constexpr int NA_TILES = NA_MAX_ROWS / UT + 1;           // vector-kernel tiles at most
constexpr int NA_PSTRIDE = NA_MAX_ROWS / RPB + NA_TILES; // >= max(SpMV blocks, tiles)
constexpr size_t NA_VEC = sizeof(double2) * NA_MAX_ROWS;
constexpr size_t NA_OFF_PART = (size_t)CGV_COUNT * NA_VEC;
constexpr size_t NA_OFF_CGS = NA_OFF_PART + sizeof(double) * 8 * NA_PSTRIDE;
constexpr size_t NA_OFF_ROWSYS = (NA_OFF_CGS + sizeof(CgScalars) * 4 + 255) & ~(size_t)255;
constexpr size_t NA_OFF_TBEG = NA_OFF_ROWSYS + sizeof(int) * NA_MAX_ROWS;
constexpr size_t NA_OFF_TSYS = NA_OFF_TBEG + sizeof(int) * (NA_TILES + 1);
constexpr size_t NA_OFF_STILE = NA_OFF_TSYS + sizeof(int) * NA_TILES;
constexpr size_t NA_OFF_ITERS = NA_OFF_STILE + sizeof(int) * 2;
constexpr size_t NA_OFF_DONE = NA_OFF_ITERS + sizeof(int) * 2;
constexpr size_t NA_BYTES = (NA_OFF_DONE + sizeof(int) + 255) & ~(size_t)255;

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu;
unsigned char *na_ptr[NA_MAX_DEVICES] = {};

// One-system layout of the vector kernels in the scratch: tiles of UT rows,
// all rows of system 0; the CG scalars start at zero.
__global__ void k_na_layout(int n, int ntiles, int *__restrict__ row_sys, int *__restrict__ tile_beg,
                            int *__restrict__ tile_sys, int *__restrict__ sys_tile, CgScalars *__restrict__ cgs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) row_sys[k] = 0;
  if (k <= ntiles) tile_beg[k] = k < ntiles ? k * UT : n;
  if (k < ntiles) tile_sys[k] = 0;
  if (k == 0) { sys_tile[0] = 0; sys_tile[1] = ntiles; }
  if (k < 4) cgs[k] = CgScalars{};
}

}  // namespace

// The scratch array itself; host code never names it after swr_noalloc_init.
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

int swr_noalloc_init(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) {
    swr_set_error_msg("device index beyond the no-allocation entry point's table");
    return SWR_ERR_ARG;
  }
  std::lock_guard<std::mutex> g(na_mu);
  if (na_ptr[device]) return SWR_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    swr_set_error_msg(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return SWR_ERR_HIP;
  }
  na_ptr[device] = (unsigned char *)p;
  return SWR_OK;
}

int swr_qeq_cg_noalloc_run(swr_system *S, const double *b, double *x, int *iters_s, int *iters_t) {
  const int n = S->n;
  if (n <= 0 || !S->H_val || !S->Hdia_inv) {
    swr_set_error_msg("swr_qeq_cg_noalloc: no H (call swr_qeq_init_matvec first)");
    return SWR_ERR_STATE;
  }
  if (n > NA_MAX_ROWS) {
    swr_set_error_msg("swr_qeq_cg_noalloc: " + std::to_string(n) + " rows, at most " +
                      std::to_string(NA_MAX_ROWS));
    return SWR_ERR_ARG;
  }
  if (S->nsys > 1 || S->mixed) {
    swr_set_error_msg("swr_qeq_cg_noalloc: one system, FP64 QEq only");
    return SWR_ERR_ARG;
  }
  if (!b || !x) return SWR_ERR_ARG;
  unsigned char *base;
  {
    std::lock_guard<std::mutex> g(na_mu);
    base = (S->device >= 0 && S->device < NA_MAX_DEVICES) ? na_ptr[S->device] : nullptr;
  }
  if (!base) {
    swr_set_error_msg("swr_qeq_cg_noalloc: scratch address of this device not set up");
    return SWR_ERR_STATE;
  }
  const int ntiles = div_up(n, UT);
  double2 *vb = (double2 *)base;
  CgVecsT<double> V;
  V.r = vb + CGV_R * (size_t)NA_MAX_ROWS; V.u = vb + CGV_U * (size_t)NA_MAX_ROWS;
  V.w = vb + CGV_W * (size_t)NA_MAX_ROWS; V.z = vb + CGV_Z * (size_t)NA_MAX_ROWS;
  V.ss = vb + CGV_SS * (size_t)NA_MAX_ROWS; V.v = vb + CGV_V * (size_t)NA_MAX_ROWS;
  V.p = vb + CGV_P * (size_t)NA_MAX_ROWS; V.m = vb + CGV_M * (size_t)NA_MAX_ROWS;
  V.d[0] = vb + CGV_D0 * (size_t)NA_MAX_ROWS; V.d[1] = vb + CGV_D1 * (size_t)NA_MAX_ROWS;
  V.q = vb + CGV_Q * (size_t)NA_MAX_ROWS;
  V.x = (double2 *)x;
  V.b = (const double2 *)b;
  V.Hd = S->Hdia_inv;
  CgLayout L;
  L.row_sys = (int *)(base + NA_OFF_ROWSYS);
  L.tile_beg = (int *)(base + NA_OFF_TBEG);
  L.tile_sys = (int *)(base + NA_OFF_TSYS);
  L.sys_tile = (int *)(base + NA_OFF_STILE);
  L.nsys = 1;
  L.ntiles = ntiles;
  L.done_count = (int *)(base + NA_OFF_DONE);
  CgScalars *cgs = (CgScalars *)(base + NA_OFF_CGS);
  hipLaunchKernelGGL(k_na_layout, dim3(div_up(std::max(n, ntiles + 1), 256)), dim3(256), 0, S->stream, n, ntiles,
                     (int *)L.row_sys, (int *)L.tile_beg, (int *)L.tile_sys, (int *)L.sys_tile, cgs);
  CgRun<double, double> R{};
  R.V = V;
  R.val = S->H_val;
  R.L = L;
  R.cgs = cgs;
  R.iters_out = (int *)(base + NA_OFF_ITERS);
  int its[2] = {0, 0};
  R.iters_host = its;
  R.tol = S->ctl.tol;
  R.imax = S->imax;
  R.kt_iter = KT_CG_ITER;
  R.kt_update = KT_CG_UPDATE;
  R.kt_setup = KT_CG_SETUP;
  R.kt_scalars = KT_CG_SCALARS;
  R.partials = (double *)(base + NA_OFF_PART);
  R.pstride = NA_PSTRIDE;
  int its0 = 0;
  for (size_t k = 0; k < S->h_sys_iters.size(); k += 2) its0 = std::max(its0, S->h_sys_iters[k]);
  const bool ktime = S->ktime;
  S->ktime = false;
  const int rc = cg_solve(S, R, its0 + 2, 0);
  S->ktime = ktime;
  if (rc) return rc;
  if (iters_s) *iters_s = its[0];
  if (iters_t) *iters_t = its[1];
  return swr_status_to_rc(S->status_host->flags);
}
