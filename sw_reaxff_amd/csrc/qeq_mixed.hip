// qeq_mixed.hip — mixed-precision QEq: H stored in FP32, inner CG in FP32,
// inside an FP64 iterative refinement (the `fix qeq/reax ... mixed` form).
//
// The reference solves H s = -chi and H t = -1 with CG_v2 in FP64
// (fix_qeq_reax_sunway.cpp:983-1167).  Here H's off-diagonal values are
// stored once in FP32 (8 B per entry with the column index instead of 12).
// For both right-hand sides together:
//
//   r = b - H x                (FP64 arithmetic on the FP32 values, k_resid)
//   stop when sqrt(u.r)/|b| <= tol/100, u = M^-1 r   (the reference's test, FP64)
//   solve H d = r in FP32 with a Jacobi-preconditioned CG in its standard
//     form (FP32 vectors, FP64 dots; no setup products, so each restart
//     costs nothing), from d = 0, to the relative tolerance that brings the
//     outer test below its target in one pass (0.9 of it), clamped to
//     [3e-4, 0.5], at most 50 iterations
//   x += d                     (FP64, k_mixed_update)
//
// In MD a solve starts from the extrapolated history, and one pass of
// ~13 inner iterations plus the two residuals brings it to the target.
//
// The outer target is a hundredth of the reference's tolerance: at tol itself
// any solver's charges carry ~1e-6..1e-5 e of truncation (measured on the
// 648-water box: FP64 CG_v2 at 1e-6 is 7.4e-6 e from the converged charges),
// and the mixed form is required to stay within 1e-6 e of the converged FP64
// charges.  At tol/10 the errors of s and t, amplified by sum(s)/sum(t) in
// q = s - (sum s/sum t) t, still reached 1.7e-6 e over MD steps.
//
// The cutoff tests that decide H's sparsity, the dot products, the residual
// norms and the charge neutrality step (calculate_Q) stay FP64.  The solve
// converges to the solution of the FP32-rounded H; on the 648-water box that
// differs from the FP64-H solution by 2.6e-7 e at most (measured against a
// dense FP64 solve).
#include "qeq_cg.h"

#include <algorithm>
#include <cmath>

namespace {

using namespace swr_cg;

constexpr int MAX_OUTER = 20;
constexpr double OUTER_TOL_FACTOR = 0.01;  // outer target = tol * this
constexpr double INNER_TOL_MIN = 3e-4;
constexpr int INNER_IMAX = 50;

// Outer state (device), read back by the host after every residual.
struct MixState {
  int done;            // both right-hand sides meet the tolerance
  int outer;           // residual evaluations so far
  int converged[2];
  int matvecs[2];      // FP64 residual SpMVs + FP32 inner iterations, per rhs
  double rho[2];       // sqrt(u.r)/|b| of the latest residual
};

// r = b - (eta x + H x) in FP64 over FP32 values; u = M^-1 r; the dots
// u.r (s, t) and r.r (s, t), b.b (s, t); r is also written as the FP32
// right-hand side of the inner solve.  Half-wave per row as in k_cg.
__global__ void __launch_bounds__(QT)
k_resid(int n, const DevParams *__restrict__ P, const int *__restrict__ type, const int *__restrict__ firstnbr,
        const int *__restrict__ numnbrs, const int *__restrict__ jlist, const float *__restrict__ val32,
        const double2 *__restrict__ x, const double2 *__restrict__ b, const double *__restrict__ Hd,
        float2 *__restrict__ r32, double *__restrict__ partials, int pstride) {
  const int w = wave_in_block(), lane = lane_id();
  const int i = blockIdx.x * RPB + w * 2 + (lane >> 5);
  const bool owner = (lane & 31) == 0 && i < n;
  const int beg = i < n ? firstnbr[i] : 0, cnt = i < n ? numnbrs[i] : 0;
  double2 q = row_spmv2_h<double, float>(beg, cnt, jlist, val32, x);
  double d[6] = {0, 0, 0, 0, 0, 0};
  if (owner) {
    const double eta = P->eta[type[i]];
    const double2 xi = x[i], bi = b[i];
    const double rs = bi.x - (eta * xi.x + q.x), rt = bi.y - (eta * xi.y + q.y);
    const double h = Hd[i];
    d[0] = h * rs * rs; d[1] = h * rt * rt;
    d[2] = rs * rs; d[3] = rt * rt;
    d[4] = bi.x * bi.x; d[5] = bi.y * bi.y;
    r32[i] = make_float2((float)rs, (float)rt);
  }
  block_partials<6, QT>(d, partials, pstride, blockIdx.x);
}

// Outer test and the inner tolerances (one workgroup).  A right-hand side
// that meets the test gets tol = -1 in the inner scalars, so its inner solve
// does nothing and its d stays 0.
__global__ void __launch_bounds__(RT)
k_resid_scalars(int nblk, const double *__restrict__ partials, int pstride, MixState *__restrict__ ms,
                CgScalars *__restrict__ inner, double tol, int first, DevStatus *st, int *__restrict__ iters) {
  double tot[6];
  reduce_partials<6>(partials, pstride, 0, nblk, tot);
  if (threadIdx.x != 0) return;
  MixState m = *ms;
  if (first) {
    m.done = 0; m.outer = 0;
    m.converged[0] = m.converged[1] = 0;
    m.matvecs[0] = m.matvecs[1] = 0;
  } else {
    for (int s = 0; s < 2; s++)
      if (!m.converged[s]) m.matvecs[s] += inner[s].iter;
  }
  m.outer += 1;
  for (int s = 0; s < 2; s++) {
    if (m.converged[s]) {
      inner[s].tol = -1.0;
      continue;
    }
    m.matvecs[s] += 1;
    const double bn = sqrt(tot[4 + s]), ur = tot[s], rn = sqrt(tot[2 + s]);
    const double rho = bn > 0 ? sqrt(fmax(ur, 0.0)) / bn : 0.0;
    m.rho[s] = rho;
    if (!(rho > tol)) {  // meets the test (a NaN does not: it is flagged below)
      m.converged[s] = isnan(rho) ? 0 : 1;
      inner[s].tol = -1.0;
      if (isnan(rho)) atomicOr(&st->flags, SWR_F_NAN);
      continue;
    }
    // the corrected x's residual is the inner one, and the inner test is
    // sqrt(u_in.r_in) <= tol_in |r| on the same preconditioned norm: it
    // meets the outer test sqrt(u.r) <= tol |b| at tol_in = tol |b| / |r|
    // (0.9 of it, for the FP32 rounding of d)
    const double tin = 0.9 * tol * bn / fmax(rn, 1e-300);
    inner[s].tol = fmin(0.5, fmax(INNER_TOL_MIN, tin));
  }
  m.done = (m.converged[0] && m.converged[1]) ? 1 : 0;
  if (!m.done && m.outer >= MAX_OUTER) {
    m.done = 1;
    atomicOr(&st->flags, SWR_F_CG_MAXIT);
  }
  if (m.done) {
    iters[0] = m.matvecs[0];
    iters[1] = m.matvecs[1];
  }
  *ms = m;
}

__global__ void k_mixed_update(int n, double2 *__restrict__ x, const float2 *__restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = d[i];
  double2 xx = x[i];
  xx.x += (double)v.x;
  xx.y += (double)v.y;
  x[i] = xx;
}

// ---------------------------------------------------------------- FP32 PCG
// The inner solve H d = r (d from 0): Jacobi-preconditioned CG in its
// standard form, FP32 vectors (s and t side by side as float2), FP64 dots
// and scalars.  Unlike the pipelined recurrence it needs no setup products
// and keeps converging at the FP64 rate over the one to two orders of
// magnitude an inner solve asks for.  Per iteration: q = H p with p.q
// (k_pcg_spmv), alpha (k_pcg_alpha), d += alpha p, r -= alpha q, z = M^-1 r
// with r.z (k_pcg_update), beta and the stopping test sqrt(r.z) <= tol |b|
// on the outer test's norm (k_pcg_beta), p = z + beta p (k_pcg_p).  A
// right-hand side that stops is frozen; launches after both stopped exit at
// once.
struct PcgState {
  double rz, alpha, beta, bnorm, tol;
  int iter, active;
};
struct PcgCtl {
  PcgState st[2];
  int done;
};

__global__ void __launch_bounds__(UT)
k_pcg_init(int n, const float2 *__restrict__ rhs, const double *__restrict__ Hd, float2 *__restrict__ d,
           float2 *__restrict__ r, float2 *__restrict__ z, float2 *__restrict__ p, double *__restrict__ partials,
           int pstride) {
  const int i = blockIdx.x * UT + threadIdx.x;
  double dd[4] = {0, 0, 0, 0};
  if (i < n) {
    const float2 b = rhs[i];
    const float h = (float)Hd[i];
    const float2 zz = make_float2(h * b.x, h * b.y);
    d[i] = make_float2(0.f, 0.f);
    r[i] = b;
    z[i] = zz;
    p[i] = zz;
    dd[0] = (double)b.x * zz.x; dd[1] = (double)b.y * zz.y;
    dd[2] = (double)b.x * b.x; dd[3] = (double)b.y * b.y;
  }
  block_partials<4, UT>(dd, partials, pstride, blockIdx.x);
}

__global__ void __launch_bounds__(RT)
k_pcg_init_scalars(int nblk, const double *__restrict__ partials, int pstride, const CgScalars *__restrict__ cgs,
                   PcgCtl *__restrict__ ctl) {
  double tot[4];
  reduce_partials<4>(partials, pstride, 0, nblk, tot);
  if (threadIdx.x != 0) return;
  for (int s = 0; s < 2; s++) {
    PcgState &c = ctl->st[s];
    c.rz = tot[s];
    c.bnorm = sqrt(tot[2 + s]);
    c.tol = cgs[s].tol;
    c.iter = 0;
    c.alpha = c.beta = 0.0;
    c.active = (c.tol > 0.0 && c.bnorm > 0.0 && sqrt(fmax(c.rz, 0.0)) > c.tol * c.bnorm) ? 1 : 0;
  }
  ctl->done = (ctl->st[0].active || ctl->st[1].active) ? 0 : 1;
}

__global__ void __launch_bounds__(QT)
k_pcg_spmv(int n, const DevParams *__restrict__ P, const int *__restrict__ type, const int *__restrict__ firstnbr,
           const int *__restrict__ numnbrs, const int *__restrict__ jlist, const float *__restrict__ val,
           const float2 *__restrict__ p, float2 *__restrict__ q, const PcgCtl *__restrict__ ctl,
           double *__restrict__ partials, int pstride) {
  const int w = wave_in_block(), lane = lane_id();
  const int i = blockIdx.x * RPB + w * 2 + (lane >> 5);
  const int beg0 = i < n ? firstnbr[i] : 0, cnt0 = i < n ? numnbrs[i] : 0;
  if (ctl->done && cnt0 >= 0) return;
  const float2 sum = row_spmv2_h<float, float>(beg0, cnt0, jlist, val, p);
  double dd[2] = {0, 0};
  if ((lane & 31) == 0 && i < n) {
    const float eta = (float)P->eta[type[i]];
    const float2 pi = p[i];
    const float2 qq = make_float2(eta * pi.x + sum.x, eta * pi.y + sum.y);
    q[i] = qq;
    dd[0] = (double)pi.x * qq.x;
    dd[1] = (double)pi.y * qq.y;
  }
  block_partials<2, QT>(dd, partials, pstride, blockIdx.x);
}

__global__ void __launch_bounds__(RT)
k_pcg_alpha(int nblk, const double *__restrict__ partials, int pstride, PcgCtl *__restrict__ ctl) {
  if (ctl->done) return;
  double tot[2];
  reduce_partials<2>(partials, pstride, 0, nblk, tot);
  if (threadIdx.x != 0) return;
  for (int s = 0; s < 2; s++) {
    PcgState &c = ctl->st[s];
    c.alpha = c.active ? c.rz / tot[s] : 0.0;
  }
}

__global__ void __launch_bounds__(UT)
k_pcg_update(int n, const double *__restrict__ Hd, float2 *__restrict__ d, float2 *__restrict__ r,
             float2 *__restrict__ z, const float2 *__restrict__ p, const float2 *__restrict__ q,
             const PcgCtl *__restrict__ ctl, double *__restrict__ partials, int pstride) {
  if (ctl->done) return;
  const int i = blockIdx.x * UT + threadIdx.x;
  const float as = (float)ctl->st[0].alpha, at = (float)ctl->st[1].alpha;
  double dd[2] = {0, 0};
  if (i < n) {
    const float2 pp = p[i], qq = q[i];
    float2 dv = d[i], rv = r[i];
    dv.x += as * pp.x; dv.y += at * pp.y;
    rv.x -= as * qq.x; rv.y -= at * qq.y;
    const float h = (float)Hd[i];
    const float2 zz = make_float2(h * rv.x, h * rv.y);
    d[i] = dv;
    r[i] = rv;
    z[i] = zz;
    dd[0] = (double)rv.x * zz.x;
    dd[1] = (double)rv.y * zz.y;
  }
  block_partials<2, UT>(dd, partials, pstride, blockIdx.x);
}

__global__ void __launch_bounds__(RT)
k_pcg_beta(int nblk, const double *__restrict__ partials, int pstride, PcgCtl *__restrict__ ctl, int imax) {
  if (ctl->done) return;
  double tot[2];
  reduce_partials<2>(partials, pstride, 0, nblk, tot);
  if (threadIdx.x != 0) return;
  for (int s = 0; s < 2; s++) {
    PcgState &c = ctl->st[s];
    if (!c.active) { c.beta = 0.0; continue; }
    c.iter += 1;
    c.beta = tot[s] / c.rz;
    c.rz = tot[s];
    c.active = (c.iter < imax && sqrt(fmax(c.rz, 0.0)) > c.tol * c.bnorm) ? 1 : 0;
  }
  ctl->done = (ctl->st[0].active || ctl->st[1].active) ? 0 : 1;
}

__global__ void __launch_bounds__(UT)
k_pcg_p(int n, const float2 *__restrict__ z, float2 *__restrict__ p, const PcgCtl *__restrict__ ctl) {
  if (ctl->done) return;
  const int i = blockIdx.x * UT + threadIdx.x;
  if (i >= n) return;
  const float bs = (float)ctl->st[0].beta, bt = (float)ctl->st[1].beta;
  const bool as = ctl->st[0].active != 0, at = ctl->st[1].active != 0;
  const float2 zz = z[i];
  float2 pp = p[i];
  if (as) pp.x = zz.x + bs * pp.x;
  if (at) pp.y = zz.y + bt * pp.y;
  p[i] = pp;
}

// Writes the iteration counts into the inner scalars the outer accounting reads.
__global__ void k_pcg_finish(const PcgCtl *__restrict__ ctl, CgScalars *__restrict__ cgs, int *__restrict__ iters) {
  if (threadIdx.x == 0) {
    cgs[0].iter = ctl->st[0].iter;
    cgs[1].iter = ctl->st[1].iter;
    iters[0] = ctl->st[0].iter;
    iters[1] = ctl->st[1].iter;
  }
}

// Inner vectors (float2) and the outer state, carved from one allocation.
enum { MV_X = 0, MV_R, MV_U, MV_W, MV_Z, MV_SS, MV_V, MV_P, MV_M, MV_D0, MV_D1, MV_B, MV_Q, MV_COUNT };

// Layout: MixState at the start, the float2 vectors from byte 256.
int ensure_mixw(swr_system *S) {
  static_assert(sizeof(MixState) <= 256, "MixState must fit its slot");
  const size_t need = 256 + sizeof(float2) * MV_COUNT * (size_t)S->ncap;
  if (S->mixw && S->mixw_bytes >= need) return SWR_OK;
  if (S->mixw) (void)hipFree(S->mixw);
  S->mixw = nullptr;
  S->mixw_bytes = 0;
  if (hipMalloc(&S->mixw, need) != hipSuccess) {
    swr_set_error_msg("hipMalloc of the mixed-precision workspace failed");
    return SWR_ERR_NOMEM;
  }
  S->mixw_bytes = need;
  return SWR_OK;
}

}  // namespace

int swr_qeq_mixed_solve(swr_system *S) {
  int rc;
  if (!S->H_val32) {
    swr_set_error_msg("mixed-precision QEq needs the FP32 H (fix qeq/reax ... mixed at setup)");
    return SWR_ERR_STATE;
  }
  if ((rc = swr_qeq_layout(S)) || (rc = ensure_mixw(S))) return rc;
  if (S->nsys != 1) {
    swr_set_error_msg("the mixed-precision solve runs one system at a time");
    return SWR_ERR_STATE;
  }
  static_assert(sizeof(MixState) <= 128 && sizeof(PcgCtl) <= 128, "mixed-precision state must fit its slots");
  const int n = S->n;
  const int nblk = div_up(n, RPB);    // half-wave-per-row kernels
  const int ublk = div_up(n, UT);     // thread-per-row kernels
  const int ps = S->cg_pstride;
  MixState *ms = (MixState *)S->mixw;
  PcgCtl *pc = (PcgCtl *)((char *)S->mixw + 128);
  float2 *base = (float2 *)((char *)S->mixw + 256);
  const size_t c = S->ncap;
  float2 *dv = base + MV_X * c, *rv = base + MV_R * c, *zv = base + MV_Z * c, *pv = base + MV_P * c,
         *qv = base + MV_Q * c;
  float2 *rhs = base + MV_B * c;  // the inner solve's b: the FP32 residual
  CgScalars *inner = S->cgs + 2;  // after the FP64 solve's s, t: the inner tolerances
  double2 *x = (double2 *)S->st;
  const double2 *b = (const double2 *)S->b_s;
  for (int outer = 0;; outer++) {
    swr_kt_begin(S, KT_MIXED_OUTER);
    hipLaunchKernelGGL(k_resid, dim3(nblk), dim3(QT), 0, S->stream, n, S->dp, S->type, S->H_firstnbr, S->H_numnbrs,
                       S->H_jlist, S->H_val32, (const double2 *)x, b, S->Hdia_inv, rhs, S->partials, ps);
    hipLaunchKernelGGL(k_resid_scalars, dim3(1), dim3(RT), 0, S->stream, nblk, S->partials, ps, ms, inner,
                       S->ctl.tol * OUTER_TOL_FACTOR, outer == 0 ? 1 : 0, S->status, S->sys_iters);
    swr_kt_end(S, KT_MIXED_OUTER);
    HIP_CHECK(hipGetLastError());
    MixState h;
    HIP_CHECK(hipMemcpyAsync(&h, ms, sizeof(h), hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
    S->mixed_outer = h.outer;
    if (outer == 0) { S->mixed_rho0[0] = h.rho[0]; S->mixed_rho0[1] = h.rho[1]; }
    if (h.done) break;
    // inner FP32 PCG from d = 0; iterations launched in batches sized from
    // the last solve, the host checking the done flag between batches
    swr_kt_begin(S, KT_MIXED_INNER);
    hipLaunchKernelGGL(k_pcg_init, dim3(ublk), dim3(UT), 0, S->stream, n, rhs, S->Hdia_inv, dv, rv, zv, pv,
                       S->partials, ps);
    hipLaunchKernelGGL(k_pcg_init_scalars, dim3(1), dim3(RT), 0, S->stream, ublk, S->partials, ps, inner, pc);
    int it = 0;
    int batch = std::max(2, S->mixed_last_inner + 1);
    while (it < INNER_IMAX) {
      for (int k = 0; k < batch && it < INNER_IMAX; k++, it++) {
        hipLaunchKernelGGL(k_pcg_spmv, dim3(nblk), dim3(QT), 0, S->stream, n, S->dp, S->type, S->H_firstnbr,
                           S->H_numnbrs, S->H_jlist, S->H_val32, (const float2 *)pv, qv, pc, S->partials, ps);
        hipLaunchKernelGGL(k_pcg_alpha, dim3(1), dim3(RT), 0, S->stream, nblk, S->partials, ps, pc);
        hipLaunchKernelGGL(k_pcg_update, dim3(ublk), dim3(UT), 0, S->stream, n, S->Hdia_inv, dv, rv, zv,
                           (const float2 *)pv, (const float2 *)qv, pc, S->partials, ps);
        hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(RT), 0, S->stream, ublk, S->partials, ps, pc, INNER_IMAX);
        hipLaunchKernelGGL(k_pcg_p, dim3(ublk), dim3(UT), 0, S->stream, n, (const float2 *)zv, pv, pc);
      }
      HIP_CHECK(hipGetLastError());
      PcgCtl hc;
      HIP_CHECK(hipMemcpyAsync(&hc, pc, sizeof(hc), hipMemcpyDeviceToHost, S->stream));
      HIP_CHECK(hipStreamSynchronize(S->stream));
      if (hc.done) {
        S->mixed_last_inner = std::max(hc.st[0].iter, hc.st[1].iter);
        break;
      }
      batch = 4;
    }
    hipLaunchKernelGGL(k_pcg_finish, dim3(1), dim3(64), 0, S->stream, pc, inner, S->status->inner_iters);
    swr_kt_end(S, KT_MIXED_INNER);
    hipLaunchKernelGGL(k_mixed_update, dim3(div_up(n, 256)), dim3(256), 0, S->stream, n, x, (const float2 *)dv);
    HIP_CHECK(hipGetLastError());
  }
  return SWR_OK;
}
