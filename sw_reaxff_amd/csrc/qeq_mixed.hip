// qeq_mixed.hip — mixed-precision QEq: an FP32 copy of H and an FP32 inner
// CG inside an FP64 iterative refinement (the `fix qeq/reax ... mixed` form).
//
// The reference solves H s = -chi and H t = -1 with CG_v2 in FP64
// (fix_qeq_reax_sunway.cpp:983-1167).  Here H's off-diagonal values are also
// stored in FP32 (8 B per entry with the column index instead of 12), and
// the inner solves run on that copy.  For both right-hand sides together:
//
//   r = b - H x                (FP64 values and arithmetic, k_resid)
//   stop when sqrt(u.r)/|b| <= tol/1000, u = M^-1 r  (the reference's test, FP64)
//   solve H32 d = r in FP32 with a Jacobi-preconditioned CG in its standard
//     form (FP32 vectors, FP64 dots; no setup products, so each restart
//     costs nothing), from d = 0, to the relative tolerance that brings the
//     outer test below its target in one pass (0.9 of it), clamped to
//     [3e-4, 0.5], at most 50 iterations
//   x += d                     (FP64, k_mixed_update)
//
// Because the residual uses the FP64 values, the refinement converges to
// the solution of the FP64 H, whatever the rounding of the FP32 copy (with
// the FP32 values in the residual it converged to the FP32 matrix's
// solution instead, 2.6e-7 e from the FP64 one on the 648-water box).
//
// The outer target is a thousandth of the reference's tolerance: at tol itself
// any solver's charges carry ~1e-6..1e-5 e of truncation (measured on the
// 648-water box: FP64 CG_v2 at 1e-6 is 7.4e-6 e from the converged charges),
// and over an MD run the warm start carries each solve's truncation into the
// next (FP64 at tol 1e-6 drifts 9e-4 e from the converged charges over 200
// steps of the 32,000-water box), while the mixed form is required to stay
// within 1e-6 e of them.  At tol/100 it reached 6.3e-6 e (32,000 waters)
// and 2.1e-5 e (256,000 waters) over 200 steps; at tol/1000 3.1e-7 e and
// 5.8e-7 e.
//
// The cutoff tests that decide H's sparsity, the dot products, the residual
// norms and the charge neutrality step (calculate_Q) stay FP64.
#include "qeq_cg.h"

#include <algorithm>
#include <cmath>
#include <cstddef>

namespace {

using namespace swr_cg;

constexpr int MAX_OUTER = 20;
constexpr double OUTER_TOL_FACTOR = 0.001;  // outer target = tol * this
constexpr double INNER_TOL_MIN = 3e-4;
constexpr int INNER_IMAX = 50;
// A pass after one whose residual norms were all at most this (relative to
// |b|) forms its residual incrementally, r -= eta d + H32 d, from the FP32
// values and the last inner correction d: its rounding, ~1e-6 |H d| with
// |H d| ~ |r| <= INC_RHO |b|, stays two orders below the outer target.
constexpr double INC_RHO = 1e-4;
// Threads of the per-system scalar kernels in a batch: most systems of a
// batch are small, and a 256-thread workgroup finds room on a busy CU far
// sooner than a 1024-thread one (one system alone uses RT).
constexpr int BT = 256;

// Outer state of one system (device).
struct MixState {
  int done;            // both right-hand sides meet the tolerance
  int outer;           // residual evaluations so far
  int converged[2];
  int matvecs[2];      // FP64 residual SpMVs + FP32 inner iterations, per rhs
  double rho[2];       // sqrt(u.r)/|b| of the latest residual
  int inc;             // the next residual may be formed incrementally
};

// Inner FP32 PCG state of one right-hand side of one system.
struct PcgState {
  double rz, alpha, beta, bnorm, tol;
  int iter, active;
};

// Counters over the batch of systems, read back by the host between launch
// batches (with system 0's MixState right after them).
struct MixCtl {
  int outer_ndone;     // systems whose refinement has finished
  int inner_ndone;     // systems whose inner s and t solves have both finished
  int inner_itmax;     // most inner iterations of a system in the current inner solve
  int pad;
};

// Workspace layout: MixCtl at 0, MixState[nsys] from byte 64, two buffers of
// PcgState[2 nsys] (the inner state alternates between them) and the float2
// vectors each from the next 256-byte boundary.
constexpr size_t MS_OFF = 64;
inline size_t pcg_off(int nsys) { return (MS_OFF + sizeof(MixState) * (size_t)nsys + 255) & ~(size_t)255; }
inline size_t vec_off(int nsys) { return (pcg_off(nsys) + sizeof(PcgState) * 4 * (size_t)nsys + 255) & ~(size_t)255; }

__device__ __forceinline__ bool pcg_live(const PcgState *__restrict__ c) { return (c[0].active | c[1].active) != 0; }

// r = b - (eta x + H x) in FP64 (FP64 values); u = M^-1 r; the dots
// u.r (s, t) and r.r (s, t), b.b (s, t); r is kept in FP64 (r64) and also
// written as the FP32 right-hand side of the inner solve.  For a system the
// previous pass flagged (ms.inc: its residual was already small), r is
// formed incrementally instead, r64 -= eta d + H d on the FP32 values and
// the last inner correction d (whose x += d has run), FP64 from there on.
// One workgroup per SpMV panel (its partial dots in slot p).  Panels of
// systems whose refinement has finished are skipped (not on the first pass,
// whose state is the previous solve's).
template <bool BATCH, int SH>
__global__ void __launch_bounds__(SpmvShape<SH>::NT)
k_resid(int n, const DevParams *__restrict__ P, const int *__restrict__ type, const int *__restrict__ firstnbr,
        const int *__restrict__ numnbrs, const int *__restrict__ jlist, const double *__restrict__ val,
        const float *__restrict__ val32, const double2 *__restrict__ x, const float2 *__restrict__ dinc,
        const double2 *__restrict__ b, const double *__restrict__ Hd, double2 *__restrict__ r64,
        float2 *__restrict__ r32, double *__restrict__ partials, int pstride, PanelArgs A,
        const int *__restrict__ row_sys, const MixState *__restrict__ ms, int first) {
  constexpr int NT = SpmvShape<SH>::NT, G = SpmvShape<SH>::G, U = SpmvShape<SH>::U;
  __shared__ double2 xl[SWR_PANEL_MAXC];
  const int p = blockIdx.x;
  int sys = 0;
  if (BATCH) {
    const int r0 = A.pbeg[p];
    sys = r0 < A.pbeg[p + 1] ? row_sys[A.prow[r0]] : 0;
    if (!first && ms[sys].done) return;
  }
  double d[6] = {0, 0, 0, 0, 0, 0};
  auto dots = [&](int i, double rs, double rt) {
    const double2 bi = b[i];
    const double h = Hd[i];
    d[0] += h * rs * rs; d[1] += h * rt * rt;
    d[2] += rs * rs; d[3] += rt * rt;
    d[4] += bi.x * bi.x; d[5] += bi.y * bi.y;
    r64[i] = make_double2(rs, rt);
    r32[i] = make_float2((float)rs, (float)rt);
  };
  if (!first && ms[sys].inc) {
    panel_rows<float, float, NT, G, U>(p, A, firstnbr, numnbrs, jlist, val32, [&](int j) { return dinc[j]; },
                                       reinterpret_cast<float2 *>(xl), [&](int i, bool live, float2 q) {
      if (lane_id() % G == 0 && live) {
        const float eta = (float)P->eta[type[i]];
        const float2 di = dinc[i];
        const double2 r = r64[i];
        dots(i, r.x - (double)(eta * di.x + q.x), r.y - (double)(eta * di.y + q.y));
      }
    });
  } else {
    panel_rows<double, double, NT, G, U>(p, A, firstnbr, numnbrs, jlist, val, [&](int j) { return x[j]; }, xl,
                                         [&](int i, bool live, double2 q) {
      if (lane_id() % G == 0 && live) {
        const double eta = P->eta[type[i]];
        const double2 xi = x[i], bi = b[i];
        dots(i, bi.x - (eta * xi.x + q.x), bi.y - (eta * xi.y + q.y));
      }
    });
  }
  block_partials<6, NT>(d, partials, pstride, p);
}

// Outer test and the inner tolerances, one workgroup per system (over its
// own blocks' partials, in a fixed order).  A right-hand side that meets the
// test gets tol = -1 in the inner scalars, so its inner solve does nothing
// and its d stays 0.
template <int NT>
__global__ void __launch_bounds__(NT)
k_resid_scalars(const double *__restrict__ partials, int pstride, const int *__restrict__ sys_panel,
                MixCtl *__restrict__ ctl, MixState *__restrict__ msa, CgScalars *__restrict__ inner_all, double tol,
                int first, DevStatus *st, int *__restrict__ iters_all) {
  const int sys = blockIdx.x;
  MixState m = msa[sys];
  if (!first && m.done) return;
  double tot[6];
  reduce_partials<6, NT>(partials, pstride, sys_panel[sys], sys_panel[sys + 1], tot);
  if (threadIdx.x != 0) return;
  CgScalars *inner = inner_all + 2 * sys;
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
  m.inc = 1;
  for (int s = 0; s < 2; s++)
    if (!m.converged[s] && !(m.rho[s] <= INC_RHO)) m.inc = 0;
  m.done = (m.converged[0] && m.converged[1]) ? 1 : 0;
  if (!m.done && m.outer >= MAX_OUTER) {
    m.done = 1;
    atomicOr(&st->flags, SWR_F_CG_MAXIT);
  }
  if (m.done) {
    // a system that stops here gets no inner solve: both tolerances are -1
    inner[0].tol = inner[1].tol = -1.0;
    iters_all[2 * sys] = m.matvecs[0];
    iters_all[2 * sys + 1] = m.matvecs[1];
    atomicAdd(&ctl->outer_ndone, 1);
  }
  msa[sys] = m;
}

// x += d over all rows (d is 0 in systems whose inner solve did not run).
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
// magnitude an inner solve asks for.  Iteration k is three launches:
// k_pcg_spmv computes q = eta p + H p with the p.q partials; k_pcg_update
// sums those partials of its system in every workgroup (the same fixed order
// everywhere, so all get the same alpha), then d += alpha p, r -= alpha q,
// z = M^-1 r with the r.z partials; k_pcg_p sums those the same way, forms
// beta and the stopping test sqrt(r.z) <= tol |b| on the outer test's norm,
// and sets p = z + beta p.  (Forming p inside the next SpMV instead, from
// two alternating buffers, made the SpMV 10 us slower than k_pcg_p takes.)
// Each system has its own scalars (its workgroups sum its own partials in a
// fixed order, so a system takes the same iterations alone and in a batch);
// a right-hand side that stops is frozen, panels and tiles of a system whose
// solves both stopped are skipped, and launches after every system stopped
// exit at once.

__global__ void __launch_bounds__(UT)
k_pcg_init(CgLayout L, const float2 *__restrict__ rhs, const double *__restrict__ Hd, float2 *__restrict__ d,
           float2 *__restrict__ r, float2 *__restrict__ z, float2 *__restrict__ p, double *__restrict__ partials,
           int pstride) {
  const int tile = blockIdx.x;
  const int i = L.tile_beg[tile] + threadIdx.x;
  double dd[4] = {0, 0, 0, 0};
  if (i < L.tile_beg[tile + 1]) {
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
  block_partials<4, UT>(dd, partials, pstride, tile);
}

template <int NT>
__global__ void __launch_bounds__(NT)
k_pcg_init_scalars(const double *__restrict__ partials, int pstride, CgLayout L,
                   const CgScalars *__restrict__ inner_all, PcgState *__restrict__ pst, MixCtl *__restrict__ ctl) {
  const int sys = blockIdx.x;
  double tot[4];
  reduce_partials<4, NT>(partials, pstride, L.sys_tile[sys], L.sys_tile[sys + 1], tot);
  if (threadIdx.x != 0) return;
  PcgState *c = pst + 2 * sys;
  for (int s = 0; s < 2; s++) {
    c[s].rz = tot[s];
    c[s].bnorm = sqrt(tot[2 + s]);
    c[s].tol = inner_all[2 * sys + s].tol;
    c[s].iter = 0;
    c[s].alpha = c[s].beta = 0.0;
    c[s].active = (c[s].tol > 0.0 && c[s].bnorm > 0.0 && sqrt(fmax(c[s].rz, 0.0)) > c[s].tol * c[s].bnorm) ? 1 : 0;
    c[2 * gridDim.x + s] = c[s];   // both state buffers (a system without tiles is never carried)
  }
  if (!pcg_live(c)) atomicAdd(&ctl->inner_ndone, 1);
}

template <bool BATCH, int SH>
__global__ void __launch_bounds__(SpmvShape<SH>::NT)
k_pcg_spmv(int n, int nsys, const DevParams *__restrict__ P, const int *__restrict__ type,
           const int *__restrict__ firstnbr, const int *__restrict__ numnbrs, const int *__restrict__ jlist,
           const float *__restrict__ val, const float2 *__restrict__ p, float2 *__restrict__ q,
           const MixCtl *__restrict__ ctl, const PcgState *__restrict__ pst, double *__restrict__ pq_partials,
           int pstride, PanelArgs A, const int *__restrict__ row_sys, int gate = 1) {
  constexpr int NT = SpmvShape<SH>::NT, G = SpmvShape<SH>::G, U = SpmvShape<SH>::U;
  __shared__ float2 xl[SWR_PANEL_MAXC];
  const int pn = blockIdx.x;   // panel and partial slot
  // (gate 0: the autotuner's timing launches, which run every panel)
  if (gate && ctl->inner_ndone >= nsys) return;
  if (gate && BATCH) {
    const int r0 = A.pbeg[pn];
    if (r0 < A.pbeg[pn + 1] && !pcg_live(pst + 2 * row_sys[A.prow[r0]])) return;
  }
  double dd[2] = {0, 0};
  panel_rows<float, float, NT, G, U>(pn, A, firstnbr, numnbrs, jlist, val, [&](int j) { return p[j]; }, xl,
                                     [&](int i, bool live, float2 sum) {
    if (lane_id() % G == 0 && live) {
      const float eta = (float)P->eta[type[i]];
      const float2 pi = p[i];
      const float2 qq = make_float2(eta * pi.x + sum.x, eta * pi.y + sum.y);
      q[i] = qq;
      dd[0] += (double)pi.x * qq.x;
      dd[1] += (double)pi.y * qq.y;
    }
  });
  block_partials<2, NT>(dd, pq_partials, pstride, pn);
}

__global__ void __launch_bounds__(UT)
k_pcg_update(CgLayout L, const double *__restrict__ Hd, float2 *__restrict__ d, float2 *__restrict__ r,
             float2 *__restrict__ z, const float2 *__restrict__ p, const float2 *__restrict__ q,
             const PcgState *__restrict__ pst, const double *__restrict__ pq_partials,
             const int *__restrict__ sys_panel, double *__restrict__ partials, int pstride) {
  __shared__ double alpha[2];
  const int tile = blockIdx.x;
  const int sys = L.tile_sys[tile];
  const PcgState *c = pst + 2 * sys;
  if (!pcg_live(c)) return;
  {  // alpha = r.z / p.q of the system (its panels' p.q partials)
    double tot[2];
    reduce_partials<2, UT>(pq_partials, pstride, sys_panel[sys], sys_panel[sys + 1], tot);
    if (threadIdx.x == 0)
      for (int s = 0; s < 2; s++) alpha[s] = c[s].active ? c[s].rz / tot[s] : 0.0;
    __syncthreads();
  }
  const int i = L.tile_beg[tile] + threadIdx.x;
  const float as = (float)alpha[0], at = (float)alpha[1];
  double dd[2] = {0, 0};
  if (i < L.tile_beg[tile + 1]) {
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
  block_partials<2, UT>(dd, partials, pstride, tile);
}

// beta, the stopping test and p = z + beta p in one launch: every workgroup
// of a system sums the system's r.z partials in the same fixed order (so
// all reach the same beta and the same stopping decision), and the system's
// first tile writes the new state into `next` while the others still read
// `cur` (the state alternates between two buffers, iteration by iteration;
// a stopped system's state is carried over unchanged).
__global__ void __launch_bounds__(UT)
k_pcg_p(CgLayout L, const float2 *__restrict__ z, float2 *__restrict__ p, const PcgState *__restrict__ cur,
        PcgState *__restrict__ next, const double *__restrict__ partials, int pstride, MixCtl *__restrict__ ctl,
        int imax) {
  __shared__ float beta[2];
  __shared__ int act[2];
  const int tile = blockIdx.x;
  const int sys = L.tile_sys[tile];
  const bool first = tile == L.sys_tile[sys];
  const PcgState *c = cur + 2 * sys;
  if (!pcg_live(c)) {
    if (first && threadIdx.x < 2) next[2 * sys + threadIdx.x] = c[threadIdx.x];
    return;
  }
  {
    double tot[2];
    reduce_partials<2, UT>(partials, pstride, L.sys_tile[sys], L.sys_tile[sys + 1], tot);
    if (threadIdx.x == 0) {
      PcgState ns[2] = {c[0], c[1]};
      for (int s = 0; s < 2; s++) {
        if (!ns[s].active) { ns[s].beta = 0.0; continue; }
        ns[s].iter += 1;
        ns[s].beta = tot[s] / ns[s].rz;
        ns[s].rz = tot[s];
        ns[s].active = (ns[s].iter < imax && sqrt(fmax(ns[s].rz, 0.0)) > ns[s].tol * ns[s].bnorm) ? 1 : 0;
      }
      for (int s = 0; s < 2; s++) {
        beta[s] = (float)ns[s].beta;
        act[s] = ns[s].active;
      }
      if (first) {
        next[2 * sys] = ns[0];
        next[2 * sys + 1] = ns[1];
        if (!pcg_live(ns)) {
          atomicAdd(&ctl->inner_ndone, 1);
          atomicMax(&ctl->inner_itmax, max(ns[0].iter, ns[1].iter));
        }
      }
    }
    __syncthreads();
  }
  const int i = L.tile_beg[tile] + threadIdx.x;
  if (i >= L.tile_beg[tile + 1]) return;
  const float2 zz = z[i];
  float2 pp = p[i];
  if (act[0]) pp.x = zz.x + beta[0] * pp.x;
  if (act[1]) pp.y = zz.y + beta[1] * pp.y;
  p[i] = pp;
}

// Writes each system's iteration counts into the inner scalars the outer
// accounting reads; the status block gets the largest.
__global__ void __launch_bounds__(256)
k_pcg_finish(int nsys, const PcgState *__restrict__ pst, CgScalars *__restrict__ inner_all,
             int *__restrict__ status_iters) {
  __shared__ int mx[2][256];
  int m0 = 0, m1 = 0;
  for (int s = threadIdx.x; s < nsys; s += 256) {
    const int i0 = pst[2 * s].iter, i1 = pst[2 * s + 1].iter;
    inner_all[2 * s].iter = i0;
    inner_all[2 * s + 1].iter = i1;
    m0 = max(m0, i0);
    m1 = max(m1, i1);
  }
  mx[0][threadIdx.x] = m0;
  mx[1][threadIdx.x] = m1;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      mx[0][threadIdx.x] = max(mx[0][threadIdx.x], mx[0][threadIdx.x + k]);
      mx[1][threadIdx.x] = max(mx[1][threadIdx.x], mx[1][threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    status_iters[0] = mx[0][0];
    status_iters[1] = mx[1][0];
  }
}

// Inner vectors (float2), carved from the workspace after the states.
enum { MV_D = 0, MV_R, MV_Z, MV_P, MV_Q, MV_B, MV_COUNT };

// The FP64 residual (double2 per row) after the float2 vectors.
inline size_t r64_off(int nsys, size_t ncap) {
  return (vec_off(nsys) + sizeof(float2) * MV_COUNT * ncap + 255) & ~(size_t)255;
}

int ensure_mixw(swr_system *S) {
  const size_t need = r64_off(S->nsys, (size_t)S->ncap) + sizeof(double2) * (size_t)S->ncap;
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
  static_assert(sizeof(MixCtl) <= MS_OFF, "the batch counters must fit their slot");
  const int n = S->n;
  const int nsys = S->nsys;
  const bool batch = nsys > 1;
  const CgLayout L = swr_qeq_cg_layout(S);
  const int nblk = S->npanel;   // SpMV kernels: one workgroup per panel
  const int ps = S->cg_pstride;
  char *w = (char *)S->mixw;
  MixCtl *ctl = (MixCtl *)w;
  MixState *ms = (MixState *)(w + MS_OFF);
  PcgState *pst = (PcgState *)(w + pcg_off(nsys));
  float2 *base = (float2 *)(w + vec_off(nsys));
  const size_t c = S->ncap;
  float2 *dv = base + MV_D * c, *rv = base + MV_R * c, *zv = base + MV_Z * c, *pv = base + MV_P * c,
         *qv = base + MV_Q * c;
  double2 *r64 = (double2 *)(w + r64_off(nsys, c));
  float2 *rhs = base + MV_B * c;      // the inner solve's b: the FP32 residual
  double *pq = S->partials + 2 * (size_t)ps;  // p.q partials (rows 2-3; r.z uses rows 0-1)
  CgScalars *inner = S->cgs + 2 * nsys;  // after the FP64 solve's s, t: the inner tolerances
  double2 *x = (double2 *)S->st;
  const double2 *b = (const double2 *)S->b_s;
  if (nblk < 1 || L.ntiles < 1) return SWR_OK;
  HIP_CHECK(hipMemsetAsync(ctl, 0, sizeof(MixCtl), S->stream));
  struct Head { MixCtl c; char pad[MS_OFF - sizeof(MixCtl)]; MixState m0; };
  static_assert(offsetof(Head, m0) == MS_OFF, "host copy mirrors the workspace");
  static_assert(sizeof(Head) + sizeof(MixCtl) <= SWR_PIN_MIX, "the pinned copy holds both reads");
  Head &h = *(Head *)S->pin_mix;                          // pinned: the copies are plain DMA
  MixCtl &hc = *(MixCtl *)((char *)S->pin_mix + sizeof(Head));
  int outer = 0;
  for (;; outer++) {
    swr_kt_begin(S, KT_MIXED_OUTER);
    spmv_dispatch(S->spmv_shape[2], [&](auto shc) {
      constexpr int SH = decltype(shc)::value;
      const dim3 tb(SpmvShape<SH>::NT);
      if (batch)
        hipLaunchKernelGGL((k_resid<true, SH>), dim3(nblk), tb, 0, S->stream, n, S->dp, S->type, S->H_firstnbr,
                           S->H_numnbrs, S->H_jlist, S->H_val, S->H_val32, (const double2 *)x, (const float2 *)dv, b,
                           S->Hdia_inv, r64, rhs, S->partials,
                           ps, panel_args(S), S->row_sys, ms, outer == 0 ? 1 : 0);
      else
        hipLaunchKernelGGL((k_resid<false, SH>), dim3(nblk), tb, 0, S->stream, n, S->dp, S->type, S->H_firstnbr,
                           S->H_numnbrs, S->H_jlist, S->H_val, S->H_val32, (const double2 *)x, (const float2 *)dv, b,
                           S->Hdia_inv, r64, rhs, S->partials,
                           ps, panel_args(S), S->row_sys, ms, outer == 0 ? 1 : 0);
    });
    if (batch)
      hipLaunchKernelGGL(k_resid_scalars<BT>, dim3(nsys), dim3(BT), 0, S->stream, S->partials, ps, S->sys_panel, ctl,
                         ms, inner, S->ctl.tol * OUTER_TOL_FACTOR, outer == 0 ? 1 : 0, S->status, S->sys_iters);
    else
      hipLaunchKernelGGL(k_resid_scalars<RT>, dim3(nsys), dim3(RT), 0, S->stream, S->partials, ps, S->sys_panel, ctl,
                         ms, inner, S->ctl.tol * OUTER_TOL_FACTOR, outer == 0 ? 1 : 0, S->status, S->sys_iters);
    swr_kt_end(S, KT_MIXED_OUTER);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
    if (outer == 0) { S->mixed_rho0[0] = h.m0.rho[0]; S->mixed_rho0[1] = h.m0.rho[1]; }
    if (h.c.outer_ndone >= nsys || outer + 1 >= MAX_OUTER) break;
    // inner FP32 PCG from d = 0; iterations launched in batches sized from
    // the last solve, the host checking the done count between batches
    swr_kt_begin(S, KT_MIXED_INNER);
    HIP_CHECK(hipMemsetAsync(&ctl->inner_ndone, 0, 2 * sizeof(int), S->stream));
    hipLaunchKernelGGL(k_pcg_init, dim3(L.ntiles), dim3(UT), 0, S->stream, L, rhs, S->Hdia_inv, dv, rv, zv, pv,
                       S->partials, ps);
    if (batch)
      hipLaunchKernelGGL(k_pcg_init_scalars<BT>, dim3(nsys), dim3(BT), 0, S->stream, S->partials, ps, L, inner, pst,
                         ctl);
    else
      hipLaunchKernelGGL(k_pcg_init_scalars<RT>, dim3(nsys), dim3(RT), 0, S->stream, S->partials, ps, L, inner, pst,
                         ctl);
    int it = 0;
    int nb = std::max(2, S->mixed_last_inner + 1);
    while (it < INNER_IMAX) {
      for (int k = 0; k < nb && it < INNER_IMAX; k++, it++) {
        PcgState *cur = pst + (it & 1) * 2 * nsys, *nxt = pst + ((it + 1) & 1) * 2 * nsys;
        spmv_dispatch(S->spmv_shape[1], [&](auto shc) {
          constexpr int SH = decltype(shc)::value;
          const dim3 tb(SpmvShape<SH>::NT);
          if (batch)
            hipLaunchKernelGGL((k_pcg_spmv<true, SH>), dim3(nblk), tb, 0, S->stream, n, nsys, S->dp, S->type,
                               S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val32, (const float2 *)pv, qv, ctl, cur,
                               pq, ps, panel_args(S), S->row_sys);
          else
            hipLaunchKernelGGL((k_pcg_spmv<false, SH>), dim3(nblk), tb, 0, S->stream, n, nsys, S->dp, S->type,
                               S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val32, (const float2 *)pv, qv, ctl, cur,
                               pq, ps, panel_args(S), S->row_sys);
        });
        hipLaunchKernelGGL(k_pcg_update, dim3(L.ntiles), dim3(UT), 0, S->stream, L, S->Hdia_inv, dv, rv, zv,
                           (const float2 *)pv, (const float2 *)qv, cur, (const double *)pq,
                           (const int *)S->sys_panel, S->partials, ps);
        hipLaunchKernelGGL(k_pcg_p, dim3(L.ntiles), dim3(UT), 0, S->stream, L, (const float2 *)zv, pv, cur, nxt,
                           (const double *)S->partials, ps, ctl, INNER_IMAX);
      }
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(&hc, ctl, sizeof(hc), hipMemcpyDeviceToHost, S->stream));
      HIP_CHECK(hipStreamSynchronize(S->stream));
      if (hc.inner_ndone >= nsys) {
        S->mixed_last_inner = hc.inner_itmax;
        break;
      }
      nb = 4;
    }
    // (the state after `it` iterations; stopped systems were carried into it)
    hipLaunchKernelGGL(k_pcg_finish, dim3(1), dim3(256), 0, S->stream, nsys, pst + (it & 1) * 2 * nsys, inner,
                       S->status->inner_iters);
    swr_kt_end(S, KT_MIXED_INNER);
    hipLaunchKernelGGL(k_mixed_update, dim3(div_up(n, 256)), dim3(256), 0, S->stream, n, x, (const float2 *)dv);
    HIP_CHECK(hipGetLastError());
  }
  S->mixed_outer = outer + 1;
  return SWR_OK;
}

// Autotuner hook: the average time of `reps` launches of the mixed solve's
// FP64 residual (which = 0) or FP32 inner SpMV (which = 1) in shape sh on the
// current H, after at least one mixed-precision solve (its workspace and
// vectors are scratch between solves).
int swr_qeq_time_mixed(swr_system *S, int which, int sh, int reps, float *ms) {
  if (!S->mixed || !S->mixw || !S->H_val32 || S->n <= 0) return SWR_ERR_STATE;
  const int n = S->n, nsys = S->nsys;
  const bool batch = nsys > 1;
  const int nblk = S->npanel;
  const int ps = S->cg_pstride;
  char *w = (char *)S->mixw;
  MixCtl *ctl = (MixCtl *)w;
  MixState *mst = (MixState *)(w + MS_OFF);
  PcgState *pst = (PcgState *)(w + pcg_off(nsys));
  float2 *base = (float2 *)(w + vec_off(nsys));
  const size_t c = S->ncap;
  float2 *pv = base + MV_P * c, *qv = base + MV_Q * c, *rhs = base + MV_B * c, *dv = base + MV_D * c;
  double2 *r64 = (double2 *)(w + r64_off(nsys, c));
  double *pq = S->partials + 2 * (size_t)ps;
  const double2 *x = (const double2 *)S->st, *b = (const double2 *)S->b_s;
  hipEvent_t e[2];
  HIP_CHECK(hipEventCreate(&e[0]));
  HIP_CHECK(hipEventCreate(&e[1]));
  spmv_dispatch(sh, [&](auto shc) {
    constexpr int SH = decltype(shc)::value;
    const dim3 tb(SpmvShape<SH>::NT);
    for (int r = -1; r < reps; r++) {
      if (r == 0) (void)hipEventRecord(e[0], S->stream);
      if (which == 0) {
        if (batch)
          hipLaunchKernelGGL((k_resid<true, SH>), dim3(nblk), tb, 0, S->stream, n, S->dp, S->type, S->H_firstnbr,
                             S->H_numnbrs, S->H_jlist, S->H_val, S->H_val32, x, (const float2 *)dv, b, S->Hdia_inv, r64,
                             rhs, S->partials, ps, panel_args(S),
                             S->row_sys, mst, 1);
        else
          hipLaunchKernelGGL((k_resid<false, SH>), dim3(nblk), tb, 0, S->stream, n, S->dp, S->type, S->H_firstnbr,
                             S->H_numnbrs, S->H_jlist, S->H_val, S->H_val32, x, (const float2 *)dv, b, S->Hdia_inv, r64,
                             rhs, S->partials, ps, panel_args(S),
                             S->row_sys, mst, 1);
      } else {
        if (batch)
          hipLaunchKernelGGL((k_pcg_spmv<true, SH>), dim3(nblk), tb, 0, S->stream, n, nsys, S->dp, S->type,
                             S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val32, (const float2 *)pv, qv, ctl, pst,
                             pq, ps, panel_args(S), S->row_sys, 0);
        else
          hipLaunchKernelGGL((k_pcg_spmv<false, SH>), dim3(nblk), tb, 0, S->stream, n, nsys, S->dp, S->type,
                             S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val32, (const float2 *)pv, qv, ctl, pst,
                             pq, ps, panel_args(S), S->row_sys, 0);
      }
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
