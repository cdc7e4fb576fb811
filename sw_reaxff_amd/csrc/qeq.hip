// qeq.hip — QEq: H matrix (a11-a13), SpMV (a14), the s/t pipelined PCG
// (a15) and the charge update (a16).
//
// H (compute_H_Full_C_para, fix_qeq_reax_sw64.c:768-911): full rows for the
// local atoms, every Verlet neighbour with r^2 <= swb^2,
//   val = Tap(r) * 14.4 / (r^3 + shld[ti][tj])^0.3333333333333
// with Tap by Horner (calculate_H, fix_qeq_reax_sunway.cpp:965-981).  Row i
// holds jlist/val[firstnbr[i] .. firstnbr[i]+numnbrs[i]); firstnbr[i] is the
// row's Verlet offset (the reference uses a 600*i stride,
// fix_qeq_reax_sw64.c:828).  Ghost columns are folded onto their owning
// local atom: on one domain the forward comm of x makes x[ghost] ==
// x[owner], so the SpMV result is unchanged and x needs only n entries.
//
// CG: CG_v2 (fix_qeq_reax_sunway.cpp:983-1167) for s and t at once.  Vectors
// are (s,t) pairs (double2), so each H entry is read once and feeds both
// solves.  One kernel per iteration does the reference's B and C steps of
// iteration i (SpMV of d, x/p/ss/v/z updates) and the A step of iteration
// i+1 (r/u/w updates, the two dots, d = M^-1 w).  Each workgroup writes its
// partial dots; a one-workgroup kernel sums them in a fixed order and
// computes the next iteration's scalars (no atomics, so results are the same
// bits on every run).  Each solve stops by the reference's test
// sqrt(u.r)/|b| <= tol (or i = imax) and is frozen while the other runs.
#include "swr_internal.h"
#include "swr_device.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr int QT = 256;                    // threads per block
constexpr int QW = QT / SWR_WAVE;          // waves per block
constexpr int ROWS_PER_WAVE = 2;         // one row per half-wave in the CG kernels
constexpr int RPB = QW * ROWS_PER_WAVE;    // rows per block (fixed: results do not depend on grid size)
constexpr double EV_TO_KCAL_PER_MOL = 14.4;

// ---------------------------------------------------------------- H build
__global__ void __launch_bounds__(QT)
k_build_H(int n, const DevParams *__restrict__ P, const double *__restrict__ x, const int *__restrict__ type,
          const int *__restrict__ owner, const long long *__restrict__ voff, const int *__restrict__ vlist,
          int *__restrict__ firstnbr, int *__restrict__ numnbrs, int *__restrict__ jlist, double *__restrict__ val,
          float *__restrict__ val32, double *__restrict__ Hdia_inv, double2 *__restrict__ b2,
          double2 *__restrict__ st, const double *__restrict__ s_hist, const double *__restrict__ t_hist) {
  const int i = xcd_block(gridDim.x) * QW + wave_in_block();
  if (i >= n) return;
  const int lane = lane_id();
  const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
  const int ti = type[i];
  const int nt1 = SWR_MAX_TYPES + 1;
  const double swbsq = P->swbsq;
  double T[8];
#pragma unroll
  for (int k = 0; k < 8; k++) T[k] = P->Tap[k];
  const long long beg = voff[i], end = voff[i + 1];
  int cnt = 0;
  for (long long base = beg; base < end; base += SWR_WAVE) {
    const long long k = base + lane;
    bool hit = false;
    int jo = 0;
    double hv = 0.0;
    if (k < end) {
      const int j = vlist[k];
      const double dx = x[3 * j] - xi, dy = x[3 * j + 1] - yi, dz = x[3 * j + 2] - zi;
      const double r_sqr = dx * dx + dy * dy + dz * dz;
      if (r_sqr <= swbsq) {
        hit = true;
        const double r = sqrt(r_sqr);
        const double gamma = P->shld[ti * nt1 + type[j]];
        double Taper = T[7] * r + T[6];
        Taper = Taper * r + T[5];
        Taper = Taper * r + T[4];
        Taper = Taper * r + T[3];
        Taper = Taper * r + T[2];
        Taper = Taper * r + T[1];
        Taper = Taper * r + T[0];
        double denom = r * r * r + gamma;
        denom = pow(denom, 0.3333333333333);
        hv = Taper * EV_TO_KCAL_PER_MOL / denom;
        jo = owner[j];
      }
    }
    const unsigned long long m = __ballot(hit);
    if (hit) {
      const long long slot = beg + cnt + ballot_rank(m);
      jlist[slot] = jo;
      val[slot] = hv;
      if (val32) val32[slot] = (float)hv;
    }
    cnt += __popcll(m);
  }
  if (lane == 0) {
    firstnbr[i] = (int)beg;
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
__global__ void __launch_bounds__(QT)
k_spmv1(int n, const DevParams *__restrict__ P, const int *__restrict__ type, const int *__restrict__ firstnbr,
        const int *__restrict__ numnbrs, const int *__restrict__ jlist, const double *__restrict__ val,
        const double *__restrict__ xv, double *__restrict__ b) {
  const int i = xcd_block(gridDim.x) * QW + wave_in_block();
  if (i >= n) return;
  const int lane = lane_id();
  const int beg = firstnbr[i], end = beg + numnbrs[i];
  double a0 = 0.0, a1 = 0.0;
  int k = beg + lane;
  for (; k + SWR_WAVE < end; k += 2 * SWR_WAVE) {
    const int j0 = jlist[k], j1 = jlist[k + SWR_WAVE];
    const double v0 = val[k], v1 = val[k + SWR_WAVE];
    a0 += v0 * xv[j0];
    a1 += v1 * xv[j1];
  }
  if (k < end) a0 += val[k] * xv[jlist[k]];
  const double a = wave_sum(a0 + a1);
  if (lane == 0) b[i] = P->eta[type[i]] * xv[i] + a;
}

constexpr int RT = 1024;  // threads of the one-workgroup reduction kernels

// Sums partials[k*pstride + b] over b < nblocks for k < ND with one
// workgroup: thread t adds b = t, t+RT, ... in order, then a fixed LDS tree
// combines the RT sums.  The order depends only on the block count.
template <int ND>
__device__ void reduce_partials(const double *__restrict__ partials, int pstride, int nblocks, double *out) {
  __shared__ double sr[ND][RT];
  double acc[ND];
#pragma unroll
  for (int k = 0; k < ND; k++) acc[k] = 0.0;
#pragma unroll 4
  for (int b = threadIdx.x; b < nblocks; b += RT)
#pragma unroll
    for (int k = 0; k < ND; k++) acc[k] += partials[(size_t)k * pstride + b];
#pragma unroll
  for (int k = 0; k < ND; k++) sr[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = RT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < ND; k++) sr[k][threadIdx.x] += sr[k][threadIdx.x + s];
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < ND; k++) out[k] = sr[k][0];
}

// Block sum of per-thread values d[k] in a fixed order (wave tree, then
// waves in order); thread 0 writes partials[k*pstride + blk].
template <int ND, int NT>
__device__ __forceinline__ void block_partials(const double *d, double *__restrict__ partials, int pstride,
                                               int blk) {
  constexpr int NW = NT / SWR_WAVE;
  __shared__ double sred[NW][ND];
  const int w = wave_in_block();
#pragma unroll
  for (int k = 0; k < ND; k++) {
    const double s = wave_sum(d[k]);
    if (lane_id() == 0) sred[w][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < ND) {
    double s = 0.0;
#pragma unroll
    for (int ww = 0; ww < NW; ww++) s += sred[ww][threadIdx.x];
    partials[(size_t)threadIdx.x * pstride + blk] = s;
  }
}

// ---------------------------------------------------------------- CG
enum { PH_S1 = 0, PH_S2, PH_S3, PH_A1, PH_K };

struct CgVecs {
  double2 *x, *r, *u, *w, *z, *ss, *v, *p, *m, *d[2];
  const double2 *b;
  const double *Hd;
};

// Scalars after the A step of an iteration reduced dots l0 = u.r, l1 = u.w
// (fix_qeq_reax_sunway.cpp:1133-1141 and :1152).
__device__ void cg_after_A(CgScalars &c, double l0, double l1, int imax, double tol) {
  const double beta = l0 / c.sig_old;
  c.heta = l1 - beta * beta * c.heta;
  c.beta = beta;
  c.alpha_x = c.alpha_a;
  c.alpha_a = l0 / c.heta;
  c.sig_old = l0;
  c.dg0 = l0;
  c.dg1 = l1;
  c.iter += 1;
  c.active_b = 1;
  c.active_a = (c.iter < imax && sqrt(l0) / c.b_norm > tol) ? 1 : 0;
}

// Half-wave (32 lanes) per row, 2 rows per wave: four independent
// index/value/gather loads in flight per lane; fixed summation order
// (per lane a0..a3, then a fixed 32-lane tree).
__device__ __forceinline__ double2 row_spmv2_h(int beg, int cnt, const int *__restrict__ jlist,
                                               const double *__restrict__ val, const double2 *__restrict__ xv) {
  const int sl = lane_id() & 31;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
  const int end = beg + cnt;
  int k = beg + sl;
  for (; k + 96 < end; k += 128) {
    const int j0 = jlist[k], j1 = jlist[k + 32], j2 = jlist[k + 64], j3 = jlist[k + 96];
    const double v0 = val[k], v1 = val[k + 32], v2 = val[k + 64], v3 = val[k + 96];
    const double2 x0 = xv[j0], x1 = xv[j1], x2 = xv[j2], x3 = xv[j3];
    a0 += v0 * x0.x; b0 += v0 * x0.y;
    a1 += v1 * x1.x; b1 += v1 * x1.y;
    a2 += v2 * x2.x; b2 += v2 * x2.y;
    a3 += v3 * x3.x; b3 += v3 * x3.y;
  }
  for (; k < end; k += 32) {
    const double2 xx = xv[jlist[k]];
    const double v = val[k];
    a0 += v * xx.x;
    b0 += v * xx.y;
  }
  double a = (a0 + a1) + (a2 + a3), b = (b0 + b1) + (b2 + b3);
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, SWR_WAVE);
    b += __shfl_xor(b, o, SWR_WAVE);
  }
  return make_double2(a, b);
}

template <int PH>
__global__ void __launch_bounds__(QT)
k_cg(int n, const DevParams *__restrict__ P, const int *__restrict__ type, const int *__restrict__ firstnbr,
     const int *__restrict__ numnbrs, const int *__restrict__ jlist, const double *__restrict__ val, CgVecs V,
     int par_in, const CgScalars *__restrict__ cgs, double *__restrict__ partials, int pstride) {
  CgScalars c0, c1;
  bool b0 = false, b1 = false, a0 = true, a1 = true;
  if (PH == PH_A1 || PH == PH_K) {
    c0 = cgs[0];
    c1 = cgs[1];
    b0 = PH == PH_K && c0.active_b != 0;
    b1 = PH == PH_K && c1.active_b != 0;
    a0 = c0.active_a != 0;
    a1 = c1.active_a != 0;
    if (!b0 && !b1 && !a0 && !a1) return;  // both solves finished
  }
  const int w = wave_in_block(), lane = lane_id();
  const int blk = blockIdx.x;
  const int i = blk * RPB + w * 2 + (lane >> 5);
  const bool owner = (lane & 31) == 0 && i < n;
  double2 q = make_double2(0.0, 0.0);
  if (PH != PH_A1) {
    const double2 *src = PH == PH_S1 ? V.x : PH == PH_S2 ? V.d[0] : PH == PH_S3 ? V.d[1] : (par_in ? V.d[1] : V.d[0]);
    const int beg = i < n ? firstnbr[i] : 0, cnt = i < n ? numnbrs[i] : 0;
    q = row_spmv2_h(beg, cnt, jlist, val, src);
    if (owner) {
      // b_i = eta_i x_i + sum_j H_ij x_j (sparse_matvec_C_para, :424-466)
      const double eta = P->eta[type[i]];
      const double2 xi = src[i];
      q.x = eta * xi.x + q.x;
      q.y = eta * xi.y + q.y;
    }
  }
  double dots[6] = {0, 0, 0, 0, 0, 0};
  if (owner) {
    const double Hd = V.Hd[i];
    if (PH == PH_S1) {
      const double2 b = V.b[i];
      const double2 r = make_double2(b.x - 1 * q.x, b.y - 1 * q.y);
      const double2 u = make_double2(r.x * Hd, r.y * Hd);
      V.r[i] = r; V.u[i] = u; V.d[0][i] = u;
    } else if (PH == PH_S2) {
      V.w[i] = q;
      const double2 m = make_double2(q.x * Hd, q.y * Hd);
      V.m[i] = m; V.d[1][i] = m;
    } else if (PH == PH_S3) {
      const double2 b = V.b[i], u = V.u[i], r = V.r[i], ww = V.w[i], m = V.m[i];
      V.z[i] = q; V.p[i] = u; V.ss[i] = ww; V.v[i] = m;
      dots[0] = b.x * b.x; dots[1] = u.x * r.x; dots[2] = u.x * ww.x;
      dots[3] = b.y * b.y; dots[4] = u.y * r.y; dots[5] = u.y * ww.y;
    } else {
      double2 xx = V.x[i], p = V.p[i], u = V.u[i], ss = V.ss[i], w2 = V.w[i], v = V.v[i], z = V.z[i], r = V.r[i];
      const double2 dc = PH == PH_K ? (par_in ? V.d[1] : V.d[0])[i] : make_double2(0.0, 0.0);
      double2 dn = dc;
      // B: x += alpha p; p = u + p beta; ss = w + ss beta; v = d + v beta   (:1135-1141)
      // C: z = q + z beta                                                   (:1147-1150)
      if (PH == PH_K && b0) {
        const double al = c0.alpha_x, be = c0.beta;
        xx.x += al * p.x; p.x = u.x + p.x * be; ss.x = w2.x + ss.x * be; v.x = dc.x + v.x * be;
        z.x = q.x + z.x * be;
      }
      if (PH == PH_K && b1) {
        const double al = c1.alpha_x, be = c1.beta;
        xx.y += al * p.y; p.y = u.y + p.y * be; ss.y = w2.y + ss.y * be; v.y = dc.y + v.y * be;
        z.y = q.y + z.y * be;
      }
      // A of the next iteration: r -= alpha ss; u -= alpha v; w -= alpha z; d = M^-1 w (:1113-1123)
      if (a0) {
        const double al = c0.alpha_a;
        r.x -= al * ss.x; u.x -= al * v.x; w2.x -= al * z.x;
        dots[0] = u.x * r.x; dots[1] = u.x * w2.x;
        dn.x = w2.x * Hd;
      }
      if (a1) {
        const double al = c1.alpha_a;
        r.y -= al * ss.y; u.y -= al * v.y; w2.y -= al * z.y;
        dots[2] = u.y * r.y; dots[3] = u.y * w2.y;
        dn.y = w2.y * Hd;
      }
      if (PH == PH_K && (b0 || b1)) {
        V.x[i] = xx; V.p[i] = p; V.ss[i] = ss; V.v[i] = v; V.z[i] = z;
      }
      if (a0 || a1) {
        V.r[i] = r; V.u[i] = u; V.w[i] = w2;
        (PH == PH_A1 ? V.d[0] : (par_in ? V.d[0] : V.d[1]))[i] = dn;
      }
    }
  }
  if (PH == PH_S1 || PH == PH_S2) return;
  block_partials<PH == PH_S3 ? 6 : 4, QT>(dots, partials, pstride, blk);
}

// Scalars of the iteration whose dots k_cg<PH> just wrote (one workgroup).
template <int PH>
__global__ void __launch_bounds__(RT)
k_cg_scalars(int nblk, const double *__restrict__ partials, int pstride, CgScalars *__restrict__ cgs, int imax,
             double tol, DevStatus *st) {
  CgScalars c0 = cgs[0], c1 = cgs[1];
  const bool b0 = PH == PH_K && c0.active_b != 0, b1 = PH == PH_K && c1.active_b != 0;
  const bool a0 = PH == PH_S3 || c0.active_a != 0, a1 = PH == PH_S3 || c1.active_a != 0;
  if (!b0 && !b1 && !a0 && !a1) return;
  constexpr int nd = PH == PH_S3 ? 6 : 4;
  double tot[nd];
  reduce_partials<nd>(partials, pstride, nblk, tot);
  if (threadIdx.x != 0) return;
  unsigned int bad = 0;
  if constexpr (PH == PH_S3) {
    auto setup = [&](CgScalars &cc, double bb, double ur, double uw, int s) {
      cc.bdot = bb; cc.udot_r = ur; cc.udot_w = uw;
      cc.b_norm = sqrt(bb);
      cc.sig_old = ur;
      cc.heta = uw;
      cc.alpha_a = cc.sig_old / cc.heta;
      cc.alpha_x = 0.0;
      cc.beta = 0.0;
      cc.dg0 = cc.sig_old;
      cc.dg1 = 0.0;
      cc.iter = 1;
      cc.active_b = 0;
      cc.active_a = (1 < imax && sqrt(cc.dg0) / cc.b_norm > tol) ? 1 : 0;
      cc.done = cc.active_a ? 0 : 1;
      if (cc.done) st->cg_iters[s] = cc.iter;
      if (!isfinite(cc.alpha_a) && cc.active_a) bad = 1;
    };
    setup(c0, tot[0], tot[1], tot[2], 0);
    setup(c1, tot[3], tot[4], tot[5], 1);
  } else {
    auto fin = [&](CgScalars &cc, bool act_a, bool act_b, double l0, double l1, int s) {
      if (act_a) {
        cg_after_A(cc, l0, l1, imax, tol);
        if (!isfinite(cc.alpha_a) || !isfinite(cc.beta)) bad = 1;
      } else if (act_b) {
        cc.active_b = 0;
        cc.done = 1;
        st->cg_iters[s] = cc.iter;
      }
    };
    fin(c0, a0, b0, tot[0], tot[1], 0);
    fin(c1, a1, b1, tot[2], tot[3], 1);
  }
  cgs[0] = c0;
  cgs[1] = c1;
  if (bad) atomicOr(&st->flags, SWR_F_NAN);
}

// ---------------------------------------------------------------- calculate_Q
// Block partial sums of s and t over the local atoms (calculate_Q,
// fix_qeq_reax_sunway.cpp:1733-1734: the two parallel_vector_acc calls).
__global__ void __launch_bounds__(QT)
k_q_sums(int n, const double2 *__restrict__ st, double *__restrict__ partials, int pstride) {
  const int i = blockIdx.x * QT + threadIdx.x;
  double d[2] = {0.0, 0.0};
  if (i < n) { const double2 v = st[i]; d[0] = v.x; d[1] = v.y; }
  block_partials<2, QT>(d, partials, pstride, blockIdx.x);
}

__global__ void __launch_bounds__(RT)
k_q_total(int nblk, const double *__restrict__ partials, int pstride, double *__restrict__ qsum) {
  double tot[2];
  reduce_partials<2>(partials, pstride, nblk, tot);
  if (threadIdx.x != 0) return;
  qsum[0] = tot[0];
  qsum[1] = tot[1];
  qsum[2] = tot[0] / tot[1];
}

// q = s - u t for local atoms, history shift, and the ghost q refresh
// (calculate_Q :1736-1754).
__global__ void k_q_apply(int n, int N, const double2 *__restrict__ st, const double *__restrict__ qsum,
                          const int *__restrict__ owner, double *__restrict__ q, double *__restrict__ s_hist,
                          double *__restrict__ t_hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double u = qsum[2];
  const int o = i < n ? i : owner[i];
  const double2 v = st[o];
  q[i] = v.x - u * v.y;
  if (i < n) {
    double *sh = s_hist + 5 * (size_t)i, *th = t_hist + 5 * (size_t)i;
    for (int k = 4; k > 0; --k) { sh[k] = sh[k - 1]; th[k] = th[k - 1]; }
    sh[0] = v.x;
    th[0] = v.y;
  }
}

CgVecs make_vecs(swr_system *S) {
  CgVecs V;
  double2 *base = (double2 *)S->cgv;
  const size_t c = S->ncap;
  V.r = base + CGV_R * c; V.u = base + CGV_U * c; V.w = base + CGV_W * c; V.z = base + CGV_Z * c;
  V.ss = base + CGV_SS * c; V.v = base + CGV_V * c; V.p = base + CGV_P * c; V.m = base + CGV_M * c;
  V.d[0] = base + CGV_D0 * c; V.d[1] = base + CGV_D1 * c;
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
                     S->voff, S->vlist, S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val,
                     S->mixed ? S->H_val32 : (float *)nullptr, S->Hdia_inv, (double2 *)S->b_s, (double2 *)S->st,
                     S->s_hist, S->t_hist);
  swr_kt_end(S, KT_BUILD_H);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_qeq_spmv(swr_system *S, const double *xv, double *b) {
  const int n = S->n;
  swr_kt_begin(S, KT_SPMV);
  hipLaunchKernelGGL(k_spmv1, dim3(div_up(n, QW)), dim3(QT), 0, S->stream, n, S->dp, S->type, S->H_firstnbr,
                     S->H_numnbrs, S->H_jlist, S->H_val, xv, b);
  swr_kt_end(S, KT_SPMV);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

static int launch_cg(swr_system *S, int ph, int par_in) {
  const int n = S->n;
  const int nblk = div_up(n, RPB);
  const int ps = S->cg_blocks;
  CgVecs V = make_vecs(S);
  const double tol = S->ctl.tol;
  const int kt = ph == PH_K ? KT_CG_ITER : KT_CG_SETUP;
  const int kts = ph == PH_K ? KT_CG_SCALARS : KT_CG_SETUP;
#define CG_ARGS n, S->dp, S->type, S->H_firstnbr, S->H_numnbrs, S->H_jlist, S->H_val, V, par_in, S->cgs, S->partials, ps
#define SC_ARGS nblk, S->partials, ps, S->cgs, S->imax, tol, S->status
  swr_kt_begin(S, kt);
  switch (ph) {
    case PH_S1: hipLaunchKernelGGL(k_cg<PH_S1>, dim3(nblk), dim3(QT), 0, S->stream, CG_ARGS); break;
    case PH_S2: hipLaunchKernelGGL(k_cg<PH_S2>, dim3(nblk), dim3(QT), 0, S->stream, CG_ARGS); break;
    case PH_S3: hipLaunchKernelGGL(k_cg<PH_S3>, dim3(nblk), dim3(QT), 0, S->stream, CG_ARGS); break;
    case PH_A1: hipLaunchKernelGGL(k_cg<PH_A1>, dim3(nblk), dim3(QT), 0, S->stream, CG_ARGS); break;
    default: hipLaunchKernelGGL(k_cg<PH_K>, dim3(nblk), dim3(QT), 0, S->stream, CG_ARGS); break;
  }
  swr_kt_end(S, kt);
  swr_kt_begin(S, kts);
  switch (ph) {
    case PH_S1: case PH_S2: break;
    case PH_S3: hipLaunchKernelGGL(k_cg_scalars<PH_S3>, dim3(1), dim3(RT), 0, S->stream, SC_ARGS); break;
    case PH_A1: hipLaunchKernelGGL(k_cg_scalars<PH_A1>, dim3(1), dim3(RT), 0, S->stream, SC_ARGS); break;
    default: hipLaunchKernelGGL(k_cg_scalars<PH_K>, dim3(1), dim3(RT), 0, S->stream, SC_ARGS); break;
  }
  swr_kt_end(S, kts);
#undef CG_ARGS
#undef SC_ARGS
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Runs both solves to convergence.  Iteration kernels are enqueued in
// batches; between batches the host reads the two done flags.
int swr_qeq_cg_run(swr_system *S, int first_batch) {
  int rc;
  if ((rc = launch_cg(S, PH_S1, 0)) || (rc = launch_cg(S, PH_S2, 0)) || (rc = launch_cg(S, PH_S3, 0)) ||
      (rc = launch_cg(S, PH_A1, 0)))
    return rc;
  int k = 1;
  int batch = std::max(2, first_batch);
  while (k < S->imax + 1) {
    for (int b = 0; b < batch && k < S->imax + 1; b++, k++)
      if ((rc = launch_cg(S, PH_K, (k - 1) & 1))) return rc;
    CgScalars h[2];
    HIP_CHECK(hipMemcpyAsync(h, S->cgs, sizeof(h), hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
    if (h[0].done && h[1].done) break;
    batch = 4;
  }
  return SWR_OK;
}

int swr_qeq_calc_q(swr_system *S) {
  const int n = S->n;
  const int nblk = div_up(n, QT);
  const int ps = S->cg_blocks;
  double *qsum = S->partials + (size_t)ps * 8;
  swr_kt_begin(S, KT_QEQ_Q);
  hipLaunchKernelGGL(k_q_sums, dim3(nblk), dim3(QT), 0, S->stream, n, (const double2 *)S->st, S->partials, ps);
  hipLaunchKernelGGL(k_q_total, dim3(1), dim3(RT), 0, S->stream, nblk, S->partials, ps, qsum);
  hipLaunchKernelGGL(k_q_apply, dim3(div_up(S->N, 256)), dim3(256), 0, S->stream, n, S->N, (const double2 *)S->st,
                     qsum, S->owner, S->q, S->s_hist, S->t_hist);
  swr_kt_end(S, KT_QEQ_Q);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Partial-sum slots per dot (the stride of the partials array): enough for
// every grid the QEq kernels launch on n rows.
int swr_qeq_blocks(int n) { return div_up(n, RPB); }
