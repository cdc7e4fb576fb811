// qeq_cg.h — the pipelined Jacobi-PCG of the QEq solve (CG_v2,
// fix_qeq_reax_sunway.cpp:983-1167) as templates over the vector type, shared
// by the FP64 solve (qeq.hip) and the FP32 inner solve of the mixed-precision
// refinement (qeq_mixed.hip).
//
// Vectors hold the s and t systems side by side (T2 = double2 or float2), so
// each H entry is read once and feeds both solves.  An iteration is three
// launches: k_cg_spmv (q = H d, G lanes per row with 16-byte loads of the
// aligned rows, nothing else so the gathers run at full occupancy; the
// shape is the autotuner's), k_cg_upd (one thread per row, coalesced:
// the reference's B and C steps of iteration i -- x/p/ss/v/z updates -- and
// the A step of iteration i+1 -- r/u/w updates, the two dots, d = M^-1 w) and
// k_cg_scalars.  Each k_cg_upd workgroup writes its partial dots; the
// one-workgroup k_cg_scalars sums them in a fixed order and computes the next
// iteration's scalars (no atomics, so results are the same bits on every
// run).  Dot products and all
// scalars are FP64 whatever T is.  Each solve stops by the reference's test
// sqrt(u.r)/|b| <= tol (or i = imax) and is frozen while the other runs.
#pragma once

#include <cstring>
#include <type_traits>

#include "swr_internal.h"
#include "swr_device.h"

namespace swr_cg {

constexpr int QT = 256;                    // threads per block
constexpr int QW = QT / SWR_WAVE;          // waves per block
constexpr int ROWS_PER_WAVE = 2;           // one row per half-wave
constexpr int RPB = QW * ROWS_PER_WAVE;    // rows per block (fixed: results do not depend on grid size)
constexpr int RT = 1024;                   // threads of the one-workgroup reduction kernels
constexpr int UT = 256;                    // rows (threads) per block of the update kernels

enum { PH_S1 = 0, PH_S2, PH_S3, PH_A1, PH_K };

// Rows of a batch of independent systems (one system when unbatched).
// System s owns local rows [row_off[s], row_off[s+1]); the vector kernels
// run over tiles of at most UT rows of one system each (tile t: rows
// [tile_beg[t], tile_beg[t+1]) of system tile_sys[t]; system s owns tiles
// [sys_tile[s], sys_tile[s+1])), so every partial sum belongs to one system
// and each system's scalars come from its own tiles in a fixed order: a
// system's iterations are the same alone and in a batch.
struct CgLayout {
  const int *row_sys, *tile_beg, *tile_sys, *sys_tile;
  int nsys, ntiles;
  int *done_count;    // systems whose s and t solves have both finished
};

template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef double2 type; };
template <> struct Vec2<float> { typedef float2 type; };

template <typename T>
struct CgVecsT {
  typedef typename Vec2<T>::type T2;
  T2 *x, *r, *u, *w, *z, *ss, *v, *p, *m, *d[2], *q;
  const T2 *b;
  const double *Hd;
};

// Sums partials[k*pstride + b] over b in [b0, b1) for k < ND with one
// workgroup of NT threads: thread t adds b = b0+t, b0+t+NT, ... in order,
// each wave combines its lanes by a fixed butterfly, then the wave sums are
// added in wave order (one barrier) by thread 0, the only thread whose out
// holds the sums.  The order depends only on the range and NT.
template <int ND, int NT = RT>
__device__ void reduce_partials(const double *__restrict__ partials, int pstride, int b0, int b1, double *out) {
  constexpr int NW = NT / SWR_WAVE;
  __shared__ double sr[NW][ND];
  double acc[ND];
#pragma unroll
  for (int k = 0; k < ND; k++) acc[k] = 0.0;
#pragma unroll 4
  for (int b = b0 + (int)threadIdx.x; b < b1; b += NT)
#pragma unroll
    for (int k = 0; k < ND; k++) acc[k] += partials[(size_t)k * pstride + b];
  const int w = wave_in_block();
#pragma unroll
  for (int k = 0; k < ND; k++) {
    const double v = wave_sum(acc[k]);
    if (lane_id() == 0) sr[w][k] = v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ND; k++) {
    double t = 0.0;
    if (threadIdx.x == 0)
#pragma unroll
      for (int ww = 0; ww < NW; ww++) t += sr[ww][k];
    out[k] = t;
  }
}

// Block sum of per-thread values d[k] in a fixed order (wave tree, then
// waves in order); thread k < ND writes partials[k*pstride + blk].
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

// Scalars after the A step of an iteration reduced dots l0 = u.r, l1 = u.w
// (fix_qeq_reax_sunway.cpp:1133-1141 and :1152).
__device__ inline void cg_after_A(CgScalars &c, double l0, double l1, int imax) {
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
  c.active_a = (c.iter < imax && sqrt(l0) / c.b_norm > c.tol) ? 1 : 0;
}

// Four consecutive stored values of an aligned row as T (two 16-byte loads
// of double pairs, or one of four floats).
template <typename T, typename VT> struct Quad;
template <> struct Quad<double, double> {
  double v[4];
  __device__ __forceinline__ void load(const double *p) {
    const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
  }
};
template <> struct Quad<float, float> {
  float v[4];
  __device__ __forceinline__ void load(const float *p) {
    const float4 a = *reinterpret_cast<const float4 *>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  }
};

// G lanes per row over a row that starts 16-byte aligned and is padded to a
// multiple of 4 entries (value 0): each lane takes groups of 4 consecutive
// entries (one index load -- four 32-bit columns of x, or four 16-bit
// positions in a panel's LDS list --, 16-byte value loads), U groups in
// flight, group g of the row on lane g % G.  J and V point at the row's
// first entry; xv(k) returns the vector entry of index k (x[k], or entry k
// of the panel's list).  Fixed summation order (per lane two accumulators
// per right-hand side, then a G-lane butterfly), so a row's sum depends
// only on the row and (G, U).  Returns the sums on every lane of the row's
// group.
// A lane's first U groups of a row (see row_sum4), requested ahead.
template <typename T, typename VT, int U, typename I4>
struct RowBatch {
  I4 jj[U];
  Quad<T, VT> vv[U];
};

// Requests the first U groups of this lane's share of a row when the row
// has that many (the first trip of row_sum4's main loop).
template <typename T, typename VT, int G, int U, typename I4>
__device__ __forceinline__ void row_batch0(int cnt, const I4 *__restrict__ J, const VT *__restrict__ V,
                                           RowBatch<T, VT, U, I4> &b) {
  const int sl = lane_id() % G;
  if (sl + (U - 1) * G < ((cnt + 3) >> 2)) {
#pragma unroll
    for (int u = 0; u < U; u++) b.jj[u] = J[sl + u * G];
#pragma unroll
    for (int u = 0; u < U; u++) b.vv[u].load(V + 4 * (sl + u * G));
  }
}

// PRE: the main loop's first trip uses the groups row_batch0 requested into
// pre (the same entries, summed in the same order).
template <typename T, typename VT, int G, int U, bool PRE = false, typename I4, class XV>
__device__ __forceinline__ typename Vec2<T>::type row_sum4(int cnt, const I4 *__restrict__ J,
                                                           const VT *__restrict__ V, XV xv,
                                                           const RowBatch<T, VT, U, I4> *pre = nullptr) {
  typedef typename Vec2<T>::type T2;
  const int sl = lane_id() % G;
  const int n4 = (cnt + 3) >> 2;
  T a = 0, b = 0, c = 0, d = 0;
  auto batch = [&](const RowBatch<T, VT, U, I4> &r) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const T2 x0 = xv(r.jj[u].x), x1 = xv(r.jj[u].y), x2 = xv(r.jj[u].z), x3 = xv(r.jj[u].w);
      a += r.vv[u].v[0] * x0.x; b += r.vv[u].v[0] * x0.y;
      c += r.vv[u].v[1] * x1.x; d += r.vv[u].v[1] * x1.y;
      a += r.vv[u].v[2] * x2.x; b += r.vv[u].v[2] * x2.y;
      c += r.vv[u].v[3] * x3.x; d += r.vv[u].v[3] * x3.y;
    }
  };
  int g = sl;
  if (PRE && g + (U - 1) * G < n4) {
    batch(*pre);
    g += U * G;
  }
  for (; g + (U - 1) * G < n4; g += U * G) {
    RowBatch<T, VT, U, I4> r;
#pragma unroll
    for (int u = 0; u < U; u++) r.jj[u] = J[g + u * G];
#pragma unroll
    for (int u = 0; u < U; u++) r.vv[u].load(V + 4 * (g + u * G));
    batch(r);
  }
  for (; g < n4; g += G) {
    const I4 jj = J[g];
    Quad<T, VT> vv;
    vv.load(V + 4 * g);
    const T2 x0 = xv(jj.x), x1 = xv(jj.y), x2 = xv(jj.z), x3 = xv(jj.w);
    a += vv.v[0] * x0.x; b += vv.v[0] * x0.y;
    c += vv.v[1] * x1.x; d += vv.v[1] * x1.y;
    a += vv.v[2] * x2.x; b += vv.v[2] * x2.y;
    c += vv.v[3] * x3.x; d += vv.v[3] * x3.y;
  }
  T s = a + c, t = b + d;
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, SWR_WAVE);
    t += __shfl_xor(t, o, SWR_WAVE);
  }
  T2 r;
  r.x = s;
  r.y = t;
  return r;
}

// The SpMV panels of swr_system (panel.hip): rows in panel order, panel
// p's rows [pbeg[p], pbeg[p+1]) of prow, its column list pcol[p][0, pcnt[p])
// (pcnt < 0: none) and each entry's position in it (jloc).
struct PanelArgs {
  const int *prow, *pbeg, *pcnt, *pcol;
  const unsigned short *jloc;
};
inline PanelArgs panel_args(const swr_system *S) { return PanelArgs{S->prow, S->pbeg, S->pcnt, S->pcol, S->jloc}; }

// One workgroup of NT threads per panel: the panel's entries of the vector
// (src(j) returns its entry j) are copied into the LDS list xl once (when
// the panel has a list), then G lanes per row sum each row of the panel,
// NT/G rows at a time, and call row(i, live, q) on every lane of the row's
// group (live: a row of the panel; q: the row's sums).  Panels without a
// list gather through src.
//
// With a list, the fill's first loads (FU list entries per thread at once),
// the first row pass's set-up and its first U groups of H values are all
// requested before the list is written, in that order (vmcnt retires in
// issue order, so writing those entries waits only for the fill's loads),
// and the barrier orders LDS only: the values stream from HBM while the
// list fills.  FU = 2 keeps the default shape within 128 registers (four
// waves per SIMD).
template <typename T, typename VT, int NT, int G, int U, class Src, class Row>
__device__ __forceinline__ void panel_rows(int p, const PanelArgs &A, const int *__restrict__ firstnbr,
                                           const int *__restrict__ numnbrs, const int *__restrict__ jlist,
                                           const VT *__restrict__ val, Src src, typename Vec2<T>::type *xl,
                                           Row row) {
  typedef typename Vec2<T>::type T2;
  const int r0 = A.pbeg[p], r1 = A.pbeg[p + 1];
  const int m = A.pcnt[p];
  const int grp = threadIdx.x / G;
  int rb = r0;
  if (m >= 0) {
    constexpr int FU = 2;
    const int *pc = A.pcol + (size_t)p * SWR_PANEL_MAXC;
    int kc[FU];
#pragma unroll
    for (int u = 0; u < FU; u++) {
      const int k = threadIdx.x + u * NT;
      kc[u] = k < m ? pc[k] : 0;
    }
    const int rr = rb + grp;
    const bool live = rr < r1;
    const int i = live ? A.prow[rr] : 0;
    T2 xk[FU];
#pragma unroll
    for (int u = 0; u < FU; u++)
      if (threadIdx.x + u * NT < m) xk[u] = src(kc[u]);
    const int beg = live ? firstnbr[i] : 0, cnt = live ? numnbrs[i] : 0;
    const ushort4 *J = reinterpret_cast<const ushort4 *>(A.jloc + beg);
    RowBatch<T, VT, U, ushort4> b0;
    row_batch0<T, VT, G, U>(cnt, J, val + beg, b0);
#pragma unroll
    for (int u = 0; u < FU; u++)
      if (threadIdx.x + u * NT < m) xl[threadIdx.x + u * NT] = xk[u];
    for (int k = threadIdx.x + FU * NT; k < m; k += NT) xl[k] = src(pc[k]);
    lds_barrier();
    const T2 q = row_sum4<T, VT, G, U, true>(cnt, J, val + beg, [&](int k) { return xl[k]; }, &b0);
    row(i, live, q);
    rb += NT / G;
  }
  for (; rb < r1; rb += NT / G) {
    const int rr = rb + grp;
    const bool live = rr < r1;
    const int i = live ? A.prow[rr] : 0;
    const int beg = live ? firstnbr[i] : 0, cnt = live ? numnbrs[i] : 0;
    T2 q;
    if (m >= 0)
      q = row_sum4<T, VT, G, U>(cnt, reinterpret_cast<const ushort4 *>(A.jloc + beg), val + beg,
                                [&](int k) { return xl[k]; });
    else   // (rare: one group in flight keeps the kernel's registers those of the list path)
      q = row_sum4<T, VT, G, 1>(cnt, reinterpret_cast<const int4 *>(jlist + beg), val + beg, src);
    row(i, live, q);
  }
}

// SpMV shapes the autotuner chooses from: NT threads per panel, G lanes
// per row, U groups of 4 entries in flight per lane.
enum { SPMV_T512G8U4 = 0, SPMV_T512G8U2, SPMV_T256G8U4, SPMV_T256G4U4, SPMV_NSHAPES };
constexpr int SPMV_DEFAULT = SPMV_T512G8U4;
template <int SH> struct SpmvShape;
template <> struct SpmvShape<SPMV_T512G8U4> { static constexpr int NT = 512, G = 8, U = 4; };
template <> struct SpmvShape<SPMV_T512G8U2> { static constexpr int NT = 512, G = 8, U = 2; };
template <> struct SpmvShape<SPMV_T256G8U4> { static constexpr int NT = 256, G = 8, U = 4; };
template <> struct SpmvShape<SPMV_T256G4U4> { static constexpr int NT = 256, G = 4, U = 4; };

// Calls f(std::integral_constant<int, SH>) for the runtime shape sh.
template <class F>
__host__ inline void spmv_dispatch(int sh, F &&f) {
  switch (sh) {
    case SPMV_T512G8U2: f(std::integral_constant<int, SPMV_T512G8U2>()); break;
    case SPMV_T256G8U4: f(std::integral_constant<int, SPMV_T256G8U4>()); break;
    case SPMV_T256G4U4: f(std::integral_constant<int, SPMV_T256G4U4>()); break;
    default: f(std::integral_constant<int, SPMV_T512G8U4>()); break;
  }
}

template <typename T2, typename T>
__device__ __forceinline__ T2 mk2(T a, T b) {
  T2 r;
  r.x = a;
  r.y = b;
  return r;
}

// Whether a system's solves still run in phase PH (PH_A1, PH_K read the
// scalars; c points at the system's s and t scalars).
__device__ __forceinline__ bool cg_any_active(int PH, const CgScalars *__restrict__ c) {
  if (PH != PH_A1 && PH != PH_K) return true;
  // all four flags requested at once (no short-circuit chain of loads)
  const int ab = c[0].active_b | c[1].active_b, aa = c[0].active_a | c[1].active_a;
  return ((PH == PH_K ? ab : 0) | aa) != 0;
}

// q = eta d + H d (sparse_matvec_C_para, fix_qeq_reax_sw64.c:424-466) of the
// vector phase PH multiplies: x (PH_S1), d0 (PH_S2), d1 (PH_S3) or the
// current d (PH_K).  One workgroup per SpMV panel (shape SH).
template <int PH, bool BATCH, typename T, typename VT, int SH>
__global__ void __launch_bounds__(SpmvShape<SH>::NT)
k_cg_spmv(int n, const DevParams *__restrict__ P, const int *__restrict__ type, const int *__restrict__ firstnbr,
          const int *__restrict__ numnbrs, const int *__restrict__ jlist, const VT *__restrict__ val, CgVecsT<T> V,
          int par_in, const CgScalars *__restrict__ cgs, const int *__restrict__ row_sys, PanelArgs A) {
  typedef typename Vec2<T>::type T2;
  constexpr int NT = SpmvShape<SH>::NT, G = SpmvShape<SH>::G, U = SpmvShape<SH>::U;
  __shared__ T2 xl[SWR_PANEL_MAXC];
  const int p = blockIdx.x;
  // skipped once both solves of the panel's system have finished (a panel
  // holds rows of one system)
  if (!BATCH && !cg_any_active(PH, cgs)) return;
  if (BATCH) {
    const int r0 = A.pbeg[p];
    if (r0 < A.pbeg[p + 1] && !cg_any_active(PH, cgs + 2 * row_sys[A.prow[r0]])) return;
  }
  const T2 *src = PH == PH_S1 ? V.x : PH == PH_S2 ? V.d[0] : PH == PH_S3 ? V.d[1] : (par_in ? V.d[1] : V.d[0]);
  panel_rows<T, VT, NT, G, U>(p, A, firstnbr, numnbrs, jlist, val, [&](int j) { return src[j]; }, xl,
                              [&](int i, bool live, T2 q) {
    if (lane_id() % G == 0 && live) {
      const T eta = (T)P->eta[type[i]];
      const T2 xi = src[i];
      q.x = eta * xi.x + q.x;
      q.y = eta * xi.y + q.y;
      V.q[i] = q;
    }
  });
}

// The vector part of phase PH, one thread per row.  PH_S1..S3 are the setup
// steps of CG_v2 (:1020-1087), PH_A1 the first A step, PH_K the fused
// iteration.  xzero (PH_S1 only): x is the zero vector (written here), so
// r = b and no SpMV ran.
template <int PH, typename T>
__global__ void __launch_bounds__(UT)
k_cg_upd(CgVecsT<T> V, int par_in, const CgScalars *__restrict__ cgs, CgLayout L, double *__restrict__ partials,
         int pstride, int xzero) {
  typedef typename Vec2<T>::type T2;
  const int tile = blockIdx.x;
  const int sys = L.tile_sys[tile];
  CgScalars c0, c1;
  bool b0 = false, b1 = false, a0 = true, a1 = true;
  if (PH == PH_A1 || PH == PH_K) {
    c0 = cgs[2 * sys];
    c1 = cgs[2 * sys + 1];
    b0 = PH == PH_K && c0.active_b != 0;
    b1 = PH == PH_K && c1.active_b != 0;
    a0 = c0.active_a != 0;
    a1 = c1.active_a != 0;
    if (!b0 && !b1 && !a0 && !a1) return;  // both solves finished
  }
  const int i = L.tile_beg[tile] + threadIdx.x;
  double dots[6] = {0, 0, 0, 0, 0, 0};
  if (i < L.tile_beg[tile + 1]) {
    const T Hd = (T)V.Hd[i];
    const T2 q = (PH == PH_A1 || (PH == PH_S1 && xzero)) ? mk2<T2>((T)0, (T)0) : V.q[i];
    if (PH == PH_S1) {
      if (xzero) V.x[i] = mk2<T2>((T)0, (T)0);
      const T2 b = V.b[i];
      const T2 r = mk2<T2>(b.x - q.x, b.y - q.y);
      const T2 u = mk2<T2>(r.x * Hd, r.y * Hd);
      V.r[i] = r; V.u[i] = u; V.d[0][i] = u;
    } else if (PH == PH_S2) {
      V.w[i] = q;
      const T2 m = mk2<T2>(q.x * Hd, q.y * Hd);
      V.m[i] = m; V.d[1][i] = m;
    } else if (PH == PH_S3) {
      const T2 b = V.b[i], u = V.u[i], r = V.r[i], ww = V.w[i], m = V.m[i];
      V.z[i] = q; V.p[i] = u; V.ss[i] = ww; V.v[i] = m;
      dots[0] = (double)b.x * b.x; dots[1] = (double)u.x * r.x; dots[2] = (double)u.x * ww.x;
      dots[3] = (double)b.y * b.y; dots[4] = (double)u.y * r.y; dots[5] = (double)u.y * ww.y;
    } else {
      T2 xx = V.x[i], p = V.p[i], u = V.u[i], ss = V.ss[i], w2 = V.w[i], v = V.v[i], z = V.z[i], r = V.r[i];
      const T2 dc = PH == PH_K ? (par_in ? V.d[1] : V.d[0])[i] : mk2<T2>((T)0, (T)0);
      T2 dn = dc;
      // B: x += alpha p; p = u + p beta; ss = w + ss beta; v = d + v beta   (:1135-1141)
      // C: z = q + z beta                                                   (:1147-1150)
      if (PH == PH_K && b0) {
        const T al = (T)c0.alpha_x, be = (T)c0.beta;
        xx.x += al * p.x; p.x = u.x + p.x * be; ss.x = w2.x + ss.x * be; v.x = dc.x + v.x * be;
        z.x = q.x + z.x * be;
      }
      if (PH == PH_K && b1) {
        const T al = (T)c1.alpha_x, be = (T)c1.beta;
        xx.y += al * p.y; p.y = u.y + p.y * be; ss.y = w2.y + ss.y * be; v.y = dc.y + v.y * be;
        z.y = q.y + z.y * be;
      }
      // A of the next iteration: r -= alpha ss; u -= alpha v; w -= alpha z; d = M^-1 w (:1113-1123)
      if (a0) {
        const T al = (T)c0.alpha_a;
        r.x -= al * ss.x; u.x -= al * v.x; w2.x -= al * z.x;
        dots[0] = (double)u.x * r.x; dots[1] = (double)u.x * w2.x;
        dn.x = w2.x * Hd;
      }
      if (a1) {
        const T al = (T)c1.alpha_a;
        r.y -= al * ss.y; u.y -= al * v.y; w2.y -= al * z.y;
        dots[2] = (double)u.y * r.y; dots[3] = (double)u.y * w2.y;
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
  block_partials<PH == PH_S3 ? 6 : 4, UT>(dots, partials, pstride, tile);
}

// Scalars of the iteration whose dots k_cg_upd<PH> just wrote: one
// workgroup per system, over the system's tiles.  tol > 0 sets the solves'
// tolerance at setup; tol <= 0 keeps the values already in cgs[].tol (a
// negative one marks a solve with nothing to do).  A finished solve's
// iteration count goes to iters_out[2 sys + rhs]; a system whose two solves
// have both finished adds one to L.done_count.
template <int PH>
__global__ void __launch_bounds__(RT)
k_cg_scalars(const double *__restrict__ partials, int pstride, CgScalars *__restrict__ cgs, CgLayout L, int imax,
             double tol, DevStatus *st, int *__restrict__ iters_out) {
  const int sys = blockIdx.x;
  CgScalars c0 = cgs[2 * sys], c1 = cgs[2 * sys + 1];
  const bool b0 = PH == PH_K && c0.active_b != 0, b1 = PH == PH_K && c1.active_b != 0;
  const bool a0 = PH == PH_S3 || c0.active_a != 0, a1 = PH == PH_S3 || c1.active_a != 0;
  if (!b0 && !b1 && !a0 && !a1) return;
  const bool was_done = PH != PH_S3 && c0.done && c1.done;
  constexpr int nd = PH == PH_S3 ? 6 : 4;
  double tot[nd];
  reduce_partials<nd>(partials, pstride, L.sys_tile[sys], L.sys_tile[sys + 1], tot);
  if (threadIdx.x != 0) return;
  unsigned int bad = 0;
  int *its = iters_out + 2 * sys;
  if constexpr (PH == PH_S3) {
    auto setup = [&](CgScalars &cc, double bb, double ur, double uw, int s) {
      if (tol > 0) cc.tol = tol;
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
      cc.active_a = (1 < imax && cc.tol > 0 && cc.b_norm > 0 && sqrt(cc.dg0) / cc.b_norm > cc.tol) ? 1 : 0;
      cc.done = cc.active_a ? 0 : 1;
      if (cc.done) its[s] = cc.iter;
      if (!isfinite(cc.alpha_a) && cc.active_a) bad = 1;
    };
    setup(c0, tot[0], tot[1], tot[2], 0);
    setup(c1, tot[3], tot[4], tot[5], 1);
  } else {
    auto fin = [&](CgScalars &cc, bool act_a, bool act_b, double l0, double l1, int s) {
      if (act_a) {
        cg_after_A(cc, l0, l1, imax);
        if (!isfinite(cc.alpha_a) || !isfinite(cc.beta)) bad = 1;
      } else if (act_b) {
        cc.active_b = 0;
        cc.done = 1;
        its[s] = cc.iter;
      }
    };
    fin(c0, a0, b0, tot[0], tot[1], 0);
    fin(c1, a1, b1, tot[2], tot[3], 1);
  }
  cgs[2 * sys] = c0;
  cgs[2 * sys + 1] = c1;
  if (!was_done && c0.done && c1.done) atomicAdd(L.done_count, 1);
  if (bad) atomicOr(&st->flags, SWR_F_NAN);
}

// Everything one CG solve needs from the system.
template <typename T, typename VT>
struct CgRun {
  CgVecsT<T> V;
  const VT *val;
  CgLayout L;
  CgScalars *cgs;     // 2 per system (s, t)
  int *iters_out;     // 2 ints per system on the device
  int *iters_host;    // when set: the host copy, refreshed at every done check
  int *pin_done;      // when set: pinned host int and iteration counts the
  int *pin_iters;     // done checks copy into (then iters_host from pin_iters)
  double tol;         // > 0: set at setup; <= 0: preset in cgs[].tol
  int imax;           // iteration limit (the reference's imax = 200 for the FP64 solve)
  int kt_iter, kt_update, kt_setup, kt_scalars;
  double *partials;   // per-workgroup partial dots, [8][pstride]
  int pstride;
};

template <typename T, typename VT>
int launch_cg(swr_system *S, const CgRun<T, VT> &R, int ph, int par_in, int xzero) {
  const int n = S->n;
  const int ps = R.pstride;
  const int kt = ph == PH_K ? R.kt_iter : R.kt_setup;
  const int ktu = ph == PH_K ? R.kt_update : R.kt_setup;
  const int kts = ph == PH_K ? R.kt_scalars : R.kt_setup;
#define SP_ARGS n, S->dp, S->type, S->H_firstnbr, S->H_numnbrs, S->H_jlist, R.val, R.V, par_in, R.cgs, R.L.row_sys, \
                panel_args(S)
#define UP_ARGS R.V, par_in, R.cgs, R.L, R.partials, ps, xzero
#define SC_ARGS R.partials, ps, R.cgs, R.L, R.imax, R.tol, S->status, R.iters_out
  swr_kt_begin(S, kt);
  spmv_dispatch(S->spmv_shape[0], [&](auto shc) {
    constexpr int SH = decltype(shc)::value;
    const dim3 gb(S->npanel), tb(SpmvShape<SH>::NT);
    if (R.L.nsys > 1) {
      switch (ph) {
        case PH_S1:
          if (!xzero) hipLaunchKernelGGL((k_cg_spmv<PH_S1, true, T, VT, SH>), gb, tb, 0, S->stream, SP_ARGS);
          break;
        case PH_S2: hipLaunchKernelGGL((k_cg_spmv<PH_S2, true, T, VT, SH>), gb, tb, 0, S->stream, SP_ARGS); break;
        case PH_S3: hipLaunchKernelGGL((k_cg_spmv<PH_S3, true, T, VT, SH>), gb, tb, 0, S->stream, SP_ARGS); break;
        case PH_A1: break;
        default: hipLaunchKernelGGL((k_cg_spmv<PH_K, true, T, VT, SH>), gb, tb, 0, S->stream, SP_ARGS); break;
      }
    } else {
      switch (ph) {
        case PH_S1:
          if (!xzero) hipLaunchKernelGGL((k_cg_spmv<PH_S1, false, T, VT, SH>), gb, tb, 0, S->stream, SP_ARGS);
          break;
        case PH_S2: hipLaunchKernelGGL((k_cg_spmv<PH_S2, false, T, VT, SH>), gb, tb, 0, S->stream, SP_ARGS); break;
        case PH_S3: hipLaunchKernelGGL((k_cg_spmv<PH_S3, false, T, VT, SH>), gb, tb, 0, S->stream, SP_ARGS); break;
        case PH_A1: break;
        default: hipLaunchKernelGGL((k_cg_spmv<PH_K, false, T, VT, SH>), gb, tb, 0, S->stream, SP_ARGS); break;
      }
    }
  });
  swr_kt_end(S, kt);
  swr_kt_begin(S, ktu);
  const dim3 gt(R.L.ntiles);
  switch (ph) {
    case PH_S1: hipLaunchKernelGGL((k_cg_upd<PH_S1, T>), gt, dim3(UT), 0, S->stream, UP_ARGS); break;
    case PH_S2: hipLaunchKernelGGL((k_cg_upd<PH_S2, T>), gt, dim3(UT), 0, S->stream, UP_ARGS); break;
    case PH_S3: hipLaunchKernelGGL((k_cg_upd<PH_S3, T>), gt, dim3(UT), 0, S->stream, UP_ARGS); break;
    case PH_A1: hipLaunchKernelGGL((k_cg_upd<PH_A1, T>), gt, dim3(UT), 0, S->stream, UP_ARGS); break;
    default: hipLaunchKernelGGL((k_cg_upd<PH_K, T>), gt, dim3(UT), 0, S->stream, UP_ARGS); break;
  }
  swr_kt_end(S, ktu);
  swr_kt_begin(S, kts);
  const dim3 gs(R.L.nsys);
  switch (ph) {
    case PH_S1: case PH_S2: break;
    case PH_S3:
      HIP_CHECK(hipMemsetAsync(R.L.done_count, 0, sizeof(int), S->stream));
      hipLaunchKernelGGL(k_cg_scalars<PH_S3>, gs, dim3(RT), 0, S->stream, SC_ARGS);
      break;
    case PH_A1: hipLaunchKernelGGL(k_cg_scalars<PH_A1>, gs, dim3(RT), 0, S->stream, SC_ARGS); break;
    default: hipLaunchKernelGGL(k_cg_scalars<PH_K>, gs, dim3(RT), 0, S->stream, SC_ARGS); break;
  }
  swr_kt_end(S, kts);
#undef SP_ARGS
#undef UP_ARGS
#undef SC_ARGS
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Runs the solves of every system of R to convergence.  Iteration kernels
// are enqueued in batches (the first of first_batch launches); between
// batches the host reads how many systems have finished.  Launches enqueued
// after a system finished skip its rows.
template <typename T, typename VT>
int cg_solve(swr_system *S, const CgRun<T, VT> &R, int first_batch, int xzero) {
  int rc;
  if ((rc = launch_cg(S, R, PH_S1, 0, xzero)) || (rc = launch_cg(S, R, PH_S2, 0, 0)) ||
      (rc = launch_cg(S, R, PH_S3, 0, 0)) || (rc = launch_cg(S, R, PH_A1, 0, 0)))
    return rc;
  int k = 1;
  int batch = first_batch > 2 ? first_batch : 2;
  while (k < R.imax + 1) {
    for (int b = 0; b < batch && k < R.imax + 1; b++, k++)
      if ((rc = launch_cg(S, R, PH_K, (k - 1) & 1, 0))) return rc;
    // one round trip: done count, status block and iteration counts (into
    // pinned memory when the run has it)
    int done = 0;
    int *dh = R.pin_done ? R.pin_done : &done;
    int *ih = R.pin_iters ? R.pin_iters : R.iters_host;
    HIP_CHECK(hipMemcpyAsync(dh, R.L.done_count, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipMemcpyAsync(S->status_host, S->status, sizeof(DevStatus), hipMemcpyDeviceToHost, S->stream));
    if (ih)
      HIP_CHECK(hipMemcpyAsync(ih, R.iters_out, sizeof(int) * 2 * R.L.nsys, hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
    done = *dh;
    if (R.iters_host && ih != R.iters_host) std::memcpy(R.iters_host, ih, sizeof(int) * 2 * R.L.nsys);
    swr_after_sync(S);
    if (done >= R.L.nsys) break;
    batch = 4;
  }
  return SWR_OK;
}

}  // namespace swr_cg

// The row/tile layout of S's systems for the CG kernels (qeq.hip).
swr_cg::CgLayout swr_qeq_cg_layout(swr_system *S);
