// panel.hip — SpMV panels of the QEq matrix H.
//
// The local rows of H are ordered along a Morton curve over the neighbour
// list's bins (grouped by system in a batch) and cut into panels of at most
// SWR_PANEL_ROWS rows of one system.  Per panel, the list of its distinct
// columns goes to pcol and each H entry gets the index of its column in that
// list (jloc, 16 bits).  The SpMV kernels (qeq_cg.h, qeq_mixed.hip) load a
// panel's entries of x into LDS once and gather from there: a compact group
// of atoms shares most of its neighbours (a 64-row panel of a 96,000-atom
// water box touches ~1,400 distinct columns for ~27,000 entries), so the
// per-entry gathers leave the L2 and each entry streams a 2-byte index
// instead of the 4-byte column.  A panel with more than SWR_PANEL_MAXC
// columns keeps pcnt = -1 and its rows gather from global memory as before.
//
// The row order changes only when the bins or the batch partition change
// (a neighbour-list build or swr_set_batch); the column lists are rebuilt
// with H every step.  Dense indices follow the hash table's slot order,
// which can differ between runs, but a row's sum runs over its entries in
// H order whatever the indices are, so the SpMV results do not depend on it.
#include <vector>

#include "swr_internal.h"
#include "swr_device.h"

namespace {

constexpr int PB_T = 1024;                // threads of the panel build
constexpr int PB_HASH_BITS = 13;
constexpr int PB_HASH = 1 << PB_HASH_BITS;  // LDS hash slots
static_assert(SWR_PANEL_MAXC + PB_T < PB_HASH, "the hash table never fills");
static_assert(SWR_PANEL_MAXC <= 65536, "16-bit local indices");

__device__ __forceinline__ unsigned spread10(unsigned v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Morton key of each local atom's bin (10 bits per axis; larger grids fold
// onto themselves, which only makes the order less compact).
__global__ void k_morton_keys(int n, const int *__restrict__ bin_of, int nx, int ny, int *__restrict__ key) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = bin_of[i];
  const unsigned bx = (unsigned)(b % nx), by = (unsigned)((b / nx) % ny), bz = (unsigned)(b / (nx * ny));
  key[i] = (int)(spread10(bx) | (spread10(by) << 1) | (spread10(bz) << 2));
}

// The system of each row of an order (the key of the stable second sort).
__global__ void k_sys_keys(int n, const int *__restrict__ order, const int *__restrict__ row_sys,
                           int *__restrict__ key) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  key[k] = row_sys[order[k]];
}

__global__ void k_panels_off(int np, int *__restrict__ pcnt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) pcnt[p] = -1;
}

__device__ __forceinline__ unsigned pb_hash(int col) {
  return ((unsigned)col * 2654435761u) >> (32 - PB_HASH_BITS);
}

// One workgroup per panel: every column of its rows' entries (padding
// entries included) into an LDS hash set, stopping once the panel has more
// distinct columns than its list holds, each entry's slot noted in jloc;
// then dense indices in slot order, the list, and each entry's slot
// replaced by its index (by the thread that wrote it).  The rows' extents
// are staged in LDS first and each lane requests PB_U columns before using
// any, so the global loads overlap instead of waiting one at a time.
constexpr int PB_U = 8;   // columns in flight per lane
__global__ void __launch_bounds__(PB_T) k_panel_build(const int *__restrict__ prow, const int *__restrict__ pbeg,
                                                      const int *__restrict__ firstnbr,
                                                      const int *__restrict__ numnbrs,
                                                      const int *__restrict__ jlist, int *__restrict__ pcol,
                                                      int *__restrict__ pcnt, unsigned short *__restrict__ jloc) {
  constexpr int NW = PB_T / SWR_WAVE;
  __shared__ int hkey[PB_HASH];
  __shared__ unsigned short hid[PB_HASH];
  __shared__ int rbeg[SWR_PANEL_ROWS], rlen[SWR_PANEL_ROWS];
  __shared__ int wsum[NW];
  __shared__ int ninserted;
  const int p = blockIdx.x;
  const int r0 = pbeg[p], nr = pbeg[p + 1] - r0;   // nr <= SWR_PANEL_ROWS
  for (int k = threadIdx.x; k < PB_HASH; k += PB_T) hkey[k] = -1;
  if (threadIdx.x < nr) {
    const int i = prow[r0 + threadIdx.x];
    rbeg[threadIdx.x] = firstnbr[i];
    rlen[threadIdx.x] = (numnbrs[i] + 3) & ~3;
  }
  if (threadIdx.x == 0) ninserted = 0;
  __syncthreads();
  const int w = wave_in_block(), lane = lane_id();
  bool over = false;
  for (int r = w; r < nr && !over; r += NW) {
    const int beg = rbeg[r], n4 = rlen[r];
    for (int e0 = 0; e0 < n4 && !over; e0 += PB_U * SWR_WAVE) {
      int col[PB_U];
#pragma unroll
      for (int u = 0; u < PB_U; u++) {
        const int e = e0 + u * SWR_WAVE + lane;
        col[u] = e < n4 ? jlist[beg + e] : -1;
      }
#pragma unroll
      for (int u = 0; u < PB_U; u++) {
        if (col[u] < 0) continue;
        // checked before every insert: at most SWR_PANEL_MAXC + PB_T keys
        // are ever inserted, so a free slot or the key is always found
        if (ninserted > SWR_PANEL_MAXC) {   // too many: the panel gathers from global memory
          over = true;
          break;
        }
        unsigned h = pb_hash(col[u]);
        for (int t = 0; t < PB_HASH; t++) {
          const int old = atomicCAS(&hkey[h], -1, col[u]);
          if (old == -1) {
            atomicAdd(&ninserted, 1);
            break;
          }
          if (old == col[u]) break;
          h = (h + 1) & (PB_HASH - 1);
        }
        jloc[beg + e0 + u * SWR_WAVE + lane] = (unsigned short)h;
      }
    }
  }
  __syncthreads();
  if (ninserted > SWR_PANEL_MAXC) {
    if (threadIdx.x == 0) pcnt[p] = -1;
    return;
  }
  // dense indices in slot order: per-thread counts, a block scan, then ids
  constexpr int PER = PB_HASH / PB_T;
  const int s0 = threadIdx.x * PER;
  int c = 0;
  for (int s = s0; s < s0 + PER; s++) c += hkey[s] >= 0 ? 1 : 0;
  const int incl = wave_incl_scan(c);
  if (lane == SWR_WAVE - 1) wsum[w] = incl;
  __syncthreads();
  int id = incl - c;
  for (int v = 0; v < w; v++) id += wsum[v];
  int *pc = pcol + (size_t)p * SWR_PANEL_MAXC;
  for (int s = s0; s < s0 + PER; s++) {
    const int key = hkey[s];
    if (key >= 0) {
      hid[s] = (unsigned short)id;
      pc[id] = key;
      id++;
    }
  }
  __syncthreads();
  // every entry's slot (written above by this same thread) to its index
  for (int r = w; r < nr; r += NW) {
    const int beg = rbeg[r], n4 = rlen[r];
    for (int e0 = 0; e0 < n4; e0 += PB_U * SWR_WAVE) {
      unsigned short sl[PB_U];
#pragma unroll
      for (int u = 0; u < PB_U; u++) {
        const int e = e0 + u * SWR_WAVE + lane;
        sl[u] = e < n4 ? jloc[beg + e] : 0;
      }
#pragma unroll
      for (int u = 0; u < PB_U; u++) {
        const int e = e0 + u * SWR_WAVE + lane;
        if (e < n4) jloc[beg + e] = hid[sl[u]];
      }
    }
  }
  if (threadIdx.x == 0) pcnt[p] = ninserted;
}

}  // namespace

int swr_panel_build(swr_system *S) {
  const int n = S->n;
  int rc;
  if ((rc = swr_qeq_layout(S))) return rc;
  if (n <= 0) return SWR_OK;
  if (S->prow_n != n || S->prow_nbr_gen != S->nbr_gen || S->prow_layout_gen != S->layout_gen) {
    if (!S->bin_of) {
      swr_set_error_msg("SpMV panels need the neighbour list's bins (build the neighbour list first)");
      return SWR_ERR_STATE;
    }
    if ((rc = swr_reserve(&S->prow, S->cap_prow, (size_t)n)) ||
        (rc = swr_reserve(&S->pkey, S->cap_pkey, 2 * (size_t)n)))
      return rc;
    swr_kt_begin(S, KT_PANEL);
    hipLaunchKernelGGL(k_morton_keys, dim3(div_up(n, 256)), dim3(256), 0, S->stream, n, (const int *)S->bin_of,
                       S->nbin[0], S->nbin[1], S->pkey);
    if (S->nsys <= 1) {
      if ((rc = swr_radix_sort_i32(S, S->pkey, nullptr, n, 30, S->prow))) return rc;
    } else {
      // Morton order, then stably by system: each system's rows together, in
      // Morton order, at the system's own row range
      int *tmp = S->pkey + n;
      if ((rc = swr_radix_sort_i32(S, S->pkey, nullptr, n, 30, tmp))) return rc;
      hipLaunchKernelGGL(k_sys_keys, dim3(div_up(n, 256)), dim3(256), 0, S->stream, n, (const int *)tmp,
                         (const int *)S->row_sys, S->pkey);
      int bits = 1;
      while ((S->nsys - 1) >> bits) bits++;
      if ((rc = swr_radix_sort_i32(S, S->pkey, tmp, n, bits, S->prow))) return rc;
    }
    swr_kt_end(S, KT_PANEL);
    std::vector<int> pbeg;
    const int nsys = (int)S->sys_row.size() - 1;
    std::vector<int> sys_panel(nsys + 1);
    for (int s = 0; s < nsys; s++) {
      sys_panel[s] = (int)pbeg.size();
      for (int r = S->sys_row[s]; r < S->sys_row[s + 1]; r += SWR_PANEL_ROWS) pbeg.push_back(r);
    }
    sys_panel[nsys] = (int)pbeg.size();
    pbeg.push_back(n);
    const int np = (int)pbeg.size() - 1;
    if ((rc = swr_reserve(&S->pbeg, S->cap_pbeg, pbeg.size())) || (rc = swr_reserve(&S->pcnt, S->cap_pcnt, (size_t)np)) ||
        (rc = swr_reserve(&S->sys_panel, S->cap_sys_panel, sys_panel.size())))
      return rc;
    HIP_CHECK(hipMemcpyAsync(S->pbeg, pbeg.data(), sizeof(int) * pbeg.size(), hipMemcpyHostToDevice, S->stream));
    HIP_CHECK(hipMemcpyAsync(S->sys_panel, sys_panel.data(), sizeof(int) * sys_panel.size(), hipMemcpyHostToDevice,
                             S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));   // host temporaries
    S->npanel = np;
    S->prow_n = n;
    S->prow_nbr_gen = S->nbr_gen;
    S->prow_layout_gen = S->layout_gen;
  }
  const int np = S->npanel;
  if ((rc = swr_reserve(&S->pcol, S->cap_pcol, (size_t)np * SWR_PANEL_MAXC)) ||
      (rc = swr_reserve(&S->jloc, S->cap_jloc, (size_t)std::max(S->Hcap, 1LL))))
    return rc;
  swr_kt_begin(S, KT_PANEL);
  if (S->panel_shape != 1)
    hipLaunchKernelGGL(k_panel_build, dim3(np), dim3(PB_T), 0, S->stream, (const int *)S->prow,
                       (const int *)S->pbeg, (const int *)S->H_firstnbr, (const int *)S->H_numnbrs,
                       (const int *)S->H_jlist, S->pcol, S->pcnt, S->jloc);
  else
    hipLaunchKernelGGL(k_panels_off, dim3(div_up(np, 256)), dim3(256), 0, S->stream, np, S->pcnt);
  swr_kt_end(S, KT_PANEL);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}
