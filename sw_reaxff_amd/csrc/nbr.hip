// nbr.hip — ghost images, cell binning and the full Verlet list (a1).
//
// The list replaces npair_full_bin_ghost_sunway_build_packed_para
// (npair_full_bin_ghost_sw5.c:80-226): for every atom i, local or ghost, all
// j != i with |x_i - x_j|^2 <= cutneigh^2, tested in FP64 with the
// reference's expression (del = x_i - x_j; rsq = delx^2+dely^2+delz^2,
// :183-193).  The set is the mathematically complete one: ghosts see every
// atom that exists within the cutoff (no bin-0 exclusion, :151).
//
// Design (MI355X): atoms are counting-sorted into bins of about cutneigh/2,
// each bin sorted by atom index so that the list order is deterministic.
// One wave owns one atom i and walks the stencil as (dy,dz) rows; along x
// the bins of a row are contiguous in the bin-sorted arrays, so each row is
// one contiguous segment that the 64 lanes stream through.  Hits are
// compacted with a wave ballot.  Pass 1 counts, a scan gives CSR offsets,
// pass 2 writes.
#include "swr_internal.h"
#include "swr_device.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr int NB_THREADS = 256;
constexpr int WAVES_PER_BLOCK = NB_THREADS / SWR_WAVE;

// ---------------------------------------------------------------- borders
// Comm::borders with one process: for dimension d, atoms [0,nlast) with
// x <= lo+cut are imaged at +L, then atoms with x >= hi-cut at -L.
__global__ void k_border_flag(const double *__restrict__ x, int nlast, int d, double thr, int side,
                              int *__restrict__ flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlast) return;
  double v = x[3 * i + d];
  flag[i] = side == 0 ? (v <= thr) : (v >= thr);
}

__global__ void k_border_scatter(double *__restrict__ x, double *__restrict__ q, int *__restrict__ type,
                                 int *__restrict__ tag, int *__restrict__ owner, signed char *__restrict__ img,
                                 int *__restrict__ num_bonds, int *__restrict__ num_hbonds,
                                 const int *__restrict__ flag, const int *__restrict__ pos, int nlast, int ncur,
                                 int d, double shift, int sgn) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlast || !flag[i]) return;
  int g = ncur + pos[i];
  for (int k = 0; k < 3; k++) x[3 * g + k] = x[3 * i + k];
  x[3 * g + d] = x[3 * i + d] + shift;
  q[g] = q[i];   // the charge travels with the border data, as atom style charge's does
  type[g] = type[i];
  tag[g] = tag[i];
  owner[g] = owner[i];
  for (int k = 0; k < 3; k++) img[3 * g + k] = img[3 * i + k];
  img[3 * g + d] = (signed char)(img[3 * i + d] + sgn);
  num_bonds[g] = num_bonds[owner[i]];
  num_hbonds[g] = num_hbonds[owner[i]];
}

// Forward comm of positions: ghost = owner + image shift (pbc add of +-L).
__global__ void k_forward_x(double *__restrict__ x, const int *__restrict__ owner,
                            const signed char *__restrict__ img, int n, int N, double Lx, double Ly, double Lz) {
  int g = n + blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  int o = owner[g];
  const double L[3] = {Lx, Ly, Lz};
  for (int d = 0; d < 3; d++) {
    double v = x[3 * o + d];
    int s = img[3 * g + d];
    // Images of images add one box length per crossing, as LAMMPS does.
    for (int k = 0; k < s; k++) v = v + L[d];
    for (int k = 0; k > s; k--) v = v - L[d];
    x[3 * g + d] = v;
  }
}

// Ghost refresh of capacities on re-neighbour steps (forward_comm_fix of
// fix REAXC, pair_reaxc_sunway.cpp:551).
__global__ void k_forward_caps(int *__restrict__ num_bonds, int *__restrict__ num_hbonds,
                               const int *__restrict__ owner, int n, int N) {
  int g = n + blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  num_bonds[g] = num_bonds[owner[g]];
  num_hbonds[g] = num_hbonds[owner[g]];
}

// NVE with the forces this path does not compute left at zero: x += dt v.
__global__ void k_nve(double *__restrict__ x, const double *__restrict__ v, int n, double dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n) return;
  x[i] = x[i] + dt * v[i];
}

// Domain::pbc remap of local atoms into [lo, hi).
__global__ void k_pbc(double *__restrict__ x, int n, double lox, double loy, double loz, double hix,
                      double hiy, double hiz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double lo[3] = {lox, loy, loz}, hi[3] = {hix, hiy, hiz};
  for (int d = 0; d < 3; d++) {
    double v = x[3 * i + d];
    double L = hi[d] - lo[d];
    if (v < lo[d]) v = v + L;
    if (v >= hi[d]) { v = v - L; v = v > lo[d] ? v : lo[d]; }
    x[3 * i + d] = v;
  }
}

__global__ void k_init_local_maps(int *__restrict__ owner, signed char *__restrict__ img, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  owner[i] = i;
  img[3 * i] = img[3 * i + 1] = img[3 * i + 2] = 0;
}

// ---------------------------------------------------------------- binning
__device__ __forceinline__ int coord_bin(double v, double lo, double inv, int nb, int *out_of_range) {
  double f = (v - lo) * inv;
  int b = (int)floor(f);
  if (b < 0) { if (b < -1 || f < -0.5) *out_of_range = 1; b = 0; }
  if (b >= nb) { if (b > nb || f > nb + 0.5) *out_of_range = 1; b = nb - 1; }
  return b;
}

// Bin of every atom (bin = (bz * ny + by) * nx + bx).
__global__ void k_bin_keys(const double *__restrict__ x, int N, double lx, double ly, double lz, double ix,
                           double iy, double iz, int nx, int ny, int nz, int *__restrict__ bin_of, DevStatus *st) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int oor = 0;
  int bx = coord_bin(x[3 * i], lx, ix, nx, &oor);
  int by = coord_bin(x[3 * i + 1], ly, iy, ny, &oor);
  int bz = coord_bin(x[3 * i + 2], lz, iz, nz, &oor);
  if (oor) atomicOr(&st->flags, SWR_F_BIN);
  bin_of[i] = (bz * ny + by) * nx + bx;
}

// ---- counting sort of the atoms by bin, staged in LDS
// A stable LSD radix sort of (bin, atom) pairs with 8-bit digits: per pass
// every block takes RS_CHUNK consecutive pairs (in the previous pass's
// order; atom index order in the first), counts their digits in an LDS
// histogram (rs_hist: hist[d * nblk + blk]), an exclusive scan over
// (digit, block) gives each block's start per digit, and rs_scatter writes
// each pair at its block's start for its digit plus its rank among the
// block's earlier pairs with that digit, ranked in LDS in input order (a
// wave ranks its lanes by matching digit bits with ballots; waves and
// rounds are added in order).  Stability makes the atoms of each bin come
// out in ascending index order, with no atomics on global memory, so the
// binning is the same on every run.
constexpr int RS_T = 256;                   // threads per block
constexpr int RS_ROUNDS = 4;                // pairs per thread
constexpr int RS_CHUNK = RS_T * RS_ROUNDS;  // pairs per block
constexpr int RS_RADIX = 256;

__global__ void __launch_bounds__(RS_T) k_rs_hist(const int *__restrict__ keys, int N, int shift, int nblk,
                                                  int *__restrict__ hist) {
  __shared__ int h[RS_RADIX];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * RS_CHUNK;
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; r++) {
    const int k = base + r * RS_T + threadIdx.x;
    if (k < N) atomicAdd(&h[(keys[k] >> shift) & (RS_RADIX - 1)], 1);
  }
  __syncthreads();
  hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// vals_in == nullptr: the values are the input positions (atom indices).
__global__ void __launch_bounds__(RS_T) k_rs_scatter(const int *__restrict__ keys_in, const int *__restrict__ vals_in,
                                                     int N, int shift, int nblk, const int *__restrict__ off,
                                                     int *__restrict__ keys_out, int *__restrict__ vals_out) {
  constexpr int NW = RS_T / SWR_WAVE;
  __shared__ int base[RS_RADIX];       // this block's next slot per digit
  __shared__ int wcnt[NW][RS_RADIX];   // per-wave digit counts of the current round
  const int w = wave_in_block(), lane = lane_id();
  base[threadIdx.x] = off[threadIdx.x * nblk + blockIdx.x];
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (SWR_WAVE - lane));
  for (int r = 0; r < RS_ROUNDS; r++) {
#pragma unroll
    for (int d = lane; d < RS_RADIX; d += SWR_WAVE) wcnt[w][d] = 0;
    const int k = blockIdx.x * RS_CHUNK + r * RS_T + threadIdx.x;
    const bool valid = k < N;
    const int key = valid ? keys_in[k] : 0;
    const int val = valid ? (vals_in ? vals_in[k] : k) : 0;
    const int dig = (key >> shift) & (RS_RADIX - 1);
    // lanes of this wave holding the same digit
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const unsigned long long m = __ballot((dig >> b) & 1);
      peers &= ((dig >> b) & 1) ? m : ~m;
    }
    const int rank = __popcll(peers & below);
    __syncthreads();   // wcnt zeroed, base of the previous round final
    if (valid && rank == 0) wcnt[w][dig] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = base[dig] + rank;
      for (int v = 0; v < w; v++) pos += wcnt[v][dig];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int v = 0; v < NW; v++) add += wcnt[v][threadIdx.x];
    base[threadIdx.x] += add;
    __syncthreads();
  }
}

// Atoms per bin from the sorted keys: each run of equal keys writes its
// length (bins without atoms keep the zero they were cleared to) and
// appends its bin to the occupied-bin list occ (occ[0] counts them; the
// Verlet build takes one wave per entry, and the rows it writes do not
// depend on the entries' order).
__global__ void k_bin_runs(const int *__restrict__ skeys, int N, int *__restrict__ bin_count, int *__restrict__ occ) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int key = skeys[k];
  if (k > 0 && skeys[k - 1] == key) return;
  int e = k + 1;
  while (e < N && skeys[e] == key) e++;
  bin_count[key] = e - k;
  occ[1 + atomicAdd(&occ[0], 1)] = key;   // the occupied bins, in no particular order
}

// Positions of the sorted atoms (x, y, z, unused).
__global__ void k_bin_pos(int N, const int *__restrict__ bp_id, const double *__restrict__ x,
                          double4 *__restrict__ bp_pos) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int j = bp_id[k];
  bp_pos[k] = make_double4(x[3 * j], x[3 * j + 1], x[3 * j + 2], 0.0);
}

// ---------------------------------------------------------------- Verlet
// The 64-bit value v of lane l (l uniform across the wave).
__device__ __forceinline__ long long readlane_ll(long long v, int l) {
  const int lo = __builtin_amdgcn_readlane((int)v, l), hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return ((long long)hi << 32) | (unsigned)lo;
}

// The full Verlet list, row i = every j != i within cutneigh, in stencil
// order (stencil rows in order, each row's candidates in bin order: the
// list one wave per atom walking its stencil would write), built one wave
// per occupied bin: the atoms of a bin share the bin's stencil, so each candidate is
// loaded once per bin, one per lane, and tested against every atom of the
// bin in turn (64 atoms at a time).  The atom's position and id are
// wave-uniform (scalar loads, one atom ahead), its row start and running
// count live in lane a % 64 of oa and ca; lanes past a row's last
// candidate test a point at infinity (no branch).  The candidates' cache
// traffic is that of one wave per atom divided by the bin occupancy (~20
// atoms in water at half-cutoff bins).
template <bool WRITE>
__global__ void __launch_bounds__(NB_THREADS)
k_verlet_bins(const int *__restrict__ occ, const int *__restrict__ bin_start, const double4 *__restrict__ bp_pos,
              const int *__restrict__ bp_id, const int4 *__restrict__ rows, int nrows, int nx, int ny, int nz,
              double cutsq, int *__restrict__ vcount, const long long *__restrict__ voff, int *__restrict__ vlist,
              int nrow) {
  // the grid is sized for the worst case (every atom in its own bin); the
  // blocks that hold occupied bins are spread over the XCDs by their own
  // count, so each XCD takes a contiguous share of them
  const int nocc = occ[0];
  const int nact = (nocc + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
  if ((int)blockIdx.x >= nact) return;
  const int t = xcd_block(nact) * WAVES_PER_BLOCK + wave_in_block();
  if (t >= nocc) return;
  const int b = occ[1 + t];   // an occupied bin (a mostly empty grid, as a batch's, costs nothing)
  const int lane = lane_id();
  const int a0 = __builtin_amdgcn_readfirstlane(bin_start[b]), a1 = __builtin_amdgcn_readfirstlane(bin_start[b + 1]);
  const int bx = b % nx, by = (b / nx) % ny, bz = b / (nx * ny);
  for (int g = a0; g < a1; g += SWR_WAVE) {
    const int ma = min(SWR_WAVE, a1 - g);
    int ca = 0;   // lane a: the running count of atom g + a
    long long oa = 0;   // lane a: its row start
    if (WRITE && lane < ma) {
      const int ia = bp_id[g + lane];
      if (ia < nrow) oa = voff[ia];
    }
    for (int r = 0; r < nrows; r++) {
      const int4 rw = rows[r];
      const int cy = by + rw.x, cz = bz + rw.y;
      if (cy < 0 || cy >= ny || cz < 0 || cz >= nz) continue;
      const int cx0 = max(bx - rw.z, 0), cx1 = min(bx + rw.z, nx - 1);
      const int rowbase = (cz * ny + cy) * nx;
      const int s = bin_start[rowbase + cx0], e = bin_start[rowbase + cx1 + 1];
      for (int base = s; base < e; base += SWR_WAVE) {
        const int k = base + lane;
        int j = -1;
        double px = HUGE_VAL, py = HUGE_VAL, pz = HUGE_VAL;
        if (k < e) {
          j = bp_id[k];
          const double4 p = bp_pos[k];
          px = p.x;
          py = p.y;
          pz = p.z;
        }
        // atom a + 1's id and position are requested while atom a is tested
        // (uniform: scalar loads)
        int inx = bp_id[g];
        double4 pnx = bp_pos[g];
        for (int a = 0; a < ma; a++) {
          const int i = inx;
          const double4 pa = pnx;
          if (a + 1 < ma) {
            inx = bp_id[g + a + 1];
            pnx = bp_pos[g + a + 1];
          }
          if (i >= nrow) continue;   // rows past nrow stay empty (the local-atom-only list)
          const double dx = pa.x - px, dy = pa.y - py, dz = pa.z - pz;
          const double rsq = dx * dx + dy * dy + dz * dz;
          const bool hit = (j != i) && (rsq <= cutsq);
          const unsigned long long m = __ballot(hit);
          if (WRITE && hit) vlist[readlane_ll(oa, a) + __builtin_amdgcn_readlane(ca, a) + ballot_rank(m)] = j;
          if (lane == a) ca += __popcll(m);
        }
      }
    }
    if (!WRITE && lane < ma) {
      const int ia = bp_id[g + lane];
      vcount[ia] = ia < nrow ? ca : 0;
    }
  }
}

}  // namespace

// Stencil rows for bins of edge bs[] and cutoff cut: every (dy,dz) whose
// nearest bin distance is within cut, with the x half-width sx reaching it.
static std::vector<int4> stencil_rows(const double bs[3], double cut) {
  std::vector<int4> rows;
  const int sy = (int)std::ceil(cut / bs[1]), sz = (int)std::ceil(cut / bs[2]);
  const int sxm = (int)std::ceil(cut / bs[0]);
  const double c2 = cut * cut;
  for (int dz = -sz; dz <= sz; dz++)
    for (int dy = -sy; dy <= sy; dy++) {
      double ey = std::max(std::abs(dy) - 1, 0) * bs[1], ez = std::max(std::abs(dz) - 1, 0) * bs[2];
      double rem = c2 - ey * ey - ez * ez;
      if (rem < 0) continue;
      int sx = std::min(sxm, (int)std::floor(std::sqrt(rem) / bs[0]) + 1);
      rows.push_back(make_int4(dy, dz, sx, 0));
    }
  return rows;
}

// ---------------------------------------------------------------- host API
namespace {
constexpr int BBOX_BLOCKS = 64;

// Per-block min and max of the atoms' coordinates (grid-stride over N).
__global__ void __launch_bounds__(256) k_bbox(const double *__restrict__ x, int N, double *__restrict__ part) {
  __shared__ double s[6][256];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256)
    for (int d = 0; d < 3; d++) {
      const double c = x[3 * (size_t)i + d];
      v[d] = fmin(v[d], c);
      v[3 + d] = fmax(v[3 + d], c);
    }
  for (int k = 0; k < 6; k++) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 6; k++)
        s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], s[k][threadIdx.x + o])
                                  : fmax(s[k][threadIdx.x], s[k][threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}
}  // namespace

// Bounding box of all N atoms (lo, hi on the host).
int swr_nbr_bbox(swr_system *S, double lo[3], double hi[3]) {
  int rc = swr_ensure_tmp(S, sizeof(double) * 6 * BBOX_BLOCKS);
  if (rc) return rc;
  double *part = (double *)S->tmp_i;
  double h[6 * BBOX_BLOCKS];
  hipLaunchKernelGGL(k_bbox, dim3(BBOX_BLOCKS), dim3(256), 0, S->stream, S->x, S->N, part);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, S->stream));
  HIP_CHECK(hipStreamSynchronize(S->stream));
  for (int d = 0; d < 3; d++) { lo[d] = 1e300; hi[d] = -1e300; }
  for (int b = 0; b < BBOX_BLOCKS; b++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], h[6 * b + d]);
      hi[d] = std::max(hi[d], h[6 * b + 3 + d]);
    }
  return SWR_OK;
}

int swr_nbr_set_grid(swr_system *S, const double lo[3], const double hi[3], double cutghost, double cutneigh) {
  // Grid covering [lo - cutghost, hi + cutghost] with bins of at least
  // cutneigh/2 (a flat or linear set of atoms gets one bin of that size
  // across its thin dimensions: the stencil then stays within +-2 bins).
  const double target = 0.5 * cutneigh;
  if (!(target > 0.0)) {
    swr_set_error_msg("neighbour cutoff must be positive");
    return SWR_ERR_ARG;
  }
  for (int d = 0; d < 3; d++) {
    double a = lo[d] - cutghost - 1e-6, b = hi[d] + cutghost + 1e-6;
    int nb = std::max(1, (int)std::floor((b - a) / target));
    S->binlo[d] = a;
    S->binsize[d] = std::max((b - a) / nb, target);
    S->bininv[d] = 1.0 / S->binsize[d];
    S->nbin[d] = nb;
  }
  S->cutneigh = cutneigh;
  int nb = S->nbin[0] * S->nbin[1] * S->nbin[2];
  if (nb + 1 > S->binscap) {
    int rc;
    if ((rc = swr_alloc(&S->bin_count, nb + 1)) || (rc = swr_alloc(&S->bin_start, nb + 2))) return rc;
    S->binscap = nb + 1;
  }
  S->nbins = nb;
  std::vector<int4> rows = stencil_rows(S->binsize, cutneigh);
  S->nrows = (int)rows.size();
  int rc = swr_reserve(&S->stencil_rows, S->cap_stencil, rows.size());
  if (rc) return rc;
  HIP_CHECK(hipMemcpyAsync(S->stencil_rows, rows.data(), rows.size() * sizeof(int4), hipMemcpyHostToDevice,
                           S->stream));
  HIP_CHECK(hipStreamSynchronize(S->stream));  // rows is a local
  return SWR_OK;
}

int swr_nbr_ensure_atoms(swr_system *S, int Ncap);

// Builds ghost images of the n local atoms (already remapped into the box).
int swr_nbr_borders(swr_system *S, const double lo[3], const double hi[3], double cutghost) {
  // one image per side and dimension (a single swap each way) covers the
  // ghost shell only when every box length is at least cutghost
  for (int d = 0; d < 3; d++)
    if (!(hi[d] - lo[d] >= cutghost)) {
      swr_set_error_msg("periodic box shorter than the ghost cutoff in a dimension");
      return SWR_ERR_ARG;
    }
  S->next_totals_known = false;  // new ghosts take their owners' capacities
  S->gcsr_valid = false;
  const int n = S->n;
  int ncur = n;
  hipLaunchKernelGGL(k_init_local_maps, dim3(div_up(n, 256)), dim3(256), 0, S->stream, S->owner, S->img, n);
  for (int d = 0; d < 3; d++) {
    const int nlast = ncur;
    for (int side = 0; side < 2; side++) {
      int rc = swr_ensure_tmp(S, sizeof(int) * (2 * (size_t)nlast + 2));
      if (rc) return rc;
      int *flag = S->tmp_i, *pos = S->tmp_i + nlast + 1;
      double thr = side == 0 ? lo[d] + cutghost : hi[d] - cutghost;
      hipLaunchKernelGGL(k_border_flag, dim3(div_up(nlast, 256)), dim3(256), 0, S->stream, S->x, nlast, d, thr,
                         side, flag);
      rc = swr_scan_i32(S, flag, pos, nlast, pos + nlast);
      if (rc) return rc;
      int add = 0;
      HIP_CHECK(hipMemcpyAsync(&add, pos + nlast, sizeof(int), hipMemcpyDeviceToHost, S->stream));
      HIP_CHECK(hipStreamSynchronize(S->stream));
      if (ncur + add > S->Ncap) {
        if ((rc = swr_nbr_ensure_atoms(S, (int)((ncur + add) * 1.3) + 64))) return rc;
        flag = S->tmp_i;
        pos = S->tmp_i + nlast + 1;
      }
      double L = hi[d] - lo[d];
      hipLaunchKernelGGL(k_border_scatter, dim3(div_up(nlast, 256)), dim3(256), 0, S->stream, S->x, S->q,
                         S->type, S->tag, S->owner, S->img, S->num_bonds, S->num_hbonds, flag, pos, nlast, ncur, d,
                         side == 0 ? L : -L, side == 0 ? 1 : -1);
      ncur += add;
    }
  }
  S->N = ncur;
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_nbr_forward_x(swr_system *S, const double lo[3], const double hi[3]) {
  if (S->N > S->n)
    hipLaunchKernelGGL(k_forward_x, dim3(div_up(S->N - S->n, 256)), dim3(256), 0, S->stream, S->x, S->owner,
                       S->img, S->n, S->N, hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_nbr_forward_caps(swr_system *S) {
  if (S->N > S->n)
    hipLaunchKernelGGL(k_forward_caps, dim3(div_up(S->N - S->n, 256)), dim3(256), 0, S->stream, S->num_bonds,
                       S->num_hbonds, S->owner, S->n, S->N);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_nbr_nve(swr_system *S, double dt) {
  hipLaunchKernelGGL(k_nve, dim3(div_up(3 * S->n, 256)), dim3(256), 0, S->stream, S->x, S->v, S->n, dt);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_nbr_pbc(swr_system *S, const double lo[3], const double hi[3]) {
  hipLaunchKernelGGL(k_pbc, dim3(div_up(S->n, 256)), dim3(256), 0, S->stream, S->x, S->n, lo[0], lo[1], lo[2],
                     hi[0], hi[1], hi[2]);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Stable sort of n (key, value) pairs by the low nbits of the key with the
// LDS radix passes above: vals_in == nullptr sorts the indices 0..n-1;
// the sorted values go to vals_out (which may alias neither input).  Uses
// the binning's scratch (rs_keys, rs_vals, rs_hist), sized for N >= n.
int swr_radix_sort_i32(swr_system *S, const int *keys, const int *vals_in, int n, int nbits, int *vals_out) {
  if (n <= 0) return SWR_OK;
  const int rs_nblk = div_up(n, RS_CHUNK);
  int rc;
  if ((rc = swr_reserve(&S->rs_keys, S->cap_rs_keys, 2 * (size_t)n)) ||
      (rc = swr_reserve(&S->rs_vals, S->cap_rs_vals, (size_t)n)) ||
      (rc = swr_reserve(&S->rs_hist, S->cap_rs_hist, 2 * (size_t)RS_RADIX * rs_nblk + 1)))
    return rc;
  const int passes = std::max(1, (nbits + 7) / 8);
  int *hist = S->rs_hist, *off = S->rs_hist + (size_t)RS_RADIX * rs_nblk;
  const int *kin = keys, *vin = vals_in;
  int *kbuf[2] = {S->rs_keys, S->rs_keys + n};
  for (int p = 0; p < passes; p++) {
    int *vout = (p == passes - 1) ? vals_out : ((passes - 1 - p) % 2 ? S->rs_vals : vals_out);
    int *kout = kbuf[p & 1];
    hipLaunchKernelGGL(k_rs_hist, dim3(rs_nblk), dim3(RS_T), 0, S->stream, kin, n, 8 * p, rs_nblk, hist);
    if ((rc = swr_scan_i32(S, hist, off, RS_RADIX * rs_nblk, nullptr))) return rc;
    hipLaunchKernelGGL(k_rs_scatter, dim3(rs_nblk), dim3(RS_T), 0, S->stream, kin, vin, n, 8 * p, rs_nblk,
                       (const int *)off, kout, vout);
    kin = kout;
    vin = vout;
  }
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Bin all N atoms and build the full Verlet list (CSR: voff[N+1], vlist).
// Full neighbour list of all N atoms (nrow = N: the ghost list) or of the
// local atoms only (nrow = n: the atom-only list, ghost rows empty).
int swr_nbr_build(swr_system *S, int local_only) {
  const int N = S->N;
  const int nrow = local_only ? S->n : N;
  int rc;
  if ((rc = swr_reserve(&S->bin_of, S->cap_bin_of, N)) || (rc = swr_reserve(&S->bp_id, S->cap_bp_id, N)) ||
      (rc = swr_reserve(&S->bp_pos, S->cap_bp_pos, N)) || (rc = swr_reserve(&S->bin_occ, S->cap_bin_occ, (size_t)N + 1)))
    return rc;
  // counting sort by bin, staged in LDS (k_rs_*): passes of 8 bits up to the
  // largest bin id; ping-pong between the key/value buffers so that the
  // last pass writes bp_id
  const int rs_nblk = div_up(N, RS_CHUNK);
  int passes = 1;
  while (passes < 4 && (S->nbins - 1) >> (8 * passes)) passes++;
  if ((rc = swr_reserve(&S->rs_keys, S->cap_rs_keys, 2 * (size_t)N)) ||
      (rc = swr_reserve(&S->rs_vals, S->cap_rs_vals, (size_t)N)) ||
      (rc = swr_reserve(&S->rs_hist, S->cap_rs_hist, 2 * (size_t)RS_RADIX * rs_nblk + 1)))
    return rc;
  swr_kt_begin(S, KT_BINS);
  hipLaunchKernelGGL(k_bin_keys, dim3(div_up(N, 256)), dim3(256), 0, S->stream, S->x, N, S->binlo[0], S->binlo[1],
                     S->binlo[2], S->bininv[0], S->bininv[1], S->bininv[2], S->nbin[0], S->nbin[1], S->nbin[2],
                     S->bin_of, S->status);
  {
    int *hist = S->rs_hist, *off = S->rs_hist + (size_t)RS_RADIX * rs_nblk;
    const int *kin = S->bin_of, *vin = nullptr;
    int *kbuf[2] = {S->rs_keys, S->rs_keys + N};
    for (int p = 0; p < passes; p++) {
      // the last pass writes the values into bp_id, earlier ones alternate
      // so that no pass reads the buffer it writes
      int *vout = (p == passes - 1) ? S->bp_id : ((passes - 1 - p) % 2 ? S->rs_vals : S->bp_id);
      int *kout = kbuf[p & 1];
      hipLaunchKernelGGL(k_rs_hist, dim3(rs_nblk), dim3(RS_T), 0, S->stream, kin, N, 8 * p, rs_nblk, hist);
      if ((rc = swr_scan_i32(S, hist, off, RS_RADIX * rs_nblk, nullptr))) return rc;
      hipLaunchKernelGGL(k_rs_scatter, dim3(rs_nblk), dim3(RS_T), 0, S->stream, kin, vin, N, 8 * p, rs_nblk,
                         (const int *)off, kout, vout);
      kin = kout;
      vin = vout;
    }
    HIP_CHECK(hipMemsetAsync(S->bin_count, 0, sizeof(int) * (S->nbins + 1), S->stream));
    HIP_CHECK(hipMemsetAsync(S->bin_occ, 0, sizeof(int), S->stream));
    hipLaunchKernelGGL(k_bin_runs, dim3(div_up(N, 256)), dim3(256), 0, S->stream, kin, N, S->bin_count, S->bin_occ);
  }
  if ((rc = swr_scan_i32(S, S->bin_count, S->bin_start, S->nbins + 1, nullptr))) return rc;
  hipLaunchKernelGGL(k_bin_pos, dim3(div_up(N, 256)), dim3(256), 0, S->stream, N, (const int *)S->bp_id, S->x,
                     S->bp_pos);
  swr_kt_end(S, KT_BINS);
  if ((rc = swr_reserve(&S->vcount, S->cap_vcount, N)) || (rc = swr_reserve(&S->voff, S->cap_voff, (size_t)N + 1)))
    return rc;
  const double cutsq = S->cutneigh * S->cutneigh;
  // one wave per occupied bin (at most min(nbins, N) of them)
  const int nblk = div_up(std::min(S->nbins, N), WAVES_PER_BLOCK);
  swr_kt_begin(S, KT_VERLET);
  hipLaunchKernelGGL(k_verlet_bins<false>, dim3(nblk), dim3(NB_THREADS), 0, S->stream, S->bin_occ, S->bin_start,
                     S->bp_pos, S->bp_id, S->stencil_rows, S->nrows, S->nbin[0], S->nbin[1], S->nbin[2], cutsq,
                     S->vcount, (const long long *)nullptr, (int *)nullptr, nrow);
  swr_kt_end(S, KT_VERLET);
  long long total = 0;
  if ((rc = swr_scan_i32_to_i64(S, S->vcount, S->voff, N, &total))) return rc;
  if (total > S->vcap) {
    long long cap = (long long)(total * 1.15) + 1024;
    if ((rc = swr_alloc(&S->vlist, (size_t)cap))) return rc;
    if ((rc = swr_alloc(&S->far, (size_t)cap))) return rc;
    S->vcap = cap;
  }
  S->vtotal = total;
  // local rows' share (the H matrix capacity), read once per build
  HIP_CHECK(hipMemcpyAsync(&S->vtotal_local, S->voff + S->n, sizeof(long long), hipMemcpyDeviceToHost, S->stream));
  HIP_CHECK(hipStreamSynchronize(S->stream));
  swr_kt_begin(S, KT_VERLET);
  hipLaunchKernelGGL(k_verlet_bins<true>, dim3(nblk), dim3(NB_THREADS), 0, S->stream, S->bin_occ, S->bin_start,
                     S->bp_pos, S->bp_id, S->stencil_rows, S->nrows, S->nbin[0], S->nbin[1], S->nbin[2], cutsq,
                     S->vcount, (const long long *)S->voff, S->vlist, nrow);
  swr_kt_end(S, KT_VERLET);
  HIP_CHECK(hipGetLastError());
  S->nbr_gen++;
  return SWR_OK;
}
