// util.hip — exclusive prefix sums (3-phase, deterministic) and small helpers.
#include "swr_internal.h"
#include "swr_device.h"

#include <mutex>

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

std::mutex g_err_mu;
std::string g_err;

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total.
template <class T>
__device__ T block_excl_scan(T v, T *lds, T *total) {
  const int w = wave_in_block(), l = lane_id();
  T inc = wave_incl_scan(v);
  if (l == SWR_WAVE - 1) lds[w] = inc;
  __syncthreads();
  T wave_off = 0, tot = 0;
  for (int k = 0; k < SCAN_THREADS / SWR_WAVE; k++) {
    if (k < w) wave_off += lds[k];
    tot += lds[k];
  }
  __syncthreads();
  *total = tot;
  return wave_off + inc - v;
}

template <class TI, class TO>
__global__ void __launch_bounds__(SCAN_THREADS) k_tile_sums(const TI *in, int n, TO *tile_sums) {
  __shared__ TO lds[SCAN_THREADS / SWR_WAVE];
  const long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  TO s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++)
    if (base + k < n) s += (TO)in[base + k];
  TO tot;
  block_excl_scan<TO>(s, lds, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// One block scans all tile sums in order (carry across chunks).
template <class TO>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(TO *tile_sums, int ntiles, TO *total) {
  __shared__ TO lds[SCAN_THREADS / SWR_WAVE];
  TO carry = 0;
  for (int c = 0; c < ntiles; c += SCAN_THREADS) {
    int i = c + threadIdx.x;
    TO v = i < ntiles ? tile_sums[i] : 0;
    TO tot;
    TO ex = block_excl_scan<TO>(v, lds, &tot);
    if (i < ntiles) tile_sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <class TI, class TO>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const TI *in, int n, const TO *tile_off, TO *out) {
  __shared__ TO lds[SCAN_THREADS / SWR_WAVE];
  const long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  TO vals[SCAN_ITEMS];
  TO s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    vals[k] = (base + k < n) ? (TO)in[base + k] : 0;
    s += vals[k];
  }
  TO tot;
  TO ex = block_excl_scan<TO>(s, lds, &tot) + tile_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n) out[base + k] = ex;
    ex += vals[k];
  }
}

// Three exclusive int scans of the same length in one set of three launches
// (blockIdx.y picks the array): the same tiles, sums and order as three
// separate scans.
struct Scan3 {
  const int *in[3];
  int *out[3];
  int *total[3];
};
template <class T>
__device__ __forceinline__ T pick3(int a, T x, T y, T z) { return a == 0 ? x : a == 1 ? y : z; }

__global__ void __launch_bounds__(SCAN_THREADS) k_tile_sums3(Scan3 A, int n, int ntiles, int *tiles) {
  __shared__ int lds[SCAN_THREADS / SWR_WAVE];
  const int a = blockIdx.y;
  const int *in = pick3(a, A.in[0], A.in[1], A.in[2]);
  const long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++)
    if (base + k < n) s += in[base + k];
  int tot;
  block_excl_scan<int>(s, lds, &tot);
  if (threadIdx.x == 0) tiles[(size_t)a * ntiles + blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles3(Scan3 A, int ntiles, int *tiles) {
  __shared__ int lds[SCAN_THREADS / SWR_WAVE];
  const int a = blockIdx.y;
  int *ts = tiles + (size_t)a * ntiles;
  int carry = 0;
  for (int c = 0; c < ntiles; c += SCAN_THREADS) {
    int i = c + threadIdx.x;
    int v = i < ntiles ? ts[i] : 0;
    int tot;
    int ex = block_excl_scan<int>(v, lds, &tot);
    if (i < ntiles) ts[i] = carry + ex;
    carry += tot;
  }
  int *total = pick3(a, A.total[0], A.total[1], A.total[2]);
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply3(Scan3 A, int n, int ntiles, const int *tiles) {
  __shared__ int lds[SCAN_THREADS / SWR_WAVE];
  const int a = blockIdx.y;
  const int *in = pick3(a, A.in[0], A.in[1], A.in[2]);
  int *out = pick3(a, A.out[0], A.out[1], A.out[2]);
  const long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int vals[SCAN_ITEMS];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    vals[k] = (base + k < n) ? in[base + k] : 0;
    s += vals[k];
  }
  int tot;
  int ex = block_excl_scan<int>(s, lds, &tot) + tiles[(size_t)a * ntiles + blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n) out[base + k] = ex;
    ex += vals[k];
  }
}

template <class TI, class TO>
int scan_impl(swr_system *S, const TI *in, TO *out, int n, TO *total_dev) {
  int ntiles = div_up(n > 0 ? n : 1, SCAN_TILE);
  size_t need = sizeof(TO) * (size_t)(ntiles + 1);
  if (need > S->scan_tmp_bytes) {
    if (S->scan_tmp) (void)hipFree(S->scan_tmp);
    S->scan_tmp = nullptr;
    S->scan_tmp_bytes = 0;
    HIP_CHECK(hipMalloc(&S->scan_tmp, need * 2));
    S->scan_tmp_bytes = need * 2;
  }
  TO *tiles = (TO *)S->scan_tmp;
  TO *tot = total_dev ? total_dev : tiles + ntiles;
  if (n > 0) {
    hipLaunchKernelGGL((k_tile_sums<TI, TO>), dim3(ntiles), dim3(SCAN_THREADS), 0, S->stream, in, n, tiles);
    hipLaunchKernelGGL((k_scan_tiles<TO>), dim3(1), dim3(SCAN_THREADS), 0, S->stream, tiles, ntiles, tot);
    hipLaunchKernelGGL((k_scan_apply<TI, TO>), dim3(ntiles), dim3(SCAN_THREADS), 0, S->stream, in, n,
                       (const TO *)tiles, out);
  } else {
    HIP_CHECK(hipMemsetAsync(tot, 0, sizeof(TO), S->stream));
  }
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

}  // namespace

void swr_set_error_msg(const std::string &m) {
  std::lock_guard<std::mutex> g(g_err_mu);
  g_err = m;
}

extern "C" const char *swr_last_error(void) {
  static thread_local std::string copy;
  std::lock_guard<std::mutex> g(g_err_mu);
  copy = g_err;
  return copy.c_str();
}

// Exclusive scan of n ints into int64 offsets; out has n+1 entries
// (out[n] = total).  If total_host is given, synchronises and returns it.
int swr_scan_i32_to_i64(swr_system *S, const int *in, long long *out, int n, long long *total_host) {
  int rc = scan_impl<int, long long>(S, in, out, n, out + n);
  if (rc) return rc;
  if (total_host) {
    HIP_CHECK(hipMemcpyAsync(total_host, out + n, sizeof(long long), hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
  }
  return SWR_OK;
}

// Three exclusive scans of n ints each (in[k] -> out[k], totals to the
// device ints total[k]) in three launches.
int swr_scan3_i32(swr_system *S, const int *const in[3], int *const out[3], int n, int *const total[3]) {
  if (n <= 0) {
    for (int k = 0; k < 3; k++)
      if (total[k]) HIP_CHECK(hipMemsetAsync(total[k], 0, sizeof(int), S->stream));
    return SWR_OK;
  }
  const int ntiles = div_up(n, SCAN_TILE);
  const size_t need = sizeof(int) * 3 * (size_t)ntiles;
  if (need > S->scan_tmp_bytes) {
    if (S->scan_tmp) (void)hipFree(S->scan_tmp);
    S->scan_tmp = nullptr;
    S->scan_tmp_bytes = 0;
    HIP_CHECK(hipMalloc(&S->scan_tmp, need * 2));
    S->scan_tmp_bytes = need * 2;
  }
  int *tiles = (int *)S->scan_tmp;
  Scan3 A;
  for (int k = 0; k < 3; k++) {
    A.in[k] = in[k];
    A.out[k] = out[k];
    A.total[k] = total[k];
  }
  hipLaunchKernelGGL(k_tile_sums3, dim3(ntiles, 3), dim3(SCAN_THREADS), 0, S->stream, A, n, ntiles, tiles);
  hipLaunchKernelGGL(k_scan_tiles3, dim3(1, 3), dim3(SCAN_THREADS), 0, S->stream, A, ntiles, tiles);
  hipLaunchKernelGGL(k_scan_apply3, dim3(ntiles, 3), dim3(SCAN_THREADS), 0, S->stream, A, n, ntiles,
                     (const int *)tiles);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Exclusive scan of n ints into ints; total written to *total_dev (device).
int swr_scan_i32(swr_system *S, const int *in, int *out, int n, int *total_dev) {
  return scan_impl<int, int>(S, in, out, n, total_dev);
}

const char *const swr_kt_names[KT_COUNT] = {"verlet", "far", "build_H", "cg_iter", "cg_setup", "bo",
                                            "validate", "spmv", "reset", "qeq_q", "bins", "mixed_inner",
                                            "mixed_outer", "bonds", "cg_scalars", "mark", "cg_update", "nonbonded",
                                            "atom_energy", "dbond", "hbonds", "valence",
                                            "torsion", "panel"};

// Records a start event for kernel slot k on the system stream.
void swr_kt_begin(swr_system *S, int k) {
  if (!S->ktime) return;
  swr_system::KTimer &t = S->kt[k];
  if ((size_t)(2 * t.used + 2) > t.ev.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    t.ev.push_back(a);
    t.ev.push_back(b);
  }
  (void)hipEventRecord(t.ev[2 * t.used], S->stream);
}

void swr_kt_end(swr_system *S, int k) {
  if (!S->ktime) return;
  swr_system::KTimer &t = S->kt[k];
  if ((size_t)(2 * t.used + 2) > t.ev.size()) return;
  (void)hipEventRecord(t.ev[2 * t.used + 1], S->stream);
  t.used++;
}

extern "C" int swr_ktimer_enable(swr_system *S, int on) {
  if (!S) return SWR_ERR_ARG;
  S->ktime = on != 0;
  return SWR_OK;
}

// Accumulates the recorded launches of slot `name` (synchronises) and
// returns the total milliseconds and launch count since the last reset.
extern "C" int swr_ktimer_read(swr_system *S, const char *name, double *ms, long long *count, int reset) {
  if (!S || !name) return SWR_ERR_ARG;
  int k = -1;
  for (int i = 0; i < KT_COUNT; i++)
    if (std::string(name) == swr_kt_names[i]) k = i;
  if (k < 0) return SWR_ERR_ARG;
  swr_system::KTimer &t = S->kt[k];
  if (t.used) {
    HIP_CHECK(hipEventSynchronize(t.ev[2 * t.used - 1]));
    for (int u = 0; u < t.used; u++) {
      float e = 0.f;
      if (hipEventElapsedTime(&e, t.ev[2 * u], t.ev[2 * u + 1]) == hipSuccess) t.ms += e;
    }
    t.count += t.used;
    t.used = 0;
  }
  if (ms) *ms = t.ms;
  if (count) *count = t.count;
  if (reset) { t.ms = 0; t.count = 0; }
  return SWR_OK;
}

int swr_ensure_tmp(swr_system *S, size_t bytes) {
  size_t count = (bytes + sizeof(int) - 1) / sizeof(int);
  if (count <= S->tmp_i_cap) return SWR_OK;
  int rc = swr_alloc(&S->tmp_i, count);
  if (rc) return rc;
  S->tmp_i_cap = count;
  return SWR_OK;
}
