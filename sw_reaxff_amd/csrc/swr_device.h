// swr_device.h — device helpers shared by the kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SWR_FULL_MASK 0xffffffffffffffffull

__device__ __forceinline__ int lane_id() { return threadIdx.x & (SWR_WAVE - 1); }
__device__ __forceinline__ int wave_in_block() { return threadIdx.x / SWR_WAVE; }

// Workgroup barrier that orders LDS accesses only: it does not wait for the
// wave's outstanding global loads and stores (a full workgroup fence, as in
// __syncthreads, adds s_waitcnt vmcnt(0), and vmcnt counts both).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Number of set bits of a wave ballot below this lane.
__device__ __forceinline__ int ballot_rank(unsigned long long m) {
  const int l = lane_id();
  unsigned long long below = (l == 0) ? 0ull : (m & ((~0ull) >> (64 - l)));
  return __popcll(below);
}

// Inclusive wave scan (Hillis-Steele over 64 lanes).
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < SWR_WAVE; o <<= 1) {
    T t = __shfl_up(v, o, SWR_WAVE);
    if (l >= o) v += t;
  }
  return v;
}

// Wave sum in a fixed tree order (the same for every call with the same
// lane values), broadcast to all lanes.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = SWR_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, SWR_WAVE);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = SWR_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, SWR_WAVE);
  return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = SWR_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, SWR_WAVE);
  return v;
}

// XCD-aware block index: MI355X dispatches workgroups round-robin over 8
// XCDs, so consecutive blockIdx.x land on different L2s.  Remapping makes
// each XCD own a contiguous range of work (spatially close atoms share an L2).
__device__ __forceinline__ int xcd_block(int nblocks) {
  const int b = blockIdx.x;
  const int nx = 8;
  const int per = (nblocks + nx - 1) / nx;
  const int xcd = b % nx, k = b / nx;
  const int full = nblocks % nx == 0 ? nx : nblocks % nx;  // XCDs with `per` blocks
  // blocks on XCD x: x < full ? per : per-1 (when nblocks % 8 != 0)
  int start = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return start + k;
}
