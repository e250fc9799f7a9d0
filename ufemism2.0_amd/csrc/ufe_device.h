// Device-side helpers shared by the engine's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace ufe {

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const std::string &w) : std::runtime_error(w), code(c) {}
};

#define UFE_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess)                                                                  \
      throw ::ufe::HipError(_e, std::string(#call) + ": " + hipGetErrorString(_e) + " at " + \
                                    __FILE__ + ":" + std::to_string(__LINE__));            \
  } while (0)

#define UFE_LAUNCH_CHECK() UFE_HIP(hipGetLastError())

constexpr int kBlock = 256;      // threads per workgroup for 1-D kernels (4 wavefronts)
constexpr int kMaxPartials = 2048;  // per-reduction partial sums (one per workgroup)

inline int blocks_for(long n, int per_block = kBlock) {
  long b = (n + per_block - 1) / per_block;
  return (int)(b < 1 ? 1 : b);
}

// Deterministic grid reduction slot: each workgroup writes its partial, the
// last workgroup to arrive (atomic ticket) sums the partials in index order
// and publishes the result.  The ticket is reset by that last workgroup, so a
// slot can be reused by the next launch on the same stream.
struct RedSlot {
  double *partials;   // [kMaxPartials * width]
  unsigned int *ticket;
  double *result;     // [width]
};

// Per-solve scalars that kernels exchange through device memory.
struct KrylovScalars {
  double rho, rho_old, alpha, omega, beta;
  double rr, bb, tol2;   // ||r||^2, ||b||^2, stop threshold on ||r||^2
  double dot_a, dot_b;   // scratch reductions
  int done;              // 0 running, 1 converged, 2 breakdown, 3 max iterations
  int its;
  int maxit;
  int pad;
};

// Wavefront-wide sum with cross-lane shuffles (wave64).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Workgroup-wide sum of up to 4 values; result valid in thread 0.
template <int W>
__device__ __forceinline__ void block_sum(double (&v)[W], double (*lds)[kBlock / 64]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    double s = wave_sum(v[i]);
    if (lane == 0) lds[i][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < W; ++i) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += lds[i][w];
      v[i] = s;
    }
  }
}

// Finish a grid reduction of W values.  Call from all threads after the
// workgroup sum; returns true in exactly one thread of the last workgroup,
// with out[] holding the grid totals.
template <int W>
__device__ __forceinline__ bool grid_finish(const RedSlot &slot, double (&v)[W], double (&out)[W]) {
  __shared__ bool is_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < W; ++i) slot.partials[i * kMaxPartials + blockIdx.x] = v[i];
    __threadfence();
    unsigned int t = atomicAdd(slot.ticket, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  // last workgroup: sum partials in fixed order (deterministic)
  __shared__ double lds[W][kBlock / 64];
  double acc[W];
  __threadfence();
#pragma unroll
  for (int i = 0; i < W; ++i) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock) {
      s += slot.partials[i * kMaxPartials + b];
    }
    acc[i] = s;
  }
  block_sum<W>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < W; ++i) out[i] = acc[i];
    *slot.ticket = 0u;
    return true;
  }
  return false;
}

}  // namespace ufe
