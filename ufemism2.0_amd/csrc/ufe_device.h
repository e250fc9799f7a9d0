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
constexpr int kMaxPartials = 8192;  // per-reduction partial sums (one per workgroup)

// MI355X dispatches the workgroups of a launch to its 8 XCDs round robin
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  Kernels that
// gather a vector through the mesh connectivity renumber their workgroups so
// that each XCD owns one contiguous range of rows: with the locality-ordered
// mesh, that XCD's L2 then holds only its own part of the gathered vector
// instead of all of it.  The grid must be a multiple of 8 (xcd_grid); the
// renumbering is a bijection, so correctness does not depend on the dispatch
// order.
constexpr int kXcds = 8;
inline int xcd_grid(int g) { return (g + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ int xcd_block() {
  return (int)(blockIdx.x % kXcds) * (int)(gridDim.x / kXcds) + (int)(blockIdx.x / kXcds);
}

inline int blocks_for(long n, int per_block = kBlock) {
  long b = (n + per_block - 1) / per_block;
  return (int)(b < 1 ? 1 : b);
}

// Deterministic grid reduction slot: per-workgroup partials and the result.
struct RedSlot {
  double *partials;   // [kMaxPartials * width]
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

// Workgroup-wide sum of W values over a BS-thread workgroup; result valid in thread 0.
template <int W, int BS = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[W], double (*lds)[BS / 64]) {
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
      for (int w = 0; w < BS / 64; ++w) s += lds[i][w];
      v[i] = s;
    }
  }
}

// Grid reductions are two-phase.  Each workgroup writes its partial sums with
// plain stores (block_partials); a one-workgroup kernel launched next on the
// same stream adds the partials in index order (partials_sum), which makes the
// result deterministic, and runs the epilogue.  The kernel boundary orders the
// two phases: no device-scope fence or ticket inside the producing kernel
// (such a fence writes back the L2 of every workgroup's XCD).
template <int W>
__device__ __forceinline__ void block_partials(const RedSlot &slot, const double (&v)[W]) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < W; ++i) slot.partials[i * kMaxPartials + blockIdx.x] = v[i];
  }
}

// Sum of the first nparts partials of each of W values; valid in thread 0 of a
// one-workgroup launch of BS threads.  The loads of all W values are issued
// together, U per thread and value in flight (U = 8 for W <= 4), so that up to
// U * BS partials cost one memory round trip.
template <int W, int BS = kBlock>
__device__ __forceinline__ void partials_sum(const RedSlot &slot, int nparts, double (&out)[W]) {
  __shared__ double lds[W][BS / 64];
  constexpr int U = W > 4 ? 2 : 8;
  double s[W][U];
#pragma unroll
  for (int i = 0; i < W; ++i)
#pragma unroll
    for (int u = 0; u < U; ++u) s[i][u] = 0.0;
  for (int b0 = 0; b0 < nparts; b0 += U * BS) {
#pragma unroll
    for (int i = 0; i < W; ++i)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int b = b0 + u * BS + (int)threadIdx.x;
        if (b < nparts) s[i][u] += slot.partials[i * kMaxPartials + b];
      }
  }
#pragma unroll
  for (int i = 0; i < W; ++i) {
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) t += s[i][u];
    out[i] = t;
  }
  block_sum<W, BS>(out, lds);
}

// Second phase writing the W totals to out[0..W).
template <int W>
__global__ void __launch_bounds__(kBlock) k_partials_out(RedSlot slot, int nparts, double *__restrict__ out) {
  double tot[W];
  partials_sum<W>(slot, nparts, tot);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < W; ++i) out[i] = tot[i];
  }
}

}  // namespace ufe
