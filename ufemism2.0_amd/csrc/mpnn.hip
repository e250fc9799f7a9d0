// Mesh message-passing network on the triangle graph: the learned Picard
// initial guess.  Inference and training run on the device with hand-written
// kernels; the dense layers use BF16 matrix-core instructions
// (gfx950's mfma_f32_16x16x32_bf16; 16x16x16bf16_1k for the 16-wide layers)
// with FP32 accumulation, everything else is FP32.
//
// Network (nodes = triangles, neighbours = TriC, mean aggregation):
//   h0  = relu(X We + be)                                   X [N,16]  We [16,64]
//   h_l = h_{l-1} + relu([h_{l-1} | mean_nb h_{l-1}] W_l + b_l)    W_l [128,64], l = 1..3
//   y   = h_3 Wd + bd                                       Wd [64,16], outputs 0..2 used
// The outputs are log10 of the speed (m/yr) and the flow direction (x, y;
// normalised when applied), so members whose speeds differ by orders of
// magnitude (A and beta_sq vary over decades) share one scale.  Training
// minimises sum_i ||y_i[0:3] - t_i||^2 / N with Adam (full batch over the
// disjoint union of the training graphs).
//
// Kernels
//   k_rowgemm<K, NO, EPI>  C[N,NO] = A[N,K] W[K,NO] (+ bias, epilogue); one
//       workgroup = 4 wavefronts = a 64-row tile; W^T (bf16) and the A tile
//       (converted to bf16) are staged in LDS with K contiguous, so each lane
//       reads its 4 consecutive bf16 operands with one 64-bit LDS load.
//   k_colgemm<K, NO>       partial sums of A^T B over the rows of each
//       workgroup (weight gradients) plus column sums of B (bias gradients);
//       k_reduce_parts adds the partials in a fixed order (deterministic).
//   gathers: k_concat_mean (forward aggregation), k_dh_update (its adjoint,
//       as a gather over the symmetric neighbour lists: no atomics).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "krylov.h"
#include "ufe_internal.h"

namespace ufe {
namespace mp {

constexpr int kFin = 16, kHid = 64, kFout = 16, kLayers = 3;
constexpr int kTargets = 3;  // trained outputs: log10 speed, direction (x, y)
constexpr int kRows = 64;  // rows per workgroup tile

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// One 16x16 FP32 tile += A (16 x KS) B (KS x 16) on the matrix cores, KS = 32
// (gfx950's v_mfma_f32_16x16x32_bf16: lane l holds A[l % 16][8 (l / 16) ..
// +7] and B[8 (l / 16) .. +7][l % 16]) or KS = 16 (the 16x16x16 shape, for
// the 16-wide layers: lane l holds 4 values from 4 (l / 16)).  p and q point
// at lane l's first A and B values in LDS (16- and 8-byte aligned).
template <int KS>
__device__ __forceinline__ f32x4 mfma_tile(const unsigned short *p, const unsigned short *q, f32x4 acc) {
  if constexpr (KS == 32) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8 *>(p),
                                                   *reinterpret_cast<const bf16x8 *>(q), acc, 0, 0, 0);
  } else {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(*reinterpret_cast<const bf16x4 *>(p),
                                                     *reinterpret_cast<const bf16x4 *>(q), acc, 0, 0, 0);
  }
}

// float32 -> bfloat16, round to nearest even
__device__ __forceinline__ unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

enum { EPI_ENC = 0, EPI_MP = 1, EPI_LIN = 2 };

// Out[N,NO] = epi(A[N,K] W + bias); Wt = W^T in bf16, [NO][K].
template <int K, int NO, int EPI>
__global__ void __launch_bounds__(256) k_rowgemm(int N, const float *__restrict__ A,
                                                 const unsigned short *__restrict__ Wt,
                                                 const float *__restrict__ bias, const float *__restrict__ Hin,
                                                 float *__restrict__ Zout, float *__restrict__ Out) {
  // K steps of 32 where K allows (gfx950's 16x16x32 shape), else of 16; row
  // pitch K + 8 keeps every lane's 16-byte operand aligned and the 16 rows of
  // a step on distinct LDS banks
  constexpr int KS = K % 32 == 0 ? 32 : 16;
  constexpr int KP = K + 8;
  __shared__ __align__(16) unsigned short sA[kRows * KP];
  __shared__ __align__(16) unsigned short sW[NO * KP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < NO * K; e += 256) sW[(e / K) * KP + (e % K)] = Wt[e];
  const int ar = wave * 16 + (lane & 15);
  const int kq = (KS / 4) * (lane >> 4);
  for (int tile = blockIdx.x; tile * kRows < N; tile += gridDim.x) {
    const int r0 = tile * kRows;
    __syncthreads();
    for (int e = tid; e < kRows * K; e += 256) {
      const int r = e / K, k = e % K, row = r0 + r;
      sA[r * KP + k] = f2bf(row < N ? A[(size_t)row * K + k] : 0.0f);
    }
    __syncthreads();
    f32x4 acc[NO / 16];
#pragma unroll
    for (int t = 0; t < NO / 16; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < K; kb += KS) {
#pragma unroll
      for (int t = 0; t < NO / 16; ++t)
        acc[t] = mfma_tile<KS>(&sA[ar * KP + kb + kq], &sW[(t * 16 + (lane & 15)) * KP + kb + kq], acc[t]);
    }
    // lane holds rows 4*(lane/16)+i, column lane%16 of each 16x16 tile
#pragma unroll
    for (int t = 0; t < NO / 16; ++t) {
      const int col = t * 16 + (lane & 15);
      const float bc = bias ? bias[col] : 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = r0 + wave * 16 + 4 * (lane >> 4) + i;
        if (row >= N) continue;
        const float z = acc[t][i] + bc;
        const size_t o = (size_t)row * NO + col;
        if (EPI == EPI_ENC) {
          Zout[o] = z;
          Out[o] = fmaxf(z, 0.0f);
        } else if (EPI == EPI_MP) {
          Zout[o] = z;
          Out[o] = Hin[o] + fmaxf(z, 0.0f);
        } else {
          Out[o] = z;
        }
      }
    }
  }
}

// part[block][K*NO + NO]: sum over this block's rows of A[r,:]^T B[r,:], then
// the column sums of B.
template <int K, int NO>
__global__ void __launch_bounds__(256) k_colgemm(int N, const float *__restrict__ A, const float *__restrict__ B,
                                                 float *__restrict__ part) {
  // the reduction runs over the tile's kRows rows in steps of 32 (16x16x32)
  static_assert(kRows % 32 == 0, "row tile must be a multiple of 32");
  constexpr int RP = kRows + 8;
  constexpr int TM = K / 16, TN = NO / 16, TT = TM * TN, PER = (TT + 3) / 4;
  __shared__ __align__(16) unsigned short sAt[K * RP];
  __shared__ __align__(16) unsigned short sBt[NO * RP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kq = 8 * (lane >> 4);
  f32x4 acc[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
  for (int tile = blockIdx.x; tile * kRows < N; tile += gridDim.x) {
    const int r0 = tile * kRows;
    __syncthreads();
    for (int e = tid; e < kRows * K; e += 256) {
      const int r = e / K, k = e % K, row = r0 + r;
      sAt[k * RP + r] = f2bf(row < N ? A[(size_t)row * K + k] : 0.0f);
    }
    for (int e = tid; e < kRows * NO; e += 256) {
      const int r = e / NO, n = e % NO, row = r0 + r;
      sBt[n * RP + r] = f2bf(row < N ? B[(size_t)row * NO + n] : 0.0f);
    }
    if (tid < NO) {
      const int rmax = min(kRows, N - r0);
      for (int r = 0; r < rmax; ++r) bsum += B[(size_t)(r0 + r) * NO + tid];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int t = wave * PER + q;
      if (t < TT) {
        const int tm = t / TN, tn = t % TN;
#pragma unroll
        for (int rb = 0; rb < kRows; rb += 32)
          acc[q] = mfma_tile<32>(&sAt[(tm * 16 + (lane & 15)) * RP + rb + kq],
                                 &sBt[(tn * 16 + (lane & 15)) * RP + rb + kq], acc[q]);
      }
    }
  }
  float *out = part + (size_t)blockIdx.x * (K * NO + NO);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int t = wave * PER + q;
    if (t < TT) {
      const int tm = t / TN, tn = t % TN;
#pragma unroll
      for (int i = 0; i < 4; ++i) out[(tm * 16 + 4 * (lane >> 4) + i) * NO + tn * 16 + (lane & 15)] = acc[q][i];
    }
  }
  if (tid < NO) out[K * NO + tid] = bsum;
}

// out[e] = sum over parts p (in order) of part[p * len + e]
__global__ void __launch_bounds__(256) k_reduce_parts(int nparts, int len, const float *__restrict__ part,
                                                      float *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  float s = 0.0f;
  for (int p = 0; p < nparts; ++p) s += part[(size_t)p * len + e];
  out[e] = s;
}

// Cat[i] = [H[i] | mean over valid neighbours of H[j]]  (H: [N,64], Cat: [N,128])
__global__ void __launch_bounds__(256) k_concat_mean(int N, const float *__restrict__ H, const int *__restrict__ nb,
                                                     float *__restrict__ Cat) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)N * kHid) return;
  const int i = (int)(e / kHid), c = (int)(e % kHid);
  float s = 0.0f;
  int d = 0;
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const int j = nb[3 * i + n];
    if (j >= 0) { s += H[(size_t)j * kHid + c]; ++d; }
  }
  Cat[(size_t)i * 2 * kHid + c] = H[e];
  Cat[(size_t)i * 2 * kHid + kHid + c] = d ? s / d : 0.0f;
}

// dZ = dH * (Z > 0)
__global__ void __launch_bounds__(256) k_relu_back(long n, const float *__restrict__ dH, const float *__restrict__ Z,
                                                   float *__restrict__ dZ) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) dZ[e] = Z[e] > 0.0f ? dH[e] : 0.0f;
}

// dH[j] += dCat[j, 0:64] + sum_{i in nb(j)} dCat[i, 64:128] / deg(i)
// (adjoint of k_concat_mean for symmetric neighbour lists)
__global__ void __launch_bounds__(256) k_dh_update(int N, const float *__restrict__ dCat, const int *__restrict__ nb,
                                                   float *__restrict__ dH) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)N * kHid) return;
  const int j = (int)(e / kHid), c = (int)(e % kHid);
  float s = dH[e] + dCat[(size_t)j * 2 * kHid + c];
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const int i = nb[3 * j + n];
    if (i < 0) continue;
    int d = 0;
#pragma unroll
    for (int m = 0; m < 3; ++m) d += nb[3 * i + m] >= 0;
    s += dCat[(size_t)i * 2 * kHid + kHid + c] / (float)d;
  }
  dH[e] = s;
}

// dY = 2 (Y[:, 0:3] - T) / N (other columns 0); per-block partial sums of the
// squared error into part[blockIdx]
__global__ void __launch_bounds__(256) k_loss_dy(int N, const float *__restrict__ Y, const float *__restrict__ T,
                                                 float *__restrict__ dY, float *__restrict__ part) {
  __shared__ float red[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float se = 0.0f;
  if (i < N) {
    const float s = 2.0f / (float)N;
#pragma unroll
    for (int c = 0; c < kTargets; ++c) {
      const float d = Y[(size_t)i * kFout + c] - T[(size_t)kTargets * i + c];
      se += d * d;
      dY[(size_t)i * kFout + c] = s * d;
    }
#pragma unroll
    for (int c = kTargets; c < kFout; ++c) dY[(size_t)i * kFout + c] = 0.0f;
  }
  red[threadIdx.x] = se;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// W [K, NO] fp32 -> Wt [NO][K] bf16 (forward operand) and Wn [K][NO] bf16
// (operand of the input-gradient product)
__global__ void __launch_bounds__(256) k_weights_bf16(int K, int NO, const float *__restrict__ W,
                                                      unsigned short *__restrict__ Wt, unsigned short *__restrict__ Wn) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= K * NO) return;
  const int k = e / NO, n = e % NO;
  const unsigned short b = f2bf(W[e]);
  Wt[n * K + k] = b;
  Wn[e] = b;
}

__global__ void __launch_bounds__(256) k_adam(int n, float *__restrict__ w, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.0f - b1) * gi;
  const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  w[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Parameter layout in one fp32 vector.
struct Layout {
  size_t We, be, W[kLayers], b[kLayers], Wd, bd, total;
  Layout() {
    size_t o = 0;
    We = o; o += kFin * kHid;
    be = o; o += kHid;
    for (int l = 0; l < kLayers; ++l) {
      W[l] = o; o += 2 * kHid * kHid;
      b[l] = o; o += kHid;
    }
    Wd = o; o += kHid * kFout;
    bd = o; o += kFout;
    total = o;
  }
};

inline int grid_rows(int N) { return std::max(1, std::min((N + kRows - 1) / kRows, 1024)); }
inline int grid_elems(long n) { return (int)std::max<long>(1, (n + 255) / 256); }

}  // namespace mp
}  // namespace ufe

using namespace ufe;
using namespace ufe::mp;

struct ufe_mpnn {
  int device = 0;
  hipStream_t s = nullptr;
  Layout L;
  DBuf w, g, m, v, wt16, wn16, part;
  int step = 0;
  // graph buffers (sized for the largest N seen)
  int cap = 0;
  DBuf X, nb, T, Ze, H[kLayers + 1], Z[kLayers], Cat[kLayers], Y, dY, dH, dZ, dCat, lpart;
  std::string err;

  size_t wt_off(size_t woff) const { return woff; }  // bf16 copies share the fp32 offsets

  void ensure(int N) {
    if (N <= cap) return;
    cap = N;
    size_t n = (size_t)N;
    X.alloc(n * kFin * 4);
    nb.alloc(n * 3 * 4);
    T.alloc(n * kTargets * 4);
    Ze.alloc(n * kHid * 4);
    for (auto &b : H) b.alloc(n * kHid * 4);
    for (auto &b : Z) b.alloc(n * kHid * 4);
    for (auto &b : Cat) b.alloc(n * 2 * kHid * 4);
    Y.alloc(n * kFout * 4);
    dY.alloc(n * kFout * 4);
    dH.alloc(n * kHid * 4);
    dZ.alloc(n * kHid * 4);
    dCat.alloc(n * 2 * kHid * 4);
    lpart.alloc(((n + 255) / 256) * 4);
  }

  void refresh_bf16() {
    auto conv = [&](size_t off, int K, int NO) {
      hipLaunchKernelGGL(k_weights_bf16, dim3(grid_elems((long)K * NO)), dim3(256), 0, s, K, NO,
                         w.as<float>() + off, wt16.as<unsigned short>() + off, wn16.as<unsigned short>() + off);
      UFE_LAUNCH_CHECK();
    };
    conv(L.We, kFin, kHid);
    for (int l = 0; l < kLayers; ++l) conv(L.W[l], 2 * kHid, kHid);
    conv(L.Wd, kHid, kFout);
  }

  const unsigned short *wt(size_t off) const { return wt16.as<unsigned short>() + off; }
  const unsigned short *wn(size_t off) const { return wn16.as<unsigned short>() + off; }
  const float *wf(size_t off) const { return w.as<float>() + off; }

  // forward over N nodes already in X / nb
  void forward(int N) {
    float *x = X.as<float>();
    const int *n3 = nb.as<int>();
    const int gr = grid_rows(N);
    hipLaunchKernelGGL((k_rowgemm<kFin, kHid, EPI_ENC>), dim3(gr), dim3(256), 0, s, N, x, wt(L.We), wf(L.be),
                       (const float *)nullptr, Ze.as<float>(), H[0].as<float>());
    UFE_LAUNCH_CHECK();
    for (int l = 0; l < kLayers; ++l) {
      hipLaunchKernelGGL(k_concat_mean, dim3(grid_elems((long)N * kHid)), dim3(256), 0, s, N, H[l].as<float>(), n3,
                         Cat[l].as<float>());
      UFE_LAUNCH_CHECK();
      hipLaunchKernelGGL((k_rowgemm<2 * kHid, kHid, EPI_MP>), dim3(gr), dim3(256), 0, s, N, Cat[l].as<float>(),
                         wt(L.W[l]), wf(L.b[l]), H[l].as<float>(), Z[l].as<float>(), H[l + 1].as<float>());
      UFE_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((k_rowgemm<kHid, kFout, EPI_LIN>), dim3(gr), dim3(256), 0, s, N, H[kLayers].as<float>(),
                       wt(L.Wd), wf(L.bd), (const float *)nullptr, (float *)nullptr, Y.as<float>());
    UFE_LAUNCH_CHECK();
  }

  template <int K, int NO>
  void weight_grad(int N, const float *A, const float *B, size_t woff, size_t boff) {
    const int gr = grid_rows(N);
    const int len = K * NO + NO;
    hipLaunchKernelGGL((k_colgemm<K, NO>), dim3(gr), dim3(256), 0, s, N, A, B, part.as<float>());
    UFE_LAUNCH_CHECK();
    // the weight block and the bias follow each other in the partials; reduce
    // them into a contiguous scratch then scatter into the gradient layout
    float *tmp = part.as<float>() + (size_t)1024 * (2 * kHid * kHid + kHid);
    hipLaunchKernelGGL(k_reduce_parts, dim3(grid_elems(len)), dim3(256), 0, s, gr, len, part.as<float>(), tmp);
    UFE_LAUNCH_CHECK();
    UFE_HIP(hipMemcpyAsync(g.as<float>() + woff, tmp, sizeof(float) * K * NO, hipMemcpyDeviceToDevice, s));
    UFE_HIP(hipMemcpyAsync(g.as<float>() + boff, tmp + K * NO, sizeof(float) * NO, hipMemcpyDeviceToDevice, s));
  }

  // loss and gradient for the N nodes in X / nb / T; returns the loss
  float loss_grad(int N) {
    forward(N);
    const int nbk = grid_elems(N);
    hipLaunchKernelGGL(k_loss_dy, dim3(nbk), dim3(256), 0, s, N, Y.as<float>(), T.as<float>(), dY.as<float>(),
                       lpart.as<float>());
    UFE_LAUNCH_CHECK();
    const int gr = grid_rows(N);
    // decoder
    weight_grad<kHid, kFout>(N, H[kLayers].as<float>(), dY.as<float>(), L.Wd, L.bd);
    hipLaunchKernelGGL((k_rowgemm<kFout, kHid, EPI_LIN>), dim3(gr), dim3(256), 0, s, N, dY.as<float>(), wn(L.Wd),
                       (const float *)nullptr, (const float *)nullptr, (float *)nullptr, dH.as<float>());
    UFE_LAUNCH_CHECK();
    for (int l = kLayers - 1; l >= 0; --l) {
      hipLaunchKernelGGL(k_relu_back, dim3(grid_elems((long)N * kHid)), dim3(256), 0, s, (long)N * kHid,
                         dH.as<float>(), Z[l].as<float>(), dZ.as<float>());
      UFE_LAUNCH_CHECK();
      weight_grad<2 * kHid, kHid>(N, Cat[l].as<float>(), dZ.as<float>(), L.W[l], L.b[l]);
      hipLaunchKernelGGL((k_rowgemm<kHid, 2 * kHid, EPI_LIN>), dim3(gr), dim3(256), 0, s, N, dZ.as<float>(),
                         wn(L.W[l]), (const float *)nullptr, (const float *)nullptr, (float *)nullptr,
                         dCat.as<float>());
      UFE_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_dh_update, dim3(grid_elems((long)N * kHid)), dim3(256), 0, s, N, dCat.as<float>(),
                         nb.as<int>(), dH.as<float>());
      UFE_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_relu_back, dim3(grid_elems((long)N * kHid)), dim3(256), 0, s, (long)N * kHid,
                       dH.as<float>(), Ze.as<float>(), dZ.as<float>());
    UFE_LAUNCH_CHECK();
    weight_grad<kFin, kHid>(N, X.as<float>(), dZ.as<float>(), L.We, L.be);
    std::vector<float> lp(nbk);
    UFE_HIP(hipMemcpyAsync(lp.data(), lpart.p, sizeof(float) * nbk, hipMemcpyDeviceToHost, s));
    UFE_HIP(hipStreamSynchronize(s));
    double tot = 0.0;
    for (float x : lp) tot += x;
    return (float)(tot / N);
  }

  void adam(float lr) {
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(grid_elems((long)L.total)), dim3(256), 0, s, (int)L.total, w.as<float>(),
                       g.as<float>(), m.as<float>(), v.as<float>(), lr, b1, b2, 1e-8f, c1, c2);
    UFE_LAUNCH_CHECK();
    refresh_bf16();
  }

  void upload_graph(int N, const float *Xh, const int *nbh, const float *Th) {
    ensure(N);
    UFE_HIP(hipMemcpyAsync(X.p, Xh, sizeof(float) * kFin * (size_t)N, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(nb.p, nbh, sizeof(int) * 3 * (size_t)N, hipMemcpyHostToDevice, s));
    if (Th) UFE_HIP(hipMemcpyAsync(T.p, Th, sizeof(float) * kTargets * (size_t)N, hipMemcpyHostToDevice, s));
  }
};

namespace {

int mp_fail(ufe_mpnn *n, int code, const std::string &msg) {
  if (n) n->err = msg;
  set_global_error(msg);
  return code;
}

template <typename F>
int mp_guard(ufe_mpnn *n, F &&f) {
  try {
    return f();
  } catch (const HipError &e) {
    return mp_fail(n, UFE_ERR_DEVICE, e.what());
  } catch (const std::exception &e) {
    return mp_fail(n, UFE_ERR_ARGUMENT, e.what());
  }
}

// Neighbour lists must be symmetric (j in nb(i) <=> i in nb(j)) and in range:
// the backward gather relies on it.
bool valid_graph(int N, const int *nb) {
  for (int i = 0; i < N; ++i)
    for (int a = 0; a < 3; ++a) {
      int j = nb[3 * i + a];
      if (j < -1 || j >= N) return false;
      if (j < 0) continue;
      bool back = false;
      for (int b = 0; b < 3; ++b) back |= nb[3 * j + b] == i;
      if (!back) return false;
    }
  return true;
}

}  // namespace

extern "C" int ufe_mpnn_num_weights(void) { return (int)Layout().total; }

extern "C" int ufe_mpnn_create(ufe_mpnn **out, int device, const float *weights) {
  if (!out) return mp_fail(nullptr, UFE_ERR_ARGUMENT, "null argument");
  auto n = std::make_unique<ufe_mpnn>();
  int rc = mp_guard(n.get(), [&]() -> int {
    n->device = device;
    UFE_HIP(hipSetDevice(device));
    UFE_HIP(hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking));
    const size_t P = n->L.total;
    n->w.alloc(P * 4);
    n->g.alloc(P * 4);
    n->m.alloc(P * 4);
    n->v.alloc(P * 4);
    n->wt16.alloc(P * 2);
    n->wn16.alloc(P * 2);
    // partials: up to 1024 workgroups x the largest block (128x64 + 64), plus
    // a reduction target after them
    n->part.alloc(sizeof(float) * (size_t)1025 * (2 * kHid * kHid + kHid));
    UFE_HIP(hipMemsetAsync(n->g.p, 0, P * 4, n->s));
    UFE_HIP(hipMemsetAsync(n->m.p, 0, P * 4, n->s));
    UFE_HIP(hipMemsetAsync(n->v.p, 0, P * 4, n->s));
    if (weights) UFE_HIP(hipMemcpyAsync(n->w.p, weights, P * 4, hipMemcpyHostToDevice, n->s));
    else UFE_HIP(hipMemsetAsync(n->w.p, 0, P * 4, n->s));
    n->refresh_bf16();
    UFE_HIP(hipStreamSynchronize(n->s));
    return UFE_OK;
  });
  if (rc != UFE_OK) return rc;
  *out = n.release();
  return UFE_OK;
}

extern "C" int ufe_mpnn_destroy(ufe_mpnn *n) {
  if (!n) return UFE_OK;
  (void)hipSetDevice(n->device);
  if (n->s) (void)hipStreamSynchronize(n->s);
  hipStream_t s = n->s;
  delete n;
  if (s) (void)hipStreamDestroy(s);
  return UFE_OK;
}

extern "C" int ufe_mpnn_get_weights(ufe_mpnn *n, float *weights) {
  if (!n || !weights) return mp_fail(n, UFE_ERR_ARGUMENT, "null argument");
  return mp_guard(n, [&]() -> int {
    UFE_HIP(hipSetDevice(n->device));
    UFE_HIP(hipMemcpyAsync(weights, n->w.p, n->L.total * 4, hipMemcpyDeviceToHost, n->s));
    UFE_HIP(hipStreamSynchronize(n->s));
    return UFE_OK;
  });
}

extern "C" int ufe_mpnn_forward(ufe_mpnn *n, int N, const float *X, const int *nb, float *Y) {
  if (!n || N < 1 || !X || !nb || !Y) return mp_fail(n, UFE_ERR_ARGUMENT, "invalid argument");
  if (!valid_graph(N, nb)) return mp_fail(n, UFE_ERR_ARGUMENT, "neighbour lists must be in range and symmetric");
  return mp_guard(n, [&]() -> int {
    UFE_HIP(hipSetDevice(n->device));
    n->upload_graph(N, X, nb, nullptr);
    n->forward(N);
    UFE_HIP(hipMemcpyAsync(Y, n->Y.p, sizeof(float) * kFout * (size_t)N, hipMemcpyDeviceToHost, n->s));
    UFE_HIP(hipStreamSynchronize(n->s));
    return UFE_OK;
  });
}

extern "C" int ufe_mpnn_loss_grad(ufe_mpnn *n, int N, const float *X, const int *nb, const float *T, float *loss,
                                  float *grad) {
  if (!n || N < 1 || !X || !nb || !T || !loss) return mp_fail(n, UFE_ERR_ARGUMENT, "invalid argument");
  if (!valid_graph(N, nb)) return mp_fail(n, UFE_ERR_ARGUMENT, "neighbour lists must be in range and symmetric");
  return mp_guard(n, [&]() -> int {
    UFE_HIP(hipSetDevice(n->device));
    n->upload_graph(N, X, nb, T);
    *loss = n->loss_grad(N);
    if (grad) {
      UFE_HIP(hipMemcpyAsync(grad, n->g.p, n->L.total * 4, hipMemcpyDeviceToHost, n->s));
      UFE_HIP(hipStreamSynchronize(n->s));
    }
    return UFE_OK;
  });
}

extern "C" int ufe_mpnn_train(ufe_mpnn *n, int N, const float *X, const int *nb, const float *T, int steps,
                              float lr, float *loss_history) {
  if (!n || N < 1 || !X || !nb || !T || steps < 0 || !(lr > 0.0f))
    return mp_fail(n, UFE_ERR_ARGUMENT, "invalid argument");
  if (!valid_graph(N, nb)) return mp_fail(n, UFE_ERR_ARGUMENT, "neighbour lists must be in range and symmetric");
  return mp_guard(n, [&]() -> int {
    UFE_HIP(hipSetDevice(n->device));
    n->upload_graph(N, X, nb, T);
    for (int it = 0; it < steps; ++it) {
      float l = n->loss_grad(N);
      if (loss_history) loss_history[it] = l;
      if (!std::isfinite(l)) return mp_fail(n, UFE_ERR_DIVERGED, "training loss is not finite");
      n->adam(lr);
    }
    UFE_HIP(hipStreamSynchronize(n->s));
    return UFE_OK;
  });
}

// ---------------------------------------------------------------------------
// Inference inside a solver context (features from the uploaded inputs).
// ---------------------------------------------------------------------------
namespace ufe {
namespace mp {
int ctx_features(ufe_ctx *c, float *X_dev, int *nb_dev, hipStream_t *s_out, int *N_out);
int ctx_set_guess(ufe_ctx *c, const float *Y_dev);
}  // namespace mp
}  // namespace ufe

extern "C" int ufe_mpnn_initial_guess(ufe_ctx *c, ufe_mpnn *n) {
  if (!c || !n) return mp_fail(n, UFE_ERR_ARGUMENT, "null argument");
  return mp_guard(n, [&]() -> int {
    UFE_HIP(hipSetDevice(n->device));
    int N = 0;
    hipStream_t cs;
    // first call sizes the buffers
    int rc = ufe::mp::ctx_features(c, nullptr, nullptr, &cs, &N);
    if (rc != UFE_OK) return rc;
    n->ensure(N);
    UFE_HIP(hipStreamSynchronize(n->s));
    rc = ufe::mp::ctx_features(c, n->X.as<float>(), n->nb.as<int>(), &cs, &N);
    if (rc != UFE_OK) return rc;
    UFE_HIP(hipStreamSynchronize(cs));
    n->forward(N);
    UFE_HIP(hipStreamSynchronize(n->s));
    return ufe::mp::ctx_set_guess(c, n->Y.as<float>());
  });
}
