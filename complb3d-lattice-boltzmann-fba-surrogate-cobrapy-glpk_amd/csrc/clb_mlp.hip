// clb_mlp.hip -- the speciation warm-start network: FP32 training kernels (forward,
// backward, Adam) and a fused BF16 MFMA inference kernel with FP32 accumulation.
//
// Network din -> H -> H -> dout, tanh, tanh, linear; inputs and outputs min-max scaled
// to [-1, 1] (the mapminmax convention of surrogate_training/trainSurrogate.py:182-201;
// tansig hidden layers as in src/complab3d_processors_surrogate.hh:74-295). Training
// minimises the mean squared error with Adam.
#include "clb_lattice.h"

namespace clb {

namespace {

// ------------------------------------------------------------------ FP32 tiled GEMM
// C[M x N] = act(op(A) op(B) + bias) + beta * C, op = transpose when the flag is set.
// A is M x K (or K x M when transA), B is K x N (or N x K when transB); row-major with
// leading dimensions lda / ldb / ldc. act: 0 none, 1 tanh. Tile 64 x 64 x 16, 256
// threads, 4 x 4 outputs per thread.
//
// Split K: block z sums k in [z kc, (z + 1) kc) into its own C + z sc (the weight
// gradients reduce over the batch rows; with one split per 32 rows a 128 x 128 gradient
// runs on 64 workgroups instead of 4, and the partials are summed in a fixed order by
// the optimizer step).
constexpr int TM = 64, TN = 64, TK = 16;

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                   const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                   int ldc, const float* __restrict__ bias, int act, float beta,
                                                   int kc, long long sc) {
  __shared__ float As[TK][TM + 4];
  __shared__ float Bs[TK][TN + 4];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int kbeg = blockIdx.z * kc, kend = min(K, kbeg + kc);
  C += blockIdx.z * sc;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    // load A tile: TM x TK elements, 1024 -> 4 per thread
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * 256;
      const int mm = e / TK, kk = e % TK;
      const int gm = m0 + mm, gk = k0 + kk;
      float v = 0.f;
      if (gm < M && gk < kend) v = TA ? A[(long long)gk * lda + gm] : A[(long long)gm * lda + gk];
      As[kk][mm] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * 256;
      const int nn = e / TK, kk = e % TK;
      const int gn = n0 + nn, gk = k0 + kk;
      float v = 0.f;
      if (gn < N && gk < kend) v = TB ? B[(long long)gn * ldb + gk] : B[(long long)gk * ldb + gn];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gm = m0 + ty * 4 + i;
    if (gm >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + tx * 4 + j;
      if (gn >= N) continue;
      float v = acc[i][j];
      if (bias) v += bias[gn];
      if (act == 1) v = tanhf(v);
      float* c = &C[(long long)gm * ldc + gn];
      *c = (beta != 0.f) ? v + beta * *c : v;
    }
  }
}

__global__ void tanh_backward_kernel(long long n, float* dz, const float* h) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) dz[i] *= (1.f - h[i] * h[i]);
}

// out[c] = sum_r A[r, c]; one block per column, shared-memory reduction. Split rows:
// block y sums rows [y kc, (y + 1) kc) into out + y so.
__global__ void colsum_kernel(int rows, int cols, const float* A, float* out, int kc, long long so) {
  const int c = blockIdx.x;
  const int r1 = min(rows, (int)blockIdx.y * kc + kc);
  out += blockIdx.y * so;
  float s = 0.f;
  for (int r = blockIdx.y * kc + threadIdx.x; r < r1; r += blockDim.x) s += A[(long long)r * cols + c];
  __shared__ float sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sh[0];
}

__global__ void mse_grad_kernel(long long n, const float* Y, const float* T, float scale, float* dY, double* loss) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (i < n) {
    const float e = Y[i] - T[i];
    dY[i] = scale * e;
    e2 = (double)e * (double)e;
  }
  if (loss) {
    __shared__ double sh[256];
    sh[threadIdx.x] = e2;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
      if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, sh[0]);
  }
}

// Adam step; the gradient is the sum of `splits` partial gradients g[z n + i], z ascending.
__global__ void adam_kernel(long long n, float* p, const float* g, int splits, float* m, float* v, float lr,
                            float b1, float b2, float eps, float b1t, float b2t) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  for (int z = 1; z < splits; ++z) gi += g[z * n + i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ------------------------------------------------------------------ BF16 MFMA inference
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);                                          // round to nearest even
  return (unsigned short)(u >> 16);
}

constexpr int HID = 128;     // hidden width of the MFMA kernel
constexpr int KIN = 16;      // padded input width
constexpr int NOUT = 32;     // padded output width
constexpr int HPAD = HID + 4;
constexpr int KPAD = KIN + 4;

// Weights staged once per block into LDS as BF16, [out][in] with padded rows.
struct MlpSmem {
  unsigned short w1[HID * KPAD];
  unsigned short w2[HID * HPAD];
  unsigned short w3[NOUT * HPAD];
  float b1[HID], b2[HID], b3[NOUT];
};

// Four consecutive BF16 values (8-byte aligned in LDS) as one 64-bit LDS read.
__device__ __forceinline__ bf16x4 lds_bf16x4(const unsigned short* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}

__device__ __forceinline__ float fast_tanh(float x) {
  // 1 - 2 / (e^{2x} + 1): saturates to +-1 (e^{2x} overflows to inf or underflows to 0);
  // its error is far below the BF16 rounding the activation gets next
  return 1.f - 2.f * __frcp_rn(__expf(2.f * x) + 1.f);
}

__device__ __forceinline__ short bf16s(float f) { return (short)f2bf(f); }

// The layers are computed transposed, Z^T = W A^T: the weights (LDS) are the A operand,
// the activations the B operand, so a 32 x 32 tile covers 32 units x 32 samples. In the
// accumulator, lane l holds sample l % 32 and units 8 (r >> 2) + 4 (l >> 5) + (r & 3) of
// its tile (r = 0..15); the B operand of k-step ks wants units 8 ks + 4 (l >> 5) + j, which
// are entries r = 4 (ks % 4) + j of tile ks / 4 on the same lane. So each layer's output,
// through tanh and BF16, feeds the next layer straight from registers.
template <int KSTEPS>
__device__ __forceinline__ f32x16 mfma_wt(const unsigned short* W, int ldw, int n0, const bf16x4* b, f32x16 acc) {
  const int lane = threadIdx.x & 63;
  const int i = lane & 31, kg = lane >> 5;
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ++ks) {
    const bf16x4 a = lds_bf16x4(W + (n0 + i) * ldw + ks * 8 + kg * 4);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b[ks], acc, 0, 0, 0);
    // keep the weight reads of later k-steps from being hoisted: they would hold registers
    // that the other wavefronts of the SIMD need
    __builtin_amdgcn_sched_barrier(0);
  }
  return acc;
}

// The three layers for 32 samples per wavefront (no LDS besides the weights, no
// barriers). bx is layer 1's B operand: lane l holds inputs 8 ks + 4 (l >> 5) + j of sample
// l % 32 in bx[ks][j]. Returns the output layer's accumulator without its bias: lane l
// holds outputs 8 (r >> 2) + 4 (l >> 5) + (r & 3) of sample l % 32 in entry r.
__device__ __forceinline__ f32x16 mlp_layers(const MlpSmem& sm, const bf16x4* bx) {
  const int lane = threadIdx.x & 63;
  const int kg = lane >> 5;
  // hidden layers: h1 / h2 hold the B operand of the next layer, 16 k-steps of 4 units;
  // each 32-unit tile goes through tanh as soon as it is done
  bf16x4 h1[HID / 8], h2[HID / 8];
#pragma unroll
  for (int t = 0; t < HID / 32; ++t) {
    f32x16 z = {0};
    z = mfma_wt<KIN / 8>(sm.w1, KPAD, t * 32, bx, z);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int unit = t * 32 + 8 * q + 4 * kg + j;
        h1[t * 4 + q][j] = bf16s(fast_tanh(z[q * 4 + j] + sm.b1[unit]));
      }
  }
  {
    // layer 2, k outer: the four 32-unit output tiles accumulate side by side
    const int i = lane & 31;
    f32x16 z[HID / 32];
#pragma unroll
    for (int t = 0; t < HID / 32; ++t) z[t] = f32x16{0};
#pragma unroll
    for (int ks = 0; ks < HID / 8; ++ks) {
#pragma unroll
      for (int t = 0; t < HID / 32; ++t) {
        const bf16x4 a = lds_bf16x4(sm.w2 + (t * 32 + i) * HPAD + ks * 8 + kg * 4);
        z[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, h1[ks], z[t], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int t = 0; t < HID / 32; ++t)
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int unit = t * 32 + 8 * q + 4 * kg + j;
          h2[t * 4 + q][j] = bf16s(fast_tanh(z[t][q * 4 + j] + sm.b2[unit]));
        }
  }
  // output layer: one tile of 32 (dout used)
  f32x16 acc = {0};
  return mfma_wt<HID / 8>(sm.w3, HPAD, 0, h2, acc);
}

// Inference for one tile: X holds din scaled inputs per sample (fp32, row-major); Y
// receives dout scaled outputs.
__device__ void mlp_tile(const MlpSmem& sm, int din, int dout, const float* X, float* Y, long long s0, long long n) {
  const int lane = threadIdx.x & 63;
  const int smp = lane & 31, kg = lane >> 5;
  const long long s = s0 + smp;
  const bool live = s < n;
  bf16x4 bx[KIN / 8];
#pragma unroll
  for (int ks = 0; ks < KIN / 8; ++ks)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = ks * 8 + kg * 4 + j;
      bx[ks][j] = bf16s((live && k < din) ? X[s * din + k] : 0.f);
    }
  const f32x16 o = mlp_layers(sm, bx);
  if (live) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int out = 8 * (r >> 2) + 4 * kg + (r & 3);
      if (out < dout) Y[s * dout + out] = o[r] + sm.b3[out];
    }
  }
}

__device__ void stage_weights(MlpSmem& sm, const MlpDims& d, const float* theta) {
  const float* W1 = theta;
  const float* b1 = W1 + HID * d.din;
  const float* W2 = b1 + HID;
  const float* b2 = W2 + HID * HID;
  const float* W3 = b2 + HID;
  const float* b3 = W3 + d.dout * HID;
  for (int e = threadIdx.x; e < HID * KPAD; e += blockDim.x) {
    const int o = e / KPAD, k = e % KPAD;
    sm.w1[e] = f2bf((k < d.din) ? W1[o * d.din + k] : 0.f);
  }
  for (int e = threadIdx.x; e < HID * HPAD; e += blockDim.x) {
    const int o = e / HPAD, k = e % HPAD;
    sm.w2[e] = f2bf((k < HID) ? W2[o * HID + k] : 0.f);
  }
  for (int e = threadIdx.x; e < NOUT * HPAD; e += blockDim.x) {
    const int o = e / HPAD, k = e % HPAD;
    sm.w3[e] = f2bf((o < d.dout && k < HID) ? W3[o * HID + k] : 0.f);
  }
  for (int e = threadIdx.x; e < HID; e += blockDim.x) { sm.b1[e] = b1[e]; sm.b2[e] = b2[e]; }
  for (int e = threadIdx.x; e < NOUT; e += blockDim.x) sm.b3[e] = (e < d.dout) ? b3[e] : 0.f;
  __syncthreads();
}

__global__ void __launch_bounds__(256) mlp_infer_kernel(MlpDims d, long long n, const float* theta, const float* X,
                                                        float* Y) {
  extern __shared__ unsigned char smraw[];
  MlpSmem& sm = *reinterpret_cast<MlpSmem*>(smraw);
  stage_weights(sm, d, theta);
  const int wave = threadIdx.x >> 6;
  const long long tiles = (n + 31) / 32;
  // each wavefront takes its own tiles (after the staging barrier, none is needed)
  for (long long t = (long long)blockIdx.x * 4 + wave; t < tiles; t += (long long)gridDim.x * 4)
    mlp_tile(sm, d.din, d.dout, X, Y, t * 32, n);
}

// Warm start for the batched Newton solve: inputs are log10 totals and the per-voxel logK
// values, scaled with the training box; outputs unscaled into log activities. A voxel with
// any input outside the box gets NaN, which the Newton kernel replaces by the reference
// rule.
//
// One tile of the fused warm start: the inputs are scaled in registers (each lane computes
// the inputs its B operand holds) and the outputs are unscaled as they leave the
// accumulator, so neither the features nor the scaled outputs go through memory. The box
// test runs in FP32 on the FP32 inputs: a voxel within FP32 rounding of the box edge may
// go either way, and both starts converge.
__device__ void mlp_warm_tile(const MlpSmem& sm, int nc, int nvar, int dout, const double* totals,
                              const double* logk_var, const double* xmin, const double* xmax, const double* ymin,
                              const double* ymax, double* logc, long long s0, long long n) {
  const int lane = threadIdx.x & 63;
  const int smp = lane & 31, kg = lane >> 5;
  const long long s = s0 + smp;
  const bool live = s < n;
  const int din = nc + nvar;
  bool ok = true;
  bf16x4 bx[KIN / 8];
#pragma unroll
  for (int ks = 0; ks < KIN / 8; ++ks)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = ks * 8 + kg * 4 + j;
      float f = 0.f;
      if (live && k < din) {
        // FP32 logarithm: the input is rounded to BF16 (8-bit mantissa) next
        const float raw = (k < nc) ? __log10f((float)fmax(totals[s * nc + k], 1e-30))
                                   : (float)logk_var[s * nvar + (k - nc)];
        const float lo = (float)xmin[k], hi = (float)xmax[k];
        if (raw < lo || raw > hi) ok = false;
        f = (hi > lo) ? (raw - lo) * (2.f / (hi - lo)) - 1.f : 0.f;
      }
      bx[ks][j] = bf16s(f);
    }
  // lanes l and l ^ 32 hold the two halves of a sample's inputs
  ok = ok && (__shfl_xor((int)ok, 32) != 0);
  const f32x16 o = mlp_layers(sm, bx);
  if (live) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int out = 8 * (r >> 2) + 4 * kg + (r & 3);
      if (out < dout) {
        const double y = (double)(o[r] + sm.b3[out]);
        logc[s * dout + out] = ok ? ymin[out] + 0.5 * (y + 1.0) * (ymax[out] - ymin[out]) : __builtin_nan("");
      }
    }
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) mlp_warm_kernel(MlpDims d, long long n, const float* theta, int nc, int nvar,
                                                       const double* totals, const double* logk_var,
                                                       const double* xmin, const double* xmax, const double* ymin,
                                                       const double* ymax, double* logc) {
  extern __shared__ unsigned char smraw[];
  MlpSmem& sm = *reinterpret_cast<MlpSmem*>(smraw);
  stage_weights(sm, d, theta);
  const int wave = threadIdx.x >> 6;
  const long long tiles = (n + 31) / 32;
  for (long long t = (long long)blockIdx.x * 4 + wave; t < tiles; t += (long long)gridDim.x * 4)
    mlp_warm_tile(sm, nc, nvar, d.dout, totals, logk_var, xmin, xmax, ymin, ymax, logc, t * 32, n);
}

// Training pairs from a solved batch: X = scaled (log10 T, logK), T = scaled log activities.
__global__ void mlp_dataset_kernel(long long n, int nc, int nvar, const double* totals, const double* logk_var,
                                   const double* logc, const double* xmin, const double* xmax, const double* ymin,
                                   const double* ymax, float* X, float* T) {
  const long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int din = nc + nvar;
  for (int k = 0; k < din; ++k) {
    const double raw = (k < nc) ? log10(fmax(totals[v * nc + k], 1e-30)) : logk_var[v * nvar + (k - nc)];
    const double span = xmax[k] - xmin[k];
    X[v * din + k] = (float)((span > 0) ? 2.0 * (raw - xmin[k]) / span - 1.0 : 0.0);
  }
  for (int k = 0; k < nc; ++k) {
    const double span = ymax[k] - ymin[k];
    T[v * nc + k] = (float)((span > 0) ? 2.0 * (logc[v * nc + k] - ymin[k]) / span - 1.0 : 0.0);
  }
}

}  // namespace

void launch_gemm_split(int transA, int transB, int M, int N, int K, const float* A, int lda, const float* B,
                       int ldb, float* C, int ldc, const float* bias, int act, float beta, int splits, long long sc,
                       hipStream_t s) {
  const int kc = (K + splits - 1) / splits;
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, splits);
#define CLB_GEMM(TA_, TB_) \
  hipLaunchKernelGGL((gemm_kernel<TA_, TB_>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, act, beta, kc, sc)
  if (!transA && !transB) CLB_GEMM(false, false);
  else if (!transA && transB) CLB_GEMM(false, true);
  else if (transA && !transB) CLB_GEMM(true, false);
  else CLB_GEMM(true, true);
#undef CLB_GEMM
}

void launch_gemm(int transA, int transB, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                 float* C, int ldc, const float* bias, int act, float beta, hipStream_t s) {
  launch_gemm_split(transA, transB, M, N, K, A, lda, B, ldb, C, ldc, bias, act, beta, 1, 0, s);
}

void launch_tanh_backward(long long n, float* dz, const float* h, hipStream_t s) {
  hipLaunchKernelGGL(tanh_backward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, dz, h);
}
void launch_colsum(int rows, int cols, const float* A, float* out, int splits, long long so, hipStream_t s) {
  const int kc = (rows + splits - 1) / splits;
  hipLaunchKernelGGL(colsum_kernel, dim3(cols, splits), dim3(256), 0, s, rows, cols, A, out, kc, so);
}
void launch_mse_grad(long long n, const float* Y, const float* T, float scale, float* dY, double* loss,
                     hipStream_t s) {
  hipLaunchKernelGGL(mse_grad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, Y, T, scale, dY, loss);
}
void launch_adam(long long n, float* p, const float* g, int splits, float* m, float* v, float lr, float b1, float b2,
                 float eps, float b1t, float b2t, hipStream_t s) {
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, p, g, splits, m, v, lr, b1,
                     b2, eps, b1t, b2t);
}

// The fused kernel stages ~82 KB of weights and activations in LDS (CDNA4: 160 KB per
// CU); raise the per-launch dynamic LDS limit once.
static void mlp_smem_attr() {
  static bool done = false;
  if (!done) {
    (void)hipFuncSetAttribute((const void*)mlp_infer_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sizeof(MlpSmem));
    done = true;
  }
}

static int mlp_grid(long long n) {
  const long long tiles = (n + 31) / 32;
  long long blocks = (tiles + 3) / 4;
  if (blocks > 2048) blocks = 2048;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}

void launch_mlp_infer_bf16(const MlpDims& d, long long n, const float* theta, const float* X, float* Y,
                           hipStream_t s) {
  mlp_smem_attr();
  hipLaunchKernelGGL(mlp_infer_kernel, dim3(mlp_grid(n)), dim3(256), sizeof(MlpSmem), s, d, n, theta, X, Y);
}

void launch_mlp_warm(const MlpDims& d, long long n, const float* theta, const double* xmin, const double* xmax,
                     const double* ymin, const double* ymax, const Tableau& tb, const SpecArgs& sa, hipStream_t s) {
  // one fused kernel: features, the three layers and the unscaling per 32-sample tile
  mlp_smem_attr();
  hipLaunchKernelGGL(mlp_warm_kernel, dim3(mlp_grid(n)), dim3(256), sizeof(MlpSmem), s, d, n, theta, tb.nc, sa.nvar,
                     sa.totals, sa.logk_var, xmin, xmax, ymin, ymax, sa.logc);
}

void launch_mlp_make_dataset(const MlpDims& d, long long n, const Tableau& tb, const SpecArgs& sa, float* X, float* T,
                             const double* xmin, const double* xmax, const double* ymin, const double* ymax,
                             hipStream_t s) {
  (void)d;
  hipLaunchKernelGGL(mlp_dataset_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, tb.nc, sa.nvar,
                     sa.totals, sa.logk_var, sa.logc, xmin, xmax, ymin, ymax, X, T);
}

}  // namespace clb
