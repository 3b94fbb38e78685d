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

// The sum of `splits` partials g[z gs + i] in a fixed order: partial z goes to accumulator
// z mod 8, and the eight are added pairwise; eight loads are in flight at a time.
__device__ __forceinline__ float sum_partials(const float* g, int splits, long long gs, long long i) {
  float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int z = 0;
  for (; z + 8 <= splits; z += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += g[(z + u) * gs + i];
  }
#pragma unroll
  for (int u = 0; u < 8; ++u)
    if (z + u < splits) a[u] += g[(z + u) * gs + i];
  return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// Adam step; the gradient is the sum of `splits` partial gradients (sum_partials).
__global__ void adam_kernel(long long n, float* p, const float* g, int splits, long long gs, float* m, float* v,
                            float lr, float b1, float b2, float eps, float b1t, float b2t) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = sum_partials(g, splits, gs, i);
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}


// ------------------------------------------------------------------ fused training step
// One minibatch's forward pass, loss gradient and backward pass in one launch, for H = 128,
// din <= 16 and dout <= 16 (the warm-start network). Workgroup z takes rows
// [z kStepRows, (z + 1) kStepRows) of the minibatch, stages W1, W2, W3 and its rows'
// activations in LDS, and writes its partial gradient, in the layout of theta, to
// grad + z gs (gs >= P, a multiple of 4); the Adam kernel sums the partials in ascending z,
// so training stays deterministic. Each workgroup adds its squared errors to *loss with one atomic.
//
// The layers are small (a workgroup's rows hold 8 x 128 activations), so the launches of a
// GEMM per layer and pass cost more than the arithmetic; this kernel replaces the 14
// launches of a step by one.
constexpr int kStepRows = 8;
constexpr int FH = 128;          // hidden width
constexpr int FI = 16;           // padded input / output width
constexpr int W2S = FH + 1;      // LDS row strides: rows read by consecutive lanes and
constexpr int W1S = FI + 1;      // columns read by consecutive lanes are both free of
constexpr int W3S = FH + 1;      // bank conflicts

struct StepSmem {
  float w2[FH * W2S];
  float w1[FH * W1S];
  float w3[FI * W3S];
  float x[kStepRows][FI];
  float h1[kStepRows][FH];
  float h2[kStepRows][FH];  // H2, then dZ1
  float dz[kStepRows][FH];  // dZ2
  float dy[kStepRows][FI];
  double red[4];
};

__global__ void __launch_bounds__(256) mlp_step_kernel(int din, int dout, long long B, long long row0,
                                                       const float* __restrict__ X, const float* __restrict__ T,
                                                       const float* __restrict__ theta, float* __restrict__ grad,
                                                       long long gs, float scale, double* loss) {
  constexpr int R = kStepRows, RH = R / 2;
  extern __shared__ __align__(16) unsigned char step_smem[];
  StepSmem& sm = *reinterpret_cast<StepSmem*>(step_smem);
  const int t = threadIdx.x;
  const long long r0 = row0 + (long long)blockIdx.x * R;  // first row (of X and T) of this workgroup
  const long long rend = row0 + B;
  const float* W1 = theta;
  const float* b1 = W1 + FH * din;
  const float* W2 = b1 + FH;
  const float* b2 = W2 + FH * FH;
  const float* W3 = b2 + FH;
  const float* b3 = W3 + dout * FH;
  float* g = grad + (long long)blockIdx.x * gs;  // gs: a multiple of 4 floats (16-byte stores)
  float* gW1 = g;
  float* gb1 = gW1 + FH * din;
  float* gW2 = gb1 + FH;
  float* gb2 = gW2 + FH * FH;
  float* gW3 = gb2 + FH;
  float* gb3 = gW3 + dout * FH;

  // Stage the weights: every load is issued before the first LDS store, so the loads'
  // latencies overlap (one load / store pair at a time would pay ~100 of them in turn).
  // W2 starts at a multiple of 4 floats (128 (din + 1)), so its float4 loads are aligned.
  {
    float4 v[FH * FH / 4 / 256];
#pragma unroll
    for (int u = 0; u < FH * FH / 4 / 256; ++u) v[u] = reinterpret_cast<const float4*>(W2)[t + 256 * u];
    float w1v[FH * FI / 256], w3v[FI * FH / 256];
#pragma unroll
    for (int u = 0; u < FH * FI / 256; ++u) {
      const int q = t + 256 * u, j = q / FI, i = q % FI;
      w1v[u] = (i < din) ? W1[j * din + i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < FI * FH / 256; ++u) {
      const int q = t + 256 * u, o = q / FH, k = q % FH;
      w3v[u] = (o < dout) ? W3[o * FH + k] : 0.f;
    }
    float xv = 0.f;
    if (t < R * FI) {
      const int r = t / FI, i = t % FI;
      if (r0 + r < rend && i < din) xv = X[(r0 + r) * din + i];
    }
#pragma unroll
    for (int u = 0; u < FH * FH / 4 / 256; ++u) {
      const int q = 4 * (t + 256 * u), j = q >> 7, k = q & (FH - 1);
      sm.w2[j * W2S + k] = v[u].x;
      sm.w2[j * W2S + k + 1] = v[u].y;
      sm.w2[j * W2S + k + 2] = v[u].z;
      sm.w2[j * W2S + k + 3] = v[u].w;
    }
#pragma unroll
    for (int u = 0; u < FH * FI / 256; ++u) {
      const int q = t + 256 * u;
      sm.w1[(q / FI) * W1S + q % FI] = w1v[u];
    }
#pragma unroll
    for (int u = 0; u < FI * FH / 256; ++u) {
      const int q = t + 256 * u;
      sm.w3[(q / FH) * W3S + q % FH] = w3v[u];
    }
    if (t < R * FI) sm.x[t / FI][t % FI] = xv;
  }
  __syncthreads();

  const int j = t & (FH - 1);  // this thread's unit
  const int rh = t >> 7;       // and rows rh, rh + 2, ...
  {
    // layer 1: h1 = tanh(W1 x + b1)
    float z[RH];
#pragma unroll
    for (int q = 0; q < RH; ++q) z[q] = b1[j];
#pragma unroll
    for (int i = 0; i < FI; ++i) {
      const float w = sm.w1[j * W1S + i];
#pragma unroll
      for (int q = 0; q < RH; ++q) z[q] = fmaf(sm.x[rh + 2 * q][i], w, z[q]);
    }
#pragma unroll
    for (int q = 0; q < RH; ++q) sm.h1[rh + 2 * q][j] = tanhf(z[q]);
  }
  __syncthreads();
  {
    // layer 2: h2 = tanh(W2 h1 + b2)
    float z[RH];
#pragma unroll
    for (int q = 0; q < RH; ++q) z[q] = b2[j];
#pragma unroll 8
    for (int k = 0; k < FH; ++k) {
      const float w = sm.w2[j * W2S + k];
#pragma unroll
      for (int q = 0; q < RH; ++q) z[q] = fmaf(sm.h1[rh + 2 * q][k], w, z[q]);
    }
#pragma unroll
    for (int q = 0; q < RH; ++q) sm.h2[rh + 2 * q][j] = tanhf(z[q]);
  }
  __syncthreads();
  // output layer and the loss gradient dY = scale (Y - T)
  double e2 = 0.0;
  if (t < R * FI) {
    const int r = t / FI, o = t % FI;
    float dyv = 0.f;
    if (r0 + r < rend && o < dout) {
      float y = b3[o];
#pragma unroll 8
      for (int k = 0; k < FH; ++k) y = fmaf(sm.h2[r][k], sm.w3[o * W3S + k], y);
      const float e = y - T[(r0 + r) * dout + o];
      dyv = scale * e;
      e2 = (double)e * (double)e;
    }
    sm.dy[r][o] = dyv;
  }
  __syncthreads();
  // output layer's gradient; dZ2 = (W3^T dY) (1 - h2^2)
  for (int q = t; q < dout * FH; q += 256) {
    const int o = q / FH, k = q % FH;
    float a = 0.f;
#pragma unroll
    for (int r = 0; r < R; ++r) a = fmaf(sm.dy[r][o], sm.h2[r][k], a);
    gW3[o * FH + k] = a;
  }
  if (t < dout) {
    float a = 0.f;
#pragma unroll
    for (int r = 0; r < R; ++r) a += sm.dy[r][t];
    gb3[t] = a;
  }
#pragma unroll
  for (int q = 0; q < RH; ++q) {
    const int r = rh + 2 * q;
    float a = 0.f;
#pragma unroll
    for (int o = 0; o < FI; ++o) a = fmaf(sm.dy[r][o], sm.w3[o * W3S + j], a);
    const float h = sm.h2[r][j];
    sm.dz[r][j] = a * (1.f - h * h);
  }
  __syncthreads();
  {
    // layer 2's gradient: thread t owns units [8 (t >> 4), + 8) x inputs [8 (t & 15), + 8)
    const int j0 = (t >> 4) * 8, k0 = (t & 15) * 8;
    float acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int b = 0; b < 8; ++b) acc[a][b] = 0.f;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const float4 a0 = *reinterpret_cast<const float4*>(&sm.dz[r][j0]);
      const float4 a1 = *reinterpret_cast<const float4*>(&sm.dz[r][j0 + 4]);
      const float4 c0 = *reinterpret_cast<const float4*>(&sm.h1[r][k0]);
      const float4 c1 = *reinterpret_cast<const float4*>(&sm.h1[r][k0 + 4]);
      const float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
      const float cv[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = fmaf(av[a], cv[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      float4* dst = reinterpret_cast<float4*>(gW2 + (j0 + a) * FH + k0);
      dst[0] = make_float4(acc[a][0], acc[a][1], acc[a][2], acc[a][3]);
      dst[1] = make_float4(acc[a][4], acc[a][5], acc[a][6], acc[a][7]);
    }
  }
  if (t < FH) {
    float a = 0.f;
#pragma unroll
    for (int r = 0; r < R; ++r) a += sm.dz[r][t];
    gb2[t] = a;
  }
  {
    // dZ1 = (W2^T dZ2) (1 - h1^2), into the h2 rows
    float a[RH];
#pragma unroll
    for (int q = 0; q < RH; ++q) a[q] = 0.f;
#pragma unroll 8
    for (int u = 0; u < FH; ++u) {
      const float w = sm.w2[u * W2S + j];
#pragma unroll
      for (int q = 0; q < RH; ++q) a[q] = fmaf(sm.dz[rh + 2 * q][u], w, a[q]);
    }
#pragma unroll
    for (int q = 0; q < RH; ++q) {
      const float h = sm.h1[rh + 2 * q][j];
      sm.h2[rh + 2 * q][j] = a[q] * (1.f - h * h);
    }
  }
  __syncthreads();
  // layer 1's gradient
  for (int q = t; q < FH * din; q += 256) {
    const int u = q / din, i = q % din;
    float a = 0.f;
#pragma unroll
    for (int r = 0; r < R; ++r) a = fmaf(sm.h2[r][u], sm.x[r][i], a);
    gW1[u * din + i] = a;
  }
  if (t < FH) {
    float a = 0.f;
#pragma unroll
    for (int r = 0; r < R; ++r) a += sm.h2[r][t];
    gb1[t] = a;
  }
  if (loss) {
    for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off);
    if ((t & 63) == 0) sm.red[t >> 6] = e2;
    __syncthreads();
    if (t == 0) atomicAdd(loss, (sm.red[0] + sm.red[1]) + (sm.red[2] + sm.red[3]));
  }
}

// out[i] (+)= sum of the partials g[z gs + i] (sum_partials).
__global__ void partial_sum_kernel(long long P, const float* g, int splits, long long gs, float* out,
                                   int accumulate) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float a = sum_partials(g, splits, gs, i);
  out[i] = accumulate ? out[i] + a : a;
}

// ------------------------------------------------------------------ BF16 MFMA inference
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);                                          // round to nearest even
  return (unsigned short)(u >> 16);
}

constexpr int HID = 128;     // hidden width of the MFMA kernel
constexpr int KIN = 16;      // padded input width (one K = 16 step)
constexpr int NOUT = 32;     // padded output width
constexpr int HPAD = HID + 8;  // row strides keep 16-byte rows and spread the rows over the banks
constexpr int KPAD = KIN + 8;

// The layers use v_mfma_f32_32x32x16_bf16 on the transposed product Z^T = W A^T: the
// weights (LDS) are the A operand, the activations the B operand, so a 32 x 32 tile covers
// 32 units x 32 samples. Lane l works on sample l % 32 with lane group g = l >> 5. The
// accumulator of a tile holds, in entry r, unit 8 (r >> 2) + 4 g + (r & 3). In k-step ks
// the MFMA pairs A[i][8 g + e] with B[8 g + e][n] (e = 0..7), and any assignment of units
// to these slots gives the same sum when A and B use the same one. Slot (ks, g, e) is given
// unit 16 ks + 8 (e >> 2) + 4 g + (e & 3): then a lane's B operand for k-step ks is entries
// 8 (ks % 2) .. 8 (ks % 2) + 7 of tile ks / 2 of the previous layer's accumulator, on the
// same lane, and each layer's output feeds the next one from registers. The weights are
// staged with their columns in slot order, so an A operand is one 16-byte LDS read.
__host__ __device__ constexpr int slot_unit(int p) {  // column p = 16 ks + 8 g + e
  return (p & ~15) + 8 * ((p & 7) >> 2) + 4 * ((p >> 3) & 1) + (p & 3);
}

// Weights staged once per block into LDS as BF16, [out][slot] with padded rows.
struct MlpSmem {
  alignas(16) unsigned short w1[HID * KPAD];
  alignas(16) unsigned short w2[HID * HPAD];
  alignas(16) unsigned short w3[NOUT * HPAD];
  float b1[HID], b2[HID], b3[NOUT];
};

__device__ __forceinline__ bf16x8 lds_bf16x8(const unsigned short* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}

__device__ __forceinline__ float fast_tanh(float x) {
  // 1 - 2 / (e^{2x} + 1) with the hardware reciprocal (1 ulp; a correctly rounded divide
  // costs ten instructions more): saturates to +-1 (e^{2x} overflows to inf or underflows
  // to 0); its error is far below the BF16 rounding the activation gets next
  return 1.f - 2.f * __builtin_amdgcn_rcpf(__expf(2.f * x) + 1.f);
}

template <int KSTEPS>
__device__ __forceinline__ f32x16 mfma_wt(const unsigned short* W, int ldw, int n0, const bf16x8* b, f32x16 acc) {
  const int lane = threadIdx.x & 63;
  const int i = lane & 31, g = lane >> 5;
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ++ks) {
    const bf16x8 a = lds_bf16x8(W + (n0 + i) * ldw + ks * 16 + g * 8);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[ks], acc, 0, 0, 0);
    // keep the weight reads of later k-steps from being hoisted: they would hold registers
    // that the other wavefronts of the SIMD need
    __builtin_amdgcn_sched_barrier(0);
  }
  return acc;
}

// Hidden activation of one 32-unit tile into the next layer's B operand (k-steps 2t, 2t+1).
__device__ __forceinline__ void tile_to_operand(const f32x16& z, const float* bias, int t, bf16x8* h) {
  const int g = (threadIdx.x & 63) >> 5;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int unit = t * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
    h[2 * t + (r >> 3)][r & 7] = (__bf16)fast_tanh(z[r] + bias[unit]);
  }
}

// The three layers for 32 samples per wavefront (no LDS besides the weights, no
// barriers). bx is layer 1's B operand: lane l holds input slot_unit(8 g + e) of sample
// l % 32 in bx[e]. Returns the output layer's accumulator without its bias: lane l holds
// outputs 8 (r >> 2) + 4 g + (r & 3) of sample l % 32 in entry r.
__device__ __forceinline__ f32x16 mlp_layers(const MlpSmem& sm, const bf16x8& bx) {
  const int lane = threadIdx.x & 63;
  const int i = lane & 31, g = lane >> 5;
  bf16x8 h1[HID / 16], h2[HID / 16];
#pragma unroll
  for (int t = 0; t < HID / 32; ++t) {
    f32x16 z = {0};
    z = mfma_wt<1>(sm.w1, KPAD, t * 32, &bx, z);
    tile_to_operand(z, sm.b1, t, h1);
  }
  {
    // layer 2, k outer: the four 32-unit output tiles accumulate side by side
    f32x16 z[HID / 32];
#pragma unroll
    for (int t = 0; t < HID / 32; ++t) z[t] = f32x16{0};
#pragma unroll
    for (int ks = 0; ks < HID / 16; ++ks) {
#pragma unroll
      for (int t = 0; t < HID / 32; ++t) {
        const bf16x8 a = lds_bf16x8(sm.w2 + (t * 32 + i) * HPAD + ks * 16 + g * 8);
        z[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, h1[ks], z[t], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int t = 0; t < HID / 32; ++t) tile_to_operand(z[t], sm.b2, t, h2);
  }
  // output layer: one tile of 32 (dout used)
  f32x16 acc = {0};
  return mfma_wt<HID / 16>(sm.w3, HPAD, 0, h2, acc);
}

// Inference for one tile: X holds din scaled inputs per sample (fp32, row-major); Y
// receives dout scaled outputs.
__device__ void mlp_tile(const MlpSmem& sm, int din, int dout, const float* X, float* Y, long long s0, long long n) {
  const int lane = threadIdx.x & 63;
  const int smp = lane & 31, g = lane >> 5;
  const long long s = s0 + smp;
  const bool live = s < n;
  bf16x8 bx;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = slot_unit(8 * g + e);
    bx[e] = (__bf16)((live && k < din) ? X[s * din + k] : 0.f);
  }
  const f32x16 o = mlp_layers(sm, bx);
  if (live) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int out = 8 * (r >> 2) + 4 * g + (r & 3);
      if (out < dout) Y[s * dout + out] = o[r] + sm.b3[out];
    }
  }
}

// Column of unit k: slot_unit swaps bits 2 and 3 of the index, so it is its own inverse
// and keeps groups of four consecutive units together.
__device__ __forceinline__ int unit_slot(int k) { return slot_unit(k); }

// Four FP32 weights of consecutive units as BF16 into LDS (one 8-byte store).
__device__ __forceinline__ void store_bf16x4(unsigned short* p, const float4& v) {
  const unsigned int lo = (unsigned int)f2bf(v.x) | ((unsigned int)f2bf(v.y) << 16);
  const unsigned int hi = (unsigned int)f2bf(v.z) | ((unsigned int)f2bf(v.w) << 16);
  *reinterpret_cast<uint2*>(p) = make_uint2(lo, hi);
}

// Stage the weights (256 threads). W2 and W3 rows are read as float4 (every row starts on
// a 16-byte boundary: theta is 16-byte aligned and the offsets are multiples of 128
// floats), all of a thread's loads issued before its stores.
__device__ void stage_weights(MlpSmem& sm, const MlpDims& d, const float* theta) {
  const float* W1 = theta;
  const float* b1 = W1 + HID * d.din;
  const float* W2 = b1 + HID;
  const float* b2 = W2 + HID * HID;
  const float* W3 = b2 + HID;
  const float* b3 = W3 + d.dout * HID;
  const int tid = threadIdx.x;
  {
    constexpr int PER = HID * HID / 4 / 256;
    float4 v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) v[q] = reinterpret_cast<const float4*>(W2)[tid + q * 256];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int idx = tid + q * 256, o = idx >> 5, k = (idx & 31) * 4;
      store_bf16x4(&sm.w2[o * HPAD + unit_slot(k)], v[q]);
    }
  }
  {
    constexpr int PER = NOUT * HID / 4 / 256;
    float4 v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int idx = tid + q * 256, o = idx >> 5;
      v[q] = (o < d.dout) ? reinterpret_cast<const float4*>(W3)[idx] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int idx = tid + q * 256, o = idx >> 5, k = (idx & 31) * 4;
      store_bf16x4(&sm.w3[o * HPAD + unit_slot(k)], v[q]);
    }
  }
  {
    constexpr int PER = HID * KIN / 256;
    float v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * 256, o = e / KIN, k = slot_unit(e % KIN);
      v[q] = (k < d.din) ? W1[o * d.din + k] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = tid + q * 256;
      sm.w1[(e / KIN) * KPAD + e % KIN] = f2bf(v[q]);
    }
  }
  if (tid < HID) {
    sm.b1[tid] = b1[tid];
    sm.b2[tid] = b2[tid];
  }
  if (tid < NOUT) sm.b3[tid] = (tid < d.dout) ? b3[tid] : 0.f;
  __syncthreads();
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
mlp_infer_kernel(MlpDims d, long long n, const float* theta, const float* X, float* Y) {
  extern __shared__ __align__(16) unsigned char smraw[];
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
  const int smp = lane & 31, g = lane >> 5;
  const long long s = s0 + smp;
  const bool live = s < n;
  const int din = nc + nvar;
  bool ok = true;
  bf16x8 bx;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = slot_unit(8 * g + e);
    float f = 0.f;
    if (live && k < din) {
      // FP32 logarithm: the input is rounded to BF16 (8-bit mantissa) next
      const float raw = (k < nc) ? __log10f((float)fmax(totals[s * nc + k], 1e-30))
                                 : (float)logk_var[s * nvar + (k - nc)];
      const float lo = (float)xmin[k], hi = (float)xmax[k];
      if (raw < lo || raw > hi) ok = false;
      f = (hi > lo) ? (raw - lo) * (2.f / (hi - lo)) - 1.f : 0.f;
    }
    bx[e] = (__bf16)f;
  }
  // lanes l and l ^ 32 hold the two halves of a sample's inputs
  ok = ok && (__shfl_xor((int)ok, 32) != 0);
  const f32x16 o = mlp_layers(sm, bx);
  if (live) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int out = 8 * (r >> 2) + 4 * g + (r & 3);
      if (out < dout) {
        const double y = (double)(o[r] + sm.b3[out]);
        logc[s * dout + out] = ok ? ymin[out] + 0.5 * (y + 1.0) * (ymax[out] - ymin[out]) : __builtin_nan("");
      }
    }
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
mlp_warm_kernel(MlpDims d, long long n, const float* theta, int nc, int nvar, const double* totals,
                const double* logk_var, const double* xmin, const double* xmax, const double* ymin,
                const double* ymax, double* logc) {
  extern __shared__ __align__(16) unsigned char smraw[];
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
void launch_adam(long long n, float* p, const float* g, int splits, long long gs, float* m, float* v, float lr,
                 float b1, float b2, float eps, float b1t, float b2t, hipStream_t s) {
  // 64-thread blocks spread the partial sums over more CUs
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, n, p, g, splits, gs, m, v, lr,
                     b1, b2, eps, b1t, b2t);
}

// The fused step's LDS (about 96 KB) is above the default limit; allow it once.
static void mlp_step_attr() {
  static const bool done = [] {
    (void)hipFuncSetAttribute((const void*)mlp_step_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sizeof(StepSmem));
    return true;
  }();
  (void)done;
}

bool mlp_step_fits(const MlpDims& d) { return d.H == FH && d.din <= FI && d.dout <= FI; }

int mlp_step_rows() { return kStepRows; }

void launch_mlp_step(const MlpDims& d, long long B, long long row0, const float* X, const float* T,
                     const float* theta, float* grad, long long gs, float scale, double* loss, hipStream_t s) {
  mlp_step_attr();
  const unsigned blocks = (unsigned)((B + kStepRows - 1) / kStepRows);
  hipLaunchKernelGGL(mlp_step_kernel, dim3(blocks), dim3(256), sizeof(StepSmem), s, d.din, d.dout, B, row0, X, T,
                     theta, grad, gs, scale, loss);
}

void launch_partial_sum(long long P, const float* g, int splits, long long gs, float* out, int accumulate,
                        hipStream_t s) {
  hipLaunchKernelGGL(partial_sum_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), 0, s, P, g, splits, gs, out,
                     accumulate);
}

// The inference kernels stage ~50 KB of BF16 weights in LDS per block (three blocks per
// CU); allow the dynamic LDS size explicitly once.
static void mlp_smem_attr() {
  // a function-local static is initialised once, thread-safely
  static const bool done = [] {
    (void)hipFuncSetAttribute((const void*)mlp_infer_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sizeof(MlpSmem));
    (void)hipFuncSetAttribute((const void*)mlp_warm_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sizeof(MlpSmem));
    return true;
  }();
  (void)done;
}

// One resident wave of blocks (three per CU): each block stages the weights once and its
// wavefronts then take tiles in turn.
static int mlp_grid(long long n) {
  // sized once (thread-safe static initialisation); any grid size is correct, since the
  // wavefronts loop over the tiles
  static const int resident = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || cus < 1)
      cus = 256;
    return 3 * cus;
  }();
  const long long tiles = (n + 31) / 32;
  long long blocks = (tiles + 3) / 4;
  if (blocks > resident) blocks = resident;
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
