// BF16 MLP kernels for the chemistry surrogate (no reference counterpart).
//
// gemm_bf16: C[M][N] = sum_k A(m,k) B(n,k) on CDNA4 matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation).  A workgroup of 4 waves
// owns a 64x64 tile of C, each wave one 32x32 MFMA tile; K advances 16 per
// MFMA.  Operands are read straight from global memory (they are small and
// L2-resident at these shapes): layout N means the reduction index is
// contiguous (A[m*lda+k], B[n*ldb+k]) and one 16-byte load fills a lane's
// fragment; layout T means the row index is contiguous (A[k*lda+m],
// B[k*ldb+n]) and a fragment is 8 coalesced 2-byte loads.
//
// Lane layout of the 32x32x16 bf16 MFMA: lane l supplies A[l%32][8*(l/32)+j]
// and B[8*(l/32)+j][l%32], j=0..7, and receives D[8*(r/4)+4*(l/32)+r%4][l%32],
// r=0..15 (checked on the GPU by tests/test_gpu_surrogate.py against numpy).
//
// Epilogues fuse what follows each product in the network:
//   EPI_BIAS_GELU : z = acc + b[n];  Z[m][n] = z (f32);  O[m][n] = gelu(z) (bf16)
//   EPI_BIAS      : O32[m][n] = acc + b[n]
//   EPI_GELU_BWD  : d = acc * gelu'(Z[m][n]);  O32 = d;  O[m][n] = d (bf16)
//   EPI_ACCUM     : atomicAdd(O32[m][n], acc)   (split-K weight gradients)
//   EPI_BIAS_GELU_INF : O[m][n] = gelu(acc + b[n]) (bf16), inference
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace brams {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

enum { LAY_N = 0, LAY_T = 1 };
enum { EPI_BIAS_GELU = 0, EPI_BIAS = 1, EPI_GELU_BWD = 2, EPI_ACCUM = 3, EPI_BIAS_GELU_INF = 4 };

__device__ __forceinline__ float gelu_f(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad_f(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * __expf(-0.5f * x * x);
  return cdf + x * pdf;
}

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

template <int LAY>
__device__ __forceinline__ bf16x8_t load_frag(const __bf16* __restrict__ p, long long ld, long long row,
                                              long long k) {
  if constexpr (LAY == LAY_N) {
    return *reinterpret_cast<const bf16x8_t*>(p + row * ld + k);
  } else {
    bf16x8_t v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = p[(k + j) * ld + row];
    return v;
  }
}

struct GemmArgs {
  int M, N, K;                 // M, N multiples of 64; K multiple of 16
  const __bf16* A; long long lda;
  const __bf16* B; long long ldb;
  const float* bias;           // [N] (EPI_BIAS*)
  float* Z; long long ldz;     // EPI_BIAS_GELU: out pre-activation; EPI_GELU_BWD: in
  float* O32; long long ldo32; // f32 output
  __bf16* O; long long ldo;    // bf16 output
  int ksplit;                  // EPI_ACCUM: number of K chunks (gridDim.z)
};

template <int LA, int LB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16(GemmArgs g) {
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const long long m0 = (long long)blockIdx.y * 64 + (wave >> 1) * 32;
  const long long n0 = (long long)blockIdx.x * 64 + (wave & 1) * 32;
  const int kchunk = g.K / g.ksplit;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = kbeg + kchunk;
  const long long arow = m0 + (lane & 31);
  const long long bcol = n0 + (lane & 31);
  const int koff = 8 * (lane >> 5);
  f32x16_t acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (int k = kbeg; k < kend; k += 16) {
    const bf16x8_t a = load_frag<LA>(g.A, g.lda, arow, k + koff);
    const bf16x8_t b = load_frag<LB>(g.B, g.ldb, bcol, k + koff);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  const long long n = bcol;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long long m = m0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const float v = acc[r];
    if constexpr (EPI == EPI_BIAS_GELU) {
      const float z = v + g.bias[n];
      g.Z[m * g.ldz + n] = z;
      g.O[m * g.ldo + n] = to_bf16(gelu_f(z));
    } else if constexpr (EPI == EPI_BIAS_GELU_INF) {
      g.O[m * g.ldo + n] = to_bf16(gelu_f(v + g.bias[n]));
    } else if constexpr (EPI == EPI_BIAS) {
      g.O32[m * g.ldo32 + n] = v + g.bias[n];
    } else if constexpr (EPI == EPI_GELU_BWD) {
      const float d = v * gelu_grad_f(g.Z[m * g.ldz + n]);
      g.O32[m * g.ldo32 + n] = d;
      g.O[m * g.ldo + n] = to_bf16(d);
    } else {
      atomicAdd(&g.O32[m * g.ldo32 + n], v);
    }
  }
}

// ---------------------------------------------------------------- helpers
__global__ void fill_f32(float* p, long long n, float v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

// column sums out[n] += sum_m X[m][n] over a 64-column x 512-row tile per
// block (out must be zeroed first); 4 row phases of 64 lanes each
__global__ __launch_bounds__(256) void colsum_f32(const float* X, long long ld, int M, int N,
                                                  float* out) {
  __shared__ float part[4][64];
  const int c = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + c;
  const int m0 = blockIdx.y * 512;
  float s = 0.f;
  if (n < N)
    for (int m = m0 + ph; m < min(M, m0 + 512); m += 4) s += X[(long long)m * ld + n];
  part[ph][c] = s;
  __syncthreads();
  if (ph == 0 && n < N) atomicAdd(&out[n], part[0][c] + part[1][c] + part[2][c] + part[3][c]);
}

__device__ __forceinline__ float block_sum_256(float v, float* sh) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return sh[0] + sh[1] + sh[2] + sh[3];
}

// MSE on the first nvalid outputs: d = 2 (o - y) / (M nvalid); padded
// columns get 0.  The loss sum is reduced per block, one atomic per block.
__global__ __launch_bounds__(256) void mse_grad(const float* out, const float* y, int M, int npad,
                                                int nvalid, float* d, __bf16* dbf, float* loss) {
  __shared__ float sh[4];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (i < (long long)M * npad) {
    const int n = (int)(i % npad);
    float v = 0.f;
    if (n < nvalid) {
      const float e = out[i] - y[i];
      v = 2.0f * e / ((float)M * (float)nvalid);
      e2 = e * e / ((float)M * (float)nvalid);
    }
    d[i] = v;
    dbf[i] = to_bf16(v);
  }
  const float s = block_sum_256(e2, sh);
  if (threadIdx.x == 0 && loss) atomicAdd(loss, s);
}

// gather rows of a dataset into a batch: Xb[b] = X[idx[b]] (bf16 rows), Yb likewise (f32)
__global__ void gather_rows(const __bf16* X, int kx, const float* Y, int ky, const long long* idx,
                            int B, __bf16* Xb, float* Yb) {
  const int b = blockIdx.x;
  if (b >= B) return;
  const long long src = idx[b];
  for (int j = threadIdx.x; j < kx; j += blockDim.x) Xb[(long long)b * kx + j] = X[src * kx + j];
  for (int j = threadIdx.x; j < ky; j += blockDim.x) Yb[(long long)b * ky + j] = Y[src * ky + j];
}

// Adam on FP32 master parameters (bias-corrected), then the bf16 copies used
// by the matrix cores: W[o][i] for the forward pass and WT[i][o] for dX.
struct AdamArgs {
  float* p; float* m; float* v; const float* g; long long n;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)
};
__global__ void adam_step(AdamArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < a.n; i += (long long)gridDim.x * blockDim.x) {
    const float g = a.g[i];
    const float m = a.b1 * a.m[i] + (1.f - a.b1) * g;
    const float v = a.b2 * a.v[i] + (1.f - a.b2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    a.p[i] -= a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
  }
}

// bf16 packing of one layer: W[o][i] = p[o*nin + i] for o < nout_valid (rows
// beyond stay zero), WT[i][o] likewise transposed.
__global__ void pack_layer(const float* p, int nout, int nin, int nout_valid, __bf16* W, __bf16* WT) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nout * nin) return;
  const int o = (int)(t / nin), i = (int)(t % nin);
  const float v = o < nout_valid ? p[t] : 0.f;
  const __bf16 b = to_bf16(v);
  W[t] = b;
  if (WT) WT[(long long)i * nout + o] = b;
}

// ---------------------------------------------------------------- inference
// Fused inference of whole networks, all layers of every ensemble member in
// one launch, activations held in registers.
//
// The products run transposed: H'^T[o][cell] = sum_k W[o][k] H^T[k][cell],
// with the weights as MFMA operand A (32 output features x 16 k) and the
// activations as operand B (16 k x 32 cells).  A wave owns 32 cells.  Lane
// l then receives, for output tile t, D rows 32t + 8*(r/4) + 4*(l/32) + r%4
// of cell l%32 (r = 0..15), and operand B of the next layer wants from the
// same lane 8 consecutive k of one 16-wide slab.  The output registers
// acc[8q..8q+7] of tile t are used as they are for slab 2t+q, and the
// reduction index inside each slab is permuted instead: logical k = 8h + j
// of lane half h holds the physical feature 16s + perm(h, j) with
// perm(h, j) = 4h + j (j < 4), 8 + 4h + (j - 4) (j >= 4).  The weights of
// every layer fed by a hidden layer are packed with that permutation;
// the first layer reads the feature rows unpermuted.
//
// Packed weights ("fragment order"): fragment f = ot * (din/16) + ks of a
// layer holds 64 lanes x 8 bf16, lane l's 16 bytes being
// W[32 ot + l%32][16 ks + kperm(l/32, j)], j = 0..7, so every fragment is
// one coalesced 1 KB load.
__device__ __forceinline__ int frag_k(int h, int j, bool permute) {
  return permute ? (j < 4 ? 4 * h + j : 4 + 4 * h + j) : 8 * h + j;
}

__global__ void pack_frag(const float* __restrict__ p, int nout, int nin, int nout_valid, int permute,
                          __bf16* __restrict__ F) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nout * nin) return;
  const int j = (int)(t & 7);
  const int lane = (int)((t >> 3) & 63);
  const long long f = t >> 9;
  const int nks = nin / 16;
  const int ot = (int)(f / nks), ks = (int)(f % nks);
  const int o = ot * 32 + (lane & 31);
  const int k = ks * 16 + frag_k(lane >> 5, j, permute != 0);
  F[t] = to_bf16(o < nout_valid ? p[(long long)o * nin + k] : 0.f);
}

constexpr int kInferMaxMembers = 4;
constexpr int kInferMaxLayers = 9;   // hidden layers + output layer

struct InferArgs {
  const __bf16* X;            // [M][ldx] bf16 feature rows (ldx = NIN_PAD)
  long long M;                // rows, a multiple of 32
  int ldx;
  int nmembers, nhl;          // members, hidden layers
  const __bf16* F[kInferMaxMembers][kInferMaxLayers];  // packed weights
  const float* bias[kInferMaxMembers][kInferMaxLayers];
  float* out[kInferMaxMembers];                        // [M][NOUT] f32
};

__device__ __forceinline__ bf16x8_t frag_load(const __bf16* __restrict__ F, int f, int lane) {
  return *reinterpret_cast<const bf16x8_t*>(F + ((long long)f * 64 + lane) * 8);
}

// one layer with bias + GELU: NI input slabs -> NO output slabs (NO/2 tiles)
template <int NI, int NO>
__device__ __forceinline__ void infer_hidden(const __bf16* __restrict__ F, const float* __restrict__ bias,
                                             const bf16x8_t (&in)[NI], bf16x8_t (&out)[NO], int lane) {
  const int h = lane >> 5;
#pragma unroll
  for (int ot = 0; ot < NO / 2; ++ot) {
    f32x16_t acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
    for (int ks = 0; ks < NI; ++ks)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(frag_load(F, ot * NI + ks, lane), in[ks], acc, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 b = *reinterpret_cast<const float4*>(bias + ot * 32 + 8 * q + 4 * h);
      const float bb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
      for (int e = 0; e < 4; ++e)
        out[2 * ot + (q >> 1)][(q & 1) * 4 + e] = to_bf16(gelu_f(acc[4 * q + e] + bb[e]));
    }
  }
}

template <int NH, int NIN_S, int NOUT_T>
__global__ __launch_bounds__(256) void mlp_infer(InferArgs a) {
  constexpr int NS = NH / 16;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5;
  const long long cell0 = ((long long)blockIdx.x * 4 + wave) * 32;
  if (cell0 >= a.M) return;
  const long long cell = cell0 + (lane & 31);
  bf16x8_t xin[NIN_S];
#pragma unroll
  for (int s = 0; s < NIN_S; ++s)
    xin[s] = *reinterpret_cast<const bf16x8_t*>(a.X + cell * a.ldx + 16 * s + 8 * h);
  for (int m = 0; m < a.nmembers; ++m) {
    bf16x8_t hb[NS];
    infer_hidden<NIN_S, NS>(a.F[m][0], a.bias[m][0], xin, hb, lane);
    for (int l = 1; l < a.nhl; ++l) {
      bf16x8_t hn[NS];
      infer_hidden<NS, NS>(a.F[m][l], a.bias[m][l], hb, hn, lane);
#pragma unroll
      for (int s = 0; s < NS; ++s) hb[s] = hn[s];
    }
    const __bf16* F = a.F[m][a.nhl];
    const float* bias = a.bias[m][a.nhl];
    float* out = a.out[m] + cell * (32 * NOUT_T);
#pragma unroll
    for (int ot = 0; ot < NOUT_T; ++ot) {
      f32x16_t acc;
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
      for (int ks = 0; ks < NS; ++ks)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(frag_load(F, ot * NS + ks, lane), hb[ks], acc, 0, 0, 0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int o = ot * 32 + 8 * q + 4 * h;
        const float4 b = *reinterpret_cast<const float4*>(bias + o);
        *reinterpret_cast<float4*>(out + o) =
            make_float4(acc[4 * q] + b.x, acc[4 * q + 1] + b.y, acc[4 * q + 2] + b.z, acc[4 * q + 3] + b.w);
      }
    }
  }
}

}  // namespace brams
