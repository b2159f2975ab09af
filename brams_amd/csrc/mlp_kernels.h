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
typedef __attribute__((__vector_size__(4 * sizeof(unsigned)))) unsigned u32x4_t;

enum { LAY_N = 0, LAY_T = 1 };
enum { EPI_BIAS_GELU = 0, EPI_BIAS = 1, EPI_GELU_BWD = 2, EPI_ACCUM = 3, EPI_BIAS_GELU_INF = 4 };

// GELU(x) = x Phi(x) = 0.5 x (1 + erf(x / sqrt 2)), with erf from
// Abramowitz & Stegun, Handbook of Mathematical Functions (1964), 7.1.26
// (Hastings' rational approximation): for z >= 0
//   erf(z) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-z^2),
//   t = 1 / (1 + p z),  |error| <= 1.5e-7,
// and erf(-z) = -erf(z).  Branch-free: about 15 VALU operations against
// some 35 with data-dependent branches for the library erff, which matters
// because every hidden activation passes through it.  Training and
// inference use the same function.
__device__ __forceinline__ float erfc_tail(float z) {   // 1 - erf(z), z >= 0
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, z, 1.0f));
  float q = fmaf(1.061405429f, t, -1.453152027f);
  q = fmaf(q, t, 1.421413741f);
  q = fmaf(q, t, -0.284496736f);
  q = fmaf(q, t, 0.254829592f);
  return q * t * __builtin_amdgcn_exp2f(-1.4426950408889634f * z * z);
}
__device__ __forceinline__ float gelu_f(float x) {
  const float h = 0.5f * x * erfc_tail(fabsf(x) * 0.70710678118654752f);
  return x >= 0.0f ? x - h : h;
}
__device__ __forceinline__ float gelu_grad_f(float x) {
  const float tail = 0.5f * erfc_tail(fabsf(x) * 0.70710678118654752f);
  const float cdf = x >= 0.0f ? 1.0f - tail : tail;
  const float pdf = 0.39894228040143268f * __builtin_amdgcn_exp2f(-0.72134752044448170f * x * x);
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
  float* out[kInferMaxMembers];                        // f32 outputs of each member
  long long ldo;              // 0: row-major [M][NOUT]; else species-major, out[o * ldo + cell]
};

// Weight tiles are staged through LDS: the 256 threads of a workgroup load
// tile t+1 (NI fragments, NI KB) from global memory into registers before
// the MFMAs of tile t and store it into the LDS tile buffer after them,
// between two barriers.  Each fragment crosses L2 once per workgroup (four
// waves, 128 cells) and reaches a wave as one 16-byte LDS load per lane.
// A layer's outputs go to a per-wave LDS buffer in the operand-B lane order
// and are read back as the next layer's input, so one activation array (not
// two) lives in registers; 80 KB of LDS per workgroup leaves room for two
// workgroups per CU.
// A fetch always moves a whole tile buffer (FMAX fragments) so that the
// prefetch registers stay registers; a tile of fewer fragments is followed
// by the next one or by the kInferPad bytes of padding behind every packed
// layer, and the MFMAs read only its own fragments.
constexpr int kInferPadElems = 16 * 512;   // bf16 elements behind each packed layer

template <int NL>
__device__ __forceinline__ void tile_fetch(const __bf16* __restrict__ src, int tid, u32x4_t (&r)[NL]) {
  const u32x4_t* s = reinterpret_cast<const u32x4_t*>(src);
#pragma unroll
  for (int i = 0; i < NL; ++i) r[i] = s[i * 256 + tid];
}

template <int NL>
__device__ __forceinline__ void tile_store(u32x4_t* dst, int tid, const u32x4_t (&r)[NL]) {
#pragma unroll
  for (int i = 0; i < NL; ++i) dst[i * 256 + tid] = r[i];
}

// the tile after (m, l, t) in the order members -> layers -> output tiles
// (after the last one: the first tile again, fetched and never used)
template <int NH, int NHL, int NIN_S, int NOUT_T>
__device__ __forceinline__ const __bf16* next_tile(const InferArgs& a, int m, int l, int t) {
  constexpr int NT_H = NH / 32, NS = NH / 16;
  const int ntl = (l == NHL) ? NOUT_T : NT_H;
  if (t + 1 < ntl) {
    ++t;
  } else if (l < NHL) {
    ++l;
    t = 0;
  } else if (m + 1 < a.nmembers) {
    ++m;
    l = 0;
    t = 0;
  } else {
    return a.F[0][0];
  }
  return a.F[m][l] + (long long)t * ((l == 0) ? NIN_S : NS) * 512;
}

template <int NI, int N>
__device__ __forceinline__ f32x16_t tile_mfma(const u32x4_t* w, const bf16x8_t (&in)[N], int lane) {
  static_assert(NI <= N, "operand slabs");
  f32x16_t acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
  for (int ks = 0; ks < NI; ++ks) {
    const bf16x8_t f = __builtin_bit_cast(bf16x8_t, w[ks * 64 + lane]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f, in[ks], acc, 0, 0, 0);
  }
  return acc;
}

// bias + GELU of output tile t -> slabs 2t, 2t+1 of the wave's LDS buffer
__device__ __forceinline__ void tile_gelu_lds(const f32x16_t& acc, const float* __restrict__ bias, int t, int h,
                                              int lane, u32x4_t* act) {
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    bf16x8_t o;
#pragma unroll
    for (int qq = 0; qq < 2; ++qq) {
      const int q = 2 * half + qq;
      const float4 b = *reinterpret_cast<const float4*>(bias + t * 32 + 8 * q + 4 * h);
      o[4 * qq + 0] = to_bf16(gelu_f(acc[4 * q + 0] + b.x));
      o[4 * qq + 1] = to_bf16(gelu_f(acc[4 * q + 1] + b.y));
      o[4 * qq + 2] = to_bf16(gelu_f(acc[4 * q + 2] + b.z));
      o[4 * qq + 3] = to_bf16(gelu_f(acc[4 * q + 3] + b.w));
    }
    act[(2 * t + half) * 64 + lane] = __builtin_bit_cast(u32x4_t, o);
  }
}

template <int NH, int NHL, int NIN_S, int NOUT_T>
__global__ __launch_bounds__(256) void mlp_infer(InferArgs a) {
  constexpr int NS = NH / 16;
  constexpr int NT_H = NH / 32;
  constexpr int FMAX = NS > NIN_S ? NS : NIN_S;
  static_assert(FMAX <= 16 && FMAX % 4 == 0, "a tile is staged by 256 threads x FMAX/4 x 16 bytes");
  static_assert(FMAX * 512 <= kInferPadElems, "padding behind each packed layer");
  __shared__ u32x4_t wbuf[FMAX * 64];
  __shared__ u32x4_t act_all[4][NS * 64];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int h = lane >> 5;
  u32x4_t* act = act_all[wave];
  const long long cell0 = ((long long)blockIdx.x * 4 + wave) * 32;
  const bool active = cell0 < a.M;                    // idle waves still stage tiles
  const long long cell = active ? cell0 + (lane & 31) : (lane & 31);
  constexpr int NL = FMAX / 4;                        // 16-byte loads per thread per tile
  u32x4_t r[NL];
  tile_fetch<NL>(a.F[0][0], tid, r);
  tile_store<NL>(wbuf, tid, r);
  __syncthreads();
  for (int m = 0; m < a.nmembers; ++m) {
    bf16x8_t in[FMAX];                                // layer input slabs
#pragma unroll
    for (int s = 0; s < NIN_S; ++s)
      in[s] = *reinterpret_cast<const bf16x8_t*>(a.X + cell * a.ldx + 16 * s + 8 * h);
#pragma unroll
    for (int l = 0; l <= NHL; ++l) {
      const int nt = (l == NHL) ? NOUT_T : NT_H;
#pragma unroll
      for (int t = 0; t < nt; ++t) {
        tile_fetch<NL>(next_tile<NH, NHL, NIN_S, NOUT_T>(a, m, l, t), tid, r);
        const f32x16_t acc = (l == 0) ? tile_mfma<NIN_S>(wbuf, in, lane) : tile_mfma<NS>(wbuf, in, lane);
        if (l < NHL) {
          tile_gelu_lds(acc, a.bias[m][l], t, h, lane, act);
        } else if (active) {
          const float* bias = a.bias[m][NHL];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int o = t * 32 + 8 * q + 4 * h;
            const float4 b = *reinterpret_cast<const float4*>(bias + o);
            const float4 v = make_float4(acc[4 * q] + b.x, acc[4 * q + 1] + b.y, acc[4 * q + 2] + b.z,
                                         acc[4 * q + 3] + b.w);
            if (a.ldo) {                              // 32 consecutive cells per store: coalesced
              float* out = a.out[m] + cell;
              out[(o + 0) * a.ldo] = v.x;
              out[(o + 1) * a.ldo] = v.y;
              out[(o + 2) * a.ldo] = v.z;
              out[(o + 3) * a.ldo] = v.w;
            } else {
              *reinterpret_cast<float4*>(a.out[m] + cell * (32 * NOUT_T) + o) = v;
            }
          }
        }
        __syncthreads();                              // every wave is done with this tile
        tile_store<NL>(wbuf, tid, r);
        __syncthreads();                              // the next tile is in place
      }
      if (l < NHL) {
#pragma unroll
        for (int s = 0; s < NS; ++s)                // this wave's own outputs
          in[s] = __builtin_bit_cast(bf16x8_t, act[s * 64 + lane]);
      }
    }
  }
}

}  // namespace brams
