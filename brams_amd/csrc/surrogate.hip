// Host side of the BF16 MLP surrogate (include/brams_surrogate.h).
//
// One brams_mlp owns FP32 master parameters, Adam moments, bf16 copies of
// every weight matrix (W[out][in] for the forward product, WT[in][out] for
// the input gradient, and the fragment-ordered copy read by the fused
// inference kernel) and the activation buffers of one training batch.
// Training products run on gemm_bf16, inference on mlp_infer (both in
// mlp_kernels.h); nothing here falls back to a library GEMM or to the CPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/brams_surrogate.h"
#include "mlp_kernels.h"

extern "C" int brams_chem_set_error(int code, const char* msg);

namespace {

int err(int code, const char* what, hipError_t e) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  return brams_chem_set_error(code, buf);
}

#define SUR_TRY(expr)                                      \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return err(-100 - (int)e_, #expr, e_); \
  } while (0)

template <class T>
hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1));
}

using brams::GemmArgs;

template <int LA, int LB, int EPI>
hipError_t gemm(const GemmArgs& g, hipStream_t st) {
  dim3 grid(g.N / 64, g.M / 64, g.ksplit);
  hipLaunchKernelGGL((brams::gemm_bf16<LA, LB, EPI>), grid, dim3(256), 0, st, g);
  return hipGetLastError();
}

hipError_t fill(float* p, long long n, float v, hipStream_t st) {
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(brams::fill_f32, dim3(blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

}  // namespace

struct brams_mlp {
  int nh = 0, nl = 0, maxb = 0, nlay = 0;
  std::vector<int> din, dout, dvalid;           // per layer: in, out (padded), valid outputs
  std::vector<long long> woff, boff;            // offsets of W and b in the flat vectors
  long long np = 0;
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  std::vector<__bf16*> W, WT;
  std::vector<__bf16*> Fr;                      // fragment-ordered weights for mlp_infer
  bool frag_dirty = true;                       // Fr older than p
  // training activations
  std::vector<float*> Z;                        // pre-activations [maxb][nh]
  std::vector<__bf16*> A;                       // activations [maxb][nh]
  std::vector<float*> dZ;                       // [maxb][nh]
  std::vector<__bf16*> dZb;
  float *out = nullptr, *dout_f = nullptr, *Yb = nullptr, *loss = nullptr;
  __bf16 *dout_b = nullptr, *Xb = nullptr;
  long long* idx = nullptr;
  long long idx_n = 0;
  long long step = 0;
};

namespace {

int pack_all(brams_mlp* m, hipStream_t st) {
  for (int l = 0; l < m->nlay; ++l) {
    const long long n = (long long)m->dout[l] * m->din[l];
    hipLaunchKernelGGL(brams::pack_layer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       m->p + m->woff[l], m->dout[l], m->din[l], m->dvalid[l], m->W[l], m->WT[l]);
    SUR_TRY(hipGetLastError());
  }
  m->frag_dirty = true;
  return 0;
}

int pack_frags(brams_mlp* m, hipStream_t st) {
  for (int l = 0; l < m->nlay; ++l) {
    const long long n = (long long)m->dout[l] * m->din[l];
    hipLaunchKernelGGL(brams::pack_frag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       m->p + m->woff[l], m->dout[l], m->din[l], m->dvalid[l], l > 0 ? 1 : 0, m->Fr[l]);
    SUR_TRY(hipGetLastError());
  }
  m->frag_dirty = false;
  return 0;
}

template <int NH, int NHL>
hipError_t launch_infer(const brams::InferArgs& a, hipStream_t st) {
  const unsigned grid = (unsigned)((a.M + 127) / 128);
  hipLaunchKernelGGL((brams::mlp_infer<NH, NHL, BRAMS_SUR_NIN_PAD / 16, BRAMS_SUR_NOUT_PAD / 32>), dim3(grid),
                     dim3(256), 0, st, a);
  return hipGetLastError();
}

template <int NH>
hipError_t launch_infer_nl(const brams::InferArgs& a, hipStream_t st) {
  switch (a.nhl) {
    case 1: return launch_infer<NH, 1>(a, st);
    case 2: return launch_infer<NH, 2>(a, st);
    case 3: return launch_infer<NH, 3>(a, st);
    case 4: return launch_infer<NH, 4>(a, st);
    default: return hipErrorInvalidValue;
  }
}

// training forward through all layers, keeping Z/A for the backward pass
int forward(brams_mlp* m, const __bf16* X, long long M, float* out, hipStream_t st) {
  const __bf16* in = X;
  long long ldin = BRAMS_SUR_NIN_PAD;
  for (int l = 0; l < m->nlay; ++l) {
    GemmArgs g{};
    g.M = (int)M; g.N = m->dout[l]; g.K = m->din[l];
    g.A = in; g.lda = ldin;
    g.B = m->W[l]; g.ldb = m->din[l];
    g.bias = m->p + m->boff[l];
    g.ksplit = 1;
    if (l == m->nlay - 1) {
      g.O32 = out; g.ldo32 = m->dout[l];
      SUR_TRY((gemm<brams::LAY_N, brams::LAY_N, brams::EPI_BIAS>(g, st)));
    } else {
      g.Z = m->Z[l]; g.ldz = m->nh;
      g.O = m->A[l]; g.ldo = m->nh;
      SUR_TRY((gemm<brams::LAY_N, brams::LAY_N, brams::EPI_BIAS_GELU>(g, st)));
      in = m->A[l];
      ldin = m->nh;
    }
  }
  return 0;
}

int choose_ksplit(int B) {
  int ks = std::max(1, B / 512);
  while (ks > 1 && (B % ks != 0 || (B / ks) % 16 != 0)) --ks;
  return ks;
}

// gradients of the loss on the batch already in Xb/Yb-shaped buffers
int backward(brams_mlp* m, const __bf16* X, const float* Y, int B, hipStream_t st) {
  SUR_TRY(fill(m->g, m->np, 0.f, st));
  const long long nout = (long long)B * BRAMS_SUR_NOUT_PAD;
  hipLaunchKernelGGL(brams::mse_grad, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, m->out,
                     Y, B, BRAMS_SUR_NOUT_PAD, BRAMS_SUR_NOUT, m->dout_f, m->dout_b, m->loss);
  SUR_TRY(hipGetLastError());
  const int ks = choose_ksplit(B);
  const float* dcur_f = m->dout_f;
  const __bf16* dcur_b = m->dout_b;
  long long ldd = BRAMS_SUR_NOUT_PAD;
  for (int l = m->nlay - 1; l >= 0; --l) {
    const __bf16* ain = (l == 0) ? X : m->A[l - 1];
    const long long lda_in = (l == 0) ? BRAMS_SUR_NIN_PAD : m->nh;
    // dW[o][i] = sum_b d[b][o] a[b][i]
    {
      GemmArgs g{};
      g.M = m->dout[l]; g.N = m->din[l]; g.K = B;
      g.A = dcur_b; g.lda = ldd;          // (row o, red b) at b*ldd + o
      g.B = ain; g.ldb = lda_in;          // (col i, red b) at b*lda + i
      g.O32 = m->g + m->woff[l]; g.ldo32 = m->din[l];
      g.ksplit = ks;
      SUR_TRY((gemm<brams::LAY_T, brams::LAY_T, brams::EPI_ACCUM>(g, st)));
    }
    // db[o] = sum_b d[b][o]
    {
      dim3 grid((m->dout[l] + 63) / 64, (B + 511) / 512);
      hipLaunchKernelGGL(brams::colsum_f32, grid, dim3(256), 0, st, dcur_f, ldd, B, m->dout[l],
                         m->g + m->boff[l]);
      SUR_TRY(hipGetLastError());
    }
    if (l == 0) break;
    // dA[b][i] = sum_o d[b][o] W[o][i], times gelu'(z) of the layer below
    {
      GemmArgs g{};
      g.M = B; g.N = m->din[l]; g.K = m->dout[l];
      g.A = dcur_b; g.lda = ldd;          // (row b, red o) at b*ldd + o
      g.B = m->WT[l]; g.ldb = m->dout[l]; // (col i, red o) at i*dout + o
      g.Z = m->Z[l - 1]; g.ldz = m->nh;
      g.O32 = m->dZ[l - 1]; g.ldo32 = m->nh;
      g.O = m->dZb[l - 1]; g.ldo = m->nh;
      g.ksplit = 1;
      SUR_TRY((gemm<brams::LAY_N, brams::LAY_N, brams::EPI_GELU_BWD>(g, st)));
    }
    dcur_f = m->dZ[l - 1];
    dcur_b = m->dZb[l - 1];
    ldd = m->nh;
  }
  return 0;
}

int adam(brams_mlp* m, float lr, hipStream_t st) {
  m->step += 1;
  brams::AdamArgs a{};
  a.p = m->p; a.m = m->m; a.v = m->v; a.g = m->g; a.n = m->np;
  a.lr = lr; a.b1 = 0.9f; a.b2 = 0.999f; a.eps = 1e-8f;
  a.c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)m->step)));
  a.c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)m->step)));
  const int blocks = (int)std::min<long long>((m->np + 255) / 256, 4096);
  hipLaunchKernelGGL(brams::adam_step, dim3(blocks), dim3(256), 0, st, a);
  SUR_TRY(hipGetLastError());
  return pack_all(m, st);
}

}  // namespace

extern "C" {

int brams_mlp_create(int nhidden, int nlayers, int max_batch, uint64_t seed, brams_mlp** out) {
  if (!out) return brams_chem_set_error(-2, "null output");
  if (nhidden < 64 || nhidden > 256 || nhidden % 64 || nlayers < 1 || nlayers > 4 ||
      max_batch < 64 || max_batch % 64)
    return brams_chem_set_error(-2, "nhidden must be 64, 128, 192 or 256, nlayers 1..4, "
                                    "max_batch a multiple of 64");
  brams_mlp* m = new brams_mlp();
  m->nh = nhidden; m->nl = nlayers; m->maxb = max_batch; m->nlay = nlayers + 1;
  for (int l = 0; l < m->nlay; ++l) {
    m->din.push_back(l == 0 ? BRAMS_SUR_NIN_PAD : nhidden);
    m->dout.push_back(l == m->nlay - 1 ? BRAMS_SUR_NOUT_PAD : nhidden);
    m->dvalid.push_back(l == m->nlay - 1 ? BRAMS_SUR_NOUT : nhidden);
  }
  long long off = 0;
  for (int l = 0; l < m->nlay; ++l) {
    m->woff.push_back(off);
    off += (long long)m->dout[l] * m->din[l];
    m->boff.push_back(off);
    off += m->dout[l];
  }
  m->np = off;
  // initial parameters: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) over the valid
  // fan-in (68 features for the first layer), zero on padded rows/columns
  std::vector<float> h(m->np, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < m->nlay; ++l) {
    const int fan = (l == 0) ? BRAMS_SUR_NIN : m->din[l];
    const float bound = 1.0f / std::sqrt((float)fan);
    std::uniform_real_distribution<float> U(-bound, bound);
    for (int o = 0; o < m->dvalid[l]; ++o) {
      for (int i = 0; i < (l == 0 ? BRAMS_SUR_NIN : m->din[l]); ++i)
        h[m->woff[l] + (long long)o * m->din[l] + i] = U(rng);
      h[m->boff[l] + o] = U(rng);
    }
  }
  hipError_t e = hipSuccess;
#define A_(x) if (e == hipSuccess) e = (x)
  A_(dalloc(&m->p, m->np));
  A_(dalloc(&m->m, m->np));
  A_(dalloc(&m->v, m->np));
  A_(dalloc(&m->g, m->np));
  A_(hipMemcpy(m->p, h.data(), sizeof(float) * m->np, hipMemcpyHostToDevice));
  std::vector<float> zero(m->np, 0.f);
  A_(hipMemcpy(m->m, zero.data(), sizeof(float) * m->np, hipMemcpyHostToDevice));
  A_(hipMemcpy(m->v, zero.data(), sizeof(float) * m->np, hipMemcpyHostToDevice));
  m->W.resize(m->nlay);
  m->WT.resize(m->nlay);
  m->Fr.resize(m->nlay);
  for (int l = 0; l < m->nlay; ++l) {
    A_(dalloc(&m->W[l], (size_t)m->dout[l] * m->din[l]));
    A_(dalloc(&m->WT[l], (size_t)m->dout[l] * m->din[l]));
    A_(dalloc(&m->Fr[l], (size_t)m->dout[l] * m->din[l] + brams::kInferPadElems));
  }
  m->Z.resize(nlayers); m->A.resize(nlayers); m->dZ.resize(nlayers); m->dZb.resize(nlayers);
  for (int l = 0; l < nlayers; ++l) {
    A_(dalloc(&m->Z[l], (size_t)max_batch * nhidden));
    A_(dalloc(&m->A[l], (size_t)max_batch * nhidden));
    A_(dalloc(&m->dZ[l], (size_t)max_batch * nhidden));
    A_(dalloc(&m->dZb[l], (size_t)max_batch * nhidden));
  }
  A_(dalloc(&m->out, (size_t)max_batch * BRAMS_SUR_NOUT_PAD));
  A_(dalloc(&m->dout_f, (size_t)max_batch * BRAMS_SUR_NOUT_PAD));
  A_(dalloc(&m->dout_b, (size_t)max_batch * BRAMS_SUR_NOUT_PAD));
  A_(dalloc(&m->Xb, (size_t)max_batch * BRAMS_SUR_NIN_PAD));
  A_(dalloc(&m->Yb, (size_t)max_batch * BRAMS_SUR_NOUT_PAD));
  A_(dalloc(&m->loss, 1));
#undef A_
  if (e != hipSuccess) {
    brams_mlp_destroy(m);
    return err(-100 - (int)e, "brams_mlp_create", e);
  }
  int rc = pack_all(m, nullptr);
  if (rc) {
    brams_mlp_destroy(m);
    return rc;
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess) {
    brams_mlp_destroy(m);
    return err(-100 - (int)e, "brams_mlp_create", e);
  }
  *out = m;
  return 0;
}

void brams_mlp_destroy(brams_mlp* m) {
  if (!m) return;
  void* ps[] = {m->p, m->m, m->v, m->g, m->out, m->dout_f, m->dout_b, m->Xb, m->Yb, m->loss,
                m->idx};
  for (void* p : ps)
    if (p) hipFree(p);
  for (auto* p : m->W) if (p) hipFree(p);
  for (auto* p : m->WT) if (p) hipFree(p);
  for (auto* p : m->Fr) if (p) hipFree(p);
  for (auto* p : m->Z) if (p) hipFree(p);
  for (auto* p : m->A) if (p) hipFree(p);
  for (auto* p : m->dZ) if (p) hipFree(p);
  for (auto* p : m->dZb) if (p) hipFree(p);
  delete m;
}

int64_t brams_mlp_num_params(const brams_mlp* m) { return m ? m->np : -1; }

int brams_mlp_get_params(const brams_mlp* m, float* host, int64_t n) {
  if (!m || n != m->np) return brams_chem_set_error(-2, "parameter count mismatch");
  SUR_TRY(hipMemcpy(host, m->p, sizeof(float) * n, hipMemcpyDeviceToHost));
  return 0;
}

int brams_mlp_set_params(brams_mlp* m, const float* host, int64_t n) {
  if (!m || n != m->np) return brams_chem_set_error(-2, "parameter count mismatch");
  SUR_TRY(hipMemcpy(m->p, host, sizeof(float) * n, hipMemcpyHostToDevice));
  int rc = pack_all(m, nullptr);
  if (rc) return rc;
  SUR_TRY(hipDeviceSynchronize());
  return 0;
}

}  // extern "C"

namespace {

int infer(brams_mlp* const* members, int n, const void* X, int64_t M, float* const* out, bool species_major,
          void* stream) {
  if (!members || n < 1 || n > brams::kInferMaxMembers || !X || !out || M <= 0 || M % 64)
    return brams_chem_set_error(-2, "bad forward arguments");
  hipStream_t st = (hipStream_t)stream;
  const brams_mlp* m0 = members[0];
  brams::InferArgs a{};
  a.X = (const __bf16*)X;
  a.M = M;
  a.ldx = BRAMS_SUR_NIN_PAD;
  a.nmembers = n;
  a.nhl = m0 ? m0->nl : 0;
  a.ldo = species_major ? M : 0;
  for (int k = 0; k < n; ++k) {
    brams_mlp* m = members[k];
    if (!m || !out[k] || m->nh != m0->nh || m->nl != m0->nl)
      return brams_chem_set_error(-2, "ensemble members differ in shape");
    if (m->frag_dirty) {
      int rc = pack_frags(m, st);
      if (rc) return rc;
    }
    for (int l = 0; l < m->nlay; ++l) {
      a.F[k][l] = m->Fr[l];
      a.bias[k][l] = m->p + m->boff[l];
    }
    a.out[k] = out[k];
  }
  if (a.nhl < 1 || a.nhl > 4) return brams_chem_set_error(-2, "fused inference takes 1..4 hidden layers");
  switch (m0->nh) {
    case 64: SUR_TRY(launch_infer_nl<64>(a, st)); break;
    case 128: SUR_TRY(launch_infer_nl<128>(a, st)); break;
    case 192: SUR_TRY(launch_infer_nl<192>(a, st)); break;
    case 256: SUR_TRY(launch_infer_nl<256>(a, st)); break;
    default: return brams_chem_set_error(-2, "unsupported hidden width");
  }
  return 0;
}

}  // namespace

extern "C" {

int brams_mlp_forward_ensemble(brams_mlp* const* members, int n, const void* X, int64_t M,
                               float* const* out, void* stream) {
  return infer(members, n, X, M, out, true, stream);
}

int brams_mlp_forward(brams_mlp* m, const void* X, int64_t M, float* out, void* stream) {
  int rc = infer(&m, 1, X, M, &out, false, stream);
  if (rc) return rc;
  SUR_TRY(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int brams_mlp_gradients(brams_mlp* m, const void* X, const float* Y, int B, float* grads, float* loss) {
  if (!m || !X || !Y || B <= 0 || B % 64 || B > m->maxb)
    return brams_chem_set_error(-2, "bad gradient arguments");
  SUR_TRY(fill(m->loss, 1, 0.f, nullptr));
  int rc = forward(m, (const __bf16*)X, B, m->out, nullptr);
  if (!rc) rc = backward(m, (const __bf16*)X, Y, B, nullptr);
  if (rc) return rc;
  SUR_TRY(hipDeviceSynchronize());
  if (grads) SUR_TRY(hipMemcpy(grads, m->g, sizeof(float) * m->np, hipMemcpyDeviceToHost));
  if (loss) SUR_TRY(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int brams_mlp_train_range(brams_mlp* m, const void* X, const float* Y, int64_t N, int batch, int epoch0,
                          int nepochs, int total_epochs, float lr0, uint64_t seed, float* epoch_loss) {
  if (!m || !X || !Y || batch <= 0 || batch % 64 || batch > m->maxb || N < batch || nepochs < 1 ||
      epoch0 < 0 || epoch0 + nepochs > total_epochs)
    return brams_chem_set_error(-2, "bad training arguments");
  if (m->idx_n < N) {
    if (m->idx) (void)hipFree(m->idx);
    m->idx = nullptr;
    SUR_TRY(dalloc(&m->idx, (size_t)N));
    m->idx_n = N;
  }
  const long long nb = N / batch;
  const long long total = nb * total_epochs;
  std::vector<long long> perm(N);
  for (int e = 0; e < nepochs; ++e) {
    const int ep = epoch0 + e;
    // the permutation of an epoch depends on (seed, epoch) only, so a
    // schedule run in several calls equals one run in a single call
    std::iota(perm.begin(), perm.end(), 0LL);
    std::mt19937_64 rng(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(ep + 1));
    std::shuffle(perm.begin(), perm.end(), rng);
    SUR_TRY(hipMemcpy(m->idx, perm.data(), sizeof(long long) * N, hipMemcpyHostToDevice));
    SUR_TRY(fill(m->loss, 1, 0.f, nullptr));
    long long t = (long long)ep * nb;
    for (long long b = 0; b < nb; ++b, ++t) {
      hipLaunchKernelGGL(brams::gather_rows, dim3(batch), dim3(128), 0, nullptr, (const __bf16*)X,
                         BRAMS_SUR_NIN_PAD, Y, BRAMS_SUR_NOUT_PAD, m->idx + b * batch, batch, m->Xb,
                         m->Yb);
      SUR_TRY(hipGetLastError());
      int rc = forward(m, m->Xb, batch, m->out, nullptr);
      if (!rc) rc = backward(m, m->Xb, m->Yb, batch, nullptr);
      const float lr = (float)(0.5 * lr0 * (1.0 + std::cos(M_PI * (double)t / (double)total)));
      if (!rc) rc = adam(m, lr, nullptr);
      if (rc) return rc;
    }
    float l = 0.f;
    SUR_TRY(hipMemcpy(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost));
    if (epoch_loss) epoch_loss[e] = l / (float)nb;
  }
  SUR_TRY(hipDeviceSynchronize());
  return 0;
}

int brams_mlp_train(brams_mlp* m, const void* X, const float* Y, int64_t N, int batch, int epochs,
                    float lr0, uint64_t seed, float* epoch_loss) {
  return brams_mlp_train_range(m, X, Y, N, batch, 0, epochs, epochs, lr0, seed, epoch_loss);
}

}  // extern "C"
