// Learned variant selector: a 3-layer MLP (16 graph features -> 64 -> 64 -> 9 predicted log
// runtimes, ReLU), BF16 weights and activations on the MFMA units with FP32 accumulation, FP32
// master weights and Adam.  No counterpart in the reference.
//
// Every product runs on v_mfma_f32_16x16x16_bf16: a 16x16 output tile per wavefront and
// instruction, K in steps of 16.  Operand layout of that instruction (lane l, i = 0..3):
// A[l % 16][4 (l / 16) + i], B[4 (l / 16) + i][l % 16], C/D[4 (l / 16) + i][l % 16].
//
// Training runs in one workgroup that owns the whole model in LDS and walks the minibatches of
// an epoch (the model has 6288 parameters; a step is a few microseconds).  Rounding points,
// mirrored by the CPU oracle (oracle/mlp.py): inputs, weights and the ReLU outputs are rounded
// to BF16 before each product; dY = 2 (Y - T) / (rows * 9) and the masked dH are rounded to BF16
// before theirs; biases, losses, gradients and the Adam state are FP32.
#include <cmath>
#include <cstring>
#include <vector>

#include "pbf_device.h"

namespace {
constexpr int IN = PBF_MLP_IN, HID = PBF_MLP_HIDDEN, OUT = PBF_MLP_OUT, OUTP = 16;
constexpr int MB = 128;  // rows per minibatch / per workgroup tile
constexpr int NT = 256;  // threads (4 wavefronts)
// parameter layout (FP32): W1[IN][HID], b1[HID], W2[HID][HID], b2[HID], W3[HID][OUTP], b3[OUTP]
constexpr int OW1 = 0, OB1 = OW1 + IN * HID, OW2 = OB1 + HID, OB2 = OW2 + HID * HID, OW3 = OB2 + HID,
              OB3 = OW3 + HID * OUTP, NPARAM = OB3 + OUTP;
static_assert(NPARAM == PBF_MLP_PARAMS, "parameter count");

typedef uint16_t bf16;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16 to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (no NaN can occur here)
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16 b) { return __uint_as_float((uint32_t)b << 16); }

// One 16x16 tile of C = A (M x K) * B (K x N), A and B in LDS as BF16.  TA: memory holds A^T
// (K x M, leading dimension lda); TB: memory holds B^T (N x K).
template <bool TA, bool TB>
__device__ __forceinline__ f32x4 mfma_tile(const bf16* A, int lda, const bf16* B, int ldb, int m0, int n0, int K) {
  const int lane = threadIdx.x & 63, r = lane & 15, kq = (lane >> 4) * 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = k0 + kq + i;
      a[i] = (short)(TA ? A[k * lda + m0 + r] : A[(m0 + r) * lda + k]);
      b[i] = (short)(TB ? B[(n0 + r) * ldb + k] : B[k * ldb + n0 + r]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

struct Lds {
  bf16 w1[IN * HID], w2[HID * HID], w3[HID * OUTP];
  float b1[HID], b2[HID], b3[OUTP];
  bf16 x[MB * IN], h1[MB * HID], h2[MB * HID], dy[MB * OUTP], dh2[MB * HID], dh1[MB * HID];
  float g[NPARAM];  // gradient accumulators
  float red[NT / 64];
};

__device__ void load_weights(Lds& s, const float* __restrict__ p) {
  for (int q = threadIdx.x; q < IN * HID; q += NT) s.w1[q] = to_bf16(p[OW1 + q]);
  for (int q = threadIdx.x; q < HID * HID; q += NT) s.w2[q] = to_bf16(p[OW2 + q]);
  for (int q = threadIdx.x; q < HID * OUTP; q += NT) s.w3[q] = to_bf16(p[OW3 + q]);
  for (int q = threadIdx.x; q < HID; q += NT) { s.b1[q] = p[OB1 + q]; s.b2[q] = p[OB2 + q]; }
  for (int q = threadIdx.x; q < OUTP; q += NT) s.b3[q] = q < OUT ? p[OB3 + q] : 0.f;
}

// Forward over the MB rows in s.x (rows >= rows are zero); Y (FP32, MB x OUTP) to yout when
// given, and dY (BF16) into s.dy with the loss when T is given.
__device__ float forward(Lds& s, int rows, const float* __restrict__ T, const int* __restrict__ idx, float* yout) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, rq = (lane >> 4) * 4;
  // layer 1: (MB x 16) * (16 x 64): 8 x 4 tiles
  for (int t = w; t < (MB / 16) * (HID / 16); t += NT / 64) {
    const int m0 = (t / (HID / 16)) * 16, n0 = (t % (HID / 16)) * 16;
    const f32x4 acc = mfma_tile<false, false>(s.x, IN, s.w1, HID, m0, n0, IN);
#pragma unroll
    for (int i = 0; i < 4; i++) s.h1[(m0 + rq + i) * HID + n0 + c] = to_bf16(fmaxf(acc[i] + s.b1[n0 + c], 0.f));
  }
  __syncthreads();
  for (int t = w; t < (MB / 16) * (HID / 16); t += NT / 64) {
    const int m0 = (t / (HID / 16)) * 16, n0 = (t % (HID / 16)) * 16;
    const f32x4 acc = mfma_tile<false, false>(s.h1, HID, s.w2, HID, m0, n0, HID);
#pragma unroll
    for (int i = 0; i < 4; i++) s.h2[(m0 + rq + i) * HID + n0 + c] = to_bf16(fmaxf(acc[i] + s.b2[n0 + c], 0.f));
  }
  __syncthreads();
  float loss = 0.f;
  for (int t = w; t < MB / 16; t += NT / 64) {
    const int m0 = t * 16;
    const f32x4 acc = mfma_tile<false, false>(s.h2, HID, s.w3, OUTP, m0, 0, HID);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = m0 + rq + i;
      const float y = acc[i] + s.b3[c];
      if (yout && row < rows && c < OUT) yout[row * OUT + c] = y;
      if (T) {
        float d = 0.f;
        if (row < rows && c < OUT) {
          const float e = y - T[(size_t)(idx ? idx[row] : row) * OUT + c];
          loss += e * e;
          d = 2.f * e / (float)(rows * OUT);
        }
        s.dy[row * OUTP + c] = to_bf16(d);
      }
    }
  }
  __syncthreads();
  return loss;
}

// Gradients of the loss for the rows in s.x, accumulated into s.g (FP32).
__device__ void backward(Lds& s) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, rq = (lane >> 4) * 4;
  // dW3 = H2^T dY (64 x 16, K = MB); db3 = column sums of dY
  for (int t = w; t < HID / 16; t += NT / 64) {
    const f32x4 acc = mfma_tile<true, false>(s.h2, HID, s.dy, OUTP, t * 16, 0, MB);
#pragma unroll
    for (int i = 0; i < 4; i++) s.g[OW3 + (t * 16 + rq + i) * OUTP + c] += acc[i];
  }
  for (int q = threadIdx.x; q < OUTP; q += NT) {
    float sum = 0.f;
    for (int r = 0; r < MB; r++) sum += from_bf16(s.dy[r * OUTP + q]);
    s.g[OB3 + q] += sum;
  }
  // dH2 = (dY W3^T) masked by H2 > 0 (MB x 64, K = 16)
  for (int t = w; t < (MB / 16) * (HID / 16); t += NT / 64) {
    const int m0 = (t / (HID / 16)) * 16, n0 = (t % (HID / 16)) * 16;
    const f32x4 acc = mfma_tile<false, true>(s.dy, OUTP, s.w3, OUTP, m0, n0, OUTP);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int q = (m0 + rq + i) * HID + n0 + c;
      s.dh2[q] = to_bf16(s.h2[q] != 0 ? acc[i] : 0.f);
    }
  }
  __syncthreads();
  // dW2 = H1^T dH2 (64 x 64, K = MB); db2
  for (int t = w; t < (HID / 16) * (HID / 16); t += NT / 64) {
    const int m0 = (t / (HID / 16)) * 16, n0 = (t % (HID / 16)) * 16;
    const f32x4 acc = mfma_tile<true, false>(s.h1, HID, s.dh2, HID, m0, n0, MB);
#pragma unroll
    for (int i = 0; i < 4; i++) s.g[OW2 + (m0 + rq + i) * HID + n0 + c] += acc[i];
  }
  for (int q = threadIdx.x; q < HID; q += NT) {
    float sum = 0.f;
    for (int r = 0; r < MB; r++) sum += from_bf16(s.dh2[r * HID + q]);
    s.g[OB2 + q] += sum;
  }
  // dH1 = (dH2 W2^T) masked by H1 > 0 (MB x 64, K = 64)
  for (int t = w; t < (MB / 16) * (HID / 16); t += NT / 64) {
    const int m0 = (t / (HID / 16)) * 16, n0 = (t % (HID / 16)) * 16;
    const f32x4 acc = mfma_tile<false, true>(s.dh2, HID, s.w2, HID, m0, n0, HID);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int q = (m0 + rq + i) * HID + n0 + c;
      s.dh1[q] = to_bf16(s.h1[q] != 0 ? acc[i] : 0.f);
    }
  }
  __syncthreads();
  // dW1 = X^T dH1 (16 x 64, K = MB); db1
  for (int t = w; t < HID / 16; t += NT / 64) {
    const f32x4 acc = mfma_tile<true, false>(s.x, IN, s.dh1, HID, 0, t * 16, MB);
#pragma unroll
    for (int i = 0; i < 4; i++) s.g[OW1 + (rq + i) * HID + t * 16 + c] += acc[i];
  }
  for (int q = threadIdx.x; q < HID; q += NT) {
    float sum = 0.f;
    for (int r = 0; r < MB; r++) sum += from_bf16(s.dh1[r * HID + q]);
    s.g[OB1 + q] += sum;
  }
  __syncthreads();
}

__device__ void load_rows(Lds& s, const float* __restrict__ X, const int* __restrict__ idx, int row0, int rows) {
  for (int q = threadIdx.x; q < MB * IN; q += NT) {
    const int r = q / IN, f = q % IN;
    s.x[q] = r < rows ? to_bf16(X[(size_t)(idx ? idx[row0 + r] : row0 + r) * IN + f]) : (bf16)0;
  }
}

__device__ float block_reduce(Lds& s, float v) {
  for (int d = 32; d; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0) s.red[threadIdx.x >> 6] = v;
  __syncthreads();
  float t = 0.f;
  for (int k = 0; k < NT / 64; k++) t += s.red[k];
  __syncthreads();
  return t;
}

// Training: one workgroup.  steps minibatches of MB rows taken in the order of idx (a
// permutation per epoch, concatenated); Adam on the FP32 master parameters p with moments
// m, v.  Mode GRAD (nsteps = 1, no update): writes the gradients of the first `rows` rows to
// gout and the loss to *loss.
template <bool UPDATE>
__global__ __launch_bounds__(NT) void k_mlp_train(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ X, const float* __restrict__ T,
                                                  const int* __restrict__ idx, int64_t nrows, int64_t nsteps, float lr,
                                                  int64_t t0, float* __restrict__ gout, float* __restrict__ loss_out) {
  __shared__ Lds s;
  float loss_sum = 0.f;
  for (int64_t st = 0; st < nsteps; st++) {
    const int64_t row0 = st * MB;
    const int rows = (int)min((int64_t)MB, nrows - row0);
    load_weights(s, p);
    for (int q = threadIdx.x; q < NPARAM; q += NT) s.g[q] = 0.f;
    load_rows(s, X, idx, (int)row0, rows);
    __syncthreads();
    const float l = block_reduce(s, forward(s, rows, T, idx ? idx + row0 : nullptr, nullptr));
    loss_sum += l / (float)(rows * OUT);
    backward(s);
    if constexpr (UPDATE) {
      const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
      const float tt = (float)(t0 + st + 1);
      const float c1 = 1.f - powf(b1, tt), c2 = 1.f - powf(b2, tt);
      for (int q = threadIdx.x; q < NPARAM; q += NT) {
        const bool pad = (q >= OW3 && q < OB3 && ((q - OW3) % OUTP) >= OUT) || (q >= OB3 + OUT);
        if (pad) continue;
        const float g = s.g[q];
        const float mq = b1 * m[q] + (1.f - b1) * g;
        const float vq = b2 * v[q] + (1.f - b2) * g * g;
        m[q] = mq;
        v[q] = vq;
        p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
      }
    } else {
      for (int q = threadIdx.x; q < NPARAM; q += NT) gout[q] = s.g[q];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss_out) *loss_out = loss_sum / (float)nsteps;
}

// Inference: one workgroup per MB rows.
__global__ __launch_bounds__(NT) void k_mlp_forward(const float* __restrict__ p, const float* __restrict__ X, int64_t nrows,
                                                    float* __restrict__ Y) {
  __shared__ Lds s;
  load_weights(s, p);
  for (int64_t row0 = (int64_t)blockIdx.x * MB; row0 < nrows; row0 += (int64_t)gridDim.x * MB) {
    const int rows = (int)min((int64_t)MB, nrows - row0);
    load_rows(s, X, nullptr, (int)row0, rows);
    __syncthreads();
    forward(s, rows, nullptr, nullptr, Y + row0 * OUT);
    __syncthreads();
  }
}

uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct DevBufs {
  float *p, *m, *v, *x, *t, *g, *l;
  int* idx;
};

int bufs(pbf_context* ctx, int64_t n, DevBufs* b) {
  void* q;
  const size_t floats = 4 * (size_t)NPARAM + (size_t)n * IN + (size_t)n * OUT + 16;
  PBF_CHECK(pbf_scratch(ctx, S_MLP, 4 * floats + 4 * (size_t)(n + 16), &q));
  float* f = static_cast<float*>(q);
  b->p = f; b->m = f + NPARAM; b->v = f + 2 * NPARAM; b->g = f + 3 * NPARAM;
  b->x = f + 4 * NPARAM; b->t = b->x + (size_t)n * IN; b->l = b->t + (size_t)n * OUT;
  b->idx = reinterpret_cast<int*>(b->l + 16);
  return PBF_OK;
}
}  // namespace

extern "C" int pbf_mlp_init(uint64_t seed, float* params) {
  if (!params) return PBF_ERR_ARG;
  memset(params, 0, sizeof(float) * NPARAM);
  uint64_t st = seed;
  auto uni = [&]() { return (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0); };
  auto fill = [&](int off, int rows, int cols, int used_cols) {
    const double a = std::sqrt(6.0 / rows);  // He-uniform for ReLU inputs
    for (int r = 0; r < rows; r++)
      for (int c = 0; c < cols; c++) params[off + r * cols + c] = c < used_cols ? (float)((2 * uni() - 1) * a) : 0.f;
  };
  fill(OW1, IN, HID, HID);
  fill(OW2, HID, HID, HID);
  fill(OW3, HID, OUTP, OUT);
  return PBF_OK;
}

extern "C" int pbf_mlp_train(pbf_context* ctx, float* params, const float* X, const float* T, int64_t n, int epochs,
                             float lr, uint64_t seed, float* final_loss) {
  if (!ctx || !params || !X || !T || n <= 0 || epochs <= 0) return PBF_ERR_ARG;
  if (n >= (int64_t(1) << 31)) return PBF_ERR_CAPACITY;
  DevBufs b;
  PBF_CHECK(bufs(ctx, n, &b));
  // permutations of all epochs, uploaded per epoch
  std::vector<int> perm((size_t)n);
  for (int64_t k = 0; k < n; k++) perm[k] = (int)k;
  PBF_CHECK(pbf_h2d(ctx, b.p, params, sizeof(float) * NPARAM));
  PBF_CHECK(pbf_fill_u32(ctx, reinterpret_cast<uint32_t*>(b.m), 0u, 2 * NPARAM));
  PBF_CHECK(pbf_h2d(ctx, b.x, X, sizeof(float) * (size_t)n * IN));
  PBF_CHECK(pbf_h2d(ctx, b.t, T, sizeof(float) * (size_t)n * OUT));
  uint64_t st = seed;
  const int64_t steps = (n + MB - 1) / MB;
  for (int e = 0; e < epochs; e++) {
    for (int64_t k = n - 1; k > 0; k--) std::swap(perm[k], perm[splitmix(st) % (uint64_t)(k + 1)]);
    PBF_CHECK(pbf_h2d(ctx, b.idx, perm.data(), sizeof(int) * (size_t)n));
    hipLaunchKernelGGL(k_mlp_train<true>, dim3(1), dim3(NT), 0, ctx->stream, b.p, b.m, b.v, b.x, b.t, b.idx, n, steps, lr,
                       (int64_t)e * steps, nullptr, b.l);
    PBF_HIP(hipGetLastError());
    PBF_HIP(hipStreamSynchronize(ctx->stream));  // the permutation buffer is reused next epoch
  }
  PBF_CHECK(pbf_d2h(ctx, params, b.p, sizeof(float) * NPARAM));
  float l = 0;
  PBF_CHECK(pbf_d2h(ctx, &l, b.l, sizeof(float)));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  if (final_loss) *final_loss = l;
  return PBF_OK;
}

extern "C" int pbf_mlp_grad(pbf_context* ctx, const float* params, const float* X, const float* T, int64_t n,
                            float* grads, float* loss) {
  if (!ctx || !params || !X || !T || !grads || n <= 0 || n > MB) return PBF_ERR_ARG;
  DevBufs b;
  PBF_CHECK(bufs(ctx, MB, &b));
  PBF_CHECK(pbf_h2d(ctx, b.p, params, sizeof(float) * NPARAM));
  PBF_CHECK(pbf_h2d(ctx, b.x, X, sizeof(float) * (size_t)n * IN));
  PBF_CHECK(pbf_h2d(ctx, b.t, T, sizeof(float) * (size_t)n * OUT));
  hipLaunchKernelGGL(k_mlp_train<false>, dim3(1), dim3(NT), 0, ctx->stream, b.p, b.m, b.v, b.x, b.t, nullptr, n, (int64_t)1,
                     0.f, (int64_t)0, b.g, b.l);
  PBF_HIP(hipGetLastError());
  PBF_CHECK(pbf_d2h(ctx, grads, b.g, sizeof(float) * NPARAM));
  float l = 0;
  PBF_CHECK(pbf_d2h(ctx, &l, b.l, sizeof(float)));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  if (loss) *loss = l;
  return PBF_OK;
}

extern "C" int pbf_mlp_predict(pbf_context* ctx, const float* params, const float* X, int64_t n, float* Y) {
  if (!ctx || !params || !X || !Y || n < 0) return PBF_ERR_ARG;
  if (n == 0) return PBF_OK;
  DevBufs b;
  PBF_CHECK(bufs(ctx, n, &b));
  PBF_CHECK(pbf_h2d(ctx, b.p, params, sizeof(float) * NPARAM));
  PBF_CHECK(pbf_h2d(ctx, b.x, X, sizeof(float) * (size_t)n * IN));
  const unsigned blocks = (unsigned)std::min<int64_t>((n + MB - 1) / MB, (int64_t)ctx->num_cus * 4);
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(NT), 0, ctx->stream, b.p, b.x, n, b.t);
  PBF_HIP(hipGetLastError());
  PBF_CHECK(pbf_d2h(ctx, Y, b.t, sizeof(float) * (size_t)n * OUT));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  return PBF_OK;
}
