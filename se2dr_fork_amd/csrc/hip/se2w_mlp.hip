// se2w_mlp.hip -- the seismogram surrogate: an MLP trained and evaluated with
// hand-written gfx950 kernels (include/se2w_mlp.h).
//
// Every product is one kernel shape, C[M][N] = sum_k A[M][k] B[N][k] with
// both operands BF16 and K-contiguous, on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation.  The three products of a layer (forward, input gradient,
// weight gradient) all take that shape because the forward and backward
// epilogues also write the transposes of the activations and of the
// pre-activation gradients, and the Adam update keeps a transposed BF16 copy
// of each weight matrix.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "se2w.h"
#include "se2w_mlp.h"

namespace se2w {
int set_error(int code, const char *msg);
}

namespace {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

int mfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return se2w::set_error(code, buf);
}

#define MCHK(call)                                                                             \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return mfail(SE2W_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { EPI_GELU = 0, EPI_LINEAR = 1, EPI_DGELU = 2, EPI_STORE = 3 };

struct GemmArgs {
  const bf16 *A;  // [M][lda]
  const bf16 *B;  // [N][ldb]
  int M, N, K, lda, ldb;
  const float *bias;   // [N]          (EPI_GELU, EPI_LINEAR)
  float *Cf;           // [M][ldc]     FP32 result (pre-activation or gradient)
  bf16 *Cb;            // [M][ldc]     BF16 activation / gradient
  bf16 *Cbt;           // [N][ldt]     its transpose
  const float *Zprev;  // [M][ldc]     pre-activation whose GELU' scales the product (EPI_DGELU)
  int ldc, ldt;
};

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

// 256 threads: four waves, each owning a 32 x 32 quarter of the 64 x 64 block
// tile as 2 x 2 MFMA tiles of 16 x 16.  MFMA 16x16x32 operand layout: lane l
// holds A[l % 16][8 (l / 16) .. +7] and B[8 (l / 16) .. +7][l % 16] (here B
// is stored [n][k], so also row l % 16, eight consecutive k); the result
// D[4 (l / 16) + r][l % 16] is its r-th accumulator.
template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[BM][LDK];
  __shared__ __attribute__((aligned(16))) bf16 Bs[BN][LDK];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int lr = tid >> 2, lk = (tid & 3) * 8;  // this thread's 16-byte piece of each operand tile
  floatx4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const bf16 *ga = g.A + (long)(m0 + lr) * g.lda + lk;
  const bf16 *gb = g.B + (long)(n0 + lr) * g.ldb + lk;
  for (int k0 = 0; k0 < g.K; k0 += BK) {
    *(uint4 *)&As[lr][lk] = *(const uint4 *)(ga + k0);
    *(uint4 *)&Bs[lr][lk] = *(const uint4 *)(gb + k0);
    __syncthreads();
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; i++) a[i] = *(const bf16x8 *)&As[wm + i * 16 + (lane & 15)][8 * (lane >> 4)];
    for (int j = 0; j < 2; j++) b[j] = *(const bf16x8 *)&Bs[wn + j * 16 + (lane & 15)][8 * (lane >> 4)];
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  for (int i = 0; i < 2; i++) {
    for (int j = 0; j < 2; j++) {
      const int n = n0 + wn + j * 16 + (lane & 15);
      const int mb = m0 + wm + i * 16 + 4 * (lane >> 4);
      bf16 t4[4];
      for (int r = 0; r < 4; r++) {
        const int m = mb + r;
        const long o = (long)m * g.ldc + n;
        const float c = acc[i][j][r];
        if constexpr (EPI == EPI_GELU) {
          const float z = c + g.bias[n];
          g.Cf[o] = z;
          const bf16 h = (bf16)gelu(z);
          g.Cb[o] = h;
          t4[r] = h;
        } else if constexpr (EPI == EPI_LINEAR) {
          g.Cf[o] = c + g.bias[n];
        } else if constexpr (EPI == EPI_DGELU) {
          const bf16 d = (bf16)(c * gelu_grad(g.Zprev[o]));
          g.Cb[o] = d;
          t4[r] = d;
        } else {
          g.Cf[o] = c;
        }
      }
      if constexpr (EPI == EPI_GELU || EPI == EPI_DGELU) {
        // rows mb..mb+3 of column n are consecutive in the transpose
        bf16 *t = g.Cbt + (long)n * g.ldt + mb;
        for (int r = 0; r < 4; r++) t[r] = t4[r];
      }
    }
  }
}

// Batch gather: rows idx[r] of the FP32 data become the BF16 input
// activations (and their transpose) and the FP32 targets of the batch.
__global__ void k_gather(const float *X, const float *T, const int *idx, int B, int in_pad, int out, bf16 *H0,
                         bf16 *H0t, float *Tb) {
  const int r = blockIdx.x;
  const long src = idx[r];
  for (int k = threadIdx.x; k < in_pad; k += blockDim.x) {
    const bf16 h = (bf16)X[src * in_pad + k];
    H0[(long)r * in_pad + k] = h;
    H0t[(long)k * B + r] = h;
  }
  if (T)
    for (int k = threadIdx.x; k < out; k += blockDim.x) Tb[(long)r * out + k] = T[src * out + k];
}

// Mean-squared error: dY = 2 (Y - T) / (B N) as BF16 (and transpose); the
// per-block sums of (Y - T)^2 go to part[] (summed in order afterwards).
__global__ __launch_bounds__(256) void k_loss(const float *Y, const float *Tb, int B, int N, float scale, bf16 *dZ,
                                              bf16 *dZt, double *part) {
  __shared__ double red[256];
  double s = 0.0;
  const long n = (long)B * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int m = (int)(i / N), c = (int)(i - (long)m * N);
    const float d = Y[i] - Tb[i];
    s += (double)d * d;
    const bf16 g = (bf16)(scale * d);
    dZ[i] = g;
    dZt[(long)c * B + m] = g;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// acc[slot] += sum of part[0..n) (one thread, fixed order), divided by count.
__global__ void k_sum_parts(const double *part, int n, double inv_count, double *acc, int slot) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += part[i];
    acc[slot] += s * inv_count;
  }
}

// Bias gradient: gb[o] = sum over the batch of dZ[b][o] (rows of the transpose).
__global__ void k_bias_grad(const bf16 *dZt, int B, int N, float *gb) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= N) return;
  const bf16 *row = dZt + (long)o * B;
  float s = 0.f;
  for (int b = 0; b < B; b++) s += (float)row[b];
  gb[o] = s;
}

struct AdamArgs {
  float lr, b1, b2, eps, bc1, bc2;  // bc = 1 - beta^t
  float wd;                         // decoupled weight decay (AdamW); 0 = Adam
};

__device__ __forceinline__ float adam_update(float &w, float &m, float &v, float g, const AdamArgs &a) {
  m = a.b1 * m + (1.f - a.b1) * g;
  v = a.b2 * v + (1.f - a.b2) * g * g;
  w -= a.lr * ((m / a.bc1) / (sqrtf(v / a.bc2) + a.eps) + a.wd * w);
  return w;
}

// Adam on a [rows][cols] weight matrix (rows, cols multiples of 32): FP32
// master weights, BF16 copy W[rows][cols] and transposed copy Wt[cols][rows].
// A 32 x 32 tile per workgroup of 32 x 8 threads; the transpose goes through LDS.
__global__ __launch_bounds__(256) void k_adam_w(float *W, float *m, float *v, const float *g, int rows, int cols,
                                                AdamArgs a, bf16 *Wb, bf16 *Wtb) {
  __shared__ bf16 tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const long o = (long)(r0 + k) * cols + c0 + tx;
    float w = W[o], mm = m[o], vv = v[o];
    adam_update(w, mm, vv, g[o], a);
    W[o] = w;
    m[o] = mm;
    v[o] = vv;
    const bf16 h = (bf16)w;
    Wb[o] = h;
    tile[k][tx] = h;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) Wtb[(long)(c0 + k) * rows + r0 + tx] = tile[tx][k];
}

__global__ void k_adam_b(float *b, float *m, float *v, const float *g, int n, AdamArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float w = b[i], mm = m[i], vv = v[i];
  adam_update(w, mm, vv, g[i], a);
  b[i] = w;
  m[i] = mm;
  v[i] = vv;
}

// BF16 copies of the FP32 master weights (after set_params / init).
__global__ void k_to_bf16(const float *W, int rows, int cols, bf16 *Wb, bf16 *Wtb) {
  const long n = (long)rows * cols;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int r = (int)(i / cols), c = (int)(i - (long)r * cols);
    const bf16 h = (bf16)W[i];
    Wb[i] = h;
    Wtb[(long)c * rows + r] = h;
  }
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }

}  // namespace

struct MlpLayer {
  int in = 0, in_pad = 0, out = 0;
  float *W = nullptr, *b = nullptr, *gW = nullptr, *gb = nullptr, *mW = nullptr, *vW = nullptr, *mb = nullptr,
        *vb = nullptr;
  bf16 *Wb = nullptr, *Wtb = nullptr;
  float *Z = nullptr;                        // pre-activation [B][out] (hidden layers)
  bf16 *H = nullptr, *Ht = nullptr;          // activation [B][out], [out][B] (hidden layers)
  bf16 *dZ = nullptr, *dZt = nullptr;        // d loss / d pre-activation [B][out], [out][B]
};

struct se2w_mlp {
  int device = 0, nl = 0, B = 0, in = 0, in_pad = 0, out = 0;
  std::vector<MlpLayer> L;
  bf16 *H0 = nullptr, *H0t = nullptr;
  float *Y = nullptr, *Tb = nullptr;
  float *Xs = nullptr, *Ts = nullptr;  // staging of one host batch
  int *idx = nullptr;                  // batch row indices
  float *X = nullptr, *T = nullptr;    // training data
  int ndata = 0;
  double *part = nullptr, *acc = nullptr;
  int nparts = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  long t = 0;  // Adam steps taken
  float weight_decay = 0.f;
  int cosine = 0;  // cosine learning-rate decay over each se2w_mlp_train call
  double train_ms = 0.0, fwd_ms = 0.0;
  std::vector<void *> allocs;
};

namespace {

template <typename T>
int malloc_zero(se2w_mlp *m, T **p, size_t n) {
  *p = nullptr;
  void *q = nullptr;
  if (hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess)
    return mfail(SE2W_ERR_MEM, "hipMalloc of %zu bytes failed", sizeof(T) * n);
  m->allocs.push_back(q);
  MCHK(hipMemsetAsync(q, 0, sizeof(T) * std::max<size_t>(n, 1), m->stream));
  *p = (T *)q;
  return SE2W_OK;
}

template <int EPI>
int gemm(se2w_mlp *m, const GemmArgs &g) {
  if (g.M % BM || g.N % BN || g.K % BK || g.lda % 8 || g.ldb % 8)
    return mfail(SE2W_ERR_ARG, "product shape %d x %d x %d not tileable", g.M, g.N, g.K);
  hipLaunchKernelGGL((gemm_nt<EPI>), dim3(g.N / BN, g.M / BM), dim3(256), 0, m->stream, g);
  return SE2W_OK;
}

// Forward pass of the batch in H0/H0t; output pre-activations in Y.
int forward(se2w_mlp *m) {
  for (int l = 0; l < m->nl; l++) {
    MlpLayer &Ly = m->L[l];
    GemmArgs g = {};
    g.A = l == 0 ? m->H0 : m->L[l - 1].H;
    g.lda = Ly.in_pad;
    g.B = Ly.Wb;
    g.ldb = Ly.in_pad;
    g.M = m->B;
    g.N = Ly.out;
    g.K = Ly.in_pad;
    g.bias = Ly.b;
    g.ldc = Ly.out;
    int ierr;
    if (l < m->nl - 1) {
      g.Cf = Ly.Z;
      g.Cb = Ly.H;
      g.Cbt = Ly.Ht;
      g.ldt = m->B;
      ierr = gemm<EPI_GELU>(m, g);
    } else {
      g.Cf = m->Y;
      ierr = gemm<EPI_LINEAR>(m, g);
    }
    if (ierr) return ierr;
  }
  MCHK(hipGetLastError());
  return SE2W_OK;
}

// Loss of the batch (added to acc[slot] / count) and all parameter gradients.
int backward(se2w_mlp *m, int slot, double inv_count) {
  MlpLayer &Lo = m->L[m->nl - 1];
  const float scale = 2.0f / ((float)m->B * (float)m->out);
  hipLaunchKernelGGL(k_loss, dim3(m->nparts), dim3(256), 0, m->stream, m->Y, m->Tb, m->B, m->out, scale, Lo.dZ,
                     Lo.dZt, m->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, m->stream, m->part, m->nparts, inv_count, m->acc, slot);
  for (int l = m->nl - 1; l >= 0; l--) {
    MlpLayer &Ly = m->L[l];
    GemmArgs g = {};
    g.A = Ly.dZt;
    g.lda = m->B;
    g.B = l == 0 ? m->H0t : m->L[l - 1].Ht;
    g.ldb = m->B;
    g.M = Ly.out;
    g.N = Ly.in_pad;
    g.K = m->B;
    g.Cf = Ly.gW;
    g.ldc = Ly.in_pad;
    if (int ierr = gemm<EPI_STORE>(m, g)) return ierr;
    hipLaunchKernelGGL(k_bias_grad, dim3((Ly.out + 255) / 256), dim3(256), 0, m->stream, Ly.dZt, m->B, Ly.out, Ly.gb);
    if (l > 0) {
      MlpLayer &Lp = m->L[l - 1];
      GemmArgs h = {};
      h.A = Ly.dZ;
      h.lda = Ly.out;
      h.B = Ly.Wtb;
      h.ldb = Ly.out;
      h.M = m->B;
      h.N = Ly.in_pad;
      h.K = Ly.out;
      h.Zprev = Lp.Z;
      h.ldc = Lp.out;
      h.Cb = Lp.dZ;
      h.Cbt = Lp.dZt;
      h.ldt = m->B;
      if (int ierr = gemm<EPI_DGELU>(m, h)) return ierr;
    }
  }
  MCHK(hipGetLastError());
  return SE2W_OK;
}

int adam(se2w_mlp *m, float lr) {
  m->t++;
  AdamArgs a;
  a.lr = lr;
  a.b1 = 0.9f;
  a.b2 = 0.999f;
  a.eps = 1e-8f;
  a.wd = m->weight_decay;
  a.bc1 = (float)(1.0 - std::pow(0.9, (double)m->t));
  a.bc2 = (float)(1.0 - std::pow(0.999, (double)m->t));
  for (MlpLayer &Ly : m->L) {
    hipLaunchKernelGGL(k_adam_w, dim3(Ly.in_pad / 32, Ly.out / 32), dim3(256), 0, m->stream, Ly.W, Ly.mW, Ly.vW,
                       (const float *)Ly.gW, Ly.out, Ly.in_pad, a, Ly.Wb, Ly.Wtb);
    hipLaunchKernelGGL(k_adam_b, dim3((Ly.out + 255) / 256), dim3(256), 0, m->stream, Ly.b, Ly.mb, Ly.vb,
                       (const float *)Ly.gb, Ly.out, a);
  }
  MCHK(hipGetLastError());
  return SE2W_OK;
}

// Uploads n <= B host rows into the staging buffers (zero rows after n) and
// gathers them as the current batch.
int stage_batch(se2w_mlp *m, int n, const float *x, const float *t) {
  std::vector<float> xp((size_t)m->B * m->in_pad, 0.f);
  for (int r = 0; r < n; r++) memcpy(&xp[(size_t)r * m->in_pad], x + (size_t)r * m->in, sizeof(float) * m->in);
  MCHK(hipMemcpyAsync(m->Xs, xp.data(), sizeof(float) * xp.size(), hipMemcpyHostToDevice, m->stream));
  if (t) MCHK(hipMemcpyAsync(m->Ts, t, sizeof(float) * (size_t)n * m->out, hipMemcpyHostToDevice, m->stream));
  std::vector<int> id(m->B);
  for (int r = 0; r < m->B; r++) id[r] = r;
  MCHK(hipMemcpyAsync(m->idx, id.data(), sizeof(int) * m->B, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_gather, dim3(m->B), dim3(128), 0, m->stream, m->Xs, t ? m->Ts : nullptr, m->idx, m->B,
                     m->in_pad, m->out, m->H0, m->H0t, m->Tb);
  MCHK(hipStreamSynchronize(m->stream));  // the host vectors above go out of scope
  return SE2W_OK;
}

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

int refresh_bf16(se2w_mlp *m) {
  for (MlpLayer &Ly : m->L)
    hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, m->stream, (const float *)Ly.W, Ly.out, Ly.in_pad, Ly.Wb,
                       Ly.Wtb);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->stream));
  return SE2W_OK;
}

}  // namespace

extern "C" {

int se2w_mlp_create(int32_t nlayers, const int32_t *widths, int32_t batch, uint64_t seed, int32_t device,
                    se2w_mlp **out) {
  if (!out || !widths || nlayers < 1 || batch < 64 || batch % 64) return mfail(SE2W_ERR_ARG, "bad network arguments");
  *out = nullptr;
  if (widths[0] < 1) return mfail(SE2W_ERR_ARG, "input width must be >= 1");
  for (int l = 1; l <= nlayers; l++)
    if (widths[l] < 64 || widths[l] % 64) return mfail(SE2W_ERR_ARG, "width %d (layer %d) must be a multiple of 64", widths[l], l);
  MCHK(hipSetDevice(device));
  se2w_mlp *m = new se2w_mlp;
  m->device = device;
  m->nl = nlayers;
  m->B = batch;
  m->in = widths[0];
  m->in_pad = round_up(widths[0], 64);
  m->out = widths[nlayers];
  int ierr = SE2W_OK;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&m->e0) != hipSuccess ||
      hipEventCreate(&m->e1) != hipSuccess) {
    se2w_mlp_destroy(m);
    return mfail(SE2W_ERR_HIP, "stream/event creation failed");
  }
  m->L.resize(nlayers);
  uint64_t rng = seed;
  for (int l = 0; l < nlayers && !ierr; l++) {
    MlpLayer &Ly = m->L[l];
    Ly.in = widths[l];
    Ly.in_pad = l == 0 ? m->in_pad : widths[l];
    Ly.out = widths[l + 1];
    const size_t nw = (size_t)Ly.out * Ly.in_pad;
    if ((ierr = malloc_zero(m, &Ly.W, nw)) || (ierr = malloc_zero(m, &Ly.b, Ly.out)) ||
        (ierr = malloc_zero(m, &Ly.gW, nw)) || (ierr = malloc_zero(m, &Ly.gb, Ly.out)) ||
        (ierr = malloc_zero(m, &Ly.mW, nw)) || (ierr = malloc_zero(m, &Ly.vW, nw)) ||
        (ierr = malloc_zero(m, &Ly.mb, Ly.out)) || (ierr = malloc_zero(m, &Ly.vb, Ly.out)) ||
        (ierr = malloc_zero(m, &Ly.Wb, nw)) || (ierr = malloc_zero(m, &Ly.Wtb, nw)) ||
        (ierr = malloc_zero(m, &Ly.dZ, (size_t)batch * Ly.out)) || (ierr = malloc_zero(m, &Ly.dZt, (size_t)batch * Ly.out)))
      break;
    if (l < nlayers - 1 &&
        ((ierr = malloc_zero(m, &Ly.Z, (size_t)batch * Ly.out)) || (ierr = malloc_zero(m, &Ly.H, (size_t)batch * Ly.out)) ||
         (ierr = malloc_zero(m, &Ly.Ht, (size_t)batch * Ly.out))))
      break;
    // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
    const double bound = 1.0 / std::sqrt((double)Ly.in);
    std::vector<float> w(nw, 0.f), b(Ly.out);
    for (int o = 0; o < Ly.out; o++)
      for (int i = 0; i < Ly.in; i++) w[(size_t)o * Ly.in_pad + i] = (float)((2.0 * uniform01(rng) - 1.0) * bound);
    for (int o = 0; o < Ly.out; o++) b[o] = (float)((2.0 * uniform01(rng) - 1.0) * bound);
    if (hipMemcpyAsync(Ly.W, w.data(), sizeof(float) * nw, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(Ly.b, b.data(), sizeof(float) * Ly.out, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
      ierr = mfail(SE2W_ERR_HIP, "weight upload failed");
  }
  m->nparts = 256;
  if (!ierr && ((ierr = malloc_zero(m, &m->H0, (size_t)batch * m->in_pad)) ||
                (ierr = malloc_zero(m, &m->H0t, (size_t)batch * m->in_pad)) ||
                (ierr = malloc_zero(m, &m->Y, (size_t)batch * m->out)) ||
                (ierr = malloc_zero(m, &m->Tb, (size_t)batch * m->out)) ||
                (ierr = malloc_zero(m, &m->Xs, (size_t)batch * m->in_pad)) ||
                (ierr = malloc_zero(m, &m->Ts, (size_t)batch * m->out)) || (ierr = malloc_zero(m, &m->idx, batch)) ||
                (ierr = malloc_zero(m, &m->part, m->nparts)) || (ierr = malloc_zero(m, &m->acc, 1))))
    ;
  if (!ierr) ierr = refresh_bf16(m);
  if (ierr) {
    se2w_mlp_destroy(m);
    return ierr;
  }
  *out = m;
  return SE2W_OK;
}

int se2w_mlp_param_count(se2w_mlp *m, int64_t *count) {
  if (!m || !count) return mfail(SE2W_ERR_ARG, "null argument");
  int64_t n = 0;
  for (const MlpLayer &Ly : m->L) n += (int64_t)Ly.out * Ly.in + Ly.out;
  *count = n;
  return SE2W_OK;
}

// copy between the packed host layout and the padded device weights
static int params_copy(se2w_mlp *m, float *host, bool to_host, const std::vector<float *> &src_w,
                       const std::vector<float *> &src_b) {
  size_t off = 0;
  for (int l = 0; l < m->nl; l++) {
    const MlpLayer &Ly = m->L[l];
    MCHK(hipMemcpy2DAsync(to_host ? (void *)(host + off) : (void *)src_w[l], to_host ? sizeof(float) * Ly.in : sizeof(float) * Ly.in_pad,
                          to_host ? (const void *)src_w[l] : (const void *)(host + off),
                          to_host ? sizeof(float) * Ly.in_pad : sizeof(float) * Ly.in, sizeof(float) * Ly.in, Ly.out,
                          to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice, m->stream));
    off += (size_t)Ly.out * Ly.in;
    MCHK(hipMemcpyAsync(to_host ? (void *)(host + off) : (void *)src_b[l], to_host ? (const void *)src_b[l] : (const void *)(host + off),
                        sizeof(float) * Ly.out, to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice, m->stream));
    off += Ly.out;
  }
  MCHK(hipStreamSynchronize(m->stream));
  return SE2W_OK;
}

int se2w_mlp_get_params(se2w_mlp *m, float *params) {
  if (!m || !params) return mfail(SE2W_ERR_ARG, "null argument");
  MCHK(hipSetDevice(m->device));
  std::vector<float *> w, b;
  for (MlpLayer &Ly : m->L) { w.push_back(Ly.W); b.push_back(Ly.b); }
  return params_copy(m, params, true, w, b);
}

int se2w_mlp_set_params(se2w_mlp *m, const float *params) {
  if (!m || !params) return mfail(SE2W_ERR_ARG, "null argument");
  MCHK(hipSetDevice(m->device));
  std::vector<float *> w, b;
  for (MlpLayer &Ly : m->L) {
    w.push_back(Ly.W);
    b.push_back(Ly.b);
    MCHK(hipMemsetAsync(Ly.mW, 0, sizeof(float) * (size_t)Ly.out * Ly.in_pad, m->stream));
    MCHK(hipMemsetAsync(Ly.vW, 0, sizeof(float) * (size_t)Ly.out * Ly.in_pad, m->stream));
    MCHK(hipMemsetAsync(Ly.mb, 0, sizeof(float) * Ly.out, m->stream));
    MCHK(hipMemsetAsync(Ly.vb, 0, sizeof(float) * Ly.out, m->stream));
  }
  m->t = 0;
  if (int ierr = params_copy(m, (float *)params, false, w, b)) return ierr;
  return refresh_bf16(m);
}

int se2w_mlp_forward(se2w_mlp *m, int32_t n, const float *x, float *y) {
  if (!m || n < 0 || (n > 0 && (!x || !y))) return mfail(SE2W_ERR_ARG, "bad forward arguments");
  MCHK(hipSetDevice(m->device));
  float total = 0.f;
  for (int r0 = 0; r0 < n; r0 += m->B) {
    const int k = std::min(m->B, n - r0);
    if (int ierr = stage_batch(m, k, x + (size_t)r0 * m->in, nullptr)) return ierr;
    MCHK(hipEventRecord(m->e0, m->stream));
    if (int ierr = forward(m)) return ierr;
    MCHK(hipEventRecord(m->e1, m->stream));
    MCHK(hipMemcpyAsync(y + (size_t)r0 * m->out, m->Y, sizeof(float) * (size_t)k * m->out, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    float ms = 0.f;
    MCHK(hipEventElapsedTime(&ms, m->e0, m->e1));
    total += ms;
  }
  m->fwd_ms = total;
  return SE2W_OK;
}

int se2w_mlp_gradient(se2w_mlp *m, int32_t n, const float *x, const float *target, float *loss, float *grad) {
  if (!m || !x || !target || !grad) return mfail(SE2W_ERR_ARG, "null argument");
  if (n != m->B) return mfail(SE2W_ERR_ARG, "gradient batch must have %d rows", m->B);
  MCHK(hipSetDevice(m->device));
  if (int ierr = stage_batch(m, n, x, target)) return ierr;
  MCHK(hipMemsetAsync(m->acc, 0, sizeof(double), m->stream));
  if (int ierr = forward(m)) return ierr;
  if (int ierr = backward(m, 0, 1.0 / ((double)m->B * m->out))) return ierr;
  double l = 0.0;
  MCHK(hipMemcpyAsync(&l, m->acc, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  std::vector<float *> w, b;
  for (MlpLayer &Ly : m->L) { w.push_back(Ly.gW); b.push_back(Ly.gb); }
  if (int ierr = params_copy(m, grad, true, w, b)) return ierr;
  if (loss) *loss = (float)l;
  return SE2W_OK;
}

int se2w_mlp_adam_step(se2w_mlp *m, const float *grad, float lr) {
  if (!m || !grad) return mfail(SE2W_ERR_ARG, "null argument");
  MCHK(hipSetDevice(m->device));
  std::vector<float *> w, b;
  for (MlpLayer &Ly : m->L) { w.push_back(Ly.gW); b.push_back(Ly.gb); }
  if (int ierr = params_copy(m, (float *)grad, false, w, b)) return ierr;
  if (int ierr = adam(m, lr)) return ierr;
  MCHK(hipStreamSynchronize(m->stream));
  return SE2W_OK;
}

int se2w_mlp_set_data(se2w_mlp *m, int32_t n, const float *x, const float *target) {
  if (!m || n < m->B || !x || !target) return mfail(SE2W_ERR_ARG, "need at least one batch of data");
  MCHK(hipSetDevice(m->device));
  if (m->X) hipFree(m->X);
  if (m->T) hipFree(m->T);
  m->X = m->T = nullptr;
  m->ndata = 0;
  MCHK(hipMalloc(&m->X, sizeof(float) * (size_t)n * m->in_pad));
  MCHK(hipMalloc(&m->T, sizeof(float) * (size_t)n * m->out));
  MCHK(hipMemset2DAsync(m->X, sizeof(float) * m->in_pad, 0, sizeof(float) * m->in_pad, n, m->stream));
  MCHK(hipMemcpy2DAsync(m->X, sizeof(float) * m->in_pad, x, sizeof(float) * m->in, sizeof(float) * m->in, n,
                        hipMemcpyHostToDevice, m->stream));
  MCHK(hipMemcpyAsync(m->T, target, sizeof(float) * (size_t)n * m->out, hipMemcpyHostToDevice, m->stream));
  MCHK(hipStreamSynchronize(m->stream));
  m->ndata = n;
  return SE2W_OK;
}

int se2w_mlp_train(se2w_mlp *m, int32_t epochs, float lr, uint64_t shuffle_seed, float *epoch_loss) {
  if (!m || epochs < 1 || !(lr > 0.f)) return mfail(SE2W_ERR_ARG, "bad training arguments");
  if (m->ndata < m->B) return mfail(SE2W_ERR_STATE, "set the training data first");
  MCHK(hipSetDevice(m->device));
  const int nb = m->ndata / m->B;
  std::vector<int> perm(m->ndata);
  int *d_perm = nullptr;
  double *d_loss = nullptr;
  MCHK(hipMalloc(&d_perm, sizeof(int) * (size_t)nb * m->B));
  if (hipMalloc(&d_loss, sizeof(double) * epochs) != hipSuccess) {
    hipFree(d_perm);
    return mfail(SE2W_ERR_MEM, "hipMalloc failed");
  }
  int ierr = SE2W_OK;
  uint64_t rng = shuffle_seed;
  if (hipMemsetAsync(d_loss, 0, sizeof(double) * epochs, m->stream) != hipSuccess ||
      hipEventRecord(m->e0, m->stream) != hipSuccess)
    ierr = mfail(SE2W_ERR_HIP, "training setup failed");
  for (int e = 0; e < epochs && !ierr; e++) {
    for (int i = 0; i < m->ndata; i++) perm[i] = i;
    for (int i = m->ndata - 1; i > 0; i--) {  // Fisher-Yates
      const int j = (int)(uniform01(rng) * (i + 1));
      std::swap(perm[i], perm[j]);
    }
    // the previous epoch's kernels may still read d_perm
    if (hipStreamSynchronize(m->stream) != hipSuccess ||
        hipMemcpy(d_perm, perm.data(), sizeof(int) * (size_t)nb * m->B, hipMemcpyHostToDevice) != hipSuccess) {
      ierr = mfail(SE2W_ERR_HIP, "permutation upload failed");
      break;
    }
    for (int k = 0; k < nb && !ierr; k++) {
      hipLaunchKernelGGL(k_gather, dim3(m->B), dim3(128), 0, m->stream, (const float *)m->X, (const float *)m->T,
                         (const int *)(d_perm + (size_t)k * m->B), m->B, m->in_pad, m->out, m->H0, m->H0t, m->Tb);
      if ((ierr = forward(m))) break;
      // each batch loss enters d_loss[e] as loss / nb
      double *acc0 = m->acc;
      m->acc = d_loss;
      ierr = backward(m, e, 1.0 / ((double)m->B * m->out * nb));
      m->acc = acc0;
      if (!ierr) {
        // cosine decay: lr_s = lr (1 + cos(pi s / S)) / 2 over the S steps of this call
        const double frac = (double)(e * nb + k) / ((double)epochs * nb);
        const float lr_s = m->cosine ? (float)(0.5 * lr * (1.0 + std::cos(3.14159265358979323846 * frac))) : lr;
        ierr = adam(m, lr_s);
      }
    }
  }
  if (!ierr) {
    if (hipEventRecord(m->e1, m->stream) != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess) {
      ierr = mfail(SE2W_ERR_HIP, "training failed: %s", hipGetErrorString(hipGetLastError()));
    } else {
      float ms = 0.f;
      hipEventElapsedTime(&ms, m->e0, m->e1);
      m->train_ms = ms;
      if (epoch_loss) {
        std::vector<double> l(epochs);
        hipMemcpy(l.data(), d_loss, sizeof(double) * epochs, hipMemcpyDeviceToHost);
        for (int e = 0; e < epochs; e++) epoch_loss[e] = (float)l[e];
      }
    }
  }
  hipStreamSynchronize(m->stream);
  hipFree(d_perm);
  hipFree(d_loss);
  return ierr;
}

int se2w_mlp_set_lr_schedule(se2w_mlp *m, int32_t cosine) {
  if (!m) return mfail(SE2W_ERR_ARG, "null argument");
  m->cosine = cosine != 0;
  return SE2W_OK;
}

int se2w_mlp_set_weight_decay(se2w_mlp *m, float wd) {
  if (!m || !(wd >= 0.f)) return mfail(SE2W_ERR_ARG, "bad weight decay");
  m->weight_decay = wd;
  return SE2W_OK;
}

int se2w_mlp_get_timing(se2w_mlp *m, double *train_ms, double *forward_ms) {
  if (!m) return mfail(SE2W_ERR_ARG, "null argument");
  if (train_ms) *train_ms = m->train_ms;
  if (forward_ms) *forward_ms = m->fwd_ms;
  return SE2W_OK;
}

int se2w_mlp_destroy(se2w_mlp *m) {
  if (!m) return SE2W_OK;
  hipSetDevice(m->device);
  if (m->stream) hipStreamSynchronize(m->stream);
  for (void *p : m->allocs) hipFree(p);
  if (m->X) hipFree(m->X);
  if (m->T) hipFree(m->T);
  if (m->e0) hipEventDestroy(m->e0);
  if (m->e1) hipEventDestroy(m->e1);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return SE2W_OK;
}

}  // extern "C"
