// mlp_device.hip — the learned warm start: a per-vertex MLP
// 12 -> 128 -> 128 -> 128 -> 3 (GELU), BF16 weights and activations, FP32
// accumulation on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).
// No reference counterpart: it predicts each step's displacement in place of
// the adaptive inertial guess (src/vbd/prediction.rs:15-41), and the step
// clips it to the OGC balls exactly as it clips that guess (src/vbd.rs:316-324).
//
// Matrix-core layout of v_mfma_f32_16x16x32_bf16 (one wavefront, checked on
// the device): lane l holds A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16]
// for j < 8, and D[4 (l / 16) + j][l % 16] for j < 4.
//
// Weights live as FP32 master parameters (training state) and as BF16 copies
// pre-swizzled per 32 x 16 B-operand tile, so a lane's 8 operands are one
// 16-byte load: forward copies W (K x N) for Z = A W, transposed copies for
// the backward products dA = dZ W^T.  Activations of a 16-row tile sit in LDS
// rows of 136 bf16 (272 bytes, 16-byte aligned, no bank aliasing of rows).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "gpu.hpp"
#include "mlp.hpp"

using namespace cvmlp;

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;

constexpr int ROW = 136;  // LDS row stride (bf16 elements)

// Round-to-nearest-even FP32 -> BF16 (bit pattern), as the CPU checker does:
// one v_cvt_pk_bf16_f32 on gfx950.
__device__ __forceinline__ uint16_t bf16_bits(float f) {
  __bf16 h = (__bf16)f;
  uint16_t u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}
__device__ __forceinline__ float bf16_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Orders this wavefront's LDS writes before its later LDS reads of other
// lanes' data.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// GELU z Phi(z) with Phi(z) = (1 + erf(z / sqrt 2)) / 2, and its derivative
// Phi(z) + z phi(z).  erf is evaluated with Abramowitz & Stegun 7.1.25
// (|error| <= 2.5e-5, two orders below the BF16 rounding (2^-9 relative) the
// activations go through): erf(x) = 1 - (a1 t + a2 t^2 + a3 t^3) exp(-x^2),
// t = 1 / (1 + p x), x >= 0; exp(-x^2) = exp(-z^2 / 2) is shared with phi(z).
// The inference kernel is bound by this VALU work, not by the matrix cores.
struct GeluParts {
  float cdf, pdf;  // Phi(z), phi(z)
};
__device__ __forceinline__ GeluParts gelu_parts(float z) {
  float x = fabsf(z) * 0.70710678118654752f;
  float t = __builtin_amdgcn_rcpf(fmaf(0.47047f, x, 1.0f));  // v_rcp_f32, 1 ulp
  float poly = fmaf(fmaf(0.7478556f, t, -0.0958798f), t, 0.3480242f) * t;
  // exp(-z^2 / 2) = 2^(-z^2 log2(e) / 2) with the bare v_exp_f32 (results
  // below 2^-126 flush to 0, where the GELU terms vanish anyway).
  float e = __builtin_amdgcn_exp2f(z * z * -0.72134752044448170f);
  float erfx = fmaf(-poly, e, 1.0f);
  GeluParts g;
  g.cdf = fmaf(0.5f, copysignf(erfx, z), 0.5f);
  g.pdf = 0.39894228040143268f * e;
  return g;
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_parts(z).cdf; }
__device__ __forceinline__ float gelu_d(float z) {
  GeluParts g = gelu_parts(z);
  return g.cdf + z * g.pdf;
}

__device__ __forceinline__ f4 mfma(bf16x8 a, bf16x8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// A operand of a 16-row tile from LDS rows (row stride ROW), K chunk kc.
__device__ __forceinline__ bf16x8 lds_a(const uint16_t* act, int kc, int lane) {
  return *(const bf16x8*)(act + (lane % 16) * ROW + kc * 32 + 8 * (lane / 16));
}
// A operand whose rows are features and columns samples (a transposed read,
// for weight gradients): A[m][k] = act[k][m0 + m].
__device__ __forceinline__ bf16x8 lds_at(const uint16_t* act, int m0, int k0, int lane) {
  uint16_t t[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) t[j] = act[(k0 + 8 * (lane / 16) + j) * ROW + m0 + (lane % 16)];
  bf16x8 r;
  __builtin_memcpy(&r, t, 16);
  return r;
}
// B operand B[k][n] = buf[k0 + k][n0 + n] from LDS rows.
__device__ __forceinline__ bf16x8 lds_b(const uint16_t* buf, int k0, int n0, int lane) {
  uint16_t t[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) t[j] = buf[(k0 + 8 * (lane / 16) + j) * ROW + n0 + (lane % 16)];
  bf16x8 r;
  __builtin_memcpy(&r, t, 16);
  return r;
}
__device__ __forceinline__ bf16x8 gl_b(const uint16_t* w, int tile, int lane) {
  return *(const bf16x8*)(w + ((size_t)tile * 64 + lane) * 8);
}

struct Weights {
  const uint16_t* f[4];  // forward tiles per layer
  const uint16_t* b[4];  // transposed tiles (layers 1..3)
  const float* bias[4];
};

// Z = A W + b for one 16-row tile; `emit(nt, acc)` receives each 16-column
// output tile (lane l: rows 4 (l/16) + j, column nt*16 + l%16).
template <class Emit>
__device__ __forceinline__ void layer(const uint16_t* act, const uint16_t* wf, int kchunks, int ntiles, int lane,
                                      Emit&& emit) {
  for (int nt = 0; nt < ntiles; ++nt) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < kchunks; ++kc) acc = mfma(lds_a(act, kc, lane), gl_b(wf, kc * ntiles + nt, lane), acc);
    emit(nt, acc);
  }
}

// Forward of one 16-row tile.  act0 holds the inputs (cols 0..11, zero to 31);
// the hidden activations go to act1..act3 (bf16), pre-activation derivatives
// to gd1..gd3 when training; y (3 floats per row) to `y`.
template <bool TRAIN>
__device__ void forward_tile(const Weights& W, uint16_t* const act[4], uint16_t* const gd[3], float* y,
                             int lane) {
  for (int l = 0; l < 3; ++l) {
    int kch = l == 0 ? 1 : 4;
    uint16_t* out = act[l + 1];
    uint16_t* g = TRAIN ? gd[l] : nullptr;
    const float* bias = W.bias[l];
    layer(act[l], W.f[l], kch, 8, lane, [&](int nt, f4 acc) {
      int c = nt * 16 + lane % 16;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int r = 4 * (lane / 16) + j;
        float z = acc[j] + bias[c];
        out[r * ROW + c] = bf16_bits(gelu(z));
        if (TRAIN) g[r * ROW + c] = bf16_bits(gelu_d(z));
      }
    });
    wave_sync();
  }
  layer(act[3], W.f[3], 4, 1, lane, [&](int, f4 acc) {
    int c = lane % 16;
    if (c < OUT)
#pragma unroll
      for (int j = 0; j < 4; ++j) y[(4 * (lane / 16) + j) * OUT + c] = acc[j] + W.bias[3][c];
  });
}

__device__ __forceinline__ void load_inputs(const float* X, uint32_t n, uint32_t row0, uint16_t* act0, int lane) {
  // 16 rows x 32 columns: lane l covers row l / 4, columns 8 (l % 4) .. + 7.
  int r = lane / 4, c0 = 8 * (lane % 4);
  uint32_t s = row0 + r;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    int c = c0 + j;
    float v = (s < n && c < IN) ? X[(size_t)s * IN + c] : 0.f;
    act0[r * ROW + c] = bf16_bits(v);
  }
}

// Inference.  The forward weight tiles (77.8 KB of pre-swizzled BF16) and
// biases are staged in LDS once per block; each of the block's 8 wavefronts
// then runs 16-row tiles on its own (no block barriers), with the output
// column tiles of a layer accumulating independently (up to 8 MFMA chains
// in flight) and two wavefronts per SIMD, so one wavefront's GELU work
// overlaps the other's matrix-core work.  The LDS footprint holds one block
// per CU; the grid strides over the row tiles.
constexpr int IW = 8;  // wavefronts per inference block
__device__ __forceinline__ bf16x8 lds_w(const uint16_t* w, int tile, int lane) {
  return *(const bf16x8*)(w + (tile * 64 + lane) * 8);
}
// One 16-row tile through a hidden layer: out = gelu(in W + b), KC K-chunks
// of 32, 8 column tiles of 16.
template <int KC>
__device__ __forceinline__ void hidden_layer(const uint16_t* in, const uint16_t* w, const float* bias, uint16_t* out,
                                             int lane) {
  f4 acc[8];
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) acc[nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kc = 0; kc < KC; ++kc) {
    bf16x8 a = lds_a(in, kc, lane);
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = mfma(a, lds_w(w, kc * 8 + nt, lane), acc[nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) {
    int c = nt * 16 + lane % 16;
    float bc = bias[c];
#pragma unroll
    for (int j = 0; j < 4; ++j) out[(4 * (lane / 16) + j) * ROW + c] = bf16_bits(gelu(acc[nt][j] + bc));
  }
}
__global__ __launch_bounds__(64 * IW) void k_mlp_infer(Weights W, const float* X, uint32_t n, float* Y) {
  __shared__ __attribute__((aligned(16))) uint16_t wl[FW_TOTAL];
  __shared__ __attribute__((aligned(16))) uint16_t buf[IW][2][16 * ROW];
  __shared__ float ybuf[IW][16 * OUT];
  __shared__ float bl[3 * H + OUT];
  {
    const uint4* src = (const uint4*)W.f[0];  // the layers' tiles are contiguous from layer 0
    uint4* dst = (uint4*)wl;
    for (uint32_t i = threadIdx.x; i < FW_TOTAL / 8; i += blockDim.x) dst[i] = src[i];
    for (uint32_t i = threadIdx.x; i < 3 * H + OUT; i += blockDim.x) bl[i] = W.bias[i / H < 3 ? i / H : 3][i % H];
  }
  __syncthreads();
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  uint16_t* a = buf[wave][0];
  uint16_t* b = buf[wave][1];
  const uint32_t tiles = (n + 15) / 16;
  for (uint32_t t = blockIdx.x * IW + wave; t < tiles; t += gridDim.x * IW) {
    load_inputs(X, n, t * 16, a, lane);
    wave_sync();
    hidden_layer<1>(a, wl + FW_OFS[0], bl, b, lane);
    wave_sync();
    hidden_layer<4>(b, wl + FW_OFS[1], bl + H, a, lane);
    wave_sync();
    hidden_layer<4>(a, wl + FW_OFS[2], bl + 2 * H, b, lane);
    wave_sync();
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) acc = mfma(lds_a(b, kc, lane), lds_w(wl + FW_OFS[3], kc, lane), acc);
    int c = lane % 16;
    if (c < OUT)
#pragma unroll
      for (int j = 0; j < 4; ++j) ybuf[wave][(4 * (lane / 16) + j) * OUT + c] = acc[j] + bl[3 * H + c];
    wave_sync();
    if (lane < 16 * OUT) {
      uint32_t s = t * 16 + lane / OUT;
      if (s < n) Y[(size_t)s * OUT + lane % OUT] = ybuf[wave][lane];
    }
    wave_sync();
  }
}

// Column sums of a 16-row accumulator tile (lane l < 16 gets column l).
__device__ __forceinline__ float col_sum(f4 v, int lane) {
  float s = v[0] + v[1] + v[2] + v[3];
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  return s;
}

// Loss and gradient of one block of 32 samples (2 wavefronts x 16 rows):
// mean squared error over 3 outputs of the batch (scale 1 / (3 B)), written as
// this block's partial gradient (parameter order of the master vector) and
// partial loss.  Deterministic: fixed reduction orders everywhere.
constexpr int GW = 2;  // wavefronts per gradient block
__global__ __launch_bounds__(64 * GW) void k_mlp_grad(Weights W, const float* X, const float* T, uint32_t n,
                                                      uint32_t row0, float inv_count, float* partial,
                                                      float* loss_partial) {
  __shared__ __attribute__((aligned(16))) uint16_t act[4][GW * 16 * ROW];  // a0..a3
  __shared__ __attribute__((aligned(16))) uint16_t gd[3][GW * 16 * ROW];   // gelu'(z1..z3)
  __shared__ __attribute__((aligned(16))) uint16_t dz[4][GW * 16 * ROW];   // dz1..dz3, dz4 = dy
  __shared__ float ybuf[GW][16 * OUT];
  __shared__ float dbias[GW][4][H];
  __shared__ float lpart[GW];
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  uint32_t base = row0 + blockIdx.x * (16 * GW) + wave * 16;
  uint16_t* a[4] = {act[0] + wave * 16 * ROW, act[1] + wave * 16 * ROW, act[2] + wave * 16 * ROW,
                    act[3] + wave * 16 * ROW};
  uint16_t* g[3] = {gd[0] + wave * 16 * ROW, gd[1] + wave * 16 * ROW, gd[2] + wave * 16 * ROW};
  uint16_t* d[4] = {dz[0] + wave * 16 * ROW, dz[1] + wave * 16 * ROW, dz[2] + wave * 16 * ROW,
                    dz[3] + wave * 16 * ROW};
  load_inputs(X, n, base, a[0], lane);
  wave_sync();
  forward_tile<true>(W, a, g, ybuf[wave], lane);
  // dy = 2 (y - t) / (3 B) in columns 0..2 of dz4 (zero to 31); loss.
  {
    int r = lane / 4, c0 = 8 * (lane % 4);
    uint32_t s = base + r;
    float l2 = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int c = c0 + j;
      float v = 0.f;
      if (c < OUT && s < n) {
        float e = ybuf[wave][r * OUT + c] - T[(size_t)s * OUT + c];
        l2 += e * e;
        v = 2.f * e * inv_count;
      }
      d[3][r * ROW + c] = bf16_bits(v);
      if (c < OUT) ybuf[wave][r * OUT + c] = v;  // keep fp32 dy for the bias gradient
    }
    for (int o = 32; o > 0; o >>= 1) l2 += __shfl_xor(l2, o, 64);
    if (lane == 0) lpart[wave] = l2;
  }
  wave_sync();
  if (lane < OUT) {
    float s = 0.f;
    for (int r = 0; r < 16; ++r) s += ybuf[wave][r * OUT + lane];
    dbias[wave][3][lane] = s;
  }
  // Backward through layers 3..1: da_l = dz_{l+1} W_l^T, dz_l = da_l * gelu'(z_l).
  for (int l = 3; l >= 1; --l) {
    int kch = l == 3 ? 1 : 4;  // K = output width of layer l (3 padded to 32, or 128)
    const uint16_t* src = d[l];
    uint16_t* dst = d[l - 1];
    const uint16_t* gl = g[l - 1];
    for (int nt = 0; nt < 8; ++nt) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int kc = 0; kc < kch; ++kc) acc = mfma(lds_a(src, kc, lane), gl_b(W.b[l], kc * 8 + nt, lane), acc);
      int c = nt * 16 + lane % 16;
      f4 dzv;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int r = 4 * (lane / 16) + j;
        dzv[j] = acc[j] * bf16_float(gl[r * ROW + c]);
        dst[r * ROW + c] = bf16_bits(dzv[j]);
      }
      float cs = col_sum(dzv, lane);
      if (lane < 16) dbias[wave][l - 1][c] = cs;
    }
    wave_sync();
  }
  __syncthreads();
  // Weight gradients dW_l = a_l^T dz_{l+1} over the block's 32 samples
  // (K = 32: one matrix-core step per 16 x 16 tile), tiles split over waves.
  float* P = partial + (size_t)blockIdx.x * NPARAM;
  const int mt_of[4] = {2, 8, 8, 8}, nt_of[4] = {8, 8, 8, 1};
  const int kin[4] = {IN, H, H, H}, nout[4] = {H, H, H, OUT};
  const int wofs[4] = {oW0, oW1, oW2, oW3};
  int tile = 0;
  for (int l = 0; l < 4; ++l) {
    for (int mt = 0; mt < mt_of[l]; ++mt)
      for (int nt = 0; nt < nt_of[l]; ++nt, ++tile) {
        if (tile % GW != wave) continue;
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(lds_at(act[l], mt * 16, 0, lane), lds_b(dz[l], 0, nt * 16, lane), acc);
        int c = nt * 16 + lane % 16;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          int m = mt * 16 + 4 * (lane / 16) + j;
          if (m < kin[l] && c < nout[l]) P[wofs[l] + m * nout[l] + c] = acc[j];
        }
      }
  }
  // Bias gradients: the waves' column sums in fixed order.
  const int bofs[4] = {ob0, ob1, ob2, ob3};
  for (int i = threadIdx.x; i < 3 * H + OUT; i += blockDim.x) {
    int l = i < 3 * H ? i / H : 3, c = i < 3 * H ? i % H : i - 3 * H;
    float s = 0.f;
    for (int w = 0; w < GW; ++w) s += dbias[w][l][c];
    P[bofs[l] + c] = s;
  }
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < GW; ++w) s += lpart[w];
    loss_partial[blockIdx.x] = s;
  }
}

// Sum of the block partials in block order.
__global__ void k_mlp_reduce(const float* partial, const float* loss_partial, uint32_t blocks, float* grad,
                             float* loss, int accumulate) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPARAM) {
    float s = 0.f;
    for (uint32_t b = 0; b < blocks; ++b) s += partial[(size_t)b * NPARAM + p];
    grad[p] = s;
  }
  if (p == 0) {
    float s = 0.f;
    for (uint32_t b = 0; b < blocks; ++b) s += loss_partial[b];
    *loss = accumulate ? *loss + s : s;
  }
}

// Adam (decoupled step count t >= 1) on the FP32 master parameters.
__global__ void k_mlp_adam(float* master, const float* grad, float* m1, float* m2, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float gr = grad[p];
  float a = b1 * m1[p] + (1.f - b1) * gr, b = b2 * m2[p] + (1.f - b2) * gr * gr;
  m1[p] = a;
  m2[p] = b;
  master[p] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 operand tiles from the master parameters: forward tiles of W_l
// (K x N, padded) and transposed tiles of W_1..W_3.
__global__ void k_mlp_pack(const float* master, uint16_t* fw, uint16_t* bw) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < FW_TOTAL) {
    int l = i < FW_OFS[1] ? 0 : i < FW_OFS[2] ? 1 : i < FW_OFS[3] ? 2 : 3;
    uint32_t r = i - FW_OFS[l];
    int ntiles = l == 3 ? 1 : 8, j = r % 8, lane = (r / 8) % 64, tile = r / 512;
    int kc = tile / ntiles, nt = tile % ntiles;
    int k = kc * 32 + 8 * (lane / 16) + j, nn = nt * 16 + lane % 16;
    const int kin[4] = {IN, H, H, H}, nout[4] = {H, H, H, OUT}, wofs[4] = {oW0, oW1, oW2, oW3};
    float v = (k < kin[l] && nn < nout[l]) ? master[wofs[l] + k * nout[l] + nn] : 0.f;
    fw[i] = bf16_bits(v);
  }
  if (i < BW_TOTAL) {
    int l = i < BW_OFS[2] ? 1 : i < BW_OFS[3] ? 2 : 3;
    uint32_t r = i - BW_OFS[l];
    int j = r % 8, lane = (r / 8) % 64, tile = r / 512;
    int kc = tile / 8, nt = tile % 8;  // k over the layer's outputs, n over its inputs
    int out = kc * 32 + 8 * (lane / 16) + j, in = nt * 16 + lane % 16;
    const int nout[4] = {H, H, H, OUT}, wofs[4] = {oW0, oW1, oW2, oW3};
    float v = out < nout[l] ? master[wofs[l] + in * nout[l] + out] : 0.f;
    bw[i] = bf16_bits(v);
  }
}

__global__ void k_gather_rows(const float* src, const uint32_t* perm, uint32_t n, int width, float* dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * (uint32_t)width) return;
  uint32_t r = i / width, c = i % width;
  dst[i] = src[(size_t)perm[r] * width + c];
}

inline unsigned nblocks(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ----------------------------------------------------------------- host side
void mlp_ensure(cv_gpu& g) {
  Mlp& m = g.mlp;
  if (m.master.n) return;
  m.master.reserve(NPARAM);
  m.grad.reserve(NPARAM);
  m.m1.reserve(NPARAM);
  m.m2.reserve(NPARAM);
  m.fw.reserve(FW_TOTAL);
  m.bw.reserve(BW_TOTAL);
  m.loss.reserve(1);
  CV_HIP(hipMemsetAsync(m.m1.p, 0, NPARAM * sizeof(float), g.stream));
  CV_HIP(hipMemsetAsync(m.m2.p, 0, NPARAM * sizeof(float), g.stream));
}

static Weights weights_of(cv_gpu& g) {
  Mlp& m = g.mlp;
  Weights w;
  for (int l = 0; l < 4; ++l) {
    w.f[l] = m.fw.p + FW_OFS[l];
    w.b[l] = l == 0 ? nullptr : m.bw.p + BW_OFS[l];
  }
  w.bias[0] = m.master.p + ob0;
  w.bias[1] = m.master.p + ob1;
  w.bias[2] = m.master.p + ob2;
  w.bias[3] = m.master.p + ob3;
  return w;
}

void mlp_set_params(cv_gpu& g, const float* params) {
  mlp_ensure(g);
  Mlp& m = g.mlp;
  CV_HIP(hipMemcpyAsync(m.master.p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, g.stream));
  CV_HIP(hipMemsetAsync(m.m1.p, 0, NPARAM * sizeof(float), g.stream));
  CV_HIP(hipMemsetAsync(m.m2.p, 0, NPARAM * sizeof(float), g.stream));
  m.adam_t = 0;
  k_mlp_pack<<<nblocks(FW_TOTAL > BW_TOTAL ? FW_TOTAL : BW_TOTAL, 256), 256, 0, g.stream>>>(m.master.p, m.fw.p,
                                                                                           m.bw.p);
  CV_HIP(hipGetLastError());
  m.ready = true;
}

void mlp_get_params(cv_gpu& g, float* params) {
  if (!g.mlp.ready) fail(CV_ERR_NOT_BUILT, "warm-start network not set up");
  CV_HIP(hipMemcpyAsync(params, g.mlp.master.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, g.stream));
  CV_HIP(hipStreamSynchronize(g.stream));
}

void mlp_infer_device(cv_gpu& g, const float* dX, uint32_t n, float* dY) {
  if (!g.mlp.ready) fail(CV_ERR_NOT_BUILT, "warm-start network not set up");
  if (n == 0) return;
  unsigned tiles = (n + 15) / 16;
  // One block per CU fits in LDS; enough blocks to cover every CU.
  unsigned blocks = std::min<unsigned>((tiles + IW - 1) / IW, 1024);
  k_mlp_infer<<<blocks, 64 * IW, 0, g.stream>>>(weights_of(g), dX, n, dY);
  CV_HIP(hipGetLastError());
}

// Gradient of the batch rows [row0, row0 + count) into m.grad; returns nothing
// (loss in m.loss).  count is a multiple of 32 or the tail.
void mlp_grad_device(cv_gpu& g, const float* dX, const float* dT, uint32_t n, uint32_t row0, uint32_t count,
                     bool accumulate_loss) {
  Mlp& m = g.mlp;
  unsigned blocks = nblocks(count, 16 * GW);
  m.partial.reserve((size_t)blocks * NPARAM);
  m.loss_partial.reserve(blocks);
  uint32_t end = std::min<uint32_t>(n, row0 + count);
  if (end <= row0) fail(CV_ERR_INVALID_ARGUMENT, "empty training batch");
  k_mlp_grad<<<blocks, 64 * GW, 0, g.stream>>>(weights_of(g), dX, dT, end, row0,
                                               1.0f / (3.0f * (float)(end - row0)), m.partial.p,
                                               m.loss_partial.p);
  CV_HIP(hipGetLastError());
  k_mlp_reduce<<<nblocks(NPARAM, 256), 256, 0, g.stream>>>(m.partial.p, m.loss_partial.p, blocks, m.grad.p,
                                                           m.loss.p, accumulate_loss ? 1 : 0);
  CV_HIP(hipGetLastError());
}

void mlp_adam_device(cv_gpu& g, float lr) {
  Mlp& m = g.mlp;
  m.adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  float c1 = 1.f - std::pow(b1, (float)m.adam_t), c2 = 1.f - std::pow(b2, (float)m.adam_t);
  k_mlp_adam<<<nblocks(NPARAM, 256), 256, 0, g.stream>>>(m.master.p, m.grad.p, m.m1.p, m.m2.p, lr, b1, b2, 1e-8f,
                                                         c1, c2);
  k_mlp_pack<<<nblocks(FW_TOTAL > BW_TOTAL ? FW_TOTAL : BW_TOTAL, 256), 256, 0, g.stream>>>(m.master.p, m.fw.p,
                                                                                           m.bw.p);
  CV_HIP(hipGetLastError());
}

void mlp_gather(cv_gpu& g, const float* src, const uint32_t* perm, uint32_t n, int width, float* dst) {
  if (n == 0) return;
  k_gather_rows<<<nblocks((uint64_t)n * width, 256), 256, 0, g.stream>>>(src, perm, n, width, dst);
  CV_HIP(hipGetLastError());
}

// ------------------------------------------------ allocation-free inference
// Staging of cv_warm_infer_noalloc: the features [12 r] and predictions
// [3 r] of up to CV_WARM_NOALLOC_MAX_ROWS rows, in one file-scope device
// array that every GPU the library runs on holds (15 MiB) until the process
// ends.  In host code `scratch` is a host-side stand-in, never a device
// address: host code uses only the per-device address warm_scratch_setup
// takes with hipGetSymbolAddress.
constexpr size_t kNoAllocRows = CV_WARM_NOALLOC_MAX_ROWS;
static __device__ __align__(256) float scratch[(size_t)(12 + 3) * kNoAllocRows];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
float* scratch_ptr[kMaxDevices] = {};
}  // namespace

void warm_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) fail(CV_ERR_DEVICE, "device index beyond the library's table of 64");
  std::lock_guard<std::mutex> l(scratch_mu);
  if (scratch_ptr[device]) return;
  void* p = nullptr;
  CV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  scratch_ptr[device] = static_cast<float*>(p);
}

static float* warm_scratch(int device) {
  std::lock_guard<std::mutex> l(scratch_mu);
  return device >= 0 && device < kMaxDevices ? scratch_ptr[device] : nullptr;
}

extern "C" cv_status cv_warm_infer_noalloc(cv_gpu* g, const float* features, uint64_t n, float* out) {
  if (!g || (n && (!features || !out))) return CV_ERR_INVALID_ARGUMENT;
  return guard(&g->last_error, [&] {
    if (n > kNoAllocRows)
      fail(CV_ERR_CAPACITY, "cv_warm_infer_noalloc: " + std::to_string(n) + " rows exceed the limit of " +
                                std::to_string(kNoAllocRows) + " (use cv_warm_infer)");
    CV_HIP(hipSetDevice(g->device));
    if (!g->mlp.ready) fail(CV_ERR_NOT_BUILT, "warm-start network not set up");
    if (!n) return;
    float* X = warm_scratch(g->device);
    if (!X) fail(CV_ERR_DEVICE, "staging array of the device not set up");
    float* Y = X + 12 * kNoAllocRows;
    // Every call writes the rows it reads: nothing needs clearing.
    CV_HIP(hipMemcpyAsync(X, features, 12 * n * sizeof(float), hipMemcpyHostToDevice, g->stream));
    mlp_infer_device(*g, X, (uint32_t)n, Y);
    CV_HIP(hipMemcpyAsync(out, Y, 3 * n * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    CV_HIP(hipStreamSynchronize(g->stream));
  });
}
