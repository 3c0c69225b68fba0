// Spectral neural operator kernels for MI355X (gfx950): a batched strided
// GEMM on bf16 MFMA with fp32 accumulation, fused MFMA kernels for the
// operator's memory-bound stages (truncated x transform, spectral-layer
// tails, projection head, weight and bias gradients), and the elementwise
// kernels (bias + GELU, spectral repacks, mode-mixing weights, masked MSE,
// AdamW). The C ABI is declared in include/swno.h.
//
// The operator itself has no counterpart in the reference's code (its design
// document, TFNO_Full_Ring_Plan.tex:200-209, describes a Fourier neural
// operator); everything here is this project's own design.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>

#include "swno.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define SWNO_HIP_CHECK(expr)                                                          \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) return fail(SWNO_EHIP, std::string(#expr) + ": " +      \
                                                         hipGetErrorString(e_));      \
    } while (0)

hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;

struct GemmP {
    int M, N, K, batch, k_split, k_chunk;
    const float *A;
    int64_t a_rs, a_cs, a_bs;
    const float *B;
    int64_t b_rs, b_cs, b_bs;
    float *C;
    int64_t c_rs, c_cs, c_bs;
    float alpha, beta;
    float *ws;
};

constexpr int kTK = 32;     // K per MFMA stage (v_mfma_f32_16x16x32_bf16)
constexpr int kPadK = 8;    // LDS row padding (bf16 elements)

// One workgroup = 4 wavefronts in a 2 x 2 arrangement computing a TM x TN
// tile of one batch entry (and one K slice). Tiles of A (TM x 32) and B
// (32 x TN, stored transposed so that a lane's 8 consecutive K values are
// contiguous) are rounded to bf16 into LDS; each wavefront then issues
// (TM/32) x (TN/32) MFMA 16x16x32 per stage. Lane l of an MFMA holds A row
// l % 16 and B column l % 16 with K values 8 (l / 16) .. 8 (l / 16) + 7, and
// receives C rows 4 (l / 16) .. 4 (l / 16) + 3 of column l % 16.
template <int TM, int TN>
__global__ void __launch_bounds__(256) gemm_kernel(GemmP p)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[TM][kTK + kPadK];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[TN][kTK + kPadK];
    constexpr int WM = TM / 2, WN = TN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int z = blockIdx.z;
    const int b = z / p.k_split;
    const int ks = z - b * p.k_split;
    const int m0 = blockIdx.x * TM;
    const int n0 = blockIdx.y * TN;
    const int k_begin = ks * p.k_chunk;
    const int k_end = min(p.K, k_begin + p.k_chunk);
    const float *Ab = p.A + (int64_t)b * p.a_bs;
    const float *Bb = p.B + (int64_t)b * p.b_bs;
    f4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const bool a_kfast = p.a_cs == 1;
    const bool b_nfast = p.b_cs == 1;
    for (int k0 = k_begin; k0 < k_end; k0 += kTK) {
        for (int i = tid; i < TM * kTK; i += 256) {
            int m, k;
            if (a_kfast) {
                m = i / kTK;
                k = i - m * kTK;
            } else {
                k = i / TM;
                m = i - k * TM;
            }
            const int gm = m0 + m, gk = k0 + k;
            float v = 0.0f;
            if (gm < p.M && gk < k_end) v = Ab[(int64_t)gm * p.a_rs + (int64_t)gk * p.a_cs];
            As[m][k] = (__bf16)v;
        }
        for (int i = tid; i < kTK * TN; i += 256) {
            int k, n;
            if (b_nfast) {
                k = i / TN;
                n = i - k * TN;
            } else {
                n = i / kTK;
                k = i - n * kTK;
            }
            const int gn = n0 + n, gk = k0 + k;
            float v = 0.0f;
            if (gn < p.N && gk < k_end) v = Bb[(int64_t)gk * p.b_rs + (int64_t)gn * p.b_cs];
            Bs[n][k] = (__bf16)v;
        }
        __syncthreads();
        const int r = lane & 15;
        const int kq = (lane >> 4) * 8;
        bf16x8 af[FM], bfr[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = *(const bf16x8 *)&As[wm * WM + i * 16 + r][kq];
#pragma unroll
        for (int j = 0; j < FN; ++j) bfr[j] = *(const bf16x8 *)&Bs[wn * WN + j * 16 + r][kq];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0,
                                                                    0, 0);
        __syncthreads();
    }
    const int col_l = lane & 15;
    const int row_l = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = m0 + wm * WM + i * 16 + row_l + e;
                const int col = n0 + wn * WN + j * 16 + col_l;
                if (row < p.M && col < p.N) {
                    const float v = acc[i][j][e];
                    if (p.k_split == 1) {
                        float *c = p.C + (int64_t)b * p.c_bs + (int64_t)row * p.c_rs +
                                   (int64_t)col * p.c_cs;
                        *c = p.beta != 0.0f ? p.alpha * v + p.beta * *c : p.alpha * v;
                    } else {
                        p.ws[(((int64_t)b * p.k_split + ks) * p.M + row) * p.N + col] = v;
                    }
                }
            }
}

// C = alpha * sum over K slices (in slice order) + beta * C.
__global__ void __launch_bounds__(256) gemm_reduce_kernel(GemmP p)
{
    const int64_t total = (int64_t)p.batch * p.M * p.N;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / ((int64_t)p.M * p.N);
        const int64_t rem = t - b * p.M * p.N;
        const int row = (int)(rem / p.N);
        const int col = (int)(rem - (int64_t)row * p.N);
        float s = 0.0f;
        for (int ks = 0; ks < p.k_split; ++ks)
            s += p.ws[((b * p.k_split + ks) * p.M + row) * p.N + col];
        float *c = p.C + b * p.c_bs + (int64_t)row * p.c_rs + (int64_t)col * p.c_cs;
        *c = p.beta != 0.0f ? p.alpha * s + p.beta * *c : p.alpha * s;
    }
}

// Weight and bias gradients of a pointwise linear layer over many rows:
// dW (M x N) = A^T B and db (N) = column sums of B, for A (rows x M) and
// B (rows x N) in row-major layout. Workgroup p takes rows
// [p rows_per_part, (p + 1) rows_per_part) in slabs of kWgRows rows: the slab
// is loaded with float4 (or scalar) loads, its A^T and B tiles rounded to
// bf16 into LDS with K (the row) contiguous, and each wavefront accumulates
// its 16 x 16 output tiles with MFMA 16x16x32 over the slab. B's column sums
// are kept in fp32 per thread (a thread's columns do not change from slab to
// slab) and combined across threads in thread order. Each workgroup writes
// one partial (M N values, then N) to the workspace; wgrad_reduce_kernel sums
// the partials in order. The next slab's loads are issued before the current
// slab's MFMA so that they are in flight meanwhile.
constexpr int kWgRows = 64;      // rows per slab (two MFMA K steps)
constexpr int kWgMaxParts = 1024;

struct WgradP {
    int64_t rows, rows_per_part;
    int M, N;
    const void *A;  // float, or bf16 codes (a_bf16)
    int64_t lda;
    const float *B;
    int64_t ldb;
    int bias;          // write column sums of B
    float *ws;         // [part][M N + N]
};

// One operand's share of a slab: item i of thread t is a float4 (V = 1), a
// scalar (V = 0) or eight bf16 codes (V = 2) number e = t + 256 i of the slab
// in row-major order. Its row and its offset in the operand do not change
// from slab to slab.
template <int CP, int V>
struct WgOperand {
    static constexpr int kVec = V == 2 ? 8 : (V ? 4 : 1);  // values per item
    static constexpr int kItems = CP * kWgRows / 256 / kVec;
    typedef typename std::conditional<V == 2, __bf16, float>::type elem_t;
    typedef typename std::conditional<V == 2, bf16x8,
                                      typename std::conditional<V == 1, f4, float>::type>::type
        item_t;
    item_t reg[kItems];
    int row[kItems];  // slab row, or kWgRows when the item is past the slab
    int64_t off[kItems];
    int col[kItems];

    __device__ void init(int cols, int64_t ld, int tid)
    {
        const int per_row = cols / kVec;
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
            const int e = tid + 256 * i;
            const int r = e / per_row, c = (e - r * per_row) * kVec;
            row[i] = e < kWgRows * per_row ? r : kWgRows;
            col[i] = c;
            off[i] = (int64_t)r * ld + c;
        }
    }
    __device__ void load(const void *Xv, int64_t ld, int64_t r0, int64_t r_end)
    {
        const elem_t *base = (const elem_t *)Xv + r0 * ld;
        // rows of this slab inside the range (items past the slab have row == kWgRows)
        const int64_t lim = min(r_end - r0, (int64_t)kWgRows);
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
            item_t v = item_t{};
            if (row[i] < lim) v = *(const item_t *)(base + off[i]);
            reg[i] = v;
        }
    }
    // LDS tile T[col][row] in bf16; rows past the range hold zeros
    __device__ void store(__bf16 (*T)[kWgRows + kPadK]) const
    {
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
            if (row[i] < kWgRows) {
                if constexpr (V) {
#pragma unroll
                    for (int j = 0; j < kVec; ++j) T[col[i] + j][row[i]] = (__bf16)reg[i][j];
                } else {
                    T[col[i]][row[i]] = (__bf16)reg[i];
                }
            }
        }
    }
};

template <int MP, int NP, int VA, int VB>
__global__ void __launch_bounds__(256) wgrad_kernel(WgradP p)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[MP][kWgRows + kPadK];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NP][kWgRows + kPadK];
    __shared__ float bpart[256][4];
    constexpr int TM = MP / 16, TN = NP / 16, T = TM * TN;
    constexpr int TPW = (T + 3) / 4;  // tiles per wavefront
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // padded rows / columns must hold finite values (their outputs are dropped)
    for (int i = tid; i < MP * (kWgRows + kPadK); i += 256) (&As[0][0])[i] = (__bf16)0.0f;
    for (int i = tid; i < NP * (kWgRows + kPadK); i += 256) (&Bs[0][0])[i] = (__bf16)0.0f;
    const int64_t r_begin = (int64_t)blockIdx.x * p.rows_per_part;
    const int64_t r_end = min(p.rows, r_begin + p.rows_per_part);
    f4 acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) acc[i] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    WgOperand<MP, VA> a;
    WgOperand<NP, VB> b;
    a.init(p.M, p.lda, tid);
    b.init(p.N, p.ldb, tid);
    typename WgOperand<NP, VB>::item_t bsum{};
    a.load(p.A, p.lda, r_begin, r_end);
    b.load(p.B, p.ldb, r_begin, r_end);
    const int r = lane & 15, kq = (lane >> 4) * 8;
    for (int64_t r0 = r_begin; r0 < r_end; r0 += kWgRows) {
        __syncthreads();
        a.store(As);
        b.store(Bs);
#pragma unroll
        for (int i = 0; i < WgOperand<NP, VB>::kItems; ++i) bsum += b.reg[i];  // fixed columns
        __syncthreads();
        if (r0 + kWgRows < r_end) {
            a.load(p.A, p.lda, r0 + kWgRows, r_end);
            b.load(p.B, p.ldb, r0 + kWgRows, r_end);
        }
#pragma unroll
        for (int ks = 0; ks < kWgRows; ks += kTK) {
#pragma unroll
            for (int i = 0; i < TPW; ++i) {
                const int t = wave + 4 * i;
                if (t < T) {
                    const int tm = t / TN, tn = t - tm * TN;
                    const bf16x8 af = *(const bf16x8 *)&As[tm * 16 + r][ks + kq];
                    const bf16x8 bf = *(const bf16x8 *)&Bs[tn * 16 + r][ks + kq];
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf, acc[i], 0, 0, 0);
                }
            }
        }
    }
    const int MN = p.M * p.N;
    const int J = MN + (p.bias ? p.N : 0);
    float *out = p.ws + (int64_t)blockIdx.x * J;
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
        const int t = wave + 4 * i;
        if (t < T) {
            const int tm = t / TN, tn = t - tm * TN;
            const int col = tn * 16 + (lane & 15);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = tm * 16 + (lane >> 4) * 4 + e;
                if (row < p.M && col < p.N) out[row * p.N + col] = acc[i][e];
            }
        }
    }
    if (p.bias) {
        // thread t holds the sums of columns c(t) .. c(t) + 3 (float4) or c(t)
        if constexpr (VB) {
#pragma unroll
            for (int j = 0; j < 4; ++j) bpart[tid][j] = bsum[j];
        } else {
            bpart[tid][0] = bsum;
        }
        __syncthreads();
        const int groups = VB ? p.N >> 2 : p.N;  // distinct column groups
        for (int c = tid; c < p.N; c += 256) {
            const int g = VB ? c >> 2 : c, j = VB ? c & 3 : 0;
            float sum = 0.0f;
            for (int u = g; u < 256; u += groups) sum += bpart[u][j];
            out[MN + c] = sum;
        }
    }
}

// out[j] = sum over parts, in part order within 16 contiguous ranges that are
// then added in range order; j < M N goes to dW, the rest to db.
__global__ void __launch_bounds__(256) wgrad_reduce_kernel(int n_part, int MN, int J,
                                                           const float *ws, float *dW, float *db)
{
    __shared__ float part[16][16];
    const int tid = threadIdx.x;
    const int jj = tid & 15, q = tid >> 4;
    const int j = blockIdx.x * 16 + jj;
    const int per = (n_part + 15) / 16;
    const int p0 = q * per, p1 = min(n_part, p0 + per);
    float s = 0.0f;
    if (j < J)
        for (int pp = p0; pp < p1; ++pp) s += ws[(int64_t)pp * J + j];
    part[q][jj] = s;
    __syncthreads();
    if (q == 0 && j < J) {
        float t = 0.0f;
#pragma unroll
        for (int u = 0; u < 16; ++u) t += part[u][jj];
        if (j < MN)
            dW[j] = t;
        else
            db[j - MN] = t;
    }
}

__device__ __forceinline__ float gelu(float x)
{
    return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// Tail of a spectral layer, fused, for kLayerC = 32 channels. A slice is one
// row y of one sample b (n pixels x 32 channels, channel-last). Forward:
//   pre = Gx Z + v W + bias, vnext = GELU(pre) (or pre for the last layer),
// with Gx (n x 2m) the inverse x transform, Z the slice's 2m x 32 spectrum,
// v the slice's input and W the layer's pointwise weights. Backward (for the
// gradient dpre of a layer's pre-activation):
//   dv = dpre W^T + Fx^T dT1, out = dv * GELU'(pre_prev) (or dv),
// with Fx (2m x n) the forward x transform and dT1 the slice's 2m x 32
// gradient of the x transform's output; pre_prev is the previous layer's
// pre-activation, so `out` is that layer's dpre. Both are two MFMA 16x16x32
// products summed in one accumulator (2m <= 32 and 32 channels: one MFMA per
// product and 16 x 16 tile), then the elementwise epilogue, so the slice is
// read once and written once. Each wavefront takes row tiles w, w + 4, ... of
// the slice (n <= 128: at most two) and both 16-column tiles; the shared
// matrices are held in registers as bf16 operands for the kernel's lifetime.
constexpr int kLayerC = 32;
constexpr int kLayerMaxN = 128;

__device__ __forceinline__ float gelu(float x);
__device__ __forceinline__ float gelu_grad(float x);

// 8 consecutive floats (two float4) at p, or zeros past `valid` of them
__device__ __forceinline__ bf16x8 bf16_row8(const float *p, int valid)
{
    bf16x8 r;
    if (valid >= 8) {
        const f4 x0 = *(const f4 *)p;
        const f4 x1 = *(const f4 *)(p + 4);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            r[j] = (__bf16)x0[j];
            r[4 + j] = (__bf16)x1[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] = (__bf16)(j < valid ? p[j] : 0.0f);
    }
    return r;
}

// 8 consecutive activations stored as bf16 (the rounding every MFMA consumer
// applies to an fp32 activation, so storing them rounded changes no result),
// or zeros past `valid` of them; p 16-byte aligned when valid >= 8
__device__ __forceinline__ bf16x8 bf16_row8(const __bf16 *p, int valid)
{
    if (valid >= 8) return *(const bf16x8 *)p;
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = j < valid ? p[j] : (__bf16)0.0f;
    return r;
}

// 8 floats p[0], p[stride], ..., or zeros from index `valid` on
__device__ __forceinline__ bf16x8 bf16_col8(const float *p, int64_t stride, int valid)
{
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = (__bf16)(j < valid ? p[j * stride] : 0.0f);
    return r;
}

// Truncated transform along x, per slice: out[s] (M x 32) = A (M x n) X[s]
// (n x 32), for M = 2 modes <= 32 and n <= 128 (A = Fx forward, or Gx^T in
// the backward pass, given by its row and column strides). A is held in
// registers as bf16 MFMA operands for the kernel's lifetime. A slice's X is
// loaded with coalesced float4 loads (the next slice's while the current one
// is multiplied), rounded to bf16 into LDS transposed (channel-major, the row
// index contiguous) in one of two buffers, so one barrier per slice suffices;
// wavefront w then computes the output tile (w / 2, w % 2).
// X is fp32 or bf16 (T); a thread loads 16 bytes per item (4 or 8 values).
template <typename T>
__global__ void __launch_bounds__(256) x_transform_kernel(int n_slices, int M, int n,
                                                          const float *A, int64_t a_rs,
                                                          int64_t a_cs, const T *X,
                                                          float *out)
{
    constexpr bool kF32 = std::is_same<T, float>::value;
    constexpr int kVec = kF32 ? 4 : 8;  // values per 16-byte item
    constexpr int kXtItems = kLayerMaxN * kLayerC / kVec / 256;  // items per thread and slice
    typedef typename std::conditional<kF32, f4, bf16x8>::type item_t;
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][kLayerC][kLayerMaxN + 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int tm = wave >> 1, tn = wave & 1;
    const bool active = 16 * tm < M;  // M <= 16: two wavefronts only load and wait
    constexpr int kSteps = kLayerMaxN / 32;
    const int steps = (n + 31) / 32;
    const int am = 16 * tm + r16;
    bf16x8 a[kSteps];
#pragma unroll
    for (int kk = 0; kk < kSteps; ++kk)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 32 * kk + 8 * g + j;
            a[kk][j] = (__bf16)(am < M && k < n ? A[am * a_rs + k * a_cs] : 0.0f);
        }
    // rows k >= n of both buffers stay zero (they pad the last 32-row step)
    for (int i = tid; i < 2 * kLayerC * (kLayerMaxN + 8); i += 256) (&Bs[0][0][0])[i] = (__bf16)0.0f;
    const int count = n * (kLayerC / kVec);
    item_t rx[kXtItems];
    auto load = [&](int sl) {
        const T *xs = X + (int64_t)sl * n * kLayerC;
#pragma unroll
        for (int i = 0; i < kXtItems; ++i) {
            const int e4 = tid + 256 * i;
            rx[i] = e4 < count ? *(const item_t *)(xs + kVec * e4) : item_t{};
        }
    };
    int sl = blockIdx.x;
    if (sl < n_slices) load(sl);
    __syncthreads();
    const int col = 16 * tn + r16;
    for (int buf = 0; sl < n_slices; sl += gridDim.x, buf ^= 1) {
#pragma unroll
        for (int i = 0; i < kXtItems; ++i) {
            const int e4 = tid + 256 * i;
            if (e4 < count) {
                const int k = e4 / (kLayerC / kVec), c4 = (e4 % (kLayerC / kVec)) * kVec;
#pragma unroll
                for (int j = 0; j < kVec; ++j) Bs[buf][c4 + j][k] = (__bf16)rx[i][j];
            }
        }
        __syncthreads();
        if (sl + (int)gridDim.x < n_slices) load(sl + gridDim.x);
        if (active) {
            f4 acc = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kk = 0; kk < kSteps; ++kk)
                if (kk < steps) {
                    const bf16x8 b = *(const bf16x8 *)&Bs[buf][col][32 * kk + 8 * g];
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kk], b, acc, 0, 0, 0);
                }
            float *os = out + (int64_t)sl * M * kLayerC;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = 16 * tm + 4 * g + e;
                if (row < M) os[row * kLayerC + col] = acc[e];
            }
        }
    }
}

// Forward tail of a layer: each wavefront's outputs (its 32 rows x 32
// channels of pre and vnext) are staged in LDS and written with float4
// stores, consecutive lanes at consecutive 16 bytes (measured faster than
// stores from the MFMA output layout, which leaves 64-byte pieces). The
// layer input v (TV) and output vnext (TN) are fp32 or bf16; pre is fp32.
template <typename TV, typename TN>
__global__ void __launch_bounds__(256) layer_forward_kernel(
    int n_slices, int n, int k1, const float *Gx, const float *Z, const TV *v,
    const float *W, const float *bias, float *pre, TN *vnext, int act)
{
    constexpr int kStride = kLayerC + 4;  // floats per staged row (16-byte aligned rows)
    __shared__ __attribute__((aligned(16))) float stage[4][2][32][kStride];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int row_tiles = n >> 4;
    const int kv = max(0, min(8, k1 - 8 * g));
    bf16x8 ga[2], wb[2];
    float bc[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int rt = wave + 4 * i;
        if (rt < row_tiles) ga[i] = bf16_row8(Gx + (int64_t)(16 * rt + r16) * k1 + 8 * g, kv);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        wb[j] = bf16_col8(W + (8 * g) * kLayerC + 16 * j + r16, kLayerC, 8);
        bc[j] = bias[16 * j + r16];
    }
    float (*sp)[kStride] = stage[wave][0];
    float (*sv)[kStride] = stage[wave][1];
    for (int sl = blockIdx.x; sl < n_slices; sl += gridDim.x) {
        const float *zs = Z + (int64_t)sl * k1 * kLayerC;
        const TV *vs = v + (int64_t)sl * n * kLayerC;
        bf16x8 zb[2], va[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) zb[j] = bf16_col8(zs + (8 * g) * kLayerC + 16 * j + r16,
                                                      kLayerC, kv);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int rt = wave + 4 * i;
            if (rt < row_tiles) va[i] = bf16_row8(vs + (16 * rt + r16) * kLayerC + 8 * g, 8);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int rt = wave + 4 * i;
            if (rt >= row_tiles) break;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                f4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    ga[i], zb[j], f4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va[i], wb[j], acc, 0, 0, 0);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int rl = 16 * i + 4 * g + e;  // row within the wavefront's stage
                    const float pv = acc[e] + bc[j];
                    sp[rl][16 * j + r16] = pv;
                    sv[rl][16 * j + r16] = act ? gelu(pv) : pv;
                }
            }
        }
        // the wavefront reads back only what it wrote (LDS is in order per wavefront)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float *ps = pre + (int64_t)sl * n * kLayerC;
        TN *ns = vnext + (int64_t)sl * n * kLayerC;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int rt = wave + 4 * i;
            if (rt >= row_tiles) break;
            // 16 rows x 32 floats = 128 float4, two per lane
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = lane + 64 * h;
                const int rr = q >> 3, c4 = (q & 7) * 4;
                const int rl = 16 * i + rr;
                const int64_t o = (int64_t)(16 * rt + rr) * kLayerC + c4;
                *(f4 *)(ps + o) = *(const f4 *)&sp[rl][c4];
                if constexpr (std::is_same<TN, float>::value)
                    *(f4 *)(ns + o) = *(const f4 *)&sv[rl][c4];
            }
            if constexpr (!std::is_same<TN, float>::value) {
                // 16 rows x 32 values in bf16 = 64 pieces of 16 bytes, one per lane
                const int rr = lane >> 2, c8 = (lane & 3) * 8;
                const int rl = 16 * i + rr;
                const f4 x0 = *(const f4 *)&sv[rl][c8], x1 = *(const f4 *)&sv[rl][c8 + 4];
                bf16x8 b;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    b[j] = (__bf16)x0[j];
                    b[4 + j] = (__bf16)x1[j];
                }
                *(bf16x8 *)(ns + (int64_t)(16 * rt + rr) * kLayerC + c8) = b;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__global__ void __launch_bounds__(256) layer_backward_kernel(int n_slices, int n, int k1,
                                                             const float *dpre, const float *W,
                                                             const float *Fx, const float *dT1,
                                                             const float *pre_prev, float *out)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int row_tiles = n >> 4;
    const int kv = max(0, min(8, k1 - 8 * g));
    bf16x8 fa[2], wtb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int rt = wave + 4 * i;
        // A[x][k] = Fx[k][x]
        if (rt < row_tiles) fa[i] = bf16_col8(Fx + (int64_t)(8 * g) * n + 16 * rt + r16, n, kv);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)  // B[k][c] = W[c][k]
        wtb[j] = bf16_row8(W + (16 * j + r16) * kLayerC + 8 * g, 8);
    // a slice's operands, loaded before any store of the slice before it (out may
    // not alias the inputs, but the compiler cannot know that): the next slice's are
    // in flight while the current one is computed
    bf16x8 tb[2], da[2], tn[2], dn[2];
    float pv[2][2][4], pn[2][2][4];
    auto load = [&](int sl, bf16x8 (&t)[2], bf16x8 (&d)[2], float (&q)[2][2][4]) {
        const float *ts = dT1 + (int64_t)sl * k1 * kLayerC;
        const float *ds = dpre + (int64_t)sl * n * kLayerC;
        const float *pp = pre_prev != nullptr ? pre_prev + (int64_t)sl * n * kLayerC : nullptr;
#pragma unroll
        for (int j = 0; j < 2; ++j) t[j] = bf16_col8(ts + (8 * g) * kLayerC + 16 * j + r16,
                                                     kLayerC, kv);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int rt = wave + 4 * i;
            if (rt < row_tiles) {
                d[i] = bf16_row8(ds + (16 * rt + r16) * kLayerC + 8 * g, 8);
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        q[i][j][e] =
                            pp != nullptr ? pp[(16 * rt + 4 * g + e) * kLayerC + 16 * j + r16]
                                          : 0.0f;
            }
        }
    };
    if (blockIdx.x < n_slices) load(blockIdx.x, tb, da, pv);
    for (int sl = blockIdx.x; sl < n_slices; sl += gridDim.x) {
        if (sl + (int)gridDim.x < n_slices) load(sl + gridDim.x, tn, dn, pn);
        const bool gp = pre_prev != nullptr;
        float *os = out + (int64_t)sl * n * kLayerC;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int rt = wave + 4 * i;
            if (rt >= row_tiles) break;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                f4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    da[i], wtb[j], f4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], tb[j], acc, 0, 0, 0);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int idx = (16 * rt + 4 * g + e) * kLayerC + 16 * j + r16;
                    os[idx] = gp ? acc[e] * gelu_grad(pv[i][j][e]) : acc[e];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            tb[i] = tn[i];
            da[i] = dn[i];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) pv[i][j][e] = pn[i][j][e];
        }
    }
}

// Projection head of the operator, fused: for each pixel row,
//   hpre = v Q1 + bq1 (H values), out = GELU(hpre) . Q2 + bq2,
// with v (rows x kHeadC) and the H = kHeadH hidden values never leaving the
// workgroup except as hpre (kept for the backward pass). Each wavefront takes
// 16 rows: one MFMA 16x16x32 per 16 hidden columns (K = the 32 channels, Q1
// held in registers as bf16 B operands), then bias, GELU and the dot with Q2
// in fp32 in the MFMA's output layout, and a fixed-order reduction over the
// 16 lanes that hold a row's columns; hpre goes out through LDS as float4.
constexpr int kHeadC = 32, kHeadH = 128;
constexpr int kHeadTiles = kHeadH / 16;

template <typename T>  // v: fp32 or bf16
__global__ void __launch_bounds__(256) head_forward_kernel(int64_t rows, const T *v,
                                                           const float *Q1, const float *bq1,
                                                           const float *Q2, const float *bq2,
                                                           float *hpre, float *out)
{
    // per wavefront: its 16 rows of hpre, staged for float4 stores
    __shared__ __attribute__((aligned(16))) float stage[4][16][kHeadH + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    float (*st)[kHeadH + 4] = stage[wave];
    bf16x8 qb[kHeadTiles];
    float bias[kHeadTiles], q2[kHeadTiles];
#pragma unroll
    for (int t = 0; t < kHeadTiles; ++t) {
#pragma unroll
        for (int j = 0; j < 8; ++j) qb[t][j] = (__bf16)Q1[(8 * g + j) * kHeadH + 16 * t + r16];
        bias[t] = bq1[16 * t + r16];
        q2[t] = Q2[16 * t + r16];
    }
    const float b2 = bq2[0];
    const int64_t n_blocks = (rows + 63) / 64;
    for (int64_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
        const int64_t row0 = blk * 64 + wave * 16;
        const int64_t ra = row0 + r16;
        const bf16x8 a = bf16_row8(v + (ra < rows ? ra : 0) * kHeadC + 8 * g, ra < rows ? 8 : 0);
        float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < kHeadTiles; ++t) {
            const f4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                a, qb[t], f4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float hp = acc[e] + bias[t];
                st[4 * g + e][16 * t + r16] = hp;
                part[e] += gelu(hp) * q2[t];
            }
        }
        // hpre of the wavefront's 16 rows from LDS with float4 stores (the wavefront
        // reads back only what it wrote; LDS is in order per wavefront)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int h = 0; h < kHeadH * 16 / 4 / 64; ++h) {
            const int q = lane + 64 * h;
            const int rr = q / (kHeadH / 4), c4 = (q % (kHeadH / 4)) * 4;
            const int64_t row = row0 + rr;
            if (row < rows) *(f4 *)(hpre + row * kHeadH + c4) = *(const f4 *)&st[rr][c4];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int e = 0; e < 4; ++e) {
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) part[e] += __shfl_xor(part[e], m);
            const int64_t row = row0 + 4 * g + e;
            if (r16 == 0 && row < rows) out[row] = part[e] + b2;
        }
    }
}

// Backward pass of the projection head for d loss / d out = dout: with
// h = GELU(hpre) recomputed and dh = dout Q2^T * GELU'(hpre),
//   dv = dh Q1^T (rows x kHeadC), and per workgroup the partial sums of
//   dQ1 = v^T dh, dbq1 = sum dh, dQ2 = h^T dout, dbq2 = sum dout
// over its rows, written as one partial [dQ1 | dbq1 | dQ2 | dbq2] to the
// workspace (head_reduce_kernel sums the partials in order). A 64-row block
// stages dh in LDS as bf16 twice (row-major for dv, transposed for dQ1) and
// v^T once; dv and dQ1 are MFMA products, the rest fp32.
constexpr int kHeadJ = kHeadC * kHeadH + 2 * kHeadH + 1;

// One 64-row block's inputs of head_backward_kernel, as one wavefront's lane
// holds them: hpre at rows 16 w + 4 g + e and columns 16 t + r16, dout at
// rows 16 w + 4 g + e, and two float4 of the block's v (for v^T in LDS).
template <typename T>  // v: fp32 (two float4 per thread) or bf16 (eight values)
struct HeadBlock {
    float hp[kHeadTiles][4];
    float d[4];
    typename std::conditional<std::is_same<T, float>::value, f4[2], bf16x8[1]>::type vx;
};

template <typename T>
__device__ __forceinline__ void head_load(HeadBlock<T> &b, int64_t blk0, int64_t r_end,
                                          const float *dout, const float *hpre, const T *v,
                                          int tid)
{
    const int lane = tid & 63, wave = tid >> 6, r16 = lane & 15, g = lane >> 4;
    const int64_t row0 = blk0 + 16 * wave + 4 * g;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const bool in = row0 + e < r_end;
        b.d[e] = in ? dout[row0 + e] : 0.0f;
#pragma unroll
        for (int t = 0; t < kHeadTiles; ++t)
            b.hp[t][e] = in ? hpre[(row0 + e) * kHeadH + 16 * t + r16] : 0.0f;
    }
    if constexpr (std::is_same<T, float>::value) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int e4 = tid + 256 * i;
            const int rr = e4 >> 3, c4 = (e4 & 7) * 4;
            b.vx[i] = blk0 + rr < r_end ? *(const f4 *)(v + (blk0 + rr) * kHeadC + c4)
                                        : f4{0.0f, 0.0f, 0.0f, 0.0f};
        }
    } else {
        const int rr = tid >> 2, c8 = (tid & 3) * 8;
        b.vx[0] = blk0 + rr < r_end ? *(const bf16x8 *)(v + (blk0 + rr) * kHeadC + c8)
                                    : bf16x8{};
    }
}

// two workgroups per CU: at most 256 VGPRs, LDS below half of the CU's
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
head_backward_kernel(int64_t rows, int64_t rows_per_part, const float *dout, const float *hpre,
                     const T *v, const float *Q1, const float *Q2, float *dv, float *ws)
{
    constexpr int kDH = 64 * (kHeadH + 8) * 2, kDHT = kHeadH * (64 + 8) * 2;
    constexpr int kVT = kHeadC * (64 + 8) * 2, kQS = kHeadC * (kHeadH + 8) * 2;
    static_assert(16 * (2 * kHeadH + 1) * 4 <= kDH + kDHT, "red overlays DH and DHT");
    __shared__ __attribute__((aligned(16))) char smem[kDH + kDHT + kVT + kQS];
    auto DH = reinterpret_cast<__bf16 (*)[kHeadH + 8]>(smem);
    auto DHT = reinterpret_cast<__bf16 (*)[64 + 8]>(smem + kDH);
    auto VT = reinterpret_cast<__bf16 (*)[64 + 8]>(smem + kDH + kDHT);
    auto QS = reinterpret_cast<__bf16 (*)[kHeadH + 8]>(smem + kDH + kDHT + kVT);  // Q1, bf16
    auto red = reinterpret_cast<float (*)[2 * kHeadH + 1]>(smem);  // after the last block
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    for (int i = tid; i < kHeadC * kHeadH; i += 256)
        QS[i / kHeadH][i % kHeadH] = (__bf16)Q1[i];
    float q2[kHeadTiles];
#pragma unroll
    for (int t = 0; t < kHeadTiles; ++t) q2[t] = Q2[16 * t + r16];
    constexpr int kGradTiles = (kHeadC / 16) * kHeadTiles;  // 16 output tiles of dQ1
    constexpr int kTilesPerWave = kGradTiles / 4;
    f4 gq1[kTilesPerWave];
#pragma unroll
    for (int i = 0; i < kTilesPerWave; ++i) gq1[i] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    float gq2[kHeadTiles], gb1[kHeadTiles];
#pragma unroll
    for (int t = 0; t < kHeadTiles; ++t) gq2[t] = gb1[t] = 0.0f;
    float gb2 = 0.0f;
    const int64_t r_begin = (int64_t)blockIdx.x * rows_per_part;
    const int64_t r_end = min(rows, r_begin + rows_per_part);
    HeadBlock<T> cur, nxt;
    head_load(cur, r_begin, r_end, dout, hpre, v, tid);
    for (int64_t blk0 = r_begin; blk0 < r_end; blk0 += 64) {
        // the next block's loads are in flight while this block is computed
        if (blk0 + 64 < r_end) head_load(nxt, blk0 + 64, r_end, dout, hpre, v, tid);
        if (r16 == 0) gb2 += (cur.d[0] + cur.d[1]) + (cur.d[2] + cur.d[3]);
#pragma unroll
        for (int t = 0; t < kHeadTiles; ++t) {
            const int col = 16 * t + r16;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float x = cur.hp[t][e];
                const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
                const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
                gq2[t] += x * cdf * cur.d[e];
                const float dh = cur.d[e] * q2[t] * (cdf + x * pdf);
                gb1[t] += dh;
                const int rl = 16 * wave + 4 * g + e;
                DH[rl][col] = (__bf16)dh;
                DHT[col][rl] = (__bf16)dh;
            }
        }
        if constexpr (std::is_same<T, float>::value) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int e4 = tid + 256 * i;
                const int rr = e4 >> 3, c4 = (e4 & 7) * 4;
#pragma unroll
                for (int j = 0; j < 4; ++j) VT[c4 + j][rr] = (__bf16)cur.vx[i][j];
            }
        } else {
            const int rr = tid >> 2, c8 = (tid & 3) * 8;
#pragma unroll
            for (int j = 0; j < 8; ++j) VT[c8 + j][rr] = cur.vx[0][j];
        }
        __syncthreads();
        const int64_t row0 = blk0 + 16 * wave;
        // dv for this wavefront's 16 rows: B[k = h][n = c] = Q1[c][h]
#pragma unroll
        for (int t = 0; t < kHeadC / 16; ++t) {
            f4 acc = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kk = 0; kk < kHeadH / 32; ++kk) {
                const bf16x8 a = *(const bf16x8 *)&DH[16 * wave + r16][32 * kk + 8 * g];
                const bf16x8 b = *(const bf16x8 *)&QS[16 * t + r16][32 * kk + 8 * g];
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t row = row0 + 4 * g + e;
                if (row < r_end) dv[row * kHeadC + 16 * t + r16] = acc[e];
            }
        }
        // dQ1 tiles of this wavefront over the block's 64 rows
#pragma unroll
        for (int i = 0; i < kTilesPerWave; ++i) {
            const int ti = wave + 4 * i;
            const int tm = ti / kHeadTiles, tn = ti - tm * kHeadTiles;
#pragma unroll
            for (int ks = 0; ks < 64; ks += 32) {
                const bf16x8 a = *(const bf16x8 *)&VT[16 * tm + r16][ks + 8 * g];
                const bf16x8 b = *(const bf16x8 *)&DHT[16 * tn + r16][ks + 8 * g];
                gq1[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, gq1[i], 0, 0, 0);
            }
        }
        __syncthreads();
        cur = nxt;
    }
    float *out = ws + (int64_t)blockIdx.x * kHeadJ;
#pragma unroll
    for (int i = 0; i < kTilesPerWave; ++i) {
        const int ti = wave + 4 * i;
        const int tm = ti / kHeadTiles, tn = ti - tm * kHeadTiles;
#pragma unroll
        for (int e = 0; e < 4; ++e)
            out[(16 * tm + 4 * g + e) * kHeadH + 16 * tn + r16] = gq1[i][e];
    }
    // column sums: 16 contributors per column (4 wavefronts x 4 lane groups),
    // combined in (wavefront, group) order
    const int slot = wave * 4 + g;
#pragma unroll
    for (int t = 0; t < kHeadTiles; ++t) {
        red[slot][16 * t + r16] = gb1[t];
        red[slot][kHeadH + 16 * t + r16] = gq2[t];
    }
    if (r16 == 0) red[slot][2 * kHeadH] = gb2;
    __syncthreads();
    for (int c = tid; c < 2 * kHeadH + 1; c += 256) {
        float sum = 0.0f;
        for (int u = 0; u < 16; ++u) sum += red[u][c];
        out[kHeadC * kHeadH + c] = sum;
    }
}

// Sum of the head's partials in part order (16 contiguous ranges added in
// range order) into dQ1 | dbq1 | dQ2 | dbq2.
__global__ void __launch_bounds__(256) head_reduce_kernel(int n_part, const float *ws,
                                                          float *dQ1, float *dbq1, float *dQ2,
                                                          float *dbq2)
{
    __shared__ float part[16][16];
    const int tid = threadIdx.x;
    const int jj = tid & 15, q = tid >> 4;
    const int j = blockIdx.x * 16 + jj;
    const int per = (n_part + 15) / 16;
    const int p0 = q * per, p1 = min(n_part, p0 + per);
    float s = 0.0f;
    if (j < kHeadJ)
        for (int pp = p0; pp < p1; ++pp) s += ws[(int64_t)pp * kHeadJ + j];
    part[q][jj] = s;
    __syncthreads();
    if (q == 0 && j < kHeadJ) {
        float t = 0.0f;
#pragma unroll
        for (int u = 0; u < 16; ++u) t += part[u][jj];
        constexpr int o1 = kHeadC * kHeadH, o2 = o1 + kHeadH, o3 = o2 + kHeadH;
        if (j < o1)
            dQ1[j] = t;
        else if (j < o2)
            dbq1[j - o1] = t;
        else if (j < o3)
            dQ2[j - o2] = t;
        else
            dbq2[0] = t;
    }
}

template <typename TY>  // y: fp32 or bf16 codes
__global__ void __launch_bounds__(256) bias_act_kernel(int64_t rows, int cols, const float *x,
                                                       const float *bias, float *pre, TY *y,
                                                       int use_gelu)
{
    const int64_t n = rows * cols;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % cols);
        const float v = x[i] + (bias != nullptr ? bias[c] : 0.0f);
        pre[i] = v;
        y[i] = (TY)(use_gelu ? gelu(v) : v);
    }
}

__global__ void __launch_bounds__(256) gelu_backward_kernel(int64_t n, const float *pre,
                                                            const float *dy, float *dx)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dx[i] = dy[i] * gelu_grad(pre[i]);
}

constexpr int kColRows = 4096;  // rows per partial of a column sum

// Partial column sums: block b sums rows [b kColRows, (b+1) kColRows) of every
// column; thread t takes column t % cols and every (256 / cols)-th row from
// t / cols, and the block combines those in lane order.
__global__ void __launch_bounds__(256) col_sum_partial_kernel(int64_t rows, int cols,
                                                              const float *x, float *ws)
{
    __shared__ float part[256];
    const int t = threadIdx.x;
    const int per = max(1, 256 / cols);
    const int c = t % cols;
    const int r0 = t / cols;
    const int64_t base = (int64_t)blockIdx.x * kColRows;
    float s = 0.0f;
    if (r0 < per && t < per * cols) {
        for (int64_t r = base + r0; r < min(rows, base + kColRows); r += per)
            s += x[r * cols + c];
    }
    part[t] = s;
    __syncthreads();
    if (t < cols) {
        float acc = 0.0f;
        for (int k = 0; k < per; ++k) acc += part[k * cols + t];
        ws[(int64_t)blockIdx.x * cols + t] = acc;
    }
}

__global__ void col_sum_final_kernel(int n_part, int cols, const float *ws, float beta,
                                     float *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int k = 0; k < n_part; ++k) s += ws[(int64_t)k * cols + c];
    out[c] = beta != 0.0f ? beta * out[c] + s : s;
}

__global__ void __launch_bounds__(256) spectral_pack_kernel(int op, int batch, int m, int ch,
                                                            const float *in, float *out)
{
    // one thread per (b, q, j, c) of the (batch, m, m, ch) mode grid
    const int64_t n = (int64_t)batch * m * m * ch;
    const int two_m = 2 * m;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(t % ch);
        int64_t r = t / ch;
        const int j = (int)(r % m);
        r /= m;
        const int q = (int)(r % m);
        const int64_t b = r / m;
        // full grid (b, 2m, 2m, ch) and mode grid (b, m, m, 2 ch) offsets
        auto full = [&](int rr, int jj) -> int64_t {
            return ((b * two_m + rr) * two_m + jj) * ch + c;
        };
        const int64_t mre = ((b * m + q) * m + j) * (2 * ch) + c;
        const int64_t mim = mre + ch;
        if (op == 0) {
            out[mre] = in[full(q, j)] + in[full(m + q, m + j)];
            out[mim] = in[full(q, m + j)] - in[full(m + q, j)];
        } else if (op == 1) {
            out[full(q, j)] = in[mre];
            out[full(m + q, m + j)] = in[mre];
            out[full(q, m + j)] = in[mim];
            out[full(m + q, j)] = -in[mim];
        } else if (op == 2) {
            out[full(q, j)] = in[mre];
            out[full(q, m + j)] = in[mim];
            out[full(m + q, j)] = in[mim];
            out[full(m + q, m + j)] = -in[mre];
        } else {
            out[mre] = in[full(q, j)] - in[full(m + q, m + j)];
            out[mim] = in[full(q, m + j)] + in[full(m + q, j)];
        }
    }
}

__global__ void __launch_bounds__(256) mix_assemble_kernel(int modes2, int ch, const float *re,
                                                           const float *im, float *rm)
{
    const int64_t n = (int64_t)modes2 * ch * ch;
    const int w = 2 * ch;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(t % ch);
        const int i = (int)((t / ch) % ch);
        const int64_t md = t / ((int64_t)ch * ch);
        float *r = rm + md * w * w;
        const float a = re[t], bb = im[t];
        r[(int64_t)i * w + o] = a;
        r[(int64_t)i * w + ch + o] = bb;
        r[(int64_t)(ch + i) * w + o] = -bb;
        r[(int64_t)(ch + i) * w + ch + o] = a;
    }
}

__global__ void __launch_bounds__(256) mix_assemble_backward_kernel(int modes2, int ch,
                                                                    const float *drm, float *dre,
                                                                    float *dim, int accumulate)
{
    const int64_t n = (int64_t)modes2 * ch * ch;
    const int w = 2 * ch;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(t % ch);
        const int i = (int)((t / ch) % ch);
        const int64_t md = t / ((int64_t)ch * ch);
        const float *r = drm + md * w * w;
        const float gre = r[(int64_t)i * w + o] + r[(int64_t)(ch + i) * w + ch + o];
        const float gim = r[(int64_t)i * w + ch + o] - r[(int64_t)(ch + i) * w + o];
        dre[t] = accumulate ? dre[t] + gre : gre;
        dim[t] = accumulate ? dim[t] + gim : gim;
    }
}

constexpr int kMseBlock = 4096;  // elements per partial of the loss

__global__ void __launch_bounds__(256) mse_partial_kernel(int64_t n, const float *pred,
                                                          const float *target, const float *mask,
                                                          float *grad, float *ws)
{
    __shared__ float s_e[256], s_w[256];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kMseBlock;
    float e = 0.0f, w = 0.0f;
    for (int64_t i = base + t; i < min(n, base + kMseBlock); i += 256) {
        const float wi = mask != nullptr ? mask[i] : 1.0f;
        const float d = pred[i] - target[i];
        e += wi * d * d;
        w += wi;
        grad[i] = 2.0f * wi * d;  // scaled by 1 / sum(w) afterwards
    }
    s_e[t] = e;
    s_w[t] = w;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            s_e[t] += s_e[t + s];
            s_w[t] += s_w[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        ws[2 * blockIdx.x] = s_e[0];
        ws[2 * blockIdx.x + 1] = s_w[0];
    }
}

__global__ void mse_final_kernel(int n_part, const float *ws, float *loss)
{
    if (threadIdx.x != 0) return;
    float e = 0.0f, w = 0.0f;
    for (int k = 0; k < n_part; ++k) {
        e += ws[2 * k];
        w += ws[2 * k + 1];
    }
    loss[0] = w > 0.0f ? e / w : 0.0f;
    loss[1] = w > 0.0f ? 1.0f / w : 0.0f;
}

__global__ void __launch_bounds__(256) scale_by_kernel(int64_t n, float *x, const float *s)
{
    const float f = s[0];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= f;
}

__global__ void __launch_bounds__(256) adamw_kernel(int64_t n, float *param, const float *grad,
                                                    float *m, float *v, float lr, float b1,
                                                    float b2, float eps, float wd, float c1,
                                                    float c2)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float g = grad[i];
        const float mi = b1 * m[i] + (1.0f - b1) * g;
        const float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / c1, vh = vi / c2;
        param[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * param[i]);
    }
}

unsigned blocks_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

int k_chunk_of(const swno_gemm_args *g)
{
    const int ks = std::max(1, g->k_split);
    int chunk = (g->K + ks - 1) / ks;
    chunk = (chunk + kTK - 1) / kTK * kTK;
    return std::max(chunk, kTK);
}

template <int MP, int NP, int VA, int VB>
void launch_wgrad(dim3 grid, hipStream_t st, const WgradP &p)
{
    hipLaunchKernelGGL((wgrad_kernel<MP, NP, VA, VB>), grid, dim3(256), 0, st, p);
}

// va: 0 scalar, 1 float4, 2 bf16 codes (8 per load) for A
template <int MP, int NP>
void launch_wgrad_v(int va, bool vb, dim3 grid, hipStream_t st, const WgradP &p)
{
    if (va == 2) {
        if constexpr (MP >= 32) {
            if (vb) {
                launch_wgrad<MP, NP, 2, 1>(grid, st, p);
            } else if constexpr (NP <= 32) {
                launch_wgrad<MP, NP, 2, 0>(grid, st, p);
            }
        }
    } else if (va) {
        if (vb) {
            launch_wgrad<MP, NP, 1, 1>(grid, st, p);
        } else if constexpr (NP <= 32) {
            launch_wgrad<MP, NP, 1, 0>(grid, st, p);
        }
    } else if constexpr (MP <= 32) {
        if (vb) {
            launch_wgrad<MP, NP, 0, 1>(grid, st, p);
        } else if constexpr (NP <= 32) {
            launch_wgrad<MP, NP, 0, 0>(grid, st, p);
        }
    }
}

template <int MP>
void launch_wgrad_n(int np, int va, bool vb, dim3 grid, hipStream_t st, const WgradP &p)
{
    switch (np) {
    case 16: launch_wgrad_v<MP, 16>(va, vb, grid, st, p); break;
    case 32: launch_wgrad_v<MP, 32>(va, vb, grid, st, p); break;
    case 64: launch_wgrad_v<MP, 64>(va, vb, grid, st, p); break;
    default: launch_wgrad_v<MP, 128>(va, vb, grid, st, p); break;
    }
}

void launch_wgrad_m(int mp, int np, int va, bool vb, dim3 grid, hipStream_t st,
                    const WgradP &p)
{
    switch (mp) {
    case 16: launch_wgrad_n<16>(np, va, vb, grid, st, p); break;
    case 32: launch_wgrad_n<32>(np, va, vb, grid, st, p); break;
    case 64: launch_wgrad_n<64>(np, va, vb, grid, st, p); break;
    default: launch_wgrad_n<128>(np, va, vb, grid, st, p); break;
    }
}

}  // namespace

extern "C" {


const char *swno_last_error(void) { return g_err.c_str(); }

size_t swno_gemm_workspace_floats(const swno_gemm_args *g)
{
    if (g == nullptr || g->k_split <= 1) return 0;
    return (size_t)g->batch * g->k_split * g->M * g->N;
}

int swno_gemm(const swno_gemm_args *g, float *workspace, void *stream)
{
    if (g == nullptr) return fail(SWNO_EINVAL, "gemm args are NULL");
    if (g->M < 1 || g->N < 1 || g->K < 1 || g->batch < 1 || g->k_split < 1)
        return fail(SWNO_EINVAL, "gemm: M, N, K, batch and k_split must be positive");
    if (g->A == nullptr || g->B == nullptr || g->C == nullptr)
        return fail(SWNO_EINVAL, "gemm: operand is NULL");
    if (g->k_split > 1 && workspace == nullptr)
        return fail(SWNO_EINVAL, "gemm: split-K needs the workspace");
    if ((int64_t)g->batch * g->k_split > 65535)
        return fail(SWNO_EINVAL, "gemm: batch * k_split must be at most 65535");
    GemmP p;
    std::memset(&p, 0, sizeof p);
    p.M = g->M;
    p.N = g->N;
    p.K = g->K;
    p.batch = g->batch;
    p.k_split = g->k_split;
    p.k_chunk = k_chunk_of(g);
    p.A = g->A;
    p.a_rs = g->a_rs;
    p.a_cs = g->a_cs;
    p.a_bs = g->a_bs;
    p.B = g->B;
    p.b_rs = g->b_rs;
    p.b_cs = g->b_cs;
    p.b_bs = g->b_bs;
    p.C = g->C;
    p.c_rs = g->c_rs;
    p.c_cs = g->c_cs;
    p.c_bs = g->c_bs;
    p.alpha = g->alpha;
    p.beta = g->beta;
    p.ws = workspace;
    hipStream_t st = as_stream(stream);
    const bool big = g->M >= 64 && g->N >= 64;
    const int tm = big ? 64 : 32, tn = big ? 64 : 32;
    const dim3 grid((unsigned)((g->M + tm - 1) / tm), (unsigned)((g->N + tn - 1) / tn),
                    (unsigned)(g->batch * g->k_split));
    if (grid.y > 65535) return fail(SWNO_EINVAL, "gemm: N too large");
    if (big)
        hipLaunchKernelGGL((gemm_kernel<64, 64>), grid, dim3(256), 0, st, p);
    else
        hipLaunchKernelGGL((gemm_kernel<32, 32>), grid, dim3(256), 0, st, p);
    SWNO_HIP_CHECK(hipGetLastError());
    if (g->k_split > 1) {
        hipLaunchKernelGGL(gemm_reduce_kernel,
                           dim3(blocks_for((int64_t)g->batch * g->M * g->N)), dim3(256), 0, st,
                           p);
        SWNO_HIP_CHECK(hipGetLastError());
    }
    return SWNO_OK;
}

int swno_bias_act(int64_t rows, int32_t cols, const float *x, const float *bias, float *pre,
                  void *y, int32_t use_gelu, int32_t y_bf16, void *stream)
{
    if (rows < 1 || cols < 1 || x == nullptr || pre == nullptr || y == nullptr)
        return fail(SWNO_EINVAL, "bias_act: bad arguments");
    if ((use_gelu || y_bf16) && y == pre)
        return fail(SWNO_EINVAL, "bias_act: y must differ from pre");
    if (y_bf16)
        hipLaunchKernelGGL(bias_act_kernel<__bf16>, dim3(blocks_for(rows * cols)), dim3(256), 0,
                           as_stream(stream), rows, cols, x, bias, pre, (__bf16 *)y, use_gelu);
    else
        hipLaunchKernelGGL(bias_act_kernel<float>, dim3(blocks_for(rows * cols)), dim3(256), 0,
                           as_stream(stream), rows, cols, x, bias, pre, (float *)y, use_gelu);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

int swno_gelu_backward(int64_t n, const float *pre, const float *dy, float *dx, void *stream)
{
    if (n < 1 || pre == nullptr || dy == nullptr || dx == nullptr)
        return fail(SWNO_EINVAL, "gelu_backward: bad arguments");
    hipLaunchKernelGGL(gelu_backward_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       as_stream(stream), n, pre, dy, dx);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

size_t swno_weight_grad_workspace_floats(int64_t rows, int32_t M, int32_t N)
{
    if (rows < 1 || M < 1 || N < 1) return 0;
    const int64_t slabs = (rows + kWgRows - 1) / kWgRows;
    const int64_t parts = std::min<int64_t>(kWgMaxParts, slabs);
    return (size_t)parts * (size_t)(M * N + N);
}

int swno_weight_grad(int64_t rows, int32_t M, int32_t N, const void *A, int64_t lda,
                     const float *B, int64_t ldb, float *dW, float *db, float *workspace,
                     int32_t a_bf16, void *stream)
{
    if (rows < 1 || M < 1 || N < 1 || M > 128 || N > 128)
        return fail(SWNO_EINVAL, "weight_grad: rows >= 1 and 1 <= M, N <= 128 required");
    if (A == nullptr || B == nullptr || dW == nullptr || workspace == nullptr)
        return fail(SWNO_EINVAL, "weight_grad: operand or workspace is NULL");
    if (lda < M || ldb < N) return fail(SWNO_EINVAL, "weight_grad: lda < M or ldb < N");
    WgradP p;
    std::memset(&p, 0, sizeof p);
    p.rows = rows;
    p.M = M;
    p.N = N;
    p.A = A;
    p.lda = lda;
    p.B = B;
    p.ldb = ldb;
    const bool va = M % 4 == 0 && lda % 4 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0;
    if (a_bf16 && !(M % 8 == 0 && M > 16 && lda % 8 == 0 &&
                    reinterpret_cast<uintptr_t>(A) % 16 == 0))
        return fail(SWNO_EINVAL, "weight_grad: a bf16 A needs M a multiple of 8 above 16, lda a "
                                 "multiple of 8 and a 16-byte aligned base");
    const bool vb = N % 4 == 0 && ldb % 4 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0;
    // scalar loads keep up to 8 values per thread: 64 rows x 32 columns
    if ((!va && M > 32) || (!vb && N > 32))
        return fail(SWNO_EINVAL, "weight_grad: an operand wider than 32 columns must allow "
                                 "float4 loads (width and row stride multiples of 4, 16-byte "
                                 "aligned)");
    p.bias = db != nullptr;
    if (p.bias && (vb ? 256 % (N / 4) : 256 % N) != 0)
        return fail(SWNO_EINVAL, "weight_grad: column sums need N (N / 4 with float4 loads) "
                                 "to divide 256");
    const int64_t slabs = (rows + kWgRows - 1) / kWgRows;
    const int64_t parts0 = std::min<int64_t>(kWgMaxParts, slabs);
    p.rows_per_part = (slabs + parts0 - 1) / parts0 * kWgRows;
    const int n_part = (int)((rows + p.rows_per_part - 1) / p.rows_per_part);
    p.ws = workspace;
    hipStream_t st = as_stream(stream);
    const int mp = M <= 16 ? 16 : M <= 32 ? 32 : M <= 64 ? 64 : 128;
    const int np = N <= 16 ? 16 : N <= 32 ? 32 : N <= 64 ? 64 : 128;
    launch_wgrad_m(mp, np, a_bf16 ? 2 : (va ? 1 : 0), vb, dim3((unsigned)n_part), st, p);
    SWNO_HIP_CHECK(hipGetLastError());
    const int MN = M * N, J = MN + (p.bias ? N : 0);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((unsigned)((J + 15) / 16)), dim3(256), 0, st,
                       n_part, MN, J, (const float *)workspace, dW, db);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

bool layer_shape_ok(int32_t slices, int32_t n, int32_t modes, int32_t ch)
{
    return slices >= 1 && ch == kLayerC && n >= 16 && n <= kLayerMaxN && n % 16 == 0 &&
           modes >= 2 && modes % 2 == 0 && 2 * modes <= 32;
}

int swno_x_transform(int32_t slices, int32_t rows_out, int32_t n, int32_t ch, const float *A,
                     int64_t a_rs, int64_t a_cs, const void *X, float *out, int32_t x_bf16,
                     void *stream)
{
    if (slices < 1 || ch != kLayerC || rows_out < 1 || rows_out > 32 || n < 1 ||
        n > kLayerMaxN)
        return fail(SWNO_EINVAL, "x_transform: needs 32 channels, 1 <= rows_out <= 32 and "
                                 "1 <= n <= 128");
    if (A == nullptr || X == nullptr || out == nullptr)
        return fail(SWNO_EINVAL, "x_transform: operand is NULL");
    if (reinterpret_cast<uintptr_t>(X) % 16 != 0)
        return fail(SWNO_EINVAL, "x_transform: X must be 16-byte aligned");
    // a few slices per workgroup, so that the next slice's loads overlap the current one
    const unsigned grid = (unsigned)std::min(slices, 2048);
    if (x_bf16)
        hipLaunchKernelGGL(x_transform_kernel<__bf16>, dim3(grid), dim3(256), 0,
                           as_stream(stream), slices, rows_out, n, A, a_rs, a_cs,
                           (const __bf16 *)X, out);
    else
        hipLaunchKernelGGL(x_transform_kernel<float>, dim3(grid), dim3(256), 0,
                           as_stream(stream), slices, rows_out, n, A, a_rs, a_cs,
                           (const float *)X, out);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

int swno_layer_forward(int32_t slices, int32_t n, int32_t modes, int32_t ch, const float *Gx,
                       const float *Z, const void *v, const float *W, const float *bias,
                       float *pre, void *vnext, int32_t gelu_act, int32_t bf16_flags,
                       void *stream)
{
    if (!layer_shape_ok(slices, n, modes, ch))
        return fail(SWNO_EINVAL, "layer_forward: needs 32 channels, n a multiple of 16 up to "
                                 "128 and an even number of modes up to 16");
    if (Gx == nullptr || Z == nullptr || v == nullptr || W == nullptr || bias == nullptr ||
        pre == nullptr || vnext == nullptr)
        return fail(SWNO_EINVAL, "layer_forward: operand is NULL");
    if (reinterpret_cast<uintptr_t>(v) % 16 != 0 || reinterpret_cast<uintptr_t>(Gx) % 16 != 0)
        return fail(SWNO_EINVAL, "layer_forward: v and Gx must be 16-byte aligned");
    if ((bf16_flags & 2) != 0 && reinterpret_cast<uintptr_t>(vnext) % 16 != 0)
        return fail(SWNO_EINVAL, "layer_forward: a bf16 vnext must be 16-byte aligned");
    const unsigned grid = (unsigned)std::min(slices, 8192);
    hipStream_t st = as_stream(stream);
#define SWNO_LAYER_FWD(TV, TN)                                                              \
    hipLaunchKernelGGL((layer_forward_kernel<TV, TN>), dim3(grid), dim3(256), 0, st, slices, n, \
                       2 * modes, Gx, Z, (const TV *)v, W, bias, pre, (TN *)vnext, gelu_act)
    switch (bf16_flags & 3) {
    case 0: SWNO_LAYER_FWD(float, float); break;
    case 1: SWNO_LAYER_FWD(__bf16, float); break;
    case 2: SWNO_LAYER_FWD(float, __bf16); break;
    default: SWNO_LAYER_FWD(__bf16, __bf16); break;
    }
#undef SWNO_LAYER_FWD
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

int swno_layer_backward(int32_t slices, int32_t n, int32_t modes, int32_t ch, const float *dpre,
                        const float *W, const float *Fx, const float *dT1,
                        const float *pre_prev, float *out, void *stream)
{
    if (!layer_shape_ok(slices, n, modes, ch))
        return fail(SWNO_EINVAL, "layer_backward: needs 32 channels, n a multiple of 16 up to "
                                 "128 and an even number of modes up to 16");
    if (dpre == nullptr || W == nullptr || Fx == nullptr || dT1 == nullptr || out == nullptr)
        return fail(SWNO_EINVAL, "layer_backward: operand is NULL");
    if (out == dpre || (pre_prev != nullptr && out == pre_prev))
        return fail(SWNO_EINVAL, "layer_backward: out must not alias dpre or pre_prev");
    if (reinterpret_cast<uintptr_t>(dpre) % 16 != 0 || reinterpret_cast<uintptr_t>(W) % 16 != 0)
        return fail(SWNO_EINVAL, "layer_backward: dpre and W must be 16-byte aligned");
    const unsigned grid = (unsigned)std::min(slices, 2048);  // a few slices each (prefetch)
    hipLaunchKernelGGL(layer_backward_kernel, dim3(grid), dim3(256), 0, as_stream(stream),
                       slices, n, 2 * modes, dpre, W, Fx, dT1, pre_prev, out);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

constexpr int kHeadMaxParts = 1024;

int head_parts(int64_t rows, int64_t *rows_per_part)
{
    const int64_t blocks = (rows + 63) / 64;
    const int64_t parts0 = std::min<int64_t>(kHeadMaxParts, blocks);
    *rows_per_part = (blocks + parts0 - 1) / parts0 * 64;
    return (int)((rows + *rows_per_part - 1) / *rows_per_part);
}

int swno_head_forward(int64_t rows, int32_t ch, int32_t hidden, const void *v, const float *Q1,
                      const float *bq1, const float *Q2, const float *bq2, float *hpre,
                      float *out, int32_t v_bf16, void *stream)
{
    if (ch != kHeadC || hidden != kHeadH)
        return fail(SWNO_EINVAL, "head_forward: built for 32 channels and 128 hidden units");
    if (rows < 1 || v == nullptr || Q1 == nullptr || bq1 == nullptr || Q2 == nullptr ||
        bq2 == nullptr || hpre == nullptr || out == nullptr)
        return fail(SWNO_EINVAL, "head_forward: bad arguments");
    if (reinterpret_cast<uintptr_t>(v) % 16 != 0)
        return fail(SWNO_EINVAL, "head_forward: v must be 16-byte aligned");
    const int64_t blocks = (rows + 63) / 64;
    const unsigned grid = (unsigned)std::min<int64_t>(blocks, 4096);
    if (v_bf16)
        hipLaunchKernelGGL(head_forward_kernel<__bf16>, dim3(grid), dim3(256), 0,
                           as_stream(stream), rows, (const __bf16 *)v, Q1, bq1, Q2, bq2, hpre,
                           out);
    else
        hipLaunchKernelGGL(head_forward_kernel<float>, dim3(grid), dim3(256), 0,
                           as_stream(stream), rows, (const float *)v, Q1, bq1, Q2, bq2, hpre,
                           out);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

size_t swno_head_backward_workspace_floats(int64_t rows)
{
    if (rows < 1) return 0;
    int64_t rpp = 0;
    return (size_t)head_parts(rows, &rpp) * kHeadJ;
}

int swno_head_backward(int64_t rows, int32_t ch, int32_t hidden, const float *dout,
                       const float *hpre, const void *v, const float *Q1, const float *Q2,
                       float *dv, float *dQ1, float *dbq1, float *dQ2, float *dbq2,
                       float *workspace, int32_t v_bf16, void *stream)
{
    if (ch != kHeadC || hidden != kHeadH)
        return fail(SWNO_EINVAL, "head_backward: built for 32 channels and 128 hidden units");
    if (rows < 1 || dout == nullptr || hpre == nullptr || v == nullptr || Q1 == nullptr ||
        Q2 == nullptr || dv == nullptr || dQ1 == nullptr || dbq1 == nullptr ||
        dQ2 == nullptr || dbq2 == nullptr || workspace == nullptr)
        return fail(SWNO_EINVAL, "head_backward: bad arguments");
    if (reinterpret_cast<uintptr_t>(v) % 16 != 0)
        return fail(SWNO_EINVAL, "head_backward: v must be 16-byte aligned");
    int64_t rpp = 0;
    const int n_part = head_parts(rows, &rpp);
    hipStream_t st = as_stream(stream);
    if (v_bf16)
        hipLaunchKernelGGL(head_backward_kernel<__bf16>, dim3((unsigned)n_part), dim3(256), 0, st,
                           rows, rpp, dout, hpre, (const __bf16 *)v, Q1, Q2, dv, workspace);
    else
        hipLaunchKernelGGL(head_backward_kernel<float>, dim3((unsigned)n_part), dim3(256), 0, st,
                           rows, rpp, dout, hpre, (const float *)v, Q1, Q2, dv, workspace);
    SWNO_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(head_reduce_kernel, dim3((unsigned)((kHeadJ + 15) / 16)), dim3(256), 0,
                       st, n_part, (const float *)workspace, dQ1, dbq1, dQ2, dbq2);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

size_t swno_col_sum_workspace_floats(int64_t rows, int32_t cols)
{
    return (size_t)((rows + kColRows - 1) / kColRows) * (size_t)cols;
}

int swno_col_sum(int64_t rows, int32_t cols, const float *x, float beta, float *out,
                 float *workspace, void *stream)
{
    if (rows < 1 || cols < 1 || cols > 256 || x == nullptr || out == nullptr ||
        workspace == nullptr)
        return fail(SWNO_EINVAL, "col_sum: bad arguments (cols must be 1..256)");
    const int n_part = (int)((rows + kColRows - 1) / kColRows);
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(col_sum_partial_kernel, dim3(n_part), dim3(256), 0, st, rows, cols, x,
                       workspace);
    SWNO_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(col_sum_final_kernel, dim3((cols + 255) / 256), dim3(256), 0, st, n_part,
                       cols, workspace, beta, out);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

int swno_spectral_pack(int32_t op, int32_t batch, int32_t modes, int32_t ch, const float *in,
                       float *out, void *stream)
{
    if (op < 0 || op > 3 || batch < 1 || modes < 1 || ch < 1 || in == nullptr ||
        out == nullptr || in == out)
        return fail(SWNO_EINVAL, "spectral_pack: bad arguments");
    const int64_t n = (int64_t)batch * modes * modes * ch;
    hipLaunchKernelGGL(spectral_pack_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       as_stream(stream), op, batch, modes, ch, in, out);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

int swno_mix_assemble(int32_t modes2, int32_t ch, const float *re, const float *im, float *rm,
                      void *stream)
{
    if (modes2 < 1 || ch < 1 || re == nullptr || im == nullptr || rm == nullptr)
        return fail(SWNO_EINVAL, "mix_assemble: bad arguments");
    hipLaunchKernelGGL(mix_assemble_kernel, dim3(blocks_for((int64_t)modes2 * ch * ch)),
                       dim3(256), 0, as_stream(stream), modes2, ch, re, im, rm);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

int swno_mix_assemble_backward(int32_t modes2, int32_t ch, const float *drm, float *d_re,
                               float *d_im, int32_t accumulate, void *stream)
{
    if (modes2 < 1 || ch < 1 || drm == nullptr || d_re == nullptr || d_im == nullptr)
        return fail(SWNO_EINVAL, "mix_assemble_backward: bad arguments");
    hipLaunchKernelGGL(mix_assemble_backward_kernel,
                       dim3(blocks_for((int64_t)modes2 * ch * ch)), dim3(256), 0,
                       as_stream(stream), modes2, ch, drm, d_re, d_im, accumulate);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

size_t swno_mse_workspace_floats(int64_t n) { return 2 * (size_t)((n + kMseBlock - 1) / kMseBlock); }

int swno_mse(int64_t n, const float *pred, const float *target, const float *mask, float *loss,
             float *grad, float *workspace, void *stream)
{
    if (n < 1 || pred == nullptr || target == nullptr || loss == nullptr || grad == nullptr ||
        workspace == nullptr)
        return fail(SWNO_EINVAL, "mse: bad arguments");
    const int n_part = (int)((n + kMseBlock - 1) / kMseBlock);
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(mse_partial_kernel, dim3(n_part), dim3(256), 0, st, n, pred, target, mask,
                       grad, workspace);
    SWNO_HIP_CHECK(hipGetLastError());
    // loss[0] = mean, loss[1] = 1 / sum(w) (used to scale the gradient)
    hipLaunchKernelGGL(mse_final_kernel, dim3(1), dim3(64), 0, st, n_part, workspace, loss);
    SWNO_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(scale_by_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, grad,
                       loss + 1);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

int swno_adamw(int64_t n, float *param, const float *grad, float *m, float *v, float lr,
               float beta1, float beta2, float eps, float weight_decay, int32_t step,
               void *stream)
{
    if (n < 1 || param == nullptr || grad == nullptr || m == nullptr || v == nullptr ||
        step < 1)
        return fail(SWNO_EINVAL, "adamw: bad arguments");
    const float c1 = 1.0f - std::pow(beta1, (float)step);
    const float c2 = 1.0f - std::pow(beta2, (float)step);
    hipLaunchKernelGGL(adamw_kernel, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), n,
                       param, grad, m, v, lr, beta1, beta2, eps, weight_decay, c1, c2);
    SWNO_HIP_CHECK(hipGetLastError());
    return SWNO_OK;
}

}  // extern "C"
