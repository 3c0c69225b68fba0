// Spectral neural operator kernels for MI355X (gfx950): a batched strided
// GEMM on bf16 MFMA with fp32 accumulation, and the elementwise kernels of
// the operator (bias + GELU, spectral repacks, mode-mixing weights, masked
// MSE, AdamW). The C ABI is declared in include/swno.h.
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

__global__ void __launch_bounds__(256) bias_act_kernel(int64_t rows, int cols, const float *x,
                                                       const float *bias, float *pre, float *y,
                                                       int use_gelu)
{
    const int64_t n = rows * cols;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % cols);
        const float v = x[i] + (bias != nullptr ? bias[c] : 0.0f);
        pre[i] = v;
        y[i] = use_gelu ? gelu(v) : v;
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
                  float *y, int32_t use_gelu, void *stream)
{
    if (rows < 1 || cols < 1 || x == nullptr || pre == nullptr || y == nullptr)
        return fail(SWNO_EINVAL, "bias_act: bad arguments");
    if (use_gelu && y == pre) return fail(SWNO_EINVAL, "bias_act: y must differ from pre");
    hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_for(rows * cols)), dim3(256), 0,
                       as_stream(stream), rows, cols, x, bias, pre, y, use_gelu);
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
