// chi_B surrogate: a small multi-layer perceptron trained and evaluated with hand-written kernels.
//
// No reference counterpart: SQAChiAnnealer::run_chi (sqa_chi_annealer.cpp:232-281) always runs the
// pilot scan; the surrogate predicts that scan's chi_B profile from instance statistics so the
// production anneal can start without it, and falls back to the pilot when it is unsure.
//
// Model: an ensemble of `members` networks, each n_in -> width (tanh) x depth -> n_out (sigmoid).
// Parameters of one member, contiguous: for each layer l, W_l [out_l][in_l] then b_l [out_l].
// Every member runs in the same launch (blockIdx.z = member). Precision: FP32 master parameters
// (Adam state in FP32); the products read bf16 copies of the weights (round to nearest even,
// refreshed after every update) times FP32 activations with FP32 accumulation; biases stay FP32.
// Gradients are those of the network with the bf16 weights.
//
// Kernels:
//   gemm_kernel     C = epilogue(op(A) op(B)) for the three products of a dense layer
//                   (forward X W^T, weight gradient dZ^T X, data gradient dZ W), 64 x 64 tiles of
//                   4 x 4 per thread through LDS, FP32 accumulation;
//   dact_kernel     dZ = dA * act'(A) from the layer's stored output;
//   colsum_kernel   bias gradient sum_b dZ[b][o];
//   loss_kernel     mean squared error per member and dL/dY;
//   adam_kernel     Adam with bias correction and decoupled weight decay;
//   gather_kernel   minibatch rows by index.
// The network is latency-bound at this size (about 0.15 M parameters per member at width 256):
// one training step is a dozen launches over a few hundred KB.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <type_traits>
#include <vector>

#include "qa_internal.hpp"

namespace qa {
namespace {

enum { ACT_NONE = 0, ACT_TANH = 1, ACT_SIGMOID = 2 };

template <class TB>
struct GemmArgs {
    const float *A;
    int64_t a_rs, a_cs, a_mz;  // A(m,k) = A[m*a_rs + k*a_cs + z*a_mz]
    const TB *B;
    int64_t b_rs, b_cs, b_mz;  // B(k,n) = B[k*b_rs + n*b_cs + z*b_mz]; uint16_t: bf16 bits
    float *C;
    int64_t c_rs, c_mz;        // C[m*c_rs + n + z*c_mz]
    const float *bias;         // bias[n + z*bias_mz] or null
    int64_t bias_mz;
    int M, N, K;
    int act;
};

constexpr int kT = 64, kBK = 16;

__device__ __forceinline__ float activate(float x, int act) {
    if (act == ACT_TANH) return tanhf(x);
    if (act == ACT_SIGMOID) return 1.0f / (1.0f + expf(-x));
    return x;
}

__device__ __forceinline__ float load_b(const float *p) { return *p; }
__device__ __forceinline__ float load_b(const uint16_t *p) { return __uint_as_float(static_cast<uint32_t>(*p) << 16); }

__device__ __forceinline__ uint16_t bf16_bits(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return static_cast<uint16_t>(u >> 16);
}

template <class TB>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs<TB> g) {
    __shared__ float As[kBK][kT + 4];
    __shared__ float Bs[kBK][kT + 4];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
    const int64_t z = blockIdx.z;
    const float *A = g.A + z * g.a_mz;
    const TB *B = g.B + z * g.b_mz;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < g.K; k0 += kBK) {
        for (int e = tid; e < kT * kBK; e += 256) {
            const int r = e / kBK, k = e - r * kBK;  // A tile row r (m), column k
            const int m = m0 + r, kk = k0 + k;
            As[k][r] = (m < g.M && kk < g.K) ? A[m * g.a_rs + kk * g.a_cs] : 0.0f;
            const int c = e % kT, kb = e / kT;        // B tile row kb (k), column c (n)
            const int n = n0 + c, kk2 = k0 + kb;
            Bs[kb][c] = (n < g.N && kk2 < g.K) ? load_b(B + kk2 * g.b_rs + n * g.b_cs) : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kBK; ++k) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[k][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    float *C = g.C + z * g.c_mz;
    const float *bias = g.bias ? g.bias + z * g.bias_mz : nullptr;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + ty * 4 + i;
        if (m >= g.M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + tx * 4 + j;
            if (n >= g.N) continue;
            const float v = acc[i][j] + (bias ? bias[n] : 0.0f);
            C[m * g.c_rs + n] = activate(v, g.act);
        }
    }
}

// The same products on the matrix cores: a 64 x 64 tile of C per workgroup of four waves (2 x 2
// waves of 32 x 32), v_mfma_f32_32x32x16_bf16 with FP32 accumulation. An FP32 operand x enters as
// two bf16 terms, hi = bf16(x) and lo = bf16(x - hi) (|x - hi - lo| <= 2^-16 |x| ulp-wise); bf16
// weights enter as they are. With bf16 B: C = A_hi B + A_lo B; with FP32 B (the weight gradient):
// C = A_hi B_hi + A_hi B_lo + A_lo B_hi (the lo x lo term, below 2^-16 of the product, dropped).
// Operand lanes: A row (lane % 32) and B column (lane % 32) over k = 8 (lane / 32) .. +7; D column
// (lane % 32), rows 8 (v / 4) + 4 (lane / 32) + v % 4 of accumulator register v.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split_bf16(float x, uint32_t &hi, uint32_t &lo) {
    hi = bf16_bits(x);
    lo = bf16_bits(x - __uint_as_float(hi << 16));
}

template <class TB>
__global__ void __launch_bounds__(256) gemm_mfma_kernel(GemmArgs<TB> g) {
    __shared__ float As[kBK][kT + 4];
    __shared__ float Bs[kBK][kT + 4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int64_t z = blockIdx.z;
    const float *A = g.A + z * g.a_mz;
    const TB *B = g.B + z * g.b_mz;
    const int kk = (lane >> 5) * 8, rl = lane & 31;
    f32x16 acc = {};
    for (int k0 = 0; k0 < g.K; k0 += kBK) {
        for (int e = tid; e < kT * kBK; e += 256) {
            const int r = e / kBK, k = e - r * kBK;  // A tile row r (m), column k
            const int m = m0 + r, kk1 = k0 + k;
            As[k][r] = (m < g.M && kk1 < g.K) ? A[m * g.a_rs + kk1 * g.a_cs] : 0.0f;
            const int c = e % kT, kb = e / kT;        // B tile row kb (k), column c (n)
            const int n = n0 + c, kk2 = k0 + kb;
            Bs[kb][c] = (n < g.N && kk2 < g.K) ? load_b(B + kk2 * g.b_rs + n * g.b_cs) : 0.0f;
        }
        __syncthreads();
        u32x4 ah, al, bh, bl;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint32_t h0, l0, h1, l1;
            split_bf16(As[kk + 2 * q][wm + rl], h0, l0);
            split_bf16(As[kk + 2 * q + 1][wm + rl], h1, l1);
            ah[q] = h0 | (h1 << 16);
            al[q] = l0 | (l1 << 16);
            split_bf16(Bs[kk + 2 * q][wn + rl], h0, l0);
            split_bf16(Bs[kk + 2 * q + 1][wn + rl], h1, l1);
            bh[q] = h0 | (h1 << 16);
            bl[q] = l0 | (l1 << 16);
        }
        const bf16x8 Ah = __builtin_bit_cast(bf16x8, ah), Al = __builtin_bit_cast(bf16x8, al);
        const bf16x8 Bh = __builtin_bit_cast(bf16x8, bh), Bl = __builtin_bit_cast(bf16x8, bl);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, Bh, acc, 0, 0, 0);
        if constexpr (!std::is_same<TB, uint16_t>::value)  // bf16 weights have no lo part
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bl, acc, 0, 0, 0);
        __syncthreads();
    }
    float *C = g.C + z * g.c_mz;
    const float *bias = g.bias ? g.bias + z * g.bias_mz : nullptr;
    const int n = n0 + wn + rl;
    if (n < g.N) {
        const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int m = m0 + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            if (m < g.M) C[m * g.c_rs + n] = activate(acc[v] + bn, g.act);
        }
    }
}

// dZ = dA * act'(Y) in place; both [rows][cols] per member (dense rows), members da_mz / y_mz apart
__global__ void dact_kernel(float *dA, int64_t da_mz, const float *Y, int64_t y_mz, int rows, int cols, int act) {
    const int64_t z = blockIdx.y;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < static_cast<int64_t>(rows) * cols;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float y = Y[z * y_mz + e];
        const float d = act == ACT_TANH ? 1.0f - y * y : (act == ACT_SIGMOID ? y * (1.0f - y) : 1.0f);
        dA[z * da_mz + e] *= d;
    }
}

// db[z][c] = sum_r dZ[z][r][c]
__global__ void colsum_kernel(const float *dZ, int rows, int cols, int64_t ld, int64_t mz, float *db, int64_t db_mz) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t z = blockIdx.y;
    if (c >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += dZ[z * mz + static_cast<int64_t>(r) * ld + c];
    db[z * db_mz + c] = s;
}

// per member: loss[z] += sum (y - t)^2 / (rows*cols); dY = 2 (y - t) / (rows*cols). Y and dY are
// [rows][cols] per member, y_mz and dy_mz apart; T is shared by the members.
__global__ void loss_kernel(const float *Y, int64_t y_mz, const float *T, int rows, int cols, float *dY, int64_t dy_mz,
                            double *loss) {
    __shared__ double red[4];
    const int64_t z = blockIdx.y;
    const int64_t total = static_cast<int64_t>(rows) * cols;
    const double inv = 1.0 / static_cast<double>(total);
    double acc = 0.0;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float d = Y[z * y_mz + e] - T[e];
        acc += static_cast<double>(d) * d;
        dY[z * dy_mz + e] = static_cast<float>(2.0 * inv) * d;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) s += red[w];
        atomicAdd(loss + z, s * inv);
    }
}

__global__ void adam_kernel(float *p, uint16_t *pbf, const float *g, float *m, float *v, int64_t count, float lr_t,
                            float b1, float b2, float eps, float wd_lr) {
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < count;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.0f - b1) * gi;
        const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        const float pn = p[e] - lr_t * mi / (sqrtf(vi) + eps) - wd_lr * p[e];
        p[e] = pn;
        pbf[e] = bf16_bits(pn);
    }
}

__global__ void to_bf16_kernel(const float *p, uint16_t *pbf, int64_t count) {
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < count;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        pbf[e] = bf16_bits(p[e]);
}

__global__ void gather_kernel(const float *src, const int32_t *idx, int rows, int cols, float *dst) {
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < static_cast<int64_t>(rows) * cols;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t r = e / cols, c = e - r * cols;
        dst[e] = src[static_cast<int64_t>(idx[r]) * cols + c];
    }
}

__global__ void fill_kernel(double *p, int64_t count) {
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < count;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[e] = 0.0;
}

unsigned blocks_for(int64_t work) { return static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>((work + 255) / 256, 1), 1024)); }

template <class TB>
void launch_gemm(const GemmArgs<TB> &g, int members, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>((g.N + kT - 1) / kT), static_cast<unsigned>((g.M + kT - 1) / kT),
                    static_cast<unsigned>(members));
    // QANNEAL_MLP_GEMM=valu: the VALU kernel (FP32 activations times bf16 weights in FMA order)
    static const bool valu = [] {
        const char *e = std::getenv("QANNEAL_MLP_GEMM");
        return e && std::strcmp(e, "valu") == 0;
    }();
    if (valu) hipLaunchKernelGGL(gemm_kernel<TB>, grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL(gemm_mfma_kernel<TB>, grid, dim3(256), 0, s, g);
    QA_HIP(hipGetLastError());
}

int in_dim(const qa_mlp *m, int l) { return l == 0 ? m->n_in : m->width; }
int out_dim(const qa_mlp *m, int l) { return l == m->depth ? m->n_out : m->width; }

// activation buffer of layer l's output for member 0 ([cap][out_l]); members are act_mz apart
float *act_of(qa_mlp *m, int l) { return m->acts.ptr + static_cast<int64_t>(l) * m->cap * m->maxdim; }
int64_t act_mz(const qa_mlp *m) { return static_cast<int64_t>(m->depth + 1) * m->cap * m->maxdim; }

void ensure_cap(qa_mlp *m, int64_t batch) {
    if (batch <= m->cap) return;
    m->cap = std::max<int64_t>(batch, 64);
    m->acts.alloc(static_cast<size_t>(m->members) * (m->depth + 1) * m->cap * m->maxdim);
    m->dA.alloc(static_cast<size_t>(m->members) * m->cap * m->maxdim);
    m->dB.alloc(static_cast<size_t>(m->members) * m->cap * m->maxdim);
    m->xin.alloc(static_cast<size_t>(m->cap) * m->n_in);
    m->tgt.alloc(static_cast<size_t>(m->cap) * m->n_out);
}

// forward of every member on xin[0..B) (normalized inputs already on the device)
void forward_dev(qa_mlp *m, int64_t B) {
    hipStream_t s = m->ctx->stream;
    for (int l = 0; l <= m->depth; ++l) {
        const int K = in_dim(m, l), N = out_dim(m, l);
        GemmArgs<uint16_t> g{};
        if (l == 0) {
            g.A = m->xin.ptr;
            g.a_rs = K;
            g.a_cs = 1;
            g.a_mz = 0;
        } else {
            g.A = act_of(m, l - 1);
            g.a_rs = K;
            g.a_cs = 1;
            g.a_mz = act_mz(m);
        }
        g.B = m->wbf.ptr + m->w_off[l];  // B(k,n) = W[n][k] (bf16)
        g.b_rs = 1;
        g.b_cs = K;
        g.b_mz = m->nparams;
        g.C = act_of(m, l);
        g.c_rs = N;
        g.c_mz = act_mz(m);
        g.bias = m->params.ptr + m->b_off[l];
        g.bias_mz = m->nparams;
        g.M = static_cast<int>(B);
        g.N = N;
        g.K = K;
        g.act = l == m->depth ? ACT_SIGMOID : ACT_TANH;
        launch_gemm(g, m->members, s);
    }
}

// loss and gradients of every member on (xin, tgt)[0..B); loss per member into m->loss
void backward_dev(qa_mlp *m, int64_t B) {
    hipStream_t s = m->ctx->stream;
    const int L = m->depth;
    const int64_t dmz = m->cap * m->maxdim;
    hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(64), 0, s, m->loss.ptr, static_cast<int64_t>(m->members));
    hipLaunchKernelGGL(loss_kernel, dim3(blocks_for(B * m->n_out), static_cast<unsigned>(m->members)), dim3(256), 0, s,
                       act_of(m, L), act_mz(m), m->tgt.ptr, static_cast<int>(B), m->n_out, m->dA.ptr, dmz, m->loss.ptr);
    QA_HIP(hipGetLastError());
    // the gradient w.r.t. each layer's output, [member][B][out_l] with member stride dmz, alternates
    // between the two gradient buffers
    float *dcur = m->dA.ptr, *dnext = m->dB.ptr;
    const int64_t cur_mz = dmz;
    for (int l = L; l >= 0; --l) {
        const int K = in_dim(m, l), N = out_dim(m, l);
        const int act = l == L ? ACT_SIGMOID : ACT_TANH;
        hipLaunchKernelGGL(dact_kernel, dim3(blocks_for(B * N), static_cast<unsigned>(m->members)), dim3(256), 0, s,
                           dcur, cur_mz, act_of(m, l), act_mz(m), static_cast<int>(B), N, act);
        QA_HIP(hipGetLastError());
        // weight gradient dW[n][k] = sum_b dZ[b][n] A_l[b][k]
        GemmArgs<float> gw{};
        gw.A = dcur;  // A(m=n, k=b) = dZ[b][n]
        gw.a_rs = 1;
        gw.a_cs = N;
        gw.a_mz = cur_mz;
        if (l == 0) {
            gw.B = m->xin.ptr;
            gw.b_mz = 0;
        } else {
            gw.B = act_of(m, l - 1);
            gw.b_mz = act_mz(m);
        }
        gw.b_rs = K;  // B(k=b, n=k) = A_l[b][k]
        gw.b_cs = 1;
        gw.C = m->grads.ptr + m->w_off[l];
        gw.c_rs = K;
        gw.c_mz = m->nparams;
        gw.M = N;
        gw.N = K;
        gw.K = static_cast<int>(B);
        gw.act = ACT_NONE;
        launch_gemm(gw, m->members, s);
        hipLaunchKernelGGL(colsum_kernel, dim3(static_cast<unsigned>((N + 255) / 256), static_cast<unsigned>(m->members)),
                           dim3(256), 0, s, dcur, static_cast<int>(B), N, static_cast<int64_t>(N), cur_mz,
                           m->grads.ptr + m->b_off[l], m->nparams);
        QA_HIP(hipGetLastError());
        if (l == 0) break;
        // data gradient dA_l[b][k] = sum_n dZ[b][n] W[n][k]
        GemmArgs<uint16_t> gd{};
        gd.A = dcur;
        gd.a_rs = N;
        gd.a_cs = 1;
        gd.a_mz = cur_mz;
        gd.B = m->wbf.ptr + m->w_off[l];
        gd.b_rs = K;
        gd.b_cs = 1;
        gd.b_mz = m->nparams;
        gd.C = dnext;
        gd.c_rs = K;
        gd.c_mz = dmz;
        gd.M = static_cast<int>(B);
        gd.N = K;
        gd.K = N;
        gd.act = ACT_NONE;
        launch_gemm(gd, m->members, s);
        std::swap(dcur, dnext);
    }
}

void normalize_rows(const qa_mlp *m, const double *x, int64_t B, std::vector<float> &out) {
    out.resize(static_cast<size_t>(B) * m->n_in);
    for (int64_t b = 0; b < B; ++b)
        for (int k = 0; k < m->n_in; ++k)
            out[b * m->n_in + k] = static_cast<float>((x[b * m->n_in + k] - m->x_mean[k]) / m->x_scale[k]);
}

}  // namespace

void mlp_sync_bf16(qa_mlp *m) {
    const int64_t total = static_cast<int64_t>(m->members) * m->nparams;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(total)), dim3(256), 0, m->ctx->stream, m->params.ptr, m->wbf.ptr,
                       total);
    QA_HIP(hipGetLastError());
}

void mlp_init(qa_mlp *m, uint64_t seed) {
    m->w_off.assign(m->depth + 1, 0);
    m->b_off.assign(m->depth + 1, 0);
    int64_t off = 0;
    for (int l = 0; l <= m->depth; ++l) {
        m->w_off[l] = off;
        off += static_cast<int64_t>(in_dim(m, l)) * out_dim(m, l);
        m->b_off[l] = off;
        off += out_dim(m, l);
    }
    m->nparams = off;
    m->maxdim = std::max(m->width, m->n_out);
    std::vector<float> host(static_cast<size_t>(m->members) * off, 0.0f);
    std::mt19937_64 rng(seed);
    for (int z = 0; z < m->members; ++z)
        for (int l = 0; l <= m->depth; ++l) {
            // Glorot-uniform weights, zero biases
            const double a = std::sqrt(6.0 / static_cast<double>(in_dim(m, l) + out_dim(m, l)));
            std::uniform_real_distribution<double> U(-a, a);
            const int64_t nw = static_cast<int64_t>(in_dim(m, l)) * out_dim(m, l);
            for (int64_t e = 0; e < nw; ++e) host[z * off + m->w_off[l] + e] = static_cast<float>(U(rng));
        }
    hipStream_t s = m->ctx->stream;
    m->params.alloc(host.size());
    m->params.upload(host.data(), host.size(), s);
    std::vector<float> zero(host.size(), 0.0f);
    m->grads.alloc(host.size());
    m->m1.alloc(host.size());
    m->m2.alloc(host.size());
    m->grads.upload(zero.data(), zero.size(), s);
    m->m1.upload(zero.data(), zero.size(), s);
    m->m2.upload(zero.data(), zero.size(), s);
    m->loss.alloc(m->members);
    m->wbf.alloc(host.size());
    mlp_sync_bf16(m);
    QA_HIP(hipStreamSynchronize(s));
    m->x_mean.assign(m->n_in, 0.0);
    m->x_scale.assign(m->n_in, 1.0);
    m->x_lo.assign(m->n_in, -1e300);
    m->x_hi.assign(m->n_in, 1e300);
    m->adam_t = 0;
}

void mlp_forward(qa_mlp *m, int64_t B, const double *x, float *y) {
    if (B <= 0) return;
    ensure_cap(m, B);
    std::vector<float> xn;
    normalize_rows(m, x, B, xn);
    hipStream_t s = m->ctx->stream;
    m->xin.upload(xn.data(), xn.size(), s);
    forward_dev(m, B);
    const int L = m->depth;
    for (int z = 0; z < m->members; ++z)
        QA_HIP(hipMemcpyAsync(y + static_cast<int64_t>(z) * B * m->n_out, act_of(m, L) + z * act_mz(m),
                              static_cast<size_t>(B) * m->n_out * sizeof(float), hipMemcpyDeviceToHost, s));
    QA_HIP(hipStreamSynchronize(s));
}

void mlp_loss_grad(qa_mlp *m, int64_t B, const double *x, const float *t, double *loss, float *grad) {
    QA_REQUIRE(B > 0, "batch must be > 0");
    ensure_cap(m, B);
    std::vector<float> xn;
    normalize_rows(m, x, B, xn);
    hipStream_t s = m->ctx->stream;
    m->xin.upload(xn.data(), xn.size(), s);
    m->tgt.upload(t, static_cast<size_t>(B) * m->n_out, s);
    forward_dev(m, B);
    backward_dev(m, B);
    if (loss) m->loss.download(loss, m->members, s);
    if (grad) m->grads.download(grad, static_cast<size_t>(m->members) * m->nparams, s);
    QA_HIP(hipStreamSynchronize(s));
}

void mlp_train(qa_mlp *m, int64_t S, const double *x, const float *t, const qa_mlp_train_params *tp, double *history) {
    QA_REQUIRE(S >= 2, "training needs at least two samples");
    QA_REQUIRE(tp->epochs > 0 && tp->batch > 0, "epochs and batch must be > 0");
    hipStream_t s = m->ctx->stream;
    // input normalization from the training set; the training range (in normalized units)
    // bounds what the surrogate is trusted on
    for (int k = 0; k < m->n_in; ++k) {
        double mu = 0.0, sq = 0.0;
        for (int64_t i = 0; i < S; ++i) mu += x[i * m->n_in + k];
        mu /= static_cast<double>(S);
        for (int64_t i = 0; i < S; ++i) sq += (x[i * m->n_in + k] - mu) * (x[i * m->n_in + k] - mu);
        const double sd = std::sqrt(sq / static_cast<double>(S));
        m->x_mean[k] = mu;
        m->x_scale[k] = sd > 1e-9 ? sd : 1.0;
        double lo = 1e300, hi = -1e300;
        for (int64_t i = 0; i < S; ++i) {
            const double v = (x[i * m->n_in + k] - mu) / m->x_scale[k];
            lo = std::min(lo, v);
            hi = std::max(hi, v);
        }
        m->x_lo[k] = lo;
        m->x_hi[k] = hi;
    }
    std::vector<float> xn;
    normalize_rows(m, x, S, xn);
    DevBuf<float> dX(xn.size()), dT(static_cast<size_t>(S) * m->n_out);
    dX.upload(xn.data(), xn.size(), s);
    dT.upload(t, static_cast<size_t>(S) * m->n_out, s);
    const int64_t B = std::min<int64_t>(tp->batch, S);
    ensure_cap(m, B);
    DevBuf<int32_t> didx(static_cast<size_t>(S));
    std::vector<int32_t> idx(static_cast<size_t>(S));
    std::iota(idx.begin(), idx.end(), 0);
    std::mt19937_64 rng(tp->seed);
    const int64_t total = static_cast<int64_t>(m->members) * m->nparams;
    const double b1 = tp->beta1 > 0 ? tp->beta1 : 0.9, b2 = tp->beta2 > 0 ? tp->beta2 : 0.999;
    const double eps = tp->eps > 0 ? tp->eps : 1e-8;
    std::vector<double> lz(m->members);
    for (int64_t ep = 0; ep < tp->epochs; ++ep) {
        std::shuffle(idx.begin(), idx.end(), rng);
        didx.upload(idx.data(), idx.size(), s);
        double ep_loss = 0.0;
        int64_t nb = 0;
        for (int64_t b0 = 0; b0 + B <= S; b0 += B) {
            hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(B * m->n_in)), dim3(256), 0, s, dX.ptr, didx.ptr + b0,
                               static_cast<int>(B), m->n_in, m->xin.ptr);
            hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(B * m->n_out)), dim3(256), 0, s, dT.ptr, didx.ptr + b0,
                               static_cast<int>(B), m->n_out, m->tgt.ptr);
            QA_HIP(hipGetLastError());
            forward_dev(m, B);
            backward_dev(m, B);
            ++m->adam_t;
            const double lr_t = tp->lr * std::sqrt(1.0 - std::pow(b2, static_cast<double>(m->adam_t))) /
                                (1.0 - std::pow(b1, static_cast<double>(m->adam_t)));
            hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(total)), dim3(256), 0, s, m->params.ptr, m->wbf.ptr, m->grads.ptr,
                               m->m1.ptr, m->m2.ptr, total, static_cast<float>(lr_t), static_cast<float>(b1),
                               static_cast<float>(b2), static_cast<float>(eps), static_cast<float>(tp->lr * tp->weight_decay));
            QA_HIP(hipGetLastError());
            m->loss.download(lz.data(), m->members, s);
            QA_HIP(hipStreamSynchronize(s));
            for (double v : lz) ep_loss += v / static_cast<double>(m->members);
            ++nb;
        }
        if (history) history[ep] = nb ? ep_loss / static_cast<double>(nb) : 0.0;
    }
    m->trained = true;
}

}  // namespace qa
