// Local fields on the matrix cores, and the FP64 drift check of running FP32 fields.
//
// f[m][j] = h[j] + sum_k J[j][k] s[m][k] for M states at once (DenseIsing::compute_local_fields,
// dense_ising.cpp:64-83, whose J_jj term is dropped: the FP32 couplings keep a cleared
// diagonal). The FP32 couplings are split into two bf16 matrices, J_hi = bf16(J) and
// J_lo = bf16(J - J_hi), both rounded to nearest even; spins are exact in bf16. Each 32x32 tile of
// F accumulates S J_hi^T + S J_lo^T in FP32 with v_mfma_f32_32x32x16_bf16 (gfx950), so the
// coupling error is |J - J_hi - J_lo| <= 2^-17 |J| per term; couplings with at most 16 significant
// bits (QUBO-derived integer or half-integer J, qubo.cpp:33-69) are carried exactly and their
// fields come out exact.
//
// Layouts. S: int8 [M][n]. J_hi / J_lo: uint16 bf16 bits [rows][ldk], rows = n rounded up to 64,
// ldk = n rounded up to 16, zero padded. Because J is symmetric on this path, the B operand of
// column j is row j of J, read along k. Operand lanes of the 32x32x16 instruction: A row
// (lane % 32), k = 8 (lane / 32) .. +7; B column (lane % 32), same k; D column (lane % 32), rows
// 8 (v / 4) + 4 (lane / 32) + v % 4 for accumulator register v.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qa_internal.hpp"

namespace qa {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kBf16One = 0x3F80u, kBf16MinusOne = 0xBF80u;

// 64 x 64 tile of F per workgroup of four waves (2 x 2 waves of 32 x 32).
__global__ void __launch_bounds__(256) fields_mfma_kernel(const int8_t *__restrict__ S, int64_t M, int n,
                                                          const uint16_t *__restrict__ Jhi,
                                                          const uint16_t *__restrict__ Jlo, int ldk,
                                                          const float *__restrict__ h, float *__restrict__ F,
                                                          int64_t ldf) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * 64 + (wave >> 1) * 32;
    const int j0 = static_cast<int>(blockIdx.y) * 64 + (wave & 1) * 32;
    const int64_t ia = i0 + (lane & 31);
    const int jb = j0 + (lane & 31);
    const int kk = (lane >> 5) * 8;
    const int8_t *srow = S + (ia < M ? ia : 0) * static_cast<int64_t>(n);
    const uint16_t *hrow = Jhi + static_cast<int64_t>(jb) * ldk;
    const uint16_t *lrow = Jlo + static_cast<int64_t>(jb) * ldk;
    f32x16 acc = {};
    for (int k0 = 0; k0 < n; k0 += 16) {
        u32x4 a;
#pragma unroll
        for (int e2 = 0; e2 < 4; ++e2) {
            uint32_t w = 0;
#pragma unroll
            for (int h2 = 0; h2 < 2; ++h2) {
                const int k = k0 + kk + 2 * e2 + h2;
                const int s = (ia < M && k < n) ? srow[k] : 0;
                const uint32_t b = s > 0 ? kBf16One : (s < 0 ? kBf16MinusOne : 0u);
                w |= b << (16 * h2);
            }
            a[e2] = w;
        }
        const u32x4 bh = *reinterpret_cast<const u32x4 *>(hrow + k0 + kk);
        const u32x4 bl = *reinterpret_cast<const u32x4 *>(lrow + k0 + kk);
        const bf16x8 av = __builtin_bit_cast(bf16x8, a);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, bh), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, bl), acc, 0, 0, 0);
    }
    if (jb < n) {
        const float hj = h[jb];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int64_t i = i0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            if (i < M) F[i * ldf + jb] = acc[v] + hj;
        }
    }
}

// FP64 fields of `count` states for a symmetric dense J, each element in the reference's order
// (DenseIsing::compute_local_fields, dense_ising.cpp:64-83): h_i, then + J_ij s_j for j = 0 .. n-1,
// then - J_ii s_i. Every product is exact (s = +-1), so each element is the same sequence of FP64
// additions as the one-state-per-workgroup fields_kernel, bit for bit. Lane = element i of a block
// of 64; with J symmetric, term j reads J[j][i], so a wave's loads of one j are contiguous, and
// kFieldsCG states share each load; their spins are staged in LDS as +-1.0, kFieldsJT rows at a time.
constexpr int kFieldsCG = 8, kFieldsJT = 256;
__global__ void __launch_bounds__(64) fields64_sym_kernel(const double *__restrict__ J64,
                                                          const double *__restrict__ h64,
                                                          const int8_t *__restrict__ S, int n, int64_t count,
                                                          double *__restrict__ out) {
    __shared__ double sd[kFieldsJT][kFieldsCG];
    const int lane = threadIdx.x;
    const int i = static_cast<int>(blockIdx.x) * 64 + lane;
    const int iv = i < n ? i : n - 1;  // loads stay in the matrix; stores are guarded
    const int64_t c0 = static_cast<int64_t>(blockIdx.y) * kFieldsCG;
    const int nc = static_cast<int>(min<int64_t>(kFieldsCG, count - c0));
    double acc[kFieldsCG];
    const double h = h64[iv];
#pragma unroll
    for (int c = 0; c < kFieldsCG; ++c) acc[c] = h;
    for (int jb = 0; jb < n; jb += kFieldsJT) {
        const int jn = min(kFieldsJT, n - jb);
        __syncthreads();
        for (int e = lane; e < kFieldsJT * kFieldsCG; e += 64) {
            const int jj = e / kFieldsCG, c = e - jj * kFieldsCG;
            sd[jj][c] = (jj < jn && c < nc) ? static_cast<double>(S[(c0 + c) * n + jb + jj]) : 0.0;
        }
        __syncthreads();
        const double *col = J64 + static_cast<int64_t>(jb) * n + iv;
#pragma unroll 8
        for (int jj = 0; jj < jn; ++jj) {
            const double v = col[static_cast<int64_t>(jj) * n];
#pragma unroll
            for (int c = 0; c < kFieldsCG; ++c) acc[c] += v * sd[jj][c];
        }
    }
    if (i >= n) return;
    const double d = J64[static_cast<int64_t>(i) * n + i];
    for (int c = 0; c < nc; ++c) {
        const double si = static_cast<double>(S[(c0 + c) * n + i]);
        out[(c0 + c) * n + i] = acc[c] - d * si;
    }
}

// max |F[m][j] - F64[m][j]| over the sampled states, as the bits of a non-negative double
// (integer order == value order), folded into *out.
__global__ void drift_kernel(const float *F, int64_t ldf, const double *F64, int n, int64_t count,
                             unsigned long long *out) {
    double mx = 0.0;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < count * n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t m = e / n, j = e - m * n;
        mx = fmax(mx, fabs(static_cast<double>(F[m * ldf + j]) - F64[e]));
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(mx)));
}

__global__ void f64_to_f32_kernel(const double *in, float *out, int64_t count) {
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < count;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x)
        out[e] = static_cast<float>(in[e]);
}

uint16_t bf16_rne(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

float bf16_value(uint16_t b) {
    const uint32_t u = static_cast<uint32_t>(b) << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

}  // namespace

void build_bf16_split(qa_problem *p, const float *J32, hipStream_t s) {
    const int64_t n = p->n;
    const int64_t rows = (n + 63) / 64 * 64, ldk = (n + 15) / 16 * 16;
    std::vector<uint16_t> hi(static_cast<size_t>(rows * ldk), 0), lo(static_cast<size_t>(rows * ldk), 0);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = 0; k < n; ++k) {
            const float v = J32[j * n + k];
            const uint16_t bh = bf16_rne(v);
            hi[j * ldk + k] = bh;
            lo[j * ldk + k] = bf16_rne(v - bf16_value(bh));
        }
    p->Jbf_hi.alloc(rows * ldk);
    p->Jbf_lo.alloc(rows * ldk);
    p->Jbf_hi.upload(hi.data(), rows * ldk, s);
    p->Jbf_lo.upload(lo.data(), rows * ldk, s);
    QA_HIP(hipStreamSynchronize(s));  // the host copies go out of scope here
    p->ldk_bf = static_cast<int>(ldk);
}

void fields_fp32_launch(const qa_problem *p, const int8_t *d_spins, int64_t count, float *d_fields, int64_t ldf,
                        double *d_scratch, hipStream_t s) {
    const int64_t n = p->n;
    if (count == 0) return;
    if (p->kind == QA_KIND_DENSE && p->symmetric) {
        QA_REQUIRE(p->Jbf_hi.ptr && p->Jbf_lo.ptr, "bf16 couplings missing");
        const dim3 grid(static_cast<unsigned>((count + 63) / 64), static_cast<unsigned>((n + 63) / 64));
        hipLaunchKernelGGL(fields_mfma_kernel, grid, dim3(256), 0, s, d_spins, count, static_cast<int>(n),
                           p->Jbf_hi.ptr, p->Jbf_lo.ptr, p->ldk_bf, p->h32.ptr, d_fields, ldf);
        QA_HIP(hipGetLastError());
    } else {
        // sparse / HUBO / asymmetric dense: FP64 fields in the reference order, rounded once
        QA_REQUIRE(ldf == n && d_scratch, "fields of non-dense problems are written densely via FP64 scratch");
        fields_fp64_launch(p, d_spins, d_scratch, count, s);
        hipLaunchKernelGGL(f64_to_f32_kernel, dim3(256), dim3(256), 0, s, d_scratch, d_fields, count * n);
        QA_HIP(hipGetLastError());
    }
}

void field_drift_launch(const qa_problem *p, const int8_t *d_spins, const float *d_fields, int64_t ldf, int64_t count,
                        double *d_scratch, unsigned long long *d_max, hipStream_t s) {
    const int64_t n = p->n;
    if (count <= 0) return;
    // QANNEAL_DRIFT_FIELDS=reference: the one-state-per-workgroup kernel (same values bit for bit;
    // a test switch)
    const char *env = std::getenv("QANNEAL_DRIFT_FIELDS");
    if (p->kind == QA_KIND_DENSE && p->symmetric && !(env && std::strcmp(env, "reference") == 0)) {
        const dim3 grid(static_cast<unsigned>((n + 63) / 64), static_cast<unsigned>((count + kFieldsCG - 1) / kFieldsCG));
        hipLaunchKernelGGL(fields64_sym_kernel, grid, dim3(64), 0, s, p->J64.ptr, p->h64.ptr, d_spins,
                           static_cast<int>(n), count, d_scratch);
        QA_HIP(hipGetLastError());
    } else {
        fields_fp64_launch(p, d_spins, d_scratch, count, s);
    }
    hipLaunchKernelGGL(drift_kernel, dim3(static_cast<unsigned>(std::min<int64_t>((count * n + 255) / 256, 1024))),
                       dim3(256), 0, s, d_fields, ldf, d_scratch, static_cast<int>(n), count, d_max);
    QA_HIP(hipGetLastError());
}

double drift_value(unsigned long long bits) {
    double d;
    std::memcpy(&d, &bits, 8);
    return d;
}

}  // namespace qa
