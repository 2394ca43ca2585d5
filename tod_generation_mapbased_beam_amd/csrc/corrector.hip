// Learned interpolation corrector: features, BF16-MFMA training kernels
// and the C-ABI that drives them (the model is defined in corrector.hpp).
//
// Matrix products run on MFMA in BF16 (v_mfma_f32_16x16x16_bf16: operands
// rounded to bfloat16, FP32 accumulation); biases, GELU, the loss and the
// Adam update are FP32 on the FP32 master weights.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include <string>

#include "healpix_ring.hpp"
#include "internal.hpp"
#include "corrector.hpp"
#include "../../include/todgpu.h"

using namespace todg;

namespace todg_corr {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

static_assert(kNParams == TODG_CORR_NPARAMS, "parameter count");
static_assert(todg_smlp::kNParams == TODG_SMLP_NPARAMS, "stencil MLP parameter count");

// C (M x N, row-major, ldc) = alpha * op(A) op(B) + beta * C with
// op(A)[i][k] = ta ? A[k*lda + i] : A[i*lda + k] (M x K) and
// op(B)[k][j] = tb ? B[j*ldb + k] : B[k*ldb + j] (K x N).  One wave per
// 16 x 16 tile of C, K in steps of 16 (zero padded).  MFMA 16x16x16 operand
// layout: lane l holds A[l%16][4(l/16) .. +3], B[4(l/16) .. +3][l%16], and
// C[4(l/16) .. +3][l%16].
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const float* __restrict__ A,
                                                   int lda, int ta, const float* __restrict__ B,
                                                   int ldb, int tb, float* __restrict__ C, int ldc,
                                                   float alpha, float beta) {
    const int lane = threadIdx.x & 63;
    const int tiles_n = (N + 15) / 16;
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t tiles = (int64_t)((M + 15) / 16) * tiles_n;
    if (tile >= tiles) return;
    const int ti = (int)(tile / tiles_n), tj = (int)(tile % tiles_n);
    const int ia = ti * 16 + (lane & 15);
    const int jb = tj * 16 + (lane & 15);
    const int kq = (lane >> 4) * 4;
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        short4_t av, bv;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + kq + q;
            float x = 0.f, y = 0.f;
            if (ia < M && k < K) x = ta ? A[(int64_t)k * lda + ia] : A[(int64_t)ia * lda + k];
            if (jb < N && k < K) y = tb ? B[(int64_t)jb * ldb + k] : B[(int64_t)k * ldb + jb];
            av[q] = bf16_bits(x);
            bv[q] = bf16_bits(y);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = ti * 16 + kq + q;
        if (i < M && jb < N) {
            float* c = C + (int64_t)i * ldc + jb;
            *c = beta == 0.f ? alpha * acc[q] : fmaf(alpha, acc[q], beta * *c);
        }
    }
}

// Split-K form of k_gemm_bf16 for products with few output tiles and a long
// K (the weight gradients, K = batch rows): wave w takes tile w % tiles and
// the K range [s kc, (s + 1) kc) of split s = w / tiles, and writes its
// 16 x 16 partial tile (MFMA layout: lane l, element q) to
// P[(s tiles + tile) 256 + 4 l + q]; k_splitk_reduce adds the splits in
// order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_gemm_bf16_splitk(int M, int N, int K, int kc, int splits,
                                                          const float* __restrict__ A, int lda, int ta,
                                                          const float* __restrict__ B, int ldb, int tb,
                                                          float* __restrict__ P) {
    const int lane = threadIdx.x & 63;
    const int tiles_n = (N + 15) / 16;
    const int tiles = ((M + 15) / 16) * tiles_n;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= (int64_t)tiles * splits) return;
    const int tile = (int)(w % tiles), sp = (int)(w / tiles);
    const int ti = tile / tiles_n, tj = tile % tiles_n;
    const int ia = ti * 16 + (lane & 15);
    const int jb = tj * 16 + (lane & 15);
    const int kq = (lane >> 4) * 4;
    const int kbeg = sp * kc;
    const int kend = kbeg + kc < K ? kbeg + kc : K;
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        short4_t av, bv;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + kq + q;
            float x = 0.f, y = 0.f;
            if (ia < M && k < kend) x = ta ? A[(int64_t)k * lda + ia] : A[(int64_t)ia * lda + k];
            if (jb < N && k < kend) y = tb ? B[(int64_t)jb * ldb + k] : B[(int64_t)k * ldb + jb];
            av[q] = bf16_bits(x);
            bv[q] = bf16_bits(y);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    float* o = P + ((int64_t)sp * tiles + tile) * 256 + 4 * lane;
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = acc[q];
}

// C[i][j] = alpha * sum_s P[s][tile(i, j)][lane(i, j), q(i, j)] + beta * C[i][j],
// splits added in order.
__global__ __launch_bounds__(256) void k_splitk_reduce(int M, int N, int splits,
                                                       const float* __restrict__ P, float* __restrict__ C,
                                                       int ldc, float alpha, float beta) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)M * N) return;
    const int i = (int)(e / N), j = (int)(e % N);
    const int tiles_n = (N + 15) / 16;
    const int tiles = ((M + 15) / 16) * tiles_n;
    const int tile = (i / 16) * tiles_n + j / 16;
    const int r = i % 16, c = j % 16;
    const int off = 4 * ((r / 4) * 16 + c) + r % 4;  // C[4(l/16) + q][l%16]
    float sum = 0.f;
    for (int sp = 0; sp < splits; ++sp) sum += P[((int64_t)sp * tiles + tile) * 256 + off];
    float* out = C + (int64_t)i * ldc + j;
    *out = beta == 0.f ? alpha * sum : fmaf(alpha, sum, beta * *out);
}

// Z[i][f] += b[f]; A = gelu(Z) (Z kept for the backward pass).
__global__ __launch_bounds__(256) void k_bias_gelu(int64_t rows, int F, const float* __restrict__ b,
                                                   float* __restrict__ Z, float* __restrict__ A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * F) return;
    const float z = Z[i] + b[i % F];
    Z[i] = z;
    A[i] = gelu(z);
}

// O[i][f] += b[f].
__global__ __launch_bounds__(256) void k_bias(int64_t rows, int F, const float* __restrict__ b,
                                              float* __restrict__ O) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * F) return;
    O[i] += b[i % F];
}

// dZ = dA * gelu'(Z) (in place on dA).
__global__ __launch_bounds__(256) void k_gelu_bwd(int64_t n, const float* __restrict__ Z,
                                                  float* __restrict__ dA) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dA[i] *= gelu_grad(Z[i]);
}

// Head: corrected weights, normalised prediction and residual per row;
// dO = (2 / N) r (m_hat - mean(m_hat)); per-block partial sums of r^2.
__global__ __launch_bounds__(256) void k_head(int64_t N, const float* __restrict__ X,
                                              const float* __restrict__ wbil,
                                              const float* __restrict__ that,
                                              const float* __restrict__ O, float* __restrict__ dO,
                                              float* __restrict__ pred, double* __restrict__ part) {
    __shared__ double red[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0.0;
    if (i < N) {
        const float* x = X + i * kIn + kGeo;
        const float* o = O + i * kOut;
        float om = 0.f, mm = 0.f;
#pragma unroll
        for (int k = 0; k < kOut; ++k) {
            om += o[k];
            mm += x[k];
        }
        om *= 1.f / kOut;
        mm *= 1.f / kOut;
        const float* wb = wbil + i * 4;
        float p = wb[0] * x[5] + wb[1] * x[6] + wb[2] * x[9] + wb[3] * x[10];
#pragma unroll
        for (int k = 0; k < kOut; ++k) p = fmaf(o[k] - om, x[k], p);
        if (pred) pred[i] = p;
        const float r = p - that[i];
        r2 = (double)r * (double)r;
        if (dO) {
            const float g = 2.f * r / (float)N;
#pragma unroll
            for (int k = 0; k < kOut; ++k) dO[i * kOut + k] = g * (x[k] - mm);
        }
    }
    red[threadIdx.x] = r2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && part) part[blockIdx.x] = red[0];
}

// out[f] = sum_i D[i][f] (bias gradients) for few rows: one block per
// column, 256 threads striding over the rows, a tree reduction.
__global__ __launch_bounds__(256) void k_colsum(int64_t rows, int F, const float* __restrict__ D,
                                                float* __restrict__ out) {
    __shared__ float red[256];
    const int f = blockIdx.x;
    float s = 0.f;
    for (int64_t i = threadIdx.x; i < rows; i += blockDim.x) s += D[i * F + f];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[f] = red[0];
}

// Column sums for many rows (the stencil MLP's millions of map pixels),
// F dividing 256, in two passes with a fixed order.  Pass 1:
// block b takes the rows [b chunk, (b + 1) chunk); thread t sums column
// t % F over every (256 / F)-th of them, so a wave's loads cover whole rows
// (coalesced); the block's 256 / F partial sums per column are added in
// order into part[b F + f].  Pass 2: one thread per column adds the blocks'
// sums in order.
__global__ __launch_bounds__(256) void k_colsum_part(int64_t rows, int F, int64_t chunk,
                                                     const float* __restrict__ D, float* __restrict__ part) {
    __shared__ float red[256];
    const int G = 256 / F;
    const int f = threadIdx.x % F, g = threadIdx.x / F;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = r0 + chunk < rows ? r0 + chunk : rows;
    float s = 0.f;
    for (int64_t i = r0 + g; i < r1; i += G) s += D[i * F + f];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < F) {
        float t = 0.f;
        for (int k = 0; k < G; ++k) t += red[k * F + threadIdx.x];
        part[(int64_t)blockIdx.x * F + threadIdx.x] = t;
    }
}

__global__ __launch_bounds__(64) void k_colsum_final(int nb, int F, const float* __restrict__ part,
                                                     float* __restrict__ out) {
    const int f = threadIdx.x;
    if (f >= F) return;
    float t = 0.f;
    for (int b = 0; b < nb; ++b) t += part[(int64_t)b * F + f];
    out[f] = t;
}

__global__ __launch_bounds__(256) void k_sum_parts(int n, const double* __restrict__ part,
                                                   double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Adam on the FP32 master weights (bias-corrected, step t >= 1).
__global__ __launch_bounds__(256) void k_adam(int n, float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Features of N query directions (z, phi) on one map field: the bilinear
// cell (numba_healpy.py:443-480: rings ira, ira+1, slots ia, ib, fractions
// fa, fb, weight v of the lower ring), the fractional slot coordinates fm, fp
// of the query in rings ira-1 and ira+2, and the 4 x 4 neighbourhood rings
// ira-1 .. ira+2 x slots j-1 .. j+2 (j = the query's slot in that ring).
// X (N x kIn), wbil (N x 4: weights of the cells 5, 6, 9, 10), mu, sigma,
// pix (N x 16) and ok (0 where the neighbourhood leaves the normal branch:
// the corrector is not applied there).
__global__ __launch_bounds__(256) void k_features(int nside, const RingEntry* __restrict__ rt,
                                                  const float* __restrict__ field, int64_t N,
                                                  const double* __restrict__ z,
                                                  const double* __restrict__ phi,
                                                  float* __restrict__ X, float* __restrict__ wbil,
                                                  float* __restrict__ mu, float* __restrict__ sigma,
                                                  int* __restrict__ pix, int* __restrict__ ok) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double zz = fmax(-1.0, fmin(1.0, z[i]));
    const double pp = phi[i];
    const int ira = ring_above(nside, zz);
    const bool good = ira >= 2 && ira <= 4 * nside - 3;
    int p16[16];
    double fa = 0.0, v = 0.0, fb = 0.0, fm = 0.0, fp = 0.0;
    if (good) {
        Nb o;
        ring_interp(nside, rt, zz, pp, o);
        const double wa = o.w[0] + o.w[1];
        v = o.w[2] + o.w[3];
        fa = wa > 0.0 ? o.w[1] / wa : 0.0;
        fb = v > 0.0 ? o.w[3] / v : 0.0;
        for (int r = 0; r < 4; ++r) {
            const RingEntry& e = rt[ira - 1 + r];
            int j;
            if (r == 1)
                j = o.pos[0];
            else if (r == 2)
                j = o.pos[2];
            else {
                double fr;
                j = ring_slot(pp, e.phi0, e.dphi, e.inv_dphi, e.n, fr);
                if (r == 0) fm = fr;
                else fp = fr;
            }
            for (int s = 0; s < 4; ++s) {
                int q = j - 1 + s;
                q = ((q % e.n) + e.n) % e.n;
                p16[4 * r + s] = (int)e.first + q;
            }
        }
    } else {
        for (int k = 0; k < 16; ++k) p16[k] = 0;
    }
    float m[16];
    double s1 = 0.0;
    for (int k = 0; k < 16; ++k) {
        TODG_CHECK((uint64_t)p16[k] < 12ull * (uint64_t)nside * (uint64_t)nside, kMapPixel);
        m[k] = field[p16[k]];
        s1 += m[k];
    }
    const double mean = s1 / 16.0;
    double s2 = 0.0;
    for (int k = 0; k < 16; ++k) s2 += ((double)m[k] - mean) * ((double)m[k] - mean);
    const double sd = sqrt(s2 / 16.0);
    const double inv = sd > 1e-30 ? 1.0 / sd : 0.0;
    float* x = X + i * kIn;
    x[0] = (float)(fa - 0.5);
    x[1] = (float)(v - 0.5);
    x[2] = (float)(fb - 0.5);
    x[3] = (float)(fm - 0.5);
    x[4] = (float)(fp - 0.5);
    for (int k = 0; k < 16; ++k) x[kGeo + k] = (float)(((double)m[k] - mean) * inv);
    wbil[i * 4 + 0] = (float)((1.0 - v) * (1.0 - fa));
    wbil[i * 4 + 1] = (float)((1.0 - v) * fa);
    wbil[i * 4 + 2] = (float)(v * (1.0 - fb));
    wbil[i * 4 + 3] = (float)(v * fb);
    mu[i] = (float)mean;
    sigma[i] = (float)sd;
    for (int k = 0; k < 16; ++k) pix[i * 16 + k] = p16[k];
    ok[i] = good ? 1 : 0;
}

// ------------------------------------------------- MLP stencil corrector
// (the model is defined in corrector.hpp, namespace todg_smlp)

// Channels from the MLP outputs O (N x 32) and the features F (N x 16):
// A0 = F_0 + sum_{j>=1} O_j F_j, A1 = sum_{j>=1} O_{16+j} F_j.
__global__ __launch_bounds__(256) void k_smlp_head(int64_t N, const float* __restrict__ F,
                                                   const float* __restrict__ O, float* __restrict__ A0,
                                                   float* __restrict__ A1) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float* f = F + i * 16;
    const float* o = O + i * todg_smlp::kOut;
    double s0 = (double)f[0], s1 = 0.0;
#pragma unroll
    for (int j = 1; j < 16; ++j) {
        s0 = fma((double)o[j], (double)f[j], s0);
        s1 = fma((double)o[16 + j], (double)f[j], s1);
    }
    A0[i] = (float)s0;
    A1[i] = (float)s1;
}

// Backward of k_smlp_head: dO_j = G0 F_j, dO_{16+j} = G1 F_j (j >= 1); the
// unused outputs 0 and 16 get 0.
__global__ __launch_bounds__(256) void k_smlp_head_bwd(int64_t N, const float* __restrict__ F,
                                                       const float* __restrict__ G0,
                                                       const float* __restrict__ G1, float* __restrict__ dO) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float* f = F + i * 16;
    float* d = dO + i * todg_smlp::kOut;
    const float g0 = G0[i], g1 = G1[i];
    d[0] = 0.f;
    d[16] = 0.f;
#pragma unroll
    for (int j = 1; j < 16; ++j) {
        d[j] = g0 * f[j];
        d[16 + j] = g1 * f[j];
    }
}

// Rows of a two-channel sparse operator (CSR, int64 row pointers, int32
// columns, binary32 values): y_r = sum_e v0_e x0[c_e] + v1_e x1[c_e] in row
// order (binary64); e_r = y_r - target_r (target may be null); resid_r =
// scale e_r (binary32, may be null); part[block] = sum of e_r^2 over the
// block (may be null); y (binary64, may be null).
__global__ __launch_bounds__(256) void k_csr_rows2(int64_t nrows, const int64_t* __restrict__ rp,
                                                   const int* __restrict__ col, const float* __restrict__ v0,
                                                   const float* __restrict__ v1, const float* __restrict__ x0,
                                                   const float* __restrict__ x1, const double* __restrict__ target,
                                                   double scale, double* __restrict__ y,
                                                   float* __restrict__ resid, double* __restrict__ part) {
    __shared__ double red[256];
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (r < nrows) {
        double acc = 0.0;
        for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
            const int c = col[e];
            acc = fma((double)v0[e], (double)x0[c], acc);
            acc = fma((double)v1[e], (double)x1[c], acc);
        }
        if (y) y[r] = acc;
        const double d = target ? acc - target[r] : acc;
        if (resid) resid[r] = (float)(scale * d);
        e2 = d * d;
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0 && part) part[blockIdx.x] = red[0];
}

// The transpose's rows (the operator stored column-major as CSR): out0_p =
// sum_e v0_e r[c_e], out1_p = sum_e v1_e r[c_e] in row order (binary64,
// stored binary32).
__global__ __launch_bounds__(256) void k_csr_adjoint2(int64_t nrows, const int64_t* __restrict__ rp,
                                                      const int* __restrict__ col, const float* __restrict__ v0,
                                                      const float* __restrict__ v1, const float* __restrict__ rv,
                                                      float* __restrict__ out0, float* __restrict__ out1) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nrows) return;
    double a0 = 0.0, a1 = 0.0;
    for (int64_t e = rp[p]; e < rp[p + 1]; ++e) {
        const double r = (double)rv[col[e]];
        a0 = fma((double)v0[e], r, a0);
        a1 = fma((double)v1[e], r, a1);
    }
    out0[p] = (float)a0;
    out1[p] = (float)a1;
}

int fail(int code, const std::string& msg) { return todg_internal::set_error(code, msg); }

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

// K per split of the split-K products (256, or more for long K so that there
// are at most ~256 splits: the partials then stay small and their ordered
// reduction cheap), and the largest tile count they take (the 64 x 64
// weight gradient).
constexpr int kSplitKc = 256;
constexpr int kSplitMax = 256;
constexpr int kSplitMaxTiles = (kH / 16) * (kH / 16);

inline int64_t split_kc(int64_t K) {
    const int64_t kc = (K + kSplitMax - 1) / kSplitMax;
    return kc <= kSplitKc ? kSplitKc : (kc + 15) / 16 * 16;
}
inline int64_t n_splits(int64_t K) { return (K + split_kc(K) - 1) / split_kc(K); }

// Column-sum workspace (floats) and the column sums themselves.
constexpr int kColsumBlocks = 512;
constexpr int64_t kColsumParts = (int64_t)kColsumBlocks * 64;

void colsum(hipStream_t st, int64_t rows, int F, const float* D, float* out, float* part) {
    if (rows <= 65536) {  // the per-column blocks are latency-bound beyond a few hundred rows per thread
        hipLaunchKernelGGL(k_colsum, dim3((unsigned)F), dim3(256), 0, st, rows, F, D, out);
        return;
    }
    const int64_t chunk0 = (rows + kColsumBlocks - 1) / kColsumBlocks;
    const int64_t chunk = chunk0 < 1024 ? 1024 : chunk0;
    const int nb = (int)((rows + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)nb), dim3(256), 0, st, rows, F, chunk, D, part);
    hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(64), 0, st, nb, F, part, out);
}

// part: workspace for the split-K partials (n_splits(K) * tiles * 256
// floats), or null for the one-wave-per-tile kernel.
void gemm(hipStream_t st, int M, int N, int K, const float* A, int lda, int ta, const float* B,
          int ldb, int tb, float* C, int ldc, float alpha, float beta, float* part = nullptr) {
    const int64_t tiles = (int64_t)((M + 15) / 16) * ((N + 15) / 16);
    if (part && K >= 4 * kSplitKc && tiles <= kSplitMaxTiles) {
        const int splits = (int)n_splits(K);
        const int64_t waves = tiles * splits;
        hipLaunchKernelGGL(k_gemm_bf16_splitk, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, M, N, K,
                           (int)split_kc(K), splits, A, lda, ta, B, ldb, tb, part);
        hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)(((int64_t)M * N + 255) / 256)), dim3(256), 0, st,
                           M, N, splits, part, C, ldc, alpha, beta);
        return;
    }
    hipLaunchKernelGGL(k_gemm_bf16, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, st, M, N, K,
                       A, lda, ta, B, ldb, tb, C, ldc, alpha, beta);
}

// Forward pass into the workspace: Z1, A1 (N x kH), Z2, A2 (N x kH), O (N x kOut).
void forward(hipStream_t st, const float* P, const float* X, int64_t N, float* Z1, float* A1,
             float* Z2, float* A2, float* O) {
    const int n = (int)N;
    gemm(st, n, kH, kIn, X, kIn, 0, P + kOffW1, kIn, 1, Z1, kH, 1.f, 0.f);
    hipLaunchKernelGGL(k_bias_gelu, dim3(nblk(N * kH)), dim3(256), 0, st, N, kH, P + kOffB1, Z1, A1);
    gemm(st, n, kH, kH, A1, kH, 0, P + kOffW2, kH, 1, Z2, kH, 1.f, 0.f);
    hipLaunchKernelGGL(k_bias_gelu, dim3(nblk(N * kH)), dim3(256), 0, st, N, kH, P + kOffB2, Z2, A2);
    gemm(st, n, kOut, kH, A2, kH, 0, P + kOffW3, kH, 1, O, kOut, 1.f, 0.f);
    hipLaunchKernelGGL(k_bias, dim3(nblk(N * kOut)), dim3(256), 0, st, N, kOut, P + kOffB3, O);
}

// The stencil MLP's forward pass: Z1, A1, Z2, A2 (N x kH), O (N x kOut).
void smlp_forward(hipStream_t st, const float* P, const float* X, int64_t N, float* Z1, float* A1,
                  float* Z2, float* A2, float* O) {
    namespace S = todg_smlp;
    const int n = (int)N;
    gemm(st, n, S::kH, S::kIn, X, S::kIn, 0, P + S::kOffW1, S::kIn, 1, Z1, S::kH, 1.f, 0.f);
    hipLaunchKernelGGL(k_bias_gelu, dim3(nblk(N * S::kH)), dim3(256), 0, st, N, S::kH, P + S::kOffB1, Z1, A1);
    gemm(st, n, S::kH, S::kH, A1, S::kH, 0, P + S::kOffW2, S::kH, 1, Z2, S::kH, 1.f, 0.f);
    hipLaunchKernelGGL(k_bias_gelu, dim3(nblk(N * S::kH)), dim3(256), 0, st, N, S::kH, P + S::kOffB2, Z2, A2);
    gemm(st, n, S::kOut, S::kH, A2, S::kH, 0, P + S::kOffW3, S::kH, 1, O, S::kOut, 1.f, 0.f);
    hipLaunchKernelGGL(k_bias, dim3(nblk(N * S::kOut)), dim3(256), 0, st, N, S::kOut, P + S::kOffB3, O);
}

}  // namespace todg_corr

using namespace todg_corr;

extern "C" {

// Z1, A1, Z2, A2 (N x kH each), O, dO (N x kOut), the head's per-block loss
// parts (doubles, within N floats), then the split-K partials of the weight
// gradients.
size_t todg_corr_workspace_floats(int64_t N) {
    return (size_t)N * (4 * kH + 2 * kOut + 1) + 1024 + (size_t)n_splits(N) * kSplitMaxTiles * 256 +
           (size_t)kColsumParts;
}

int todg_corr_features(int64_t nside, const float* field, int64_t N, const double* z,
                       const double* phi, float* X, float* wbil, float* mu, float* sigma, int* pix,
                       int* ok, void* stream) {
    if (N < 0 || nside < 1 || nside > 8192 || (nside & (nside - 1)))
        return fail(TODG_ERR_ARG, "bad nside or N");
    if (N == 0) return TODG_OK;
    const RingEntry* rt = nullptr;
    int rc = todg_internal::ring_table(nside, &rt);
    if (rc) return rc;
    hipLaunchKernelGGL(k_features, dim3(nblk(N)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       (int)nside, rt, field, N, z, phi, X, wbil, mu, sigma, pix, ok);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_corr_forward(const float* params, const float* X, int64_t N, float* ws, float* O,
                      void* stream) {
    if (N <= 0 || N > (1 << 30)) return fail(TODG_ERR_ARG, "N out of range");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    float* Z1 = ws;
    float* A1 = Z1 + N * kH;
    float* Z2 = A1 + N * kH;
    float* A2 = Z2 + N * kH;
    forward(st, params, X, N, Z1, A1, Z2, A2, O);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_corr_loss_grad(const float* params, const float* X, const float* wbil, const float* that,
                        int64_t N, float* ws, float* grad, double* loss, float* pred,
                        void* stream) {
    if (N <= 0 || N > (1 << 30)) return fail(TODG_ERR_ARG, "N out of range");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int n = (int)N;
    float* Z1 = ws;
    float* A1 = Z1 + N * kH;
    float* Z2 = A1 + N * kH;
    float* A2 = Z2 + N * kH;
    float* O = A2 + N * kH;
    float* dO = O + N * kOut;
    double* part = reinterpret_cast<double*>(dO + N * kOut);  // 8-byte aligned: 288 N floats
    float* spl = dO + N * kOut + N + 1024;                     // split-K partials
    float* csp = spl + n_splits(N) * kSplitMaxTiles * 256;      // column-sum partials
    forward(st, params, X, N, Z1, A1, Z2, A2, O);
    const unsigned hb = nblk(N);
    hipLaunchKernelGGL(k_head, dim3(hb), dim3(256), 0, st, N, X, wbil, that, O, grad ? dO : nullptr,
                       pred, part);
    if (loss) hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (int)hb, part, loss);
    if (grad) {
        // layer 3: dW3 = dO^T A2, db3 = colsum dO, dA2 = dO W3
        gemm(st, kOut, kH, n, dO, kOut, 1, A2, kH, 0, grad + kOffW3, kH, 1.f, 0.f, spl);
        colsum(st, N, kOut, dO, grad + kOffB3, csp);
        gemm(st, n, kH, kOut, dO, kOut, 0, params + kOffW3, kH, 0, A2, kH, 1.f, 0.f);
        hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(N * kH)), dim3(256), 0, st, N * kH, Z2, A2);
        // layer 2 (dZ2 in A2): dW2 = dZ2^T A1, db2, dA1 = dZ2 W2
        gemm(st, kH, kH, n, A2, kH, 1, A1, kH, 0, grad + kOffW2, kH, 1.f, 0.f, spl);
        colsum(st, N, kH, A2, grad + kOffB2, csp);
        gemm(st, n, kH, kH, A2, kH, 0, params + kOffW2, kH, 0, A1, kH, 1.f, 0.f);
        hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(N * kH)), dim3(256), 0, st, N * kH, Z1, A1);
        // layer 1 (dZ1 in A1): dW1 = dZ1^T X, db1
        gemm(st, kH, kIn, n, A1, kH, 1, X, kIn, 0, grad + kOffW1, kIn, 1.f, 0.f, spl);
        colsum(st, N, kH, A1, grad + kOffB1, csp);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_corr_adam(float* params, const float* grad, float* m, float* v, int step, float lr,
                   float beta1, float beta2, float eps, void* stream) {
    if (step < 1) return fail(TODG_ERR_ARG, "step must be >= 1");
    const float c1 = 1.f - powf(beta1, (float)step);
    const float c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(kNParams)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       kNParams, params, grad, m, v, lr, beta1, beta2, eps, c1, c2);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_corr_gemm_bf16(int M, int N, int K, const float* A, int lda, int ta, const float* B,
                        int ldb, int tb, float* C, int ldc, float alpha, float beta,
                        void* stream) {
    if (M <= 0 || N <= 0 || K <= 0) return fail(TODG_ERR_ARG, "empty product");
    gemm(reinterpret_cast<hipStream_t>(stream), M, N, K, A, lda, ta, B, ldb, tb, C, ldc, alpha, beta);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

// ------------------------------------------------- MLP stencil corrector

// Z1, A1, Z2, A2 (N x kH each), then the split-K partials of the weight
// gradients.
size_t todg_smlp_workspace_floats(int64_t N) {
    return (size_t)N * 4 * todg_smlp::kH + 1024 + (size_t)n_splits(N) * kSplitMaxTiles * 256 +
           (size_t)kColsumParts;
}

int todg_smlp_forward(const float* params, const float* X, int64_t N, float* ws, float* O, void* stream) {
    namespace S = todg_smlp;
    if (N <= 0 || N > (1 << 30)) return fail(TODG_ERR_ARG, "N out of range");
    if (!params || !X || !ws || !O) return fail(TODG_ERR_ARG, "null pointer argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    float* Z1 = ws;
    float* A1 = Z1 + N * S::kH;
    float* Z2 = A1 + N * S::kH;
    float* A2 = Z2 + N * S::kH;
    smlp_forward(st, params, X, N, Z1, A1, Z2, A2, O);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_smlp_backward(const float* params, const float* X, int64_t N, float* ws, const float* dO,
                       float* grad, void* stream) {
    namespace S = todg_smlp;
    if (N <= 0 || N > (1 << 30)) return fail(TODG_ERR_ARG, "N out of range");
    if (!params || !X || !ws || !dO || !grad) return fail(TODG_ERR_ARG, "null pointer argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int n = (int)N;
    float* Z1 = ws;
    float* A1 = Z1 + N * S::kH;
    float* Z2 = A1 + N * S::kH;
    float* A2 = Z2 + N * S::kH;
    float* spl = A2 + N * S::kH + 1024;
    float* csp = spl + n_splits(N) * kSplitMaxTiles * 256;  // column-sum partials
    // layer 3: dW3 = dO^T A2, db3 = colsum dO, dA2 = dO W3
    gemm(st, S::kOut, S::kH, n, dO, S::kOut, 1, A2, S::kH, 0, grad + S::kOffW3, S::kH, 1.f, 0.f, spl);
    colsum(st, N, S::kOut, dO, grad + S::kOffB3, csp);
    gemm(st, n, S::kH, S::kOut, dO, S::kOut, 0, params + S::kOffW3, S::kH, 0, A2, S::kH, 1.f, 0.f);
    hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(N * S::kH)), dim3(256), 0, st, N * S::kH, Z2, A2);
    // layer 2 (dZ2 in A2): dW2 = dZ2^T A1, db2, dA1 = dZ2 W2
    gemm(st, S::kH, S::kH, n, A2, S::kH, 1, A1, S::kH, 0, grad + S::kOffW2, S::kH, 1.f, 0.f, spl);
    colsum(st, N, S::kH, A2, grad + S::kOffB2, csp);
    gemm(st, n, S::kH, S::kH, A2, S::kH, 0, params + S::kOffW2, S::kH, 0, A1, S::kH, 1.f, 0.f);
    hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(N * S::kH)), dim3(256), 0, st, N * S::kH, Z1, A1);
    // layer 1 (dZ1 in A1): dW1 = dZ1^T X, db1
    gemm(st, S::kH, S::kIn, n, A1, S::kH, 1, X, S::kIn, 0, grad + S::kOffW1, S::kIn, 1.f, 0.f, spl);
    colsum(st, N, S::kH, A1, grad + S::kOffB1, csp);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_smlp_head(int64_t N, const float* F, const float* O, float* A0, float* A1, void* stream) {
    if (N < 0) return fail(TODG_ERR_ARG, "N out of range");
    if (N == 0) return TODG_OK;
    if (!F || !O || !A0 || !A1) return fail(TODG_ERR_ARG, "null pointer argument");
    hipLaunchKernelGGL(k_smlp_head, dim3(nblk(N)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), N, F, O,
                       A0, A1);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_smlp_head_backward(int64_t N, const float* F, const float* G0, const float* G1, float* dO,
                            void* stream) {
    if (N < 0) return fail(TODG_ERR_ARG, "N out of range");
    if (N == 0) return TODG_OK;
    if (!F || !G0 || !G1 || !dO) return fail(TODG_ERR_ARG, "null pointer argument");
    hipLaunchKernelGGL(k_smlp_head_bwd, dim3(nblk(N)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), N,
                       F, G0, G1, dO);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

size_t todg_csr_rows2_parts(int64_t nrows) { return (size_t)nblk(nrows); }

int todg_csr_rows2(int64_t nrows, const int64_t* rowptr, const int* col, const float* v0, const float* v1,
                   const float* x0, const float* x1, const double* target, double scale, double* y,
                   float* resid, double* part, double* sumsq, void* stream) {
    if (nrows < 0) return fail(TODG_ERR_ARG, "nrows out of range");
    if (!rowptr || !col || !v0 || !v1 || !x0 || !x1) return fail(TODG_ERR_ARG, "null pointer argument");
    if (sumsq && !part) return fail(TODG_ERR_ARG, "sumsq needs the part workspace");
    if (nrows == 0) return TODG_OK;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const unsigned nb = nblk(nrows);
    hipLaunchKernelGGL(k_csr_rows2, dim3(nb), dim3(256), 0, st, nrows, rowptr, col, v0, v1, x0, x1, target, scale,
                       y, resid, part);
    if (sumsq) hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (int)nb, part, sumsq);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_csr_adjoint2(int64_t nrows, const int64_t* rowptr, const int* col, const float* v0, const float* v1,
                      const float* r, float* out0, float* out1, void* stream) {
    if (nrows < 0) return fail(TODG_ERR_ARG, "nrows out of range");
    if (nrows == 0) return TODG_OK;
    if (!rowptr || !col || !v0 || !v1 || !r || !out0 || !out1) return fail(TODG_ERR_ARG, "null pointer argument");
    hipLaunchKernelGGL(k_csr_adjoint2, dim3(nblk(nrows)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       nrows, rowptr, col, v0, v1, r, out0, out1);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_smlp_adam(float* params, const float* grad, float* m, float* v, int step, float lr, float lr_out,
                   float beta1, float beta2, float eps, void* stream) {
    namespace S = todg_smlp;
    if (step < 1) return fail(TODG_ERR_ARG, "step must be >= 1");
    if (!params || !grad || !m || !v) return fail(TODG_ERR_ARG, "null pointer argument");
    const float c1 = 1.f - powf(beta1, (float)step);
    const float c2 = 1.f - powf(beta2, (float)step);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // hidden layers at lr, the output layer (W3, b3) at lr_out
    hipLaunchKernelGGL(k_adam, dim3(nblk(S::kOffW3)), dim3(256), 0, st, S::kOffW3, params, grad, m, v, lr, beta1,
                       beta2, eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(nblk(S::kNParams - S::kOffW3)), dim3(256), 0, st, S::kNParams - S::kOffW3,
                       params + S::kOffW3, grad + S::kOffW3, m + S::kOffW3, v + S::kOffW3, lr_out, beta1, beta2,
                       eps, c1, c2);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : fail(TODG_ERR_HIP, hipGetErrorString(e));
}

}  // extern "C"
