// Learned interpolation corrector: features, BF16-MFMA training kernels
// and the C-ABI that drives them.
//
// The corrector is a small MLP.  For a query direction it sees the
// sub-pixel position of the direction in its bilinear cell and the 4 x 4
// map neighbourhood around that cell (4 rings x 4 slots), normalised by the
// neighbourhood's mean and RMS; it emits 16 weight corrections whose mean is
// removed, so that the corrected weights (bilinear + correction) still sum to
// one and reproduce constant maps exactly.  It is trained to turn bilinear
// interpolation into the bicubic judge (bicubic_visit).
//
//   x   = [fa - 1/2, v - 1/2, (m_k - mu) / sigma for k < 16]     (18)
//   h1  = gelu(W1 x + b1)        (64)    gelu(z) = z sigmoid(1.702 z)
//   h2  = gelu(W2 h1 + b2)       (64)
//   o   = W3 h2 + b3             (16)
//   w   = w_bilinear + o - mean(o)
//   prediction = mu + sigma * sum_k w_k (m_k - mu) / sigma
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
#include "../../include/todgpu.h"

using namespace todg;

namespace todg_corr {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kIn = 18, kH = 64, kOut = 16;
// flat parameter layout (FP32): W1 (kH x kIn), b1, W2 (kH x kH), b2, W3 (kOut x kH), b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kIn;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kOut * kH;
constexpr int kNParams = kOffB3 + kOut;
static_assert(kNParams == TODG_CORR_NPARAMS, "parameter count");

// fp32 -> bf16, round to nearest even (finite inputs).
__device__ __forceinline__ short bf16_bits(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// GELU in its sigmoid form, z * sigmoid(1.702 z) (Hendrycks & Gimpel 2016),
// and its derivative.
constexpr float kGeluA = 1.702f;
__device__ __forceinline__ float gelu(float z) {
    const float s = 1.f / (1.f + __expf(-kGeluA * z));
    return z * s;
}
__device__ __forceinline__ float gelu_grad(float z) {
    const float s = 1.f / (1.f + __expf(-kGeluA * z));
    return fmaf(kGeluA * z * s, 1.f - s, s);
}

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
        const float* x = X + i * kIn + 2;
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

// out[f] = sum_i D[i][f] (bias gradients), one thread per column.
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
// fa, fb, weight v of the lower ring) and the 4 x 4 neighbourhood rings
// ira-1 .. ira+2 x slots j-1 .. j+2 (j = the ring's bilinear slot).
// X (N x 18), wbil (N x 4: weights of the cells 5, 6, 9, 10), mu, sigma,
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
    double fa = 0.0, v = 0.0, fb = 0.0;
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
    for (int k = 0; k < 16; ++k) x[2 + k] = (float)(((double)m[k] - mean) * inv);
    wbil[i * 4 + 0] = (float)((1.0 - v) * (1.0 - fa));
    wbil[i * 4 + 1] = (float)((1.0 - v) * fa);
    wbil[i * 4 + 2] = (float)(v * (1.0 - fb));
    wbil[i * 4 + 3] = (float)(v * fb);
    mu[i] = (float)mean;
    sigma[i] = (float)sd;
    for (int k = 0; k < 16; ++k) pix[i * 16 + k] = p16[k];
    ok[i] = good ? 1 : 0;
}


// ------------------------------------------------------ fused TOD inference

// Beam-convolution TOD with the corrector applied to every beam pixel:
// tod[c, b] += sum_s beam_val[s] * corrected interpolation of field c at
// R_b vec_orig[s].  Q/U are rotated per neighbour pixel by the reference's
// spin-2 angle first (tod_bilinear.py:403-432), so the corrector sees the
// rotated fields and, with zero correction, the result is the bilinear
// kernel's.  Where the corrector does not apply (pole neighbourhoods) or its
// correction exceeds `bound` (in units of the neighbourhood RMS), the plain
// bilinear weights are used.
struct LearnedArgs {
    const float *vec_orig, *beam_vals, *axes, *cos_a, *sin_a, *ax_pts, *cos_p, *sin_p;
    const float* map;  // packed (npix, CS)
    const RingEntry* rings;
    const float* params;
    double* tod;
    unsigned long long* stats;  // [0] corrected beam-pixel samples, [1] fallbacks (or null)
    int64_t B, S, npix;
    int nside, nslots;
    int row[4];
    float bound;
};

constexpr int kLThreads = 256;
constexpr int kLWaves = kLThreads / 64;
constexpr int kKp = 32;  // layer-1 inputs padded to two MFMA k-steps
// dynamic LDS layout (bytes)
constexpr int kLW1 = 0;                                // bf16 [kH][kKp]
constexpr int kLW2 = kLW1 + kH * kKp * 2;              // bf16 [kH][kH]
constexpr int kLW3 = kLW2 + kH * kH * 2;               // bf16 [kOut][kH]
constexpr int kLB = kLW3 + kOut * kH * 2;              // f32 b1[kH] b2[kH] b3[kOut]
constexpr int kLWave = kLB + (2 * kH + kOut) * 4;      // per-wave areas
constexpr int kLX = 0;                                 // bf16 [64][kKp]
constexpr int kLH1 = kLX + 64 * kKp * 2;               // bf16 [16][kH]
constexpr int kLH2 = kLH1 + 16 * kH * 2;               // bf16 [16][kH]
constexpr int kLO = kLH2 + 16 * kH * 2;                // f32 [64][kOut]
constexpr int kLWaveBytes = kLO + 64 * kOut * 4;
constexpr int kLRed = kLWave + kLWaves * kLWaveBytes;  // f64 [kLWaves][4]
constexpr int kLSmem = kLRed + kLWaves * 4 * 8;

extern __shared__ __align__(16) unsigned char lr_smem[];

template <typename T>
__device__ __forceinline__ T* lr_at(int off) {
    return reinterpret_cast<T*>(lr_smem + off);
}

__device__ __forceinline__ float bf16_to_f(unsigned short h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// One MFMA layer for a 16-row tile: out[16][16*ncol] = act(in[16][16*nk] W^T + b)
// with W (bf16, row-major [out][in], row stride wld) in LDS.  `in` rows have
// stride ild (bf16 elements).  GELU when act, else raw f32 into outf.
__device__ __forceinline__ void lr_layer(const unsigned short* in, int ild, int nk,
                                         const unsigned short* W, int wld, const float* bias,
                                         int ncol, bool act, unsigned short* outb, int old,
                                         float* outf, int ofld) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    for (int ct = 0; ct < ncol; ++ct) {
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < nk; ++ks) {
            const short4_t a = *reinterpret_cast<const short4_t*>(in + r * ild + ks * 16 + kq);
            const short4_t b =
                *reinterpret_cast<const short4_t*>(W + (ct * 16 + r) * wld + ks * 16 + kq);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        const int col = ct * 16 + r;
        const float bb = bias[col];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = kq + q;
            const float z = acc[q] + bb;
            if (act)
                outb[row * old + col] = (unsigned short)bf16_bits(gelu(z));
            else
                outf[row * ofld + col] = z;
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// cos 2delta, sin 2delta of pixel centre (z_n, sth_n, phi_n) against the
// boresight (tod_bilinear.py:84-145) with a light sin/cos of the half gap.
__device__ __forceinline__ void lr_spin2(double z_n, double sth_n, double phi_n, double z_pts,
                                         double sth_pts, double phi_pts, double& c2d, double& s2d) {
    if (phi_pts == phi_n && sth_pts == sth_n && z_pts == z_n) {
        c2d = 1.0;
        s2d = 0.0;
        return;
    }
    double h = (phi_pts - phi_n) * 0.5;
    if (h > 0.5 * kPi) h -= kPi;
    if (h < -0.5 * kPi) h += kPi;
    double sh, ch;
    sincos_any(h, sh, ch);
    spin2_from_half(spin2_ring(z_n, sth_n, z_pts, sth_pts), sh, ch, c2d, s2d);
}

template <bool QU, int CS>
__global__ __launch_bounds__(kLThreads, 2) void k_tod_learned(const LearnedArgs a) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // parameters -> LDS (bf16 weights, f32 biases)
    {
        unsigned short* W1 = lr_at<unsigned short>(kLW1);
        unsigned short* W2 = lr_at<unsigned short>(kLW2);
        unsigned short* W3 = lr_at<unsigned short>(kLW3);
        float* bs = lr_at<float>(kLB);
        for (int i = tid; i < kH * kKp; i += kLThreads) {
            const int o = i / kKp, k = i % kKp;
            W1[i] = k < kIn ? (unsigned short)bf16_bits(a.params[kOffW1 + o * kIn + k]) : 0;
        }
        for (int i = tid; i < kH * kH; i += kLThreads) W2[i] = (unsigned short)bf16_bits(a.params[kOffW2 + i]);
        for (int i = tid; i < kOut * kH; i += kLThreads) W3[i] = (unsigned short)bf16_bits(a.params[kOffW3 + i]);
        for (int i = tid; i < kH; i += kLThreads) {
            bs[i] = a.params[kOffB1 + i];
            bs[kH + i] = a.params[kOffB2 + i];
        }
        for (int i = tid; i < kOut; i += kLThreads) bs[2 * kH + i] = a.params[kOffB3 + i];
    }
    __syncthreads();
    const unsigned short* W1 = lr_at<unsigned short>(kLW1);
    const unsigned short* W2 = lr_at<unsigned short>(kLW2);
    const unsigned short* W3 = lr_at<unsigned short>(kLW3);
    const float* b1 = lr_at<float>(kLB);
    const float* b2 = b1 + kH;
    const float* b3 = b2 + kH;
    unsigned char* wb = lr_smem + kLWave + wave * kLWaveBytes;
    unsigned short* Xw = reinterpret_cast<unsigned short*>(wb + kLX);
    unsigned short* H1 = reinterpret_cast<unsigned short*>(wb + kLH1);
    unsigned short* H2 = reinterpret_cast<unsigned short*>(wb + kLH2);
    float* Ow = reinterpret_cast<float*>(wb + kLO);
    double* red = lr_at<double>(kLRed);
    const int nside = a.nside;
    const int F = a.nslots < CS ? a.nslots : CS;
    unsigned long long n_corr = 0, n_fb = 0;

    // XCD-aware sample order (workgroup bid runs on XCD bid % 8)
    const int G = gridDim.x, xcd = blockIdx.x & 7, per = G >> 3, kx = blockIdx.x >> 3;
    const int64_t chunk = (a.B + 7) / 8;
    const int64_t bbeg = (int64_t)xcd * chunk;
    const int64_t bend = (bbeg + chunk < a.B) ? bbeg + chunk : a.B;
    for (int64_t b = bbeg + kx; b < bend; b += per) {
        double qq[12];
        qq[0] = a.axes[3 * b]; qq[1] = a.axes[3 * b + 1]; qq[2] = a.axes[3 * b + 2];
        qq[3] = a.cos_a[b]; qq[4] = a.sin_a[b]; qq[5] = 1.0 - qq[3];
        qq[6] = a.ax_pts[3 * b]; qq[7] = a.ax_pts[3 * b + 1]; qq[8] = a.ax_pts[3 * b + 2];
        qq[9] = a.cos_p[b]; qq[10] = a.sin_p[b]; qq[11] = 1.0 - qq[9];
        const double z_pts = fmax(-1.0, fmin(1.0, qq[8]));
        const double sth_pts = sqrt(fmax(0.0, 1.0 - qq[8] * qq[8]));
        double phi_pts = atan2(qq[7], qq[6]);
        if (phi_pts < 0.0) phi_pts += kTwoPi;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int64_t s0 = 0; s0 < a.S; s0 += kLThreads) {
            const int64_t s = s0 + tid;
            const bool live = s < a.S;
            double bv = 0.0;
            float vals[4][16];
            float mu[4], sg[4];
            float fa = 0.f, fv = 0.f, fb = 0.f;
            bool good = false;
            Nb o;
            double z = 0.0, phi_w = 0.0;
            if (live) {
                double vx, vy, vz;
                rodrigues((double)a.vec_orig[3 * s], (double)a.vec_orig[3 * s + 1],
                          (double)a.vec_orig[3 * s + 2], qq[0], qq[1], qq[2], qq[3], qq[4], qq[5],
                          qq[6], qq[7], qq[8], qq[9], qq[10], qq[11], vx, vy, vz);
                phi_w = atan2(vy, vx);
                if (phi_w < 0.0) phi_w += kTwoPi;
                z = fmax(-1.0, fmin(1.0, vz));
                bv = (double)a.beam_vals[s];
                ring_interp(nside, a.rings, z, phi_w, o);
                const int ira = ring_above(nside, z);
                good = ira >= 2 && ira <= 4 * nside - 3;
                if (good) {
                    const double wa = o.w[0] + o.w[1];
                    const double v = o.w[2] + o.w[3];
                    fa = (float)(wa > 0.0 ? o.w[1] / wa : 0.0);
                    fb = (float)(v > 0.0 ? o.w[3] / v : 0.0);
                    fv = (float)v;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const RingEntry& e = a.rings[ira - 1 + r];
                        int j;
                        if (r == 1)
                            j = o.pos[0];
                        else if (r == 2)
                            j = o.pos[2];
                        else {
                            double fr;
                            j = ring_slot(phi_w, e.phi0, e.dphi, e.inv_dphi, e.n, fr);
                        }
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            int qp = j - 1 + t;
                            qp = ((qp % e.n) + e.n) % e.n;
                            const float4 m = load_px<CS>(a.map, e.first + qp, a.npix);
                            float m0 = m.x, m1 = m.y;
                            if constexpr (QU) {
                                double c2d, s2d;
                                lr_spin2(e.z, e.sth, e.phi0 + (double)qp * e.dphi, z_pts, sth_pts,
                                         phi_pts, c2d, s2d);
                                m0 = (float)fma((double)m.x, c2d, (double)m.y * s2d);
                                m1 = (float)fma((double)m.y, c2d, -((double)m.x * s2d));
                            }
                            vals[0][4 * r + t] = m0;
                            vals[1][4 * r + t] = m1;
                            vals[2][4 * r + t] = m.z;
                            vals[3][4 * r + t] = m.w;
                        }
                    }
#pragma unroll
                    for (int f = 0; f < 4; ++f) {
                        float sm = 0.f;
                        for (int k = 0; k < 16; ++k) sm += vals[f][k];
                        const float m = sm * (1.f / 16.f);
                        float s2 = 0.f;
                        for (int k = 0; k < 16; ++k) s2 = fmaf(vals[f][k] - m, vals[f][k] - m, s2);
                        mu[f] = m;
                        sg[f] = sqrtf(s2 * (1.f / 16.f));
                    }
                }
            }
            // the MLP, one field at a time over the wave's 64 rows
            for (int f = 0; f < F; ++f) {
                // this field's values (selects keep every register index static)
                float xv[16];
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    xv[k] = f == 0 ? vals[0][k] : (f == 1 ? vals[1][k] : (f == 2 ? vals[2][k] : vals[3][k]));
                const float muf = f == 0 ? mu[0] : (f == 1 ? mu[1] : (f == 2 ? mu[2] : mu[3]));
                const float sgf = f == 0 ? sg[0] : (f == 1 ? sg[1] : (f == 2 ? sg[2] : sg[3]));
                const float inv = good && sgf > 1e-30f ? 1.f / sgf : 0.f;
                unsigned short* xr = Xw + lane * kKp;
                xr[0] = (unsigned short)bf16_bits(good ? fa - 0.5f : 0.f);
                xr[1] = (unsigned short)bf16_bits(good ? fv - 0.5f : 0.f);
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    xr[2 + k] = (unsigned short)bf16_bits(good ? (xv[k] - muf) * inv : 0.f);
                for (int k = kIn; k < kKp; ++k) xr[k] = 0;
                __builtin_amdgcn_wave_barrier();
                for (int rt = 0; rt < 4; ++rt) {
                    lr_layer(Xw + rt * 16 * kKp, kKp, 2, W1, kKp, b1, 4, true, H1, kH, nullptr, 0);
                    lr_layer(H1, kH, 4, W2, kH, b2, 4, true, H2, kH, nullptr, 0);
                    lr_layer(H2, kH, 4, W3, kH, b3, 1, false, nullptr, 0, Ow + rt * 16 * kOut, kOut);
                }
                if (live) {
                    double val;
                    bool corrected = false;
                    if (good) {
                        const float* orow = Ow + lane * kOut;
                        float om = 0.f;
#pragma unroll
                        for (int k = 0; k < kOut; ++k) om += orow[k];
                        om *= 1.f / kOut;
                        float corr = 0.f;
#pragma unroll
                        for (int k = 0; k < kOut; ++k) corr = fmaf(orow[k] - om, (xv[k] - muf) * inv, corr);
                        const float bil = (1.f - fv) * ((1.f - fa) * xv[5] + fa * xv[6]) +
                                          fv * ((1.f - fb) * xv[9] + fb * xv[10]);
                        if (fabsf(corr) <= a.bound) {
                            val = (double)bil + (double)sgf * (double)corr;
                            corrected = true;
                        } else {
                            val = (double)bil;
                        }
                    } else {  // plain bilinear (pole neighbourhoods)
                        val = 0.0;
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const RingEntry& e = a.rings[o.ring[k]];
                            const float4 m = load_px<CS>(a.map, e.first + o.pos[k], a.npix);
                            double x = (double)(f == 0 ? m.x : (f == 1 ? m.y : (f == 2 ? m.z : m.w)));
                            if (QU && f < 2) {
                                double c2d, s2d;
                                lr_spin2(e.z, e.sth, e.phi0 + (double)o.pos[k] * e.dphi, z_pts, sth_pts,
                                         phi_pts, c2d, s2d);
                                x = f == 0 ? fma((double)m.x, c2d, (double)m.y * s2d)
                                           : fma((double)m.y, c2d, -((double)m.x * s2d));
                            }
                            val = fma(o.w[k], x, val);
                        }
                    }
                    const double add = bv * val;
                    acc[0] += f == 0 ? add : 0.0;
                    acc[1] += f == 1 ? add : 0.0;
                    acc[2] += f == 2 ? add : 0.0;
                    acc[3] += f == 3 ? add : 0.0;
                    if (f == 0) {
                        if (corrected) ++n_corr;
                        else ++n_fb;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        // block reduction, thread 0 adds to tod
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double v = acc[f];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) red[wave * 4 + f] = v;
        }
        __syncthreads();
        if (tid == 0) {
            for (int f = 0; f < F; ++f) {
                double t = 0.0;
                for (int w = 0; w < kLWaves; ++w) t += red[w * 4 + f];
                atomicAdd(&a.tod[(int64_t)a.row[f] * a.B + b], t);
            }
        }
        __syncthreads();
    }
    if (a.stats) {
        unsigned long long c = n_corr, fb = n_fb;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            c += __shfl_xor(c, off, 64);
            fb += __shfl_xor(fb, off, 64);
        }
        if (lane == 0) {
            atomicAdd(&a.stats[0], c);
            atomicAdd(&a.stats[1], fb);
        }
    }
}

int fail(int code, const std::string& msg) { return todg_internal::set_error(code, msg); }

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

void gemm(hipStream_t st, int M, int N, int K, const float* A, int lda, int ta, const float* B,
          int ldb, int tb, float* C, int ldc, float alpha, float beta) {
    const int64_t tiles = (int64_t)((M + 15) / 16) * ((N + 15) / 16);
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

}  // namespace todg_corr

using namespace todg_corr;

extern "C" {

size_t todg_corr_workspace_floats(int64_t N) { return (size_t)N * (4 * kH + 2 * kOut + 1) + 1024; }

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
    forward(st, params, X, N, Z1, A1, Z2, A2, O);
    const unsigned hb = nblk(N);
    hipLaunchKernelGGL(k_head, dim3(hb), dim3(256), 0, st, N, X, wbil, that, O, grad ? dO : nullptr,
                       pred, part);
    if (loss) hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, (int)hb, part, loss);
    if (grad) {
        // layer 3: dW3 = dO^T A2, db3 = colsum dO, dA2 = dO W3
        gemm(st, kOut, kH, n, dO, kOut, 1, A2, kH, 0, grad + kOffW3, kH, 1.f, 0.f);
        hipLaunchKernelGGL(k_colsum, dim3(kOut), dim3(256), 0, st, N, kOut, dO, grad + kOffB3);
        gemm(st, n, kH, kOut, dO, kOut, 0, params + kOffW3, kH, 0, A2, kH, 1.f, 0.f);
        hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(N * kH)), dim3(256), 0, st, N * kH, Z2, A2);
        // layer 2 (dZ2 in A2): dW2 = dZ2^T A1, db2, dA1 = dZ2 W2
        gemm(st, kH, kH, n, A2, kH, 1, A1, kH, 0, grad + kOffW2, kH, 1.f, 0.f);
        hipLaunchKernelGGL(k_colsum, dim3(kH), dim3(256), 0, st, N, kH, A2, grad + kOffB2);
        gemm(st, n, kH, kH, A2, kH, 0, params + kOffW2, kH, 0, A1, kH, 1.f, 0.f);
        hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(N * kH)), dim3(256), 0, st, N * kH, Z1, A1);
        // layer 1 (dZ1 in A1): dW1 = dZ1^T X, db1
        gemm(st, kH, kIn, n, A1, kH, 1, X, kIn, 0, grad + kOffW1, kIn, 1.f, 0.f);
        hipLaunchKernelGGL(k_colsum, dim3(kH), dim3(256), 0, st, N, kH, A1, grad + kOffB1);
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

int todg_gather_accum_learned(const float* vec_orig, const float* axes, const float* cos_a,
                              const float* sin_a, const float* ax_pts, const float* cos_p,
                              const float* sin_p, int64_t nside, const float* map_packed, int C,
                              const float* beam_vals, int64_t B, int64_t S, double* tod, int c_q,
                              int c_u, const float* params, float bound,
                              unsigned long long* stats, void* stream) {
    if (nside < 1 || nside > 8192 || (nside & (nside - 1)))
        return fail(TODG_ERR_ARG, "nside must be a power of two <= 8192");
    if (B < 0 || S < 0) return fail(TODG_ERR_ARG, "B and S must be non-negative");
    bool qu;
    int perm[4], nslots = 0, CS = 1;
    LearnedArgs a;
    int rc = todg_internal::map_slots(C, c_q, c_u, &qu, perm, a.row, &nslots, &CS);
    if (rc) return rc;
    if (B == 0 || S == 0) return TODG_OK;
    if (!vec_orig || !axes || !cos_a || !sin_a || !ax_pts || !cos_p || !sin_p || !map_packed ||
        !beam_vals || !tod || !params)
        return fail(TODG_ERR_ARG, "null pointer argument");
    const RingEntry* rt = nullptr;
    rc = todg_internal::ring_table(nside, &rt);
    if (rc) return rc;
    a.vec_orig = vec_orig;
    a.beam_vals = beam_vals;
    a.axes = axes;
    a.cos_a = cos_a;
    a.sin_a = sin_a;
    a.ax_pts = ax_pts;
    a.cos_p = cos_p;
    a.sin_p = sin_p;
    a.map = map_packed;
    a.rings = rt;
    a.params = params;
    a.tod = tod;
    a.stats = stats;
    a.B = B;
    a.S = S;
    a.npix = 12 * nside * nside;
    a.nside = (int)nside;
    a.nslots = nslots;
    a.bound = bound;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int64_t blocks = (int64_t)todg_internal::n_cu() * 2;
    const int64_t need = ((B + 7) / 8) * 8;
    if (blocks > need) blocks = need;
    blocks = ((blocks + 7) / 8) * 8;
    auto go = [&](auto kern) {
        (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, kLSmem);
        hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kLThreads), kLSmem, st, a);
    };
    if (qu) {
        if (CS == 2) go(k_tod_learned<true, 2>);
        else go(k_tod_learned<true, 4>);
    } else {
        if (CS == 1) go(k_tod_learned<false, 1>);
        else if (CS == 2) go(k_tod_learned<false, 2>);
        else go(k_tod_learned<false, 4>);
    }
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

}  // extern "C"
