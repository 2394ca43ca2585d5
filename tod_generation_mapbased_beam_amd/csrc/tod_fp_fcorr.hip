// The factorised corrector (footprint mode 4): its launcher and the
// map-side stencil kernel that builds its two channels.
//
// The corrected TOD of a sample is
//   sum_s bv_s sum_k w_k(s) [A0(p_k) + (1 - g_s) g_s A1(p_k)]
// over the four bilinear neighbours p_k of each rotated beam pixel, w_k its
// bilinear weights and g_s the lower ring's weight.  A0 and A1 are linear
// stencils of the map, built once per map (k_fcorr_channels): a learned
// prefilter and the channel that the beam pixels' ring positions modulate.
// The footprint kernel then deposits both weight sets into two LDS grids
// and gathers both channels once per footprint pixel.
#include "tod_device.hpp"

namespace todg_fp {

hipError_t launch_m4(const TodArgs& a, bool qu, int CS, hipStream_t st) {
    return launch_mode<4>(a, qu, CS, 0, st);
}

// Ring (1 .. 4 nside - 1) of RING pixel p.
__device__ __forceinline__ int pixel_ring(int nside, int64_t p, int64_t npix) {
    const int64_t ncap = 2 * (int64_t)nside * (nside - 1);
    if (p < ncap) {
        int64_t s = (int64_t)sqrt((double)(1 + 2 * p));
        while (s * s > 1 + 2 * p) --s;
        while ((s + 1) * (s + 1) <= 1 + 2 * p) ++s;
        return (int)((1 + s) >> 1);
    }
    if (p < npix - ncap) return (int)((p - ncap) / (4 * (int64_t)nside)) + nside;
    const int64_t q = npix - 1 - p;
    int64_t s = (int64_t)sqrt((double)(1 + 2 * q));
    while (s * s > 1 + 2 * q) --s;
    while ((s + 1) * (s + 1) <= 1 + 2 * q) ++s;
    return 4 * nside - (int)((1 + s) >> 1);
}

// Value at slot coordinate t of a ring (n pixels from mb): linear
// interpolation between the two pixels whose centres bracket t.  The
// callers' coordinates lie within a few slots of [0, n) (a pixel's own
// longitude plus at most 3 of its ring's slots, which are at most 1.75
// slots of a ring 3 rings away), so one conditional wrap replaces the
// 64-bit modulo; anything else takes the general remainder.
__device__ __forceinline__ double ring_lerp_t(const float* __restrict__ mb, int n, double t) {
    const double jf = floor(t);
    const double f = t - jf;
    int j = (int)jf;
    if (j < 0)
        j += n;
    else if (j >= n)
        j -= n;
    if ((unsigned)j >= (unsigned)n) j = (int)(((int64_t)jf % n + n) % n);
    const int j1 = j + 1 == n ? 0 : j + 1;
    TODG_CHECK((unsigned)j < (unsigned)n, kRingSlot);
    return (1.0 - f) * (double)mb[j] + f * (double)mb[j1];
}

// The 16 stencil features of pixel p on ring r (longitude phi): F_0 = m(p)
// and, for (dr, ds) in {0..3}^2 minus (0, 0) in row-major order (feature
// j = 4 dr + ds), the mean of the ring-interpolated values at rings r -+ dr,
// longitudes phi -+ ds dphi_r, minus m(p) (features with dr > 0 are 0
// within 3 rings of a pole).  On ring r itself those longitudes are pixel
// centres, read directly; on each neighbour ring the slot coordinate of
// phi is computed once and advances by dphi_r / dphi_ring per step.  Sums
// in the order (a- + a+) + b- + b+ over rings a = r - dr, b = r + dr.
__device__ __forceinline__ void fcorr_features(const RingEntry* __restrict__ rt, int nside,
                                               const float* __restrict__ m, int64_t p, int r,
                                               double phi, double f[16]) {
    const RingEntry& e = rt[r];
    const double mp = (double)m[p];
    const bool ok = r >= 4 && r <= 4 * nside - 4;
    f[0] = mp;
    {
        const int n = e.n;
        const int i = (int)(p - e.first);
        const float* __restrict__ mb = m + e.first;
#pragma unroll
        for (int ds = 1; ds < 4; ++ds) {
            int il = i - ds, ih = i + ds;
            if (il < 0) il += n;
            if (ih >= n) ih -= n;
            // (rings have >= 4 pixels, so one wrap suffices for ds <= 3)
            TODG_CHECK((unsigned)il < (unsigned)n && (unsigned)ih < (unsigned)n, kRingSlot);
            f[ds] = 0.5 * ((double)mb[il] + (double)mb[ih]) - mp;
        }
    }
#pragma unroll
    for (int dr = 1; dr < 4; ++dr) {
        if (ok) {
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const RingEntry& ea = rt[side == 0 ? r - dr : r + dr];
                const double t0 = (phi - ea.phi0) * ea.inv_dphi;
                const double step = e.dphi * ea.inv_dphi;
                const float* __restrict__ mb = m + ea.first;
                const int n = ea.n;
                const double v0 = ring_lerp_t(mb, n, t0);
                f[4 * dr] = side == 0 ? v0 : f[4 * dr] + v0;
#pragma unroll
                for (int ds = 1; ds < 4; ++ds) {
                    const double vl = ring_lerp_t(mb, n, fma(-(double)ds, step, t0));
                    const double vh = ring_lerp_t(mb, n, fma((double)ds, step, t0));
                    f[4 * dr + ds] = side == 0 ? vl + vh : (f[4 * dr + ds] + vl) + vh;
                }
            }
            f[4 * dr] = 0.5 * f[4 * dr] - mp;
#pragma unroll
            for (int ds = 1; ds < 4; ++ds) f[4 * dr + ds] = 0.25 * f[4 * dr + ds] - mp;
        } else {
#pragma unroll
            for (int ds = 0; ds < 4; ++ds) f[4 * dr + ds] = 0.0;
        }
    }
}

// One thread per pixel: out0 = sum_j c[j] F_j, out1 = sum_j c[16 + j] F_j
// (binary64 sums in feature order, stored binary32 at out[p * stride]).
__global__ __launch_bounds__(256, 4) void k_fcorr_channels(const RingEntry* __restrict__ rt, int nside,
                                                         const float* __restrict__ m, FcorrCoef c,
                                                         float* __restrict__ out0, float* __restrict__ out1,
                                                         int64_t stride) {
    const int64_t npix = 12 * (int64_t)nside * nside;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int r = pixel_ring(nside, p, npix);
        const RingEntry& e = rt[r];
        const double phi = e.phi0 + (double)(p - e.first) * e.dphi;
        double f[16];
        fcorr_features(rt, nside, m, p, r, phi, f);
        double s0 = c.c[0] * f[0], s1 = c.c[16] * f[0];
#pragma unroll
        for (int j = 1; j < 16; ++j) {
            s0 = fma(c.c[j], f[j], s0);
            s1 = fma(c.c[16 + j], f[j], s1);
        }
        out0[p * stride] = (float)s0;
        out1[p * stride] = (float)s1;
    }
}

// ------------------------------------------------- MLP stencil corrector

// The MLP's inputs of pixel p from its features: 2 |z_p| - 1 and the
// features 1..15 over their local RMS (0 where the RMS is 0).
__device__ __forceinline__ void smlp_inputs(const double f[16], double zp, float x[16]) {
    double s2 = 0.0;
#pragma unroll
    for (int j = 1; j < 16; ++j) s2 = fma(f[j], f[j], s2);
    const double sg = sqrt(s2 * (1.0 / 15.0));
    const double inv = sg > 0.0 ? 1.0 / sg : 0.0;
    x[0] = (float)(2.0 * fabs(zp) - 1.0);
#pragma unroll
    for (int j = 1; j < 16; ++j) x[j] = (float)(f[j] * inv);
}

// Training side: the features F (npix x 16, binary32) and the MLP inputs X
// (npix x 16) of every pixel.
__global__ __launch_bounds__(256) void k_smlp_features(const RingEntry* __restrict__ rt, int nside,
                                                        const float* __restrict__ m, float* __restrict__ F,
                                                        float* __restrict__ X) {
    const int64_t npix = 12 * (int64_t)nside * nside;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int r = pixel_ring(nside, p, npix);
        const RingEntry& e = rt[r];
        const double phi = e.phi0 + (double)(p - e.first) * e.dphi;
        double f[16];
        fcorr_features(rt, nside, m, p, r, phi, f);
        float x[16];
        smlp_inputs(f, e.z, x);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            F[p * 16 + j] = (float)f[j];
            X[p * 16 + j] = x[j];
        }
    }
}

// Inference: features, the MLP on v_mfma_f32_16x16x32_bf16 and the two
// channels, fused, for 64 pixels per wave and round (one pixel per lane).
// Each layer is computed transposed, out^T = W in^T (lane l: A[l%16][8(l/16)
// .. +7], B[8(l/16) .. +7][l%16], D[4(l/16) .. +3][l%16]); the biases are the
// accumulators' initial values (FP32), so the arithmetic is the training
// path's: BF16 operands (inputs, weights, activations), FP32 accumulation.
// A layer's output tile leaves lane l holding features 4(l/16) .. +3 of row
// l%16, which is the next layer's B operand once that layer's k order is
// permuted (smlp_perm).  Inputs go through a per-wave LDS tile (64 rows x
// 32 bf16, natural k order), outputs come back through it (64 rows x 32
// floats).
namespace smlp = todg_smlp;
constexpr int kSmlpHT = smlp::kH / 16;    // hidden output tiles of 16
constexpr int kSmlpHB = smlp::kH / 32;    // hidden k-blocks of 32
constexpr int kSmlpOT = smlp::kOut / 16;  // output tiles of 16
constexpr int kSmlpTile = 64 * smlp::kOut * 4;  // bytes per wave
static_assert(smlp::kIn <= 32 && smlp::kH % 32 == 0 && smlp::kOut % 16 == 0, "stencil MLP tiling");

__device__ __forceinline__ int smlp_perm(int kb, int g, int j) {
    return 16 * (2 * kb + (j >> 2)) + 4 * g + (j & 3);
}

// The weights and biases live in LDS as per-lane MFMA fragments, read at
// each use, rather than in ~100 VGPRs per lane: three workgroups (12 waves)
// then fit on a CU instead of two, which hides more of the feature gathers'
// latency.
constexpr int kSmlpW1 = 0;                                 // [ft][lane]
constexpr int kSmlpW2 = kSmlpW1 + kSmlpHT * 64;            // [ft][kb][lane]
constexpr int kSmlpW3 = kSmlpW2 + kSmlpHT * kSmlpHB * 64;  // [ot][kb][lane]
constexpr int kSmlpWFrags = kSmlpW3 + kSmlpOT * kSmlpHB * 64;
constexpr int kSmlpB1 = 0, kSmlpB2 = kSmlpHT * 4, kSmlpB3 = 2 * kSmlpHT * 4;  // [tile][g]
constexpr int kSmlpBFrags = (2 * kSmlpHT + kSmlpOT) * 4;

__global__ __launch_bounds__(256, 3) void k_smlp_channels(const RingEntry* __restrict__ rt, int nside,
                                                        const float* __restrict__ m,
                                                        const float* __restrict__ P, float* __restrict__ out0,
                                                        float* __restrict__ out1, int64_t stride) {
    __shared__ __attribute__((aligned(16))) unsigned char tiles[4 * kSmlpTile];
    // weights as MFMA A operands, biases as accumulator rows 4g .. 4g + 3
    __shared__ bf16x8 wf[kSmlpWFrags];
    __shared__ f32x4 bfr[kSmlpBFrags];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    unsigned char* tile = tiles + wave * kSmlpTile;
    // wave w lays out weight fragments w, w + 4, ... (fragment = one tile
    // and k-block of one layer, 16 bytes per lane)
    for (int fi = wave; fi < kSmlpWFrags / 64; fi += 4) {
        bf16x8 v;
        if (fi < kSmlpW2 / 64) {
            const int o = 16 * fi + r16;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 8 * g + j;
                v[j] = (__bf16)(k < smlp::kIn ? P[smlp::kOffW1 + o * smlp::kIn + k] : 0.f);
            }
        } else if (fi < kSmlpW3 / 64) {
            const int ft = (fi - kSmlpW2 / 64) / kSmlpHB, kb = (fi - kSmlpW2 / 64) % kSmlpHB;
            const int o = 16 * ft + r16;
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = (__bf16)P[smlp::kOffW2 + o * smlp::kH + smlp_perm(kb, g, j)];
        } else {
            const int ot = (fi - kSmlpW3 / 64) / kSmlpHB, kb = (fi - kSmlpW3 / 64) % kSmlpHB;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                v[j] = (__bf16)P[smlp::kOffW3 + (16 * ot + r16) * smlp::kH + smlp_perm(kb, g, j)];
        }
        wf[64 * fi + lane] = v;
    }
    if (threadIdx.x < kSmlpBFrags) {
        const int t = threadIdx.x / 4, gg = threadIdx.x % 4;
        const int off = t < kSmlpHT ? smlp::kOffB1 + 16 * t
                                    : (t < 2 * kSmlpHT ? smlp::kOffB2 + 16 * (t - kSmlpHT)
                                                       : smlp::kOffB3 + 16 * (t - 2 * kSmlpHT));
        f32x4 v;
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = P[off + 4 * gg + q];
        bfr[threadIdx.x] = v;
    }
    __syncthreads();
    auto w1 = [&](int ft) { return wf[kSmlpW1 + 64 * ft + lane]; };
    auto w2 = [&](int ft, int kb) { return wf[kSmlpW2 + 64 * (ft * kSmlpHB + kb) + lane]; };
    auto w3 = [&](int ot, int kb) { return wf[kSmlpW3 + 64 * (ot * kSmlpHB + kb) + lane]; };
    auto b1 = [&](int ft) { return bfr[kSmlpB1 + 4 * ft + g]; };
    auto b2 = [&](int ft) { return bfr[kSmlpB2 + 4 * ft + g]; };
    auto b3 = [&](int ot) { return bfr[kSmlpB3 + 4 * ot + g]; };
    const int64_t npix = 12 * (int64_t)nside * nside;
    const int64_t nchunks = (npix + 63) / 64;
    const int64_t wstride = (int64_t)gridDim.x * 4;
    for (int64_t ch = (int64_t)blockIdx.x * 4 + wave; ch < nchunks; ch += wstride) {
        const int64_t p = ch * 64 + lane;
        const bool on = p < npix;
        double f[16];
        float x[16];
        if (on) {
            const int r = pixel_ring(nside, p, npix);
            const RingEntry& e = rt[r];
            const double phi = e.phi0 + (double)(p - e.first) * e.dphi;
            fcorr_features(rt, nside, m, p, r, phi, f);
            smlp_inputs(f, e.z, x);
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                f[j] = 0.0;
                x[j] = 0.f;
            }
        }
        // this lane's input row: 16 inputs, 16 zeros
        {
            bf16x8 v[4];
#pragma unroll
            for (int k = 0; k < 32; ++k) v[k >> 3][k & 7] = (__bf16)(k < smlp::kIn ? x[k] : 0.f);
            bf16x8* dst = reinterpret_cast<bf16x8*>(tile + lane * (smlp::kOut * 4));
#pragma unroll
            for (int q = 0; q < 4; ++q) dst[q] = v[q];
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            unsigned char* row = tile + (16 * t + r16) * (smlp::kOut * 4);
            const bf16x8 xin = *reinterpret_cast<const bf16x8*>(row + 16 * g);
            f32x4 h[kSmlpHT];
            bf16x8 a[kSmlpHB];
#pragma unroll
            for (int ft = 0; ft < kSmlpHT; ++ft) h[ft] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1(ft), xin, b1(ft), 0, 0, 0);
#pragma unroll
            for (int kb = 0; kb < kSmlpHB; ++kb) a[kb] = corr_act(h[2 * kb], h[2 * kb + 1]);
#pragma unroll
            for (int ft = 0; ft < kSmlpHT; ++ft) {
                h[ft] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2(ft, 0), a[0], b2(ft), 0, 0, 0);
#pragma unroll
                for (int kb = 1; kb < kSmlpHB; ++kb)
                    h[ft] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2(ft, kb), a[kb], h[ft], 0, 0, 0);
            }
#pragma unroll
            for (int kb = 0; kb < kSmlpHB; ++kb) a[kb] = corr_act(h[2 * kb], h[2 * kb + 1]);
            // all lanes have read their B operand of this row tile before
            // the outputs overwrite it
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int ot = 0; ot < kSmlpOT; ++ot) {
                f32x4 o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3(ot, 0), a[0], b3(ot), 0, 0, 0);
#pragma unroll
                for (int kb = 1; kb < kSmlpHB; ++kb)
                    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3(ot, kb), a[kb], o, 0, 0, 0);
                // row 16t + r16, outputs 16 ot + 4g .. +3
                *reinterpret_cast<f32x4*>(row + 4 * (16 * ot + 4 * g)) = o;
            }
        }
        __builtin_amdgcn_wave_barrier();
        const float* mine = reinterpret_cast<const float*>(tile + lane * (smlp::kOut * 4));
        double s0 = f[0], s1 = 0.0;
#pragma unroll
        for (int j = 1; j < 16; ++j) {
            s0 = fma((double)mine[j], f[j], s0);
            s1 = fma((double)mine[16 + j], f[j], s1);
        }
        if (on) {
            out0[p * stride] = (float)s0;
            out1[p * stride] = (float)s1;
        }
        // the next round's input rows overwrite this round's outputs
        __builtin_amdgcn_wave_barrier();
    }
}

hipError_t launch_smlp_features(const RingEntry* rings, int nside, const float* field, float* F, float* X,
                                hipStream_t st) {
    const int64_t npix = 12 * (int64_t)nside * nside;
    int64_t blocks = (npix + 255) / 256;
    const int64_t cap = (int64_t)todg_internal::n_cu() * 64;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_smlp_features, dim3((unsigned)blocks), dim3(256), 0, st, rings, nside, field, F, X);
    return hipGetLastError();
}

hipError_t launch_smlp_channels(const RingEntry* rings, int nside, const float* field, const float* params,
                                float* out0, float* out1, int64_t stride, hipStream_t st) {
    const int64_t npix = 12 * (int64_t)nside * nside;
    int64_t blocks = (npix + 255) / 256;
    const int64_t cap = (int64_t)todg_internal::n_cu() * 16;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_smlp_channels, dim3((unsigned)blocks), dim3(256), 0, st, rings, nside, field, params,
                       out0, out1, stride);
    return hipGetLastError();
}

hipError_t launch_fcorr_channels(const RingEntry* rings, int nside, const float* field, const FcorrCoef& c,
                                 float* out0, float* out1, int64_t stride, hipStream_t st) {
    const int64_t npix = 12 * (int64_t)nside * nside;
    int64_t blocks = (npix + 255) / 256;
    const int64_t cap = (int64_t)todg_internal::n_cu() * 64;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_fcorr_channels, dim3((unsigned)blocks), dim3(256), 0, st, rings, nside, field, c, out0,
                       out1, stride);
    return hipGetLastError();
}

}  // namespace todg_fp
