// Learned interpolation corrector: definitions shared by the training
// kernels (corrector.hip) and the fused TOD kernel (tod_kernels.hip).
//
// For a query direction the corrector sees the sub-pixel position of the
// direction in the 4 x 4 ring neighbourhood of its bilinear cell (rings
// ira-1 .. ira+2, four slots j-1 .. j+2 around the query in each ring) and
// the neighbourhood's 16 map values normalised by their mean mu and RMS
// sigma; it emits 16 weight corrections whose mean is removed, so that the
// corrected weights (bilinear + correction) still sum to one:
//
//   x  = [fa - 1/2, v - 1/2, fb - 1/2, fm - 1/2, fp - 1/2,
//         (m_k - mu) / sigma for k < 16]                          (21)
//   h1 = gelu(W1 x + b1)        (64)    gelu: see below
//   h2 = gelu(W2 h1 + b2)       (64)
//   o  = W3 h2 + b3             (16)
//   w  = w_bilinear + o - mean(o)
//   prediction = mu + sigma * sum_k w_k (m_k - mu) / sigma
//
// fa, fb, fm, fp are the query's fractional slot coordinates in rings ira,
// ira+1, ira-1, ira+2 and v the bilinear ring weight (numba_healpy.py:443-480
// for rings ira, ira+1).  Neighbour k = 4 r + t sits in ring ira-1+r, slot
// j_r - 1 + t; the bilinear weights sit on k = 5, 6, 9, 10.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace todg_corr {

constexpr int kGeo = 5, kVals = 16, kIn = kGeo + kVals, kH = 64, kOut = 16;
// flat parameter layout (FP32): W1 (kH x kIn), b1, W2 (kH x kH), b2, W3 (kOut x kH), b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kIn;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kOut * kH;
constexpr int kNParams = kOffB3 + kOut;

// fp32 -> bf16, round to nearest even (finite inputs).
__device__ __forceinline__ short bf16_bits(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// GELU with a polynomial gate: gelu(z) = z s(z), where the Gaussian CDF of
// GELU (Hendrycks & Gimpel 2016) is replaced by the quintic smoothstep
// s = 1/2 + (15/16) t - (5/8) t^3 + (3/16) t^5, t = clamp(z / c, -1, 1),
// c = 2.64 (max |gelu(z) - z Phi(z)| = 0.019; C^2, and s is exactly 0 or 1
// beyond +-c).  Six VALU operations, no transcendental.
constexpr float kGateC = 2.64f;
constexpr float kGateA1 = 15.f / (16.f * kGateC);
constexpr float kGateA3 = -5.f / (8.f * kGateC * kGateC * kGateC);
constexpr float kGateA5 = 3.f / (16.f * kGateC * kGateC * kGateC * kGateC * kGateC);

__device__ __forceinline__ float gelu(float z) {
    const float u = __builtin_amdgcn_fmed3f(z, -kGateC, kGateC);
    const float u2 = u * u;
    const float p = fmaf(fmaf(kGateA5, u2, kGateA3), u2, kGateA1);
    return z * fmaf(u, p, 0.5f);
}
__device__ __forceinline__ float gelu_grad(float z) {
    const float u = __builtin_amdgcn_fmed3f(z, -kGateC, kGateC);
    const float u2 = u * u;
    const float p = fmaf(fmaf(kGateA5, u2, kGateA3), u2, kGateA1);
    const float s = fmaf(u, p, 0.5f);
    const float ds = fabsf(z) < kGateC ? fmaf(fmaf(5.f * kGateA5, u2, 3.f * kGateA3), u2, kGateA1) : 0.f;
    return fmaf(z, ds, s);
}

}  // namespace todg_corr

// MLP stencil corrector: the factorised corrector's two channels with
// per-pixel stencil weights emitted by an MLP evaluated once per map pixel.
//
// For pixel p of a map field, F_0 = m(p) and F_1 .. F_15 are the stencil
// features of the factorised corrector (ring-interpolated means at ring
// offsets dr = j / 4 and ring-step offsets ds = j % 4, minus m(p));
// sigma = sqrt(sum_{j>=1} F_j^2 / 15) is their local RMS.
//
//   x  = [2 |z_p| - 1, F_1 / sigma, .., F_15 / sigma]            (16)
//   h1 = gelu(W1 x + b1)        (64)
//   h2 = gelu(W2 h1 + b2)       (64)
//   o  = W3 h2 + b3             (32)
//   A0 = F_0 + sum_{j>=1} o_j F_j,   A1 = sum_{j>=1} o_{16+j} F_j
//
// (o_0 and o_16 are not used; x is 0 where sigma is 0.)  A constant map has
// F_j = 0 for j >= 1, so A0 = m and A1 = 0 there; the channels scale with
// the map.  The TOD kernel is the factorised corrector's (footprint mode 4).
namespace todg_smlp {

constexpr int kIn = 16, kH = 64, kOut = 32;
// flat parameter layout (FP32): W1 (kH x kIn), b1, W2 (kH x kH), b2, W3 (kOut x kH), b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kIn;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kOut * kH;
constexpr int kNParams = kOffB3 + kOut;

}  // namespace todg_smlp
