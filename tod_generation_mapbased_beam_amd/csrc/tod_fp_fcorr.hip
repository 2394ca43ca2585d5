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

// Value of ring r2 at longitude phi by linear interpolation between the two
// ring pixels whose centres bracket phi.
__device__ __forceinline__ double ring_lerp(const RingEntry* __restrict__ rt, const float* __restrict__ m,
                                            int r2, double phi) {
    const RingEntry& e = rt[r2];
    const double t = (phi - e.phi0) * e.inv_dphi;
    const double jf = floor(t);
    const double f = t - jf;
    int64_t j = (int64_t)jf % e.n;
    if (j < 0) j += e.n;
    const int64_t j1 = j + 1 == e.n ? 0 : j + 1;
    return (1.0 - f) * (double)m[e.first + j] + f * (double)m[e.first + j1];
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

// One thread per pixel: the 16 stencil features F_0 = m(p) and, for
// (dr, ds) in {0..3}^2 minus (0, 0) in row-major order, the mean of the
// ring-interpolated values at rings r -+ dr, longitudes phi_p -+ ds dphi_r,
// minus m(p) (features with dr > 0 are 0 within 3 rings of a pole); then
// out0 = sum_j c[j] F_j, out1 = sum_j c[16 + j] F_j (binary64 sums, stored
// binary32 at out[p * stride]).
__global__ __launch_bounds__(256) void k_fcorr_channels(const RingEntry* __restrict__ rt, int nside,
                                                         const float* __restrict__ m, FcorrCoef c,
                                                         float* __restrict__ out0, float* __restrict__ out1,
                                                         int64_t stride) {
    const int64_t npix = 12 * (int64_t)nside * nside;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int r = pixel_ring(nside, p, npix);
        const RingEntry& e = rt[r];
        const double phi = e.phi0 + (double)(p - e.first) * e.dphi;
        const double mp = (double)m[p];
        const bool ok = r >= 4 && r <= 4 * nside - 4;
        double s0 = c.c[0] * mp, s1 = c.c[16] * mp;
        int j = 1;
#pragma unroll
        for (int dr = 0; dr < 4; ++dr) {
#pragma unroll
            for (int ds = 0; ds < 4; ++ds) {
                if (dr == 0 && ds == 0) continue;
                double f = 0.0;
                if (dr == 0) {
                    f = 0.5 * (ring_lerp(rt, m, r, phi - ds * e.dphi) + ring_lerp(rt, m, r, phi + ds * e.dphi)) - mp;
                } else if (ok) {
                    const int ra = r - dr, rb = r + dr;
                    if (ds == 0) {
                        f = 0.5 * (ring_lerp(rt, m, ra, phi) + ring_lerp(rt, m, rb, phi)) - mp;
                    } else {
                        const double pl = phi - ds * e.dphi, ph = phi + ds * e.dphi;
                        f = 0.25 * (ring_lerp(rt, m, ra, pl) + ring_lerp(rt, m, ra, ph) + ring_lerp(rt, m, rb, pl) +
                                    ring_lerp(rt, m, rb, ph)) -
                            mp;
                    }
                }
                s0 = fma(c.c[j], f, s0);
                s1 = fma(c.c[16 + j], f, s1);
                ++j;
            }
        }
        out0[p * stride] = (float)s0;
        out1[p * stride] = (float)s1;
    }
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
