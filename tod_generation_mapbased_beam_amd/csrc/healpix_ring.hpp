// HEALPix RING-scheme geometry on the device (gfx950).
//
// The index path (ring_above, per-ring phi slot, truncation) is evaluated in
// binary64 with the reference's operation order and no FMA contraction (this
// file is compiled with -ffp-contract=off), so for the same rotated direction
// it selects the same pixels as numba_healpy.py:53-114 and :358-499.  Per-ring
// constants (z, acos(z), phi0, dphi, first pixel, pixel count) come from a
// table built on the host with the same formulas and glibc's acos, so they
// are bit-identical to what the reference computes inline.  Everything that
// only feeds weights or spin-2 factors (a tolerance path, not an index path)
// uses explicit fma() and cheaper formulations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "check.hpp"

namespace todg {

constexpr double kTwoPi = 2.0 * 3.141592653589793;  // numba_healpy.py:42
constexpr double kPi = 3.141592653589793;
constexpr double kTwoThirds = 2.0 / 3.0;            // numba_healpy.py:44
// Reference longitudes lam_j = j * 2 pi / kLamN of the footprint kernel's
// longitude frame (their cosines and sines live in a device table).
constexpr int kLamN = 512;

// One row per ring index ir in [0, 4*nside]; rows 0 and 4*nside are padding
// so that ir indexes directly.  80 bytes.
struct alignas(16) RingEntry {
    double z;        // ring centre cos(theta)                 numba_healpy.py:104-114
    double theta;    // acos(z), glibc                          numba_healpy.py:331-332
    double phi0;     // longitude of the first pixel            numba_healpy.py:99-100
    double dphi;     // 2*pi / n                                numba_healpy.py:99
    double inv_dphi; // 1 / dphi, correctly rounded             (exact slot division)
    double sth;      // sqrt(max(0, 1 - z*z))                   tod_bilinear.py:194
    double inv_dth;  // 1 / (theta[ir+1] - theta[ir])           (weights only)
    int64_t first;   // first pixel index                       numba_healpy.py:85-98
    int32_t n;       // pixels in the ring                      numba_healpy.py:85-98
    int32_t pad;
};

// CPython float modulo (as lowered by Numba) for a positive modulus n and the
// range the lookup produces, x in (-n, 2n): exact, no libm call.
//   0 <= x < n : fmod = x (and -0.0 becomes +0.0)
//   n <= x < 2n: fmod = x - n, exact by Sterbenz
//   -n < x < 0 : fmod = x, sign differs from n, so x + n (rounded)
__device__ __forceinline__ double py_fmod_pos(double x, double n) {
    if (x >= n) {
        if (x < 2.0 * n) return x - n;
        double r = fmod(x, n);
        return r + 0.0;
    }
    if (x < 0.0) {
        double r = (x > -n) ? x : fmod(x, n);
        return (r != 0.0) ? r + n : 0.0;
    }
    return x + 0.0;
}

// numba_healpy.py:53-70
__device__ __forceinline__ int ring_above(int nside, double z) {
    double az = fabs(z);
    int ir;
    if (az > kTwoThirds) {
        double tp = (double)nside * sqrt(3.0 * (1.0 - az));
        ir = (int)tp;
        if (z < 0.0) ir = 4 * nside - ir - 1;
    } else {
        ir = (int)((double)nside * (2.0 - 1.5 * z));
    }
    return ir;
}

// x / d, correctly rounded, from the correctly rounded reciprocal inv = 1/d:
// q0 = x*inv, residual r = x - q0*d exactly by FMA, q = q0 + r*inv.
// Markstein's theorem makes q the correctly rounded quotient when nothing
// underflows; tiny |x| takes the IEEE division.  (Also checked exhaustively
// near every quotient boundary for all 8192 ring divisors 2*pi/n.)
__device__ __forceinline__ double div_by(double x, double d, double inv) {
    if (fabs(x) < 0x1p-960) return x / d;
    const double q0 = x * inv;
    const double r = fma(-q0, d, x);
    return fma(r, inv, q0);
}

// Position of phi_w in ring r: tw = ((phi_w - phi0)/dphi) % n, i = int(tw),
// frac = tw - i (numba_healpy.py:454-456).  i is in [0, n]; i == n only in
// the reference's 1-ulp float-mod corner, which is reproduced.
__device__ __forceinline__ int ring_slot(double phi_w, double phi0, double dphi,
                                         double inv_dphi, int n, double& frac) {
    const double tw = py_fmod_pos(div_by(phi_w - phi0, dphi, inv_dphi), (double)n);
    const int i = (int)tw;
    frac = tw - (double)i;
    return i;
}

// (i + 1) % n for i in [0, n] without an integer division.
__device__ __forceinline__ int next_slot(int i, int n) {
    const int j = i + 1;
    return j >= n ? j - n : j;
}

// ------------------------------------------------ tolerance-path helpers

// sin and cos of |x| <= 0.3 by Taylor series (truncation < 1e-19).
__device__ __forceinline__ void sincos_small(double x, double& s, double& c) {
    const double x2 = x * x;
    double ps = fma(x2, 1.0 / 6227020800.0, -1.0 / 39916800.0);
    ps = fma(x2, ps, 1.0 / 362880.0);
    ps = fma(x2, ps, -1.0 / 5040.0);
    ps = fma(x2, ps, 1.0 / 120.0);
    ps = fma(x2, ps, -1.0 / 6.0);
    s = fma(x * x2, ps, x);
    double pc = fma(x2, 1.0 / 479001600.0, -1.0 / 3628800.0);
    pc = fma(x2, pc, 1.0 / 40320.0);
    pc = fma(x2, pc, -1.0 / 720.0);
    pc = fma(x2, pc, 1.0 / 24.0);
    pc = fma(x2, pc, -0.5);
    c = fma(x2, pc, 1.0);
}

// sin and cos of x, |x| < 2 pi: quadrant reduction by a two-part pi/2 and
// Taylor series on |r| <= pi/4 (truncation < 1e-19).
__device__ __forceinline__ void sincos_red(double x, double& s, double& c) {
    const double k = rint(x * (2.0 / kPi));
    double r = fma(-k, 1.5707963267948966, x);
    r = fma(-k, 6.123233995736766e-17, r);
    const double r2 = r * r;
    double ps = fma(r2, -1.0 / 355687428096000.0, 1.0 / 1307674368000.0);
    ps = fma(r2, ps, -1.0 / 6227020800.0);
    ps = fma(r2, ps, 1.0 / 39916800.0);
    ps = fma(r2, ps, -1.0 / 362880.0);
    ps = fma(r2, ps, 1.0 / 5040.0);
    ps = fma(r2, ps, -1.0 / 120.0);
    ps = fma(r2, ps, 1.0 / 6.0);
    const double sr = fma(-(r * r2), ps, r);
    double pc = fma(r2, 1.0 / 6402373705728000.0, -1.0 / 20922789888000.0);
    pc = fma(r2, pc, 1.0 / 87178291200.0);
    pc = fma(r2, pc, -1.0 / 479001600.0);
    pc = fma(r2, pc, 1.0 / 3628800.0);
    pc = fma(r2, pc, -1.0 / 40320.0);
    pc = fma(r2, pc, 1.0 / 720.0);
    pc = fma(r2, pc, -1.0 / 24.0);
    pc = fma(r2, pc, 0.5);
    const double cr = fma(-r2, pc, 1.0);
    const int q = ((int)k) & 3;
    s = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
    c = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
}

// sin and cos of |x| <= 0.05 by Taylor series (truncation < 1e-20).
__device__ __forceinline__ void sincos_tiny(double x, double& s, double& c) {
    const double x2 = x * x;
    double ps = fma(x2, -1.0 / 5040.0, 1.0 / 120.0);
    ps = fma(x2, ps, -1.0 / 6.0);
    s = fma(x * x2, ps, x);
    double pc = fma(x2, 1.0 / 40320.0, -1.0 / 720.0);
    pc = fma(x2, pc, 1.0 / 24.0);
    pc = fma(x2, pc, -0.5);
    c = fma(x2, pc, 1.0);
}

// sin and cos of 0.05 < |x| < 2 pi, out of line so that the hot callers do
// not carry its coefficients.
__device__ __noinline__ double2 sincos_wide(double x) {
    double s, c;
    if (fabs(x) <= 0.3)
        sincos_small(x, s, c);
    else
        sincos_red(x, s, c);
    return make_double2(s, c);
}

// sin and cos of |x| < 2 pi: Taylor when small, else quadrant-reduced.
__device__ __forceinline__ void sincos_any(double x, double& s, double& c) {
    if (fabs(x) <= 0.05) {
        sincos_tiny(x, s, c);
    } else {
        const double2 r = sincos_wide(x);
        s = r.x;
        c = r.y;
    }
}

// sin and cos of the half angle x; only products of pairs are used
// downstream, so (s, c) and (-s, -c) are interchangeable and x may be
// reduced by pi.
__device__ __forceinline__ void sincos_half(double x, double& s, double& c) {
    if (fabs(x) <= 0.3) {
        sincos_small(x, s, c);
    } else {
        sincos(x, &s, &c);
    }
}

// 1/x from the hardware estimate plus one Newton step: the estimate's
// relative error e is below 2^-23 (measured 2^-24.4 on gfx950), the step's
// below e^2 + 2^-52 < 2e-14 (measured 2^-48.7).
__device__ __forceinline__ double rcp_nr1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// 1/x from the hardware estimate plus two Newton steps (< 1 ulp error).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Boresight-side and ring-side constants of the spin-2 formula
// (tod_bilinear.py:111-137) that do not depend on the pixel's longitude.
struct Spin2Ring {
    double st2, S, s2dt, zsum, SC4;
};

__device__ __forceinline__ Spin2Ring spin2_ring(double z_pix, double sth_pix, double z_pts,
                                                double sth_pts) {
    Spin2Ring r;
    const double dz = z_pts - z_pix;
    const double ds = sth_pts - sth_pix;
    r.st2 = 0.25 * fma(dz, dz, ds * ds);
    r.S = sth_pix * sth_pts;
    r.s2dt = 4.0 * r.st2 * (1.0 - r.st2);
    r.zsum = z_pts + z_pix;
    r.SC4 = 4.0 * r.S * (z_pts * z_pix);
    return r;
}

// cos(2 delta), sin(2 delta) given (sin, cos) of half the longitude gap
// (tod_bilinear.py:112-145).  The reference's exact-coincidence guard
// returns (1, 0); for coincident points N = D = 0, which is tested here.
__device__ __forceinline__ void spin2_from_half(const Spin2Ring& r, double sh, double ch,
                                                double& c2d, double& s2d) {
    const double sin_dphi = 2.0 * ch * sh;
    const double sh2 = sh * sh;
    const double cos_dphi = fma(-2.0, sh2, 1.0);
    const double h = fma(r.S, sh2, r.st2);
    const double N = 2.0 * sin_dphi * r.zsum * h;
    const double D = fma(r.S * sin_dphi, sin_dphi, fma(r.s2dt, cos_dphi, -(r.SC4 * sh2 * sh2)));
    const double den = fma(N, N, D * D);
    if (den == 0.0) {
        c2d = 1.0;
        s2d = 0.0;
        return;
    }
    const double u = rcp_nr(den);
    c2d = fma(D, D, -(N * N)) * u;
    s2d = 2.0 * N * D * u;
}

// Full spin-2 for one pixel (generic longitude gap).
__device__ __forceinline__ void spin2(double z_pix, double sth_pix, double phi_pix,
                                      double z_pts, double sth_pts, double phi_pts,
                                      double& c2d, double& s2d) {
    if (phi_pts == phi_pix && sth_pts == sth_pix && z_pts == z_pix) {
        c2d = 1.0;
        s2d = 0.0;
        return;
    }
    double sh, ch;
    sincos_half((phi_pts - phi_pix) * 0.5, sh, ch);
    spin2_from_half(spin2_ring(z_pix, sth_pix, z_pts, sth_pts), sh, ch, c2d, s2d);
}

// Binary32 spin-2 for the footprint sweep: the same formula from the
// per-ring constants rounded to binary32 (each term keeps ~1e-7 relative
// accuracy: near the boresight all of N and D are sums of like-signed small
// terms).  cos 2delta and sin 2delta come out with absolute error ~1e-7,
// well inside the tolerance of the tod they rotate.
struct Spin2RingF {
    float st2, S, s2dt, zsum, SC4;
};

__device__ __forceinline__ void spin2_from_half_f(const Spin2RingF& r, float sh, float ch, float& c2d,
                                                  float& s2d) {
    const float sin_dphi = 2.f * ch * sh;
    const float sh2 = sh * sh;
    const float cos_dphi = fmaf(-2.f, sh2, 1.f);
    const float h = fmaf(r.S, sh2, r.st2);
    float N = 2.f * sin_dphi * r.zsum * h;
    float D = fmaf(r.S * sin_dphi, sin_dphi, fmaf(r.s2dt, cos_dphi, -(r.SC4 * sh2 * sh2)));
    const float mx = fmaxf(fabsf(N), fabsf(D));
    if (mx == 0.f) {
        c2d = 1.f;
        s2d = 0.f;
        return;
    }
    const float im = __builtin_amdgcn_rcpf(mx);  // scale to [1, 2] before squaring
    N *= im;
    D *= im;
    const float u = __builtin_amdgcn_rcpf(fmaf(N, N, D * D));
    c2d = fmaf(D, D, -(N * N)) * u;
    s2d = 2.f * N * D * u;
}

// sin and cos in binary32: Taylor for |x| <= 0.05 (truncation < 3e-15),
// the library's sincosf beyond (out of line).
__device__ __noinline__ float2 sincosf_wide(float x) {
    float s, c;
    sincosf(x, &s, &c);
    return make_float2(s, c);
}

__device__ __forceinline__ void sincos_any_f(float x, float& s, float& c) {
    if (fabsf(x) <= 0.05f) {
        const float x2 = x * x;
        s = fmaf(x * x2, fmaf(x2, 1.f / 120.f, -1.f / 6.f), x);
        c = fmaf(x2, fmaf(x2, fmaf(x2, -1.f / 720.f, 1.f / 24.f), -0.5f), 1.f);
    } else {
        const float2 r = sincosf_wide(x);
        s = r.x;
        c = r.y;
    }
}

// theta - theta_a for a query z below ring a (z <= za): asin(sin(theta -
// theta_a)) with sin(theta - theta_a) = sqrt(1-z^2) za - z sqrt(1-za^2),
// by its series when small, else via acos.  Weights only.
__device__ __forceinline__ double dtheta_from(double z, double za, double sa, double theta_a) {
    const double sq = sqrt(fmax(0.0, (1.0 - z) * (1.0 + z)));
    const double x = fma(sq, za, -(z * sa));
    if (fabs(x) < 1e-2) {
        const double x2 = x * x;
        // asin(x) = x + x^3/6 + 3x^5/40 + 15x^7/336 + 105x^9/3456
        double p = fma(x2, 105.0 / 3456.0, 15.0 / 336.0);
        p = fma(x2, p, 3.0 / 40.0);
        p = fma(x2, p, 1.0 / 6.0);
        return fma(x * x2, p, x);
    }
    return acos(z) - theta_a;
}

// Double Rodrigues rotation, tod_rotations.py:122-133, binary64, same order.
// oma = 1 - ca and omp = 1 - cp_ are computed by the caller once per sample.
__device__ __forceinline__ void rodrigues(double vx_in, double vy_in, double vz_in,
                                          double kx, double ky, double kz, double ca,
                                          double sa, double oma, double px, double py,
                                          double pz, double cp_, double sp_, double omp,
                                          double& vx, double& vy, double& vz) {
    const double dkv = kx * vx_in + ky * vy_in + kz * vz_in;
    const double rx = vx_in * ca + (ky * vz_in - kz * vy_in) * sa + kx * dkv * oma;
    const double ry = vy_in * ca + (kz * vx_in - kx * vz_in) * sa + ky * dkv * oma;
    const double rz = vz_in * ca + (kx * vy_in - ky * vx_in) * sa + kz * dkv * oma;
    const double dpr = px * rx + py * ry + pz * rz;
    vx = rx * cp_ + (py * rz - pz * ry) * sp_ + px * dpr * omp;
    vy = ry * cp_ + (pz * rx - px * rz) * sp_ + py * dpr * omp;
    vz = rz * cp_ + (px * ry - py * rx) * sp_ + pz * dpr * omp;
}

// One bilinear neighbour: ring index, slot in the ring, global pixel, weight
// and the pixel centre (z, phi) used for spin-2.
struct Nb {
    int ring[4];
    int pos[4];
    double w[4];
};

// Bilinear neighbours (numba_healpy.py:358-499) from the global ring table:
// rings, slots and weights.  Pixel index = rt[ring].first + pos; pixel
// centre = (rt[ring].z, rt[ring].phi0 + pos * rt[ring].dphi).
__device__ __forceinline__ void ring_interp(int nside, const RingEntry* __restrict__ rt,
                                            double z, double phi_w, Nb& o) {
    const int ir_above = ring_above(nside, z);
    const int ir_below = ir_above + 1;
    if (ir_above != 0 && ir_below != 4 * nside) {
        const RingEntry& ra = rt[ir_above];
        const RingEntry& rb = rt[ir_below];
        const double dth = dtheta_from(z, ra.z, ra.sth, ra.theta);
        const double w_below = dth * ra.inv_dth;
        const double w_above = 1.0 - w_below;
        double fa, fb;
        const int ia = ring_slot(phi_w, ra.phi0, ra.dphi, ra.inv_dphi, ra.n, fa);
        const int ib = ring_slot(phi_w, rb.phi0, rb.dphi, rb.inv_dphi, rb.n, fb);
        o.ring[0] = o.ring[1] = ir_above;
        o.ring[2] = o.ring[3] = ir_below;
        o.pos[0] = ia;
        o.pos[1] = next_slot(ia, ra.n);
        o.pos[2] = ib;
        o.pos[3] = next_slot(ib, rb.n);
        o.w[0] = w_above * (1.0 - fa);
        o.w[1] = w_above * fa;
        o.w[2] = w_below * (1.0 - fb);
        o.w[3] = w_below * fb;
        return;
    }
    const double theta = acos(z);
    if (ir_above == 0) {
        // North-pole boundary: numba_healpy.py:384-411
        const RingEntry& r1 = rt[1];
        double frac;
        const int ip = ring_slot(phi_w, r1.phi0, r1.dphi, r1.inv_dphi, r1.n, frac);
        const int ip2 = next_slot(ip, r1.n);
        o.ring[0] = o.ring[1] = o.ring[2] = o.ring[3] = 1;
        o.pos[0] = (ip + 2) % r1.n;
        o.pos[1] = (ip2 + 2) % r1.n;
        o.pos[2] = ip;
        o.pos[3] = ip2;
        const double w_theta = theta / r1.theta;
        const double nf = (1.0 - w_theta) * 0.25;
        o.w[0] = nf;
        o.w[1] = nf;
        o.w[2] = (1.0 - frac) * w_theta + nf;
        o.w[3] = frac * w_theta + nf;
    } else {
        // South-pole boundary: numba_healpy.py:413-441
        const int irl = 4 * nside - 1;
        const RingEntry& rl = rt[irl];
        double frac;
        const int ip = ring_slot(phi_w, rl.phi0, rl.dphi, rl.inv_dphi, rl.n, frac);
        const int ip2 = next_slot(ip, rl.n);
        o.ring[0] = o.ring[1] = o.ring[2] = o.ring[3] = irl;
        o.pos[0] = ip;
        o.pos[1] = ip2;
        o.pos[2] = (ip + 2) % rl.n;
        o.pos[3] = (ip2 + 2) % rl.n;
        const double wts = (theta - rl.theta) / (kPi - rl.theta);
        const double sf = wts * 0.25;
        o.w[0] = (1.0 - frac) * (1.0 - wts) + sf;
        o.w[1] = frac * (1.0 - wts) + sf;
        o.w[2] = sf;
        o.w[3] = sf;
    }
}

// The nearest kernel's 4-candidate search: tod_nearest.py:112-141.
// z is not clamped here (tod_nearest.py:112); sin_th comes from the caller.
__device__ __forceinline__ void nearest_pix(int nside, const RingEntry* __restrict__ rt,
                                            double z, double sin_th, double phi_w,
                                            int& best_ring, int& best_pos) {
    int ir_above = ring_above(nside, z);
    if (ir_above < 1)
        ir_above = 1;
    else if (ir_above > 4 * nside - 2)
        ir_above = 4 * nside - 2;
    double bcos = -2.0;
    best_ring = ir_above;
    best_pos = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int ir = ir_above + k;
        const RingEntry& r = rt[ir];
        const double sin_z_c = sqrt(fmax(0.0, 1.0 - r.z * r.z));
        int ip_base = (int)(phi_w * (double)r.n / kTwoPi);
        if (ip_base >= r.n) ip_base -= r.n;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int ip_try = j == 0 ? ip_base : next_slot(ip_base, r.n);
            const double phi_c = r.phi0 + (double)ip_try * r.dphi;
            const double cos_d = sin_th * sin_z_c * cos(phi_w - phi_c) + z * r.z;
            if (cos_d > bcos) {
                bcos = cos_d;
                best_ring = ir;
                best_pos = ip_try;
            }
        }
    }
}


// ------------------------------------------------------------ map gathers

// One pixel of the packed pixel-major map (npix, CS).  The reference never
// bounds-checks; an index can only leave [0, npix) through the 1-ulp
// "tw == n" float-mod corner at the last ring, or from a non-finite
// pointing.  The address is clamped (the unsigned compare also catches
// negatives) so such cases read a valid pixel, never fault.
template <int CS>
__device__ __forceinline__ float4 load_px(const float* __restrict__ map, int64_t p, int64_t npix) {
    TODG_CHECK((uint64_t)p < (uint64_t)npix, kMapPixel);
    p = (uint64_t)p < (uint64_t)npix ? p : ((p < 0) ? 0 : npix - 1);
    if constexpr (CS == 4) {
        return *(reinterpret_cast<const float4*>(map) + p);
    } else if constexpr (CS == 2) {
        float2 v = *reinterpret_cast<const float2*>(map + 2 * p);
        return make_float4(v.x, v.y, 0.f, 0.f);
    } else {
        return make_float4(map[p], 0.f, 0.f, 0.f);
    }
}

// One pixel of the packed map by a 32-bit index (npix < 2^31), clamped to
// pmax = npix - 1 by one unsigned min: for indices that are in range by
// construction (the footprint sweep's laid-out cells), where the 64-bit
// clamp above would cost its compare-and-select chain per pixel.
template <int CS>
__device__ __forceinline__ float4 load_px_u32(const float* __restrict__ map, uint32_t p, uint32_t pmax) {
    TODG_CHECK(p <= pmax, kMapPixel);
    p = min(p, pmax);
    if constexpr (CS == 4) {
        return *(reinterpret_cast<const float4*>(map) + (uint64_t)p);
    } else if constexpr (CS == 2) {
        float2 v = *(reinterpret_cast<const float2*>(map) + (uint64_t)p);
        return make_float4(v.x, v.y, 0.f, 0.f);
    } else {
        return make_float4(map[(uint64_t)p], 0.f, 0.f, 0.f);
    }
}

// ------------------------------------------------------------ bicubic judge

// Keys cubic convolution kernel, a = -0.5 (Catmull-Rom), support |t| < 2.
__device__ __forceinline__ double keys_kernel(double t) {
    t = fabs(t);
    if (t < 1.0) return (1.5 * t - 2.5) * t * t + 1.0;
    if (t < 2.0) return ((-0.5 * t + 2.5) * t - 4.0) * t + 2.0;
    return 0.0;
}

// The bicubic judge's candidates and unnormalised weights for a query
// (z, phi): the reference's 7-ring x 5-pixel stencil (numba_healpy.py:938-976)
// around the query, each candidate weighted by K(x) K(y), where (x, y) are its
// gnomonic coordinates on the tangent plane at the query, x in units of the
// candidate's own ring step dphi*sin(theta), y in units of the ring spacing
// theta(ira+1) - theta(ira) at the query (README.md:120: Keys/Catmull-Rom
// weights via gnomonic projection; the reference's module itself is not in
// its snapshot).  Calls f(ring, slot, w) for every candidate with w != 0 and
// returns the sum of the weights.  sin/cos of the longitude gap come from
// sincos_any, so no absolute-angle trigonometry is needed.
template <class F>
__device__ __forceinline__ double bicubic_visit(int nside, const RingEntry* __restrict__ rt,
                                                double z, double phi, F&& f) {
    const double st = sqrt(fmax(0.0, (1.0 - z) * (1.0 + z)));
    int ira = ring_above(nside, z);
    ira = ira < 1 ? 1 : (ira > 4 * nside - 2 ? 4 * nside - 2 : ira);
    const double dth = rt[ira + 1].theta - rt[ira].theta;
    int irc = ring_above(nside, z);
    irc = irc < 1 ? 1 : (irc > 4 * nside - 1 ? 4 * nside - 1 : irc);
    const int r_lo = irc - 3 > 1 ? irc - 3 : 1;
    const int r_hi = irc + 3 < 4 * nside - 1 ? irc + 3 : 4 * nside - 1;
    double sum = 0.0;
    for (int ir = r_lo; ir <= r_hi; ++ir) {
        const RingEntry& e = rt[ir];
        const int n = e.n;
        const double step = e.dphi * e.sth;
        int cnt = n <= 4 ? n : 5;
        int ipc = 0;
        if (n > 4) {
            const int64_t fl = (int64_t)floor((phi - e.phi0) / e.dphi + 0.5);
            ipc = (int)(((fl % n) + n) % n);
        }
        for (int j = 0; j < cnt; ++j) {
            int ip = j;
            if (n > 4) {
                ip = ipc + j - 2;
                ip = ip < 0 ? ip + n : (ip >= n ? ip - n : ip);
            }
            double g = e.phi0 + (double)ip * e.dphi - phi;  // longitude gap
            g = g > kPi ? g - kTwoPi : (g < -kPi ? g + kTwoPi : g);
            double sg, cg;
            sincos_any(g, sg, cg);
            const double d = fma(e.sth * st, cg, e.z * z);
            if (!(d > 0.0 && step > 0.0 && dth > 0.0)) continue;
            const double x = e.sth * sg / d / step;
            const double y = fma(-z * e.sth, cg, e.z * st) / d / dth;
            const double w = keys_kernel(x) * keys_kernel(y);
            if (w != 0.0) {
                f(ir, ip, w);
                sum += w;
            }
        }
    }
    return sum;
}

}  // namespace todg
