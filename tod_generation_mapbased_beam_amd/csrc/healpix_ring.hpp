// HEALPix RING-scheme geometry on the device (gfx950).
//
// The index path (ring_above, per-ring phi slot, truncation) is evaluated in
// binary64 with the reference's operation order and no FMA contraction (this
// file is compiled with -ffp-contract=off), so for the same rotated direction
// it selects the same pixels as numba_healpy.py:53-114 and :358-499.  Per-ring
// constants (z, acos(z), phi0, dphi, first pixel, pixel count) come from a
// table built on the host with the same formulas and glibc's acos, so they
// are bit-identical to what the reference computes inline.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace todg {

constexpr double kTwoPi = 2.0 * 3.141592653589793;  // numba_healpy.py:42
constexpr double kPi = 3.141592653589793;
constexpr double kTwoThirds = 2.0 / 3.0;            // numba_healpy.py:44

// One row per ring index ir in [0, 4*nside]; rows 0 and 4*nside are unused
// padding so that ir can index directly.  48 bytes, 16-byte aligned.
struct alignas(16) RingEntry {
    double z;       // ring centre cos(theta)         numba_healpy.py:104-114
    double theta;   // acos(z), glibc                  numba_healpy.py:331-332
    double phi0;    // longitude of first pixel        numba_healpy.py:99-100
    double dphi;    // 2*pi / n                        numba_healpy.py:99
    int64_t first;  // first pixel index               numba_healpy.py:85-98
    int64_t n;      // pixels in the ring              numba_healpy.py:85-98
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

struct Bilin {
    int64_t p[4];
    double w[4];
    double zn[4];
    double phin[4];
};

// Bilinear neighbours, weights and neighbour (z, phi): numba_healpy.py:358-499.
// theta_q = acos(z) only feeds the weights (tolerance path).
__device__ __forceinline__ void ring_interp(int nside, const RingEntry* __restrict__ rt,
                                            double z, double phi_w, Bilin& o) {
    const int ir_above = ring_above(nside, z);
    const int ir_below = ir_above + 1;
    if (ir_above != 0 && ir_below != 4 * nside) {
        const RingEntry ra = rt[ir_above];
        const RingEntry rb = rt[ir_below];
        const double theta = acos(z);
        const double w_below = (theta - ra.theta) / (rb.theta - ra.theta);
        const double w_above = 1.0 - w_below;

        double tw = py_fmod_pos((phi_w - ra.phi0) / ra.dphi, (double)ra.n);
        int64_t ia = (int64_t)tw;
        double fa = tw - (double)ia;
        int64_t ia1 = (ia + 1) % ra.n;
        o.p[0] = ra.first + ia;
        o.p[1] = ra.first + ia1;
        o.w[0] = w_above * (1.0 - fa);
        o.w[1] = w_above * fa;

        tw = py_fmod_pos((phi_w - rb.phi0) / rb.dphi, (double)rb.n);
        int64_t ib = (int64_t)tw;
        double fb = tw - (double)ib;
        int64_t ib1 = (ib + 1) % rb.n;
        o.p[2] = rb.first + ib;
        o.p[3] = rb.first + ib1;
        o.w[2] = w_below * (1.0 - fb);
        o.w[3] = w_below * fb;

        o.zn[0] = ra.z;
        o.zn[1] = ra.z;
        o.zn[2] = rb.z;
        o.zn[3] = rb.z;
        o.phin[0] = ra.phi0 + (double)ia * ra.dphi;
        o.phin[1] = ra.phi0 + (double)ia1 * ra.dphi;
        o.phin[2] = rb.phi0 + (double)ib * rb.dphi;
        o.phin[3] = rb.phi0 + (double)ib1 * rb.dphi;
        return;
    }
    const double theta = acos(z);
    if (ir_above == 0) {
        // North-pole boundary: numba_healpy.py:384-411
        const RingEntry r1 = rt[1];
        double tw = py_fmod_pos((phi_w - r1.phi0) / r1.dphi, (double)r1.n);
        int64_t ip = (int64_t)tw;
        double frac = tw - (double)ip;
        int64_t ip2 = (ip + 1) % r1.n;
        int64_t q0 = (ip + 2) % r1.n, q1 = (ip2 + 2) % r1.n;
        o.p[0] = r1.first + q0;
        o.p[1] = r1.first + q1;
        o.p[2] = r1.first + ip;
        o.p[3] = r1.first + ip2;
        double w_theta = theta / r1.theta;
        double nf = (1.0 - w_theta) * 0.25;
        o.w[0] = nf;
        o.w[1] = nf;
        o.w[2] = (1.0 - frac) * w_theta + nf;
        o.w[3] = frac * w_theta + nf;
        o.zn[0] = o.zn[1] = o.zn[2] = o.zn[3] = r1.z;
        o.phin[0] = r1.phi0 + (double)q0 * r1.dphi;
        o.phin[1] = r1.phi0 + (double)q1 * r1.dphi;
        o.phin[2] = r1.phi0 + (double)ip * r1.dphi;
        o.phin[3] = r1.phi0 + (double)ip2 * r1.dphi;
    } else {
        // South-pole boundary: numba_healpy.py:413-441
        const RingEntry rl = rt[4 * nside - 1];
        double tw = py_fmod_pos((phi_w - rl.phi0) / rl.dphi, (double)rl.n);
        int64_t ip = (int64_t)tw;
        double frac = tw - (double)ip;
        int64_t ip2 = (ip + 1) % rl.n;
        int64_t q2 = (ip + 2) % rl.n, q3 = (ip2 + 2) % rl.n;
        o.p[0] = rl.first + ip;
        o.p[1] = rl.first + ip2;
        o.p[2] = rl.first + q2;
        o.p[3] = rl.first + q3;
        double wts = (theta - rl.theta) / (kPi - rl.theta);
        double sf = wts * 0.25;
        o.w[0] = (1.0 - frac) * (1.0 - wts) + sf;
        o.w[1] = frac * (1.0 - wts) + sf;
        o.w[2] = sf;
        o.w[3] = sf;
        o.zn[0] = o.zn[1] = o.zn[2] = o.zn[3] = rl.z;
        o.phin[0] = rl.phi0 + (double)ip * rl.dphi;
        o.phin[1] = rl.phi0 + (double)ip2 * rl.dphi;
        o.phin[2] = rl.phi0 + (double)q2 * rl.dphi;
        o.phin[3] = rl.phi0 + (double)q3 * rl.dphi;
    }
}

// The nearest kernel's 4-candidate search: tod_nearest.py:112-141.
// z is not clamped here (tod_nearest.py:112); sin_th comes from the caller.
__device__ __forceinline__ void nearest_pix(int nside, const RingEntry* __restrict__ rt,
                                            double z, double sin_th, double phi_w,
                                            int64_t& best_pix, double& best_z,
                                            double& best_phi) {
    int ir_above = ring_above(nside, z);
    if (ir_above < 1)
        ir_above = 1;
    else if (ir_above > 4 * nside - 2)
        ir_above = 4 * nside - 2;
    double bcos = -2.0;
    best_pix = 0;
    best_z = 0.0;
    best_phi = 0.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int ir = ir_above + k;
        const RingEntry r = rt[ir];
        const double sin_z_c = sqrt(fmax(0.0, 1.0 - r.z * r.z));
        const int64_t ip_base = ((int64_t)(phi_w * (double)r.n / kTwoPi)) % r.n;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t ip_try = j == 0 ? ip_base : (ip_base + 1) % r.n;
            const double phi_c = r.phi0 + (double)ip_try * r.dphi;
            const double cos_d = sin_th * sin_z_c * cos(phi_w - phi_c) + z * r.z;
            if (cos_d > bcos) {
                bcos = cos_d;
                best_pix = r.first + ip_try;
                best_z = r.z;
                best_phi = phi_c;
            }
        }
    }
}

// cos(2 delta), sin(2 delta) of the spin-2 frame rotation between a map pixel
// and the boresight: tod_bilinear.py:84-145, same formula and order.
__device__ __forceinline__ void spin2(double z_pix, double sth_pix, double phi_pix,
                                      double z_pts, double sth_pts, double phi_pts,
                                      double& c2d, double& s2d) {
    if (phi_pts == phi_pix && sth_pts == sth_pix && z_pts == z_pix) {
        c2d = 1.0;
        s2d = 0.0;
        return;
    }
    const double dphi = phi_pts - phi_pix;
    double sh, ch;
    sincos(dphi * 0.5, &sh, &ch);
    const double sin_dphi = 2.0 * ch * sh;
    const double cos_dphi = 1.0 - 2.0 * sh * sh;
    const double sh2 = sh * sh;
    const double dz = z_pts - z_pix;
    const double ds = sth_pts - sth_pix;
    const double st2 = 0.25 * (dz * dz + ds * ds);
    const double S = sth_pix * sth_pts;
    const double h = st2 + S * sh2;
    const double sin2_dtheta = 4.0 * st2 * (1.0 - st2);
    const double z_sum = z_pts + z_pix;
    const double N = 2.0 * sin_dphi * z_sum * h;
    const double C = z_pts * z_pix;
    const double term1 = sin2_dtheta * cos_dphi;
    const double term2 = 4.0 * S * C * sh2 * sh2;
    const double term3 = S * sin_dphi * sin_dphi;
    const double D = term1 - term2 + term3;
    const double u = 1.0 / (N * N + D * D);
    c2d = (D * D - N * N) * u;
    s2d = 2.0 * N * D * u;
}

// Double Rodrigues rotation, tod_rotations.py:122-133, binary64, same order.
__device__ __forceinline__ void rodrigues(double vx_in, double vy_in, double vz_in,
                                          double kx, double ky, double kz, double ca,
                                          double sa, double px, double py, double pz,
                                          double cp_, double sp_, double& vx, double& vy,
                                          double& vz) {
    const double oma = 1.0 - ca;
    const double omp = 1.0 - cp_;
    const double dkv = kx * vx_in + ky * vy_in + kz * vz_in;
    const double rx = vx_in * ca + (ky * vz_in - kz * vy_in) * sa + kx * dkv * oma;
    const double ry = vy_in * ca + (kz * vx_in - kx * vz_in) * sa + ky * dkv * oma;
    const double rz = vz_in * ca + (kx * vy_in - ky * vx_in) * sa + kz * dkv * oma;
    const double dpr = px * rx + py * ry + pz * rz;
    vx = rx * cp_ + (py * rz - pz * ry) * sp_ + px * dpr * omp;
    vy = ry * cp_ + (pz * rx - px * rz) * sp_ + py * dpr * omp;
    vz = rz * cp_ + (px * ry - py * rx) * sp_ + pz * dpr * omp;
}

}  // namespace todg
