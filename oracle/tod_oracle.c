/*
 * tod_oracle.c -- CPU restatement of the reference's beam-convolution TOD hot
 * path.  TEST INFRASTRUCTURE ONLY: nothing in the product links, loads or
 * calls this file.  It is the checker that the HIP kernels are compared with
 * (tests/, __graft_entry__.smoke(), and the cpu_baseline leg of bench.py).
 *
 * Arithmetic contract (what makes it a faithful restatement of the Numba
 * code): IEEE binary64, glibc libm (acos, atan2, cos, sin, sqrt, fmod), built
 * with -O2 -ffp-contract=off and no fast-math, every expression evaluated in
 * the reference's source order (Python evaluates a+b+c as (a+b)+c).  Python
 * semantics reproduced exactly:
 *   - float % : r = fmod(a, n); if r != 0 and sign(r) != sign(n): r += n;
 *               r == 0 -> +0.0 (Numba lowers it as CPython's float_divmod);
 *   - int(x)  : truncation toward zero;
 *   - max(-1, min(1, v)) : CPython's first-wins comparison order.
 *
 * The reference's per-sample spin-2 cache (tod_bilinear.py:151-199) is keyed
 * by pixel index, and the stored factors are a deterministic function of the
 * neighbour's (z, phi), which the pixel index fixes -- except in the float-mod
 * corner where the slot coordinate rounds to exactly n (see py_fmod below):
 * the out-of-ring index first + n is then the next ring's first pixel, whose
 * cached factors (if that pixel was met first in the sample) the reference
 * would return for a neighbour that carries the current ring's (z, phi).
 * The cache is left out here: every factor is recomputed from the
 * neighbour's own (z, phi), which equals the reference bit for bit except in
 * that ~1-ulp corner.
 *
 * Citations are to the reference repository (file:line).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_TWO_PI (2.0 * 3.141592653589793)   /* numba_healpy.py:42 */
#define OR_PI 3.141592653589793
#define OR_TWO_THIRDS (2.0 / 3.0)             /* numba_healpy.py:44 */

/* ---------------------------------------------------------------- helpers */

/* Python float modulo as lowered by Numba (CPython float_divmod). */
static inline double py_fmod(double a, double n)
{
    double r = fmod(a, n);
    if (r != 0.0) {
        if ((n < 0.0) != (r < 0.0))
            r += n;
    } else {
        r = (n < 0.0) ? -0.0 : 0.0;
    }
    return r;
}

static inline double py_min(double a, double b) { return (b < a) ? b : a; }
static inline double py_max(double a, double b) { return (b > a) ? b : a; }

/* --------------------------------------------- ring geometry (RING scheme) */

/* numba_healpy.py:53-70 */
int64_t or_ring_above(int64_t nside, double z)
{
    double az = fabs(z);
    int64_t ir;
    if (az > OR_TWO_THIRDS) {
        double tp = (double)nside * sqrt(3.0 * (1.0 - az));
        ir = (int64_t)tp;
        if (z < 0.0)
            ir = 4 * nside - ir - 1;
    } else {
        ir = (int64_t)((double)nside * (2.0 - 1.5 * z));
    }
    return ir;
}

/* numba_healpy.py:73-101 */
void or_ring_info(int64_t nside, int64_t ir, int64_t npix_total, int64_t *n_pix,
                  int64_t *first_pix, double *phi0, double *dphi)
{
    int64_t n, fp, s;
    if (ir < nside) {
        n = 4 * ir;
        fp = 2 * ir * (ir - 1);
        s = 1;
    } else if (ir <= 3 * nside) {
        n = 4 * nside;
        fp = 2 * nside * (nside - 1) + (ir - nside) * 4 * nside;
        s = ((ir - nside) % 2 == 0) ? 1 : 0;
    } else {
        int64_t i2 = 4 * nside - ir;
        n = 4 * i2;
        fp = npix_total - 2 * i2 * (i2 + 1);
        s = 1;
    }
    double dp = OR_TWO_PI / (double)n;
    *n_pix = n;
    *first_pix = fp;
    *dphi = dp;
    *phi0 = (double)s * dp * 0.5;
}

/* numba_healpy.py:104-114 */
double or_ring_z(int64_t nside, int64_t ir)
{
    if (ir < nside) {
        double tmp = (double)ir;
        return 1.0 - tmp * tmp / (3.0 * (double)nside * (double)nside);
    } else if (ir <= 3 * nside) {
        return (2.0 / 3.0) * (2.0 - (double)ir / (double)nside);
    } else {
        double tmp = (double)(4 * nside - ir);
        return -(1.0 - tmp * tmp / (3.0 * (double)nside * (double)nside));
    }
}

/* ------------------------------------------- bilinear neighbour lookup */

/* numba_healpy.py:358-499 (_ring_interp_with_angles_jit).  The pixels and
 * weights are those of _ring_interp_single_jit (numba_healpy.py:259-355); the
 * reference asserts the two agree bit for bit. */
void or_ring_interp_with_angles(int64_t nside, double z, double phi_w,
                                int64_t npix_total, int64_t *p, double *w,
                                double *zn, double *phin)
{
    int64_t ir_above = or_ring_above(nside, z);
    int64_t ir_below = ir_above + 1;
    int64_t na, fpa, nb, fpb;
    double phi0a, dphia, phi0b, dphib;

    if (ir_above == 0) {
        or_ring_info(nside, 1, npix_total, &na, &fpa, &phi0a, &dphia);
        double tw = py_fmod((phi_w - phi0a) / dphia, (double)na);
        int64_t ip_a = (int64_t)tw;
        double frac = tw - (double)ip_a;
        int64_t ip_a2 = (ip_a + 1) % na;
        p[0] = fpa + (ip_a + 2) % na;
        p[1] = fpa + (ip_a2 + 2) % na;
        p[2] = fpa + ip_a;
        p[3] = fpa + ip_a2;
        double za = or_ring_z(nside, 1);
        double ta = acos(za);
        double theta = acos(z);
        double w_theta = theta / ta;
        double nf = (1.0 - w_theta) * 0.25;
        w[0] = nf;
        w[1] = nf;
        w[2] = (1.0 - frac) * w_theta + nf;
        w[3] = frac * w_theta + nf;
        zn[0] = zn[1] = zn[2] = zn[3] = za;
        phin[0] = phi0a + (double)((ip_a + 2) % na) * dphia;
        phin[1] = phi0a + (double)((ip_a2 + 2) % na) * dphia;
        phin[2] = phi0a + (double)ip_a * dphia;
        phin[3] = phi0a + (double)ip_a2 * dphia;
    } else if (ir_below == 4 * nside) {
        int64_t ir_last = 4 * nside - 1;
        or_ring_info(nside, ir_last, npix_total, &na, &fpa, &phi0a, &dphia);
        double tw = py_fmod((phi_w - phi0a) / dphia, (double)na);
        int64_t ip_a = (int64_t)tw;
        double frac = tw - (double)ip_a;
        int64_t ip_a2 = (ip_a + 1) % na;
        p[0] = fpa + ip_a;
        p[1] = fpa + ip_a2;
        p[2] = fpa + (ip_a + 2) % na;
        p[3] = fpa + (ip_a2 + 2) % na;
        double za = or_ring_z(nside, ir_last);
        double ta = acos(za);
        double theta = acos(z);
        double wts = (theta - ta) / (OR_PI - ta);
        double sf = wts * 0.25;
        w[0] = (1.0 - frac) * (1.0 - wts) + sf;
        w[1] = frac * (1.0 - wts) + sf;
        w[2] = sf;
        w[3] = sf;
        zn[0] = zn[1] = zn[2] = zn[3] = za;
        phin[0] = phi0a + (double)ip_a * dphia;
        phin[1] = phi0a + (double)ip_a2 * dphia;
        phin[2] = phi0a + (double)((ip_a + 2) % na) * dphia;
        phin[3] = phi0a + (double)((ip_a2 + 2) % na) * dphia;
    } else {
        double za = or_ring_z(nside, ir_above);
        double zb = or_ring_z(nside, ir_below);
        double ta = acos(za);
        double tb = acos(zb);
        double theta = acos(z);
        double w_below = (theta - ta) / (tb - ta);
        double w_above = 1.0 - w_below;

        or_ring_info(nside, ir_above, npix_total, &na, &fpa, &phi0a, &dphia);
        double tw = py_fmod((phi_w - phi0a) / dphia, (double)na);
        int64_t iphia = (int64_t)tw;
        double fphia = tw - (double)iphia;
        int64_t iphia1 = (iphia + 1) % na;
        p[0] = fpa + iphia;
        p[1] = fpa + iphia1;
        w[0] = w_above * (1.0 - fphia);
        w[1] = w_above * fphia;

        or_ring_info(nside, ir_below, npix_total, &nb, &fpb, &phi0b, &dphib);
        tw = py_fmod((phi_w - phi0b) / dphib, (double)nb);
        int64_t iphib = (int64_t)tw;
        double fphib = tw - (double)iphib;
        int64_t iphib1 = (iphib + 1) % nb;
        p[2] = fpb + iphib;
        p[3] = fpb + iphib1;
        w[2] = w_below * (1.0 - fphib);
        w[3] = w_below * fphib;

        zn[0] = za;
        zn[1] = za;
        zn[2] = zb;
        zn[3] = zb;
        phin[0] = phi0a + (double)iphia * dphia;
        phin[1] = phi0a + (double)iphia1 * dphia;
        phin[2] = phi0b + (double)iphib * dphib;
        phin[3] = phi0b + (double)iphib1 * dphib;
    }
}

/* numba_healpy.py:120-236 (_get_interp_weights_jit): theta-driven variant of
 * the same lookup, as used by the reference's split pipeline. */
void or_get_interp_weights(int64_t nside, int64_t N, const double *theta_arr,
                           const double *phi_arr, int64_t *pix_out, double *wgt_out)
{
    int64_t npix_total = 12 * nside * nside;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double theta = theta_arr[i];
        double phi = phi_arr[i];
        double z = cos(theta);
        int64_t ir_above = or_ring_above(nside, z);
        int64_t ir_below = ir_above + 1;
        int64_t na, fpa, nb, fpb;
        double phi0a, dphia, phi0b, dphib;
        if (ir_above == 0) {
            or_ring_info(nside, 1, npix_total, &na, &fpa, &phi0a, &dphia);
            double tw = py_fmod((phi - phi0a) / dphia, (double)na);
            int64_t ip = (int64_t)tw;
            double frac = tw - (double)ip;
            int64_t ip2 = (ip + 1) % na;
            double za = or_ring_z(nside, 1);
            double ta = acos(za);
            double w_theta = theta / ta;
            double nf = (1.0 - w_theta) * 0.25;
            pix_out[0 * N + i] = (ip + 2) % na;
            pix_out[1 * N + i] = (ip2 + 2) % na;
            pix_out[2 * N + i] = fpa + ip;
            pix_out[3 * N + i] = fpa + ip2;
            wgt_out[0 * N + i] = nf;
            wgt_out[1 * N + i] = nf;
            wgt_out[2 * N + i] = (1.0 - frac) * w_theta + nf;
            wgt_out[3 * N + i] = frac * w_theta + nf;
        } else if (ir_below == 4 * nside) {
            int64_t ir_last = 4 * nside - 1;
            or_ring_info(nside, ir_last, npix_total, &na, &fpa, &phi0a, &dphia);
            double tw = py_fmod((phi - phi0a) / dphia, (double)na);
            int64_t ip = (int64_t)tw;
            double frac = tw - (double)ip;
            int64_t ip2 = (ip + 1) % na;
            double za = or_ring_z(nside, ir_last);
            double ta = acos(za);
            double wts = (theta - ta) / (OR_PI - ta);
            double sf = wts * 0.25;
            pix_out[0 * N + i] = fpa + ip;
            pix_out[1 * N + i] = fpa + ip2;
            pix_out[2 * N + i] = (ip + 2) % na + fpa;
            pix_out[3 * N + i] = (ip2 + 2) % na + fpa;
            wgt_out[0 * N + i] = (1.0 - frac) * (1.0 - wts) + sf;
            wgt_out[1 * N + i] = frac * (1.0 - wts) + sf;
            wgt_out[2 * N + i] = sf;
            wgt_out[3 * N + i] = sf;
        } else {
            double za = or_ring_z(nside, ir_above);
            double zb = or_ring_z(nside, ir_below);
            double ta = acos(za);
            double tb = acos(zb);
            double w_below = (theta - ta) / (tb - ta);
            double w_above = 1.0 - w_below;
            or_ring_info(nside, ir_above, npix_total, &na, &fpa, &phi0a, &dphia);
            double tw = py_fmod((phi - phi0a) / dphia, (double)na);
            int64_t iphia = (int64_t)tw;
            double fphia = tw - (double)iphia;
            pix_out[0 * N + i] = fpa + iphia;
            pix_out[1 * N + i] = fpa + (iphia + 1) % na;
            wgt_out[0 * N + i] = w_above * (1.0 - fphia);
            wgt_out[1 * N + i] = w_above * fphia;
            or_ring_info(nside, ir_below, npix_total, &nb, &fpb, &phi0b, &dphib);
            tw = py_fmod((phi - phi0b) / dphib, (double)nb);
            int64_t iphib = (int64_t)tw;
            double fphib = tw - (double)iphib;
            pix_out[2 * N + i] = fpb + iphib;
            pix_out[3 * N + i] = fpb + (iphib + 1) % nb;
            wgt_out[2 * N + i] = w_below * (1.0 - fphib);
            wgt_out[3 * N + i] = w_below * fphib;
        }
    }
}

/* ------------------------------------------------ pix2ang / nearest pixel */

/* numba_healpy.py:553-580 (_pix2zphi_ring_jit) */
void or_pix2zphi(int64_t nside, int64_t pix, double *z, double *phi)
{
    int64_t npix = 12 * nside * nside;
    int64_t ncap = 2 * nside * (nside - 1);
    int64_t iring, ip_in;
    if (pix < ncap) {
        iring = (int64_t)(0.5 * (1.0 + sqrt(1.0 + 2.0 * (double)pix)));
        ip_in = pix - 2 * iring * (iring - 1);
    } else if (pix < npix - ncap) {
        int64_t ip = pix - ncap;
        iring = ip / (4 * nside) + nside;
        ip_in = ip % (4 * nside);
    } else {
        int64_t ip_s = npix - pix - 1;
        int64_t iring_s = (int64_t)(0.5 * (1.0 + sqrt(1.0 + 2.0 * (double)ip_s)));
        iring = 4 * nside - iring_s;
        int64_t first_s = npix - 2 * iring_s * (iring_s + 1);
        ip_in = pix - first_s;
    }
    int64_t n, fp;
    double phi0, dphi;
    or_ring_info(nside, iring, npix, &n, &fp, &phi0, &dphi);
    *phi = phi0 + (double)ip_in * dphi;
    *z = or_ring_z(nside, iring);
}

/* The reference's 4-candidate nearest-pixel search as inlined in
 * tod_nearest.py:112-141: two rings bracketing z (ring_above clamped to
 * [1, 4n-2]), slots ip and ip+1 with ip = int(phi*n/2pi) % n, maximum of
 * cos(angular distance), strict '>' so the first candidate wins ties.  z is
 * NOT clamped on this path and sin(theta) comes from the caller. */
void or_nearest_pix(int64_t nside, double z, double sin_th, double phi_w,
                    int64_t npix_total, int64_t *best_pix, double *best_z_c,
                    double *best_phi_c)
{
    int64_t ir_above = or_ring_above(nside, z);
    if (ir_above < 1)
        ir_above = 1;
    else if (ir_above > 4 * nside - 2)
        ir_above = 4 * nside - 2;
    int64_t rings[2] = {ir_above, ir_above + 1};
    int64_t bp = 0;
    double bcos = -2.0, bz = 0.0, bphi = 0.0;
    for (int k = 0; k < 2; ++k) {
        int64_t ir_g = rings[k];
        if (ir_g < 1 || ir_g > 4 * nside - 1)
            continue;
        int64_t n_pix, first_pix;
        double phi0, dphi_r;
        or_ring_info(nside, ir_g, npix_total, &n_pix, &first_pix, &phi0, &dphi_r);
        double z_c = or_ring_z(nside, ir_g);
        double sin_z_c = sqrt(py_max(0.0, 1.0 - z_c * z_c));
        int64_t ip_base = ((int64_t)(phi_w * (double)n_pix / OR_TWO_PI)) % n_pix;
        int64_t tries[2] = {ip_base, (ip_base + 1) % n_pix};
        for (int j = 0; j < 2; ++j) {
            int64_t ip_try = tries[j];
            double phi_c = phi0 + (double)ip_try * dphi_r;
            double cos_d = sin_th * sin_z_c * cos(phi_w - phi_c) + z * z_c;
            if (cos_d > bcos) {
                bcos = cos_d;
                bp = first_pix + ip_try;
                bz = z_c;
                bphi = phi_c;
            }
        }
    }
    *best_pix = bp;
    *best_z_c = bz;
    *best_phi_c = bphi;
}

/* numba_healpy.py:892-977 (_gather_ring_stencil_jit): up to 7 rings x 5
 * pixels around the query (all 4 pixels of a ring with n <= 4). */
int64_t or_gather_ring_stencil(int64_t nside, double vz, double ph, int64_t *out_buf,
                               double *z_buf, double *phi_buf)
{
    int64_t npix_total = 12 * nside * nside;
    int64_t ir_center = or_ring_above(nside, vz);
    if (ir_center < 1)
        ir_center = 1;
    else if (ir_center > 4 * nside - 1)
        ir_center = 4 * nside - 1;
    int64_t ir_lo = ir_center - 3 > 1 ? ir_center - 3 : 1;
    int64_t ir_hi = ir_center + 3 < 4 * nside - 1 ? ir_center + 3 : 4 * nside - 1;
    int64_t count = 0;
    for (int64_t ir = ir_lo; ir <= ir_hi; ++ir) {
        int64_t n_p, fp;
        double phi0, dphi;
        or_ring_info(nside, ir, npix_total, &n_p, &fp, &phi0, &dphi);
        double z_ring = or_ring_z(nside, ir);
        if (n_p <= 4) {
            for (int64_t ip = 0; ip < n_p; ++ip) {
                out_buf[count] = fp + ip;
                z_buf[count] = z_ring;
                phi_buf[count] = phi0 + (double)ip * dphi;
                ++count;
            }
        } else {
            int64_t f = (int64_t)floor((ph - phi0) / dphi + 0.5);
            int64_t ip_center = ((f % n_p) + n_p) % n_p; /* Python int % */
            for (int64_t dip = -2; dip <= 2; ++dip) {
                int64_t ip_in = (((ip_center + dip) % n_p) + n_p) % n_p;
                out_buf[count] = fp + ip_in;
                z_buf[count] = z_ring;
                phi_buf[count] = phi0 + (double)ip_in * dphi;
                ++count;
            }
        }
    }
    return count;
}

/* ----------------------------------------------------------- rotations */

/* tod_rotations.py:83-134 (_rodrigues_apply_one_jit), all binary64. */
static inline void rodrigues_one(double vx_in, double vy_in, double vz_in, double kx,
                                 double ky, double kz, double ca, double sa, double px,
                                 double py, double pz, double cp_, double sp_, double *vx,
                                 double *vy, double *vz)
{
    double oma = 1.0 - ca;
    double omp = 1.0 - cp_;
    double dkv = kx * vx_in + ky * vy_in + kz * vz_in;
    double rx = vx_in * ca + (ky * vz_in - kz * vy_in) * sa + kx * dkv * oma;
    double ry = vy_in * ca + (kz * vx_in - kx * vz_in) * sa + ky * dkv * oma;
    double rz = vz_in * ca + (kx * vy_in - ky * vx_in) * sa + kz * dkv * oma;
    double dpr = px * rx + py * ry + pz * rz;
    *vx = rx * cp_ + (py * rz - pz * ry) * sp_ + px * dpr * omp;
    *vy = ry * cp_ + (pz * rx - px * rz) * sp_ + py * dpr * omp;
    *vz = rz * cp_ + (px * ry - py * rx) * sp_ + pz * dpr * omp;
}

void or_rodrigues_apply_one(const double *in13, double *out3)
{
    rodrigues_one(in13[0], in13[1], in13[2], in13[3], in13[4], in13[5], in13[6], in13[7],
                  in13[8], in13[9], in13[10], in13[11], in13[12], &out3[0], &out3[1],
                  &out3[2]);
}

/* tod_rotations.py:33-80 (_rodrigues_jit).  Numba types: float32 arrays,
 * float64 literals, so 1.0 - ca is binary64, kx*dkv is binary32 and any
 * binary32 x binary64 product is binary64; the output rounds to binary32. */
void or_rodrigues_batch(const float *vec_orig, const float *axes, const float *cos_a,
                        const float *sin_a, const float *ax_pts, const float *cos_p,
                        const float *sin_p, int64_t B, int64_t S, float *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < B; ++b) {
        float kx = axes[3 * b], ky = axes[3 * b + 1], kz = axes[3 * b + 2];
        float ca = cos_a[b], sa = sin_a[b];
        double oma = 1.0 - (double)ca;
        float px = ax_pts[3 * b], py = ax_pts[3 * b + 1], pz = ax_pts[3 * b + 2];
        float cp_ = cos_p[b], sp_ = sin_p[b];
        double omp = 1.0 - (double)cp_;
        for (int64_t s = 0; s < S; ++s) {
            float vx = vec_orig[3 * s], vy = vec_orig[3 * s + 1], vz = vec_orig[3 * s + 2];
            float dkv = kx * vx + ky * vy + kz * vz;
            float t0 = vx * ca;
            float t1 = (ky * vz - kz * vy) * sa;
            float t2 = kx * dkv;
            double rx = (double)(t0 + t1) + (double)t2 * oma;
            t0 = vy * ca;
            t1 = (kz * vx - kx * vz) * sa;
            t2 = ky * dkv;
            double ry = (double)(t0 + t1) + (double)t2 * oma;
            t0 = vz * ca;
            t1 = (kx * vy - ky * vx) * sa;
            t2 = kz * dkv;
            double rz = (double)(t0 + t1) + (double)t2 * oma;
            double dpr = (double)px * rx + (double)py * ry + (double)pz * rz;
            double ox = rx * (double)cp_ + ((double)py * rz - (double)pz * ry) * (double)sp_ +
                        (double)px * dpr * omp;
            double oy = ry * (double)cp_ + ((double)pz * rx - (double)px * rz) * (double)sp_ +
                        (double)py * dpr * omp;
            double oz = rz * (double)cp_ + ((double)px * ry - (double)py * rx) * (double)sp_ +
                        (double)pz * dpr * omp;
            float *o = out + 3 * (b * S + s);
            o[0] = (float)ox;
            o[1] = (float)oy;
            o[2] = (float)oz;
        }
    }
}

/* ----------------------------------------------------------------- spin-2 */

/* tod_bilinear.py:84-145 (_spin2_cos2d_sin2d_jit) */
void or_spin2(double z_pix, double sth_pix, double phi_pix, double z_pts, double sth_pts,
              double phi_pts, double *c2d, double *s2d)
{
    if (phi_pts == phi_pix && sth_pts == sth_pix && z_pts == z_pix) {
        *c2d = 1.0;
        *s2d = 0.0;
        return;
    }
    double dphi = phi_pts - phi_pix;
    double ch = cos(dphi * 0.5);
    double sh = sin(dphi * 0.5);
    double sin_dphi = 2.0 * ch * sh;
    double cos_dphi = 1.0 - 2.0 * sh * sh;
    double sh2 = sh * sh;
    double dz = z_pts - z_pix;
    double ds = sth_pts - sth_pix;
    double st2 = 0.25 * (dz * dz + ds * ds);
    double S = sth_pix * sth_pts;
    double h = st2 + S * sh2;
    double sin2_dtheta = 4.0 * st2 * (1.0 - st2);
    double z_sum = z_pts + z_pix;
    double N = 2.0 * sin_dphi * z_sum * h;
    double C = z_pts * z_pix;
    double term1 = sin2_dtheta * cos_dphi;
    double term2 = 4.0 * S * C * sh2 * sh2;
    double term3 = S * sin_dphi * sin_dphi;
    double D = term1 - term2 + term3;
    double u = 1.0 / (N * N + D * D);
    *c2d = (D * D - N * N) * u;
    *s2d = 2.0 * N * D * u;
}

static inline void spin2_neighbour(double z_n, double phi_n, double z_pts, double sth_pts,
                                   double phi_pts, double *c2d, double *s2d)
{
    /* tod_bilinear.py:194-195 */
    double sth_n = sqrt(py_max(0.0, 1.0 - z_n * z_n));
    or_spin2(z_n, sth_n, phi_n, z_pts, sth_pts, phi_pts, c2d, s2d);
}

/* ------------------------------------------------------ fused kernels */

typedef struct {
    double kx, ky, kz, ca, sa, bx, by, bz, cp, sp;
} rot_params;

static inline rot_params load_rot(const float *axes, const float *cos_a, const float *sin_a,
                                  const float *ax_pts, const float *cos_p, const float *sin_p,
                                  int64_t b)
{
    rot_params r;
    r.kx = axes[3 * b];
    r.ky = axes[3 * b + 1];
    r.kz = axes[3 * b + 2];
    r.ca = cos_a[b];
    r.sa = sin_a[b];
    r.bx = ax_pts[3 * b];
    r.by = ax_pts[3 * b + 1];
    r.bz = ax_pts[3 * b + 2];
    r.cp = cos_p[b];
    r.sp = sin_p[b];
    return r;
}

static inline void rotate_sample(const rot_params *r, const float *vec_orig, int64_t s,
                                 double *vx, double *vy, double *vz)
{
    rodrigues_one((double)vec_orig[3 * s], (double)vec_orig[3 * s + 1],
                  (double)vec_orig[3 * s + 2], r->kx, r->ky, r->kz, r->ca, r->sa, r->bx,
                  r->by, r->bz, r->cp, r->sp, vx, vy, vz);
}

static inline double wrap_atan2(double y, double x)
{
    double p = atan2(y, x);
    if (p < 0.0)
        p += OR_TWO_PI;
    return p;
}

/* tod_bilinear.py:205-477 (_gather_accum_fused_jit).  mp_stacked is the
 * reference's (C, npix) field-major float32 map; tod (C, B) is accumulated
 * in place.  OpenMP over b mirrors numba.prange; each b is summed over s in
 * order, so the result does not depend on the thread count. */
void or_gather_accum_fused(const float *vec_orig, const float *axes, const float *cos_a,
                           const float *sin_a, const float *ax_pts, const float *cos_p,
                           const float *sin_p, int64_t nside, const float *mp_stacked,
                           int64_t C, const float *beam_vals, int64_t B, int64_t S,
                           double *tod, int64_t c_q, int64_t c_u, int nthreads)
{
    int has_qu = (c_q >= 0 && c_u >= 0);
    int64_t npix_total = 12 * nside * nside;
    int64_t other[8];
    int64_t n_other = 0;
    for (int64_t c = 0; c < C && c < 8; ++c)
        if (c != c_q && c != c_u)
            other[n_other++] = c;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t b = 0; b < B; ++b) {
        rot_params r = load_rot(axes, cos_a, sin_a, ax_pts, cos_p, sin_p, b);
        if (has_qu) {
            double z_pts = py_max(-1.0, py_min(1.0, r.bz));
            double sth_pts = sqrt(py_max(0.0, 1.0 - r.bz * r.bz));
            double phi_pts = wrap_atan2(r.by, r.bx);
            /* acc[c] stands for tod[c, b]; indexing by channel keeps the
             * reference's semantics when c_q == c_u (both += on one row). */
            double acc[8];
            for (int64_t c = 0; c < C && c < 8; ++c)
                acc[c] = tod[c * B + b];
            for (int64_t s = 0; s < S; ++s) {
                double vx, vy, vz;
                rotate_sample(&r, vec_orig, s, &vx, &vy, &vz);
                double z = py_max(-1.0, py_min(1.0, vz));
                double phi_w = wrap_atan2(vy, vx);
                double bv = (double)beam_vals[s];
                int64_t p[4];
                double w[4], zn[4], phin[4], c2[4], s2[4];
                or_ring_interp_with_angles(nside, z, phi_w, npix_total, p, w, zn, phin);
                for (int k = 0; k < 4; ++k)
                    spin2_neighbour(zn[k], phin[k], z_pts, sth_pts, phi_pts, &c2[k], &s2[k]);
                double q[4], u[4];
                for (int k = 0; k < 4; ++k) {
                    q[k] = (double)mp_stacked[c_q * npix_total + p[k]];
                    u[k] = (double)mp_stacked[c_u * npix_total + p[k]];
                }
                acc[c_q] += (w[0] * (q[0] * c2[0] + u[0] * s2[0]) + w[1] * (q[1] * c2[1] + u[1] * s2[1]) +
                             w[2] * (q[2] * c2[2] + u[2] * s2[2]) + w[3] * (q[3] * c2[3] + u[3] * s2[3])) *
                            bv;
                acc[c_u] += (w[0] * (-q[0] * s2[0] + u[0] * c2[0]) +
                             w[1] * (-q[1] * s2[1] + u[1] * c2[1]) +
                             w[2] * (-q[2] * s2[2] + u[2] * c2[2]) +
                             w[3] * (-q[3] * s2[3] + u[3] * c2[3])) *
                            bv;
                for (int64_t oi = 0; oi < n_other; ++oi) {
                    const float *m = mp_stacked + other[oi] * npix_total;
                    acc[other[oi]] += (w[0] * (double)m[p[0]] + w[1] * (double)m[p[1]] +
                                       w[2] * (double)m[p[2]] + w[3] * (double)m[p[3]]) *
                                      bv;
                }
            }
            for (int64_t c = 0; c < C && c < 8; ++c)
                tod[c * B + b] = acc[c];
        } else {
            double acc[8];
            for (int64_t c = 0; c < C && c < 8; ++c)
                acc[c] = tod[c * B + b];
            for (int64_t s = 0; s < S; ++s) {
                double vx, vy, vz;
                rotate_sample(&r, vec_orig, s, &vx, &vy, &vz);
                double z = py_max(-1.0, py_min(1.0, vz));
                double phi_w = wrap_atan2(vy, vx);
                int64_t p[4];
                double w[4], zn[4], phin[4];
                or_ring_interp_with_angles(nside, z, phi_w, npix_total, p, w, zn, phin);
                double bv = (double)beam_vals[s];
                for (int64_t c = 0; c < C && c < 8; ++c) {
                    const float *m = mp_stacked + c * npix_total;
                    acc[c] += (w[0] * (double)m[p[0]] + w[1] * (double)m[p[1]] +
                               w[2] * (double)m[p[2]] + w[3] * (double)m[p[3]]) *
                              bv;
                }
            }
            for (int64_t c = 0; c < C && c < 8; ++c)
                tod[c * B + b] = acc[c];
        }
    }
}

/* tod_nearest.py:21-158 (_gather_accum_nearest_jit) */
void or_gather_accum_nearest(const float *vec_orig, const float *axes, const float *cos_a,
                             const float *sin_a, const float *ax_pts, const float *cos_p,
                             const float *sin_p, int64_t nside, const float *mp_stacked,
                             int64_t C, const float *beam_vals, int64_t B, int64_t S,
                             double *tod, int64_t c_q, int64_t c_u, int nthreads)
{
    int has_qu = (c_q >= 0 && c_u >= 0);
    int64_t npix_total = 12 * nside * nside;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t b = 0; b < B; ++b) {
        rot_params r = load_rot(axes, cos_a, sin_a, ax_pts, cos_p, sin_p, b);
        double bz_pts = py_max(-1.0, py_min(1.0, r.bz));
        double bsth_pts = sqrt(py_max(0.0, 1.0 - r.bz * r.bz));
        double bphi_pts = wrap_atan2(r.by, r.bx);
        double acc[8];
        for (int64_t c = 0; c < C && c < 8; ++c)
            acc[c] = tod[c * B + b];
        for (int64_t s = 0; s < S; ++s) {
            double vx, vy, vz;
            rotate_sample(&r, vec_orig, s, &vx, &vy, &vz);
            double phi_w = wrap_atan2(vy, vx);
            double z = vz;
            double sin_th = sqrt(py_max(0.0, vx * vx + vy * vy));
            int64_t bp;
            double bzc, bphic;
            or_nearest_pix(nside, z, sin_th, phi_w, npix_total, &bp, &bzc, &bphic);
            double bv = (double)beam_vals[s];
            if (!has_qu) {
                for (int64_t c = 0; c < C && c < 8; ++c)
                    acc[c] += (double)mp_stacked[c * npix_total + bp] * bv;
            } else {
                double sth_n = sqrt(py_max(0.0, 1.0 - bzc * bzc));
                double c2d, s2d;
                or_spin2(bzc, sth_n, bphic, bz_pts, bsth_pts, bphi_pts, &c2d, &s2d);
                double qv = (double)mp_stacked[c_q * npix_total + bp];
                double uv = (double)mp_stacked[c_u * npix_total + bp];
                acc[c_q] += (qv * c2d + uv * s2d) * bv;
                acc[c_u] += (-qv * s2d + uv * c2d) * bv;
                for (int64_t c = 0; c < C && c < 8; ++c)
                    if (c != c_q && c != c_u)
                        acc[c] += (double)mp_stacked[c * npix_total + bp] * bv;
            }
        }
        for (int64_t c = 0; c < C && c < 8; ++c)
            tod[c * B + b] = acc[c];
    }
}

/* tod_bilinear.py:41-78 (_gather_accum_jit): split-pipeline gather from
 * precomputed (4, B*Sc) pixels and weights. */
void or_gather_accum(const int64_t *pixels, const double *weights, const float *beam_vals,
                     const float *mp_stacked, int64_t C, int64_t npix, int64_t B, int64_t Sc,
                     double *tod)
{
    int64_t N = B * Sc;
    for (int64_t b = 0; b < B; ++b) {
        for (int64_t s = 0; s < Sc; ++s) {
            int64_t n = b * Sc + s;
            double bv = (double)beam_vals[s];
            int64_t p0 = pixels[n], p1 = pixels[N + n], p2 = pixels[2 * N + n],
                    p3 = pixels[3 * N + n];
            double w0 = weights[n], w1 = weights[N + n], w2 = weights[2 * N + n],
                   w3 = weights[3 * N + n];
            for (int64_t c = 0; c < C; ++c) {
                const float *m = mp_stacked + c * npix;
                tod[c * B + b] += ((double)m[p0] * w0 + (double)m[p1] * w1 +
                                   (double)m[p2] * w2 + (double)m[p3] * w3) *
                                  bv;
            }
        }
    }
}

/* ----------------------------------------------- per-(b, s) index dumps */

/* Neighbour pixels and weights of the fused bilinear kernel for samples
 * [b0, b0+nb), written as pix[(b-b0)*S*4 + s*4 + k], wgt likewise. */
void or_fused_indices(const float *vec_orig, const float *axes, const float *cos_a,
                      const float *sin_a, const float *ax_pts, const float *cos_p,
                      const float *sin_p, int64_t nside, int64_t S, int64_t b0, int64_t nb,
                      int64_t *pix, double *wgt)
{
    int64_t npix_total = 12 * nside * nside;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i) {
        int64_t b = b0 + i;
        rot_params r = load_rot(axes, cos_a, sin_a, ax_pts, cos_p, sin_p, b);
        for (int64_t s = 0; s < S; ++s) {
            double vx, vy, vz;
            rotate_sample(&r, vec_orig, s, &vx, &vy, &vz);
            double z = py_max(-1.0, py_min(1.0, vz));
            double phi_w = wrap_atan2(vy, vx);
            double zn[4], phin[4];
            or_ring_interp_with_angles(nside, z, phi_w, npix_total, pix + (i * S + s) * 4,
                                       wgt + (i * S + s) * 4, zn, phin);
        }
    }
}

/* Nearest pixel of the nearest kernel for samples [b0, b0+nb). */
void or_nearest_indices(const float *vec_orig, const float *axes, const float *cos_a,
                        const float *sin_a, const float *ax_pts, const float *cos_p,
                        const float *sin_p, int64_t nside, int64_t S, int64_t b0, int64_t nb,
                        int64_t *pix)
{
    int64_t npix_total = 12 * nside * nside;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i) {
        int64_t b = b0 + i;
        rot_params r = load_rot(axes, cos_a, sin_a, ax_pts, cos_p, sin_p, b);
        for (int64_t s = 0; s < S; ++s) {
            double vx, vy, vz;
            rotate_sample(&r, vec_orig, s, &vx, &vy, &vz);
            double phi_w = wrap_atan2(vy, vx);
            double sin_th = sqrt(py_max(0.0, vx * vx + vy * vy));
            double bzc, bphic;
            or_nearest_pix(nside, vz, sin_th, phi_w, npix_total, pix + i * S + s, &bzc, &bphic);
        }
    }
}

/* Rotated directions (binary64) of the fused kernels, for [b0, b0+nb). */
void or_rotated_dirs(const float *vec_orig, const float *axes, const float *cos_a,
                     const float *sin_a, const float *ax_pts, const float *cos_p,
                     const float *sin_p, int64_t S, int64_t b0, int64_t nb, double *out)
{
    for (int64_t i = 0; i < nb; ++i) {
        rot_params r = load_rot(axes, cos_a, sin_a, ax_pts, cos_p, sin_p, b0 + i);
        for (int64_t s = 0; s < S; ++s) {
            double *o = out + 3 * (i * S + s);
            rotate_sample(&r, vec_orig, s, &o[0], &o[1], &o[2]);
        }
    }
}

int or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* --------------------------------------------- bicubic (reconstructed) */

/* Keys cubic convolution kernel, a = -0.5 (Catmull-Rom), support |t| < 2. */
static inline double keys_kernel(double t)
{
    const double a = -0.5;
    t = fabs(t);
    if (t < 1.0)
        return ((a + 2.0) * t - (a + 3.0)) * t * t + 1.0;
    if (t < 2.0)
        return ((a * t - 5.0 * a) * t + 8.0 * a) * t - 4.0 * a;
    return 0.0;
}

/* Bicubic judge weights for a query (z, phi).  Reconstructed from the
 * reference's description (README.md:120: Keys/Catmull-Rom kernel via
 * gnomonic projection) and its stencil (numba_healpy.py:892-977, support
 * notes :909-919); the reference module itself is not in the snapshot, so
 * this is the build's definition:
 *   - candidates: the 7-ring x 5-pixel stencil around the query;
 *   - each pixel centre is projected gnomonically onto the tangent plane at
 *     the query: x (east), y (north);
 *   - x is measured in units of that pixel's own ring step (dphi * sin
 *     theta), y in units of the local ring spacing (theta of the two rings
 *     bracketing the query), so K(x) K(y) is a cubic-convolution
 *     interpolator on the ring lattice: exact at pixel centres and exact for
 *     quadratics along a ring;
 *   - weights K(x) K(y), Keys a = -0.5, normalised to unit sum.
 * Returns the stencil size (entries with zero weight are kept). */
int64_t or_bicubic_weights(int64_t nside, double z, double phi, int64_t *pix, double *w)
{
    int64_t npix_total = 12 * nside * nside;
    double st = sqrt(py_max(0.0, (1.0 - z) * (1.0 + z)));
    double qx = st * cos(phi), qy = st * sin(phi), qz = z;
    double ex = -sin(phi), ey = cos(phi);                  /* east  */
    double nx = -z * cos(phi), ny = -z * sin(phi), nz = st; /* north */
    int64_t ira = or_ring_above(nside, z);
    if (ira < 1)
        ira = 1;
    if (ira > 4 * nside - 2)
        ira = 4 * nside - 2;
    double dth = acos(or_ring_z(nside, ira + 1)) - acos(or_ring_z(nside, ira));
    /* stencil rings and slots as in numba_healpy.py:938-976 */
    int64_t ir_center = or_ring_above(nside, z);
    if (ir_center < 1)
        ir_center = 1;
    else if (ir_center > 4 * nside - 1)
        ir_center = 4 * nside - 1;
    int64_t ir_lo = ir_center - 3 > 1 ? ir_center - 3 : 1;
    int64_t ir_hi = ir_center + 3 < 4 * nside - 1 ? ir_center + 3 : 4 * nside - 1;
    int64_t m = 0;
    double sum = 0.0;
    for (int64_t ir = ir_lo; ir <= ir_hi; ++ir) {
        int64_t n_p, fp;
        double phi0, dphi;
        or_ring_info(nside, ir, npix_total, &n_p, &fp, &phi0, &dphi);
        double zr = or_ring_z(nside, ir);
        double sr = sqrt(py_max(0.0, (1.0 - zr) * (1.0 + zr)));
        double step = dphi * sr;
        int64_t cnt = n_p <= 4 ? n_p : 5;
        int64_t ipc = 0;
        if (n_p > 4) {
            int64_t f = (int64_t)floor((phi - phi0) / dphi + 0.5);
            ipc = ((f % n_p) + n_p) % n_p;
        }
        for (int64_t j = 0; j < cnt; ++j) {
            int64_t ip = n_p <= 4 ? j : ((((ipc + j - 2) % n_p) + n_p) % n_p);
            double ph = phi0 + (double)ip * dphi;
            double px = sr * cos(ph), py = sr * sin(ph), pz = zr;
            double d = px * qx + py * qy + pz * qz;
            double wk = 0.0;
            if (d > 0.0 && step > 0.0 && dth > 0.0) {
                double x = (px * ex + py * ey) / d / step;
                double y = (px * nx + py * ny + pz * nz) / d / dth;
                wk = keys_kernel(x) * keys_kernel(y);
            }
            pix[m] = fp + ip;
            w[m] = wk;
            sum += wk;
            ++m;
        }
    }
    if (sum != 0.0) {
        for (int64_t k = 0; k < m; ++k)
            w[k] /= sum;
        return m;
    }
    /* degenerate geometry (no candidate inside the support): bilinear */
    double w4[4], zn[4], pn[4];
    or_ring_interp_with_angles(nside, z, phi, npix_total, pix, w4, zn, pn);
    for (int k = 0; k < 4; ++k)
        w[k] = w4[k];
    return 4;
}

/* Beam TOD with the bicubic judge: same rotation and spin-2 as the fused
 * bilinear kernel, stencil weights instead of the 4 bilinear ones. */
void or_gather_accum_bicubic(const float *vec_orig, const float *axes, const float *cos_a,
                             const float *sin_a, const float *ax_pts, const float *cos_p,
                             const float *sin_p, int64_t nside, const float *mp_stacked,
                             int64_t C, const float *beam_vals, int64_t B, int64_t S,
                             double *tod, int64_t c_q, int64_t c_u, int nthreads)
{
    int has_qu = (c_q >= 0 && c_u >= 0);
    int64_t npix_total = 12 * nside * nside;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t b = 0; b < B; ++b) {
        rot_params r = load_rot(axes, cos_a, sin_a, ax_pts, cos_p, sin_p, b);
        double z_pts = py_max(-1.0, py_min(1.0, r.bz));
        double sth_pts = sqrt(py_max(0.0, 1.0 - r.bz * r.bz));
        double phi_pts = wrap_atan2(r.by, r.bx);
        double acc[8];
        for (int64_t c = 0; c < C && c < 8; ++c)
            acc[c] = tod[c * B + b];
        for (int64_t s = 0; s < S; ++s) {
            double vx, vy, vz;
            rotate_sample(&r, vec_orig, s, &vx, &vy, &vz);
            double z = py_max(-1.0, py_min(1.0, vz));
            double phi_w = wrap_atan2(vy, vx);
            int64_t pix[64];
            double w[64];
            int64_t m = or_bicubic_weights(nside, z, phi_w, pix, w);
            double bv = (double)beam_vals[s];
            for (int64_t k = 0; k < m; ++k) {
                if (w[k] == 0.0)
                    continue;
                double wb = w[k] * bv;
                int64_t p = pix[k];
                if (has_qu) {
                    double zc, pc, c2, s2;
                    or_pix2zphi(nside, p, &zc, &pc);
                    spin2_neighbour(zc, pc, z_pts, sth_pts, phi_pts, &c2, &s2);
                    double q = (double)mp_stacked[c_q * npix_total + p];
                    double u = (double)mp_stacked[c_u * npix_total + p];
                    acc[c_q] += wb * (q * c2 + u * s2);
                    acc[c_u] += wb * (-q * s2 + u * c2);
                    for (int64_t c = 0; c < C && c < 8; ++c)
                        if (c != c_q && c != c_u)
                            acc[c] += wb * (double)mp_stacked[c * npix_total + p];
                } else {
                    for (int64_t c = 0; c < C && c < 8; ++c)
                        acc[c] += wb * (double)mp_stacked[c * npix_total + p];
                }
            }
        }
        for (int64_t c = 0; c < C && c < 8; ++c)
            tod[c * B + b] = acc[c];
    }
}

/* Interpolated map values at N directions: mode 0 bilinear, 1 bicubic. */
void or_interp_values(int64_t nside, const float *mp, int64_t npix, int64_t N, const double *z,
                      const double *phi, int mode, double *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double acc = 0.0;
        if (mode == 0) {
            int64_t p[4];
            double w[4], zn[4], pn[4];
            or_ring_interp_with_angles(nside, z[i], phi[i], 12 * nside * nside, p, w, zn, pn);
            for (int k = 0; k < 4; ++k)
                acc += w[k] * (double)mp[p[k] < npix ? p[k] : npix - 1];
        } else {
            int64_t p[64];
            double w[64];
            int64_t m = or_bicubic_weights(nside, z[i], phi[i], p, w);
            for (int64_t k = 0; k < m; ++k)
                acc += w[k] * (double)mp[p[k]];
        }
        out[i] = acc;
    }
}

/* ------------------------------------------ factorised corrector (own) */

/* The factorised corrector is this library's own construction (the fast
 * form of its learned interpolation corrector; the reference has no
 * counterpart).  These restate the definitions the library documents for
 * todg_fcorr_channels and todg_gather_accum_fcorr, independently of the GPU
 * code: ring of a pixel by exact integer square roots, ring values by
 * division instead of the GPU's reciprocal multiplication, per-neighbour
 * combination instead of the GPU's per-pixel weight sums. */

static int64_t fc_isqrt(int64_t x)
{
    int64_t s = (int64_t)sqrt((double)x);
    while (s * s > x)
        --s;
    while ((s + 1) * (s + 1) <= x)
        ++s;
    return s;
}

/* RING scheme: ring (1 .. 4 nside - 1) of pixel p. */
int64_t or_pixel_ring(int64_t nside, int64_t p)
{
    int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    if (p < ncap)
        return (1 + fc_isqrt(1 + 2 * p)) / 2;
    if (p < npix - ncap)
        return (p - ncap) / (4 * nside) + nside;
    return 4 * nside - (1 + fc_isqrt(1 + 2 * (npix - 1 - p))) / 2;
}

static double fc_ring_value(int64_t nside, const float *m, int64_t r, double phi)
{
    int64_t n, fp;
    double phi0, dphi;
    or_ring_info(nside, r, 12 * nside * nside, &n, &fp, &phi0, &dphi);
    double t = (phi - phi0) / dphi;
    double jf = floor(t);
    double f = t - jf;
    int64_t j = ((int64_t)jf % n + n) % n;
    int64_t j1 = (j + 1) % n;
    return (1.0 - f) * (double)m[fp + j] + f * (double)m[fp + j1];
}

/* The 16 stencil features of pixel p (F[0] = m(p); F[1..15] for (dr, ds) in
 * {0..3}^2 minus (0, 0), row-major). */
void or_fcorr_features(int64_t nside, const float *m, int64_t p, double *F)
{
    int64_t r = or_pixel_ring(nside, p);
    int64_t n, fp;
    double phi0, dphi;
    or_ring_info(nside, r, 12 * nside * nside, &n, &fp, &phi0, &dphi);
    double phi = phi0 + (double)(p - fp) * dphi;
    double mp = (double)m[p];
    int ok = r >= 4 && r <= 4 * nside - 4;
    int j = 0;
    F[j++] = mp;
    for (int dr = 0; dr < 4; ++dr) {
        for (int ds = 0; ds < 4; ++ds) {
            if (dr == 0 && ds == 0)
                continue;
            double v = 0.0;
            double pl = phi - ds * dphi, ph = phi + ds * dphi;
            if (dr == 0) {
                v = 0.5 * (fc_ring_value(nside, m, r, pl) + fc_ring_value(nside, m, r, ph)) - mp;
            } else if (ok) {
                if (ds == 0)
                    v = 0.5 * (fc_ring_value(nside, m, r - dr, phi) + fc_ring_value(nside, m, r + dr, phi)) - mp;
                else
                    v = 0.25 * (fc_ring_value(nside, m, r - dr, pl) + fc_ring_value(nside, m, r - dr, ph) +
                                fc_ring_value(nside, m, r + dr, pl) + fc_ring_value(nside, m, r + dr, ph)) -
                        mp;
            }
            F[j++] = v;
        }
    }
}

/* out0[p] = sum_j coef[j] F_j(p), out1[p] = sum_j coef[16 + j] F_j(p). */
void or_fcorr_channels(int64_t nside, const float *m, const double *coef, double *out0, double *out1)
{
    int64_t npix = 12 * nside * nside;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npix; ++p) {
        double F[16];
        or_fcorr_features(nside, m, p, F);
        double s0 = 0.0, s1 = 0.0;
        for (int j = 0; j < 16; ++j) {
            s0 += coef[j] * F[j];
            s1 += coef[16 + j] * F[j];
        }
        out0[p] = s0;
        out1[p] = s1;
    }
}

/* Corrected beam TOD: or_gather_accum_fused with each neighbour's value
 * A0(p_k) + (1 - g) g A1(p_k), g = w_2 + w_3, per field (A0, A1: (C, npix)
 * field-major float32), spin-2 applied to the combined Q/U values. */
void or_gather_accum_fcorr(const float *vec_orig, const float *axes, const float *cos_a,
                           const float *sin_a, const float *ax_pts, const float *cos_p,
                           const float *sin_p, int64_t nside, const float *A0, const float *A1,
                           int64_t C, const float *beam_vals, int64_t B, int64_t S, double *tod,
                           int64_t c_q, int64_t c_u, int nthreads)
{
    int has_qu = (c_q >= 0 && c_u >= 0);
    int64_t npix_total = 12 * nside * nside;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t b = 0; b < B; ++b) {
        rot_params r = load_rot(axes, cos_a, sin_a, ax_pts, cos_p, sin_p, b);
        double z_pts = py_max(-1.0, py_min(1.0, r.bz));
        double sth_pts = sqrt(py_max(0.0, 1.0 - r.bz * r.bz));
        double phi_pts = wrap_atan2(r.by, r.bx);
        double acc[8];
        for (int64_t c = 0; c < C && c < 8; ++c)
            acc[c] = tod[c * B + b];
        for (int64_t s = 0; s < S; ++s) {
            double vx, vy, vz;
            rotate_sample(&r, vec_orig, s, &vx, &vy, &vz);
            double z = py_max(-1.0, py_min(1.0, vz));
            double phi_w = wrap_atan2(vy, vx);
            int64_t p[4];
            double w[4], zn[4], phin[4];
            or_ring_interp_with_angles(nside, z, phi_w, npix_total, p, w, zn, phin);
            double uc = (w[0] + w[1]) * (w[2] + w[3]);
            double bv = (double)beam_vals[s];
            for (int k = 0; k < 4; ++k) {
                double wk = w[k] * bv;
                if (has_qu) {
                    double c2, s2;
                    spin2_neighbour(zn[k], phin[k], z_pts, sth_pts, phi_pts, &c2, &s2);
                    double q = (double)A0[c_q * npix_total + p[k]] + uc * (double)A1[c_q * npix_total + p[k]];
                    double u = (double)A0[c_u * npix_total + p[k]] + uc * (double)A1[c_u * npix_total + p[k]];
                    acc[c_q] += wk * (q * c2 + u * s2);
                    acc[c_u] += wk * (-q * s2 + u * c2);
                }
                for (int64_t c = 0; c < C && c < 8; ++c) {
                    if (has_qu && (c == c_q || c == c_u))
                        continue;
                    acc[c] += wk * ((double)A0[c * npix_total + p[k]] + uc * (double)A1[c * npix_total + p[k]]);
                }
            }
        }
        for (int64_t c = 0; c < C && c < 8; ++c)
            tod[c * B + b] = acc[c];
    }
}
