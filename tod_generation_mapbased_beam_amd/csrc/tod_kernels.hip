// Beam-convolution TOD kernels for gfx950 and their C-ABI entry points.
//
//   tod[c, b] += sum_s beam_vals[s] * spin2_c( map[ R_b * vec_orig[s] ] )
//
// Work decomposition: one wavefront owns a tile of T consecutive samples and
// sweeps the beam with its 64 lanes (lane l handles beam pixels l, l+64, ...).
// The per-sample rotation scalars are wave-uniform; the rotated direction,
// the HEALPix ring lookup (binary64, reference operation order) and the
// spin-2 factors are computed per lane; the I/Q/U values of each neighbour
// come from one 16-byte load of a pixel-major packed map.  Each lane keeps
// binary64 partial sums (binary64 add issues at the full VALU rate on CDNA4)
// and the wave reduces them across lanes once per sample, so each sample's
// tod entry is written by exactly one lane: no atomics, deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <math.h>

#include <mutex>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "healpix_ring.hpp"
#include "../../include/todgpu.h"

using namespace todg;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define TODG_HIP_CHECK(expr)                                                         \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess)                                                        \
            return fail(TODG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------- ring table

// Host construction of the per-ring constants with the reference formulas
// (numba_healpy.py:73-114) and glibc acos (numba_healpy.py:331-332).
std::vector<RingEntry> build_ring_table(int64_t nside) {
    const double two_pi = 2.0 * 3.141592653589793;
    const int64_t npix = 12 * nside * nside;
    std::vector<RingEntry> t((size_t)(4 * nside + 1));
    memset(t.data(), 0, t.size() * sizeof(RingEntry));
    for (int64_t ir = 1; ir <= 4 * nside - 1; ++ir) {
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
            fp = npix - 2 * i2 * (i2 + 1);
            s = 1;
        }
        volatile double dphi = two_pi / (double)n;
        volatile double phi0 = (double)s * dphi * 0.5;
        double z;
        if (ir < nside) {
            volatile double tmp = (double)ir;
            volatile double den = 3.0 * (double)nside * (double)nside;
            z = 1.0 - tmp * tmp / den;
        } else if (ir <= 3 * nside) {
            volatile double q = (double)ir / (double)nside;
            z = (2.0 / 3.0) * (2.0 - q);
        } else {
            volatile double tmp = (double)(4 * nside - ir);
            volatile double den = 3.0 * (double)nside * (double)nside;
            z = -(1.0 - tmp * tmp / den);
        }
        RingEntry& e = t[(size_t)ir];
        e.z = z;
        e.theta = acos(z);
        e.phi0 = phi0;
        e.dphi = dphi;
        e.first = fp;
        e.n = n;
    }
    return t;
}

struct RingKey {
    int device;
    int64_t nside;
    bool operator<(const RingKey& o) const {
        return device != o.device ? device < o.device : nside < o.nside;
    }
};

std::mutex g_ring_mu;
std::map<RingKey, RingEntry*> g_rings;

int get_ring_table(int64_t nside, const RingEntry** out) {
    int dev = 0;
    TODG_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ring_mu);
    auto it = g_rings.find({dev, nside});
    if (it != g_rings.end()) {
        *out = it->second;
        return TODG_OK;
    }
    std::vector<RingEntry> h = build_ring_table(nside);
    RingEntry* d = nullptr;
    TODG_HIP_CHECK(hipMalloc(&d, h.size() * sizeof(RingEntry)));
    TODG_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(RingEntry), hipMemcpyHostToDevice));
    g_rings[{dev, nside}] = d;
    *out = d;
    return TODG_OK;
}

int cu_count() {
    static int cached = 0;
    if (cached) return cached;
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        n = 256;
    cached = n;
    return n;
}

// ------------------------------------------------------------ kernel args

constexpr int kMaxSlots = 4;

struct TodArgs {
    const float* vec_orig;   // (S, 3)
    const float* beam_vals;  // (S,)
    const float* axes;       // (B, 3)
    const float* cos_a;      // (B,)
    const float* sin_a;      // (B,)
    const float* ax_pts;     // (B, 3)
    const float* cos_p;      // (B,)
    const float* sin_p;      // (B,)
    const float* map;        // (npix, CS) pixel-major, slots permuted
    const RingEntry* rings;  // (4*nside+1,)
    double* tod;             // (C, B), accumulated
    int64_t B;
    int64_t S;
    int64_t npix;
    int nside;
    int nslots;
    int row[kMaxSlots];      // tod row of each packed slot
};

struct Sample {
    double kx, ky, kz, ca, sa, bx, by, bz, cp, sp;
    double z_pts, sth_pts, phi_pts;
};

__device__ __forceinline__ void load_sample(const TodArgs& a, int64_t b, Sample& P) {
    P.kx = (double)a.axes[3 * b];
    P.ky = (double)a.axes[3 * b + 1];
    P.kz = (double)a.axes[3 * b + 2];
    P.ca = (double)a.cos_a[b];
    P.sa = (double)a.sin_a[b];
    P.bx = (double)a.ax_pts[3 * b];
    P.by = (double)a.ax_pts[3 * b + 1];
    P.bz = (double)a.ax_pts[3 * b + 2];
    P.cp = (double)a.cos_p[b];
    P.sp = (double)a.sin_p[b];
    // Boresight (z, sin theta, phi) for spin-2: tod_bilinear.py:306-310.
    P.z_pts = fmax(-1.0, fmin(1.0, P.bz));
    P.sth_pts = sqrt(fmax(0.0, 1.0 - P.bz * P.bz));
    double ph = atan2(P.by, P.bx);
    if (ph < 0.0) ph += kTwoPi;
    P.phi_pts = ph;
}

template <int CS>
__device__ __forceinline__ float4 load_px(const float* __restrict__ map, int64_t p, int64_t npix) {
    // The reference never bounds-checks; an index can only leave [0, npix)
    // through the 1-ulp "tw == n" float-mod corner at the last ring, or from
    // a non-finite pointing.  Clamp the address (unsigned compare also
    // catches negatives) so such cases read a valid pixel, never fault.
    p = (uint64_t)p < (uint64_t)npix ? p : ((p < 0) ? 0 : npix - 1);
    if constexpr (CS == 4) {
        return *reinterpret_cast<const float4*>(map + 4 * p);
    } else if constexpr (CS == 2) {
        float2 v = *reinterpret_cast<const float2*>(map + 2 * p);
        return make_float4(v.x, v.y, 0.f, 0.f);
    } else {
        return make_float4(map[p], 0.f, 0.f, 0.f);
    }
}

__device__ __forceinline__ float slot(const float4& v, int k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// MODE 0 = bilinear (tod_bilinear.py:205-477), 1 = nearest (tod_nearest.py:21-158).
// QU: slots 0 and 1 hold Q and U and get the spin-2 rotation; other slots
// are accumulated plainly.  Without QU every slot is accumulated plainly.
template <int MODE, bool QU, int CS, int T>
__global__ __launch_bounds__(256) void k_tod(const TodArgs a) {
    const int lane = threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t wpb = blockDim.x >> 6;
    const int64_t wave0 = (int64_t)blockIdx.x * wpb + wib;
    const int64_t nwaves = (int64_t)gridDim.x * wpb;
    const int64_t ntiles = (a.B + T - 1) / T;
    const int nslots = a.nslots;

    for (int64_t tile = wave0; tile < ntiles; tile += nwaves) {
        const int64_t b0 = tile * T;
        Sample P[T];
        double acc[T][kMaxSlots];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int64_t b = (b0 + t < a.B) ? b0 + t : a.B - 1;
            load_sample(a, b, P[t]);
#pragma unroll
            for (int k = 0; k < kMaxSlots; ++k) acc[t][k] = 0.0;
        }

        for (int64_t s = lane; s < a.S; s += 64) {
            const double v0 = (double)a.vec_orig[3 * s];
            const double v1 = (double)a.vec_orig[3 * s + 1];
            const double v2 = (double)a.vec_orig[3 * s + 2];
            const double bv = (double)a.beam_vals[s];
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const Sample& q = P[t];
                double vx, vy, vz;
                rodrigues(v0, v1, v2, q.kx, q.ky, q.kz, q.ca, q.sa, q.bx, q.by, q.bz, q.cp,
                          q.sp, vx, vy, vz);
                double phi_w = atan2(vy, vx);
                if (phi_w < 0.0) phi_w += kTwoPi;
                if constexpr (MODE == 0) {
                    const double z = fmax(-1.0, fmin(1.0, vz));
                    Bilin o;
                    ring_interp(a.nside, a.rings, z, phi_w, o);
                    float4 m[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) m[k] = load_px<CS>(a.map, o.p[k], a.npix);
                    if constexpr (QU) {
                        double c2[4], s2[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const double sth_n = sqrt(fmax(0.0, 1.0 - o.zn[k] * o.zn[k]));
                            spin2(o.zn[k], sth_n, o.phin[k], q.z_pts, q.sth_pts, q.phi_pts,
                                  c2[k], s2[k]);
                        }
                        double qq[4], uu[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            qq[k] = (double)m[k].x;
                            uu[k] = (double)m[k].y;
                        }
                        acc[t][0] += (o.w[0] * (qq[0] * c2[0] + uu[0] * s2[0]) +
                                      o.w[1] * (qq[1] * c2[1] + uu[1] * s2[1]) +
                                      o.w[2] * (qq[2] * c2[2] + uu[2] * s2[2]) +
                                      o.w[3] * (qq[3] * c2[3] + uu[3] * s2[3])) *
                                     bv;
                        acc[t][1] += (o.w[0] * (-qq[0] * s2[0] + uu[0] * c2[0]) +
                                      o.w[1] * (-qq[1] * s2[1] + uu[1] * c2[1]) +
                                      o.w[2] * (-qq[2] * s2[2] + uu[2] * c2[2]) +
                                      o.w[3] * (-qq[3] * s2[3] + uu[3] * c2[3])) *
                                     bv;
#pragma unroll
                        for (int k = 2; k < CS; ++k) {
                            if (k < nslots)
                                acc[t][k] += (o.w[0] * (double)slot(m[0], k) +
                                              o.w[1] * (double)slot(m[1], k) +
                                              o.w[2] * (double)slot(m[2], k) +
                                              o.w[3] * (double)slot(m[3], k)) *
                                             bv;
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < CS; ++k) {
                            if (k < nslots)
                                acc[t][k] += (o.w[0] * (double)slot(m[0], k) +
                                              o.w[1] * (double)slot(m[1], k) +
                                              o.w[2] * (double)slot(m[2], k) +
                                              o.w[3] * (double)slot(m[3], k)) *
                                             bv;
                        }
                    }
                } else {
                    const double sin_th = sqrt(fmax(0.0, vx * vx + vy * vy));
                    int64_t bp;
                    double bzc, bphic;
                    nearest_pix(a.nside, a.rings, vz, sin_th, phi_w, bp, bzc, bphic);
                    const float4 m = load_px<CS>(a.map, bp, a.npix);
                    if constexpr (QU) {
                        const double sth_n = sqrt(fmax(0.0, 1.0 - bzc * bzc));
                        double c2d, s2d;
                        spin2(bzc, sth_n, bphic, q.z_pts, q.sth_pts, q.phi_pts, c2d, s2d);
                        const double qv = (double)m.x, uv = (double)m.y;
                        acc[t][0] += (qv * c2d + uv * s2d) * bv;
                        acc[t][1] += (-qv * s2d + uv * c2d) * bv;
#pragma unroll
                        for (int k = 2; k < CS; ++k)
                            if (k < nslots) acc[t][k] += (double)slot(m, k) * bv;
                    } else {
#pragma unroll
                        for (int k = 0; k < CS; ++k)
                            if (k < nslots) acc[t][k] += (double)slot(m, k) * bv;
                    }
                }
            }
        }

#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int64_t b = b0 + t;
#pragma unroll
            for (int k = 0; k < kMaxSlots; ++k) {
                if (k < nslots && k < CS) {
                    const double sum = wave_sum(acc[t][k]);
                    if (lane == 0 && b < a.B) a.tod[(int64_t)a.row[k] * a.B + b] += sum;
                }
            }
        }
    }
}

// Per-(b, s) dumps of the index path for bit-exact comparisons.
__global__ __launch_bounds__(256) void k_bilinear_indices(const TodArgs a, int64_t b0,
                                                           int64_t nb, int64_t* pix,
                                                           double* wgt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * a.S) return;
    const int64_t b = b0 + i / a.S, s = i % a.S;
    Sample q;
    load_sample(a, b, q);
    double vx, vy, vz;
    rodrigues((double)a.vec_orig[3 * s], (double)a.vec_orig[3 * s + 1],
              (double)a.vec_orig[3 * s + 2], q.kx, q.ky, q.kz, q.ca, q.sa, q.bx, q.by, q.bz,
              q.cp, q.sp, vx, vy, vz);
    double phi_w = atan2(vy, vx);
    if (phi_w < 0.0) phi_w += kTwoPi;
    const double z = fmax(-1.0, fmin(1.0, vz));
    Bilin o;
    ring_interp(a.nside, a.rings, z, phi_w, o);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pix[4 * i + k] = o.p[k];
        wgt[4 * i + k] = o.w[k];
    }
}

__global__ __launch_bounds__(256) void k_nearest_indices(const TodArgs a, int64_t b0,
                                                          int64_t nb, int64_t* pix) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * a.S) return;
    const int64_t b = b0 + i / a.S, s = i % a.S;
    Sample q;
    load_sample(a, b, q);
    double vx, vy, vz;
    rodrigues((double)a.vec_orig[3 * s], (double)a.vec_orig[3 * s + 1],
              (double)a.vec_orig[3 * s + 2], q.kx, q.ky, q.kz, q.ca, q.sa, q.bx, q.by, q.bz,
              q.cp, q.sp, vx, vy, vz);
    double phi_w = atan2(vy, vx);
    if (phi_w < 0.0) phi_w += kTwoPi;
    const double sin_th = sqrt(fmax(0.0, vx * vx + vy * vy));
    int64_t bp;
    double bzc, bphic;
    nearest_pix(a.nside, a.rings, vz, sin_th, phi_w, bp, bzc, bphic);
    pix[i] = bp;
}

// Field-major (C, npix) -> pixel-major (npix, CS) with slot k taken from
// source channel perm[k] (-1 = zero fill).
__global__ __launch_bounds__(256) void k_pack_map(const float* __restrict__ src, int64_t npix,
                                                  int CS, int p0, int p1, int p2, int p3,
                                                  float* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const int perm[4] = {p0, p1, p2, p3};
    for (int k = 0; k < CS; ++k)
        dst[CS * i + k] = perm[k] >= 0 ? src[(int64_t)perm[k] * npix + i] : 0.f;
}

int slots_stride(int nslots) { return nslots <= 1 ? 1 : (nslots == 2 ? 2 : 4); }

bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

// Derive the packed-slot order and the tod row of each slot from (C, c_q, c_u)
// the way the reference picks its channels (tod_bilinear.py:275-285): with
// both Q and U present, slots 0/1 are Q/U (spin-2 rotated) and the remaining
// channels follow in order; otherwise slot c is channel c.  When c_q == c_u
// the reference still runs the spin-2 branch and both updates land on the
// same row, so that channel occupies two slots.
int slot_layout(int C, int c_q, int c_u, bool* qu, int perm[kMaxSlots], int row[kMaxSlots],
                int* nslots) {
    if (C < 1 || C > kMaxSlots) return fail(TODG_ERR_ARG, "C must be in [1, 4]");
    if (c_q >= C || c_u >= C || c_q < -1 || c_u < -1)
        return fail(TODG_ERR_ARG, "c_q / c_u must be -1 or a channel index < C");
    *qu = (c_q >= 0 && c_u >= 0);
    for (int k = 0; k < kMaxSlots; ++k) perm[k] = -1, row[k] = 0;
    if (*qu) {
        int k = 2;
        for (int c = 0; c < C; ++c)
            if (c != c_q && c != c_u) ++k;
        if (k > kMaxSlots) return fail(TODG_ERR_ARG, "too many channels for c_q == c_u");
        perm[0] = c_q;
        perm[1] = c_u;
        row[0] = c_q;
        row[1] = c_u;
        k = 2;
        for (int c = 0; c < C; ++c)
            if (c != c_q && c != c_u) {
                perm[k] = c;
                row[k] = c;
                ++k;
            }
        *nslots = k;
    } else {
        for (int c = 0; c < C; ++c) perm[c] = c, row[c] = c;
        *nslots = C;
    }
    return TODG_OK;
}

template <int MODE, bool QU, int CS>
hipError_t launch_cs(const TodArgs& a, int T, int blocks, hipStream_t st) {
    switch (T) {
        case 1: hipLaunchKernelGGL((k_tod<MODE, QU, CS, 1>), dim3(blocks), dim3(256), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_tod<MODE, QU, CS, 2>), dim3(blocks), dim3(256), 0, st, a); break;
        default: hipLaunchKernelGGL((k_tod<MODE, QU, CS, 4>), dim3(blocks), dim3(256), 0, st, a); break;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_mode(const TodArgs& a, bool qu, int CS, int T, int blocks, hipStream_t st) {
    if (qu) {
        if (CS == 2) return launch_cs<MODE, true, 2>(a, T, blocks, st);
        return launch_cs<MODE, true, 4>(a, T, blocks, st);
    }
    if (CS == 1) return launch_cs<MODE, false, 1>(a, T, blocks, st);
    if (CS == 2) return launch_cs<MODE, false, 2>(a, T, blocks, st);
    return launch_cs<MODE, false, 4>(a, T, blocks, st);
}

int fill_args(const float* vec_orig, const float* axes, const float* cos_a, const float* sin_a,
              const float* ax_pts, const float* cos_p, const float* sin_p, int64_t nside,
              const float* map_packed, int C, const float* beam_vals, int64_t B, int64_t S,
              double* tod, int c_q, int c_u, TodArgs* a, bool* qu, int* CS) {
    if (!is_pow2(nside) || nside > 8192)
        return fail(TODG_ERR_ARG, "nside must be a power of two <= 8192");
    if (B < 0 || S < 0) return fail(TODG_ERR_ARG, "B and S must be non-negative");
    int perm[kMaxSlots];
    int nslots = 0;
    int rc = slot_layout(C, c_q, c_u, qu, perm, a->row, &nslots);
    if (rc) return rc;
    if (B > 0 && S > 0 &&
        (!vec_orig || !axes || !cos_a || !sin_a || !ax_pts || !cos_p || !sin_p ||
         !map_packed || !beam_vals || !tod))
        return fail(TODG_ERR_ARG, "null pointer argument");
    const RingEntry* rings = nullptr;
    rc = get_ring_table(nside, &rings);
    if (rc) return rc;
    a->vec_orig = vec_orig;
    a->beam_vals = beam_vals;
    a->axes = axes;
    a->cos_a = cos_a;
    a->sin_a = sin_a;
    a->ax_pts = ax_pts;
    a->cos_p = cos_p;
    a->sin_p = sin_p;
    a->map = map_packed;
    a->rings = rings;
    a->tod = tod;
    a->B = B;
    a->S = S;
    a->npix = 12 * nside * nside;
    a->nside = (int)nside;
    a->nslots = nslots;
    *CS = slots_stride(nslots);
    return TODG_OK;
}

int run_tod(int mode, const float* vec_orig, const float* axes, const float* cos_a,
            const float* sin_a, const float* ax_pts, const float* cos_p, const float* sin_p,
            int64_t nside, const float* map_packed, int C, const float* beam_vals, int64_t B,
            int64_t S, double* tod, int c_q, int c_u, int tile, void* stream) {
    TodArgs a;
    bool qu;
    int CS;
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                       beam_vals, B, S, tod, c_q, c_u, &a, &qu, &CS);
    if (rc) return rc;
    if (B == 0 || S == 0) return TODG_OK;
    const int T = tile == 1 || tile == 2 || tile == 4 ? tile : 2;
    const int64_t ntiles = (B + T - 1) / T;
    const int64_t waves_cap = (int64_t)cu_count() * 32;  // 8 waves/SIMD-ish residency
    int64_t waves = ntiles < waves_cap ? ntiles : waves_cap;
    const int blocks = (int)((waves + 3) / 4);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = mode == 0 ? launch_mode<0>(a, qu, CS, T, blocks, st)
                             : launch_mode<1>(a, qu, CS, T, blocks, st);
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TODG_OK;
}

}  // namespace

// ================================================================ C-ABI

extern "C" {

int todg_version(void) { return TODG_ABI_VERSION; }

int todg_last_error(char* buf, size_t n) {
    if (!buf || n == 0) return (int)g_last_error.size();
    size_t k = g_last_error.size() < n - 1 ? g_last_error.size() : n - 1;
    memcpy(buf, g_last_error.data(), k);
    buf[k] = 0;
    return (int)k;
}

int todg_map_slots(int C, int c_q, int c_u, int* perm4, int* row4) {
    bool qu;
    int perm[kMaxSlots], row[kMaxSlots], nslots = 0;
    int rc = slot_layout(C, c_q, c_u, &qu, perm, row, &nslots);
    if (rc) return rc;
    for (int k = 0; k < kMaxSlots; ++k) {
        if (perm4) perm4[k] = perm[k];
        if (row4) row4[k] = row[k];
    }
    return slots_stride(nslots);
}

int todg_pack_map(const float* mp_stacked, int C, int64_t npix, int c_q, int c_u, float* packed,
                  void* stream) {
    bool qu;
    int perm[kMaxSlots], row[kMaxSlots], nslots = 0;
    int rc = slot_layout(C, c_q, c_u, &qu, perm, row, &nslots);
    if (rc) return rc;
    if (npix <= 0 || !mp_stacked || !packed) return fail(TODG_ERR_ARG, "bad map arguments");
    const int CS = slots_stride(nslots);
    const int blocks = (int)((npix + 255) / 256);
    hipLaunchKernelGGL(k_pack_map, dim3(blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), mp_stacked, npix, CS, perm[0],
                       perm[1], perm[2], perm[3], packed);
    TODG_HIP_CHECK(hipGetLastError());
    return TODG_OK;
}

int todg_gather_accum_fused(const float* vec_orig, const float* axes, const float* cos_a,
                            const float* sin_a, const float* ax_pts, const float* cos_p,
                            const float* sin_p, int64_t nside, const float* map_packed, int C,
                            const float* beam_vals, int64_t B, int64_t S, double* tod, int c_q,
                            int c_u, int tile, void* stream) {
    return run_tod(0, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                   beam_vals, B, S, tod, c_q, c_u, tile, stream);
}

int todg_gather_accum_nearest(const float* vec_orig, const float* axes, const float* cos_a,
                              const float* sin_a, const float* ax_pts, const float* cos_p,
                              const float* sin_p, int64_t nside, const float* map_packed, int C,
                              const float* beam_vals, int64_t B, int64_t S, double* tod,
                              int c_q, int c_u, int tile, void* stream) {
    return run_tod(1, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                   beam_vals, B, S, tod, c_q, c_u, tile, stream);
}

int todg_bilinear_indices(const float* vec_orig, const float* axes, const float* cos_a,
                          const float* sin_a, const float* ax_pts, const float* cos_p,
                          const float* sin_p, int64_t nside, int64_t B, int64_t S, int64_t b0,
                          int64_t nb, int64_t* pix, double* wgt, void* stream) {
    if (b0 < 0 || nb < 0 || b0 + nb > B) return fail(TODG_ERR_ARG, "sample range out of bounds");
    TodArgs a;
    bool qu;
    int CS;
    static const float dummy_map[4] = {0, 0, 0, 0};
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, dummy_map, 1,
                       vec_orig, B, S, reinterpret_cast<double*>(pix), -1, -1, &a, &qu, &CS);
    if (rc) return rc;
    const int64_t n = nb * S;
    if (n == 0) return TODG_OK;
    if (!pix || !wgt) return fail(TODG_ERR_ARG, "null output");
    hipLaunchKernelGGL(k_bilinear_indices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), a, b0, nb, pix, wgt);
    TODG_HIP_CHECK(hipGetLastError());
    return TODG_OK;
}

int todg_nearest_indices(const float* vec_orig, const float* axes, const float* cos_a,
                         const float* sin_a, const float* ax_pts, const float* cos_p,
                         const float* sin_p, int64_t nside, int64_t B, int64_t S, int64_t b0,
                         int64_t nb, int64_t* pix, void* stream) {
    if (b0 < 0 || nb < 0 || b0 + nb > B) return fail(TODG_ERR_ARG, "sample range out of bounds");
    TodArgs a;
    bool qu;
    int CS;
    static const float dummy_map[4] = {0, 0, 0, 0};
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, dummy_map, 1,
                       vec_orig, B, S, reinterpret_cast<double*>(pix), -1, -1, &a, &qu, &CS);
    if (rc) return rc;
    const int64_t n = nb * S;
    if (n == 0) return TODG_OK;
    if (!pix) return fail(TODG_ERR_ARG, "null output");
    hipLaunchKernelGGL(k_nearest_indices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), a, b0, nb, pix);
    TODG_HIP_CHECK(hipGetLastError());
    return TODG_OK;
}

int todg_ring_table(int64_t nside, double* z, double* theta, double* phi0, double* dphi,
                    int64_t* first, int64_t* n) {
    if (!is_pow2(nside) || nside > 8192)
        return fail(TODG_ERR_ARG, "nside must be a power of two <= 8192");
    std::vector<RingEntry> t = build_ring_table(nside);
    for (size_t i = 0; i < t.size(); ++i) {
        if (z) z[i] = t[i].z;
        if (theta) theta[i] = t[i].theta;
        if (phi0) phi0[i] = t[i].phi0;
        if (dphi) dphi[i] = t[i].dphi;
        if (first) first[i] = t[i].first;
        if (n) n[i] = t[i].n;
    }
    return TODG_OK;
}

}  // extern "C"
