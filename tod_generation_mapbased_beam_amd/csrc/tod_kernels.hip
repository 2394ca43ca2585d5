// Host side of the beam-convolution TOD library for gfx950: the C-ABI
// entry points of include/todgpu.h, argument checks and launch dispatch, the
// per-device ring and reference-longitude tables, and the small kernels that
// are not per-mode (index-path dumps, point interpolation, map packing, the
// rotation operands).
//
//   tod[c, b] += sum_s beam_vals[s] * spin2_c( map[ R_b * vec_orig[s] ] )
//
// The TOD kernels themselves -- the footprint kernel k_tod_fp (per-sample
// footprint and weight grid in LDS, certified fast index path, one gather
// per footprint pixel), its batched persistent form k_tod_fp_batch and the
// direct kernel k_tod_direct -- are in tod_device.hpp; each mode's template
// instantiations are compiled in their own unit (tod_fp_bilinear.hip,
// tod_fp_nearest.hip, tod_fp_bicubic.hip, tod_fp_learned.hip,
// tod_fp_fcorr.hip) and reached through the launch_m* functions.
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
#include "internal.hpp"
#include "corrector.hpp"
#include "../../include/todgpu.h"
#include "tod_device.hpp"

using namespace todg;
using namespace todg_fp;

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
// (numba_healpy.py:73-114), glibc acos (numba_healpy.py:331-332) and sqrt
// (tod_bilinear.py:194).  volatile keeps each rounding step in source order.
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
            volatile double q = tmp * tmp / den;
            z = 1.0 - q;
        } else if (ir <= 3 * nside) {
            volatile double q = (double)ir / (double)nside;
            z = (2.0 / 3.0) * (2.0 - q);
        } else {
            volatile double tmp = (double)(4 * nside - ir);
            volatile double den = 3.0 * (double)nside * (double)nside;
            volatile double q = tmp * tmp / den;
            z = -(1.0 - q);
        }
        RingEntry& e = t[(size_t)ir];
        e.z = z;
        e.theta = acos(z);
        e.phi0 = phi0;
        e.dphi = dphi;
        e.inv_dphi = 1.0 / dphi;
        volatile double zz = z * z;
        e.sth = sqrt(1.0 - zz > 0.0 ? 1.0 - zz : 0.0);
        e.first = fp;
        e.n = (int32_t)n;
    }
    for (int64_t ir = 1; ir < 4 * nside - 1; ++ir) {
        volatile double d = t[(size_t)ir + 1].theta - t[(size_t)ir].theta;
        t[(size_t)ir].inv_dth = 1.0 / d;
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

std::mutex g_trig_mu;
std::map<int, double*> g_trig;

// cos and sin of lam_j = j * (2 pi / kLamN), glibc, per device.
int get_lam_table(const double** out) {
    int dev = 0;
    TODG_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_trig_mu);
    auto it = g_trig.find(dev);
    if (it != g_trig.end()) {
        *out = it->second;
        return TODG_OK;
    }
    std::vector<double> h(2 * kLamN);
    for (int j = 0; j < kLamN; ++j) {
        const double lam = (double)j * (kTwoPi / kLamN);
        h[2 * j] = cos(lam);
        h[2 * j + 1] = sin(lam);
    }
    double* d = nullptr;
    TODG_HIP_CHECK(hipMalloc(&d, h.size() * sizeof(double)));
    TODG_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    g_trig[dev] = d;
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

// ================================================== index-path dump kernels

__global__ __launch_bounds__(256) void k_bilinear_indices(const TodArgs a, int64_t b0,
                                                           int64_t nb, int64_t* pix,
                                                           double* wgt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * a.S) return;
    const int64_t b = b0 + i / a.S, s = i % a.S;
    Sample q;
    load_sample(a, b, q);
    double vx, vy, vz;
    rotate(q, (double)a.vec_orig[3 * s], (double)a.vec_orig[3 * s + 1],
           (double)a.vec_orig[3 * s + 2], vx, vy, vz);
    const double phi_w = wrap_phi(vy, vx);
    Nb o;
    ring_interp(a.nside, a.rings, fmax(-1.0, fmin(1.0, vz)), phi_w, o);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pix[4 * i + k] = a.rings[o.ring[k]].first + o.pos[k];
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
    rotate(q, (double)a.vec_orig[3 * s], (double)a.vec_orig[3 * s + 1],
           (double)a.vec_orig[3 * s + 2], vx, vy, vz);
    const double phi_w = wrap_phi(vy, vx);
    const double sin_th = sqrt(fmax(0.0, vx * vx + vy * vy));
    int ring, pos;
    nearest_pix(a.nside, a.rings, vz, sin_th, phi_w, ring, pos);
    pix[i] = a.rings[ring].first + pos;
}

// Point values of one map field at N directions (z, phi): mode 0 = the
// reference's bilinear interpolation (numba_healpy.py:358-499), 1 = the
// bicubic judge (bicubic_visit, normalised; bilinear where no candidate
// falls inside the support).
__global__ __launch_bounds__(256) void k_interp_points(int mode, int nside,
                                                       const RingEntry* __restrict__ rt,
                                                       const float* __restrict__ map, int64_t npix,
                                                       int64_t N, const double* __restrict__ z,
                                                       const double* __restrict__ phi,
                                                       double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double zz = z[i], pp = phi[i];
    double acc = 0.0;
    double sum = 0.0;
    if (mode == 1) {
        sum = bicubic_visit(nside, rt, zz, pp, [&](int ring, int pos, double w) {
            TODG_CHECK((unsigned)pos < (unsigned)rt[ring].n, kRingSlot);
            acc += w * (double)map[rt[ring].first + pos];
        });
    }
    if (sum != 0.0) {
        acc /= sum;
    } else {
        Nb o;
        ring_interp(nside, rt, zz, pp, o);
        acc = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int64_t p = rt[o.ring[k]].first + o.pos[k];
            // (p == npix in the float-mod corner at the last ring)
            TODG_CHECK(p < npix, kMapPixel);
            p = p < npix ? p : npix - 1;
            acc += o.w[k] * (double)map[p];
        }
    }
    out[i] = acc;
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

// Beam centre c (unit vector) and its local north n_c, binary64.
struct RotCentre {
    double c[3];
    double n[3];
};

// Binary32 sin / cos as the binary64 function rounded once (the host's
// float32 routines differ from this by at most one ulp).
__device__ __forceinline__ float sin_f32(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cos_f32(float x) { return (float)cos((double)x); }

// Per-sample rotation operands from boresight pointings, one thread per
// sample: the driver's casts (sample_based_tod_generation_gridint.py:277-281:
// pointings to binary32, psi_b - beta), the Rodrigues vector and parallactic
// offset of precompute_rotation_vector_batch (tod_rotations.py:219-303) and
// the kernel operands of _rotation_params (tod_rotations.py:306-333), with
// each step in the precision those take: binary32 unit vectors of the
// binary32 angles, binary64 cross products, arccos, rotation and arctan2,
// then the binary32 angle, axis, boresight and roll.  ctr = beam centre c
// and its local north n_c (binary64, from the host).
__global__ __launch_bounds__(256) void k_rotation_params(const double* __restrict__ theta,
                                                         const double* __restrict__ phi,
                                                         const double* __restrict__ psi, int64_t B,
                                                         RotCentre ctr, float* __restrict__ axes,
                                                         float* __restrict__ cos_a, float* __restrict__ sin_a,
                                                         float* __restrict__ ax_pts, float* __restrict__ cos_p,
                                                         float* __restrict__ sin_p) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float th = (float)theta[b], ph = (float)phi[b], ps = (float)psi[b];
    const float st = sin_f32(th), ct = cos_f32(th), sp = sin_f32(ph), cp = cos_f32(ph);
    // boresight t, its north n_t and east e_t (binary32)
    const float tx = st * cp, ty = st * sp, tz = ct;
    const float nx = cp * ct, ny = sp * ct, nz = -st;
    const float ex = -sp, ey = cp;
    const double cx = ctr.c[0], cy = ctr.c[1], cz = ctr.c[2];
    double kx = cy * (double)tz - cz * (double)ty;
    double ky = cz * (double)tx - cx * (double)tz;
    double kz = cx * (double)ty - cy * (double)tx;
    const double kn = sqrt(kx * kx + ky * ky + kz * kz);
    if (kn > 1e-10) {
        kx /= kn;
        ky /= kn;
        kz /= kn;
    } else {
        kx = ky = kz = 0.0;
    }
    const double dot = cx * (double)tx + cy * (double)ty + cz * (double)tz;
    const double ang = acos(fmax(-1.0, fmin(1.0, dot)));
    const double rx = kx * ang, ry = ky * ang, rz = kz * ang;
    // the centre's north rotated by the same Rodrigues rotation, and its
    // bearing in the boresight's (north, east) frame
    const double ca = cos(ang), sa = sin(ang);
    const double vx = ctr.n[0], vy = ctr.n[1], vz = ctr.n[2];
    const double kv = kx * vx + ky * vy + kz * vz;
    const double wx = vx * ca + (ky * vz - kz * vy) * sa + kx * kv * (1.0 - ca);
    const double wy = vy * ca + (kz * vx - kx * vz) * sa + ky * kv * (1.0 - ca);
    const double wz = vz * ca + (kx * vy - ky * vx) * sa + kz * kv * (1.0 - ca);
    double beta = atan2(wx * (double)ex + wy * (double)ey + wz * 0.0,
                        wx * (double)nx + wy * (double)ny + wz * (double)nz);
    if (beta < 0.0) beta += 2.0 * M_PI;
    const double psis = -beta + (double)ps;
    // kernel operands
    const float a32 = (float)sqrt(rx * rx + ry * ry + rz * rz);
    const bool ok = a32 > 1e-10f;
    const double den = ok ? (double)a32 : 1.0;
    axes[3 * b] = ok ? (float)(rx / den) : 0.f;
    axes[3 * b + 1] = ok ? (float)(ry / den) : 0.f;
    axes[3 * b + 2] = ok ? (float)(rz / den) : 0.f;
    cos_a[b] = cos_f32(a32);
    sin_a[b] = sin_f32(a32);
    ax_pts[3 * b] = st * cp;
    ax_pts[3 * b + 1] = st * sp;
    ax_pts[3 * b + 2] = ct;
    const float p32 = (float)psis;
    cos_p[b] = cos_f32(p32);
    sin_p[b] = sin_f32(p32);
}

// ============================================================ host helpers

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

int fill_args(const float* vec_orig, const float* axes, const float* cos_a, const float* sin_a,
              const float* ax_pts, const float* cos_p, const float* sin_p, int64_t nside,
              const float* map_packed, int C, const float* beam_vals, int64_t B, int64_t S,
              double* tod, int c_q, int c_u, bool need_map, TodArgs* a, bool* qu, int* CS) {
    if (!is_pow2(nside) || nside > 8192)
        return fail(TODG_ERR_ARG, "nside must be a power of two <= 8192");
    if (B < 0 || S < 0) return fail(TODG_ERR_ARG, "B and S must be non-negative");
    int perm[kMaxSlots];
    int nslots = 0;
    int rc = slot_layout(C, c_q, c_u, qu, perm, a->row, &nslots);
    if (rc) return rc;
    if (S > (int64_t(1) << 28)) return fail(TODG_ERR_ARG, "S must be <= 2^28 beam pixels");
    if (B > 0 && S > 0 &&
        (!vec_orig || !axes || !cos_a || !sin_a || !ax_pts || !cos_p || !sin_p ||
         (need_map && (!map_packed || !beam_vals || !tod))))
        return fail(TODG_ERR_ARG, "null pointer argument");
    const RingEntry* rings = nullptr;
    rc = get_ring_table(nside, &rings);
    if (rc) return rc;
    const double* trig = nullptr;
    rc = get_lam_table(&trig);
    if (rc) return rc;
    a->lam_trig = trig;
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
    a->corr = nullptr;
    a->corr_stats = nullptr;
    a->corr_bound = 0.f;
    a->corr_slot = (*qu && nslots > 2) ? 2 : 0;  // the first field outside the Q/U pair
    a->corr_count = 0;
    a->status = nullptr;
    a->dump = nullptr;
    a->map1 = nullptr;
    *CS = slots_stride(nslots);
    return TODG_OK;
}

bool known_variant(int mode, int variant) {
    if (variant == 0 || variant == 8) return true;
    if (mode < 2 && (variant == 1 || variant == 2 || variant == 4)) return true;
    return mode < 2 && variant >= kTileFirst && variant < kTileFirst + kTileCount;
}

int run_tod(int mode, const float* vec_orig, const float* axes, const float* cos_a,
            const float* sin_a, const float* ax_pts, const float* cos_p, const float* sin_p,
            int64_t nside, const float* map_packed, int C, const float* beam_vals, int64_t B,
            int64_t S, double* tod, int c_q, int c_u, int variant, int* status, void* stream) {
    TodArgs a;
    bool qu;
    int CS;
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                       beam_vals, B, S, tod, c_q, c_u, true, &a, &qu, &CS);
    if (rc) return rc;
    if (!known_variant(mode, variant))
        return fail(TODG_ERR_ARG, "unknown kernel variant " + std::to_string(variant));
    a.status = status;
    if (B == 0 || S == 0) return TODG_OK;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = mode == 0   ? launch_m0(a, qu, CS, variant, st)
                   : mode == 1 ? launch_m1(a, qu, CS, variant, st)
                               : launch_m2(a, qu, CS, variant, st);
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TODG_OK;
}

int run_learned(const float* vec_orig, const float* axes, const float* cos_a, const float* sin_a,
                const float* ax_pts, const float* cos_p, const float* sin_p, int64_t nside,
                const float* map_packed, int C, const float* beam_vals, int64_t B, int64_t S,
                double* tod, int c_q, int c_u, const float* params, float bound,
                int64_t n_corrected, unsigned long long* stats, void* stream) {
    TodArgs a;
    bool qu;
    int CS;
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                       beam_vals, B, S, tod, c_q, c_u, true, &a, &qu, &CS);
    if (rc) return rc;
    if (!(bound >= 0.f)) return fail(TODG_ERR_ARG, "bound must be >= 0 (inf: no bound)");
    if (B == 0 || S == 0) return TODG_OK;
    if (!params) return fail(TODG_ERR_ARG, "null corrector parameters");
    if (n_corrected < 0) return fail(TODG_ERR_ARG, "n_corrected must be >= 0");
    a.corr = params;
    a.corr_stats = stats;
    a.corr_bound = bound;
    a.corr_count = n_corrected < S ? n_corrected : S;
    hipError_t e = launch_m3(a, qu, CS, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TODG_OK;
}

int run_fcorr(const float* vec_orig, const float* axes, const float* cos_a, const float* sin_a,
              const float* ax_pts, const float* cos_p, const float* sin_p, int64_t nside,
              const float* map_packed, const float* map1_packed, int C, const float* beam_vals,
              int64_t B, int64_t S, double* tod, int c_q, int c_u, int* status, void* stream) {
    TodArgs a;
    bool qu;
    int CS;
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                       beam_vals, B, S, tod, c_q, c_u, true, &a, &qu, &CS);
    if (rc) return rc;
    if (B == 0 || S == 0) return TODG_OK;
    if (!map1_packed) return fail(TODG_ERR_ARG, "null correction channel");
    a.map1 = map1_packed;
    a.status = status;
    hipError_t e = launch_m4(a, qu, CS, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TODG_OK;
}

}  // namespace

namespace todg_internal {

int set_error(int code, const std::string& msg) { return fail(code, msg); }

int ring_table(int64_t nside, const RingEntry** out) { return get_ring_table(nside, out); }

int map_slots(int C, int c_q, int c_u, bool* qu, int perm[4], int row[4], int* nslots, int* cs) {
    int rc = slot_layout(C, c_q, c_u, qu, perm, row, nslots);
    if (rc) return rc;
    *cs = slots_stride(*nslots);
    return TODG_OK;
}

int n_cu() { return cu_count(); }

}  // namespace todg_internal

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

int todg_check_counts(unsigned long long* counts, int n, int reset) {
    if (!counts || n < 0) return fail(TODG_ERR_ARG, "todg_check_counts: counts must be a host array, n >= 0");
#ifdef TODG_CHECKED
    unsigned long long acc[todg_check::kKinds] = {};
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("todg_check_counts: ") + hipGetErrorString(e));
    for (auto rd : todg_check::readers()) {
        const int rc = rd(acc, reset != 0);
        if (rc != 0)
            return fail(TODG_ERR_HIP, std::string("todg_check_counts: ") +
                                               hipGetErrorString((hipError_t)rc));
    }
    for (int k = 0; k < n; ++k) counts[k] = k < todg_check::kKinds ? acc[k] : 0ull;
    return todg_check::kKinds;
#else
    (void)reset;
    for (int k = 0; k < n; ++k) counts[k] = 0ull;
    return fail(TODG_ERR_ARG, "todg_check_counts: library built without index checks (TODG_CHECKED)");
#endif
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
                            int c_u, int variant, void* stream) {
    return run_tod(0, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                   beam_vals, B, S, tod, c_q, c_u, variant, nullptr, stream);
}

int todg_gather_accum_nearest(const float* vec_orig, const float* axes, const float* cos_a,
                              const float* sin_a, const float* ax_pts, const float* cos_p,
                              const float* sin_p, int64_t nside, const float* map_packed, int C,
                              const float* beam_vals, int64_t B, int64_t S, double* tod,
                              int c_q, int c_u, int variant, void* stream) {
    return run_tod(1, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                   beam_vals, B, S, tod, c_q, c_u, variant, nullptr, stream);
}

int todg_gather_accum_bicubic(const float* vec_orig, const float* axes, const float* cos_a,
                              const float* sin_a, const float* ax_pts, const float* cos_p,
                              const float* sin_p, int64_t nside, const float* map_packed, int C,
                              const float* beam_vals, int64_t B, int64_t S, double* tod,
                              int c_q, int c_u, int variant, void* stream) {
    return run_tod(2, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                   beam_vals, B, S, tod, c_q, c_u, variant, nullptr, stream);
}

int todg_gather_accum(int mode, const float* vec_orig, const float* axes, const float* cos_a,
                      const float* sin_a, const float* ax_pts, const float* cos_p,
                      const float* sin_p, int64_t nside, const float* map_packed, int C,
                      const float* beam_vals, int64_t B, int64_t S, double* tod, int c_q, int c_u,
                      int variant, int* status, void* stream) {
    if (mode < 0 || mode > 2)
        return fail(TODG_ERR_ARG, "mode must be 0 (bilinear), 1 (nearest) or 2 (bicubic)");
    return run_tod(mode, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                   beam_vals, B, S, tod, c_q, c_u, variant, status, stream);
}

size_t todg_fp_dump_bytes(int64_t B, int64_t S) {
    if (B < 0 || S < 0) return 0;
    return (size_t)(65 * B * S);
}

int todg_fp_dump(int mode, const float* vec_orig, const float* axes, const float* cos_a,
                 const float* sin_a, const float* ax_pts, const float* cos_p, const float* sin_p,
                 int64_t nside, const float* map_packed, int C, const float* beam_vals, int64_t B,
                 int64_t S, double* tod, int c_q, int c_u, int64_t* dump, void* stream) {
    if (mode != 0 && mode != 1) return fail(TODG_ERR_ARG, "mode must be 0 (bilinear) or 1 (nearest)");
    TodArgs a;
    bool qu;
    int CS;
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                       beam_vals, B, S, tod, c_q, c_u, true, &a, &qu, &CS);
    if (rc) return rc;
    if (!((qu && CS == 4) || (!qu && CS == 1)))
        return fail(TODG_ERR_ARG, "index dumps run for I/Q/U with Q/U (C = 3) or one field (C = 1)");
    if (B * S > ((int64_t)1 << 31)) return fail(TODG_ERR_ARG, "index dumps hold at most 2^31 (b, s) pairs");
    if (B == 0 || S == 0) return TODG_OK;
    if (!dump) return fail(TODG_ERR_ARG, "null dump buffer");
    a.dump = dump;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // every (b, s) is written by the kernel: the fill marks nothing as missing
    TODG_HIP_CHECK(hipMemsetAsync(dump, 0xff, (size_t)(65 * B * S), st));
    const hipError_t e = mode == 0 ? launch_dump_m0(a, qu, st) : launch_dump_m1(a, qu, st);
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TODG_OK;
}

int todg_gather_accum_learned(const float* vec_orig, const float* axes, const float* cos_a,
                              const float* sin_a, const float* ax_pts, const float* cos_p,
                              const float* sin_p, int64_t nside, const float* map_packed, int C,
                              const float* beam_vals, int64_t B, int64_t S, double* tod, int c_q,
                              int c_u, const float* params, float bound, int64_t n_corrected,
                              unsigned long long* stats, void* stream) {
    return run_learned(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                       beam_vals, B, S, tod, c_q, c_u, params, bound, n_corrected, stats, stream);
}

int todg_rotation_params(const double* theta, const double* phi, const double* psi, int64_t B,
                         const double* centre, float* axes, float* cos_a, float* sin_a, float* ax_pts,
                         float* cos_p, float* sin_p, void* stream) {
    if (B < 0) return fail(TODG_ERR_ARG, "B must be non-negative");
    if (B == 0) return TODG_OK;
    if (!theta || !phi || !psi || !centre || !axes || !cos_a || !sin_a || !ax_pts || !cos_p || !sin_p)
        return fail(TODG_ERR_ARG, "null pointer argument");
    RotCentre ctr;
    for (int k = 0; k < 3; ++k) {
        ctr.c[k] = centre[k];
        ctr.n[k] = centre[3 + k];
    }
    for (int k = 0; k < 3; ++k)
        if (!isfinite(ctr.c[k]) || !isfinite(ctr.n[k])) return fail(TODG_ERR_ARG, "non-finite beam centre");
    hipLaunchKernelGGL(k_rotation_params, dim3((unsigned)((B + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), theta, phi, psi, B, ctr, axes, cos_a, sin_a,
                       ax_pts, cos_p, sin_p);
    TODG_HIP_CHECK(hipGetLastError());
    return TODG_OK;
}

int todg_gather_accum_fcorr(const float* vec_orig, const float* axes, const float* cos_a,
                            const float* sin_a, const float* ax_pts, const float* cos_p,
                            const float* sin_p, int64_t nside, const float* map_packed,
                            const float* map1_packed, int C, const float* beam_vals, int64_t B,
                            int64_t S, double* tod, int c_q, int c_u, int* status, void* stream) {
    return run_fcorr(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed,
                     map1_packed, C, beam_vals, B, S, tod, c_q, c_u, status, stream);
}

int todg_fcorr_channels(int64_t nside, const float* field, const double* coef, float* out0,
                        float* out1, int64_t stride, void* stream) {
    if (!is_pow2(nside) || nside > 8192)
        return fail(TODG_ERR_ARG, "nside must be a power of two <= 8192");
    if (!field || !coef || !out0 || !out1) return fail(TODG_ERR_ARG, "null pointer argument");
    if (stride < 1) return fail(TODG_ERR_ARG, "stride must be >= 1");
    FcorrCoef c;
    for (int k = 0; k < 32; ++k) {
        if (!isfinite(coef[k])) return fail(TODG_ERR_ARG, "non-finite stencil coefficient");
        c.c[k] = coef[k];
    }
    const RingEntry* rings = nullptr;
    int rc = get_ring_table(nside, &rings);
    if (rc) return rc;
    hipError_t e = launch_fcorr_channels(rings, (int)nside, field, c, out0, out1, stride,
                                         reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TODG_OK;
}

int todg_smlp_features(int64_t nside, const float* field, float* F, float* X, void* stream) {
    if (!is_pow2(nside) || nside > 8192)
        return fail(TODG_ERR_ARG, "nside must be a power of two <= 8192");
    if (!field || !F || !X) return fail(TODG_ERR_ARG, "null pointer argument");
    const RingEntry* rings = nullptr;
    int rc = get_ring_table(nside, &rings);
    if (rc) return rc;
    hipError_t e = launch_smlp_features(rings, (int)nside, field, F, X, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TODG_OK;
}

int todg_smlp_channels(int64_t nside, const float* field, const float* params, float* out0, float* out1,
                       int64_t stride, void* stream) {
    if (!is_pow2(nside) || nside > 8192)
        return fail(TODG_ERR_ARG, "nside must be a power of two <= 8192");
    if (!field || !params || !out0 || !out1) return fail(TODG_ERR_ARG, "null pointer argument");
    if (stride < 1) return fail(TODG_ERR_ARG, "stride must be >= 1");
    const RingEntry* rings = nullptr;
    int rc = get_ring_table(nside, &rings);
    if (rc) return rc;
    hipError_t e = launch_smlp_channels(rings, (int)nside, field, params, out0, out1, stride,
                                        reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TODG_OK;
}

int todg_interp_points(int mode, int64_t nside, const float* map, int64_t N, const double* z,
                       const double* phi, double* out, void* stream) {
    if (mode != 0 && mode != 1) return fail(TODG_ERR_ARG, "mode must be 0 (bilinear) or 1 (bicubic)");
    if (!is_pow2(nside) || nside > 8192)
        return fail(TODG_ERR_ARG, "nside must be a power of two <= 8192");
    if (N < 0) return fail(TODG_ERR_ARG, "N must be non-negative");
    if (N == 0) return TODG_OK;
    if (!map || !z || !phi || !out) return fail(TODG_ERR_ARG, "null pointer argument");
    const RingEntry* rings = nullptr;
    int rc = get_ring_table(nside, &rings);
    if (rc) return rc;
    const int64_t blocks = (N + 255) / 256;
    hipLaunchKernelGGL(k_interp_points, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), mode, (int)nside, rings, map,
                       12 * nside * nside, N, z, phi, out);
    TODG_HIP_CHECK(hipGetLastError());
    return TODG_OK;
}

int todg_gather_accum_batch(int mode, const todg_work_item* items, int n_items,
                            int64_t total_samples, int64_t nside, const float* map_packed, int C,
                            int c_q, int c_u, int chunk, unsigned int* queue, int* item_status,
                            int variant, void* stream) {
    if (mode < 0 || mode > 2)
        return fail(TODG_ERR_ARG, "mode must be 0 (bilinear), 1 (nearest) or 2 (bicubic)");
    if (!is_pow2(nside) || nside > 8192)
        return fail(TODG_ERR_ARG, "nside must be a power of two <= 8192");
    if (n_items < 0 || total_samples < 0) return fail(TODG_ERR_ARG, "negative sizes");
    if (n_items == 0 || total_samples == 0) return TODG_OK;
    if (!items || !map_packed || !queue) return fail(TODG_ERR_ARG, "null pointer argument");
    // claim counters stay below 2^31: <= 2^30 samples per queue plus the
    // overshoot of the final claims (blocks x chunk <= 2^12 x 2^16)
    if (chunk < 1 || chunk > (1 << 16)) return fail(TODG_ERR_ARG, "chunk must be in [1, 65536]");
    if (total_samples > ((int64_t)8 << 30)) return fail(TODG_ERR_ARG, "too many samples per launch");
    bool qu;
    int perm[kMaxSlots], nslots = 0;
    BatchArgs ba;
    int rc = slot_layout(C, c_q, c_u, &qu, perm, ba.row, &nslots);
    if (rc) return rc;
    const RingEntry* rings = nullptr;
    rc = get_ring_table(nside, &rings);
    if (rc) return rc;
    const double* trig = nullptr;
    rc = get_lam_table(&trig);
    if (rc) return rc;
    ba.lam_trig = trig;
    ba.items = items;
    ba.n_items = n_items;
    ba.total = total_samples;
    ba.map = map_packed;
    ba.rings = rings;
    ba.npix = 12 * nside * nside;
    ba.nside = (int)nside;
    ba.nslots = nslots;
    ba.queue = queue;
    ba.status = item_status;
    ba.chunk = chunk;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int CS = slots_stride(nslots);
    hipError_t e = mode == 0   ? launch_batch_m0(ba, qu, CS, variant, st)
                   : mode == 1 ? launch_batch_m1(ba, qu, CS, variant, st)
                               : launch_batch_m2(ba, qu, CS, variant, st);
    if (e != hipSuccess) return fail(TODG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return TODG_OK;
}

int todg_bilinear_indices(const float* vec_orig, const float* axes, const float* cos_a,
                          const float* sin_a, const float* ax_pts, const float* cos_p,
                          const float* sin_p, int64_t nside, int64_t B, int64_t S, int64_t b0,
                          int64_t nb, int64_t* pix, double* wgt, void* stream) {
    if (b0 < 0 || nb < 0 || b0 + nb > B) return fail(TODG_ERR_ARG, "sample range out of bounds");
    TodArgs a;
    bool qu;
    int CS;
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, nullptr, 1,
                       nullptr, B, S, nullptr, -1, -1, false, &a, &qu, &CS);
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
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, nullptr, 1,
                       nullptr, B, S, nullptr, -1, -1, false, &a, &qu, &CS);
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
