// Beam-convolution TOD kernels for gfx950 and their C-ABI entry points.
//
//   tod[c, b] += sum_s beam_vals[s] * spin2_c( map[ R_b * vec_orig[s] ] )
//
// Two kernels share one device index path (binary64, reference operation
// order, see healpix_ring.hpp):
//
// * k_tod_fp (default): one workgroup per sample at a time.  The sample's
//   footprint on the map -- the rings and ring slots that the rotated beam
//   and its bilinear neighbours can touch -- is laid out in LDS.  Phase 1
//   sweeps the beam: for every beam pixel it evaluates the rotated direction
//   and its neighbours and adds beam_val * weight into a per-pixel LDS
//   weight grid (ds_add_f64).  Phase 2 sweeps the footprint once: per map
//   pixel it computes the spin-2 factors and gathers I/Q/U with one 16-byte
//   load.  Spin-2 and the map gather therefore run once per footprint pixel
//   (~2.3k per sample at nside 2048 / 30') instead of once per neighbour
//   (4 x S ~ 25k), which is the per-sample spin-2 memo of the reference
//   (tod_bilinear.py:151-199) taken to its limit.  A neighbour that falls
//   outside the laid-out footprint is accumulated directly (spin-2 + gather
//   on the spot), so the result never depends on the footprint estimate.
//
// * k_tod_direct: one wavefront owns T consecutive samples and sweeps the
//   beam with its lanes, evaluating spin-2 and gathering per neighbour.
//   Kept as a cross-check and for callers that pick it explicitly.
//
// Both keep binary64 sums per lane (binary64 adds issue at the full VALU
// rate on CDNA4) and reduce across lanes once per sample; each sample's tod
// entry is written by exactly one thread, so there are no atomics on tod.
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
    double kx, ky, kz, ca, sa, oma, bx, by, bz, cp, sp, omp;
    double z_pts, sth_pts, phi_pts;
};

__device__ __forceinline__ void load_sample(const TodArgs& a, int64_t b, Sample& P) {
    P.kx = (double)a.axes[3 * b];
    P.ky = (double)a.axes[3 * b + 1];
    P.kz = (double)a.axes[3 * b + 2];
    P.ca = (double)a.cos_a[b];
    P.sa = (double)a.sin_a[b];
    P.oma = 1.0 - P.ca;
    P.bx = (double)a.ax_pts[3 * b];
    P.by = (double)a.ax_pts[3 * b + 1];
    P.bz = (double)a.ax_pts[3 * b + 2];
    P.cp = (double)a.cos_p[b];
    P.sp = (double)a.sin_p[b];
    P.omp = 1.0 - P.cp;
    // Boresight (z, sin theta, phi) for spin-2: tod_bilinear.py:306-310.
    P.z_pts = fmax(-1.0, fmin(1.0, P.bz));
    P.sth_pts = sqrt(fmax(0.0, 1.0 - P.bz * P.bz));
    double ph = atan2(P.by, P.bx);
    if (ph < 0.0) ph += kTwoPi;
    P.phi_pts = ph;
}

// Broadcast lane 0's value to the wave (an SGPR value for the compiler).
__device__ __forceinline__ double uniform(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ void make_uniform(Sample& P) {
    P.kx = uniform(P.kx);
    P.ky = uniform(P.ky);
    P.kz = uniform(P.kz);
    P.ca = uniform(P.ca);
    P.sa = uniform(P.sa);
    P.oma = uniform(P.oma);
    P.bx = uniform(P.bx);
    P.by = uniform(P.by);
    P.bz = uniform(P.bz);
    P.cp = uniform(P.cp);
    P.sp = uniform(P.sp);
    P.omp = uniform(P.omp);
    P.z_pts = uniform(P.z_pts);
    P.sth_pts = uniform(P.sth_pts);
    P.phi_pts = uniform(P.phi_pts);
}

__device__ __forceinline__ void rotate(const Sample& q, double v0, double v1, double v2,
                                       double& vx, double& vy, double& vz) {
    rodrigues(v0, v1, v2, q.kx, q.ky, q.kz, q.ca, q.sa, q.oma, q.bx, q.by, q.bz, q.cp, q.sp,
              q.omp, vx, vy, vz);
}

__device__ __forceinline__ double wrap_phi(double y, double x) {
    double p = atan2(y, x);
    if (p < 0.0) p += kTwoPi;
    return p;
}

template <int CS>
__device__ __forceinline__ float4 load_px(const float* __restrict__ map, int64_t p, int64_t npix) {
    // The reference never bounds-checks; an index can only leave [0, npix)
    // through the 1-ulp "tw == n" float-mod corner at the last ring, or from
    // a non-finite pointing.  Clamp the address (unsigned compare also
    // catches negatives) so such cases read a valid pixel, never fault.
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

__device__ __forceinline__ float slot(const float4& v, int k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Accumulate w * map-pixel (with spin-2 rotation of slots 0/1 when QU) into
// acc.  Pixel centre (z_n, sth_n, phi_n) feeds spin-2.
template <bool QU, int CS>
__device__ __forceinline__ void accum_pixel(const TodArgs& a, const Sample& q, int64_t p,
                                            double z_n, double sth_n, double phi_n, double w,
                                            double* acc) {
    const float4 m = load_px<CS>(a.map, p, a.npix);
    if constexpr (QU) {
        double c2d, s2d;
        spin2(z_n, sth_n, phi_n, q.z_pts, q.sth_pts, q.phi_pts, c2d, s2d);
        const double qv = (double)m.x, uv = (double)m.y;
        acc[0] = fma(w, fma(qv, c2d, uv * s2d), acc[0]);
        acc[1] = fma(w, fma(uv, c2d, -(qv * s2d)), acc[1]);
#pragma unroll
        for (int k = 2; k < CS; ++k)
            if (k < a.nslots) acc[k] = fma(w, (double)slot(m, k), acc[k]);
    } else {
#pragma unroll
        for (int k = 0; k < CS; ++k)
            if (k < a.nslots) acc[k] = fma(w, (double)slot(m, k), acc[k]);
    }
}

template <bool QU, int CS>
__device__ __forceinline__ void accum_ring_pos(const TodArgs& a, const Sample& q, int ring,
                                               int pos, double w, double* acc) {
    const RingEntry& e = a.rings[ring];
    accum_pixel<QU, CS>(a, q, e.first + pos, e.z, e.sth, e.phi0 + (double)pos * e.dphi, w, acc);
}

// ============================================================ direct kernel

// MODE 0 = bilinear (tod_bilinear.py:205-477), 1 = nearest (tod_nearest.py:21-158).
template <int MODE, bool QU, int CS, int T>
__global__ __launch_bounds__(256) void k_tod_direct(const TodArgs a) {
    const int lane = threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t wpb = blockDim.x >> 6;
    const int64_t wave0 = (int64_t)blockIdx.x * wpb + wib;
    const int64_t nwaves = (int64_t)gridDim.x * wpb;
    const int64_t ntiles = (a.B + T - 1) / T;

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
                double vx, vy, vz;
                rotate(P[t], v0, v1, v2, vx, vy, vz);
                const double phi_w = wrap_phi(vy, vx);
                if constexpr (MODE == 0) {
                    Nb o;
                    ring_interp(a.nside, a.rings, fmax(-1.0, fmin(1.0, vz)), phi_w, o);
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        accum_ring_pos<QU, CS>(a, P[t], o.ring[k], o.pos[k], o.w[k] * bv, acc[t]);
                } else {
                    const double sin_th = sqrt(fmax(0.0, vx * vx + vy * vy));
                    int ring, pos;
                    nearest_pix(a.nside, a.rings, vz, sin_th, phi_w, ring, pos);
                    accum_ring_pos<QU, CS>(a, P[t], ring, pos, bv, acc[t]);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int64_t b = b0 + t;
#pragma unroll
            for (int k = 0; k < kMaxSlots; ++k) {
                if (k < a.nslots && k < CS) {
                    const double sum = wave_sum(acc[t][k]);
                    if (lane == 0 && b < a.B) a.tod[(int64_t)a.row[k] * a.B + b] += sum;
                }
            }
        }
    }
}

// ================================================ out-of-line rare paths

// Four accumulator slots returned by value (no stack array).
struct Acc4 {
    double v[kMaxSlots];
};

// Direct contribution of one map pixel with weight w (spin-2 on slots 0/1
// when QU).  Kept out of line: the footprint kernel only takes it for the
// rare neighbour that falls outside its LDS grid.
template <bool QU, int CS>
__device__ __forceinline__ Acc4 direct_one(const float* __restrict__ map, int64_t npix, int nslots,
                                        int64_t p, double z_n, double sth_n, double phi_n,
                                        double w, double z_pts, double sth_pts,
                                        double phi_pts) {
    Acc4 r = {{0.0, 0.0, 0.0, 0.0}};
    const float4 m = load_px<CS>(map, p, npix);
    if constexpr (QU) {
        double c2d, s2d;
        spin2(z_n, sth_n, phi_n, z_pts, sth_pts, phi_pts, c2d, s2d);
        const double qv = (double)m.x, uv = (double)m.y;
        r.v[0] = w * fma(qv, c2d, uv * s2d);
        r.v[1] = w * fma(uv, c2d, -(qv * s2d));
#pragma unroll
        for (int k = 2; k < CS; ++k)
            if (k < nslots) r.v[k] = w * (double)slot(m, k);
    } else {
#pragma unroll
        for (int k = 0; k < CS; ++k)
            if (k < nslots) r.v[k] = w * (double)slot(m, k);
    }
    return r;
}

// ========================================================= footprint kernel

constexpr int kFpThreads = 512;
constexpr int kFpWaves = kFpThreads / 64;

// Certified fast index path (fp_phase1).  A ring or slot coordinate computed
// there differs from the reference's binary64 value by far less than 1e-10
// (see fp_phase1); its integer part is taken only when the coordinate is at
// least kCertSlot (kCertPolar for the polar-cap ring formula, whose slope
// grows near the pole) away from an integer, otherwise the beam pixel is
// recomputed by the exact path.
constexpr double kCertSlot = 1e-9;
constexpr double kCertPolar = 1e-6;
// Largest |tan| of the longitude offset from the footprint centre that the
// fast path's atan series takes (truncation t^13/13 < 1e-21 at 0.03).
constexpr double kMaxTan = 0.03;

// One footprint ring in LDS.  The fast loop reads the first 48 bytes.
struct alignas(16) FpRing {
    double base;      // slot coordinate of the footprint centre, window-relative
    double inv_dphi;  // 1 / dphi
    double z, sth, inv_dth;
    int width, off;   // window width in slots, first weight-grid cell
    double theta, phi0, dphi;
    int64_t first;
    int n, i0;        // ring pixels, first slot of the window
};

// Per-sample values shared by the workgroup (written by wave 0).
struct alignas(16) FpFrame {
    double m[9];   // m[3*i + j]: row i, column j of Rz(-phi_f) * R2 * R1
    double q[12];  // Sample rotation fields: kx ky kz ca sa oma bx by bz cp sp omp
    double z_pts, sth_pts, phi_pts;
    int ir_lo, nr, total, pad;
};

struct FpConf {
    int cap;        // weight-grid entries
    int max_rings;  // FpRing entries per buffer
};

// Local index of (ring entry e, slot pos in [0, n]) in the weight grid, or -1.
__device__ __forceinline__ int fp_local(const FpRing& e, int pos) {
    int d = pos - e.i0;
    if (d < 0) d += e.n;
    if (d >= e.n) d -= e.n;
    return d < e.width ? e.off + d : -1;
}

// Bilinear neighbours through the global ring table (pole caps, rings
// outside the layout, uncertified coordinates): LDS cell when laid out,
// direct contribution otherwise.
template <bool QU, int CS>
__device__ __forceinline__ Acc4 fp_slow_bilinear(const TodArgs* __restrict__ ap, double z,
                                                 double phi_w, double bv, const FpRing* rl,
                                                 int ir_lo, int nr, double* W, double z_pts,
                                                 double sth_pts, double phi_pts) {
    const TodArgs& a = *ap;
    Nb o;
    ring_interp(a.nside, a.rings, z, phi_w, o);
    Acc4 r = {{0.0, 0.0, 0.0, 0.0}};
    for (int k = 0; k < 4; ++k) {
        const int lr = o.ring[k] - ir_lo;
        const int li = (lr >= 0 && lr < nr) ? fp_local(rl[lr], o.pos[k]) : -1;
        if (li >= 0) {
            atomicAdd(&W[li], o.w[k] * bv);
        } else {
            const RingEntry& e = a.rings[o.ring[k]];
            const Acc4 d = direct_one<QU, CS>(a.map, a.npix, a.nslots, e.first + o.pos[k], e.z,
                                              e.sth, e.phi0 + (double)o.pos[k] * e.dphi,
                                              o.w[k] * bv, z_pts, sth_pts, phi_pts);
            for (int j = 0; j < kMaxSlots; ++j) r.v[j] += d.v[j];
        }
    }
    return r;
}

// LDS carve-up of the footprint kernels.  Ring descriptors and frames are
// double-buffered so that wave 0 can lay out sample b+1 while the other
// waves still read sample b's.
struct FpLds {
    FpRing* rings[2];  // fc.max_rings descriptors each
    FpFrame* frame[2];
    double* W;         // fc.cap weight-grid cells
    double* red;       // kFpWaves * kMaxSlots reduction slots
    double* scratch;   // kFpWaves doubles for fp_beam_frame
};

__device__ __forceinline__ FpLds fp_lds(unsigned char* smem, const FpConf& fc) {
    FpLds L;
    unsigned char* p = smem;
    L.rings[0] = reinterpret_cast<FpRing*>(p);
    p += (size_t)fc.max_rings * sizeof(FpRing);
    L.rings[1] = reinterpret_cast<FpRing*>(p);
    p += (size_t)fc.max_rings * sizeof(FpRing);
    L.frame[0] = reinterpret_cast<FpFrame*>(p);
    p += sizeof(FpFrame);
    L.frame[1] = reinterpret_cast<FpFrame*>(p);
    p += sizeof(FpFrame);
    L.W = reinterpret_cast<double*>(p);
    p += (size_t)fc.cap * sizeof(double);
    L.red = reinterpret_cast<double*>(p);
    p += (size_t)kFpWaves * kMaxSlots * sizeof(double);
    L.scratch = reinterpret_cast<double*>(p);
    return L;
}

// Beam centre c (unit vector in the beam frame) and the footprint disc
// radius R around its image R_b c.
struct BeamFrame {
    double cx, cy, cz, R, cosR;
};

// All threads of the block: the beam centre c = R1^-1(boresight) of sample
// b0 and the beam radius about it.  For a beam prepared around its centre
// every sample's rotation maps c onto its boresight; any c yields a correct
// footprint disc around R_b c (rotations keep angles), a good c a small one.
__device__ __forceinline__ BeamFrame fp_beam_frame(const TodArgs& a, int64_t b0, const FpLds& L) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    double* red = L.scratch;
    double cx, cy, cz, cosr;
    {
        Sample q;
        load_sample(a, b0, q);
        const double px = q.bx, py = q.by, pz = q.bz;
        const double dkv = q.kx * px + q.ky * py + q.kz * pz;
        cx = px * q.ca - (q.ky * pz - q.kz * py) * q.sa + q.kx * dkv * q.oma;
        cy = py * q.ca - (q.kz * px - q.kx * pz) * q.sa + q.ky * dkv * q.oma;
        cz = pz * q.ca - (q.kx * py - q.ky * px) * q.sa + q.kz * dkv * q.oma;
        const double nn = fma(cx, cx, fma(cy, cy, cz * cz));
        const double nrm = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
        cx *= nrm;
        cy *= nrm;
        cz = nn > 0.0 ? cz * nrm : 1.0;
        double mn = 2.0;
        for (int64_t s = tid; s < a.S; s += kFpThreads) {
            const double d = (double)a.vec_orig[3 * s] * cx + (double)a.vec_orig[3 * s + 1] * cy +
                             (double)a.vec_orig[3 * s + 2] * cz;
            mn = fmin(mn, d);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) mn = fmin(mn, __shfl_xor(mn, off, 64));
        if (lane == 0) red[wave] = mn;
        __syncthreads();
        mn = red[0];
        for (int w = 1; w < kFpWaves; ++w) mn = fmin(mn, red[w]);
        __syncthreads();
        cosr = fmax(-1.0, fmin(1.0, mn));
    }
    // Disc radius for the ring range and the per-ring longitude half-width:
    // beam radius + 2 pixel sizes (bilinear neighbours lie within ~1.5 of the
    // query) + a rounding allowance.
    BeamFrame f;
    f.cx = cx;
    f.cy = cy;
    f.cz = cz;
    f.R = fmin(kPi, acos(cosr) + 2.0 / (double)a.nside + 1e-6);
    f.cosR = cos(f.R);
    return f;
}

// Wave 0 only: sample b's frame and footprint layout -- every ring and ring
// slot that the rotated beam's bilinear neighbours can touch, as per-ring
// slot windows packed into the weight grid.
__device__ __forceinline__ void fp_layout(const TodArgs& a, int64_t b, const BeamFrame& bf,
                                          const FpConf& fc, FpFrame* fr, FpRing* rings) {
    const int lane = threadIdx.x & 63;
    const int nside = a.nside;
    Sample q;
    load_sample(a, b, q);
    make_uniform(q);
    double fx, fy, fz;  // footprint centre = this sample's image of c
    rotate(q, bf.cx, bf.cy, bf.cz, fx, fy, fz);
    const double fzc = fmax(-1.0, fmin(1.0, fz));
    const double fth = acos(fzc);
    const double fst = sqrt(fmax(0.0, (1.0 - fzc) * (1.0 + fzc)));
    double fph = atan2(fy, fx);
    if (fph < 0.0) fph += kTwoPi;
    const double rho = sqrt(fx * fx + fy * fy);
    const double cf = rho > 0.0 ? fx / rho : 1.0;
    const double sf = rho > 0.0 ? fy / rho : 0.0;

    const double th_lo = fth - bf.R, th_hi = fth + bf.R;
    int ir_lo = th_lo <= 0.0 ? 1 : ring_above(nside, cos(th_lo));
    int ir_hi = th_hi >= kPi ? 4 * nside - 1 : ring_above(nside, cos(th_hi)) + 1;
    ir_lo = ir_lo < 1 ? 1 : ir_lo;
    ir_hi = ir_hi > 4 * nside - 1 ? 4 * nside - 1 : ir_hi;
    int nr = ir_hi - ir_lo + 1;
    if (nr > fc.max_rings || nr < 0) nr = 0;  // no layout: everything goes direct

    // Composite rotation: column j of R2 R1 is the reference rotation of the
    // unit vector e_j; the rows are then turned by -phi_f about z, so that
    // the rotated direction's longitude is phi_f + atan2(row1.v, row0.v).
    if (lane < 3) {
        double mx, my, mz;
        rotate(q, lane == 0 ? 1.0 : 0.0, lane == 1 ? 1.0 : 0.0, lane == 2 ? 1.0 : 0.0, mx, my,
               mz);
        fr->m[lane] = fma(cf, mx, sf * my);
        fr->m[3 + lane] = fma(cf, my, -(sf * mx));
        fr->m[6 + lane] = mz;
    }

    int carry = 0, laid = 0;
    for (int r0 = 0; r0 < nr; r0 += 64) {
        const int r = r0 + lane;
        int width = 0, i0 = 0;
        double base = 0.0;
        FpRing e;
        if (r < nr) {
            const RingEntry& g = a.rings[ir_lo + r];
            e.inv_dphi = g.inv_dphi;
            e.z = g.z;
            e.sth = g.sth;
            e.inv_dth = g.inv_dth;
            e.theta = g.theta;
            e.phi0 = g.phi0;
            e.dphi = g.dphi;
            e.first = g.first;
            e.n = g.n;
            const double tc = (fph - g.phi0) * g.inv_dphi;  // centre's slot coordinate
            const double den = fst * g.sth;
            const double cd = den > 1e-12 ? (bf.cosR - fzc * g.z) / den : -2.0;
            bool full = false;
            if (cd <= -1.0) {
                full = true;
            } else if (cd < 1.0) {
                const double dp = acos(cd);
                const double wf = 2.0 * dp / g.dphi + 6.0;
                if (wf >= (double)g.n) {
                    full = true;
                } else {
                    width = (int)wf;
                    const int f = (int)floor((fph - dp - g.phi0) / g.dphi) - 2;
                    base = tc - (double)f;
                    const int fm = f % g.n;
                    i0 = fm < 0 ? fm + g.n : fm;
                }
            }
            if (full) {
                width = g.n;
                i0 = 0;
                base = tc >= (double)g.n ? tc - (double)g.n : (tc < 0.0 ? tc + (double)g.n : tc);
            }
        }
        int incl = width;  // inclusive scan of widths across the wave
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        const int off = carry + incl - width;
        const bool fits = off + width <= fc.cap;
        int end = (r < nr && fits) ? off + width : 0;
        if (r < nr) {
            e.base = base;
            e.i0 = i0;
            e.width = fits ? width : 0;  // beyond capacity: accumulated directly
            e.off = fits ? off : 0;
            rings[r] = e;
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) end = max(end, __shfl_xor(end, d, 64));
        laid = max(laid, end);
        carry += __shfl(incl, 63, 64);
    }
    if (lane == 0) {
        fr->q[0] = q.kx;
        fr->q[1] = q.ky;
        fr->q[2] = q.kz;
        fr->q[3] = q.ca;
        fr->q[4] = q.sa;
        fr->q[5] = q.oma;
        fr->q[6] = q.bx;
        fr->q[7] = q.by;
        fr->q[8] = q.bz;
        fr->q[9] = q.cp;
        fr->q[10] = q.sp;
        fr->q[11] = q.omp;
        fr->z_pts = q.z_pts;
        fr->sth_pts = q.sth_pts;
        fr->phi_pts = q.phi_pts;
        fr->ir_lo = ir_lo;
        fr->nr = nr;
        fr->total = laid;  // grid cells [0, laid) belong to rings
    }
}

__device__ __forceinline__ void frame_sample(const FpFrame* fr, Sample& q) {
    q.kx = fr->q[0];
    q.ky = fr->q[1];
    q.kz = fr->q[2];
    q.ca = fr->q[3];
    q.sa = fr->q[4];
    q.oma = fr->q[5];
    q.bx = fr->q[6];
    q.by = fr->q[7];
    q.bz = fr->q[8];
    q.cp = fr->q[9];
    q.sp = fr->q[10];
    q.omp = fr->q[11];
    q.z_pts = fr->z_pts;
    q.sth_pts = fr->sth_pts;
    q.phi_pts = fr->phi_pts;
    make_uniform(q);
}

// sqrt(x) for x > 0 from the hardware reciprocal square root and one
// Goldschmidt step (relative error ~1e-14).  Weights only.
__device__ __forceinline__ double sqrt_fast(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    const double s = x * r;
    const double h = 0.5 * r;
    const double e = fma(-s, h, 0.5);
    return fma(s, e, s);
}

// Bilinear fast path for one beam pixel (unit vector u, beam value bv):
// returns false, touching nothing, when the exact path must take it.
//
// Index certification.  The rotated direction comes from the composite
// matrix instead of two sequential Rodrigues steps, and its longitude as
// phi_f + atan(t) instead of atan2; both differ from the reference's binary64
// values by a few ulp (|dz| < 1e-15, |dphi| < 1e-15 for sin(theta) > 0.05).
// A ring coordinate nside*(2 - 1.5 z) then differs by < 1e-11, a polar one
// nside*sqrt(3(1-|z|)) by < 1e-7, and a slot coordinate (phi - phi0)/dphi by
// < 1e-11 (including the reference's own rounding of its division), so a
// coordinate whose fractional part lies in [eps, 1 - eps] has the
// reference's integer part: ring_above, the two ring slots and hence all four
// pixels are the reference's.  Everything else -- margins, pole caps, rings
// or slots outside the laid-out windows, large longitude offsets -- goes to
// the exact path.  Weights only need tolerance accuracy.
__device__ __forceinline__ bool fp_fast_bilinear(const double* __restrict__ m, int nside,
                                                 int ir_lo, int nr, const FpRing* rings,
                                                 double* W, double u0, double u1, double u2,
                                                 double bv) {
    const double xp = fma(m[0], u0, fma(m[1], u1, m[2] * u2));
    const double yp = fma(m[3], u0, fma(m[4], u1, m[5] * u2));
    const double zr = fma(m[6], u0, fma(m[7], u1, m[8] * u2));
    const double z = fmax(-1.0, fmin(1.0, zr));
    const double az = fabs(z);
    double tr;
    int ira;
    bool ok;
    if (az <= kTwoThirds) {
        tr = fma(z, -1.5 * (double)nside, 2.0 * (double)nside);
        ira = (int)tr;
        const double fr = tr - (double)ira;
        ok = fr > kCertSlot && fr < 1.0 - kCertSlot;
    } else {
        tr = (double)nside * sqrt_fast(3.0 * (1.0 - az));
        const int it = (int)tr;
        ira = z < 0.0 ? 4 * nside - it - 1 : it;
        const double fr = tr - (double)it;
        ok = fr > kCertPolar && fr < 1.0 - kCertPolar;
    }
    ok = ok && fabs(az - kTwoThirds) > 1e-12;
    const int la = ira - ir_lo;
    ok = ok && ira >= 1 && ira <= 4 * nside - 2 && la >= 0 && la + 1 < nr;
    ok = ok && xp > 0.0 && fabs(yp) < kMaxTan * xp;
    if (!ok) return false;
    const FpRing& A = rings[la];
    const FpRing& Bn = rings[la + 1];
    // longitude offset from the footprint centre: atan(t), |t| < 0.03
    const double t = yp * rcp_nr(xp);
    const double t2 = t * t;
    double p = fma(t2, -1.0 / 11.0, 1.0 / 9.0);
    p = fma(t2, p, -1.0 / 7.0);
    p = fma(t2, p, 1.0 / 5.0);
    p = fma(t2, p, -1.0 / 3.0);
    const double dl = fma(t * t2, p, t);
    const double ta = fma(dl, A.inv_dphi, A.base);
    const double tb = fma(dl, Bn.inv_dphi, Bn.base);
    const int ia = (int)ta, ib = (int)tb;
    const double fa = ta - (double)ia, fb = tb - (double)ib;
    ok = ta >= 0.0 && tb >= 0.0 && ia + 1 < A.width && ib + 1 < Bn.width;
    ok = ok && fa > kCertSlot && fa < 1.0 - kCertSlot && fb > kCertSlot && fb < 1.0 - kCertSlot;
    // theta - theta_above = asin(sin(theta) z_a - z sin(theta_a))
    const double sq = sqrt_fast((1.0 - z) * (1.0 + z));
    const double x = fma(sq, A.z, -(z * A.sth));
    ok = ok && fabs(x) < 1e-2;
    if (!ok) return false;
    const double x2 = x * x;
    double pp = fma(x2, 105.0 / 3456.0, 15.0 / 336.0);
    pp = fma(x2, pp, 3.0 / 40.0);
    pp = fma(x2, pp, 1.0 / 6.0);
    const double wb = fma(x * x2, pp, x) * A.inv_dth * bv;
    const double wa = bv - wb;
    double* ca = W + A.off + ia;
    double* cb = W + Bn.off + ib;
    atomicAdd(ca, wa - wa * fa);
    atomicAdd(ca + 1, wa * fa);
    atomicAdd(cb, wb - wb * fb);
    atomicAdd(cb + 1, wb * fb);
    return true;
}

// All threads: phase 1 of one sample -- sweep the beam and scatter
// beam_val * weight into the LDS weight grid; returns the outside-grid sums.
template <int MODE, bool QU, int CS>
__device__ __forceinline__ Acc4 fp_phase1(const TodArgs& a, const FpFrame* fr,
                                          const FpRing* rings, double* W) {
    const int tid = threadIdx.x;
    const int nside = a.nside;
    const int ir_lo = __builtin_amdgcn_readfirstlane(fr->ir_lo);
    const int nr = __builtin_amdgcn_readfirstlane(fr->nr);
    double m[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = uniform(fr->m[k]);
    Sample qn;  // nearest mode rotates with the exact path
    if constexpr (MODE != 0) frame_sample(fr, qn);
    Acc4 accd = {{0.0, 0.0, 0.0, 0.0}};  // outside-grid sums
    for (int64_t s0 = 0; s0 < a.S; s0 += (int64_t)kFpThreads * 64) {
        uint64_t slow = 0;
        float nv0 = 0.f, nv1 = 0.f, nv2 = 0.f, nbv = 0.f;
        int64_t s = s0 + tid;
        if (s < a.S) {
            nv0 = a.vec_orig[3 * s];
            nv1 = a.vec_orig[3 * s + 1];
            nv2 = a.vec_orig[3 * s + 2];
            nbv = a.beam_vals[s];
        }
        for (int it = 0; it < 64 && s < a.S; ++it, s += kFpThreads) {
            const double v0 = (double)nv0, v1 = (double)nv1, v2 = (double)nv2;
            const double bv = (double)nbv;
            const int64_t sn = s + kFpThreads;
            if (it + 1 < 64 && sn < a.S) {  // fetch the next beam pixel early
                nv0 = a.vec_orig[3 * sn];
                nv1 = a.vec_orig[3 * sn + 1];
                nv2 = a.vec_orig[3 * sn + 2];
                nbv = a.beam_vals[sn];
            }
            bool ok = false;
            if constexpr (MODE == 0) {
                ok = fp_fast_bilinear(m, nside, ir_lo, nr, rings, W, v0, v1, v2, bv);
            } else {
                double vx, vy, vz;
                rotate(qn, v0, v1, v2, vx, vy, vz);
                const double phi_w = wrap_phi(vy, vx);
                const double sin_th = sqrt(fmax(0.0, vx * vx + vy * vy));
                int ring, pos;
                nearest_pix(nside, a.rings, vz, sin_th, phi_w, ring, pos);
                const int lr = ring - ir_lo;
                const int li = (lr >= 0 && lr < nr) ? fp_local(rings[lr], pos) : -1;
                if (li >= 0) {
                    atomicAdd(&W[li], bv);
                    ok = true;
                }
            }
            if (!ok) slow |= 1ull << it;
        }
        // exact path for the flagged beam pixels
        if (slow) {
            Sample q;
            frame_sample(fr, q);
            while (slow) {
                const int it = __builtin_ctzll(slow);
                slow &= slow - 1;
                const int64_t ss = s0 + (int64_t)it * kFpThreads + tid;
                const double v0 = (double)a.vec_orig[3 * ss];
                const double v1 = (double)a.vec_orig[3 * ss + 1];
                const double v2 = (double)a.vec_orig[3 * ss + 2];
                const double bv = (double)a.beam_vals[ss];
                double vx, vy, vz;
                rotate(q, v0, v1, v2, vx, vy, vz);
                const double phi_w = wrap_phi(vy, vx);
                if constexpr (MODE == 0) {
                    const Acc4 d = fp_slow_bilinear<QU, CS>(&a, fmax(-1.0, fmin(1.0, vz)), phi_w,
                                                            bv, rings, ir_lo, nr, W, q.z_pts,
                                                            q.sth_pts, q.phi_pts);
#pragma unroll
                    for (int j = 0; j < kMaxSlots; ++j) accd.v[j] += d.v[j];
                } else {
                    const double sin_th = sqrt(fmax(0.0, vx * vx + vy * vy));
                    int ring, pos;
                    nearest_pix(nside, a.rings, vz, sin_th, phi_w, ring, pos);
                    const RingEntry& e = a.rings[ring];
                    const Acc4 d = direct_one<QU, CS>(a.map, a.npix, a.nslots, e.first + pos, e.z,
                                                      e.sth, e.phi0 + (double)pos * e.dphi, bv,
                                                      q.z_pts, q.sth_pts, q.phi_pts);
#pragma unroll
                    for (int j = 0; j < kMaxSlots; ++j) accd.v[j] += d.v[j];
                }
            }
        }
    }
    return accd;
}

// All threads: phase 2 -- sweep the laid-out footprint once: spin-2 and one
// gather per map pixel with nonzero weight; the cell is cleared for the next
// sample.  Returns the per-thread sums (outside-grid sums included).
template <bool QU, int CS>
__device__ __forceinline__ Acc4 fp_phase2(const TodArgs& a, const FpFrame* fr,
                                          const FpRing* rings, double* W, Acc4 acc) {
    const int tid = threadIdx.x;
    const int nr = __builtin_amdgcn_readfirstlane(fr->nr);
    const int total = __builtin_amdgcn_readfirstlane(fr->total);
    const double z_pts = uniform(fr->z_pts), sth_pts = uniform(fr->sth_pts);
    const double phi_pts = uniform(fr->phi_pts);
    int r = 0;
    for (int j = tid; j < total; j += kFpThreads) {
        while (r + 1 < nr && j >= rings[r].off + rings[r].width) ++r;
        const double wsum = W[j];
        if (wsum == 0.0) continue;
        W[j] = 0.0;
        const FpRing& e = rings[r];
        int pos = e.i0 + (j - e.off);
        if (pos >= e.n) pos -= e.n;
        const float4 mv = load_px<CS>(a.map, e.first + pos, a.npix);
        if constexpr (QU) {
            const double phi_n = e.phi0 + (double)pos * e.dphi;
            double c2d, s2d;
            if (phi_pts == phi_n && sth_pts == e.sth && z_pts == e.z) {
                c2d = 1.0;
                s2d = 0.0;
            } else {
                double h = (phi_pts - phi_n) * 0.5;
                if (h > 0.5 * kPi) h -= kPi;
                if (h < -0.5 * kPi) h += kPi;
                double sh, ch;
                sincos_half(h, sh, ch);
                spin2_from_half(spin2_ring(e.z, e.sth, z_pts, sth_pts), sh, ch, c2d, s2d);
            }
            const double qv = (double)mv.x, uv = (double)mv.y;
            acc.v[0] = fma(wsum, fma(qv, c2d, uv * s2d), acc.v[0]);
            acc.v[1] = fma(wsum, fma(uv, c2d, -(qv * s2d)), acc.v[1]);
#pragma unroll
            for (int k = 2; k < CS; ++k)
                if (k < a.nslots) acc.v[k] = fma(wsum, (double)slot(mv, k), acc.v[k]);
        } else {
#pragma unroll
            for (int k = 0; k < CS; ++k)
                if (k < a.nslots) acc.v[k] = fma(wsum, (double)slot(mv, k), acc.v[k]);
        }
    }
    return acc;
}

// A sample whose block sums sit in red[] until thread 0 writes them.
struct FpPending {
    double* tod;
    int64_t B, b;
};

// Thread 0: add the pending sample's block sums to its tod entries, slots in
// order (two slots may share a row when c_q == c_u).  Callers guarantee that
// every wave's red[] entries are complete (a barrier since they were made).
__device__ __forceinline__ void fp_flush(const TodArgs& a, const FpLds& L, FpPending& pend) {
    if (threadIdx.x == 0 && pend.tod) {
        for (int k = 0; k < kMaxSlots && k < a.nslots; ++k) {
            double sum = 0.0;
            for (int w = 0; w < kFpWaves; ++w) sum += L.red[w * kMaxSlots + k];
            pend.tod[(int64_t)a.row[k] * pend.B + pend.b] += sum;
        }
    }
    pend.tod = nullptr;
}

// All threads of the block: one sample through layout, both phases and the
// block reduction; two barriers per sample.  The tod entries of sample b are
// written after the next sample's first barrier (or by fp_drain).
template <int MODE, bool QU, int CS>
__device__ __forceinline__ void fp_sample(const TodArgs& a, int64_t b, const BeamFrame& bf,
                                          const FpConf& fc, const FpLds& L, int& buf,
                                          FpPending& pend) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    FpFrame* fr = L.frame[buf];
    FpRing* rings = L.rings[buf];
    if (wave == 0) fp_layout(a, b, bf, fc, fr, rings);
    __syncthreads();
    fp_flush(a, L, pend);
    Acc4 acc = fp_phase1<MODE, QU, CS>(a, fr, rings, L.W);
    __syncthreads();
    acc = fp_phase2<QU, CS>(a, fr, rings, L.W, acc);
#pragma unroll
    for (int k = 0; k < kMaxSlots; ++k) {
        if (k < CS) {
            const double v = wave_sum(acc.v[k]);
            if (lane == 0) L.red[wave * kMaxSlots + k] = v;
        }
    }
    pend.tod = a.tod;
    pend.B = a.B;
    pend.b = b;
    buf ^= 1;
}

__device__ __forceinline__ void fp_drain(const TodArgs& a, const FpLds& L, FpPending& pend) {
    __syncthreads();
    fp_flush(a, L, pend);
    __syncthreads();
}

__device__ __forceinline__ void fp_clear_grid(const FpConf& fc, const FpLds& L) {
    for (int j = threadIdx.x; j < fc.cap; j += kFpThreads) L.W[j] = 0.0;
    __syncthreads();
}

template <int MODE, bool QU, int CS, int MINW>
__global__ __launch_bounds__(kFpThreads, MINW) void k_tod_fp(const TodArgs a, const FpConf fc) {
    extern __shared__ __align__(16) unsigned char smem[];
    const FpLds L = fp_lds(smem, fc);
    // XCD-aware sample order: workgroups are dispatched round-robin over the
    // 8 XCDs, so block bid runs on XCD bid % 8.  Each XCD gets a contiguous
    // eighth of the samples, so concurrently processed samples on one XCD
    // have overlapping footprints and share that XCD's L2.
    const int G = gridDim.x;
    const int xcd = blockIdx.x & 7;
    const int per = G >> 3;
    const int kx = blockIdx.x >> 3;
    const int64_t chunk = (a.B + 7) / 8;
    const int64_t bbeg = (int64_t)xcd * chunk;
    const int64_t bend = (bbeg + chunk < a.B) ? bbeg + chunk : a.B;
    if (bbeg + kx >= bend) return;  // uniform over the block
    fp_clear_grid(fc, L);
    const BeamFrame bf = fp_beam_frame(a, bbeg + kx, L);
    int buf = 0;
    FpPending pend = {nullptr, 0, 0};
    for (int64_t b = bbeg + kx; b < bend; b += per)
        fp_sample<MODE, QU, CS>(a, b, bf, fc, L, buf, pend);
    fp_drain(a, L, pend);
}

// ======================================================= batched kernel

// Shared fields of a batch; per-detector fields come from todg_work_item.
struct BatchArgs {
    const todg_work_item* items;  // device array
    int n_items;
    int64_t total;                // sum of items[i].B
    const float* map;
    const RingEntry* rings;
    int64_t npix;
    int nside;
    int nslots;
    int row[kMaxSlots];
    unsigned int* queue;          // 8 per-XCD claim counters, zeroed before launch
    int* status;                  // per-item status words or null
    int chunk;                    // samples per claim
};

__device__ __forceinline__ int find_item(const todg_work_item* items, int n, int64_t g) {
    int lo = 0, hi = n - 1;  // last item with first <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].first <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ bool finite_sample(const TodArgs& a, int64_t b) {
    const float v[10] = {a.axes[3 * b], a.axes[3 * b + 1], a.axes[3 * b + 2], a.cos_a[b], a.sin_a[b],
                         a.ax_pts[3 * b], a.ax_pts[3 * b + 1], a.ax_pts[3 * b + 2], a.cos_p[b],
                         a.sin_p[b]};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 10; ++k) ok = ok && isfinite(v[k]);
    return ok;
}

// Persistent multi-detector kernel.  The global sample range (all items'
// samples back to back) is split into 8 contiguous parts, one work queue per
// XCD; a workgroup claims `chunk` samples at a time from its own XCD's
// queue and, once that is drained, from the others, so uneven beams and
// chunk lengths balance themselves.  Every workgroup exits when all eight
// queues are drained.  A sample with a non-finite pointing operand is
// skipped and flags its item's status word.
template <int MODE, bool QU, int CS, int MINW>
__global__ __launch_bounds__(kFpThreads, MINW) void k_tod_fp_batch(const BatchArgs ba,
                                                                   const FpConf fc) {
    extern __shared__ __align__(16) unsigned char smem[];
    const FpLds L = fp_lds(smem, fc);
    const int tid = threadIdx.x;
    const int xcd = blockIdx.x & 7;
    const int64_t qlen = (ba.total + 7) / 8;
    int cur = -1;
    TodArgs a;
    a.map = ba.map;
    a.rings = ba.rings;
    a.npix = ba.npix;
    a.nside = ba.nside;
    a.nslots = ba.nslots;
#pragma unroll
    for (int k = 0; k < kMaxSlots; ++k) a.row[k] = ba.row[k];
    BeamFrame bf = {0.0, 0.0, 1.0, 0.0, 1.0};
    bool have_frame = false;
    int buf = 0;
    FpPending pend = {nullptr, 0, 0};
    fp_clear_grid(fc, L);
    int* claim = reinterpret_cast<int*>(L.scratch + kFpWaves);
    for (int att = 0; att < 8;) {
        const int qi = (xcd + att) & 7;
        const int64_t qbeg = (int64_t)qi * qlen;
        const int64_t qend = (qbeg + qlen < ba.total) ? qbeg + qlen : ba.total;
        if (tid == 0) *claim = (int)atomicAdd(&ba.queue[qi], (unsigned int)ba.chunk);
        __syncthreads();
        const int64_t g0 = qbeg + (int64_t)(unsigned int)*claim;
        __syncthreads();
        if (g0 >= qend) {
            ++att;
            continue;
        }
        const int64_t g1 = (g0 + ba.chunk < qend) ? g0 + ba.chunk : qend;
        for (int64_t g = g0; g < g1; ++g) {
            const int it = find_item(ba.items, ba.n_items, g);
            const todg_work_item& w = ba.items[it];
            if (it != cur) {
                a.vec_orig = w.vec_orig;
                a.beam_vals = w.beam_vals;
                a.axes = w.axes;
                a.cos_a = w.cos_a;
                a.sin_a = w.sin_a;
                a.ax_pts = w.ax_pts;
                a.cos_p = w.cos_p;
                a.sin_p = w.sin_p;
                a.tod = w.tod;
                a.B = w.B;
                a.S = w.S;
                cur = it;
                have_frame = false;
            }
            const int64_t b = g - w.first;
            if (a.S <= 0) continue;
            if (!finite_sample(a, b)) {
                if (tid == 0 && ba.status) atomicOr(&ba.status[it], 1);
                continue;
            }
            if (!have_frame) {  // beam frame from the item's first finite sample seen here
                bf = fp_beam_frame(a, b, L);
                have_frame = true;
            }
            fp_sample<MODE, QU, CS>(a, b, bf, fc, L, buf, pend);
        }
    }
    fp_drain(a, L, pend);
}

__global__ void k_zero_u32(unsigned int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
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

// Footprint-kernel LDS configuration: two buffers of ring descriptors and
// frames, the weight grid, reduction and scratch slots (~62 KB per
// workgroup: two workgroups of 8 waves per CU).
constexpr int kFpMaxRings = 128;
constexpr int kFpCap = 4608;

size_t fp_smem_bytes(const FpConf& fc) {
    return 2 * (size_t)fc.max_rings * sizeof(FpRing) + 2 * sizeof(FpFrame) +
           (size_t)fc.cap * sizeof(double) + (size_t)kFpWaves * kMaxSlots * sizeof(double) +
           (size_t)(kFpWaves + 2) * sizeof(double);
}

template <int MODE, bool QU, int CS>
hipError_t launch_direct_cs(const TodArgs& a, int T, hipStream_t st) {
    const int64_t ntiles = (a.B + T - 1) / T;
    const int64_t waves_cap = (int64_t)cu_count() * 32;
    const int64_t waves = ntiles < waves_cap ? ntiles : waves_cap;
    const int blocks = (int)((waves + 3) / 4);
    switch (T) {
        case 1: hipLaunchKernelGGL((k_tod_direct<MODE, QU, CS, 1>), dim3(blocks), dim3(256), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_tod_direct<MODE, QU, CS, 2>), dim3(blocks), dim3(256), 0, st, a); break;
        default: hipLaunchKernelGGL((k_tod_direct<MODE, QU, CS, 4>), dim3(blocks), dim3(256), 0, st, a); break;
    }
    return hipGetLastError();
}

// MINW = waves per SIMD the kernel is compiled for: 2 (256 VGPRs, no spills,
// one workgroup per CU) or 4 (128 VGPRs, two workgroups per CU).
template <int MODE, bool QU, int CS, int MINW>
hipError_t launch_fp_cs(const TodArgs& a, hipStream_t st) {
    const FpConf fc{kFpCap, kFpMaxRings};
    const size_t smem = fp_smem_bytes(fc);
    static std::once_flag once;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)k_tod_fp<MODE, QU, CS, MINW>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    });
    // Workgroups resident per CU (LDS allows two), a multiple of 8 for the
    // XCD mapping, no more than the samples need.
    int64_t blocks = (int64_t)cu_count() * (MINW >= 4 ? 2 : 1) * 2;
    const int64_t need = ((a.B + 7) / 8) * 8;
    if (blocks > need) blocks = need;
    blocks = ((blocks + 7) / 8) * 8;
    hipLaunchKernelGGL((k_tod_fp<MODE, QU, CS, MINW>), dim3((unsigned)blocks), dim3(kFpThreads),
                       smem, st, a, fc);
    return hipGetLastError();
}

template <int MODE, bool QU, int CS, int MINW>
hipError_t launch_batch_cs(const BatchArgs& ba, hipStream_t st) {
    const FpConf fc{kFpCap, kFpMaxRings};
    const size_t smem = fp_smem_bytes(fc);
    static std::once_flag once;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)k_tod_fp_batch<MODE, QU, CS, MINW>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    });
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, ba.queue, 8);
    int64_t blocks = (int64_t)cu_count() * (MINW >= 4 ? 2 : 1) * 2;
    blocks = ((blocks + 7) / 8) * 8;
    hipLaunchKernelGGL((k_tod_fp_batch<MODE, QU, CS, MINW>), dim3((unsigned)blocks),
                       dim3(kFpThreads), smem, st, ba, fc);
    return hipGetLastError();
}

template <int MODE, bool QU, int CS>
hipError_t launch_batch_v(const BatchArgs& ba, int variant, hipStream_t st) {
    if (variant == 8) return launch_batch_cs<MODE, QU, CS, 4>(ba, st);
    return launch_batch_cs<MODE, QU, CS, 2>(ba, st);
}

template <int MODE>
hipError_t launch_batch_mode(const BatchArgs& ba, bool qu, int CS, int variant, hipStream_t st) {
    if (qu) {
        if (CS == 2) return launch_batch_v<MODE, true, 2>(ba, variant, st);
        return launch_batch_v<MODE, true, 4>(ba, variant, st);
    }
    if (CS == 1) return launch_batch_v<MODE, false, 1>(ba, variant, st);
    if (CS == 2) return launch_batch_v<MODE, false, 2>(ba, variant, st);
    return launch_batch_v<MODE, false, 4>(ba, variant, st);
}

template <int MODE, bool QU, int CS>
hipError_t launch_cs(const TodArgs& a, int variant, hipStream_t st) {
    if (variant == 1 || variant == 2 || variant == 4)
        return launch_direct_cs<MODE, QU, CS>(a, variant, st);
    if (variant == 8) return launch_fp_cs<MODE, QU, CS, 4>(a, st);
    return launch_fp_cs<MODE, QU, CS, 2>(a, st);
}

template <int MODE>
hipError_t launch_mode(const TodArgs& a, bool qu, int CS, int variant, hipStream_t st) {
    if (qu) {
        if (CS == 2) return launch_cs<MODE, true, 2>(a, variant, st);
        return launch_cs<MODE, true, 4>(a, variant, st);
    }
    if (CS == 1) return launch_cs<MODE, false, 1>(a, variant, st);
    if (CS == 2) return launch_cs<MODE, false, 2>(a, variant, st);
    return launch_cs<MODE, false, 4>(a, variant, st);
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
    if (B > 0 && S > 0 &&
        (!vec_orig || !axes || !cos_a || !sin_a || !ax_pts || !cos_p || !sin_p ||
         (need_map && (!map_packed || !beam_vals || !tod))))
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
            int64_t S, double* tod, int c_q, int c_u, int variant, void* stream) {
    TodArgs a;
    bool qu;
    int CS;
    int rc = fill_args(vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                       beam_vals, B, S, tod, c_q, c_u, true, &a, &qu, &CS);
    if (rc) return rc;
    if (B == 0 || S == 0) return TODG_OK;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = mode == 0 ? launch_mode<0>(a, qu, CS, variant, st)
                             : launch_mode<1>(a, qu, CS, variant, st);
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
                            int c_u, int variant, void* stream) {
    return run_tod(0, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                   beam_vals, B, S, tod, c_q, c_u, variant, stream);
}

int todg_gather_accum_nearest(const float* vec_orig, const float* axes, const float* cos_a,
                              const float* sin_a, const float* ax_pts, const float* cos_p,
                              const float* sin_p, int64_t nside, const float* map_packed, int C,
                              const float* beam_vals, int64_t B, int64_t S, double* tod,
                              int c_q, int c_u, int variant, void* stream) {
    return run_tod(1, vec_orig, axes, cos_a, sin_a, ax_pts, cos_p, sin_p, nside, map_packed, C,
                   beam_vals, B, S, tod, c_q, c_u, variant, stream);
}

int todg_gather_accum_batch(int mode, const todg_work_item* items, int n_items,
                            int64_t total_samples, int64_t nside, const float* map_packed, int C,
                            int c_q, int c_u, int chunk, unsigned int* queue, int* item_status,
                            int variant, void* stream) {
    if (mode != 0 && mode != 1) return fail(TODG_ERR_ARG, "mode must be 0 (bilinear) or 1 (nearest)");
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
    hipError_t e = mode == 0 ? launch_batch_mode<0>(ba, qu, CS, variant, st)
                             : launch_batch_mode<1>(ba, qu, CS, variant, st);
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
