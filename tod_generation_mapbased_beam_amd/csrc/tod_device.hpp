// Device code of the beam-convolution TOD kernels for gfx950: kernel
// arguments, the shared index path, the direct, footprint and batched
// kernels, and the launch templates.  Included by the library's translation
// units: tod_kernels.hip (host side and C ABI) and one unit per kernel mode
// (tod_fp_bilinear.hip, tod_fp_nearest.hip, tod_fp_bicubic.hip,
// tod_fp_learned.hip), which hipcc compiles in parallel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include <mutex>
#include <type_traits>

#include "healpix_ring.hpp"
#include "internal.hpp"
#include "corrector.hpp"
#include "../../include/todgpu.h"

namespace todg_fp {

using namespace todg;

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
    const double* lam_trig;  // (kLamN, 2): cos, sin of lam_j
    double* tod;             // (C, B), accumulated
    int64_t B;
    int64_t S;
    int64_t npix;
    int nside;
    int nslots;
    int row[kMaxSlots];      // tod row of each packed slot
    // learned corrector (footprint mode 3 only)
    const float* corr;                 // FP32 parameters (corrector.hpp layout)
    unsigned long long* corr_stats;    // [0] corrected, [1] plain bilinear beam-pixel samples, or null
    float corr_bound;                  // largest |correction| applied, in neighbourhood RMS units
    int corr_slot;                     // packed slot whose values drive the corrector
    int64_t corr_count;                // beam pixels [0, corr_count) take the corrector, the rest bilinear
    int* status;                       // bit 0 set when a sample had a non-finite operand, or null
    int64_t* dump;                     // index-dump builds only: see FpDump
    // factorised corrector (footprint mode 4 only): the correction channel,
    // packed like map; each beam pixel's bilinear weights times its ring
    // weight product (1 - g) g go with it (FcorrWeights)
    const float* map1;
};

// Index dump of the footprint kernel (its DUMP builds): for every (b, s) of
// the launch, i = b S + s, the pixels the kernel actually deposited into --
// pix[4 i + k] (bilinear neighbours k = 0..3 in the reference's order; nearest:
// k = 0, the rest -1) -- the deposited value w[4 i + k] (beam value x weight,
// as rounded onto the weight grid) and path[i] (0 = certified fast path,
// 1 = exact path).  One allocation: pix, then w, then path.
struct FpDump {
    int64_t* pix;
    double* w;
    unsigned char* path;
};

__device__ __forceinline__ FpDump fp_dump_of(int64_t* base, int64_t B, int64_t S) {
    const int64_t n = B * S;
    return FpDump{base, reinterpret_cast<double*>(base + 4 * n),
                  reinterpret_cast<unsigned char*>(base + 8 * n)};
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

// Weight-grid deposits are rounded to multiples of a quantum q = 2^(E - 52),
// 2^E >= 64 sum_s |beam_vals[s]|, by (x + M) - M with M = 1.5 2^E (x + M lies
// in [2^E, 2^(E+1)), whose spacing is q).  Every partial sum of a cell is then
// a multiple of q below 2^53 q in magnitude: the grid keeps it as an exact
// integer count of quanta (WCell below), so the cell's sum does not depend on
// the order of the atomic adds.  q is 2^-46 of the beam's total or finer: far
// below the binary32 weights' precision.
__device__ __forceinline__ double grid_magic(double sum_abs_beam) {
    if (!(sum_abs_beam > 0.0) || !(sum_abs_beam < 1e300)) return 0.0;  // no rounding
    int e = 0;
    (void)frexp(64.0 * sum_abs_beam, &e);  // 64 sum = f 2^e, f in [0.5, 1)
    return 1.5 * ldexp(1.0, e);
}

__device__ __forceinline__ double qround(double x, double M) { return (x + M) - M; }

// A weight-grid cell holds its sum as an integer number of quanta q (two's
// complement): x rounded to the grid is bits(x + M) - bits(M) quanta, and a
// cell of n quanta is worth bits^-1(bits(M) + n) - M = n q exactly, since
// |n q| stays below 2^E / 64 (at most the beam's sum of |values|).  Integer
// sums are exact and order-free like the binary64 sums of grid multiples,
// with the same values; LDS adds them with ds_add_u64, far cheaper than
// ds_add_f64.  A beam with M = 0 (no grid) gets no layout (fp_layout).
typedef unsigned long long WCell;
__device__ __forceinline__ WCell wbits(double x) { return __builtin_bit_cast(WCell, x); }
__device__ __forceinline__ WCell wq(double x, double M) { return wbits(x + M) - wbits(M); }
// -bits(M) for a uniform M, opaque to the compiler: x + neg_bits(M) then stays
// an add (one v_lshl_add_u64 with a scalar operand) instead of being folded
// back into a 64-bit subtraction (a subtract-with-borrow pair and a move).
// The empty asm emits nothing.
__device__ __forceinline__ WCell neg_bits(double M) {
    WCell n = 0ull - wbits(M);
    asm("" : "+s"(n));
    return n;
}
__device__ __forceinline__ double wval(WCell n, double M) { return __builtin_bit_cast(double, wbits(M) + n) - M; }

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
        bool live[T];  // in range with finite operands (others are skipped)
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int64_t b = (b0 + t < a.B) ? b0 + t : a.B - 1;
            live[t] = b0 + t < a.B && finite_sample(a, b);
            if (b0 + t < a.B && !live[t] && lane == 0 && a.status) atomicOr(a.status, 1);
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
                if (!live[t]) continue;  // uniform over the wave
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
                    if (lane == 0 && live[t]) a.tod[(int64_t)a.row[k] * a.B + b] += sum;
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

// ========================================================= footprint kernel

constexpr int kFpThreads = 512;
constexpr int kFpWaves = kFpThreads / 64;

// Certified fast index path (fp_fast_bilinear).  A ring or slot coordinate
// computed there differs from the reference's binary64 value by far less
// than 1e-10; its integer part is taken only when the coordinate is at least
// kCertSlot (kCertPolar for the polar-cap ring formula, whose slope grows
// near the pole) away from an integer, otherwise the beam pixel is
// recomputed by the exact path.
constexpr double kCertSlot = 1e-9;
constexpr double kCertPolar = 1e-6;
// Largest |tan| of the longitude offset from the reference longitude that
// the fast path's atan series takes (five terms: truncation t^11/11 < 2e-18
// at 0.03; the offsets of beams up to ~1 deg across stay below it).
constexpr double kMaxTan = 0.03;

// atan(t) for |t| < kMaxTan.
__device__ __forceinline__ double atan_small(double t) {
    const double t2 = t * t;
    double p = fma(t2, 1.0 / 9.0, -1.0 / 7.0);
    p = fma(t2, p, 1.0 / 5.0);
    p = fma(t2, p, -1.0 / 3.0);
    return fma(t * t2, p, t);
}

// Per-sample values shared by the workgroup (written by wave 0).
struct alignas(16) FpFrame {
    double m[9];      // m[3*i + j]: row i, column j of Rz(-lam) * R2 * R1
    double q[12];     // exact-path rotation: kx ky kz ca sa oma bx by bz cp sp omp
    double lam;       // reference longitude
    double z_pts, sth_pts;  // boresight z and sin(theta)  (tod_bilinear.py:306-310)
    double sha, cha;  // sin, cos of half the boresight longitude minus lam
    int ir_lo, nr, total;
    int coinc_ring, coinc_pos;  // the map pixel whose centre IS the boresight, or -1
    int coinc_cell;             // its weight-grid cell, or -1
    int direct;                 // 1: no layout, or rings beyond the grid's capacity
                                // (neighbours outside accumulate directly, fixed claims)
};

// Map and geometry operands of the exact path (passed by value to the
// out-of-line functions).
struct FpGlobals {
    const float* map;
    const RingEntry* rings;
    int64_t npix;
    int nside;
    int nslots;
    const float* map1;  // mode 4's correction channel
};

__device__ __forceinline__ FpGlobals fp_globals(const TodArgs& a) {
    return FpGlobals{a.map, a.rings, a.npix, a.nside, a.nslots, a.map1};
}

// Spin-2 inputs of one sample (uniform).
struct FpSpin {
    double z_pts, sth_pts, sha, cha, lam;
    int coinc_ring, coinc_pos;
};

// LDS layout of the footprint kernels (dynamic shared memory, compile-time
// offsets, so every access is one address register plus an immediate).
// Ring descriptors and frames are double-buffered so that wave 0 can lay
// out sample b+1 while the other waves still read sample b's.  Ring r of
// the footprint is global ring ir_lo + r; its window covers slots
// i0 .. i0 + width - 1 (mod n) and grid cells off .. off + width - 1.
extern __shared__ __align__(16) unsigned char fp_smem[];

constexpr int kFpCap = 4608;       // weight-grid cells
constexpr int kFpMaxRings = 128;   // ring descriptors per buffer
constexpr int kRingD = 14;        // 8-byte arrays per ring buffer
constexpr int kRingI = 5;         // int arrays per ring buffer
constexpr int kRingBuf = kFpMaxRings * (kRingD * 8 + kRingI * 4);
constexpr int kOffRings = 0;
constexpr int kOffFrame = kOffRings + 2 * kRingBuf;
constexpr int kOffRed = kOffFrame + 2 * (int)sizeof(FpFrame);
constexpr int kOffScratch = kOffRed + kFpWaves * kMaxSlots * 8;
// scratch: two per-wave reduction rows (fp_beam_frame), then the batched
// kernel's claim word
constexpr int kOffJob = kOffScratch + (2 * kFpWaves + 2) * 8;
// the batched kernel's current item (TodArgs) and beam frame, kept in LDS
// rather than in registers across a step
constexpr int kOffItem = kOffJob + 2 * 64;
constexpr int kItemBytes = 256;
constexpr int kOffW = kOffItem + kItemBytes;
// weight-grid cell type (beam value x interpolation weight sums)
// ring of each weight-grid cell (footprint-relative, < kFpMaxRings <= 256),
// one table per buffer
constexpr int kOffRtab = kOffW + kFpCap * (int)sizeof(WCell);
// Side list of one sample: neighbour pixels outside the laid-out footprint
// with their (grid-rounded) weights, kept so that they can be summed in pixel
// order (see fp_side_flush): a count, kSideCap keys (ring << 16 | slot) and
// kSideCap weights.
constexpr int kSideCap = 32;
constexpr int kOffSide = kOffRtab + 2 * kFpCap;
constexpr int kOffSideKey = kOffSide + 8;
constexpr int kOffSideW = kOffSideKey + 4 * kSideCap;
constexpr int kFpSmemBytes = ((kOffSideW + 8 * kSideCap) + 15) & ~15;
static_assert(kOffSide % 8 == 0 && kOffSideW % 8 == 0, "side list alignment");
// mode 3 (learned corrector) adds the driving field's values of the
// footprint cells, one 64 x 64-byte staging tile per wave and the biases
constexpr int kOffIv = kFpSmemBytes;
constexpr int kOffStage = kOffIv + kFpCap * 4;
constexpr int kStageBytes = 64 * 64;
constexpr int kOffBias = kOffStage + kFpWaves * kStageBytes;  // b2 (kH), b3 (kOut) as float
constexpr int kFpSmemBytesCorr = kOffBias + (todg_corr::kH + todg_corr::kOut) * 4;
static_assert(kOffStage % 16 == 0 && kOffBias % 16 == 0, "16-byte LDS vectors");
static_assert(kFpSmemBytesCorr <= 160 * 1024, "one workgroup per CU");
static_assert(kFpSmemBytes <= 80 * 1024, "two workgroups per CU");
// mode 4 (factorised corrector) splits the weight grid: cells [0,
// kFpCapFc) hold the main channel's weights, cells [kFpCapFc, kFpCap) the
// correction channel's, so its footprints are laid out up to kFpCapFc cells
// (a 30' beam at nside 2048 needs ~2000); it adds the correction weights of
// the side list
constexpr int kFpCapFc = kFpCap / 2;
constexpr int kOffSideW1 = kFpSmemBytes;
constexpr int kFpSmemBytesFc = kOffSideW1 + 8 * kSideCap;
static_assert(kFpSmemBytesFc <= 80 * 1024, "two workgroups per CU");

template <typename T>
__device__ __forceinline__ T& lds_at(unsigned off) {
    return *reinterpret_cast<T*>(fp_smem + off);
}

// One buffer of footprint ring descriptors, structure of arrays.
struct FpRings {
    unsigned o;  // byte offset of the buffer
    __device__ double& d(int k, int r) const { return lds_at<double>(o + 8 * (k * kFpMaxRings + r)); }
    __device__ int& i(int k, int r) const { return lds_at<int>(o + 8 * kRingD * kFpMaxRings + 4 * (k * kFpMaxRings + r)); }
    __device__ double& base(int r) const { return d(0, r); }  // slot coordinate of lam, window-relative
    __device__ double& inv_dphi(int r) const { return d(1, r); }
    __device__ double& z(int r) const { return d(2, r); }
    __device__ double& sth(int r) const { return d(3, r); }
    __device__ double& inv_dth(int r) const { return d(4, r); }
    __device__ double& phi0(int r) const { return d(5, r); }
    __device__ double& dphi(int r) const { return d(6, r); }
    // spin-2 constants of the ring against the sample's boresight
    // (Spin2Ring), rounded to binary32 for the footprint sweep; each sits in
    // the first half of an 8-byte slot
    __device__ float& f(int k, int r) const { return lds_at<float>(o + 8 * (k * kFpMaxRings + r)); }
    __device__ float& st2(int r) const { return f(7, r); }
    __device__ float& sS(int r) const { return f(8, r); }
    __device__ float& s2dt(int r) const { return f(9, r); }
    __device__ float& zsum(int r) const { return f(10, r); }
    __device__ float& sc4(int r) const { return f(11, r); }
    // more binary32 values in the second halves of those slots:
    // 1 / (theta_below - theta_above), and the sweep's spin-2 longitudes --
    // half the longitude gap between window slot k and the boresight is
    // (k - kb) hdp + gb, kb the window slot nearest the boresight
    __device__ float& g(int k, int r) const { return lds_at<float>(o + 8 * (k * kFpMaxRings + r) + 4); }
    __device__ float& inv_dth_f(int r) const { return g(7, r); }
    __device__ float& gb(int r) const { return g(8, r); }
    __device__ float& hdp(int r) const { return g(9, r); }
    __device__ float& kb(int r) const { return g(10, r); }
    // base + off: lam's weight-grid cell coordinate (binary64; off < 2^13, so
    // the sum rounds by < 5e-13)
    __device__ double& baseo(int r) const { return d(12, r); }
    // {off, max(width - 1, 0)}: a bilinear pair (c, c + 1) lies in the
    // window iff (unsigned)(c - off) < max(width - 1, 0)
    __device__ int2& ow(int r) const { return lds_at<int2>(o + 8 * (13 * kFpMaxRings + r)); }
    __device__ int& width(int r) const { return i(0, r); }
    __device__ int& off(int r) const { return i(1, r); }    // non-decreasing in r
    __device__ int& i0(int r) const { return i(2, r); }
    __device__ int& n(int r) const { return i(3, r); }
    __device__ int& first(int r) const { return i(4, r); }  // npix < 2^31
};

__device__ __forceinline__ FpRings fp_rings(int buf) { return FpRings{(unsigned)(kOffRings + buf * kRingBuf)}; }
__device__ __forceinline__ FpFrame* fp_frame(int buf) {
    return &lds_at<FpFrame>(kOffFrame + buf * (int)sizeof(FpFrame));
}
__device__ __forceinline__ unsigned char* fp_rtab(int buf) { return fp_smem + kOffRtab + buf * kFpCap; }
__device__ __forceinline__ double* fp_red() { return &lds_at<double>(kOffRed); }
__device__ __forceinline__ double* fp_scratch() { return &lds_at<double>(kOffScratch); }
__device__ __forceinline__ WCell* fp_W() { return &lds_at<WCell>(kOffW); }
__device__ __forceinline__ WCell* fp_W1() { return &lds_at<WCell>(kOffW + kFpCapFc * (int)sizeof(WCell)); }

// The per-sample operands a stage needs after the sample's layout: the
// beam, the tod destination, the sample index and phase 1's claim counter
// (one slot per buffer).
struct alignas(16) FpJob {
    const float* vec_orig;
    const float* beam_vals;
    double* tod;
    int64_t S, B, b;
    int claim;  // next chunk of beam pixels (LDS atomic counter)
    double M;   // weight-grid rounding constant of the beam (grid_magic)
};
static_assert(sizeof(FpJob) <= 64, "job slot");
__device__ __forceinline__ FpJob* fp_job(int buf) { return &lds_at<FpJob>(kOffJob + 64 * buf); }

template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    return reinterpret_cast<T*>(__builtin_bit_cast(uint64_t, uniform(__builtin_bit_cast(double, reinterpret_cast<uint64_t>(p)))));
}

// A pointer read back from LDS carries no address space, so loads through
// it compile to FLAT instructions, which also count against the LDS wait
// counter: every LDS wait would then wait for those global loads too.
// These casts restore the global address space.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T>
__device__ __forceinline__ gptr<const T> global_in(const T* p) {
    return (gptr<const T>)uniform_ptr(p);
}

// Local index of (footprint ring r, slot pos in [0, n]) in the weight grid, or -1.
__device__ __forceinline__ int fp_local(const FpRings& R, int r, int pos) {
    const int n = R.n(r);
    int d = pos - R.i0(r);
    if (d < 0) d += n;
    if (d >= n) d -= n;
    return d < R.width(r) ? R.off(r) + d : -1;
}

// Exact boresight longitude (tod_bilinear.py:306-310), for the rare exact
// coincidence test of the spin-2 guard.
static __device__ __noinline__ double boresight_phi(double bx, double by) {
    double ph = atan2(by, bx);
    if (ph < 0.0) ph += kTwoPi;
    return ph;
}

// atan2(y, x), out of line (keeps the library's registers out of the caller).
static __device__ __noinline__ double atan2_wide(double y, double x) { return atan2(y, x); }

// cos 2delta, sin 2delta of map pixel (ring, pos) with centre (z_n, sth_n,
// phi_n) against the boresight (tod_bilinear.py:84-145).  The longitude gap
// enters through half angles: phi_pts - phi_n = (phi_pts - lam) - (phi_n -
// lam), with the boresight half angle precomputed per sample, so no atan2 is
// needed.  The reference's exact-coincidence guard (tod_bilinear.py:108-109)
// compares with the one pixel, found per sample, whose centre equals the
// boresight bit for bit.
__device__ __forceinline__ void fp_spin2(const FpSpin& sp, int ring, int pos, double z_n,
                                         double sth_n, double phi_n, double& c2d, double& s2d) {
    if (ring == sp.coinc_ring && pos == sp.coinc_pos) {
        c2d = 1.0;
        s2d = 0.0;
        return;
    }
    const double hb = (phi_n - sp.lam) * 0.5;
    double sb, cb;
    sincos_any(hb, sb, cb);
    const double sh = fma(sp.sha, cb, -(sp.cha * sb));
    const double ch = fma(sp.cha, cb, sp.sha * sb);
    spin2_from_half(spin2_ring(z_n, sth_n, sp.z_pts, sp.sth_pts), sh, ch, c2d, s2d);
}

// Direct contribution of one map pixel with weight w (spin-2 on slots 0/1
// when QU), for neighbours outside the laid-out footprint.
template <bool QU, int CS>
__device__ __forceinline__ Acc4 fp_direct(const FpGlobals& a, const FpSpin& sp, int ring, int pos,
                                          double w) {
    Acc4 r = {{0.0, 0.0, 0.0, 0.0}};
    const RingEntry& e = a.rings[ring];
    const float4 m = load_px<CS>(a.map, e.first + pos, a.npix);
    if constexpr (QU) {
        double c2d, s2d;
        fp_spin2(sp, ring, pos, e.z, e.sth, e.phi0 + (double)pos * e.dphi, c2d, s2d);
        const double qv = (double)m.x, uv = (double)m.y;
        r.v[0] = w * fma(qv, c2d, uv * s2d);
        r.v[1] = w * fma(uv, c2d, -(qv * s2d));
#pragma unroll
        for (int k = 2; k < CS; ++k)
            if (k < a.nslots) r.v[k] = w * (double)slot(m, k);
    } else {
#pragma unroll
        for (int k = 0; k < CS; ++k)
            if (k < a.nslots) r.v[k] = w * (double)slot(m, k);
    }
    return r;
}

// Mode 4's direct contribution of one map pixel: the main channel with
// weight w plus the correction channel with weight w1, then spin-2 on slots
// 0/1 when QU.
template <bool QU, int CS>
__device__ __forceinline__ Acc4 fp_direct_fc(const FpGlobals& a, const FpSpin& sp, int ring, int pos,
                                             double w, double w1) {
    Acc4 r = {{0.0, 0.0, 0.0, 0.0}};
    const RingEntry& e = a.rings[ring];
    const float4 m = load_px<CS>(a.map, e.first + pos, a.npix);
    const float4 m1 = load_px<CS>(a.map1, e.first + pos, a.npix);
    double v[kMaxSlots] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < CS; ++k) v[k] = fma(w1, (double)slot(m1, k), w * (double)slot(m, k));
    if constexpr (QU) {
        double c2d, s2d;
        fp_spin2(sp, ring, pos, e.z, e.sth, e.phi0 + (double)pos * e.dphi, c2d, s2d);
        r.v[0] = fma(v[0], c2d, v[1] * s2d);
        r.v[1] = fma(v[1], c2d, -(v[0] * s2d));
#pragma unroll
        for (int k = 2; k < CS; ++k)
            if (k < a.nslots) r.v[k] = v[k];
    } else {
#pragma unroll
        for (int k = 0; k < CS; ++k)
            if (k < a.nslots) r.v[k] = v[k];
    }
    return r;
}

template <bool QU, int CS, bool FC>
__device__ __forceinline__ Acc4 fp_direct_any(const FpGlobals& a, const FpSpin& sp, int ring, int pos,
                                              double w, double w1) {
    if constexpr (FC)
        return fp_direct_fc<QU, CS>(a, sp, ring, pos, w, w1);
    else
        return fp_direct<QU, CS>(a, sp, ring, pos, w);
}

__device__ __forceinline__ int* fp_side_count() { return &lds_at<int>(kOffSide); }
__device__ __forceinline__ unsigned* fp_side_key() { return &lds_at<unsigned>(kOffSideKey); }
__device__ __forceinline__ double* fp_side_w() { return &lds_at<double>(kOffSideW); }
__device__ __forceinline__ double* fp_side_w1() { return &lds_at<double>(kOffSideW1); }

// A neighbour pixel (ring, slot pos <= n <= 2^15) outside the laid-out
// footprint with grid-rounded weight w (and, in mode 4, correction weight
// w1): appended to the sample's side list, which fp_side_flush sums in pixel
// order.  When the list is full the pixel is accumulated on the spot
// (returned), and that sample's last bits then depend on which thread took
// it.
template <bool QU, int CS, bool FC = false>
__device__ __forceinline__ Acc4 fp_outside(const FpGlobals& a, const FpSpin& sp, int ring, int pos,
                                           double w, double w1 = 0.0) {
    const int i = atomicAdd(fp_side_count(), 1);
    if (i < kSideCap) {
        fp_side_key()[i] = ((unsigned)ring << 16) | (unsigned)pos;
        fp_side_w()[i] = w;
        if constexpr (FC) fp_side_w1()[i] = w1;
        return Acc4{{0.0, 0.0, 0.0, 0.0}};
    }
    return fp_direct_any<QU, CS, FC>(a, sp, ring, pos, w, w1);
}

__device__ __forceinline__ FpSpin fp_spin_of(const FpFrame* fr);

// One thread, after phase 2: the side list's pixels in key order (equal keys
// merged; their grid-rounded weights add exactly), each with its spin-2 and
// one gather, added to acc in that order; the list is emptied.
template <bool QU, int CS, bool FC = false>
__device__ __noinline__ Acc4 fp_side_flush(FpGlobals a, const FpFrame* fr, Acc4 acc) {
    int* cnt = fp_side_count();
    unsigned* key = fp_side_key();
    double* w = fp_side_w();
    double* w1 = fp_side_w1();
    const int n = min(*cnt, kSideCap);
    for (int i = 1; i < n; ++i) {  // insertion sort by key
        const unsigned k = key[i];
        const double x = w[i];
        const double x1 = FC ? w1[i] : 0.0;
        int j = i - 1;
        while (j >= 0 && key[j] > k) {
            key[j + 1] = key[j];
            w[j + 1] = w[j];
            if constexpr (FC) w1[j + 1] = w1[j];
            --j;
        }
        key[j + 1] = k;
        w[j + 1] = x;
        if constexpr (FC) w1[j + 1] = x1;
    }
    const FpSpin sp = fp_spin_of(fr);
    for (int i = 0; i < n;) {
        const unsigned k = key[i];
        double ws = 0.0, ws1 = 0.0;
        for (; i < n && key[i] == k; ++i) {
            ws += w[i];
            if constexpr (FC) ws1 += w1[i];
        }
        const Acc4 d = fp_direct_any<QU, CS, FC>(a, sp, (int)(k >> 16), (int)(k & 0xffffu), ws, ws1);
        for (int j = 0; j < kMaxSlots; ++j) acc.v[j] += d.v[j];
    }
    *cnt = 0;
    return acc;
}

__device__ __forceinline__ FpSpin fp_spin_of(const FpFrame* fr) {
    FpSpin sp;
    sp.z_pts = uniform(fr->z_pts);
    sp.sth_pts = uniform(fr->sth_pts);
    sp.sha = uniform(fr->sha);
    sp.cha = uniform(fr->cha);
    sp.lam = uniform(fr->lam);
    sp.coinc_ring = __builtin_amdgcn_readfirstlane(fr->coinc_ring);
    sp.coinc_pos = __builtin_amdgcn_readfirstlane(fr->coinc_pos);
    return sp;
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
    make_uniform(q);
}

// Exact path for one beam pixel u (beam value bv): the reference rotation,
// atan2 and neighbour lookup; neighbours inside the layout go to their LDS
// cell, the others are accumulated directly.  Out of line: it runs only for
// the rare beam pixels the fast path does not certify.
template <int MODE, bool QU, int CS, bool DUMP = false>
__device__ __noinline__ Acc4 fp_exact_pixel(FpGlobals a, int buf, float u0, float u1, float u2,
                                            float bvf, FpDump dd = FpDump{}, int64_t di = 0) {
    const FpRings R = fp_rings(buf);
    const FpFrame* fr = fp_frame(buf);
    const double M = fp_job(buf)->M;
    WCell* W = fp_W();
    Sample q;
    frame_sample(fr, q);
    const FpSpin sp = fp_spin_of(fr);
    const int ir_lo = fr->ir_lo, nr = fr->nr;
    const double bv = (double)bvf;
    double vx, vy, vz;
    rotate(q, (double)u0, (double)u1, (double)u2, vx, vy, vz);
    const double phi_w = wrap_phi(vy, vx);
    Acc4 acc = {{0.0, 0.0, 0.0, 0.0}};
    // one grid-rounded deposit: into its cell inside the layout, else to the
    // side list (a sample without a complete layout accumulates directly, in
    // its threads' fixed claim order)
    auto deposit = [&](int ring, int pos, double w) {
        const int lr = ring - ir_lo;
        const int li = (lr >= 0 && lr < nr) ? fp_local(R, lr, pos) : -1;
        if (li >= 0) {
            TODG_CHECK(li < fr->total, kGridCell);
            atomicAdd(&W[li], wq(w, M));
        } else {
            const Acc4 d = fr->direct ? fp_direct<QU, CS>(a, sp, ring, pos, w)
                                      : fp_outside<QU, CS>(a, sp, ring, pos, w);
            for (int j = 0; j < kMaxSlots; ++j) acc.v[j] += d.v[j];
        }
    };
    if constexpr (MODE == 4) {
        // bilinear deposits, each also into the correction grid times the
        // ring weight product (1 - g) g, g = the lower ring's weight
        Nb o;
        ring_interp(a.nside, a.rings, fmax(-1.0, fmin(1.0, vz)), phi_w, o);
        const double uc = (o.w[0] + o.w[1]) * (o.w[2] + o.w[3]);
        WCell* W1 = fp_W1();
        for (int k = 0; k < 4; ++k) {
            const double w = qround(o.w[k] * bv, M);
            const double w1 = qround(w * uc, M);
            const int lr = o.ring[k] - ir_lo;
            const int li = (lr >= 0 && lr < nr) ? fp_local(R, lr, o.pos[k]) : -1;
            if (li >= 0) {
                TODG_CHECK(li < fr->total && fr->total <= kFpCapFc, kGridCell);
                atomicAdd(&W[li], wq(w, M));
                atomicAdd(&W1[li], wq(w1, M));
            } else {
                const Acc4 d = fr->direct ? fp_direct_fc<QU, CS>(a, sp, o.ring[k], o.pos[k], w, w1)
                                          : fp_outside<QU, CS, true>(a, sp, o.ring[k], o.pos[k], w, w1);
                for (int j = 0; j < kMaxSlots; ++j) acc.v[j] += d.v[j];
            }
        }
    } else if constexpr (MODE == 0) {
        Nb o;
        ring_interp(a.nside, a.rings, fmax(-1.0, fmin(1.0, vz)), phi_w, o);
        for (int k = 0; k < 4; ++k) {
            const double w = qround(o.w[k] * bv, M);
            deposit(o.ring[k], o.pos[k], w);
            if constexpr (DUMP) {
                dd.pix[4 * di + k] = a.rings[o.ring[k]].first + o.pos[k];
                dd.w[4 * di + k] = w;
            }
        }
        if constexpr (DUMP) dd.path[di] = 1;
    } else if constexpr (MODE == 1) {
        const double sin_th = sqrt(fmax(0.0, vx * vx + vy * vy));
        int ring, pos;
        nearest_pix(a.nside, a.rings, vz, sin_th, phi_w, ring, pos);
        const double w = qround(bv, M);
        deposit(ring, pos, w);
        if constexpr (DUMP) {
            dd.pix[4 * di] = a.rings[ring].first + pos;
            dd.w[4 * di] = w;
            dd.path[di] = 1;
        }
    } else {
        // bicubic judge: weights summed first, then scattered normalised
        const double z = fmax(-1.0, fmin(1.0, vz));
        const double sum = bicubic_visit(a.nside, a.rings, z, phi_w, [](int, int, double) {});
        if (sum != 0.0) {
            const double scale = bv / sum;
            bicubic_visit(a.nside, a.rings, z, phi_w, [&](int ring, int pos, double w) {
                deposit(ring, pos, qround(w * scale, M));
            });
        } else {  // no candidate inside the support: bilinear (as the oracle)
            Nb o;
            ring_interp(a.nside, a.rings, z, phi_w, o);
            for (int k = 0; k < 4; ++k) deposit(o.ring[k], o.pos[k], qround(o.w[k] * bv, M));
        }
    }
    return acc;
}

// Beam centre c (unit vector in the beam frame) and the footprint disc
// radius R around its image R_b c.
struct BeamFrame {
    double cx, cy, cz, R, cosR;
    double M;  // weight-grid rounding constant (grid_magic of sum |beam_vals|)
};

// All threads of the block: the beam centre c = R1^-1(boresight) of sample
// b0 and the beam radius about it.  For a beam prepared around its centre
// every sample's rotation maps c onto its boresight; any c yields a correct
// footprint disc around R_b c (rotations keep angles), a good c a small one.
struct BeamSrc {
    const float *vec_orig, *beam_vals, *axes, *cos_a, *sin_a, *ax_pts, *cos_p, *sin_p;
    int64_t S;
    int nside;
};

__device__ __forceinline__ BeamSrc beam_src(const TodArgs& a) {
    return BeamSrc{a.vec_orig, a.beam_vals, a.axes, a.cos_a, a.sin_a, a.ax_pts, a.cos_p, a.sin_p,
                   a.S, a.nside};
}

static __device__ __noinline__ BeamFrame fp_beam_frame(BeamSrc a, int64_t b0, double* red,
                                                double margin_px) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    double cx, cy, cz, cosr, sabs;
    {
        Sample q;
        q.kx = (double)a.axes[3 * b0];
        q.ky = (double)a.axes[3 * b0 + 1];
        q.kz = (double)a.axes[3 * b0 + 2];
        q.ca = (double)a.cos_a[b0];
        q.sa = (double)a.sin_a[b0];
        q.oma = 1.0 - q.ca;
        q.bx = (double)a.ax_pts[3 * b0];
        q.by = (double)a.ax_pts[3 * b0 + 1];
        q.bz = (double)a.ax_pts[3 * b0 + 2];
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
        if (!(isfinite(cx) && isfinite(cy) && isfinite(cz))) {  // no finite sample to centre on
            cx = 0.0;
            cy = 0.0;
            cz = 1.0;
        }
        double mn = 2.0, sa = 0.0;
        for (int64_t s = tid; s < a.S; s += kFpThreads) {
            const double d = (double)a.vec_orig[3 * s] * cx + (double)a.vec_orig[3 * s + 1] * cy +
                             (double)a.vec_orig[3 * s + 2] * cz;
            mn = fmin(mn, d);
            sa += fabs((double)a.beam_vals[s]);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, off, 64));
            sa += __shfl_xor(sa, off, 64);
        }
        if (lane == 0) {
            red[wave] = mn;
            red[kFpWaves + wave] = sa;
        }
        __syncthreads();
        mn = red[0];
        sa = red[kFpWaves];
        for (int w = 1; w < kFpWaves; ++w) {
            mn = fmin(mn, red[w]);
            sa += red[kFpWaves + w];
        }
        __syncthreads();
        cosr = fmax(-1.0, fmin(1.0, mn));
        sabs = sa;
    }
    // Disc radius for the ring range and the per-ring longitude half-width:
    // beam radius + margin_px pixel sizes + a rounding allowance.  The
    // margin covers the neighbour rings (one ring spacing, at most 0.9 pixel
    // sizes, from the query for bilinear and nearest; three spacings and the
    // stencil's slots for bicubic); the slot windows add the longitude
    // neighbours.
    BeamFrame f;
    f.cx = cx;
    f.cy = cy;
    f.cz = cz;
    f.R = fmin(kPi, acos(cosr) + margin_px / (double)a.nside + 1e-6);
    f.cosR = cos(f.R);
    f.M = grid_magic(sabs);
    return f;
}

// Wave 0 only: sample b's frame and footprint layout -- every ring and ring
// slot that the rotated beam's bilinear neighbours can touch, as per-ring
// slot windows packed into the weight grid.  The windows only steer work
// (a neighbour outside them takes the exact path), so they are estimated in
// binary32: on each ring the disc's chord [L, R] in slot coordinates, as
// slots floor(L) - 1 .. at least floor(R) + 2 (the right neighbour of the
// last slot and one more); the reference longitude lam and everything the
// index path derives from it are binary64.
// Ring table entries [k0, k1) = r: single bytes up to a 4-byte boundary and
// after the last one, whole words between (the table base is 4-byte
// aligned; a neighbouring ring's bytes in a shared word are written by
// their own byte stores).
__device__ __forceinline__ void fill_ring_cells(unsigned char* rtab, int k0, int k1, int r) {
    const unsigned rep4 = (unsigned)r * 0x01010101u;
    int k = k0;
    for (; k < k1 && (k & 3) != 0; ++k) rtab[k] = (unsigned char)r;
    for (; k + 4 <= k1; k += 4) *reinterpret_cast<unsigned*>(rtab + k) = rep4;
    for (; k < k1; ++k) rtab[k] = (unsigned char)r;
}

__device__ __forceinline__ bool fp_layout(const TodArgs& a, int64_t b, const BeamFrame& bf,
                                          FpFrame* fr, FpRings R, unsigned char* rtab, int cap) {
    const int lane = threadIdx.x & 63;
    const int nside = a.nside;
    Sample q;
    {
        const gptr<const float> axes = global_in(a.axes), ax_pts = global_in(a.ax_pts);
        q.kx = (double)axes[3 * b];
        q.ky = (double)axes[3 * b + 1];
        q.kz = (double)axes[3 * b + 2];
        q.ca = (double)global_in(a.cos_a)[b];
        q.sa = (double)global_in(a.sin_a)[b];
        q.oma = 1.0 - q.ca;
        q.bx = (double)ax_pts[3 * b];
        q.by = (double)ax_pts[3 * b + 1];
        q.bz = (double)ax_pts[3 * b + 2];
        q.cp = (double)global_in(a.cos_p)[b];
        q.sp = (double)global_in(a.sin_p)[b];
        q.omp = 1.0 - q.cp;
    }
    make_uniform(q);
    if (!(isfinite(q.kx) && isfinite(q.ky) && isfinite(q.kz) && isfinite(q.ca) && isfinite(q.sa) &&
          isfinite(q.bx) && isfinite(q.by) && isfinite(q.bz) && isfinite(q.cp) && isfinite(q.sp))) {
        // a non-finite operand: nothing is laid out (the job holds no beam
        // pixels, fp_set_job), the sample's tod is left as it is, and the
        // status word says so
        if (lane == 0) {
            for (int k = 0; k < 9; ++k) fr->m[k] = 0.0;
            for (int k = 0; k < 12; ++k) fr->q[k] = 0.0;
            fr->lam = fr->z_pts = fr->sth_pts = fr->sha = 0.0;
            fr->cha = 1.0;
            fr->ir_lo = 1;
            fr->nr = 0;
            fr->total = 0;
            fr->coinc_ring = fr->coinc_pos = fr->coinc_cell = -1;
            fr->direct = 0;
            if (a.status) atomicOr(a.status, 1);
        }
        return false;
    }
    double fx, fy, fz;  // footprint centre = this sample's image of c
    rotate(q, bf.cx, bf.cy, bf.cz, fx, fy, fz);
    const float fzf = fmaxf(-1.f, fminf(1.f, (float)fz));
    const float fthf = acosf(fzf);
    const float fstf = sqrtf(fmaxf(0.f, (1.f - fzf) * (1.f + fzf)));
    float fphf = atan2f((float)fy, (float)fx);
    if (fphf < 0.f) fphf += 6.2831855f;
    int jl = (int)rintf(fphf * (float)(kLamN / kTwoPi));
    if (jl >= kLamN) {
        // within half a grid step below 2 pi: lam = 0, and the windows'
        // longitudes are measured from the same origin (a window start
        // below 0 wraps like any other)
        jl -= kLamN;
        fphf -= 6.2831855f;
    }
    const double lam = (double)jl * (kTwoPi / kLamN);
    const gptr<const double> trig = global_in(a.lam_trig);
    const double cl = trig[2 * jl], sl = trig[2 * jl + 1];

    const float Rf = (float)bf.R, cosRf = (float)bf.cosR;
    const float th_lo = fthf - Rf, th_hi = fthf + Rf;
    int ir_lo = th_lo <= 0.f ? 1 : ring_above(nside, (double)cosf(th_lo));
    int ir_hi = th_hi >= 3.1415925f ? 4 * nside - 1 : ring_above(nside, (double)cosf(th_hi)) + 1;
    ir_lo = ir_lo < 1 ? 1 : ir_lo;
    ir_hi = ir_hi > 4 * nside - 1 ? 4 * nside - 1 : ir_hi;
    int nr = ir_hi - ir_lo + 1;
    // no layout (everything goes direct): too many rings, or a beam without a
    // weight grid (sum of |values| zero or not finite: M = 0)
    if (nr > kFpMaxRings || nr < 0 || !(bf.M > 0.0)) nr = 0;

    // Composite rotation: column j of R2 R1 is the reference rotation of the
    // unit vector e_j; its rows are then turned by -lam about z, so that the
    // rotated direction's longitude is lam + atan2(row1.v, row0.v).
    if (lane < 3) {
        double mx, my, mz;
        rotate(q, lane == 0 ? 1.0 : 0.0, lane == 1 ? 1.0 : 0.0, lane == 2 ? 1.0 : 0.0, mx, my,
               mz);
        fr->m[lane] = fma(cl, mx, sl * my);
        fr->m[3 + lane] = fma(cl, my, -(sl * mx));
        fr->m[6 + lane] = mz;
    }

    // boresight z and sin(theta) (tod_bilinear.py:306-310)
    const double z_pts = fmax(-1.0, fmin(1.0, q.bz));
    const double sth_pts = sqrt(fmax(0.0, 1.0 - q.bz * q.bz));
    // the boresight relative to lam: (x, y) = Rz(-lam) (bx, by), longitude
    // offset phb = atan2(y, x) (0 on the axis, as the reference's atan2(0, 0))
    const double xb = fma(cl, q.bx, sl * q.by);
    const double yb = fma(cl, q.by, -(sl * q.bx));
    // (the five-term series is within 5e-16 up to |t| = 0.05; the boresight
    // sits within ~2 deg of the footprint centre for focal-plane offsets up
    // to that, the general case goes out of line)
    const double phb = (xb > 0.0 && fabs(yb) < 0.05 * xb) ? atan_small(yb * rcp_nr(xb))
                                                           : atan2_wide(yb, xb);
    int carry = 0, laid = 0;
    bool unfit = false;  // a ring with a window beyond the grid's capacity
    for (int r0 = 0; r0 < nr; r0 += 64) {
        const int r = r0 + lane;
        int width = 0, i0 = 0;
        double base = 0.0;
        if (r < nr) {
            // (a copy: all of the row's loads are issued before the LDS
            // stores below, which the compiler would not move them past)
            const RingEntry g = a.rings[ir_lo + r];
            const float nf = (float)g.n;
            const double tc = (lam - g.phi0) * g.inv_dphi;  // lam's slot coordinate
            const float den = fstf * (float)g.sth;
            const float cd = den > 1e-6f ? (cosRf - fzf * (float)g.z) / den : -2.f;
            bool full = false;
            if (cd <= -1.f) {
                full = true;
            } else if (cd < 1.f) {
                const float dp = acosf(cd);
                const float spr = nf * (float)(1.0 / kTwoPi);  // slots per radian
                const float wf = 2.f * dp * spr + 5.f;
                if (wf >= nf) {
                    full = true;
                } else {
                    width = (int)wf;
                    const int f = (int)floorf((fphf - dp - (float)g.phi0) * spr) - 1;
                    base = tc - (double)f;
                    const int fm = f % g.n;
                    i0 = fm < 0 ? fm + g.n : fm;
                }
            }
            if (full) {
                width = g.n;
                base = tc >= (double)g.n ? tc - (double)g.n : (tc < 0.0 ? tc + (double)g.n : tc);
            }
            const Spin2Ring sr = spin2_ring(g.z, g.sth, z_pts, sth_pts);
            R.st2(r) = (float)sr.st2;
            R.sS(r) = (float)sr.S;
            R.s2dt(r) = (float)sr.s2dt;
            R.zsum(r) = (float)sr.zsum;
            R.sc4(r) = (float)sr.SC4;
            R.base(r) = base;
            R.inv_dphi(r) = g.inv_dphi;
            // the sweep's half longitude gaps to the boresight, window slot kb
            // nearest it (window slots are unwrapped: a gap never jumps by pi)
            const double kbd = rint(fma(phb, g.inv_dphi, base));
            R.gb(r) = (float)(0.5 * fma(kbd - base, g.dphi, -phb));
            R.hdp(r) = (float)(0.5 * g.dphi);
            R.kb(r) = (float)kbd;
            R.z(r) = g.z;
            R.sth(r) = g.sth;
            R.inv_dth(r) = g.inv_dth;
            R.inv_dth_f(r) = (float)g.inv_dth;
            R.phi0(r) = g.phi0;
            R.dphi(r) = g.dphi;
            R.n(r) = g.n;
            R.first(r) = (int)g.first;
        }
        int incl = width;  // inclusive scan of widths across the wave
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        const int off = carry + incl - width;
        const bool fits = off + width <= cap;
        int end = (r < nr && fits) ? off + width : 0;
        if (r < nr) {
            R.i0(r) = i0;
            R.width(r) = fits ? width : 0;  // beyond capacity: accumulated directly
            R.off(r) = off;
            R.baseo(r) = base + (double)off;
            R.ow(r) = make_int2(off, fits && width > 1 ? width - 1 : 0);
            if (fits)
                fill_ring_cells(rtab, off, off + width, r);
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) end = max(end, __shfl_xor(end, d, 64));
        laid = max(laid, end);
        carry += __shfl(incl, 63, 64);
        unfit = unfit || __any(r < nr && !fits && width > 0);
    }
    if (lane == 0) {
        // boresight half angle relative to lam, from the rotated boresight
        // (x, y) = Rz(-lam) (bx, by): cos a = x/rho, sin a = y/rho
        const double xb = fma(cl, q.bx, sl * q.by);
        const double yb = fma(cl, q.by, -(sl * q.bx));
        const double rho = sqrt(fma(xb, xb, yb * yb));
        double sha, cha;
        if (rho > 0.0) {
            const double ca = xb / rho, sa = yb / rho;
            if (ca >= 0.0) {
                cha = sqrt(0.5 * (1.0 + ca));
                sha = sa / (2.0 * cha);
            } else {
                sha = sqrt(0.5 * (1.0 - ca));
                if (sa < 0.0) sha = -sha;
                cha = sa / (2.0 * sha);
            }
        } else {  // boresight on the axis: the reference's atan2(0, 0) = 0
            sincos_red(-0.5 * lam, sha, cha);
        }
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
        // A map pixel centre can equal the boresight only on a ring whose z
        // and sin(theta) equal the boresight's bit for bit (e.g. z = 0).
        int cr = -1, cp = -1;
        const int ic = ring_above(nside, z_pts);
        const RingEntry* rings = a.rings;
        double cz[3], cs[3];  // the candidate rings' z and sin(theta), loaded together
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int ir = min(max(ic - 1 + k, 1), 4 * nside - 1);
            cz[k] = rings[ir].z;
            cs[k] = rings[ir].sth;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int ir = ic - 1 + k;
            if (ir < 1 || ir > 4 * nside - 1) continue;
            if (cz[k] == z_pts && cs[k] == sth_pts) {
                const RingEntry& g = a.rings[ir];
                const double ph = boresight_phi(q.bx, q.by);
                const double kp = rint((ph - g.phi0) / g.dphi);
                if (kp >= 0.0 && kp < (double)g.n && g.phi0 + kp * g.dphi == ph) {
                    cr = ir;
                    cp = (int)kp;
                }
            }
        }
        int cc = -1;  // its weight-grid cell
        if (cr >= ir_lo && cr - ir_lo < nr) {
            const int r = cr - ir_lo;
            int d = cp - R.i0(r);
            if (d < 0) d += R.n(r);
            if (d < R.width(r)) cc = R.off(r) + d;
        }
        fr->coinc_ring = cr;
        fr->coinc_pos = cp;
        fr->coinc_cell = cc;
        fr->lam = lam;
        fr->z_pts = z_pts;
        fr->sth_pts = sth_pts;
        fr->sha = sha;
        fr->cha = cha;
        fr->ir_lo = ir_lo;
        fr->nr = nr;
        TODG_CHECK(laid >= 0 && laid <= cap, kGridCell);
        fr->total = laid;  // grid cells [0, laid) belong to rings
        fr->direct = (nr == 0 || unfit) ? 1 : 0;
    }
    return true;
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
// lam + atan(t) instead of atan2; both differ from the reference's binary64
// values by a few ulp (|dz| < 1e-15, |dphi| < 1e-15 for sin(theta) > 0.05).
// A ring coordinate nside*(2 - 1.5 z) then differs by < 1e-11, a polar one
// nside*sqrt(3(1-|z|)) by < 1e-7, and a slot coordinate (phi - phi0)/dphi by
// < 1e-11 (including the reference's own rounding of its division), so a
// coordinate whose fractional part lies in [eps, 1 - eps] has the
// reference's integer part: ring_above, the two ring slots and hence all four
// pixels are the reference's.  The belt/cap choice needs no test of its own:
// z can fall on the other side of |z| = 2/3 from the reference's only within
// ~1e-15 of it, where both ring formulas give nside to within 2e-11, which the
// certification rejects.  Everything else -- margins, pole caps, rings or
// slots outside the laid-out windows, large longitude offsets -- goes to the
// exact path.  Weights only need tolerance accuracy (the ring weight is
// binary32); the four deposits are rounded onto the beam's weight grid
// (grid_magic) as round(wa) - round(wa fa), round(wa fa), and likewise on
// ring la + 1, so each pair sums to round(wa) and every cell's sum is exact
// and independent of the order of the atomic adds.
// Pixel of weight-grid cell c on footprint ring r (for the index dumps).
__device__ __forceinline__ int64_t fp_cell_pixel(const FpRings& R, int r, int c) {
    int pos = R.i0(r) + (c - R.off(r));
    const int n = R.n(r);
    if (pos >= n) pos -= n;
    return (int64_t)R.first(r) + pos;
}

template <bool DUMP = false, bool FC = false>
__device__ __forceinline__ bool fp_fast_bilinear(const double* __restrict__ m, int nside,
                                                 int ir_lo, int nr, const FpRings& R, WCell* W,
                                                 double u0, double u1, double u2, float bvf,
                                                 double M, FpDump dd = FpDump{}, int64_t di = 0) {
    // (conditions are combined with & rather than && so that they compile to
    // masks instead of branches; every LDS index is clamped into range first)
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
        ok = (fr > kCertSlot) & (fr < 1.0 - kCertSlot);
    } else {
        tr = (double)nside * sqrt_fast(3.0 * (1.0 - az));
        const int it = (int)tr;
        ira = z < 0.0 ? 4 * nside - it - 1 : it;
        const double fr = tr - (double)it;
        ok = (fr > kCertPolar) & (fr < 1.0 - kCertPolar);
    }
    // rings la and la + 1 inside the footprint (ir_lo >= 1 and
    // ir_lo + nr <= 4 nside, so then also 1 <= ira <= 4 nside - 2; nr = 0
    // when there is no layout)
    int la = ira - ir_lo;
    ok &= (unsigned)la < (unsigned)(nr > 1 ? nr - 1 : 0);
    // longitude offset from lam: atan(t), |t| < kMaxTan (t's relative error
    // < 2e-14 moves a slot coordinate by < 1e-12)
    const double t = yp * rcp_nr1(xp);
    ok &= (xp > 0.0) & (fabs(t) < kMaxTan);
    la = ok ? la : 0;
    const double dl = atan_small(t);
    // weight-grid cell coordinates on rings la, la + 1
    const double ta = fma(dl, R.inv_dphi(la), R.baseo(la));
    const double tb = fma(dl, R.inv_dphi(la + 1), R.baseo(la + 1));
    const int ia = (int)ta, ib = (int)tb;
    // (a negative coordinate truncates towards zero, leaving a fraction
    // <= 0, which the certification rejects)
    const double fa = ta - (double)ia, fb = tb - (double)ib;
    const int2 owa = R.ow(la), owb = R.ow(la + 1);
    ok &= ((unsigned)(ia - owa.x) < (unsigned)owa.y) & ((unsigned)(ib - owb.x) < (unsigned)owb.y);
    ok &= (fa > kCertSlot) & (fa < 1.0 - kCertSlot) & (fb > kCertSlot) & (fb < 1.0 - kCertSlot);
    // theta - theta_above = asin(sin(theta) z_a - z sin(theta_a))
    const double sq = sqrt_fast((1.0 - z) * (1.0 + z));
    const double x = fma(sq, R.z(la), -(z * R.sth(la)));
    ok &= fabs(x) < 1e-2;
    if (!ok) return false;
    // weights in binary32 (relative error ~1e-7): asin(x) by its series,
    // |x| < 1e-2 (x itself needed binary64: it is a difference of O(1) terms)
    const float xf = (float)x;
    const float x2 = xf * xf;
    float pp = fmaf(x2, 105.f / 3456.f, 15.f / 336.f);
    pp = fmaf(x2, pp, 3.f / 40.f);
    pp = fmaf(x2, pp, 1.f / 6.f);
    const float gfr = fmaf(xf * x2, pp, xf) * R.inv_dth_f(la);  // lower ring's weight g
    const float wb = gfr * bvf;
    const float wa = bvf - wb;
    const double wad = (double)wa, wbd = (double)wb;
    // M + round(wa), M + round(wa fa): both in [2^E, 2^(E+1)), so their
    // difference and the subtraction of M are exact
    const double ua = wad + M, va = fma(wad, fa, M);
    const double ub = wbd + M, vb = fma(wbd, fb, M);
    const double d0 = ua - va, d1 = va - M, d2 = ub - vb, d3 = vb - M;
    // the same four values as integer quanta: differences of the bit
    // patterns of ua, va, ub, vb and M, which share one binade
    WCell* ca = W + ia;
    WCell* cb = W + ib;
    TODG_CHECK(ia >= 0 && ib >= 0 && ia + 1 < (FC ? kFpCapFc : kFpCap) && ib + 1 < (FC ? kFpCapFc : kFpCap),
               kGridCell);
    const WCell nM = neg_bits(M), bva = wbits(va), bvb = wbits(vb);
    atomicAdd(ca, wbits(ua) - bva);
    atomicAdd(ca + 1, bva + nM);
    atomicAdd(cb, wbits(ub) - bvb);
    atomicAdd(cb + 1, bvb + nM);
    if constexpr (FC) {
        // the correction grid: each deposit times (1 - g) g, grid-rounded
        const double uc = (double)fmaf(-gfr, gfr, gfr);
        WCell* W1 = fp_W1();
        // (d uc + M by one fma: the product rounded onto the grid once)
        atomicAdd(W1 + ia, wbits(fma(d0, uc, M)) + nM);
        atomicAdd(W1 + ia + 1, wbits(fma(d1, uc, M)) + nM);
        atomicAdd(W1 + ib, wbits(fma(d2, uc, M)) + nM);
        atomicAdd(W1 + ib + 1, wbits(fma(d3, uc, M)) + nM);
    }
    if constexpr (DUMP) {
        dd.pix[4 * di] = fp_cell_pixel(R, la, ia);
        dd.pix[4 * di + 1] = fp_cell_pixel(R, la, ia + 1);
        dd.pix[4 * di + 2] = fp_cell_pixel(R, la + 1, ib);
        dd.pix[4 * di + 3] = fp_cell_pixel(R, la + 1, ib + 1);
        dd.w[4 * di] = d0;
        dd.w[4 * di + 1] = d1;
        dd.w[4 * di + 2] = d2;
        dd.w[4 * di + 3] = d3;
        dd.path[di] = 0;
    }
    return true;
}

// cos x for |x| <= 0.05 by Taylor series (truncation x^10/10! < 3e-20).
__device__ __forceinline__ double cos_tiny(double x) {
    const double x2 = x * x;
    double p = fma(x2, 1.0 / 40320.0, -1.0 / 720.0);
    p = fma(x2, p, 1.0 / 24.0);
    p = fma(x2, p, -0.5);
    return fma(x2, p, 1.0);
}

// Nearest fast path for one beam pixel: the reference's candidate set
// (rings ir, ir+1 x slots int(phi n / 2 pi), +1; tod_nearest.py:115-141)
// with the candidates' cos-distances evaluated from the composite rotation.
// Certification: the ring and both candidate slot coordinates are taken
// only with fractional parts in [kCertSlot, 1 - kCertSlot] (kCertPolar for
// the polar ring formula), so the candidates are the reference's; the
// winner is taken only when its cos-distance exceeds every other
// candidate's by more than kCertNearest, far above the few-ulp differences
// between these cos-distances and the reference's, so it is the reference's
// winner (ties, where the reference keeps the first candidate, fall to the
// exact path).  The winner's cell gets the beam value.
//
// The two candidates of one ring share sin(theta_r) and z_r, so their
// cos-distances ss cos(g) + zz order as |g| does: only each ring's nearer
// candidate needs its cosine, and its lead over the farther one,
// ss (cos g_w - cos g_l) = ss (g_l^2 - g_w^2) / 2 (1 - (g_l^2 + g_w^2) / 12
// + ...), is at least kHalfLow ss (g_l^2 - g_w^2) for |g| <= 0.05 (the
// series alternates with falling terms; the bracket is >= 1 - 0.005 / 12).
constexpr double kCertNearest = 1e-12;
constexpr double kHalfLow = 0.4995;

template <bool DUMP = false>
__device__ __forceinline__ bool fp_fast_nearest(const double* __restrict__ m, int nside, int ir_lo,
                                                int nr, const FpRings& R, WCell* W, double u0,
                                                double u1, double u2, double bv, double M,
                                                FpDump dd = FpDump{}, int64_t di = 0) {
    const double xp = fma(m[0], u0, fma(m[1], u1, m[2] * u2));
    const double yp = fma(m[3], u0, fma(m[4], u1, m[5] * u2));
    const double z = fma(m[6], u0, fma(m[7], u1, m[8] * u2));  // not clamped (tod_nearest.py:112)
    const double az = fabs(z);
    double tr;
    int ira;
    bool ok;
    if (az <= kTwoThirds) {
        tr = fma(z, -1.5 * (double)nside, 2.0 * (double)nside);
        ira = (int)tr;
        const double fr = tr - (double)ira;
        ok = (fr > kCertSlot) & (fr < 1.0 - kCertSlot);
    } else {
        tr = (double)nside * sqrt_fast(fmax(3.0 * (1.0 - az), 1e-300));
        const int it = (int)tr;
        ira = z < 0.0 ? 4 * nside - it - 1 : it;
        const double fr = tr - (double)it;
        ok = (fr > kCertPolar) & (fr < 1.0 - kCertPolar) & (az < 1.0);
    }
    // (the belt/cap choice and the ring range as in fp_fast_bilinear)
    int la = ira - ir_lo;
    ok &= (unsigned)la < (unsigned)(nr > 1 ? nr - 1 : 0);
    const double t = yp * rcp_nr1(xp);
    ok &= (xp > 0.0) & (fabs(t) < kMaxTan);
    la = ok ? la : 0;
    const double dl = atan_small(t);
    const double sth = sqrt_fast(fmax(fma(xp, xp, yp * yp), 1e-300));
    double d[2], lead[2];
    int c[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int r = la + k;
        // slot coordinate of phi n / 2 pi: the ring's (phi - phi0)/dphi
        // plus 1/2 on rings whose first pixel is shifted by half a slot
        const double b0 = R.base(r);
        const double tc = fma(dl, R.inv_dphi(r), R.phi0(r) > 0.0 ? b0 + 0.5 : b0);
        const int j = (int)tc;
        const double f = tc - (double)j;  // <= 0 for tc < 0: rejected below
        ok &= (j + 1 < R.width(r)) & (f > kCertSlot) & (f < 1.0 - kCertSlot);
        const double ss = sth * R.sth(r), zz = z * R.z(r), dphi = R.dphi(r);
        // phi_w - phi_c = dl + dphi (b0 - (j + q)) in window-relative slots
        // (rounding here moves the distances by ulps, far inside the
        // certification margins)
        const double bj = b0 - (double)j;
        const double g0 = fma(dphi, bj, dl);
        const double g1 = fma(dphi, bj - 1.0, dl);
        const bool far0 = fabs(g1) < fabs(g0);
        const double gw = far0 ? g1 : g0, gl = far0 ? g0 : g1;
        ok &= fabs(gl) <= 0.05;
        d[k] = fma(ss, cos_tiny(gw), zz);
        lead[k] = kHalfLow * ss * fma(gl, gl, -gw * gw);
        c[k] = R.off(r) + j + (far0 ? 1 : 0);
    }
    // best of the two rings' nearer candidates; the runner-up is the other
    // ring's nearer candidate or the best ring's farther one
    const bool w1 = d[1] > d[0];
    const double best = w1 ? d[1] : d[0], other = w1 ? d[0] : d[1];
    const int cell = w1 ? c[1] : c[0];
    ok &= (best - other > kCertNearest) & ((w1 ? lead[1] : lead[0]) > kCertNearest);
    if (!ok) return false;
    const double w = qround(bv, M);
    TODG_CHECK((unsigned)cell < (unsigned)kFpCap, kGridCell);
    atomicAdd(W + cell, wq(bv, M));
    if constexpr (DUMP) {
        dd.pix[4 * di] = fp_cell_pixel(R, la + (w1 ? 1 : 0), cell);
        dd.w[4 * di] = w;
        dd.path[di] = 0;
    }
    return true;
}

// Phase 1 of one sample -- sweep the beam and scatter beam_val * weight
// into the LDS weight grid; returns the outside-grid sums.  Waves claim
// kClaim beam pixels at a time from the sample's counter; wave 0 joins once
// it has laid out the next sample.
constexpr int kClaim = 256;  // beam pixels per claim (4 per lane)
constexpr int kClaimSub = 2;  // consecutive beam pixels per lane in a claim's block
constexpr int kTailPixels = 1024;  // the beam's last >= 1024 pixels go in one-block claims
static_assert((kClaim / 64) % kClaimSub == 0, "whole blocks per claim");

// A claim is kClaim / (64 kClaimSub) blocks of 64 kClaimSub beam pixels; in
// block k lane l takes pixels 64 kClaimSub k + kClaimSub l .. + kClaimSub - 1,
// so in one iteration the wave's lanes sit kClaimSub pixels apart: few of
// them scatter into the same weight-grid cell (same-address LDS atomics
// serialise) or, at the ~1 cell per 2 pixels of the beams here, into the
// same LDS bank.  Claim-relative index of a lane's it-th beam pixel, minus
// kClaimSub lane:
__device__ __forceinline__ constexpr int claim_offset(int it) {
    return (it / kClaimSub) * 64 * kClaimSub + it % kClaimSub;
}

template <int MODE, bool QU, int CS, bool DUMP>
__device__ __forceinline__ Acc4 fp_phase1_fixed(FpGlobals g, const float* vec, const float* bvs, int S,
                                                int buf, FpDump dd, int64_t dbase);

template <int MODE, bool QU, int CS, int CLAIM = kClaim, bool DUMP = false>
__device__ __forceinline__ Acc4 fp_phase1(const TodArgs& a, FpJob* J, const FpFrame* fr,
                                          const FpRings& R, int buf, WCell* W) {
    static_assert((CLAIM / 64) % kClaimSub == 0 && CLAIM >= 64 * kClaimSub, "whole blocks per claim");
    const int lane = threadIdx.x & 63;
    const int nside = a.nside;
    const gptr<const float> __restrict__ vec = global_in(J->vec_orig);
    const gptr<const float> __restrict__ bvs = global_in(J->beam_vals);
    // S <= 2^28 (checked on the host; clamped here so that 3 S stays in int range)
    const int64_t S64 = (int64_t)__builtin_bit_cast(uint64_t, uniform(__builtin_bit_cast(double, J->S)));
    const int S = (int)(S64 < (int64_t(1) << 28) ? S64 : (int64_t(1) << 28));
    const int ir_lo = __builtin_amdgcn_readfirstlane(fr->ir_lo);
    const int nr = __builtin_amdgcn_readfirstlane(fr->nr);
    const double M = uniform(J->M);
    double m[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = uniform(fr->m[k]);
    FpDump dd{};
    int64_t dbase = 0;
    if constexpr (DUMP) {
        dd = fp_dump_of(a.dump, a.B, a.S);
        dbase = J->b * a.S;
    }
    Acc4 accd = {{0.0, 0.0, 0.0, 0.0}};  // outside-grid sums
    // A sample without a (complete) layout -- its footprint exceeds the grid
    // -- accumulates the neighbours outside the layout directly into
    // per-thread sums; in nearest, bicubic and factorised-corrector mode it
    // runs with a fixed beam pixel per thread, which keeps those sums, and
    // the result, independent of timing.  (Bilinear keeps the dynamic claims for such samples: the
    // test alone costs its fast loop 3 %, measured; those samples' sums then
    // agree between runs to rounding.)
    if constexpr (MODE != 0) {
        if (__builtin_expect(__builtin_amdgcn_readfirstlane(fr->direct) != 0, 0))
            return fp_phase1_fixed<MODE, QU, CS, DUMP>(fp_globals(a), J->vec_orig, J->beam_vals, S, buf,
                                                       dd, dbase);
    }
    // claims are pipelined: the next chunk is claimed, and its first beam
    // pixel fetched, while the current chunk is swept.  Claims 0 .. nbig - 1
    // hold CLAIM beam pixels each; the last kTailPixels or more of the beam
    // go in claims of one 64 kClaimSub block, so that the waves reach the
    // barrier after phase 1 within a block's sweep of each other.
    constexpr int kTail = 64 * kClaimSub;
    const int nbig = S > kTailPixels + CLAIM ? (S - kTailPixels) / CLAIM : 0;
    auto claim = [&]() {
        int c = 0;
        if (lane == 0) c = atomicAdd(&J->claim, 1);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
        return c < nbig ? c * CLAIM : nbig * CLAIM + (c - nbig) * kTail;
    };
    int s0 = claim();
    float nv0 = 0.f, nv1 = 0.f, nv2 = 0.f, nbv = 0.f;
    constexpr int kPer = CLAIM / 64;
    constexpr int kPerTail = kTail / 64;
    constexpr int kSub = kClaimSub;
    auto in_claim = [](int it) { return claim_offset(it); };
    // beam pixel sb + in_claim(it) + kSub lane of a claim starting at sb
    // (uniform), clamped to the last one so that the loads need no branch,
    // through 32-bit byte offsets from the uniform bases (12 (S + kClaim) < 2^32)
    const uint32_t lane12 = 12u * kSub * lane, lane4 = 4u * kSub * lane;
    const uint32_t last12 = 12u * (uint32_t)(S - 1), last4 = 4u * (uint32_t)(S - 1);
    auto fetch = [&](int sb, int it) {
        const uint32_t o12 = min(12u * (uint32_t)(sb + in_claim(it)) + lane12, last12);
        const uint32_t o4 = min(4u * (uint32_t)(sb + in_claim(it)) + lane4, last4);
        const gptr<const float> vp = (gptr<const float>)((gptr<const char>)vec + o12);
        nv0 = vp[0];
        nv1 = vp[1];
        nv2 = vp[2];
        nbv = *(gptr<const float>)((gptr<const char>)bvs + o4);
    };
    // one claim of NPER iterations starting at s0 (uniform); returns the
    // next claim's start, whose first beam pixel it has fetched
    auto sweep = [&](auto nper) {
        constexpr int NPER = decltype(nper)::value;
        const int s1 = claim();
        uint32_t slow = 0;
#pragma unroll
        for (int it = 0; it < NPER; ++it) {
            const int s = s0 + in_claim(it) + lane * kSub;
            const double v0 = (double)nv0, v1 = (double)nv1, v2 = (double)nv2;
            const float bvf = nbv;
            // fetch the next beam pixel early (the next chunk's first one
            // after the last iteration)
            if (it + 1 < NPER)
                fetch(s0, it + 1);
            else
                fetch(s1, 0);
            if (s < S) {
                bool ok = false;
                if constexpr (MODE == 0 || MODE == 4)
                    ok = fp_fast_bilinear<DUMP, MODE == 4>(m, nside, ir_lo, nr, R, W, v0, v1, v2, bvf, M, dd,
                                                           dbase + s);
                if constexpr (MODE == 1)
                    ok = fp_fast_nearest<DUMP>(m, nside, ir_lo, nr, R, W, v0, v1, v2, (double)bvf, M, dd,
                                               dbase + s);
                if (!ok) slow |= 1u << it;
            }
        }
        // exact path for the beam pixels not certified above
        while (slow) {
            const int it = __builtin_ctz(slow);
            slow &= slow - 1;
            const int ss = s0 + in_claim(it) + lane * kSub;
            const Acc4 d = fp_exact_pixel<MODE, QU, CS, DUMP>(fp_globals(a), buf, vec[3 * ss],
                                                              vec[3 * ss + 1], vec[3 * ss + 2], bvs[ss],
                                                              dd, dbase + ss);
#pragma unroll
            for (int j = 0; j < kMaxSlots; ++j) accd.v[j] += d.v[j];
        }
        return s1;
    };
    if (S > 0) fetch(s0, 0);
    // (a big claim starts below nbig CLAIM <= S - kTailPixels, a tail claim
    // at or above it)
    while (s0 < nbig * CLAIM) s0 = sweep(std::integral_constant<int, kPer>{});
    while (s0 < S) s0 = sweep(std::integral_constant<int, kPerTail>{});
    return accd;
}

// Phase 1 of a sample without a complete layout: beam pixel s on thread
// s mod 512, every one on the exact path (its neighbours go into the laid-out
// cells where there are any, the others straight into the thread's sums).
template <int MODE, bool QU, int CS, bool DUMP>
__device__ __forceinline__ Acc4 fp_phase1_fixed(FpGlobals g, const float* vec, const float* bvs, int S,
                                                int buf, FpDump dd, int64_t dbase) {
    Acc4 acc = {{0.0, 0.0, 0.0, 0.0}};
    const gptr<const float> v = global_in(vec), b = global_in(bvs);
    for (int s = (int)threadIdx.x; s < S; s += kFpThreads) {
        const Acc4 d = fp_exact_pixel<MODE, QU, CS, DUMP>(g, buf, v[3 * s], v[3 * s + 1], v[3 * s + 2],
                                                          b[s], dd, dbase + s);
#pragma unroll
        for (int j = 0; j < kMaxSlots; ++j) acc.v[j] += d.v[j];
    }
    return acc;
}

// ============================================ footprint mode 3: corrector

// The corrector's matrices as MFMA A operands, held in registers for the
// life of the kernel.  Each layer is computed transposed, out^T = W in^T,
// with v_mfma_f32_16x16x32_bf16 (lane l: A[l%16][8(l/16) .. +7],
// B[8(l/16) .. +7][l%16], D[4(l/16) .. +3][l%16]): a layer's output tile
// leaves lane l holding features 4(l/16) .. +3 of row l%16, which is the B
// operand of the next layer once the next layer's k order is permuted so
// that k-block kb, lane group g holds features 16(2kb) + 4g .. +3 and
// 16(2kb+1) + 4g .. +3.  Layer 1 reads its inputs through a per-wave LDS
// tile in natural k order, with b1 folded in as input 21 = 1.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kCorrHT = todg_corr::kH / 16;  // hidden features in output tiles of 16
constexpr int kCorrHB = todg_corr::kH / 32;  // hidden features in k-blocks of 32
static_assert(todg_corr::kH % 32 == 0 && todg_corr::kOut == 16, "corrector tiling");

struct CorrRegs {
    bf16x8 w1[kCorrHT];
    bf16x8 w2[kCorrHT][kCorrHB];
    bf16x8 w3[kCorrHB];
};

__device__ __forceinline__ int corr_perm(int kb, int g, int j) {
    return 16 * (2 * kb + (j >> 2)) + 4 * g + (j & 3);
}

__device__ __forceinline__ void corr_load(const float* __restrict__ P, CorrRegs& c) {
    using namespace todg_corr;
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int ft = 0; ft < kCorrHT; ++ft) {
        const int o = 16 * ft + r;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 8 * g + j;
            const float w = k < kIn ? P[kOffW1 + o * kIn + k] : (k == kIn ? P[kOffB1 + o] : 0.f);
            c.w1[ft][j] = (__bf16)w;
        }
#pragma unroll
        for (int kb = 0; kb < kCorrHB; ++kb)
#pragma unroll
            for (int j = 0; j < 8; ++j) c.w2[ft][kb][j] = (__bf16)P[kOffW2 + o * kH + corr_perm(kb, g, j)];
    }
#pragma unroll
    for (int kb = 0; kb < kCorrHB; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j) c.w3[kb][j] = (__bf16)P[kOffW3 + r * kH + corr_perm(kb, g, j)];
    // b2, b3 (FP32) go to LDS: each lane reads the four of its output rows
    float* bias = &lds_at<float>(kOffBias);
    for (int i = threadIdx.x; i < kH + kOut; i += kFpThreads)
        bias[i] = i < kH ? P[kOffB2 + i] : P[kOffB3 + i - kH];
}

__device__ __forceinline__ bf16x8 corr_act(const f32x4& lo, const f32x4& hi) {
    bf16x8 b;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        b[q] = (__bf16)todg_corr::gelu(lo[q]);
        b[4 + q] = (__bf16)todg_corr::gelu(hi[q]);
    }
    return b;
}

// The MLP over the wave's 64 rows.  On entry every lane has written its
// input row (32 bf16: 21 inputs, the constant 1, zeros) to the wave's tile;
// on return o[] holds the lane's 16 outputs (the tile is reused for them).
__device__ __forceinline__ void corr_mlp(const CorrRegs& c, unsigned char* tile, float o[16]) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    // b2 as [ft * 4 + g], b3 at [kH / 4 + g]
    const f32x4* bias = reinterpret_cast<const f32x4*>(&lds_at<float>(kOffBias));
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        unsigned char* row = tile + (16 * t + r) * 64 + 16 * g;
        const bf16x8 x = *reinterpret_cast<const bf16x8*>(row);
        const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
        f32x4 h[kCorrHT];
        bf16x8 a[kCorrHB];
#pragma unroll
        for (int ft = 0; ft < kCorrHT; ++ft)
            h[ft] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(c.w1[ft], x, zero, 0, 0, 0);
#pragma unroll
        for (int kb = 0; kb < kCorrHB; ++kb) a[kb] = corr_act(h[2 * kb], h[2 * kb + 1]);
#pragma unroll
        for (int ft = 0; ft < kCorrHT; ++ft) {
            h[ft] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(c.w2[ft][0], a[0], bias[ft * 4 + g], 0, 0, 0);
#pragma unroll
            for (int kb = 1; kb < kCorrHB; ++kb)
                h[ft] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(c.w2[ft][kb], a[kb], h[ft], 0, 0, 0);
        }
#pragma unroll
        for (int kb = 0; kb < kCorrHB; ++kb) a[kb] = corr_act(h[2 * kb], h[2 * kb + 1]);
        f32x4 out = __builtin_amdgcn_mfma_f32_16x16x32_bf16(c.w3[0], a[0], bias[todg_corr::kH / 4 + g], 0, 0, 0);
#pragma unroll
        for (int kb = 1; kb < kCorrHB; ++kb)
            out = __builtin_amdgcn_mfma_f32_16x16x32_bf16(c.w3[kb], a[kb], out, 0, 0, 0);
        *reinterpret_cast<f32x4*>(row) = out;  // row 16t + r, outputs 4g .. 4g + 3
    }
    const f32x4* mine = reinterpret_cast<const f32x4*>(tile + lane * 64);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const f32x4 v = mine[q];
        o[4 * q] = v[0];
        o[4 * q + 1] = v[1];
        o[4 * q + 2] = v[2];
        o[4 * q + 3] = v[3];
    }
}

// Writes one input row: geometry, normalised values, the constant 1.
__device__ __forceinline__ void corr_put_row(unsigned char* tile, const float geo[5],
                                             const float xh[16], bool on) {
    bf16x8 v[4];
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        float x = 0.f;
        if (k < 5) x = geo[k];
        else if (k < 21) x = xh[k - 5];
        else if (k == 21) x = 1.f;
        v[k >> 3][k & 7] = (__bf16)(on ? x : 0.f);
    }
    bf16x8* dst = reinterpret_cast<bf16x8*>(tile + (threadIdx.x & 63) * 64);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = v[q];
}

__device__ __forceinline__ float* fp_Iv() { return &lds_at<float>(kOffIv); }

// All threads: the driving field's value of every laid-out cell of a
// layout (waves take rings, lanes take slots).
template <int CS>
__device__ __forceinline__ void fp_load_values(const TodArgs& a, const FpFrame* fr, const FpRings& R) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nr = __builtin_amdgcn_readfirstlane(fr->nr);
    float* Iv = fp_Iv();
    for (int r = wave; r < nr; r += kFpWaves) {
        const int w = R.width(r), off = R.off(r), i0 = R.i0(r), n = R.n(r);
        const int64_t first = R.first(r);
        for (int d = lane; d < w; d += 64) {
            int pos = i0 + d;
            if (pos >= n) pos -= n;
            Iv[off + d] = slot(load_px<CS>(a.map, first + pos, a.npix), a.corr_slot);
        }
    }
}

// Corrector geometry of one beam pixel on the certified fast path (the same
// index certification as fp_fast_bilinear, for all four rings of the 4 x 4
// neighbourhood, which must lie inside the laid-out windows).  Returns
// false, touching nothing, when the exact path must take it.
struct CorrGeo {
    int cell[4];          // grid cell of slot j_r - 1 in footprint ring la - 1 + r
    double fa, fb, v;     // bilinear fractions and ring weight
    float fm, fp;         // fractions in rings ira - 1, ira + 2
};

__device__ __forceinline__ bool fp_fast_corr(const double* __restrict__ m, int nside, int ir_lo,
                                             int nr, const FpRings& R, double u0, double u1,
                                             double u2, CorrGeo& G) {
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
        ok = (fr > kCertSlot) & (fr < 1.0 - kCertSlot);
    } else {
        tr = (double)nside * sqrt_fast(3.0 * (1.0 - az));
        const int it = (int)tr;
        ira = z < 0.0 ? 4 * nside - it - 1 : it;
        const double fr = tr - (double)it;
        ok = (fr > kCertPolar) & (fr < 1.0 - kCertPolar);
    }
    ok &= fabs(az - kTwoThirds) > 1e-12;
    int la = ira - ir_lo;
    ok &= (ira >= 2) & (ira <= 4 * nside - 3) & (la >= 1) & (la + 2 < nr);
    ok &= (xp > 0.0) & (fabs(yp) < kMaxTan * xp);
    la = ok ? la : 1;
    const double dl = atan_small(yp * rcp_nr(xp));
    double fr4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int rr = la - 1 + k;
        const double tc = fma(dl, R.inv_dphi(rr), R.base(rr));
        const int j = (int)tc;
        const double f = tc - (double)j;
        ok &= (tc >= 1.0) & (j + 2 < R.width(rr)) & (f > kCertSlot) & (f < 1.0 - kCertSlot);
        fr4[k] = f;
        G.cell[k] = R.off(rr) + j - 1;
    }
    const double sq = sqrt_fast((1.0 - z) * (1.0 + z));
    const double x = fma(sq, R.z(la), -(z * R.sth(la)));
    ok &= fabs(x) < 1e-2;
    if (!ok) return false;
    const double x2 = x * x;
    double pp = fma(x2, 105.0 / 3456.0, 15.0 / 336.0);
    pp = fma(x2, pp, 3.0 / 40.0);
    pp = fma(x2, pp, 1.0 / 6.0);
    G.v = fma(x * x2, pp, x) * R.inv_dth(la);
    G.fa = fr4[1];
    G.fb = fr4[2];
    G.fm = (float)fr4[0];
    G.fp = (float)fr4[3];
    return true;
}

// Exact-path corrector inputs of one beam pixel (the reference rotation,
// atan2, ring_interp): geometry, the 16 values (from the value grid inside
// the layout, from the map outside it) and whether the corrector applies.
struct CorrIn {
    float geo[5];
    float val[16];
    int applies;
};

__device__ __forceinline__ bool corr_neighbourhood(const FpGlobals& a, double z, double phi_w,
                                                   Nb& o, double& fa, double& fb, double& v,
                                                   double& fm, double& fp, int ring[4], int js[4]) {
    const int ira = ring_above(a.nside, z);
    if (!(ira >= 2 && ira <= 4 * a.nside - 3)) return false;
    ring_interp(a.nside, a.rings, z, phi_w, o);
    const double wa = o.w[0] + o.w[1];
    v = o.w[2] + o.w[3];
    fa = wa > 0.0 ? o.w[1] / wa : 0.0;
    fb = v > 0.0 ? o.w[3] / v : 0.0;
    fm = fp = 0.0;
    for (int r = 0; r < 4; ++r) {
        const RingEntry& e = a.rings[ira - 1 + r];
        ring[r] = ira - 1 + r;
        if (r == 1) {
            js[r] = o.pos[0];
        } else if (r == 2) {
            js[r] = o.pos[2];
        } else {
            double fr;
            js[r] = ring_slot(phi_w, e.phi0, e.dphi, e.inv_dphi, e.n, fr);
            if (r == 0) fm = fr;
            else fp = fr;
        }
    }
    return true;
}

template <int CS>
__device__ __noinline__ CorrIn fp_corr_exact_inputs(FpGlobals a, int buf, int cslot, float u0,
                                                     float u1, float u2) {
    const FpRings R = fp_rings(buf);
    const FpFrame* fr = fp_frame(buf);
    Sample q;
    frame_sample(fr, q);
    double vx, vy, vz;
    rotate(q, (double)u0, (double)u1, (double)u2, vx, vy, vz);
    const double phi_w = wrap_phi(vy, vx);
    const double z = fmax(-1.0, fmin(1.0, vz));
    CorrIn ci;
    Nb o;
    double fa, fb, v, fm, fp;
    int ring[4], js[4];
    ci.applies = corr_neighbourhood(a, z, phi_w, o, fa, fb, v, fm, fp, ring, js) ? 1 : 0;
    for (int k = 0; k < 16; ++k) ci.val[k] = 0.f;
    ci.geo[0] = (float)(fa - 0.5);
    ci.geo[1] = (float)(v - 0.5);
    ci.geo[2] = (float)(fb - 0.5);
    ci.geo[3] = (float)(fm - 0.5);
    ci.geo[4] = (float)(fp - 0.5);
    if (!ci.applies) return ci;
    const int ir_lo = fr->ir_lo, nr = fr->nr;
    for (int r = 0; r < 4; ++r) {
        const RingEntry& e = a.rings[ring[r]];
        for (int t = 0; t < 4; ++t) {
            int pos = js[r] - 1 + t;
            pos = ((pos % e.n) + e.n) % e.n;
            const int lr = ring[r] - ir_lo;
            const int li = (lr >= 0 && lr < nr) ? fp_local(R, lr, pos) : -1;
            ci.val[4 * r + t] = li >= 0 ? fp_Iv()[li]
                                        : slot(load_px<CS>(a.map, e.first + pos, a.npix), cslot);
        }
    }
    return ci;
}

// Per-beam-pixel output of the MLP handed to the exact deposit.
struct CorrW {
    float dw[16];  // o_k - mean(o)
    int apply;
};

// Exact-path deposit of one beam pixel: corrected weights (bilinear +
// correction) when `apply`, the reference's bilinear weights otherwise;
// cells inside the layout go to the weight grid, the others are accumulated
// directly.  Out of line: only beam pixels off the fast path come here.
template <bool QU, int CS>
__device__ __noinline__ Acc4 fp_corr_exact_deposit(FpGlobals a, int buf, float u0, float u1,
                                                   float u2, float bvf, CorrW cw) {
    if (!cw.apply) return fp_exact_pixel<0, QU, CS>(a, buf, u0, u1, u2, bvf);
    const FpRings R = fp_rings(buf);
    const FpFrame* fr = fp_frame(buf);
    WCell* W = fp_W();
    Sample q;
    frame_sample(fr, q);
    const FpSpin sp = fp_spin_of(fr);
    double vx, vy, vz;
    rotate(q, (double)u0, (double)u1, (double)u2, vx, vy, vz);
    const double phi_w = wrap_phi(vy, vx);
    const double z = fmax(-1.0, fmin(1.0, vz));
    Nb o;
    double fa, fb, v, fm, fp;
    int ring[4], js[4];
    Acc4 acc = {{0.0, 0.0, 0.0, 0.0}};
    if (!corr_neighbourhood(a, z, phi_w, o, fa, fb, v, fm, fp, ring, js)) return acc;
    const double bv = (double)bvf;
    const double M = fp_job(buf)->M;
    const int ir_lo = fr->ir_lo, nr = fr->nr;
    for (int r = 0; r < 4; ++r) {
        const RingEntry& e = a.rings[ring[r]];
        for (int t = 0; t < 4; ++t) {
            const int k = 4 * r + t;
            double w = (double)cw.dw[k];
            if (k == 5) w += o.w[0];
            else if (k == 6) w += o.w[1];
            else if (k == 9) w += o.w[2];
            else if (k == 10) w += o.w[3];
            w = qround(w * bv, M);
            int pos = js[r] - 1 + t;
            pos = ((pos % e.n) + e.n) % e.n;
            const int lr = ring[r] - ir_lo;
            const int li = (lr >= 0 && lr < nr) ? fp_local(R, lr, pos) : -1;
            if (li >= 0) {
                TODG_CHECK(li < fr->total, kGridCell);
                atomicAdd(&W[li], wq(w, M));
            } else {
                const Acc4 d = fr->direct ? fp_direct<QU, CS>(a, sp, ring[r], pos, w)
                                          : fp_outside<QU, CS>(a, sp, ring[r], pos, w);
                for (int j = 0; j < kMaxSlots; ++j) acc.v[j] += d.v[j];
            }
        }
    }
    return acc;
}

// Phase 1 of one sample with the corrector, by every wave: wave 0 joins
// when it has laid out the next sample.  Waves claim 64 beam pixels at a
// time (one per lane) from the sample's LDS counter: geometry and the 16
// neighbourhood values (fast path from the value grid, else the exact
// path), the MLP on MFMA over the wave's 64 rows, then the corrected
// weights times the beam value into the weight grid.  Where the corrector
// does not apply (the rings next to the poles) or its correction exceeds
// the bound, the plain bilinear weights are used.
template <bool QU, int CS>
__device__ __forceinline__ Acc4 fp_phase1_corr(const TodArgs& a, FpJob* J, const FpFrame* fr,
                                               const FpRings& R, int buf, WCell* W,
                                               const CorrRegs& C, unsigned long long& n_corr,
                                               unsigned long long& n_fb) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nside = a.nside;
    unsigned char* tile = fp_smem + kOffStage + wave * kStageBytes;
    int* cnt = &J->claim;
    const gptr<const float> __restrict__ vec = global_in(J->vec_orig);
    const gptr<const float> __restrict__ bvs = global_in(J->beam_vals);
    const int64_t S = (int64_t)__builtin_bit_cast(uint64_t, uniform(__builtin_bit_cast(double, J->S)));
    const int ir_lo = __builtin_amdgcn_readfirstlane(fr->ir_lo);
    const int nr = __builtin_amdgcn_readfirstlane(fr->nr);
    const float bound = a.corr_bound;
    const int64_t K = a.corr_count;
    const float* Iv = fp_Iv();
    const double M = uniform(J->M);
    double m[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = uniform(fr->m[k]);
    Acc4 accd = {{0.0, 0.0, 0.0, 0.0}};
    constexpr int kPer = kClaim / 64;
    int it = kPer;
    int64_t s0 = 0;
    for (;;) {
        if (it == kPer) {  // claim the next kClaim beam pixels
            int c = 0;
            if (lane == 0) c = atomicAdd(cnt, 1);
            s0 = (int64_t)__builtin_amdgcn_readfirstlane(__shfl(c, 0, 64)) * kClaim;
            it = 0;
        }
        if (s0 + claim_offset(it) >= S) break;  // no lane has a beam pixel left (uniform)
        if (s0 >= K) {
            // the whole claim lies beyond the corrected beam pixels: the
            // plain bilinear path, no MLP (uniform branch)
            for (; it < kPer; ++it) {
                const int64_t s = s0 + (int64_t)lane * kClaimSub + claim_offset(it);
                if (s < S) {
                    const float u0 = vec[3 * s], u1 = vec[3 * s + 1], u2 = vec[3 * s + 2], bvf = bvs[s];
                    ++n_fb;
                    if (!fp_fast_bilinear(m, nside, ir_lo, nr, R, W, (double)u0, (double)u1, (double)u2,
                                          bvf, M)) {
                        const Acc4 d = fp_exact_pixel<0, QU, CS>(fp_globals(a), buf, u0, u1, u2, bvf);
#pragma unroll
                        for (int j = 0; j < kMaxSlots; ++j) accd.v[j] += d.v[j];
                    }
                }
            }
            continue;
        }
        // lanes kClaimSub beam pixels apart (claim_offset)
        const int64_t s = s0 + (int64_t)lane * kClaimSub + claim_offset(it);
        ++it;
        const bool live = s < S;
        float u0 = 0.f, u1 = 0.f, u2 = 0.f, bvf = 0.f;
        if (live) {
            u0 = vec[3 * s];
            u1 = vec[3 * s + 1];
            u2 = vec[3 * s + 2];
            bvf = bvs[s];
        }
        CorrGeo G;
        int kind = 0;  // 0 no beam pixel, 1 fast path, 2 exact path (corrector applies), 3 exact, bilinear
        float geo[5], val[16];
        if (live && fp_fast_corr(m, nside, ir_lo, nr, R, (double)u0, (double)u1, (double)u2, G)) {
            kind = 1;
            geo[0] = (float)(G.fa - 0.5);
            geo[1] = (float)(G.v - 0.5);
            geo[2] = (float)(G.fb - 0.5);
            geo[3] = G.fm - 0.5f;
            geo[4] = G.fp - 0.5f;
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int t = 0; t < 4; ++t) val[4 * r + t] = Iv[G.cell[r] + t];
        } else {
#pragma unroll
            for (int k = 0; k < 5; ++k) geo[k] = 0.f;
#pragma unroll
            for (int k = 0; k < 16; ++k) val[k] = 0.f;
            if (live) {
                const CorrIn ci = fp_corr_exact_inputs<CS>(fp_globals(a), buf, a.corr_slot, u0, u1, u2);
                kind = ci.applies ? 2 : 3;
#pragma unroll
                for (int k = 0; k < 5; ++k) geo[k] = ci.geo[k];
#pragma unroll
                for (int k = 0; k < 16; ++k) val[k] = ci.val[k];
            }
        }
        const bool on = (kind == 1 || kind == 2) && s < K;
        float mean = 0.f;
#pragma unroll
        for (int k = 0; k < 16; ++k) mean += val[k];
        mean *= 1.f / 16.f;
        float var = 0.f;
#pragma unroll
        for (int k = 0; k < 16; ++k) var = fmaf(val[k] - mean, val[k] - mean, var);
        const float sd = sqrtf(var * (1.f / 16.f));
        const float inv = sd > 1e-30f ? 1.f / sd : 0.f;
        float xh[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) xh[k] = (val[k] - mean) * inv;
        corr_put_row(tile, geo, xh, on);
        float o[16];
        corr_mlp(C, tile, o);
        float om = 0.f;
#pragma unroll
        for (int k = 0; k < 16; ++k) om += o[k];
        om *= 1.f / 16.f;
        float corr = 0.f;
#pragma unroll
        for (int k = 0; k < 16; ++k) corr = fmaf(o[k] - om, xh[k], corr);
        const bool apply = on && fabsf(corr) <= bound;
        if (live) {
            if (apply) ++n_corr;
            else ++n_fb;
        }
        if (kind == 1) {
            const double bv = (double)bvf;
            const double wb = G.v * bv;
            const double wa = bv - wb;
            double w16[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) w16[k] = apply ? (double)(o[k] - om) * bv : 0.0;
            w16[5] += wa - wa * G.fa;
            w16[6] += wa * G.fa;
            w16[9] += wb - wb * G.fb;
            w16[10] += wb * G.fb;
            TODG_CHECK(G.cell[0] >= 0 && G.cell[1] >= 0 && G.cell[2] >= 0 && G.cell[3] >= 0 &&
                           G.cell[0] + 3 < kFpCap && G.cell[1] + 3 < kFpCap && G.cell[2] + 3 < kFpCap &&
                           G.cell[3] + 3 < kFpCap,
                       kGridCell);
            if (apply) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        atomicAdd(W + G.cell[r] + t, wq(w16[4 * r + t], M));
            } else {
                atomicAdd(W + G.cell[1] + 1, wq(w16[5], M));
                atomicAdd(W + G.cell[1] + 2, wq(w16[6], M));
                atomicAdd(W + G.cell[2] + 1, wq(w16[9], M));
                atomicAdd(W + G.cell[2] + 2, wq(w16[10], M));
            }
        } else if (kind >= 2) {
            CorrW cw;
#pragma unroll
            for (int k = 0; k < 16; ++k) cw.dw[k] = o[k] - om;
            cw.apply = apply ? 1 : 0;
            const Acc4 d = fp_corr_exact_deposit<QU, CS>(fp_globals(a), buf, u0, u1, u2, bvf, cw);
#pragma unroll
            for (int j = 0; j < kMaxSlots; ++j) accd.v[j] += d.v[j];
        }
    }
    return accd;
}

// Footprint pixels whose ring lies within sin^2(d/2) < kSpinF32St2 of the
// boresight (d < ~2 deg) and whose half longitude gap is small take the
// binary32 spin-2 (absolute error ~1e-7); the rest, where its terms can
// cancel, the binary64 one.
constexpr float kSpinF32St2 = 3e-4f;

// All threads: phase 2 -- sweep the laid-out footprint once: spin-2 and one
// gather per map pixel with nonzero weight; the cell is cleared for the next
// sample.  Cell j of the grid lies on footprint ring rtab[j]; a wave takes
// 64 consecutive cells at a time, kP2Cells such chunks per round, so that
// each lane has kP2Cells independent weight reads and gathers in flight.
// Returns the per-thread sums (outside-grid sums included).  Every slot of
// the packed map is accumulated, padding included (its values are 0, and
// the flush writes the first nslots sums only): a per-slot test would cost
// each cell two selects per slot.
constexpr int kP2Cells = 4;

template <bool QU, int CS, int P2 = kP2Cells, bool FC = false>
__device__ __forceinline__ Acc4 fp_phase2(const TodArgs& a, const FpFrame* fr, const FpRings& R,
                                          const unsigned char* rtab, WCell* W, double M, Acc4 acc) {
    constexpr int kP2 = P2;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int total = __builtin_amdgcn_readfirstlane(fr->total);
    const int coinc = __builtin_amdgcn_readfirstlane(fr->coinc_cell);
    const uint32_t pmax = (uint32_t)(a.npix - 1);
    const FpSpin sp = fp_spin_of(fr);
    for (int base = wave * 64; base < total; base += kFpThreads * kP2) {
        double wsum[kP2], w1sum[kP2];
        int rr[kP2], kk[kP2];
        float4 mv[kP2], mv1[kP2];
        WCell* W1 = fp_W1();
#pragma unroll
        for (int u = 0; u < kP2; ++u) {
            const int j = base + u * kFpThreads + lane;
            const bool in = j < total;
            // (a cell of 0 quanta converts to exactly 0)
            wsum[u] = in ? wval(W[j], M) : 0.0;
            w1sum[u] = (FC && in) ? wval(W1[j], M) : 0.0;
            rr[u] = in ? (int)rtab[j] : 0;
        }
        // (mode 4: a cell is live when either grid holds weight)
        auto live = [&](int u) { return FC ? ((wsum[u] != 0.0) | (w1sum[u] != 0.0)) : wsum[u] != 0.0; };
#pragma unroll
        for (int u = 0; u < kP2; ++u) {
            const int j = base + u * kFpThreads + lane;
            kk[u] = 0;
            if (live(u)) {
                W[j] = (WCell)0;
                if constexpr (FC) W1[j] = (WCell)0;
                const int r = rr[u];
                const int k = j - R.off(r);  // window slot
                int pos = R.i0(r) + k;
                const int n = R.n(r);
                if (pos >= n) pos -= n;
                kk[u] = k;
                TODG_CHECK((unsigned)k < (unsigned)R.width(r) && (unsigned)pos < (unsigned)n, kRingSlot);
                // cell j < total lies in ring r's window: 0 <= k < width <= n
                // and 0 <= i0 < n, so pos is in [0, n) and the pixel in the
                // ring (npix < 2^31); the clamp only backs that up
                const uint32_t pix = (uint32_t)(R.first(r) + pos);
                mv[u] = load_px_u32<CS>(a.map, pix, pmax);
                if constexpr (FC) mv1[u] = load_px_u32<CS>(a.map1, pix, pmax);
            }
        }
#pragma unroll
        for (int u = 0; u < kP2; ++u) {
            if (!live(u)) continue;
            const int j = base + u * kFpThreads + lane;
            const int r = rr[u], k = kk[u];
            const double wsum_u = wsum[u];
            const float4 mv_u = mv[u];
            if constexpr (QU) {
                double c2d, s2d;
                // half the longitude gap from the boresight to the pixel
                const float gh = fmaf((float)k - R.kb(r), R.hdp(r), R.gb(r));
                const float st2 = R.st2(r);
                if (j == coinc) {  // tod_bilinear.py:108-109
                    c2d = 1.0;
                    s2d = 0.0;
                } else if (st2 < kSpinF32St2 && fabsf(gh) < 0.05f) {
                    // pixel within ~2 deg of the boresight: binary32, sin and
                    // cos of the half gap by Taylor series (|gh| < 0.05)
                    const float g2 = gh * gh;
                    const float sg = fmaf(gh * g2, fmaf(g2, 1.f / 120.f, -1.f / 6.f), gh);
                    const float cg = fmaf(g2, fmaf(g2, fmaf(g2, -1.f / 720.f, 1.f / 24.f), -0.5f), 1.f);
                    Spin2RingF g;
                    g.st2 = st2;
                    g.S = R.sS(r);
                    g.s2dt = R.s2dt(r);
                    g.zsum = R.zsum(r);
                    g.SC4 = R.sc4(r);
                    float cf, sf;
                    spin2_from_half_f(g, -sg, cg, cf, sf);
                    c2d = (double)cf;
                    s2d = (double)sf;
                } else {
                    int pos = R.i0(r) + k;
                    const int n = R.n(r);
                    if (pos >= n) pos -= n;
                    double sb, cb;
                    // the half longitude gap relative to lam (a difference of angles)
                    const double hd = (R.phi0(r) + (double)pos * R.dphi(r) - sp.lam) * 0.5;
                    sincos_any(hd, sb, cb);
                    const double sh = fma(sp.sha, cb, -(sp.cha * sb));
                    const double ch = fma(sp.cha, cb, sp.sha * sb);
                    const Spin2Ring g = spin2_ring(R.z(r), R.sth(r), sp.z_pts, sp.sth_pts);
                    spin2_from_half(g, sh, ch, c2d, s2d);
                }
                if constexpr (FC) {
                    // both channels weighted first, then the spin-2 rotation
                    const double w1 = w1sum[u];
                    const float4 m1 = mv1[u];
                    const double qv = fma(w1, (double)m1.x, wsum_u * (double)mv_u.x);
                    const double uv = fma(w1, (double)m1.y, wsum_u * (double)mv_u.y);
                    acc.v[0] += fma(qv, c2d, uv * s2d);
                    acc.v[1] += fma(uv, c2d, -(qv * s2d));
#pragma unroll
                    for (int k2 = 2; k2 < CS; ++k2)
                        acc.v[k2] += fma(w1, (double)slot(m1, k2), wsum_u * (double)slot(mv_u, k2));
                } else {
                    const double qv = (double)mv_u.x, uv = (double)mv_u.y;
                    acc.v[0] = fma(wsum_u, fma(qv, c2d, uv * s2d), acc.v[0]);
                    acc.v[1] = fma(wsum_u, fma(uv, c2d, -(qv * s2d)), acc.v[1]);
#pragma unroll
                    for (int k2 = 2; k2 < CS; ++k2) acc.v[k2] = fma(wsum_u, (double)slot(mv_u, k2), acc.v[k2]);
                }
            } else {
#pragma unroll
                for (int k2 = 0; k2 < CS; ++k2) {
                    if constexpr (FC)
                        acc.v[k2] += fma(w1sum[u], (double)slot(mv1[u], k2), wsum_u * (double)slot(mv_u, k2));
                    else
                        acc.v[k2] = fma(wsum_u, (double)slot(mv_u, k2), acc.v[k2]);
                }
            }
        }
    }
    return acc;
}

// The wave's sums of the four slots into red[0..3]: a butterfly that
// halves the values per lane at each step (reduce-scatter: 7 exchanges of a
// double instead of 4 x 6), leaving slot 2 (lane >= 32) + ((lane >> 4) & 1)
// summed over each group of 16 lanes.
// v of the lane that DPP control CTRL selects (every lane has a source for
// the quad permutes and mirrors used here), both halves of the double.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// x + the partner lane's x across the halves of the wave for lanes 0..31
// (partner l + 32), y + partner's y for lanes 32..63, by v_permlane32_swap:
// it exchanges the upper half of one register with the lower half of the
// other, and the two results add up to the pair sum whichever comes first.
__device__ __forceinline__ double pair_sum32(double x, double y) {
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(y), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(y), false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}

// The same across rows of 16 lanes (v_permlane16_swap): x + partner's x in
// even rows (partner l + 16), y + partner's y in odd rows.
__device__ __forceinline__ double pair_sum16(double x, double y) {
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(y), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(y), false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}

__device__ __forceinline__ void fp_wave_reduce4(const Acc4& acc, double* red) {
    const int lane = threadIdx.x & 63;
    const bool hi = lane >= 32;
    const bool b4 = (lane >> 4) & 1;
    // lanes 0..31: slots 0 and 1 summed over l, l + 32; lanes 32..63: slots 2 and 3
    const double a0 = pair_sum32(acc.v[0], acc.v[2]);
    const double a1 = pair_sum32(acc.v[1], acc.v[3]);
    // even rows: a0 summed over two rows, odd rows: a1
    double v = pair_sum16(a0, a1);
    // the sum over each row of 16 lanes by DPP, without LDS round trips:
    // pairs, quads (quad_perm), half-rows (row_half_mirror), rows (row_mirror)
    v += dpp_f64<0xb1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4e>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    if ((lane & 15) == 0) red[(hi ? 2 : 0) + (b4 ? 1 : 0)] = v;
}

// Thread 0: add a finished sample's block sums (red[], complete since the
// last barrier) to its tod entries, slots in order (two slots may share a
// row when c_q == c_u).  Each tod entry belongs to one sample and each sample
// to one workgroup, so the atomic adds have no contention and leave the
// result deterministic; they need no reply, so nothing waits on them.
__device__ __forceinline__ void fp_flush(const TodArgs& a, const FpJob* J) {
    if (threadIdx.x == 0 && J->S > 0) {  // (S = 0: a skipped non-finite sample)
        const gptr<double> tod = (gptr<double>)J->tod;
        const int64_t B = J->B, b = J->b;
        for (int k = 0; k < kMaxSlots && k < a.nslots; ++k) {
            double sum = 0.0;
            for (int w = 0; w < kFpWaves; ++w) sum += fp_red()[w * kMaxSlots + k];
            __builtin_amdgcn_global_atomic_fadd_f64(tod + (int64_t)a.row[k] * B + b, sum);
        }
    }
}

// Zero the weight grid (both channels' halves in mode 4) and the side list.
__device__ __forceinline__ void fp_init() {
    for (int j = threadIdx.x; j < kFpCap; j += kFpThreads) fp_W()[j] = (WCell)0;
    if (threadIdx.x == 0) *fp_side_count() = 0;
    __syncthreads();
}

// Thread 0: record sample b of item a in job slot buf (M: the beam's
// weight-grid rounding constant).
__device__ __forceinline__ void fp_set_job(const TodArgs& a, int64_t b, int buf, double M) {
    if (threadIdx.x == 0) {
        FpJob* J = fp_job(buf);
        J->vec_orig = a.vec_orig;
        J->beam_vals = a.beam_vals;
        J->tod = a.tod;
        J->S = a.S;
        J->B = a.B;
        J->b = b;
        J->claim = 0;
        J->M = M;
    }
}

// Wave 0: job and layout of sample b into buffer buf; a sample with a
// non-finite operand gets an empty job (no beam pixels, no tod update).
__device__ __forceinline__ void fp_prepare(const TodArgs& a, int64_t b, const BeamFrame& bf, int buf,
                                           int cap = kFpCap) {
    fp_set_job(a, b, buf, bf.M);
    const bool ok = fp_layout(a, b, bf, fp_frame(buf), fp_rings(buf), fp_rtab(buf), cap);
    if (!ok && threadIdx.x == 0) fp_job(buf)->S = 0;
}

// The software pipeline of the footprint kernels, one step: wave 0 lays out
// the next sample (if any, item `an`, index bn, beam frame bfn) into the
// other buffer while waves 1..7 run phase 1 of the current sample; after a
// barrier all waves run phase 2 and the block reduction; after a second
// barrier thread 0 adds the sums to tod.  Two barriers per sample, and the
// layout's latency hides behind phase 1.
template <int MODE, bool QU, int CS, int CLAIM = kClaim, int P2 = kP2Cells, bool DUMP = false>
__device__ __forceinline__ void fp_step(const TodArgs& an, bool have_next, int64_t bn,
                                        const BeamFrame& bfn, int buf, const CorrRegs* C = nullptr,
                                        unsigned long long* n_corr = nullptr,
                                        unsigned long long* n_fb = nullptr) {
    const int wave = threadIdx.x >> 6;
    const FpRings R = fp_rings(buf);
    const FpFrame* fr = fp_frame(buf);
    Acc4 acc = {{0.0, 0.0, 0.0, 0.0}};
    if (wave == 0 && have_next) fp_prepare(an, bn, bfn, buf ^ 1, MODE == 4 ? kFpCapFc : kFpCap);
    // every wave, wave 0 once its layout is done: the job's claim counter
    // was zeroed when the job was set
    if constexpr (MODE == 3)
        acc = fp_phase1_corr<QU, CS>(an, fp_job(buf), fr, R, buf, fp_W(), *C, *n_corr, *n_fb);
    else
        acc = fp_phase1<MODE, QU, CS, CLAIM, DUMP>(an, fp_job(buf), fr, R, buf, fp_W());
    __syncthreads();
    acc = fp_phase2<QU, CS, P2, MODE == 4>(an, fr, R, fp_rtab(buf), fp_W(), uniform(fp_job(buf)->M), acc);
    // the pixels outside the layout, in pixel order (one thread)
    if (threadIdx.x == 0 && *fp_side_count() != 0)
        acc = fp_side_flush<QU, CS, MODE == 4>(fp_globals(an), fr, acc);
    // mode 3: the next sample's values go into the value grid, which no
    // wave reads between the barrier above and the one below
    if constexpr (MODE == 3)
        if (have_next) fp_load_values<CS>(an, fp_frame(buf ^ 1), fp_rings(buf ^ 1));
    fp_wave_reduce4(acc, fp_red() + wave * kMaxSlots);
    __syncthreads();
    fp_flush(an, fp_job(buf));
}

template <int MODE, bool QU, int CS, int MINW, int CLAIM = kClaim, int P2 = kP2Cells, bool DUMP = false>
__global__ __launch_bounds__(kFpThreads, MINW) void k_tod_fp(const TodArgs a) {
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
    fp_init();
    // footprint margin in pixel sizes: bilinear and nearest neighbour rings
    // lie within one ring spacing (<= 0.9) of the query (the factorised
    // corrector deposits into the bilinear neighbours too), 4 x 4 corrector
    // neighbourhoods within ~3.6, bicubic candidates within ~4
    const double margin = MODE == 2 ? 5.0 : (MODE == 3 ? 4.5 : 1.25);
    // the beam frame from the block's first sample with finite operands
    int64_t bfs = bbeg + kx;
    while (bfs + per < bend && !finite_sample(a, bfs)) bfs += per;
    const BeamFrame bf = fp_beam_frame(beam_src(a), bfs, fp_scratch(), margin);
    int64_t b = bbeg + kx;
    if ((threadIdx.x >> 6) == 0) fp_prepare(a, b, bf, 0, MODE == 4 ? kFpCapFc : kFpCap);
    __syncthreads();
    CorrRegs C;
    unsigned long long n_corr = 0, n_fb = 0;
    if constexpr (MODE == 3) {
        corr_load(a.corr, C);
        fp_load_values<CS>(a, fp_frame(0), fp_rings(0));
        __syncthreads();
    }
    int buf = 0;
    for (;;) {
        const int64_t bn = b + per;
        const bool have_next = bn < bend;
        fp_step<MODE, QU, CS, CLAIM, P2, DUMP>(a, have_next, bn, bf, buf, &C, &n_corr, &n_fb);
        if (!have_next) break;
        b = bn;
        buf ^= 1;
    }
    if constexpr (MODE == 3) {
        if (a.corr_stats) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                n_corr += __shfl_xor(n_corr, off, 64);
                n_fb += __shfl_xor(n_fb, off, 64);
            }
            if ((threadIdx.x & 63) == 0) {
                atomicAdd(&a.corr_stats[0], n_corr);
                atomicAdd(&a.corr_stats[1], n_fb);
            }
        }
    }
}

// ======================================================= batched kernel

// Shared fields of a batch; per-detector fields come from todg_work_item.
struct BatchArgs {
    const todg_work_item* items;  // device array
    int n_items;
    int64_t total;                // sum of items[i].B
    const float* map;
    const RingEntry* rings;
    const double* lam_trig;
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

// Persistent multi-detector kernel.  The global sample range (all items'
// samples back to back) is split into 8 contiguous parts, one work queue per
// XCD; a workgroup claims `chunk` samples at a time from its own XCD's
// queue and, once that is drained, from the others, so uneven beams and
// chunk lengths balance themselves.  Every workgroup exits when all eight
// queues are drained.  A sample with a non-finite pointing operand is
// skipped and flags its item's status word.
static_assert(((sizeof(TodArgs) + 15) & ~15) + sizeof(BeamFrame) <= kItemBytes, "item slot");

struct BatchCursor {
    int att;          // queues tried so far
    int64_t g, g1;    // current claim [g, g1)
    int cur;          // item of `a`
    bool have_frame;
};

// All threads of the block (uniform): advance to the next sample to process;
// loads its item into `a`, its beam frame into `bf`.  Returns false when
// every queue is drained.
__device__ __forceinline__ bool batch_next(const BatchArgs& ba, BatchCursor& c, TodArgs& a,
                                           BeamFrame& bf, int64_t& b, double margin_px) {
    const int tid = threadIdx.x;
    const int xcd = blockIdx.x & 7;
    const int64_t qlen = (ba.total + 7) / 8;
    int* claim = reinterpret_cast<int*>(fp_scratch() + 2 * kFpWaves);
    for (;;) {
        while (c.g >= c.g1) {
            if (c.att >= 8) return false;
            const int qi = (xcd + c.att) & 7;
            const int64_t qbeg = (int64_t)qi * qlen;
            const int64_t qend = (qbeg + qlen < ba.total) ? qbeg + qlen : ba.total;
            if (tid == 0) *claim = (int)atomicAdd(&ba.queue[qi], (unsigned int)ba.chunk);
            __syncthreads();
            const int64_t g0 = qbeg + (int64_t)(unsigned int)*claim;
            __syncthreads();
            if (g0 >= qend) {
                ++c.att;
                continue;
            }
            c.g = g0;
            c.g1 = (g0 + ba.chunk < qend) ? g0 + ba.chunk : qend;
        }
        const int64_t g = c.g++;
        const int it = find_item(ba.items, ba.n_items, g);
        const todg_work_item& w = ba.items[it];
        if (it != c.cur) {
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
            c.cur = it;
            c.have_frame = false;
        }
        b = g - w.first;
        if (a.S <= 0) continue;
        if (!finite_sample(a, b)) {
            if (tid == 0 && ba.status) atomicOr(&ba.status[it], 1);
            continue;
        }
        if (!c.have_frame) {
            // beam frame from the item's first finite sample, whichever
            // workgroup computes it: every workgroup lays out the item's
            // samples the same way, so results do not depend on the claims
            int64_t b0 = 0;
            while (b0 < b && !finite_sample(a, b0)) ++b0;
            const BeamFrame f = fp_beam_frame(beam_src(a), b0, fp_scratch(), margin_px);
            bf = f;
            c.have_frame = true;
        }
        return true;
    }
}

template <int MODE, bool QU, int CS, int MINW>
__global__ __launch_bounds__(kFpThreads, MINW) void k_tod_fp_batch(const BatchArgs ba,
                                                                   int unused) {
    // item and beam frame in LDS (every thread writes the same values)
    TodArgs& a = lds_at<TodArgs>(kOffItem);
    BeamFrame& bf = lds_at<BeamFrame>(kOffItem + (((int)sizeof(TodArgs) + 15) & ~15));
    a.map = ba.map;
    a.rings = ba.rings;
    a.lam_trig = ba.lam_trig;
    a.npix = ba.npix;
    a.nside = ba.nside;
    a.nslots = ba.nslots;
    a.status = nullptr;  // (non-finite samples are skipped in batch_next)
    a.dump = nullptr;
#pragma unroll
    for (int k = 0; k < kMaxSlots; ++k) a.row[k] = ba.row[k];
    fp_init();
    BatchCursor c = {0, 0, 0, -1, false};
    int64_t b = 0;
    const double margin_px = MODE == 2 ? 5.0 : 1.25;
    if (!batch_next(ba, c, a, bf, b, margin_px)) return;
    if ((threadIdx.x >> 6) == 0) fp_prepare(a, b, bf, 0);
    __syncthreads();
    int buf = 0;
    for (;;) {
        int64_t bn = 0;
        const bool have_next = batch_next(ba, c, a, bf, bn, margin_px);
        fp_step<MODE, QU, CS>(a, have_next, bn, bf, buf);
        if (!have_next) break;
        buf ^= 1;
    }
}

static __global__ void k_zero_u32(unsigned int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <int MODE, bool QU, int CS>
hipError_t launch_direct_cs(const TodArgs& a, int T, hipStream_t st) {
    const int64_t ntiles = (a.B + T - 1) / T;
    const int64_t waves_cap = (int64_t)todg_internal::n_cu() * 32;
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
// one workgroup per CU) or 4 (128 VGPRs, two workgroups per CU).  CLAIM =
// beam pixels per phase-1 claim, P2 = 64-cell chunks per lane and phase-2
// round, DUMP = the index-dump build.
template <int MODE, bool QU, int CS, int MINW, int CLAIM = kClaim, int P2 = kP2Cells, bool DUMP = false>
hipError_t launch_fp_cs(const TodArgs& a, hipStream_t st) {
    const size_t smem = MODE == 3 ? kFpSmemBytesCorr : (MODE == 4 ? kFpSmemBytesFc : kFpSmemBytes);
    static std::once_flag once;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)k_tod_fp<MODE, QU, CS, MINW, CLAIM, P2, DUMP>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    });
    // Workgroups resident per CU (LDS allows two), a multiple of 8 for the
    // XCD mapping, no more than the samples need.
    int64_t blocks = (int64_t)todg_internal::n_cu() * (MINW >= 4 ? 2 : 1) * 2;
    const int64_t need = ((a.B + 7) / 8) * 8;
    if (blocks > need) blocks = need;
    blocks = ((blocks + 7) / 8) * 8;
    hipLaunchKernelGGL((k_tod_fp<MODE, QU, CS, MINW, CLAIM, P2, DUMP>), dim3((unsigned)blocks),
                       dim3(kFpThreads), smem, st, a);
    return hipGetLastError();
}

// Tile variants of the footprint kernel (4 waves/SIMD): variant 16 + 3 i + j
// takes claims of {128, 256, 512}[i] beam pixels and {3, 4, 5}[j] chunks per
// phase-2 round.  Built for the two production channel layouts (I/Q/U with
// spin-2, and one field); other layouts run the default tile.
constexpr int kTileFirst = 16;
constexpr int kTileCount = 9;

template <int MODE, bool QU, int CS>
hipError_t launch_fp_tile(const TodArgs& a, int tile, hipStream_t st) {
    if constexpr (MODE < 2 && ((QU && CS == 4) || (!QU && CS == 1))) {
        switch (tile) {
            case 0: return launch_fp_cs<MODE, QU, CS, 4, 128, 3>(a, st);
            case 1: return launch_fp_cs<MODE, QU, CS, 4, 128, 4>(a, st);
            case 2: return launch_fp_cs<MODE, QU, CS, 4, 128, 5>(a, st);
            case 3: return launch_fp_cs<MODE, QU, CS, 4, 256, 3>(a, st);
            case 4: return launch_fp_cs<MODE, QU, CS, 4, 256, 4>(a, st);
            case 5: return launch_fp_cs<MODE, QU, CS, 4, 256, 5>(a, st);
            case 6: return launch_fp_cs<MODE, QU, CS, 4, 512, 3>(a, st);
            case 7: return launch_fp_cs<MODE, QU, CS, 4, 512, 4>(a, st);
            default: return launch_fp_cs<MODE, QU, CS, 4, 512, 5>(a, st);
        }
    } else {
        (void)tile;
        return launch_fp_cs<MODE, QU, CS, 4>(a, st);
    }
}

template <int MODE, bool QU, int CS, int MINW>
hipError_t launch_batch_cs(const BatchArgs& ba, hipStream_t st) {
    const size_t smem = kFpSmemBytes;
    static std::once_flag once;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)k_tod_fp_batch<MODE, QU, CS, MINW>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    });
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, ba.queue, 8);
    int64_t blocks = (int64_t)todg_internal::n_cu() * (MINW >= 4 ? 2 : 1) * 2;
    blocks = ((blocks + 7) / 8) * 8;
    hipLaunchKernelGGL((k_tod_fp_batch<MODE, QU, CS, MINW>), dim3((unsigned)blocks),
                       dim3(kFpThreads), smem, st, ba, 0);
    return hipGetLastError();
}

template <int MODE, bool QU, int CS>
hipError_t launch_batch_v(const BatchArgs& ba, int variant, hipStream_t st) {
    if (variant == 8) return launch_batch_cs<MODE, QU, CS, 2>(ba, st);
    return launch_batch_cs<MODE, QU, CS, 4>(ba, st);
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
    if constexpr (MODE == 3) {
        // the corrector keeps its matrices in registers: the 2-waves build
        return launch_fp_cs<MODE, QU, CS, 2>(a, st);
    } else if constexpr (MODE == 4) {
        (void)variant;
        return launch_fp_cs<MODE, QU, CS, 4>(a, st);
    } else {
        if constexpr (MODE < 2) {  // the bicubic judge runs on the footprint kernel only
            if (variant == 1 || variant == 2 || variant == 4)
                return launch_direct_cs<MODE, QU, CS>(a, variant, st);
            if (variant >= kTileFirst && variant < kTileFirst + kTileCount)
                return launch_fp_tile<MODE, QU, CS>(a, variant - kTileFirst, st);
        }
        if (variant == 8) return launch_fp_cs<MODE, QU, CS, 2>(a, st);
        return launch_fp_cs<MODE, QU, CS, 4>(a, st);
    }
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

// Per-mode launchers (defined in tod_fp_<mode>.hip): qu / CS from the map's
// slot layout, variant as documented in todgpu.h.
hipError_t launch_m0(const TodArgs& a, bool qu, int CS, int variant, hipStream_t st);
hipError_t launch_m1(const TodArgs& a, bool qu, int CS, int variant, hipStream_t st);
hipError_t launch_m2(const TodArgs& a, bool qu, int CS, int variant, hipStream_t st);
hipError_t launch_m3(const TodArgs& a, bool qu, int CS, hipStream_t st);
hipError_t launch_m4(const TodArgs& a, bool qu, int CS, hipStream_t st);
// The factorised corrector's stencil coefficients: c[0..15] build the main
// channel, c[16..31] the correction channel (k_fcorr_channels, in
// tod_fp_fcorr.hip).
struct FcorrCoef {
    double c[32];
};
hipError_t launch_fcorr_channels(const RingEntry* rings, int nside, const float* field, const FcorrCoef& c,
                                 float* out0, float* out1, int64_t stride, hipStream_t st);
// The MLP stencil corrector's map side (tod_fp_fcorr.hip): per-pixel
// features and MLP inputs (training), and the fused channel builder.
hipError_t launch_smlp_features(const RingEntry* rings, int nside, const float* field, float* F, float* X,
                                hipStream_t st);
hipError_t launch_smlp_channels(const RingEntry* rings, int nside, const float* field, const float* params,
                                float* out0, float* out1, int64_t stride, hipStream_t st);
hipError_t launch_batch_m0(const BatchArgs& ba, bool qu, int CS, int variant, hipStream_t st);
hipError_t launch_batch_m1(const BatchArgs& ba, bool qu, int CS, int variant, hipStream_t st);
hipError_t launch_batch_m2(const BatchArgs& ba, bool qu, int CS, int variant, hipStream_t st);
// index-dump builds (C = 3 with Q/U, or one field)
hipError_t launch_dump_m0(const TodArgs& a, bool qu, hipStream_t st);
hipError_t launch_dump_m1(const TodArgs& a, bool qu, hipStream_t st);

}  // namespace todg_fp
