// Harmonic tools on the GPU: the beam transfer function and the ring sums of
// a spherical-harmonic synthesis.
//
// Beam transfer function:
//
//   B_ell = sum_i w_i P_ell(x_i),   ell = 0 .. lmax
//
// with the Legendre three-term recurrence per beam pixel
// (tod_utils.py:147-251: P_{l+1} = ((2l+1) x P_l - l P_{l-1}) / (l+1)).
// Each thread carries one pixel's recurrence; every ell's partial sum over a
// block is reduced with wavefront shuffles and LDS, 64 multipoles per
// barrier; a second kernel adds the blocks' partials.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "internal.hpp"
#include "../../include/todgpu.h"

namespace todg_beam {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kChunk = 64;  // multipoles per LDS round

__global__ __launch_bounds__(kThreads) void k_bell_partial(const double* __restrict__ x,
                                                           const double* __restrict__ w,
                                                           int64_t N, int lmax,
                                                           double* __restrict__ part) {
    __shared__ double red[kWaves][kChunk];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t i = (int64_t)blockIdx.x * kThreads + tid;
    const double xi = i < N ? x[i] : 0.0;
    const double wi = i < N ? w[i] : 0.0;
    double p2 = 1.0, p1 = xi;  // P_{l-1}, P_l as the loop advances
    for (int l0 = 0; l0 <= lmax; l0 += kChunk) {
        for (int k = 0; k < kChunk; ++k) {
            const int l = l0 + k;
            double p;
            if (l == 0) {
                p = 1.0;
            } else if (l == 1) {
                p = xi;
            } else {
                const double lm = (double)(l - 1);
                p = ((2.0 * lm + 1.0) * xi * p1 - lm * p2) / (lm + 1.0);
                p2 = p1;
                p1 = p;
            }
            double v = wi * p;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) red[wave][k] = v;
        }
        __syncthreads();
        if (tid < kChunk && l0 + tid <= lmax) {
            double s = 0.0;
            for (int wv = 0; wv < kWaves; ++wv) s += red[wv][tid];
            part[(int64_t)blockIdx.x * (lmax + 1) + l0 + tid] = s;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_bell_sum(const double* __restrict__ part, int nblocks,
                                                  int lmax, double* __restrict__ bell) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > lmax) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[(int64_t)b * (lmax + 1) + l];
    bell[l] = s;
}

__global__ __launch_bounds__(256) void k_fill_zero(double* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Scratch of todg_beam_bell_noalloc: the per-block partial sums, at most
// kScratchDoubles of them (8 MiB per device, held until the process ends).
constexpr int64_t kScratchDoubles = int64_t(1) << 20;

}  // namespace todg_beam

// Ring sums of a spin-0 spherical-harmonic synthesis (what healpy's
// alm2map / synfast computes before the per-ring FFT; the synthetic skies of
// the workloads and the corrector's training and judging are made this way):
//
//   F_m(theta_r) = sum_{l=m}^{lmax} a_lm lambda_lm(theta_r),
//
// lambda_lm the orthonormal associated Legendre functions, by their
// three-term recurrence in l at fixed m.  One lane per ring, one grid row per
// m: the wave's lanes share m and l, so the recurrence coefficients and the
// a_lm read are uniform.  lambda_mm = (-1)^m exp(logmm_m) sin^m theta is
// carried as a signed mantissa and an exponent E (value = mantissa exp(E));
// whenever the mantissa exceeds `big` it is divided by big and E grows by
// log(big), so polar rings at high m neither underflow nor lose precision,
// and terms whose exp(E) underflows add zero.  Every operation is a
// correctly rounded binary64 one in a fixed order (no contraction), so the
// sums do not depend on the launch shape.
namespace todg_sht {

constexpr int kThreads = 256;

template <int K>
__global__ __launch_bounds__(kThreads) void k_ring_sums(const double2* __restrict__ alm, int lmax,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ st, int64_t R,
                                                        const double* __restrict__ logmm, double big,
                                                        double logbig, double2* __restrict__ F) {
    const int m = blockIdx.y;
    const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= R) return;  // no barriers below
    const int64_t M = lmax + 1;
    const double xr = x[r];
    const double dm = (double)m;
    double E = dm * log(fmax(st[r], 1e-300)) + logmm[m];
    double p1 = (m & 1) ? -1.0 : 1.0;  // mantissa of lambda_{l-1,m} (lambda_mm before the loop)
    double p2 = 0.0;                   // mantissa of lambda_{l-2,m}
    double scale = exp(fmin(E, 700.0));
    double2 acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = make_double2(0.0, 0.0);
    const double2* __restrict__ a = alm + (int64_t)m * M;  // a_lm of field k at a[k M^2 + l]
    for (int l = m; l <= lmax; ++l) {
        const double dl = (double)l;
        double cur;
        if (l == m) {
            cur = p1;
        } else if (l == m + 1) {
            cur = xr * sqrt(2.0 * dm + 3.0) * p1;
        } else {
            const double ca = sqrt((4.0 * dl * dl - 1.0) / fmax(dl * dl - dm * dm, 1.0));
            const double l1 = dl - 1.0;
            const double cb = sqrt(fmax((l1 * l1 - dm * dm) / (4.0 * (l1 * l1) - 1.0), 0.0));
            cur = ca * (xr * p1 - cb * p2);
        }
        const double v = cur * scale;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double2 c = a[(int64_t)k * M * M + l];
            acc[k].x += c.x * v;
            acc[k].y += c.y * v;
        }
        if (l > m) {
            p2 = p1;
            p1 = cur;
        }
        if (fabs(p1) > big) {
            p1 /= big;
            p2 /= big;
            E += logbig;
            scale = exp(fmin(E, 700.0));
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) F[((int64_t)k * M + m) * R + r] = acc[k];
}

}  // namespace todg_sht

// The no-allocation entry's scratch.  In host code this name stands for the
// array but is not its device address: kernels get the address that
// hipGetSymbolAddress returns for the current device (scratch_address).
static __device__ __align__(256) double scratch[todg_beam::kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
double* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

// The current device's address of `scratch`, looked up once per device (with
// that device current) and kept.  No allocation: the array lives in the
// library's code object on each device.
int scratch_address(double** out) {
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return todg_internal::set_error(TODG_ERR_HIP, hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return todg_internal::set_error(TODG_ERR_ARG, "device index beyond 63");
    std::call_once(g_scratch_once[dev], [dev] {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    });
    if (g_scratch_err[dev] != hipSuccess)
        return todg_internal::set_error(TODG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                                          hipGetErrorString(g_scratch_err[dev]));
    *out = g_scratch_ptr[dev];
    return TODG_OK;
}

}  // namespace

extern "C" {

size_t todg_beam_bell_workspace(int64_t N, int lmax) {
    const int64_t nb = (N + todg_beam::kThreads - 1) / todg_beam::kThreads;
    return (size_t)(nb > 0 ? nb : 1) * (size_t)(lmax + 1);
}

int todg_beam_bell(const double* x, const double* w, int64_t N, int lmax, double* ws, double* bell,
                   void* stream) {
    if (N < 0 || lmax < 0) return todg_internal::set_error(TODG_ERR_ARG, "N and lmax must be >= 0");
    if (!bell || (N > 0 && (!x || !w || !ws)))
        return todg_internal::set_error(TODG_ERR_ARG, "null pointer argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (N == 0) {
        hipError_t e = hipMemsetAsync(bell, 0, (size_t)(lmax + 1) * sizeof(double), st);
        return e == hipSuccess ? TODG_OK : todg_internal::set_error(TODG_ERR_HIP, hipGetErrorString(e));
    }
    const int64_t nb = (N + todg_beam::kThreads - 1) / todg_beam::kThreads;
    if (nb > (1 << 24)) return todg_internal::set_error(TODG_ERR_ARG, "too many beam pixels");
    hipLaunchKernelGGL(todg_beam::k_bell_partial, dim3((unsigned)nb), dim3(todg_beam::kThreads), 0, st,
                       x, w, N, lmax, ws);
    hipLaunchKernelGGL(todg_beam::k_bell_sum, dim3((unsigned)((lmax + 256) / 256)), dim3(256), 0, st,
                       ws, (int)nb, lmax, bell);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : todg_internal::set_error(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_beam_bell_noalloc(const double* x, const double* w, int64_t N, int lmax, double* bell,
                           void* stream) {
    if (N < 0 || lmax < 0) return todg_internal::set_error(TODG_ERR_ARG, "N and lmax must be >= 0");
    if (!bell || (N > 0 && (!x || !w)))
        return todg_internal::set_error(TODG_ERR_ARG, "null pointer argument");
    const int64_t nb = (N + todg_beam::kThreads - 1) / todg_beam::kThreads;
    const int64_t need = (nb > 0 ? nb : 1) * (int64_t)(lmax + 1);
    if (need > todg_beam::kScratchDoubles)
        return todg_internal::set_error(
            TODG_ERR_ARG, "todg_beam_bell_noalloc: N = " + std::to_string(N) + " and lmax = " +
                              std::to_string(lmax) + " need " + std::to_string(need) +
                              " doubles of scratch, more than its " +
                              std::to_string(todg_beam::kScratchDoubles) +
                              " (ceil(N / 256) * (lmax + 1) must stay within it; todg_beam_bell "
                              "takes a caller workspace of any size)");
    double* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (N == 0) {  // a fill kernel rather than a memset node, so graph capture replays it
        hipLaunchKernelGGL(todg_beam::k_fill_zero, dim3((unsigned)((lmax + 256) / 256)), dim3(256), 0, st,
                           bell, lmax + 1);
    } else {
        // k_bell_partial writes every scratch entry that k_bell_sum reads:
        // nothing needs clearing first
        hipLaunchKernelGGL(todg_beam::k_bell_partial, dim3((unsigned)nb), dim3(todg_beam::kThreads), 0, st,
                           x, w, N, lmax, ws);
        hipLaunchKernelGGL(todg_beam::k_bell_sum, dim3((unsigned)((lmax + 256) / 256)), dim3(256), 0, st,
                           ws, (int)nb, lmax, bell);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : todg_internal::set_error(TODG_ERR_HIP, hipGetErrorString(e));
}

int todg_ring_sums(const double* alm, int K, int lmax, const double* x, const double* st, int64_t R,
                   const double* logmm, double big, double logbig, double* F, void* stream) {
    if (K < 1 || K > 4) return todg_internal::set_error(TODG_ERR_ARG, "K must be 1..4 fields");
    if (lmax < 0 || lmax > 65534) return todg_internal::set_error(TODG_ERR_ARG, "lmax must be 0..65534");
    if (R < 0 || R > (int64_t(1) << 31)) return todg_internal::set_error(TODG_ERR_ARG, "R must be 0..2^31 rings");
    if (!(big > 1.0) || !(logbig > 0.0))
        return todg_internal::set_error(TODG_ERR_ARG, "big must exceed 1 and logbig be its logarithm");
    if (R == 0) return TODG_OK;
    if (!alm || !x || !st || !logmm || !F) return todg_internal::set_error(TODG_ERR_ARG, "null pointer argument");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)((R + todg_sht::kThreads - 1) / todg_sht::kThreads), (unsigned)(lmax + 1));
    const double2* a = reinterpret_cast<const double2*>(alm);
    double2* f = reinterpret_cast<double2*>(F);
    switch (K) {
        case 1: hipLaunchKernelGGL(todg_sht::k_ring_sums<1>, grid, dim3(todg_sht::kThreads), 0, s, a, lmax, x, st, R, logmm, big, logbig, f); break;
        case 2: hipLaunchKernelGGL(todg_sht::k_ring_sums<2>, grid, dim3(todg_sht::kThreads), 0, s, a, lmax, x, st, R, logmm, big, logbig, f); break;
        case 3: hipLaunchKernelGGL(todg_sht::k_ring_sums<3>, grid, dim3(todg_sht::kThreads), 0, s, a, lmax, x, st, R, logmm, big, logbig, f); break;
        default: hipLaunchKernelGGL(todg_sht::k_ring_sums<4>, grid, dim3(todg_sht::kThreads), 0, s, a, lmax, x, st, R, logmm, big, logbig, f); break;
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? TODG_OK : todg_internal::set_error(TODG_ERR_HIP, hipGetErrorString(e));
}

}  // extern "C"
