// fh_metrics.hip -- FP64 evidence of lab-frame spectra, one workgroup per
// spectrum: the quantities _field_metrics computes for metrics()/verify()
// (propagation.jl:176-202).
//
// Per spectrum of M modes x N samples (FFT order):
//   mode energy E_m   = dt sum_t |fft(u_m)|^2 = dt N sum_w |u_m|^2   (Parseval)
//   energy            = sum_m E_m
//   peak power        = max_t sum_m |fft(u_m)[t]|^2
//   temporal edges    = edge fraction of sum_m |fft(u_m)|^2          (_edge_fraction)
//   spectral edges    = edge fraction of fftshift(sum_m |u_m|^2)
//   photon number     = dt sum_{w,m} |u_m(w)|^2 / |omega(w)|          (_photon_number)
// _edge_fraction takes e = max(1, floor(0.05 N)) samples at each end of the
// array.  In FFT order the ends of the fftshift-ed spectrum are the bins
// around Nyquist: shifted index (q + N/2) mod N.  The time-domain power uses
// the block FFT of the plan (its precision and twiddles); every sum is FP64
// through deterministic workgroup trees.
#include <hip/hip_runtime.h>

#include "fh_launch.hpp"
#include "fh_single.hpp"

namespace fh {

template <typename R> struct MetricArgs {
    const C2<R>* spec;     // [count][modes][N]
    const C2<R>* tw;       // Stockham twiddles of FftPlan<N, T>
    const double* omega;   // [N] |omega_s| in FFT order, rad/ps
    double dt;             // sample spacing, ps
    int modes;
    double* out;           // [count][modes + 5]
};

template <int T> __device__ __forceinline__ double block_max(double x, void* lds_raw)
{
    double* red = reinterpret_cast<double*>(lds_raw);
    const int tid = threadIdx.x;
    __syncthreads();
    red[tid] = x;
    __syncthreads();
#pragma unroll 1
    for (int w = T / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

template <typename R, int N, int T>
__global__ __launch_bounds__(T) void k_field_metrics(MetricArgs<R> a)
{
    using V = C2<R>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = N / T;
    constexpr bool PLANAR = SmTraits<R, N, T>::PLANAR;
    __shared__ __align__(16) unsigned char lds[SmLayout<R, N, T>::LDS_BYTES];
    const size_t blk = blockIdx.x;
    const Buf tw(a.tw, FftPlan<N, T>::TW_SIZE * sizeof(V));
    const Buf om(a.omega, N * sizeof(double));
    const Buf spec(a.spec + blk * (size_t)a.modes * N, (uint32_t)((size_t)a.modes * N * sizeof(V)));
    double* out = a.out + blk * (size_t)(a.modes + 5);
    const int e = max(1, (int)floor(0.05 * N));
    double tp[P], sp[P];
#pragma unroll
    for (int k = 0; k < P; ++k) tp[k] = sp[k] = 0.0;
    double photon = 0.0, etot = 0.0;
#pragma unroll 1
    for (int m = 0; m < a.modes; ++m) {
        V v[P];
        double em = 0.0;
        const uint32_t moff = (uint32_t)m * (uint32_t)(N * sizeof(V));
#pragma unroll
        for (int k = 0; k < P; ++k) {
            v[k] = spec.ld<V>(CV::voff() + moff, CV::soff(k));
            const double p2 = (double)v[k].x * v[k].x + (double)v[k].y * v[k].y;
            em += p2;
            sp[k] += p2;
            photon += p2 / om.ld<double>(CD::voff(), CD::soff(k));
        }
        em = block_sum<T>(em, lds) * a.dt * (double)N;
        if (threadIdx.x == 0) out[m] = em;
        etot += em;
        block_fft<R, N, T, -1, PLANAR>(v, tw, lds);
#pragma unroll
        for (int k = 0; k < P; ++k) tp[k] += (double)v[k].x * v[k].x + (double)v[k].y * v[k].y;
    }
    double tmax = 0.0, tedge = 0.0, ttot = 0.0, sedge = 0.0, stot = 0.0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int q = threadIdx.x + T * k;
        const int sh = (q + N / 2) & (N - 1);
        ttot += tp[k];
        tmax = fmax(tmax, tp[k]);
        if (q < e || q >= N - e) tedge += tp[k];
        stot += sp[k];
        if (sh < e || sh >= N - e) sedge += sp[k];
    }
    ttot = block_sum<T>(ttot, lds);
    tedge = block_sum<T>(tedge, lds);
    stot = block_sum<T>(stot, lds);
    sedge = block_sum<T>(sedge, lds);
    photon = block_sum<T>(photon, lds);
    tmax = block_max<T>(tmax, lds);
    if (threadIdx.x == 0) {
        out[a.modes + 0] = etot;
        out[a.modes + 1] = tmax;
        out[a.modes + 2] = tedge / ttot;
        out[a.modes + 3] = sedge / stot;
        out[a.modes + 4] = a.dt * photon;
    }
}

// The transform must match the plan's twiddle table: single-mode plans use
// the single-mode workgroup shape, multimode plans the multimode one.
template <typename R, int N> hipError_t metrics_go(const MetricArgs<R>& a, int count, hipStream_t s)
{
    constexpr int TS = SingleShape<R, N>::T, TM = MultiShape<R, N>::T;
    if (a.modes == 1 || TS == TM) {
        hipLaunchKernelGGL((k_field_metrics<R, N, TS>), dim3(count), dim3(TS), 0, s, a);
    } else {
        hipLaunchKernelGGL((k_field_metrics<R, N, TM>), dim3(count), dim3(TM), 0, s, a);
    }
    return hipGetLastError();
}

template <typename R>
hipError_t launch_field_metrics_t(int nt, const void* spec, const void* tw, const double* omega, double dt, int modes,
                                  double* out, int count, hipStream_t s)
{
    MetricArgs<R> a{(const C2<R>*)spec, (const C2<R>*)tw, omega, dt, modes, out};
    switch (nt) {
    case 64: return metrics_go<R, 64>(a, count, s);
    case 128: return metrics_go<R, 128>(a, count, s);
    case 256: return metrics_go<R, 256>(a, count, s);
    case 512: return metrics_go<R, 512>(a, count, s);
    case 1024: return metrics_go<R, 1024>(a, count, s);
    case 2048: return metrics_go<R, 2048>(a, count, s);
    case 4096: return metrics_go<R, 4096>(a, count, s);
    case 8192: return metrics_go<R, 8192>(a, count, s);
    case 16384: return metrics_go<R, 16384>(a, count, s);
    case 32768:
        if constexpr (sizeof(R) == 4) return metrics_go<R, 32768>(a, count, s);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_field_metrics(int nt, bool fp64, const void* spec, const void* tw, const double* omega, double dt,
                                int modes, double* out, int count, hipStream_t s)
{
    return fp64 ? launch_field_metrics_t<double>(nt, spec, tw, omega, dt, modes, out, count, s)
                : launch_field_metrics_t<float>(nt, spec, tw, omega, dt, modes, out, count, s);
}

}  // namespace fh
