// fh_tsit5.hip -- instantiations of the adaptive Tsit5 kernels (fh_tsit5.hpp)
// for single-mode grids of 256..16384 samples (FP32 also 32768).
#include <hip/hip_runtime.h>

#include "fh_launch.hpp"
#include "fh_tsit5.hpp"

namespace {

template <typename R, int N> hipError_t go(bool grad, const fh::TsArgs<R>& a, int batch, hipStream_t s)
{
    constexpr int T = fh::SingleShape<R, N>::T;
    if (grad)
        hipLaunchKernelGGL((fh::k_grad_tsit5<R, N, T>), dim3(batch), dim3(T), 0, s, a);
    else if (a.nz > 0)
        hipLaunchKernelGGL((fh::k_tsit5_single<R, N, T, true>), dim3(batch), dim3(T), 0, s, a);
    else
        hipLaunchKernelGGL((fh::k_tsit5_single<R, N, T, false>), dim3(batch), dim3(T), 0, s, a);
    return hipGetLastError();
}

template <typename R> hipError_t by_size(int nt, bool grad, const fh::TsArgs<R>& a, int batch, hipStream_t s)
{
    switch (nt) {
    case 256: return go<R, 256>(grad, a, batch, s);
    case 512: return go<R, 512>(grad, a, batch, s);
    case 1024: return go<R, 1024>(grad, a, batch, s);
    case 2048: return go<R, 2048>(grad, a, batch, s);
    case 4096: return go<R, 4096>(grad, a, batch, s);
    case 8192: return go<R, 8192>(grad, a, batch, s);
    case 16384: return go<R, 16384>(grad, a, batch, s);
    case 32768:
        if constexpr (sizeof(R) == 4) return go<R, 32768>(grad, a, batch, s);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

template <typename R, int N> int fields() { return fh::TsLayout<R, N, fh::SingleShape<R, N>::T>::NW_GRAD; }

template <typename R> int fields_by_size(int nt)
{
    switch (nt) {
    case 256: return fields<R, 256>();
    case 512: return fields<R, 512>();
    case 1024: return fields<R, 1024>();
    case 2048: return fields<R, 2048>();
    case 4096: return fields<R, 4096>();
    case 8192: return fields<R, 8192>();
    case 16384: return fields<R, 16384>();
    case 32768:
        if constexpr (sizeof(R) == 4) return fields<R, 32768>();
        return 0;
    default: return 0;
    }
}

}  // namespace

namespace fh {

bool tsit5_supported(int nt, int m, bool fp64)
{
    if (m > 1) return multi_supported(nt, m, fp64);
    // FP64 at 32768 samples runs on the multimode kernels with split transforms
    return m == 1 && nt >= 256 && nt <= 32768 && (nt & (nt - 1)) == 0;
}

int tsit5_fields(int nt, bool fp64) { return fp64 ? fields_by_size<double>(nt) : fields_by_size<float>(nt); }

hipError_t launch_tsit5(int nt, bool grad, const TsArgs<double>& a, int batch, hipStream_t s)
{
    return by_size<double>(nt, grad, a, batch, s);
}

hipError_t launch_tsit5(int nt, bool grad, const TsArgs<float>& a, int batch, hipStream_t s)
{
    return by_size<float>(nt, grad, a, batch, s);
}

}  // namespace fh
