// fh_gain.hip -- instantiations of the pump + signal (gain) Tsit5 kernel
// (k_tsit5_gain, fh_tsit5.hpp) for single-mode grids of 256..16384 samples.
#include <hip/hip_runtime.h>

#include "fh_launch.hpp"
#include "fh_tsit5.hpp"

namespace {

template <typename R, int N> hipError_t go(const fh::TsArgs<R>& a, int batch, hipStream_t s)
{
    constexpr int T = fh::SingleShape<R, N>::T;
    hipLaunchKernelGGL((fh::k_tsit5_gain<R, N, T>), dim3(batch), dim3(T), 0, s, a);
    return hipGetLastError();
}

template <typename R> hipError_t by_size(int nt, const fh::TsArgs<R>& a, int batch, hipStream_t s)
{
    switch (nt) {
    case 256: return go<R, 256>(a, batch, s);
    case 512: return go<R, 512>(a, batch, s);
    case 1024: return go<R, 1024>(a, batch, s);
    case 2048: return go<R, 2048>(a, batch, s);
    case 4096: return go<R, 4096>(a, batch, s);
    case 8192: return go<R, 8192>(a, batch, s);
    case 16384: return go<R, 16384>(a, batch, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R, int N> int fields() { return fh::TsLayout<R, N, fh::SingleShape<R, N>::T>::NW_FWD + 1; }

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
    default: return 0;
    }
}

}  // namespace

namespace fh {

bool gain_supported(int nt) { return nt >= 256 && nt <= 16384 && (nt & (nt - 1)) == 0; }

int gain_fields(int nt, bool fp64) { return fp64 ? fields_by_size<double>(nt) : fields_by_size<float>(nt); }

hipError_t launch_tsit5_gain(int nt, const TsArgs<double>& a, int batch, hipStream_t s)
{
    return by_size<double>(nt, a, batch, s);
}

hipError_t launch_tsit5_gain(int nt, const TsArgs<float>& a, int batch, hipStream_t s)
{
    return by_size<float>(nt, a, batch, s);
}

}  // namespace fh
