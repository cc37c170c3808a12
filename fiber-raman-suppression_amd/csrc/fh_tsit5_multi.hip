// fh_tsit5_multi.hip -- instantiations of the adaptive multimode kernel
// (fh_tsit5_multi.hpp) for the grids and mode counts of the fixed-step
// multimode kernels.
#include <hip/hip_runtime.h>

#include "fh_launch.hpp"
#include "fh_tsit5_multi.hpp"

namespace {

template <typename R, int N, int M> hipError_t go(const fh::MtArgs<R>& a, int batch, hipStream_t s)
{
    if constexpr (sizeof(R) == 8 && M == 2) {
        return hipErrorInvalidValue;  // not offered (multi_supported)
    } else {
        constexpr int T = fh::SingleShape<R, N>::T;
        hipLaunchKernelGGL((fh::k_tsit5_multi<R, N, T, M>), dim3(batch), dim3(T), 0, s, a);
        return hipGetLastError();
    }
}

template <typename R, int N> hipError_t by_modes(int m, const fh::MtArgs<R>& a, int batch, hipStream_t s)
{
    switch (m) {
    case 2: return go<R, N, 2>(a, batch, s);
    case 3: return go<R, N, 3>(a, batch, s);
    case 4: return go<R, N, 4>(a, batch, s);
    case 6: return go<R, N, 6>(a, batch, s);
    case 10: return go<R, N, 10>(a, batch, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R> hipError_t by_size(int nt, int m, const fh::MtArgs<R>& a, int batch, hipStream_t s)
{
    switch (nt) {
    case 256: return by_modes<R, 256>(m, a, batch, s);
    case 512: return by_modes<R, 512>(m, a, batch, s);
    case 1024: return by_modes<R, 1024>(m, a, batch, s);
    case 2048: return by_modes<R, 2048>(m, a, batch, s);
    case 4096: return by_modes<R, 4096>(m, a, batch, s);
    case 8192: return by_modes<R, 8192>(m, a, batch, s);
    case 16384:
        if constexpr (sizeof(R) == 4) return by_modes<R, 16384>(m, a, batch, s);
        return hipErrorInvalidValue;
    case 32768:
        if constexpr (sizeof(R) == 4) return by_modes<R, 32768>(m, a, batch, s);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

namespace fh {

int tsit5_multi_fields(int m) { return 10 * m + (m * m + 1) / 2; }

hipError_t launch_tsit5_multi(int nt, int m, const MtArgs<double>& a, int batch, hipStream_t s)
{
    return by_size<double>(nt, m, a, batch, s);
}

hipError_t launch_tsit5_multi(int nt, int m, const MtArgs<float>& a, int batch, hipStream_t s)
{
    return by_size<float>(nt, m, a, batch, s);
}

}  // namespace fh
