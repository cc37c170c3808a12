// fh_tsit5_multi_inst.hpp -- per-mode-count instantiation of the adaptive
// multimode kernels.  Each fh_tm_<fwd|grad>_m<M>.hip defines one launcher
// pair from by_size, so the large kernels compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "fh_launch.hpp"
#include "fh_tsit5_multi.hpp"

namespace fh {
namespace tm {

template <typename R, int N, int M, bool GRAD> hipError_t go(const MtArgs<R>& a, int batch, hipStream_t s)
{
    if constexpr (M == 1 && !mm_split<R, N>()) {
        // one mode runs here only where the single-mode kernels have no
        // shape (split transforms)
        return hipErrorInvalidValue;
    } else {
        constexpr int T = MultiShape<R, N>::T;
        if constexpr (GRAD)
            hipLaunchKernelGGL((k_grad_tsit5_multi<R, N, T, M>), dim3(batch), dim3(T), 0, s, a);
        else
            hipLaunchKernelGGL((k_tsit5_multi<R, N, T, M>), dim3(batch), dim3(T), 0, s, a);
        return hipGetLastError();
    }
}

template <typename R, int M, bool GRAD> hipError_t by_size(int nt, const MtArgs<R>& a, int batch, hipStream_t s)
{
    switch (nt) {
    case 256: return go<R, 256, M, GRAD>(a, batch, s);
    case 512: return go<R, 512, M, GRAD>(a, batch, s);
    case 1024: return go<R, 1024, M, GRAD>(a, batch, s);
    case 2048: return go<R, 2048, M, GRAD>(a, batch, s);
    case 4096: return go<R, 4096, M, GRAD>(a, batch, s);
    case 8192: return go<R, 8192, M, GRAD>(a, batch, s);
    case 16384: return go<R, 16384, M, GRAD>(a, batch, s);
    case 32768: return go<R, 32768, M, GRAD>(a, batch, s);
    default: return hipErrorInvalidValue;
    }
}

// Launchers, one pair per unit fh_tm_<fwd|grad>_m<M>.hip.
hipError_t launch_fwd_m1(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m1(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_grad_m1(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_grad_m1(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m2(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m2(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_grad_m2(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_grad_m2(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m3(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m3(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_grad_m3(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_grad_m3(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m4(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m4(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_grad_m4(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_grad_m4(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m6(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m6(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_grad_m6(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_grad_m6(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m10(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_fwd_m10(int nt, const MtArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_grad_m10(int nt, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_grad_m10(int nt, const MtArgs<float>& a, int batch, hipStream_t s);

}  // namespace tm
}  // namespace fh
