// fh_multi.hip -- instantiations of the multimode propagation and gradient
// kernels (kept in their own translation unit so they compile in parallel
// with the single-mode ones).
#include <hip/hip_runtime.h>

#include "fh_launch.hpp"
#include "fh_multi.hpp"

namespace {

template <typename R, int N, int M> hipError_t go(bool grad, const fh::MmArgs<R>& a, int batch, hipStream_t s)
{
    if constexpr (M == 1 && !fh::mm_split<R, N>()) {
        // one mode runs here only where the single-mode kernels have no
        // shape (split transforms)
        return hipErrorInvalidValue;
    } else {
        constexpr int T = fh::MultiShape<R, N>::T;
        if (grad)
            hipLaunchKernelGGL((fh::k_grad_multi<R, N, T, M>), dim3(batch), dim3(T), 0, s, a);
        else
            hipLaunchKernelGGL((fh::k_rk4ip_multi<R, N, T, M>), dim3(batch), dim3(T), 0, s, a);
        return hipGetLastError();
    }
}

template <typename R, int N> hipError_t by_modes(bool grad, int m, const fh::MmArgs<R>& a, int batch, hipStream_t s)
{
    switch (m) {
    case 1: return go<R, N, 1>(grad, a, batch, s);
    case 2: return go<R, N, 2>(grad, a, batch, s);
    case 3: return go<R, N, 3>(grad, a, batch, s);
    case 4: return go<R, N, 4>(grad, a, batch, s);
    case 6: return go<R, N, 6>(grad, a, batch, s);
    case 10: return go<R, N, 10>(grad, a, batch, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R> hipError_t by_size(bool grad, int nt, int m, const fh::MmArgs<R>& a, int batch, hipStream_t s)
{
    switch (nt) {
    case 256: return by_modes<R, 256>(grad, m, a, batch, s);
    case 512: return by_modes<R, 512>(grad, m, a, batch, s);
    case 1024: return by_modes<R, 1024>(grad, m, a, batch, s);
    case 2048: return by_modes<R, 2048>(grad, m, a, batch, s);
    case 4096: return by_modes<R, 4096>(grad, m, a, batch, s);
    case 8192: return by_modes<R, 8192>(grad, m, a, batch, s);
    case 16384: return by_modes<R, 16384>(grad, m, a, batch, s);
    case 32768: return by_modes<R, 32768>(grad, m, a, batch, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

namespace fh {

#if FH_MM_PART != 64
bool multi_supported(int nt, int m, bool fp64)
{
    // One mode runs on these kernels only where transforms are split.
    const bool mok = m == 2 || m == 3 || m == 4 || m == 6 || m == 10 || (m == 1 && mm_split_rt(nt, fp64));
    const bool nok = nt >= 256 && nt <= 32768 && (nt & (nt - 1)) == 0;
    return mok && nok;
}

// Slice fields of the fixed-step multimode kernels (MmTraits::FIELDS,
// FIELDS_GRAD), including the scratch of split transforms.
int multi_fields(int nt, int m, bool fp64) { return 4 * m + (m * m + 1) / 2 + (mm_split_rt(nt, fp64) ? 2 : 0); }

int multi_grad_fields(int nt, int m, bool fp64)
{
    return 9 * m + 2 * ((m * m + 1) / 2) + (mm_split_rt(nt, fp64) ? 2 : 0);
}

hipError_t launch_multi(int nt, int m, const MmArgs<float>& a, int batch, hipStream_t s)
{
    return by_size<float>(false, nt, m, a, batch, s);
}

hipError_t launch_multi_grad(int nt, int m, const MmArgs<float>& a, int batch, hipStream_t s)
{
    return by_size<float>(true, nt, m, a, batch, s);
}
#endif

#if FH_MM_PART != 32
hipError_t launch_multi(int nt, int m, const MmArgs<double>& a, int batch, hipStream_t s)
{
    return by_size<double>(false, nt, m, a, batch, s);
}

hipError_t launch_multi_grad(int nt, int m, const MmArgs<double>& a, int batch, hipStream_t s)
{
    return by_size<double>(true, nt, m, a, batch, s);
}
#endif

}  // namespace fh
