// fh_tsit5_multi.hip -- dispatch of the adaptive multimode kernels
// (fh_tsit5_multi.hpp) over the mode counts of the fixed-step multimode
// kernels; the instantiations live in one unit per mode count and kind
// (fh_tm_<fwd|grad>_m<M>.hip) so that they compile in parallel.
#include <hip/hip_runtime.h>

#include "fh_launch.hpp"
#include "fh_tsit5_multi_inst.hpp"

namespace {

template <typename R> hipError_t dispatch(int nt, int m, bool grad, const fh::MtArgs<R>& a, int batch, hipStream_t s)
{
    using namespace fh::tm;
    switch (m) {
    case 1: return grad ? launch_grad_m1(nt, a, batch, s) : launch_fwd_m1(nt, a, batch, s);
    case 2: return grad ? launch_grad_m2(nt, a, batch, s) : launch_fwd_m2(nt, a, batch, s);
    case 3: return grad ? launch_grad_m3(nt, a, batch, s) : launch_fwd_m3(nt, a, batch, s);
    case 4: return grad ? launch_grad_m4(nt, a, batch, s) : launch_fwd_m4(nt, a, batch, s);
    case 6: return grad ? launch_grad_m6(nt, a, batch, s) : launch_fwd_m6(nt, a, batch, s);
    case 10: return grad ? launch_grad_m10(nt, a, batch, s) : launch_fwd_m10(nt, a, batch, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

namespace fh {

// Slice fields of the adaptive multimode kernels (MtLayout, MtgLayout), plus
// the scratch of split transforms after them.
int tsit5_multi_fields(int nt, int m, bool fp64)
{
    return 10 * m + (m * m + 1) / 2 + (mm_split_rt(nt, fp64) ? 2 : 0);
}

int tsit5_multi_grad_fields(int nt, int m, bool fp64)
{
    return 15 * m + 2 * ((m * m + 1) / 2) + (mm_split_rt(nt, fp64) ? 4 : 0);
}

hipError_t launch_tsit5_multi(int nt, int m, bool grad, const MtArgs<double>& a, int batch, hipStream_t s)
{
    return dispatch<double>(nt, m, grad, a, batch, s);
}

hipError_t launch_tsit5_multi(int nt, int m, bool grad, const MtArgs<float>& a, int batch, hipStream_t s)
{
    return dispatch<float>(nt, m, grad, a, batch, s);
}

}  // namespace fh
