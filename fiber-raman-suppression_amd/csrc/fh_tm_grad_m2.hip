// Adaptive multimode gradient kernels for M = 2 (see fh_tsit5_multi_inst.hpp).
#include "fh_tsit5_multi_inst.hpp"

namespace fh {
namespace tm {

#if FH_MM_PART != 32
hipError_t launch_grad_m2(int nt, const MtArgs<double>& a, int batch, hipStream_t s)
{
    return by_size<double, 2, true>(nt, a, batch, s);
}
#endif

#if FH_MM_PART != 64
hipError_t launch_grad_m2(int nt, const MtArgs<float>& a, int batch, hipStream_t s)
{
    return by_size<float, 2, true>(nt, a, batch, s);
}
#endif

}  // namespace tm
}  // namespace fh
