// Adaptive multimode forward kernels for M = 4 (see fh_tsit5_multi_inst.hpp).
#include "fh_tsit5_multi_inst.hpp"

namespace fh {
namespace tm {

#if FH_MM_PART != 32
hipError_t launch_fwd_m4(int nt, const MtArgs<double>& a, int batch, hipStream_t s)
{
    return by_size<double, 4, false>(nt, a, batch, s);
}
#endif

#if FH_MM_PART != 64
hipError_t launch_fwd_m4(int nt, const MtArgs<float>& a, int batch, hipStream_t s)
{
    return by_size<float, 4, false>(nt, a, batch, s);
}
#endif

}  // namespace tm
}  // namespace fh
