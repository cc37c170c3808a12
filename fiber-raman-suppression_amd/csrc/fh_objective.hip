// fh_objective.hip -- terminal objectives and physical gradients on device
// fields, with FP64 reductions in a fixed order.
//
// Objectives (the terminal cost and its Wirtinger seed dJ/dconj(uL) that the
// backward sweep starts from):
//   FH_OBJ_BAND_SUM        J = sum_m E_band_m / sum_m E_m,
//                          seed_m = u_m (mask - J) / sum_m E_m
//                          (_field_spectral_band_cost, physics_helpers.jl:222-232;
//                          mmf_cost_sum, mmf_cost.jl:49-62)
//   FH_OBJ_FUNDAMENTAL     J = E_band_1 / E_1, seed only on mode 1
//                          (mmf_cost_fundamental, mmf_cost.jl:81-98)
//   FH_OBJ_WORST_MODE      J = r_max + (log sum_active exp(tau (r_m - r_max)) - log K) / tau,
//                          r_m = E_band_m / E_m over the K modes with E_m > 0,
//                          seed_m = w_m u_m (mask - r_m) / E_m, w = softmax(tau r)
//                          (mmf_cost_worst_mode, mmf_cost.jl:124-166)
// With log_cost the cost is 10 log10(max(J, 1e-15)) and the seed is scaled by
// 10 / (ln 10 J) (0 when J <= 1e-15), _maybe_log_cost of physics_helpers.jl:320-326.
//
// Physical gradients of the spectral-shaper launch alpha A cis(phi) uw0
// (spectral_shaper_model, physics_models.jl:1042-1047) from lam0 and the
// shaped launch u0 = alpha A cis(phi) uw0:
//   phase     2 Re(conj(lam0) i u0)
//   amplitude 2 Re(conj(lam0) u0)          (the caller divides by A)
//   energy    Re sum conj(lam0) u0         (the caller divides by the energy)
#include <hip/hip_runtime.h>

#include "../../include/fiberhip.h"
#include "fh_device.hpp"
#include "fh_launch.hpp"

namespace fh {
namespace {

constexpr int OT = 256;

// Workgroup sum of one value per thread, fixed tree order.
__device__ double wg_sum(double x, double* red)
{
    const int tid = threadIdx.x;
    __syncthreads();
    red[tid] = x;
    __syncthreads();
#pragma unroll 1
    for (int w = OT / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

template <typename R>
__global__ __launch_bounds__(OT) void k_objective(const C2<R>* __restrict__ uL, const unsigned char* __restrict__ mask,
                                                  int nt, int modes, int kind, double tau, int log_cost,
                                                  C2<R>* __restrict__ seed, double* __restrict__ cost)
{
    __shared__ double red[OT];
    __shared__ double et[kObjectiveMaxModes], eb[kObjectiveMaxModes], rm[kObjectiveMaxModes],
        cm[kObjectiveMaxModes];
    const int tid = threadIdx.x;
    const size_t sol = blockIdx.x;
    const C2<R>* u = uL + sol * (size_t)modes * nt;
    C2<R>* sd = seed + sol * (size_t)modes * nt;
#pragma unroll 1
    for (int m = 0; m < modes; ++m) {
        double st = 0.0, sb = 0.0;
        for (int t = tid; t < nt; t += OT) {
            const C2<R> x = u[(size_t)m * nt + t];
            const double e = (double)x.x * x.x + (double)x.y * x.y;
            st += e;
            if (mask[t]) sb += e;
        }
        st = wg_sum(st, red);
        sb = wg_sum(sb, red);
        if (tid == 0) {
            et[m] = st;
            eb[m] = sb;
        }
    }
    __syncthreads();
    if (tid == 0) {
        double J = 0.0;
        if (kind == FH_OBJ_FUNDAMENTAL) {
            J = eb[0] / et[0];
            for (int m = 0; m < modes; ++m) {
                rm[m] = J;
                cm[m] = m == 0 ? 1.0 / et[0] : 0.0;
            }
        } else if (kind == FH_OBJ_WORST_MODE) {
            double rmax = -1.0;
            int active = 0;
            for (int m = 0; m < modes; ++m) {
                rm[m] = et[m] > 0 ? eb[m] / et[m] : 0.0;
                if (et[m] > 0) {
                    rmax = active ? fmax(rmax, rm[m]) : rm[m];
                    ++active;
                }
            }
            double denom = 0.0;
            for (int m = 0; m < modes; ++m)
                if (et[m] > 0) denom += exp(tau * (rm[m] - rmax));
            J = rmax + (log(denom) - log((double)active)) / tau;
            for (int m = 0; m < modes; ++m) cm[m] = et[m] > 0 ? exp(tau * (rm[m] - rmax)) / denom / et[m] : 0.0;
        } else {
            double st = 0.0, sb = 0.0;
            for (int m = 0; m < modes; ++m) {
                st += et[m];
                sb += eb[m];
            }
            J = sb / st;
            for (int m = 0; m < modes; ++m) {
                rm[m] = J;
                cm[m] = 1.0 / st;
            }
        }
        if (log_cost) {
            const double s = J <= 1e-15 ? 0.0 : 10.0 / (2.302585092994045684 * J);
            for (int m = 0; m < modes; ++m) cm[m] *= s;
            J = 10.0 * log10(fmax(J, 1e-15));
        }
        cost[sol] = J;
    }
    __syncthreads();
#pragma unroll 1
    for (int m = 0; m < modes; ++m) {
        const double r = rm[m], c = cm[m];
        for (int t = tid; t < nt; t += OT) {
            const C2<R> x = u[(size_t)m * nt + t];
            const R f = (R)(((mask[t] ? 1.0 : 0.0) - r) * c);
            sd[(size_t)m * nt + t] = C2<R>{x.x * f, x.y * f};
        }
    }
}

template <typename R>
__global__ __launch_bounds__(OT) void k_physical_gradients(const C2<R>* __restrict__ lam0, const C2<R>* __restrict__ u0,
                                                           int n, double* __restrict__ gphase,
                                                           double* __restrict__ gamp, double* __restrict__ edot)
{
    __shared__ double red[OT];
    const size_t base = (size_t)blockIdx.x * n;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += OT) {
        const C2<R> l = lam0[base + i], u = u0[base + i];
        const double re = (double)l.x * u.x + (double)l.y * u.y;
        if (gphase) gphase[base + i] = 2.0 * ((double)l.y * u.x - (double)l.x * u.y);
        if (gamp) gamp[base + i] = 2.0 * re;
        acc += re;
    }
    acc = wg_sum(acc, red);
    if (edot && threadIdx.x == 0) edot[blockIdx.x] = acc;
}

}  // namespace

hipError_t launch_objective(bool fp64, const void* uL, const unsigned char* mask, int nt, int modes, int kind,
                            double tau, int log_cost, void* seed, double* cost, int batch, hipStream_t s)
{
    if (modes < 1 || modes > kObjectiveMaxModes) return hipErrorInvalidValue;
    if (fp64)
        hipLaunchKernelGGL(k_objective<double>, dim3(batch), dim3(OT), 0, s, (const C2<double>*)uL, mask, nt, modes,
                           kind, tau, log_cost, (C2<double>*)seed, cost);
    else
        hipLaunchKernelGGL(k_objective<float>, dim3(batch), dim3(OT), 0, s, (const C2<float>*)uL, mask, nt, modes,
                           kind, tau, log_cost, (C2<float>*)seed, cost);
    return hipGetLastError();
}

hipError_t launch_physical_gradients(bool fp64, const void* lam0, const void* u0, int n, double* gphase,
                                     double* gamp, double* edot, int batch, hipStream_t s)
{
    if (fp64)
        hipLaunchKernelGGL(k_physical_gradients<double>, dim3(batch), dim3(OT), 0, s, (const C2<double>*)lam0,
                           (const C2<double>*)u0, n, gphase, gamp, edot);
    else
        hipLaunchKernelGGL(k_physical_gradients<float>, dim3(batch), dim3(OT), 0, s, (const C2<float>*)lam0,
                           (const C2<float>*)u0, n, gphase, gamp, edot);
    return hipGetLastError();
}

}  // namespace fh
