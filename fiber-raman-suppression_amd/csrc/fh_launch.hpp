// fh_launch.hpp -- launch shapes and cross-unit launchers.
#pragma once
#include <hip/hip_runtime.h>

namespace fh {

template <typename R> struct MmArgs;

// Workgroup size per precision and transform length: 8 samples per thread
// for FP64 (1024 threads from 8192 samples up) keeps the radix-8 working set
// in registers; FP32 uses 16 from 4096 samples up.
// Units that instantiate multimode kernels are built once per precision
// (FH_MM_PART = 32 or 64); without the define a unit builds both.  The FP64
// part is compiled with every memory wait forced to zero: without it the
// FP64 multimode kernels returned results that differed between identical
// solves of one launch.
#ifndef FH_MM_PART
#define FH_MM_PART 0
#endif

template <typename R, int N> struct SingleShape {
    static constexpr int T = sizeof(R) == 8 ? (N >= 8192 ? 1024 : (N >= 256 ? N / 8 : N / 4))
                                            : (N >= 32768 ? 1024 : (N >= 4096 ? N / 16 : (N >= 512 ? N / 8 : N / 4)));
};

// Multimode kernels: FP32 at 4096 samples takes 8 samples per thread (512
// threads, twice the single-mode workgroup).  A multimode solve keeps its
// per-sample work in global scratch and needs more waves per CU to hide that
// latency; 512 threads still leave 256 VGPRs per wave for the M = 10
// contraction.  Other sizes and FP64 use the single-mode shape.
template <typename R, int N> struct MultiShape {
    static constexpr int T = (sizeof(R) == 4 && N == 4096) ? 512 : SingleShape<R, N>::T;
};

template <typename R> struct TsArgs;
bool tsit5_supported(int nt, int m, bool fp64);
int tsit5_fields(int nt, bool fp64);
hipError_t launch_tsit5(int nt, bool grad, const TsArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_tsit5(int nt, bool grad, const TsArgs<float>& a, int batch, hipStream_t s);

// Pump + signal (gain) solves (fh_gain.hip): single mode, 256..16384 samples.
bool gain_supported(int nt);
int gain_fields(int nt, bool fp64);
hipError_t launch_tsit5_gain(int nt, const TsArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_tsit5_gain(int nt, const TsArgs<float>& a, int batch, hipStream_t s);

// Terminal objectives and physical gradients on device fields (fh_objective.hip).
constexpr int kObjectiveMaxModes = 16;
hipError_t launch_objective(bool fp64, const void* uL, const unsigned char* mask, int nt, int modes, int kind,
                            double tau, int log_cost, void* seed, double* cost, int batch, hipStream_t s);
hipError_t launch_physical_gradients(bool fp64, const void* lam0, const void* u0, int n, double* gphase,
                                     double* gamp, double* edot, int batch, hipStream_t s);

hipError_t launch_field_metrics(int nt, bool fp64, const void* spec, const void* tw, const double* omega, double dt,
                                int modes, double* out, int count, hipStream_t s);

template <typename R> struct MtArgs;
int tsit5_multi_fields(int nt, int m, bool fp64);
int tsit5_multi_grad_fields(int nt, int m, bool fp64);
hipError_t launch_tsit5_multi(int nt, int m, bool grad, const MtArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_tsit5_multi(int nt, int m, bool grad, const MtArgs<float>& a, int batch, hipStream_t s);

bool multi_supported(int nt, int m, bool fp64);
int multi_fields(int nt, int m, bool fp64);
int multi_grad_fields(int nt, int m, bool fp64);
hipError_t launch_multi(int nt, int m, const MmArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_multi(int nt, int m, const MmArgs<float>& a, int batch, hipStream_t s);
hipError_t launch_multi_grad(int nt, int m, const MmArgs<double>& a, int batch, hipStream_t s);
hipError_t launch_multi_grad(int nt, int m, const MmArgs<float>& a, int batch, hipStream_t s);

}  // namespace fh
