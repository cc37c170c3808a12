// fh_internal.hpp -- definitions shared by the ABI units of libfiberhip
// (fiberhip.hip, fh_server.hip); not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/fiberhip.h"

struct fh_context_s {
    int device;
    hipStream_t stream;
};

namespace fhi {
bool plan_fp64(fh_plan p);
int plan_nt(fh_plan p);
int plan_modes(fh_plan p);
int plan_max_batch(fh_plan p);
bool plan_adaptive(fh_plan p);
void* plan_u_in(fh_plan p);
void* plan_u_out(fh_plan p);
int64_t plan_bytes(fh_plan p);
int set_error(int code, const std::string& msg);
int plan_prepare_gradient(fh_plan p);
int enqueue_batch(fh_plan p, int batch, bool grad, const double* d_lengths, const int* d_nsteps, hipStream_t s,
                  void* h_uL, double* h_grad, double* h_cost, int* h_status, long long* h_stats);
}  // namespace fhi
