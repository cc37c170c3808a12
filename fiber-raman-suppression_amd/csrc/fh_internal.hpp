// fh_internal.hpp -- definitions shared by the ABI units of libfiberhip
// (fiberhip.hip, fh_server.hip); not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
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
// Device memory taken with one hipMalloc and handed out to plans first-fit
// (256-byte granules, free neighbours coalesced).  Plans created on an arena
// allocate and release without hipMalloc/hipFree, which under load can take
// seconds and (hipFree) wait for every running launch.
// With `host`, the arena is pinned host memory (hipHostMalloc): the staging
// of plan tables, so their uploads are asynchronous.
struct Arena {
    std::mutex mu;
    char* base = nullptr;
    size_t size = 0;
    bool host = false;
    std::map<size_t, size_t> free_;  // offset -> bytes
    int init(size_t bytes, bool pinned_host = false);
    void destroy();
    void* alloc(size_t bytes);
    void release(void* p, size_t bytes);
};
// fh_plan_create with the table uploads on copy_stream (null: the context
// stream) and device buffers from `arena` (null: hipMalloc).  With a pinned
// host arena `stage`, the tables are staged there and their uploads are left
// in flight on copy_stream (work enqueued after them on that stream sees the
// tables); the staging is released with the plan.
int plan_create_on(fh_context ctx, const fh_problem_desc* desc, hipStream_t copy_stream, Arena* arena, Arena* stage,
                   fh_plan* out);
int enqueue_batch(fh_plan p, int batch, bool grad, const double* d_lengths, const int* d_nsteps, hipStream_t s,
                  void* h_uL, double* h_grad, double* h_cost, int* h_status, long long* h_stats);
}  // namespace fhi
