// fh_server.hip -- request server of libfiberhip: continuous batching of
// independent forward / forward+adjoint solves (see fh_server_* in
// include/fiberhip.h).
//
// One scheduler thread owns every HIP call of the server and releases the
// request lock while it makes them (plan creation and release, staging,
// launches, output scatter), so submissions never wait behind device work.
// Requests that
// share their operators form a batch (one fused launch, one workgroup per
// solve, each solve with its own length); up to `lanes` batches run at once
// on their own default-priority streams, each with an exclusive plan
// instance (plans hold the per-solve scratch).  Plans are created on demand,
// kept while idle for reuse, and released least-recently-used first when a
// new plan would exceed the memory budget.  Inputs are staged through pinned
// host buffers; outputs come back asynchronously on the lane's stream and are
// scattered to the callers' buffers when the lane's completion event fires.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fh_device.hpp"
#include "fh_internal.hpp"
#include "fh_launch.hpp"

namespace {

int64_t now_ns()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n)
    {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    }
    template <typename T> void val(T v) { add(&v, sizeof(v)); }
};

// Server-owned copy of one operator set (the batching key).
struct OpSet {
    int nt = 0, modes = 0;
    bool grad = false;
    int capacity = 0;  // solves per launch of this class (0 = not sized yet)
    std::vector<double> disp, gamma, raman, shock;
    std::vector<uint8_t> mask;
    double one_m_fr = 0;
};

struct Req {
    int kind = 0;
    uint64_t key = 0;
    double length = 0;
    int nsteps = 0;
    bool heavy = false;  // estimated work above kHeavyFlops
    bool solo = false;   // re-run alone after non-finite fields inside a batch
    std::vector<double> u0;
    double *uL = nullptr, *grad = nullptr, *cost = nullptr;
    int state = 0;  // 0 queued, 1 running, 2 finished
    std::string msg;  // failure message
    fh_request_result res{};
};

struct PlanInst {
    fh_plan plan = nullptr;
    uint64_t key = 0;
    int capacity = 1;
    bool busy = false;
    int64_t last_used = 0, bytes = 0;
};

// Pinned staging buffer of a lane.  Growing it retires the old buffer
// instead of freeing it: hipHostFree waits for all work on the device, which
// would stall the scheduler behind every running launch.  Retired buffers
// (at most about the final size in total, growth is geometric) are freed
// when the server is destroyed.
template <typename T> struct Pinned {
    T* p = nullptr;
    size_t n = 0;
    std::vector<T*> retired;
    int ensure(size_t want)
    {
        if (want <= n) return FH_OK;
        const size_t grow = std::max(want, 2 * n);
        T* q = nullptr;
        if (hipHostMalloc((void**)&q, sizeof(T) * grow, hipHostMallocDefault) != hipSuccess)
            return fhi::set_error(FH_ENOMEM, "hipHostMalloc of a staging buffer failed");
        if (p) retired.push_back(p);
        p = q;
        n = grow;
        return FH_OK;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        for (T* q : retired) (void)hipHostFree(q);
        retired.clear();
        p = nullptr;
        n = 0;
    }
};

// Opt-in trace (environment FH_SERVER_TRACE set), to stderr: scheduler steps
// that took longer than 20 ms, and every launch that returned a non-finite
// solve, member by member.
bool trace_on()
{
    static const bool on = std::getenv("FH_SERVER_TRACE") != nullptr;
    return on;
}

void trace_slow(const char* what, int64_t t0, int detail)
{
    if (!trace_on()) return;
    const int64_t dt = std::chrono::duration_cast<std::chrono::nanoseconds>(
                           std::chrono::steady_clock::now().time_since_epoch()).count() - t0;
    if (dt > 20000000) std::fprintf(stderr, "fh_server: %s took %.1f ms (%d)\n", what, dt / 1e6, detail);
}

struct Lane {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    bool heavy = false;  // running a batch of heavy requests
    bool grad = false;
    int modes = 1, nt = 0;
    PlanInst* pi = nullptr;
    std::vector<int64_t> ids;
    Pinned<char> h_in, h_out;
    Pinned<double> h_grad, h_cost, h_len;
    Pinned<int> h_status, h_nst;
    Pinned<long long> h_stats;
    double* d_len = nullptr;
    int* d_nst = nullptr;
    // host staging of the running batch: a block of the server's pinned
    // staging arena, or the lane's own buffers above when it does not fit
    char* io = nullptr;
    size_t io_bytes = 0;
    char* in = nullptr;
    char* out = nullptr;
    double *grad_h = nullptr, *cost_h = nullptr, *len_h = nullptr;
    int *status_h = nullptr, *nst_h = nullptr;
    long long* stats_h = nullptr;
};

}  // namespace

struct fh_server_s {
    fh_context ctx = nullptr;
    fh_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, Req> reqs;
    std::unordered_map<uint64_t, OpSet> ops;
    std::list<PlanInst> plans;
    std::vector<Lane> lanes;
    fhi::Arena arena;  // the memory budget, allocated once; plans sub-allocate from it
    fhi::Arena stage;  // pinned host staging of plan tables, so plan creation does not wait for uploads
    fhi::Arena io;     // pinned host staging of batch inputs and outputs
    int64_t next_id = 1, plan_bytes = 0;
    bool stop = false;
    std::thread worker;
};

namespace {

// Largest fixed-step count of a multimode request.
constexpr int kMaxMultimodeSteps = 1 << 22;

// A launch holds its lane until its slowest solve ends.  Requests whose
// estimated work exceeds kHeavyFlops (about ten seconds of one workgroup)
// may occupy at most lanes - lanes/4 lanes, so light requests keep lanes to
// run on while heavy ones are in flight.
constexpr double kHeavyFlops = 1.5e12;

// Pinned host staging of plan tables (a plan of 32768 samples and 10 modes
// stages about 4 MB); when it is full, uploads fall back to synchronous ones.
constexpr size_t kStageBytes = 256ull << 20;

// Pinned host staging of batch inputs and outputs: a sixteenth of the memory
// budget, between 256 MiB and 4 GiB.  A batch that does not fit stages
// through its lane's own buffers.
constexpr size_t kIoBytes = 4ull << 30;

// Algorithmic flops of a request: the disp_mmf! structure per right-hand
// side (5 N log2 N (2M + 2M^2) + 2 M^4 N + 17 M^2 N + 16 M N), times the
// fixed step count (4 per step) or, for adaptive single-mode solves, about
// 100 accepted steps per metre (7 evaluations each), times 4.5 for a
// forward + adjoint request.
double request_flops(int nt, int m, int nsteps, double length_m, bool grad)
{
    const double N = nt, M = m;
    const double rhs = 5 * N * std::log2(N) * (2 * M + 2 * M * M) + 2 * M * M * M * M * N + 17 * M * M * N + 16 * M * N;
    const double evals = m > 1 ? 4.0 * nsteps : 7.0 * std::max(20.0, 100.0 * length_m);
    return rhs * evals * (grad ? 4.5 : 1.0);
}

int mm_grad_cap(const fh_server_s* S) { return S->cfg.mm_grad_max_steps > 0 ? S->cfg.mm_grad_max_steps : 2000; }

// Finish a request; a failure keeps its message for the caller's fh_server_wait.
void finish(fh_server_s* S, int64_t id, int code, const std::string& msg = std::string())
{
    Req& r = S->reqs[id];
    r.res.retcode = code;
    r.res.finish_ns = now_ns();
    r.state = 2;
    if (code) r.msg = msg;
}

// Release idle plans, least recently used first, until `need` more bytes fit.
void evict_for(fh_server_s* S, int64_t need)
{
    while (S->plan_bytes + need > S->cfg.memory_budget_bytes) {
        auto victim = S->plans.end();
        for (auto it = S->plans.begin(); it != S->plans.end(); ++it)
            if (!it->busy && (victim == S->plans.end() || it->last_used < victim->last_used)) victim = it;
        if (victim == S->plans.end()) return;
        S->plan_bytes -= victim->bytes;
        const int64_t t0 = now_ns();
        fh_plan_destroy(victim->plan);
        trace_slow("plan release", t0, (int)(victim->bytes >> 20));
        S->plans.erase(victim);
    }
}

// An idle plan instance for `key`, created if needed.  Returns nullptr with
// *wait = true when the budget is taken by running launches.
// Called without the request lock: S->plans, S->plan_bytes and the lanes
// belong to the scheduler thread, and `op` (an element of S->ops, which
// submissions only insert into) stays valid.
PlanInst* acquire_plan(fh_server_s* S, uint64_t key, OpSet* op, hipStream_t lane_stream, int* rc, bool* wait)
{
    *rc = FH_OK;
    *wait = false;
    for (auto& pi : S->plans)
        if (pi.key == key && !pi.busy) return &pi;
    const OpSet& o = *op;
    fh_problem_desc d{};
    d.nt = o.nt;
    d.modes = o.modes;
    d.precision = S->cfg.precision;
    d.scheme = o.modes == 1 ? FH_SCHEME_TSIT5 : FH_SCHEME_RK4IP_FIXED;
    d.steps = 1;  // multimode: per-solve step counts override
    d.checkpoint_every = 50;
    if (o.modes > 1 && o.grad) {
        // step capacity; checkpoint spacing minimising checkpoints (cap/every M fields)
        // plus tape (4 every (M + QC) fields) per solve
        const int cap = mm_grad_cap(S), M = o.modes, qc = (M * (M + 1) / 2 + 1) / 2;
        d.steps = cap;
        d.checkpoint_every = std::max(1, (int)std::lround(std::sqrt((double)cap * M / (4.0 * (M + qc)))));
    }
    d.length_m = 1e-3;  // per-solve lengths override; the phase bound is checked per request
    d.dispersion = o.disp.data();
    d.gamma = o.gamma.data();
    d.raman = o.raman.data();
    d.shock = o.shock.data();
    d.one_m_fr = o.one_m_fr;
    d.band_mask = o.grad ? o.mask.data() : nullptr;
    d.rtol = S->cfg.rtol;
    d.atol = S->cfg.atol;
    d.max_steps = S->cfg.max_steps;
    if (o.modes == 1 && o.grad) d.tape_segments = S->cfg.tape_segments;
    if (op->capacity == 0) {
        // one launch of a class holds at most a sixteenth of the memory
        // budget, so plans of many classes stay resident together (releasing
        // one synchronises the device): classes with large per-solve
        // footprints (multimode gradients, FP64 at 32768 samples) get smaller
        // batches instead of no admission
        int64_t b1 = 0, b2 = 0;
        d.max_batch = 1;
        int crc = fh_plan_check(&d, o.grad, &b1);
        d.max_batch = 2;
        if (!crc) crc = fh_plan_check(&d, o.grad, &b2);
        int cap = 1;
        if (!crc && b2 > b1) {
            const int64_t per = b2 - b1, fixed = b1 - per, share = S->cfg.memory_budget_bytes / 16;
            cap = (int)std::max<int64_t>(1, std::min<int64_t>(S->cfg.max_batch, (share - fixed) / per));
        } else if (!crc) {
            cap = S->cfg.max_batch;
        }
        op->capacity = cap;
    }
    d.max_batch = op->capacity;
    int64_t bytes = 0;
    if ((*rc = fh_plan_check(&d, o.grad, &bytes))) return nullptr;
    if (bytes > S->cfg.memory_budget_bytes) {
        *rc = fhi::set_error(FH_ENOMEM, "one launch of this request class needs " + std::to_string(bytes) +
                                            " device bytes, above the server's memory budget");
        return nullptr;
    }
    auto any_busy = [&] {
        for (auto& l : S->lanes)
            if (l.busy) return true;
        return false;
    };
    evict_for(S, bytes);
    if (S->plan_bytes + bytes > S->cfg.memory_budget_bytes) {
        if (any_busy()) *wait = true;  // running launches hold the rest of the budget
        else *rc = fhi::set_error(FH_ENOMEM, "the server's memory budget is held by plans");
        return nullptr;
    }
    fh_plan p = nullptr;
    for (int attempt = 0; attempt < 2; ++attempt) {
        const int64_t tp = now_ns();
        *rc = fhi::plan_create_on(S->ctx, &d, lane_stream, &S->arena, &S->stage, &p);
        if (*rc == FH_OK && o.grad) *rc = fhi::plan_prepare_gradient(p);
        trace_slow("plan creation", tp, d.max_batch);
        if (*rc == FH_OK) break;
        if (p) fh_plan_destroy(p);
        p = nullptr;
        if (*rc != FH_ENOMEM) return nullptr;
        // the arena is fragmented: release every idle plan and try once more
        evict_for(S, S->cfg.memory_budget_bytes + 1);
    }
    if (!p) {
        if (any_busy()) {  // fragmentation: retry after running work drains
            *wait = true;
            *rc = FH_OK;
        }
        return nullptr;
    }
    bytes = fhi::plan_bytes(p);
    S->plans.push_back(PlanInst{p, key, d.max_batch, false, now_ns(), bytes});
    S->plan_bytes += bytes;
    return &S->plans.back();
}

// Start one batch on an idle lane.  Entered and left with the request lock
// held; released around the HIP calls.  Returns true if work was dispatched
// or failed requests were finished, false when it has to wait.
bool dispatch(fh_server_s* S, Lane& L, std::unique_lock<std::mutex>& lk)
{
    // the oldest request that may run now: heavy requests only while fewer
    // than lanes - lanes/4 lanes run heavy batches
    int heavy_busy = 0;
    for (const auto& l : S->lanes) heavy_busy += l.busy && l.heavy;
    const int heavy_cap = std::max(1, (int)S->lanes.size() - (int)S->lanes.size() / 4);
    const Req* first = nullptr;
    int64_t first_id = -1;
    for (int64_t id : S->queue) {
        const Req& r = S->reqs[id];
        if (!r.heavy || heavy_busy < heavy_cap) {
            first = &r;
            first_id = id;
            break;
        }
    }
    if (!first) return false;
    const uint64_t key = first->key;
    const bool heavy = first->heavy;
    const bool solo = first->solo;
    OpSet* op = &S->ops.at(key);
    int rc;
    bool wait;
    lk.unlock();
    const int64_t ta = now_ns();
    PlanInst* pi = acquire_plan(S, key, op, L.s, &rc, &wait);
    trace_slow("acquire_plan", ta, op->nt * op->modes);
    lk.lock();
    if (wait) return false;
    // only this thread removes from the queue: the class's requests are still there
    std::vector<int64_t> ids;
    std::vector<Req*> rq;
    const int capacity = solo ? 1 : (pi ? pi->capacity : S->cfg.max_batch);
    for (auto it = S->queue.begin(); it != S->queue.end() && (int)ids.size() < capacity;) {
        const Req& c = S->reqs[*it];
        if (c.key == key && c.heavy == heavy && c.solo == solo && (!solo || *it == first_id)) {
            ids.push_back(*it);
            rq.push_back(&S->reqs[*it]);
            it = S->queue.erase(it);
        } else {
            ++it;
        }
    }
    if (!pi) {
        const std::string m = fh_last_error();
        for (int64_t id : ids) finish(S, id, rc, "class setup: " + m);
        S->cv_done.notify_all();
        return true;
    }
    const int64_t t0 = now_ns();
    for (Req* r : rq) {
        r->state = 1;
        r->res.start_ns = t0;
        r->res.batch = (int)ids.size();
        r->res.attempts += 1;
    }
    // requests in state 1 are neither erased nor modified by other threads
    lk.unlock();
    const OpSet& o = *op;
    const int B = (int)ids.size(), nf = o.nt * o.modes;
    const int64_t ts = now_ns();
    const bool fp64 = S->cfg.precision == FH_PREC_FP64;
    const size_t es = fp64 ? sizeof(double) : sizeof(float);
    int e = FH_OK;
    {
        // host staging from the pinned io arena (no allocation while serving);
        // the lane's own buffers only when the block does not fit
        const size_t a256 = 256;
        auto up = [&](size_t n) { return (n + a256 - 1) / a256 * a256; };
        const size_t s_f = up((size_t)B * nf * 2 * es), s_g = up(sizeof(double) * (size_t)B * o.nt),
                     s_d = up(sizeof(double) * B), s_i = up(sizeof(int) * B), s_l = up(sizeof(long long) * 3 * B);
        const size_t total = 2 * s_f + s_g + 2 * s_d + 2 * s_i + s_l;
        char* blk = (char*)S->io.alloc(total);
        L.io = blk;
        L.io_bytes = blk ? total : 0;
        if (blk) {
            char* q = blk;
            L.in = q; q += s_f;
            L.out = q; q += s_f;
            L.grad_h = (double*)q; q += s_g;
            L.cost_h = (double*)q; q += s_d;
            L.len_h = (double*)q; q += s_d;
            L.status_h = (int*)q; q += s_i;
            L.nst_h = (int*)q; q += s_i;
            L.stats_h = (long long*)q;
        } else {
            if (!e) e = L.h_in.ensure((size_t)B * nf * 2 * es);
            if (!e) e = L.h_out.ensure((size_t)B * nf * 2 * es);
            if (!e) e = L.h_grad.ensure((size_t)B * o.nt);
            if (!e) e = L.h_cost.ensure(B);
            if (!e) e = L.h_len.ensure(B);
            if (!e) e = L.h_status.ensure(B);
            if (!e) e = L.h_nst.ensure(B);
            if (!e) e = L.h_stats.ensure(3 * (size_t)B);
            L.in = L.h_in.p;
            L.out = L.h_out.p;
            L.grad_h = L.h_grad.p;
            L.cost_h = L.h_cost.p;
            L.len_h = L.h_len.p;
            L.status_h = L.h_status.p;
            L.nst_h = L.h_nst.p;
            L.stats_h = L.h_stats.p;
        }
    }
    for (int b = 0; e == FH_OK && b < B; ++b) {
        const Req& r = *rq[b];
        if (fp64) {
            std::memcpy(L.in + (size_t)b * nf * 2 * es, r.u0.data(), (size_t)nf * 2 * es);
        } else {
            float* dst = (float*)(L.in + (size_t)b * nf * 2 * es);
            for (int i = 0; i < 2 * nf; ++i) dst[i] = (float)r.u0[i];
        }
        L.len_h[b] = r.length;
        L.nst_h[b] = r.nsteps;
    }
    if (e == FH_OK && hipMemcpyAsync(fhi::plan_u_in(pi->plan), L.in, (size_t)B * nf * 2 * es,
                                     hipMemcpyHostToDevice, L.s) != hipSuccess)
        e = fhi::set_error(FH_EDEVICE, "staging copy failed");
    if (e == FH_OK && (hipMemcpyAsync(L.d_len, L.len_h, sizeof(double) * B, hipMemcpyHostToDevice, L.s) != hipSuccess ||
                       hipMemcpyAsync(L.d_nst, L.nst_h, sizeof(int) * B, hipMemcpyHostToDevice, L.s) != hipSuccess))
        e = fhi::set_error(FH_EDEVICE, "staging copy failed");
    if (e == FH_OK)
        e = fhi::enqueue_batch(pi->plan, B, o.grad, L.d_len, o.modes > 1 ? L.d_nst : nullptr, L.s, L.out,
                               L.grad_h, L.cost_h, L.status_h, L.stats_h);
    if (e == FH_OK && hipEventRecord(L.ev, L.s) != hipSuccess) e = fhi::set_error(FH_EDEVICE, "hipEventRecord failed");
    trace_slow("staging and launch", ts, B);
    if (e != FH_OK) {
        (void)hipStreamSynchronize(L.s);
        if (L.io) S->io.release(L.io, L.io_bytes);
        L.io = nullptr;
        pi->last_used = now_ns();
        lk.lock();
        const std::string m = fh_last_error();
        for (int64_t id : ids) finish(S, id, e, "launch: " + m);
        S->cv_done.notify_all();
        return true;
    }
    pi->busy = true;
    L.busy = true;
    L.heavy = heavy;
    L.pi = pi;
    L.ids = std::move(ids);
    L.grad = o.grad;
    L.modes = o.modes;
    L.nt = o.nt;
    lk.lock();
    return true;
}

// Scatter a finished lane's outputs to the callers' buffers (without the
// request lock: the requests are in state 1) and finish them.  Entered and
// left with the lock held.
void complete(fh_server_s* S, Lane& L, int device_rc, std::unique_lock<std::mutex>& lk)
{
    std::vector<Req*> rq;
    for (int64_t id : L.ids) rq.push_back(&S->reqs[id]);
    lk.unlock();
    const int64_t tc = now_ns();
    const bool fp64 = S->cfg.precision == FH_PREC_FP64;
    const size_t es = fp64 ? sizeof(double) : sizeof(float);
    const int nf = L.nt * L.modes;
    const bool adaptive = fhi::plan_adaptive(L.pi->plan);
    std::vector<int> codes(rq.size(), device_rc);
    for (size_t b = 0; b < rq.size(); ++b) {
        Req& r = *rq[b];
        if (codes[b] == FH_OK) {
            const int st = L.status_h[b];
            codes[b] = (st & 2) ? FH_EMAXSTEPS : (st ? FH_ENONFINITE : FH_OK);
            if (fp64) {
                std::memcpy(r.uL, L.out + b * nf * 2 * es, (size_t)nf * 2 * es);
            } else {
                const float* src = (const float*)(L.out + b * nf * 2 * es);
                for (int i = 0; i < 2 * nf; ++i) r.uL[i] = src[i];
            }
            if (L.grad) {
                if (r.grad) std::memcpy(r.grad, L.grad_h + b * L.nt, sizeof(double) * L.nt);
                if (r.cost) r.cost[0] = L.cost_h[b];
            }
            if (adaptive) {
                r.res.accepted = L.stats_h[3 * b];
                r.res.rejected = L.stats_h[3 * b + 1];
                r.res.rhs_evals = L.stats_h[3 * b + 2];
            } else {
                r.res.accepted = r.nsteps;
                r.res.rhs_evals = 4LL * r.nsteps;
            }
        }
    }
    if (trace_on()) {
        // one line per launch that returned a non-finite solve: every member's
        // status, length and count of non-finite output entries
        bool any = false;
        for (int c : codes) any |= c == FH_ENONFINITE;
        if (any) {
            std::string line = "fh_server: launch of " + std::to_string(rq.size()) + " (nt " + std::to_string(L.nt) +
                               ", modes " + std::to_string(L.modes) + (L.grad ? ", gradient" : ", forward") + "):";
            for (size_t b = 0; b < rq.size(); ++b) {
                const Req& r = *rq[b];
                int bad_u = 0, bad_g = 0;
                for (int i = 0; i < 2 * nf; ++i) bad_u += !std::isfinite(r.uL[i]);
                if (L.grad && r.grad)
                    for (int i = 0; i < L.nt; ++i) bad_g += !std::isfinite(r.grad[i]);
                char buf[160];
                std::snprintf(buf, sizeof buf, " [id %lld L %.3f status %d uL-nonfinite %d grad-nonfinite %d attempt %d]",
                              (long long)L.ids[b], r.length, L.status_h[b], bad_u, bad_g, r.res.attempts);
                line += buf;
            }
            std::fprintf(stderr, "%s\n", line.c_str());
        }
    }
    // A plan instance whose launch returned a non-finite solve is released
    // instead of reused: its re-runs and later requests of the class get a
    // plan with freshly uploaded tables.
    bool quarantine = false;
    for (int c : codes) quarantine |= c == FH_ENONFINITE;
    if (quarantine) {
        for (auto it = S->plans.begin(); it != S->plans.end(); ++it)
            if (&*it == L.pi) {
                S->plan_bytes -= it->bytes;
                fh_plan_destroy(it->plan);
                S->plans.erase(it);
                break;
            }
    } else {
        L.pi->busy = false;
        L.pi->last_used = now_ns();
    }
    L.pi = nullptr;
    trace_slow("output scatter", tc, (int)rq.size());
    if (L.io) S->io.release(L.io, L.io_bytes);
    L.io = nullptr;
    lk.lock();
    // Non-finite fields from a solve inside a batch are re-run once, alone:
    // every such failure seen so far came from FP32 multimode batches under
    // heavy concurrent load, and the same requests solved alone or in small
    // batches returned finite results.  A second failure is reported.
    for (size_t b = 0; b < L.ids.size(); ++b) {
        Req& r = S->reqs[L.ids[b]];
        if (codes[b] == FH_ENONFINITE && L.ids.size() > 1 && r.res.attempts == 1) {
            r.solo = true;
            r.state = 0;
            S->queue.push_front(L.ids[b]);
            codes[b] = -1;  // re-queued
        }
    }
    for (size_t b = 0; b < L.ids.size(); ++b)
        if (codes[b] != -1)
        finish(S, L.ids[b], codes[b],
               codes[b] == FH_EMAXSTEPS ? "adaptive solve reached max_steps"
                                        : (codes[b] == FH_ENONFINITE ? "solve produced non-finite fields"
                                                                     : "device error: " + std::string(fh_last_error())));
    L.ids.clear();
    L.busy = false;
    S->cv_done.notify_all();
}

void worker_main(fh_server_s* S)
{
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        bool progressed = false;
        for (auto& L : S->lanes) {
            if (!L.busy) continue;
            const hipError_t q = hipEventQuery(L.ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) (void)hipStreamSynchronize(L.s);
            complete(S, L, q == hipSuccess ? FH_OK : fhi::set_error(FH_EDEVICE, hipGetErrorString(q)), lk);
            progressed = true;
        }
        for (auto& L : S->lanes) {
            if (L.busy || S->queue.empty()) continue;
            if (!dispatch(S, L, lk)) break;
            progressed = true;
        }
        bool any_busy = false;
        for (auto& L : S->lanes) any_busy |= L.busy;
        if (S->stop && S->queue.empty() && !any_busy) break;
        if (!progressed) S->cv_work.wait_for(lk, std::chrono::microseconds(any_busy ? 50 : 2000));
    }
}

int validate(const fh_server_s* S, const fh_request* r, double* dmax)
{
    if (!r) return fhi::set_error(FH_EINVAL, "request is NULL");
    if (r->kind != FH_REQ_FORWARD && r->kind != FH_REQ_GRADIENT) return fhi::set_error(FH_EINVAL, "unknown request kind");
    if (r->nt < 4 || (r->nt & (r->nt - 1)) || r->modes < 1) return fhi::set_error(FH_EINVAL, "nt must be a power of two, modes >= 1");
    if (!r->dispersion || !r->gamma || !r->raman || !r->shock || !r->u0 || !r->uL)
        return fhi::set_error(FH_EINVAL, "operator, launch or output pointer is NULL");
    if (!(std::isfinite(r->length_m) && r->length_m > 0)) return fhi::set_error(FH_EINVAL, "fiber length must be positive and finite");
    // multimode requests run ceil(L / multimode_step_m) fixed steps: bound the
    // count before it is converted to an int
    if (r->modes > 1 && !(std::ceil(r->length_m / S->cfg.multimode_step_m) <= (double)kMaxMultimodeSteps))
        return fhi::set_error(FH_EUNSUPPORTED, "multimode request needs more than " +
                                                   std::to_string(kMaxMultimodeSteps) + " fixed steps");
    if (r->kind == FH_REQ_GRADIENT) {
        if (!r->band_mask) return fhi::set_error(FH_EINVAL, "gradient requests need a band mask");
        // the step count submitted below (at least 16) must fit the plan's checkpoint capacity
        if (r->modes > 1 && std::max(16.0, std::ceil(r->length_m / S->cfg.multimode_step_m)) > mm_grad_cap(S))
            return fhi::set_error(FH_EUNSUPPORTED, "multimode gradient needs more steps than the server's capacity");
    }
    const bool fp64 = S->cfg.precision == FH_PREC_FP64;
    if (r->modes == 1 ? !fh::tsit5_supported(r->nt, 1, fp64) : !fh::multi_supported(r->nt, r->modes, fp64))
        return fhi::set_error(FH_EUNSUPPORTED, "no kernel for nt=" + std::to_string(r->nt) +
                                                   " with " + std::to_string(r->modes) + " modes");
    const size_t nf = (size_t)r->nt * r->modes;
    double m = 0;
    for (size_t i = 0; i < nf; ++i) {
        if (!std::isfinite(r->dispersion[i])) return fhi::set_error(FH_EINVAL, "dispersion must be finite");
        m = std::fmax(m, std::fabs(r->dispersion[i]));
    }
    if (m * r->length_m >= FH_MAX_PHASE) return fhi::set_error(FH_EUNSUPPORTED, "dispersion phase |D|*L exceeds the bound");
    for (size_t i = 0; i < 2 * nf; ++i)
        if (!std::isfinite(r->u0[i])) return fhi::set_error(FH_EINVAL, "input spectrum must be finite");
    *dmax = m;
    return FH_OK;
}

}  // namespace

extern "C" {

int fh_server_create(fh_context ctx, const fh_server_config* cfg, fh_server* out)
{
    if (!ctx || !cfg || !out) return fhi::set_error(FH_EINVAL, "context, config or out is NULL");
    if (cfg->lanes < 1 || cfg->lanes > 32) return fhi::set_error(FH_EINVAL, "lanes must lie in [1, 32]");
    if (cfg->max_batch < 1) return fhi::set_error(FH_EINVAL, "max_batch must be positive");
    if (cfg->memory_budget_bytes <= 0) return fhi::set_error(FH_EINVAL, "memory budget must be positive");
    if (cfg->precision != FH_PREC_FP64 && cfg->precision != FH_PREC_FP32_BULK) return fhi::set_error(FH_EINVAL, "unknown precision");
    if (!(cfg->rtol > 0) || !(cfg->atol >= 0) || !(cfg->multimode_step_m > 0))
        return fhi::set_error(FH_EINVAL, "tolerances and the multimode step must be positive");
    if (cfg->tape_segments < 0) return fhi::set_error(FH_EINVAL, "tape_segments must be >= 0");
    if (hipSetDevice(ctx->device) != hipSuccess) return fhi::set_error(FH_EDEVICE, "hipSetDevice failed");
    fh_server_s* S = new fh_server_s;
    S->ctx = ctx;
    S->cfg = *cfg;
    if (int rc = S->arena.init((size_t)cfg->memory_budget_bytes)) {
        delete S;
        return rc;
    }
    // Diagnostic (environment FH_SERVER_FILL=<byte>): fill the device arena
    // with that byte, so a kernel that reads plan memory it has not written
    // sees it (0xff reads as NaN in FP32 and FP64).
    if (const char* f = std::getenv("FH_SERVER_FILL")) {
        const int byte = std::atoi(f) & 0xff;
        if (hipMemset(S->arena.base, byte, S->arena.size) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            S->arena.destroy();
            delete S;
            return fhi::set_error(FH_EDEVICE, "filling the server's device memory failed");
        }
    }
    if (int rc = S->stage.init(kStageBytes, true)) {
        S->arena.destroy();
        delete S;
        return rc;
    }
    if (int rc = S->io.init(std::min<size_t>(kIoBytes, std::max<size_t>(kStageBytes, cfg->memory_budget_bytes / 16)),
                            true)) {
        S->stage.destroy();
        S->arena.destroy();
        delete S;
        return rc;
    }
    S->lanes.resize(cfg->lanes);
    for (auto& L : S->lanes) {
        // non-blocking: plan uploads on the null stream do not wait for running lanes
        bool ok = hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking) == hipSuccess && hipEventCreateWithFlags(&L.ev, hipEventDisableTiming) == hipSuccess &&
                  hipMalloc((void**)&L.d_len, sizeof(double) * cfg->max_batch) == hipSuccess &&
                  hipMalloc((void**)&L.d_nst, sizeof(int) * cfg->max_batch) == hipSuccess;
        if (!ok) {
            fh_server_destroy(S);
            return fhi::set_error(FH_EDEVICE, "lane setup failed");
        }
    }
    S->worker = std::thread(worker_main, S);
    *out = S;
    return FH_OK;
}

int fh_server_submit(fh_server S, const fh_request* r, int64_t* id)
{
    if (!S || !id) return fhi::set_error(FH_EINVAL, "server or id is NULL");
    double dmax = 0;
    int rc = validate(S, r, &dmax);
    if (rc) return rc;
    const bool grad = r->kind == FH_REQ_GRADIENT;
    const size_t nf = (size_t)r->nt * r->modes, m4 = (size_t)r->modes * r->modes * r->modes * r->modes;
    Fnv h;
    h.val(grad);
    h.val(r->nt);
    h.val(r->modes);
    h.val(r->one_m_fr);
    h.add(r->dispersion, sizeof(double) * nf);
    h.add(r->gamma, sizeof(double) * m4);
    h.add(r->raman, sizeof(double) * 2 * r->nt);
    h.add(r->shock, sizeof(double) * r->nt);
    if (grad) h.add(r->band_mask, (size_t)r->nt);
    Req q;
    q.kind = r->kind;
    q.key = h.h;
    q.length = r->length_m;
    q.nsteps = r->modes > 1 ? std::max(16, (int)std::ceil(r->length_m / S->cfg.multimode_step_m)) : 0;
    q.heavy = request_flops(r->nt, r->modes, q.nsteps, r->length_m, grad) > kHeavyFlops;
    q.u0.assign(r->u0, r->u0 + 2 * nf);
    q.uL = r->uL;
    q.grad = r->grad;
    q.cost = r->cost;
    q.res.submit_ns = now_ns();
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fhi::set_error(FH_EINVAL, "server is shutting down");
    if (!S->ops.count(q.key)) {
        OpSet o;
        o.nt = r->nt;
        o.modes = r->modes;
        o.grad = grad;
        o.disp.assign(r->dispersion, r->dispersion + nf);
        o.gamma.assign(r->gamma, r->gamma + m4);
        o.raman.assign(r->raman, r->raman + 2 * r->nt);
        o.shock.assign(r->shock, r->shock + r->nt);
        if (grad) o.mask.assign(r->band_mask, r->band_mask + r->nt);
        o.one_m_fr = r->one_m_fr;
        S->ops.emplace(q.key, std::move(o));
    }
    *id = S->next_id++;
    S->reqs.emplace(*id, std::move(q));
    S->queue.push_back(*id);
    S->cv_work.notify_one();
    return FH_OK;
}

int fh_server_poll(fh_server S, int64_t id, int32_t* state)
{
    if (!S || !state) return fhi::set_error(FH_EINVAL, "server or state is NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fhi::set_error(FH_EINVAL, "unknown request id");
    *state = it->second.state;
    return FH_OK;
}

int fh_server_wait(fh_server S, int64_t id, fh_request_result* res)
{
    if (!S || !res) return fhi::set_error(FH_EINVAL, "server or result is NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    if (!S->reqs.count(id)) return fhi::set_error(FH_EINVAL, "unknown request id");
    S->cv_done.wait(lk, [&] { return S->reqs[id].state == 2; });
    *res = S->reqs[id].res;
    const std::string msg = S->reqs[id].msg;
    S->reqs.erase(id);
    if (res->retcode) return fhi::set_error(res->retcode, "request " + std::to_string(id) + " failed: " + msg);
    return FH_OK;
}

int fh_server_destroy(fh_server S)
{
    if (!S) return FH_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->ctx->device);
    for (auto& L : S->lanes)
        if (L.s) (void)hipStreamSynchronize(L.s);
    for (auto& pi : S->plans) fh_plan_destroy(pi.plan);
    S->plans.clear();
    S->arena.destroy();
    S->stage.destroy();
    S->io.destroy();
    for (auto& L : S->lanes) {
        if (L.s) (void)hipStreamSynchronize(L.s);
        L.h_in.release();
        L.h_out.release();
        L.h_grad.release();
        L.h_cost.release();
        L.h_len.release();
        L.h_status.release();
        L.h_nst.release();
        L.h_stats.release();
        if (L.d_len) (void)hipFree(L.d_len);
        if (L.d_nst) (void)hipFree(L.d_nst);
        if (L.ev) (void)hipEventDestroy(L.ev);
        if (L.s) (void)hipStreamDestroy(L.s);
    }
    delete S;
    return FH_OK;
}

}  // extern "C"
