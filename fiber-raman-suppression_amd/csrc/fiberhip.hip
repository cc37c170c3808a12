// fiberhip.hip -- C ABI of libfiberhip (see include/fiberhip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <mutex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fiberhip.h"
#include "fh_internal.hpp"
#include "fh_launch.hpp"
#include "fh_multi.hpp"
#include "fh_single.hpp"
#include "fh_tsit5.hpp"
#include "fh_tsit5_multi.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

#define FH_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(e_ == hipErrorOutOfMemory ? FH_ENOMEM : FH_EDEVICE,                \
                        std::string(#call) + ": " + hipGetErrorString(e_));                \
    } while (0)

}  // namespace

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    fhi::Arena* arena = nullptr;  // owner of p when it came from an arena
};

struct fh_plan_s {
    fh_context ctx;
    fh_problem_desc d;
    hipStream_t copy_stream = nullptr;  // stream of the table uploads at creation (null: the context stream)
    fhi::Arena* arena = nullptr;        // source of the plan's device buffers (null: hipMalloc)
    fhi::Arena* stage = nullptr;        // pinned host staging of the table uploads (null: synchronous uploads)
    std::vector<std::pair<void*, size_t>> staged;  // staging chunks held until the plan is destroyed
    int threads;  // workgroup size of the plan's kernels (single-mode or multimode shape)
    DevBuf disp, hn, shn, tw, work, status, u_in, u_out;
    DevBuf mask, ckpt, tape_u, tape_h, g_lam0, g_grad, g_cost, g_seed;  // adjoint, allocated on first use
    int every = 50, nseg = 0;
    int ntaped = 0;  // segments whose stage tape the forward sweep keeps (single mode)
    bool has_mask = false;
    DevBuf gr, qidx;  // multimode: reduced gamma [Q][P] and output-pair index [M][M]
    int q = 0;
    size_t work_per_solve = 0;  // complex entries
    int64_t total_bytes = 0;
    double gamma0 = 0.0;        // gamma[1,1,1,1] of a single-mode plan
    bool adaptive = false;      // FH_SCHEME_TSIT5
    bool gamma_full_sym = true; // gamma invariant under i<->j, j<->k, k<->l (sensitivity_disp_mmf.jl:103-112)
    int max_steps = 0;
    DevBuf hs, zs, sstats;      // adaptive: step sizes, positions [max_batch][max_steps]; counters [max_batch][3]
    size_t grad_work_per_solve = 0;  // multimode gradient slice (complex entries per solve)
    size_t ckpt_per_solve = 0, tape_u_per_solve = 0, tape_h_per_solve = 0;  // bytes
    std::string grad_error;          // why this plan offers no gradient (empty: offered)
    fh_tape spare = nullptr;         // storage of a freed tape, reused by the next fh_forward_tape
    int64_t gradient_bytes = 0;      // plan bytes once the adjoint buffers exist (layout)
    bool grad_work = false;          // multimode: the work slice has the gradient kernels' size
    bool mm = false;     // runs on the multimode kernels (M > 1, or one mode with split transforms)
    bool split = false;  // split transforms (FP64 at 32768 samples): Raman table index-reversed
};

// A kept forward solve (fh_forward_tape): the shaped launch, the output, the
// checkpoints and kept stage tape, and the accepted steps of adaptive plans.
struct fh_tape_s {
    fh_plan plan = nullptr;
    int batch = 0, capacity = 0;  // solves kept; solves the buffers hold
    bool swept = false;           // a single-mode backward sweep has consumed the kept tape
    DevBuf u0, uL, ckpt, tape_u, tape_h, hs, zs, sstats, status, seed, lam0, kgrad, gphase, gamp, edot, cost;
    int64_t bytes = 0;
};

namespace {

int dev_alloc(fh_plan p, DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (p->arena) {
        b.p = p->arena->alloc(bytes);
        if (!b.p) return fail(FH_ENOMEM, "the server's device memory has no free range of " + std::to_string(bytes) + " bytes");
        b.arena = p->arena;
        b.bytes = bytes;
        p->total_bytes += (int64_t)bytes;
        return FH_OK;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(e == hipErrorOutOfMemory ? FH_ENOMEM : FH_EDEVICE,
                    std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    p->total_bytes += (int64_t)bytes;
    return FH_OK;
}

void dev_free(DevBuf& b)
{
    if (b.p && b.arena)
        b.arena->release(b.p, b.bytes);
    else if (b.p)
        (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    b.arena = nullptr;
}

// Free a plan buffer and take it off the plan's byte count.
void total_release(fh_plan p, DevBuf& b)
{
    p->total_bytes -= (int64_t)b.bytes;
    dev_free(b);
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

template <int N, int T> void fill_twiddles(std::vector<std::complex<long double>>& out)
{
    using PL = fh::FftPlan<N, T>;
    out.assign(PL::TW_SIZE, std::complex<long double>(1, 0));
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int q = 0; q < PL::NPASS; ++q) {
        const int rad = 1 << PL::log_radix(q), ns = 1 << PL::log_ns(q);
        const int off = PL::tw_offset(q);
        for (int r = 1; r < rad; ++r)
            for (int jm = 0; jm < ns; ++jm) {
                long double a = -two_pi * (long double)jm * (long double)r / ((long double)ns * rad);
                out[off + (r - 1) * ns + jm] = std::complex<long double>(cosl(a), sinl(a));
            }
    }
}

// ------------------------------------------------------------ dispatch

using fh::SingleShape;

template <typename R, int N> struct SingleKernel {
    static constexpr int T = SingleShape<R, N>::T;
    using LY = fh::SmLayout<R, N, T>;
    static size_t work_per_solve() { return (size_t)(LY::NW > 0 ? LY::NW : 1) * N; }
    static void twiddles(std::vector<std::complex<long double>>& tw) { fill_twiddles<N, T>(tw); }
    static hipError_t launch(const fh::SmArgs<R>& a, int batch, hipStream_t s)
    {
        hipLaunchKernelGGL((fh::k_rk4ip_single<R, N, T>), dim3(batch), dim3(T), 0, s, a);
        return hipGetLastError();
    }
};

template <typename R, int N> struct GradKernel {
    static constexpr int T = SingleShape<R, N>::T;
    static hipError_t launch(const fh::SmArgs<R>& a, int batch, hipStream_t s)
    {
        hipLaunchKernelGGL((fh::k_grad_single<R, N, T>), dim3(batch), dim3(T), 0, s, a);
        return hipGetLastError();
    }
};

template <typename R, template <typename, int> class Op, typename... A>
auto with_size(int nt, A&&... args) -> decltype(Op<R, 1024>::run(args...))
{
    switch (nt) {
    case 64: return Op<R, 64>::run(args...);
    case 128: return Op<R, 128>::run(args...);
    case 256: return Op<R, 256>::run(args...);
    case 512: return Op<R, 512>::run(args...);
    case 1024: return Op<R, 1024>::run(args...);
    case 2048: return Op<R, 2048>::run(args...);
    case 4096: return Op<R, 4096>::run(args...);
    case 8192: return Op<R, 8192>::run(args...);
    case 16384: return Op<R, 16384>::run(args...);
    case 32768:
        if constexpr (sizeof(R) == 4) return Op<R, 32768>::run(args...);  // FP32 only: LDS budget
        return Op<R, 1024>::unsupported();
    default: return Op<R, 1024>::unsupported();
    }
}

template <typename R, int N> struct OpWork {
    static size_t run() { return SingleKernel<R, N>::work_per_solve(); }
    static size_t unsupported() { return 0; }
};
template <typename R, int N> struct OpTw {
    static int run(std::vector<std::complex<long double>>& tw)
    {
        SingleKernel<R, N>::twiddles(tw);
        return SingleKernel<R, N>::T;
    }
    static int unsupported() { return 0; }
};
template <typename R, int N> struct OpTwMulti {
    static int run(std::vector<std::complex<long double>>& tw)
    {
        fill_twiddles<N, fh::MultiShape<R, N>::T>(tw);
        return fh::MultiShape<R, N>::T;
    }
    static int unsupported() { return 0; }
};
template <typename R, int N> struct OpShape {
    // workgroup size and twiddle-table entries of the single-mode or multimode kernels
    static std::pair<int, int> run(bool multi)
    {
        constexpr int TS = fh::SingleShape<R, N>::T, TM = fh::MultiShape<R, N>::T;
        return multi ? std::make_pair(TM, (int)fh::FftPlan<N, TM>::TW_SIZE)
                     : std::make_pair(TS, (int)fh::FftPlan<N, TS>::TW_SIZE);
    }
    static std::pair<int, int> unsupported() { return {0, 0}; }
};
template <typename R, int N> struct OpGrad {
    static hipError_t run(const fh::SmArgs<R>& a, int batch, hipStream_t s) { return GradKernel<R, N>::launch(a, batch, s); }
    static hipError_t unsupported() { return hipErrorInvalidValue; }
};
template <typename R, int N> struct OpLaunch {
    static hipError_t run(const fh::SmArgs<R>& a, int batch, hipStream_t s)
    {
        return SingleKernel<R, N>::launch(a, batch, s);
    }
    static hipError_t unsupported() { return hipErrorInvalidValue; }
};

// -------------------------------------------------------- energy kernel

// One workgroup per (field, mode): FP64 sum of |u|^2 in a fixed tree order.
template <typename R>
__global__ __launch_bounds__(256) void k_energy(const fh::C2<R>* __restrict__ u, int nt, double* __restrict__ e)
{
    __shared__ double red[256];
    const fh::C2<R>* col = u + (size_t)blockIdx.x * nt;
    double s = 0.0;
    for (int t = threadIdx.x; t < nt; t += 256) {
        const double x = (double)col[t].x, y = (double)col[t].y;
        s += x * x + y * y;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) e[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

int check_desc(const fh_problem_desc* d)
{
    if (!d) return fail(FH_EINVAL, "problem descriptor is NULL");
    if (!is_pow2(d->nt) || d->nt < 4) return fail(FH_EINVAL, "nt must be a power of two >= 4");
    if (d->modes < 1) return fail(FH_EINVAL, "modes must be positive");
    if (d->precision != FH_PREC_FP64 && d->precision != FH_PREC_FP32_BULK)
        return fail(FH_EINVAL, "unknown precision");
    if (d->scheme != FH_SCHEME_RK4IP_FIXED && d->scheme != FH_SCHEME_TSIT5) return fail(FH_EINVAL, "unknown scheme");
    if (d->scheme == FH_SCHEME_RK4IP_FIXED && d->steps <= 0) return fail(FH_EINVAL, "steps must be a positive integer");
    if (d->scheme == FH_SCHEME_TSIT5) {
        if (!(std::isfinite(d->rtol) && d->rtol > 0)) return fail(FH_EINVAL, "rtol must be positive and finite");
        if (!(std::isfinite(d->atol) && d->atol >= 0)) return fail(FH_EINVAL, "atol must be nonnegative and finite");
        if (d->max_steps < 0) return fail(FH_EINVAL, "max_steps must be >= 0");
        if (!fh::tsit5_supported(d->nt, d->modes, d->precision == FH_PREC_FP64))
            return fail(FH_EUNSUPPORTED, "the adaptive scheme covers single-mode grids of 256..32768 samples "
                                         "and the multimode grids and mode counts of the fixed-step scheme");
    }
    if (d->max_batch <= 0) return fail(FH_EINVAL, "max_batch must be positive");
    if (!(std::isfinite(d->length_m) && d->length_m > 0)) return fail(FH_EINVAL, "fiber length must be positive and finite");
    if (!d->dispersion || !d->gamma || !d->raman || !d->shock) return fail(FH_EINVAL, "operator array is NULL");
    const size_t nt = (size_t)d->nt, m = (size_t)d->modes;
    for (size_t i = 0; i < nt * m; ++i)
        if (!std::isfinite(d->dispersion[i])) return fail(FH_EINVAL, "dispersion must be finite");
    for (size_t i = 0; i < m * m * m * m; ++i)
        if (!std::isfinite(d->gamma[i])) return fail(FH_EINVAL, "gamma tensor must be finite");
    for (size_t i = 0; i < 2 * nt; ++i)
        if (!std::isfinite(d->raman[i])) return fail(FH_EINVAL, "Raman response must be finite");
    for (size_t i = 0; i < nt; ++i) {
        if (!std::isfinite(d->shock[i])) return fail(FH_EINVAL, "shock factor must be finite");
        if (!(d->shock[i] > 0))
            return fail(FH_EGRID, "simulation bandwidth reaches nonpositive absolute optical frequencies "
                                  "(shock factor w/w0 <= 0); increase the time window or use fewer samples");
    }
    if (!std::isfinite(d->one_m_fr) || d->one_m_fr < 0 || d->one_m_fr > 1)
        return fail(FH_EINVAL, "one_m_fR must lie in [0, 1]");
    double dmax = 0.0;
    for (size_t i = 0; i < nt * m; ++i) dmax = std::max(dmax, std::fabs(d->dispersion[i]));
    if (dmax * d->length_m >= FH_MAX_PHASE)
        return fail(FH_EUNSUPPORTED, "dispersion phase |D|*L exceeds " + std::to_string(FH_MAX_PHASE) + " rad");
    if (d->checkpoint_every < 0) return fail(FH_EINVAL, "checkpoint_every must be >= 0");
    if (d->tape_segments < 0) return fail(FH_EINVAL, "tape_segments must be >= 0");
    if (d->objective != FH_OBJ_BAND_SUM && d->objective != FH_OBJ_FUNDAMENTAL && d->objective != FH_OBJ_WORST_MODE)
        return fail(FH_EINVAL, "unknown objective");
    if (!(std::isfinite(d->objective_tau) && d->objective_tau >= 0))
        return fail(FH_EINVAL, "objective_tau must be finite and >= 0 (0 = 50)");
    if (d->modes == 1) {
        if (d->nt < 64 || d->nt > 32768) return fail(FH_EUNSUPPORTED, "nt must lie in [64, 32768]");
    } else if (!fh::multi_supported(d->nt, d->modes, d->precision == FH_PREC_FP64)) {
        return fail(FH_EUNSUPPORTED, "multimode kernels cover M in {2,3,4,6,10} with nt in [256, 32768]");
    }
    return FH_OK;
}

// Everything a plan allocates, derived from the descriptor alone (no device
// needed): fh_plan_create allocates these sizes, the first gradient call the
// adjoint part, and fh_plan_check reports them.
struct PlanLayout {
    int threads = 0, tw_entries = 0;
    bool adaptive = false;
    bool mm = false;     // runs on the multimode kernels (M > 1, or one mode with split transforms)
    bool split = false;  // FP64 fields of 32 samples per thread: transforms in two halves
    int every = 50, nseg = 0, ntaped = 0, max_steps = 0, q = 0;
    bool gamma_full_sym = true;
    std::vector<std::pair<int, int>> outp, inp;  // multimode output / input pairs (i, j)
    std::vector<int> qidx;                        // [M][M] output-pair index
    size_t work_per_solve = 0;       // complex entries of the forward kernels' slice
    size_t grad_work_per_solve = 0;  // multimode gradient slice (the forward slice serves single mode)
    size_t ckpt_per_solve = 0, tape_u_per_solve = 0, tape_h_per_solve = 0;  // bytes
    std::string grad_error;          // why a gradient is not offered (empty: offered)
    int64_t forward_bytes = 0, gradient_bytes = 0;  // device bytes of the plan, without / with adjoint buffers
};

size_t alloc_size(size_t bytes) { return bytes == 0 ? 16 : bytes; }  // dev_alloc's rounding

template <typename R> int compute_layout_t(const fh_problem_desc& d, PlanLayout& L)
{
    const size_t nt = (size_t)d.nt, M = (size_t)d.modes, mb = (size_t)d.max_batch, V = 2 * sizeof(R);
    const bool fp64 = sizeof(R) == 8;
    L.split = fh::mm_split_rt(d.nt, fp64);
    L.mm = M > 1 || L.split;
    const auto shape = L.split ? std::make_pair(1024, fh::mm_twiddle_entries<double, 32768, 1024>())
                               : with_size<R, OpShape>(d.nt, L.mm);
    if (shape.first == 0) return fail(FH_EUNSUPPORTED, "no kernel for nt=" + std::to_string(d.nt));
    if (L.split && d.nt != 32768) return fail(FH_EUNSUPPORTED, "no kernel for nt=" + std::to_string(d.nt));
    L.threads = shape.first;
    L.tw_entries = shape.second;
    L.adaptive = d.scheme == FH_SCHEME_TSIT5;
    L.every = d.checkpoint_every > 0 ? d.checkpoint_every : 50;
    if (L.adaptive) {
        L.max_steps = d.max_steps > 0 ? d.max_steps : 10000;
        if (L.every > L.max_steps) L.every = L.max_steps;
        L.nseg = (L.max_steps + L.every - 1) / L.every;
        L.work_per_solve = !L.mm ? (size_t)fh::tsit5_fields(d.nt, fp64) * nt
                                 : (size_t)fh::tsit5_multi_fields(d.nt, d.modes, fp64) * nt;
    } else {
        if (L.every > d.steps) L.every = d.steps;
        L.nseg = (d.steps + L.every - 1) / L.every;
        L.work_per_solve = !L.mm ? with_size<R, OpWork>(d.nt) : (size_t)fh::multi_fields(d.nt, d.modes, fp64) * nt;
    }
    L.ntaped = std::min(d.tape_segments, L.nseg);
    if (L.mm) {
        const int m = d.modes;
        auto G = [&](int i, int j, int k, int l) { return d.gamma[i + m * (j + m * (k + m * l))]; };
        double gmax = 0;
        for (size_t i = 0; i < M * M * M * M; ++i) gmax = std::max(gmax, std::fabs(d.gamma[i]));
        bool sym = true;
        for (int i = 0; i < m && sym; ++i)
            for (int j = 0; j < m && sym; ++j)
                for (int k = 0; k < m && sym; ++k)
                    for (int l = 0; l < m; ++l)
                        if (std::fabs(G(i, j, k, l) - G(j, i, k, l)) > 1e-12 * std::max(1.0, gmax)) {
                            sym = false;
                            break;
                        }
        for (int i = 0; i < m; ++i)
            for (int j = i; j < m; ++j) L.inp.push_back({i, j});
        L.qidx.assign(M * M, 0);
        for (int i = 0; i < m; ++i)
            for (int j = sym ? i : 0; j < m; ++j) {
                L.qidx[i * m + j] = (int)L.outp.size();
                if (sym) L.qidx[j * m + i] = (int)L.outp.size();
                L.outp.push_back({i, j});
            }
        L.q = (int)L.outp.size();
        const double tol = 1e-12 * std::max(1.0, gmax);
        auto close = [&](double x, double y) { return std::fabs(x - y) <= tol + 1e-12 * std::fabs(y); };
        for (int i = 0; i < m && L.gamma_full_sym; ++i)
            for (int j = 0; j < m && L.gamma_full_sym; ++j)
                for (int k = 0; k < m && L.gamma_full_sym; ++k)
                    for (int l = 0; l < m; ++l) {
                        const double g = G(i, j, k, l);
                        if (!close(g, G(j, i, k, l)) || !close(g, G(i, k, j, l)) || !close(g, G(i, j, l, k))) {
                            L.gamma_full_sym = false;
                            break;
                        }
                    }
    }
    // forward: operators, twiddles, mask, reduced gamma, slice, status, fields, step records
    const size_t field = V * nt * M * mb;
    size_t fwd = alloc_size(8 * nt * M) + alloc_size(sizeof(R) * 2 * nt) + alloc_size(sizeof(R) * nt) +
                 alloc_size(sizeof(R) * 2 * (size_t)L.tw_entries);
    if (d.band_mask) fwd += alloc_size(nt);
    if (L.mm) fwd += alloc_size(sizeof(R) * L.outp.size() * L.inp.size()) + alloc_size(sizeof(int) * M * M);
    const size_t work = alloc_size(V * L.work_per_solve * mb);
    fwd += work + alloc_size(sizeof(int) * mb) + 2 * alloc_size(field);
    if (L.adaptive) fwd += 2 * alloc_size(sizeof(double) * mb * (size_t)L.max_steps) + alloc_size(sizeof(long long) * 3 * mb);
    L.forward_bytes = (int64_t)fwd;
    // gradient: checkpoints, stage tape (one region per kept segment plus the
    // shared region of recomputed segments: fixed step only when some
    // segments are recomputed; adaptive always, as the segment count is known
    // only at the end of a solve), adjoint fields
    if (L.mm)
        L.grad_work_per_solve = (size_t)(L.adaptive ? fh::tsit5_multi_grad_fields(d.nt, d.modes, fp64)
                                                    : fh::multi_grad_fields(d.nt, d.modes, fp64)) * nt;
    L.ckpt_per_solve = V * nt * M * (size_t)L.nseg;
    const size_t stages = L.adaptive ? 6 : 4;
    const size_t hbytes = !L.mm ? sizeof(R) * nt : V * nt * (size_t)((L.q + 1) / 2);
    const size_t regions = (size_t)L.ntaped + ((L.adaptive || L.ntaped < L.nseg) ? 1 : 0);
    L.tape_u_per_solve = V * nt * M * stages * (size_t)L.every * regions;
    L.tape_h_per_solve = hbytes * stages * (size_t)L.every * regions;
    if (L.ckpt_per_solve > 0xffffffffull)
        L.grad_error = "the checkpoints of one solve exceed 4 GiB (" + std::to_string(L.ckpt_per_solve) +
                       " bytes); raise checkpoint_every or lower steps / max_steps";
    else if (L.tape_u_per_solve > 0xffffffffull || L.tape_h_per_solve > 0xffffffffull)
        L.grad_error = "the stage tape of one solve exceeds 4 GiB; lower tape_segments or checkpoint_every";
    else if (L.mm && V * L.grad_work_per_solve > 0xffffffffull)
        L.grad_error = "the gradient work slice of one solve exceeds 4 GiB";
    size_t grad = fwd;
    if (L.mm) grad = grad - work + alloc_size(V * L.grad_work_per_solve * mb);
    grad += alloc_size(L.ckpt_per_solve * mb) + alloc_size(L.tape_u_per_solve * mb) +
            alloc_size(L.tape_h_per_solve * mb) + 2 * alloc_size(field) + alloc_size(sizeof(double) * nt * mb) +
            alloc_size(sizeof(double) * mb);
    L.gradient_bytes = (int64_t)grad;
    return FH_OK;
}

int compute_layout(const fh_problem_desc& d, PlanLayout& L)
{
    return d.precision == FH_PREC_FP64 ? compute_layout_t<double>(d, L) : compute_layout_t<float>(d, L);
}

template <typename R> int upload_tables(fh_plan p, const PlanLayout& L)
{
    const fh_problem_desc& d = p->d;
    const int nt = d.nt;
    // uploads are asynchronous on one stream and synchronised before the host
    // tables go out of scope: a synchronous copy runs on the null stream,
    // whose hardware queue may be shared with a stream running a long launch
    const hipStream_t cs = p->copy_stream ? p->copy_stream : p->ctx->stream;
    // one table upload: through the plan's pinned staging, left in flight;
    // otherwise from the caller's memory, synchronised before returning
    auto upload = [&](void* dst, const void* src, size_t bytes) -> int {
        if (p->stage && p->copy_stream) {
            if (void* st = p->stage->alloc(bytes)) {
                std::memcpy(st, src, bytes);
                p->staged.emplace_back(st, bytes);
                FH_HIP(hipMemcpyAsync(dst, st, bytes, hipMemcpyHostToDevice, cs));
                return FH_OK;
            }
        }
        FH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, cs));
        FH_HIP(hipStreamSynchronize(cs));
        return FH_OK;
    };
    std::vector<std::complex<long double>> tw;
    int T = 0;
    if (!L.mm) {
        T = with_size<R, OpTw>(nt, tw);
    } else if (!L.split) {
        T = with_size<R, OpTwMulti>(nt, tw);
    } else {
        // split transforms: the N/2-point block plan, then w^n = exp(-2 pi i n / N), n < N/2
        T = 1024;
        fill_twiddles<16384, 1024>(tw);
        const long double two_pi = 6.283185307179586476925286766559005768L;
        for (int n = 0; n < nt / 2; ++n) {
            const long double a = -two_pi * (long double)n / (long double)nt;
            tw.emplace_back(cosl(a), sinl(a));
        }
    }
    if (T != L.threads || (int)tw.size() != L.tw_entries)
        return fail(FH_EINVAL, "internal: kernel shape differs from the plan layout");
    p->threads = T;
    p->mm = L.mm;
    p->split = L.split;
    std::vector<R> twr(2 * tw.size()), hn(2 * (size_t)nt), shn(nt);
    for (size_t i = 0; i < tw.size(); ++i) {
        twr[2 * i] = (R)tw[i].real();
        twr[2 * i + 1] = (R)tw[i].imag();
    }
    // hR (-1)^k / N; split transforms convolve through the +1 transform
    // first and take the table at index -k
    for (int e = 0; e < nt; ++e) {
        const int k = L.split ? (nt - e) % nt : e;
        const double sgn = (k & 1) ? -1.0 : 1.0;
        hn[2 * e] = (R)(d.raman[2 * k] * sgn / nt);
        hn[2 * e + 1] = (R)(d.raman[2 * k + 1] * sgn / nt);
        shn[e] = (R)(d.shock[e] / nt);
    }
    int rc;
    if ((rc = dev_alloc(p, p->disp, sizeof(double) * (size_t)nt * d.modes))) return rc;
    if ((rc = dev_alloc(p, p->hn, sizeof(R) * hn.size()))) return rc;
    if ((rc = dev_alloc(p, p->shn, sizeof(R) * shn.size()))) return rc;
    if ((rc = dev_alloc(p, p->tw, sizeof(R) * twr.size()))) return rc;
    if ((rc = upload(p->disp.p, d.dispersion, p->disp.bytes))) return rc;
    if ((rc = upload(p->hn.p, hn.data(), p->hn.bytes))) return rc;
    if ((rc = upload(p->shn.p, shn.data(), p->shn.bytes))) return rc;
    if ((rc = upload(p->tw.p, twr.data(), p->tw.bytes))) return rc;
    if (d.band_mask) {
        bool any = false;
        for (int k = 0; k < nt; ++k) any |= d.band_mask[k] != 0;
        if (!any) return fail(FH_EINVAL, "band mask must contain at least one bin");
        if ((rc = dev_alloc(p, p->mask, (size_t)nt))) return rc;
        if ((rc = upload(p->mask.p, d.band_mask, (size_t)nt))) return rc;
        p->has_mask = true;
    }
    p->every = L.every;
    p->nseg = L.nseg;
    p->ntaped = L.ntaped;
    p->max_steps = L.max_steps;
    p->work_per_solve = L.work_per_solve;
    p->grad_work_per_solve = L.grad_work_per_solve;
    p->ckpt_per_solve = L.ckpt_per_solve;
    p->tape_u_per_solve = L.tape_u_per_solve;
    p->tape_h_per_solve = L.tape_h_per_solve;
    p->grad_error = L.grad_error;
    p->q = L.q;
    p->gamma_full_sym = L.gamma_full_sym;
    if (L.mm) {
        const int M = d.modes;
        auto G = [&](int i, int j, int k, int l) { return d.gamma[i + M * (j + M * (k + M * l))]; };
        const bool sym = L.q != M * M;
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < M; ++j)
                if (L.qidx[i * M + j] != fh::pair_index(M, i, j, sym))
                    return fail(FH_EINVAL, "internal: output-pair order differs from the kernels' pair_index");
        std::vector<R> gr(L.outp.size() * L.inp.size());
        for (size_t qq = 0; qq < L.outp.size(); ++qq)
            for (size_t pp = 0; pp < L.inp.size(); ++pp) {
                const int i = L.outp[qq].first, j = L.outp[qq].second, k = L.inp[pp].first, l = L.inp[pp].second;
                gr[qq * L.inp.size() + pp] = (R)(k == l ? G(i, j, k, k) : G(i, j, k, l) + G(i, j, l, k));
            }
        if ((rc = dev_alloc(p, p->gr, sizeof(R) * gr.size()))) return rc;
        if ((rc = dev_alloc(p, p->qidx, sizeof(int) * L.qidx.size()))) return rc;
        if ((rc = upload(p->gr.p, gr.data(), p->gr.bytes))) return rc;
        if ((rc = upload(p->qidx.p, L.qidx.data(), p->qidx.bytes))) return rc;
    }
    const size_t mb = (size_t)d.max_batch;
    if ((rc = dev_alloc(p, p->work, 2 * sizeof(R) * p->work_per_solve * mb))) return rc;
    if ((rc = dev_alloc(p, p->status, sizeof(int) * mb))) return rc;
    const size_t field = 2 * sizeof(R) * (size_t)nt * d.modes * mb;
    if ((rc = dev_alloc(p, p->u_in, field))) return rc;
    if ((rc = dev_alloc(p, p->u_out, field))) return rc;
    if (p->adaptive) {
        if ((rc = dev_alloc(p, p->hs, sizeof(double) * mb * (size_t)p->max_steps))) return rc;
        if ((rc = dev_alloc(p, p->zs, sizeof(double) * mb * (size_t)p->max_steps))) return rc;
        if ((rc = dev_alloc(p, p->sstats, sizeof(long long) * 3 * mb))) return rc;
    }
    if (p->total_bytes != L.forward_bytes) return fail(FH_EINVAL, "internal: plan bytes differ from the plan layout");
    return FH_OK;
}

template <typename R> fh::TsArgs<R> ts_args(fh_plan p, const fh::SmArgs<R>& s)
{
    fh::TsArgs<R> a{};
    a.s = s;
    a.rtol = p->d.rtol;
    a.atol = p->d.atol;
    a.max_steps = p->max_steps;
    a.hs = (double*)p->hs.p;
    a.zs = (double*)p->zs.p;
    a.stats = (long long*)p->sstats.p;
    return a;
}

template <typename R> fh::SmArgs<R> base_args(fh_plan p)
{
    const fh_problem_desc& d = p->d;
    fh::SmArgs<R> a{};
    a.work = (fh::C2<R>*)p->work.p;
    a.disp = (const double*)p->disp.p;
    a.hn = (const fh::C2<R>*)p->hn.p;
    a.shn = (const R*)p->shn.p;
    a.tw = (const fh::C2<R>*)p->tw.p;
    a.status = (int*)p->status.p;
    a.gamma = (R)p->gamma0;
    a.one_m_fr = (R)d.one_m_fr;
    a.h = d.length_m / d.steps;
    a.length = d.length_m;
    a.steps = d.steps;
    a.every = p->every;
    a.nseg = p->nseg;
    a.ntaped = p->ntaped;
    a.log_cost = d.log_cost;
    a.mask = (const unsigned char*)p->mask.p;
    a.phase = 3;
    return a;
}

// Multimode gradient kernels keep the adjoint fields in the work slice too,
// so the slice grows to their size before the first gradient launch.
template <typename R> int ensure_grad_work(fh_plan p)
{
    if (!p->grad_error.empty()) return fail(FH_EINVAL, p->grad_error);
    if (p->grad_work || !p->mm) return FH_OK;
    total_release(p, p->work);
    int rc = dev_alloc(p, p->work, 2 * sizeof(R) * p->grad_work_per_solve * (size_t)p->d.max_batch);
    if (rc) return rc;
    p->grad_work = true;
    return FH_OK;
}

template <typename R> int ensure_adjoint_buffers(fh_plan p)
{
    if (p->ckpt.p) return FH_OK;
    const size_t nt = (size_t)p->d.nt, mb = (size_t)p->d.max_batch, V = 2 * sizeof(R), M = (size_t)p->d.modes;
    int rc = ensure_grad_work<R>(p);
    if (rc) return rc;
    if ((rc = dev_alloc(p, p->ckpt, p->ckpt_per_solve * mb))) return rc;
    if ((rc = dev_alloc(p, p->tape_u, p->tape_u_per_solve * mb))) return rc;
    if ((rc = dev_alloc(p, p->tape_h, p->tape_h_per_solve * mb))) return rc;
    if ((rc = dev_alloc(p, p->g_lam0, V * nt * M * mb))) return rc;
    if ((rc = dev_alloc(p, p->g_grad, sizeof(double) * nt * mb))) return rc;
    if ((rc = dev_alloc(p, p->g_cost, sizeof(double) * mb))) return rc;
    if ((rc = dev_alloc(p, p->g_seed, V * nt * M * mb))) return rc;
    if (p->total_bytes != p->gradient_bytes)
        return fail(FH_EINVAL, "internal: gradient plan bytes differ from the plan layout");
    return FH_OK;
}

// Device buffers of one gradient launch beyond its inputs and outputs: the
// plan's own for fh_gradient*, a tape's for fh_forward_tape / fh_adjoint.
struct GradBufs {
    void *ckpt, *tape_u, *tape_h;
    double *hs, *zs;
    long long* stats;
    int* status;
    int phase;                // SmArgs::phase
    void* work = nullptr;     // workgroup slices (null: the plan's)
    bool fill_status = false; // clear status with a fill kernel instead of hipMemsetAsync
};

GradBufs plan_grad_bufs(fh_plan p)
{
    return GradBufs{p->ckpt.p, p->tape_u.p, p->tape_h.p, (double*)p->hs.p, (double*)p->zs.p,
                    (long long*)p->sstats.p, (int*)p->status.p, 3};
}

__global__ __launch_bounds__(256) void k_zero_words(int* __restrict__ p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

int clear_status(const GradBufs& gb, int batch, hipStream_t s)
{
    if (gb.fill_status) {
        hipLaunchKernelGGL(k_zero_words, dim3((batch + 255) / 256), dim3(256), 0, s, gb.status, batch);
        FH_HIP(hipGetLastError());
    } else {
        FH_HIP(hipMemsetAsync(gb.status, 0, sizeof(int) * (size_t)batch, s));
    }
    return FH_OK;
}

template <typename R>
int launch_gradient_with(fh_plan p, const GradBufs& gb, int batch, const void* d_u0, const void* d_seed,
                         void* d_uL, void* d_lam0, double* d_grad, double* d_cost, hipStream_t s,
                         const double* d_lengths = nullptr, const int* d_nsteps = nullptr)
{
    if (!d_seed && !p->has_mask)
        return fail(FH_EINVAL, "gradient needs a terminal seed or a plan band mask");
    if (p->mm) {
        if (!p->gamma_full_sym)
            return fail(FH_ENOTSYM, "fiber adjoint requires a fully permutation-symmetric gamma tensor");
        if (!p->adaptive && !d_lengths != !d_nsteps)
            return fail(FH_EINVAL, "multimode per-solve lengths need per-solve step counts");
    }
    if (p->mm) {
        const fh_problem_desc& d = p->d;
        fh::MmArgs<R> m{};
        m.u0 = (const fh::C2<R>*)d_u0;
        m.out = (fh::C2<R>*)d_uL;
        m.work = (fh::C2<R>*)(gb.work ? gb.work : p->work.p);
        m.disp = (const double*)p->disp.p;
        m.hn = (const fh::C2<R>*)p->hn.p;
        m.shn = (const R*)p->shn.p;
        m.tw = (const fh::C2<R>*)p->tw.p;
        m.gr = (const R*)p->gr.p;
        m.qidx = (const int*)p->qidx.p;
        m.status = gb.status;
        m.one_m_fr = (R)d.one_m_fr;
        m.h = d.length_m / d.steps;
        m.length = d.length_m;
        m.steps = d.steps;
        m.q = p->q;
        m.seed = (const fh::C2<R>*)d_seed;
        m.mask = (const unsigned char*)p->mask.p;
        m.log_cost = d.log_cost;
        m.lam0 = (fh::C2<R>*)d_lam0;
        m.grad = d_grad;
        m.cost = d_cost;
        m.ckpt = (fh::C2<R>*)gb.ckpt;
        m.tape_u = (fh::C2<R>*)gb.tape_u;
        m.tape_h = (fh::C2<R>*)gb.tape_h;
        m.phase = gb.phase;
        m.every = p->every;
        m.nseg = p->nseg;
        m.ntaped = p->ntaped;
        m.lengths = d_lengths;
        m.nsteps = d_nsteps;  // each at most the plan's steps (checkpoint capacity)
        if ((gb.phase & 1) && clear_status(gb, batch, s)) return FH_EDEVICE;
        hipError_t e;
        if (p->adaptive) {
            m.h = 0.0;
            m.steps = 0;
            m.nsteps = nullptr;
            fh::MtArgs<R> t{};
            t.m = m;
            t.rtol = d.rtol;
            t.atol = d.atol;
            t.max_steps = p->max_steps;
            t.hs = gb.hs;
            t.zs = gb.zs;
            t.stats = gb.stats;
            e = fh::launch_tsit5_multi(d.nt, d.modes, true, t, batch, s);
        } else {
            e = fh::launch_multi_grad(d.nt, d.modes, m, batch, s);
        }
        if (e != hipSuccess) return fail(FH_EDEVICE, std::string("multimode gradient launch: ") + hipGetErrorString(e));
        return FH_OK;
    }
    fh::SmArgs<R> a = base_args<R>(p);
    a.u0 = (const fh::C2<R>*)d_u0;
    a.seed = (const fh::C2<R>*)d_seed;
    a.out = (fh::C2<R>*)d_uL;
    a.lam0 = (fh::C2<R>*)d_lam0;
    a.grad = d_grad;
    a.cost = d_cost;
    a.ckpt = (fh::C2<R>*)gb.ckpt;
    a.tape_u = (fh::C2<R>*)gb.tape_u;
    a.tape_h = (R*)gb.tape_h;
    a.status = gb.status;
    a.phase = gb.phase;
    if (gb.work) a.work = (fh::C2<R>*)gb.work;
    if (d_lengths && !p->adaptive) return fail(FH_EINVAL, "per-solve lengths need the adaptive scheme");
    if ((gb.phase & 1) && clear_status(gb, batch, s)) return FH_EDEVICE;
    fh::TsArgs<R> ta = ts_args<R>(p, a);
    ta.hs = gb.hs;
    ta.zs = gb.zs;
    ta.stats = gb.stats;
    ta.lengths = d_lengths;
    hipError_t e = p->adaptive ? fh::launch_tsit5(p->d.nt, true, ta, batch, s)
                               : with_size<R, OpGrad>(p->d.nt, a, batch, s);
    if (e != hipSuccess) return fail(FH_EDEVICE, std::string("gradient launch: ") + hipGetErrorString(e));
    return FH_OK;
}

template <typename R>
int launch_gradient(fh_plan p, int batch, const void* d_u0, const void* d_seed, void* d_uL, void* d_lam0,
                    double* d_grad, double* d_cost, hipStream_t s, const double* d_lengths = nullptr,
                    const int* d_nsteps = nullptr)
{
    if (!d_seed && !p->has_mask)
        return fail(FH_EINVAL, "gradient needs a terminal seed or a plan band mask");
    int rc = ensure_adjoint_buffers<R>(p);
    if (rc) return rc;
    GradBufs gb = plan_grad_bufs(p);
    if (d_seed || p->d.modes == 1 || p->d.objective == FH_OBJ_BAND_SUM)
        return launch_gradient_with<R>(p, gb, batch, d_u0, d_seed, d_uL, d_lam0, d_grad, d_cost, s, d_lengths,
                                       d_nsteps);
    // another multimode objective: forward sweep, its seed on the device from
    // the output, backward sweep
    if ((p->d.modes > 1 && !p->gamma_full_sym))
        return fail(FH_ENOTSYM, "fiber adjoint requires a fully permutation-symmetric gamma tensor");
    gb.phase = 1;
    if ((rc = launch_gradient_with<R>(p, gb, batch, d_u0, p->g_seed.p, d_uL, d_lam0, d_grad, d_cost, s, d_lengths,
                                      d_nsteps)))
        return rc;
    const double tau = p->d.objective_tau > 0 ? p->d.objective_tau : 50.0;
    hipError_t e = fh::launch_objective(sizeof(R) == 8, d_uL, (const unsigned char*)p->mask.p, p->d.nt, p->d.modes,
                                        p->d.objective, tau, p->d.log_cost, p->g_seed.p, d_cost, batch, s);
    if (e != hipSuccess) return fail(FH_EDEVICE, std::string("objective launch: ") + hipGetErrorString(e));
    gb.phase = 2;
    return launch_gradient_with<R>(p, gb, batch, d_u0, p->g_seed.p, d_uL, d_lam0, d_grad, d_cost, s, d_lengths,
                                   d_nsteps);
}

template <typename R>
int launch_forward_multi(fh_plan p, int batch, const void* d_u0, void* d_uL, hipStream_t s,
                         const double* d_lengths = nullptr, const int* d_nsteps = nullptr,
                         const double* d_zsave = nullptr, int nz = 0, void* d_saved = nullptr)
{
    const fh_problem_desc& d = p->d;
    fh::MmArgs<R> a{};
    a.u0 = (const fh::C2<R>*)d_u0;
    a.out = (fh::C2<R>*)d_uL;
    a.work = (fh::C2<R>*)p->work.p;
    a.disp = (const double*)p->disp.p;
    a.hn = (const fh::C2<R>*)p->hn.p;
    a.shn = (const R*)p->shn.p;
    a.tw = (const fh::C2<R>*)p->tw.p;
    a.gr = (const R*)p->gr.p;
    a.qidx = (const int*)p->qidx.p;
    a.status = (int*)p->status.p;
    a.one_m_fr = (R)d.one_m_fr;
    a.h = d.length_m / d.steps;
    a.length = d.length_m;
    a.steps = d.steps;
    a.q = p->q;
    a.lengths = d_lengths;
    a.nsteps = d_nsteps;
    FH_HIP(hipMemsetAsync(p->status.p, 0, sizeof(int) * (size_t)batch, s));
    if (p->adaptive) {
        fh::MtArgs<R> t{};
        t.m = a;
        t.rtol = d.rtol;
        t.atol = d.atol;
        t.max_steps = p->max_steps;
        t.hs = (double*)p->hs.p;
        t.zs = (double*)p->zs.p;
        t.stats = (long long*)p->sstats.p;
        t.zsave = d_zsave;
        t.nz = nz;
        t.saved = (fh::C2<R>*)d_saved;
        hipError_t e = fh::launch_tsit5_multi(d.nt, d.modes, false, t, batch, s);
        if (e != hipSuccess) return fail(FH_EDEVICE, std::string("multimode launch: ") + hipGetErrorString(e));
        return FH_OK;
    }
    hipError_t e = fh::launch_multi(d.nt, d.modes, a, batch, s);
    if (e != hipSuccess) return fail(FH_EDEVICE, std::string("multimode launch: ") + hipGetErrorString(e));
    return FH_OK;
}

template <typename R>
int launch_forward(fh_plan p, int batch, const void* d_u0, void* d_uL, hipStream_t s,
                   const double* d_lengths = nullptr, const int* d_nsteps = nullptr, const double* d_zsave = nullptr,
                   int nz = 0, void* d_saved = nullptr)
{
    const fh_problem_desc& d = p->d;
    if (p->mm)
        return launch_forward_multi<R>(p, batch, d_u0, d_uL, s, d_lengths, d_nsteps, d_zsave, nz, d_saved);
    if (d_lengths && !p->adaptive) return fail(FH_EINVAL, "per-solve lengths need the adaptive scheme");
    fh::SmArgs<R> a = base_args<R>(p);
    a.u0 = (const fh::C2<R>*)d_u0;
    a.out = (fh::C2<R>*)d_uL;
    FH_HIP(hipMemsetAsync(p->status.p, 0, sizeof(int) * (size_t)batch, s));
    fh::TsArgs<R> ta = ts_args<R>(p, a);
    ta.lengths = d_lengths;
    ta.zsave = d_zsave;
    ta.nz = nz;
    ta.saved = (fh::C2<R>*)d_saved;
    hipError_t e = p->adaptive ? fh::launch_tsit5(d.nt, false, ta, batch, s)
                               : with_size<R, OpLaunch>(d.nt, a, batch, s);
    if (e != hipSuccess) return fail(FH_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return FH_OK;
}

// Per-solve counters of the last call (device counters for the adaptive
// scheme, the fixed step count otherwise).
int read_solve_stats(fh_plan p, int batch, std::vector<long long>& st)
{
    st.assign(3 * (size_t)batch, 0);
    if (p->adaptive) {
        FH_HIP(hipMemcpy(st.data(), p->sstats.p, sizeof(long long) * 3 * (size_t)batch, hipMemcpyDeviceToHost));
    } else {
        for (int b = 0; b < batch; ++b) {
            st[3 * b] = p->d.steps;
            st[3 * b + 2] = 4LL * p->d.steps;
        }
    }
    return FH_OK;
}

int finish_call(fh_plan p, int batch, const std::vector<int>& status, fh_stats* stats, bool grad)
{
    int bad = 0;
    for (int b = 0; b < batch; ++b) bad |= status[b];
    std::vector<long long> st;
    int rc = read_solve_stats(p, batch, st);
    if (rc) return rc;
    const int code = (bad & 2) ? FH_EMAXSTEPS : (bad ? FH_ENONFINITE : FH_OK);
    if (stats) {
        stats->accepted = stats->rejected = stats->rhs_evals = 0;
        for (int b = 0; b < batch; ++b) {
            stats->accepted += st[3 * b];
            stats->rejected += st[3 * b + 1];
            stats->rhs_evals += st[3 * b + 2];
        }
        stats->retcode = code;
    }
    if (code == FH_EMAXSTEPS)
        return fail(code, "adaptive propagation reached max_steps=" + std::to_string(p->max_steps) +
                              " accepted steps before the fiber end");
    if (code)
        return fail(code, grad ? "adjoint propagation produced non-finite fields"
                               : "propagation produced non-finite fields");
    return FH_OK;
}

// ------------------------------------------------------------ tapes

void release_tape_storage(fh_tape t)
{
    for (DevBuf* b : {&t->u0, &t->uL, &t->ckpt, &t->tape_u, &t->tape_h, &t->hs, &t->zs, &t->sstats, &t->status,
                      &t->seed, &t->lam0, &t->kgrad, &t->gphase, &t->gamp, &t->edot, &t->cost})
        dev_free(*b);
    delete t;
}

int tape_alloc(fh_tape t, DevBuf& b, size_t bytes)
{
    bytes = alloc_size(bytes);
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(e == hipErrorOutOfMemory ? FH_ENOMEM : FH_EDEVICE,
                    std::string("hipMalloc(") + std::to_string(bytes) + ") for a tape: " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    t->bytes += (int64_t)bytes;
    return FH_OK;
}

// Storage for `batch` kept solves: the plan's spare when it is large enough.
int new_tape(fh_plan p, int batch, fh_tape* out)
{
    if (p->spare && p->spare->capacity >= batch) {
        *out = p->spare;
        p->spare = nullptr;
    } else {
        if (p->spare) release_tape_storage(p->spare);
        p->spare = nullptr;
        fh_tape t = new fh_tape_s;
        const size_t nt = (size_t)p->d.nt, M = (size_t)p->d.modes, B = (size_t)batch;
        const size_t V = p->d.precision == FH_PREC_FP64 ? 16 : 8, field = V * nt * M * B;
        const size_t steps = p->adaptive ? (size_t)p->max_steps : 0;
        int rc = FH_OK;
        for (auto& pr : std::vector<std::pair<DevBuf*, size_t>>{
                 {&t->u0, field}, {&t->uL, field}, {&t->seed, field}, {&t->lam0, field},
                 {&t->ckpt, p->ckpt_per_solve * B}, {&t->tape_u, p->tape_u_per_solve * B},
                 {&t->tape_h, p->tape_h_per_solve * B}, {&t->hs, sizeof(double) * steps * B},
                 {&t->zs, sizeof(double) * steps * B}, {&t->sstats, sizeof(long long) * 3 * B},
                 {&t->status, sizeof(int) * B}, {&t->kgrad, sizeof(double) * nt * B},
                 {&t->gphase, sizeof(double) * nt * M * B}, {&t->gamp, sizeof(double) * nt * M * B},
                 {&t->edot, sizeof(double) * B}, {&t->cost, sizeof(double) * B}})
            if (rc == FH_OK) rc = tape_alloc(t, *pr.first, pr.second);
        if (rc) {
            release_tape_storage(t);
            return rc;
        }
        t->capacity = batch;
        *out = t;
    }
    (*out)->plan = p;
    (*out)->batch = batch;
    (*out)->swept = false;
    return FH_OK;
}

GradBufs tape_grad_bufs(fh_tape t, int phase)
{
    return GradBufs{t->ckpt.p, t->tape_u.p, t->tape_h.p, (double*)t->hs.p, (double*)t->zs.p,
                    (long long*)t->sstats.p, (int*)t->status.p, phase};
}

// Forward sweep of a tape from its kept launch (phase 1 on the tape's buffers).
int tape_forward(fh_tape t, hipStream_t s)
{
    fh_plan p = t->plan;
    const GradBufs gb = tape_grad_bufs(t, 1);
    // the seed pointer is not read by a forward-only launch; passing the tape's
    // seed buffer keeps the band-energy accumulation off the (optional) mask
    return p->d.precision == FH_PREC_FP64
               ? launch_gradient_with<double>(p, gb, t->batch, t->u0.p, t->seed.p, t->uL.p, t->lam0.p,
                                              (double*)t->kgrad.p, (double*)t->cost.p, s)
               : launch_gradient_with<float>(p, gb, t->batch, t->u0.p, t->seed.p, t->uL.p, t->lam0.p,
                                             (double*)t->kgrad.p, (double*)t->cost.p, s);
}

// Status and counters of a tape's forward sweep.
int tape_finish(fh_tape t, fh_stats* stats, bool grad)
{
    fh_plan p = t->plan;
    const int batch = t->batch;
    std::vector<int> status(batch);
    std::vector<long long> st(3 * (size_t)batch, 0);
    FH_HIP(hipMemcpy(status.data(), t->status.p, sizeof(int) * batch, hipMemcpyDeviceToHost));
    if (p->adaptive) {
        FH_HIP(hipMemcpy(st.data(), t->sstats.p, sizeof(long long) * 3 * batch, hipMemcpyDeviceToHost));
    } else {
        for (int b = 0; b < batch; ++b) {
            st[3 * b] = p->d.steps;
            st[3 * b + 2] = 4LL * p->d.steps;
        }
    }
    int bad = 0;
    for (int b = 0; b < batch; ++b) bad |= status[b];
    const int code = (bad & 2) ? FH_EMAXSTEPS : (bad ? FH_ENONFINITE : FH_OK);
    if (stats) {
        stats->accepted = stats->rejected = stats->rhs_evals = 0;
        for (int b = 0; b < batch; ++b) {
            stats->accepted += st[3 * b];
            stats->rejected += st[3 * b + 1];
            stats->rhs_evals += st[3 * b + 2];
        }
        stats->retcode = code;
    }
    if (code == FH_EMAXSTEPS)
        return fail(code, "adaptive propagation reached max_steps=" + std::to_string(p->max_steps) +
                              " accepted steps before the fiber end");
    if (code)
        return fail(code, grad ? "adjoint propagation produced non-finite fields"
                               : "propagation produced non-finite fields");
    return FH_OK;
}

}  // namespace

// Internal interface for the request server (fh_server.hip); C++ linkage.
namespace fhi {

bool plan_fp64(fh_plan p) { return p->d.precision == FH_PREC_FP64; }
int plan_nt(fh_plan p) { return p->d.nt; }
int plan_modes(fh_plan p) { return p->d.modes; }
int plan_max_batch(fh_plan p) { return p->d.max_batch; }
bool plan_adaptive(fh_plan p) { return p->adaptive; }
void* plan_u_in(fh_plan p) { return p->u_in.p; }
void* plan_u_out(fh_plan p) { return p->u_out.p; }
int64_t plan_bytes(fh_plan p) { return p->total_bytes; }
int set_error(int code, const std::string& msg) { return fail(code, msg); }

// Allocate the adjoint buffers now, so the plan's footprint is final.
int plan_prepare_gradient(fh_plan p)
{
    FH_HIP(hipSetDevice(p->ctx->device));
    return plan_fp64(p) ? ensure_adjoint_buffers<double>(p) : ensure_adjoint_buffers<float>(p);
}

// Enqueue one batch on `s` from the plan's own input buffer (u_in) with
// per-solve lengths / step counts (device arrays, may be null), followed by
// asynchronous copies of the outputs and counters into host (pinned) memory:
// uL [batch][modes][nt] in the plan's precision, grad [batch][nt] and cost
// [batch] (gradient), status [batch] and stats [batch][3].
int enqueue_batch(fh_plan p, int batch, bool grad, const double* d_lengths, const int* d_nsteps, hipStream_t s,
                  void* h_uL, double* h_grad, double* h_cost, int* h_status, long long* h_stats)
{
    const bool fp64 = plan_fp64(p);
    int rc;
    if (grad) {
        if ((rc = plan_prepare_gradient(p))) return rc;
        rc = fp64 ? launch_gradient<double>(p, batch, p->u_in.p, nullptr, p->u_out.p, p->g_lam0.p,
                                            (double*)p->g_grad.p, (double*)p->g_cost.p, s, d_lengths, d_nsteps)
                  : launch_gradient<float>(p, batch, p->u_in.p, nullptr, p->u_out.p, p->g_lam0.p,
                                           (double*)p->g_grad.p, (double*)p->g_cost.p, s, d_lengths, d_nsteps);
    } else {
        rc = fp64 ? launch_forward<double>(p, batch, p->u_in.p, p->u_out.p, s, d_lengths, d_nsteps)
                  : launch_forward<float>(p, batch, p->u_in.p, p->u_out.p, s, d_lengths, d_nsteps);
    }
    if (rc) return rc;
    const size_t n = (size_t)batch * p->d.nt * p->d.modes;
    FH_HIP(hipMemcpyAsync(h_uL, p->u_out.p, n * 2 * (fp64 ? sizeof(double) : sizeof(float)), hipMemcpyDeviceToHost, s));
    if (grad) {
        FH_HIP(hipMemcpyAsync(h_grad, p->g_grad.p, sizeof(double) * (size_t)batch * p->d.nt, hipMemcpyDeviceToHost, s));
        FH_HIP(hipMemcpyAsync(h_cost, p->g_cost.p, sizeof(double) * (size_t)batch, hipMemcpyDeviceToHost, s));
    }
    FH_HIP(hipMemcpyAsync(h_status, p->status.p, sizeof(int) * (size_t)batch, hipMemcpyDeviceToHost, s));
    if (p->adaptive)
        FH_HIP(hipMemcpyAsync(h_stats, p->sstats.p, sizeof(long long) * 3 * (size_t)batch, hipMemcpyDeviceToHost, s));
    return FH_OK;
}

}  // namespace fhi

// ------------------------------------------- allocation-free gradient
//
// fh_gradient_device_noalloc keeps every buffer its kernels need beyond the
// caller's inputs and outputs (workgroup slices, status words, checkpoints,
// stage tape, adaptive step records, a device objective's seed) in this
// fixed array instead of plan memory.  Each GPU holds its own copy for the
// life of the process; fh_context_create resolves its device address once per
// device.  In host code `scratch` names a host-side stand-in, so only the
// resolved address is ever passed to kernels.
constexpr size_t kNoallocScratchBytes = 256ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), resolved on first use.
int resolve_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(FH_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return FH_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(FH_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return FH_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Regions of the scratch array for `batch` solves of plan p, laid out like the
// plan's own buffers (per-solve strides of its layout).
struct ScratchLayout {
    size_t status, work, ckpt, tape_u, tape_h, hs, zs, stats, seed, total;
};

ScratchLayout scratch_layout(fh_plan p, int batch)
{
    const size_t B = (size_t)batch, nt = (size_t)p->d.nt, M = (size_t)p->d.modes;
    const size_t V = p->d.precision == FH_PREC_FP64 ? 16 : 8;
    const size_t work = p->mm ? p->grad_work_per_solve : p->work_per_solve;
    const size_t steps = p->adaptive ? (size_t)p->max_steps : 0;
    ScratchLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += align256(bytes);
        return at;
    };
    L.status = take(sizeof(int) * B);
    L.work = take(V * work * B);
    L.ckpt = take(p->ckpt_per_solve * B);
    L.tape_u = take(p->tape_u_per_solve * B);
    L.tape_h = take(p->tape_h_per_solve * B);
    L.hs = take(sizeof(double) * steps * B);
    L.zs = take(sizeof(double) * steps * B);
    L.stats = take(sizeof(long long) * 3 * B);
    L.seed = take(M > 1 && p->d.objective != FH_OBJ_BAND_SUM ? V * nt * M * B : 0);
    L.total = off;
    return L;
}

template <typename R>
int gradient_noalloc(fh_plan p, int batch, const void* d_u0, const void* d_seed, void* d_uL, void* d_lam0,
                     double* d_grad, double* d_cost, hipStream_t s, unsigned char* base, const ScratchLayout& L)
{
    GradBufs gb{base + L.ckpt, base + L.tape_u, base + L.tape_h, (double*)(base + L.hs), (double*)(base + L.zs),
                (long long*)(base + L.stats), (int*)(base + L.status), 3, base + L.work, true};
    if (d_seed || p->d.modes == 1 || p->d.objective == FH_OBJ_BAND_SUM)
        return launch_gradient_with<R>(p, gb, batch, d_u0, d_seed, d_uL, d_lam0, d_grad, d_cost, s);
    void* seed = base + L.seed;  // multimode objective other than the band sum: seed from the output
    gb.phase = 1;
    int rc = launch_gradient_with<R>(p, gb, batch, d_u0, seed, d_uL, d_lam0, d_grad, d_cost, s);
    if (rc) return rc;
    const double tau = p->d.objective_tau > 0 ? p->d.objective_tau : 50.0;
    hipError_t e = fh::launch_objective(sizeof(R) == 8, d_uL, (const unsigned char*)p->mask.p, p->d.nt, p->d.modes,
                                        p->d.objective, tau, p->d.log_cost, seed, d_cost, batch, s);
    if (e != hipSuccess) return fail(FH_EDEVICE, std::string("objective launch: ") + hipGetErrorString(e));
    gb.phase = 2;
    return launch_gradient_with<R>(p, gb, batch, d_u0, seed, d_uL, d_lam0, d_grad, d_cost, s);
}

}  // namespace

extern "C" {

const char* fh_last_error(void) { return g_last_error.c_str(); }

int fh_gradient_device_noalloc(fh_plan p, int32_t batch, const void* d_u0, const void* d_seed, void* d_uL,
                               void* d_lam0, double* d_grad, double* d_cost, void* stream)
{
    if (!p || !d_u0 || !d_uL || !d_lam0 || !d_grad || !d_cost) return fail(FH_EINVAL, "NULL argument");
    if (batch < 1 || batch > p->d.max_batch) return fail(FH_EINVAL, "batch must lie in [1, max_batch]");
    if (!d_seed && !p->has_mask) return fail(FH_EINVAL, "gradient needs a terminal seed or a plan band mask");
    if (!p->grad_error.empty()) return fail(FH_EINVAL, p->grad_error);
    if (p->d.modes > 1 && !p->gamma_full_sym)
        return fail(FH_ENOTSYM, "fiber adjoint requires a fully permutation-symmetric gamma tensor");
    const int dev = p->ctx->device;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = dev >= 0 && dev < kMaxDevices ? g_scratch_ptr[dev] : nullptr;
    }
    if (!base) return fail(FH_EDEVICE, "scratch of this device is not resolved (create its context first)");
    const ScratchLayout L = scratch_layout(p, batch);
    if (L.total > kNoallocScratchBytes)
        return fail(FH_EINVAL, "this gradient needs " + std::to_string(L.total) + " bytes of scratch for " +
                                   std::to_string(batch) + " solves, above the " +
                                   std::to_string(kNoallocScratchBytes) + " of the allocation-free entry point; "
                                   "lower the batch, checkpoint_every or tape_segments");
    FH_HIP(hipSetDevice(dev));
    hipStream_t s = stream ? (hipStream_t)stream : p->ctx->stream;
    unsigned char* b = (unsigned char*)base;
    return p->d.precision == FH_PREC_FP64
               ? gradient_noalloc<double>(p, batch, d_u0, d_seed, d_uL, d_lam0, d_grad, d_cost, s, b, L)
               : gradient_noalloc<float>(p, batch, d_u0, d_seed, d_uL, d_lam0, d_grad, d_cost, s, b, L);
}

int fh_context_create(int device, fh_context* out)
{
    if (!out) return fail(FH_EINVAL, "out is NULL");
    int n = 0;
    FH_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(FH_EINVAL, "no such device");
    FH_HIP(hipSetDevice(device));
    int rc = resolve_scratch(device);  // the allocation-free entry point's array on this device
    if (rc) return rc;
    fh_context c = new fh_context_s;
    c->device = device;
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        delete c;
        return fail(FH_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return FH_OK;
}

int fh_context_destroy(fh_context ctx)
{
    if (!ctx) return FH_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FH_OK;
}

}  // extern "C"

namespace fhi {

namespace {
constexpr size_t kGranule = 256;
size_t granules(size_t bytes) { return (bytes + kGranule - 1) / kGranule * kGranule; }
}  // namespace

int Arena::init(size_t bytes, bool pinned_host)
{
    size = granules(bytes);
    host = pinned_host;
    const hipError_t e = host ? hipHostMalloc((void**)&base, size, hipHostMallocDefault) : hipMalloc((void**)&base, size);
    if (e != hipSuccess) {
        base = nullptr;
        size = 0;
        return fail(e == hipErrorOutOfMemory ? FH_ENOMEM : FH_EDEVICE,
                    "hipMalloc of the server's device memory (" + std::to_string(bytes) + " bytes): " +
                        hipGetErrorString(e));
    }
    free_.clear();
    free_[0] = size;
    return FH_OK;
}

void Arena::destroy()
{
    if (base) (void)(host ? hipHostFree(base) : hipFree(base));
    base = nullptr;
    size = 0;
    free_.clear();
}

void* Arena::alloc(size_t bytes)
{
    const size_t n = granules(bytes);
    std::lock_guard<std::mutex> lk(mu);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < n) continue;
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return base + off;
    }
    return nullptr;
}

void Arena::release(void* p, size_t bytes)
{
    if (!p) return;
    size_t off = (size_t)((char*)p - base), n = granules(bytes);
    std::lock_guard<std::mutex> lk(mu);
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + n == next->first) {  // merge with the following free range
        n += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {  // merge with the preceding free range
            prev->second += n;
            return;
        }
    }
    free_[off] = n;
}

// fh_plan_create with the table uploads on `copy_stream` (null: the context
// stream) and the device buffers from `arena` (null: hipMalloc).  The request
// server passes the idle lane that will run the plan and its arena.
int plan_create_on(fh_context ctx, const fh_problem_desc* desc, hipStream_t copy_stream, Arena* arena, Arena* stage,
                   fh_plan* out)
{
    if (!ctx || !out) return fail(FH_EINVAL, "context or out is NULL");
    int rc = check_desc(desc);
    if (rc) return rc;
    PlanLayout L;
    if ((rc = compute_layout(*desc, L))) return rc;
    FH_HIP(hipSetDevice(ctx->device));
    fh_plan p = new fh_plan_s;
    p->ctx = ctx;
    p->d = *desc;  // array pointers are only read during creation
    p->gamma0 = desc->gamma[0];
    p->adaptive = desc->scheme == FH_SCHEME_TSIT5;
    p->gradient_bytes = L.gradient_bytes;
    p->copy_stream = copy_stream;
    p->arena = arena;
    p->stage = stage;
    rc = desc->precision == FH_PREC_FP64 ? upload_tables<double>(p, L) : upload_tables<float>(p, L);
    if (rc && !p->staged.empty()) (void)hipStreamSynchronize(copy_stream);  // no upload left in flight
    p->copy_stream = nullptr;
    p->d.dispersion = p->d.gamma = p->d.raman = p->d.shock = nullptr;
    p->d.band_mask = nullptr;
    if (rc) {
        fh_plan_destroy(p);
        return rc;
    }
    *out = p;
    return FH_OK;
}

}  // namespace fhi

extern "C" {

int fh_plan_create(fh_context ctx, const fh_problem_desc* desc, fh_plan* out)
{
    return fhi::plan_create_on(ctx, desc, nullptr, nullptr, nullptr, out);
}

int fh_plan_check(const fh_problem_desc* desc, int32_t gradient, int64_t* bytes)
{
    if (!bytes) return fail(FH_EINVAL, "bytes is NULL");
    int rc = check_desc(desc);
    if (rc) return rc;
    PlanLayout L;
    if ((rc = compute_layout(*desc, L))) return rc;
    if (gradient && !L.grad_error.empty()) return fail(FH_EINVAL, L.grad_error);
    *bytes = gradient ? L.gradient_bytes : L.forward_bytes;
    return FH_OK;
}

int fh_plan_destroy(fh_plan p)
{
    if (!p) return FH_OK;
    (void)hipSetDevice(p->ctx->device);
    if (p->spare) release_tape_storage(p->spare);
    for (DevBuf* b : {&p->disp, &p->hn, &p->shn, &p->tw, &p->work, &p->status, &p->u_in, &p->u_out, &p->mask,
                      &p->ckpt, &p->tape_u, &p->tape_h, &p->g_lam0, &p->g_grad, &p->g_cost, &p->g_seed, &p->gr, &p->qidx, &p->hs, &p->zs, &p->sstats})
        dev_free(*b);
    // staged uploads completed before any launch that followed them on the
    // copy stream; a plan is released only after its launches completed
    for (auto& c : p->staged) p->stage->release(c.first, c.second);
    delete p;
    return FH_OK;
}

int fh_plan_device_bytes(fh_plan p, int64_t* bytes)
{
    if (!p || !bytes) return fail(FH_EINVAL, "plan or bytes is NULL");
    *bytes = p->total_bytes;
    return FH_OK;
}

int fh_forward_device(fh_plan p, int32_t batch, const void* d_u0, void* d_uL, void* stream)
{
    if (!p || !d_u0 || !d_uL) return fail(FH_EINVAL, "plan or field pointer is NULL");
    if (batch < 1 || batch > p->d.max_batch)
        return fail(FH_EINVAL, "batch must lie in [1, max_batch=" + std::to_string(p->d.max_batch) + "]");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : p->ctx->stream;
    return p->d.precision == FH_PREC_FP64 ? launch_forward<double>(p, batch, d_u0, d_uL, s)
                                          : launch_forward<float>(p, batch, d_u0, d_uL, s);
}

int fh_forward(fh_plan p, int32_t batch, const double* u0, double* uL, fh_stats* stats)
{
    if (!p || !u0 || !uL) return fail(FH_EINVAL, "plan or field pointer is NULL");
    if (batch < 1 || batch > p->d.max_batch)
        return fail(FH_EINVAL, "batch must lie in [1, max_batch=" + std::to_string(p->d.max_batch) + "]");
    const size_t n = (size_t)batch * p->d.nt * p->d.modes;
    double norm2 = 0.0;
    for (size_t i = 0; i < 2 * n; ++i) {
        if (!std::isfinite(u0[i])) return fail(FH_EINVAL, "input spectrum must be finite");
        norm2 += u0[i] * u0[i];
    }
    if (!(norm2 > 0)) return fail(FH_EINVAL, "input spectrum must be nonzero");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    const bool fp64 = p->d.precision == FH_PREC_FP64;
    if (fp64) {
        FH_HIP(hipMemcpyAsync(p->u_in.p, u0, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
    } else {
        std::vector<float> tmp(2 * n);
        for (size_t i = 0; i < 2 * n; ++i) tmp[i] = (float)u0[i];
        FH_HIP(hipMemcpyAsync(p->u_in.p, tmp.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, s));
        FH_HIP(hipStreamSynchronize(s));
    }
    int rc = fp64 ? launch_forward<double>(p, batch, p->u_in.p, p->u_out.p, s)
                  : launch_forward<float>(p, batch, p->u_in.p, p->u_out.p, s);
    if (rc) return rc;
    std::vector<int> status(batch);
    FH_HIP(hipMemcpyAsync(status.data(), p->status.p, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
    if (fp64) {
        FH_HIP(hipMemcpyAsync(uL, p->u_out.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
        FH_HIP(hipStreamSynchronize(s));
    } else {
        std::vector<float> tmp(2 * n);
        FH_HIP(hipMemcpyAsync(tmp.data(), p->u_out.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, s));
        FH_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < 2 * n; ++i) uL[i] = tmp[i];
    }
    return finish_call(p, batch, status, stats, false);
}

namespace {

// Device memory owned by one call.
struct TmpBuf {
    void* p = nullptr;
    ~TmpBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// Host complex128 -> device field of the plan's precision (count complex entries).
int upload_field(bool fp64, void* dst, const double* src, size_t count, hipStream_t s)
{
    if (fp64) {
        FH_HIP(hipMemcpyAsync(dst, src, sizeof(double) * 2 * count, hipMemcpyHostToDevice, s));
        FH_HIP(hipStreamSynchronize(s));
    } else {
        std::vector<float> tmp(2 * count);
        for (size_t i = 0; i < 2 * count; ++i) tmp[i] = (float)src[i];
        FH_HIP(hipMemcpyAsync(dst, tmp.data(), sizeof(float) * 2 * count, hipMemcpyHostToDevice, s));
        FH_HIP(hipStreamSynchronize(s));
    }
    return FH_OK;
}

int download_field(bool fp64, double* dst, const void* src, size_t count, hipStream_t s)
{
    if (fp64) {
        FH_HIP(hipMemcpyAsync(dst, src, sizeof(double) * 2 * count, hipMemcpyDeviceToHost, s));
        FH_HIP(hipStreamSynchronize(s));
    } else {
        std::vector<float> tmp(2 * count);
        FH_HIP(hipMemcpyAsync(tmp.data(), src, sizeof(float) * 2 * count, hipMemcpyDeviceToHost, s));
        FH_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < 2 * count; ++i) dst[i] = tmp[i];
    }
    return FH_OK;
}

}  // namespace

int fh_forward_saveat(fh_plan p, int32_t batch, const double* u0, double* uL, int32_t nz, const double* zsave,
                      double* saved, fh_stats* stats)
{
    if (!p || !u0 || !uL || !zsave || !saved) return fail(FH_EINVAL, "plan, field or save pointer is NULL");
    if (!p->adaptive) return fail(FH_EUNSUPPORTED, "dense output (saveat) needs an adaptive (FH_SCHEME_TSIT5) plan");
    if (batch < 1 || batch > p->d.max_batch)
        return fail(FH_EINVAL, "batch must lie in [1, max_batch=" + std::to_string(p->d.max_batch) + "]");
    const size_t nt = (size_t)p->d.nt, nm = (size_t)p->d.modes * nt;
    const bool fp64 = p->d.precision == FH_PREC_FP64;
    const size_t esz = fp64 ? 16 : 8;
    if (nz < 1 || (size_t)nz * nm * esz >= ((size_t)1 << 31))
        return fail(FH_EINVAL, "nz must lie in [1, 2^31 / (modes * nt * element size))");
    for (int j = 0; j < nz; ++j) {
        if (!std::isfinite(zsave[j])) return fail(FH_EINVAL, "saveat distances must be finite");
        if (zsave[j] < 0 || zsave[j] > p->d.length_m)
            return fail(FH_EINVAL, "saveat distances must lie in [0, L]");
        if (j > 0 && !(zsave[j] > zsave[j - 1])) return fail(FH_EINVAL, "saveat distances must be sorted and unique");
    }
    const size_t n = (size_t)batch * nm;
    double norm2 = 0.0;
    for (size_t i = 0; i < 2 * n; ++i) {
        if (!std::isfinite(u0[i])) return fail(FH_EINVAL, "input spectrum must be finite");
        norm2 += u0[i] * u0[i];
    }
    if (!(norm2 > 0)) return fail(FH_EINVAL, "input spectrum must be nonzero");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    TmpBuf dz, dsv;
    FH_HIP(hipMalloc(&dz.p, sizeof(double) * (size_t)nz));
    FH_HIP(hipMalloc(&dsv.p, esz * (size_t)nz * n));
    FH_HIP(hipMemcpyAsync(dz.p, zsave, sizeof(double) * (size_t)nz, hipMemcpyHostToDevice, s));
    int rc = upload_field(fp64, p->u_in.p, u0, n, s);
    if (rc) return rc;
    rc = fp64 ? launch_forward<double>(p, batch, p->u_in.p, p->u_out.p, s, nullptr, nullptr, (const double*)dz.p,
                                       nz, dsv.p)
              : launch_forward<float>(p, batch, p->u_in.p, p->u_out.p, s, nullptr, nullptr, (const double*)dz.p, nz,
                                      dsv.p);
    if (rc) return rc;
    std::vector<int> status(batch);
    FH_HIP(hipMemcpyAsync(status.data(), p->status.p, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
    if ((rc = download_field(fp64, uL, p->u_out.p, n, s))) return rc;
    if ((rc = download_field(fp64, saved, dsv.p, (size_t)nz * n, s))) return rc;
    return finish_call(p, batch, status, stats, false);
}

int fh_forward_gain(fh_plan p, int32_t batch, const fh_gain_desc* g, const double* pump0, const double* u0,
                    double* uL, double* pumpL, fh_stats* stats)
{
    if (!p || !g || !pump0 || !u0 || !uL || !pumpL) return fail(FH_EINVAL, "plan, gain, pump or field pointer is NULL");
    if (!g->w12 || !g->w21 || !g->ga || !g->ge) return fail(FH_EINVAL, "gain coefficient array is NULL");
    if (!p->adaptive || p->d.modes != 1 || p->mm || !fh::gain_supported(p->d.nt))
        return fail(FH_EUNSUPPORTED, "pump + signal solves need a single-mode adaptive (FH_SCHEME_TSIT5) plan with "
                                     "256 <= nt <= 16384");
    if (batch < 1 || batch > p->d.max_batch)
        return fail(FH_EINVAL, "batch must lie in [1, max_batch=" + std::to_string(p->d.max_batch) + "]");
    const size_t nt = (size_t)p->d.nt, n = (size_t)batch * nt;
    const bool fp64 = p->d.precision == FH_PREC_FP64;
    if ((size_t)fh::gain_fields((int)nt, fp64) * nt > p->work_per_solve)
        return fail(FH_EINVAL, "internal: plan slice too small for the gain slot");
    std::vector<double> coef(4 * nt);
    for (size_t k = 0; k < nt; ++k) {
        coef[k] = g->w12[k];
        coef[nt + k] = g->w21[k];
        coef[2 * nt + k] = g->ga[k];
        coef[3 * nt + k] = g->ge[k];
    }
    for (double c : coef)
        if (!std::isfinite(c)) return fail(FH_EINVAL, "gain coefficients must be finite");
    for (double c : {g->r12, g->r21, g->inv_tau, g->gpa, g->gpe})
        if (!std::isfinite(c)) return fail(FH_EINVAL, "gain coefficients must be finite");
    for (int b = 0; b < batch; ++b)
        if (!std::isfinite(pump0[b])) return fail(FH_EINVAL, "pump power must be finite");
    double norm2 = 0.0;
    for (size_t i = 0; i < 2 * n; ++i) {
        if (!std::isfinite(u0[i])) return fail(FH_EINVAL, "input spectrum must be finite");
        norm2 += u0[i] * u0[i];
    }
    if (!(norm2 > 0)) return fail(FH_EINVAL, "input spectrum must be nonzero");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    TmpBuf dcoef, dp0, dpL;
    FH_HIP(hipMalloc(&dcoef.p, sizeof(double) * 4 * nt));
    FH_HIP(hipMalloc(&dp0.p, sizeof(double) * (size_t)batch));
    FH_HIP(hipMalloc(&dpL.p, sizeof(double) * (size_t)batch));
    FH_HIP(hipMemcpyAsync(dcoef.p, coef.data(), sizeof(double) * 4 * nt, hipMemcpyHostToDevice, s));
    FH_HIP(hipMemcpyAsync(dp0.p, pump0, sizeof(double) * (size_t)batch, hipMemcpyHostToDevice, s));
    int rc = upload_field(fp64, p->u_in.p, u0, n, s);
    if (rc) return rc;
    FH_HIP(hipMemsetAsync(p->status.p, 0, sizeof(int) * (size_t)batch, s));
    fh::GainDev gd;
    gd.coef = (const double*)dcoef.p;
    gd.r12 = g->r12;
    gd.r21 = g->r21;
    gd.inv_tau = g->inv_tau;
    gd.gpa = g->gpa;
    gd.gpe = g->gpe;
    gd.pump0 = (const double*)dp0.p;
    gd.pumpL = (double*)dpL.p;
    hipError_t e;
    if (fp64) {
        fh::SmArgs<double> a = base_args<double>(p);
        a.u0 = (const fh::C2<double>*)p->u_in.p;
        a.out = (fh::C2<double>*)p->u_out.p;
        fh::TsArgs<double> ta = ts_args<double>(p, a);
        ta.gain = gd;
        e = fh::launch_tsit5_gain((int)nt, ta, batch, s);
    } else {
        fh::SmArgs<float> a = base_args<float>(p);
        a.u0 = (const fh::C2<float>*)p->u_in.p;
        a.out = (fh::C2<float>*)p->u_out.p;
        fh::TsArgs<float> ta = ts_args<float>(p, a);
        ta.gain = gd;
        e = fh::launch_tsit5_gain((int)nt, ta, batch, s);
    }
    if (e != hipSuccess) return fail(FH_EDEVICE, std::string("gain launch: ") + hipGetErrorString(e));
    std::vector<int> status(batch);
    FH_HIP(hipMemcpyAsync(status.data(), p->status.p, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
    FH_HIP(hipMemcpyAsync(pumpL, dpL.p, sizeof(double) * (size_t)batch, hipMemcpyDeviceToHost, s));
    if ((rc = download_field(fp64, uL, p->u_out.p, n, s))) return rc;
    return finish_call(p, batch, status, stats, false);
}

int fh_field_metrics(fh_plan p, int32_t count, const double* spectra, const double* omega, double dt_ps, double* out)
{
    if (!p || !spectra || !omega || !out) return fail(FH_EINVAL, "plan, spectra, omega or out pointer is NULL");
    if (count < 1) return fail(FH_EINVAL, "count must be >= 1");
    if (!(std::isfinite(dt_ps) && dt_ps > 0)) return fail(FH_EINVAL, "dt must be positive and finite");
    const size_t nt = (size_t)p->d.nt, m = (size_t)p->d.modes;
    const bool fp64 = p->d.precision == FH_PREC_FP64;
    const size_t esz = fp64 ? 16 : 8;
    if (p->split) return fail(FH_EUNSUPPORTED, "field metrics: no FP64 kernel at nt=" + std::to_string(nt));
    if (m * nt * esz >= ((size_t)1 << 31)) return fail(FH_EINVAL, "spectrum too large");
    for (size_t k = 0; k < nt; ++k)
        if (!(std::isfinite(omega[k]) && omega[k] > 0)) return fail(FH_EINVAL, "omega must be positive and finite");
    const size_t per = m * nt;
    for (int c = 0; c < count; ++c) {
        double norm2 = 0.0;
        for (size_t i = 0; i < 2 * per; ++i) {
            const double x = spectra[(size_t)c * 2 * per + i];
            if (!std::isfinite(x)) return fail(FH_EINVAL, "spectra must be finite");
            norm2 += x * x;
        }
        if (!(norm2 > 0)) return fail(FH_EINVAL, "field power must be nonzero");
    }
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    TmpBuf dsp, dom, dout;
    FH_HIP(hipMalloc(&dsp.p, esz * per * (size_t)count));
    FH_HIP(hipMalloc(&dom.p, sizeof(double) * nt));
    FH_HIP(hipMalloc(&dout.p, sizeof(double) * (m + 5) * (size_t)count));
    FH_HIP(hipMemcpyAsync(dom.p, omega, sizeof(double) * nt, hipMemcpyHostToDevice, s));
    int rc = upload_field(fp64, dsp.p, spectra, per * (size_t)count, s);
    if (rc) return rc;
    hipError_t e = fh::launch_field_metrics((int)nt, fp64, dsp.p, p->tw.p, (const double*)dom.p, dt_ps, (int)m,
                                            (double*)dout.p, count, s);
    if (e != hipSuccess) return fail(FH_EDEVICE, std::string("metrics launch: ") + hipGetErrorString(e));
    FH_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * (m + 5) * (size_t)count, hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    return FH_OK;
}

int fh_device_alloc(fh_context ctx, int64_t bytes, void** ptr)
{
    if (!ctx || !ptr || bytes <= 0) return fail(FH_EINVAL, "context, ptr or size invalid");
    FH_HIP(hipSetDevice(ctx->device));
    FH_HIP(hipMalloc(ptr, (size_t)bytes));
    return FH_OK;
}

int fh_device_free(fh_context ctx, void* ptr)
{
    if (!ctx) return fail(FH_EINVAL, "context is NULL");
    if (!ptr) return FH_OK;
    FH_HIP(hipSetDevice(ctx->device));
    FH_HIP(hipFree(ptr));
    return FH_OK;
}

int fh_copy_to_device(fh_context ctx, void* dst, const void* src, int64_t bytes)
{
    if (!ctx || !dst || !src || bytes < 0) return fail(FH_EINVAL, "invalid copy arguments");
    FH_HIP(hipSetDevice(ctx->device));
    FH_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    FH_HIP(hipStreamSynchronize(ctx->stream));
    return FH_OK;
}

int fh_copy_to_host(fh_context ctx, void* dst, const void* src, int64_t bytes)
{
    if (!ctx || !dst || !src || bytes < 0) return fail(FH_EINVAL, "invalid copy arguments");
    FH_HIP(hipSetDevice(ctx->device));
    FH_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    FH_HIP(hipStreamSynchronize(ctx->stream));
    return FH_OK;
}

int fh_device_copy(fh_context ctx, void* dst, const void* src, int64_t bytes)
{
    if (!ctx || !dst || !src || bytes < 0 || (bytes % 16) != 0 || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
        return fail(FH_EINVAL, "copy needs 16-byte aligned pointers and a multiple of 16 bytes");
    FH_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>(4096, std::max<size_t>(1, (n + 255) / 256));
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4*)src, (uint4*)dst, n);
    FH_HIP(hipGetLastError());
    FH_HIP(hipStreamSynchronize(ctx->stream));
    return FH_OK;
}

int fh_synchronize(fh_context ctx)
{
    if (!ctx) return fail(FH_EINVAL, "context is NULL");
    FH_HIP(hipSetDevice(ctx->device));
    FH_HIP(hipStreamSynchronize(ctx->stream));
    return FH_OK;
}

int fh_time_forward_device(fh_plan p, int32_t batch, const void* d_u0, void* d_uL, int32_t launches, float* ms)
{
    if (!p || !d_u0 || !d_uL || !ms || launches < 1) return fail(FH_EINVAL, "invalid timing arguments");
    if (batch < 1 || batch > p->d.max_batch) return fail(FH_EINVAL, "batch must lie in [1, max_batch]");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    std::vector<hipEvent_t> ev(2 * (size_t)launches, nullptr);
    int rc = FH_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = fail(FH_EDEVICE, "hipEventCreate failed");
    for (int i = 0; rc == FH_OK && i < launches; ++i) {
        if (hipEventRecord(ev[2 * i], s) != hipSuccess) rc = fail(FH_EDEVICE, "hipEventRecord failed");
        if (rc == FH_OK)
            rc = p->d.precision == FH_PREC_FP64 ? launch_forward<double>(p, batch, d_u0, d_uL, s)
                                                : launch_forward<float>(p, batch, d_u0, d_uL, s);
        if (rc == FH_OK && hipEventRecord(ev[2 * i + 1], s) != hipSuccess)
            rc = fail(FH_EDEVICE, "hipEventRecord failed");
    }
    if (rc == FH_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(FH_EDEVICE, "hipStreamSynchronize failed");
    for (int i = 0; rc == FH_OK && i < launches; ++i)
        if (hipEventElapsedTime(&ms[i], ev[2 * i], ev[2 * i + 1]) != hipSuccess)
            rc = fail(FH_EDEVICE, "hipEventElapsedTime failed");
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int fh_gradient_device(fh_plan p, int32_t batch, const void* d_u0, const void* d_seed, void* d_uL, void* d_lam0,
                       double* d_grad, double* d_cost, void* stream)
{
    if (!p || !d_u0 || !d_uL || !d_lam0 || !d_grad || !d_cost) return fail(FH_EINVAL, "NULL argument");
    if (batch < 1 || batch > p->d.max_batch) return fail(FH_EINVAL, "batch must lie in [1, max_batch]");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : p->ctx->stream;
    return p->d.precision == FH_PREC_FP64
               ? launch_gradient<double>(p, batch, d_u0, d_seed, d_uL, d_lam0, d_grad, d_cost, s)
               : launch_gradient<float>(p, batch, d_u0, d_seed, d_uL, d_lam0, d_grad, d_cost, s);
}

int fh_time_gradient_device(fh_plan p, int32_t batch, const void* d_u0, void* d_uL, void* d_lam0, double* d_grad,
                            double* d_cost, int32_t launches, float* ms)
{
    if (!p || !ms || launches < 1) return fail(FH_EINVAL, "invalid timing arguments");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    std::vector<hipEvent_t> ev(2 * (size_t)launches, nullptr);
    int rc = FH_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = fail(FH_EDEVICE, "hipEventCreate failed");
    for (int i = 0; rc == FH_OK && i < launches; ++i) {
        if (hipEventRecord(ev[2 * i], s) != hipSuccess) rc = fail(FH_EDEVICE, "hipEventRecord failed");
        if (rc == FH_OK) rc = fh_gradient_device(p, batch, d_u0, nullptr, d_uL, d_lam0, d_grad, d_cost, s);
        if (rc == FH_OK && hipEventRecord(ev[2 * i + 1], s) != hipSuccess)
            rc = fail(FH_EDEVICE, "hipEventRecord failed");
    }
    if (rc == FH_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(FH_EDEVICE, "hipStreamSynchronize failed");
    for (int i = 0; rc == FH_OK && i < launches; ++i)
        if (hipEventElapsedTime(&ms[i], ev[2 * i], ev[2 * i + 1]) != hipSuccess)
            rc = fail(FH_EDEVICE, "hipEventElapsedTime failed");
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int fh_gradient(fh_plan p, int32_t batch, const double* u0, const double* seed, double* uL, double* lam0,
                double* grad, double* cost, fh_stats* stats)
{
    if (!p || !u0) return fail(FH_EINVAL, "plan or launch pointer is NULL");
    if (batch < 1 || batch > p->d.max_batch)
        return fail(FH_EINVAL, "batch must lie in [1, max_batch=" + std::to_string(p->d.max_batch) + "]");
    const size_t n = (size_t)batch * p->d.nt * p->d.modes;
    double norm2 = 0.0;
    for (size_t i = 0; i < 2 * n; ++i) {
        if (!std::isfinite(u0[i])) return fail(FH_EINVAL, "input spectrum must be finite");
        norm2 += u0[i] * u0[i];
    }
    if (!(norm2 > 0)) return fail(FH_EINVAL, "input spectrum must be nonzero");
    if (seed)
        for (size_t i = 0; i < 2 * n; ++i)
            if (!std::isfinite(seed[i])) return fail(FH_EINVAL, "adjoint terminal condition must be finite");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    const bool fp64 = p->d.precision == FH_PREC_FP64;
    int rc = fp64 ? ensure_adjoint_buffers<double>(p) : ensure_adjoint_buffers<float>(p);
    if (rc) return rc;
    const size_t es = fp64 ? sizeof(double) : sizeof(float);
    auto to_dev = [&](void* dst, const double* src) -> int {
        if (fp64) {
            FH_HIP(hipMemcpyAsync(dst, src, 2 * es * n, hipMemcpyHostToDevice, s));
        } else {
            std::vector<float> tmp(2 * n);
            for (size_t i = 0; i < 2 * n; ++i) tmp[i] = (float)src[i];
            FH_HIP(hipMemcpyAsync(dst, tmp.data(), 2 * es * n, hipMemcpyHostToDevice, s));
            FH_HIP(hipStreamSynchronize(s));
        }
        return FH_OK;
    };
    auto to_host = [&](double* dst, const void* src) -> int {
        if (fp64) {
            FH_HIP(hipMemcpyAsync(dst, src, 2 * es * n, hipMemcpyDeviceToHost, s));
            FH_HIP(hipStreamSynchronize(s));
        } else {
            std::vector<float> tmp(2 * n);
            FH_HIP(hipMemcpyAsync(tmp.data(), src, 2 * es * n, hipMemcpyDeviceToHost, s));
            FH_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < 2 * n; ++i) dst[i] = tmp[i];
        }
        return FH_OK;
    };
    if ((rc = to_dev(p->u_in.p, u0))) return rc;
    if (seed && (rc = to_dev(p->g_seed.p, seed))) return rc;
    rc = fh_gradient_device(p, batch, p->u_in.p, seed ? p->g_seed.p : nullptr, p->u_out.p, p->g_lam0.p,
                            (double*)p->g_grad.p, (double*)p->g_cost.p, s);
    if (rc) return rc;
    std::vector<int> status(batch);
    FH_HIP(hipMemcpyAsync(status.data(), p->status.p, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    if (uL && (rc = to_host(uL, p->u_out.p))) return rc;
    if (lam0 && (rc = to_host(lam0, p->g_lam0.p))) return rc;
    const size_t nb = (size_t)batch * p->d.nt;
    if (grad) FH_HIP(hipMemcpy(grad, p->g_grad.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
    if (cost) {
        if (seed)
            for (int b = 0; b < batch; ++b) cost[b] = NAN;
        else
            FH_HIP(hipMemcpy(cost, p->g_cost.p, sizeof(double) * batch, hipMemcpyDeviceToHost));
    }
    return finish_call(p, batch, status, stats, true);
}

int fh_forward_tape(fh_plan p, int32_t batch, const double* u0, double* uL, fh_tape* tape, fh_stats* stats)
{
    if (!p || !u0 || !tape) return fail(FH_EINVAL, "plan, launch or tape pointer is NULL");
    *tape = nullptr;
    if (batch < 1 || batch > p->d.max_batch)
        return fail(FH_EINVAL, "batch must lie in [1, max_batch=" + std::to_string(p->d.max_batch) + "]");
    if (p->d.modes > 1 && !p->gamma_full_sym)
        return fail(FH_ENOTSYM, "fiber adjoint requires a fully permutation-symmetric gamma tensor");
    const size_t n = (size_t)batch * p->d.nt * p->d.modes;
    double norm2 = 0.0;
    for (size_t i = 0; i < 2 * n; ++i) {
        if (!std::isfinite(u0[i])) return fail(FH_EINVAL, "input spectrum must be finite");
        norm2 += u0[i] * u0[i];
    }
    if (!(norm2 > 0)) return fail(FH_EINVAL, "input spectrum must be nonzero");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    const bool fp64 = p->d.precision == FH_PREC_FP64;
    int rc = fp64 ? ensure_grad_work<double>(p) : ensure_grad_work<float>(p);
    if (rc) return rc;
    fh_tape t = nullptr;
    if ((rc = new_tape(p, batch, &t))) return rc;
    if (!(rc = upload_field(fp64, t->u0.p, u0, n, s)) && !(rc = tape_forward(t, s))) {
        FH_HIP(hipStreamSynchronize(s));
        rc = tape_finish(t, stats, false);
    }
    if (!rc && uL) rc = download_field(fp64, uL, t->uL.p, n, s);
    if (rc) {
        release_tape_storage(t);
        return rc;
    }
    *tape = t;
    return FH_OK;
}

int fh_adjoint(fh_tape t, const double* seed, double* lam0, double* grad_phase, double* grad_amp, double* energy_dot,
               double* cost, fh_stats* stats)
{
    if (!t || !t->plan) return fail(FH_EINVAL, "tape is NULL");
    fh_plan p = t->plan;
    const int batch = t->batch;
    const size_t nt = (size_t)p->d.nt, M = (size_t)p->d.modes, n = (size_t)batch * nt * M;
    if (!seed && !p->has_mask) return fail(FH_EINVAL, "adjoint needs a terminal seed or a plan band mask");
    if (seed)
        for (size_t i = 0; i < 2 * n; ++i)
            if (!std::isfinite(seed[i])) return fail(FH_EINVAL, "adjoint terminal condition must be finite");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    const bool fp64 = p->d.precision == FH_PREC_FP64;
    int rc;
    // a single-mode backward sweep overwrites the kept tape (each slot is
    // pulled back once): sweeping again rebuilds it from the kept launch,
    // which reproduces the same bits
    if (t->swept && !p->mm && p->ntaped > 0) {
        if ((rc = tape_forward(t, s))) return rc;
    }
    if (seed) {
        if ((rc = upload_field(fp64, t->seed.p, seed, n, s))) return rc;
    } else {
        const double tau = p->d.objective_tau > 0 ? p->d.objective_tau : 50.0;
        hipError_t e = fh::launch_objective(fp64, t->uL.p, (const unsigned char*)p->mask.p, (int)nt, (int)M,
                                            p->d.objective, tau, p->d.log_cost, t->seed.p, (double*)t->cost.p, batch,
                                            s);
        if (e != hipSuccess) return fail(FH_EDEVICE, std::string("objective launch: ") + hipGetErrorString(e));
    }
    const GradBufs gb = tape_grad_bufs(t, 2);
    rc = fp64 ? launch_gradient_with<double>(p, gb, batch, t->u0.p, t->seed.p, t->uL.p, t->lam0.p,
                                             (double*)t->kgrad.p, (double*)t->cost.p, s)
              : launch_gradient_with<float>(p, gb, batch, t->u0.p, t->seed.p, t->uL.p, t->lam0.p,
                                            (double*)t->kgrad.p, (double*)t->cost.p, s);
    if (rc) return rc;
    t->swept = true;
    hipError_t e = fh::launch_physical_gradients(fp64, t->lam0.p, t->u0.p, (int)(nt * M), (double*)t->gphase.p,
                                                 (double*)t->gamp.p, (double*)t->edot.p, batch, s);
    if (e != hipSuccess) return fail(FH_EDEVICE, std::string("physical gradient launch: ") + hipGetErrorString(e));
    FH_HIP(hipStreamSynchronize(s));
    if ((rc = tape_finish(t, stats, true))) return rc;
    if (lam0 && (rc = download_field(fp64, lam0, t->lam0.p, n, s))) return rc;
    if (grad_phase) FH_HIP(hipMemcpy(grad_phase, t->gphase.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (grad_amp) FH_HIP(hipMemcpy(grad_amp, t->gamp.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (energy_dot) FH_HIP(hipMemcpy(energy_dot, t->edot.p, sizeof(double) * batch, hipMemcpyDeviceToHost));
    if (cost) {
        if (seed) {
            for (int b = 0; b < batch; ++b) cost[b] = NAN;
        } else {
            FH_HIP(hipMemcpy(cost, t->cost.p, sizeof(double) * batch, hipMemcpyDeviceToHost));
            for (int b = 0; b < batch; ++b)
                if (!std::isfinite(cost[b]))
                    return fail(FH_EINVAL, "objective is undefined for this output (a mode it measures has "
                                           "zero energy)");
        }
    }
    return FH_OK;
}

int fh_tape_device_bytes(fh_tape t, int64_t* bytes)
{
    if (!t || !bytes) return fail(FH_EINVAL, "tape or bytes is NULL");
    *bytes = t->bytes;
    return FH_OK;
}

int fh_tape_free(fh_tape t)
{
    if (!t) return FH_OK;
    fh_plan p = t->plan;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    if (!p->spare) {
        t->plan = nullptr;
        p->spare = t;
    } else {
        release_tape_storage(t);
    }
    return FH_OK;
}

int fh_solve_stats(fh_plan p, int32_t batch, int64_t* stats)
{
    if (!p || !stats) return fail(FH_EINVAL, "plan or stats is NULL");
    if (batch < 1 || batch > p->d.max_batch) return fail(FH_EINVAL, "batch must lie in [1, max_batch]");
    FH_HIP(hipSetDevice(p->ctx->device));
    std::vector<long long> st;
    int rc = read_solve_stats(p, batch, st);
    if (rc) return rc;
    for (size_t i = 0; i < st.size(); ++i) stats[i] = st[i];
    return FH_OK;
}

int fh_step_sizes(fh_plan p, int32_t batch, double* hs)
{
    if (!p || !hs) return fail(FH_EINVAL, "plan or hs is NULL");
    if (!p->adaptive) return fail(FH_EINVAL, "step sizes are recorded by the adaptive scheme only");
    if (batch < 1 || batch > p->d.max_batch) return fail(FH_EINVAL, "batch must lie in [1, max_batch]");
    FH_HIP(hipSetDevice(p->ctx->device));
    FH_HIP(hipMemcpy(hs, p->hs.p, sizeof(double) * (size_t)batch * p->max_steps, hipMemcpyDeviceToHost));
    return FH_OK;
}

int fh_energy_device(fh_plan p, int32_t batch, const void* d_u, double* energies)
{
    if (!p || !d_u || !energies) return fail(FH_EINVAL, "NULL argument");
    if (batch < 1) return fail(FH_EINVAL, "batch must be positive");
    FH_HIP(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    const int cols = batch * p->d.modes;
    double* d_e = nullptr;
    FH_HIP(hipMallocAsync((void**)&d_e, sizeof(double) * cols, s));
    if (p->d.precision == FH_PREC_FP64)
        hipLaunchKernelGGL(k_energy<double>, dim3(cols), dim3(256), 0, s, (const fh::C2<double>*)d_u, p->d.nt, d_e);
    else
        hipLaunchKernelGGL(k_energy<float>, dim3(cols), dim3(256), 0, s, (const fh::C2<float>*)d_u, p->d.nt, d_e);
    FH_HIP(hipGetLastError());
    FH_HIP(hipMemcpyAsync(energies, d_e, sizeof(double) * cols, hipMemcpyDeviceToHost, s));
    FH_HIP(hipFreeAsync(d_e, s));
    FH_HIP(hipStreamSynchronize(s));
    return FH_OK;
}

}  // extern "C"
