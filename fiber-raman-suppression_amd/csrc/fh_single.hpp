// fh_single.hpp -- fused single-mode RK4IP propagation and its discrete
// adjoint, one workgroup per solve.
//
// The whole fixed-step integration of one solve runs inside one workgroup:
// the field stays in registers (cyclic layout, see fh_device.hpp) or in a
// workgroup-private slice of global scratch, every FFT goes through LDS, and
// HBM only sees the launch, the outputs, checkpoints/tape of the adjoint and
// the shared operator tables (D, hR, shock, twiddles), which stay L2-resident
// across the batch.  A launch of B solves uses B workgroups.
//
// Right-hand side, restating disp_mmf! for M = 1 in the operation order of
// _scalar_rk4_output's closure (scalar_crosscheck.jl:38-48):
//   u   = fft(cis(D z) .* x)
//   q   = gamma |u|^2
//   dR  = real(fftshift(ifft(hR .* fft(q))))   -- fftshift folded into hR as (-1)^k
//   h   = (1-fR) q + dR
//   eta = ifft(h .* u) .* shock
//   dx  = i cis(-D z) .* eta
// and the classic global-frame RK4 of scalar_crosscheck.jl:53-60.
//
// Pullback of the right-hand side (real-linear map; cotangent gy of dx):
//   G  = fft(-i cis(D z) (shock/N) gy)
//   gu = h G + 2 gamma b u,   a = Re(conj(G) u),  b = (1-fR) a + real(ifft(conj(hR)(-1)^k fft(a)))
//   gx = conj(cis(D z)) N ifft(gu)
// with u and h of the stage taken from the tape recorded by the segment
// recomputation.  conj(hR) and the (-1)^k alternation are the same operators
// adjoint_disp_mmf! uses (sensitivity_disp_mmf.jl:96,127,241-242).
#pragma once
#include "fh_device.hpp"

namespace fh {

template <typename R> struct SmArgs {
    const C2<R>* u0;     // [batch][N] interaction-picture launch (= uw0)
    C2<R>* out;          // [batch][N] lab-frame output cis(D L) x(L)
    C2<R>* work;         // [batch][NW][N] workgroup-private scratch
    const double* disp;  // [N] D(w), rad/m, FFT order
    const C2<R>* hn;     // [N] hR(w) (-1)^k / N
    const R* shn;        // [N] shock / N
    const C2<R>* tw;     // Stockham twiddles of FftPlan<N,T>
    int* status;         // [batch] 0 ok, 1 non-finite output
    R gamma, one_m_fr;
    double h, length;
    int steps;
    // adjoint only
    const C2<R>* seed;          // [batch][N] lab-frame Wirtinger seed, or null for the band objective
    const unsigned char* mask;  // [N] band mask of the band objective
    int log_cost;               // 10 log10 J (physics_helpers.jl:320-326)
    C2<R>* lam0;                // [batch][N] dJ/dconj(u0)
    double* grad;               // [batch][N] 2 Re(conj(lam0) i u0)
    double* cost;               // [batch]
    C2<R>* ckpt;                // [batch][nseg][N]
    C2<R>* tape_u;              // [batch][regions][4*every][N]
    R* tape_h;                  // [batch][regions][4*every][N]
    int every, nseg;
    int ntaped;  // the last ntaped segments keep the forward sweep's tape (k_grad_single)
    // gradient kernels: bit 1 runs the forward sweep (checkpoints, kept tape,
    // output), bit 2 the terminal seed and backward sweep; 3 runs both in one
    // launch.  A backward-only launch (2) needs a caller seed and the
    // checkpoints and kept tape of an earlier forward-only launch (1).
    int phase;
};

template <typename R, int N, int T> struct SmTraits;
// FP64: state/accumulators in global scratch; time field in registers when a
// workgroup has at most 512 threads and 8 samples per thread; the rotation
// cis(D z) is re-evaluated for the back-rotation instead of being held.
template <int N, int T> struct SmTraits<double, N, T> {
    static constexpr bool STATE_REGS = false;
    static constexpr bool UT_REGS = T <= 512 && N / T <= 8;
    static constexpr bool KEEP_PHASE = false;
    static constexpr bool PLANAR = (size_t)LdsSize<N>::value * 16 > LdsBudget<double>::bytes;
};
template <int N, int T> struct SmTraits<float, N, T> {
    static constexpr bool STATE_REGS = T <= 512;
    static constexpr bool UT_REGS = N / T <= 16;     // 32 samples per thread keep u in global scratch
    // the rotation is held across the RHS only where registers allow: in
    // 1024-thread workgroups (128 registers) re-evaluating it is cheaper than
    // the spills 32 held registers cause (measured 3.6 % on the 16384-sample
    // gradient)
    static constexpr bool KEEP_PHASE = N / T <= 16 && T <= 512;
    static constexpr bool PLANAR = (size_t)LdsSize<N>::value * 8 > LdsBudget<float>::bytes;
};

template <typename R, int N, int T> struct SmLayout {
    using TR = SmTraits<R, N, T>;
    // st, ac, lam, lacc when not in registers; ut, gu when not in registers
    static constexpr int NW = (TR::STATE_REGS ? 0 : 4) + (TR::UT_REGS ? 0 : 2);
    static constexpr int NWA = NW > 0 ? NW : 1;
    static constexpr size_t LDS_BYTES = (size_t)LdsSize<N>::value * (TR::PLANAR ? sizeof(R) : 2 * sizeof(R));
};

// A cyclic field of this thread, in registers or in a global slice.
template <typename V, int T, int P, bool REGS> struct Field;
template <typename V, int T, int P> struct Field<V, T, P, true> {
    V r[P];
    __device__ __forceinline__ explicit Field(const Buf&) {}
    __device__ __forceinline__ V get(int k) const { return r[k]; }
    __device__ __forceinline__ void set(int k, V x) { r[k] = x; }
};
template <typename V, int T, int P> struct Field<V, T, P, false> {
    Buf b;
    __device__ __forceinline__ explicit Field(const Buf& bb) : b(bb) {}
    __device__ __forceinline__ V get(int k) const { return b.ld<V>(Cyc<V, T>::voff(), Cyc<V, T>::soff(k)); }
    __device__ __forceinline__ void set(int k, V x) { b.st(x, Cyc<V, T>::voff(), Cyc<V, T>::soff(k)); }
};

// Buffer views of the shared operator tables.
template <typename R, int N> struct SmTables {
    Buf disp, hn, shn, tw;
    __device__ __forceinline__ SmTables(const SmArgs<R>& a, int tw_entries)
        : disp(a.disp, N * 8), hn(a.hn, N * 2 * sizeof(R)), shn(a.shn, N * sizeof(R)),
          tw(a.tw, tw_entries * 2 * sizeof(R))
    {
    }
};

template <typename R, int N, int T> struct SmCtx {
    using V = C2<R>;
    using TR = SmTraits<R, N, T>;
    using LY = SmLayout<R, N, T>;
    static constexpr int P = N / T;
    using FS = Field<V, T, P, TR::STATE_REGS>;
    using FU = Field<V, T, P, TR::UT_REGS>;
};

// One right-hand-side evaluation.  On entry v holds the stage input y
// (interaction picture, cyclic layout); on exit v holds f(zs, y).  With TAPE
// the stage's time field u and weight h are stored to tu / th.
template <typename R, int N, int T, bool TAPE>
__device__ __forceinline__ void sm_rhs(C2<R>* v, double zs, const SmArgs<R>& a, const SmTables<R, N>& tb,
                                       typename SmCtx<R, N, T>::FU& ut, const Fld& tu, const Fld& th, void* lds)
{
    using V = C2<R>;
    using TR = SmTraits<R, N, T>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    using CR = Cyc<R, T>;
    constexpr int P = N / T;
    R cs[TR::KEEP_PHASE ? P : 1], sn[TR::KEEP_PHASE ? P : 1];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        R c, s;
        cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), zs, c, s);
        if constexpr (TR::KEEP_PHASE) {
            cs[k] = c;
            sn[k] = s;
        }
        v[k] = V{v[k].x * c - v[k].y * s, v[k].x * s + v[k].y * c};
    }
    block_fft<R, N, T, -1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        ut.set(k, v[k]);
        if constexpr (TAPE) tu.st(v[k], CV::voff(), CV::soff(k));
        const R q = a.gamma * (v[k].x * v[k].x + v[k].y * v[k].y);
        v[k] = V{q, (R)0};
    }
    block_fft<R, N, T, -1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = cmul(v[k], tb.hn.template ld<V>(CV::voff(), CV::soff(k)));
    block_fft<R, N, T, +1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const V u = ut.get(k);
        const R q = a.gamma * (u.x * u.x + u.y * u.y);
        const R f = a.one_m_fr * q + v[k].x;
        if constexpr (TAPE) th.st(f, CR::voff(), CR::soff(k));
        v[k] = V{u.x * f, u.y * f};
    }
    block_fft<R, N, T, +1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        R c, s;
        if constexpr (TR::KEEP_PHASE) {
            c = cs[k];
            s = sn[k];
        } else {
            cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), zs, c, s);
        }
        const R sc = tb.shn.template ld<R>(CR::voff(), CR::soff(k));
        const V e = V{v[k].x * sc, v[k].y * sc};
        // (i cis(-x)) * e = (sin x + i cos x) * e
        v[k] = V{s * e.x - c * e.y, s * e.y + c * e.x};
    }
}

// One RK4 step from state st (v holds st on entry and the new state on exit).
// tape_u / tape_h: the solve's tape regions (fixed descriptors); the stage's
// slot is addressed through the per-lane offset (slot0 + s) * N * element.
template <typename R, int N, int T, bool TAPE>
__device__ __forceinline__ void sm_step(C2<R>* v, int step, const SmArgs<R>& a, const SmTables<R, N>& tb,
                                        typename SmCtx<R, N, T>::FS& st, typename SmCtx<R, N, T>::FS& ac,
                                        typename SmCtx<R, N, T>::FU& ut, const Buf& tape_u, const Buf& tape_h,
                                        int slot0, void* lds)
{
    using V = C2<R>;
    constexpr int P = N / T;
    const double h = a.h;
    const R h2 = (R)(h / 2), h1 = (R)h, h6 = (R)(h / 6);
    const double z = step * h;
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        const double zs = s == 0 ? z : (s == 3 ? z + h : z + h / 2);
        const Fld tu{tape_u, (uint32_t)(slot0 + s) * (uint32_t)(N * sizeof(V))};
        const Fld th{tape_h, (uint32_t)(slot0 + s) * (uint32_t)(N * sizeof(R))};
        sm_rhs<R, N, T, TAPE>(v, zs, a, tb, ut, tu, th, lds);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V x = st.get(k);
            V c = s != 0 ? ac.get(k) : V{(R)0, (R)0};
            const V kk = v[k];
            if (s == 0) {
                c = kk;
                v[k] = V{x.x + h2 * kk.x, x.y + h2 * kk.y};
            } else if (s == 1) {
                c = V{c.x + 2 * kk.x, c.y + 2 * kk.y};
                v[k] = V{x.x + h2 * kk.x, x.y + h2 * kk.y};
            } else if (s == 2) {
                c = V{c.x + 2 * kk.x, c.y + 2 * kk.y};
                v[k] = V{x.x + h1 * kk.x, x.y + h1 * kk.y};
            } else {
                const V t = V{c.x + kk.x, c.y + kk.y};
                const V xn = V{x.x + h6 * t.x, x.y + h6 * t.y};
                v[k] = xn;
                st.set(k, xn);
            }
            if (s != 3) ac.set(k, c);
        }
    }
}

template <typename R, int N, int T>
__global__ __launch_bounds__(T) void k_rk4ip_single(SmArgs<R> a)
{
    using CX = SmCtx<R, N, T>;
    using V = C2<R>;
    using LY = typename CX::LY;
    using CV = Cyc<V, T>;
    constexpr int P = N / T;
    __shared__ __align__(16) unsigned char lds[LY::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const SmTables<R, N> tb(a, FftPlan<N, T>::TW_SIZE);
    const Buf u0(a.u0 + sol * N, N * sizeof(V));
    V* gwork = a.work + sol * (size_t)LY::NWA * N;
    typename CX::FS st(Buf(gwork, N * sizeof(V))), ac(Buf(gwork + N, N * sizeof(V)));
    typename CX::FU ut(Buf(gwork + (size_t)(CX::TR::STATE_REGS ? 0 : 4) * N, N * sizeof(V)));

    V v[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        v[k] = u0.ld<V>(CV::voff(), CV::soff(k));
        st.set(k, v[k]);
    }
    const Buf none(a.u0, 0);
    for (int step = 0; step < a.steps; ++step) sm_step<R, N, T, false>(v, step, a, tb, st, ac, ut, none, none, 0, lds);
    int bad = 0;
    const Buf out(a.out + sol * N, N * sizeof(V));
    using CD = Cyc<double, T>;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        R c, s;
        cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), a.length, c, s);
        const V x = v[k];
        const V o = V{x.x * c - x.y * s, x.x * s + x.y * c};
        bad |= !(isfinite(o.x) && isfinite(o.y));
        out.st(o, CV::voff(), CV::soff(k));
    }
    if (bad) atomicOr(a.status + sol, 1);
}

// Pullback of one right-hand-side evaluation.  On entry v holds the cotangent
// gy of the stage output, on exit the cotangent gx of the stage input.
// Each tape slot is pulled back exactly once, so the slot's h entry is
// overwritten with the real field a = Re(conj(G) u) after h is used.
template <typename R, int N, int T>
__device__ __forceinline__ void sm_vjp(C2<R>* v, double zs, const SmArgs<R>& a, const SmTables<R, N>& tb,
                                       typename SmCtx<R, N, T>::FU& gu, const Fld& tu, const Fld& th, void* lds)
{
    using V = C2<R>;
    using TR = SmTraits<R, N, T>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    using CR = Cyc<R, T>;
    constexpr int P = N / T;
    R cs[TR::KEEP_PHASE ? P : 1], sn[TR::KEEP_PHASE ? P : 1];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        R c, s;
        cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), zs, c, s);
        if constexpr (TR::KEEP_PHASE) {
            cs[k] = c;
            sn[k] = s;
        }
        const R sc = tb.shn.template ld<R>(CR::voff(), CR::soff(k));
        const V e = V{v[k].x * sc, v[k].y * sc};
        const V w = V{e.x * c - e.y * s, e.x * s + e.y * c};
        v[k] = V{w.y, -w.x};  // -i w
    }
    block_fft<R, N, T, -1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const V u = tu.template ld<V>(CV::voff(), CV::soff(k));
        const R hh = th.template ld<R>(CR::voff(), CR::soff(k));
        const V G = v[k];
        gu.set(k, V{hh * G.x, hh * G.y});
        const R ak = G.x * u.x + G.y * u.y;
        th.st(ak, CR::voff(), CR::soff(k));  // h of this slot is consumed; the slot now holds a
        v[k] = V{ak, (R)0};
    }
    block_fft<R, N, T, -1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = cmul(v[k], cconj(tb.hn.template ld<V>(CV::voff(), CV::soff(k))));
    block_fft<R, N, T, +1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const V u = tu.template ld<V>(CV::voff(), CV::soff(k));
        const R b = v[k].x + a.one_m_fr * th.template ld<R>(CR::voff(), CR::soff(k));
        const R f = 2 * a.gamma * b;
        const V g = gu.get(k);
        v[k] = V{g.x + f * u.x, g.y + f * u.y};
    }
    block_fft<R, N, T, +1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        R c, s;
        if constexpr (TR::KEEP_PHASE) {
            c = cs[k];
            s = sn[k];
        } else {
            cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), zs, c, s);
        }
        v[k] = V{c * v[k].x + s * v[k].y, c * v[k].y - s * v[k].x};  // conj(cis) * v
    }
}

// Workgroup FP64 sum (deterministic tree) through the LDS buffer.
template <int T> __device__ __forceinline__ double block_sum(double x, void* lds_raw)
{
    double* red = reinterpret_cast<double*>(lds_raw);
    const int tid = threadIdx.x;
    __syncthreads();
    red[tid] = x;
    __syncthreads();
#pragma unroll 1
    for (int w = T / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Forward solve with checkpoints, terminal seed (caller-supplied or the band
// objective of _field_spectral_band_cost, physics_helpers.jl:222-232),
// backward sweep of the discrete adjoint with per-segment recomputation, and
// the phase chain rule g = 2 Re(conj(lam0) i u0) (physics_models.jl:1042) in
// FP64.
//
// The tape of a solve has one region of 4*every stage slots per kept segment
// (the last ntaped segments, written by the forward sweep) plus one shared
// region for the segments the backward sweep recomputes.  The recomputation
// repeats the forward sweep's arithmetic from the checkpoint, so a kept
// segment's tape holds the same values and the result does not depend on
// ntaped.
template <typename R, int N, int T>
__global__ __launch_bounds__(T) void k_grad_single(SmArgs<R> a)
{
    using CX = SmCtx<R, N, T>;
    using V = C2<R>;
    using LY = typename CX::LY;
    using TR = typename CX::TR;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = N / T;
    __shared__ __align__(16) unsigned char lds[LY::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const SmTables<R, N> tb(a, FftPlan<N, T>::TW_SIZE);
    const Buf u0(a.u0 + sol * N, N * sizeof(V));
    V* gwork = a.work + sol * (size_t)LY::NWA * N;
    constexpr int US = TR::STATE_REGS ? 0 : 4;
    typename CX::FS st(Buf(gwork, N * sizeof(V))), ac(Buf(gwork + N, N * sizeof(V)));
    typename CX::FS lam(Buf(gwork + 2 * N, N * sizeof(V))), lacc(Buf(gwork + 3 * N, N * sizeof(V)));
    typename CX::FU ut(Buf(gwork + (size_t)US * N, N * sizeof(V)));
    typename CX::FU gu(Buf(gwork + (size_t)(US + 1) * N, N * sizeof(V)));
    // Fixed descriptors per solve region; checkpoints and tape slots are
    // addressed through the per-lane offset.
    const Buf ckpt(a.ckpt + sol * (size_t)a.nseg * N, (uint32_t)((size_t)a.nseg * N * sizeof(V)));
    const int first_kept = a.nseg - a.ntaped;  // segments first_kept.. keep their tape
    const int seg_slots = 4 * a.every;
    const size_t tape_slots = (size_t)(a.ntaped + (first_kept > 0 ? 1 : 0)) * seg_slots;
    const Buf tape_u(a.tape_u + sol * tape_slots * N, (uint32_t)(tape_slots * N * sizeof(V)));
    const Buf tape_h(a.tape_h + sol * tape_slots * N, (uint32_t)(tape_slots * N * sizeof(R)));
    // first tape slot of segment sg: its own region, or the shared one
    auto seg_base = [&](int sg) { return (sg >= first_kept ? sg - first_kept : a.ntaped) * seg_slots; };
    const Buf none(a.u0, 0);
    const double h = a.h;

    // forward with checkpoints
    V v[P];
    double etot = 0.0, eband = 0.0;
    int bad = 0;
    if (a.phase & 1) {
#pragma unroll
        for (int k = 0; k < P; ++k) {
            v[k] = u0.ld<V>(CV::voff(), CV::soff(k));
            st.set(k, v[k]);
        }
        const int kept_from = min(first_kept * a.every, a.steps);
        for (int step = 0; step < kept_from; ++step) {
            if (step % a.every == 0) {
                const Fld ck{ckpt, (uint32_t)(step / a.every) * (uint32_t)(N * sizeof(V))};
#pragma unroll
                for (int k = 0; k < P; ++k) ck.st(v[k], CV::voff(), CV::soff(k));
            }
            sm_step<R, N, T, false>(v, step, a, tb, st, ac, ut, none, none, 0, lds);
        }
        for (int step = kept_from; step < a.steps; ++step) {
            const int sg = step / a.every;
            sm_step<R, N, T, true>(v, step, a, tb, st, ac, ut, tape_u, tape_h, seg_base(sg) + 4 * (step - sg * a.every),
                                   lds);
        }
        // output and terminal seed
        const Buf out(a.out + sol * N, N * sizeof(V));
#pragma unroll
        for (int k = 0; k < P; ++k) {
            R c, s;
            cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), a.length, c, s);
            const V x = v[k];
            const V o = V{x.x * c - x.y * s, x.x * s + x.y * c};
            bad |= !(isfinite(o.x) && isfinite(o.y));
            out.st(o, CV::voff(), CV::soff(k));
            v[k] = o;
            const double e = (double)o.x * o.x + (double)o.y * o.y;
            etot += e;
            if (!a.seed && a.mask[threadIdx.x + T * k]) eband += e;
        }
        if (!(a.phase & 2)) {  // forward sweep only: checkpoints and kept tape stay for a later backward launch
            if (bad) atomicOr(a.status + sol, 1);
            return;
        }
    }
    if (!a.seed) {
        etot = block_sum<T>(etot, lds);
        eband = block_sum<T>(eband, lds);
        const double J = eband / etot;
        double scale = 1.0 / etot;
        if (a.log_cost) scale *= J <= 1e-15 ? 0.0 : 10.0 / (2.302585092994045684 * J);
        if (threadIdx.x == 0) a.cost[sol] = a.log_cost ? 10.0 * log10(fmax(J, 1e-15)) : J;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const R f = (R)(((a.mask[threadIdx.x + T * k] ? 1.0 : 0.0) - J) * scale);
            v[k] = V{v[k].x * f, v[k].y * f};
        }
    } else {
        const Buf sd(a.seed + sol * N, N * sizeof(V));
#pragma unroll
        for (int k = 0; k < P; ++k) v[k] = sd.ld<V>(CV::voff(), CV::soff(k));
    }
    // interaction picture: lam = cis(-D L) lamL
#pragma unroll
    for (int k = 0; k < P; ++k) {
        R c, s;
        cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), a.length, c, s);
        lam.set(k, V{c * v[k].x + s * v[k].y, c * v[k].y - s * v[k].x});
    }
    const R c6 = (R)(h / 6), c3 = (R)(h / 3), c2 = (R)(h / 2), c1 = (R)h;
    for (int sg = a.nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * a.every;
        const int s1 = min(s0 + a.every, a.steps);
        const int base = seg_base(sg);
        if (sg < first_kept) {  // recompute the segment's tape from its checkpoint
            const Fld ck{ckpt, (uint32_t)sg * (uint32_t)(N * sizeof(V))};
#pragma unroll
            for (int k = 0; k < P; ++k) {
                v[k] = ck.template ld<V>(CV::voff(), CV::soff(k));
                st.set(k, v[k]);
            }
            for (int step = s0; step < s1; ++step)
                sm_step<R, N, T, true>(v, step, a, tb, st, ac, ut, tape_u, tape_h, base + 4 * (step - s0), lds);
        }
        for (int step = s1 - 1; step >= s0; --step) {
            const double z = step * h;
#pragma unroll 1
            for (int s = 3; s >= 0; --s) {
                drain_memory();  // lam / lacc were stored by the previous stage
                const double zs = s == 0 ? z : (s == 3 ? z + h : z + h / 2);
                const R cl = (s == 0 || s == 3) ? c6 : c3;   // weight of k_s in the update
                const R cg = s == 3 ? (R)0 : (s == 2 ? c1 : c2);  // dy_{s+1}/dk_s
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    const V l = lam.get(k);
                    if (s == 3) lacc.set(k, l);
                    v[k] = V{cl * l.x + cg * v[k].x, cl * l.y + cg * v[k].y};
                }
                const int slot = base + 4 * (step - s0) + s;
                const Fld tu{tape_u, (uint32_t)slot * (uint32_t)(N * sizeof(V))};
                const Fld th{tape_h, (uint32_t)slot * (uint32_t)(N * sizeof(R))};
                sm_vjp<R, N, T>(v, zs, a, tb, gu, tu, th, lds);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    const V c = lacc.get(k);
                    const V nc = V{c.x + v[k].x, c.y + v[k].y};
                    if (s == 0)
                        lam.set(k, nc);
                    else
                        lacc.set(k, nc);
                }
            }
        }
    }
    // lam0 and the phase chain rule in FP64
    const Buf l0(a.lam0 + sol * N, N * sizeof(V));
    const Buf gr(a.grad + sol * N, N * sizeof(double));
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const V l = lam.get(k);
        const V u = u0.ld<V>(CV::voff(), CV::soff(k));
        l0.st(l, CV::voff(), CV::soff(k));
        const double g = 2.0 * ((double)l.y * (double)u.x - (double)l.x * (double)u.y);
        bad |= !(isfinite(l.x) && isfinite(l.y));
        gr.st(g, CD::voff(), CD::soff(k));
    }
    if (bad) atomicOr(a.status + sol, 1);
}

}  // namespace fh
