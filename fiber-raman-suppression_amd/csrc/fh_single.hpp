// fh_single.hpp -- fused single-mode RK4IP propagation (one workgroup per solve).
//
// The whole fixed-step integration of one solve runs inside one workgroup:
// the field stays in registers (cyclic layout, see fh_device.hpp) and in a
// workgroup-private slice of global scratch, every FFT goes through LDS, and
// HBM only sees the launch spectrum, the output spectrum and the shared
// operator tables (D, hR, shock, twiddles), which stay L2-resident across the
// batch.  A launch of B solves uses B workgroups.
//
// Right-hand side, restating disp_mmf! for M = 1 in the operation order of
// _scalar_rk4_output's closure (scalar_crosscheck.jl:38-48):
//   u   = fft(cis(D z) .* x)
//   q   = gamma |u|^2
//   dR  = real(fftshift(ifft(hR .* fft(q))))   -- fftshift folded into hR as (-1)^k
//   eta = ifft(((1-fR) q + dR) .* u) .* shock
//   dx  = i cis(-D z) .* eta
// and the classic global-frame RK4 of scalar_crosscheck.jl:53-60.
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
};

template <typename R, int N, int T> struct SmTraits;
// FP64: state/accumulator in global scratch; time field in registers when a
// workgroup has at most 512 threads (2 waves per SIMD -> 256 VGPRs each);
// the rotation cis(D z) is re-evaluated for the back-rotation instead of
// being held across the four transforms.
template <int N, int T> struct SmTraits<double, N, T> {
    static constexpr bool STATE_REGS = false;
    static constexpr bool UT_REGS = T <= 512 && N / T <= 8;
    static constexpr bool KEEP_PHASE = false;
    static constexpr bool PLANAR = (size_t)LdsSize<N>::value * 16 > 160 * 1024;
};
template <int N, int T> struct SmTraits<float, N, T> {
    static constexpr bool STATE_REGS = T <= 512;
    static constexpr bool UT_REGS = true;
    static constexpr bool KEEP_PHASE = true;
    static constexpr bool PLANAR = (size_t)LdsSize<N>::value * 8 > 160 * 1024;
};

template <typename R, int N, int T> struct SmLayout {
    using TR = SmTraits<R, N, T>;
    static constexpr int NW = (TR::STATE_REGS ? 0 : 2) + (TR::UT_REGS ? 0 : 1);
    static constexpr size_t LDS_BYTES = (size_t)LdsSize<N>::value * (TR::PLANAR ? sizeof(R) : 2 * sizeof(R));
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

// One right-hand-side evaluation.  On entry v holds the stage input y
// (interaction picture, cyclic layout); on exit v holds f(zs, y).
template <typename R, int N, int T>
__device__ __forceinline__ void sm_rhs(C2<R>* v, double zs, const SmArgs<R>& a, const SmTables<R, N>& tb,
                                       const Buf& gut, void* lds)
{
    using V = C2<R>;
    using TR = SmTraits<R, N, T>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    using CR = Cyc<R, T>;
    constexpr int P = N / T;
    R cs[TR::KEEP_PHASE ? P : 1], sn[TR::KEEP_PHASE ? P : 1];
    V ut[TR::UT_REGS ? P : 1];
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
        if constexpr (TR::UT_REGS)
            ut[k] = v[k];
        else
            gut.st(v[k], CV::voff(), CV::soff(k));
        const R q = a.gamma * (v[k].x * v[k].x + v[k].y * v[k].y);
        v[k] = V{q, (R)0};
    }
    block_fft<R, N, T, -1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = cmul(v[k], tb.hn.template ld<V>(CV::voff(), CV::soff(k)));
    block_fft<R, N, T, +1, TR::PLANAR>(v, tb.tw, lds);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        V u;
        if constexpr (TR::UT_REGS)
            u = ut[k];
        else
            u = gut.ld<V>(CV::voff(), CV::soff(k));
        const R q = a.gamma * (u.x * u.x + u.y * u.y);
        const R f = a.one_m_fr * q + v[k].x;
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

template <typename R, int N, int T>
__global__ __launch_bounds__(T) void k_rk4ip_single(SmArgs<R> a)
{
    using V = C2<R>;
    using TR = SmTraits<R, N, T>;
    using LY = SmLayout<R, N, T>;
    using CV = Cyc<V, T>;
    constexpr int P = N / T;
    __shared__ __align__(16) unsigned char lds[LY::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const SmTables<R, N> tb(a, FftPlan<N, T>::TW_SIZE);
    const Buf u0(a.u0 + sol * N, N * sizeof(V));
    V* gwork = a.work + sol * (size_t)(LY::NW > 0 ? LY::NW : 1) * N;
    const Buf gst(gwork, N * sizeof(V));
    const Buf gac(gwork + N, N * sizeof(V));
    const Buf gut(gwork + (TR::STATE_REGS ? 0 : 2 * N), N * sizeof(V));

    V v[P];
    V st[TR::STATE_REGS ? P : 1], ac[TR::STATE_REGS ? P : 1];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const V x = u0.ld<V>(CV::voff(), CV::soff(k));
        v[k] = x;
        if constexpr (TR::STATE_REGS)
            st[k] = x;
        else
            gst.st(x, CV::voff(), CV::soff(k));
    }
    const double h = a.h;
    const R h2 = (R)(h / 2), h1 = (R)h, h6 = (R)(h / 6);
    for (int step = 0; step < a.steps; ++step) {
        const double z = step * h;
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
            const double zs = s == 0 ? z : (s == 3 ? z + h : z + h / 2);
            sm_rhs<R, N, T>(v, zs, a, tb, gut, lds);
#pragma unroll
            for (int k = 0; k < P; ++k) {
                V x, c;
                if constexpr (TR::STATE_REGS) {
                    x = st[k];
                    c = ac[k];
                } else {
                    x = gst.ld<V>(CV::voff(), CV::soff(k));
                    c = s != 0 ? gac.ld<V>(CV::voff(), CV::soff(k)) : V{(R)0, (R)0};
                }
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
                    x = V{x.x + h6 * t.x, x.y + h6 * t.y};
                    v[k] = x;
                }
                if constexpr (TR::STATE_REGS) {
                    if (s == 3) st[k] = x;
                    ac[k] = c;
                } else {
                    if (s == 3)
                        gst.st(x, CV::voff(), CV::soff(k));
                    else
                        gac.st(c, CV::voff(), CV::soff(k));
                }
            }
        }
    }
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

}  // namespace fh
