// fh_multi.hpp -- fused multimode RK4IP propagation, one workgroup per solve.
//
// Right-hand side, restating disp_mmf! (simulate_disp_mmf.jl:25-60) for M modes:
//   u_m   = fft(cis(D_m z) .* y_m)                                    M transforms
//   dK_ij = sum_kl gamma[i,j,k,l] Re(u_k conj(u_l))                   per sample
//   dR_ij = real(fftshift(ifft(hR .* fft(dK_ij))))                    Q real signals
//   h_ij  = (1-fR) dK_ij + dR_ij
//   k_i   = i cis(-D_i z) .* shock .* ifft(sum_j h_ij u_j)            M transforms
//
// Reductions against the reference's operation count (same values):
// * Re(u_k conj u_l) is symmetric in (k,l): the contraction runs over the
//   P = M(M+1)/2 pairs k <= l with Gr[q][p] = gamma[i,j,k,l] + gamma[i,j,l,k]
//   (k < l) or gamma[i,j,k,k];
// * when gamma is symmetric in (i,j), dK_ij = dK_ji and only the Q = M(M+1)/2
//   output pairs i <= j are formed (otherwise Q = M^2);
// * dK is real and hR is Hermitian, so two real signals share one complex
//   transform pair: ifft(H fft(a + i b)) = R(a) + i R(b).
// Per right-hand side: 2M + 2 ceil(Q/2) transforms instead of 2M + 2M^2.
//
// All per-sample data (stage inputs, state, accumulator, time fields, h
// pairs) lives in a workgroup-private global slice addressed in the cyclic
// layout, so outside the transforms every thread only touches its own
// samples and no barrier is needed.
#pragma once
#include "fh_device.hpp"

namespace fh {

template <typename R> struct MmArgs {
    const C2<R>* u0;     // [batch][M][N]
    C2<R>* out;          // [batch][M][N]
    C2<R>* work;         // [batch][4M + QC][N]
    const double* disp;  // [M][N]
    const C2<R>* hn;     // [N] hR (-1)^k / N
    const R* shn;        // [N] shock / N
    const C2<R>* tw;     // Stockham twiddles
    const R* gr;         // [Q][P] reduced gamma
    const int* qidx;     // [M][M] output-pair index of (i,j)
    int* status;
    R one_m_fr;
    double h, length;
    int steps;
    int q;  // number of output pairs (M(M+1)/2 or M^2)
    const double* lengths;  // [batch] per-solve fiber length, or null for `length`
    const int* nsteps;      // [batch] per-solve step count, or null for `steps`
};

template <typename R, int N, int T, int M> struct MmTraits {
    static constexpr int P = N / T;
    static constexpr int NP = M * (M + 1) / 2;  // input pairs k <= l
    static constexpr bool PLANAR = (size_t)LdsSize<N>::value * 2 * sizeof(R) > LdsBudget<R>::bytes;
    static constexpr size_t LDS_BYTES = (size_t)LdsSize<N>::value * (PLANAR ? sizeof(R) : 2 * sizeof(R));
    static constexpr int QMAX = M * M;
    static constexpr int QCMAX = (QMAX + 1) / 2;
    // slice layout: y[M], st[M], ac[M], u[M], hq[QC]
    static constexpr int FIELDS = 4 * M + QCMAX;
};

template <typename R, int N, int T, int M>
__global__ __launch_bounds__(T) void k_rk4ip_multi(MmArgs<R> a)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    using CR = Cyc<R, T>;
    constexpr int P = TR::P;
    constexpr int NP = TR::NP;
    __shared__ __align__(16) unsigned char lds[TR::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const int Q = a.q, QC = (a.q + 1) / 2;
    // One descriptor for the whole workgroup slice, fixed for the kernel's
    // lifetime; field f is addressed through the per-lane offset
    // vo + f * N * sizeof(V) (see Fld).
    V* slice = a.work + sol * (size_t)TR::FIELDS * N;
    const Buf sl(slice, (uint32_t)((size_t)TR::FIELDS * N * sizeof(V)));
    auto fld = [&](int f) { return Fld{sl, (uint32_t)f * (uint32_t)(N * sizeof(V))}; };
    const Buf disp(a.disp, M * N * 8), hn(a.hn, N * sizeof(V)), shn(a.shn, N * sizeof(R));
    const Buf tw(a.tw, FftPlan<N, T>::TW_SIZE * sizeof(V));
    const Buf in(a.u0 + sol * M * N, M * N * sizeof(V));
    const R* __restrict__ gr = a.gr;
    const uint32_t vo = CV::voff();

    // y = st = u0
    for (int m = 0; m < M; ++m) {
        const Fld y = fld(m), st = fld(M + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V x = in.ld<V>(vo, (uint32_t)(m * N * sizeof(V)) + CV::soff(k));
            y.st(x, vo, CV::soff(k));
            st.st(x, vo, CV::soff(k));
        }
    }
    const double Lb = a.lengths ? a.lengths[sol] : a.length;
    const int Sb = a.nsteps ? a.nsteps[sol] : a.steps;
    const double h = a.lengths ? Lb / Sb : a.h;
    const R h2 = (R)(h / 2), h1 = (R)h, h6 = (R)(h / 6);
    V v[P];
    for (int step = 0; step < Sb; ++step) {
        const double z = step * h;
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
            const double zs = s == 0 ? z : (s == 3 ? z + h : z + h / 2);
            drain_memory();  // re-reads y / st / ac of the previous stage
            // 1. u_m = fft(cis(D_m zs) y_m)
#pragma unroll 1
            for (int m = 0; m < M; ++m) {
                const Fld y = fld(m), u = fld(3 * M + m);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    R c, sn;
                    cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), zs, c, sn);
                    const V x = y.template ld<V>(vo, CV::soff(k));
                    v[k] = V{x.x * c - x.y * sn, x.x * sn + x.y * c};
                }
                block_fft<R, N, T, -1, TR::PLANAR>(v, tw, lds);
#pragma unroll
                for (int k = 0; k < P; ++k) u.st(v[k], vo, CV::soff(k));
            }
            drain_memory();  // re-reads u of step 1
            // 2. Kerr contraction per sample: dK_q = sum_p Gr[q][p] B_p
#pragma unroll 1
            for (int k = 0; k < P; ++k) {
                V uk[M];
#pragma unroll
                for (int m = 0; m < M; ++m) uk[m] = fld(3 * M + m).template ld<V>(vo, CV::soff(k));
                R b[NP];
                {
                    int p = 0;
#pragma unroll
                    for (int i = 0; i < M; ++i)
#pragma unroll
                        for (int j = i; j < M; ++j) b[p++] = uk[i].x * uk[j].x + uk[i].y * uk[j].y;
                }
#pragma unroll 1
                for (int c = 0; c < QC; ++c) {
                    R d0 = 0, d1 = 0;
                    const R* g0 = gr + (size_t)(2 * c) * NP;
#pragma unroll
                    for (int p = 0; p < NP; ++p) d0 += g0[p] * b[p];
                    if (2 * c + 1 < Q) {
                        const R* g1 = g0 + NP;
#pragma unroll
                        for (int p = 0; p < NP; ++p) d1 += g1[p] * b[p];
                    }
                    fld(4 * M + c).st(V{d0, d1}, vo, CV::soff(k));
                }
            }
            drain_memory();  // re-reads hq of step 2
            // 3. Raman on packed pairs: h = (1-fR) dK + real(ifft(H fft(dK)))
#pragma unroll 1
            for (int c = 0; c < QC; ++c) {
                const Fld hq = fld(4 * M + c);
#pragma unroll
                for (int k = 0; k < P; ++k) v[k] = hq.template ld<V>(vo, CV::soff(k));
                block_fft<R, N, T, -1, TR::PLANAR>(v, tw, lds);
#pragma unroll
                for (int k = 0; k < P; ++k) v[k] = cmul(v[k], hn.ld<V>(vo, CV::soff(k)));
                block_fft<R, N, T, +1, TR::PLANAR>(v, tw, lds);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    const V d = hq.template ld<V>(vo, CV::soff(k));
                    hq.st(V{a.one_m_fr * d.x + v[k].x, a.one_m_fr * d.y + v[k].y}, vo, CV::soff(k));
                }
            }
            drain_memory();  // re-reads hq of step 3
            // 4. per mode: eta_i = sum_j h_ij u_j; k_i; RK4 bookkeeping
#pragma unroll 1
            for (int i = 0; i < M; ++i) {
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    V acc = V{(R)0, (R)0};
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        const int q = a.qidx[i * M + j];
                        const V hp = fld(4 * M + (q >> 1)).template ld<V>(vo, CV::soff(k));
                        const R hij = (q & 1) ? hp.y : hp.x;
                        const V uj = fld(3 * M + j).template ld<V>(vo, CV::soff(k));
                        acc = V{acc.x + hij * uj.x, acc.y + hij * uj.y};
                    }
                    v[k] = acc;
                }
                block_fft<R, N, T, +1, TR::PLANAR>(v, tw, lds);
                const Fld y = fld(i), st = fld(M + i), ac = fld(2 * M + i);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    R c, sn;
                    cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(i * N * 8) + CD::soff(k)), zs, c, sn);
                    const R sc = shn.ld<R>(CR::voff(), CR::soff(k));
                    const V e = V{v[k].x * sc, v[k].y * sc};
                    const V kk = V{sn * e.x - c * e.y, sn * e.y + c * e.x};
                    const V x = st.template ld<V>(vo, CV::soff(k));
                    V cc = s != 0 ? ac.template ld<V>(vo, CV::soff(k)) : V{(R)0, (R)0};
                    V yn;
                    if (s == 0) {
                        cc = kk;
                        yn = V{x.x + h2 * kk.x, x.y + h2 * kk.y};
                    } else if (s == 1) {
                        cc = V{cc.x + 2 * kk.x, cc.y + 2 * kk.y};
                        yn = V{x.x + h2 * kk.x, x.y + h2 * kk.y};
                    } else if (s == 2) {
                        cc = V{cc.x + 2 * kk.x, cc.y + 2 * kk.y};
                        yn = V{x.x + h1 * kk.x, x.y + h1 * kk.y};
                    } else {
                        const V t = V{cc.x + kk.x, cc.y + kk.y};
                        yn = V{x.x + h6 * t.x, x.y + h6 * t.y};
                        st.st(yn, vo, CV::soff(k));
                    }
                    if (s != 3) ac.st(cc, vo, CV::soff(k));
                    y.st(yn, vo, CV::soff(k));
                }
            }
        }
    }
    int bad = 0;
    const Buf out(a.out + sol * M * N, M * N * sizeof(V));
    for (int m = 0; m < M; ++m) {
        const Fld st = fld(M + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            R c, sn;
            cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), Lb, c, sn);
            const V x = st.template ld<V>(vo, CV::soff(k));
            const V o = V{x.x * c - x.y * sn, x.x * sn + x.y * c};
            bad |= !(isfinite(o.x) && isfinite(o.y));
            out.st(o, vo, (uint32_t)(m * N * sizeof(V)) + CV::soff(k));
        }
    }
    if (bad) atomicOr(a.status + sol, 1);
}

}  // namespace fh
