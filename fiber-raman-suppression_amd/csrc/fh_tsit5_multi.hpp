// fh_tsit5_multi.hpp -- adaptive Tsit5 propagation for M modes, one
// workgroup per solve: solve_disp_mmf (simulate_disp_mmf.jl:183-207) on the
// multimode right-hand side disp_mmf! (mm_rhs of fh_multi.hpp).
//
// The tableau, error control and step controller are those of the
// single-mode driver (fh_tsit5.hpp).  The error norm is the RMS over all
// N*M complex entries of e / (atol + max(|y|, |ynew|) rtol), accumulated in
// FP64 over the workgroup, which is OrdinaryDiffEq's norm on the (N, M)
// state array.  Every solve of a launch runs its own step sequence and
// fiber length.
//
// Slice layout (fields of N samples, groups of M): y, ynew (also the stage
// input of stages 1..5), k0..k6, u (time fields), hq (packed h pairs).  k0
// and k6 trade group slots on acceptance (first same as last), as y and ynew
// do, so no field is copied.
#pragma once
#include "fh_multi.hpp"
#include "fh_tsit5.hpp"

namespace fh {

template <typename R> struct MtArgs {
    MmArgs<R> m;
    double rtol, atol;
    int max_steps;
    double* hs;        // [batch][max_steps] accepted step sizes
    double* zs;        // [batch][max_steps] position before each accepted step
    long long* stats;  // [batch][3] accepted, rejected, rhs evaluations
    const double* zsave;  // [nz] sorted save positions in [0, L], shared by the batch
    int nz;               // 0: no dense output
    C2<R>* saved;         // [batch][nz][M][N] lab-frame spectra cis(D_m z) y_m(z)
};

template <int M> struct MtLayout {
    static constexpr int Y = 0, YN = M, K0 = 2 * M, U = 9 * M, HQ = 10 * M;
    static constexpr int FIELDS = 10 * M + (M * M + 1) / 2;  // split transforms add MmTraits::XTRA after these
};

// Gradient slice: the forward phase uses MtLayout; the backward phase keeps
// y, u, hq below the adjoint fields of MmTraits (lam, lacc, gx, G, gu, bq,
// which mm_vjp addresses) and the six stage slopes k0..k5 after them.  The
// stage slopes double as the stage-input cotangents gY_1..gY_5 of the
// backward sweep; gx holds the stage inputs during the recomputation.
template <typename R, int N, int T, int M> struct MtgLayout {
    using TR = MmTraits<R, N, T, M>;
    static constexpr int Y = 0, U = M, HQ = 2 * M, K0 = TR::FIELDS_GRAD;
    static constexpr int FIELDS = TR::FIELDS_GRAD + 6 * M;  // split transforms add TR::XTRA after these
    static_assert(HQ + TR::QCMAX <= TR::LAM, "backward scratch must stay below the adjoint fields");
    static_assert(FIELDS >= MtLayout<M>::FIELDS, "the forward phase shares the slice");
};

// Adaptive integration from (y group of MtLayout, z = 0) to this solve's L.
// On return *ys_out is the group holding y(L).  Accepted step sizes and
// positions are recorded; with CKPT the state at accepted steps 0, every,
// 2 every, ... is stored to ckpt (M fields per checkpoint); without CKPT the
// dense-output samples of a.zsave are written.  Returns the workgroup-uniform
// status: 0 ok, 1 non-finite, 2 step limit.
// With CKPT every stage evaluation also stores its u and h pairs to the stage
// tape: slot 6 (n mod every) + i of accepted step n, in the region of segment
// n / every when that segment is one of the first ntaped (kept), else in the
// shared region the backward sweep recomputes into.  A rejected attempt is
// overwritten by the next one; the seventh stage is stage 0 of step n + 1.
template <typename R, int N, int T, int M, bool CKPT>
__device__ __forceinline__ int mt_forward(const MtArgs<R>& a, const MmCtx<R, N, T, M>& c, const Buf& ckpt,
                                          const Buf& tape_u, const Buf& tape_h, int* ys_out, long long* nacc_out,
                                          long long* nrej_out, long long* nfe_out, C2<R>* v, void* lds)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using LY = MtLayout<M>;
    using TB = Tsit5Tab;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = TR::P;
    const size_t sol = blockIdx.x;
    const MmArgs<R>& ma = a.m;
    const uint32_t vo = CV::voff();
    const uint32_t FB = (uint32_t)(N * sizeof(V));
    const double L = ma.lengths ? ma.lengths[sol] : ma.length;
    const double rtol = a.rtol, atol = a.atol;
    const double nm = (double)N * M;
    const Buf none(ma.u0, 0);
    int sy = LY::Y, syn = LY::YN, sk0 = LY::K0, sk6 = LY::K0 + 6 * M;
    auto kslot = [&](int j) { return j == 0 ? sk0 : (j == 6 ? sk6 : LY::K0 + j * M); };
    auto tape_slot = [&](long long n, int i) {
        const int sg = (int)(n / ma.every);
        const int region = sg < ma.ntaped ? sg : ma.ntaped;
        return (uint32_t)(region * 6 * ma.every + 6 * (int)(n % ma.every) + i);
    };
    long long nfe = 0, nacc = 0, nrej = 0;

    // dense output: lab-frame sample of every mode at zj (see ts_save)
    auto save = [&](int j, double zj, double z0, double hh, double zend) {
        const Buf sv(a.saved + sol * (size_t)a.nz * M * N, (uint32_t)((size_t)a.nz * M * FB));
        const bool at_start = zj == z0, inside = zj != z0 && zj != zend;
        R bw[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) bw[i] = inside ? (R)TB::dense(i, (zj - z0) / hh) : (R)0;
        const R hr = (R)hh;
#pragma unroll 1
        for (int m = 0; m < M; ++m) {
            const Fld y = c.fld(sy + m), yn = c.fld(syn + m), dst{sv, (uint32_t)(j * M + m) * FB};
#pragma unroll
            for (int k = 0; k < P; ++k) {
                V val;
                if (at_start) {
                    val = y.template ld<V>(vo, CV::soff(k));
                } else if (!inside) {
                    val = yn.template ld<V>(vo, CV::soff(k));
                } else {
                    V acc = V{(R)0, (R)0};
#pragma unroll
                    for (int i = 0; i < 7; ++i) {
                        const V ki = c.fld(kslot(i) + m).template ld<V>(vo, CV::soff(k));
                        acc = V{acc.x + bw[i] * ki.x, acc.y + bw[i] * ki.y};
                    }
                    const V u = y.template ld<V>(vo, CV::soff(k));
                    val = V{u.x + hr * acc.x, u.y + hr * acc.y};
                }
                R cs, sn;
                cis_arg(c.disp.template ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), zj, cs, sn);
                dst.st(V{val.x * cs - val.y * sn, val.x * sn + val.y * cs}, CV::voff(), CV::soff(k));
            }
        }
    };

    // k0 = f(0, y0) and the starting-step norms (Hairer-Norsett-Wanner II.4)
    mm_rhs<R, N, T, M, CKPT>(c, 0.0, sy, sk0, LY::U, LY::HQ, CKPT ? tape_u : none, CKPT ? tape_h : none,
                             CKPT ? tape_slot(0, 0) : 0u, v, lds);
    nfe++;
    drain_memory();
    double s0 = 0, s1 = 0;
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld y = c.fld(sy + m), k0 = c.fld(sk0 + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V u = y.template ld<V>(vo, CV::soff(k));
            const V f0 = k0.template ld<V>(vo, CV::soff(k));
            const V z0 = V{(R)0, (R)0};
            s0 += scaled_sq(u, u, z0, 1.0, rtol, atol);
            s1 += scaled_sq(f0, u, z0, 1.0, rtol, atol);
        }
    }
    const double d0 = sqrt(block_sum<T>(s0, lds) / nm), d1 = sqrt(block_sum<T>(s1, lds) / nm);
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    dt0 = fmin(dt0, L);
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld y = c.fld(sy + m), k0 = c.fld(sk0 + m), yn = c.fld(syn + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V u = y.template ld<V>(vo, CV::soff(k));
            const V f0 = k0.template ld<V>(vo, CV::soff(k));
            yn.st(V{u.x + (R)dt0 * f0.x, u.y + (R)dt0 * f0.y}, vo, CV::soff(k));
        }
    }
    mm_rhs<R, N, T, M>(c, dt0, syn, kslot(1), LY::U, LY::HQ, none, none, 0, v, lds);
    nfe++;
    drain_memory();
    double s2 = 0;
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld y = c.fld(sy + m), k0 = c.fld(sk0 + m), k1 = c.fld(kslot(1) + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V u = y.template ld<V>(vo, CV::soff(k));
            const V f0 = k0.template ld<V>(vo, CV::soff(k));
            const V f1 = k1.template ld<V>(vo, CV::soff(k));
            s2 += scaled_sq(V{f1.x - f0.x, f1.y - f0.y}, u, V{(R)0, (R)0}, 1.0, rtol, atol);
        }
    }
    const double d2 = sqrt(block_sum<T>(s2, lds) / nm) / dt0;
    const double mx = fmax(d1, d2);
    const double dt1 = mx <= 1e-15 ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / mx, 1.0 / 6.0);
    double h = fmin(fmin(100 * dt0, dt1), L);

    double z = 0.0, qold = TB::qoldinit;
    int status = 0, js = 0, last_ck = -1;
    if constexpr (!CKPT) {
        while (js < a.nz && a.zsave[js] == 0.0) {
            save(js, 0.0, 0.0, 0.0, 0.0);
            ++js;
        }
    }
    bool done = false;
    while (!done) {
        if (nacc >= a.max_steps || nrej >= a.max_steps) {  // bounded for every wave
            status = 2;
            break;
        }
        if (CKPT && nacc % ma.every == 0 && last_ck != (int)nacc) {
            drain_memory();
#pragma unroll 1
            for (int m = 0; m < M; ++m) {
                const Fld y = c.fld(sy + m);
                const Fld ck{ckpt, ((uint32_t)(nacc / ma.every) * (uint32_t)M + (uint32_t)m) * FB};
#pragma unroll
                for (int k = 0; k < P; ++k) ck.st(y.template ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
            }
            last_ck = (int)nacc;
        }
        bool last = false;
        if (z + h >= L) {
            h = L - z;
            last = true;
        }
        const R hr = (R)h;
#pragma unroll 1
        for (int i = 1; i < 7; ++i) {
            drain_memory();
#pragma unroll 1
            for (int m = 0; m < M; ++m) {
                const Fld y = c.fld(sy + m), yn = c.fld(syn + m);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    V s = V{(R)0, (R)0};
#pragma unroll
                    for (int j = 0; j < 6; ++j) {
                        if (j < i) {
                            const R aij = (R)TB::a_at(i, j);
                            const V kj = c.fld(kslot(j) + m).template ld<V>(vo, CV::soff(k));
                            s = V{s.x + aij * kj.x, s.y + aij * kj.y};
                        }
                    }
                    const V u = y.template ld<V>(vo, CV::soff(k));
                    yn.st(V{u.x + hr * s.x, u.y + hr * s.y}, vo, CV::soff(k));
                }
            }
            const uint32_t ts = CKPT ? (i < 6 ? tape_slot(nacc, i) : tape_slot(nacc + 1, 0)) : 0u;
            mm_rhs<R, N, T, M, CKPT>(c, z + TB::c_at(i) * h, syn, kslot(i), LY::U, LY::HQ, CKPT ? tape_u : none,
                                     CKPT ? tape_h : none, ts, v, lds);
        }
        nfe += 6;
        drain_memory();
        double se = 0;
#pragma unroll 1
        for (int m = 0; m < M; ++m) {
            const Fld y = c.fld(sy + m), yn = c.fld(syn + m);
#pragma unroll
            for (int k = 0; k < P; ++k) {
                V e = V{(R)0, (R)0};
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    const V kj = c.fld(kslot(j) + m).template ld<V>(vo, CV::soff(k));
                    e = V{e.x + (R)TB::bt[j] * kj.x, e.y + (R)TB::bt[j] * kj.y};
                }
                se += scaled_sq(e, y.template ld<V>(vo, CV::soff(k)), yn.template ld<V>(vo, CV::soff(k)), fabs(h),
                                rtol, atol);
            }
        }
        const double eest = sqrt(block_sum<T>(se, lds) / nm);
        if (!isfinite(eest)) {
            status = 1;
            break;
        }
        const double q11 = pow(eest, TB::beta1);
        if (eest <= 1.0) {
            double q = q11 / pow(qold, TB::beta2) / TB::safety;
            q = fmin(fmax(q, 1.0 / TB::qmax), 1.0 / TB::qmin);
            qold = fmax(eest, TB::qoldinit);
            if (threadIdx.x == 0) {
                a.hs[sol * a.max_steps + nacc] = h;
                a.zs[sol * a.max_steps + nacc] = z;
            }
            if constexpr (!CKPT) {
                const double zend = last ? L : z + h;
                while (js < a.nz && a.zsave[js] <= zend) {
                    save(js, a.zsave[js], z, h, zend);
                    ++js;
                }
            }
            nacc++;
            z += h;
            const int t0 = sy;
            sy = syn;
            syn = t0;
            const int t1 = sk0;
            sk0 = sk6;
            sk6 = t1;
            h = h / q;
            done = last;
        } else {
            nrej++;
            h = h / fmin(1.0 / TB::qmin, q11 / TB::safety);
        }
    }
    *ys_out = sy;
    *nacc_out = nacc;
    *nrej_out = nrej;
    *nfe_out = nfe;
    return status;
}

template <typename R, int N, int T, int M>
__global__ __launch_bounds__(T) void k_tsit5_multi(MtArgs<R> a)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using LY = MtLayout<M>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = TR::P;
    __shared__ __align__(16) unsigned char lds[TR::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const MmArgs<R>& ma = a.m;
    V* slice = ma.work + sol * (size_t)(LY::FIELDS + TR::XTRA) * N;
    const Buf sl(slice, (uint32_t)((size_t)(LY::FIELDS + TR::XTRA) * N * sizeof(V)));
    const Buf disp(ma.disp, M * N * 8), hn(ma.hn, N * sizeof(V)), shn(ma.shn, N * sizeof(R));
    const Buf tw(ma.tw, mm_twiddle_entries<R, N, T>() * sizeof(V));
    const Buf in(ma.u0 + sol * M * N, M * N * sizeof(V));
    const MmCtx<R, N, T, M> c{ma, sl, disp, hn, shn, tw, ma.q, (ma.q + 1) / 2, LY::FIELDS};
    const uint32_t vo = CV::voff();
    const uint32_t FB = (uint32_t)(N * sizeof(V));
    const double L = ma.lengths ? ma.lengths[sol] : ma.length;
    V v[TR::PV];
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld y = c.fld(LY::Y + m);
#pragma unroll
        for (int k = 0; k < P; ++k) y.st(in.ld<V>(CV::voff(), (uint32_t)m * FB + CV::soff(k)), vo, CV::soff(k));
    }
    int ys;
    long long nacc, nrej, nfe;
    const Buf none(ma.u0, 0);
    const int status = mt_forward<R, N, T, M, false>(a, c, none, none, none, &ys, &nacc, &nrej, &nfe, v, lds);
    if (threadIdx.x == 0) {
        a.stats[sol * 3 + 0] = nacc;
        a.stats[sol * 3 + 1] = nrej;
        a.stats[sol * 3 + 2] = nfe;
    }
    drain_memory();
    int bad = status;
    const Buf out(ma.out + sol * M * N, M * N * sizeof(V));
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld y = c.fld(ys + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            R cs, sn;
            cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), L, cs, sn);
            const V x = y.template ld<V>(vo, CV::soff(k));
            const V o = V{x.x * cs - x.y * sn, x.x * sn + x.y * cs};
            if (!(isfinite(o.x) && isfinite(o.y))) bad |= 1;
            out.st(o, CV::voff(), (uint32_t)m * FB + CV::soff(k));
        }
    }
    if (bad) atomicOr(ma.status + sol, bad);
}

// Gradient of the terminal objective through the discrete adjoint of the
// accepted-step map (the multimode counterpart of k_grad_tsit5): forward with
// checkpoints, output and seed (band objective summed over modes, mmf_cost_sum,
// or the caller's seed), then per segment a recomputation with the six stage
// tapes (u and h pairs) and the backward sweep
//   gk_i = h b_i lam' + h sum_{m > i} a_mi gY_m,   gY_i = J_i^T gk_i,
//   lam  = lam' + sum_i gY_i,
// with J_i^T the multimode pullback mm_vjp; finally lam0 and the shared-phase
// gradient sum_m 2 Re(conj(lam0_m) i u0_m) in FP64.
template <typename R, int N, int T, int M>
__global__ __launch_bounds__(T) void k_grad_tsit5_multi(MtArgs<R> a)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using LF = MtLayout<M>;
    using LB = MtgLayout<R, N, T, M>;
    using TB = Tsit5Tab;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = TR::P;
    __shared__ __align__(16) unsigned char lds[TR::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const MmArgs<R>& ma = a.m;
    const uint32_t FB = (uint32_t)(N * sizeof(V));
    V* slice = ma.work + sol * (size_t)(LB::FIELDS + TR::XTRA) * N;
    const Buf sl(slice, (uint32_t)((size_t)(LB::FIELDS + TR::XTRA) * N * sizeof(V)));
    const Buf disp(ma.disp, M * N * 8), hn(ma.hn, N * sizeof(V)), shn(ma.shn, N * sizeof(R));
    const Buf tw(ma.tw, mm_twiddle_entries<R, N, T>() * sizeof(V));
    const Buf in(ma.u0 + sol * M * N, M * N * sizeof(V));
    const int Q = ma.q, QC = (ma.q + 1) / 2;
    const MmCtx<R, N, T, M> c{ma, sl, disp, hn, shn, tw, Q, QC, LB::FIELDS};
    const Buf ckpt(ma.ckpt + sol * (size_t)ma.nseg * M * N, (uint32_t)((size_t)ma.nseg * M * N * sizeof(V)));
    // stage tape: one region of 6*every slots per kept segment (the first
    // ntaped) plus the shared region of the recomputed segments
    const size_t tslots = (size_t)(ma.ntaped + 1) * 6 * ma.every;
    const Buf tape_u(ma.tape_u + sol * tslots * M * N, (uint32_t)(tslots * M * N * sizeof(V)));
    const Buf tape_h(ma.tape_h + sol * tslots * QC * N, (uint32_t)(tslots * QC * N * sizeof(V)));
    const uint32_t vo = CV::voff();
    const double L = ma.lengths ? ma.lengths[sol] : ma.length;
    V v[TR::PV];
    int ys = LF::Y, bad = 0;
    long long nacc, nrej, nfe;
    if (ma.phase & 1) {
#pragma unroll 1
        for (int m = 0; m < M; ++m) {
            const Fld y = c.fld(LF::Y + m);
#pragma unroll
            for (int k = 0; k < P; ++k) y.st(in.ld<V>(CV::voff(), (uint32_t)m * FB + CV::soff(k)), vo, CV::soff(k));
        }
        bad = mt_forward<R, N, T, M, true>(a, c, ckpt, tape_u, tape_h, &ys, &nacc, &nrej, &nfe, v, lds);
        if (threadIdx.x == 0) {
            a.stats[sol * 3 + 0] = nacc;
            a.stats[sol * 3 + 1] = nrej;
            a.stats[sol * 3 + 2] = nfe;
        }
        if (bad) {
            if (threadIdx.x == 0) atomicOr(ma.status + sol, bad);
            return;  // workgroup-uniform
        }
    } else {
        nacc = a.stats[sol * 3 + 0];  // accepted steps of the forward-only launch
    }
    drain_memory();
    // output and terminal seed (band energy over all modes)
    const Buf out(ma.out + sol * M * N, M * N * sizeof(V));
    double etot = 0.0, eband = 0.0;
    if (ma.phase & 1) {
#pragma unroll 1
        for (int m = 0; m < M; ++m) {
            const Fld y = c.fld(ys + m), lam = c.fld(TR::LAM + m);
#pragma unroll
            for (int k = 0; k < P; ++k) {
                R cs, sn;
                cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), L, cs, sn);
                const V x = y.template ld<V>(vo, CV::soff(k));
                const V o = V{x.x * cs - x.y * sn, x.x * sn + x.y * cs};
                bad |= !(isfinite(o.x) && isfinite(o.y));
                out.st(o, CV::voff(), (uint32_t)m * FB + CV::soff(k));
                lam.st(o, vo, CV::soff(k));  // lab-frame output, scaled into the seed below
                const double e = (double)o.x * o.x + (double)o.y * o.y;
                etot += e;
                if (!ma.seed && ma.mask[threadIdx.x + T * k]) eband += e;
            }
        }
        if (!(ma.phase & 2)) {  // forward sweep only: checkpoints, kept tape and steps stay for the backward launch
            if (bad) atomicOr(ma.status + sol, bad);
            return;
        }
    }
    double J = 0.0, scale = 0.0;
    if (!ma.seed) {
        etot = mm_block_sum<T>(etot, lds);
        eband = mm_block_sum<T>(eband, lds);
        J = eband / etot;
        scale = 1.0 / etot;
        if (ma.log_cost) scale *= J <= 1e-15 ? 0.0 : 10.0 / (2.302585092994045684 * J);
        if (threadIdx.x == 0) ma.cost[sol] = ma.log_cost ? 10.0 * log10(fmax(J, 1e-15)) : J;
    }
    drain_memory();
    // lam = cis(-D L) seed
    const Buf sd(ma.seed ? ma.seed + sol * M * N : ma.u0, ma.seed ? (uint32_t)(M * N * sizeof(V)) : 0u);
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld lam = c.fld(TR::LAM + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            V s;
            if (ma.seed) {
                s = sd.ld<V>(CV::voff(), (uint32_t)m * FB + CV::soff(k));
            } else {
                const V o = lam.template ld<V>(vo, CV::soff(k));
                const R f = (R)(((ma.mask[threadIdx.x + T * k] ? 1.0 : 0.0) - J) * scale);
                s = V{o.x * f, o.y * f};
            }
            R cs, sn;
            cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), L, cs, sn);
            lam.st(V{cs * s.x + sn * s.y, cs * s.y - sn * s.x}, vo, CV::soff(k));
        }
    }
    const int nseg = (int)((nacc + ma.every - 1) / ma.every);
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * ma.every;
        const int s1 = (long long)s0 + ma.every < nacc ? s0 + ma.every : (int)nacc;
        const bool kept = sg < ma.ntaped;  // the forward sweep left this segment's tape
        const uint32_t rbase = (uint32_t)((kept ? sg : ma.ntaped) * 6 * ma.every);
        drain_memory();
#pragma unroll 1
        for (int m = 0; m < (kept ? 0 : M); ++m) {
            const Fld ck{ckpt, ((uint32_t)sg * (uint32_t)M + (uint32_t)m) * FB};
            const Fld y = c.fld(LB::Y + m);
#pragma unroll
            for (int k = 0; k < P; ++k) y.st(ck.template ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
        }
        // recompute the segment with stage tapes; stage inputs in gx
        for (int s = s0; s < (kept ? s0 : s1); ++s) {
            const double h = a.hs[sol * a.max_steps + s], z = a.zs[sol * a.max_steps + s];
            const R hr = (R)h;
#pragma unroll 1
            for (int i = 0; i < 6; ++i) {
                if (i > 0) {
                    drain_memory();
#pragma unroll 1
                    for (int m = 0; m < M; ++m) {
                        const Fld y = c.fld(LB::Y + m), st = c.fld(TR::GX + m);
#pragma unroll
                        for (int k = 0; k < P; ++k) {
                            V acc = V{(R)0, (R)0};
#pragma unroll
                            for (int j = 0; j < 5; ++j) {
                                if (j < i) {
                                    const R aij = (R)TB::a_at(i, j);
                                    const V kj = c.fld(LB::K0 + j * M + m).template ld<V>(vo, CV::soff(k));
                                    acc = V{acc.x + aij * kj.x, acc.y + aij * kj.y};
                                }
                            }
                            const V u = y.template ld<V>(vo, CV::soff(k));
                            st.st(V{u.x + hr * acc.x, u.y + hr * acc.y}, vo, CV::soff(k));
                        }
                    }
                }
                mm_rhs<R, N, T, M, true>(c, z + TB::c_at(i) * h, i == 0 ? LB::Y : TR::GX, LB::K0 + i * M, LB::U,
                                         LB::HQ, tape_u, tape_h, rbase + (uint32_t)(6 * (s - s0) + i), v, lds);
            }
            if (s + 1 < s1) {
                drain_memory();
#pragma unroll 1
                for (int m = 0; m < M; ++m) {
                    const Fld y = c.fld(LB::Y + m);
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        V acc = V{(R)0, (R)0};
#pragma unroll
                        for (int j = 0; j < 6; ++j) {
                            const V kj = c.fld(LB::K0 + j * M + m).template ld<V>(vo, CV::soff(k));
                            acc = V{acc.x + (R)TB::a[6][j] * kj.x, acc.y + (R)TB::a[6][j] * kj.y};
                        }
                        const V u = y.template ld<V>(vo, CV::soff(k));
                        y.st(V{u.x + hr * acc.x, u.y + hr * acc.y}, vo, CV::soff(k));
                    }
                }
            }
        }
        // backward over the segment: gY_i kept in the k slots (i >= 1)
        for (int s = s1 - 1; s >= s0; --s) {
            const double h = a.hs[sol * a.max_steps + s], z = a.zs[sol * a.max_steps + s];
            const R hr = (R)h;
            drain_memory();
#pragma unroll 1
            for (int m = 0; m < M; ++m) {
                const Fld lam = c.fld(TR::LAM + m), lacc = c.fld(TR::LACC + m);
#pragma unroll
                for (int k = 0; k < P; ++k) lacc.st(lam.template ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
            }
#pragma unroll 1
            for (int i = 5; i >= 0; --i) {
                // gx = h sum_{m > i} a_mi gY_m (the h b_i lam' part enters through cl)
                if (i < 5) {
                    drain_memory();
#pragma unroll 1
                    for (int m = 0; m < M; ++m) {
                        const Fld gx = c.fld(TR::GX + m);
#pragma unroll
                        for (int k = 0; k < P; ++k) {
                            V g = V{(R)0, (R)0};
#pragma unroll
                            for (int r = 1; r < 6; ++r) {
                                if (r > i) {
                                    const R ari = hr * (R)TB::a_at(r, i);
                                    const V gy = c.fld(LB::K0 + r * M + m).template ld<V>(vo, CV::soff(k));
                                    g = V{g.x + ari * gy.x, g.y + ari * gy.y};
                                }
                            }
                            gx.st(g, vo, CV::soff(k));
                        }
                    }
                }
                mm_vjp<R, N, T, M>(c, z + TB::c_at(i) * h, hr * (R)TB::a_at(6, i), i < 5 ? (R)1 : (R)0, tape_u, tape_h,
                                   rbase + (uint32_t)(6 * (s - s0) + i), v, lds);
                drain_memory();
#pragma unroll 1
                for (int m = 0; m < M; ++m) {
                    const Fld gx = c.fld(TR::GX + m), lacc = c.fld(TR::LACC + m);
                    const Fld gi = c.fld(LB::K0 + i * M + m);
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const V g = gx.template ld<V>(vo, CV::soff(k));
                        if (i >= 1) gi.st(g, vo, CV::soff(k));
                        const V acc = lacc.template ld<V>(vo, CV::soff(k));
                        lacc.st(V{acc.x + g.x, acc.y + g.y}, vo, CV::soff(k));
                    }
                }
            }
            drain_memory();
#pragma unroll 1
            for (int m = 0; m < M; ++m) {
                const Fld lam = c.fld(TR::LAM + m), lacc = c.fld(TR::LACC + m);
#pragma unroll
                for (int k = 0; k < P; ++k) lam.st(lacc.template ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
            }
        }
    }
    drain_memory();
    // lam0 and the shared-phase gradient in FP64
    const Buf l0(ma.lam0 + sol * M * N, M * N * sizeof(V));
    const Buf gb(ma.grad + sol * N, N * sizeof(double));
    double g[P];
#pragma unroll
    for (int k = 0; k < P; ++k) g[k] = 0.0;
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld lam = c.fld(TR::LAM + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V l = lam.template ld<V>(vo, CV::soff(k));
            const V u = in.ld<V>(CV::voff(), (uint32_t)m * FB + CV::soff(k));
            l0.st(l, CV::voff(), (uint32_t)m * FB + CV::soff(k));
            g[k] += 2.0 * ((double)l.y * (double)u.x - (double)l.x * (double)u.y);
            bad |= !(isfinite(l.x) && isfinite(l.y));
        }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) gb.st(g[k], CD::voff(), CD::soff(k));
    if (bad) atomicOr(ma.status + sol, 1);
}

}  // namespace fh
