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
    static constexpr int FIELDS = 10 * M + (M * M + 1) / 2;
};

template <typename R, int N, int T, int M>
__global__ __launch_bounds__(T) void k_tsit5_multi(MtArgs<R> a)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using LY = MtLayout<M>;
    using TB = Tsit5Tab;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = TR::P;
    __shared__ __align__(16) unsigned char lds[TR::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const MmArgs<R>& ma = a.m;
    V* slice = ma.work + sol * (size_t)LY::FIELDS * N;
    const Buf sl(slice, (uint32_t)((size_t)LY::FIELDS * N * sizeof(V)));
    const Buf disp(ma.disp, M * N * 8), hn(ma.hn, N * sizeof(V)), shn(ma.shn, N * sizeof(R));
    const Buf tw(ma.tw, FftPlan<N, T>::TW_SIZE * sizeof(V));
    const Buf in(ma.u0 + sol * M * N, M * N * sizeof(V));
    const MmCtx<R, N, T, M> c{ma, sl, disp, hn, shn, tw, ma.q, (ma.q + 1) / 2};
    const uint32_t vo = CV::voff();
    const uint32_t FB = (uint32_t)(N * sizeof(V));
    const double L = ma.lengths ? ma.lengths[sol] : ma.length;
    const double rtol = a.rtol, atol = a.atol;
    const double nm = (double)N * M;
    int sy = LY::Y, syn = LY::YN, sk0 = LY::K0, sk6 = LY::K0 + 6 * M;
    auto kslot = [&](int j) { return j == 0 ? sk0 : (j == 6 ? sk6 : LY::K0 + j * M); };
    V v[P];
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
                cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), zj, cs, sn);
                dst.st(V{val.x * cs - val.y * sn, val.x * sn + val.y * cs}, vo, CV::soff(k));
            }
        }
    };

#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld y = c.fld(sy + m);
#pragma unroll
        for (int k = 0; k < P; ++k) y.st(in.ld<V>(vo, (uint32_t)m * FB + CV::soff(k)), vo, CV::soff(k));
    }

    // k0 = f(0, y0) and the starting-step norms (Hairer-Norsett-Wanner II.4)
    mm_rhs<R, N, T, M>(c, 0.0, sy, sk0, LY::U, LY::HQ, v, lds);
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
    mm_rhs<R, N, T, M>(c, dt0, syn, kslot(1), LY::U, LY::HQ, v, lds);
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
    int status = 0, js = 0;
    while (js < a.nz && a.zsave[js] == 0.0) {
        save(js, 0.0, 0.0, 0.0, 0.0);
        ++js;
    }
    bool done = false;
    while (!done) {
        if (nacc >= a.max_steps || nrej >= a.max_steps) {  // bounded for every wave
            status = 2;
            break;
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
                            const R aij = (R)TB::a[i][j];
                            const V kj = c.fld(kslot(j) + m).template ld<V>(vo, CV::soff(k));
                            s = V{s.x + aij * kj.x, s.y + aij * kj.y};
                        }
                    }
                    const V u = y.template ld<V>(vo, CV::soff(k));
                    yn.st(V{u.x + hr * s.x, u.y + hr * s.y}, vo, CV::soff(k));
                }
            }
            mm_rhs<R, N, T, M>(c, z + TB::c[i] * h, syn, kslot(i), LY::U, LY::HQ, v, lds);
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
            const double zend = last ? L : z + h;
            while (js < a.nz && a.zsave[js] <= zend) {
                save(js, a.zsave[js], z, h, zend);
                ++js;
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
        const Fld y = c.fld(sy + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            R cs, sn;
            cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), L, cs, sn);
            const V x = y.template ld<V>(vo, CV::soff(k));
            const V o = V{x.x * cs - x.y * sn, x.x * sn + x.y * cs};
            if (!(isfinite(o.x) && isfinite(o.y))) bad |= 1;
            out.st(o, vo, (uint32_t)m * FB + CV::soff(k));
        }
    }
    if (bad) atomicOr(ma.status + sol, bad);
}

}  // namespace fh
