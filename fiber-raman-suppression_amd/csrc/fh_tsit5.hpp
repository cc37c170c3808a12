// fh_tsit5.hpp -- adaptive Tsit5 propagation and the discrete adjoint of its
// accepted-step map, single mode, one workgroup per solve.
//
// This is the integrator of solve_disp_mmf (simulate_disp_mmf.jl:183-207:
// OrdinaryDiffEq Tsit5 on the interaction-picture RHS disp_mmf!) and of the
// backward sweep of solve_adjoint_disp_mmf (sensitivity_disp_mmf.jl:309-332).
// Every workgroup runs its own step-size sequence; the right-hand side and
// its pullback are the single-mode routines of fh_single.hpp.
//
// Tableau: Tsitouras 5(4) pair (Comput. Math. Appl. 62 (2011) 770-775), FSAL.
// Error control: e = h sum_j btilde_j k_j, EEst = sqrt(mean |e_i / (atol +
// max(|y_i|, |ynew_i|) rtol)|^2) accumulated in FP64 over the workgroup;
// accept when EEst <= 1.  PI controller beta1 = 7/50, beta2 = 2/25, safety
// 0.9, next step between h/5 and 10 h; rejected steps shrink by
// min(5, EEst^beta1 / 0.9).  Starting step: Hairer-Norsett-Wanner estimate
// (Solving ODEs I, II.4) from f(0, y0) and one trial evaluation.
//
// Workgroup-private fields live in one global slice addressed through a
// fixed descriptor; the y/ynew and k1/k7 roles are swapped by slot index on
// acceptance (first same as last), so no field is copied.
#pragma once
#include "fh_single.hpp"

namespace fh {

struct Tsit5Tab {
    static constexpr double c[7] = {0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0};
    static constexpr double a[7][6] = {
        {0, 0, 0, 0, 0, 0},
        {0.161, 0, 0, 0, 0, 0},
        {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
        {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
        {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
        {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0},
        {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774}};
    static constexpr double bt[7] = {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                                     -0.1447110071732629, 0.5823571654525552, -0.45808210592918697,
                                     0.015151515151515152};
    static constexpr double beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, safety = 0.9, qmin = 0.2, qmax = 10.0,
                            qoldinit = 1e-4;
    // Continuous extension (Tsitouras 2011; the interpolant OrdinaryDiffEq
    // evaluates for solution(t) inside a step): y(z + th h) = y + h sum_i
    // b_i(th) k_i with b_i(th) = th (r_i0 + th (r_i1 + th (r_i2 + th r_i3))).
    static constexpr double r[7][4] = {{1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
                                       {0.0, 0.13169999999999998, -0.2234, 0.1017},
                                       {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
                                       {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
                                       {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
                                       {0.0, -27.896526289197286, 65.09189467479366, -34.87065786149661},
                                       {0.0, 1.5, -4.0, 2.5}};
    // a[i][j] and c[i] at a workgroup-uniform runtime stage index, as a
    // select over immediates: the FP64 multimode units (built with forced
    // memory waits) must not address constant tables through PC-relative
    // sequences, which those waits would split.
    static __device__ __forceinline__ double a_at(int i, int j)
    {
        double r = 0.0;
#pragma unroll
        for (int ii = 0; ii < 7; ++ii)
#pragma unroll
            for (int jj = 0; jj < 6; ++jj)
                if (ii == i && jj == j) r = a[ii][jj];
        return r;
    }
    static __device__ __forceinline__ double c_at(int i)
    {
        double r = 0.0;
#pragma unroll
        for (int ii = 0; ii < 7; ++ii)
            if (ii == i) r = c[ii];
        return r;
    }
    static __device__ __forceinline__ double dense(int i, double th)
    {
        return th * (r[i][0] + th * (r[i][1] + th * (r[i][2] + th * r[i][3])));
    }
};

// Gain model of the pump + signal solve (disp_gain_smf!,
// simulate_disp_gain_mmf.jl:21-63, with calculate_gain_YDFA, gain.jl:148-167)
// as coefficients on |x_k|^2 in FFT order: coef rows w12, w21, ga, ge of N
// doubles each; W12 = sum w12 |x|^2, W21 = sum w21 |x|^2, R12 = r12 Pp,
// R21 = r21 Pp, n2 = (R12 + W12) / (R12 + R21 + W12 + W21 + inv_tau),
// g_k = ge_k n2 - ga_k n1, gP = gpe n2 - gpa n1.
struct GainDev {
    const double* coef = nullptr;
    double r12 = 0, r21 = 0, inv_tau = 0, gpa = 0, gpe = 0;
    const double* pump0 = nullptr;  // [batch] Pp(0)
    double* pumpL = nullptr;        // [batch] Pp(L)
};

template <typename R> struct TsArgs {
    SmArgs<R> s;
    double rtol, atol;
    int max_steps;
    double* hs;        // [batch][max_steps] accepted step sizes
    double* zs;        // [batch][max_steps] position before each accepted step
    long long* stats;  // [batch][3] accepted, rejected, rhs evaluations
    const double* lengths;  // [batch] fiber length per solve, or null for s.length
    // dense output of the forward solve (propagate(problem; saveat))
    const double* zsave;  // [nz] sorted save positions in [0, L], shared by the batch
    int nz;               // 0: no dense output
    C2<R>* saved;         // [batch][nz][N] lab-frame spectra cis(D z) y(z)
    GainDev gain;         // pump + signal solves only
    __device__ __forceinline__ double length(size_t sol) const { return lengths ? lengths[sol] : s.length; }
};

// Lab-frame sample at save position zj of the accepted step [z, zend]: the
// stored state at the step's ends (ode_interpolation returns stored states
// at step boundaries) and the continuous extension inside it.
template <typename R, int N, int T, typename KS>
__device__ __forceinline__ void ts_save(const SmTables<R, N>& tb, const Fld& dst, double zj, double z, double h,
                                        double zend, const Fld& y, const Fld& yn, const Fld& slot0, KS kslot)
{
    using V = C2<R>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = N / T;
    const uint32_t vo = CV::voff(), FB = (uint32_t)(N * sizeof(V));
    const bool at_start = zj == z, inside = zj != z && zj != zend;
    R bw[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) bw[i] = inside ? (R)Tsit5Tab::dense(i, (zj - z) / h) : (R)0;
    const R hr = (R)h;
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
                const V ki = Fld{slot0.b, slot0.off + (uint32_t)kslot(i) * FB}.template ld<V>(vo, CV::soff(k));
                acc = V{acc.x + bw[i] * ki.x, acc.y + bw[i] * ki.y};
            }
            const V u = y.template ld<V>(vo, CV::soff(k));
            val = V{u.x + hr * acc.x, u.y + hr * acc.y};
        }
        R c, s;
        cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), zj, c, s);
        dst.st(V{val.x * c - val.y * s, val.x * s + val.y * c}, vo, CV::soff(k));
    }
}

// Slot map of the workgroup slice.  Forward: y, ynew, k0..k6.  Gradient adds
// lam, lacc; the backward sweep reuses k1..k5 for the stage cotangents.
template <typename R, int N, int T> struct TsLayout {
    using TR = SmTraits<R, N, T>;
    static constexpr int Y = 0, YN = 1, K0 = 2, LAM = 9, LACC = 10;
    static constexpr int CORE_FWD = 9, CORE_GRAD = 11;
    static constexpr int EXTRA = TR::UT_REGS ? 0 : 2;  // ut, gu
    static constexpr int NW_FWD = CORE_FWD + EXTRA, NW_GRAD = CORE_GRAD + EXTRA;
};

// Gain terms of disp_gain_smf! at one stage evaluation (before the
// right-hand side): the population n2 from the stage input x (W12, W21 summed
// in FP64 over the workgroup) and the stage pump pin; x is kept in slot gs
// for the term g/2 .* x that gain_post adds to the right-hand side.  Returns
// n2; *kp = gP pin, the pump's derivative.
template <typename R, int N, int T>
__device__ __forceinline__ double gain_pre(const GainDev& g, const C2<R>* v, const Fld& gs, double pin, double* kp,
                                           void* lds)
{
    using V = C2<R>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = N / T;
    const Buf cf(g.coef, (uint32_t)(4 * N * sizeof(double)));
    constexpr uint32_t ROW = N * sizeof(double);
    double s12 = 0, s21 = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const double a2 = (double)v[k].x * v[k].x + (double)v[k].y * v[k].y;
        s12 += cf.ld<double>(CD::voff(), CD::soff(k)) * a2;
        s21 += cf.ld<double>(CD::voff(), CD::soff(k) + ROW) * a2;
        gs.st(v[k], CV::voff(), CV::soff(k));
    }
    const double w12 = block_sum<T>(s12, lds), w21 = block_sum<T>(s21, lds);
    const double r12 = g.r12 * pin, r21 = g.r21 * pin;
    const double n2 = (r12 + w12) / (r12 + r21 + w12 + w21 + g.inv_tau);
    *kp = (g.gpe * n2 - g.gpa * (1.0 - n2)) * pin;
    return n2;
}

// Adds g/2 .* x (simulate_disp_gain_mmf.jl:62) to the right-hand side v of
// the stage input x kept by gain_pre.
template <typename R, int N, int T>
__device__ __forceinline__ void gain_post(const GainDev& g, C2<R>* v, const Fld& gs, double n2)
{
    using V = C2<R>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = N / T;
    const Buf cf(g.coef, (uint32_t)(4 * N * sizeof(double)));
    constexpr uint32_t ROW = N * sizeof(double);
    const double n1 = 1.0 - n2;
    drain_memory();
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const double ga = cf.ld<double>(CD::voff(), CD::soff(k) + 2 * ROW);
        const double ge = cf.ld<double>(CD::voff(), CD::soff(k) + 3 * ROW);
        const R hg = (R)(0.5 * (ge * n2 - ga * n1));
        const V x = gs.template ld<V>(CV::voff(), CV::soff(k));
        v[k] = V{v[k].x + hg * x.x, v[k].y + hg * x.y};
    }
}

// |e| s / (atol + max(|u|, |w|) rtol), squared, for a real scalar entry.
__device__ __forceinline__ double scaled_sq_scalar(double e, double u, double w, double s, double rtol, double atol)
{
    const double r = fabs(e) * s / (atol + fmax(fabs(u), fabs(w)) * rtol);
    return r * r;
}

// Per-thread partial of sum_i |e_i s / (atol + max(|u_i|, |w_i|) rtol)|^2 in FP64.
template <typename V>
__device__ __forceinline__ double scaled_sq(V e, V u, V w, double s, double rtol, double atol)
{
    const double au = sqrt((double)u.x * u.x + (double)u.y * u.y);
    const double aw = sqrt((double)w.x * w.x + (double)w.y * w.y);
    const double r = sqrt((double)e.x * e.x + (double)e.y * e.y) * s / (atol + fmax(au, aw) * rtol);
    return r * r;
}

// Adaptive forward integration from (y slot, z = 0) to L.  On return the
// slot holding y(L) is in *ys; step sizes and positions are recorded; the
// workgroup-uniform status is 0 (ok), 1 (non-finite) or 2 (max_steps accepted
// or max_steps rejected steps).
// With CKPT, the state and position at accepted steps 0, every, 2 every, ...
// are stored to ckpt, and every stage evaluation stores its u and h to the
// stage tape: slot 6 (n mod every) + i of accepted-step index n, in the
// region of segment n / every when that segment is one of the first ntaped
// (kept), else in the shared region the backward sweep recomputes into.  An
// attempt that is rejected is overwritten by the next one; the seventh stage
// (first same as last) is stage 0 of step n + 1.
//
// With GAIN the state is (Pp, x) (the pump + signal solve): *pump holds Pp(0)
// on entry and Pp(L) on return, every stage adds the gain terms of
// gain_pre/gain_post, and the error norms run over the N + 1 entries.
template <typename R, int N, int T, bool CKPT, bool SAVE = false, bool GAIN = false>
__device__ int ts_forward(const TsArgs<R>& a, const SmTables<R, N>& tb, const Fld& slot0,
                          typename SmCtx<R, N, T>::FU& ut, const Buf& ckpt, const Buf& tape_u, const Buf& tape_h,
                          int* ys_out, long long* nacc_out, long long* nrej_out, long long* nfe_out, void* lds,
                          double* pump = nullptr)
{
    using V = C2<R>;
    using CV = Cyc<V, T>;
    using TB = Tsit5Tab;
    using LY = TsLayout<R, N, T>;
    constexpr int P = N / T;
    const uint32_t vo = CV::voff();
    const uint32_t FB = (uint32_t)(N * sizeof(V));
    auto fld = [&](int f) { return Fld{slot0.b, slot0.off + (uint32_t)f * FB}; };
    const Buf none(a.s.u0, 0);
    const Fld nf{none, 0};
    const size_t sol = blockIdx.x;
    const double L = a.length(blockIdx.x), rtol = a.rtol, atol = a.atol;
    int sy = LY::Y, syn = LY::YN;
    // k0 and k6 trade slots on acceptance (first same as last); k1..k5 are fixed
    int sk0 = LY::K0, sk6 = LY::K0 + 6;
    auto kslot = [&](int j) { return j == 0 ? sk0 : (j == 6 ? sk6 : LY::K0 + j); };
    auto tape_slot = [&](long long n, int i) {
        const int sg = (int)(n / a.s.every);
        const int region = sg < a.s.ntaped ? sg : a.s.ntaped;
        return (uint32_t)(region * 6 * a.s.every + 6 * (int)(n % a.s.every) + i);
    };
    auto tu_at = [&](uint32_t slot) { return CKPT ? Fld{tape_u, slot * (uint32_t)(N * sizeof(V))} : nf; };
    auto th_at = [&](uint32_t slot) { return CKPT ? Fld{tape_h, slot * (uint32_t)(N * sizeof(R))} : nf; };
    V v[P];
    long long nfe = 0, nacc = 0, nrej = 0;
    // pump + signal solves: pump state, its stage derivatives, the gain slot
    // and the entry count of the error norms
    [[maybe_unused]] double pp = GAIN ? *pump : 0.0, ppn = 0.0, kp[7] = {0, 0, 0, 0, 0, 0, 0};
    [[maybe_unused]] const Fld gs = fld(LY::NW_FWD);
    const bool lead = threadIdx.x == 0;
    constexpr double NN = GAIN ? (double)(N + 1) : (double)N;

    // k0 = f(0, y0) and the starting-step norms
    {
        const Fld y = fld(sy);
#pragma unroll
        for (int k = 0; k < P; ++k) v[k] = y.template ld<V>(vo, CV::soff(k));
    }
    {
        [[maybe_unused]] double n2 = 0;
        if constexpr (GAIN) n2 = gain_pre<R, N, T>(a.gain, v, gs, pp, &kp[0], lds);
        sm_rhs<R, N, T, CKPT>(v, 0.0, a.s, tb, ut, tu_at(tape_slot(0, 0)), th_at(tape_slot(0, 0)), lds);
        if constexpr (GAIN) gain_post<R, N, T>(a.gain, v, gs, n2);
    }
    nfe++;
    double s0 = 0, s1 = 0;
    {
        const Fld y = fld(sy), k0 = fld(sk0);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            k0.st(v[k], vo, CV::soff(k));
            const V u = y.template ld<V>(vo, CV::soff(k));
            const V z0 = V{(R)0, (R)0};
            s0 += scaled_sq(u, u, z0, 1.0, rtol, atol);
            s1 += scaled_sq(v[k], u, z0, 1.0, rtol, atol);
        }
        if (GAIN && lead) {
            s0 += scaled_sq_scalar(pp, pp, 0.0, 1.0, rtol, atol);
            s1 += scaled_sq_scalar(kp[0], pp, 0.0, 1.0, rtol, atol);
        }
    }
    const double d0 = sqrt(block_sum<T>(s0, lds) / NN), d1 = sqrt(block_sum<T>(s1, lds) / NN);
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    dt0 = fmin(dt0, L);
    {
        const Fld y = fld(sy), k0 = fld(sk0);
        drain_memory();
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V u = y.template ld<V>(vo, CV::soff(k));
            const V f0 = k0.template ld<V>(vo, CV::soff(k));
            v[k] = V{u.x + (R)dt0 * f0.x, u.y + (R)dt0 * f0.y};
        }
    }
    [[maybe_unused]] double kp1 = 0;
    {
        [[maybe_unused]] double n2 = 0;
        if constexpr (GAIN) n2 = gain_pre<R, N, T>(a.gain, v, gs, pp + dt0 * kp[0], &kp1, lds);
        sm_rhs<R, N, T, false>(v, dt0, a.s, tb, ut, nf, nf, lds);
        if constexpr (GAIN) gain_post<R, N, T>(a.gain, v, gs, n2);
    }
    nfe++;
    double s2 = 0;
    {
        const Fld y = fld(sy), k0 = fld(sk0);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V u = y.template ld<V>(vo, CV::soff(k));
            const V f0 = k0.template ld<V>(vo, CV::soff(k));
            s2 += scaled_sq(V{v[k].x - f0.x, v[k].y - f0.y}, u, V{(R)0, (R)0}, 1.0, rtol, atol);
        }
        if (GAIN && lead) s2 += scaled_sq_scalar(kp1 - kp[0], pp, 0.0, 1.0, rtol, atol);
    }
    const double d2 = sqrt(block_sum<T>(s2, lds) / NN) / dt0;
    const double mx = fmax(d1, d2);
    const double dt1 = mx <= 1e-15 ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / mx, 1.0 / 6.0);
    double h = fmin(fmin(100 * dt0, dt1), L);

    double z = 0.0, qold = TB::qoldinit;
    int status = 0, last_ck = -1;
    bool done = false;
    // Dense output (SAVE): samples at a.zsave[js..] as the accepted steps
    // pass them, through one fixed descriptor over this solve's saved region.
    // Kept out of the plain forward instantiation, whose arithmetic then
    // stays identical to the gradient kernel's forward sweep.
    int js = 0;
    if constexpr (SAVE) {
        const Buf sv(a.saved + sol * (size_t)a.nz * N, (uint32_t)((size_t)a.nz * FB));
        drain_memory();
        while (js < a.nz && a.zsave[js] == 0.0) {
            ts_save<R, N, T>(tb, Fld{sv, (uint32_t)js * FB}, 0.0, 0.0, 0.0, 0.0, fld(sy), fld(sy), slot0, kslot);
            ++js;
        }
    }
    while (!done) {
        if (nacc >= a.max_steps || nrej >= a.max_steps) {  // bounded for every wave
            status = 2;
            break;
        }
        if (CKPT && nacc % a.s.every == 0 && last_ck != (int)nacc) {
            const Fld y = fld(sy);
            const Fld ck{ckpt, (uint32_t)(nacc / a.s.every) * FB};
            drain_memory();
#pragma unroll
            for (int k = 0; k < P; ++k) ck.st(y.template ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
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
            const Fld y = fld(sy);
#pragma unroll
            for (int k = 0; k < P; ++k) {
                V s = V{(R)0, (R)0};
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    if (j < i) {
                        const R aij = (R)TB::a[i][j];
                        const V kj = fld(kslot(j)).template ld<V>(vo, CV::soff(k));
                        s = V{s.x + aij * kj.x, s.y + aij * kj.y};
                    }
                }
                const V u = y.template ld<V>(vo, CV::soff(k));
                v[k] = V{u.x + hr * s.x, u.y + hr * s.y};
            }
            if (i == 6) {
                const Fld yn = fld(syn);
#pragma unroll
                for (int k = 0; k < P; ++k) yn.st(v[k], vo, CV::soff(k));
            }
            const uint32_t ts = i < 6 ? tape_slot(nacc, i) : tape_slot(nacc + 1, 0);
            [[maybe_unused]] double n2 = 0;
            if constexpr (GAIN) {
                double sp = 0;
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    if (j < i) sp += TB::a_at(i, j) * kp[j];
                const double pin = pp + h * sp;
                if (i == 6) ppn = pin;
                double kpi = 0;
                n2 = gain_pre<R, N, T>(a.gain, v, gs, pin, &kpi, lds);
#pragma unroll
                for (int j = 1; j < 7; ++j)
                    if (j == i) kp[j] = kpi;
            }
            sm_rhs<R, N, T, CKPT>(v, z + TB::c[i] * h, a.s, tb, ut, tu_at(ts), th_at(ts), lds);
            if constexpr (GAIN) gain_post<R, N, T>(a.gain, v, gs, n2);
            const Fld ki = fld(kslot(i));
#pragma unroll
            for (int k = 0; k < P; ++k) ki.st(v[k], vo, CV::soff(k));
        }
        nfe += 6;
        drain_memory();
        double se = 0;
        {
            const Fld y = fld(sy), yn = fld(syn);
#pragma unroll
            for (int k = 0; k < P; ++k) {
                V e = V{(R)0, (R)0};
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    const V kj = fld(kslot(j)).template ld<V>(vo, CV::soff(k));
                    e = V{e.x + (R)TB::bt[j] * kj.x, e.y + (R)TB::bt[j] * kj.y};
                }
                se += scaled_sq(e, y.template ld<V>(vo, CV::soff(k)), yn.template ld<V>(vo, CV::soff(k)), fabs(h),
                                rtol, atol);
            }
            if (GAIN && lead) {
                double ep = 0;
#pragma unroll
                for (int j = 0; j < 7; ++j) ep += TB::bt[j] * kp[j];
                se += scaled_sq_scalar(ep, pp, ppn, fabs(h), rtol, atol);
            }
        }
        const double eest = sqrt(block_sum<T>(se, lds) / NN);
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
            if constexpr (SAVE) {
                const Buf sv(a.saved + sol * (size_t)a.nz * N, (uint32_t)((size_t)a.nz * FB));
                const double zend = last ? L : z + h;
                while (js < a.nz && a.zsave[js] <= zend) {
                    ts_save<R, N, T>(tb, Fld{sv, (uint32_t)js * FB}, a.zsave[js], z, h, zend, fld(sy), fld(syn),
                                     slot0, kslot);
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
            if constexpr (GAIN) {
                pp = ppn;
                kp[0] = kp[6];
            }
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
    if constexpr (GAIN) {
        *pump = pp;
        if (!isfinite(pp) && status == 0) status = 1;
    }
    return status;
}

template <typename R, int N, int T, bool SAVE>
__global__ __launch_bounds__(T) void k_tsit5_single(TsArgs<R> a)
{
    using CX = SmCtx<R, N, T>;
    using V = C2<R>;
    using LY = TsLayout<R, N, T>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = N / T;
    __shared__ __align__(16) unsigned char lds[SmLayout<R, N, T>::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const SmTables<R, N> tb(a.s, FftPlan<N, T>::TW_SIZE);
    V* slice = a.s.work + sol * (size_t)LY::NW_FWD * N;
    const Buf sl(slice, (uint32_t)((size_t)LY::NW_FWD * N * sizeof(V)));
    const Fld base{sl, 0};
    typename CX::FU ut(Buf(slice + (size_t)LY::CORE_FWD * N, N * sizeof(V)));
    const uint32_t vo = CV::voff();
    {
        const Buf u0(a.s.u0 + sol * N, N * sizeof(V));
        const Fld y{sl, (uint32_t)LY::Y * (uint32_t)(N * sizeof(V))};
#pragma unroll
        for (int k = 0; k < P; ++k) y.st(u0.ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
    }
    drain_memory();
    int ys;
    long long nacc, nrej, nfe;
    const Buf none(a.s.u0, 0);
    const int status = ts_forward<R, N, T, false, SAVE>(a, tb, base, ut, none, none, none, &ys, &nacc, &nrej, &nfe,
                                                        lds);
    if (threadIdx.x == 0) {
        a.stats[sol * 3 + 0] = nacc;
        a.stats[sol * 3 + 1] = nrej;
        a.stats[sol * 3 + 2] = nfe;
    }
    drain_memory();
    int bad = status;
    const Fld y{sl, (uint32_t)ys * (uint32_t)(N * sizeof(V))};
    const Buf out(a.s.out + sol * N, N * sizeof(V));
#pragma unroll
    for (int k = 0; k < P; ++k) {
        R c, s;
        cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), a.length(sol), c, s);
        const V x = y.template ld<V>(vo, CV::soff(k));
        const V o = V{x.x * c - x.y * s, x.x * s + x.y * c};
        if (!(isfinite(o.x) && isfinite(o.y))) bad |= 1;
        out.st(o, vo, CV::soff(k));
    }
    if (bad) atomicOr(a.s.status + sol, bad);
}

// Pump + signal solve (solve_disp_gain_smf, simulate_disp_gain_mmf.jl:200-239):
// one workgroup per solve, state (Pp, x), output cis(D L) x(L) and Pp(L).
// The slice holds the forward fields plus the gain slot.
template <typename R, int N, int T>
__global__ __launch_bounds__(T) void k_tsit5_gain(TsArgs<R> a)
{
    using CX = SmCtx<R, N, T>;
    using V = C2<R>;
    using LY = TsLayout<R, N, T>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = N / T;
    constexpr int NW = LY::NW_FWD + 1;
    __shared__ __align__(16) unsigned char lds[SmLayout<R, N, T>::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const SmTables<R, N> tb(a.s, FftPlan<N, T>::TW_SIZE);
    V* slice = a.s.work + sol * (size_t)NW * N;
    const Buf sl(slice, (uint32_t)((size_t)NW * N * sizeof(V)));
    const Fld base{sl, 0};
    typename CX::FU ut(Buf(slice + (size_t)LY::CORE_FWD * N, N * sizeof(V)));
    const uint32_t vo = CV::voff();
    {
        const Buf u0(a.s.u0 + sol * N, N * sizeof(V));
        const Fld y{sl, (uint32_t)LY::Y * (uint32_t)(N * sizeof(V))};
#pragma unroll
        for (int k = 0; k < P; ++k) y.st(u0.ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
    }
    drain_memory();
    int ys;
    long long nacc, nrej, nfe;
    double pump = a.gain.pump0[sol];
    const Buf none(a.s.u0, 0);
    const int status = ts_forward<R, N, T, false, false, true>(a, tb, base, ut, none, none, none, &ys, &nacc, &nrej,
                                                               &nfe, lds, &pump);
    if (threadIdx.x == 0) {
        a.stats[sol * 3 + 0] = nacc;
        a.stats[sol * 3 + 1] = nrej;
        a.stats[sol * 3 + 2] = nfe;
        a.gain.pumpL[sol] = pump;
    }
    drain_memory();
    int bad = status;
    const Fld y{sl, (uint32_t)ys * (uint32_t)(N * sizeof(V))};
    const Buf out(a.s.out + sol * N, N * sizeof(V));
#pragma unroll
    for (int k = 0; k < P; ++k) {
        R c, s;
        cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), a.length(sol), c, s);
        const V x = y.template ld<V>(vo, CV::soff(k));
        const V o = V{x.x * c - x.y * s, x.x * s + x.y * c};
        if (!(isfinite(o.x) && isfinite(o.y))) bad |= 1;
        out.st(o, vo, CV::soff(k));
    }
    if (bad) atomicOr(a.s.status + sol, bad);
}

// Forward with checkpoints, terminal seed (band objective or caller seed),
// backward sweep of the accepted-step map with per-segment recomputation
// (6 stage tapes per step), lam0 and the FP64 phase gradient.
template <typename R, int N, int T>
__global__ __launch_bounds__(T) void k_grad_tsit5(TsArgs<R> a)
{
    using CX = SmCtx<R, N, T>;
    using V = C2<R>;
    using LY = TsLayout<R, N, T>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    using TB = Tsit5Tab;
    constexpr int P = N / T;
    __shared__ __align__(16) unsigned char lds[SmLayout<R, N, T>::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const SmArgs<R>& sa = a.s;
    const SmTables<R, N> tb(sa, FftPlan<N, T>::TW_SIZE);
    V* slice = sa.work + sol * (size_t)LY::NW_GRAD * N;
    const uint32_t FB = (uint32_t)(N * sizeof(V));
    const Buf sl(slice, (uint32_t)((size_t)LY::NW_GRAD * N * sizeof(V)));
    const Fld base{sl, 0};
    auto fld = [&](int f) { return Fld{sl, (uint32_t)f * FB}; };
    typename CX::FU ut(Buf(slice + (size_t)LY::CORE_GRAD * N, N * sizeof(V)));
    typename CX::FU gu(Buf(slice + (size_t)(LY::CORE_GRAD + 1) * N, N * sizeof(V)));
    const Buf ckpt(sa.ckpt + sol * (size_t)sa.nseg * N, (uint32_t)((size_t)sa.nseg * N * sizeof(V)));
    // stage tape: one region of 6*every slots per kept segment (the first
    // ntaped) plus the shared region of the recomputed segments
    const size_t tslots = (size_t)(sa.ntaped + 1) * 6 * sa.every;
    const Buf tape_u(sa.tape_u + sol * tslots * N, (uint32_t)(tslots * N * sizeof(V)));
    const Buf tape_h(sa.tape_h + sol * tslots * N, (uint32_t)(tslots * N * sizeof(R)));
    const Buf u0(sa.u0 + sol * N, N * sizeof(V));
    const uint32_t vo = CV::voff();
    int ys = LY::Y, bad = 0;
    long long nacc, nrej, nfe;
    if (sa.phase & 1) {
        {
            const Fld y = fld(LY::Y);
#pragma unroll
            for (int k = 0; k < P; ++k) y.st(u0.ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
        }
        drain_memory();
        bad = ts_forward<R, N, T, true>(a, tb, base, ut, ckpt, tape_u, tape_h, &ys, &nacc, &nrej, &nfe, lds);
        if (threadIdx.x == 0) {
            a.stats[sol * 3 + 0] = nacc;
            a.stats[sol * 3 + 1] = nrej;
            a.stats[sol * 3 + 2] = nfe;
        }
        if (bad) {
            if (threadIdx.x == 0) atomicOr(sa.status + sol, bad);
            return;  // workgroup-uniform
        }
    } else {
        nacc = a.stats[sol * 3 + 0];  // accepted steps of the forward-only launch
    }
    drain_memory();
    // output and terminal seed
    V v[P];
    const Buf out(sa.out + sol * N, N * sizeof(V));
    double etot = 0.0, eband = 0.0;
    if (sa.phase & 1) {
        const Fld y = fld(ys);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            R c, s;
            cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), a.length(sol), c, s);
            const V x = y.template ld<V>(vo, CV::soff(k));
            const V o = V{x.x * c - x.y * s, x.x * s + x.y * c};
            if (!(isfinite(o.x) && isfinite(o.y))) bad |= 1;
            out.st(o, vo, CV::soff(k));
            v[k] = o;
            const double e = (double)o.x * o.x + (double)o.y * o.y;
            etot += e;
            if (!sa.seed && sa.mask[threadIdx.x + T * k]) eband += e;
        }
        if (!(sa.phase & 2)) {  // forward sweep only: checkpoints, kept tape and steps stay for the backward launch
            if (bad) atomicOr(sa.status + sol, bad);
            return;
        }
    }
    if (!sa.seed) {
        etot = block_sum<T>(etot, lds);
        eband = block_sum<T>(eband, lds);
        const double J = eband / etot;
        double scale = 1.0 / etot;
        if (sa.log_cost) scale *= J <= 1e-15 ? 0.0 : 10.0 / (2.302585092994045684 * J);
        if (threadIdx.x == 0) sa.cost[sol] = sa.log_cost ? 10.0 * log10(fmax(J, 1e-15)) : J;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const R f = (R)(((sa.mask[threadIdx.x + T * k] ? 1.0 : 0.0) - J) * scale);
            v[k] = V{v[k].x * f, v[k].y * f};
        }
    } else {
        const Buf sd(sa.seed + sol * N, N * sizeof(V));
#pragma unroll
        for (int k = 0; k < P; ++k) v[k] = sd.ld<V>(vo, CV::soff(k));
    }
    {
        const Fld lam = fld(LY::LAM);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            R c, s;
            cis_arg(tb.disp.template ld<double>(CD::voff(), CD::soff(k)), a.length(sol), c, s);
            lam.st(V{c * v[k].x + s * v[k].y, c * v[k].y - s * v[k].x}, vo, CV::soff(k));
        }
    }
    const Buf none(sa.u0, 0);
    const Fld nf{none, 0};
    const int nseg = (int)((nacc + sa.every - 1) / sa.every);
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * sa.every;
        const int s1 = (long long)s0 + sa.every < nacc ? s0 + sa.every : (int)nacc;
        const bool kept = sg < sa.ntaped;  // the forward sweep left this segment's tape
        const uint32_t rbase = (uint32_t)((kept ? sg : sa.ntaped) * 6 * sa.every);
        // recompute the segment with stage tapes; y lives in slot Y
        if (!kept) {
            const Fld ck{ckpt, (uint32_t)sg * FB};
            const Fld y = fld(LY::Y);
            drain_memory();
#pragma unroll
            for (int k = 0; k < P; ++k) y.st(ck.template ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
        }
        for (int s = s0; s < (kept ? s0 : s1); ++s) {
            const double h = a.hs[sol * a.max_steps + s], z = a.zs[sol * a.max_steps + s];
            const R hr = (R)h;
#pragma unroll 1
            for (int i = 0; i < 6; ++i) {
                drain_memory();
                const Fld y = fld(LY::Y);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    V acc = V{(R)0, (R)0};
#pragma unroll
                    for (int j = 0; j < 5; ++j) {
                        if (j < i) {
                            const R aij = (R)TB::a[i][j];
                            const V kj = fld(LY::K0 + j).template ld<V>(vo, CV::soff(k));
                            acc = V{acc.x + aij * kj.x, acc.y + aij * kj.y};
                        }
                    }
                    const V u = y.template ld<V>(vo, CV::soff(k));
                    v[k] = V{u.x + hr * acc.x, u.y + hr * acc.y};
                }
                const uint32_t slot = rbase + (uint32_t)(6 * (s - s0) + i);
                const Fld tu{tape_u, slot * (uint32_t)(N * sizeof(V))};
                const Fld th{tape_h, slot * (uint32_t)(N * sizeof(R))};
                sm_rhs<R, N, T, true>(v, z + TB::c[i] * h, sa, tb, ut, tu, th, lds);
                const Fld ki = fld(LY::K0 + i);
#pragma unroll
                for (int k = 0; k < P; ++k) ki.st(v[k], vo, CV::soff(k));
            }
            if (s + 1 < s1) {
                drain_memory();
                const Fld y = fld(LY::Y);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    V acc = V{(R)0, (R)0};
#pragma unroll
                    for (int j = 0; j < 6; ++j) {
                        const V kj = fld(LY::K0 + j).template ld<V>(vo, CV::soff(k));
                        acc = V{acc.x + (R)TB::a[6][j] * kj.x, acc.y + (R)TB::a[6][j] * kj.y};
                    }
                    const V u = y.template ld<V>(vo, CV::soff(k));
                    y.st(V{u.x + hr * acc.x, u.y + hr * acc.y}, vo, CV::soff(k));
                }
            }
        }
        // backward over the segment: stage cotangents gY_i in slots K0+i (i >= 1)
        for (int s = s1 - 1; s >= s0; --s) {
            const double h = a.hs[sol * a.max_steps + s], z = a.zs[sol * a.max_steps + s];
            const R hr = (R)h;
            drain_memory();
            {
                const Fld lam = fld(LY::LAM), lacc = fld(LY::LACC);
#pragma unroll
                for (int k = 0; k < P; ++k) lacc.st(lam.template ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
            }
#pragma unroll 1
            for (int i = 5; i >= 0; --i) {
                drain_memory();
                const Fld lam = fld(LY::LAM);
                const R bi = hr * (R)TB::a[6][i];
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    const V l = lam.template ld<V>(vo, CV::soff(k));
                    V g = V{bi * l.x, bi * l.y};
#pragma unroll
                    for (int m = 1; m < 6; ++m) {
                        if (m > i) {
                            const R ami = hr * (R)TB::a[m][i];
                            const V gy = fld(LY::K0 + m).template ld<V>(vo, CV::soff(k));
                            g = V{g.x + ami * gy.x, g.y + ami * gy.y};
                        }
                    }
                    v[k] = g;
                }
                const uint32_t slot = rbase + (uint32_t)(6 * (s - s0) + i);
                const Fld tu{tape_u, slot * (uint32_t)(N * sizeof(V))};
                const Fld th{tape_h, slot * (uint32_t)(N * sizeof(R))};
                sm_vjp<R, N, T>(v, z + TB::c[i] * h, sa, tb, gu, tu, th, lds);
                const Fld lacc = fld(LY::LACC);
                if (i >= 1) {
                    const Fld gi = fld(LY::K0 + i);
#pragma unroll
                    for (int k = 0; k < P; ++k) gi.st(v[k], vo, CV::soff(k));
                }
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    const V c = lacc.template ld<V>(vo, CV::soff(k));
                    lacc.st(V{c.x + v[k].x, c.y + v[k].y}, vo, CV::soff(k));
                }
            }
            drain_memory();
            {
                const Fld lam = fld(LY::LAM), lacc = fld(LY::LACC);
#pragma unroll
                for (int k = 0; k < P; ++k) lam.st(lacc.template ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
            }
        }
    }
    drain_memory();
    const Buf l0(sa.lam0 + sol * N, N * sizeof(V));
    const Buf gr(sa.grad + sol * N, N * sizeof(double));
    {
        const Fld lam = fld(LY::LAM);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V l = lam.template ld<V>(vo, CV::soff(k));
            const V u = u0.ld<V>(vo, CV::soff(k));
            l0.st(l, vo, CV::soff(k));
            const double g = 2.0 * ((double)l.y * (double)u.x - (double)l.x * (double)u.y);
            if (!(isfinite(l.x) && isfinite(l.y))) bad |= 1;
            gr.st(g, CD::voff(), CD::soff(k));
        }
    }
    if (bad) atomicOr(sa.status + sol, bad);
}

}  // namespace fh
