// fh_multi.hpp -- fused multimode RK4IP propagation and its discrete adjoint,
// one workgroup per solve.
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
// Pullback of the right-hand side (cotangent gy_i of k_i; real-linear map):
//   G_i   = fft(-i cis(D_i z) (shock/N) gy_i)                          M transforms
//   A_q   = sum_{(i,j) -> q} Re(conj(G_i) u_j)        cotangent of h_q
//   B_q   = (1-fR) A_q + R^T(A_q),  R^T(a) = real(ifft(conj(hR)(-1)^k fft(a)))
//   C_p   = sum_q Gr[q][p] B_q                        cotangent of Re(u_k conj u_l)
//   gu_j  = sum_i h_ij G_i + sum_{p=(k,l)} C_p d(Re u_k conj u_l)/du_j
//   gx_m  = conj(cis(D_m z)) N ifft(gu_m)                             M transforms
// with u and h of the stage taken from the tape of the segment
// recomputation; the pair reductions give the forward's transform count.
//
// All per-sample data (stage inputs, state, accumulator, time fields, h
// pairs, adjoint fields) lives in a workgroup-private global slice addressed
// in the cyclic layout through one fixed descriptor, so outside the
// transforms every thread only touches its own samples.
#pragma once
#include <type_traits>

#include "fh_device.hpp"

namespace fh {

template <typename R> struct MmArgs {
    const C2<R>* u0;     // [batch][M][N]
    C2<R>* out;          // [batch][M][N]
    C2<R>* work;         // [batch][fields][N]
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
    // adjoint only
    const C2<R>* seed;          // [batch][M][N] lab-frame Wirtinger seed, or null for the band objective
    const unsigned char* mask;  // [N] band mask (band energy summed over modes, mmf_cost_sum)
    int log_cost;
    C2<R>* lam0;    // [batch][M][N]
    double* grad;   // [batch][N] sum_m 2 Re(conj(lam0_m) i u0_m)
    double* cost;   // [batch]
    C2<R>* ckpt;    // [batch][nseg][M][N]
    C2<R>* tape_u;  // [batch][regions][4 every][M][N]
    C2<R>* tape_h;  // [batch][regions][4 every][QC][N]
    int every, nseg;
    int ntaped;  // the last ntaped segments of each solve keep the forward sweep's tape (k_grad_multi)
    int phase;   // gradient kernels: 1 forward sweep, 2 seed and backward sweep, 3 both (SmArgs::phase)
};

// Multimode fields transformed in two halves (mm_xform): FP64 with more than
// 16 samples per thread at the 1024-thread workgroup shape.
template <typename R, int N> constexpr bool mm_split() { return sizeof(R) == 8 && N > 16384; }
inline bool mm_split_rt(int nt, bool fp64) { return fp64 && nt > 16384; }

template <typename R, int N, int T, int M> struct MmTraits {
    static constexpr int P = N / T;  // field elements per thread
    // FP64 fields of more than 16 samples per thread are transformed in two
    // halves (mm_xform): NX-point block transforms of PV registers each
    static constexpr bool SPLIT = mm_split<R, N>();
    static constexpr int NX = SPLIT ? N / 2 : N;
    static constexpr int PV = NX / T;
    static constexpr int XTRA = SPLIT ? 2 : 0;  // scratch fields of split transforms
    static constexpr int NP = M * (M + 1) / 2;  // input pairs k <= l
    static constexpr bool PLANAR = (size_t)LdsSize<NX>::value * 2 * sizeof(R) > LdsBudget<R>::bytes;
    static constexpr size_t LDS_BYTES = (size_t)LdsSize<NX>::value * (PLANAR ? sizeof(R) : 2 * sizeof(R));
    static constexpr int QMAX = M * M;
    static constexpr int QCMAX = (QMAX + 1) / 2;
    // forward slice: y[M], st[M], ac[M], u[M], hq[QC], split scratch[XTRA]
    static constexpr int XA = 4 * M + QCMAX;
    static constexpr int FIELDS = 4 * M + QCMAX + XTRA;
    // gradient slice adds lam[M], lacc[M], gx[M], G[M], gu[M], bq[QC]
    static constexpr int LAM = FIELDS, LACC = LAM + M, GX = LACC + M, GG = GX + M, GU = GG + M, BQ = GU + M;
    static constexpr int FIELDS_GRAD = BQ + QCMAX;
};

// Output-pair index q of (i, j), in the order the plan builds qidx: the
// pairs i <= j in row order when gamma is symmetric in (i, j) (h_ij = h_ji),
// otherwise i M + j.  With i, j unrolled constants it indexes register
// arrays at compile time.
__host__ __device__ constexpr int pair_index(int M, int i, int j, bool sym)
{
    return sym ? (i <= j ? i * M - i * (i - 1) / 2 + (j - i) : j * M - j * (j - 1) / 2 + (i - j)) : i * M + j;
}

// Input pairs p of the symmetric products Re(u_i conj u_j), i <= j in row
// order (the column order of the reduced gamma Gr[q][p]), as a constant table
// that folds to immediates when indexed from unrolled loops.
template <int M> struct InputPairs {
    static constexpr int NP = M * (M + 1) / 2;
    int i[NP], j[NP];
    constexpr InputPairs() : i(), j()
    {
        int p = 0;
        for (int a = 0; a < M; ++a)
            for (int b = a; b < M; ++b) {
                i[p] = a;
                j[p] = b;
                ++p;
            }
    }
};

// Per-workgroup views shared by the stage and pullback routines.  xa: the
// first of the two slice fields split transforms use as scratch.
template <typename R, int N, int T, int M> struct MmCtx {
    using V = C2<R>;
    const MmArgs<R>& a;
    const Buf& sl;
    const Buf &disp, &hn, &shn, &tw;
    int Q, QC;
    int xa;
    __device__ __forceinline__ Fld fld(int f) const { return Fld{sl, (uint32_t)f * (uint32_t)(N * sizeof(V))}; }
};

// Entries of the twiddle table of a multimode plan: the block transform's,
// and for split transforms the N/2 radix-2 twiddles w^n after them.
template <typename R, int N, int T> constexpr int mm_twiddle_entries()
{
    return mm_split<R, N>() ? FftPlan<N / 2, T>::TW_SIZE + N / 2 : FftPlan<N, T>::TW_SIZE;
}

// Block transform of one field-sized signal: this thread's elements k < P are
// v_k = pre(k), transformed (DIR -1: exp(-2 pi i n k / N), +1: its
// conjugate, both unnormalised), and handed to post(k, x_k).  Element k is
// index t + T k on both sides.
//
// Split transforms (FP64 with 32 samples per thread): the N-point transform
// is one radix-2 step that stays inside each thread and two N/2-point block
// transforms of P/2 registers each.  The -1 transform then takes its input in
// natural order (element k is index t + T k) and returns its output in split
// order (element k < P/2 is index 2 (t + T k), element k + P/2 the index after
// it); the +1 transform takes split order and returns natural order:
//   -1:  a_n = x_n + x_{n+N/2},  b_n = (x_n - x_{n+N/2}) w^n,  w = exp(-2 pi i / N),
//        X_2m = FFT(a)_m,  X_2m+1 = FFT(b)_m
//   +1:  e = IFFT(X_2m),  o = IFFT(X_2m+1),  x_n = e_n + w^-n o_n,  x_{n+N/2} = e_n - w^-n o_n
// The state transforms map the frequency-domain fields (-1 input, +1 output)
// to the time-domain fields, so frequency-domain fields, their tables and the
// caller's arrays keep natural order and time-domain fields are in split
// order (pointwise work there does not depend on the order; mm_convolve
// handles the Raman convolution).  Half of the signal waits in scratch field
// c.xa meanwhile.  Every pre(k) is read before any post(k) is written, so a
// transform may work in place.
template <typename R, int N, int T, int M, int DIR, class Pre, class Post>
__device__ __forceinline__ void mm_xform(const MmCtx<R, N, T, M>& c, Pre&& pre, Post&& post, C2<R>* v, void* lds)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    constexpr int P = TR::P;
    if constexpr (!TR::SPLIT) {
#pragma unroll
        for (int k = 0; k < P; ++k) v[k] = pre(k);
        block_fft<R, N, T, DIR, TR::PLANAR>(v, c.tw, lds);
#pragma unroll
        for (int k = 0; k < P; ++k) post(k, v[k]);
    } else {
        using CV = Cyc<V, T>;
        constexpr int PH = P / 2, NH = N / 2;
        const uint32_t vo = CV::voff();
        const uint32_t wb = vo + (uint32_t)(FftPlan<NH, T>::TW_SIZE * sizeof(V));
        const Fld xs = c.fld(c.xa);
        if constexpr (DIR < 0) {
#pragma unroll
            for (int j = 0; j < PH; ++j) {
                const V a = pre(j), b = pre(j + PH);
                v[j] = cadd(a, b);
                xs.st(cmul(csub(a, b), c.tw.template ld<V>(wb, CV::soff(j))), vo, CV::soff(j));
            }
            block_fft<R, NH, T, -1, TR::PLANAR>(v, c.tw, lds);
#pragma unroll
            for (int j = 0; j < PH; ++j) post(j, v[j]);
            drain_memory();  // re-reads the odd half from scratch
#pragma unroll
            for (int j = 0; j < PH; ++j) v[j] = xs.template ld<V>(vo, CV::soff(j));
            block_fft<R, NH, T, -1, TR::PLANAR>(v, c.tw, lds);
#pragma unroll
            for (int j = 0; j < PH; ++j) post(j + PH, v[j]);
        } else {
#pragma unroll
            for (int j = 0; j < PH; ++j) v[j] = pre(j);
            block_fft<R, NH, T, +1, TR::PLANAR>(v, c.tw, lds);
#pragma unroll
            for (int j = 0; j < PH; ++j) xs.st(v[j], vo, CV::soff(j));
#pragma unroll
            for (int j = 0; j < PH; ++j) v[j] = pre(j + PH);
            block_fft<R, NH, T, +1, TR::PLANAR>(v, c.tw, lds);
            drain_memory();  // re-reads the even half from scratch
#pragma unroll
            for (int j = 0; j < PH; ++j) {
                const V e = xs.template ld<V>(vo, CV::soff(j));
                const V o = cmul(v[j], cconj(c.tw.template ld<V>(wb, CV::soff(j))));
                post(j, cadd(e, o));
                post(j + PH, csub(e, o));
            }
        }
    }
}

// Raman convolution of one packed pair: pre(k) -> -1 transform -> times hn
// (conj(hn) with CONJ) -> +1 transform -> post(k, x_k).  With split
// transforms the time-domain input is in split order, so the convolution runs
// the other way round: +1 transform (split in, natural out), times the plan's
// index-reversed table hn[-k], -1 transform (natural in, split out), which is
// the same circular convolution; the spectrum waits in scratch field c.xa + 1
// between the two.
template <typename R, int N, int T, int M, bool CONJ, class Pre, class Post>
__device__ __forceinline__ void mm_convolve(const MmCtx<R, N, T, M>& c, Pre&& pre, Post&& post, C2<R>* v,
                                            void* lds)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    constexpr int P = TR::P;
    const uint32_t vo = CV::voff();
    auto H = [&](int k) {
        const V h = c.hn.template ld<V>(vo, CV::soff(k));
        return CONJ ? cconj(h) : h;
    };
    if constexpr (!TR::SPLIT) {
#pragma unroll
        for (int k = 0; k < P; ++k) v[k] = pre(k);
        block_fft<R, N, T, -1, TR::PLANAR>(v, c.tw, lds);
#pragma unroll
        for (int k = 0; k < P; ++k) v[k] = cmul(v[k], H(k));
        block_fft<R, N, T, +1, TR::PLANAR>(v, c.tw, lds);
#pragma unroll
        for (int k = 0; k < P; ++k) post(k, v[k]);
    } else {
        const Fld xb = c.fld(c.xa + 1);
        mm_xform<R, N, T, M, +1>(c, pre, [&](int k, V x) { xb.st(cmul(x, H(k)), vo, CV::soff(k)); }, v, lds);
        drain_memory();  // re-reads the spectrum
        mm_xform<R, N, T, M, -1>(c, [&](int k) { return xb.template ld<V>(vo, CV::soff(k)); }, post, v, lds);
    }
}

typedef float fh_floatx16 __attribute__((ext_vector_type(16)));

// Step 2 on the matrix cores (FP32, gamma symmetric in (i, j)): the Kerr
// contraction dK_q(t) = sum_p Gr[q][p] b_p(t) is a [Q x P] x [P x N] product
// computed with v_mfma_f32_32x32x2f32.  A wave takes 32 of its samples at a
// time with two lanes per sample: lane l supplies the products b_p,
// p = 2r + l/32, of sample l%32 (row l/32 of the 2 x 32 B tile) and the
// reduced gamma Gr[qb 32 + l%32][2r + l/32] (the A tile, held in registers
// for the whole call).  Accumulator r of lane l then holds output row
// 8(r/4) + 4(l/32) + r%4 and column l%32, so each lane stores whole h pairs.
// The samples of a chunk belong to lanes of the same wave; the caller's
// drain_memory() orders their u stores before these loads.
template <typename R, int N, int T, int M>
__device__ __forceinline__ void mm_contract_mfma(const MmCtx<R, N, T, M>& c, int U, int HQ)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    constexpr int P = TR::P;
    constexpr int NP = TR::NP;
    constexpr int KS = (NP + 1) / 2;  // k-steps of two products
    constexpr int QB = (NP + 31) / 32;  // 32-row blocks of output pairs (Q = NP)
    static_assert(T % 64 == 0, "whole waves");
    const int QC = c.QC;
    const float* __restrict__ gr = c.a.gr;
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, kh = l >> 5, j = l & 31;
    float ar[QB][KS];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb)
#pragma unroll
        for (int r = 0; r < KS; ++r) {
            const int q = qb * 32 + j, p = 2 * r + kh;
            ar[qb][r] = (q < NP && p < NP) ? gr[q * NP + p] : 0.0f;
        }
#pragma unroll 1
    for (int kk = 0; kk < P; ++kk) {
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            const uint32_t vs = (uint32_t)(w * 64 + 32 * h + j) * (uint32_t)sizeof(V);
            V u[M];
#pragma unroll
            for (int m = 0; m < M; ++m) u[m] = c.fld(U + m).template ld<V>(vs, CV::soff(kk));
            constexpr InputPairs<M> ip{};
#pragma unroll
            for (int qb = 0; qb < QB; ++qb) {
                fh_floatx16 acc = {0};
#pragma unroll
                for (int r = 0; r < KS; ++r) {
                    // this lane's product: pair 2r + kh (zero past the last pair)
                    const float ve = u[ip.i[2 * r]].x * u[ip.j[2 * r]].x + u[ip.i[2 * r]].y * u[ip.j[2 * r]].y;
                    float vo = 0.0f;
                    if (2 * r + 1 < NP)
                        vo = u[ip.i[2 * r + 1]].x * u[ip.j[2 * r + 1]].x + u[ip.i[2 * r + 1]].y * u[ip.j[2 * r + 1]].y;
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[qb][r], kh ? vo : ve, acc, 0, 0, 0);
                }
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int ii = 0; ii < 4; ii += 2) {
                        const int cc = (qb * 32 + 8 * g + 4 * kh + ii) >> 1;
                        if (cc < QC) c.fld(HQ + cc).st(V{acc[4 * g + ii], acc[4 * g + ii + 1]}, vs, CV::soff(kk));
                    }
            }
        }
    }
}

// Steps 1-3 of the right-hand side at zs: time fields u_m = fft(cis(D_m zs)
// y_m) of the input fields y_m (slice fields IN + m) into U + m, and the
// pair fields h (Kerr contraction plus packed Raman) into HQ + cc.  With TAPE
// the stage's u_m and final h pairs are also stored to tape slot `tslot`
// (tape_u: M fields per slot, tape_h: QC fields per slot).
template <typename R, int N, int T, int M, bool TAPE>
__device__ __forceinline__ void mm_rhs_head(const MmCtx<R, N, T, M>& c, double zs, int IN, int U, int HQ,
                                            const Buf& tape_u, const Buf& tape_h, uint32_t tslot, C2<R>* v,
                                            void* lds)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = TR::P;
    constexpr int NP = TR::NP;
    const MmArgs<R>& a = c.a;
    const uint32_t vo = CV::voff();
    const uint32_t FB = (uint32_t)(N * sizeof(V));
    const int Q = c.Q, QC = c.QC;
    const R* __restrict__ gr = a.gr;
    drain_memory();  // re-reads the input fields written by the previous stage
    // 1. u_m = fft(cis(D_m zs) y_m)
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld y = c.fld(IN + m), u = c.fld(U + m);
        const Fld tu{tape_u, (tslot * (uint32_t)M + (uint32_t)m) * FB};
        mm_xform<R, N, T, M, -1>(
            c,
            [&](int k) {
                R cs, sn;
                cis_arg(c.disp.template ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), zs, cs, sn);
                const V x = y.template ld<V>(vo, CV::soff(k));
                return V{x.x * cs - x.y * sn, x.x * sn + x.y * cs};
            },
            [&](int k, V x) {
                u.st(x, vo, CV::soff(k));
                if constexpr (TAPE) tu.st(x, vo, CV::soff(k));
            },
            v, lds);
    }
    drain_memory();  // re-reads u of step 1
    // 2. Kerr contraction per sample: dK_q = sum_p Gr[q][p] B_p (matrix cores
    // for FP32 with symmetric gamma from M = 6 up, VALU otherwise; at most
    // 512 threads leave room for the A tile of Gr in registers)
    constexpr bool MFMA_OK = sizeof(R) == 4 && T <= 512 && M >= 6;
    if (MFMA_OK && Q != M * M) {
        if constexpr (MFMA_OK) mm_contract_mfma<R, N, T, M>(c, U, HQ);
    } else {
#pragma unroll 1
    for (int k = 0; k < P; ++k) {
        V uk[M];
#pragma unroll
        for (int m = 0; m < M; ++m) uk[m] = c.fld(U + m).template ld<V>(vo, CV::soff(k));
        R b[NP];
        {
            int p = 0;
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = i; j < M; ++j) b[p++] = uk[i].x * uk[j].x + uk[i].y * uk[j].y;
        }
#pragma unroll 1
        for (int cc = 0; cc < QC; ++cc) {
            R d0 = 0, d1 = 0;
            const R* g0 = gr + (size_t)(2 * cc) * NP;
#pragma unroll
            for (int p = 0; p < NP; ++p) d0 += g0[p] * b[p];
            if (2 * cc + 1 < Q) {
                const R* g1 = g0 + NP;
#pragma unroll
                for (int p = 0; p < NP; ++p) d1 += g1[p] * b[p];
            }
            c.fld(HQ + cc).st(V{d0, d1}, vo, CV::soff(k));
        }
    }
    }
    drain_memory();  // re-reads hq of step 2
    // 3. Raman on packed pairs: h = (1-fR) dK + real(ifft(H fft(dK)))
#pragma unroll 1
    for (int cc = 0; cc < QC; ++cc) {
        const Fld hq = c.fld(HQ + cc);
        const Fld th{tape_h, (tslot * (uint32_t)QC + (uint32_t)cc) * FB};
        mm_convolve<R, N, T, M, false>(
            c, [&](int k) { return hq.template ld<V>(vo, CV::soff(k)); },
            [&](int k, V x) {
                const V d = hq.template ld<V>(vo, CV::soff(k));
                const V hv = V{a.one_m_fr * d.x + x.x, a.one_m_fr * d.y + x.y};
                hq.st(hv, vo, CV::soff(k));
                if constexpr (TAPE) th.st(hv, vo, CV::soff(k));
            },
            v, lds);
    }
}

// eta_i = sum_j h_ij u_j for every mode, per sample: u and the h pairs are
// loaded once per sample (M + QC loads instead of 2 M^2) and eta_i overwrites
// u_i in place (u is not read again by the stage; the tape holds its copy).
template <typename R, int N, int T, int M, bool SYM>
__device__ __forceinline__ void mm_eta_body(const MmCtx<R, N, T, M>& c, int U, int HQ)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    constexpr int P = TR::P;
    constexpr int QS = SYM ? (TR::NP + 1) / 2 : TR::QCMAX;
    const uint32_t vo = CV::voff();
#pragma unroll 1
    for (int k = 0; k < P; ++k) {
        V u[M], hp[QS];
#pragma unroll
        for (int m = 0; m < M; ++m) u[m] = c.fld(U + m).template ld<V>(vo, CV::soff(k));
#pragma unroll
        for (int cc = 0; cc < QS; ++cc) hp[cc] = c.fld(HQ + cc).template ld<V>(vo, CV::soff(k));
#pragma unroll
        for (int i = 0; i < M; ++i) {
            V acc = V{(R)0, (R)0};
#pragma unroll
            for (int j = 0; j < M; ++j) {
                const int q = pair_index(M, i, j, SYM);
                const R hij = (q & 1) ? hp[q >> 1].y : hp[q >> 1].x;
                acc = V{acc.x + hij * u[j].x, acc.y + hij * u[j].y};
            }
            c.fld(U + i).st(acc, vo, CV::soff(k));
        }
    }
}

template <typename R, int N, int T, int M>
__device__ __forceinline__ void mm_eta_inplace(const MmCtx<R, N, T, M>& c, int U, int HQ)
{
    if (c.Q != M * M)
        mm_eta_body<R, N, T, M, true>(c, U, HQ);
    else
        mm_eta_body<R, N, T, M, false>(c, U, HQ);
}

// One RK4 stage s of the step from state st at position zs (the right-hand
// side fused with the RK4 bookkeeping of y/st/ac).  Tape as in mm_rhs_head.
template <typename R, int N, int T, int M, bool TAPE>
__device__ __forceinline__ void mm_stage(const MmCtx<R, N, T, M>& c, int s, double zs, double h, const Buf& tape_u,
                                         const Buf& tape_h, uint32_t tslot, C2<R>* v, void* lds)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    using CR = Cyc<R, T>;
    constexpr int P = TR::P;
    const MmArgs<R>& a = c.a;
    const uint32_t vo = CV::voff();
    const R h2 = (R)(h / 2), h1 = (R)h, h6 = (R)(h / 6);
    mm_rhs_head<R, N, T, M, TAPE>(c, zs, 0, 3 * M, 4 * M, tape_u, tape_h, tslot, v, lds);
    drain_memory();  // re-reads hq of step 3
    // 4. eta_i = sum_j h_ij u_j (into the u fields); per mode: k_i; RK4 bookkeeping
    mm_eta_inplace<R, N, T, M>(c, 3 * M, 4 * M);
    drain_memory();  // re-reads eta
#pragma unroll 1
    for (int i = 0; i < M; ++i) {
        const Fld eta = c.fld(3 * M + i);
        const Fld y = c.fld(i), st = c.fld(M + i), ac = c.fld(2 * M + i);
        mm_xform<R, N, T, M, +1>(
            c, [&](int k) { return eta.template ld<V>(vo, CV::soff(k)); },
            [&](int k, V w) {
                R cs, sn;
                cis_arg(c.disp.template ld<double>(CD::voff(), (uint32_t)(i * N * 8) + CD::soff(k)), zs, cs, sn);
                const R sc = c.shn.template ld<R>(CR::voff(), CR::soff(k));
                const V e = V{w.x * sc, w.y * sc};
                const V kk = V{sn * e.x - cs * e.y, sn * e.y + cs * e.x};
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
            },
            v, lds);
    }
}

// Full right-hand side at zs for the adaptive driver: k_i = f_i(zs, y) of the
// input fields IN + i into the output fields OUT + i (U, HQ: scratch fields
// of mm_rhs_head; with TAPE, u and h also go to tape slot `tslot`).
template <typename R, int N, int T, int M, bool TAPE = false>
__device__ __forceinline__ void mm_rhs(const MmCtx<R, N, T, M>& c, double zs, int IN, int OUT, int U, int HQ,
                                       const Buf& tape_u, const Buf& tape_h, uint32_t tslot, C2<R>* v, void* lds)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    using CR = Cyc<R, T>;
    constexpr int P = TR::P;
    const MmArgs<R>& a = c.a;
    const uint32_t vo = CV::voff();
    mm_rhs_head<R, N, T, M, TAPE>(c, zs, IN, U, HQ, tape_u, tape_h, tslot, v, lds);
    drain_memory();  // re-reads hq of step 3
    // 4. eta_i = sum_j h_ij u_j (into the u fields), k_i = i cis(-D_i zs) shock ifft(eta_i)
    mm_eta_inplace<R, N, T, M>(c, U, HQ);
    drain_memory();  // re-reads eta
#pragma unroll 1
    for (int i = 0; i < M; ++i) {
        const Fld eta = c.fld(U + i);
        const Fld out = c.fld(OUT + i);
        mm_xform<R, N, T, M, +1>(
            c, [&](int k) { return eta.template ld<V>(vo, CV::soff(k)); },
            [&](int k, V w) {
                R cs, sn;
                cis_arg(c.disp.template ld<double>(CD::voff(), (uint32_t)(i * N * 8) + CD::soff(k)), zs, cs, sn);
                const R sc = c.shn.template ld<R>(CR::voff(), CR::soff(k));
                const V e = V{w.x * sc, w.y * sc};
                out.st(V{sn * e.x - cs * e.y, sn * e.y + cs * e.x}, vo, CV::soff(k));
            },
            v, lds);
    }
}

template <typename R, int N, int T, int M>
__global__ __launch_bounds__(T) void k_rk4ip_multi(MmArgs<R> a)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = TR::P;
    __shared__ __align__(16) unsigned char lds[TR::LDS_BYTES];
    const size_t sol = blockIdx.x;
    // One descriptor for the whole workgroup slice, fixed for the kernel's
    // lifetime; field f is addressed through the per-lane offset
    // vo + f * N * sizeof(V) (see Fld).
    V* slice = a.work + sol * (size_t)TR::FIELDS * N;
    const Buf sl(slice, (uint32_t)((size_t)TR::FIELDS * N * sizeof(V)));
    const Buf disp(a.disp, M * N * 8), hn(a.hn, N * sizeof(V)), shn(a.shn, N * sizeof(R));
    const Buf tw(a.tw, mm_twiddle_entries<R, N, T>() * sizeof(V));
    const Buf in(a.u0 + sol * M * N, M * N * sizeof(V));
    const MmCtx<R, N, T, M> c{a, sl, disp, hn, shn, tw, a.q, (a.q + 1) / 2, TR::XA};
    const uint32_t vo = CV::voff();

    // y = st = u0
    for (int m = 0; m < M; ++m) {
        const Fld y = c.fld(m), st = c.fld(M + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V x = in.ld<V>(CV::voff(), (uint32_t)(m * N * sizeof(V)) + CV::soff(k));
            y.st(x, vo, CV::soff(k));
            st.st(x, vo, CV::soff(k));
        }
    }
    const double Lb = a.lengths ? a.lengths[sol] : a.length;
    const int Sb = a.nsteps ? a.nsteps[sol] : a.steps;
    const double h = a.lengths ? Lb / Sb : a.h;
    V v[TR::PV];
    const Buf none(a.u0, 0);
    for (int step = 0; step < Sb; ++step) {
        const double z = step * h;
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
            const double zs = s == 0 ? z : (s == 3 ? z + h : z + h / 2);
            mm_stage<R, N, T, M, false>(c, s, zs, h, none, none, 0, v, lds);
        }
    }
    drain_memory();
    int bad = 0;
    const Buf out(a.out + sol * M * N, M * N * sizeof(V));
    for (int m = 0; m < M; ++m) {
        const Fld st = c.fld(M + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            R cs, sn;
            cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), Lb, cs, sn);
            const V x = st.template ld<V>(vo, CV::soff(k));
            const V o = V{x.x * cs - x.y * sn, x.x * sn + x.y * cs};
            bad |= !(isfinite(o.x) && isfinite(o.y));
            out.st(o, CV::voff(), (uint32_t)(m * N * sizeof(V)) + CV::soff(k));
        }
    }
    if (bad) atomicOr(a.status + sol, 1);
}

// Pullback of one right-hand-side evaluation at zs.  The cotangent of stage
// output k_i is cl lam_i + cg gx_i (gx_i: the previous pullback's output, not
// read when cg == 0); on exit gx_i holds the cotangent of the stage input.
// u and h of the stage come from tape slot `tslot`.
template <typename R, int N, int T, int M>
__device__ __forceinline__ void mm_vjp(const MmCtx<R, N, T, M>& c, double zs, R cl, R cg, const Buf& tape_u,
                                       const Buf& tape_h, uint32_t tslot, C2<R>* v, void* lds)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    using CR = Cyc<R, T>;
    constexpr int P = TR::P;
    constexpr int NP = TR::NP;
    constexpr int QMAX = TR::QMAX;
    constexpr int QCMAX = TR::QCMAX;
    const MmArgs<R>& a = c.a;
    const uint32_t vo = CV::voff();
    const uint32_t FB = (uint32_t)(N * sizeof(V));
    const int Q = c.Q, QC = c.QC;
    const bool sym = Q != M * M;  // output pairs i <= j (h_ij = h_ji)
    const R* __restrict__ gr = a.gr;
    auto tu = [&](int m) { return Fld{tape_u, (tslot * (uint32_t)M + (uint32_t)m) * FB}; };
    auto th = [&](int cc) { return Fld{tape_h, (tslot * (uint32_t)QC + (uint32_t)cc) * FB}; };
    drain_memory();  // lam / gx were stored by the previous stage
    // V1: G_i = fft(-i cis(D_i zs) (shock/N) gy_i)
#pragma unroll 1
    for (int i = 0; i < M; ++i) {
        const Fld lam = c.fld(TR::LAM + i), gx = c.fld(TR::GX + i), G = c.fld(TR::GG + i);
        mm_xform<R, N, T, M, -1>(
            c,
            [&](int k) {
                const V l = lam.template ld<V>(vo, CV::soff(k));
                V g = V{cl * l.x, cl * l.y};
                if (cg != (R)0) {
                    const V x = gx.template ld<V>(vo, CV::soff(k));
                    g = V{g.x + cg * x.x, g.y + cg * x.y};
                }
                R cs, sn;
                cis_arg(c.disp.template ld<double>(CD::voff(), (uint32_t)(i * N * 8) + CD::soff(k)), zs, cs, sn);
                const R sc = c.shn.template ld<R>(CR::voff(), CR::soff(k));
                const V e = V{g.x * sc, g.y * sc};
                const V w = V{e.x * cs - e.y * sn, e.x * sn + e.y * cs};
                return V{w.y, -w.x};  // -i w
            },
            [&](int k, V x) { G.st(x, vo, CV::soff(k)); }, v, lds);
    }
    drain_memory();  // re-reads G
    // V2: gu_j = sum_i h_ij G_i; A_q in output-pair order, packed into bq
#pragma unroll 1
    for (int k = 0; k < P; ++k) {
        V G[M], u[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            G[m] = c.fld(TR::GG + m).template ld<V>(vo, CV::soff(k));
            u[m] = tu(m).template ld<V>(vo, CV::soff(k));
        }
        // gu_j = sum_i h_ij G_i with the stage's h pairs loaded once per sample
        auto gu_from = [&](auto symc) {
            constexpr bool SYM = decltype(symc)::value;
            constexpr int QS = SYM ? (NP + 1) / 2 : QCMAX;
            V hq[QS];
#pragma unroll
            for (int cc = 0; cc < QS; ++cc) hq[cc] = th(cc).template ld<V>(vo, CV::soff(k));
#pragma unroll
            for (int j = 0; j < M; ++j) {
                V g = V{(R)0, (R)0};
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    const int q = pair_index(M, i, j, SYM);
                    const R hij = (q & 1) ? hq[q >> 1].y : hq[q >> 1].x;
                    g = V{g.x + hij * G[i].x, g.y + hij * G[i].y};
                }
                c.fld(TR::GU + j).st(g, vo, CV::soff(k));
            }
        };
        if (sym)
            gu_from(std::true_type{});
        else
            gu_from(std::false_type{});
        R A[QMAX + 1];
        A[QMAX] = (R)0;
        if (sym) {
            int q = 0;
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = i; j < M; ++j) {
                    R s = G[i].x * u[j].x + G[i].y * u[j].y;
                    if (j != i) s += G[j].x * u[i].x + G[j].y * u[i].y;
                    A[q++] = s;
                }
#pragma unroll
            for (int r = TR::NP; r < QMAX; ++r) A[r] = (R)0;
        } else {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) A[i * M + j] = G[i].x * u[j].x + G[i].y * u[j].y;
        }
#pragma unroll
        for (int cc = 0; cc < QCMAX; ++cc)
            if (cc < QC) c.fld(TR::BQ + cc).st(V{A[2 * cc], 2 * cc + 1 < Q ? A[2 * cc + 1] : (R)0}, vo, CV::soff(k));
    }
    drain_memory();  // re-reads bq
    // V3: B_q = (1-fR) A_q + R^T(A_q) on packed pairs
#pragma unroll 1
    for (int cc = 0; cc < QC; ++cc) {
        const Fld bq = c.fld(TR::BQ + cc);
        mm_convolve<R, N, T, M, true>(
            c, [&](int k) { return bq.template ld<V>(vo, CV::soff(k)); },
            [&](int k, V x) {
                const V d = bq.template ld<V>(vo, CV::soff(k));
                bq.st(V{a.one_m_fr * d.x + x.x, a.one_m_fr * d.y + x.y}, vo, CV::soff(k));
            },
            v, lds);
    }
    drain_memory();  // re-reads bq and gu
    // V4: C_p = sum_q Gr[q][p] B_q; gu += C_p d(Re u_k conj u_l)/du
#pragma unroll 1
    for (int k = 0; k < P; ++k) {
        R C[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) C[p] = (R)0;
#pragma unroll 1
        for (int cc = 0; cc < QC; ++cc) {
            const V b = c.fld(TR::BQ + cc).template ld<V>(vo, CV::soff(k));
            const R* g0 = gr + (size_t)(2 * cc) * NP;
#pragma unroll
            for (int p = 0; p < NP; ++p) C[p] += g0[p] * b.x;
            if (2 * cc + 1 < Q) {
                const R* g1 = g0 + NP;
#pragma unroll
                for (int p = 0; p < NP; ++p) C[p] += g1[p] * b.y;
            }
        }
        V u[M], gu[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            u[m] = tu(m).template ld<V>(vo, CV::soff(k));
            gu[m] = c.fld(TR::GU + m).template ld<V>(vo, CV::soff(k));
        }
        int p = 0;
#pragma unroll
        for (int kk = 0; kk < M; ++kk)
#pragma unroll
            for (int ll = kk; ll < M; ++ll) {
                const R cp = C[p++];
                if (kk == ll) {
                    gu[kk] = V{gu[kk].x + 2 * cp * u[kk].x, gu[kk].y + 2 * cp * u[kk].y};
                } else {
                    gu[kk] = V{gu[kk].x + cp * u[ll].x, gu[kk].y + cp * u[ll].y};
                    gu[ll] = V{gu[ll].x + cp * u[kk].x, gu[ll].y + cp * u[kk].y};
                }
            }
#pragma unroll
        for (int m = 0; m < M; ++m) c.fld(TR::GU + m).st(gu[m], vo, CV::soff(k));
    }
    drain_memory();  // re-reads gu
    // V5: gx_m = conj(cis(D_m zs)) N ifft(gu_m)
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
        const Fld gu = c.fld(TR::GU + m), gx = c.fld(TR::GX + m);
        mm_xform<R, N, T, M, +1>(
            c, [&](int k) { return gu.template ld<V>(vo, CV::soff(k)); },
            [&](int k, V w) {
                R cs, sn;
                cis_arg(c.disp.template ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), zs, cs, sn);
                gx.st(V{cs * w.x + sn * w.y, cs * w.y - sn * w.x}, vo, CV::soff(k));
            },
            v, lds);
    }
}

// Workgroup FP64 sum (fixed tree order) through the LDS buffer.
template <int T> __device__ __forceinline__ double mm_block_sum(double x, void* lds_raw)
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

// Forward with checkpoints, terminal seed (band energy summed over modes,
// mmf_cost_sum of mmf_cost.jl:49-62, or a caller seed), backward sweep of the
// discrete adjoint with per-segment recomputation, lam0 and the shared-phase
// gradient sum_m 2 Re(conj(lam0_m) i u0_m) in FP64
// (mmf_raman_optimization.jl:268-271).
template <typename R, int N, int T, int M>
__global__ __launch_bounds__(T) void k_grad_multi(MmArgs<R> a)
{
    using V = C2<R>;
    using TR = MmTraits<R, N, T, M>;
    using CV = Cyc<V, T>;
    using CD = Cyc<double, T>;
    constexpr int P = TR::P;
    __shared__ __align__(16) unsigned char lds[TR::LDS_BYTES];
    const size_t sol = blockIdx.x;
    const uint32_t FB = (uint32_t)(N * sizeof(V));
    V* slice = a.work + sol * (size_t)TR::FIELDS_GRAD * N;
    const Buf sl(slice, (uint32_t)((size_t)TR::FIELDS_GRAD * N * sizeof(V)));
    const Buf disp(a.disp, M * N * 8), hn(a.hn, N * sizeof(V)), shn(a.shn, N * sizeof(R));
    const Buf tw(a.tw, mm_twiddle_entries<R, N, T>() * sizeof(V));
    const Buf in(a.u0 + sol * M * N, M * N * sizeof(V));
    const int Q = a.q, QC = (a.q + 1) / 2;
    const MmCtx<R, N, T, M> c{a, sl, disp, hn, shn, tw, Q, QC, TR::XA};
    const Buf ckpt(a.ckpt + sol * (size_t)a.nseg * M * N, (uint32_t)((size_t)a.nseg * M * N * sizeof(V)));
    // stage tape: one region of 4*every slots per kept segment plus the shared
    // region of the recomputed segments (when the plan recomputes any)
    const size_t tslots = (size_t)(a.ntaped + (a.ntaped < a.nseg ? 1 : 0)) * 4 * a.every;
    const Buf tape_u(a.tape_u + sol * tslots * M * N, (uint32_t)(tslots * M * N * sizeof(V)));
    const Buf tape_h(a.tape_h + sol * tslots * QC * N, (uint32_t)(tslots * QC * N * sizeof(V)));
    const Buf none(a.u0, 0);
    const uint32_t vo = CV::voff();
    // per-solve length and step count (the plan's steps bound the checkpoint capacity)
    const double Lb = a.lengths ? a.lengths[sol] : a.length;
    const int Sb = a.nsteps ? a.nsteps[sol] : a.steps;
    const double h = a.lengths ? Lb / Sb : a.h;
    const int nseg = (Sb + a.every - 1) / a.every;
    // the last ntaped segments of this solve keep the forward sweep's tape
    const int first_kept = nseg - min(a.ntaped, nseg);
    auto seg_base = [&](int sg) { return (uint32_t)((sg >= first_kept ? sg - first_kept : a.ntaped) * 4 * a.every); };
    V v[TR::PV];
    int bad = 0;
    double etot = 0.0, eband = 0.0;
    if (a.phase & 1) {
        for (int m = 0; m < M; ++m) {
            const Fld y = c.fld(m), st = c.fld(M + m);
#pragma unroll
            for (int k = 0; k < P; ++k) {
                const V x = in.ld<V>(CV::voff(), (uint32_t)(m * N * sizeof(V)) + CV::soff(k));
                y.st(x, vo, CV::soff(k));
                st.st(x, vo, CV::soff(k));
            }
        }
        // forward with checkpoints of the state; kept segments tape their stages
        const int kept_from = min(first_kept * a.every, Sb);
        for (int step = 0; step < kept_from; ++step) {
            if (step % a.every == 0) {
                drain_memory();
                for (int m = 0; m < M; ++m) {
                    const Fld st = c.fld(M + m);
                    const Fld ck{ckpt, ((uint32_t)(step / a.every) * (uint32_t)M + (uint32_t)m) * FB};
#pragma unroll
                    for (int k = 0; k < P; ++k) ck.st(st.template ld<V>(vo, CV::soff(k)), vo, CV::soff(k));
                }
            }
            const double z = step * h;
#pragma unroll 1
            for (int s = 0; s < 4; ++s)
                mm_stage<R, N, T, M, false>(c, s, s == 0 ? z : (s == 3 ? z + h : z + h / 2), h, none, none, 0, v, lds);
        }
        for (int step = kept_from; step < Sb; ++step) {
            const double z = step * h;
            const int sg = step / a.every;
#pragma unroll 1
            for (int s = 0; s < 4; ++s)
                mm_stage<R, N, T, M, true>(c, s, s == 0 ? z : (s == 3 ? z + h : z + h / 2), h, tape_u, tape_h,
                                           seg_base(sg) + (uint32_t)(4 * (step - sg * a.every) + s), v, lds);
        }
        drain_memory();
        // output and terminal seed (band energy over all modes)
        const Buf out(a.out + sol * M * N, M * N * sizeof(V));
        for (int m = 0; m < M; ++m) {
            const Fld st = c.fld(M + m), lam = c.fld(TR::LAM + m);
#pragma unroll
            for (int k = 0; k < P; ++k) {
                R cs, sn;
                cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), Lb, cs, sn);
                const V x = st.template ld<V>(vo, CV::soff(k));
                const V o = V{x.x * cs - x.y * sn, x.x * sn + x.y * cs};
                bad |= !(isfinite(o.x) && isfinite(o.y));
                out.st(o, CV::voff(), (uint32_t)(m * N * sizeof(V)) + CV::soff(k));
                lam.st(o, vo, CV::soff(k));  // lab-frame output, scaled into the seed below
                const double e = (double)o.x * o.x + (double)o.y * o.y;
                etot += e;
                if (!a.seed && a.mask[threadIdx.x + T * k]) eband += e;
            }
        }
        if (!(a.phase & 2)) {  // forward sweep only: checkpoints and kept tape stay for the backward launch
            if (bad) atomicOr(a.status + sol, 1);
            return;
        }
    }
    double J = 0.0, scale = 0.0;
    if (!a.seed) {
        etot = mm_block_sum<T>(etot, lds);
        eband = mm_block_sum<T>(eband, lds);
        J = eband / etot;
        scale = 1.0 / etot;
        if (a.log_cost) scale *= J <= 1e-15 ? 0.0 : 10.0 / (2.302585092994045684 * J);
        if (threadIdx.x == 0) a.cost[sol] = a.log_cost ? 10.0 * log10(fmax(J, 1e-15)) : J;
    }
    drain_memory();
    // lam = cis(-D L) seed
    const Buf sd(a.seed ? a.seed + sol * M * N : a.u0, a.seed ? (uint32_t)(M * N * sizeof(V)) : 0u);
    for (int m = 0; m < M; ++m) {
        const Fld lam = c.fld(TR::LAM + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            V s;
            if (a.seed) {
                s = sd.ld<V>(CV::voff(), (uint32_t)(m * N * sizeof(V)) + CV::soff(k));
            } else {
                const V o = lam.template ld<V>(vo, CV::soff(k));
                const R f = (R)(((a.mask[threadIdx.x + T * k] ? 1.0 : 0.0) - J) * scale);
                s = V{o.x * f, o.y * f};
            }
            R cs, sn;
            cis_arg(disp.ld<double>(CD::voff(), (uint32_t)(m * N * 8) + CD::soff(k)), Lb, cs, sn);
            lam.st(V{cs * s.x + sn * s.y, cs * s.y - sn * s.x}, vo, CV::soff(k));
        }
    }
    const R c6 = (R)(h / 6), c3 = (R)(h / 3), c2 = (R)(h / 2), c1 = (R)h;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * a.every;
        const int s1 = s0 + a.every < Sb ? s0 + a.every : Sb;
        const uint32_t base = seg_base(sg);
        drain_memory();
        if (sg < first_kept) {  // recompute the segment's tape from its checkpoint
            for (int m = 0; m < M; ++m) {
                const Fld ck{ckpt, ((uint32_t)sg * (uint32_t)M + (uint32_t)m) * FB};
                const Fld y = c.fld(m), st = c.fld(M + m);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    const V x = ck.template ld<V>(vo, CV::soff(k));
                    y.st(x, vo, CV::soff(k));
                    st.st(x, vo, CV::soff(k));
                }
            }
            for (int step = s0; step < s1; ++step) {
                const double z = step * h;
#pragma unroll 1
                for (int s = 0; s < 4; ++s)
                    mm_stage<R, N, T, M, true>(c, s, s == 0 ? z : (s == 3 ? z + h : z + h / 2), h, tape_u, tape_h,
                                               base + (uint32_t)(4 * (step - s0) + s), v, lds);
            }
        }
        for (int step = s1 - 1; step >= s0; --step) {
            const double z = step * h;
#pragma unroll 1
            for (int s = 3; s >= 0; --s) {
                const double zs = s == 0 ? z : (s == 3 ? z + h : z + h / 2);
                const R cl = (s == 0 || s == 3) ? c6 : c3;
                const R cg = s == 3 ? (R)0 : (s == 2 ? c1 : c2);
                mm_vjp<R, N, T, M>(c, zs, cl, cg, tape_u, tape_h, base + (uint32_t)(4 * (step - s0) + s), v, lds);
                drain_memory();
                for (int m = 0; m < M; ++m) {
                    const Fld lam = c.fld(TR::LAM + m), lacc = c.fld(TR::LACC + m), gx = c.fld(TR::GX + m);
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const V g = gx.template ld<V>(vo, CV::soff(k));
                        const V base = s == 3 ? lam.template ld<V>(vo, CV::soff(k)) : lacc.template ld<V>(vo, CV::soff(k));
                        const V nc = V{base.x + g.x, base.y + g.y};
                        if (s == 0)
                            lam.st(nc, vo, CV::soff(k));
                        else
                            lacc.st(nc, vo, CV::soff(k));
                    }
                }
            }
        }
    }
    drain_memory();
    // lam0 and the shared-phase gradient in FP64
    const Buf l0(a.lam0 + sol * M * N, M * N * sizeof(V));
    const Buf gb(a.grad + sol * N, N * sizeof(double));
    double g[P];
#pragma unroll
    for (int k = 0; k < P; ++k) g[k] = 0.0;
    for (int m = 0; m < M; ++m) {
        const Fld lam = c.fld(TR::LAM + m);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const V l = lam.template ld<V>(vo, CV::soff(k));
            const V u = in.ld<V>(CV::voff(), (uint32_t)(m * N * sizeof(V)) + CV::soff(k));
            l0.st(l, CV::voff(), (uint32_t)(m * N * sizeof(V)) + CV::soff(k));
            g[k] += 2.0 * ((double)l.y * (double)u.x - (double)l.x * (double)u.y);
            bad |= !(isfinite(l.x) && isfinite(l.y));
        }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) gb.st(g[k], CD::voff(), CD::soff(k));
    if (bad) atomicOr(a.status + sol, 1);
}

}  // namespace fh
