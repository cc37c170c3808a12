// fh_device.hpp -- device building blocks for libfiberhip (gfx950 / CDNA4).
//
// * complex helpers on double2 / float2
// * small register DFTs (radix 2, 4, 8, 16) with compile-time twiddles
// * a block-wide Stockham FFT over LDS.  A field of N samples is held by a
//   workgroup of T threads in a *cyclic* layout: thread t owns samples
//   t, t+T, t+2T, ... (P = N/T of them, in registers).  Every Stockham pass
//   reads its butterfly inputs from that same cyclic set, and the last pass
//   writes its outputs back into it, so a transform needs one LDS exchange
//   per pass boundary and none before the first or after the last pass.
//   Global loads/stores in this layout are fully coalesced.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

template <typename R> struct C2T;
template <> struct C2T<double> { using type = double2; };
template <> struct C2T<float> { using type = float2; };
template <typename R> using C2 = typename C2T<R>::type;

template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return V{a.x - b.x, a.y - b.y}; }
template <typename V> __device__ __forceinline__ V cmul(V a, V b)
{
    return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename V> __device__ __forceinline__ V cconj(V a) { return V{a.x, -a.y}; }
template <typename V, typename R> __device__ __forceinline__ V cscale(V a, R s) { return V{a.x * s, a.y * s}; }
// a * (-i) for DIR=-1 (forward), a * (+i) for DIR=+1 (inverse)
template <int DIR, typename V> __device__ __forceinline__ V mul_i(V a)
{
    if constexpr (DIR < 0)
        return V{a.y, -a.x};
    else
        return V{-a.y, a.x};
}

// cos(2 pi m / 16) and sin(2 pi m / 16), m = 0..15
// cos and sin of 2 pi m / 32 (radix <= 32 butterfly constants)
__device__ constexpr double kCos32[32] = {
    1.0, 0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
    0.70710678118654752440, 0.55557023301960222474, 0.38268343236508977173, 0.19509032201612826785,
    0.0, -0.19509032201612826785, -0.38268343236508977173, -0.55557023301960222474,
    -0.70710678118654752440, -0.83146961230254523708, -0.92387953251128675613, -0.98078528040323044913,
    -1.0, -0.98078528040323044913, -0.92387953251128675613, -0.83146961230254523708,
    -0.70710678118654752440, -0.55557023301960222474, -0.38268343236508977173, -0.19509032201612826785,
    0.0, 0.19509032201612826785, 0.38268343236508977173, 0.55557023301960222474,
    0.70710678118654752440, 0.83146961230254523708, 0.92387953251128675613, 0.98078528040323044913};
__device__ constexpr double kSin32[32] = {
    0.0, 0.19509032201612826785, 0.38268343236508977173, 0.55557023301960222474,
    0.70710678118654752440, 0.83146961230254523708, 0.92387953251128675613, 0.98078528040323044913,
    1.0, 0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
    0.70710678118654752440, 0.55557023301960222474, 0.38268343236508977173, 0.19509032201612826785,
    0.0, -0.19509032201612826785, -0.38268343236508977173, -0.55557023301960222474,
    -0.70710678118654752440, -0.83146961230254523708, -0.92387953251128675613, -0.98078528040323044913,
    -1.0, -0.98078528040323044913, -0.92387953251128675613, -0.83146961230254523708,
    -0.70710678118654752440, -0.55557023301960222474, -0.38268343236508977173, -0.19509032201612826785};

// multiply by w_R^m = exp(DIR * 2 pi i m / R), R | 16
template <int R, int DIR, int M, typename V> __device__ __forceinline__ V twiddle_const(V a)
{
    static_assert(R <= 32 && 32 % R == 0, "butterfly constants cover radices up to 32");
    constexpr int m = (M % R) * (32 / R);
    if constexpr (m == 0) {
        return a;
    } else if constexpr (m == 16) {
        return V{-a.x, -a.y};
    } else if constexpr (m == 8) {
        return mul_i<DIR>(a);
    } else if constexpr (m == 24) {
        return mul_i<-DIR>(a);
    } else {
        using Rt = decltype(a.x);
        const Rt c = (Rt)kCos32[m];
        const Rt s = (Rt)(DIR < 0 ? -kSin32[m] : kSin32[m]);
        return V{a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

// In-place DFT of R values held in v[0], v[S], v[2S], ... (stride S in the
// register array).  Output k lands in slot k (natural order).
template <int R, int DIR, int S, typename V> struct Dft;

template <int DIR, int S, typename V> struct Dft<1, DIR, S, V> {
    static __device__ __forceinline__ void run(V*) {}
};

template <int DIR, int S, typename V> struct Dft<2, DIR, S, V> {
    static __device__ __forceinline__ void run(V* v)
    {
        V a = v[0], b = v[S];
        v[0] = cadd(a, b);
        v[S] = csub(a, b);
    }
};

template <int DIR, int S, typename V> struct Dft<4, DIR, S, V> {
    static __device__ __forceinline__ void run(V* v)
    {
        V t0 = cadd(v[0], v[2 * S]), t1 = csub(v[0], v[2 * S]);
        V t2 = cadd(v[S], v[3 * S]), t3 = mul_i<DIR>(csub(v[S], v[3 * S]));
        v[0] = cadd(t0, t2);
        v[2 * S] = csub(t0, t2);
        v[S] = cadd(t1, t3);
        v[3 * S] = csub(t1, t3);
    }
};

// R = R1 * R2 decomposition (decimation in time): input n = R2*n1 + n2,
// output k = k1 + R1*k2.
template <int R1, int R2, int DIR, int S, typename V> struct DftSplit {
    static __device__ __forceinline__ void run(V* v)
    {
        constexpr int R = R1 * R2;
        V y[R];
#pragma unroll
        for (int n = 0; n < R; ++n) y[n] = v[n * S];
        // step 1: DFT_R1 over n1 for each n2 (elements y[R2*n1 + n2], stride R2)
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) Dft<R1, DIR, R2, V>::run(y + n2);
        // now y[R2*k1 + n2] = Y[k1][n2]; step 2: twiddle w_R^(n2*k1)
        TwApply<0>::run(y);
        // step 3: DFT_R2 over n2 for each k1 (elements y[R2*k1 + n2], stride 1)
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) Dft<R2, DIR, 1, V>::run(y + R2 * k1);
        // y[R2*k1 + k2] = X[k1 + R1*k2]
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1)
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) v[(k1 + R1 * k2) * S] = y[R2 * k1 + k2];
    }
    template <int I> struct TwApply {
        static __device__ __forceinline__ void run(V* y)
        {
            if constexpr (I < R1 * R2) {
                constexpr int k1 = I / R2, n2 = I % R2;
                y[I] = twiddle_const<R1 * R2, DIR, k1 * n2>(y[I]);
                TwApply<I + 1>::run(y);
            }
        }
    };
};

template <int DIR, int S, typename V> struct Dft<8, DIR, S, V> : DftSplit<2, 4, DIR, S, V> {};
template <int DIR, int S, typename V> struct Dft<16, DIR, S, V> : DftSplit<4, 4, DIR, S, V> {};
template <int DIR, int S, typename V> struct Dft<32, DIR, S, V> : DftSplit<2, 16, DIR, S, V> {};

// ---------------------------------------------------------------- plans

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// Radix schedule: log2(N) = r0 + p*(npass-1) where p = log2(P); the first
// pass takes the remainder radix (it has NS = 1 and needs no twiddles).
template <int N, int T> struct FftPlan {
    static constexpr int P = N / T;
    static constexpr int LOGN = ilog2c(N);
    static constexpr int LOGP = ilog2c(P);
    static constexpr int NPASS = (LOGN + LOGP - 1) / LOGP;
    static constexpr int LOG_R0 = LOGN - LOGP * (NPASS - 1);
    static __host__ __device__ constexpr int log_radix(int pass) { return pass == 0 ? LOG_R0 : LOGP; }
    static __host__ __device__ constexpr int log_ns(int pass) { return pass == 0 ? 0 : LOG_R0 + LOGP * (pass - 1); }
    // offset (in complex entries) of the twiddle table of a pass
    static __host__ __device__ constexpr int tw_offset(int pass)
    {
        int off = 0;
        for (int q = 0; q < pass; ++q) off += ((1 << log_radix(q)) - 1) * (1 << log_ns(q));
        return off;
    }
    static constexpr int TW_SIZE = tw_offset(NPASS);
    static_assert((1 << LOGN) == N && (1 << LOGP) == P, "power-of-two sizes only");
    static_assert(P >= 2 && P <= 32, "2..32 points per thread");
    static_assert(LOG_R0 >= 1, "radix schedule");
};

// LDS padding: one extra slot every 16 entries breaks the stride-16 write
// pattern of the first Stockham passes into distinct banks.
__host__ __device__ constexpr int lds_pad(int i) { return i + (i >> 4); }
template <int N> struct LdsSize { static constexpr int value = N + N / 16; };

// Ordering point: wait until every memory operation this wave has issued
// (vector, scratch, LDS, scalar) has completed.  Kernels place one at each
// LDS exchange and wherever a lane re-reads global scratch it has just
// stored, before any transform separates the two.  Without these points the
// FP64 adjoint kernels returned results that differed between identical
// solves of one launch (relative 1e-7 in lam0, varying from run to run),
// although a dataflow check of the compiler's s_waitcnt placement finds no
// missing wait; compiling with every wait forced to zero also removes the
// difference.  Measured cost: within run-to-run noise.  The wait intrinsic
// orders the hardware only; the empty asm with a memory clobber also keeps
// the compiler from moving loads and stores across the point.
__device__ __forceinline__ void drain_memory()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0);
    asm volatile("" ::: "memory");
}

// Largest complex (interleaved) exchange buffer a kernel uses before it
// switches to the planar exchange.  FP64 kernels stay at or below 128 KiB:
// the 1024-thread FP64 kernels with a 139264-byte interleaved buffer
// (8192 points) gave run-to-run differences in the 1e-8..1e-6 range between
// identical solves, which the planar layout (69632 bytes) does not show.
template <typename R> struct LdsBudget { static constexpr size_t bytes = sizeof(R) == 8 ? 128 * 1024 : 160 * 1024; };

// Raw buffer access with a per-lane byte offset (VGPR) and a uniform byte
// offset (SGPR / immediate).  Keeping the k*T part of a cyclic index in the
// uniform offset means a whole cyclic set needs one address VGPR instead of
// one 64-bit pointer per element.  Out-of-range loads return zero.
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Buf(const void* p, uint32_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000))
    {
    }
    template <typename V, int AUX = 0> __device__ __forceinline__ V ld(uint32_t voff, uint32_t soff) const
    {
        if constexpr (sizeof(V) == 16)
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, AUX));
        else if constexpr (sizeof(V) == 8)
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
        else
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
    }
    template <typename V> __device__ __forceinline__ void st(V v, uint32_t voff, uint32_t soff) const
    {
        if constexpr (sizeof(V) == 16) {
            using W = __attribute__((ext_vector_type(4))) unsigned int;
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(W, v), r, voff, soff, 0);
        } else if constexpr (sizeof(V) == 8) {
            using W = __attribute__((ext_vector_type(2))) unsigned int;
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(W, v), r, voff, soff, 0);
        } else {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
        }
    }
};

// A field inside a larger buffer: the per-lane byte offset is shifted by
// `off`.  Kernels keep one descriptor per buffer region for their whole
// lifetime and address sub-arrays this way: rebuilding descriptors per field
// with a runtime index lets the compiler recycle the descriptor SGPRs right
// after a load is issued, which made solves lose determinism.
struct Fld {
    const Buf& b;
    uint32_t off;
    template <typename W> __device__ __forceinline__ W ld(uint32_t v, uint32_t so) const
    {
        return b.template ld<W>(v + off, so);
    }
    template <typename W> __device__ __forceinline__ void st(W x, uint32_t v, uint32_t so) const
    {
        b.st(x, v + off, so);
    }
};

// Cyclic-set element k of this thread: byte offsets into a [N] array of V.
template <typename V, int T> struct Cyc {
    static __device__ __forceinline__ uint32_t voff() { return threadIdx.x * (uint32_t)sizeof(V); }
    static constexpr uint32_t soff(int k) { return (uint32_t)(k * T * sizeof(V)); }
};

// Exchange a cyclic register set through LDS.  Butterfly b of this thread
// wrote Stockham outputs r to slots base_b + r*NS; afterwards the thread
// reads back its cyclic set tid + T*k.  PLANAR exchanges the real and
// imaginary halves one after the other through a real-valued buffer (used
// when a complex buffer would not fit in LDS).
template <typename R, int N, int T, int RAD, int NS, bool PLANAR> struct Exchange {
    using V = C2<R>;
    static constexpr int P = N / T;
    static constexpr int B = P / RAD;
    static_assert(T % 16 == 0, "workgroup size must be a multiple of 16");
    static constexpr int RSTRIDE = T + T / 16;  // padded distance between k and k+1
    static constexpr int woff(int r) { return r * NS + ((NS * RAD >= 16) ? ((r * NS) >> 4) : 0); }

    static __device__ __forceinline__ void run(V* v, const int* base, void* lds_raw)
    {
        const int tid = threadIdx.x;
        drain_memory();
        if constexpr (!PLANAR) {
            V* lds = reinterpret_cast<V*>(lds_raw);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                V* pb = lds + lds_pad(base[b]);
#pragma unroll
                for (int r = 0; r < RAD; ++r) pb[woff(r)] = v[b + r * B];
            }
            __syncthreads();
            const V* rb = lds + lds_pad(tid);
#pragma unroll
            for (int k = 0; k < P; ++k) v[k] = rb[k * RSTRIDE];
            __syncthreads();
        } else {
            R* lds = reinterpret_cast<R*>(lds_raw);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                R* pb = lds + lds_pad(base[b]);
#pragma unroll
                for (int r = 0; r < RAD; ++r) pb[woff(r)] = v[b + r * B].x;
            }
            __syncthreads();
            const R* rb = lds + lds_pad(tid);
#pragma unroll
            for (int k = 0; k < P; ++k) v[k].x = rb[k * RSTRIDE];
            __syncthreads();
#pragma unroll
            for (int b = 0; b < B; ++b) {
                R* pb = lds + lds_pad(base[b]);
#pragma unroll
                for (int r = 0; r < RAD; ++r) pb[woff(r)] = v[b + r * B].y;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < P; ++k) v[k].y = rb[k * RSTRIDE];
            __syncthreads();
        }
    }
};

// One Stockham pass (radix RAD = 2^LR, NS = 2^LNS).  Butterfly b of this
// thread is j = tid + T*b and owns registers v[b + r*B], r < RAD.  Input r is
// scaled by w_N^(r * (j mod NS) * N/(NS*RAD)) (table slot (r-1)*NS + j mod NS),
// transformed, and output r goes to slot expand(j) + r*NS.
template <typename R, int N, int T, int DIR, int PASS, bool PLANAR> struct StockhamPass {
    using V = C2<R>;
    using PL = FftPlan<N, T>;
    static constexpr int P = PL::P;
    static constexpr int LR = PL::log_radix(PASS);
    static constexpr int RAD = 1 << LR;
    static constexpr int LNS = PL::log_ns(PASS);
    static constexpr int NS = 1 << LNS;
    static constexpr int B = P / RAD;
    static constexpr bool LAST = PASS == PL::NPASS - 1;
    static constexpr bool CHAIN = sizeof(R) == 8;

    static __device__ __forceinline__ void run(V* v, const Buf& tw, void* lds)
    {
        const int tid = threadIdx.x;
        int base[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = tid + T * b;
            const int jm = j & (NS - 1);
            if constexpr (NS > 1) {
                const uint32_t vo = (uint32_t)jm * sizeof(V);
                if constexpr (CHAIN) {
                    // FP64: w^r by repeated multiplication from one table entry
                    // (error grows to ~RAD ulp, far below the FP64 budget) keeps
                    // one twiddle live instead of RAD-1.
                    V w1 = tw.ld<V>(vo, (uint32_t)(PL::tw_offset(PASS) * sizeof(V)));
                    if constexpr (DIR > 0) w1 = cconj(w1);
                    V w = w1;
#pragma unroll
                    for (int r = 1; r < RAD; ++r) {
                        v[b + r * B] = cmul(v[b + r * B], w);
                        if (r + 1 < RAD) w = cmul(w, w1);
                    }
                } else {
                    // FP32: with r = Q*qa + qb, w^r = w^(Q qa) w^qb.  Loading
                    // w^1..w^(Q-1) and w^Q, w^2Q, ... (6 loads for radix 16
                    // instead of 15) adds one rounding to the mixed exponents;
                    // vector-memory issue, not arithmetic, bounds these passes.
                    constexpr int Q = RAD >= 16 ? 4 : RAD;
                    constexpr int A = RAD / Q;
                    V wb[Q], wa[A];
#pragma unroll
                    for (int q = 1; q < Q; ++q) {
                        wb[q] = tw.ld<V>(vo, (uint32_t)((PL::tw_offset(PASS) + (q - 1) * NS) * sizeof(V)));
                        if constexpr (DIR > 0) wb[q] = cconj(wb[q]);
                    }
#pragma unroll
                    for (int q = 1; q < A; ++q) {
                        wa[q] = tw.ld<V>(vo, (uint32_t)((PL::tw_offset(PASS) + (Q * q - 1) * NS) * sizeof(V)));
                        if constexpr (DIR > 0) wa[q] = cconj(wa[q]);
                    }
#pragma unroll
                    for (int r = 1; r < RAD; ++r) {
                        const int qa = r / Q, qb = r % Q;
                        const V w = qa == 0 ? wb[qb] : (qb == 0 ? wa[qa] : cmul(wa[qa], wb[qb]));
                        v[b + r * B] = cmul(v[b + r * B], w);
                    }
                }
            }
            Dft<RAD, DIR, B, V>::run(v + b);
            base[b] = ((j >> LNS) << (LNS + LR)) + jm;
        }
        if constexpr (!LAST) Exchange<R, N, T, RAD, NS, PLANAR>::run(v, base, lds);
        (void)lds;
        (void)base;
    }
};

template <typename R, int N, int T, int DIR, bool PLANAR, int PASS = 0> struct BlockFft {
    static __device__ __forceinline__ void run(C2<R>* v, const Buf& tw, void* lds)
    {
        if constexpr (PASS < FftPlan<N, T>::NPASS) {
            StockhamPass<R, N, T, DIR, PASS, PLANAR>::run(v, tw, lds);
            BlockFft<R, N, T, DIR, PLANAR, PASS + 1>::run(v, tw, lds);
        }
    }
};

// Unnormalised forward (DIR=-1, kernel exp(-2 pi i k n / N)) or inverse
// (DIR=+1) transform of the block's cyclic register set.
template <typename R, int N, int T, int DIR, bool PLANAR>
__device__ __forceinline__ void block_fft(C2<R>* v, const Buf& tw, void* lds)
{
    BlockFft<R, N, T, DIR, PLANAR>::run(v, tw, lds);
}

// ------------------------------------------------------------- phases
//
// cis(D z) for the dispersion rotation.  Arguments reach |D z| ~ 1e5 rad, so
// the reduction is done in FP64 for both precisions: Cody-Waite by pi/2 with
// a three-part constant (the first two parts carry 33 significant bits, so
// q*P1 and q*P2 are exact for |q| < 2^20; plans reject |D z| beyond
// FH_MAX_PHASE).  sin/cos of the reduced argument |r| <= pi/4 use their
// Taylor series: to r^15 / r^16 in FP64 (truncation < 7e-17) and to r^9 /
// r^10 in FP32 (truncation < 2e-9).

#define FH_MAX_PHASE 8.0e5

__host__ __device__ __forceinline__ double reduce_pio2(double x, int& quadrant)
{
    const double q = rint(x * 0.63661977236758134308);
    double r = fma(-q, 1.5707963267341256, x);
    r = fma(-q, 6.077100506303966e-11, r);
    r = fma(-q, 2.0222662487959506e-21, r);
    quadrant = (int)(long long)q & 3;
    return r;
}

template <typename R> __host__ __device__ __forceinline__ void quadrant_select(int q, R sr, R cr, R& c, R& s)
{
    const R s0 = (q & 1) ? cr : sr;
    const R c0 = (q & 1) ? sr : cr;
    s = (q & 2) ? -s0 : s0;
    c = ((q + 1) & 2) ? -c0 : c0;
}

__host__ __device__ __forceinline__ void cis_arg(double d, double z, double& c, double& s)
{
    int q;
    const double r = reduce_pio2(d * z, q);
    const double r2 = r * r;
    double ts = -1.0 / 1307674368000.0;
    ts = fma(ts, r2, 1.0 / 6227020800.0);
    ts = fma(ts, r2, -1.0 / 39916800.0);
    ts = fma(ts, r2, 1.0 / 362880.0);
    ts = fma(ts, r2, -1.0 / 5040.0);
    ts = fma(ts, r2, 1.0 / 120.0);
    ts = fma(ts, r2, -1.0 / 6.0);
    const double sr = fma(r * r2, ts, r);
    double tc = 1.0 / 20922789888000.0;
    tc = fma(tc, r2, -1.0 / 87178291200.0);
    tc = fma(tc, r2, 1.0 / 479001600.0);
    tc = fma(tc, r2, -1.0 / 3628800.0);
    tc = fma(tc, r2, 1.0 / 40320.0);
    tc = fma(tc, r2, -1.0 / 720.0);
    tc = fma(tc, r2, 1.0 / 24.0);
    const double cr = fma(r2 * r2, tc, fma(-0.5, r2, 1.0));
    quadrant_select<double>(q, sr, cr, c, s);
}

__host__ __device__ __forceinline__ void cis_arg(double d, double z, float& c, float& s)
{
    int q;
    const float r = (float)reduce_pio2(d * z, q);
    const float r2 = r * r;
    float ts = 1.0f / 362880.0f;
    ts = fmaf(ts, r2, -1.0f / 5040.0f);
    ts = fmaf(ts, r2, 1.0f / 120.0f);
    ts = fmaf(ts, r2, -1.0f / 6.0f);
    const float sr = fmaf(r * r2, ts, r);
    float tc = -1.0f / 3628800.0f;
    tc = fmaf(tc, r2, 1.0f / 40320.0f);
    tc = fmaf(tc, r2, -1.0f / 720.0f);
    tc = fmaf(tc, r2, 1.0f / 24.0f);
    const float cr = fmaf(r2 * r2, tc, fmaf(-0.5f, r2, 1.0f));
    quadrant_select<float>(q, sr, cr, c, s);
}

}  // namespace fh
