// fh_grin.hip -- GRIN modal setup on the device, FP64: the guided modes of
// the scalar wave equation on a finite-difference grid
// (solve_for_fiber_modes, src/simulation/fibers.jl:159-197) and the overlap
// tensor of sampled modes (compute_overlap_tensor, fibers.jl:214-222).
//
// Modes: the nmodes largest eigenpairs of the symmetric operator
//   A = k0^2 diag(eps) + (5-point Laplacian) / dx^2, zero field outside the grid,
// by Chebyshev-filtered subspace iteration (Zhou and Saad, "Self-consistent-
// field calculations using Chebyshev-filtered subspace iteration", J. Comput.
// Phys. 219 (2006)): a block of nmodes + 8 vectors is filtered by a degree-40
// Chebyshev polynomial that damps the spectrum below the block's lowest Ritz
// value, orthonormalised (column-scaled Cholesky QR, twice) and rotated onto
// the Ritz vectors of A.  The reference finds the same eigenpairs with
// Arnoldi (eigs, which = :LR).  The block operations are device kernels (the
// stencil sweep, tall-skinny Gram products, small-matrix combinations); the
// p x p Cholesky factor and Jacobi eigensolve run on the host.  Every
// reduction runs in a fixed order, so results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fh_internal.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kRows = 64;        // rows per tile of the Gram kernels
constexpr int kMaxVecs = 48;     // subspace size bound: nmodes + kGuard
constexpr int kGuard = 8;        // vectors beyond the wanted modes
constexpr int kDegree = 40;      // Chebyshev filter degree
constexpr int kGramBlocks = 256; // workgroups of the overlap reduction
constexpr int kParts = 256;      // workgroups (partial sums) of the subspace Gram and residual kernels
constexpr int kMaxOwnGram = (kMaxVecs * kMaxVecs + kBlock - 1) / kBlock;
constexpr int kMaxOverlapModes = 16;
constexpr int kMaxOwn = (kMaxOverlapModes * (kMaxOverlapModes + 1) / 2 * kMaxOverlapModes * (kMaxOverlapModes + 1) / 2 +
                         kBlock - 1) / kBlock;

int fail(int code, const std::string& m) { return fhi::set_error(code, m); }

#define GR_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(e_ == hipErrorOutOfMemory ? FH_ENOMEM : FH_EDEVICE,            \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

struct DevMem {
    double* p = nullptr;
    ~DevMem()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t count)
    {
        GR_HIP(hipMalloc((void**)&p, sizeof(double) * std::max<size_t>(count, 1)));
        return FH_OK;
    }
};

// y_v = a ((A - cen) x_v) + b z_v for vector v = blockIdx.y, with
// A x = d0 x + c (sum of the four grid neighbours inside the grid).  Vectors
// are [p][n], grid index i + nx j.  z is read only when b != 0.
__global__ __launch_bounds__(kBlock) void k_cheb(const double* __restrict__ x, const double* __restrict__ z,
                                                 double* __restrict__ y, const double* __restrict__ d0, int nx,
                                                 int ny, double c, double cen, double a, double b)
{
    const int64_t n = (int64_t)nx * ny;
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const int64_t off = (int64_t)blockIdx.y * n;
    const double* xv = x + off;
    const int i = (int)(r % nx);
    const int64_t j = r / nx;
    double s = 0.0;
    if (i > 0) s += xv[r - 1];
    if (i < nx - 1) s += xv[r + 1];
    if (j > 0) s += xv[r - nx];
    if (j < ny - 1) s += xv[r + nx];
    double v = a * ((d0[r] - cen) * xv[r] + c * s);
    if (b != 0.0) v += b * z[off + r];
    y[off + r] = v;
}

// partial[b][i q + j] = sum of X_i(r) Y_j(r) (i < p, j < q) over the row
// tiles b, b + G, ... of workgroup b (G = gridDim.x), accumulated in tile
// order; thread t owns the entries t + kBlock e.
__global__ __launch_bounds__(kBlock) void k_gram_partial(const double* __restrict__ X, int p,
                                                         const double* __restrict__ Y, int q, int64_t n,
                                                         double* __restrict__ partial)
{
    __shared__ double xs[kMaxVecs * kRows], ys[kMaxVecs * kRows];
    const int pq = p * q;
    double acc[kMaxOwnGram];
#pragma unroll
    for (int e = 0; e < kMaxOwnGram; ++e) acc[e] = 0.0;
    const int64_t tiles = (n + kRows - 1) / kRows;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t r0 = tile * kRows;
        __syncthreads();
        for (int t = threadIdx.x; t < p * kRows; t += kBlock) {
            const int v = t / kRows, rr = t % kRows;
            xs[t] = r0 + rr < n ? X[(int64_t)v * n + r0 + rr] : 0.0;
        }
        for (int t = threadIdx.x; t < q * kRows; t += kBlock) {
            const int v = t / kRows, rr = t % kRows;
            ys[t] = r0 + rr < n ? Y[(int64_t)v * n + r0 + rr] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < kMaxOwnGram; ++e) {
            const int t = threadIdx.x + kBlock * e;
            if (t < pq) {
                const int i = t / q, j = t % q;
                double s = 0.0;
                for (int rr = 0; rr < kRows; ++rr) s += xs[i * kRows + rr] * ys[j * kRows + rr];
                acc[e] += s;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < kMaxOwnGram; ++e) {
        const int t = threadIdx.x + kBlock * e;
        if (t < pq) partial[(int64_t)blockIdx.x * pq + t] = acc[e];
    }
}

// out[k] = sum over t < parts of partial[t][k] for k = blockIdx.x: lane l
// adds the partials l, l + kBlock, ... in order, then a fixed tree adds the
// lanes.
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* __restrict__ partial, int64_t parts, int m,
                                                         double* __restrict__ out)
{
    __shared__ double red[kBlock];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int64_t t = threadIdx.x; t < parts; t += kBlock) s += partial[t * m + k];
    red[threadIdx.x] = s;
    __syncthreads();
#pragma unroll 1
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
}

// X_j(r) = sum_i Y_i(r) C[i q + j] for j < q (X and Y distinct).
__global__ __launch_bounds__(kBlock) void k_combine(const double* __restrict__ Y, int p, const double* __restrict__ C,
                                                    int q, int64_t n, double* __restrict__ X)
{
    __shared__ double cs[kMaxVecs * kMaxVecs];
    for (int t = threadIdx.x; t < p * q; t += kBlock) cs[t] = C[t];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    for (int j = 0; j < q; ++j) {
        double s = 0.0;
        for (int i = 0; i < p; ++i) s += Y[(int64_t)i * n + r] * cs[i * q + j];
        X[(int64_t)j * n + r] = s;
    }
}

// partial[b][j] = sum of (AX_j - w_j X_j)^2 (j < m) over the rows of the row
// tiles b, b + G, ... of workgroup b, in order; lane r of the workgroup takes
// row r of each tile and the lanes' sums are added in lane order.
__global__ __launch_bounds__(kBlock) void k_resid_partial(const double* __restrict__ AX, const double* __restrict__ X,
                                                          const double* __restrict__ w, int m, int64_t n,
                                                          double* __restrict__ partial)
{
    __shared__ double red[kBlock];
    const int64_t tiles = (n + kBlock - 1) / kBlock;
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
            const int64_t r = tile * kBlock + threadIdx.x;
            if (r < n) {
                const int64_t k = (int64_t)j * n + r;
                const double e = AX[k] - w[j] * X[k];
                s += e * e;
            }
        }
        red[threadIdx.x] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int l = 0; l < kBlock; ++l) t += red[l];
            partial[(int64_t)blockIdx.x * m + j] = t;
        }
        __syncthreads();
    }
}

// Deterministic start block: splitmix64 of the element index, uniform in [-1, 1).
__global__ __launch_bounds__(kBlock) void k_random(double* __restrict__ X, int64_t total, uint64_t seed)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= total) return;
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(k + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    X[k] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// partial[b][a q' + b'] over pair products: for the workgroup's row tiles
// b, b + G, ... (G = gridDim.x), sum phi_i phi_j phi_k phi_l with (i, j) the
// a-th and (k, l) the b'-th pair of `pairs` (i <= j).  Each thread owns the
// entries t + kBlock e and accumulates them in tile order.
__global__ __launch_bounds__(kBlock) void k_overlap_partial(const double* __restrict__ phi, int M, int64_t n,
                                                            const int* __restrict__ pairs, int Q,
                                                            double* __restrict__ partial)
{
    __shared__ double ph[kMaxOverlapModes * kRows];
    __shared__ double pp[kMaxOverlapModes * (kMaxOverlapModes + 1) / 2 * kRows];
    const int QQ = Q * Q;
    double acc[kMaxOwn];
#pragma unroll
    for (int e = 0; e < kMaxOwn; ++e) acc[e] = 0.0;
    const int64_t tiles = (n + kRows - 1) / kRows;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t r0 = tile * kRows;
        __syncthreads();
        for (int t = threadIdx.x; t < M * kRows; t += kBlock) {
            const int v = t / kRows, rr = t % kRows;
            ph[t] = r0 + rr < n ? phi[(int64_t)v * n + r0 + rr] : 0.0;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < Q * kRows; t += kBlock) {
            const int a = t / kRows, rr = t % kRows;
            pp[t] = ph[pairs[2 * a] * kRows + rr] * ph[pairs[2 * a + 1] * kRows + rr];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < kMaxOwn; ++e) {
            const int t = threadIdx.x + kBlock * e;
            if (t < QQ) {
                const int a = t / Q, b = t % Q;
                double s = 0.0;
                for (int rr = 0; rr < kRows; ++rr) s += pp[a * kRows + rr] * pp[b * kRows + rr];
                acc[e] += s;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < kMaxOwn; ++e) {
        const int t = threadIdx.x + kBlock * e;
        if (t < QQ) partial[(int64_t)blockIdx.x * QQ + t] = acc[e];
    }
}

// ---------------------------------------------------------------- host side

// Lower Cholesky factor of the symmetric positive definite g (p x p, row-major) in place.
bool cholesky(std::vector<double>& g, int p)
{
    for (int j = 0; j < p; ++j) {
        double d = g[j * p + j];
        for (int k = 0; k < j; ++k) d -= g[j * p + k] * g[j * p + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        g[j * p + j] = d;
        for (int i = j + 1; i < p; ++i) {
            double s = g[i * p + j];
            for (int k = 0; k < j; ++k) s -= g[i * p + k] * g[j * p + k];
            g[i * p + j] = s / d;
        }
        for (int k = j + 1; k < p; ++k) g[j * p + k] = 0.0;
    }
    return true;
}

// Eigen-decomposition of the symmetric h (p x p, row-major) by cyclic Jacobi
// rotations (Golub and Van Loan, Matrix Computations, 4th ed., sec. 8.5):
// eigenvalues w descending, eigenvectors as the columns of v (row-major).
void jacobi_eigen(std::vector<double> h, int p, std::vector<double>& w, std::vector<double>& v)
{
    std::vector<double> q(p * p, 0.0);
    for (int i = 0; i < p; ++i) q[i * p + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < p; ++i)
            for (int j = 0; j < p; ++j) {
                tot += h[i * p + j] * h[i * p + j];
                if (i != j) off += h[i * p + j] * h[i * p + j];
            }
        if (off <= 1e-30 * tot) break;
        for (int a = 0; a < p - 1; ++a)
            for (int b = a + 1; b < p; ++b) {
                const double hab = h[a * p + b];
                if (hab == 0.0) continue;
                const double tau = (h[b * p + b] - h[a * p + a]) / (2.0 * hab);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < p; ++k) {  // h <- h J (columns a, b)
                    const double x = h[k * p + a], y = h[k * p + b];
                    h[k * p + a] = c * x - s * y;
                    h[k * p + b] = s * x + c * y;
                }
                for (int k = 0; k < p; ++k) {  // h <- J^T h (rows a, b)
                    const double x = h[a * p + k], y = h[b * p + k];
                    h[a * p + k] = c * x - s * y;
                    h[b * p + k] = s * x + c * y;
                }
                for (int k = 0; k < p; ++k) {  // q <- q J
                    const double x = q[k * p + a], y = q[k * p + b];
                    q[k * p + a] = c * x - s * y;
                    q[k * p + b] = s * x + c * y;
                }
            }
    }
    std::vector<int> order(p);
    for (int i = 0; i < p; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h[a * p + a] > h[b * p + b]; });
    w.assign(p, 0.0);
    v.assign(p * p, 0.0);
    for (int j = 0; j < p; ++j) {
        w[j] = h[order[j] * p + order[j]];
        for (int i = 0; i < p; ++i) v[i * p + j] = q[i * p + order[j]];
    }
}

struct Block {
    hipStream_t s;
    int64_t n;
    int nx, ny;
    const double* d0;
    double c;
    double* partial;  // [parts][p^2]
    double* small;    // [kMaxVecs^2] reduced Gram / coefficients
    int64_t parts() const { return std::min<int64_t>(kParts, (n + kRows - 1) / kRows); }
    dim3 rows_grid() const { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

    int apply(const double* x, const double* z, double* y, int p, double cen, double a, double b)
    {
        hipLaunchKernelGGL(k_cheb, dim3((unsigned)((n + kBlock - 1) / kBlock), (unsigned)p), dim3(kBlock), 0, s, x,
                           z, y, d0, nx, ny, c, cen, a, b);
        GR_HIP(hipGetLastError());
        return FH_OK;
    }
    // g = X^T Y (p x q, row-major) on the host
    int gram(const double* X, int p, const double* Y, int q, std::vector<double>& g)
    {
        hipLaunchKernelGGL(k_gram_partial, dim3((unsigned)parts()), dim3(kBlock), 0, s, X, p, Y, q, n, partial);
        GR_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)(p * q)), dim3(kBlock), 0, s, partial, parts(), p * q,
                           small);
        GR_HIP(hipGetLastError());
        g.resize((size_t)p * q);
        GR_HIP(hipMemcpyAsync(g.data(), small, sizeof(double) * p * q, hipMemcpyDeviceToHost, s));
        GR_HIP(hipStreamSynchronize(s));
        return FH_OK;
    }
    // X = Y C (C p x q, row-major, host)
    int combine(const double* Y, int p, const std::vector<double>& C, int q, double* X)
    {
        GR_HIP(hipMemcpyAsync(small, C.data(), sizeof(double) * p * q, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_combine, rows_grid(), dim3(kBlock), 0, s, Y, p, small, q, n, X);
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(s));  // `small` is reused by the next call
        return FH_OK;
    }
    // out = Y with orthonormal columns spanning the columns of Y (column-
    // scaled Cholesky QR, twice); tmp is scratch of the same size.
    int orthonormalize(double*& Y, double*& tmp, int p)
    {
        std::vector<double> g, C(p * p);
        for (int pass = 0; pass < 2; ++pass) {
            int rc = gram(Y, p, Y, p, g);
            if (rc) return rc;
            std::vector<double> d(p);
            for (int i = 0; i < p; ++i) {
                d[i] = std::sqrt(g[i * p + i]);
                if (!(d[i] > 0.0) || !std::isfinite(d[i]))
                    return fail(FH_ENONFINITE, "mode solve: the subspace lost a vector");
            }
            for (int i = 0; i < p; ++i)
                for (int j = 0; j < p; ++j) g[i * p + j] /= d[i] * d[j];
            if (!cholesky(g, p)) return fail(FH_ENONFINITE, "mode solve: the subspace lost rank");
            // C = D^-1 L^-T: solve L^T U = I column by column (U upper triangular)
            std::vector<double> U(p * p, 0.0);
            for (int col = 0; col < p; ++col)
                for (int i = col; i >= 0; --i) {
                    double s = i == col ? 1.0 : 0.0;
                    for (int k = i + 1; k <= col; ++k) s -= g[k * p + i] * U[k * p + col];
                    U[i * p + col] = s / g[i * p + i];
                }
            for (int i = 0; i < p; ++i)
                for (int j = 0; j < p; ++j) C[i * p + j] = U[i * p + j] / d[i];
            if ((rc = combine(Y, p, C, p, tmp))) return rc;
            std::swap(Y, tmp);
        }
        return FH_OK;
    }
};

}  // namespace

extern "C" {

int fh_fd_modes(fh_context ctx, int32_t nx, int32_t ny, double dx, double k0, const double* eps, int32_t nmodes,
                double tol, int32_t max_iter, double* beta2, double* modes, int32_t* iterations)
{
    if (!ctx || !eps || !beta2 || !modes) return fail(FH_EINVAL, "context, eps or an output is NULL");
    if (nx < 3 || ny < 3) return fail(FH_EINVAL, "the grid needs at least 3 x 3 points");
    if (nmodes < 1 || nmodes + kGuard > kMaxVecs)
        return fail(FH_EINVAL, "nmodes must lie in 1.." + std::to_string(kMaxVecs - kGuard));
    const int64_t n = (int64_t)nx * ny;
    const int p = nmodes + kGuard;
    if (n < 2 * p) return fail(FH_EINVAL, "the grid has too few points for the requested modes");
    if (!(std::isfinite(dx) && dx > 0) || !(std::isfinite(k0) && k0 > 0))
        return fail(FH_EINVAL, "dx and k0 must be positive and finite");
    if (!(tol > 0)) tol = 1e-10;
    if (max_iter <= 0) max_iter = 200;
    const double c = 1.0 / (dx * dx);
    std::vector<double> d0(n);
    double dmin = INFINITY, dmax = -INFINITY;
    for (int64_t r = 0; r < n; ++r) {
        if (!std::isfinite(eps[r])) return fail(FH_EINVAL, "eps must be finite");
        d0[r] = eps[r] * k0 * k0 - 4.0 * c;
        dmin = std::min(dmin, d0[r]);
        dmax = std::max(dmax, d0[r]);
    }
    // Gershgorin bounds of the spectrum
    const double lo = dmin - 4.0 * c, hi = dmax + 4.0 * c;
    GR_HIP(hipSetDevice(ctx->device));
    DevMem dd0, buf[5], part, small;
    int rc;
    if ((rc = dd0.alloc(n))) return rc;
    for (auto& b : buf)
        if ((rc = b.alloc((size_t)p * n))) return rc;
    if ((rc = part.alloc((size_t)kParts * p * p))) return rc;
    if ((rc = small.alloc(kMaxVecs * kMaxVecs))) return rc;
    Block B{ctx->stream, n, nx, ny, dd0.p, c, part.p, small.p};
    GR_HIP(hipMemcpyAsync(dd0.p, d0.data(), sizeof(double) * n, hipMemcpyHostToDevice, B.s));
    double* X = buf[0].p;
    double* T1 = buf[1].p;
    double* T2 = buf[2].p;
    double* AY = buf[3].p;
    double* AX = buf[4].p;
    const int64_t total = (int64_t)p * n;
    hipLaunchKernelGGL(k_random, dim3((unsigned)((total + kBlock - 1) / kBlock)), dim3(kBlock), 0, B.s, X, total,
                       (uint64_t)0x5eed0f1b2c3d4e5full);
    GR_HIP(hipGetLastError());
    if ((rc = B.orthonormalize(X, T1, p))) return rc;
    double cut = lo + 0.5 * (hi - lo);
    std::vector<double> w, Q, H, res(nmodes);
    int it = 0;
    bool converged = false;
    while (it < max_iter && !converged) {
        ++it;
        // Chebyshev filter of degree kDegree damping [lo, cut], scaled at hi:
        // Y_1 = (s1/e)(A - cen) X, Y_{k+1} = (2 s_{k+1}/e)(A - cen) Y_k - s_k s_{k+1} Y_{k-1}
        const double e = 0.5 * (cut - lo), cen = 0.5 * (cut + lo), s1 = e / (hi - cen);
        if (!(e > 0) || !(s1 > 0)) return fail(FH_ENONFINITE, "mode solve: degenerate filter interval");
        double* prev = X;
        double* cur = T1;
        double* nxt = T2;
        if ((rc = B.apply(prev, nullptr, cur, p, cen, s1 / e, 0.0))) return rc;
        double sig = s1;
        for (int k = 2; k <= kDegree; ++k) {
            const double sn = 1.0 / (2.0 / s1 - sig);
            if ((rc = B.apply(cur, prev, nxt, p, cen, 2.0 * sn / e, -sig * sn))) return rc;
            double* old = prev;
            prev = cur;
            cur = nxt;
            nxt = old;
            sig = sn;
        }
        // the filtered block is in cur; prev and nxt are free
        double* Y = cur;
        double* tmp = prev;
        if ((rc = B.orthonormalize(Y, tmp, p))) return rc;
        // Rayleigh-Ritz: H = Y^T A Y, X = Y V, AX = AY V
        if ((rc = B.apply(Y, nullptr, AY, p, 0.0, 1.0, 0.0))) return rc;
        if ((rc = B.gram(Y, p, AY, p, H))) return rc;
        for (int i = 0; i < p; ++i)
            for (int j = i + 1; j < p; ++j) H[i * p + j] = H[j * p + i] = 0.5 * (H[i * p + j] + H[j * p + i]);
        jacobi_eigen(H, p, w, Q);
        double* Xn = tmp;  // free after the orthonormalisation
        if ((rc = B.combine(Y, p, Q, p, Xn))) return rc;
        if ((rc = B.combine(AY, p, Q, p, AX))) return rc;
        // the new basis goes to buf[0]'s role: keep X pointing at it, and the
        // other two of {buf0, buf1, buf2} as T1, T2
        double* all[3] = {buf[0].p, buf[1].p, buf[2].p};
        X = Xn;
        int o = 0;
        for (double* b : all)
            if (b != X) (o++ == 0 ? T1 : T2) = b;
        // residuals of the wanted pairs
        GR_HIP(hipMemcpyAsync(small.p, w.data(), sizeof(double) * nmodes, hipMemcpyHostToDevice, B.s));
        hipLaunchKernelGGL(k_resid_partial, dim3((unsigned)B.parts()), dim3(kBlock), 0, B.s, AX, X, small.p, nmodes,
                           n, part.p);
        GR_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)nmodes), dim3(kBlock), 0, B.s, part.p, B.parts(), nmodes,
                           small.p + kMaxVecs);
        GR_HIP(hipGetLastError());
        GR_HIP(hipMemcpyAsync(res.data(), small.p + kMaxVecs, sizeof(double) * nmodes, hipMemcpyDeviceToHost, B.s));
        GR_HIP(hipStreamSynchronize(B.s));
        double worst = 0.0;
        for (int j = 0; j < nmodes; ++j) {
            const double r = std::sqrt(res[j]) / std::fabs(w[j]);
            if (!std::isfinite(r)) return fail(FH_ENONFINITE, "mode solve produced non-finite residuals");
            worst = std::max(worst, r);
        }
        converged = worst <= tol;
        cut = w[p - 1];
    }
    // outputs: eigenvalues descending, unit vectors with their largest entry positive
    std::vector<double> host((size_t)nmodes * n);
    GR_HIP(hipMemcpyAsync(host.data(), X, sizeof(double) * nmodes * n, hipMemcpyDeviceToHost, B.s));
    GR_HIP(hipStreamSynchronize(B.s));
    for (int j = 0; j < nmodes; ++j) {
        double* v = host.data() + (size_t)j * n;
        double nrm = 0.0, big = 0.0;
        for (int64_t r = 0; r < n; ++r) {
            nrm += v[r] * v[r];
            if (std::fabs(v[r]) > std::fabs(big)) big = v[r];
        }
        const double scale = (big < 0 ? -1.0 : 1.0) / std::sqrt(nrm);
        for (int64_t r = 0; r < n; ++r) v[r] *= scale;
        beta2[j] = w[j];
    }
    std::memcpy(modes, host.data(), sizeof(double) * nmodes * n);
    if (iterations) *iterations = it;
    if (!converged)
        return fail(FH_EMAXSTEPS, "mode solve did not reach the residual tolerance in " + std::to_string(max_iter) +
                                      " iterations");
    return FH_OK;
}

int fh_overlap_tensor(fh_context ctx, int64_t points, int32_t modes, const double* phi, double scale, double* sk)
{
    if (!ctx || !phi || !sk) return fail(FH_EINVAL, "context, phi or sk is NULL");
    if (modes < 1 || modes > kMaxOverlapModes)
        return fail(FH_EINVAL, "modes must lie in 1.." + std::to_string(kMaxOverlapModes));
    if (points < 1) return fail(FH_EINVAL, "points must be positive");
    if (!std::isfinite(scale)) return fail(FH_EINVAL, "scale must be finite");
    const int M = modes, Q = M * (M + 1) / 2, QQ = Q * Q;
    std::vector<int> pairs, index(M * M);
    for (int i = 0; i < M; ++i)
        for (int j = i; j < M; ++j) {
            index[i * M + j] = index[j * M + i] = (int)pairs.size() / 2;
            pairs.push_back(i);
            pairs.push_back(j);
        }
    GR_HIP(hipSetDevice(ctx->device));
    DevMem dphi, dpart, dsum;
    int* dpairs = nullptr;
    int rc;
    if ((rc = dphi.alloc((size_t)M * points))) return rc;
    if ((rc = dpart.alloc((size_t)kGramBlocks * QQ))) return rc;
    if ((rc = dsum.alloc(QQ))) return rc;
    GR_HIP(hipMalloc((void**)&dpairs, sizeof(int) * pairs.size()));
    const hipStream_t s = ctx->stream;
    std::vector<double> g(QQ);
    const hipError_t e = [&] {
        hipError_t x = hipMemcpyAsync(dphi.p, phi, sizeof(double) * M * points, hipMemcpyHostToDevice, s);
        if (x == hipSuccess) x = hipMemcpyAsync(dpairs, pairs.data(), sizeof(int) * pairs.size(), hipMemcpyHostToDevice, s);
        if (x != hipSuccess) return x;
        hipLaunchKernelGGL(k_overlap_partial, dim3(kGramBlocks), dim3(kBlock), 0, s, dphi.p, M, (int64_t)points,
                           dpairs, Q, dpart.p);
        if ((x = hipGetLastError()) != hipSuccess) return x;
        hipLaunchKernelGGL(k_sum_partials, dim3(QQ), dim3(kBlock), 0, s, dpart.p, (int64_t)kGramBlocks, QQ,
                           dsum.p);
        if ((x = hipGetLastError()) != hipSuccess) return x;
        if ((x = hipMemcpyAsync(g.data(), dsum.p, sizeof(double) * QQ, hipMemcpyDeviceToHost, s)) != hipSuccess)
            return x;
        return hipStreamSynchronize(s);
    }();
    (void)hipFree(dpairs);
    if (e != hipSuccess) return fail(FH_EDEVICE, std::string("overlap tensor: ") + hipGetErrorString(e));
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j)
            for (int k = 0; k < M; ++k)
                for (int l = 0; l < M; ++l)
                    sk[i + M * (j + M * (k + M * l))] = scale * g[index[i * M + j] * Q + index[k * M + l]];
    return FH_OK;
}

}  // extern "C"
