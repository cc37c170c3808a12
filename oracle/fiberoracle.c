/*
 * fiberoracle.c -- CPU FP64 restatement of FiberLab's GMMNLSE propagation
 * hot path.  TEST INFRASTRUCTURE ONLY: this file is the parity checker for
 * libfiberhip.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function cites the reference file:line it restates.  The reference
 * itself is Julia (FFTW, Tullio, OrdinaryDiffEq); Julia is not installed in
 * this image, so the arithmetic below is a line-by-line restatement with a
 * self-written radix-2 FFT in place of FFTW 3.3.11.
 *
 * Layout: Julia column-major, time/frequency index fastest.  A field of
 * nt samples and m modes is an array of m*nt complex doubles, sample t of
 * mode k at [k*nt + t].  gamma[i,j,k,l] (1-based in Julia) is stored at
 * i + m*(j + m*(k + m*l)) with 0-based indices.
 *
 * Conventions (restated from the source):
 *   fft  = unnormalised, kernel exp(-2 pi i k n / N)      (FFTW.fft)
 *   ifft = exp(+2 pi i k n / N) scaled by 1/N              (FFTW.ifft)
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------ FFT */

typedef struct {
    int n, logn;
    int *rev;
    cplx *tw; /* exp(-2 pi i k / n), k < n/2, from long-double sin/cos */
} orc_fft_plan;

static int ilog2(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }

static int fft_plan_init(orc_fft_plan *p, int n)
{
    p->n = n;
    p->logn = ilog2(n);
    if ((1 << p->logn) != n || n < 2) return -1;
    p->rev = (int *)malloc(sizeof(int) * (size_t)n);
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n / 2 > 0 ? n / 2 : 1));
    if (!p->rev || !p->tw) return -1;
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < p->logn; ++b) r |= ((i >> b) & 1) << (p->logn - 1 - b);
        p->rev[i] = r;
    }
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n / 2; ++k) {
        long double a = two_pi * (long double)k / (long double)n;
        p->tw[k] = (double)cosl(a) - I * (double)sinl(a);
    }
    return 0;
}

static void fft_plan_free(orc_fft_plan *p)
{
    free(p->rev);
    free(p->tw);
    p->rev = NULL;
    p->tw = NULL;
}

/* In-place radix-2 decimation-in-time transform.  inverse != 0 gives
 * FFTW.ifft semantics (conjugate kernel, 1/N scaling). */
static void fft_exec(const orc_fft_plan *p, cplx *x, int inverse)
{
    const int n = p->n;
    for (int i = 0; i < n; ++i) {
        int j = p->rev[i];
        if (i < j) { cplx t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, stride = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; ++k) {
                cplx w = p->tw[k * stride];
                if (inverse) w = conj(w);
                cplx a = x[i + k], b = x[i + k + half] * w;
                x[i + k] = a + b;
                x[i + k + half] = a - b;
            }
        }
    }
    if (inverse) {
        const double s = 1.0 / (double)n;
        for (int i = 0; i < n; ++i) x[i] *= s;
    }
}

/* Exported for convention tests: x has 2*n doubles (interleaved). */
int orc_fft(double *x, int n, int inverse)
{
    orc_fft_plan p;
    if (fft_plan_init(&p, n)) { fft_plan_free(&p); return -1; }
    fft_exec(&p, (cplx *)x, inverse);
    fft_plan_free(&p);
    return 0;
}

/* ------------------------------------------------------------ the fiber */

typedef struct {
    int nt, m;
    const double *disp;  /* D(w) [m*nt] rad/m, FFT order            */
    const double *gamma; /* [m^4] column-major gamma[i,j,k,l]        */
    const double *hr;    /* hR(w) interleaved complex [2*nt]          */
    const double *shock; /* fftshift(ws/w0) [nt]                      */
    double one_m_fr;
} orc_fiber;

#define G4(f, i, j, k, l) ((f)->gamma[(i) + (f)->m * ((j) + (f)->m * ((k) + (f)->m * (l)))])

typedef struct {
    orc_fft_plan plan;
    int nt, m;
    cplx *expp, *expm, *uw, *eta, *dkc, *gu, *gw, *tmp;
    double *v, *w, *dk, *dr, *h, *a, *b, *c;
} orc_ws;

static void ws_free(orc_ws *ws)
{
    fft_plan_free(&ws->plan);
    free(ws->expp); free(ws->expm); free(ws->uw); free(ws->eta); free(ws->dkc);
    free(ws->gu); free(ws->gw); free(ws->tmp);
    free(ws->v); free(ws->w); free(ws->dk); free(ws->dr); free(ws->h);
    free(ws->a); free(ws->b); free(ws->c);
    memset(ws, 0, sizeof(*ws));
}

static int ws_init(orc_ws *ws, int nt, int m)
{
    memset(ws, 0, sizeof(*ws));
    if (fft_plan_init(&ws->plan, nt)) { ws_free(ws); return -1; }
    ws->nt = nt;
    ws->m = m;
    size_t f = (size_t)nt * (size_t)m, ff = (size_t)nt * (size_t)m * (size_t)m;
    ws->expp = malloc(sizeof(cplx) * f);
    ws->expm = malloc(sizeof(cplx) * f);
    ws->uw = malloc(sizeof(cplx) * f);
    ws->eta = malloc(sizeof(cplx) * f);
    ws->gu = malloc(sizeof(cplx) * f);
    ws->gw = malloc(sizeof(cplx) * f);
    ws->tmp = malloc(sizeof(cplx) * f);
    ws->dkc = malloc(sizeof(cplx) * ff);
    ws->v = malloc(sizeof(double) * f);
    ws->w = malloc(sizeof(double) * f);
    ws->dk = malloc(sizeof(double) * ff);
    ws->dr = malloc(sizeof(double) * ff);
    ws->h = malloc(sizeof(double) * ff);
    ws->a = malloc(sizeof(double) * ff);
    ws->b = malloc(sizeof(double) * ff);
    ws->c = malloc(sizeof(double) * (size_t)nt * (size_t)m * (size_t)m);
    if (!ws->expp || !ws->expm || !ws->uw || !ws->eta || !ws->gu || !ws->gw || !ws->tmp ||
        !ws->dkc || !ws->v || !ws->w || !ws->dk || !ws->dr || !ws->h || !ws->a || !ws->b || !ws->c) {
        ws_free(ws);
        return -1;
    }
    return 0;
}

static inline cplx cis(double x) { return cos(x) + I * sin(x); }

/* fftshift along t for even nt: out[t] = in[(t + nt/2) mod nt]. */
static void shift_real_part(const cplx *in, double *out, int nt)
{
    const int h = nt / 2;
    for (int t = 0; t < nt; ++t) out[t] = creal(in[(t + h) % nt]);
}

/* Raman filter of one real column, literal sequence of
 * simulate_disp_mmf.jl:45-50: real(fftshift(ifft(hR .* fft(complex(q))))).
 * conj_filter != 0 uses conj(hR) (transposed operator, see below). */
static void raman_apply(orc_ws *ws, const cplx *hr, const double *q, double *out, int conj_filter,
                        int shift_before)
{
    const int nt = ws->nt;
    cplx *x = ws->tmp;
    if (shift_before) {
        for (int t = 0; t < nt; ++t) x[t] = q[(t + nt / 2) % nt];
    } else {
        for (int t = 0; t < nt; ++t) x[t] = q[t];
    }
    fft_exec(&ws->plan, x, 0);
    for (int k = 0; k < nt; ++k) x[k] *= conj_filter ? conj(hr[k]) : hr[k];
    fft_exec(&ws->plan, x, 1);
    if (shift_before) {
        for (int t = 0; t < nt; ++t) out[t] = creal(x[t]);
    } else {
        shift_real_part(x, out, nt);
    }
}

/* Forward RHS, restating disp_mmf! (simulate_disp_mmf.jl:25-60):
 *   u = fft(cis(Dz) .* x)                                   :28-35
 *   dK[t,i,j] = sum_kl gamma[i,j,k,l](v_k v_l + w_k w_l)     :38
 *   etaK = (1-fR)(dK u)                                     :39-42
 *   dR = real(fftshift(ifft(hR .* fft(dK))))                :45-50
 *   etaR = dR u                                             :51-53
 *   dx = i cis(-Dz) .* shock .* ifft(etaK + etaR)           :56-59  */
static void rhs_eval(const orc_fiber *f, orc_ws *ws, double z, const cplx *x, cplx *dx)
{
    const int nt = f->nt, m = f->m;
    const cplx *hr = (const cplx *)f->hr;
    for (int k = 0; k < m; ++k)
        for (int t = 0; t < nt; ++t) {
            size_t o = (size_t)k * nt + t;
            ws->expp[o] = cis(f->disp[o] * z);
            ws->expm[o] = cis(-f->disp[o] * z);
            ws->uw[o] = ws->expp[o] * x[o];
        }
    for (int k = 0; k < m; ++k) fft_exec(&ws->plan, ws->uw + (size_t)k * nt, 0);
    for (size_t o = 0; o < (size_t)nt * m; ++o) {
        ws->v[o] = creal(ws->uw[o]);
        ws->w[o] = cimag(ws->uw[o]);
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double *col = ws->dk + ((size_t)i + (size_t)m * j) * nt;
            for (int t = 0; t < nt; ++t) {
                double s = 0.0;
                for (int k = 0; k < m; ++k)
                    for (int l = 0; l < m; ++l)
                        s += G4(f, i, j, k, l) *
                             (ws->v[(size_t)k * nt + t] * ws->v[(size_t)l * nt + t] +
                              ws->w[(size_t)k * nt + t] * ws->w[(size_t)l * nt + t]);
                col[t] = s;
            }
        }
    for (int p = 0; p < m * m; ++p)
        raman_apply(ws, hr, ws->dk + (size_t)p * nt, ws->dr + (size_t)p * nt, 0, 0);
    for (int i = 0; i < m; ++i)
        for (int t = 0; t < nt; ++t) {
            double ak = 0, bk = 0, ar = 0, br = 0;
            for (int j = 0; j < m; ++j) {
                double d = ws->dk[((size_t)i + (size_t)m * j) * nt + t];
                double r = ws->dr[((size_t)i + (size_t)m * j) * nt + t];
                double vj = ws->v[(size_t)j * nt + t], wj = ws->w[(size_t)j * nt + t];
                ak += d * vj;
                bk += d * wj;
                ar += r * vj;
                br += r * wj;
            }
            cplx etak = (ak + I * bk) * f->one_m_fr;
            cplx etar = ar + I * br;
            ws->eta[(size_t)i * nt + t] = etak + etar;
        }
    for (int k = 0; k < m; ++k) fft_exec(&ws->plan, ws->eta + (size_t)k * nt, 1);
    for (int k = 0; k < m; ++k)
        for (int t = 0; t < nt; ++t) {
            size_t o = (size_t)k * nt + t;
            dx[o] = I * ws->expm[o] * (ws->eta[o] * f->shock[t]);
        }
}

int orc_rhs(const orc_fiber *f, double z, const double *x, double *dx)
{
    orc_ws ws;
    if (ws_init(&ws, f->nt, f->m)) return -1;
    rhs_eval(f, &ws, z, (const cplx *)x, (cplx *)dx);
    ws_free(&ws);
    return 0;
}

/* ------------------------------------------------- fixed-step RK4 (IP) */

/* One classic RK4 step in the global frame, restating the loop body of
 * _scalar_rk4_output (scalar_crosscheck.jl:53-60), generalised to M modes
 * with the disp_mmf! right-hand side.  z = index*h, stages at z, z+h/2,
 * z+h/2, z+h.  Optional stage outputs y2..y4 (stage inputs) for the adjoint. */
static void rk4_step(const orc_fiber *f, orc_ws *ws, int index, double h, cplx *x, cplx *k1, cplx *k2,
                     cplx *k3, cplx *k4, cplx *y)
{
    const size_t nf = (size_t)f->nt * f->m;
    const double z = index * h;
    rhs_eval(f, ws, z, x, k1);
    for (size_t o = 0; o < nf; ++o) y[o] = x[o] + h / 2 * k1[o];
    rhs_eval(f, ws, z + h / 2, y, k2);
    for (size_t o = 0; o < nf; ++o) y[o] = x[o] + h / 2 * k2[o];
    rhs_eval(f, ws, z + h / 2, y, k3);
    for (size_t o = 0; o < nf; ++o) y[o] = x[o] + h * k3[o];
    rhs_eval(f, ws, z + h, y, k4);
    for (size_t o = 0; o < nf; ++o) x[o] += h / 6 * (k1[o] + 2 * k2[o] + 2 * k3[o] + k4[o]);
}

/* _scalar_rk4_output (scalar_crosscheck.jl:8-68) for M >= 1.  u0 is the
 * interaction-picture launch (= uw0), out is cis(D L) .* x(L) (:61-65).
 * If traj != NULL it receives x after every step (steps*nt*m values). */
int orc_rk4_forward(const orc_fiber *f, double length_m, int steps, const double *u0, double *out,
                    double *traj)
{
    if (steps <= 0) return -1;
    const size_t nf = (size_t)f->nt * f->m;
    orc_ws ws;
    if (ws_init(&ws, f->nt, f->m)) return -1;
    cplx *x = malloc(sizeof(cplx) * nf), *k = malloc(sizeof(cplx) * nf * 5);
    if (!x || !k) { free(x); free(k); ws_free(&ws); return -1; }
    memcpy(x, u0, sizeof(cplx) * nf);
    const double h = length_m / steps;
    for (int s = 0; s < steps; ++s) {
        rk4_step(f, &ws, s, h, x, k, k + nf, k + 2 * nf, k + 3 * nf, k + 4 * nf);
        if (traj) memcpy(traj + 2 * nf * (size_t)s, x, sizeof(cplx) * nf);
    }
    cplx *o = (cplx *)out;
    int finite = 1;
    for (size_t i = 0; i < nf; ++i) {
        o[i] = cis(f->disp[i] * length_m) * x[i];
        if (!isfinite(creal(o[i])) || !isfinite(cimag(o[i]))) finite = 0;
    }
    free(x);
    free(k);
    ws_free(&ws);
    return finite ? 0 : -2;
}

/* ------------------------------------------ discrete adjoint (our VJP) */

/* Vector-Jacobian product of the RHS above.  The RHS is real-linear in x,
 * so for a real objective J with cotangent g = dJ/dRe + i dJ/dIm (or the
 * Wirtinger seed dJ/dconj, which differs by the exact factor 2) the
 * pullback is
 *   G_i     = fft(-i cis(D_i z) S gy_i) / N
 *   gu_j    = sum_i h_ij G_i,        h_ij = (1-fR) dK_ij + dR_ij
 *   A_ij    = Re(conj(G_i) u_j)
 *   B_ij    = (1-fR) A_ij + R^T(A_ij),  R^T(a) = real(ifft(conj(hR) fft(shift(a))))
 *   C_kl    = sum_ij B_ij gamma[i,j,k,l]
 *   gu_k   += sum_l (C_kl + C_lk) u_l
 *   gx_m    = conj(cis(D_m z)) N ifft(gu_m)
 * The continuous counterpart is adjoint_disp_mmf! (sensitivity_disp_mmf.jl:32-79),
 * which uses the same conj(hR) and sign-alternation (sigma1, :241-242). */
static void rhs_vjp(const orc_fiber *f, orc_ws *ws, double z, const cplx *x, const cplx *gy, cplx *gx)
{
    const int nt = f->nt, m = f->m;
    const size_t nf = (size_t)nt * m;
    const cplx *hr = (const cplx *)f->hr;
    for (size_t o = 0; o < nf; ++o) {
        ws->expp[o] = cis(f->disp[o] * z);
        ws->uw[o] = ws->expp[o] * x[o];
    }
    for (int k = 0; k < m; ++k) fft_exec(&ws->plan, ws->uw + (size_t)k * nt, 0);
    for (size_t o = 0; o < nf; ++o) {
        ws->v[o] = creal(ws->uw[o]);
        ws->w[o] = cimag(ws->uw[o]);
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double *col = ws->dk + ((size_t)i + (size_t)m * j) * nt;
            for (int t = 0; t < nt; ++t) {
                double s = 0.0;
                for (int k = 0; k < m; ++k)
                    for (int l = 0; l < m; ++l)
                        s += G4(f, i, j, k, l) *
                             (ws->v[(size_t)k * nt + t] * ws->v[(size_t)l * nt + t] +
                              ws->w[(size_t)k * nt + t] * ws->w[(size_t)l * nt + t]);
                col[t] = s;
            }
        }
    for (int p = 0; p < m * m; ++p) {
        raman_apply(ws, hr, ws->dk + (size_t)p * nt, ws->dr + (size_t)p * nt, 0, 0);
        for (int t = 0; t < nt; ++t)
            ws->h[(size_t)p * nt + t] = f->one_m_fr * ws->dk[(size_t)p * nt + t] + ws->dr[(size_t)p * nt + t];
    }
    /* G = fft(-i P S gy)/N */
    for (int k = 0; k < m; ++k)
        for (int t = 0; t < nt; ++t) {
            size_t o = (size_t)k * nt + t;
            ws->gw[o] = -I * ws->expp[o] * (f->shock[t] * gy[o]);
        }
    for (int k = 0; k < m; ++k) {
        fft_exec(&ws->plan, ws->gw + (size_t)k * nt, 0);
        for (int t = 0; t < nt; ++t) ws->gw[(size_t)k * nt + t] /= (double)nt;
    }
    for (int j = 0; j < m; ++j)
        for (int t = 0; t < nt; ++t) {
            cplx s = 0;
            for (int i = 0; i < m; ++i)
                s += ws->h[((size_t)i + (size_t)m * j) * nt + t] * ws->gw[(size_t)i * nt + t];
            ws->gu[(size_t)j * nt + t] = s;
        }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            for (int t = 0; t < nt; ++t)
                ws->a[((size_t)i + (size_t)m * j) * nt + t] =
                    creal(conj(ws->gw[(size_t)i * nt + t]) * ws->uw[(size_t)j * nt + t]);
    for (int p = 0; p < m * m; ++p) {
        raman_apply(ws, hr, ws->a + (size_t)p * nt, ws->b + (size_t)p * nt, 1, 1);
        for (int t = 0; t < nt; ++t)
            ws->b[(size_t)p * nt + t] += f->one_m_fr * ws->a[(size_t)p * nt + t];
    }
    for (int k = 0; k < m; ++k)
        for (int l = 0; l < m; ++l)
            for (int t = 0; t < nt; ++t) {
                double s = 0;
                for (int i = 0; i < m; ++i)
                    for (int j = 0; j < m; ++j)
                        s += ws->b[((size_t)i + (size_t)m * j) * nt + t] * G4(f, i, j, k, l);
                ws->c[((size_t)k + (size_t)m * l) * nt + t] = s;
            }
    for (int k = 0; k < m; ++k)
        for (int t = 0; t < nt; ++t) {
            cplx s = 0;
            for (int l = 0; l < m; ++l)
                s += (ws->c[((size_t)k + (size_t)m * l) * nt + t] + ws->c[((size_t)l + (size_t)m * k) * nt + t]) *
                     ws->uw[(size_t)l * nt + t];
            ws->gu[(size_t)k * nt + t] += s;
        }
    for (int k = 0; k < m; ++k) {
        cplx *col = ws->gu + (size_t)k * nt;
        fft_exec(&ws->plan, col, 1);
        for (int t = 0; t < nt; ++t) {
            size_t o = (size_t)k * nt + t;
            gx[o] = conj(ws->expp[o]) * (col[t] * (double)nt);
        }
    }
}

int orc_rhs_vjp(const orc_fiber *f, double z, const double *x, const double *gy, double *gx)
{
    orc_ws ws;
    if (ws_init(&ws, f->nt, f->m)) return -1;
    rhs_vjp(f, &ws, z, (const cplx *)x, (const cplx *)gy, (cplx *)gx);
    ws_free(&ws);
    return 0;
}

/* Pull a cotangent lam (of x_{n+1}) back through one RK4 step from x_n. */
static void rk4_step_vjp(const orc_fiber *f, orc_ws *ws, int index, double h, const cplx *x, cplx *lam,
                         cplx *k1, cplx *k2, cplx *k3, cplx *y2, cplx *y3, cplx *y4, cplx *g, cplx *gk)
{
    const size_t nf = (size_t)f->nt * f->m;
    const double z = index * h;
    /* recompute stage inputs */
    rhs_eval(f, ws, z, x, k1);
    for (size_t o = 0; o < nf; ++o) y2[o] = x[o] + h / 2 * k1[o];
    rhs_eval(f, ws, z + h / 2, y2, k2);
    for (size_t o = 0; o < nf; ++o) y3[o] = x[o] + h / 2 * k2[o];
    rhs_eval(f, ws, z + h / 2, y3, k3);
    for (size_t o = 0; o < nf; ++o) y4[o] = x[o] + h * k3[o];
    /* cotangents of k_s: x_{n+1} = x + h/6 (k1 + 2k2 + 2k3 + k4) */
    cplx *acc = k1; /* reuse k1 storage for the running cotangent of x */
    for (size_t o = 0; o < nf; ++o) acc[o] = lam[o];
    /* stage 4: gk4 = h/6 lam */
    for (size_t o = 0; o < nf; ++o) gk[o] = h / 6 * lam[o];
    rhs_vjp(f, ws, z + h, y4, gk, g);
    for (size_t o = 0; o < nf; ++o) { acc[o] += g[o]; k2[o] = h * g[o]; } /* k2 := extra cotangent on k3 */
    /* stage 3: gk3 = h/3 lam + h g4 */
    for (size_t o = 0; o < nf; ++o) gk[o] = h / 3 * lam[o] + k2[o];
    rhs_vjp(f, ws, z + h / 2, y3, gk, g);
    for (size_t o = 0; o < nf; ++o) { acc[o] += g[o]; k2[o] = h / 2 * g[o]; }
    /* stage 2: gk2 = h/3 lam + h/2 g3 */
    for (size_t o = 0; o < nf; ++o) gk[o] = h / 3 * lam[o] + k2[o];
    rhs_vjp(f, ws, z + h / 2, y2, gk, g);
    for (size_t o = 0; o < nf; ++o) { acc[o] += g[o]; k2[o] = h / 2 * g[o]; }
    /* stage 1: gk1 = h/6 lam + h/2 g2 */
    for (size_t o = 0; o < nf; ++o) gk[o] = h / 6 * lam[o] + k2[o];
    rhs_vjp(f, ws, z, x, gk, g);
    for (size_t o = 0; o < nf; ++o) lam[o] = acc[o] + g[o];
    (void)k3;
}

/* Discrete adjoint of orc_rk4_forward.  lam_out is the Wirtinger seed
 * dJ/dconj(u_lab(L)) (terminal_adjoint convention, adjoints.jl:723-739);
 * lam0 receives dJ/dconj(u0).  The terminal rotation exp(-i D L) restates
 * solve_adjoint_disp_mmf (sensitivity_disp_mmf.jl:311).  Checkpoints every
 * `every` steps; each segment is recomputed once. */
int orc_rk4_adjoint(const orc_fiber *f, double length_m, int steps, int every, const double *u0,
                    const double *lam_out, double *lam0)
{
    if (steps <= 0 || every <= 0) return -1;
    const size_t nf = (size_t)f->nt * f->m;
    const int nseg = (steps + every - 1) / every;
    orc_ws ws;
    if (ws_init(&ws, f->nt, f->m)) return -1;
    cplx *ck = malloc(sizeof(cplx) * nf * (size_t)nseg);
    cplx *seg = malloc(sizeof(cplx) * nf * (size_t)every);
    cplx *k = malloc(sizeof(cplx) * nf * 9);
    cplx *x = malloc(sizeof(cplx) * nf), *lam = malloc(sizeof(cplx) * nf);
    if (!ck || !seg || !k || !x || !lam) {
        free(ck); free(seg); free(k); free(x); free(lam); ws_free(&ws);
        return -1;
    }
    const double h = length_m / steps;
    memcpy(x, u0, sizeof(cplx) * nf);
    for (int s = 0; s < steps; ++s) {
        if (s % every == 0) memcpy(ck + nf * (size_t)(s / every), x, sizeof(cplx) * nf);
        rk4_step(f, &ws, s, h, x, k, k + nf, k + 2 * nf, k + 3 * nf, k + 4 * nf);
    }
    const cplx *lo = (const cplx *)lam_out;
    for (size_t i = 0; i < nf; ++i) lam[i] = cis(-f->disp[i] * length_m) * lo[i];
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * every, s1 = (s0 + every < steps) ? s0 + every : steps;
        memcpy(x, ck + nf * (size_t)sg, sizeof(cplx) * nf);
        for (int s = s0; s < s1; ++s) {
            memcpy(seg + nf * (size_t)(s - s0), x, sizeof(cplx) * nf);
            if (s + 1 < s1) rk4_step(f, &ws, s, h, x, k, k + nf, k + 2 * nf, k + 3 * nf, k + 4 * nf);
        }
        for (int s = s1 - 1; s >= s0; --s)
            rk4_step_vjp(f, &ws, s, h, seg + nf * (size_t)(s - s0), lam, k, k + nf, k + 2 * nf, k + 3 * nf,
                         k + 4 * nf, k + 5 * nf, k + 6 * nf, k + 7 * nf);
    }
    memcpy(lam0, lam, sizeof(cplx) * nf);
    int finite = 1;
    for (size_t i = 0; i < nf; ++i)
        if (!isfinite(creal(lam[i])) || !isfinite(cimag(lam[i]))) finite = 0;
    free(ck); free(seg); free(k); free(x); free(lam);
    ws_free(&ws);
    return finite ? 0 : -2;
}

/* --------------------------------- continuous adjoint RHS (reference) */

/* Restatement of adjoint_disp_mmf! (sensitivity_disp_mmf.jl:32-79) with its
 * helpers calc_lambda_dfKR1c_duc! (:141-145), calc_lambdac_dfK_duc! (:153-157),
 * calc_lambda_dfR2c_duc! (:93-101), calc_lambdac_dfR_duc! (:123-132),
 * calc_gamma_a_b! (:169-184) and calc_deltas! (:199-215).  tau = shock,
 * sigma1[k] = (-1)^k (:241-242).  ufwd is the forward state x(z). */
int orc_adjoint_rhs(const orc_fiber *f, double z, const double *lam_in, const double *ufwd, double *dlam_out)
{
    const int nt = f->nt, m = f->m;
    const size_t nf = (size_t)nt * m, ff = nf * (size_t)m;
    const cplx *lamt = (const cplx *)lam_in, *uf = (const cplx *)ufwd;
    const cplx *hr = (const cplx *)f->hr;
    cplx *dl = (cplx *)dlam_out;
    orc_fft_plan p;
    if (fft_plan_init(&p, nt)) { fft_plan_free(&p); return -1; }
    cplx *ep = malloc(sizeof(cplx) * nf), *em = malloc(sizeof(cplx) * nf);
    cplx *lw = malloc(sizeof(cplx) * nf), *lt = malloc(sizeof(cplx) * nf), *lwc = malloc(sizeof(cplx) * nf);
    cplx *uw = malloc(sizeof(cplx) * nf), *ut = malloc(sizeof(cplx) * nf), *utc = malloc(sizeof(cplx) * nf);
    double *d1 = malloc(sizeof(double) * ff), *dr1 = malloc(sizeof(double) * ff), *dkr = malloc(sizeof(double) * ff);
    cplx *d2 = malloc(sizeof(cplx) * ff), *tmpm = malloc(sizeof(cplx) * ff);
    cplx *sr = malloc(sizeof(cplx) * nf), *t1 = malloc(sizeof(cplx) * nf), *t2 = malloc(sizeof(cplx) * nf);
    cplx *t3 = malloc(sizeof(cplx) * nf), *t4 = malloc(sizeof(cplx) * nf), *col = malloc(sizeof(cplx) * nt);
    int rc = 0;
    if (!ep || !em || !lw || !lt || !lwc || !uw || !ut || !utc || !d1 || !dr1 || !dkr || !d2 || !tmpm || !sr ||
        !t1 || !t2 || !t3 || !t4 || !col) {
        rc = -1;
        goto done;
    }
    for (size_t o = 0; o < nf; ++o) {
        ep[o] = cis(f->disp[o] * z);
        em[o] = cis(-f->disp[o] * z);
    }
    /* :43-47 */
    for (int k = 0; k < m; ++k)
        for (int t = 0; t < nt; ++t) {
            size_t o = (size_t)k * nt + t;
            lw[o] = ep[o] * lamt[o] * f->shock[t];
            lt[o] = lw[o];
            lwc[o] = conj(lw[o]);
        }
    for (int k = 0; k < m; ++k) {
        fft_exec(&p, lt + (size_t)k * nt, 0);
        fft_exec(&p, lwc + (size_t)k * nt, 1);
    }
    /* :50-54 */
    for (size_t o = 0; o < nf; ++o) {
        uw[o] = ep[o] * uf[o];
        ut[o] = uw[o];
    }
    for (int k = 0; k < m; ++k) fft_exec(&p, ut + (size_t)k * nt, 0);
    for (size_t o = 0; o < nf; ++o) utc[o] = conj(ut[o]);
    for (int k = 0; k < m; ++k) fft_exec(&p, uw + (size_t)k * nt, 1);
    /* calc_deltas! :199-215: d1[t,i,j] = sum_kl (v_k v_l + w_k w_l) gamma[l,k,i,j]
     * d2 = sum_kl (v_k v_l - w_k w_l + 2i v_k w_l) gamma[l,k,i,j] */
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            for (int t = 0; t < nt; ++t) {
                double s1 = 0, s2r = 0, s2i = 0;
                for (int k = 0; k < m; ++k)
                    for (int l = 0; l < m; ++l) {
                        double vk = creal(ut[(size_t)k * nt + t]), wk = cimag(ut[(size_t)k * nt + t]);
                        double vl = creal(ut[(size_t)l * nt + t]), wl = cimag(ut[(size_t)l * nt + t]);
                        double g = G4(f, l, k, i, j);
                        s1 += (vk * vl + wk * wl) * g;
                        s2r += (vk * vl - wk * wl) * g;
                        s2i += (2 * vk * wl) * g;
                    }
                size_t o = ((size_t)i + (size_t)m * j) * nt + t;
                d1[o] = s1;
                d2[o] = s2r + I * s2i;
            }
    /* :60-69 */
    for (int q = 0; q < m * m; ++q) {
        for (int t = 0; t < nt; ++t) col[t] = d1[(size_t)q * nt + t];
        fft_exec(&p, col, 0);
        for (int t = 0; t < nt; ++t) col[t] *= hr[t];
        fft_exec(&p, col, 1);
        for (int t = 0; t < nt; ++t) dr1[(size_t)q * nt + t] = creal(col[(t + nt / 2) % nt]);
        for (int t = 0; t < nt; ++t)
            dkr[(size_t)q * nt + t] = 2 * f->one_m_fr * d1[(size_t)q * nt + t] + dr1[(size_t)q * nt + t];
    }
    /* term 1 :141-145: sum_res[t,j] = lt[t,i] dkr[t,i,j]; ifft; 1i*em */
    for (int j = 0; j < m; ++j)
        for (int t = 0; t < nt; ++t) {
            cplx s = 0;
            for (int i = 0; i < m; ++i) s += lt[(size_t)i * nt + t] * dkr[((size_t)i + (size_t)m * j) * nt + t];
            sr[(size_t)j * nt + t] = s;
        }
    for (int k = 0; k < m; ++k) fft_exec(&p, sr + (size_t)k * nt, 1);
    for (size_t o = 0; o < nf; ++o) t1[o] = I * em[o] * sr[o];
    /* term 2 :153-157: sum_res[t,j] = lwc[t,i] d2[t,i,j]; ifft; -1i*Nt*em */
    for (int j = 0; j < m; ++j)
        for (int t = 0; t < nt; ++t) {
            cplx s = 0;
            for (int i = 0; i < m; ++i) s += lwc[(size_t)i * nt + t] * d2[((size_t)i + (size_t)m * j) * nt + t];
            sr[(size_t)j * nt + t] = s;
        }
    for (int k = 0; k < m; ++k) fft_exec(&p, sr + (size_t)k * nt, 1);
    for (size_t o = 0; o < nf; ++o) t2[o] = -I * (double)nt * em[o] * sr[o];
    /* term 3 :93-101: gab = gamma(lt, utc); fft; *= conj(hR) sigma; fft; contract with ifft(uw) (uw already
     * ifft'd at :54); fft; 1i/Nt em */
    for (int pass = 0; pass < 2; ++pass) {
        const cplx *a = pass == 0 ? lt : lwc;
        const cplx *b = pass == 0 ? utc : ut;
        /* calc_gamma_a_b! :169-184: gab[t,i,j] = sum_lk (a_l b_k) gamma[l,k,i,j] */
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j)
                for (int t = 0; t < nt; ++t) {
                    double sre = 0, sim = 0;
                    for (int l = 0; l < m; ++l)
                        for (int k = 0; k < m; ++k) {
                            cplx ab = a[(size_t)l * nt + t] * b[(size_t)k * nt + t];
                            double g = G4(f, l, k, i, j);
                            sre += creal(ab) * g;
                            sim += cimag(ab) * g;
                        }
                    tmpm[((size_t)i + (size_t)m * j) * nt + t] = sre + I * sim;
                }
        for (int q = 0; q < m * m; ++q) {
            cplx *c = tmpm + (size_t)q * nt;
            if (pass == 0) {
                fft_exec(&p, c, 0);
                for (int t = 0; t < nt; ++t) c[t] *= conj(hr[t]) * ((t & 1) ? -1.0 : 1.0);
                fft_exec(&p, c, 0);
            } else {
                fft_exec(&p, c, 1);
                for (int t = 0; t < nt; ++t) c[t] *= hr[t] * ((t & 1) ? -1.0 : 1.0);
                fft_exec(&p, c, 0);
            }
        }
        const cplx *rhsf = pass == 0 ? uw : ut;
        for (int i = 0; i < m; ++i)
            for (int t = 0; t < nt; ++t) {
                cplx s = 0;
                for (int j = 0; j < m; ++j) s += tmpm[((size_t)i + (size_t)m * j) * nt + t] * rhsf[(size_t)j * nt + t];
                sr[(size_t)i * nt + t] = s;
            }
        for (int k = 0; k < m; ++k) fft_exec(&p, sr + (size_t)k * nt, pass == 0 ? 0 : 1);
        if (pass == 0)
            for (size_t o = 0; o < nf; ++o) t3[o] = I / (double)nt * em[o] * sr[o];
        else
            for (size_t o = 0; o < nf; ++o) t4[o] = -I * (double)nt * em[o] * sr[o];
    }
    /* :78 */
    for (size_t o = 0; o < nf; ++o) dl[o] = t1[o] + f->one_m_fr * t2[o] + t3[o] + t4[o];
done:
    free(ep); free(em); free(lw); free(lt); free(lwc); free(uw); free(ut); free(utc);
    free(d1); free(dr1); free(dkr); free(d2); free(tmpm); free(sr); free(t1); free(t2);
    free(t3); free(t4); free(col);
    fft_plan_free(&p);
    return rc;
}

/* ------------------------------------------------ adaptive Tsit5 (a3) */

/* Tsitouras 5(4) pair (Ch. Tsitouras, "Runge-Kutta pairs of order 5(4)
 * satisfying only the first column simplifying assumption", Comput. Math.
 * Appl. 62 (2011) 770-775): the Tsit5 method OrdinaryDiffEq 7.17 runs for
 * solve_disp_mmf (simulate_disp_mmf.jl:185-195) and
 * solve_adjoint_disp_mmf (sensitivity_disp_mmf.jl:315-329).  Row 7 of A is
 * the 5th-order weight vector b (first same as last); BT is b - bhat, the
 * embedded error weights. */
static const double TS_C[7] = {0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0};
static const double TS_A[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774}};
static const double TS_BT[7] = {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                                -0.1447110071732629, 0.5823571654525552, -0.45808210592918697,
                                0.015151515151515152};

int orc_tsit5_tableau(double *c, double *a, double *bt)
{
    memcpy(c, TS_C, sizeof(TS_C));
    memcpy(a, TS_A, sizeof(TS_A));
    memcpy(bt, TS_BT, sizeof(TS_BT));
    return 0;
}

/* Continuous extension of the pair (Tsitouras 2011, the free 4th-order
 * interpolant OrdinaryDiffEq evaluates for solution(t) between accepted
 * steps, which propagate(problem; saveat) samples through _lab_spectrum,
 * propagation.jl:34-36,106-108): y(z + th h) = y + h sum_i b_i(th) k_i with
 * b_i(th) = sum_p TS_R[i][p] th^(p+1).  b_i(1) equals row 7 of TS_A. */
static const double TS_R[7][4] = {
    {1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
    {0.0, 0.13169999999999998, -0.2234, 0.1017},
    {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
    {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
    {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
    {0.0, -27.896526289197286, 65.09189467479366, -34.87065786149661},
    {0.0, 1.5, -4.0, 2.5}};

static void dense_weights(double th, double *b)
{
    for (int i = 0; i < 7; ++i)
        b[i] = th * (TS_R[i][0] + th * (TS_R[i][1] + th * (TS_R[i][2] + th * TS_R[i][3])));
}

int orc_tsit5_dense_weights(double theta, double *b)
{
    dense_weights(theta, b);
    return 0;
}

/* Step-size control, restating OrdinaryDiffEq's defaults for a 5th-order
 * explicit pair: error norm sqrt(mean |e|^2) with e = utilde / (atol +
 * max(|u_prev|, |u_new|) rtol) per complex entry; PI controller with
 * beta1 = 7/(10*5), beta2 = 2/(5*5), safety 0.9, qmin 1/5, qmax 10,
 * qold initialised to 1e-4 and set to max(EEst, 1e-4) on acceptance; a
 * rejected step shrinks by min(1/qmin, EEst^beta1 / 0.9).  The starting step
 * is the Hairer-Norsett-Wanner estimate (Solving ODEs I, II.4) with exponent
 * 1/(order+1).  The constants below are shared with libfiberhip's kernels. */
#define TS_BETA1 (7.0 / 50.0)
#define TS_BETA2 (2.0 / 25.0)
#define TS_SAFETY 0.9
#define TS_QMIN 0.2
#define TS_QMAX 10.0
#define TS_QOLDINIT 1e-4

static double scaled_norm(const cplx *e, const cplx *u0, const cplx *u1, size_t n, double rtol, double atol,
                          double scale)
{
    double s = 0;
    for (size_t i = 0; i < n; ++i) {
        double au = cabs(u0[i]), bu = u1 ? cabs(u1[i]) : 0.0;
        double sk = atol + (au > bu ? au : bu) * rtol;
        double r = cabs(e[i]) * scale / sk;
        s += r * r;
    }
    return sqrt(s / (double)n);
}

/* One Tsit5 attempt from (z, y, k[0] = f(z, y)) with step h: fills k[1..6]
 * (k[6] = f(z+h, ynew)), ynew and returns EEst. */
static double tsit5_attempt(const orc_fiber *f, orc_ws *ws, double z, double h, const cplx *y, cplx *k,
                            cplx *ynew, cplx *tmp, double rtol, double atol)
{
    const size_t nf = (size_t)f->nt * f->m;
    for (int i = 1; i < 7; ++i) {
        cplx *dst = i == 6 ? ynew : tmp;
        for (size_t o = 0; o < nf; ++o) {
            cplx s = 0;
            for (int j = 0; j < i; ++j) s += TS_A[i][j] * k[(size_t)j * nf + o];
            dst[o] = y[o] + h * s;
        }
        rhs_eval(f, ws, z + TS_C[i] * h, dst, k + (size_t)i * nf);
    }
    for (size_t o = 0; o < nf; ++o) {
        cplx s = 0;
        for (int j = 0; j < 7; ++j) s += TS_BT[j] * k[(size_t)j * nf + o];
        tmp[o] = s;
    }
    return scaled_norm(tmp, y, ynew, nf, rtol, atol, fabs(h));
}

/* Lab-frame sample at a save position: cis(D zj) times the state at zj, which
 * is y (zj = z), ynew (zj = the step's end) or the continuous extension
 * y + h sum_i b_i(th) k_i in between (ode_interpolation returns stored
 * states at step boundaries and interpolates inside a step). */
static void save_sample(const orc_fiber *f, double zj, double z, double h, double zend, const cplx *y,
                        const cplx *ynew, const cplx *k, cplx *dst)
{
    const size_t nf = (size_t)f->nt * f->m;
    double b[7];
    const int inside = zj != z && zj != zend;
    if (inside) dense_weights((zj - z) / h, b);
    for (size_t o = 0; o < nf; ++o) {
        cplx val;
        if (zj == z) {
            val = y[o];
        } else if (!inside) {
            val = ynew[o];
        } else {
            cplx s = 0;
            for (int i = 0; i < 7; ++i) s += b[i] * k[(size_t)i * nf + o];
            val = y[o] + h * s;
        }
        dst[o] = cis(f->disp[o] * zj) * val;
    }
}

/* Adaptive Tsit5 from z = 0 to L (solve_disp_mmf with reltol/abstol,
 * simulate_disp_mmf.jl:183-207).  out = cis(D L) y(L).  hs (may be NULL,
 * max_steps entries) receives the accepted step sizes; stats = {accepted,
 * rejected, rhs evaluations}.  With nz > 0, saved[j] (nz fields) receives
 * the lab-frame spectrum at zsave[j] (sorted, in [0, L]), the samples
 * propagate(problem; saveat) stores (propagation.jl:99-110).  Returns 0, -1
 * (bad args / memory), -2 (non-finite) or -3 (max_steps reached). */
int orc_tsit5_forward_save(const orc_fiber *f, double length_m, double rtol, double atol, int max_steps,
                           const double *u0, double *out, double *hs, int64_t *stats, int nz,
                           const double *zsave, double *saved)
{
    if (length_m <= 0 || rtol <= 0 || atol < 0 || max_steps <= 0 || nz < 0 || (nz > 0 && (!zsave || !saved)))
        return -1;
    for (int j = 0; j < nz; ++j)
        if (!(zsave[j] >= 0 && zsave[j] <= length_m) || (j > 0 && !(zsave[j] > zsave[j - 1]))) return -1;
    const size_t nf = (size_t)f->nt * f->m;
    orc_ws ws;
    if (ws_init(&ws, f->nt, f->m)) return -1;
    cplx *k = malloc(sizeof(cplx) * nf * 7), *y = malloc(sizeof(cplx) * nf);
    cplx *ynew = malloc(sizeof(cplx) * nf), *tmp = malloc(sizeof(cplx) * nf);
    if (!k || !y || !ynew || !tmp) { free(k); free(y); free(ynew); free(tmp); ws_free(&ws); return -1; }
    memcpy(y, u0, sizeof(cplx) * nf);
    int64_t nacc = 0, nrej = 0, nfe = 0;
    int rc = 0, js = 0;
    cplx *sv = (cplx *)saved;
    while (js < nz && zsave[js] == 0.0) {
        save_sample(f, 0.0, 0.0, 0.0, 0.0, y, y, NULL, sv + (size_t)js * nf);
        ++js;
    }
    /* starting step */
    rhs_eval(f, &ws, 0.0, y, k);
    nfe++;
    {
        double d0 = scaled_norm(y, y, NULL, nf, rtol, atol, 1.0);
        double d1 = scaled_norm(k, y, NULL, nf, rtol, atol, 1.0);
        double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
        if (dt0 > length_m) dt0 = length_m;
        for (size_t o = 0; o < nf; ++o) tmp[o] = y[o] + dt0 * k[o];
        rhs_eval(f, &ws, dt0, tmp, ynew);
        nfe++;
        for (size_t o = 0; o < nf; ++o) tmp[o] = ynew[o] - k[o];
        double d2 = scaled_norm(tmp, y, NULL, nf, rtol, atol, 1.0) / dt0;
        double mx = d1 > d2 ? d1 : d2;
        double dt1 = mx <= 1e-15 ? (dt0 * 1e-3 > 1e-6 ? dt0 * 1e-3 : 1e-6) : pow(0.01 / mx, 1.0 / 6.0);
        double h = 100 * dt0;
        if (dt1 < h) h = dt1;
        if (length_m < h) h = length_m;
        double z = 0.0, qold = TS_QOLDINIT;
        int done = 0;
        while (!done) {
            if (nacc >= max_steps) { rc = -3; break; }
            int last = 0;
            if (z + h >= length_m) { h = length_m - z; last = 1; }
            double e = tsit5_attempt(f, &ws, z, h, y, k, ynew, tmp, rtol, atol);
            nfe += 6;
            if (!isfinite(e)) { rc = -2; break; }
            double q11 = pow(e, TS_BETA1);
            if (e <= 1.0) {
                double q = q11 / pow(qold, TS_BETA2);
                q = q / TS_SAFETY;
                if (q < 1.0 / TS_QMAX) q = 1.0 / TS_QMAX;
                if (q > 1.0 / TS_QMIN) q = 1.0 / TS_QMIN;
                qold = e > TS_QOLDINIT ? e : TS_QOLDINIT;
                if (hs) hs[nacc] = h;
                const double zend = last ? length_m : z + h;
                while (js < nz && zsave[js] <= zend) {
                    save_sample(f, zsave[js], z, h, zend, y, ynew, k, sv + (size_t)js * nf);
                    ++js;
                }
                nacc++;
                z += h;
                memcpy(y, ynew, sizeof(cplx) * nf);
                memcpy(k, k + 6 * nf, sizeof(cplx) * nf);
                h = h / q;
                done = last;
            } else {
                nrej++;
                double sh = q11 / TS_SAFETY;
                if (sh > 1.0 / TS_QMIN) sh = 1.0 / TS_QMIN;
                h = h / sh;
            }
        }
    }
    if (stats) { stats[0] = nacc; stats[1] = nrej; stats[2] = nfe; }
    if (rc == 0) {
        cplx *o = (cplx *)out;
        for (size_t i = 0; i < nf; ++i) {
            o[i] = cis(f->disp[i] * length_m) * y[i];
            if (!isfinite(creal(o[i])) || !isfinite(cimag(o[i]))) rc = -2;
        }
    }
    free(k); free(y); free(ynew); free(tmp);
    ws_free(&ws);
    return rc;
}

int orc_tsit5_forward(const orc_fiber *f, double length_m, double rtol, double atol, int max_steps,
                      const double *u0, double *out, double *hs, int64_t *stats)
{
    return orc_tsit5_forward_save(f, length_m, rtol, atol, max_steps, u0, out, hs, stats, 0, NULL, NULL);
}

/* ------------------------------------------------ pump + signal with gain */

/* Gain model of disp_gain_smf! (simulate_disp_gain_mmf.jl:21-63) with the
 * rate-equation gain of calculate_gain_YDFA (gain.jl:150-166), its inputs
 * folded into per-bin and scalar coefficients by the caller:
 *   W12 = sum_k w12[k] |u_k|^2,  W21 = sum_k w21[k] |u_k|^2   (gain.jl:154-155
 *         with Ps = fftshift(|u Nt dt|^2) rep, psd_from_uω :115-120)
 *   R12 = r12 Pp, R21 = r21 Pp                                  (:151-152)
 *   n2 = (R12 + W12) / (R12 + R21 + W12 + W21 + inv_tau), n1 = 1 - n2  (:157-158)
 *   g_k = ge[k] n2 - ga[k] n1   (Gamma_s (sigma_es n2 - sigma_as n1) rho, :160)
 *   gP  = gpe n2 - gpa n1       (Gamma_p (sigma_ep n2 - sigma_ap n1) rho, :162)
 * Arrays in FFT order (the reference computes on the shifted grid and
 * ifftshifts the gain back, simulate_disp_gain_mmf.jl:98-100).  The
 * constant model (compute_gain(uω, pGain::Number, Pp), :76-80: g = pGain,
 * gP = -1) is ga = -pGain, ge = 0, gpa = 1, gpe = 0, every rate 0 and
 * inv_tau = 1 (n2 = 0). */
typedef struct {
    const double *w12, *w21, *ga, *ge;
    double r12, r21, inv_tau, gpa, gpe;
} orc_gain;

/* disp_gain_smf! (simulate_disp_gain_mmf.jl:21-63) for M = 1: state
 * (Pp, x).  dx = disp_mmf!(x) + g/2 .* x (:62), dPp = gP Pp (:61).  |u|^2 of
 * the lab spectrum cis(Dz) x equals |x|^2. */
static void gain_rhs_eval(const orc_fiber *f, const orc_gain *g, orc_ws *ws, double z, cplx pp, const cplx *x,
                          cplx *dpp, cplx *dx)
{
    const int nt = f->nt;
    rhs_eval(f, ws, z, x, dx);
    double w12 = 0, w21 = 0;
    for (int k = 0; k < nt; ++k) {
        const double a2 = creal(x[k]) * creal(x[k]) + cimag(x[k]) * cimag(x[k]);
        w12 += g->w12[k] * a2;
        w21 += g->w21[k] * a2;
    }
    const cplx r12 = g->r12 * pp, r21 = g->r21 * pp;
    const cplx n2 = (r12 + w12) / (r12 + r21 + w12 + w21 + g->inv_tau), n1 = 1.0 - n2;
    for (int k = 0; k < nt; ++k) dx[k] += 0.5 * (g->ge[k] * n2 - g->ga[k] * n1) * x[k];
    *dpp = (g->gpe * n2 - g->gpa * n1) * pp;
}

/* sum over the state (Pp, x) of |e s / (atol + max(|u0|, |u1|) rtol)|^2 */
static double gain_scaled_sq(cplx e, cplx u0, cplx u1, double rtol, double atol, double scale)
{
    const double au = cabs(u0), bu = cabs(u1);
    const double r = cabs(e) * scale / (atol + (au > bu ? au : bu) * rtol);
    return r * r;
}

/* solve_disp_gain_smf (simulate_disp_gain_mmf.jl:200-239): Tsit5 on the
 * state [Pp; x] of nt + 1 entries from 0 to L, with the controller and
 * starting step of orc_tsit5_forward_save and the RMS norm over all nt + 1
 * entries (OrdinaryDiffEq's norm of the concatenated state, :212-213).
 * out = cis(D L) x(L), *pump_out = real(Pp(L)).  Returns as
 * orc_tsit5_forward_save. */
int orc_gain_forward(const orc_fiber *f, const orc_gain *g, double length_m, double rtol, double atol,
                     int max_steps, double pump0, const double *u0, double *out, double *pump_out, double *hs,
                     int64_t *stats)
{
    if (f->m != 1 || length_m <= 0 || rtol <= 0 || atol < 0 || max_steps <= 0) return -1;
    const size_t nf = (size_t)f->nt;
    const double nn = (double)(nf + 1);
    orc_ws ws;
    if (ws_init(&ws, f->nt, 1)) return -1;
    cplx *k = malloc(sizeof(cplx) * nf * 7), *y = malloc(sizeof(cplx) * nf);
    cplx *ynew = malloc(sizeof(cplx) * nf), *tmp = malloc(sizeof(cplx) * nf);
    if (!k || !y || !ynew || !tmp) { free(k); free(y); free(ynew); free(tmp); ws_free(&ws); return -1; }
    memcpy(y, u0, sizeof(cplx) * nf);
    cplx p = pump0, pnew = 0, kp[7];
    int64_t nacc = 0, nrej = 0, nfe = 0;
    int rc = 0;
    gain_rhs_eval(f, g, &ws, 0.0, p, y, &kp[0], k);
    nfe++;
    double s0 = gain_scaled_sq(p, p, 0, rtol, atol, 1.0), s1 = gain_scaled_sq(kp[0], p, 0, rtol, atol, 1.0);
    for (size_t o = 0; o < nf; ++o) {
        s0 += gain_scaled_sq(y[o], y[o], 0, rtol, atol, 1.0);
        s1 += gain_scaled_sq(k[o], y[o], 0, rtol, atol, 1.0);
    }
    const double d0 = sqrt(s0 / nn), d1 = sqrt(s1 / nn);
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    if (dt0 > length_m) dt0 = length_m;
    for (size_t o = 0; o < nf; ++o) tmp[o] = y[o] + dt0 * k[o];
    cplx p1 = p + dt0 * kp[0], kp1;
    gain_rhs_eval(f, g, &ws, dt0, p1, tmp, &kp1, ynew);
    nfe++;
    double s2 = gain_scaled_sq(kp1 - kp[0], p, 0, rtol, atol, 1.0);
    for (size_t o = 0; o < nf; ++o) s2 += gain_scaled_sq(ynew[o] - k[o], y[o], 0, rtol, atol, 1.0);
    const double d2 = sqrt(s2 / nn) / dt0;
    const double mx = d1 > d2 ? d1 : d2;
    const double dt1 = mx <= 1e-15 ? (dt0 * 1e-3 > 1e-6 ? dt0 * 1e-3 : 1e-6) : pow(0.01 / mx, 1.0 / 6.0);
    double h = 100 * dt0;
    if (dt1 < h) h = dt1;
    if (length_m < h) h = length_m;
    double z = 0.0, qold = TS_QOLDINIT;
    int done = 0;
    while (!done) {
        if (nacc >= max_steps) { rc = -3; break; }
        int last = 0;
        if (z + h >= length_m) { h = length_m - z; last = 1; }
        for (int i = 1; i < 7; ++i) {
            cplx *dst = i == 6 ? ynew : tmp;
            cplx sp = 0;
            for (int j = 0; j < i; ++j) sp += TS_A[i][j] * kp[j];
            const cplx pin = p + h * sp;
            if (i == 6) pnew = pin;
            for (size_t o = 0; o < nf; ++o) {
                cplx s = 0;
                for (int j = 0; j < i; ++j) s += TS_A[i][j] * k[(size_t)j * nf + o];
                dst[o] = y[o] + h * s;
            }
            gain_rhs_eval(f, g, &ws, z + TS_C[i] * h, pin, dst, &kp[i], k + (size_t)i * nf);
        }
        nfe += 6;
        cplx ep = 0;
        for (int j = 0; j < 7; ++j) ep += TS_BT[j] * kp[j];
        double se = gain_scaled_sq(ep, p, pnew, rtol, atol, fabs(h));
        for (size_t o = 0; o < nf; ++o) {
            cplx s = 0;
            for (int j = 0; j < 7; ++j) s += TS_BT[j] * k[(size_t)j * nf + o];
            se += gain_scaled_sq(s, y[o], ynew[o], rtol, atol, fabs(h));
        }
        const double e = sqrt(se / nn);
        if (!isfinite(e)) { rc = -2; break; }
        const double q11 = pow(e, TS_BETA1);
        if (e <= 1.0) {
            double q = q11 / pow(qold, TS_BETA2);
            q = q / TS_SAFETY;
            if (q < 1.0 / TS_QMAX) q = 1.0 / TS_QMAX;
            if (q > 1.0 / TS_QMIN) q = 1.0 / TS_QMIN;
            qold = e > TS_QOLDINIT ? e : TS_QOLDINIT;
            if (hs) hs[nacc] = h;
            nacc++;
            z += h;
            memcpy(y, ynew, sizeof(cplx) * nf);
            memcpy(k, k + 6 * nf, sizeof(cplx) * nf);
            p = pnew;
            kp[0] = kp[6];
            h = h / q;
            done = last;
        } else {
            nrej++;
            double sh = q11 / TS_SAFETY;
            if (sh > 1.0 / TS_QMIN) sh = 1.0 / TS_QMIN;
            h = h / sh;
        }
    }
    if (stats) { stats[0] = nacc; stats[1] = nrej; stats[2] = nfe; }
    if (rc == 0) {
        cplx *o = (cplx *)out;
        for (size_t i = 0; i < nf; ++i) {
            o[i] = cis(f->disp[i] * length_m) * y[i];
            if (!isfinite(creal(o[i])) || !isfinite(cimag(o[i]))) rc = -2;
        }
        *pump_out = creal(p);
        if (!isfinite(*pump_out)) rc = -2;
    }
    free(k); free(y); free(ynew); free(tmp);
    ws_free(&ws);
    return rc;
}

/* One evaluation of the pump + signal right-hand side, for tests. */
int orc_gain_rhs(const orc_fiber *f, const orc_gain *g, double z, double pump, const double *x, double *dpump,
                 double *dx)
{
    orc_ws ws;
    if (f->m != 1 || ws_init(&ws, f->nt, 1)) return -1;
    cplx dp;
    gain_rhs_eval(f, g, &ws, z, pump, (const cplx *)x, &dp, (cplx *)dx);
    dpump[0] = creal(dp);
    dpump[1] = cimag(dp);
    ws_free(&ws);
    return 0;
}

/* One Tsit5 step with a given h (no error control), 6 stages; b = row 7 of A.
 * If ys != NULL the stage inputs Y_1..Y_6 are stored there. */
static void tsit5_step_fixed(const orc_fiber *f, orc_ws *ws, double z, double h, cplx *y, cplx *k, cplx *tmp,
                             cplx *ys)
{
    const size_t nf = (size_t)f->nt * f->m;
    if (ys) memcpy(ys, y, sizeof(cplx) * nf);
    rhs_eval(f, ws, z, y, k);
    for (int i = 1; i < 6; ++i) {
        cplx *dst = ys ? ys + (size_t)i * nf : tmp;
        for (size_t o = 0; o < nf; ++o) {
            cplx s = 0;
            for (int j = 0; j < i; ++j) s += TS_A[i][j] * k[(size_t)j * nf + o];
            dst[o] = y[o] + h * s;
        }
        rhs_eval(f, ws, z + TS_C[i] * h, dst, k + (size_t)i * nf);
    }
    for (size_t o = 0; o < nf; ++o) {
        cplx s = 0;
        for (int j = 0; j < 6; ++j) s += TS_A[6][j] * k[(size_t)j * nf + o];
        y[o] += h * s;
    }
}

/* Forward over a recorded step sequence hs[0..n-1] (z accumulated as in
 * orc_tsit5_forward); out = cis(D L) y, L = length_m. */
int orc_tsit5_replay(const orc_fiber *f, double length_m, const double *hs, int n, const double *u0, double *out)
{
    const size_t nf = (size_t)f->nt * f->m;
    orc_ws ws;
    if (n <= 0 || ws_init(&ws, f->nt, f->m)) return -1;
    cplx *k = malloc(sizeof(cplx) * nf * 6), *y = malloc(sizeof(cplx) * nf), *tmp = malloc(sizeof(cplx) * nf);
    if (!k || !y || !tmp) { free(k); free(y); free(tmp); ws_free(&ws); return -1; }
    memcpy(y, u0, sizeof(cplx) * nf);
    double z = 0.0;
    for (int s = 0; s < n; ++s) {
        tsit5_step_fixed(f, &ws, z, hs[s], y, k, tmp, NULL);
        z += hs[s];
    }
    cplx *o = (cplx *)out;
    int rc = 0;
    for (size_t i = 0; i < nf; ++i) {
        o[i] = cis(f->disp[i] * length_m) * y[i];
        if (!isfinite(creal(o[i])) || !isfinite(cimag(o[i]))) rc = -2;
    }
    free(k); free(y); free(tmp);
    ws_free(&ws);
    return rc;
}

/* Discrete adjoint of orc_tsit5_replay (the accepted-step map of the
 * adaptive solve, step sizes held fixed).  Stage cotangents, for i = 6..1:
 *   gk_i = h b_i lam' + h sum_{l>i} a_li gY_l,  gY_i = VJP_f(z + c_i h, Y_i)(gk_i)
 * and lam = lam' + sum_i gY_i.  Checkpoints every `every` steps. */
int orc_tsit5_adjoint(const orc_fiber *f, double length_m, const double *hs, int n, int every, const double *u0,
                      const double *lam_out, double *lam0)
{
    if (n <= 0 || every <= 0) return -1;
    const size_t nf = (size_t)f->nt * f->m;
    const int nseg = (n + every - 1) / every;
    orc_ws ws;
    if (ws_init(&ws, f->nt, f->m)) return -1;
    cplx *ck = malloc(sizeof(cplx) * nf * (size_t)nseg), *seg = malloc(sizeof(cplx) * nf * (size_t)every);
    cplx *k = malloc(sizeof(cplx) * nf * 6), *ys = malloc(sizeof(cplx) * nf * 6), *gy = malloc(sizeof(cplx) * nf * 6);
    cplx *y = malloc(sizeof(cplx) * nf), *lam = malloc(sizeof(cplx) * nf), *tmp = malloc(sizeof(cplx) * nf);
    cplx *gk = malloc(sizeof(cplx) * nf);
    double *zs = malloc(sizeof(double) * (size_t)n);
    int rc = 0;
    if (!ck || !seg || !k || !ys || !gy || !y || !lam || !tmp || !gk || !zs) { rc = -1; goto out; }
    memcpy(y, u0, sizeof(cplx) * nf);
    double z = 0.0;
    for (int s = 0; s < n; ++s) {
        zs[s] = z;
        if (s % every == 0) memcpy(ck + nf * (size_t)(s / every), y, sizeof(cplx) * nf);
        tsit5_step_fixed(f, &ws, z, hs[s], y, k, tmp, NULL);
        z += hs[s];
    }
    const cplx *lo = (const cplx *)lam_out;
    for (size_t i = 0; i < nf; ++i) lam[i] = cis(-f->disp[i] * length_m) * lo[i];
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * every, s1 = (s0 + every < n) ? s0 + every : n;
        memcpy(y, ck + nf * (size_t)sg, sizeof(cplx) * nf);
        for (int s = s0; s < s1; ++s) {
            memcpy(seg + nf * (size_t)(s - s0), y, sizeof(cplx) * nf);
            if (s + 1 < s1) tsit5_step_fixed(f, &ws, zs[s], hs[s], y, k, tmp, NULL);
        }
        for (int s = s1 - 1; s >= s0; --s) {
            const double h = hs[s];
            memcpy(y, seg + nf * (size_t)(s - s0), sizeof(cplx) * nf);
            tsit5_step_fixed(f, &ws, zs[s], h, y, k, tmp, ys);
            for (int i = 5; i >= 0; --i) {
                for (size_t o = 0; o < nf; ++o) {
                    cplx g = h * TS_A[6][i] * lam[o];
                    for (int l = i + 1; l < 6; ++l) g += h * TS_A[l][i] * gy[(size_t)l * nf + o];
                    gk[o] = g;
                }
                rhs_vjp(f, &ws, zs[s] + TS_C[i] * h, ys + (size_t)i * nf, gk, gy + (size_t)i * nf);
            }
            for (size_t o = 0; o < nf; ++o) {
                cplx acc = lam[o];
                for (int i = 0; i < 6; ++i) acc += gy[(size_t)i * nf + o];
                lam[o] = acc;
            }
        }
    }
    memcpy(lam0, lam, sizeof(cplx) * nf);
    for (size_t i = 0; i < nf; ++i)
        if (!isfinite(creal(lam[i])) || !isfinite(cimag(lam[i]))) rc = -2;
out:
    free(ck); free(seg); free(k); free(ys); free(gy); free(y); free(lam); free(tmp); free(gk); free(zs);
    ws_free(&ws);
    return rc;
}
