/*
 * fiberhip.h -- C ABI of libfiberhip, the MI355X (gfx950) GMMNLSE propagation
 * and adjoint engine behind FiberLab's propagation seam.
 *
 * The reference (Nacho2027/fiber-raman-suppression, Julia) has no FFI; the
 * replaceable calls are
 *   solve_disp_mmf(uw0, fiber, sim)          src/simulation/simulate_disp_mmf.jl:183-207
 *   solve_adjoint_disp_mmf(lwL, sol, f, s)   src/simulation/sensitivity_disp_mmf.jl:309-332
 *   _scalar_rk4_output(problem; steps)       src/fiberlab/scalar_crosscheck.jl:8-68
 * and the operator arrays they consume are built by
 *   get_disp_sim_params / get_disp_fiber_params_user_defined /
 *   _single_oscillator_raman_fields          src/helpers/helpers.jl:56-84,245-282,97-130
 *   get_p_disp_mmf (selfsteep)               src/simulation/simulate_disp_mmf.jl:87-114
 * A Julia caller binds the entry points below with `ccall`; arrays are
 * Julia-native column-major, so they pass without copies: a field of nt
 * samples and m modes is m*nt complex doubles stored as interleaved (re, im)
 * pairs, sample t of mode k at element k*nt + t, and a
 * batch of fields is batch such blocks back to back.
 *
 * All functions return FH_OK (0) or one of the FH_E* codes; the message of
 * the last failure on the calling thread is returned by fh_last_error().
 */
#ifndef FIBERHIP_H
#define FIBERHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes.  Each mirrors a check of the reference:
 *   FH_EINVAL     shape / finiteness / argument checks
 *                 (propagation.jl:25-32, physics_models.jl:459-577)
 *   FH_EGRID      non-positive absolute optical frequency on the grid
 *                 (helpers.jl:14-22: shock factor w/w0 must stay > 0)
 *   FH_ENOTSYM    adjoint needs a permutation-symmetric gamma
 *                 (sensitivity_disp_mmf.jl:103-112,239-240)
 *   FH_ENONFINITE solver produced NaN/Inf (scalar_crosscheck.jl:66,
 *                 propagation.jl:48-53)
 *   FH_ENOMEM     device allocation failed
 *   FH_EDEVICE    HIP runtime error
 *   FH_EUNSUPPORTED a size or mode count this build has no kernel for
 *   FH_EMAXSTEPS  adaptive solve hit max_steps accepted steps (the
 *                 MaxIters retcode that propagation.jl:48-53 rejects) */
enum {
    FH_OK = 0,
    FH_EINVAL = 1,
    FH_EGRID = 2,
    FH_ENOTSYM = 3,
    FH_ENONFINITE = 4,
    FH_ENOMEM = 5,
    FH_EDEVICE = 6,
    FH_EUNSUPPORTED = 7,
    FH_EMAXSTEPS = 8
};

/* Arithmetic of the bulk field work.  FH_PREC_FP32_BULK keeps fields, FFTs
 * and contractions in FP32 and computes the dispersion phase D*z (argument
 * reduction) and all norms/energies/gradient dot products in FP64. */
enum { FH_PREC_FP64 = 0, FH_PREC_FP32_BULK = 1 };

/* Integrator.  FH_SCHEME_RK4IP_FIXED is the global-frame interaction-picture
 * RK4 of _scalar_rk4_output (scalar_crosscheck.jl:53-60), z_i = i*h, with
 * `steps` steps.  FH_SCHEME_TSIT5 is the adaptive Tsit5 (Tsitouras 5(4)) of
 * solve_disp_mmf (simulate_disp_mmf.jl:183-207) with relative/absolute
 * tolerances rtol/atol, an FP64 RMS error norm over all nt*modes entries and
 * PI step control; every solve of a batch runs its own step sequence.  Both
 * schemes cover the single-mode grids and the multimode grids and mode
 * counts of the library (FH_EUNSUPPORTED otherwise). */
enum { FH_SCHEME_RK4IP_FIXED = 0, FH_SCHEME_TSIT5 = 1 };

/* Terminal objective of a gradient without a caller seed (all use the plan's
 * band mask; each returns J and its Wirtinger seed dJ/dconj(uL)):
 *   FH_OBJ_BAND_SUM     sum_m E_band_m / sum_m E_m  (_field_spectral_band_cost,
 *                       physics_helpers.jl:222-232; mmf_cost_sum, mmf_cost.jl:49-62)
 *   FH_OBJ_FUNDAMENTAL  E_band_1 / E_1 of mode 1 (mmf_cost_fundamental, mmf_cost.jl:81-98)
 *   FH_OBJ_WORST_MODE   log-sum-exp proxy of max_m E_band_m / E_m with
 *                       temperature objective_tau (mmf_cost_worst_mode,
 *                       mmf_cost.jl:124-166)
 * For one mode the three coincide. */
enum { FH_OBJ_BAND_SUM = 0, FH_OBJ_FUNDAMENTAL = 1, FH_OBJ_WORST_MODE = 2 };

typedef struct fh_context_s *fh_context;
typedef struct fh_plan_s *fh_plan;

/* Resolved problem: what solve_disp_mmf reads from its fiber/sim dicts. */
typedef struct {
    int32_t nt;          /* samples, power of two                           */
    int32_t modes;       /* M                                               */
    int32_t precision;   /* FH_PREC_*                                       */
    int32_t scheme;      /* FH_SCHEME_*                                     */
    int32_t steps;       /* fixed-step count                                */
    int32_t max_batch;   /* largest batch passed to one call                */
    double length_m;     /* fiber["L"]                                      */
    const double *dispersion; /* fiber["Dw"], nt*modes, rad/m, FFT order    */
    const double *gamma;      /* fiber["gamma"], modes^4, column-major      */
    const double *raman;      /* fiber["hRw"], nt complex (2*nt doubles)    */
    const double *shock;      /* fftshift(sim["ws"]/sim["w0"]), nt          */
    double one_m_fr;          /* fiber["one_m_fR"]                          */
    int32_t checkpoint_every; /* adjoint: forward states kept every K steps (0 = 50) */
    int32_t log_cost;         /* band objective in dB (_maybe_log_cost)      */
    const uint8_t *band_mask; /* nt bins of the band objective, or NULL      */
    double rtol;              /* FH_SCHEME_TSIT5: relative tolerance (reltol) */
    double atol;              /* FH_SCHEME_TSIT5: absolute tolerance (abstol) */
    int32_t max_steps;        /* FH_SCHEME_TSIT5: accepted-step limit (0 = 10000) */
    int32_t tape_segments;    /* adjoint: K checkpoint segments keep the stage tape of
                               * the forward sweep, so the backward sweep recomputes only
                               * the others: the last K of a fixed-step solve, the first
                               * K accepted-step segments of an adaptive one (0 = none;
                               * clamped to the segment count of steps or max_steps).
                               * Device memory grows by K *
                               * checkpoint_every * stages (4 fixed, 6 adaptive) * the
                               * stage's tape (single mode: nt complex + nt real;
                               * multimode: (M + ceil(Q/2)) * nt complex) per solve of
                               * max_batch.  Results are bit-identical for every K. */
    int32_t objective;        /* FH_OBJ_* of gradients without a caller seed     */
    double objective_tau;     /* FH_OBJ_WORST_MODE temperature (0 = 50)          */
} fh_problem_desc;

/* Per-solve device regions of a gradient (checkpoints, each stage tape) are
 * addressed with 32-bit offsets, so each must stay below 4 GiB per solve;
 * plans whose gradient would exceed that are rejected with FH_EINVAL at the
 * first gradient call (or by fh_plan_check). */

/* Counters of one call, summed over its batch (per-solve values: fh_solve_stats). */
typedef struct {
    int64_t accepted;   /* accepted steps (sol.destats.naccept)            */
    int64_t rejected;   /* rejected steps                                  */
    int64_t rhs_evals;  /* right-hand-side evaluations (sol.destats.nf)    */
    int32_t retcode;    /* FH_OK or the failure code                       */
    int32_t reserved;
} fh_stats;

/* A context binds one HIP device and one stream (default priority). */
int fh_context_create(int device, fh_context *out);
int fh_context_destroy(fh_context ctx);

/* Validates the problem, copies the operators to device memory in the plan's
 * precision and sizes the scratch for max_batch solves.  The caller keeps
 * ownership of every array in the descriptor. */
int fh_plan_create(fh_context ctx, const fh_problem_desc *desc, fh_plan *out);
int fh_plan_destroy(fh_plan plan);

/* Device bytes the plan holds (operators plus scratch for max_batch). */
int fh_plan_device_bytes(fh_plan plan, int64_t *bytes);

/* Validates a descriptor on the host alone (no device is touched) and returns
 * the device bytes a plan of it holds: *bytes for forward use, and with
 * `gradient` != 0 including the adjoint buffers of its first gradient call
 * (checkpoints, stage tape, adjoint fields).  Returns the error fh_plan_create
 * (or, with `gradient`, the first fh_gradient) would return.  The server uses
 * it for admission by memory footprint. */
int fh_plan_check(const fh_problem_desc *desc, int32_t gradient, int64_t *bytes);

/* Forward solve of `batch` independent launches sharing the plan's fiber.
 * u0: batch*modes*nt complex128 launch spectra (interaction picture at z=0,
 * i.e. problem.uw0); uL: the lab-frame output spectra cis(D L) .* x(L), the
 * output of _scalar_rk4_output / _lab_spectrum (propagation.jl:34-36).
 * Host pointers; blocks until done. */
int fh_forward(fh_plan plan, int32_t batch, const double *u0, double *uL, fh_stats *stats);

/* fh_forward plus dense output: the lab-frame spectra cis(D z) y(z) at the
 * nz sorted, unique positions zsave (0 <= z <= L), which
 * propagate(problem; saveat) stores (propagation.jl:89-123, samples taken by
 * _lab_spectrum from the ODE solution, :34-36).  Positions on an accepted
 * step boundary return the stored state; positions inside a step use the
 * solver's 4th-order continuous extension.  saved: batch*nz*modes*nt
 * complex128, ordered [batch][z][mode][t].  Adaptive (FH_SCHEME_TSIT5)
 * plans only, any supported mode count (FH_EUNSUPPORTED otherwise). */
int fh_forward_saveat(fh_plan plan, int32_t batch, const double *u0, double *uL, int32_t nz, const double *zsave,
                      double *saved, fh_stats *stats);

/* Gain model of the pump + signal propagation (disp_gain_smf!,
 * simulate_disp_gain_mmf.jl:21-63, with calculate_gain_YDFA,
 * gain.jl:148-167), folded into coefficients on |u_k|^2 of the field in FFT
 * order (nt entries each):
 *   W12 = sum_k w12[k] |u_k|^2,  W21 = sum_k w21[k] |u_k|^2
 *   R12 = r12 Pp,  R21 = r21 Pp
 *   n2 = (R12 + W12) / (R12 + R21 + W12 + W21 + inv_tau),  n1 = 1 - n2
 *   gain g_k = ge[k] n2 - ga[k] n1 [1/m],  pump gain gP = gpe n2 - gpa n1.
 * The constant model (compute_gain(uω, pGain::Number, Pp), :76-80) is
 * ga = -pGain, ge = w12 = w21 = 0, r12 = r21 = 0, inv_tau = 1, gpa = 1,
 * gpe = 0.  Host arrays, read during the call. */
typedef struct {
    const double *w12, *w21, *ga, *ge;
    double r12, r21, inv_tau, gpa, gpe;
} fh_gain_desc;

/* solve_disp_gain_smf (simulate_disp_gain_mmf.jl:200-239): adaptive Tsit5
 * on the state [Pp; x] (nt + 1 entries, the error norm over all of them)
 * with dx = disp_mmf!(x) + g/2 .* x and dPp = gP Pp, from 0 to the plan's
 * length, for `batch` launches.  Single-mode FH_SCHEME_TSIT5 plans with
 * 256 <= nt <= 16384 (FH_EUNSUPPORTED otherwise); the plan's rtol/atol
 * (the reference uses reltol 1e-5 and the default abstol 1e-6).  pump0,
 * pumpL: batch pump powers [W] at 0 and L; u0, uL as in fh_forward. */
int fh_forward_gain(fh_plan plan, int32_t batch, const fh_gain_desc *gain, const double *pump0, const double *u0,
                    double *uL, double *pumpL, fh_stats *stats);

/* Same solve on device-resident arrays in the plan's precision (complex128
 * for FH_PREC_FP64, complex64 for FH_PREC_FP32_BULK), enqueued on `stream`
 * (a hipStream_t; NULL = the context stream).  Does not synchronise. */
int fh_forward_device(fh_plan plan, int32_t batch, const void *d_u0, void *d_uL, void *stream);

/* Gradient of a terminal objective through the discrete adjoint of the
 * plan's integrator -- the fixed-step RK4IP map, or the adaptive solve's
 * accepted-step map with its step sizes held fixed -- (the role of
 * solve_adjoint_disp_mmf plus the chain
 * rule of spectral_shaper_model, sensitivity_disp_mmf.jl:309-332,
 * physics_models.jl:1042-1046).  For each of `batch` launches u0 (host,
 * complex128, the already-shaped spectrum alpha*A*cis(phi)*uw0):
 *   seed != NULL : lamL = seed, the caller's Wirtinger terminal adjoint
 *                  dJ/dconj(uL) (terminal_adjoint, adjoints.jl:723-739);
 *   seed == NULL : the band objective J = E_band/E_total of the plan's
 *                  band_mask and its seed uL (mask - J)/E_total
 *                  (_field_spectral_band_cost, physics_helpers.jl:222-232),
 *                  in dB when the plan's log_cost is set.
 * Outputs (any may be NULL): uL (lab-frame output), lam0 = dJ/dconj(u0),
 * grad = 2 Re(conj(lam0) i u0) per bin summed over modes (the FullGridPhase
 * gradient, accumulated in FP64), cost (J per launch; NaN with a seed).
 * Forward states are checkpointed every checkpoint_every steps in device
 * memory and each segment is recomputed once during the backward sweep,
 * except the tape_segments segments whose stage tape the forward sweep
 * keeps. */
int fh_gradient(fh_plan plan, int32_t batch, const double *u0, const double *seed, double *uL, double *lam0,
                double *grad, double *cost, fh_stats *stats);

/* Same on device-resident arrays in the plan's precision (grad and cost are
 * double); enqueued on `stream` (NULL = context stream), no synchronisation. */
int fh_gradient_device(fh_plan plan, int32_t batch, const void *d_u0, const void *d_seed, void *d_uL,
                       void *d_lam0, double *d_grad, double *d_cost, void *stream);

/* fh_gradient_device that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The plan is created beforehand (plan creation allocates); the call itself
 * keeps every buffer beyond its inputs and outputs (workgroup slices, status,
 * checkpoints, stage tape, step records, a device objective's seed) in one
 * fixed 256 MiB array per GPU, held until the process ends, and clears what
 * must start at zero with a fill kernel.  A call whose solves need more
 * (fh_plan_check sizes the parts) fails with FH_EINVAL.  Calls on one GPU
 * share that array: calls that overlap on one GPU race, so keep them on one
 * stream or otherwise ordered.  Results equal fh_gradient_device's. */
int fh_gradient_device_noalloc(fh_plan plan, int32_t batch, const void *d_u0, const void *d_seed, void *d_uL,
                               void *d_lam0, double *d_grad, double *d_cost, void *stream);

/* Event-timed back-to-back gradient launches (see fh_time_forward_device). */
int fh_time_gradient_device(fh_plan plan, int32_t batch, const void *d_u0, void *d_uL, void *d_lam0,
                            double *d_grad, double *d_cost, int32_t launches, float *ms_out);

/* ---------------------------------------------------------- kept forward
 *
 * The reference's adjoint contract runs the forward solve once and the
 * adjoint later on the cached solution (_cached_forward! and the
 * spectral_shaper_model physical_gradient, physics_models.jl:799-819,
 * 1032-1047; solve_adjoint_disp_mmf on the forward ODESolution,
 * sensitivity_disp_mmf.jl:309-332).  A tape is that cached forward on the
 * device: the shaped launch, the output, the checkpoints, the kept stage tape
 * of tape_segments segments and (adaptive plans) the accepted steps.
 * fh_forward_tape + fh_adjoint with a seed give the bits of fh_gradient with
 * the same seed, with one forward sweep in total.  A tape belongs to its plan
 * (free it before the plan) and its calls run on the plan's stream. */

typedef struct fh_tape_s *fh_tape;

/* Forward solve of `batch` launches (host complex128, as fh_forward) that
 * keeps the tape; uL (may be NULL) receives the lab-frame outputs. */
int fh_forward_tape(fh_plan plan, int32_t batch, const double *u0, double *uL, fh_tape *tape, fh_stats *stats);

/* Backward sweep over a tape.  seed: host complex128 [batch][modes][nt]
 * Wirtinger terminal seeds dJ/dconj(uL) (terminal_adjoint), or NULL for the
 * plan's objective (FH_OBJ_*, computed on the device from the kept output;
 * needs a band mask).  Outputs (each may be NULL), FP64 accumulation:
 *   lam0         [batch][modes][nt] complex128  dJ/dconj(u0)
 *   grad_phase   [batch][modes][nt]  2 Re(conj(lam0) i u0)  (phase gradient)
 *   grad_amp     [batch][modes][nt]  2 Re(conj(lam0) u0)    (amplitude gradient
 *                                    times the amplitude A of the launch)
 *   energy_dot   [batch]             Re sum conj(lam0) u0   (energy gradient
 *                                    times the launch energy)
 *   cost         [batch]             J of the plan's objective (NaN with a seed)
 * physics_models.jl:1042-1047 forms the three physical gradients from lam0
 * this way.  A tape may be swept more than once (e.g. with different seeds);
 * single-mode sweeps consume the kept tape, so a repeated single-mode sweep
 * first re-runs the forward sweep from the kept launch on the device. */
int fh_adjoint(fh_tape tape, const double *seed, double *lam0, double *grad_phase, double *grad_amp,
               double *energy_dot, double *cost, fh_stats *stats);

/* Device bytes the tape holds. */
int fh_tape_device_bytes(fh_tape tape, int64_t *bytes);

/* Releases a tape (its memory may be kept by the plan for the next tape). */
int fh_tape_free(fh_tape tape);

/* FP64 per-mode spectral energy E_m = sum_w |u(w,m)|^2 of `batch` device
 * fields in the plan's precision, written to host energies[batch*modes]
 * (the mode_energy of _field_metrics, propagation.jl:189-202, up to the
 * Parseval factor nt*dt). */
int fh_energy_device(fh_plan plan, int32_t batch, const void *d_u, double *energies);

/* FP64 evidence of `count` lab-frame spectra (host, complex128,
 * [count][mode][t], nt and modes of the plan): the per-sample quantities of
 * _field_metrics / _photon_number / _edge_fraction (propagation.jl:176-202)
 * that metrics() and verify() gate on.  omega: |omega_s| in FFT order
 * (rad/ps, fftshift(abs.(sim["ws"]))); dt_ps: sim["dt"].  out[count][modes+5]:
 * mode energies (pJ), total energy (pJ), peak power (W), temporal edge
 * fraction, spectral edge fraction, photon number.  The time-domain power
 * uses the plan's transform and precision; every sum is FP64. */
int fh_field_metrics(fh_plan plan, int32_t count, const double *spectra, const double *omega, double dt_ps,
                     double *out);

/* ------------------------------------------------------- GRIN modal setup */

/* Guided modes of the scalar wave equation on an nx x ny grid, FP64 on the
 * device (replaces solve_for_fiber_modes, src/simulation/fibers.jl:159-197):
 * the nmodes largest eigenpairs of A = k0^2 diag(eps) + (5-point Laplacian)/dx^2
 * with zero field outside the grid (eigs(A; nev = nmodes, which = :LR)), by
 * Chebyshev-filtered subspace iteration.  eps: [ny][nx] host, first grid index
 * fastest (Julia's column-major eps); dx and 1/k0 in the same length unit.
 * Outputs: beta2[nmodes] descending (the eigenvalues beta^2), modes
 * [nmodes][nx*ny] (unit 2-norm, same grid order, largest entry positive; within
 * a degenerate group the basis is arbitrary, as with eigs), iterations (may be
 * NULL).  tol: relative residual ||A v - beta^2 v|| / beta^2 (<= 0: 1e-10);
 * max_iter <= 0: 200.  nmodes <= 40.  FH_EMAXSTEPS, with the outputs written,
 * when the tolerance is not met. */
int fh_fd_modes(fh_context ctx, int32_t nx, int32_t ny, double dx, double k0, const double *eps, int32_t nmodes,
                double tol, int32_t max_iter, double *beta2, double *modes, int32_t *iterations);

/* Overlap tensor of `modes` real sampled modes, FP64 on the device (replaces
 * compute_overlap_tensor, src/simulation/fibers.jl:214-222):
 * sk[i + M (j + M (k + M l))] = scale * sum_r phi_i(r) phi_j(r) phi_k(r) phi_l(r)
 * for phi [modes][points] (host); scale = 1 / dx^2 there.  modes <= 16. */
int fh_overlap_tensor(fh_context ctx, int64_t points, int32_t modes, const double *phi, double scale, double *sk);

/* Device buffers on the context's device (plain hipMalloc / hipMemcpy on the
 * context stream).  Callers without their own HIP allocator use these to
 * keep fields resident between calls. */
int fh_device_alloc(fh_context ctx, int64_t bytes, void **ptr);
int fh_device_free(fh_context ctx, void *ptr);
int fh_copy_to_device(fh_context ctx, void *dst, const void *src, int64_t bytes);
int fh_copy_to_host(fh_context ctx, void *dst, const void *src, int64_t bytes);
int fh_synchronize(fh_context ctx);
/* Device-to-device copy by a plain streaming kernel (16-byte vector loads and
 * stores, no reuse); also the reference point for HBM byte counters. */
int fh_device_copy(fh_context ctx, void *dst, const void *src, int64_t bytes);

/* Enqueue `launches` back-to-back forward solves of `batch` device-resident
 * launches on the context stream, each bracketed by a hipEvent pair on that
 * stream, wait for them, and return each launch's duration in ms_out[i]. */
int fh_time_forward_device(fh_plan plan, int32_t batch, const void *d_u0, void *d_uL, int32_t launches,
                           float *ms_out);

/* Per-solve counters of the last fh_forward* / fh_gradient* call on this
 * plan (after it completed): stats[3*b + {0,1,2}] = accepted steps, rejected
 * steps, right-hand-side evaluations of solve b (sol.destats.naccept /
 * .nreject / .nf).  For the fixed-step scheme these are steps, 0, 4*steps
 * (the forward sweep of a gradient). */
int fh_solve_stats(fh_plan plan, int32_t batch, int64_t *stats);

/* Accepted step sizes of the last adaptive call on this plan:
 * hs[b*max_steps + i], i < accepted steps of solve b (FH_SCHEME_TSIT5). */
int fh_step_sizes(fh_plan plan, int32_t batch, double *hs);

/* ------------------------------------------------------------ serving
 *
 * A persistent request server on one context: independent forward or
 * forward+adjoint solves of uneven nt, mode count and fiber length are
 * submitted from any thread; a scheduler thread batches requests that share
 * their operators (same dispersion, gamma, Raman response, shock and band
 * mask) into one fused launch, runs up to `lanes` launches concurrently on
 * default-priority streams, and admits a launch only while the device memory
 * of the plans it needs stays within `memory_budget_bytes` (idle plans are
 * released least-recently-used first).  Each solve's result depends only on
 * its own inputs, so it is the same alone and inside any batch.
 * Single-mode requests run the adaptive Tsit5 scheme (config rtol/atol) with
 * their own length; multimode requests run the fixed-step RK4IP with
 * ceil(length / multimode_step_m) steps (at least 16); multimode gradients
 * are offered up to mm_grad_max_steps steps (their checkpoint capacity) for
 * fully permutation-symmetric gamma.  The budget is allocated once, at
 * creation, as one device arena the plans sub-allocate from, beside a
 * 256 MiB pinned host arena that stages plan tables.  Requests whose
 * estimated work exceeds about 1.5e12 flops occupy at most three quarters
 * of the lanes.  Lanes run on their own streams: give the process lanes + 4
 * hardware queues (GPU_MAX_HW_QUEUES, read when HIP initialises) within the
 * device's compute queues, or lanes sharing a queue run one after another.
 * On MI355X, launches of the 1024-thread FP32 kernels under many concurrent
 * lanes occasionally came back non-finite (per 10,000-request trace: 0-5
 * launches at 12 lanes, 2-3 at 16, about 15 at 20, at equal throughput).
 * Such a request, if its launch held others, is re-run once alone, and the
 * plan instance of that launch is released rather than reused.
 * No counterpart exists in the reference (its solves run one at a time in
 * the calling Julia task); callers are batch optimisation drivers and
 * services that issue many solves. */

typedef struct fh_server_s *fh_server;

typedef struct {
    int64_t memory_budget_bytes; /* device memory the server's plans may hold      */
    int32_t lanes;               /* concurrent launches (streams), 1..32           */
    int32_t max_batch;           /* solves per launch (at most; a class whose launch
                                  * would exceed a sixteenth of the budget gets fewer) */
    int32_t precision;           /* FH_PREC_* of every request                     */
    int32_t max_steps;           /* adaptive accepted-step limit (0 = 10000)       */
    double rtol, atol;           /* adaptive tolerances of single-mode requests    */
    double multimode_step_m;     /* fixed step of multimode requests               */
    int32_t mm_grad_max_steps;   /* step capacity of multimode gradients (0 = 2000)*/
    int32_t tape_segments;       /* single-mode gradient plans: accepted-step segments
                                  * of 50 whose stage tape the forward sweep keeps
                                  * (fh_problem_desc.tape_segments; 0 = none)      */
} fh_server_config;

enum { FH_REQ_FORWARD = 0, FH_REQ_GRADIENT = 1 };

/* One request.  Operators as in fh_problem_desc (host, copied at submit);
 * u0 (nt*modes complex128) is copied at submit.  Outputs are written at
 * completion and must stay valid until fh_server_wait returns: uL
 * (nt*modes complex128, lab frame), and for FH_REQ_GRADIENT the band
 * objective's grad (nt, FP64 phase gradient) and cost (1). */
typedef struct {
    int32_t kind;  /* FH_REQ_*                  */
    int32_t nt, modes;
    int32_t reserved;
    double length_m;
    const double *dispersion, *gamma, *raman, *shock;
    double one_m_fr;
    const uint8_t *band_mask; /* FH_REQ_GRADIENT: nt bins */
    const double *u0;
    double *uL, *grad, *cost;
} fh_request;

typedef struct {
    int64_t submit_ns, start_ns, finish_ns; /* steady-clock times of submission, launch, completion */
    int64_t accepted, rejected, rhs_evals;  /* this solve's counters                                */
    int32_t retcode;                        /* FH_OK or the solve's error code                      */
    int32_t batch;                          /* solves in the launch that carried it                 */
    int32_t attempts;                       /* launches the request took: a solve that returned     *
                                             * non-finite fields inside a batch is re-run once, alone */
    int32_t reserved;
} fh_request_result;

int fh_server_create(fh_context ctx, const fh_server_config *cfg, fh_server *out);
/* Validates and queues one request; returns its id.  Thread-safe. */
int fh_server_submit(fh_server srv, const fh_request *req, int64_t *id);
/* state: 0 queued, 1 running, 2 finished.  Thread-safe. */
int fh_server_poll(fh_server srv, int64_t id, int32_t *state);
/* Blocks until the request finished, fills `res` and forgets the id. */
int fh_server_wait(fh_server srv, int64_t id, fh_request_result *res);
/* Finishes every queued request, then stops the scheduler and frees the plans. */
int fh_server_destroy(fh_server srv);

/* Message of the last failure on this thread ("" if none). */
const char *fh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FIBERHIP_H */
