"""bench.py -- libfiberhip throughput on one MI355X.

A "step" is one launch of the fused RK4IP kernel over a batch of B
independent solves of the workload (synthetic launches: the workload's pulse
with seeded Taylor spectral phases, as in the candidate sweep).  Inputs are
resident in HBM before the timed region; the timed region is K launches
bracketed by synchronisations, timed with HIP events on the launch stream.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import pathlib
import sys
import threading
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "fiber-raman-suppression_amd"), str(ROOT)]

# The serving leg runs one launch per lane concurrently, each on its own
# stream.  HIP maps streams onto hardware queues per process (4 by default,
# and environments may pin that value), and work of streams sharing a queue
# runs one after another: a short request on a free lane waited 4.5 s behind
# a long launch with 4 queues and 4 lanes, and 1-2 ms with 20 queues and 16
# lanes.  The four extra queues keep the null and context streams apart.
# Many more active queues oversubscribe the hardware scheduler: with 28
# lanes on 32 queues a plan's table upload waited up to 160 s behind running
# launches.  HIP reads the value once, when it initialises, so main() sets it
# before any HIP call, overriding an inherited value.  On the 10,000-request
# trace, 12, 16 and 20 lanes give the same throughput, but launches of the
# 1024-thread FP32 kernels came back non-finite in 0-5, 2-3 and about 15
# launches per run.
SERVE_LANES = 12

import numpy as np  # noqa: E402

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense vector/matrix peaks
# Matched error of the headline metric: relative output spectral energy of the
# FP32-bulk solve against the FP64 solve of the same launch.
MATCHED_ENERGY_TOL = 1e-5
# Normwise bound on the FP32-bulk phase gradient against FP64 that the GPU tests
# assert on this workload at full size (tests/test_gpu_smf28_2m_gradient.py).
GRADIENT_NORMWISE_BOUND = 5e-4
# The GPU box's CPU share for one GPU: the all-core CPU baseline uses at most
# this many threads, whatever the affinity mask shows.
HOST_CORE_SHARE = 16
# HBM bytes per unit of work measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled x2.0 / x1.0 by a plain copy kernel of known size); the summaries and
# commands are in profiles/.
MEASURED_HBM_BYTES = {
    "k_grad_single<float,16384,1024>": (1.1217e10, "profiles/k_grad_single_smf28_16384.md"),
    "k_rk4ip_single<float,8192,512>": (1.3921e5, "profiles/k_rk4ip_single_smf28_8192.md"),
    "k_rk4ip_multi<float,4096,512,6>": (4.6848e10, "profiles/k_rk4ip_multi_grin6_4096.md"),
}


def rhs_flops(nt: int, m: int) -> float:
    """Algorithmic flops of one disp_mmf! evaluation (FFTs 5 N log2 N each,
    2M + 2M^2 of them; gamma contraction 2M^4 N; pointwise work)."""
    return 5 * nt * math.log2(nt) * (2 * m + 2 * m * m) + 2 * m ** 4 * nt + 17 * m * m * nt + 16 * m * nt


def forward_flops(nt: int, m: int, steps: int) -> float:
    """Fixed-step RK4IP solve: 4 RHS per step plus 28 MN for the stage updates."""
    return steps * (4 * rhs_flops(nt, m) + 28 * m * nt)


def adjoint_rhs_flops(nt: int, m: int) -> float:
    """Algorithmic flops of one adjoint_disp_mmf! evaluation (8M + 6M^2 FFTs,
    7 gamma contractions, pointwise work)."""
    return 5 * nt * math.log2(nt) * (8 * m + 6 * m * m) + 14 * m ** 4 * nt + 60 * m * m * nt + 40 * m * nt


def gradient_flops(nt: int, m: int, steps: int) -> float:
    """Forward solve + adjoint sweep + chain rule; checkpoint recomputation not counted."""
    return forward_flops(nt, m, steps) + steps * (4 * adjoint_rhs_flops(nt, m) + 28 * m * nt) + 8 * m * nt


def workload(name: str):
    import fiberhip

    if name == "smf28-fwd-8192":
        fib = fiberhip.Fiber(length_m=0.5, power_w=0.1, preset="SMF28", beta_order=3)
        problem = fiberhip.fiber_problem(fib, grid=fiberhip.Grid(nt=8192, time_window_ps=10.0),
                                         raman_threshold_thz=None)
        return problem, 200
    if name == "grin6-fwd-4096":
        from fiberhip import grin

        return grin.grin_problem(6), 4000
    if name == "smf28-sweep-8192":
        fib = fiberhip.Fiber(length_m=1.0, power_w=0.1, preset="SMF28", beta_order=3)
        problem = fiberhip.fiber_problem(fib, grid=fiberhip.Grid(nt=8192, time_window_ps=23.531, policy="exact"),
                                         raman_threshold_thz=None)
        return problem, 0
    if name == "smf28-grad-16384":
        fib = fiberhip.Fiber(length_m=2.0, power_w=0.2, preset="SMF28", beta_order=3)
        problem = fiberhip.fiber_problem(fib, grid=fiberhip.Grid(nt=16384, time_window_ps=47.061, policy="exact"),
                                         raman_threshold_thz=None)
        return problem, 2000
    raise SystemExit(f"unknown workload {name}")


def launches(problem, batch: int, seed: int = 3, orders=(2, 3, 4)) -> np.ndarray:
    """Launch spectra with seeded Taylor phases: GDD ~ U(-1e4, 1e4) fs^2,
    TOD ~ U(-5e5, 5e5) fs^3 and, when order 4 is asked for, FOD ~ U(-1e7, 1e7) fs^4."""
    import fiberhip

    rng = np.random.default_rng(seed)
    basis = fiberhip.taylor_phase_basis(problem, orders)
    span = {2: 1e4, 3: 5e5, 4: 1e7}
    coef = np.stack([rng.uniform(-span[o], span[o], batch) for o in orders], axis=1)
    phases = coef @ basis.T  # (batch, nt)
    return problem.uω0[None, :, 0] * np.exp(1j * phases)  # (batch, nt)


def time_path(problem, steps, precision, batch, k, warmup, u_host=None):
    """K event-timed launches (hipEvents on the launch stream, inside libfiberhip)."""
    import fiberhip

    plan = fiberhip.Plan.from_problem(problem, steps=steps, precision=precision, max_batch=batch)
    u0 = plan.to_device(launches(problem, batch)[:, :, None] if u_host is None else u_host)
    out = fiberhip.DeviceArray(plan.context, u0.nbytes)
    for _ in range(warmup):
        plan.forward_device(batch, u0.ptr, out.ptr)
    plan.synchronize()
    print(f"bench: {precision} warmup done, timing {k} launches of {batch} solves", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ms = plan.time_forward_device(batch, u0.ptr, out.ptr, k)
    wall = time.perf_counter() - t0
    energies = plan.energy_device(batch, out.ptr)
    return dict(plan=plan, wall_s=wall, kernel_ms=float(np.mean(ms)), kernel_ms_p50=float(np.median(ms)),
                energies=energies)


def grin_launches(batch: int) -> np.ndarray:
    """GRIN M=6 requests: seeded Dirichlet(1) mode weights and uniform phases per request."""
    from fiberhip import grin

    out = []
    for b in range(batch):
        out.append(grin.grin_problem(6, seed=1000 + b).uω0)
    return np.stack(out)


def kept_segments(nt, steps, every, precision, batch, budget_bytes):
    """Tape segments the gradient plan keeps within a device-memory budget:
    each kept segment holds 4*every stage slots of u (complex) and h (real)
    per solve, and one more region serves the recomputed segments."""
    esz = 8 if precision == "fp64" else 4
    per_segment = 4 * every * nt * 3 * esz * batch
    nseg = -(-steps // every)
    k = int(budget_bytes // per_segment) - 1
    return max(0, min(nseg, k))


def time_gradient(problem, steps, precision, batch, k, warmup, mask, tape_segments=0):
    """K event-timed launches of the fused gradient kernel (band objective on device)."""
    import fiberhip

    plan = fiberhip.Plan.from_problem(problem, steps=steps, precision=precision, max_batch=batch, band_mask=mask,
                                      checkpoint_every=50, tape_segments=tape_segments)
    u0 = plan.to_device(launches(problem, batch, seed=2, orders=(2, 3))[:, :, None])
    out = fiberhip.DeviceArray(plan.context, u0.nbytes)
    lam0 = fiberhip.DeviceArray(plan.context, u0.nbytes)
    grad = fiberhip.DeviceArray(plan.context, 8 * batch * problem.nt)
    cost = fiberhip.DeviceArray(plan.context, 8 * batch)
    for _ in range(warmup):
        plan.time_gradient_device(batch, u0.ptr, out.ptr, lam0.ptr, grad.ptr, cost.ptr, 1)
    print(f"bench: {precision} gradient warmup done, timing {k} launches of {batch}", file=sys.stderr, flush=True)
    ms = plan.time_gradient_device(batch, u0.ptr, out.ptr, lam0.ptr, grad.ptr, cost.ptr, k)
    energies = plan.energy_device(batch, out.ptr)
    res = dict(kernel_ms=float(np.mean(ms)), ms=ms, energies=energies,
               grad=grad.to_host(np.float64, (batch, problem.nt)), cost=cost.to_host(np.float64, (batch,)),
               device_bytes=plan.device_bytes)
    for a in (u0, out, lam0, grad, cost):
        a.free()
    plan.close()
    return res


def time_sweep(problem, precision, candidates, per_launch, k, warmup, mask, rtol=1e-6, atol=1e-6, tape_segments=0):
    """Adaptive candidate sweep: `candidates` forward + adjoint gradients of the band
    objective, launched `per_launch` at a time back to back on one stream; each
    candidate runs its own Tsit5 step sequence.  Returns the per-sweep time and
    the per-candidate step statistics of the last sweep."""
    import fiberhip

    plan = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=rtol, atol=atol, precision=precision,
                                      max_batch=per_launch, band_mask=mask, checkpoint_every=50, max_steps=4000,
                                      tape_segments=tape_segments)
    host = launches(problem, candidates, seed=4)[:, :, None]
    nl = candidates // per_launch
    u0 = [plan.to_device(host[i * per_launch:(i + 1) * per_launch]) for i in range(nl)]
    out = fiberhip.DeviceArray(plan.context, u0[0].nbytes)
    lam0 = fiberhip.DeviceArray(plan.context, u0[0].nbytes)
    grad = fiberhip.DeviceArray(plan.context, 8 * per_launch * problem.nt)
    cost = fiberhip.DeviceArray(plan.context, 8 * per_launch)
    for _ in range(warmup):
        for i in range(nl):
            plan.time_gradient_device(per_launch, u0[i].ptr, out.ptr, lam0.ptr, grad.ptr, cost.ptr, 1)
    sweeps, stats, costs = [], [], []
    for _ in range(k):
        t = 0.0
        for i in range(nl):
            t += float(plan.time_gradient_device(per_launch, u0[i].ptr, out.ptr, lam0.ptr, grad.ptr, cost.ptr, 1)[0])
            if len(stats) < nl:
                stats.append(plan.solve_stats(per_launch))
                costs.append(cost.to_host(np.float64, (per_launch,)))
        sweeps.append(t)
    return dict(ms=float(np.mean(sweeps)), stats=np.concatenate(stats), cost=np.concatenate(costs))


def gain_workload(nt: int = 4096):
    """YDFA amplifier (solve_disp_gain_smf): 1030 nm, Nt samples over 20 ps,
    1 m of Yb fiber with gamma 4.7e-3 /W/m, beta2 20 ps^2/km, fR 0.18, a
    300 fs Gaussian at 50 mW average (80 MHz) and a 0.5 W co-propagating pump
    at 976 nm; the reference's YDFA parameters (gain.jl:30-75) with its Yb
    cross-section tables (tests/golden/yb_cross_sections.json, data)."""
    from fiberhip import gain as G
    from fiberhip import problem as P

    xs = json.loads((ROOT / "tests" / "golden" / "yb_cross_sections.json").read_text())
    sim = P.get_disp_sim_params(1030e-9, 1, nt, 20.0 * nt / 4096, 3)
    cs = G.get_ydfa_cross_sections(sim["fs"], xs["absorption"]["wavelength_nm"], xs["absorption"]["intensity"],
                                   xs["emission"]["wavelength_nm"], xs["emission"]["intensity"], xs["scale_m2"])
    p = G.get_YDFAParams(sim, cs)
    _, u0 = P.get_initial_state([1.0], 0.05, 300e-15, 80e6, "gauss", sim)
    fib = P.get_disp_fiber_params_user_defined(1.0, sim, fR=0.18, gamma_user=4.7e-3, betas_user=[20e-27, 0.0])
    fib["gain_parameters"] = p
    return sim, fib, u0, 0.5


def time_gain(precision: str, batch: int, k: int, nt: int = 4096):
    """Wall time of fh_forward_gain over `batch` launches (seeded GDD/TOD
    phases on the signal), including host staging; returns per-call seconds
    and the per-solve step counts of the last call."""
    from fiberhip import gain as G

    sim, fib, u0, pump = gain_workload(nt)
    rng = np.random.default_rng(6)
    w = 2 * np.pi * np.fft.fftfreq(nt, sim["Δt"]) * 1e-3  # rad/fs
    launches_ = np.stack([u0 * np.exp(1j * (rng.uniform(-1e4, 1e4) * w ** 2 / 2
                                            + rng.uniform(-5e5, 5e5) * w ** 3 / 6))[:, None]
                          for _ in range(batch)])  # (batch, nt, 1)
    G.solve_disp_gain_smf(launches_, fib, sim, pump_power=pump, precision=precision)  # warm-up
    ts = []
    for _ in range(k):
        t = time.perf_counter()
        res = G.solve_disp_gain_smf(launches_, fib, sim, pump_power=pump, precision=precision)
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), res


SERVE_WINDOW_PS = {2048: 5.883, 4096: 11.766, 8192: 23.531, 16384: 47.061, 32768: 94.122}


def serving_trace(n: int, seed: int = 5):
    """Request mix: Nt uniform over 2^11..2^15 (carrier-margin windows), M uniform
    over {1, 3, 6, 10} (SMF-28 for M = 1, 50 um GRIN with HG modes otherwise),
    length log-uniform in [0.1, 5] m, peak power log-uniform in [1, 20] kW, launch
    phase from the GDD/TOD/FOD distributions, half forward and half forward+adjoint."""
    rng = np.random.default_rng(seed)
    nts = rng.choice(list(SERVE_WINDOW_PS), n)
    ms = rng.choice([1, 3, 6, 10], n)
    lengths = np.exp(rng.uniform(np.log(0.1), np.log(5.0), n))
    peaks = np.exp(rng.uniform(np.log(1e3), np.log(2e4), n))
    coef = np.stack([rng.uniform(-1e4, 1e4, n), rng.uniform(-5e5, 5e5, n), rng.uniform(-1e7, 1e7, n)], axis=1)
    kinds = np.where(rng.uniform(size=n) < 0.5, "forward", "gradient")
    return [dict(nt=int(a), modes=int(b), length=float(c), peak=float(d), coef=e, kind=str(k))
            for a, b, c, d, e, k in zip(nts, ms, lengths, peaks, coef, kinds)]


def serving_problem(cache: dict, nt: int, modes: int):
    import fiberhip

    key = (nt, modes)
    if key not in cache:
        if modes == 1:
            fib = fiberhip.Fiber(length_m=1.0, power_w=0.1, preset="SMF28", beta_order=3)
            p = fiberhip.fiber_problem(fib, grid=fiberhip.Grid(nt=nt, time_window_ps=SERVE_WINDOW_PS[nt], policy="exact"),
                                       raman_threshold_thz=None)
        else:
            from fiberhip import grin

            p = grin.grin_problem(modes, nt=nt, time_window_ps=SERVE_WINDOW_PS[nt], length_m=1.0)
        ut = np.fft.fft(p.uω0, axis=0)
        cache[key] = (p, float(np.max(np.sum(np.abs(ut) ** 2, axis=1))), fiberhip.taylor_phase_basis(p, (2, 3, 4)),
                      fiberhip.frequency_band_mask(p, (-16.0, -10.0)))
    return cache[key]


def run_serving(trace, precision, rate, cache, lanes=SERVE_LANES, max_batch=64, budget_gib=160, drop_fields=False):
    """Submit the trace with Poisson arrivals at `rate` requests/s from a feeder
    thread and collect every request as it finishes; returns per-request results."""
    import fiberhip

    rng = np.random.default_rng(7)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(trace)))
    inputs = []
    for r in trace:
        p, peak, basis, mask = serving_problem(cache, r["nt"], r["modes"])
        u0 = p.uω0 * np.sqrt(r["peak"] / peak) * np.exp(1j * (basis @ r["coef"]))[:, None]
        inputs.append((p, u0, mask))
    print(f"bench: serving {precision}: {len(trace)} launches prepared", file=sys.stderr, flush=True)
    ids, rejected = [None] * len(trace), {}
    submitted = [0]
    srv = fiberhip.Server(memory_budget_bytes=int(budget_gib * (1 << 30)), lanes=lanes, max_batch=max_batch,
                          precision=precision,
                          rtol=1e-6, atol=1e-6, multimode_step_m=5e-3, max_steps=20000)
    # latency runs from each request's scheduled arrival (steady clock, the
    # server's clock) to its completion, so time a submission waits counts
    t0_ns = time.monotonic_ns()

    def feed():
        t0 = time.perf_counter()
        for i, (r, (p, u0, mask)) in enumerate(zip(trace, inputs)):
            delay = arrivals[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            try:
                ids[i] = srv.submit(p, kind=r["kind"], u0=u0, length_m=r["length"],
                                    band_mask=mask if r["kind"] == "gradient" else None)
            except fiberhip.FiberHIPError as exc:
                cls = f"M={r['modes']} {r['kind']} nt={r['nt']}" if exc.code == fiberhip._lib.FH_EUNSUPPORTED else str(exc)
                rejected[cls] = rejected.get(cls, 0) + 1
            submitted[0] = i + 1
            if (i + 1) % max(1, len(trace) // 10) == 0:
                print(f"bench: serving {precision}: {i + 1} of {len(trace)} requests submitted, "
                      f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    t_start = time.perf_counter()
    th = threading.Thread(target=feed)
    th.start()
    results = [None] * len(trace)
    open_ = set()
    collected, next_seen, last_print = 0, 0, time.perf_counter()
    # collect in completion order: poll the submitted requests, wait on the finished ones
    while collected < len(trace):
        while next_seen < submitted[0]:
            if ids[next_seen] is not None:
                open_.add(next_seen)
            else:
                collected += 1  # rejected at submission
            next_seen += 1
        finished = [i for i in open_ if srv.poll(ids[i]) == 2]
        for i in finished:
            open_.discard(i)
            try:
                results[i] = srv.wait(ids[i])
                results[i]["latency_ms"] = (results[i]["finish_ns"] - t0_ns) / 1e6 - arrivals[i] * 1e3
                results[i]["energy"] = np.sum(np.abs(results[i]["uL"]) ** 2, axis=0)
                if drop_fields:
                    del results[i]["uL"]
                    results[i].pop("grad", None)
            except fiberhip.FiberHIPError as exc:
                results[i] = dict(error=str(exc), code=exc.code)
            collected += 1
            if collected % max(1, len(trace) // 10) == 0 or time.perf_counter() - last_print > 30:
                last_print = time.perf_counter()
                print(f"bench: serving {precision}: {collected} of {len(trace)} requests finished, "
                      f"{time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)
        if not finished:
            time.sleep(0.05)
    th.join()
    wall = time.perf_counter() - t_start
    srv.close()
    return results, rejected, wall


def _host_description() -> dict:
    """CPU model, core counts and the oracle's compiler flags, for the baseline's record."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    flags = ""
    try:
        for ln in (ROOT / "oracle" / "Makefile").read_text().splitlines():
            if ln.startswith("CFLAGS"):
                flags = ln.split("=", 1)[1].strip()
    except OSError:
        pass
    import subprocess

    try:
        cc = "gcc " + subprocess.run(["gcc", "-dumpfullversion"], capture_output=True, text=True,
                                     timeout=10).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        cc = "gcc"
    return dict(cpu_model=model, nproc=os.cpu_count(), affinity_cores=len(os.sched_getaffinity(0)),
                compiler=cc, cflags=flags)


def _median_of_runs(fn, runs: int = 5):
    """One untimed warm-up, then the median wall time of `runs` timed calls."""
    fn()
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), ts


def _parallel(threads: int, work) -> None:
    ths = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()


def cpu_baseline_gradient(problem, steps, mask, slice_steps=50, runs=5):
    """The oracle's C discrete adjoint (FP64, one thread per gradient, as the
    reference pins one FFT thread per solve) on the host cores.  Bounded
    sample: each gradient runs over the first `slice_steps` of the workload's
    RK4 steps (same grid, step size and checkpoint spacing); the per-step cost
    is uniform, so full-length figures are the sample's scaled by
    steps/slice_steps.  Protocol: one warm-up, median of `runs` timed runs, for
    (i) one gradient on one thread and (ii) one gradient per thread on every
    core of the GPU's host share in parallel."""
    import fiberhip
    import oracle
    from oracle import np_ref

    f = problem.fiber
    fib = oracle.Fiber(f["Dω"], f["γ"], f["hRω"], fiberhip.self_steepening(problem.sim), f["one_m_fR"])
    threads = max(1, min(HOST_CORE_SHARE, len(os.sched_getaffinity(0))))
    u = launches(problem, threads, seed=2, orders=(2, 3))
    h = f["L"] / steps
    length = h * slice_steps

    def work(i):
        u0 = u[i % threads][:, None]
        uL = oracle.rk4_forward(fib, length, slice_steps, u0)
        _, seed = np_ref.band_cost(uL, mask)
        oracle.rk4_adjoint(fib, length, slice_steps, u0, seed, every=50)

    single, _ = _median_of_runs(lambda: work(0), runs)
    allcore, _ = _median_of_runs(lambda: _parallel(threads, work), runs)
    scale = steps / slice_steps
    value = threads / (allcore * scale)
    return dict(value=value, unit="grads/s", cores=threads, kind="port",
                single_core_latency_s=single * scale, all_core_grads_per_s=value,
                protocol=f"one warm-up, median of {runs} timed runs; sample of {slice_steps} of {steps} RK4 steps "
                         f"per gradient, scaled by {steps}/{slice_steps}",
                sample=f"SMF-28 2 m gradients (forward + discrete adjoint, checkpoint every 50) over the first "
                       f"{slice_steps} of {steps} RK4 steps at Nt={problem.nt}: {threads} in parallel, one per "
                       f"thread (median {allcore:.2f} s), and one alone on one thread (median {single:.2f} s); "
                       f"rates scaled by {slice_steps}/{steps}",
                **_host_description())


def cpu_baseline(problem, steps, runs=5):
    """The oracle's C restatement of the fixed-step forward (FP64, one thread
    per solve) on the host cores: one warm-up, then the median of `runs` timed
    runs for (i) one solve on one thread and (ii) one solve per thread on every
    core of the GPU's host share in parallel (the reference parallelises
    independent solves, not transforms)."""
    import fiberhip
    import oracle

    f = problem.fiber
    fib = oracle.Fiber(f["Dω"], f["γ"], f["hRω"], fiberhip.self_steepening(problem.sim), f["one_m_fR"])
    threads = max(1, min(HOST_CORE_SHARE, len(os.sched_getaffinity(0))))
    u = launches(problem, threads)

    def work(i):
        oracle.rk4_forward(fib, f["L"], steps, u[i % threads][:, None])

    single, _ = _median_of_runs(lambda: work(0), runs)
    allcore, _ = _median_of_runs(lambda: _parallel(threads, work), runs)
    return dict(value=threads / allcore, unit="solves/s", cores=threads, kind="port",
                single_core_latency_s=single, all_core_solves_per_s=threads / allcore,
                protocol=f"one warm-up, median of {runs} timed runs",
                sample=f"SMF-28 0.5 m solves ({steps} RK4 steps, Nt={problem.nt}): {threads} in parallel, one per "
                       f"thread (median {allcore:.3f} s), and one alone on one thread (median {single:.3f} s)",
                **_host_description())


def serving_report(prefix: str) -> dict:
    """Latency, throughput and FP32-vs-FP64 matched error of a serving trace
    from the per-request records of both precisions."""
    recs = {p: np.load(f"{prefix}_{p}.npz") for p in ("fp32", "fp64")}
    out = {}
    for p, r in recs.items():
        ok = r["ok"]
        lat = r["latency_ms"][ok]
        grad = ok & (r["kind"] == "gradient")
        wall = float(r["wall_s"])
        out[p] = {"served": int(ok.sum()), "wall_s": wall, "solves_per_s": float(ok.sum() / wall),
                  "gradients_per_s": float(grad.sum() / wall), "forwards_per_s": float((ok & ~grad).sum() / wall),
                  "p50_ms": float(np.percentile(lat, 50)), "p90_ms": float(np.percentile(lat, 90)),
                  "p99_ms": float(np.percentile(lat, 99)), "mean_batch": float(np.mean(r["batch"][ok])),
                  "rate_per_s": float(r["rate"])}
    a, b = recs["fp32"], recs["fp64"]
    n = min(a["ok"].size, b["ok"].size)  # one precision may have served a prefix of the trace
    a = {k: (a[k][:n] if a[k].ndim else a[k]) for k in a.files}
    b = {k: (b[k][:n] if b[k].ndim else b[k]) for k in b.files}
    both = a["ok"] & b["ok"]
    err = np.nanmax(np.abs(a["energy"][both] / b["energy"][both] - 1), axis=1)
    out["matched_requests"] = int(both.sum())
    out["fp32_vs_fp64_rel_mode_energy_err"] = {"max": float(err.max()), "p99": float(np.percentile(err, 99)),
                                               "within_1e-5": int((err <= 1e-5).sum())}
    out["fp32_over_fp64_gradients_per_s"] = out["fp32"]["gradients_per_s"] / out["fp64"]["gradients_per_s"]
    out["fp32_over_fp64_solves_per_s"] = out["fp32"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["note"] = (f"fp32 served {out['fp32']['served']} and fp64 {out['fp64']['served']} requests of the same trace "
                   f"(same arrivals); the error compares the {n} requests both were given")
    by = {}
    full = recs["fp32"]
    for m in (1, 3, 6, 10):
        for kind in ("forward", "gradient"):
            sel = full["ok"] & (full["modes"] == m) & (full["kind"] == kind)
            if sel.any():
                by[f"M={m} {kind}"] = {"fp32_p50_ms": float(np.percentile(full["latency_ms"][sel], 50)),
                                       "fp32_p99_ms": float(np.percentile(full["latency_ms"][sel], 99))}
    out["fp32_latency_by_class"] = by
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed launches per path")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256, help="gradients per launch (SMF-28 2 m gradient workload)")
    ap.add_argument("--tape-budget-gib", type=float, default=200.0,
                    help="device memory the gradient plans may spend on kept stage tape (same budget for FP32 "
                         "and FP64; 0 recomputes every segment)")
    ap.add_argument("--fwd-batch", type=int, default=1024, help="solves per launch (SMF-28 0.5 m forward workload)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["all", "gradient", "forward", "multimode", "sweep", "gain", "serving"],
                    default="all")
    ap.add_argument("--gain-batch", type=int, default=256, help="amplifier solves per call (YDFA gain workload)")
    ap.add_argument("--sweep-candidates", type=int, default=256, help="adaptive candidate sweep size")
    ap.add_argument("--sweep-tape-segments", type=int, default=3,
                    help="accepted-step segments of 50 whose stage tape the sweep's forward pass keeps (0 recomputes "
                         "every segment)")
    ap.add_argument("--serve-requests", type=int, default=100, help="requests in the serving trace")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-budget-gib", type=float, default=160.0, help="device memory of the server's plans")
    ap.add_argument("--serve-lanes", type=int, default=SERVE_LANES,
                    help="server lanes (concurrent launches); the process gets lanes + 4 hardware queues")
    ap.add_argument("--serve-precisions", default="fp32,fp64", help="server precisions to run, comma-separated")
    ap.add_argument("--serve-first", type=int, default=0,
                    help="serve only the first K requests of the --serve-requests trace (same requests and arrivals)")
    ap.add_argument("--serve-save", default="", help="write per-request records (npz) of each precision to "
                    "this path prefix; output fields are not kept in memory")
    ap.add_argument("--mm-batch", type=int, default=256, help="solves per launch (GRIN M=6 forward workload)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run the plain copy kernel over this many MiB (HBM counter calibration)")
    ap.add_argument("--serve-report", default="",
                    help="summarise the records PREFIX_fp32.npz and PREFIX_fp64.npz written by --serve-save (no GPU)")
    args = ap.parse_args()
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.serve_lanes + 4)
    if args.serve_report:
        print(json.dumps(serving_report(args.serve_report)), flush=True)
        return
    if args.copy_calibration_mib:
        import ctypes

        import fiberhip

        ctx = fiberhip.default_context()
        nbytes = args.copy_calibration_mib << 20
        a, b = fiberhip.DeviceArray(ctx, nbytes), fiberhip.DeviceArray(ctx, nbytes)
        fiberhip._lib.check(fiberhip.lib().fh_device_copy(ctx.handle, ctypes.c_void_p(b.ptr),
                                                          ctypes.c_void_p(a.ptr), nbytes))
        print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (independent solves shard across GPUs without "
              "communication); run with --gpus 1", file=sys.stderr)
        raise SystemExit(2)
    import fiberhip

    line = {}
    if args.only in ("all", "gradient"):
        problem, steps = workload("smf28-grad-16384")
        nt, m = problem.nt, problem.modes
        mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
        budget = int(args.tape_budget_gib * (1 << 30))
        kept = {p: kept_segments(nt, steps, 50, p, args.batch, budget) for p in ("fp32", "fp64")}
        g = {p: time_gradient(problem, steps, p, args.batch, args.steps, args.warmup, mask, kept[p])
             for p in ("fp32", "fp64")}
        ms = g["fp32"]["kernel_ms"]
        value = args.batch / (ms / 1e3)
        fp64_value = args.batch / (g["fp64"]["kernel_ms"] / 1e3)
        flops = gradient_flops(nt, m, steps)
        achieved = args.batch * flops / (ms / 1e3) / 1e12
        e32, e64 = g["fp32"]["energies"], g["fp64"]["energies"]
        gerr = np.linalg.norm(g["fp32"]["grad"] - g["fp64"]["grad"], axis=1) / np.linalg.norm(g["fp64"]["grad"], axis=1)
        # the matched-error condition of the metric: output spectral energy of every
        # FP32 solve within 1e-5 of the FP64 solve of the same launch
        energy_err = float(np.max(np.abs(e32 - e64) / e64))
        matched = energy_err <= MATCHED_ENERGY_TOL
        metric = "adjoint grads/sec (FP32-bulk batched path) vs plain FP64 GPU @<=1e-5 rel spectral-energy err"
        if not matched:
            metric += f" (tolerance NOT met: {energy_err:.2e})"
        # batch-1 latency: one gradient alone on the GPU (the L-BFGS loop's use), every segment's tape kept
        lat = {}
        for p in ("fp32", "fp64"):
            r1 = time_gradient(problem, steps, p, 1, 2, 1, mask, 40)
            lat[f"{p}_ms"] = r1["kernel_ms"]
        print(f"bench: batch-1 gradient {lat['fp32_ms']:.1f} ms fp32, {lat['fp64_ms']:.1f} ms fp64",
              file=sys.stderr, flush=True)
        line = {
            "metric": metric,
            "value": value,
            "unit": "grads/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": ms,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"SMF-28 2 m, 0.2 W, 185 fs sech2, Nt={nt}, T={problem.sim['time_window']} ps, "
                                   f"{steps} RK4IP steps, checkpoint every 50, band -16..-10 THz objective, "
                                   f"gradient w.r.t. the full-grid spectral phase",
                       "batch_gradients_per_step": args.batch, "modes": m, "nt": nt,
                       "kept_tape_segments": kept, "tape_budget_gib": args.tape_budget_gib},
            "fp64_value": fp64_value,
            "fp32_over_fp64": value / fp64_value,
            "fp32_vs_fp64": {"max_rel_output_spectral_energy_err": energy_err,
                             "tolerance": MATCHED_ENERGY_TOL, "within_tolerance": matched,
                             "max_rel_cost_err": float(np.max(np.abs(g["fp32"]["cost"] - g["fp64"]["cost"])
                                                              / np.abs(g["fp64"]["cost"]))),
                             "max_rel_gradient_err_normwise": float(np.max(gerr)),
                             "gradient_normwise_bound": GRADIENT_NORMWISE_BOUND},
            "batch1_gradient_latency_ms": lat,
            "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                         "frac": achieved / PEAK_TFLOPS["fp32"],
                         "limiter": "vector ALU and LDS: the block transforms' LDS exchanges run at 86% of the "
                                    "measured LDS throughput; no matrix-core instructions in this kernel",
                         "traffic": MEASURED_HBM_BYTES["k_grad_single<float,16384,1024>"][0] * args.batch,
                         "traffic_source": MEASURED_HBM_BYTES["k_grad_single<float,16384,1024>"][1],
                         "kernel": "k_grad_single<float,16384,1024>", "flops_per_gradient": flops},
            "fp64_roofline": {"achieved": args.batch * flops / (g["fp64"]["kernel_ms"] / 1e3) / 1e12,
                              "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s", "kernel": "k_grad_single<double,16384,1024>"},
            "device_bytes_fp32_plan": g["fp32"]["device_bytes"],
            "device_bytes_fp64_plan": g["fp64"]["device_bytes"],
        }
        if not args.no_cpu:
            line["cpu_baseline"] = cpu_baseline_gradient(problem, steps, mask)
    if args.only in ("all", "forward"):
        problem, steps = workload("smf28-fwd-8192")
        nt, m = problem.nt, problem.modes
        res = {p: time_path(problem, steps, p, args.fwd_batch, args.steps, args.warmup) for p in ("fp32", "fp64")}
        e32, e64 = res["fp32"]["energies"], res["fp64"]["energies"]
        flops = forward_flops(nt, m, steps)
        fwd = {
            "workload": f"SMF-28 0.5 m, 0.1 W, Nt={nt}, T={problem.sim['time_window']} ps, {steps} RK4IP steps",
            "batch_solves_per_step": args.fwd_batch,
            "fp32_solves_per_s": args.fwd_batch / (res["fp32"]["kernel_ms"] / 1e3),
            "fp64_solves_per_s": args.fwd_batch / (res["fp64"]["kernel_ms"] / 1e3),
            "fp32_ms_per_launch": res["fp32"]["kernel_ms"],
            "fp64_ms_per_launch": res["fp64"]["kernel_ms"],
            "fp32_vs_fp64_max_rel_energy_err": float(np.max(np.abs(e32 - e64) / e64)),
            "fp32_roofline_frac": args.fwd_batch * flops / (res["fp32"]["kernel_ms"] / 1e3) / 1e12 / PEAK_TFLOPS["fp32"],
            "fp64_roofline_frac": args.fwd_batch * flops / (res["fp64"]["kernel_ms"] / 1e3) / 1e12 / PEAK_TFLOPS["fp64"],
        }
        # batch-1 latency: one solve alone on the GPU
        fwd["batch1_latency_ms"] = {f"{p}_ms": time_path(problem, steps, p, 1, 5, 1)["kernel_ms"]
                                    for p in ("fp32", "fp64")}
        if not args.no_cpu:
            fwd["cpu_baseline"] = cpu_baseline(problem, steps)
        if args.only == "forward":
            line = {"metric": "fwd solves/sec (FP32-bulk batched path)", "value": fwd["fp32_solves_per_s"],
                    "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                    "ms_per_step": fwd["fp32_ms_per_launch"], "higher_is_better": True, "scaling": "weak",
                    "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                    "config": {"workload": fwd["workload"]}, "forward": fwd}
        else:
            line["forward_smf28_8192"] = fwd
    if args.only in ("all", "multimode"):
        problem, steps = workload("grin6-fwd-4096")
        nt, m = problem.nt, problem.modes
        uh = grin_launches(args.mm_batch)
        res = {p: time_path(problem, steps, p, args.mm_batch, 1, 0, u_host=uh) for p in ("fp32", "fp64")}
        e32, e64 = res["fp32"]["energies"], res["fp64"]["energies"]
        per_mode = np.max(np.abs(e32 - e64) / e64, axis=0)
        flops = forward_flops(nt, m, steps)
        # Kerr contraction on the matrix cores (FP32, M >= 6): per right-hand side
        # and 32-sample chunk, ceil(Q/32) output tiles of ceil(P/2)
        # v_mfma_f32_32x32x2f32 (4096 flops each), Q = P = M(M+1)/2 pairs; the
        # useful part is the 2 Q P flops per sample of the reduced product
        q = m * (m + 1) // 2
        mfma_inst = 4 * steps * (nt // 32) * (-(-q // 32)) * (-(-q // 2))
        mfma_rate = args.mm_batch * mfma_inst * 4096 / (res["fp32"]["kernel_ms"] / 1e3) / 1e12
        mm = {
            "workload": f"GRIN 50 um / NA 0.2, 0.3 m, M=6 HG modes (256x256), Nt={nt}, 12 ps, 10 nJ sech2 "
                        f"100 fs, seeded Dirichlet launches, {steps} RK4IP steps",
            "batch_solves_per_step": args.mm_batch,
            "fp32_solves_per_s": args.mm_batch / (res["fp32"]["kernel_ms"] / 1e3),
            "fp64_solves_per_s": args.mm_batch / (res["fp64"]["kernel_ms"] / 1e3),
            "fp32_ms_per_launch": res["fp32"]["kernel_ms"],
            "fp64_ms_per_launch": res["fp64"]["kernel_ms"],
            "fp32_vs_fp64_max_rel_energy_err_per_mode": [float(x) for x in per_mode],
            "fp32_roofline_frac": args.mm_batch * flops / (res["fp32"]["kernel_ms"] / 1e3) / 1e12 / PEAK_TFLOPS["fp32"],
            "fp64_roofline_frac": args.mm_batch * flops / (res["fp64"]["kernel_ms"] / 1e3) / 1e12 / PEAK_TFLOPS["fp64"],
            "flops_per_solve": flops,
            "skc_mfma": {"instructions_per_solve": mfma_inst, "issued_tflops": mfma_rate,
                         "frac_of_fp32_matrix_peak": mfma_rate / PEAK_TFLOPS["fp32"],
                         "useful_fraction_of_issued": 2 * q * q * 32 / (4096 * (-(-q // 32)) * (-(-q // 2))),
                         "source": "v_mfma_f32_32x32x2f32 count per solve, equal to SQ_INSTS_VALU_MFMA_F32 in "
                                   "profiles/k_rk4ip_multi_grin6_4096.md"},
        }
        if args.only == "multimode":
            line = {"metric": "fwd solves/sec (FP32-bulk batched path, GRIN M=6)", "value": mm["fp32_solves_per_s"],
                    "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": mm["fp32_ms_per_launch"],
                    "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                    "data": "synthetic", "config": {"workload": mm["workload"]}, "multimode": mm}
        else:
            line["forward_grin6_4096"] = mm
    if args.only in ("all", "sweep"):
        problem, _ = workload("smf28-sweep-8192")
        nt = problem.nt
        mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
        nc = args.sweep_candidates
        # schedule autotuning: candidates per launch, fastest FP32 sweep wins
        tune = {}
        for per in (nc, nc // 2, nc // 4):
            if per >= 1 and nc % per == 0:
                tune[per] = time_sweep(problem, "fp32", nc, per, 1, 1, mask, tape_segments=args.sweep_tape_segments)["ms"]
                print(f"bench: sweep schedule {nc // per} x {per}: {tune[per]:.1f} ms", file=sys.stderr, flush=True)
        per = min(tune, key=tune.get)
        res = {p: time_sweep(problem, p, nc, per, args.steps, args.warmup, mask, tape_segments=args.sweep_tape_segments)
               for p in ("fp32", "fp64")}
        st = res["fp64"]["stats"]
        # algorithmic flops from the actual evaluation counts: forward RHS (accepted +
        # rejected + starting step) and 6 stage pullbacks per accepted step
        flops = float(np.sum(st[:, 2]) * rhs_flops(nt, 1) + np.sum(6 * st[:, 0]) * adjoint_rhs_flops(nt, 1)) / nc
        ms32, ms64 = res["fp32"]["ms"], res["fp64"]["ms"]
        sweep = {
            "workload": f"candidate sweep: SMF-28 1 m, 0.1 W, Nt={nt}, T={problem.sim['time_window']} ps, {nc} "
                        f"GDD/TOD/FOD candidates, adaptive Tsit5 rtol=atol=1e-6, forward + adjoint gradient of the "
                        f"-16..-10 THz band objective",
            "candidates": nc, "schedule": f"{nc // per} launches x {per} candidates", "autotune_ms": tune,
            "kept_tape_segments": args.sweep_tape_segments,
            "fp32_grads_per_s": nc / (ms32 / 1e3), "fp64_grads_per_s": nc / (ms64 / 1e3),
            "fp32_over_fp64": ms64 / ms32,
            "fp64_accepted_steps_mean": float(np.mean(st[:, 0])), "fp64_rejected_steps_mean": float(np.mean(st[:, 1])),
            "fp32_accepted_steps_mean": float(np.mean(res["fp32"]["stats"][:, 0])),
            "max_rel_cost_err_fp32_vs_fp64": float(np.max(np.abs(res["fp32"]["cost"] / res["fp64"]["cost"] - 1))),
            "flops_per_gradient": flops,
            "fp32_roofline_frac": nc * flops / (ms32 / 1e3) / 1e12 / PEAK_TFLOPS["fp32"],
            "fp64_roofline_frac": nc * flops / (ms64 / 1e3) / 1e12 / PEAK_TFLOPS["fp64"],
        }
        if args.only == "sweep":
            line = {"metric": "adjoint grads/sec (adaptive candidate sweep, FP32-bulk)", "value": sweep["fp32_grads_per_s"],
                    "unit": "grads/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms32,
                    "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                    "data": "synthetic", "config": {"workload": sweep["workload"]}, "sweep": sweep}
        else:
            line["sweep_smf28_8192"] = sweep
    if args.only in ("all", "gain"):
        nt = 4096
        res = {p: time_gain(p, args.gain_batch, max(1, args.steps), nt) for p in ("fp32", "fp64")}
        # algorithmic flops from the counted right-hand-side evaluations: the
        # disp_mmf! structure plus 12 N for the gain (|x|^2, two rate sums,
        # g_k and the g/2 x term)
        nrhs = {p: res[p][1].stats["rhs_evals"] for p in res}
        flops = {p: nrhs[p] * (rhs_flops(nt, 1) + 12 * nt) / args.gain_batch for p in res}
        e32 = np.sum(np.abs(res["fp32"][1].output_spectrum) ** 2, axis=(1, 2))
        e64 = np.sum(np.abs(res["fp64"][1].output_spectrum) ** 2, axis=(1, 2))
        gain_line = {
            "workload": f"YDFA amplifier (solve_disp_gain_smf): 1030 nm, Nt={nt}, 20 ps, 1 m Yb fiber, 300 fs "
                        f"Gaussian 50 mW + 0.5 W pump, GDD/TOD launch phases, Tsit5 reltol 1e-5 on [Pp; x]",
            "batch_solves_per_call": args.gain_batch,
            "timing": "median wall time of the host call fh_forward_gain, host staging and copies included",
            **{f"{p}_solves_per_s": args.gain_batch / res[p][0] for p in res},
            **{f"{p}_ms_per_call": res[p][0] * 1e3 for p in res},
            **{f"{p}_rhs_evals_per_solve": nrhs[p] / args.gain_batch for p in res},
            **{f"{p}_roofline_frac": args.gain_batch * flops[p] / res[p][0] / 1e12 / PEAK_TFLOPS[p] for p in res},
            "fp32_vs_fp64_max_rel_energy_err": float(np.max(np.abs(e32 / e64 - 1))),
            "fp32_vs_fp64_max_rel_pump_err": float(np.max(np.abs(res["fp32"][1].pump_power /
                                                                 res["fp64"][1].pump_power - 1))),
        }
        if args.only == "gain":
            line = {"metric": "amplifier solves/sec (FP32-bulk, YDFA gain)", "value": gain_line["fp32_solves_per_s"],
                    "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
                    "ms_per_step": gain_line["fp32_ms_per_call"], "higher_is_better": True, "scaling": "weak",
                    "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                    "config": {"workload": gain_line["workload"]}, "gain": gain_line}
        else:
            line["gain_ydfa_4096"] = gain_line
    if args.only in ("all", "serving"):
        trace = serving_trace(args.serve_requests)
        if args.serve_first:
            trace = trace[:args.serve_first]
        cache = {}
        runs = {}
        for precision in args.serve_precisions.split(","):
            res, rejected, wall = run_serving(trace, precision, args.serve_rate, cache, lanes=args.serve_lanes,
                                              budget_gib=args.serve_budget_gib,
                                              drop_fields=bool(args.serve_save))
            if args.serve_save:
                ok_ = [r is not None and "error" not in r for r in res]
                energy = np.full((len(trace), 10), np.nan)
                for i, r in enumerate(res):
                    if ok_[i]:
                        energy[i, :r["energy"].size] = r["energy"]
                np.savez(f"{args.serve_save}_{precision}.npz", ok=np.array(ok_), energy=energy,
                         code=np.array([r.get("code", -1) if r is not None and "error" in r else 0 for r in res]),
                         error=np.array([r["error"] if r is not None and "error" in r else "" for r in res]),
                         latency_ms=np.array([r["latency_ms"] if o else np.nan for r, o in zip(res, ok_)]),
                         batch=np.array([r["batch"] if o else 0 for r, o in zip(res, ok_)]),
                         attempts=np.array([r["attempts"] if o else 0 for r, o in zip(res, ok_)]),
                         nt=np.array([r["nt"] for r in trace]), modes=np.array([r["modes"] for r in trace]),
                         kind=np.array([r["kind"] for r in trace]), length=np.array([r["length"] for r in trace]),
                         wall_s=wall, rate=args.serve_rate)
            ok = [r for r in res if r is not None and "error" not in r]
            lat = np.array([r["latency_ms"] for r in ok])
            ngrad = sum(1 for r, t in zip(res, trace) if r is not None and "error" not in r and t["kind"] == "gradient")
            runs[precision] = dict(results=res, rejected=rejected, wall_s=wall, served=len(ok),
                                   gradients_per_s=ngrad / wall, forwards_per_s=(len(ok) - ngrad) / wall,
                                   failed=sum(1 for r in res if r is not None and "error" in r),
                                   rerun_alone=sum(1 for r in ok if r.get("attempts", 1) > 1),
                                   p50_ms=float(np.percentile(lat, 50)) if lat.size else None,
                                   p99_ms=float(np.percentile(lat, 99)) if lat.size else None,
                                   solves_per_s=len(ok) / wall,
                                   mean_batch=float(np.mean([r["batch"] for r in ok])) if ok else None)
            print(f"bench: serving {precision}: {len(ok)} served in {wall:.1f} s, p50 {runs[precision]['p50_ms']} ms, "
                  f"p99 {runs[precision]['p99_ms']} ms", file=sys.stderr, flush=True)
        errs = []
        for a, b in zip(runs.get("fp32", {}).get("results", []), runs.get("fp64", {}).get("results", [])):
            if a and b and "error" not in a and "error" not in b:
                errs.append(float(np.max(np.abs(a["energy"] / b["energy"] - 1))))
        serving = {
            "workload": f"serving trace: {args.serve_requests} requests, Poisson arrivals at {args.serve_rate}/s, "
                        f"Nt in 2^11..2^15, M in {{1,3,6,10}} (SMF-28 / 50 um GRIN HG), L log-uniform 0.1-5 m, "
                        f"peak 1-20 kW, GDD/TOD/FOD phases, 50% forward / 50% forward+adjoint; single mode adaptive "
                        f"Tsit5 rtol=atol=1e-6, multimode fixed 5 mm RK4IP steps",
            **{f"{p}_{k}": runs[p][k] for p in runs
               for k in ("served", "failed", "rerun_alone", "p50_ms", "p99_ms", "solves_per_s", "gradients_per_s", "forwards_per_s",
                         "wall_s", "mean_batch", "rejected")},
            **({"fp32_over_fp64_gradients_per_s": runs["fp32"]["gradients_per_s"] / runs["fp64"]["gradients_per_s"]}
               if "fp32" in runs and "fp64" in runs and runs["fp64"]["gradients_per_s"] > 0 else {}),
            "lanes": args.serve_lanes, "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
            "memory_budget_gib": args.serve_budget_gib,
            "latency": "from each request's scheduled Poisson arrival to its completion",
            "fp32_vs_fp64_matched_requests": len(errs),
            "fp32_vs_fp64_max_rel_mode_energy_err": max(errs) if errs else None,
            "fp32_vs_fp64_p99_rel_mode_energy_err": float(np.percentile(errs, 99)) if errs else None,
        }
        if args.only == "serving":
            lead = "fp32" if "fp32" in runs else next(iter(runs))
            line = {"metric": f"serving solves/sec with p50/p99 latency ({'FP32-bulk' if lead == 'fp32' else 'FP64'})",
                    "value": serving[f"{lead}_solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": 1,
                    "warmup": 0, "ms_per_step": serving[f"{lead}_wall_s"] * 1e3, "higher_is_better": True,
                    "scaling": "weak", "vs_baseline": None, "dtype": "f32" if lead == "fp32" else "f64",
                    "data": "synthetic", "config": {"workload": serving["workload"]}, "serving": serving}
        else:
            line["serving_trace"] = serving
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
