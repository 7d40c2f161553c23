"""Benchmark of the MI355X Crank-Nicolson step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--accuracy-steps A]

Workload (one GPU): binary q=0.5, R=50 on a fixed orbit with both s-wave
absorbers on, boson mass mu=0.15, N=16384 radial nodes over L=4000, l_max=20
(231 modes), dt=0.25, damping mask (0.85, 1, 6), r_ex=0.8 L; initial state the
2p atom on the primary translated to the centre of mass plus a seeded 1e-3
complex Gaussian perturbation (synthetic inputs).

A step is one CrankNicolsonModes::step: the CN solve with its BiCGSTAB, the
mask and every diagnostic. The metric is mode.node-steps/s = K*N*steps/time,
reported for the mixed-precision step (FP32 coupling in the Krylov applies,
everything else FP64) -- the JSON value -- and for the plain FP64 step, with
the relative L2 difference of the two states after A untimed steps from the
same initial state. The line also carries the dominant kernels' rooflines
(HIP-event kernel times on the solver's stream) and a CPU baseline timed on
the host: the FP64 oracle on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

# Hardware queues of this process (read when the HIP runtime starts, so set
# before anything touches the device): the ensemble and the served stream run
# one stream per solver, and 8 queues let more of their kernels overlap than
# HIP's default of 4 (the single-solver step is unaffected).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "boson-cloud-binary_amd"))

import gm  # noqa: E402
from gm import initial  # noqa: E402

METRIC = "CN mode·node-steps/s, mixed vs FP64 (≤1e-8 rel err); p99 served step latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # dense FP64 matrix (and vector) peak of one MI355X
FP32_PEAK_TFLOPS = 157.3  # dense FP32 matrix peak

WORKLOAD = dict(Nr=16384, L=4000.0, dt=0.25, epsilon=0.0, l_max=20, mu=0.15, M_c=0.01, q=0.5, a=0.0, R=50.0,
                counter=False, absorption=True, save_every=50, ramp_periods=0.0, epsilon_smear=2.0)
MASK = (0.85, 1.0, 6)


def initial_states(w, seeds):
    """The translated 2p atom with one perturbation per seed (the atom built once)."""
    dr = w["L"] / (w["Nr"] - 1)
    r = w["epsilon"] + dr * np.arange(w["Nr"])
    base = initial.translated_atom(r, w["l_max"], w["mu"], w["q"] / (1 + w["q"]) * w["R"])
    base /= math.sqrt(np.vdot(base, base).real * dr)
    K = (w["l_max"] + 1) * (w["l_max"] + 2) // 2
    return [initial.perturbed_state(base, w["Nr"], K, rel=1e-3, seed=sd, dr=dr) for sd in seeds]


def headline_batch(w, members, steps, warmup, fp64_value, graphs=True):
    """The mixed-precision batched step at the headline shape: `members`
    binaries of the workload (perturbations seeded 0, 1, ...) stepped as
    batched ragged launches, one lane per member, against the plain FP64
    single step."""
    states = initial_states(w, range(members))
    sol = []
    for psi in states:
        s = gm.CrankNicolsonModes(**w)
        s.set_flags(evolve_orbit=0, precision=gm.GM_PRECISION_MIXED)
        if not graphs:
            s.set_flags(use_graphs=0)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(*MASK)
        s.begin_run(0.8 * w["L"], 100000)
        sol.append(s)
    batch = gm.Batch(sol, members)
    k = 0
    for _ in range(warmup):
        k += 1
        batch.step(k, diag=True)
    device_sync()
    t0 = time.perf_counter()
    iters = []
    for _ in range(steps):
        k += 1
        batch.step(k, diag=True)
        iters.append(float(np.mean([s.last_solver_iterations for s in sol])))
    device_sync()
    el = time.perf_counter() - t0
    K, N = sol[0].n_modes, sol[0].Nr
    value = K * N * members * steps / el
    batch.close()
    for s in sol:
        s.close()
    return {"members": members, "lanes": members, "steps": steps, "warmup": warmup,
            "ms_per_batch_step": el / steps * 1e3, "ms_per_member_step": el / steps / members * 1e3,
            "mode_node_steps_per_s": value, "over_fp64_single_step": value / fp64_value,
            "iterations_per_step": float(np.mean(iters)),
            "config": "%d binaries of the bench workload (perturbations seeded 0..%d), mixed precision, every "
                      "launch of the step a ragged grid over one member's lane, graph-captured"
                      % (members, members - 1)}


def initial_state(w, seed=0):
    """Translated 2p atom + perturbation, cached in /tmp by parameters."""
    key = "gm_init_%d_%g_%d_%g_%g_%g_%d.npy" % (w["Nr"], w["L"], w["l_max"], w["mu"], w["q"], w["R"], seed)
    path = os.path.join("/tmp", key)
    if os.path.exists(path):
        return np.load(path)
    dr = w["L"] / (w["Nr"] - 1)
    r = w["epsilon"] + dr * np.arange(w["Nr"])
    R1 = w["q"] / (1 + w["q"]) * w["R"]
    base = initial.translated_atom(r, w["l_max"], w["mu"], R1)
    base /= math.sqrt(np.vdot(base, base).real * dr)
    K = (w["l_max"] + 1) * (w["l_max"] + 2) // 2
    psi = initial.perturbed_state(base, w["Nr"], K, rel=1e-3, seed=seed, dr=dr)
    try:
        np.save(path, psi)
    except OSError:
        pass
    return psi


def coupling_terms(l_max):
    """Gaunt terms with |m*| <= l* (the multiply-adds the coupling needs per node)."""
    n = 0
    modes = initial.modes_for(l_max)
    ls_m = {}
    for (l, m) in modes:
        for (l1, m1) in modes:
            lo = max(abs(l - l1), abs(m1 - m))
            if (lo + l + l1) % 2:
                lo += 1
            if lo <= l + l1:
                n += (l + l1 - lo) // 2 + 1
    return n


def progress(msg):
    """One line per finished leg on stderr (the JSON line alone goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def make_solver(w, precision, evolve=False):
    s = gm.CrankNicolsonModes(**w)
    if evolve:
        # back-reaction on; dt = P / steps_per_period stays at the workload's dt
        s.set_flags(evolve_orbit=1, precision=precision, steps_per_period=s.P / w["dt"])
    else:
        s.set_flags(evolve_orbit=0, precision=precision)
    s.set_initial_state_full(initial_state(w))
    s.initialize()
    s.build_damping_mask(*MASK)
    s.begin_run(0.8 * w["L"], 1000)
    return s


def device_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def profile(s, on):
    s.profile(on)


def profile_read(s):
    return s.profile_read()


def run_gpu(w, precision, steps, warmup, prof_steps=5, graphs=True, tile=0):
    """Warmup, then `steps` timed steps (graph replay on unless `graphs` is
    false, no per-kernel events), then `prof_steps` steps with HIP-event kernel
    timing (graphs off) for the kernel table and the roofline. `tile` fixes
    the apply tile (nodes; 0 autotunes)."""
    s = make_solver(w, precision)
    if not graphs:
        s.set_flags(use_graphs=0)
    if tile:
        s.set_flags(apply_tile=tile)
    k = 0
    for _ in range(warmup):
        k += 1
        s.step(k, diag=True)
    iters = []
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        k += 1
        s.step(k, diag=True)
        iters.append(s.last_solver_iterations)
    device_sync()
    t1 = time.perf_counter()
    profile(s, True)
    piters = []
    for _ in range(prof_steps):
        k += 1
        s.step(k, diag=True)
        piters.append(s.last_solver_iterations)
    prof = profile_read(s)
    profile(s, False)
    return s, t1 - t0, prof, iters, piters


def evolving(w, steps, warmup, fixed, graphs=True, tile=0):
    """The workload with the orbit evolving (GW and cloud back-reaction, the
    bodies accreting through the absorbers): ms per step per precision and the
    ratio to the fixed-orbit step measured above."""
    out = {}
    for name, precision in (("fp64", gm.GM_PRECISION_FP64), ("mixed", gm.GM_PRECISION_MIXED)):
        s = make_solver(w, precision, evolve=True)
        if not graphs:
            s.set_flags(use_graphs=0)
        if tile:
            s.set_flags(apply_tile=tile)
        k = 0
        for _ in range(warmup):
            k += 1
            s.step(k, diag=True)
        iters = []
        device_sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            k += 1
            s.step(k, diag=True)
            iters.append(s.last_solver_iterations)
        device_sync()
        el = time.perf_counter() - t0
        sc = s.scalars()
        out[name] = {"ms_per_step": el / steps * 1e3, "iterations_per_step": float(np.mean(iters)),
                     "over_fixed_orbit_step": el / steps * 1e3 / fixed[name]["ms_per_step"],
                     "R_after": sc["R"], "dt_after": sc["dt"]}
        del s
    out["steps"] = steps
    return out


def cpu_baseline(w, runs=3, steps_per_run=8):
    """The FP64 oracle on the host cores: the same workload on a radial
    sub-grid of the same spacing (all 231 modes, 2048 nodes), one warm-up
    step, then `runs` timed runs of `steps_per_run` steps each; the median
    run's rate is reported (about 10 s of CPU work on 16 threads)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # sub-grid: keep dr, l_max and the physics, cut the node count
    Nsub = 2048
    ws = dict(w)
    ws["Nr"] = Nsub
    ws["L"] = w["L"] / (w["Nr"] - 1) * (Nsub - 1)
    o = oracle.Oracle(**ws)
    o.set_flags(evolve_orbit=0)
    psi = initial_state(w).reshape(-1, w["Nr"])[:, :Nsub].copy()
    psi[:, -1] = 0
    o.set_initial_state_full(psi.ravel())
    o.initialize()
    o.build_damping_mask(*MASK)
    o.begin_run(0.8 * ws["L"], 1 + runs * steps_per_run)
    o.step(1, diag=True)
    K = o.n_modes
    n, rates, iters = 1, [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        for _ in range(steps_per_run):
            n += 1
            o.step(n, diag=True)
            iters.append(o.last_solver_iterations)
        rates.append(K * Nsub * steps_per_run / (time.perf_counter() - t0))
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": float(np.median(rates)), "unit": "mode·node-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (restatement, not the reference binary), all {K} modes on a {Nsub}-node "
                      f"sub-grid of the same dr: 1 warm-up step, then the median of {runs} runs of "
                      f"{steps_per_run} steps",
            "runs": rates, "cpu_model": cpu, "iterations_per_step": float(np.mean(iters))}


# Algorithmic HBM bytes per mode-node of the bandwidth-bound kernel classes
# (complex128 = 16 B per vector element; tables and per-node profiles excluded):
#   precond  read r, p, v, pivots, write p, y (p-update variant, 6 vectors;
#            s-update variant 5: r, v, pivots / s, z) -> 5.5 on average (the
#            forward results stay on chip; the segment halos' extra reads are
#            not algorithmic)
#   update   read x, y, z, s, t, r0; write x, r -> 8
#   diag     read psi, V x; per-tile sums only -> 2
#   apply    read x (+ r0 / b), write y (+ A psi) -> 3 (coupling excluded)
HBM_PASSES = {"precond": 5.5, "precond_setup": 3.0, "update": 8.0, "diag": 2.0}


def algorithmic_flops(K, N):
    """Algorithmic flops of one operator apply: a real symmetric coupling
    block V0 (K x K per node) times a complex vector is 4 K flops per
    mode-node, plus 12 per mode-node for the three-point stencil, the
    centrifugal term and the phases."""
    return 4.0 * K * K * N + 12.0 * K * N


def executed_flops(terms, pairs, K, N):
    """Flops this build's factored apply executes: one multiply-add per Gaunt
    term kept by the multipole cut (forming W = sum G E), two per coupled mode
    pair (real W times complex x), about 20 per mode-node for the stencil,
    phases and the A = 1 + i dt/2 H form."""
    return 2.0 * terms + 4.0 * pairs + 20.0 * K * N


def mfma_macs(kernel):
    """Matrix-core multiply-adds per launch from the committed SQ counter
    summary (the first table of the file: SQ_INSTS_MFMA x 1024 for the Gaunt
    form's 16x16x4 MFMAs, x 256 for the polar form's 4x4x4 ones), or None."""
    path = os.path.join(REPO, "profiles", "k_apply_sq_counters_l20_n16384.txt")
    per = 256.0 if kernel.startswith("k_tapply<") else 1024.0
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(kernel + " "):
                    return float(line[len(kernel):].split()[2]) * per
    except (OSError, ValueError, IndexError):
        pass
    return None


def apply_kernel(s, precision):
    """The traced name of the operator apply the solver runs in `precision`:
    the polar-quadrature form k_tapply<class> (8-node tiles, FP64 in both
    precisions) or the Gaunt form k_apply<ng, T> (16 ng-node tiles), as the
    autotuner chose."""
    tile = s.apply_config(precision)[0]
    t = "float" if precision == gm.GM_PRECISION_MIXED else "double"
    if tile == 8:
        lmax = int(round((math.sqrt(8 * s.n_modes + 1) - 3) / 2))  # K = (l_max + 1)(l_max + 2) / 2
        cls = max(4, (lmax + 3) // 4 * 4)
        return "k_tapply<%d>" % cls, True
    return "k_apply<%d, %s>" % (max(tile, 16) // 16, t), False


def rooflines(s, prof, K, N, iters, mixed):
    """The dominant kernel's roofline and the achieved rates of the other
    kernel classes, from HIP-event times on the solver's stream.

    Plain FP64: the coupling apply k_apply, whose working launches (two per
    BiCGSTAB iteration plus the rhs and acceptance applies of every step) all
    do the same FP64 coupling work. Mixed: the FP32 Krylov applies (two per
    iteration) against the FP32 matrix peak. BiCGSTAB iterations are enqueued
    in batches, so a few Krylov launches after convergence return at once
    (about 5 us); they are not counted."""
    terms, pairs = s.coupling_work()
    krylov = min(prof["apply_krylov"][1], 2 * int(sum(iters)))
    kernel, polar = apply_kernel(s, gm.GM_PRECISION_MIXED if mixed else gm.GM_PRECISION_FP64)
    # the Gaunt form's FP32 apply runs on the FP32 matrix core; the polar
    # form runs in FP64 in both precisions (its transforms on the FP64 matrix
    # core, its banded product on the VALU)
    if mixed:
        ms, cnt = prof["apply_krylov"][0], krylov
        peak = FP64_PEAK_TFLOPS if polar else FP32_PEAK_TFLOPS
        name, dt = kernel + " (Krylov applies)", "f64" if polar else "f32"
    else:
        ms = prof["apply_krylov"][0] + prof["apply_other"][0]
        cnt = krylov + prof["apply_other"][1]
        peak, name, dt = FP64_PEAK_TFLOPS, kernel, "f64"
    avg_s = ms / max(cnt, 1) * 1e-3
    flops = algorithmic_flops(K, N)
    roof = {"bound": "mfma", "achieved": flops / avg_s / 1e12, "peak": peak, "unit": "TFLOP/s",
            "kernel": name, "dtype": dt, "per_launch": flops, "per_launch_unit": "flop",
            "per_launch_formula": "4 K^2 N + 12 K N (real symmetric K x K coupling times a complex vector, "
                                  "plus stencil and phases)",
            "avg_launch_ms": avg_s * 1e3, "launches": cnt}
    roof["frac"] = roof["achieved"] / roof["peak"]
    # what the apply executes instead: the Gaunt form pads the Gaunt terms
    # kept by the multipole cut onto the matrix core; the polar form runs its
    # transforms there (from the committed SQ counters)
    macs = mfma_macs(kernel)
    if polar:
        roof["executed"] = {"form": "polar quadrature", "mfma_macs_per_launch": macs,
                            "mfma_TFLOPs": (2 * macs / avg_s / 1e12) if macs else None,
                            "gaunt_terms_per_launch": terms}
    else:
        executed = executed_flops(terms, pairs, K, N)
        roof["executed"] = {"form": "Gaunt", "flops_per_launch": executed, "gaunt_terms_per_launch": terms,
                            "mode_pairs_per_launch": pairs, "achieved_TFLOPs": executed / avg_s / 1e12,
                            "mfma_macs_per_launch": macs,
                            "mfma_padding": (macs / terms) if macs else None}
    roof["traffic"] = measured_traffic("k_tapply" if polar else "k_apply" + ("_fp32" if mixed else ""))
    # whole step: the sum over kernels of max(bytes / HBM peak, flops / MFMA
    # peak) against the measured time of the profiled steps
    nsteps = max(len(iters), 1)
    nk = int(sum(iters))
    vec = 16.0 * K * N
    apply_bytes = 3.0 * vec
    b64 = max(apply_bytes / (HBM_PEAK_GBS * 1e9), flops / (FP64_PEAK_TFLOPS * 1e12))
    bk = max(apply_bytes / (HBM_PEAK_GBS * 1e9), flops / (peak * 1e12))
    bound_s = 2 * nk * bk + 2 * nsteps * b64
    for kname, passes, launches in (("precond", HBM_PASSES["precond"], 2 * nk), ("update", HBM_PASSES["update"], nk),
                                    ("diag", HBM_PASSES["diag"], nsteps), ("init", 2.0, nsteps),
                                    ("accept", 3.0, nsteps), ("phase", 2.0, 2 * nsteps)):
        bound_s += launches * passes * vec / (HBM_PEAK_GBS * 1e9)
    measured_s = sum(v[0] for v in prof.values()) * 1e-3
    roof["step"] = {"bound_ms_per_step": bound_s / nsteps * 1e3, "kernel_ms_per_step": measured_s / nsteps * 1e3,
                    "frac": bound_s / measured_s if measured_s > 0 else None,
                    "model": "sum over kernels of max(algorithmic bytes / 8 TB/s, coupling flops / MFMA peak)"}
    kernels = {}
    working = {"precond": min(prof["precond"][1], 2 * int(sum(iters))),
               "update": min(prof["update"][1], int(sum(iters)))}
    for kname, passes in HBM_PASSES.items():
        kms, kn = prof[kname][0], working.get(kname, prof[kname][1])
        if kn == 0:
            continue
        byts = passes * 16.0 * K * N
        a = byts / (kms / kn * 1e-3) / 1e9
        kernels[kname] = {"bound": "hbm", "avg_launch_ms": kms / kn, "bytes_per_launch": byts,
                          "achieved_GBs": a, "frac": a / HBM_PEAK_GBS}
    return roof, kernels


def accuracy(w, steps):
    """Mixed precision against the plain FP64 GPU step after `steps` steps
    from the same initial state, both advanced in lockstep: relative L2
    difference of psi and of each U_lm, the drift of the cloud mass, the flux
    and energy-flux error at r_ex, the energy, the binary power, the absorbed
    mass, and the BiCGSTAB iteration counts of both."""
    f64, mix = make_solver(w, gm.GM_PRECISION_FP64), make_solver(w, gm.GM_PRECISION_MIXED)
    it64, itmx = [], []
    d64 = dmx = None
    for k in range(1, steps + 1):
        last = k == steps
        d64 = f64.step(k, diag=last)
        dmx = mix.step(k, diag=last)
        it64.append(f64.last_solver_iterations)
        itmx.append(mix.last_solver_iterations)
    p64, pmx = f64.psi.reshape(f64.n_modes, -1), mix.psi.reshape(mix.n_modes, -1)
    del f64, mix
    n64 = np.linalg.norm(p64, axis=1)
    big = n64 > 1e-6 * n64.max()
    per_mode = np.linalg.norm(pmx - p64, axis=1)[big] / n64[big]

    def vec_err(key):
        a, b = np.asarray(dmx[key]), np.asarray(d64[key])
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    def rel_err(key):
        return float(abs(dmx[key] - d64[key]) / max(abs(d64[key]), 1e-300))

    return {"steps": steps,
            "psi_rel_l2": float(np.linalg.norm(pmx - p64) / np.linalg.norm(p64)),
            "U_lm_rel_l2_max": float(per_mode.max()), "U_lm_rel_l2_median": float(np.median(per_mode)),
            "modes_compared": int(big.sum()),
            "cloud_mass_rel": rel_err("cloud_mass"), "energy_rel": rel_err("energy"),
            "p_bin_inst_rel": rel_err("p_bin_inst"), "absorbed_rel": rel_err("absorbed"),
            "flux_err_rel_max": vec_err("flux"), "energy_flux_err_rel_max": vec_err("energy_flux"),
            "norms_err_rel_max": vec_err("norms"),
            "iterations_fp64_mean": float(np.mean(it64)), "iterations_mixed_mean": float(np.mean(itmx)),
            "iterations_max_abs_diff": int(np.max(np.abs(np.array(it64) - np.array(itmx)))),
            "reference": "plain FP64 GPU step (itself within 1e-10 of the FP64 oracle per step)"}


def large_case(steps, warmup, accuracy_steps):
    """The second size of the mixed-precision study: l_max 30 (496 modes),
    N 32768, L 8000, otherwise the l_max 20 workload (same perturbation
    recipe). Timed steps of both precisions, then the lockstep comparison."""
    w = dict(WORKLOAD, Nr=32768, L=8000.0, l_max=30)
    out = {"config": "l_max=30 (K=496), N=32768, L=8000, q=0.5, R=50, mu=0.15, dt=0.25, fixed orbit, "
                     "absorbers on, mask"}
    for name, precision in (("fp64", gm.GM_PRECISION_FP64), ("mixed", gm.GM_PRECISION_MIXED)):
        s, elapsed, _, iters, _ = run_gpu(w, precision, steps, warmup, prof_steps=0)
        out[name] = {"value": s.n_modes * s.Nr * steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
                     "iterations_per_step": float(np.mean(iters)), "apply_tile": s.apply_config(precision)[0]}
        del s
    out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    if accuracy_steps > 0:
        out["mixed_vs_fp64"] = accuracy(w, accuracy_steps)
    return out


SERVE_KEYS = ("submitted", "completed", "steps", "p50_ms", "p99_ms", "max_ms", "req_p50_ms", "req_p99_ms",
              "req_max_ms", "mode_node_steps_per_s", "active", "queued", "wall_seconds")


def serving(seconds):
    """The served stream: Poisson arrivals of binaries drawn as in the
    ensemble, back-reaction on, cloud mass drawn, two-period startup ramp,
    200-2000 steps, mixed precision, admission by HBM footprint (160 GB, at
    most 16 active). First at 20 requests/s as specified, which is beyond
    one GPU: per-step latency and throughput under saturation. Then at about
    half of the measured capacity, drained, so that requests finish: request
    latency from submission (queueing included) to completion, and every
    20th request (5 %, among those whose FP64 oracle fits a bounded host
    budget) spot-checked against the oracle on its state after two steps."""
    from gm import serve
    sat = serve.served_stream(duration=seconds, rate=20.0, seed=0, hbm_budget=160e9, max_active=16,
                              probe_every=0)
    out = {"config": "Poisson arrivals; q log-uniform [0.05, 1], R [20, 60], mu [0.1, 0.3], N in {4096..32768} "
                     "(L = N/4), l_max in {8..24}, M_c [0.001, 0.05], back-reaction on, 2-period ramp, "
                     "200-2000 steps of dt 0.25, mixed precision; HBM budget 160 GB, 16 active at most",
           "as_specified": dict({k: sat[k] for k in SERVE_KEYS}, arrivals_per_s=20.0, seconds=seconds)}
    # capacity: mode-node-steps/s of the saturated stream over the mean work of a request
    rng = np.random.default_rng(123)
    works = []
    for k in range(2000):
        params, n, _ = serve.draw_request(rng, k)
        works.append((params["l_max"] + 1) * (params["l_max"] + 2) / 2 * params["Nr"] * n)
    capacity = sat["mode_node_steps_per_s"] / float(np.mean(works))
    rate = 0.5 * capacity
    dur = max(seconds, 6.0 / rate) if rate > 0 else seconds
    sub = serve.served_stream(duration=dur, rate=rate, seed=1, hbm_budget=160e9, max_active=16, drain=True,
                              drain_limit=90.0, probe_every=20, probe_step=2, probe_max_work=6e8)
    out["below_capacity"] = dict({k: sub[k] for k in SERVE_KEYS}, arrivals_per_s=rate,
                                 capacity_requests_per_s=capacity, seconds=dur)
    out["below_capacity"]["oracle_spot_checks"] = spot_checks(sub["probes"])
    return out


def spot_checks(probes):
    """The FP64 oracle on the inputs of each probed request, two steps,
    against the state the server reported after its second step."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    from gm import serve
    res = []
    for rec in probes:
        if rec["psi"] is None:
            res.append({"ticket": rec["ticket"], "checked": False, "reason": "probe step not reached"})
            continue
        p = rec["params"]
        t0 = time.perf_counter()
        o = oracle.Oracle(**p)
        f = serve.default_flags()
        o.set_flags(**{k: getattr(f, k) for k, _ in gm.capi.Flags._fields_})
        o.set_initial_state_full(rec["psi0"])
        o.initialize()
        o.build_damping_mask(0.85, 1.0, 6)
        o.begin_run(0.8 * p["L"], rec["n_steps"] + 1)
        for n in range(1, rec["probe_step"] + 1):
            o.step(n, diag=False)
        ref = o.psi
        del o
        res.append({"ticket": rec["ticket"], "checked": True, "N": p["Nr"], "l_max": p["l_max"],
                    "steps": rec["probe_step"], "psi_rel_l2": float(np.linalg.norm(rec["psi"] - ref) /
                                                                   np.linalg.norm(ref)),
                    "oracle_seconds": time.perf_counter() - t0})
    return res


def measured_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (scaled by a copy-kernel calibration), or None."""
    path = os.path.join(REPO, "profiles", "k_apply_traffic_l20_n16384.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch " + kernel + " "):
                    return float(line.split()[2])
    except OSError:
        pass
    return None


def copy_calibration(nbytes=1 << 29, reps=5):
    """A plain device copy of known size for scaling the FETCH_SIZE and
    WRITE_SIZE counters (run under rocprofv3 --pmc)."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps, "fetch_kib_expected": nbytes / 1024,
                      "write_kib_expected": nbytes / 1024}))


# ---------------------------------------------------------------- rocprofv3 summaries

def _short_kernel(name):
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    name = name.replace("gmk::", "")
    i = name.find("(")
    return name[:i] if i > 0 else name


def _read_csv(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def trace_launches(path):
    """kernel -> list of launch durations (us) from a --kernel-trace csv."""
    out = {}
    for r in _read_csv(path):
        k = _short_kernel(r["Kernel_Name"])
        out.setdefault(k, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    return out


def working_us(durs, floor_us=50.0):
    w = [d for d in durs if d >= floor_us]
    return w


def counters(path):
    """kernel -> list of {counter: value} per dispatch from a --pmc csv."""
    per = {}
    for r in _read_csv(path):
        key = (r.get("Dispatch_Id") or r.get("Correlation_Id"), _short_kernel(r["Kernel_Name"]))
        per.setdefault(key, {})[r["Counter_Name"]] = float(r["Counter_Value"])
    out = {}
    for (_, k), d in per.items():
        out.setdefault(k, []).append(d)
    return out


def _mean(v):
    return sum(v) / len(v) if v else float("nan")


def summary_trace(trace_csv):
    t = trace_launches(trace_csv)
    total = sum(sum(v) for v in t.values())
    lines = ["kernel                                 calls     total_ms       avg_us     pct |  working  avg_work_us"]
    for k, v in sorted(t.items(), key=lambda kv: -sum(kv[1])):
        w = working_us(v)
        lines.append(f"{k:38s} {len(v):6d} {sum(v) / 1e3:12.3f} {_mean(v):12.1f} {100 * sum(v) / total:7.2f} | "
                     f"{len(w):8d} {_mean(w):12.1f}")
    lines.append("")
    lines.append("working = launches of at least 50 us (launches enqueued after BiCGSTAB convergence return at once)")
    return "\n".join(lines) + "\n"


def _traffic(fetch_csv, write_csv, cfetch_csv, cwrite_csv):
    f, w = counters(fetch_csv), counters(write_csv)
    cf, cw = counters(cfetch_csv), counters(cwrite_csv)
    copy = [k for k in cf if "copyBuffer" in k]
    moved_kib = (1 << 29) / 1024.0
    fs = moved_kib / _mean([d["FETCH_SIZE"] for d in cf[copy[0]]]) if copy else 1.0
    ws = moved_kib / _mean([d["WRITE_SIZE"] for d in cw[copy[0]]]) if copy else 1.0
    rows = {}
    for k in f:
        fv = [d["FETCH_SIZE"] for d in f[k] if d["FETCH_SIZE"] > 1024.0]
        wv = [d["WRITE_SIZE"] for d in w.get(k, [])]
        wv = sorted(wv)[-len(fv):] if fv and wv else wv
        if not fv:
            continue
        fk, wk = _mean(fv), _mean(wv) if wv else 0.0
        rows[k] = (len(fv), fk, wk, (fk * fs + wk * ws) * 1024.0)
    return rows, fs, ws, (_mean([d["FETCH_SIZE"] for d in cf[copy[0]]]) if copy else None), \
        (_mean([d["WRITE_SIZE"] for d in cw[copy[0]]]) if copy else None)


def _main_variant(rows, prefix, dtype):
    cands = [(v[0], k) for k, v in rows.items() if k.startswith(prefix + "<") and k.endswith(dtype + ">")]
    return max(cands)[1] if cands else None


def summary_traffic(fetch_csv, write_csv, cfetch_csv, cwrite_csv):
    rows, fs, ws, cfk, cwk = _traffic(fetch_csv, write_csv, cfetch_csv, cwrite_csv)
    lines = []
    if cfk is not None:
        lines.append(f"calibration: a 512 MiB device copy (__amd_rocclr_copyBuffer) reads FETCH_SIZE {cfk:.1f} KiB and "
                     f"WRITE_SIZE {cwk:.1f} KiB per launch against 524288 KiB moved each way")
    lines.append(f"scale factors: fetch x{fs:.3f}, write x{ws:.3f}")
    lines.append("")
    lines.append("kernel                                 working      fetch_KiB      write_KiB  bytes(scaled)")
    for k, (n, fk, wk, b) in sorted(rows.items(), key=lambda kv: -kv[1][3] * kv[1][0]):
        lines.append(f"{k:38s} {n:8d} {fk:14.1f} {wk:14.1f} {b:14.4g}")
    lines.append("")
    lines.append("working = launches that fetch more than 1 MiB (launches after convergence return at once)")
    for key, prefix, dt in (("k_apply", "k_apply", "double"), ("k_apply_fp32", "k_apply", "float"),
                            ("k_tapply", "k_tapply", "")):
        k = _main_variant(rows, prefix, dt)
        if k:
            lines.append(f"traffic_bytes_per_launch {key} {rows[k][3]:.6g}")
    return "\n".join(lines) + "\n"


def summary_rates(trace_csv, fetch_csv, write_csv, cfetch_csv, cwrite_csv):
    t = trace_launches(trace_csv)
    rows, fs, ws, _, _ = _traffic(fetch_csv, write_csv, cfetch_csv, cwrite_csv)
    lines = ["kernel                                    bytes    time_us       GB/s  of_peak"]
    out = []
    for k, (n, fk, wk, b) in rows.items():
        if k not in t:
            continue
        w = working_us(t[k], 10.0) or t[k]
        us = _mean(w)
        out.append((b / us / 1e3, k, b, us))
    for gbs, k, b, us in sorted(out, reverse=True):
        lines.append(f"{k:38s} {b:10.4g} {us:10.1f} {gbs:10.0f} {100 * gbs / HBM_PEAK_GBS:7.1f}%")
    return "\n".join(lines) + "\n"


SQ1 = ("SQ_WAVES", "SQ_INSTS_MFMA", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD", "SQ_WAVE_CYCLES",
       "SQ_WAIT_INST_ANY")
SQ2 = ("GRBM_GUI_ACTIVE", "SQ_VALU_MFMA_BUSY_CYCLES", "SQ_BUSY_CU_CYCLES", "SQ_INSTS_LDS", "SQ_INSTS_SMEM")


def summary_sq(pass1_csv, pass2_csv, trace_csv):
    t = trace_launches(trace_csv)
    tabs = []
    vals = {}
    for path, names, weight in ((pass1_csv, SQ1, "SQ_INSTS_VALU"), (pass2_csv, SQ2, "SQ_BUSY_CU_CYCLES")):
        c = counters(path)
        head = "kernel                                      ms " + "".join(f"{n[3:] if n.startswith('SQ_') else n:>22s}"
                                                                             for n in names)
        lines = [head]
        for k in sorted(c, key=lambda k: -_mean([d.get(weight, 0.0) for d in c[k]])):
            ds = [d for d in c[k] if weight in d]
            if not ds:
                continue
            mx = max(d[weight] for d in ds)
            ds = [d for d in ds if d[weight] >= 0.25 * mx]
            if mx <= 0 or not (k.startswith("k_apply<") or k.startswith("k_tapply<") or k.startswith("k_psweep<") or
                               k in ("k_update", "k_diag_apply")):
                continue
            w = working_us(t.get(k, []), 10.0)
            ms = _mean(w) / 1e3 if w else float("nan")
            row = {n: _mean([d.get(n, 0.0) for d in ds]) for n in names}
            vals.setdefault(k, {}).update(row)
            lines.append(f"{k:38s} {ms:10.4f}" + "".join(f"{row[n]:22.4g}" for n in names))
        tabs.append("\n".join(lines))
    notes = []
    simds, xcds = 1024, 8
    for k, v in sorted(vals.items()):
        if not (k.startswith("k_apply<") or k.startswith("k_tapply<")) or "GRBM_GUI_ACTIVE" not in v or \
                "SQ_INSTS_MFMA" not in v:
            continue
        cyc = v["GRBM_GUI_ACTIVE"] / xcds
        # issue cycles per MFMA: 16x16x4 f64 64, 16x16x4 f32 32, 4x4x4 f64 (polar form) 16
        per = 16 if k.startswith("k_tapply<") else (64 if k.endswith("double>") else 32)
        notes.append(f"{k}: {cyc:.0f} cycles per XCD; matrix core busy {100 * v['SQ_INSTS_MFMA'] * per / (cyc * simds):.0f}% "
                     f"of each SIMD's cycles ({per} cycles per MFMA); VALU at 4 cycles per wave64 instruction "
                     f"{100 * v['SQ_INSTS_VALU'] * 4 / (cyc * simds):.0f}%; mean resident waves per SIMD "
                     f"{v['SQ_WAVE_CYCLES'] * 4 / (cyc * simds):.1f}; waves waiting on s_waitcnt "
                     f"{100 * v['SQ_WAIT_INST_ANY'] / v['SQ_WAVE_CYCLES']:.0f}% of their cycles")
    return "\n\n".join(tabs) + "\n\n" + "\n".join(notes) + "\n"


def rocprof_summary(argv):
    kind, rest = argv[0], argv[1:]
    if kind == "trace":
        return summary_trace(*rest)
    if kind == "traffic":
        return summary_traffic(*rest)
    if kind == "rates":
        return summary_rates(*rest)
    if kind == "sq":
        return summary_sq(*rest)
    raise SystemExit("unknown summary kind " + kind)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="untimed steps of the mixed-vs-FP64 comparison (0 skips it)")
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the batched ragged ensemble measurement (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=500,
                    help="timed steps of the batched ensemble (after 3 warmup steps); the per-member-stream "
                         "comparison runs a tenth of them")
    ap.add_argument("--ensemble-lanes", type=int, default=4,
                    help="member subsets of the batched ensemble, each on its own stream")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="length of the served Poisson stream for p50/p99 step latency (0 skips it)")
    ap.add_argument("--large-steps", type=int, default=10,
                    help="timed steps per precision at l_max 30, N 32768 (0 skips the second size)")
    ap.add_argument("--large-accuracy-steps", type=int, default=100,
                    help="lockstep mixed-vs-FP64 steps at l_max 30, N 32768 (0 skips them)")
    ap.add_argument("--evolving-steps", type=int, default=50,
                    help="timed steps per precision with the orbit evolving (0 skips them)")
    ap.add_argument("--apply-tile", type=int, default=0, choices=(0, 8, 16, 32, 64),
                    help="fix the coupling apply's tile in nodes (0: autotune; profiler runs fix the tuned one, "
                         "since counter collection serialises the tuner's timings)")
    ap.add_argument("--headline-batch", type=int, default=8,
                    help="binaries of the bench workload in the mixed-precision batched step (0 skips it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="timed steps without graph replay (for kernel-trace profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 512 MiB device copy (counter calibration)")
    ap.add_argument("--rocprof-summary", nargs="+", metavar="ARG",
                    help="summarise rocprofv3 csv output instead of measuring: trace TRACE_CSV | "
                         "traffic FETCH_CSV WRITE_CSV COPY_FETCH_CSV COPY_WRITE_CSV | sq PASS1_CSV PASS2_CSV "
                         "TRACE_CSV | rates TRACE_CSV FETCH_CSV WRITE_CSV COPY_FETCH_CSV COPY_WRITE_CSV")
    args = ap.parse_args()
    if args.rocprof_summary:
        print(rocprof_summary(args.rocprof_summary), end="")
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: every box here has one MI355X and this build "
              "has no multi-GPU path (ensembles would shard whole members per GPU).", file=sys.stderr)
        sys.exit(2)
    w = dict(WORKLOAD)
    runs = {}
    for name, precision in (("fp64", gm.GM_PRECISION_FP64), ("mixed", gm.GM_PRECISION_MIXED)):
        s, elapsed, prof, iters, piters = run_gpu(w, precision, args.steps, args.warmup, graphs=not args.no_graphs,
                                                  tile=args.apply_tile)
        K, N = s.n_modes, s.Nr
        roof, kernels = rooflines(s, prof, K, N, piters, name == "mixed")
        runs[name] = {"value": K * N * args.steps / elapsed, "ms_per_step": elapsed / args.steps * 1e3,
                      "iterations_per_step": float(np.mean(iters)) if iters else None, "roofline": roof,
                      "kernels": kernels, "profiled_steps": len(piters),
                      "kernel_ms": {k: round(v[0], 4) for k, v in prof.items()},
                      "kernel_launches": {k: int(v[1]) for k, v in prof.items()}}
        tile, tms = s.apply_config(precision)
        runs[name]["apply_tile"] = {"nodes": tile, "tuning_ms": tms}
        progress("%s leg: %.4f ms per step" % (name, runs[name]["ms_per_step"]))
        terms, pairs = s.coupling_work()
        del s
    mx, f64 = runs["mixed"], runs["fp64"]
    out = {
        "metric": METRIC, "value": mx["value"], "unit": "mode·node-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": mx["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "CN step, l_max=20 (K=231), N=16384, L=4000, q=0.5, R=50, mu=0.15, dt=0.25, "
                               "fixed orbit, absorbers on, mask; value = mixed precision (FP32 preconditioner "
                               "sweeps; the Krylov applies' coupling in FP32 in a Gaunt tile, in FP64 in the "
                               "polar form the autotuner picks here), fp64 = plain FP64 drop-in", "K": K, "N": N,
                   "parallelism": "single GPU"},
        "roofline": mx["roofline"],
        "fp64": f64,
        "mixed": {k: v for k, v in mx.items() if k != "roofline"},
        "mixed_over_fp64": mx["value"] / f64["value"],
        "p99_served_step_latency_ms": None,
        "coupling_work": {"terms_per_apply": terms, "pairs_per_apply": pairs,
                          "untruncated_terms_per_apply": float(coupling_terms(w["l_max"])) * (N - 2)},
    }
    if args.headline_batch > 0:
        # the single legs' window: the same warmup and timed steps
        out["batched"] = headline_batch(w, args.headline_batch, args.steps, args.warmup, f64["value"],
                                        graphs=not args.no_graphs)
        progress("batched headline step: %.1f ms per batch step" % out["batched"]["ms_per_batch_step"])
    if args.evolving_steps > 0:
        out["evolving_orbit"] = evolving(w, args.evolving_steps, 5, runs, graphs=not args.no_graphs,
                                         tile=args.apply_tile)
        progress("evolving-orbit leg done")
    if args.accuracy_steps > 0:
        out["mixed_vs_fp64"] = accuracy(w, args.accuracy_steps)
        progress("mixed-vs-FP64 comparison done")
    if args.large_steps > 0:
        out["second_size"] = large_case(args.large_steps, 2, args.large_accuracy_steps)
        progress("second size done")
    if args.ensemble_members > 0:
        from gm import ensemble
        cfg = ("ragged fixed-orbit binaries from seed 0: q log-uniform [0.05, 1], R [20, 60], mu [0.1, 0.3], "
               "N in {4096..32768} (L = N/4), l_max in {8..24}, 2p atom on the primary + 1e-3 perturbation; "
               "mixed precision")
        ens = ensemble.run(n=args.ensemble_members, steps=args.ensemble_steps, warmup=3, seed=0, mode="batch",
                           lanes=args.ensemble_lanes)
        ens["config"] = cfg + ("; every launch of the step one ragged grid over the members of a lane (%d lanes "
                               "of balanced work, each on its own stream), graph-captured" % args.ensemble_lanes)
        ens["over_fp64_single_step"] = ens["mode_node_steps_per_s"] / f64["value"]
        progress("batched ensemble: %d steps, %.1f ms per ensemble step" % (args.ensemble_steps,
                                                                            ens["ms_per_ensemble_step"]))
        streams = ensemble.run(n=args.ensemble_members, steps=max(1, args.ensemble_steps // 10), warmup=3, seed=0,
                               mode="streams")
        streams["over_fp64_single_step"] = streams["mode_node_steps_per_s"] / f64["value"]
        ens["per_member_streams"] = streams
        progress("per-member-stream ensemble done")
        out["ensemble"] = ens
    if args.serve_seconds > 0:
        out["serving"] = serving(args.serve_seconds)
        progress("serving done")
        out["p99_served_step_latency_ms"] = out["serving"]["as_specified"]["p99_ms"]
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(w)
        except Exception as e:  # the baseline never hides the GPU number
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
