"""Benchmark: voxel FOD solves/s of the mFOD multi-tissue fit (lmax 8, 3 shells) on one MI355X.

Workload (synthetic, seeded): 96 x 96 x 60 voxels with an ellipsoidal mask holding ~40% of
them; 10 b0 plus 64 directions at each of b = 1000, 2000, 3000 s/mm^2; white matter with 1-3
crossing fibres (tensor response 1.7 / 0.3 um^2/ms), grey matter D = 0.8e-3 and CSF
D = 3.0e-3 mixed by Dirichlet(2, 1, 0.5) fractions; Rician noise at SNR 25.

One step = the whole masked volume through the hot path with the normalised signals already in
HBM: the low-resolution FOD <-> fraction rounds, the high-resolution FOD and final fractions
and the projection of the amplitudes onto SH up to lmax 8 (all liblarmor_sd.so kernels).

Prints one JSON line.  `roofline` prices the dominant kernel (the one with the most time per
step) at its algorithmic FP64 flops against MI355X's FP64 peak, from the kernel's own launch
durations measured with HIP events on its launch stream.  The algorithmic flops of a kernel:
- fod_lr_corr: 2 m n_lr per voxel and round (the correlations K^T a of the low-resolution
  design);
- fod_lr: the active-set work of its NNLS solves, which the kernel counts by formula as they
  run (per append k^2 + 4k + 4, per solve k^2 plus the forward part, 2 n per non-zero
  coefficient of each gradient, 3 n per inner step, 8 + 6 rows per Givens rotation; k the
  factor size) and reports through the library's profiling hook;
- fraction_loop: 2 m n_lr per voxel and round (the group column K x);
- fod_hr_corr: 2 m n_hr per voxel (the correlations of the high-resolution design);
- fod_hr: its counted active-set work;
- fraction_final: 2 m n_hr per voxel.
The low-resolution kernels run once per voxel and round, so their unit count is the sum of the
voxels' rounds.  `traffic` is the kernel's HBM bytes per launch from the committed rocprofv3
counter summary (null when it has no entry).
`latency_under_load` is the served form's p50/p99 request latency: 32 requests of the batched
workload (own acquisitions, 5k-400k voxels each) arriving as a Poisson stream at 5 jobs/s,
cut into 64k-row slices fitted in batched launch sequences across acquisitions (--no-serve
leaves it out; --serve reports every phase and rate: plain per-job, batched, served).
`cpu_baseline` is the CPU
FP64 oracle (a C++ restatement of the reference algorithm, `kind: port`) on a bounded random
sample of the same voxels, on the host's cores.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector
HBM_PEAK_GBS = 8000.0


WORKLOADS = {
    # name: (config, lmax, description)
    "mfod3": ("MFOD_CONFIG", 8, "mFOD multi-tissue fit, 96x96x60 volume, ellipsoid mask ~40% "
              "({n} voxels), 10 b0 + 3x64 directions (b 1000/2000/3000), WM 1-3 fibres + GM 0.8e-3 "
              "+ CSF 3.0e-3, lmax 8, 300 HR / 44 LR directions"),
    "mfod4": ("PRECISION_CONFIG", 10, "mFOD multi-tissue fit, the masked voxels of a 140x140x96 "
              "volume with an ellipsoid mask ~45% ({n} voxels), 12 b0 + 4x90 directions "
              "(b 700/1000/2000/3000), WM 1-3 fibres + GM 0.8e-3 + CSF 3.0e-3, lmax 10, 300 HR / "
              "44 LR directions"),
}


def build_workload(seed_override=None, workload="mfod3"):
    """The signals (normalised, float64), shell weights and the LR / HR kernels of a workload.
    mfod3 is the seeded phantom volume itself; mfod4's 846,728 masked voxels are drawn directly
    from the same voxel model (1-3 fibres, Dirichlet fractions, Rician noise) on the device."""
    from larmor_amd import synthetic
    from larmor_amd.sd.kernel import Kernel
    from larmor_amd.sd.model import compute_shell_weights, low_resolution_dirs
    from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
    from larmor_amd.sd.sphere import DIRS_300

    cfg = dict(getattr(synthetic, WORKLOADS[workload][0]))
    if seed_override is not None:
        cfg["seed"] = seed_override
    if workload == "mfod3":
        ph = synthetic.phantom(**cfg)
        signals = synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)
        bvals, bvecs = ph.bvals, ph.bvecs
    else:
        import torch

        rng = np.random.default_rng(cfg["seed"])
        bvals, bvecs = synthetic.acquisition(cfg["n_b0"], cfg["shells"], rng)
        n_vox = int(synthetic.ellipsoid_mask(cfg["shape"], cfg["mask_share"]).sum())
        device = "cuda" if torch.cuda.is_available() else "cpu"
        signals = np.ascontiguousarray(
            synthetic.tissue_signals(n_vox, bvals, bvecs, rng, snr=cfg["snr"], device=device),
            dtype=np.float64)
    rfs = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3.0e-3)]
    w = compute_shell_weights(bvals, 1.0)
    hr = Kernel(rfs, bvals, bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, bvals, bvecs, dirs=low_resolution_dirs(rfs, len(bvals)), weights=w)
    return cfg, signals, w, lr, hr


def host_cores() -> int:
    """The host cores this process may run on: its CPU affinity set, capped by OMP_NUM_THREADS
    when the job's CPU share is set that way (a shared GPU host grants each one-GPU job a share
    of its cores)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cores = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        cores = min(cores, int(share))
    return max(1, cores)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(signals, w, lr, hr, target_seconds=15.0):
    """Times the oracle on a random sample sized for about target_seconds of CPU work, on every
    host core this process may use."""
    import oracle

    threads = host_cores()
    rng = np.random.default_rng(0)
    probe = np.ascontiguousarray(signals[rng.choice(len(signals), size=min(1000, len(signals)),
                                                    replace=False)])
    t0 = time.perf_counter()
    oracle.fit_multi_tissue(probe, w, lr, hr, 0.0, 0.1, 1e-6, n_threads=threads)
    rate = len(probe) / (time.perf_counter() - t0)
    n = int(min(len(signals), max(2000, rate * target_seconds)))
    sample = np.ascontiguousarray(signals[rng.choice(len(signals), size=n, replace=False)])
    t0 = time.perf_counter()
    oracle.fit_multi_tissue(sample, w, lr, hr, 0.0, 0.1, 1e-6, n_threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": n / dt,
        "unit": "voxel solves/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} random masked voxels of {len(signals)} ({dt:.1f} s)",
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
    }


TRAFFIC_PROFILES = {
    ("mfod3", "fp64"): "mfod_3shell_traffic.json",
    ("mfod3", "mixed"): "mfod_3shell_mixed_traffic.json",
    ("mfod3", "fp32"): "mfod_3shell_fp32_traffic.json",
    ("mfod4", "fp64"): "mfod_4shell_traffic.json",
    ("mfod4", "mixed"): "mfod_4shell_mixed_traffic.json",
}


def measured_traffic(kernel: str, workload: str = "mfod3", form: str = "fp64"):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of the workload and form (profiles/*_traffic.json, counters scaled by a 1 GiB copy),
    or None."""
    name = TRAFFIC_PROFILES.get((workload, form))
    if name is None:
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(path) as fh:
            entry = json.load(fh)["kernels"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if entry is None else entry["hbm_bytes_per_dispatch"]


def copy_probe(mib: int) -> int:
    """Copies mib MiB device to device three times; each copy reads and writes mib MiB."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x 3", file=sys.stderr, flush=True)
    return 0


def _cpu_rate(fn, n, threads):
    t0 = time.perf_counter()
    fn()
    return n / (time.perf_counter() - t0), threads


class ServeBench:
    """The batched and served forms' workload: n_jobs seeded requests, each a volume's masked
    voxels (log-uniform 5k-400k) with its own acquisition (10 b0 + 1-4 shells of 30-90
    directions, b 700-3000) and lmax (6, 8 or 10; lowered when the top shell is too small), mFOD
    (WM + GM + CSF), float32 normalised signals in host memory as a request would bring them;
    one FodServer with every request's design made resident first (prepare), warmed on the four
    smallest requests."""

    def __init__(self, n_jobs: int, slots: int, form: str = "fp64"):
        from larmor_amd import synthetic
        from larmor_amd.sd.model import SDModel
        from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
        from larmor_amd.sd.serving import FodServer
        from larmor_amd.types import GradientTable

        cfg = dict(synthetic.BATCHED_CONFIG)
        cfg["n_jobs"] = n_jobs
        t = time.perf_counter()
        self.jobs = synthetic.batched_jobs(device="cuda", **cfg)
        self.n_vox = sum(len(j.signals) for j in self.jobs)
        print(f"serving workload: {n_jobs} requests, {self.n_vox} voxels, generated in "
              f"{time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
        model = SDModel([AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3),
                         IsotropicRF(3.0e-3)], method="mFOD", lmax=8)
        self.gtabs = [GradientTable(j.bvals, j.bvecs) for j in self.jobs]
        self.slots = slots
        self.form = form
        self.server = FodServer(model, slots=slots, hbm_budget=200e9, form=form)
        t = time.perf_counter()
        for job, gtab in zip(self.jobs, self.gtabs):
            self.server.prepare(gtab, lmax=job.lmax)
        print(f"designs resident: {n_jobs} in {time.perf_counter() - t:.1f} s", file=sys.stderr,
              flush=True)
        smallest = sorted(range(n_jobs), key=lambda i: len(self.jobs[i].signals))[:4]
        for f in [self._submit(i) for i in smallest]:
            f.result()

    def _submit(self, i):
        return self.server.submit(self.jobs[i].signals, self.gtabs[i], lmax=self.jobs[i].lmax)

    def configure(self, *, batch_jobs: int, slice_voxels: int, merge_voxels: int) -> dict:
        """Sets the server's fit policy for the next phase (read at each admission)."""
        if self.form != "fp64":
            batch_jobs = 1  # the mixed form is never batched across acquisitions
        self.server.batch_jobs = batch_jobs
        self.server.slice_voxels = slice_voxels
        self.server.merge_voxels = merge_voxels
        self.server.max_merged = 1
        return {"batch_jobs": batch_jobs, "slice_voxels": slice_voxels,
                "merge_voxels": merge_voxels}

    def run(self, name, arrivals, serial=False) -> dict:
        """Submits request i at each (offset seconds, i); returns the latency/throughput summary."""
        from larmor_amd.sd.serving import summarize

        futures = []
        t0 = time.perf_counter()
        for offset, i in arrivals:
            delay = t0 + offset - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futures.append(self._submit(i))
            if serial:
                futures[-1].result()
        summary = summarize([f.result() for f in futures])
        print(f"{name}: {json.dumps(summary)}", file=sys.stderr, flush=True)
        return summary

    def poisson(self, rate: float) -> dict:
        n = len(self.jobs)
        rng = np.random.default_rng(1000 + int(rate))
        offsets = np.concatenate([[0.0], np.cumsum(rng.exponential(1.0 / rate, n - 1))])
        return {"rate_jobs_per_s": rate,
                **self.run(f"served at {rate}/s", list(zip(offsets, rng.permutation(n))))}

    def workload(self) -> str:
        return (f"{len(self.jobs)} mFOD requests, voxels log-uniform 5k-400k ({self.n_vox} in "
                "all), each with its own acquisition (10 b0 + 1-4 shells of 30-90 directions, "
                "b 700-3000) and lmax 6/8/10, WM + GM 0.8e-3 + CSF 3.0e-3, float32 signals in "
                f"host memory; {self.slots} streams, 200 GB HBM budget; {self.form} form")

    def close(self) -> None:
        self.server.close()


def serve_report(n_jobs: int, rates: list[float], slots: int, form: str = "fp64") -> int:
    """The batched and served forms (--serve) on ServeBench's workload.  Phases: plain per-job
    (each request submitted when the previous one's results are back and fitted alone), batched
    (all submitted at once, fitted in batched launch sequences across acquisitions, the best of
    the tuned fit/slice policies) and served at each Poisson arrival rate (64k-row slices).
    Prints one JSON line with voxel solves/s and p50/p99 latency per phase."""
    bench = ServeBench(n_jobs, slots, form)
    order = list(range(n_jobs))
    # plain per-job: one request at a time, each fitted alone.
    policy = bench.configure(batch_jobs=1, slice_voxels=0, merge_voxels=1 << 18)
    plain = {**bench.run("plain per-job", [(0.0, i) for i in order], serial=True),
             "policy": policy}
    # batched: every request queued at once; the requests of each row-slot class fitted in
    # batched launch sequences of up to 64 acquisitions.  The fit size and the slice size are
    # tuned on the workload itself: each candidate policy serves the whole workload once and the
    # fastest is reported as the batched phase.
    # concurrent: every request queued at once, each fitted alone on one of the streams (no
    # batched launches): what the batched launches are measured against.
    policy = bench.configure(batch_jobs=1, slice_voxels=0, merge_voxels=1 << 18)
    concurrent = {**bench.run("concurrent", [(0.0, i) for i in order]), "policy": policy}
    candidates = []
    for merge_voxels in (1 << 19, 1 << 20, 1 << 21):
        for slice_voxels in (0, 1 << 16):
            policy = bench.configure(batch_jobs=64, slice_voxels=slice_voxels,
                                     merge_voxels=merge_voxels)
            run = bench.run(f"batched {policy}", [(0.0, i) for i in order])
            candidates.append({**run, "policy": policy, "largest_batch": bench.server.max_merged})
    batched = max(candidates, key=lambda c: c["voxel_solves_per_s"])
    tuning = [{**c["policy"], "voxel_solves_per_s": c["voxel_solves_per_s"]} for c in candidates]
    # Kernel autotuning per bucket (row class, low-resolution directions): the low-resolution
    # factor capacity timed on a sample of the bucket's first request, then the best batched
    # policy again with every fit at its bucket's choice.
    autotuned = None
    if form == "fp64":
        t = time.perf_counter()
        for job, gtab in zip(bench.jobs, bench.gtabs):
            bench.server.autotune(gtab, job.signals, lmax=job.lmax)
        tune_s = time.perf_counter() - t
        policy = bench.configure(**batched["policy"])
        autotuned = {**bench.run(f"batched, autotuned {policy}", [(0.0, i) for i in order]),
                     "policy": policy, "largest_batch": bench.server.max_merged,
                     "tuning_seconds": tune_s, "table": bench.server.tuning.as_rows()}
        if autotuned["voxel_solves_per_s"] < batched["voxel_solves_per_s"]:
            # Not faster on this run: the served phases keep the default shapes.
            from larmor_amd.sd import tuning as lr_tuning

            bench.server.tuning = None
            for rows in (256, 384, 512):
                lr_tuning.set_capacity(rows, 0)
    # served: Poisson arrivals; requests cut into slices of 64k rows, fits of up to 512k rows.
    policy = bench.configure(batch_jobs=64, slice_voxels=1 << 16, merge_voxels=1 << 19)
    served = [{**bench.poisson(rate), "policy": policy, "largest_batch": bench.server.max_merged}
              for rate in rates]
    bench.close()
    protocol = shared_protocol_phase(slots)
    result = {
        "metric": "voxel FOD solves/s (mFOD lmax=8, 3-shell); p50/p99 latency under load",
        "unit": "voxel solves/s",
        "dtype": {"fp64": "f64", "mixed": "bf16+f64"}[form],
        "data": "synthetic (seeded requests, Rician SNR 25)",
        "config": {"workload": bench.workload(), "slots": slots, "hbm_budget_bytes": 200e9},
        "plain_per_job": plain,
        "concurrent": concurrent,
        "batched": batched,
        "batched_tuning": tuning,
        "batched_autotuned": autotuned,
        "served": served,
        "batched_vs_plain": batched["voxel_solves_per_s"] / plain["voxel_solves_per_s"],
        "batched_vs_concurrent": batched["voxel_solves_per_s"] / concurrent["voxel_solves_per_s"],
        "shared_protocol": protocol,
    }
    print(json.dumps(result), flush=True)
    return 0


def shared_protocol_phase(slots: int) -> dict:
    """64 small requests (2k-20k voxels, log-uniform) on one acquisition (10 b0 + 3 x 64
    directions at b 1000/2000/3000), lmax 8, Poisson arrivals at 400 jobs/s, served with and
    without merging queued requests on the same acquisition into one fit."""
    import torch

    from larmor_amd import synthetic
    from larmor_amd.sd.model import SDModel
    from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
    from larmor_amd.sd.serving import FodServer, summarize
    from larmor_amd.types import GradientTable

    rng = np.random.default_rng(7)
    bvals, bvecs = synthetic.acquisition(10, [(1000.0, 64), (2000.0, 64), (3000.0, 64)], rng)
    sizes = np.exp(rng.uniform(np.log(2000), np.log(20000), 64)).astype(int)
    signals = [synthetic.tissue_signals(int(n), bvals, bvecs, rng, snr=25.0, device="cuda")
               for n in sizes]
    gtab = GradientTable(bvals, bvecs)
    offsets = np.concatenate([[0.0], np.cumsum(rng.exponential(1.0 / 400.0, 63))])
    model = SDModel([AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3),
                     IsotropicRF(3.0e-3)], method="mFOD", lmax=8)
    out = {"workload": f"64 requests of 2k-20k voxels ({int(sizes.sum())} in all) on one "
                       "acquisition (10 b0 + 3 x 64, b 1000/2000/3000), Poisson 400 jobs/s"}
    for name, merge in (("merged", 1 << 18), ("one_per_fit", 0)):
        with FodServer(model, slots=slots, merge_voxels=merge) as server:
            server.prepare(gtab, lmax=8)
            server.submit(signals[0][:500], gtab).result()
            futures = []
            t0 = time.perf_counter()
            for offset, sig in zip(offsets, signals):
                delay = t0 + offset - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                futures.append(server.submit(sig, gtab))
            out[name] = {**summarize([f.result() for f in futures]),
                         "largest_merge": server.max_merged}
        print(f"shared protocol, {name}: {json.dumps(out[name])}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    return out


def rows_report() -> int:
    """Times each row of the path beside its CPU oracle on a sample (--rows).  Prints one line per
    row: GPU voxels/s on the stated workload, the oracle's voxels/s on its sample with the host's
    threads, and the counted or dense FP64 rate of the row's kernels."""
    import torch

    import oracle
    from larmor_amd import _core, _lib, synthetic
    from larmor_amd.sd.csd_solver import CSDSolver, csd_problem
    from larmor_amd.sd.kernel import Kernel
    from larmor_amd.sd.model import compute_shell_weights, low_resolution_dirs
    from larmor_amd.sd.msmt import constraint_matrix, forward_matrix, msmt_ridge
    from larmor_amd.sd.peaks import _build_icosphere, _device_sphere, _get_amplitudes
    from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF, ZonalSHRF
    from larmor_amd.sd.sh import projection_matrix
    from larmor_amd.sd.solvers import (DRL, FINAL_FRACTION_TOL_X, MAX_ROUNDS, Lasso, Stage,
                                       _deconv_kernel)
    from larmor_amd.sd.sphere import DIRS_300

    threads = host_cores()
    rng = np.random.default_rng(0)
    lines = []

    def report(row, workload, n, gpu_s, cpu, kernels):
        profile = ", ".join(
            f"{k} {v[0]:.1f} ms" + (f" ({v[2] / (v[0] / 1e3) / 1e12:.2f} TFLOP/s counted)" if v[2] else "")
            for k, v in kernels.items())
        line = (f"| {row} | {workload} | {n / gpu_s:,.0f} | {cpu[0]:,.0f} ({cpu[1]} threads) | "
                f"{(n / gpu_s) / cpu[0]:.0f}x | {profile} |")
        lines.append(line)
        print(line, flush=True)

    def profiled(fn):
        """Wall time and per-kernel times of one run of fn after an unprofiled warmup."""
        fn()
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        _lib.profile_collect()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        kt = _lib.profile_collect()
        _lib.profile_enable(False)
        return t, out, kt

    cfg = dict(synthetic.MFOD_CONFIG)
    ph = synthetic.phantom(**cfg)
    signals = synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)
    n_vox, m = signals.shape
    dev = torch.from_numpy(signals).to("cuda")
    sample = np.ascontiguousarray(signals[rng.choice(n_vox, 2048, replace=False)])
    rfs = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3.0e-3)]
    vol = f"{n_vox} voxels x {m} measurements (the bench volume)"

    # a2/a4: single-tissue ridge NNLS on the 300 HR directions.
    k1 = Kernel(rfs[:1], ph.bvals, ph.bvecs, dirs=DIRS_300)
    solver = _core.FodSolver.l2nnls(np.ascontiguousarray(k1.w_k_aniso), 0.1, 1e-6)
    t, _, kt = profiled(lambda: _core.fit_single_tissue(dev, solver, 1))
    cpu = _cpu_rate(lambda: oracle.fit_single_tissue(sample, k1.w_k_aniso, 0.1, 1e-6,
                                                     n_threads=threads), len(sample), threads)
    report("single-tissue L2NNLS (a2, a4)", vol + ", 300 directions", n_vox, t, cpu, kt)

    # GRL: damped Richardson-Lucy multi-tissue (SDModel's default method).
    w = compute_shell_weights(ph.bvals, 0.2)
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=low_resolution_dirs(rfs, len(ph.bvals)), weights=w)
    lr_s = DRL()._prepare(lr, ph.bvals, Stage.LOW_RES)
    hr_s = DRL()._prepare(hr, ph.bvals, Stage.HIGH_RES)
    wt = torch.from_numpy(np.ascontiguousarray(w)).to("cuda")
    lk, hk = _deconv_kernel(lr), _deconv_kernel(hr)
    t, _, kt = profiled(lambda: _core.fit_multi_tissue(dev, wt, lk, lr_s, hk, hr_s, MAX_ROUNDS,
                                                       1e-3, FINAL_FRACTION_TOL_X, 1))

    def stage(kernel, n_iter):
        k = kernel.w_k_aniso
        iso = np.full((1, m), np.exp(-ph.bvals.max() * 0.7e-3))
        kk = kernel.k
        nu = 0.2 * oracle.fit_single_tissue_drl(iso, kk.T.copy(), kk.T @ kk, 0.0, 200,
                                                1.0 / kk.shape[1]).max()
        return (k.T.copy(), k.T @ k, nu, n_iter - 1, 1e-6)

    small = sample[:96]
    cpu = _cpu_rate(lambda: oracle.fit_multi_tissue_drl(small, w, lr, hr, stage(lr, 200),
                                                        stage(hr, 600), n_threads=threads),
                    len(small), threads)
    report("GRL (damped Richardson-Lucy)", vol + ", 44 LR / 300 HR directions", n_vox, t, cpu, kt)

    # Elastic net (Lasso, FodSolver.lasso): multi-tissue, equal shell weights, default settings.
    wl = compute_shell_weights(ph.bvals, 1.0)
    hrl = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=wl)
    lrl = Kernel(rfs, ph.bvals, ph.bvecs, dirs=low_resolution_dirs(rfs, len(ph.bvals)),
                 weights=wl)
    net = Lasso()
    lr_n = net._prepare(lrl, ph.bvals, Stage.LOW_RES)
    hr_n = net._prepare(hrl, ph.bvals, Stage.HIGH_RES)
    wlt = torch.from_numpy(np.ascontiguousarray(wl)).to("cuda")
    lkn, hkn = _deconv_kernel(lrl), _deconv_kernel(hrl)
    t, _, kt = profiled(lambda: _core.fit_multi_tissue(dev, wlt, lkn, lr_n, hkn, hr_n, MAX_ROUNDS,
                                                       1e-7, FINAL_FRACTION_TOL_X, 1))
    nsample = sample[:128]
    cpu = _cpu_rate(lambda: oracle.fit_multi_tissue_lasso(nsample, wl, lrl, hrl, net.reg, net.alpha,
                                                          net.tol, net.max_iter, n_threads=threads),
                    len(nsample), threads)
    report("Lasso (elastic net), multi-tissue", vol + ", 44 LR / 300 HR directions, reg 0.01, "
           "alpha 0.9", n_vox, t, cpu, kt)

    # a8: SH projection of the 300 amplitudes to lmax 8.
    amps = torch.rand((n_vox, 300), dtype=torch.float64, device="cuda")
    proj = torch.as_tensor(projection_matrix(DIRS_300, 8).T * (300 / (4 * np.pi)),
                           dtype=torch.float64, device="cuda")
    t, _, kt = profiled(lambda: _core.sh_project(amps, proj, exact_integral=True))
    host_amps = amps[:50000].cpu().numpy()
    host_proj = proj.cpu().numpy()
    cpu = _cpu_rate(lambda: host_amps @ host_proj, 50000, threads)
    report("SH projection (a8)", f"{n_vox} x 300 amplitudes to 45 coefficients (numpy as the "
           "CPU side)", n_vox, t, cpu, kt)

    # a9: CSD, lmax 8, the single-shell CSD phantom at 4 x its volume.
    ccfg = dict(synthetic.CSD_CONFIG)
    ccfg["shape"] = (64, 64, 40)
    cph = synthetic.phantom(**ccfg)
    on_shell = cph.bvals > 50
    cs = np.ascontiguousarray(
        np.maximum(cph.data.reshape(-1, len(cph.bvals))[:, on_shell], 0.0), dtype=np.float64)
    p = csd_problem(cph.bvecs[on_shell], synthetic.tensor_zonal(1000.0), 8, CSDSolver())
    cdev = torch.from_numpy(cs).to("cuda")
    t, _, kt = profiled(lambda: _core.fit_csd(cdev, p.initial, p.forward, p.constraint,
                                              p.norm_weight, p.threshold, p.max_iter, 1))
    csample = np.ascontiguousarray(cs[rng.choice(len(cs), 4096, replace=False)])
    cpu = _cpu_rate(lambda: oracle.fit_csd(csample, p.initial, p.forward, p.constraint,
                                           p.norm_weight, p.threshold, p.max_iter,
                                           n_threads=threads), len(csample), threads)
    report("CSD (a9)", f"{len(cs)} voxels x {cs.shape[1]} measurements, lmax 8, 300 constraints",
           len(cs), t, cpu, kt)

    # a10: MSMT-CSD (LSI): WM lmax 8 + GM + CSF on the bench volume's shells.
    shells = np.array([0.0, 1000.0, 2000.0, 3000.0])

    def zonal_tensor(par, perp):
        return ZonalSHRF(bvals=shells, coefficients=np.array(
            [synthetic.tensor_zonal(b, par, perp) for b in shells]))

    def zonal_iso(d):
        return ZonalSHRF(bvals=shells, coefficients=np.sqrt(4 * np.pi) * np.exp(-shells * d)[:, None])

    mrfs = [zonal_tensor(1.7e-3, 0.3e-3), zonal_iso(0.8e-3), zonal_iso(3.0e-3)]
    forward = forward_matrix(mrfs, [8, 0, 0], ph.bvals, ph.bvecs)
    g = constraint_matrix([8, 0, 0], DIRS_300)
    design = _core.LsiDesign(np.ascontiguousarray(forward), np.ascontiguousarray(g),
                             msmt_ridge(forward))
    t, _, kt = profiled(lambda: _core.fit_lsi(dev, design, 1))
    lsample = sample[:1024]
    cpu = _cpu_rate(lambda: oracle.fit_lsi(lsample, forward, g, msmt_ridge(forward),
                                           n_threads=threads), len(lsample), threads)
    report("MSMT-CSD / LSI (a10)", vol + ", 47 coefficients, 302 constraints", n_vox, t, cpu, kt)

    # a11: peaks of lmax-8 FODs (3 per voxel) on the 642-vertex icosphere.
    coeffs = torch.from_numpy(np.ascontiguousarray(
        (rng.random((n_vox, 300)) ** 8) @ (projection_matrix(DIRS_300, 8).T * (300 / (4 * np.pi)))
    )).to("cuda")
    sphere = _device_sphere(8, False)
    t, _, kt = profiled(lambda: _core.find_peaks_device(coeffs, 8, sphere, 3, 0.1, 0.0))
    psample = coeffs[:4096].cpu().numpy()
    vertices, neighbours = _build_icosphere()
    cpu = _cpu_rate(lambda: oracle.peaks_chunk(psample, 8, _get_amplitudes(psample, 8, False),
                                               vertices, neighbours, 3, 0.1, 0.0,
                                               n_threads=threads), len(psample), threads)
    report("peaks (a11)", f"{n_vox} lmax-8 FODs, 3 peaks, 642-vertex sphere", n_vox, t, cpu, kt)

    # a9 past lmax 8: CSD at lmax 12 (91 coefficients, the system's triangle in LDS) on a b = 3000
    # shell of 90 directions.
    hph = synthetic.phantom(shape=(64, 64, 40), n_b0=5, shells=[(3000.0, 90)], seed=1, snr=20.0,
                            mask_share=None, tissues=False)
    h_on = hph.bvals > 50
    hs = np.ascontiguousarray(
        np.maximum(hph.data.reshape(-1, len(hph.bvals))[:, h_on], 0.0), dtype=np.float64)
    hp = csd_problem(hph.bvecs[h_on], synthetic.tensor_zonal(3000.0, lmax=12), 12, CSDSolver())
    hdev = torch.from_numpy(hs).to("cuda")
    t, _, kt = profiled(lambda: _core.fit_csd(hdev, hp.initial, hp.forward, hp.constraint,
                                              hp.norm_weight, hp.threshold, hp.max_iter, 1))
    hsample = np.ascontiguousarray(hs[rng.choice(len(hs), 1024, replace=False)])
    cpu = _cpu_rate(lambda: oracle.fit_csd(hsample, hp.initial, hp.forward, hp.constraint,
                                           hp.norm_weight, hp.threshold, hp.max_iter,
                                           n_threads=threads), len(hsample), threads)
    report("CSD lmax 12 (a9, wide kernel)", f"{len(hs)} voxels x {hs.shape[1]} measurements, "
           "lmax 12 (91 coefficients), 300 constraints", len(hs), t, cpu, kt)

    # a10 past 63 parameters: MSMT-CSD with WM lmax 10 (66 + 2 parameters, two lane slots).
    wrfs = [ZonalSHRF(bvals=shells, coefficients=np.array(
                [synthetic.tensor_zonal(b, 1.7e-3, 0.3e-3, lmax=10) for b in shells])),
            zonal_iso(0.8e-3), zonal_iso(3.0e-3)]
    wforward = forward_matrix(wrfs, [10, 0, 0], ph.bvals, ph.bvecs)
    wg = constraint_matrix([10, 0, 0], DIRS_300)
    wdesign = _core.LsiDesign(np.ascontiguousarray(wforward), np.ascontiguousarray(wg),
                              msmt_ridge(wforward))
    t, _, kt = profiled(lambda: _core.fit_lsi(dev, wdesign, 1))
    wsample = sample[:512]
    cpu = _cpu_rate(lambda: oracle.fit_lsi(wsample, wforward, wg, msmt_ridge(wforward),
                                           n_threads=threads), len(wsample), threads)
    report("MSMT-CSD WM lmax 10 (a10, 68 parameters)", vol + ", 68 parameters, 302 constraints",
           n_vox, t, cpu, kt)

    # a5 past 512 measurements: the multi-tissue fit at 10 b0 + 3 x 200 (12 row slots).
    mph = synthetic.phantom(shape=(64, 64, 40), n_b0=10,
                            shells=[(1000.0, 200), (2000.0, 200), (3000.0, 200)], seed=4,
                            snr=25.0, mask_share=0.40, tissues=True)
    msig = synthetic.normalized_signals(mph.data, mph.bvals, mph.mask)
    mw = np.ones(len(mph.bvals))
    mhr = Kernel(rfs, mph.bvals, mph.bvecs, dirs=DIRS_300, weights=mw)
    mlr = Kernel(rfs, mph.bvals, mph.bvecs, dirs=low_resolution_dirs(rfs, len(mph.bvals)),
                 weights=mw)
    mlr_s = _core.FodSolver.l2nnls(np.ascontiguousarray(mlr.w_k_aniso), 0.0, 1e-6)
    mhr_s = _core.FodSolver.l2nnls(np.ascontiguousarray(mhr.w_k_aniso), 0.1, 1e-6)
    mdev = torch.from_numpy(msig).to("cuda")
    mwt = torch.from_numpy(mw).to("cuda")
    t, _, kt = profiled(lambda: _core.fit_multi_tissue(
        mdev, mwt, _deconv_kernel(mlr), mlr_s, _deconv_kernel(mhr), mhr_s, 50, 1e-7, 1e-2, 1))
    msample = np.ascontiguousarray(msig[rng.choice(len(msig), 1024, replace=False)])
    cpu = _cpu_rate(lambda: oracle.fit_multi_tissue(msample, mw, mlr, mhr, 0.0, 0.1, 1e-6,
                                                    n_threads=threads), len(msample), threads)
    report("mFOD at 610 measurements (a5, 12 row slots)",
           f"{len(msig)} voxels x {msig.shape[1]} measurements, 44 LR / 300 HR directions",
           len(msig), t, cpu, kt)

    # Response estimation (estimate_csd_rf): the CSD phantom above, every voxel searched. The
    # CPU side runs the same host code with the oracle's CSD fits and peak searches.
    from larmor_amd.dwi.inputs import signal_rows
    from larmor_amd.sd import csd_rf
    from larmor_amd.sd.sphere import get_theta_phi
    from larmor_amd.types import GradientTable, Volume

    rvol = Volume(cph.data, np.eye(4))
    rgtab = GradientTable(cph.bvals, cph.bvecs)
    rmask = np.ones(cph.data.shape[:3], dtype=bool)
    _lib.profile_enable(True)
    _lib.profile_collect()
    t0 = time.perf_counter()
    est = csd_rf.estimate_csd_rf(rvol, rgtab, rmask, n_jobs=threads)
    t = time.perf_counter() - t0
    kt = _lib.profile_collect()
    _lib.profile_enable(False)

    def oracle_fods(volume, candidates, on_shell, bvecs, response, lmax, _threads):
        problem = csd_problem(bvecs, response, lmax, CSDSolver())
        sig = np.maximum(signal_rows(volume, np.nonzero(candidates), slice(None), on_shell), 0.0)
        return oracle.fit_csd(np.ascontiguousarray(sig), problem.initial, problem.forward,
                              problem.constraint, problem.norm_weight, problem.threshold,
                              problem.max_iter, n_threads=threads)[0]

    def oracle_peaks(fods, lmax):
        verts, nbrs = csd_rf._sphere()
        th, ph_ = get_theta_phi(verts)
        amps = np.ascontiguousarray(fods @ csd_rf.build_basis_matrix(th, ph_, lmax).T)
        _, found = oracle.peaks_chunk(np.ascontiguousarray(fods), lmax, amps, verts, nbrs, 2, 0.0,
                                      0.0, n_threads=threads)
        found = np.nan_to_num(found, nan=0.0)
        found[fods[:, 0] <= 0] = 0.0
        return found

    product = (csd_rf._candidate_fods, csd_rf._two_largest_peaks)
    csd_rf._candidate_fods, csd_rf._two_largest_peaks = oracle_fods, oracle_peaks
    try:
        t0 = time.perf_counter()
        ref = csd_rf.estimate_csd_rf(rvol, rgtab, rmask, n_jobs=threads)
        cpu = (rmask.sum() / (time.perf_counter() - t0), threads)
    finally:
        csd_rf._candidate_fods, csd_rf._two_largest_peaks = product
    same = np.array_equal(est.voxels, ref.voxels) and est.num_iter == ref.num_iter
    report("CSD response estimate (estimate_csd_rf)",
           f"{int(rmask.sum())} voxels x {int((cph.bvals > 50).sum())} measurements, 300 voxels, "
           f"lmax 10, {est.num_iter} iterations (same selection as the oracle pipeline: {same})",
           int(rmask.sum()), t, cpu, kt)

    # Three-tissue response estimate (estimate_tissue_responses) on the three-tissue phantom at
    # 64 x 64 x 40; the CPU side runs the same host code with the oracle's LSI fits and peaks.
    import types

    from larmor_amd.sd import tissue_responses as tresp
    from larmor_amd.sd.msmt import constraint_matrix as cmat, forward_matrix as fmat
    from larmor_amd.sd.peaks import Peaks

    tdata, tb, tv, tmask, _, _ = synthetic.three_tissue_phantom(shape=(64, 64, 40))
    tvol = Volume(tdata, np.diag([2.5, 2.5, 2.5, 1.0]))
    tgtab = GradientTable(tb, tv)
    _lib.profile_enable(True)
    _lib.profile_collect()
    t0 = time.perf_counter()
    test = tresp.estimate_tissue_responses(tvol, tgtab, tmask, n_jobs=threads)
    t = time.perf_counter() - t0
    kt = _lib.profile_collect()
    _lib.profile_enable(False)

    def o_prepare(rfs_, lmaxes, bvals, bvecs, dirs):
        fw = fmat(rfs_, lmaxes, bvals, bvecs)
        return fw, cmat(lmaxes, dirs), msmt_ridge(fw)

    def o_fit(sig, design_, _threads):
        return oracle.fit_lsi(np.ascontiguousarray(sig, dtype=np.float64), *design_,
                              n_threads=threads)

    def o_peaks(cf, num_peaks, n_jobs):
        lm = int(round((np.sqrt(8 * cf.shape[1] + 1) - 3) / 2))
        verts, nbrs = _build_icosphere()
        d_, a_ = oracle.peaks_chunk(np.ascontiguousarray(cf), lm,
                                    np.ascontiguousarray(_get_amplitudes(cf, lm, False)), verts,
                                    nbrs, num_peaks, 0.1, 0.0, n_threads=threads)
        return Peaks(directions=d_, amplitudes=a_)

    saved = (tresp.prepare_msmt, tresp.fit_msmt, tresp.find_peaks, tresp._core)
    tresp.prepare_msmt, tresp.fit_msmt, tresp.find_peaks = o_prepare, o_fit, o_peaks
    tresp._core = types.SimpleNamespace(
        LsiDesign=lambda a_, g_, r_: (a_, g_, r_),
        fit_lsi=lambda sig, des, n_threads: oracle.fit_lsi(sig, *des))
    try:
        t0 = time.perf_counter()
        tref = tresp.estimate_tissue_responses(tvol, tgtab, tmask, n_jobs=threads)
        cpu = (tmask.sum() / (time.perf_counter() - t0), threads)
    finally:
        tresp.prepare_msmt, tresp.fit_msmt, tresp.find_peaks, tresp._core = saved
    same = all(np.array_equal(getattr(test, k), getattr(tref, k))
               for k in ("wm_voxels", "gm_voxels", "csf_voxels"))
    report("three-tissue response estimate (estimate_tissue_responses)",
           f"{int(tmask.sum())} masked voxels of 64x64x40 x {len(tb)} measurements (4 b0 + 3 x 30), "
           f"{int(test.wm_voxels.sum())} WM / {int(test.gm_voxels.sum())} GM / "
           f"{int(test.csf_voxels.sum())} CSF voxels (same as the oracle pipeline: {same})",
           int(tmask.sum()), t, cpu, kt)
    return 0


def track_report(n_seeds: int, cpu_seconds: float) -> int:
    """Streamline tracking (--track): det and prob over a brain-sized synthetic FOD field
    (96 x 96 x 60 voxels of 2 mm, lmax 8, float32; coherent arcs, a vertical and an
    anterior-posterior bundle, crossing), n_seeds seeds drawn in a 40% ellipsoid mask without
    directions (each searches its own), the tracker's defaults at this voxel size (step 0.2 mm,
    max length 200 mm = 1000 steps, det's angle 8.8 degrees, prob's 15) and cutoff 0.1.  The FOD,
    the mask, the seeds and the points buffer are resident in HBM; one call tracks every seed.
    Prints one JSON line per algorithm: streamlines/s and steps/s on the GPU, the kernel's mean
    launch time from HIP events, and the oracle on a bounded sample of the same seeds on the
    host's cores."""
    import math

    import torch

    import oracle
    from larmor_amd import _core, _lib, synthetic
    from larmor_amd.tractography.seed import MaskSeeds
    from larmor_amd.types import Volume

    fod, affine, mask = synthetic.tracking_field()
    inv = np.linalg.inv(affine)
    mask_f = mask.astype(np.float32)
    seeds = MaskSeeds(Volume(mask_f, affine)).draw(n_seeds, np.random.default_rng(1)).positions
    idx = np.arange(n_seeds, dtype=np.int64)
    d_fod = torch.from_numpy(fod).to("cuda")
    d_mask = (torch.from_numpy(mask_f).to("cuda"), inv)
    d_seeds = torch.from_numpy(seeds).to("cuda")
    d_idx = torch.from_numpy(idx).to("cuda")
    step, max_steps, cutoff = 0.2, 1000, 0.1
    det_angle = math.degrees(2 * math.asin(step / (2 * 0.65 * 2.0)))
    d_pos = torch.empty((n_seeds, max_steps + 1, 3), dtype=torch.float32, device="cuda")
    threads = host_cores()
    rng = np.random.default_rng(0)
    for name, algo, ref in [
        ("det", _core.DetParams(cutoff, det_angle),
         oracle.TrackParams(algorithm=0, cutoff=cutoff, max_angle_deg=det_angle)),
        ("prob", _core.ProbParams(cutoff, 15.0),
         oracle.TrackParams(algorithm=1, cutoff=cutoff, max_angle_deg=15.0, trials=1000,
                            attempts=1000)),
    ]:
        def run(n=n_seeds):
            return _core.track_batch(d_fod, inv, d_mask, d_seeds[:n], None, d_idx[:n], 1, step,
                                     max_steps, True, algo, d_pos, 1)

        run(min(n_seeds, 4096))  # warm-up (module load, calibration)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        counts, _, _, stats = run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        _lib.profile_enable(True)
        _lib.profile_collect()
        run()
        torch.cuda.synchronize()
        kt = _lib.profile_collect()
        _lib.profile_enable(False)
        c = counts.cpu().numpy()
        started = int((c > 0).sum())
        steps = int((c[c > 0] - 1).sum())
        # The oracle on a sample sized for about cpu_seconds of host work.
        probe = rng.choice(n_seeds, size=min(200, n_seeds), replace=False)
        t1 = time.perf_counter()
        oracle.track_batch(fod, inv, (mask_f, inv), seeds[probe], None, idx[probe], 1, step,
                           max_steps, True, ref, n_threads=threads)
        rate = len(probe) / (time.perf_counter() - t1)
        n_cpu = int(min(n_seeds, max(500, rate * cpu_seconds)))
        sample = rng.choice(n_seeds, size=n_cpu, replace=False)
        t1 = time.perf_counter()
        r = oracle.track_batch(fod, inv, (mask_f, inv), seeds[sample], None, idx[sample], 1,
                               step, max_steps, True, ref, n_threads=threads)
        cpu_dt = time.perf_counter() - t1
        same = float(np.mean(r[1] == c[sample]))
        started_cpu = r[1] > 0
        ms, launches = kt.get("track", (0.0, 0, 0.0))[:2]
        line = {
            "metric": f"streamlines/s ({name} tracking)",
            "value": n_seeds / dt,
            "unit": "seeds/s",
            "started_per_s": started / dt,
            "steps_per_s": steps / dt,
            "ms": dt * 1e3,
            "kernel_ms": ms / max(launches, 1),
            "dtype": "f64 (float32 FOD and points)",
            "config": {"workload": "synthetic FOD field 96x96x60, 2 mm, lmax 8, 40% ellipsoid "
                                   "mask; seeds without directions",
                       "seeds": n_seeds, "started": started, "steps": steps,
                       "mean_points": float(c[c > 0].mean()) if started else 0.0,
                       "step_mm": step, "max_steps": max_steps, "cutoff": cutoff,
                       "max_angle": det_angle if name == "det" else 15.0,
                       "sampling": stats},
            "cpu_baseline": {"value": n_cpu / cpu_dt, "unit": "seeds/s", "cores": threads,
                             "kind": "port",
                             "sample": f"{n_cpu} random seeds of {n_seeds} ({cpu_dt:.1f} s)",
                             "cpu_model": cpu_model(),
                             "same_counts_as_gpu": same,
                             "mean_points": float(r[1][started_cpu].mean()),
                             "gpu_mean_points_on_sample": float(c[sample][started_cpu].mean())},
        }
        print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=5)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--cpu-seconds", type=float, default=15.0)
    parser.add_argument("--rows", action="store_true",
                        help="instead of the JSON line, time every row of the path (single-tissue "
                             "NNLS, GRL, Lasso, SH projection, CSD, MSMT-CSD, peaks, the CSD and "
                             "three-tissue response estimates) beside its CPU oracle")
    parser.add_argument("--track", action="store_true",
                        help="instead of the JSON line, time det and prob streamline tracking "
                             "over a brain-sized synthetic FOD field beside the CPU oracle")
    parser.add_argument("--track-seeds", type=int, default=200_000)
    parser.add_argument("--serve", action="store_true",
                        help="instead of the JSON line above, measure the batched and served forms "
                             "(many requests with their own acquisitions; p50/p99 latency)")
    parser.add_argument("--serve-jobs", type=int, default=64)
    parser.add_argument("--serve-rates", default="5,20,50", help="Poisson arrival rates, jobs/s")
    parser.add_argument("--slots", type=int, default=4, help="requests in flight at once")
    parser.add_argument("--no-serve", action="store_true",
                        help="leave the latency-under-load phase out of the JSON line")
    parser.add_argument("--form", choices=["fp64", "fp32", "mixed"], default="fp64",
                        help="fp64: the reference's FP64 computation (bit-identical to the oracle; "
                             "the default line); fp32: the plain FP32 form; mixed: the "
                             "mixed-precision form (BF16 matrix-core correlations, FP64 "
                             "refinement rounds and solves)")
    parser.add_argument("--workload", choices=sorted(WORKLOADS), default="mfod3",
                        help="mfod3: the 96x96x60 3-shell volume (lmax 8); mfod4: the 140x140x96 "
                             "4-shell volume (lmax 10)")
    parser.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                        help="only run three device-to-device copies of MIB MiB (the known-size "
                             "kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = parser.parse_args()
    if args.serve:
        if args.form == "fp32":
            print("the request server runs the fp64 and mixed forms", file=sys.stderr)
            return 2
        return serve_report(args.serve_jobs, [float(r) for r in args.serve_rates.split(",")],
                            args.slots, args.form)
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.rows:
        return rows_report()
    if args.track:
        return track_report(args.track_seeds, args.cpu_seconds)
    if args.gpus != 1:
        print("bench.py measures one GPU; the path shards voxels across GPUs with no exchange, "
              "and runs on more than one GPU are not part of this build.", file=sys.stderr)
        return 2

    import torch

    from larmor_amd import _core, _lib
    from larmor_amd.sd.sh import projection_matrix
    from larmor_amd.sd.solvers import FINAL_FRACTION_TOL_X, MAX_ROUNDS, _deconv_kernel
    from larmor_amd.sd.sphere import DIRS_300

    t_build = time.perf_counter()
    cfg, signals, w, lr, hr = build_workload(workload=args.workload)
    n_vox, m = signals.shape
    print(f"workload: {n_vox} voxels x {m} measurements, built in "
          f"{time.perf_counter() - t_build:.1f} s; form {args.form}", file=sys.stderr, flush=True)

    dev = torch.from_numpy(signals).to("cuda")
    weights = torch.from_numpy(np.ascontiguousarray(w)).to("cuda")
    lr_k, hr_k = _deconv_kernel(lr), _deconv_kernel(hr)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    lmax = WORKLOADS[args.workload][1]
    projection = torch.as_tensor(
        projection_matrix(DIRS_300, lmax).T * (len(DIRS_300) / (4.0 * np.pi)),
        dtype=torch.float64, device="cuda")
    if args.form == "fp32":
        dev = dev.float()
        weights = weights.float()
        projection = projection.float()

    def step():
        if args.form == "fp32":
            # The plain FP32 form; its SH projection is a single-precision library GEMM.
            fods, fractions, rounds, status = _core.fit_multi_tissue_fp32(
                dev, weights, lr_k, lr_solver, hr_k, hr_solver, MAX_ROUNDS, 1e-7,
                FINAL_FRACTION_TOL_X, 1, return_diagnostics=True)
            return torch.matmul(fods, projection), fractions, rounds, status
        fods, fractions, rounds, status = _core.fit_multi_tissue(
            dev, weights, lr_k, lr_solver, hr_k, hr_solver, MAX_ROUNDS, 1e-7,
            FINAL_FRACTION_TOL_X, 1, return_diagnostics=True, form=args.form)
        sh = _core.sh_project(fods, projection, exact_integral=True)
        return sh, fractions, rounds, status

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # The timed steps run with the library's per-launch timing off.
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"timed: {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    # Per-kernel times and counted work from separate profiled steps (HIP events around each
    # launch on its own stream).
    profiled_steps = max(1, min(args.steps, 3))
    _lib.profile_enable(True)
    _lib.profile_collect()
    t_prof = time.perf_counter()
    for _ in range(profiled_steps):
        step()
    torch.cuda.synchronize()
    profiled_elapsed = time.perf_counter() - t_prof
    kernel_times = _lib.profile_collect()
    _lib.profile_enable(False)
    sh, fractions, rounds, status = out
    total_rounds = int(rounds.sum().item())

    value = n_vox * args.steps / elapsed
    n_lr = lr.w_k_aniso.shape[1]
    n_hr = hr.w_k_aniso.shape[1]
    # Algorithmic FP64 flops of each kernel per step.
    def counted(name):
        return kernel_times.get(name, (0.0, 0, 0.0))[2] / profiled_steps

    # The FOD kernels count their own work; the others are priced per voxel (and round).  In the
    # mixed form a round's correlations come from one of two passes, so those two are not priced.
    step_flops = {
        "fraction_loop": total_rounds * 2.0 * m * n_lr,
        "fod_hr_corr": n_vox * 2.0 * m * n_hr,
        "fraction_final": n_vox * 2.0 * m * n_hr,
    }
    if args.form != "mixed":
        step_flops["fod_lr_corr"] = total_rounds * 2.0 * m * n_lr
    for name, (_, _, work) in kernel_times.items():
        if work > 0:
            step_flops[name] = counted(name)
    dominant = max((k for k in step_flops if k in kernel_times),
                   key=lambda k: kernel_times[k][0])
    dom_ms, dom_launches, _ = kernel_times[dominant]
    launches_per_step = dom_launches / profiled_steps
    avg_launch_s = dom_ms / dom_launches / 1e3
    flops_per_launch = step_flops[dominant] / launches_per_step
    achieved = flops_per_launch / avg_launch_s / 1e12
    peak = FP32_PEAK_TFLOPS if args.form == "fp32" else FP64_PEAK_TFLOPS
    print("kernel ms per step: " + ", ".join(
        f"{k} {v[0] / profiled_steps:.2f} ({v[1] // profiled_steps} launches)"
        for k, v in kernel_times.items()), file=sys.stderr, flush=True)
    print("counted FP64 work per step: " + ", ".join(
        f"{k} {v[2] / profiled_steps / 1e9:.2f} GFLOP" for k, v in kernel_times.items()
        if v[2] > 0), file=sys.stderr, flush=True)
    # The whole step against the roofs, with the dense work and bytes per voxel of the path:
    # F = 2 m n_hr (K^T s, high resolution) + 2 n_hr C (SH projection) + R 4 m n_lr (per round
    # K^T a and the group column K f), R the mean rounds run; B = 4 m + 4 C + 4 T (an FP32 signal
    # in, FP32 SH and fractions out).
    n_coeffs = projection.shape[1]
    n_tissues = fractions.shape[1]
    if args.form != "fp64":
        # Error against the FP64 form on the same voxels (one more fit, outside the timing).
        ref_fods, ref_fractions = _core.fit_multi_tissue(
            torch.from_numpy(signals).to("cuda"), torch.from_numpy(np.ascontiguousarray(w)).to(
                "cuda"), lr_k, lr_solver, hr_k, hr_solver, MAX_ROUNDS, 1e-7,
            FINAL_FRACTION_TOL_X, 1)
        ref_sh = _core.sh_project(ref_fods, projection.double(), exact_integral=True)
        got_sh = sh.double()
        sh_err = float(torch.linalg.norm(got_sh - ref_sh) / torch.linalg.norm(ref_sh))
        frac_err = (fractions.double() - ref_fractions).abs()
        error_vs_fp64 = {
            "sh_rel_l2": sh_err,
            "fractions_max_abs": [float(x) for x in frac_err.max(dim=0).values],
            "fractions_p99_abs": float(torch.quantile(frac_err.max(dim=1).values.float(), 0.99)),
        }
        print(f"error against the FP64 form: {error_vs_fp64}", file=sys.stderr, flush=True)
    mean_rounds = total_rounds / n_vox
    dense_per_voxel = 2.0 * m * n_hr + 2.0 * n_hr * n_coeffs + mean_rounds * 4.0 * m * n_lr
    bytes_per_voxel = 4.0 * (m + n_coeffs + n_tissues)
    print(f"rounds: mean {total_rounds / n_vox:.2f}; status bits: "
          f"{torch.bincount(status.to(torch.int64)).tolist()}", file=sys.stderr, flush=True)

    result = {
        "metric": "voxel FOD solves/s (mFOD lmax=8, 3-shell); p50/p99 latency under load",
        "value": value,
        "unit": "voxel solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "mixed": "bf16+f64"}[args.form],
        "data": "synthetic (seeded phantom, Rician SNR 25)",
        "config": {
            "workload": WORKLOADS[args.workload][2].format(n=n_vox),
            "form": args.form,
            "voxels": n_vox,
            "measurements": m,
            "lmax": lmax,
            "mean_rounds": total_rounds / n_vox,
        },
        "roofline": {
            # The dominant kernel issues no MFMA: its active-set solves are chains of dependent
            # FP64 vector and scalar instructions, so FP64 VALU issue bounds it.
            "bound": "fp64-valu",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": measured_traffic(dominant, args.workload, args.form),
            "kernel": dominant,
            "kernel_avg_launch_ms": avg_launch_s * 1e3,
            "launches_per_step": launches_per_step,
            "flops_per_launch": flops_per_launch,
            "peak_kind": ("FP32 vector (157.3 TFLOP/s on MI355X)" if args.form == "fp32"
                          else "FP64 vector (78.6 TFLOP/s on MI355X)"),
            "profiled_steps": profiled_steps,
            "profiled_ms_per_step": profiled_elapsed / profiled_steps * 1e3,
        },
        "step_roofline": {
            "dense_flops_per_voxel": dense_per_voxel,
            "dense_tflops": value * dense_per_voxel / 1e12,
            "dense_frac_fp64": value * dense_per_voxel / 1e12 / FP64_PEAK_TFLOPS,
            "hbm_bytes_per_voxel": bytes_per_voxel,
            "hbm_gbs": value * bytes_per_voxel / 1e9,
            "hbm_frac": value * bytes_per_voxel / 1e9 / HBM_PEAK_GBS,
        },
    }
    if args.form != "fp64":
        result["error_vs_fp64"] = error_vs_fp64
    if args.workload != "mfod3":
        result["metric"] = ("voxel FOD solves/s (mFOD lmax=10, 4-shell, 140x140x96); "
                            "p50/p99 latency under load")
    if not args.no_serve and args.form != "fp32":
        # p50/p99 latency under load: requests of the served form arriving at 5 jobs/s.
        serve = ServeBench(32, args.slots, "mixed" if args.form == "mixed" else "fp64")
        policy = serve.configure(batch_jobs=64, slice_voxels=1 << 16, merge_voxels=1 << 19)
        result["latency_under_load"] = {"workload": serve.workload(), **serve.poisson(5.0),
                                        "policy": policy}
        serve.close()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(signals, w, lr, hr, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
