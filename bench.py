#!/usr/bin/env python3
"""Benchmark: G1 MSM points/s at 2^20 on one MI355X (the standalone-MSM workload).

A step is one multilinear-KZG commit of a resident 2^20-entry scalar table
against layer 0 of a device-resident prepared SRS (N = 20, seeded trapdoor):
digits, bucket sort, accumulation, bucket reduction on the device and the
final affine normalisation on the host, i.e. a complete MSM returning the
affine point.  Inputs are resident in HBM before the timed region.

Prints one JSON line (the driver contract).  `roofline` describes the dominant
kernel (bucket accumulation) measured live with HIP events on the stream the
library launches on; `cpu_baseline` times the CPU oracle's Pippenger on a
bounded sample (an MSM of 2^18 points) on this host's cores.

Beside the headline the line carries the other workloads of the metric, each
run by the native prover benchmark (_build/prove_bench, one child process per
workload): `fold` (HBM rate of the fix_variables fold, HIP events), `proofs`
(plain one-at-a-time HyperPlonk prove, vanilla and jellyfish, nv 12, with the
oracle prover as its CPU baseline), `sumcheck` (n = 20 whole proofs, degree 2
and 5), `batch` (13 uneven commits + 64 sumchecks, serial vs batched, and the
standalone MSM sizes 2^16..2^22), `srs` (device SRS generation with the
canonical file write, then the file read back and decompressed on the device,
nv 20) and `serving` (the 2000-request, nv 10..20 Poisson prover-service
stream at 0.8 x the measured batched capacity, served with continuous batching
under the measured HBM budget and again serially: proofs/s and p50/p99
latency; a 300-request nv 10..16 trace with --serve-short).
"""
import argparse
import json
import os
import subprocess
import tempfile
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Reference-published G1 MSM 2^20 (prepared SRS, Ascend910B3): 1.113355917 s.
REFERENCE_MSM_2_20_SECONDS = 1.113355917
# MI355X published figures: 256 CUs at up to 2.4 GHz, each CU issuing 64 lanes of
# one full-rate VALU operation per cycle, so 256 x 64 x 2.4e9 u32 x u32 -> u64
# multiply-adds (v_mad_u64_u32) per second; HBM3E 8 TB/s.  The line also carries
# both rates measured on the device (copy kernel, multiply-add chain kernel).
VALU_MUL_PEAK_T = 256 * 64 * 2.4e9 / 1e12
HBM_PEAK_BYTES = 8e12
# Fq Montgomery product on 12 x 32-bit limbs: 2 * 12^2 + 12 limb products.
LIMB_PRODUCTS_PER_FQM = 300
# Fixed work convention for one (point, window): a mixed addition counted as
# 7M + 4S = 11 Fq products (the executed XYZZ formula takes 8M + 2S, so this
# convention credits the kernel with 10 % more work than it does).
FQM_PER_MIXED_ADD = 11
# HBM bytes of one 2^20 k_accumulate launch by (window bits, task cap), from
# the FETCH_SIZE and WRITE_SIZE passes calibrated on bench.py's own 1 GiB copy
# kernel (profiles/k_accumulate_msm_2_20_pmc.md); rocprofv3 counters cannot be
# read from inside this process.  The autotuner picks c = 17, cap 32 here.
ACCUMULATE_TRAFFIC_2_20 = {(17, 32): 2.560e9}
# VALU lane instructions k_accumulate issues per non-zero digit: SQ_INSTS_VALU
# x 64 / digits from the SQ pass of the same profile (1 276 594 302 x 64 /
# 15 728 519); with the measured multiply-add peak it gives the fraction of
# the VALU issue rate the kernel runs at.
ACCUMULATE_VALU_INSTS_PER_DIGIT = 5194
# task cap of the built-in plan at 2^20
DEFAULT_CAP_2_20 = 32


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # the GPU clock ramps up under sustained load: 20 commits (~70 ms) reach it
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--log-size", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log", type=int, default=18)
    ap.add_argument("--extra-sizes", default="16,18")
    ap.add_argument("--skip-workloads", action="store_true", help="MSM headline only")
    ap.add_argument("--no-tune", action="store_true", help="use the built-in window rule, no autotuning")
    ap.add_argument("--cap", type=int, default=0,
                    help="lane task cap of the headline size with the built-in window rule (no autotuning)")
    ap.add_argument("--serve-short", action="store_true",
                    help="serve a 300-request nv 10..16 trace instead of the 2000-request nv 10..20 stream")
    return ap.parse_args()


PROVE_BENCH = os.path.join(ROOT, "_build", "prove_bench")
PROTOCOL_TEST = os.path.join(ROOT, "_build", "protocol_test")


def run_json(cmd, timeout):
    """Run a native benchmark child; its last stdout line is one JSON object."""
    print(f"[bench] {' '.join(os.path.basename(c) if i == 0 else c for i, c in enumerate(cmd))}",
          file=sys.stderr, flush=True)
    try:
        # stderr (progress lines, errors) passes through; stdout carries the JSON
        res = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"error": f"timeout after {timeout}s"}
    if res.returncode != 0:
        return {"error": f"exit {res.returncode} (its stderr is in the log)"}
    return json.loads(res.stdout.strip().splitlines()[-1])


def workloads(args, be, rng, peaks):
    out = {}
    # fold (fix_variables by the lowest variable) of a 2^24 table: 2 reads + 1 write of 32 B per output
    from hyperplonk_amd.fields import fr_from_bytes

    lg = 24
    table = be.mle_from_evaluations(lg, rng.fr_bytes(1 << lg))
    ms = be.fold_profile(table, fr_from_bytes(rng.fr_bytes(1).tobytes())[0], reps=20)
    moved = 96 * (1 << (lg - 1))
    out["fold"] = {
        "table": f"2^{lg}",
        "ms": round(ms, 4),
        "bytes": moved,
        "hbm_GB_per_s": round(moved / (ms * 1e-3) / 1e9, 1),
        "hbm_peak_GB_per_s": HBM_PEAK_BYTES / 1e9,
        "frac": round(moved / (ms * 1e-3) / HBM_PEAK_BYTES, 4),
        "hbm_copy_GB_per_s_measured": round(peaks["hbm_bytes_per_s"] / 1e9, 1),
        "frac_of_measured_copy": round(moved / (ms * 1e-3) / peaks["hbm_bytes_per_s"], 4),
    }
    del table
    if not os.path.exists(PROVE_BENCH):
        out["error"] = "native prover benchmark not built (make -C hyperplonk_amd/host)"
        return out
    out["proofs"] = {
        "vanilla_nv12": run_json([PROVE_BENCH, "prove", "12", "vanilla", "40", "3", "1"], 300),
        "jellyfish_nv12": run_json([PROVE_BENCH, "prove", "12", "jellyfish", "20", "3", "2"], 300),
    }
    if os.path.exists(PROTOCOL_TEST):
        threads = host_threads()
        cpu = run_json([PROTOCOL_TEST, "cpu-prove", "12", "vanilla", "3", str(threads)], 600)
        cpu["kind"] = "port"
        cpu["sample"] = "3 proofs of the nv-12 vanilla circuit on the CPU oracle backend (MSMs threaded)"
        out["proofs"]["cpu_baseline"] = cpu
    out["sumcheck"] = {
        "degree2_n20": run_json([PROVE_BENCH, "sumcheck", "20", "2", "7", "7"], 300),
        "degree5_n20": run_json([PROVE_BENCH, "sumcheck", "20", "5", "7", "7"], 300),
    }
    out["batch"] = run_json([PROVE_BENCH, "batch", "3", "4"], 600)
    # SRS: device generation + canonical file write, then file read + device decompression (nv 20)
    with tempfile.TemporaryDirectory() as tmp:
        out["srs"] = run_json([PROVE_BENCH, "srs", "20", os.path.join(tmp, "srs_nv20.bin")], 300)
    if args.serve_short:
        out["serving"] = run_json([PROVE_BENCH, "serve", "300", "auto", "10", "16", "0.3", "16", "5"], 600)
        out["serving"]["trace"] = "reduced: 300 requests, nv 10..16"
    else:
        # the prover-service stream: 2000 Poisson requests (30 % jellyfish, nv 10..20) at
        # 0.8 x the measured batched capacity, SRS resident for nv <= 22, admission by
        # slots and the measured HBM budget, batched vs serial FIFO on the same trace
        out["serving"] = run_json([PROVE_BENCH, "serve", "2000", "cap", "10", "20", "0.3", "16", "5",
                                   "auto", "22", "single"], 900)
        out["serving"]["trace"] = "2000 requests, nv 10..20, 30 % jellyfish, SRS nv 22"
    return out


def host_threads() -> int:
    """Worker threads for the CPU baselines: the cores this process may run on,
    at most 16 (one GPU's share of the host on these machines)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def host_identity() -> dict:
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cores": avail}


def reference_conventional_fqm(log_n: int) -> float:
    """FqM(N) = W (11 N + 32 2^c), c = l - floor(l/3) + 2, W = ceil(255 / c)."""
    c = log_n - log_n // 3 + 2
    w = -(-255 // c)
    return w * (11 * (1 << log_n) + 32 * (1 << c))


def main():
    args = parse()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU only: this build runs on a single MI355X and has no "
            "multi-GPU path (an MSM would shard by splitting its points across GPUs and "
            "adding the per-GPU results).",
            file=sys.stderr,
        )
        sys.exit(2)
    from hyperplonk_amd.backend import Mi355xBackend
    from hyperplonk_amd.fields import fr_from_bytes
    from hyperplonk_amd.synth import SplitMix64Stream

    log_n = args.log_size
    n = 1 << log_n
    be = Mi355xBackend(0)
    rng = SplitMix64Stream(0xB1512381)
    tuning = None
    if args.cap:
        be.msm_set_tuning([{"log_n": log_n, "window": 0, "cap": args.cap}])
    elif not args.no_tune:
        # window bits and task cap for the headline size, timed on this device
        s = time.time()
        tuning = be.msm_autotune([log_n])[0]
        tuning["seconds"] = round(time.time() - s, 2)
        print(f"[bench] autotune 2^{log_n}: {tuning}", file=sys.stderr, flush=True)
    cap_used = tuning["cap"] if tuning else (args.cap or (DEFAULT_CAP_2_20 if log_n == 20 else None))
    t = fr_from_bytes(rng.fr_bytes(log_n).tobytes())
    t0 = time.time()
    pp = be.generate_prover_param(log_n, t)
    srs_seconds = time.time() - t0
    scalars = rng.fr_bytes(n)
    mle = be.mle_from_evaluations(log_n, scalars)
    print(f"[bench] SRS N={log_n} generated+prepared in {srs_seconds:.2f}s", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        be.commit_packet(pp, mle)
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        result = be.commit_packet(pp, mle)
        times.append(time.perf_counter() - s)
    total = time.perf_counter() - t0
    ms_per_step = 1e3 * total / args.steps
    value = n / (total / args.steps)
    print(f"[bench] MSM 2^{log_n}: {ms_per_step:.3f} ms/step ({value / 1e6:.2f} M pts/s)", file=sys.stderr, flush=True)

    # live per-stage HIP-event times of the same MSM (dominant kernel: accumulate)
    stages, plan = None, None
    for _ in range(3):
        st, plan = be.msm_profile(pp, mle)
        if stages is None or st["device_total"] < stages["device_total"]:
            stages = st
    acc_ms = stages["accumulate"]
    acc_products = plan["entries"] * FQM_PER_MIXED_ADD * LIMB_PRODUCTS_PER_FQM
    achieved = acc_products / (acc_ms * 1e-3) / 1e12
    # roofline denominators measured on this device
    peaks = be.device_peaks()
    print(f"[bench] measured peaks: {peaks}", file=sys.stderr, flush=True)
    mad_peak_t = peaks["mad_per_s"] / 1e12
    conventional = reference_conventional_fqm(log_n) * LIMB_PRODUCTS_PER_FQM
    whole_msm_rate = conventional / (total / args.steps) / 1e12

    extra = {}
    for lg in [int(x) for x in args.extra_sizes.split(",") if x]:
        if lg >= log_n:
            continue
        m = be.mle_from_evaluations(lg, rng.fr_bytes(1 << lg))
        be.commit_packet(pp, m)
        s = time.perf_counter()
        reps = 10
        for _ in range(reps):
            be.commit_packet(pp, m)
        dt = (time.perf_counter() - s) / reps
        extra[f"2^{lg}"] = {"ms": round(dt * 1e3, 4), "pts_per_s": round((1 << lg) / dt, 1)}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle

        lg = args.cpu_sample_log
        threads = host_threads()
        layer = be.download_layer(pp, log_n - lg)
        sample = rng.fr_bytes(1 << lg)
        s = time.perf_counter()
        want = oracle.msm_pippenger_packets(layer, sample.tobytes(), 1 << lg, threads)
        dt = time.perf_counter() - s
        got = be.commit_packet(pp, be.mle_from_evaluations(lg, sample))
        cpu = {
            "value": round((1 << lg) / dt, 1),
            "unit": "pts/s",
            "cores": threads,
            "kind": "port",
            "sample": f"one G1 MSM of 2^{lg} points (SRS layer {log_n - lg}), oracle Pippenger, "
            f"{threads} threads, {dt:.2f} s; GPU result byte-identical: {got == want}",
            "host": host_identity(),
        }

    extra_workloads = None if args.skip_workloads else workloads(args, be, rng, peaks)

    line = {
        "metric": "MSM pts/s at 2^20",
        "value": round(value, 1),
        "unit": "pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / ((1 << 20) / REFERENCE_MSM_2_20_SECONDS), 2) if log_n == 20 else None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"BLS12-381 G1 MSM of 2^{log_n} uniform random Fr scalars against layer 0 of a "
            f"device-generated prepared SRS (N={log_n}, seeded trapdoor), affine result",
            "msm_log_size": log_n,
            "window_bits": plan["c"],
            "task_cap": cap_used,
            "windows": plan["windows"],
            "buckets": plan["buckets"],
            "nonzero_digits": plan["entries"],
            "parallelism": "single GPU",
            "vs_baseline_ref": "reference Ascend910B3 prepared-SRS MSM 2^20, 1.113 s (README.md:624)",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_accumulate (bucket accumulation, XYZZ += affine)",
            "achieved": round(achieved, 3),
            "peak": round(VALU_MUL_PEAK_T, 2),
            "unit": "T u32-limb-products/s",
            "frac": round(achieved / VALU_MUL_PEAK_T, 4),
            "work_convention": f"non-zero digits x {FQM_PER_MIXED_ADD} Fq products (7M + 4S mixed addition) "
            f"x {LIMB_PRODUCTS_PER_FQM} u32 limb products",
            "peak_measured": round(mad_peak_t, 2),
            "frac_of_measured": round(achieved / mad_peak_t, 4),
            "peak_measured_kernel": "k_mad_u64: 16 independent v_mad_u64_u32 chains per lane on every CU",
            "hbm_copy_bytes_per_s_measured": round(peaks["hbm_bytes_per_s"], 1),
            "traffic": (ACCUMULATE_TRAFFIC_2_20.get((plan["c"], cap_used)) if log_n == 20 else None),
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE at (c, cap) = (17, 32), "
            "profiles/k_accumulate_msm_2_20_pmc.md",
            "valu_issue_frac": round(plan["entries"] * ACCUMULATE_VALU_INSTS_PER_DIGIT
                                     / (acc_ms * 1e-3) / peaks["mad_per_s"], 4),
            "valu_issue_note": f"{ACCUMULATE_VALU_INSTS_PER_DIGIT} VALU lane instructions per digit "
            "(SQ_INSTS_VALU pass) over the measured multiply-add issue rate",
            "kernel_ms": round(acc_ms, 4),
            "stage_ms": {k: round(v, 4) for k, v in stages.items()},
            "whole_msm_conventional_T_per_s": round(whole_msm_rate, 3),
        },
        "cpu_baseline": cpu,
        "extra_sizes": extra,
        "workloads": extra_workloads,
        "srs_prepare_s": round(srs_seconds, 3),
        "msm_tuning": tuning,
        "result_prefix": result[:8].hex(),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
