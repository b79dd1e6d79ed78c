"""Benchmark of the MI355X Debye hot path (one GPU).

Workload (default): 250,000-atom ellipsoid (semi-axes 120 x 90 x 70 A, C/N/O/S
62/17/20/1 %) with 60,000 waters 3 A outside its surface, FF-explicit
histograms with 0.1 A weighted bins, and I(q) on 1,000 q points up to
1.5 1/A. Synthetic, deterministic inputs.

One step = HistogramManagerMTFFExplicit::calculate_all on coordinates
resident in HBM (pair kernels, merge, tail strips, histogram read-back and
bookkeeping) + the Debye transform (sinc table, bf16-split MFMA GEMM,
intensity profiles) -> one I(q). The timed region brackets exactly --steps
steps with device synchronisation on both sides.

Prints one JSON line (metric from the run's headline: pair evaluations per
second), with per-step latency percentiles, the pair kernel's roofline
against the FP32 VALU peak and the CPU oracle (a port of the reference's
multithreaded managers) timed on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (AMD published)
FLOP_PER_PAIR = 11  # 3 sub + 3 mul + 2 add + sqrt + mul(inv_w) + add(sum d)

CONFIGS = {
    1: dict(name="2k-atom globule, FF-avg, 0.25 A bins, 100 q in 0.01-0.5", n=2000, axes=(20, 20, 20),
            r_excl=2.0, waters=0, bw=0.25, mode="ffavg", q=(0.01, 0.5, 100)),
    3: dict(name="250k-atom ellipsoid + 60k waters, FF-explicit, 0.1 A bins, 1000 q to 1.5", n=250000,
            axes=(120, 90, 70), r_excl=1.8, waters=60000, bw=0.1, mode="ffexplicit", q=(0.0015, 1.5, 1000)),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(cfg, seed):
    import ausaxs_amd as A

    atoms = A.synth_globule(seed, cfg["n"], cfg["axes"], cfg["r_excl"])
    waters = A.synth_shell(seed + 1, cfg["waters"], cfg["axes"], 3.0) if cfg["waters"] else None
    return atoms, waters


def traffic_from_profiles():
    f = ROOT / "profiles" / "pair_hist_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def cpu_baseline(cfg, atoms, waters, q, budget_s=15.0):
    """The oracle's multithreaded FF managers (800-row tasks, thread-private
    24-byte entries, merge) + FP64 Debye on this host, on the full workload if
    it fits the time budget, else on a prefix sample grown until one run takes
    at least 40 % of the budget."""
    import oracle as O

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(cores, 16))
    mode = O.FFAVG if cfg["mode"] == "ffavg" else O.FFEXPLICIT
    w = waters if waters is not None else np.zeros((0, 4), np.float32)

    def pairs_of(n, m):
        return n * (n - 1) / 2 + n * m + m * (m - 1) / 2

    frac = min(1.0, 12000 / max(1, len(atoms)))
    while True:
        ns, ms = int(len(atoms) * frac), int(len(w) * frac)
        t = time.perf_counter()
        h = O.hist_ff(mode, atoms[:ns], w[:ms], cfg["bw"], 8000, 1.0, threads)
        O.debye_ff(mode, h, q, cfg["bw"])
        dt = time.perf_counter() - t
        if frac >= 1.0 or dt >= 0.4 * budget_s:
            break
        # pair work grows with the square of the prefix
        frac = min(1.0, frac * float(np.sqrt(0.8 * budget_s / max(dt, 1e-3))))
    pairs = pairs_of(ns, ms)
    sample = "full workload" if frac >= 1.0 else f"first {ns} atoms + {ms} waters of the workload"
    return dict(value=pairs / dt / 1e9, unit="Gpair-evals/s", cores=threads, kind="port",
                sample=f"{sample} ({pairs:.3g} pairs, histogram + I(q), {dt:.2f} s)",
                seconds=dt, pairs=pairs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, choices=sorted(CONFIGS))
    ap.add_argument("--debye-terms", type=int, default=3)
    ap.add_argument("--tile-j", type=int, default=8000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (traffic-counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import ctypes

        import ausaxs_amd as A

        ms = ctypes.c_double()
        A._check(A.lib.ausaxs_copy_probe(ctypes.c_size_t(args.copy_probe), 3, ctypes.byref(ms)))
        print(json.dumps({"copy_bytes": args.copy_probe, "ms": ms.value,
                          "GB/s": 2 * args.copy_probe / ms.value / 1e6}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding is design-only).",
              file=sys.stderr)
        sys.exit(2)

    import ausaxs_amd as A

    cfg = CONFIGS[args.config]
    log("device:", A.device_info())
    atoms, waters = make_inputs(cfg, args.config)
    q = np.linspace(*cfg["q"])
    mode = A.MODE_FFAVG if cfg["mode"] == "ffavg" else A.MODE_FFEXPLICIT
    settings = A.default_settings(bin_width=cfg["bw"], debye_terms=args.debye_terms, tile_j=args.tile_j)
    t0 = time.perf_counter()
    st = A.Structure(mode, atoms, waters, settings, z_exv=1.0)
    log(f"resident structure built in {time.perf_counter() - t0:.2f} s")
    st.set_timing(True)

    def step():
        h = st.calculate()
        I = h.debye_transform(q)
        return h, I

    for _ in range(args.warmup):
        step()
    lat, pair_ms, total_ms = [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        h, I = step()
        lat.append(time.perf_counter() - ts)
        tm = st.last_timing()
        pair_ms.append(tm["pair_ms"])
        total_ms.append(tm["total_ms"])
    elapsed = time.perf_counter() - t_start
    tm = st.last_timing()
    pairs = tm["pairs"]
    value = pairs * args.steps / elapsed / 1e9
    kern_s = float(np.mean(pair_ms)) / 1e3
    achieved = pairs * FLOP_PER_PAIR / kern_s / 1e12
    log(f"steps {args.steps}: {elapsed:.3f} s, pair kernel {np.mean(pair_ms):.3f} ms, "
        f"pairs+merge+strips {np.mean(total_ms):.3f} ms, items {tm['n_items']}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, atoms, waters if waters is not None else None, q)
        log(f"cpu baseline: {cpu['value']:.3f} Gpair-evals/s on {cpu['cores']} threads ({cpu['sample']})")
        cpu = {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")}

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "Gpair-evals/s",
        "value": round(value, 4),
        "unit": "Gpair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "atoms": len(atoms), "waters": 0 if waters is None else len(waters),
                   "pairs_per_step": pairs, "q_points": len(q), "debye_terms": args.debye_terms,
                   "parallelism": "single-gpu"},
        "profiles_per_s": round(args.steps / elapsed, 4),
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 4),
                       "p99": round(float(np.percentile(lat_ms, 99)), 4)},
        "roofline": {"bound": "mfma", "engine": "fp32 VALU (no matrix op on this path; 157.3 TF is the FP32 peak)",
                     "kernel": "pair_hist_kernel", "achieved": round(achieved, 4),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
                     "traffic": traffic_from_profiles(), "kernel_ms": round(kern_s * 1e3, 4),
                     "pairs_per_s": round(pairs / kern_s, 1)},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
