"""Benchmark of the MI355X Debye hot path (one GPU).

Default workload (--config 3): 250,000-atom ellipsoid (semi-axes 120 x 90 x
70 A, C/N/O/S 62/17/20/1 %) with 60,000 waters 3 A outside its surface,
FF-explicit histograms with 0.1 A weighted bins, and I(q) on 1,000 q points
up to 1.5 1/A. Synthetic, deterministic inputs. One step =
HistogramManagerMTFFExplicit::calculate_all on coordinates resident in HBM
(pair kernels, merge, tail strips, histogram read-back and bookkeeping) + the
Debye transform (sinc table, bf16-split MFMA GEMM, intensity profiles).

Other workloads: --config 1 (2k-atom globule, FF-avg), --config 2 (20k-atom
globule + 6k waters + 1 A excluded-volume grid, FF-grid), --config 4 (rigid-
body scan: bodies of 3k/5k/8k/12k atoms, 512 conformations per step in one
graph launch), --config 5 (serving: Poisson arrivals of mixed requests at 70 %
of the measured service rate; one step = one request).

In the grid workloads the dummy-dummy pairs (integer lattice points) are
counted by FFT correlation of the dummy grid rather than one by one, with
identical histograms; there "pair evaluations" counts the pairs the
histograms cover, and the roofline uses only the pairs the pair kernel itself
evaluated.

The timed region brackets exactly --steps steps with device synchronisation
on both sides. Prints one JSON line: pair evaluations per second, profiles
per second, per-step latency percentiles, the pair kernel's roofline against
the FP32 VALU peak, and the CPU oracle (a restatement of the reference's
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
METRIC = "Gpair-evals/s + I(q) profiles/s, p50/p99 latency; GPU vs CPU at ≤1e-4 rel err"

CONFIGS = {
    1: dict(kind="structure", name="2k-atom globule, FF-avg, 0.25 A bins, 100 q in 0.01-0.5", n=2000,
            axes=(20, 20, 20), r_excl=2.0, waters=0, bw=0.25, mode="ffavg", q=(0.01, 0.5, 100)),
    2: dict(kind="grid", name="20k-atom globule + 6k waters, FF-grid (1 A excluded-volume grid), 0.25 A bins, "
            "200 q in 0-0.5", n=20000, axes=(45, 45, 45), r_excl=2.0, waters=6000, bw=0.25, q=(0.0, 0.5, 200)),
    3: dict(kind="structure", name="250k-atom ellipsoid + 60k waters, FF-explicit, 0.1 A bins, 1000 q to 1.5",
            n=250000, axes=(120, 90, 70), r_excl=1.8, waters=60000, bw=0.1, mode="ffexplicit",
            q=(0.0015, 1.5, 1000)),
    4: dict(kind="rigid", name="rigid-body scan: bodies of 3k/5k/8k/12k atoms, 512 conformations per step, "
            "XYZW weights, 0.25 A bins, default q axis", sizes=(3000, 5000, 8000, 12000), n_conf=512, bw=0.25),
    5: dict(kind="serve", name="serving: Poisson arrivals at 70 % of the measured service rate, structures "
            "log-uniform 1k-100k atoms (+30 % waters), modes uniform over FF-avg/FF-explicit/FF-grid, 200 q in 0-0.5",
            pool=48, n_min=1000, n_max=100000, load=0.7, slots=4, budget_gb=200.0),
}
I0 = {3: 6.0, 7: 7.0, 11: 8.0, 12: 16.0}  # XYZW weights: electron counts of C, N, O, S


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


def traffic_from_profiles():
    f = ROOT / "profiles" / "pair_hist_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def pairs_of(n, m):
    return n * (n - 1) / 2 + n * m + m * (m - 1) / 2


def grow_sample(run, total_units, first_frac, budget_s):
    """Run run(frac) on growing prefixes of the workload until one run takes
    at least 40 % of the budget (or the whole workload ran); pair work grows
    with the square of the prefix."""
    frac = min(1.0, first_frac)
    while True:
        t = time.perf_counter()
        info = run(frac)
        dt = time.perf_counter() - t
        if frac >= 1.0 or dt >= 0.4 * budget_s:
            return frac, dt, info
        frac = min(1.0, frac * float(np.sqrt(0.8 * budget_s / max(dt, 1e-3))))


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def bench_structure(cfg, args):
    import ausaxs_amd as A

    atoms = A.synth_globule(args.config, cfg["n"], cfg["axes"], cfg["r_excl"])
    waters = A.synth_shell(args.config + 1, cfg["waters"], cfg["axes"], 3.0) if cfg["waters"] else None
    q = np.linspace(*cfg["q"])
    mode = A.MODE_FFAVG if cfg["mode"] == "ffavg" else A.MODE_FFEXPLICIT
    settings = A.default_settings(bin_width=cfg["bw"], debye_terms=args.debye_terms, tile_j=args.tile_j)
    t0 = time.perf_counter()
    st = A.Structure(mode, atoms, waters, settings, z_exv=1.0)
    log(f"resident structure built in {time.perf_counter() - t0:.2f} s")
    st.set_timing(True)
    last = {}

    def step():
        h = st.calculate()
        last["h"] = h
        return h.debye_transform(q)

    res = timed_steps(step, args)
    tm = st.last_timing()
    res.update(pairs=tm["pairs"], kernel_pairs=tm["kernel_pairs"], profiles_per_step=1, kernel_ms=tm["pair_ms"],
               items=tm["n_items"],
               extra={"atoms": len(atoms), "waters": 0 if waters is None else len(waters), "q_points": len(q),
                      "debye_terms": args.debye_terms})

    def cpu(budget):
        import oracle as O

        th = host_threads()
        omode = O.FFAVG if cfg["mode"] == "ffavg" else O.FFEXPLICIT
        w = waters if waters is not None else np.zeros((0, 4), np.float32)

        def run(frac):
            ns, ms = int(len(atoms) * frac), int(len(w) * frac)
            h = O.hist_ff(omode, atoms[:ns], w[:ms], cfg["bw"], 8000, 1.0, th)
            I, _ = O.debye_ff(omode, h, q, cfg["bw"])
            return ns, ms, h, I

        frac, dt, (ns, ms, h_cpu, I_cpu) = grow_sample(run, len(atoms), 12000 / len(atoms), budget)
        pairs = pairs_of(ns, ms)
        sample = "full workload" if frac >= 1.0 else f"first {ns} atoms + {ms} waters of the workload"
        out = dict(value=pairs / dt / 1e9, unit="Gpair-evals/s", cores=th, kind="port",
                   sample=f"{sample} ({pairs:.3g} pairs, histogram + I(q), {dt:.2f} s)")
        if frac >= 1.0:
            # the GPU step's result against the CPU restatement on the same input:
            # pair counts per bin and form-factor row, and I(q) (bf16-split Debye)
            try:
                h = last["h"]
                same = h.max_bin == h_cpu["max_bin"] and all(
                    np.array_equal(h.get(which)[1], h_cpu[key]["count"])
                    for which, key in ((A.AA, "aa"), (A.AW, "aw"), (A.WW, "ww")))
                I_gpu = h.debye_transform(q)
                err = float(np.max(np.abs(I_gpu - I_cpu) / np.abs(I_cpu)))
                out["parity"] = {"counts_identical": bool(same), "iq_max_rel_err": err, "iq_tolerance": 1e-4}
                log(f"GPU vs CPU: counts identical {same}, I(q) max relative error {err:.2e}")
            except Exception as e:  # the baseline line must still be printed
                out["parity"] = {"error": str(e)}
        return out

    return res, cpu


def bench_grid(cfg, args):
    import ausaxs_amd as A

    g = A.synth_globule(args.config, cfg["n"], cfg["axes"], cfg["r_excl"])
    w = A.synth_shell(args.config + 1, cfg["waters"], cfg["axes"], 3.0)
    xyz, ff = g[:, :3].astype(np.float64), g[:, 3].view(np.int32).copy()
    wxyz = w[:, :3].astype(np.float64)
    q = np.linspace(*cfg["q"])
    settings = A.default_settings(bin_width=cfg["bw"], debye_terms=args.debye_terms, tile_j=args.tile_j)
    t0 = time.perf_counter()
    st = A.GridStructure(xyz, ff, wxyz, settings)
    log(f"resident grid structure built in {time.perf_counter() - t0:.2f} s (dummies from the 1 A grid)")
    st.set_timing(True)

    def step():
        return st.calculate().debye_transform(q)

    res = timed_steps(step, args)
    tm = st.last_timing()
    res.update(pairs=tm["pairs"], kernel_pairs=tm["kernel_pairs"], profiles_per_step=1, kernel_ms=tm["pair_ms"],
               items=tm["n_items"],
               extra={"atoms": len(xyz), "waters": len(wxyz), "q_points": len(q), "debye_terms": args.debye_terms,
                      "lattice_pairs_per_step": tm["lattice_pairs"]})

    def cpu(budget):
        import oracle as O

        th = host_threads()
        n_x = len(O.grid_exv(xyz, ff, wxyz)[0])
        t = time.perf_counter()
        h = O.hist_ffgrid(xyz, ff, wxyz, cfg["bw"], 8000, th)
        O.debye_ffgrid(h, q, cfg["bw"])
        dt = time.perf_counter() - t
        pairs = pairs_of(len(xyz) + len(wxyz) + n_x, 0)
        return dict(value=pairs / dt / 1e9, unit="Gpair-evals/s", cores=th, kind="port",
                    sample=f"full workload incl. grid ({pairs:.3g} pairs, histogram + I(q), {dt:.2f} s)")

    return res, cpu


def rigid_inputs(cfg, seed):
    import ausaxs_amd as A

    bodies, weights, x0 = [], [], 0.0
    radii = [20.0 * (n / 2000.0) ** (1 / 3) for n in cfg["sizes"]]
    for k, (n, r) in enumerate(zip(cfg["sizes"], radii)):
        if k:
            x0 += radii[k - 1] + r + 8.0
        gl = A.synth_globule(seed + k, n, (r, r, r), 2.0)
        bodies.append(gl[:, :3].astype(np.float64) + np.array([x0, 0.0, 0.0]))
        weights.append(np.array([I0[int(t)] for t in gl[:, 3].view(np.int32)], np.float32))
    rng = np.random.default_rng(seed)
    T = np.zeros((cfg["n_conf"], len(bodies), 12))
    for c in range(cfg["n_conf"]):
        for b, xyz in enumerate(bodies):
            R = A.random_rotation(rng)
            cen = xyz.mean(0)
            T[c, b] = A.rigid_transform(R, cen - R @ cen + rng.uniform(-10, 10, 3))
    return bodies, weights, T


def bench_rigid(cfg, args):
    import ausaxs_amd as A

    bodies, weights, T = rigid_inputs(cfg, args.config)
    q = A.default_q_axis()
    settings = A.default_settings(bin_width=cfg["bw"])
    t0 = time.perf_counter()
    if args.fixed_tiles:
        settings.tile_j = args.tile_j
    scan = A.RigidScan(bodies, weights, settings=settings, autotune=not args.fixed_tiles)
    log(f"bodies resident, self terms cached in {time.perf_counter() - t0:.2f} s")
    t0 = time.perf_counter()
    scan.scan(T, q)
    info = scan.info()
    log(f"first scan (plan, autotune, graph capture) {time.perf_counter() - t0:.2f} s: {info}")

    def step():
        return scan.scan(T, q)

    res = timed_steps(step, args)
    info = scan.info()
    res.update(pairs=info["pairs"], profiles_per_step=len(T), kernel_ms=None, items=info["n_items"],
               extra={"bodies": [len(b) for b in bodies], "conformations": len(T), "q_points": len(q),
                      "tile_j": info["tile_j"], "persistent_grid": info["grid"], "lds_copies": info["copies"],
                      "graph_ms": info["scan_ms"]})

    def cpu(budget):
        import oracle as O

        th = host_threads()
        sizes = [len(b) for b in bodies]
        ident = A.rigid_transform(np.eye(3), np.zeros(3))
        self_pts = np.concatenate([O.rigid_transform(b, w, ident) for b, w in zip(bodies, weights)])
        n = 0
        t = time.perf_counter()
        while n < len(T) and (time.perf_counter() - t) < 0.5 * budget:
            moved = np.concatenate([O.rigid_transform(b, w, T[n, k]) for k, (b, w) in enumerate(zip(bodies, weights))])
            h = O.hist_partial(self_pts, moved, sizes, None, cfg["bw"], 8000, th)
            O.debye_simple(h, q, cfg["bw"])
            n += 1
        dt = time.perf_counter() - t
        pairs = n * info["pairs"] / len(T)
        return dict(value=pairs / dt / 1e9, unit="Gpair-evals/s", cores=th, kind="port",
                    sample=f"first {n} of {len(T)} conformations ({pairs:.3g} inter-body pairs, partial "
                           f"histograms + I(q), {dt:.2f} s)")

    return res, cpu


def serve_pool(cfg, seed):
    import ausaxs_amd as A

    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(cfg["n_min"]), np.log(cfg["n_max"]), cfg["pool"])).astype(int)
    pool = []
    for k, n in enumerate(sizes):
        r = 20.0 * (n / 2000.0) ** (1 / 3)
        g = A.synth_globule(seed * 1000 + k, int(n), (r, r, r), 2.0)
        w = A.synth_shell(seed * 1000 + 500 + k, max(1, int(0.3 * n)), (r, r, r), 3.0)
        pool.append((g[:, :3].astype(np.float64), g[:, 3].view(np.int32).copy(), w[:, :3].astype(np.float64)))
    return pool


def bench_serve(cfg, args):
    import threading

    import ausaxs_amd as A

    modes = (A.MODE_FFAVG, A.MODE_FFEXPLICIT, A.MODE_FFGRID)
    t0 = time.perf_counter()
    pool = serve_pool(cfg, args.config)
    log(f"request pool of {len(pool)} structures built in {time.perf_counter() - t0:.1f} s")
    q = np.linspace(0.0, 0.5, 200)
    rng = np.random.default_rng(args.config + 1)
    n_req = args.requests

    def draw(count):
        return [(int(rng.integers(len(pool))), modes[int(rng.integers(3))]) for _ in range(count)]

    srv = A.Server(slots=cfg["slots"], budget_gb=cfg["budget_gb"])

    def submit(req):
        xyz, ff, w = pool[req[0]]
        return srv.submit(req[1], xyz, ff, w, q)

    # warm-up, then the closed-loop service rate on a burst
    for i in [submit(r) for r in draw(max(args.warmup, 1) * 4)]:
        srv.wait(i)
    burst = draw(max(30, n_req // 8))
    s0 = srv.stats()["pairs"]
    t = time.perf_counter()
    for i in [submit(r) for r in burst]:
        srv.wait(i)
    mu = len(burst) / (time.perf_counter() - t)
    lam = cfg["load"] * mu
    log(f"service rate {mu:.2f} requests/s on a burst of {len(burst)} "
        f"({(srv.stats()['pairs'] - s0) / len(burst):.3g} pairs per request); arrivals at {lam:.2f}/s")

    # open loop: Poisson arrivals at lam
    reqs = draw(n_req)
    gaps = rng.exponential(1.0 / lam, n_req)
    ids = [None] * n_req
    s0 = srv.stats()["pairs"]
    k0 = srv.stats()["kernel_pairs"]
    t_start = time.perf_counter()
    t_next = t_start

    def submitter():
        nonlocal t_next
        for k, r in enumerate(reqs):
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids[k] = submit(r)
            t_next += gaps[k]

    th = threading.Thread(target=submitter)
    th.start()
    lat, queue = [], []
    for k in range(n_req):
        while ids[k] is None:
            time.sleep(1e-4)
        _, l_ms, q_ms = srv.wait(ids[k])
        lat.append(l_ms)
        queue.append(q_ms)
    th.join()
    elapsed = time.perf_counter() - t_start
    st = srv.stats()
    pairs = st["pairs"] - s0
    kernel_pairs = st["kernel_pairs"] - k0
    log(f"{n_req} requests in {elapsed:.2f} s; peak in flight {st['peak_in_flight']}, peak admitted "
        f"{st['peak_gb']:.1f} GiB; queue p50 {np.percentile(queue, 50):.1f} ms")
    res = dict(elapsed=elapsed, lat_ms=np.array(lat), pairs=pairs / n_req, kernel_pairs=kernel_pairs / n_req,
               profiles_per_step=1, kernel_ms=None, items=0, steps_override=n_req,
               extra={"requests": n_req, "arrival_rate_per_s": round(lam, 3), "service_rate_per_s": round(mu, 3),
                      "slots": cfg["slots"], "memory_budget_gb": cfg["budget_gb"],
                      "peak_in_flight": st["peak_in_flight"], "peak_admitted_gb": round(st["peak_gb"], 2),
                      "queue_ms_p50": round(float(np.percentile(queue, 50)), 3), "q_points": len(q)})

    def cpu(budget):
        import oracle as O

        th_ = host_threads()
        done, cpu_pairs = 0, 0.0
        t = time.perf_counter()
        for r in reqs:
            if time.perf_counter() - t > 0.7 * budget:
                break
            xyz, ff, w = pool[r[0]]
            if r[1] == A.MODE_FFGRID and len(xyz) > 20000:
                continue  # bounded sample: large grid requests take minutes on the host
            if r[1] == A.MODE_FFGRID:
                nx = len(O.grid_exv(xyz, ff, w)[0])
                h = O.hist_ffgrid(xyz, ff, w, 0.25, 8000, th_)
                O.debye_ffgrid(h, q, 0.25)
                n = len(xyz) + len(w) + nx
                cpu_pairs += n * (n - 1) / 2
            else:
                om = O.FFAVG if r[1] == A.MODE_FFAVG else O.FFEXPLICIT
                h = O.hist_ff(om, O.pack_ff(xyz, ff), O.pack_ff(w, np.ones(len(w), np.int32)), 0.25, 8000, 1.0, th_)
                O.debye_ff(om, h, q, 0.25)
                cpu_pairs += pairs_of(len(xyz), len(w))
            done += 1
        dt = time.perf_counter() - t
        return dict(value=cpu_pairs / dt / 1e9, unit="Gpair-evals/s", cores=th_, kind="port",
                    sample=f"{done} requests of the arrival sequence (grid requests above 20k atoms skipped), "
                           f"{cpu_pairs:.3g} pairs, {dt:.2f} s, {done / dt:.2f} profiles/s")

    return res, cpu


def timed_steps(step, args):
    for _ in range(args.warmup):
        step()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        lat.append(time.perf_counter() - ts)
    elapsed = time.perf_counter() - t_start
    return dict(elapsed=elapsed, lat_ms=np.array(lat) * 1e3)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, choices=sorted(CONFIGS))
    ap.add_argument("--debye-terms", type=int, default=3)
    ap.add_argument("--tile-j", type=int, default=8000, help="columns per work item")
    ap.add_argument("--fixed-tiles", action="store_true",
                    help="rigid scan: use --tile-j and one workgroup per item instead of the autotuner")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU work for the baseline sample")
    ap.add_argument("--requests", type=int, default=400, help="serving: requests in the timed open-loop run")
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
    fn = {"structure": bench_structure, "grid": bench_grid, "rigid": bench_rigid, "serve": bench_serve}[cfg["kind"]]
    res, cpu_fn = fn(cfg, args)
    elapsed, lat_ms, pairs = res["elapsed"], res["lat_ms"], res["pairs"]
    if "steps_override" in res:  # serving: one step = one request of the open-loop run
        args.steps = res["steps_override"]
    value = pairs * args.steps / elapsed / 1e9
    log(f"steps {args.steps}: {elapsed:.3f} s, {pairs:.4g} pairs per step, items {res['items']}")

    roofline = None
    if res["kernel_ms"] and cfg["kind"] != "serve":
        # the pairs pair_hist_kernel itself evaluated in its launch (without the
        # strips' tail pairs and the dummy pairs counted by lattice correlation)
        kern_s = res["kernel_ms"] / 1e3
        kpairs = res.get("kernel_pairs", pairs)
        achieved = kpairs * FLOP_PER_PAIR / kern_s / 1e12
        roofline = {"bound": "mfma", "engine": "fp32 VALU (no matrix op on this path; 157.3 TF is the FP32 peak)",
                    "kernel": "pair_hist_kernel", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
                    "traffic": traffic_from_profiles() if args.config == 3 else None,
                    "kernel_ms": round(res["kernel_ms"], 4), "kernel_pairs": kpairs,
                    "pairs_per_s": round(kpairs / kern_s, 1)}
    elif cfg["kind"] == "serve":
        # pairs the pair kernel evaluated over the window (grid requests count
        # their dummy-dummy pairs through lattice correlation, not here)
        achieved = res["kernel_pairs"] * args.steps / elapsed * FLOP_PER_PAIR / 1e12
        roofline = {"bound": "mfma", "engine": "fp32 VALU; whole serving window at the offered load, idle included",
                    "kernel": "pair_hist_kernel (serving)", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 5), "traffic": None}
    else:
        # whole graph launch (transform, pair, merge, strips, finalize, Debye) per step
        g_s = res["extra"]["graph_ms"] / 1e3
        achieved = pairs * FLOP_PER_PAIR / g_s / 1e12
        roofline = {"bound": "mfma", "engine": "fp32 VALU; timed over the whole graph launch",
                    "kernel": "pair_hist_kernel (graph)", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 5), "traffic": None,
                    "graph_ms": round(res["extra"]["graph_ms"], 4), "pairs_per_s": round(pairs / g_s, 1)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_fn(args.cpu_budget)
        log(f"cpu baseline: {cpu['value']:.3f} {cpu['unit']} on {cpu['cores']} threads ({cpu['sample']})")

    config = {"workload": cfg["name"], "pairs_per_step": pairs, "parallelism": "single-gpu"}
    config.update(res["extra"])
    out = {
        "metric": METRIC,
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
        "config": config,
        "profiles_per_s": round(res["profiles_per_step"] * args.steps / elapsed, 4),
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 4),
                       "p99": round(float(np.percentile(lat_ms, 99)), 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
