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
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E (AMD published)
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
            pool=48, n_min=1000, n_max=100000, load=0.7, slots=5, budget_gb=200.0),
}
I0 = {3: 6.0, 7: 7.0, 11: 8.0, 12: 16.0}  # XYZW weights: electron counts of C, N, O, S


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """(threads the CPU baseline uses, facts about the host): every logical
    core this process may run on, capped by OMP_NUM_THREADS where the host
    sets it (a GPU box's CPU share per GPU)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(affinity, int(share)) if share.isdigit() and int(share) > 0 else affinity
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    facts = {"nproc": os.cpu_count(), "affinity_cores": affinity, "cpu_model": model,
             "thread_limit": f"OMP_NUM_THREADS={share}" if share else "none"}
    return max(1, threads), facts


CPU_PROTOCOL = "1 warm-up run, then the median of 5 timed runs"
CPU_BUILD = ("oracle restatement built on this host with the reference's release flags "
             "(-O3 -ffast-math -fno-finite-math-only -march=native; reassociation and FMA contraction off)")


def median_runs(fn, reps=5):
    """One warm-up call, then `reps` timed calls: (median seconds, all times, last result)."""
    fn()
    times, info = [], None
    for _ in range(reps):
        t = time.perf_counter()
        info = fn()
        times.append(time.perf_counter() - t)
    return float(np.median(times)), times, info


def pick_fraction(run, target_s, first_frac):
    """Prefix fraction of a workload whose run takes about target_s: probe runs
    on growing prefixes (pair work grows with the square of the prefix; a
    small prefix also pays fixed costs such as the thread-private histograms)
    until one takes half the target, then the whole workload if it would take
    at most 1.3 times the target."""
    frac = min(1.0, first_frac)
    while True:
        t = time.perf_counter()
        run(frac)
        dt = time.perf_counter() - t
        grow = float(np.sqrt(target_s / max(dt, 1e-3)))
        if frac >= 1.0 or dt >= 0.5 * target_s:
            return 1.0 if frac * grow * 1.14 >= 1.0 else frac * grow
        frac = min(1.0, frac * grow)


def baseline(value, unit, threads, facts, sample, times):
    out = dict(value=value, unit=unit, cores=threads, kind="port", sample=sample, protocol=CPU_PROTOCOL,
               build=CPU_BUILD, runs_s=[round(t, 4) for t in times])
    out.update(facts)
    return out


def traffic_from_profiles():
    f = ROOT / "profiles" / "pair_hist_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def tile_kw(args):
    """The work-item width option, if given (else the library default)."""
    return {} if args.tile_j is None else {"tile_j": args.tile_j}


def pairs_of(n, m):
    return n * (n - 1) / 2 + n * m + m * (m - 1) / 2


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def bench_structure(cfg, args):
    import ausaxs_amd as A

    atoms = A.synth_globule(args.config, cfg["n"], cfg["axes"], cfg["r_excl"])
    waters = A.synth_shell(args.config + 1, cfg["waters"], cfg["axes"], 3.0) if cfg["waters"] else None
    q = np.linspace(*cfg["q"])
    mode = A.MODE_FFAVG if cfg["mode"] == "ffavg" else A.MODE_FFEXPLICIT
    settings = A.default_settings(bin_width=cfg["bw"], debye_terms=args.debye_terms, **tile_kw(args))
    t0 = time.perf_counter()
    st = A.Structure(mode, atoms, waters, settings, z_exv=1.0)
    log(f"resident structure built in {time.perf_counter() - t0:.2f} s")
    st.set_timing(True)
    last = {}

    kernel_ms, merge_ms = [], []

    def step():
        h = st.calculate()
        last["h"] = h
        # the pair kernel's and the merge's launch times (HIP events on the
        # structure's stream; calculate() has already synchronised on its read-back)
        tm = st.last_timing()
        kernel_ms.append(tm["pair_ms"])
        merge_ms.append(tm["merge_ms"])
        return h.debye_transform(q)

    res = timed_steps(step, args)
    tm = st.last_timing()
    timed_ms = kernel_ms[-args.steps:]  # the launches of the timed steps
    res.update(pairs=tm["pairs"], kernel_pairs=tm["kernel_pairs"], profiles_per_step=1,
               kernel_ms=float(np.mean(timed_ms)),
               merge_ms=float(np.mean(merge_ms[-args.steps:])), merge_bytes=tm["merge_bytes"],
               items=tm["n_items"],
               extra={"atoms": len(atoms), "waters": 0 if waters is None else len(waters), "q_points": len(q),
                      "debye_terms": args.debye_terms})

    def cpu(budget):
        import oracle as O

        th, facts = host_cpu()
        omode = O.FFAVG if cfg["mode"] == "ffavg" else O.FFEXPLICIT
        w = waters if waters is not None else np.zeros((0, 4), np.float32)

        def run(frac):
            ns, ms = max(2, int(len(atoms) * frac)), int(len(w) * frac)
            h = O.hist_ff(omode, atoms[:ns], w[:ms], cfg["bw"], 8000, 1.0, th)
            O.debye_ff(omode, h, q, cfg["bw"])
            return ns, ms

        with O.release_build():
            # about budget / 6 per run: 1 warm-up + 5 timed runs fill the budget
            frac = pick_fraction(run, budget / 6, 12000 / len(atoms))
            med, times, (ns, ms) = median_runs(lambda: run(frac))
        pairs = pairs_of(ns, ms)
        sample = "full workload" if ns == len(atoms) else f"first {ns} atoms + {ms} waters of the workload"
        out = baseline(pairs / med / 1e9, "Gpair-evals/s", th, facts,
                       f"{sample} ({pairs:.3g} pairs, histogram + I(q); median {med:.3f} s)", times)
        # the GPU step's result against the checker (the oracle's IEEE build) on
        # the full workload, outside the timing: pair counts per bin and
        # form-factor row, and I(q) (bf16-split Debye) against the FP64 oracle
        try:
            h = last["h"]
            h_cpu = O.hist_ff(omode, atoms, w, cfg["bw"], 8000, 1.0, th)
            I_cpu, _ = O.debye_ff(omode, h_cpu, q, cfg["bw"])
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
    settings = A.default_settings(bin_width=cfg["bw"], debye_terms=args.debye_terms, **tile_kw(args))
    t0 = time.perf_counter()
    st = A.GridStructure(xyz, ff, wxyz, settings)
    log(f"resident grid structure built in {time.perf_counter() - t0:.2f} s (dummies from the 1 A grid)")
    st.set_timing(True)

    kernel_ms, merge_ms = [], []

    def step():
        h = st.calculate()
        tm = st.last_timing()
        kernel_ms.append(tm["pair_ms"])
        merge_ms.append(tm["merge_ms"])
        return h.debye_transform(q)

    res = timed_steps(step, args)
    tm = st.last_timing()
    res.update(pairs=tm["pairs"], kernel_pairs=tm["kernel_pairs"], profiles_per_step=1,
               kernel_ms=float(np.mean(kernel_ms[-args.steps:])),
               merge_ms=float(np.mean(merge_ms[-args.steps:])), merge_bytes=tm["merge_bytes"],
               items=tm["n_items"],
               extra={"atoms": len(xyz), "waters": len(wxyz), "q_points": len(q), "debye_terms": args.debye_terms,
                      "lattice_pairs_per_step": tm["lattice_pairs"]})

    def cpu(budget):
        import oracle as O

        th, facts = host_cpu()

        def run(frac):
            # a prefix of the atoms with its own grid and waters (the same generator's structure, smaller)
            na, nw = max(2, int(len(xyz) * frac)), max(1, int(len(wxyz) * frac))
            h = O.hist_ffgrid(xyz[:na], ff[:na], wxyz[:nw], cfg["bw"], 8000, th)
            O.debye_ffgrid(h, q, cfg["bw"])
            return na, nw

        with O.release_build():
            frac = pick_fraction(run, budget / 6, 0.15)
            med, times, (na, nw) = median_runs(lambda: run(frac))
        n_x = len(O.grid_exv(xyz[:na], ff[:na], wxyz[:nw])[0])
        pairs = pairs_of(na + nw + n_x, 0)
        sample = "full workload" if na == len(xyz) else f"first {na} atoms + {nw} waters with their grid"
        return baseline(pairs / med / 1e9, "Gpair-evals/s", th, facts,
                        f"{sample} ({n_x} dummies, {pairs:.3g} pairs incl. grid, histogram + I(q); "
                        f"median {med:.3f} s)", times)

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
        if args.tile_j is not None:
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

        th, facts = host_cpu()
        sizes = [len(b) for b in bodies]
        ident = A.rigid_transform(np.eye(3), np.zeros(3))
        self_pts = np.concatenate([O.rigid_transform(b, w, ident) for b, w in zip(bodies, weights)])
        k = [0]

        def run():
            # one conformation: moved bodies, partial histograms (inter-body blocks) and I(q)
            n = k[0] % len(T)
            k[0] += 1
            moved = np.concatenate([O.rigid_transform(b, w, T[n, j]) for j, (b, w) in enumerate(zip(bodies, weights))])
            h = O.hist_partial(self_pts, moved, sizes, None, cfg["bw"], 8000, th)
            O.debye_simple(h, q, cfg["bw"])

        with O.release_build():
            med, times, _ = median_runs(run)
        pairs = info["pairs"] / len(T)
        return baseline(pairs / med / 1e9, "Gpair-evals/s", th, facts,
                        f"one conformation per run ({pairs:.3g} inter-body pairs, partial histograms + I(q); "
                        f"median {med:.3f} s)", times)

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
    lam = args.arrival_rate if args.arrival_rate > 0 else cfg["load"] * mu
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
                      "slots": cfg["slots"], "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                      "memory_budget_gb": cfg["budget_gb"],
                      "peak_in_flight": st["peak_in_flight"], "peak_admitted_gb": round(st["peak_gb"], 2),
                      "queue_ms_p50": round(float(np.percentile(queue, 50)), 3), "q_points": len(q)})

    def cpu(budget):
        import oracle as O

        th_, facts = host_cpu()

        def one(r):
            xyz, ff, w = pool[r[0]]
            if r[1] == A.MODE_FFGRID:
                nx = len(O.grid_exv(xyz, ff, w)[0])
                h = O.hist_ffgrid(xyz, ff, w, 0.25, 8000, th_)
                O.debye_ffgrid(h, q, 0.25)
                n = len(xyz) + len(w) + nx
                return n * (n - 1) / 2
            om = O.FFAVG if r[1] == A.MODE_FFAVG else O.FFEXPLICIT
            h = O.hist_ff(om, O.pack_ff(xyz, ff), O.pack_ff(w, np.ones(len(w), np.int32)), 0.25, 8000, 1.0, th_)
            O.debye_ff(om, h, q, 0.25)
            return pairs_of(len(xyz), len(w))

        # requests of the arrival sequence in order (grid requests above 20k
        # atoms skipped: they take minutes on the host); the first one is the
        # warm-up, then the aggregate rate over the budget
        seq = [r for r in reqs if not (r[1] == A.MODE_FFGRID and len(pool[r[0]][0]) > 20000)]
        with O.release_build():
            one(seq[0])
            done, cpu_pairs, times = 0, 0.0, []
            t = time.perf_counter()
            for r in seq[1:]:
                if done >= 5 and time.perf_counter() - t > 0.7 * budget:
                    break
                t1 = time.perf_counter()
                cpu_pairs += one(r)
                times.append(time.perf_counter() - t1)
                done += 1
            dt = time.perf_counter() - t
        out = baseline(cpu_pairs / dt / 1e9, "Gpair-evals/s", th_, facts,
                       f"{done} requests of the arrival sequence after one warm-up request (grid requests above "
                       f"20k atoms skipped), {cpu_pairs:.3g} pairs, {dt:.2f} s, {done / dt:.2f} profiles/s", times)
        out["protocol"] = "1 warm-up request, then the aggregate rate over the requests that follow"
        return out

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


def measured_ceilings(A, pairs_per_s, achieved_tflops):
    """The device's own issue ceilings (ausaxs_issue_probe, run after the timed
    steps) for the pair kernel's two co-limiting pipes: FP32 VALU issue with
    vector-register operands, as TFLOP/s at 2 flop per FMA, and ds_add_u64 at
    hashed cells of a 4,926-cell LDS histogram (two rows of the 250k-atom
    workload's 2,463), as pair accumulations per second (one atomic per pair)."""
    cus = A.device_info()["compute_units"]
    r = A.issue_probe(4926, 5)
    fp32 = r["fp32_add_vgpr"] * cus * 64 * 2 / 1e3
    lds = r["lds_add_u64"] * cus * 64 * 1e9
    return {"probe": "ausaxs_issue_probe (wave instructions per CU per ns)",
            "rates": {k: round(v, 4) for k, v in r.items()},
            "fp32_tflops": round(fp32, 2), "frac_of_fp32": round(achieved_tflops / fp32, 4),
            "lds_atomic_pairs_per_s": round(lds, 1), "frac_of_lds_atomic": round(pairs_per_s / lds, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, choices=sorted(CONFIGS))
    ap.add_argument("--debye-terms", type=int, default=3)
    ap.add_argument("--tile-j", type=int, default=None, help="columns per work item (default: the library's)")
    ap.add_argument("--fixed-tiles", action="store_true",
                    help="rigid scan: use --tile-j and one workgroup per item instead of the autotuner")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds of CPU work for the baseline (1 warm-up + 5 timed runs of a sample)")
    ap.add_argument("--requests", type=int, default=400, help="serving: requests in the timed open-loop run")
    ap.add_argument("--arrival-rate", type=float, default=0.0,
                    help="serving: Poisson arrivals per second (default: 70 %% of the measured service rate)")
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

    cfg = CONFIGS[args.config]
    if cfg["kind"] == "serve":
        # a hardware queue for every stream with work: the server's slots plus
        # the library's default stream and HIP's null stream, with room to
        # spare. Streams beyond the process's queues share one, in order, so a
        # small request's kernels would wait behind a large request's on the
        # same queue. HIP reads the setting when it starts: set it before the
        # library loads.
        need = cfg["slots"] + 4
        if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < need:
            os.environ["GPU_MAX_HW_QUEUES"] = str(need)

    import ausaxs_amd as A

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
        roofline = {"bound": "valu", "engine": "fp32 VALU (no matrix op on this path; 157.3 TF is the FP32 peak)",
                    "kernel": "pair_hist_kernel", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
                    "traffic": traffic_from_profiles() if args.config == 3 else None,
                    "kernel_ms": round(res["kernel_ms"], 4), "kernel_pairs": kpairs,
                    "pairs_per_s": round(kpairs / kern_s, 1)}
        roofline["measured_ceilings"] = measured_ceilings(A, kpairs / kern_s, achieved)
        if res.get("merge_ms"):
            # the merge of the pair kernel's partial histograms: HBM-bound
            gbs = res["merge_bytes"] / (res["merge_ms"] / 1e3) / 1e9
            roofline["merge"] = {"bound": "hbm", "kernels": "merge_chunks_kernel + merge_rows_kernel",
                                 "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes": res["merge_bytes"],
                                 "ms": round(res["merge_ms"], 4)}
    elif cfg["kind"] == "serve":
        # pairs the pair kernel evaluated over the window (grid requests count
        # their dummy-dummy pairs through lattice correlation, not here)
        achieved = res["kernel_pairs"] * args.steps / elapsed * FLOP_PER_PAIR / 1e12
        roofline = {"bound": "valu", "engine": "fp32 VALU; whole serving window at the offered load, idle included",
                    "kernel": "pair_hist_kernel (serving)", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 5), "traffic": None}
    else:
        # whole graph launch (transform, pair, merge, strips, finalize, Debye) per step
        g_s = res["extra"]["graph_ms"] / 1e3
        achieved = pairs * FLOP_PER_PAIR / g_s / 1e12
        roofline = {"bound": "valu", "engine": "fp32 VALU; timed over the whole graph launch",
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
