"""Benchmark of the LOS transfer stage on one MI355X.

Workloads (seeded synthetic sources: A ~ N(0,1) per (field,L,m,k1,k2) x (k3/0.01)^-1.5 x
[Gaussian at 280 Mpc, sigma 20 Mpc + 10% late-time ramp]; 600 tau log-spaced in
[200, 14000] Mpc, tau0 = 14200 Mpc; J tables on x step 0.3):

  c2  (default) plain FP64 GPU drop-in: T+E(+B), m = 0,1,2, sources L = 0..6; 32x32 log
      (k1,k2) grid in [1e-5, 0.1]/Mpc (528 pairs k2 <= k1), 60 linear coarse k3 per pair,
      fine k3 step 1.5e-4/Mpc; 96 l in [2, 2000]; x in [0, 15000].
  c3  batched: 2048 (k1,k2) pairs out of 64 k drawn log-uniformly in [1e-5, 0.2]/Mpc,
      8 coarse k3 per pair, fine k3 step 3.5e-4/Mpc, so the triangle-limited fine k3 counts
      run from 8 to about 400 per pair; T+E(+B), m = 0,1,2, L = 0..6, 96 l in [2, 2000];
      every pair in one batch, the run captured in one hipGraph, tile shapes autotuned.
  c4  c2 extended to l_max = 3000 with 128 l values (precision study of the FP32 form).
  k3int  the bispectrum module's k3 integral of the c2 transfers: the T rows of every l3
      (M3 = 0) over the 528 pairs' fine k3 grids, against j_l3(k3 r) on 200 r in
      [13600, 14400] Mpc (x step 0.3), plain FP64, bit-identical to the CPU code.
  cl2  the power-spectrum module's (k1, k2, k3) sum of the c2 transfers: C_l^TT and C_l^EE
      for every l and m = 0, 1, 2 (rescaled sources), power-law primordial spectrum.

A step is one pass of the LOS stage over every pair of the workload (source re-layout,
k3 spline, LOS integral) with the source and projection tables already resident in HBM.
The JSON line reports elements/s (one element = one ptr2->transfer[tt][k1][k2][k3] value)
of the --form (default: the plain FP64 form, bit-identical to the FP64 CPU code), the other
forms timed on the same resident inputs with their errors against it, the HBM roofline of
the dominant kernel (algorithmic bytes / HIP-event duration) and the CPU oracle timed with
OpenMP on a bounded sample of pairs of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from song_amd import lib as L  # noqa: E402

METRIC = ("LOS transfer elements/s (k1,k2,k3,l,m), FP32-comp vs FP64 at 1e-5 rel; p99 lat")
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = {"fp64": 78.6, "fp64fast": 78.6, "fp32c": 157.3}  # MI355X vector FP64 / FP32

_C2 = dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=6, l_max_los_p=6, k_size=32,
           k_min=1e-5, k_max=0.1, k3_size=60, tau_size=600, tau_min=200.0, tau_max=14000.0,
           tau0=14200.0, l_size=96, l_max=2000, dk_fine=1.5e-4, xx_max=15000.0, xx_step=0.3)
WORKLOADS = {
    # the CPU reference configuration: T only, m = 0, L = 0..4, 16x16 log (k1,k2) in
    # [1e-4, 0.05], 400 tau, 48 l in [2, 1000], fine k3 step 2e-4
    "c1": dict(has_T=1, has_E=0, m_list=[0], l_max_los_t=4, l_max_los_p=4, k_size=16,
               k_min=1e-4, k_max=0.05, k3_size=60, tau_size=400, tau_min=200.0,
               tau_max=14000.0, tau0=14200.0, l_size=48, l_max=1000, dk_fine=2e-4,
               xx_max=15000.0, xx_step=0.3),
    "c2": _C2,
    "c3": dict(_C2, k_size=64, k_log=2, k_max=0.2, k3_size=8, dk_fine=3.5e-4, xx_max=6000.0),
    "c4": dict(_C2, l_size=128, l_max=3000),
    # serving: the pairs of the c3 grid with l up to 3000 (jobs pick l_max prefixes)
    "c5": dict(_C2, k_size=64, k_log=2, k_max=0.2, k3_size=8, dk_fine=3.5e-4, xx_max=6000.0,
               l_size=128, l_max=3000),
    "k3int": _C2,
    "cl2": _C2,
    # small workload for quick checks
    "tiny": dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=4, l_max_los_p=4, k_size=6,
                 k3_size=16, tau_size=128, l_size=12, l_max=400, dk_fine=1e-3, xx_max=1500.0),
}
PAIRS = {"c3": 2048}  # workloads that run an explicit list of pairs
GRAPH = {"c3"}        # workloads whose steps replay a captured graph

DESCR = {
    "c1": ("plain FP64 LOS, T only m=0 L=0..4, 16x16 log (k1,k2) in [1e-4,0.05], 60 coarse "
           "k3/pair, dk=2e-4, 400 tau, 48 l in [2,1000]"),
    "c2": ("plain FP64 LOS, T+E+B m=0,1,2 L=0..6, 32x32 log (k1,k2) in [1e-5,0.1], 60 coarse "
           "k3/pair, dk=1.5e-4, 600 tau, 96 l in [2,2000]"),
    "c3": ("batched: 2048 pairs of 64 log-uniform k in [1e-5,0.2], fine k3 8..400 per pair, "
           "T+E+B m=0,1,2 L=0..6, 600 tau, 96 l in [2,2000], one captured graph per step"),
    "c4": "c2 with 128 l in [2,3000]",
    "c5": ("serving: Poisson jobs of 1-64 pairs of the c3 grid, l_max in {500,1000,2000,3000}, "
           "T+E+B m=0,1,2, resident tables, HBM-budget admission, 3 slots"),
    "tiny": "small check workload",
    "k3int": ("k3 integral of the c2 T transfers: 96 l3 (M3=0) x 200 r in [13600,14400] x 528 "
              "pairs, 41,759 fine k3, j_l3 on x step 0.3, plain FP64"),
    "cl2": ("C_l^TT and C_l^EE of the c2 transfers: 96 l x m=0,1,2 x 528 pairs x 41,759 fine "
            "k3, rescaled sources, power-law P(k), FP64"),
}

FORMS = {"fp64": L.FORM_FP64, "fp64fast": L.FORM_FP64_FAST, "fp32c": L.FORM_FP32C}
# error bars of the fast forms against the oracle: |T - T_oracle| / max |T_oracle| over each
# (l,m) field; the row-normalised figure (per (tt, pair) row) is reported, not barred
FIELD_BAR = {"fp64fast": 1e-12, "fp32c": 1e-5}
POLICY = {"fp64fast": "StreamFP64", "fp32c": "StreamFP32C"}
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a copy
# kernel of known size), keyed by kernel name; the bench reports them for the kernel it ran
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "los_{workload}_hbm_traffic.json")


def kernel_name(form: str, shape, nl: int) -> str:
    """The LOS kernel a tuned shape launches, named as rocprofv3 lists it."""
    kind, nk, ch, pers = shape
    flag = "true" if pers else "false"
    if form == "fp64":
        return (f"los_sweep_exact<{nk}, {nl}, {flag}>" if kind == 1 else
                f"los_klane_exact<{nl}, {ch}>" if kind == 2 else f"los_integrate_exact<{nl}>")
    pol = POLICY[form]
    if kind == 1:
        pairs = "true" if (form == "fp64fast" or ch == 1) else "false"
        return f"los_sweep<{pol}, {nk}, {nl}, {flag}, {pairs}>"
    if kind == 2:
        return f"los_klane<{pol}, {nl}, {ch}>"
    return f"los_integrate_stream{'_q' if pers else ''}<{pol}, {nk}, {ch}>"


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter table, or None."""
    try:
        with open(TRAFFIC_TABLE.format(workload=workload)) as fh:
            tab = json.load(fh)
    except (OSError, ValueError):
        return None
    ent = tab.get("kernels", {}).get(kernel)
    return None if ent is None else float(ent["traffic_bytes"])


def select_pairs(cfg: L.Config, n: int):
    """n distinct (k1, k2) pairs of the grid, a seeded random choice, in grid order."""
    allp = [(i1, i2) for i1 in range(cfg.k_size) for i2 in range(i1 + 1)]
    if n >= len(allp):
        return allp
    pick = np.sort(np.random.default_rng(cfg.seed).choice(len(allp), n, replace=False))
    return [allp[i] for i in pick]


def cpu_info():
    """CPU model and CPU counts of this host (lscpu), and the CPUs this process may use."""
    info = {"model": None, "logical_cpus": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0))}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
    except (OSError, ValueError):
        pass
    return info


def _oracle_lib(debug: bool):
    name = "liblos_oracle.so" if debug else "liblos_oracle_nodebug.so"
    o = C.CDLL(os.path.join(ROOT, "oracle", name))
    o.los_ref_run_pairs.argtypes = [C.c_void_p] * 8 + [C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                       C.c_int, C.POINTER(C.c_long),
                                                       C.POINTER(C.c_double)]
    o.los_ref_run_pairs.restype = C.c_int
    o.los_ref_sizeof_transfers2.restype = C.c_long
    o.los_ref_transfer2_free.argtypes = [C.c_void_p] * 3
    return o


def _oracle_pass(o, scen, chunk, out=None):
    """One pass of the oracle over the pairs `chunk` (its transfer2_init loop body per pair:
    k3 list, k3 interpolation, OpenMP static loop over k3); (elements, seconds)."""
    st = scen.s
    size = o.los_ref_sizeof_transfers2()
    tr2 = C.create_string_buffer(size)
    # the stage settings the oracle reads come from the scenario's transfers2
    C.memmove(tr2, st["ptr2"].value, size)
    k1 = (C.c_int * len(chunk))(*[p[0] for p in chunk])
    k2 = (C.c_int * len(chunk))(*[p[1] for p in chunk])
    n = C.c_long(0)
    optr = out.ctypes.data_as(C.POINTER(C.c_double)) if out is not None else None
    t0 = time.perf_counter()
    rc = o.los_ref_run_pairs(st["ppr"], st["ppr2"], st["pba"], st["pth"], st["ppt2"], st["pbs"],
                             st["pbs2"], C.cast(tr2, C.c_void_p), k1, k2, len(chunk), C.byref(n),
                             optr)
    dt = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError("oracle sample failed")
    o.los_ref_transfer2_free(st["ppr2"], st["ppt2"], C.cast(tr2, C.c_void_p))
    return n.value, dt


def cpu_baseline(cfg: L.Config, pairs, target_s: float, tt2: int, sizes_of):
    """The CPU oracle timed on a bounded sample of the workload's pairs.

    Protocol: OpenMP on every CPU this process may use (OMP_PROC_BIND=close,
    OMP_PLACES=cores, set before any OpenMP runtime starts), the sample grown in pair
    order (spread over the pair list) until one pass takes about target_s / 12, then for
    each build (-DDEBUG, the reference's default, and without) one warm-up pass and the
    median of 5 timed passes. Returns the baseline record and the sample's oracle rows
    {pair: [tt][k3]} (from the DEBUG build) for the parity check."""
    scen = L.Scenario(cfg)
    stride = 24
    order = pairs[::stride] + [p for q, p in enumerate(pairs) if q % stride]
    per_pass = target_s / 12.0
    o = _oracle_lib(True)
    sample, t_grow, done = [], 0.0, 0
    while done < len(order) and t_grow < per_pass:
        chunk = order[done:done + 4]
        done += len(chunk)
        t_grow += _oracle_pass(o, scen, chunk)[1]
        sample += chunk
    sample = sorted(sample)
    n_el = sum(tt2 * sizes_of[p] for p in sample)
    rows = np.zeros(n_el)
    times = {}
    for debug in (True, False):
        lib = o if debug else _oracle_lib(False)
        _oracle_pass(lib, scen, sample)  # warm-up
        ts = []
        for rep in range(5):
            n, dt = _oracle_pass(lib, scen, sample, rows if (debug and rep == 0) else None)
            if n != n_el:
                raise RuntimeError(f"oracle sample: {n} elements, expected {n_el}")
            ts.append(dt)
        times["debug" if debug else "nodebug"] = float(np.median(ts))
        print(f"# cpu baseline ({'DEBUG' if debug else 'no DEBUG'}): median {np.median(ts):.2f} s"
              f" over {len(sample)} pairs", file=sys.stderr, flush=True)
    scen.close()
    ref, off = {}, 0
    for p in sample:
        k = sizes_of[p]
        ref[p] = rows[off:off + tt2 * k].reshape(tt2, k)
        off += tt2 * k
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    rec = {"value": n_el / times["nodebug"], "unit": "elements/s", "cores": threads,
           "kind": "port",
           "value_debug": n_el / times["debug"],
           "median_s": times, "cpu": cpu_info(),
           "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND",
                                                   "OMP_PLACES")},
           "sample": (f"{len(sample)} of {len(pairs)} (k1,k2) pairs of the same workload, "
                      f"{n_el} elements; oracle built gcc -O3 -fopenmp (value: without "
                      f"-DDEBUG, value_debug: with); 1 warm-up + median of 5 passes; "
                      "OpenMP static over k3 per pair")}
    return rec, ref


def relative_errors(got, ref, sizes, tt2_size):
    """Two normalisations of |got - ref| for each element of a [tt][pair][k3] array.

    field: divided by max|ref| over the whole (l,m) transfer function of the same tt
           (all pairs and k3), i.e. the sup-norm relative error of T_lm;
    row:   divided by max|ref| over the k3 row of the same (tt, pair); stricter, since
           rows whose values cancel to far below the field's scale are judged on
           their own scale.
    Single elements are never used as scales: the transfers oscillate in k3."""
    F = int(sum(sizes))
    g = np.asarray(got).reshape(tt2_size, F)
    r = np.asarray(ref).reshape(tt2_size, F)
    d = np.abs(g - r)
    fscale = np.abs(r).max(axis=1, keepdims=True)
    field = np.divide(d, fscale, out=np.zeros_like(d), where=fscale > 0)
    row = np.zeros_like(d)
    o = 0
    for n in sizes:
        sc = np.abs(r[:, o:o + n]).max(axis=1, keepdims=True)
        row[:, o:o + n] = np.divide(d[:, o:o + n], sc, out=np.zeros_like(d[:, o:o + n]),
                                    where=sc > 0)
        o += n
    return field, row


def loaded_columns(scen: L.Scenario, pairs):
    """Indices, inside one transfer type of gather(), of the k3 values of `pairs`, and the
    k3 count of each of those pairs."""
    offs = scen.pair_offsets()
    cols, sizes = [], []
    for p in sorted(pairs):
        o, n = offs[p]
        cols.append(np.arange(o, o + n))
        sizes.append(n)
    return np.concatenate(cols), sizes


def serve_bench(args, cfg):
    """Serving workload: Poisson arrivals at three rates, per form."""
    import threading

    scen = L.Scenario(cfg)
    scen.prepare()
    lm_l = scen.lm_l()
    fields = 3  # T, E and B (E implies B for m > 0 in this workload)
    allp = [(i1, i2) for i1 in range(cfg.k_size) for i2 in range(i1 + 1)]
    offs = scen.pair_offsets()
    rng = np.random.default_rng(cfg.seed + 5)

    def make_jobs(n):
        jobs = []
        for _ in range(n):
            k = int(rng.integers(1, min(64, len(allp)) + 1))
            pick = [allp[i] for i in rng.choice(len(allp), k, replace=False)]
            lmax = int(rng.choice([500, 1000, 2000, 3000]))
            n_lm = sum(1 for l in lm_l if l <= lmax)
            jobs.append((pick, lmax, fields * n_lm * sum(offs[p][1] for p in pick)))
        return jobs

    forms = [args.form] + [f for f in args.also.split(",") if f and f != args.form]
    out = {}
    for form in forms:
        t0 = time.time()
        srv = L.Server(scen, FORMS[form], budget_bytes=args.serve_budget * 2.0 ** 30, slots=3)
        print(f"# {form}: server ready in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
        # capacity: a burst admitted together, closed loop
        burst = make_jobs(48)
        srv.hold(True)
        ids = [srv.submit(p, lm) for p, lm, _ in burst]
        t0 = time.perf_counter()
        srv.hold(False)
        for i in ids:
            srv.wait(i)
        cap = len(burst) / (time.perf_counter() - t0)
        res = {"capacity_jobs_per_s": cap, "rates": []}
        for frac in (0.5, 0.8, 1.1):
            rate = frac * cap
            n = int(min(args.serve_jobs, max(40, rate * args.serve_seconds)))
            jobs = make_jobs(n)
            gaps = rng.exponential(1.0 / rate, n)
            ids = [None] * n

            def feed():
                t = time.perf_counter()
                for q, (p, lm, _) in enumerate(jobs):
                    t += gaps[q]
                    d = t - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    ids[q] = srv.submit(p, lm)

            st0 = srv.stats()
            th = threading.Thread(target=feed)
            th.start()
            lat, t_first, t_last = [], None, 0.0
            for q in range(n):
                while ids[q] is None:
                    time.sleep(1e-4)
                _, (ts, ta, td) = srv.wait(ids[q])
                lat.append(1e3 * (td - ts))
                t_first = ts if t_first is None else min(t_first, ts)
                t_last = max(t_last, td)
            th.join()
            st = srv.stats()
            el = sum(j[2] for j in jobs)
            lat = np.array(lat)
            r = {"offered_jobs_per_s": rate, "jobs": n, "p50_ms": float(np.percentile(lat, 50)),
                 "p99_ms": float(np.percentile(lat, 99)),
                 "p99_samples": int(np.count_nonzero(lat >= np.percentile(lat, 99))),
                 "elements_per_s": el / (t_last - t_first),
                 "launches": st["launches"] - st0["launches"],
                 "max_busy_slots": st["max_busy_slots"]}
            print(f"# {form} rate {rate:.1f}/s: p50 {r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms "
                  f"{r['elements_per_s']:.3g} el/s", file=sys.stderr, flush=True)
            res["rates"].append(r)
        res["resident_bytes"] = srv.stats()["resident_bytes"]
        res["budget_bytes"] = srv.stats()["budget_bytes"]
        srv.close()
        out[form] = res
    scen.close()
    head = out[args.form]["rates"][1]
    line = {
        "metric": METRIC, "value": head["elements_per_s"], "unit": "elements/s",
        "n_gpus": 1, "steps": 0, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.form != "fp32c" else "f32-compensated",
        "data": "synthetic (seeded source model, FP64 J tables built on the host)",
        "config": {"workload": f"c5: {DESCR['c5']}", "form": args.form,
                   "value_at": "0.8 x measured capacity"},
        "latency_ms": {"p50": head["p50_ms"], "p99": head["p99_ms"]},
        "serving": out,
    }
    print(json.dumps(line), flush=True)
    return 0


def _pair_k3_grids(lib, scen, cfg):
    """Fine k3 grid of every pair (k1-major, k2 <= k1) from transfer2_get_k3_list."""
    sizes = scen.k3_sizes()[:, 0].astype(np.int64)
    pair_off = np.concatenate([[0], np.cumsum(sizes)])
    k3 = np.empty(int(pair_off[-1]))
    st = scen.s
    q = 0
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            buf = np.zeros(int(sizes[q]) + 4)
            last = C.c_int()
            rc = lib.transfer2_get_k3_list(st["ppr"], st["ppr2"], st["ppt2"], st["pbs"],
                                           st["pbs2"], st["ptr2"], i1, i2,
                                           buf.ctypes.data_as(C.POINTER(C.c_double)),
                                           C.byref(last))
            if rc != 0:
                raise RuntimeError("transfer2_get_k3_list failed: " + scen.error_message())
            k3[pair_off[q]:pair_off[q + 1]] = buf[:sizes[q]]
            q += 1
    return sizes, pair_off, k3


def _copy_probe(args):
    """A plain device copy of --copy-probe-mib MiB (HBM counter calibration), if asked."""
    if args.copy_probe_mib > 0:
        eng = L.Engine(0)
        eng.copy_probe(args.copy_probe_mib * 2.0 ** 20)
        eng.close()


def cl2_bench(args, cfg):
    """The second-order C_l sum (spectra2.c:241-590) over the c2 transfers."""
    lib = L.load()
    t0 = time.time()
    scen = L.Scenario(cfg)
    scen.transfer2_init()
    flat = scen.gather()
    sizes, pair_off, k3 = _pair_k3_grids(lib, scen, cfg)
    n_k3 = int(pair_off[-1])
    k = scen.grids()[1]
    scen.close()
    n_l, m = cfg.l_size, list(cfg.m_list)
    n_lm = n_l * len(m)  # every l >= 2 carries m = 0, 1, 2
    fields = flat.reshape(-1, n_k3)
    T = fields[:n_lm].reshape(n_l, len(m), n_k3)
    E = fields[n_lm:2 * n_lm].reshape(n_l, len(m), n_k3)
    pk = 2.1e-9 * (k / 0.05) ** (0.965 - 1)
    used = np.ones((n_l, len(m)), dtype=np.int32)
    print(f"# cl2 inputs ({n_l} l x {len(m)} m, {len(sizes)} pairs, {n_k3} k3) in "
          f"{time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    _copy_probe(args)
    res, ms_tot = {}, 0.0
    for name, t in (("TT", T), ("EE", E)):
        h = L.Cl2(k, pk, pair_off, k3, m, t, t, used, 1)
        for _ in range(max(args.warmup, 1)):
            h.run(1)
        ms = h.run(args.steps)
        res[name] = (h.result(), ms)
        ms_tot += ms
        h.close()
    terms = 2 * n_l * len(m) * n_k3
    # per term: 8 B of each of the two transfer rows (k3 grid and weights are shared by rows)
    alg_bytes = 2 * (2 * 8.0 * n_l * len(m) * n_k3 + 8.0 * n_k3)
    achieved = alg_bytes / (ms_tot * 1e-3) / 1e9
    line = {
        "metric": "second-order C_l (k1,k2,k3) terms/s", "value": terms / (ms_tot * 1e-3),
        "unit": "terms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_tot, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (c2 transfers from the LOS stage on the GPU)",
        "config": {"workload": f"cl2: {DESCR['cl2']}", "terms": terms,
                   "ms_TT": res["TT"][1], "ms_EE": res["EE"][1]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("cl2", "los_cl2_partial + los_cl2_finish"),
                     "kernel": "los_cl2_partial + los_cl2_finish",
                     "kernel_ms": ms_tot / 2, "alg_bytes": alg_bytes / 2},
    }
    if not args.no_cpu:
        o = C.CDLL(os.path.join(ROOT, "oracle", "liblos_oracle.so"))
        f = o.los_ref_cl2
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        f.argtypes = [C.c_int, dp, dp, lp, dp, C.c_int, C.c_int, ip, dp, dp, ip, C.c_int, dp, dp]
        po = np.ascontiguousarray(pair_off, dtype=np.int64)
        mm = np.array(m, dtype=np.int32)
        worst, tcpu = 0.0, 0.0
        for name, t in (("TT", T), ("EE", E)):
            cl = np.zeros(n_l)
            cl_m = np.zeros((n_l, len(m)))
            t1 = time.perf_counter()
            f(k.size, k.ctypes.data_as(dp), pk.ctypes.data_as(dp), po.ctypes.data_as(lp),
              k3.ctypes.data_as(dp), n_l, len(m), mm.ctypes.data_as(ip), t.ctypes.data_as(dp),
              t.ctypes.data_as(dp), used.ctypes.data_as(ip), 1, cl.ctypes.data_as(dp),
              cl_m.ctypes.data_as(dp))
            tcpu += time.perf_counter() - t1
            g = res[name][0]
            fin = np.isfinite(cl_m) & np.isfinite(g)
            nonfin = int(np.count_nonzero(~np.isfinite(cl_m)))
            if nonfin or not np.array_equal(np.isfinite(g), np.isfinite(cl_m)):
                raise RuntimeError(f"cl2 {name}: {nonfin} non-finite C_l values")
            worst = max(worst, float(np.max(np.abs(g[fin] - cl_m[fin])) / np.max(np.abs(cl_m))))
        # |GPU - CPU| over the largest |C_l^m| of the spectrum (sum orders differ)
        line["max_rel_err_vs_cpu"] = worst
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        line["cpu_baseline"] = {"value": terms / tcpu, "unit": "terms/s", "cores": threads,
                                "kind": "port",
                                "sample": f"the whole workload ({tcpu:.2f} s, OpenMP dynamic "
                                          "over l, sequential sums)"}
    print(json.dumps(line), flush=True)
    return 0


def k3int_bench(args, cfg):
    """The k3 integral of the transfers (bispectra2.c:1182-1410) on the c2 transfers."""
    lib = L.load()
    t0 = time.time()
    scen = L.Scenario(cfg)
    scen.transfer2_init()  # the LOS stage on the GPU gives the transfers
    flat = scen.gather()
    sizes = scen.k3_sizes()[:, 0].astype(np.int64)
    pair_off = np.concatenate([[0], np.cumsum(sizes)])
    n_k3 = int(pair_off[-1])
    k3 = np.empty(n_k3)
    st = scen.s
    q = 0
    for i1 in range(cfg.k_size):
        for i2 in range(i1 + 1):
            buf = np.zeros(int(sizes[q]) + 4)
            last = C.c_int()
            rc = lib.transfer2_get_k3_list(st["ppr"], st["ppr2"], st["ppt2"], st["pbs"],
                                           st["pbs2"], st["ptr2"], i1, i2,
                                           buf.ctypes.data_as(C.POINTER(C.c_double)),
                                           C.byref(last))
            if rc != 0:
                raise RuntimeError("transfer2_get_k3_list failed: " + scen.error_message())
            k3[pair_off[q]:pair_off[q + 1]] = buf[:sizes[q]]
            q += 1
    n_l = cfg.l_size
    T = np.ascontiguousarray(flat[: n_l * n_k3].reshape(n_l, n_k3))  # T field, m = 0 rows
    ls = [int(v) for v in scen.grids()[0][:n_l]]
    scen.close()
    r = np.linspace(13600.0, 14400.0, 200)
    step = cfg.xx_step
    nx = int(k3.max() * r.max() / step) + 3
    lmax = max(ls)
    jl = np.empty((nx, lmax + 1))
    row = np.empty(lmax + 1)
    for ix in range(nx):
        lib.los_sph_bessel_array(lmax, ix * step, row.ctypes.data_as(C.POINTER(C.c_double)))
        jl[ix] = row
    j_off, vals, x_min = [0], [], []
    for l in ls:
        col = jl[:, l]
        i0 = 0 if l == 0 else max(int(np.argmax(np.abs(col) >= 1e-12)) - 1, 0)
        vals.append(col[i0:])
        x_min.append(i0 * step)
        j_off.append(j_off[-1] + nx - i0)
    j = np.concatenate(vals)
    print(f"# k3int inputs ({n_l} rows, {len(sizes)} pairs, {n_k3} k3, {r.size} r) in "
          f"{time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    _copy_probe(args)
    h = L.K3Integral(pair_off, k3, T, np.arange(n_l), j_off, j, x_min, step, r, 1.0)
    for _ in range(max(args.warmup, 1)):
        h.run(1)
    ms = h.run(args.steps)
    got = h.result()
    cnt = h.counts()
    h.close()
    n_val = got.size
    flop = 16.0 * cnt["live_terms"]
    achieved = cnt["alg_bytes"] / (ms * 1e-3) / 1e9
    line = {
        "metric": "k3-integral values I_l3(r,k1,k2)/s", "value": n_val / (ms * 1e-3),
        "unit": "values/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (c2 transfers from the LOS stage on the GPU)",
        "config": {"workload": f"k3int: {DESCR['k3int']}", "values": n_val,
                   "live_terms": cnt["live_terms"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("k3int", "los_k3int_kernel"),
                     "kernel": "los_k3int_kernel", "kernel_ms": ms,
                     "alg_bytes": cnt["alg_bytes"],
                     "valu": {"achieved": flop / (ms * 1e-3) / 1e12, "peak": 78.6,
                              "unit": "TFLOP/s", "frac": flop / (ms * 1e-3) / 1e12 / 78.6,
                              "flop_alg": flop}},
    }
    if not args.no_cpu:
        o = C.CDLL(os.path.join(ROOT, "oracle", "liblos_oracle.so"))
        f = o.los_ref_k3_integral
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        f.argtypes = [C.c_int, lp, dp, C.c_int, dp, ip, lp, dp, dp, C.c_double, C.c_int, dp,
                      C.c_int, C.c_double, dp]
        po = np.ascontiguousarray(pair_off, dtype=np.int64)
        jo = np.ascontiguousarray(j_off, dtype=np.int64)
        xm = np.ascontiguousarray(x_min)
        done, tcpu = 0, 0.0
        while done < n_l and tcpu < args.cpu_seconds:
            il = np.array([done], dtype=np.int32)
            out = np.zeros((1, r.size, len(sizes)))
            t1 = time.perf_counter()
            f(len(sizes), po.ctypes.data_as(lp), k3.ctypes.data_as(dp), 1,
              T[done].ctypes.data_as(dp), il.ctypes.data_as(ip), jo.ctypes.data_as(lp),
              j.ctypes.data_as(dp), xm.ctypes.data_as(dp), step, r.size, r.ctypes.data_as(dp),
              0, 1.0, out.ctypes.data_as(dp))
            tcpu += time.perf_counter() - t1
            if not np.array_equal(out[0], got[done]):
                raise RuntimeError(f"k3int: GPU row {done} differs from the CPU code")
            done += 1
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        line["cpu_baseline"] = {"value": done * r.size * len(sizes) / tcpu, "unit": "values/s",
                                "cores": threads, "kind": "port",
                                "sample": f"{done} of {n_l} l3 rows ({tcpu:.1f} s, OpenMP "
                                          "dynamic over pairs); each equal to the GPU's"}
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--form", default="fp64", choices=sorted(FORMS),
                    help="form reported as the headline value")
    ap.add_argument("--also", default="fp32c,fp64fast",
                    help="other forms timed on the same resident inputs and checked "
                         "against the fp64 form ('' for none)")
    ap.add_argument("--graph", choices=["auto", "on", "off"], default="auto",
                    help="replay each step as one captured hipGraph (auto: batched workloads)")
    ap.add_argument("--cpu-seconds", type=float, default=24.0,
                    help="CPU baseline: about this many seconds of oracle passes in all")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip timing one transfer2_init on host arrays")
    ap.add_argument("--no-jbuild", action="store_true",
                    help="skip timing the J-table builds (host and GPU)")
    ap.add_argument("--errors-out", default="", help="write per-(l,m) errors as JSON here")
    ap.add_argument("--copy-probe-mib", type=float, default=0.0,
                    help="run a plain device copy of this size once (HBM counter calibration)")
    ap.add_argument("--serve-budget", type=float, default=8.0,
                    help="c5: HBM budget of the admitted jobs, GiB")
    ap.add_argument("--serve-seconds", type=float, default=15.0,
                    help="c5: arrival span per rate, seconds")
    ap.add_argument("--serve-jobs", type=int, default=1000,
                    help="c5: at most this many jobs per rate (p99 rests on 1% of them)")
    args = ap.parse_args(argv)
    # CPU-baseline protocol: threads bound close, one per core (read by the OpenMP runtime
    # when it starts, which is when the first library that uses it loads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    if args.gpus != 1:
        print(f"bench.py: this build runs the LOS stage on one GPU; --gpus {args.gpus} is not "
              "supported (the (k1,k2) pairs would shard across GPUs with no communication)",
              file=sys.stderr)
        return 2
    graph = args.graph == "on" or (args.graph == "auto" and args.workload in GRAPH)

    cfg = L.Config(**WORKLOADS[args.workload])
    if args.workload == "c5":
        return serve_bench(args, cfg)
    if args.workload == "k3int":
        return k3int_bench(args, cfg)
    if args.workload == "cl2":
        return cl2_bench(args, cfg)
    t0 = time.time()
    scen = L.Scenario(cfg)
    print(f"# inputs built in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    jtab = None
    if not args.no_jbuild:
        # the workload's J tables built again on the GPU (twice: the first call includes the
        # device's start-up) and on the host, which leaves the host-built tables in place
        scen.rebuild_J(gpu=True)
        t_gpu = scen.rebuild_J(gpu=True)
        t_host = scen.rebuild_J(gpu=False)
        jtab = {"gpu_s": t_gpu, "host_s": t_host,
                "host_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                "x_points": int(cfg.xx_max / cfg.xx_step + 1), "l_max": cfg.l_max}
        print(f"# J tables: GPU build {t_gpu:.3f} s, host build {t_host:.3f} s", file=sys.stderr,
              flush=True)
    scen.prepare()
    all_pairs = [(i1, i2) for i1 in range(cfg.k_size) for i2 in range(i1 + 1)]
    pairs = select_pairs(cfg, PAIRS[args.workload]) if args.workload in PAIRS else all_pairs
    eng = L.Engine(0)
    t0 = time.time()
    if args.workload in PAIRS:
        eng.load_pairs(scen, pairs)
    else:
        eng.load(scen)
    print(f"# tables uploaded in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    n_el = eng.n_elements()
    flop_live, flop_dense = eng.flops()
    probe = None
    if args.copy_probe_mib > 0:
        nb = args.copy_probe_mib * 2.0 ** 20
        pms = eng.copy_probe(nb)
        probe = {"bytes_read": nb, "bytes_written": nb, "ms": pms,
                 "gbs": 2 * nb / (pms * 1e-3) / 1e9 if pms > 0 else None}
        print(f"# copy probe {args.copy_probe_mib} MiB: {pms:.3f} ms", file=sys.stderr, flush=True)
    cols, sizes = loaded_columns(scen, pairs)

    def run_form(form):
        eng.set_form(FORMS[form])
        eng.autotune(2)
        kern = []
        for _ in range(max(1, args.warmup)):
            eng.run()
            kern.append(eng.last_kernel_ms())
        los_ms = min(k[0] for k in kern)
        prep_ms = min(k[1] for k in kern)
        if graph:
            eng.capture()
        eng.run()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng.run()
        wall = time.perf_counter() - t0
        eng.store(scen)
        tt_out = scen.gather()
        tt2 = tt_out.size // int(scen.k3_sizes()[:, 0].sum())
        out = tt_out.reshape(tt2, -1)[:, cols].ravel()
        ms_step = 1e3 * wall / args.steps
        print(f"# {form}: {ms_step:.2f} ms/step, LOS kernel {los_ms:.2f} ms, prep {prep_ms:.2f} ms",
              file=sys.stderr, flush=True)
        res = {"ms_per_step": ms_step, "value": n_el / (ms_step * 1e-3), "kernel_ms": los_ms,
               "prep_ms": prep_ms, "graph": graph, "alg_bytes": eng.alg_bytes()}
        if los_ms > 0:
            res["hbm_gbs"] = res["alg_bytes"] / (los_ms * 1e-3) / 1e9
            res["valu_tflops"] = flop_live / (los_ms * 1e-3) / 1e12
        kind, nk, ch, pers = eng.get_shape(FORMS[form])
        res["kernel"] = kernel_name(form, (kind, nk, ch, pers), max(cfg.l_max_los_t, cfg.l_max_los_p) + 1)
        res["tile_shape"] = ({"kernel": "sweep", "nk": nk, "persistent": pers,
                              "j_pairs": bool(form != "fp32c" or ch == 1)} if kind == 1 else
                             {"kernel": "k-lane", "l_per_lane": ch} if kind == 2 else
                             {"kernel": "workgroup per k3"} if form == "fp64" else
                             {"kernel": "stream", "nk": nk, "ch": ch, "persistent": pers})
        return res, out, tt2, tt_out

    forms = [args.form] + [f for f in args.also.split(",") if f and f != args.form]
    if "fp64" not in forms:
        forms.append("fp64")
    results, outs, full = {}, {}, {}
    tt2 = 0
    for f in forms:
        results[f], outs[f], tt2, full[f] = run_form(f)
    eng.close()
    offs = scen.pair_offsets()
    total_k3 = sum(n for _, n in offs.values())

    def rows_of(flat, p):
        o, n = offs[p]
        return flat.reshape(-1, total_k3)[:, o:o + n]

    per_lm = {}
    for f in forms:
        if f == "fp64":
            continue
        field, row = relative_errors(outs[f], outs["fp64"], sizes, tt2)
        results[f]["speedup_vs_fp64"] = results[f]["value"] / results["fp64"]["value"]
        results[f]["max_rel_err_vs_fp64"] = float(field.max())
        results[f]["rms_rel_err_vs_fp64"] = float(np.sqrt(np.mean(field ** 2)))
        results[f]["max_row_rel_err_vs_fp64"] = float(row.max())
        results[f]["rms_row_rel_err_vs_fp64"] = float(np.sqrt(np.mean(row ** 2)))
        results[f]["worst_tt"] = int(np.argmax(field.max(axis=1)))
        per_lm[f] = {"max": field.max(axis=1).tolist(),
                     "rms": np.sqrt(np.mean(field ** 2, axis=1)).tolist(),
                     "row_max": row.max(axis=1).tolist()}
    scen.close()

    cpu = parity = None
    per_lm_oracle = {}
    if not args.no_cpu:
        cpu, oref = cpu_baseline(cfg, pairs, args.cpu_seconds, tt2,
                                 {p: offs[p][1] for p in pairs})
        sample = sorted(oref)
        r = np.concatenate([oref[p] for p in sample], axis=1)
        g64 = np.concatenate([rows_of(full["fp64"], p) for p in sample], axis=1)
        parity = {"pairs": len(sample), "compared": int(r.size),
                  "fp64_mismatches": int(np.count_nonzero(g64 != r)), "forms": {}}
        fscale = np.abs(r).max(axis=1, keepdims=True)
        for f in forms:
            if f == "fp64":
                continue
            g = np.concatenate([rows_of(full[f], p) for p in sample], axis=1)
            d = np.abs(g - r)
            fe = np.divide(d, fscale, out=np.zeros_like(d), where=fscale > 0)
            re_ = []
            for p in sample:
                rr = oref[p]
                sc = np.abs(rr).max(axis=1, keepdims=True)
                dd = np.abs(rows_of(full[f], p) - rr)
                re_.append(np.divide(dd, sc, out=np.zeros_like(dd), where=sc > 0))
            re_ = np.concatenate(re_, axis=1)
            bar = FIELD_BAR[f]
            parity["forms"][f] = {"max_rel_err": float(fe.max()),
                                  "rms_rel_err": float(np.sqrt(np.mean(fe ** 2))),
                                  "max_rel_err_worst_tt": int(np.argmax(fe.max(axis=1))),
                                  "max_row_rel_err": float(re_.max()),
                                  "field_bar": bar, "within_field_bar": bool(fe.max() < bar)}
            if not fe.max() < bar:
                print(f"# WARNING: {f} exceeds its bar against the oracle: {fe.max():.3e} >= "
                      f"{bar:.0e}", file=sys.stderr, flush=True)
            per_lm_oracle[f] = {"max": fe.max(axis=1).tolist(),
                                "rms": np.sqrt(np.mean(fe ** 2, axis=1)).tolist(),
                                "row_max": re_.max(axis=1).tolist()}
        print(f"# parity vs the oracle on {len(sample)} pairs: {parity['fp64_mismatches']} of "
              f"{r.size} fp64 values differ; " +
              "; ".join(f"{f} max {v['max_rel_err']:.2e} rms {v['rms_rel_err']:.2e}"
                        for f, v in parity["forms"].items()), file=sys.stderr, flush=True)
    if args.errors_out:
        with open(args.errors_out, "w") as fh:
            json.dump({"workload": args.workload, "tt2_size": tt2,
                       "lm_of_tt": "tt = field * n_lm + lm, fields T, E, B, (l,m) l-major",
                       "vs_fp64_gpu": per_lm, "vs_oracle": per_lm_oracle,
                       "oracle_pairs": parity["pairs"] if parity else 0}, fh)

    head = results[args.form]
    alg_bytes = head["alg_bytes"]
    achieved = head.get("hbm_gbs", 0.0)
    traffic = measured_traffic(args.workload, head["kernel"])
    valu = {"achieved": head.get("valu_tflops", 0.0), "peak": VALU_PEAK_TFLOPS[args.form],
            "unit": "TFLOP/s",
            "frac": head.get("valu_tflops", 0.0) / VALU_PEAK_TFLOPS[args.form],
            "flop_alg": flop_live, "flop_dense": flop_dense}
    hbm = {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS}
    # the bound is whichever roofline the kernel sits closer to
    top = valu if valu["frac"] > hbm["frac"] else hbm
    line = {
        "metric": METRIC, "value": head["value"], "unit": "elements/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.form != "fp32c" else "f32-compensated",
        "data": "synthetic (seeded source model, FP64 J tables built on the host)",
        "config": {"workload": f"{args.workload}: {DESCR[args.workload]}", "form": args.form,
                   "elements": n_el, "pairs": len(pairs)},
        "roofline": {"bound": "valu" if top is valu else "hbm", "achieved": top["achieved"],
                     "peak": top["peak"], "unit": top["unit"], "frac": top["frac"],
                     "traffic": traffic,
                     "traffic_over_alg_bytes": traffic / alg_bytes if traffic else None,
                     "kernel": head["kernel"], "kernel_ms": head["kernel_ms"],
                     "alg_bytes": alg_bytes, "hbm": hbm, "valu": valu},
        "forms": results,
    }
    if probe:
        line["copy_probe"] = probe
    if jtab:
        line["j_tables"] = jtab
    if cpu:
        line["cpu_baseline"] = cpu
        line["parity"] = parity
    if args.workload not in PAIRS and not args.no_e2e:
        # the drop-in end to end: transfer2_init on host-resident SONG structures (sources
        # uploaded over PCIe, J tables laid out on the device, the run, the scatter back into
        # ptr2->transfer); not the headline, which has its inputs resident in HBM
        e2e_scen = L.Scenario(cfg)
        t0 = time.perf_counter()
        e2e_scen.transfer2_init()
        sec = time.perf_counter() - t0
        e2e_scen.close()
        line["end_to_end"] = {"seconds": sec, "elements_per_s": n_el / sec,
                              "what": "transfer2_init on host arrays: upload, device tables, "
                                      "kernel shapes tuned, run, scatter to ptr2->transfer"}
        print(f"# transfer2_init end to end: {sec:.2f} s", file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
