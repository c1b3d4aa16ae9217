"""Benchmark of the LOS transfer stage on one MI355X.

Workload c2 (the plain FP64 GPU drop-in): T+E(+B) with m = 0,1,2 and
sources L = 0..6; 32x32 log (k1,k2) grid in [1e-5, 0.1]/Mpc (528 pairs k2 <= k1), 60
linear coarse k3 per pair, fine k3 step 1.5e-4/Mpc; 600 tau log-spaced in [200, 14000] Mpc,
tau0 = 14200 Mpc; 96 l in [2, 2000]; J tables on x in [0, 15000] step 0.3. Seeded synthetic
sources (A ~ N(0,1) x (k3/0.01)^-1.5 x [Gaussian at 280 Mpc, sigma 20 Mpc + 10% late ramp]).

A step is one pass of the LOS kernels over every pair of that workload, with the source and
projection tables already resident in HBM (k3 spline coefficients + LOS integral). The
JSON line reports elements/s (one element = one ptr2->transfer[tt][k1][k2][k3] value), the
HBM roofline of the dominant kernel (algorithmic bytes / HIP-event duration) and the CPU
oracle timed with OpenMP on a bounded sample of pairs of the same workload.
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

WORKLOADS = {
    "c2": dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=6, l_max_los_p=6, k_size=32,
               k_min=1e-5, k_max=0.1, k3_size=60, tau_size=600, tau_min=200.0, tau_max=14000.0,
               tau0=14200.0, l_size=96, l_max=2000, dk_fine=1.5e-4, xx_max=15000.0,
               xx_step=0.3),
    # small workload for quick checks
    "tiny": dict(has_T=1, has_E=1, m_list=[0, 1, 2], l_max_los_t=4, l_max_los_p=4, k_size=6,
                 k3_size=16, tau_size=128, l_size=12, l_max=400, dk_fine=1e-3, xx_max=1500.0),
}

DESCR = {
    "c2": ("plain FP64 LOS, T+E+B m=0,1,2 L=0..6, 32x32 log (k1,k2) in [1e-5,0.1], 60 coarse "
           "k3/pair, dk=1.5e-4, 600 tau, 96 l in [2,2000]"),
    "tiny": "small check workload",
}

FORMS = {"fp64": L.FORM_FP64, "fp64fast": L.FORM_FP64_FAST, "fp32c": L.FORM_FP32C}


def _err():
    return C.create_string_buffer(L.ERRSZ)


def cpu_baseline(cfg: L.Config, target_s: float):
    """Time the CPU oracle on a bounded sample of pairs of the same workload."""
    opath = os.path.join(ROOT, "oracle", "liblos_oracle.so")
    o = C.CDLL(opath)
    o.los_ref_run_pairs.argtypes = [C.c_void_p] * 8 + [C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                       C.c_int, C.POINTER(C.c_long)]
    o.los_ref_run_pairs.restype = C.c_int
    o.los_ref_sizeof_transfers2.restype = C.c_long
    o.los_ref_transfer2_free.argtypes = [C.c_void_p] * 3
    scen = L.Scenario(cfg)
    s = scen.s
    nk = cfg.k_size
    pairs = [(i1, i2) for i1 in range(nk) for i2 in range(i1 + 1)]
    # spread the sample over the whole pair list; grow it until target_s is used
    total_el, total_t, done, stride = 0, 0.0, 0, 24
    order = pairs[::stride] + [p for q, p in enumerate(pairs) if q % stride]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    batch = 4
    while done < len(order) and total_t < target_s:
        chunk = order[done:done + batch]
        done += len(chunk)
        tr2 = C.create_string_buffer(o.los_ref_sizeof_transfers2())
        # copy the stage settings the oracle reads from the scenario's transfers2
        C.memmove(tr2, s["ptr2"].value, o.los_ref_sizeof_transfers2())
        k1 = (C.c_int * len(chunk))(*[p[0] for p in chunk])
        k2 = (C.c_int * len(chunk))(*[p[1] for p in chunk])
        n = C.c_long(0)
        t0 = time.perf_counter()
        rc = o.los_ref_run_pairs(s["ppr"], s["ppr2"], s["pba"], s["pth"], s["ppt2"], s["pbs"],
                                 s["pbs2"], C.cast(tr2, C.c_void_p), k1, k2, len(chunk), C.byref(n))
        dt = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError("oracle sample failed")
        o.los_ref_transfer2_free(s["ppr2"], s["ppt2"], C.cast(tr2, C.c_void_p))
        total_el += n.value
        total_t += dt
    scen.close()
    return {"value": total_el / total_t, "unit": "elements/s", "cores": threads, "kind": "port",
            "sample": f"{done} of {len(pairs)} (k1,k2) pairs of the same workload "
                      f"({total_el} elements, {total_t:.1f} s, OpenMP static over k3 per pair)"}


def row_normalised_errors(got, ref, sizes, tt2_size):
    """|got - ref| / max|ref| over the k3 row of the same (tt, pair) (the transfer
    functions oscillate in k3, so single zero crossings are not used as scales)."""
    err = np.zeros_like(ref)
    o = 0
    for _tt in range(tt2_size):
        for n in sizes:
            r = ref[o:o + n]
            scale = np.max(np.abs(r))
            if scale > 0:
                err[o:o + n] = np.abs(got[o:o + n] - r) / scale
            o += n
    return err


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
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--errors-out", default="", help="write per-(l,m) errors as JSON here")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build runs the LOS stage on one GPU; --gpus {args.gpus} is not "
              "supported (the (k1,k2) pairs would shard across GPUs with no communication)",
              file=sys.stderr)
        return 2

    lib = L.load()
    cfg = L.Config(**WORKLOADS[args.workload])
    t0 = time.time()
    scen = L.Scenario(cfg)
    print(f"# inputs built in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    s = scen.s
    err = _err()
    if lib.los_transfer2_prepare(*scen.structs()) != L.SUCCESS:
        raise L.LosError("los_transfer2_prepare: " + scen.error_message())
    scen._inited = True
    eng = C.c_void_p()
    L.check(lib.los_engine_create(0, C.byref(eng), err), err, "engine_create")
    t0 = time.time()
    L.check(lib.los_engine_load(eng, s["ppr"], s["ppr2"], s["ppt2"], s["pbs"], s["pbs2"],
                                s["ptr2"], 0, cfg.k_size, err), err, "engine_load")
    print(f"# tables uploaded in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    n_el = lib.los_engine_n_elements(eng)
    alg_bytes = lib.los_engine_alg_bytes(eng)

    def run_form(form):
        L.check(lib.los_engine_set_form(eng, FORMS[form], err), err, "set_form")
        ms = C.c_float()
        for _ in range(args.warmup):
            L.check(lib.los_engine_run(eng, C.byref(ms), err), err, "engine_run")
        los_ms, spl_ms = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            L.check(lib.los_engine_run(eng, C.byref(ms), err), err, "engine_run")
            a, b = C.c_float(), C.c_float()
            lib.los_engine_last_kernel_ms(eng, C.byref(a), C.byref(b))
            los_ms.append(a.value)
            spl_ms.append(b.value)
        wall = time.perf_counter() - t0
        L.check(lib.los_engine_store(eng, s["ptr2"], err), err, "engine_store")
        out = scen.gather()
        ms_step = 1e3 * wall / args.steps
        los_avg = sum(los_ms) / len(los_ms)
        print(f"# {form}: {ms_step:.2f} ms/step, LOS kernel {los_avg:.2f} ms", file=sys.stderr,
              flush=True)
        return {"ms_per_step": ms_step, "value": n_el / (ms_step * 1e-3), "kernel_ms": los_avg,
                "prep_ms": sum(spl_ms) / len(spl_ms)}, out

    forms = [args.form] + [f for f in args.also.split(",") if f and f != args.form]
    if "fp64" not in forms:
        forms.append("fp64")
    results, outs = {}, {}
    for f in forms:
        results[f], outs[f] = run_form(f)
    lib.los_engine_destroy(eng)

    sizes = [int(r[0]) for r in scen.k3_sizes()]
    tt2 = len(outs["fp64"]) // sum(sizes)
    lm_count = tt2 // (2 * cfg.has_E * (1 if max(cfg.m_list) > 0 else 0) + cfg.has_E + cfg.has_T)
    per_lm = {}
    for f in forms:
        if f == "fp64":
            continue
        e = row_normalised_errors(outs[f], outs["fp64"], sizes, tt2)
        results[f]["max_rel_err_vs_fp64"] = float(e.max())
        results[f]["rms_rel_err_vs_fp64"] = float(np.sqrt(np.mean(e ** 2)))
        per_tt = e.reshape(tt2, -1)
        worst = int(np.argmax(per_tt.max(axis=1)))
        results[f]["worst_tt"] = worst
        per_lm[f] = {"max": per_tt.max(axis=1).tolist(),
                     "rms": np.sqrt(np.mean(per_tt ** 2, axis=1)).tolist()}
    scen.close()
    if args.errors_out:
        with open(args.errors_out, "w") as fh:
            json.dump({"per_tt": per_lm, "tt2_size": tt2, "n_lm": lm_count}, fh)

    head = results[args.form]
    achieved = alg_bytes / (head["kernel_ms"] * 1e-3) / 1e9
    kernel_name = {"fp64": "los_integrate_exact", "fp64fast": "los_integrate_stream<StreamFP64>",
                   "fp32c": "los_integrate_stream<StreamFP32C>"}[args.form]
    line = {
        "metric": METRIC, "value": head["value"], "unit": "elements/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.form != "fp32c" else "f32-compensated",
        "data": "synthetic (seeded source model, FP64 J tables built on the host)",
        "config": {"workload": f"{args.workload}: {DESCR[args.workload]}", "form": args.form,
                   "elements": n_el, "pairs": cfg.k_size * (cfg.k_size + 1) // 2},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel_name,
                     "kernel_ms": head["kernel_ms"], "prep_ms": head["prep_ms"],
                     "alg_bytes": alg_bytes},
        "forms": results,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
