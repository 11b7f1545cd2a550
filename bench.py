#!/usr/bin/env python3
"""Full-chain benchmark: 10 keV primaries/s through transport + prechemistry, IRT
chemistry to 1 us and SSB/DSB scoring on the 21x21 fiber grid, on one MI355X.

Workload (one step): 4096 primary electrons at 10 keV, reference geometry and seeding
(isotropic start at the origin, seed = floor(E)), FP32 transport with fixed-point dose,
canonical radical ordering, FP64 IRT with the 7 reference checkpoints, direct-event SSB
scoring + DSB clustering with the reference LCG.  All intermediate data stays on the GPU;
per step only per-primary results (128 KB) and the tallies come back.

Prints one JSON line (see the "metric" of the reference's headline benchmark).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "webgpu-dna_amd"))

CK = [0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0]
N_PRIM, E_EV = 4096, 10000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def chain_step(ctx, p, dna, seed):
    from dna_amd.abi import DNA_COPY_RESULTS
    t = ctx.transport(p, copy=DNA_COPY_RESULTS)
    n_therm = int((t["results"]["esc"] == 0).sum())
    c = ctx.irt(None, t["rad_n"], n_therm, E_EV, CK, seed=seed)
    s = ctx.score(dna, None, t["rad_n"], lcg_state=0x12345678)
    return t, c, s, n_therm


def cpu_baseline(sample_prims, threads):
    """The CPU restatement (oracle/) timed on this host: the same chain on a bounded
    sample of the same workload (primaries 0..sample-1 of the 10 keV run)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from dna_amd import DNATarget, default_params
    dna = DNATarget()
    p = default_params(E_EV, N_PRIM, dna=dna)
    O.lib()
    t0 = time.perf_counter()
    t = O.transport(p, first=0, count=sample_prims, n_threads=threads)
    t1 = time.perf_counter()
    n_therm = int((t["results"]["esc"] == 0).sum())
    O.irt(t["rad"], n_therm, E_EV, CK, seed=1, n_threads=threads)
    t2 = time.perf_counter()
    O.score(dna, t["rad"], lcg_state=0x12345678)
    t3 = time.perf_counter()
    dt = t3 - t0
    return {"value": sample_prims / dt, "unit": "10 keV primaries/s", "cores": threads,
            "kind": "port",
            "sample": f"{sample_prims} of the {N_PRIM} x 10 keV primaries, full chain "
                      f"(transport {t1 - t0:.2f} s, IRT {t2 - t1:.2f} s, scoring {t3 - t2:.2f} s)",
            "seconds": dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; multi-GPU scaling is not part of it "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from dna_amd import Context, DNATarget, default_params
    ctx = Context(0)
    dna = DNATarget()
    p = default_params(E_EV, N_PRIM, dna=dna)
    for w in range(args.warmup):
        chain_step(ctx, p, dna, seed=1)
        print(f"warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    times, k_irt, k_sec, k_prim, k_score = [], [], [], [], []
    last = None
    t_all0 = time.perf_counter()
    for s in range(args.steps):
        t0 = time.perf_counter()
        last = chain_step(ctx, p, dna, seed=1)
        times.append(time.perf_counter() - t0)
        t, c, sc, n_therm = last
        k_irt.append(c.t_kernel_ms)
        k_sec.append(t["t_sec_ms"])
        k_prim.append(t["t_prim_ms"])
        k_score.append(sc["t_ms"])
        print(f"step {s + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    total = time.perf_counter() - t_all0
    ms_per_step = 1e3 * total / args.steps
    value = N_PRIM * args.steps / total
    t, c, sc, n_therm = last
    # roofline of the dominant kernel (IRT, one workgroup per track): algorithmic FP64
    # work = candidate pair evaluations x FP64 operations per evaluation
    irt_ms = sum(k_irt) / len(k_irt)
    flops = c.algo_flops
    achieved = flops / (irt_ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "kernel": "k_irt", "achieved": round(achieved, 3),
            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
            "traffic": None,
            "note": "FP64 VALU kernel; MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s",
            "avg_launch_ms": round(irt_ms, 4)}
    out = {
        "metric": "10 keV primaries/s, full chain; surrogate vs IRT speedup at <=3% G error",
        "value": round(value, 2), "unit": "10 keV primaries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (reference geometry and seeding, no external data)",
        "config": {"workload": "full reference chain: 4096 primaries at 10 keV, transport + prechemistry, "
                               "IRT to 1 us (FP64), SSB/DSB on the 21x21 B-DNA fiber grid",
                   "primaries": N_PRIM, "energy_eV": E_EV},
        "roofline": roof,
        "breakdown_ms": {"phase_a": round(sum(k_prim) / len(k_prim), 3),
                         "phase_b": round(sum(k_sec) / len(k_sec), 3),
                         "irt_pipeline": round(c.t_ms, 3), "irt_kernel": round(irt_ms, 3),
                         "score": round(sum(k_score) / len(k_score), 3)},
        "chain_check": {"n_therm": n_therm, "rad_n": t["rad_n"],
                        "G_1us": [round(c.G[6][s], 4) for s in range(5)],
                        "ssb_dir": sc["ssb_count"], "dsb": sc["dsb"]},
    }
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, threads)
        out["gpu_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
