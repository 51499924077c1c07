#!/usr/bin/env python3
"""Benchmark: candidates swept per second at a bit-exact survivor stream.

Headline workload (one step = one seeded window): the plain GPU engine on
the Dickson ladder sweep, n = 10, every prime below 2^22 sieved, a window of
2^32 consecutive j near k = 2310 j = 1e15, and every survivor classified on
the device (its full seven-base ladder depth, cap 18).  value = j swept per
second over the timed steps, inputs resident on the device (the window is
two integers; nothing is copied in per step).

Beside it, in the same JSON line:
  roofline      the sieve kernel the window runs (its lines of 695 words
                take the short-line sieve, k_sieve_short), timed with HIP
                events on the engine's stream
  cpu_baseline  the oracle's restatement of the reference CPU sieve
                (dickson-ladders/ladder_search.py:92-112) on a bounded sample
                of the same window on this host's cores, plus the
                classification of its survivors, with the sample's survivor
                stream checked equal to the GPU's
  batched       256 mixed windows of the three families (lengths 2^20..2^30,
                n 8..14), one persistent batched launch vs the plain engine
                window by window (two in flight), with the two streams
                checked equal
  serving       Poisson arrivals of requests of the three families (spans
                2^18..2^32, n 8..14, Dickson k to 1e22), continuous batching
                over 4 slots: p50/p99 completion latency and candidates/s

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import math
import os
import pathlib
import random
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "first-knowns_amd"))

METRIC = ("candidates swept/sec at bit-exact survivor stream; "
          "batched vs plain GPU engine")
LDS_PEAK = 256 * 128 * 2.4e9          # B/s: 256 CUs x 128 B/clk x 2.4 GHz
HBM_PEAK = 8.0e12                      # B/s
MASK64 = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


# ------------------------------------ headline: the Dickson ladder sweep

SWEEP_N, SWEEP_Q2, SWEEP_SPAN, SWEEP_CAP = 10, 1 << 22, 1 << 32, 18


def sweep_window(seed):
    j_lo = 432_900_432_900 + splitmix64(seed) % 10**11
    return j_lo, j_lo + SWEEP_SPAN


def run_plain(steps, warmup):
    """The timed loop: `steps` windows swept in order on the plain engine,
    two in flight (Engine.sweep_many: each window one engine call on its
    own stream and captured graph, the next enqueued before the last is
    read back, as the reference's launch loop does).  device_ms is one
    window's device time alone: two more windows swept one at a time
    (graph replays) after the timed loop.  The stage split comes from
    three more windows swept eagerly with HIP events between the stages."""
    import fk_native as N
    import ladder_gpu
    eng = ladder_gpu.GpuEngine(SWEEP_N, q2=SWEEP_Q2)
    for _ in eng._eng.sweep_many([sweep_window(1000 + s) for s in range(warmup)],
                                 N.CLASSIFY_FULL, SWEEP_CAP):
        pass
    wins = [sweep_window(s) for s in range(steps)]
    t0 = time.perf_counter()
    out = list(eng._eng.sweep_many(wins, N.CLASSIFY_FULL, SWEEP_CAP))
    wall = time.perf_counter() - t0
    for s in range(3):                            # alone, graph replays
        eng._eng.sweep(*sweep_window(2000 + s), N.CLASSIFY_FULL, SWEEP_CAP)
    device_ms = eng._eng.last_timing()[0][4]
    surv = sum(int(r.size) for r in out)
    deepest = (0, None)
    for recs in out:
        if recs.size and int(recs["run"].max()) > deepest[0]:
            i = int(recs["run"].argmax())
            deepest = (int(recs["run"][i]), int(recs["v0"][i]))
    # per-stage split and sieve candidates, eager with stage events (not
    # timed), on the first windows of the timed loop
    eng._eng.stage_timing(True)
    stage = [0.0] * 5
    nst = min(3, steps)
    cand = 0
    for s in range(nst):
        lo, hi = sweep_window(s)
        eng._eng.sweep(lo, hi, N.CLASSIFY_FULL, SWEEP_CAP)
        ms, _, c = eng._eng.last_timing()
        cand += c
        for i in range(5):
            stage[i] += ms[i] / nst
    eng._eng.stage_timing(False)
    return eng, wall, surv, stage, cand / nst, deepest, device_ms


def cpu_baseline(eng, sample_log2, threads):
    import fk_native as N
    sys.path.insert(0, str(ROOT / "oracle"))
    import fk_oracle as O
    lo, _ = sweep_window(0)
    hi = lo + (1 << sample_log2)
    sieve = O.DicksonSieve(SWEEP_N, SWEEP_Q2)            # tables outside the clock
    t0 = time.perf_counter()
    js = sieve.survivors(lo, hi, nthreads=threads)
    runs = [O.dickson_run(j * eng.W, cap=SWEEP_CAP) for j in js]
    wall = time.perf_counter() - t0
    recs = eng._eng.sweep(lo, hi, N.CLASSIFY_FULL, SWEEP_CAP)
    same = (N.record_values(recs) == js and
            [int(r) for r in recs["run"].tolist()] == runs)
    return {"value": (hi - lo) / wall, "unit": "j-candidates/s",
            "cores": threads, "kind": "port",
            "sample": f"2^{sample_log2} j from the step-0 window, n={SWEEP_N}, "
                      f"q2=2^22, survivors classified (cap {SWEEP_CAP}); "
                      f"{len(js)} survivors, {wall:.1f} s",
            "sample_stream_equal": bool(same)}


# ------------------------------------ 256 mixed windows, one batch

FAMILY_Q2 = {"euler": 1 << 16, "dickson": 1 << 22, "ap": 1 << 24}


def batch_windows(seed=3, count=256):
    """256 windows: family uniform, length 2^(20 + 10 U) on the family's
    line, n uniform in 8..14, heights spread so the strong tests run on
    64-, 128- and 192-bit values."""
    rng = random.Random(seed)
    out = []
    for _ in range(count):
        fam = rng.choice(["euler", "dickson", "ap"])
        n = rng.randint(8, 14)
        length = int(2 ** (20 + 10 * rng.random()))
        wide = rng.random() < 0.5
        if fam == "euler":
            lo = (rng.randrange(10**12, 10**13) if not wide
                  else rng.randrange(10**20, 10**22))
        elif fam == "dickson":
            W = 30030 if n >= 12 else 2310
            k = (rng.randrange(10**14, 10**16) if not wide
                 else rng.randrange(10**21, 10**22))
            lo = k // W
        else:
            lo = (rng.randrange(10**15, 10**16) if not wide
                  else rng.randrange(10**19, 10**20))
            lo -= lo % 30
            length -= length % 30
        out.append((fam, n, lo, lo + length))
    return out


def batch_spec(fam, n):
    import fk_native as N
    if fam == "euler":
        return (N.CLASSIFY_FULL, n + 3, 0)
    if fam == "dickson":
        return (N.CLASSIFY_FULL, n + 8, 0)
    return (N.CLASSIFY_TWO_PASS, n, 6)


def run_batched():
    import fk_native as N
    import fk_families as F
    wins = batch_windows()
    engines, index = [], {}
    fam_id = {"euler": N.FAMILY_EULER, "dickson": N.FAMILY_DICKSON,
              "ap": N.FAMILY_AP}
    for fam, n, lo, hi in wins:
        if (fam, n) not in index:
            t = F.tuning(fam)
            kw = {}
            if fam == "euler":
                kw["wheel_primes"] = [2, 3, 5, 7, 11, 13]
            index[(fam, n)] = len(engines)
            engines.append(N.Engine(fam_id[fam], n, FAMILY_Q2[fam], **kw,
                                    **t))
    eo = [index[(f, n)] for f, n, _, _ in wins]
    ws = [(lo, hi) for _, _, lo, hi in wins]
    specs = [batch_spec(f, n) for f, n, _, _ in wins]
    cand = sum(hi - lo for lo, hi in ws)
    # plain: one engine call per window, two in flight (the plain engine's
    # own pipelining, as in the headline loop)
    pipe = N.Pipeline()
    jobs = [(engines[eo[i]], *ws[i], specs[i]) for i in range(len(ws))]
    for _ in pipe.run(jobs[:3]):
        pass
    t0 = time.perf_counter()
    plain = [r for _, r in pipe.run(jobs)]
    t_plain = time.perf_counter() - t0
    pipe.close()
    # batched: one submission
    b = N.Batch()
    b.submit(engines, eo, ws, specs)
    b.results()
    t0 = time.perf_counter()
    b.submit(engines, eo, ws, specs)
    recs, counts = b.results()
    t_batch = time.perf_counter() - t0
    same = True
    pos = 0
    for i, p in enumerate(plain):
        mine = recs[pos:pos + int(counts[i])]
        pos += int(counts[i])
        if (N.record_values(mine) != N.record_values(p) or
                mine["run"].tolist() != p["run"].tolist()):
            same = False
    return {"windows": len(ws), "engines": len(engines),
            "candidates": cand, "survivors": int(recs.size),
            "plain_s": t_plain, "batched_s": t_batch,
            "plain_rate": cand / t_plain, "batched_rate": cand / t_batch,
            "speedup": t_plain / t_batch, "streams_equal": same}


# ------------------------------------ serving: Poisson arrivals

def run_serving(count, rate, seed=4):
    """Poisson arrivals of range-sweep requests of all three families,
    spans log-uniform in [2^18, 2^32] line values, depths 8..14, Dickson k
    up to about 1e22 (192-bit values); continuous batching over 4 slots."""
    import fk_serve as S
    reqs = S.poisson_requests(seed, count, rate)
    srv = S.Server(slots=4, budget=64 << 30)
    # engines are built (resident tables) before the clock starts
    for r in reqs:
        srv.engine(r.family, r.n, r.q2)
    st = srv.run(reqs)
    st["rate_per_s"] = rate
    st["slots"] = 4
    st["submissions"] = len(srv.submissions)
    st["mean_requests_per_submission"] = (
        sum(srv.submissions) / len(srv.submissions) if srv.submissions else 0)
    return st


TRAFFIC_JSON = ROOT / "profiles" / "k_sieve_short_dickson_sweep_traffic.json"


def rocprof_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed PMC summary
    (None when the summary covers another kernel)."""
    if TRAFFIC_JSON.exists():
        try:
            js = json.loads(TRAFFIC_JSON.read_text())
        except (OSError, ValueError):
            return None
        if kernel in js.get("kernel", ""):
            return js.get("bytes_per_dispatch")
    return None


def sieve_kernel(info, span):
    """The sieve kernel a window of `span` line values runs on this engine
    and its algorithmic LDS bytes per wheel candidate.  Lines of at most
    short_words 64-position words take k_sieve_short: one 64-bit window of
    a kill bitmap (8 B) per word and LDS prime.  Longer lines take k_sieve:
    8 B per pattern word read per pattern prime, 8 B (read + write) per
    stride mark."""
    words = (span // info.line_step + 1 + 63) // 64
    if words <= info.short_words:
        return "k_sieve_short", 8.0 * info.n_lds_primes / 64
    return "k_sieve", (8.0 * info.n_pattern_primes / 64
                       + 8.0 * info.mark_density)


def copy_probe(mib):
    """One device-to-device copy of a known size: the calibration run for
    the FETCH_SIZE / WRITE_SIZE counters (bytes read and written are both
    mib MiB)."""
    import torch
    n = mib << 20
    x = torch.ones(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-sample-log2", type=int, default=31)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=400.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device copy of MIB MiB (counter "
                         "calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path "
              f"(windows shard across GPUs by range; see DESIGN.md). "
              f"--gpus {args.gpus} is not supported.", file=sys.stderr)
        return 2

    eng, wall, surv, stage, cand, deepest, device_ms = run_plain(args.steps,
                                                                args.warmup)
    j_total = args.steps * SWEEP_SPAN
    value = j_total / wall
    info = eng._eng.info
    kernel, per_cand = sieve_kernel(info, SWEEP_SPAN)
    # one sieve launch per window: its duration is the eager sieve stage
    launch_s = stage[1] / 1e3
    per_launch = cand                    # wheel candidates per window
    achieved = per_launch * per_cand / launch_s if launch_s > 0 else 0.0
    traffic = rocprof_traffic(kernel)
    out = {
        "metric": METRIC, "value": value, "unit": "j-candidates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "dickson-ladder sweep, plain GPU engine: "
                               "2^32 consecutive j near k=1e15 per step, "
                               "n=10, primes < 2^22 sieved, full-bases "
                               "ladder depth of every survivor; windows "
                               "in order, two in flight",
                   "n": SWEEP_N, "q2": SWEEP_Q2, "window_j": SWEEP_SPAN,
                   "classify_cap": SWEEP_CAP,
                   "wheel_lines": int(info.nlines),
                   "wheel_period": int(info.line_step),
                   "q_pat": int(info.q_pat), "q_med": int(info.q_med),
                   "tile_bits": None},
        "survivors": surv,
        "deepest_ladder": {"run": deepest[0], "j": deepest[1]},
        "device_ms_per_window_alone": device_ms,
        "stages_ms_per_step": dict(zip(("prep", "sieve", "rounds",
                                        "classify", "whole"), stage)),
        "stages_note": "per-stage HIP events on three windows swept eagerly "
                       "after the timed loop; the timed sweeps keep two "
                       "windows in flight, each replaying its slot's captured "
                       "launch sequence; device_ms_per_window_alone is one "
                       "window swept by itself (graph replay, event-timed)",
        "roofline": {"bound": "lds", "kernel": kernel,
                     "achieved": achieved / 1e9, "peak": LDS_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / LDS_PEAK,
                     "traffic": traffic,
                     "algorithmic_bytes_per_candidate": per_cand,
                     "candidates_per_launch": per_launch,
                     "hbm_GBps": (traffic / launch_s / 1e9
                                  if traffic and launch_s > 0 else None),
                     "hbm_frac": (traffic / launch_s / HBM_PEAK
                                  if traffic and launch_s > 0 else None)},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(eng, args.cpu_sample_log2,
                                           threads)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    if not args.no_batched:
        out["batched"] = run_batched()
    if not args.no_serving:
        out["serving"] = run_serving(args.serve_requests, args.serve_rate)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
