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
                window by window (four in flight), with the two streams
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
import statistics
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "first-knowns_amd"))

METRIC = ("candidates swept/sec at bit-exact survivor stream; "
          "batched vs plain GPU engine")
# LDS read bandwidth: 256 B/clk/CU for 16-byte reads on gfx950 (128 B/clk
# was the previous generation's); fk_probe_peaks measures about 131 TB/s
# of it on the box, which 128 B/clk could not deliver
LDS_PEAK = 256 * 256 * 2.4e9          # B/s: 256 CUs x 256 B/clk x 2.4 GHz
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
N_STEP_SIEVE, N_STEP_ROUND_C, N_STEP_ROUND_P, N_STEP_COMPACT = 1, 2, 3, 4


def sweep_window(seed):
    j_lo = 432_900_432_900 + splitmix64(seed) % 10**11
    return j_lo, j_lo + SWEEP_SPAN


def run_plain(steps, warmup):
    """The timed loop: `steps` windows swept in order on the plain engine,
    four in flight (Engine.sweep_many: each window one engine call on its
    own stream, hardware queue and captured graph, the next ones enqueued
    before the last is read back: the reference's two-stream launch loop,
    deeper).  The warm-up sweeps at least one window per pipeline slot, so
    that every slot has captured its launch sequence before the timed loop
    (a capture is host work of milliseconds).  device_ms is one window's
    device time alone: more windows swept one at a time (graph replays)
    after the timed loop.  The stage split comes from three more windows
    swept eagerly with HIP events between the stages."""
    import fk_native as N
    import ladder_gpu
    eng = ladder_gpu.GpuEngine(SWEEP_N, q2=SWEEP_Q2)
    nwarm = max(warmup, N.PIPELINE_DEPTH)
    for _ in eng._eng.sweep_many([sweep_window(1000 + s) for s in range(nwarm)],
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
    fp = 0
    for recs in out:
        for v in N.record_values(recs):
            fp ^= int(v)
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
    launches = []                  # per window: [(kind, ms)] of its launches
    counts = []
    for s in range(nst):
        lo, hi = sweep_window(s)
        eng._eng.sweep(lo, hi, N.CLASSIFY_FULL, SWEEP_CAP)
        ms, _, c = eng._eng.last_timing()
        cand += c
        for i in range(5):
            stage[i] += ms[i] / nst
        launches.append(eng._eng.last_step_times())
        counts.append(eng._eng.last_counts())
    eng._eng.stage_timing(False)
    return {"eng": eng, "wall": wall, "surv": surv, "fp": fp, "stage": stage,
            "cand": cand / nst, "deepest": deepest, "device_ms": device_ms,
            "launches": launches, "counts": counts, "warm_windows": nwarm}


def cpu_baseline(eng, sample_log2, threads, one_core_log2):
    """The oracle's C restatement of the reference CPU sieve plus the
    Python-integer classification of its survivors, on `threads` threads
    over 2^sample_log2 j of the step-0 window, and on one thread over
    2^one_core_log2 j; each sample's stream is checked against the GPU's."""
    import fk_native as N
    sys.path.insert(0, str(ROOT / "oracle"))
    import fk_oracle as O
    sieve = O.DicksonSieve(SWEEP_N, SWEEP_Q2)            # tables outside the clock

    def leg(log2, nthreads):
        lo, _ = sweep_window(0)
        hi = lo + (1 << log2)
        t0 = time.perf_counter()
        js = sieve.survivors(lo, hi, nthreads=nthreads)
        runs = [O.dickson_run(j * eng.W, cap=SWEEP_CAP) for j in js]
        wall = time.perf_counter() - t0
        recs = eng._eng.sweep(lo, hi, N.CLASSIFY_FULL, SWEEP_CAP)
        same = (N.record_values(recs) == js and
                [int(r) for r in recs["run"].tolist()] == runs)
        return (hi - lo) / wall, len(js), wall, same

    rate, nsurv, wall, same = leg(sample_log2, threads)
    rate1, nsurv1, wall1, same1 = leg(one_core_log2, 1)
    return {"value": rate, "unit": "j-candidates/s",
            "cores": threads, "kind": "port",
            "host_cpu_count": os.cpu_count(),
            "sample": f"2^{sample_log2} j from the step-0 window, n={SWEEP_N}, "
                      f"q2=2^22, survivors classified (cap {SWEEP_CAP}); "
                      f"{nsurv} survivors, {wall:.1f} s on {threads} threads",
            "sample_stream_equal": bool(same),
            "value_one_core": rate1,
            "sample_one_core": f"2^{one_core_log2} j of the same window on one "
                               f"thread; {nsurv1} survivors, {wall1:.1f} s",
            "sample_one_core_stream_equal": bool(same1)}


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


def run_batched(repeats=3):
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
    # plain: one engine call per window, four in flight (the plain engine's
    # own pipelining, as in the headline loop); median of three passes
    pipe = N.Pipeline()
    jobs = [(engines[eo[i]], *ws[i], specs[i]) for i in range(len(ws))]
    for _ in pipe.run(jobs[:3]):
        pass
    t_plain = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        plain = [r for _, r in pipe.run(jobs)]
        t_plain.append(time.perf_counter() - t0)
    pipe.close()
    # batched: one submission, its launch sequence captured as a HIP graph
    # on the first submission and replayed after; median of three replays
    b = N.Batch(graph=2)
    b.submit(engines, eo, ws, specs)
    b.results()
    t_batch = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        b.submit(engines, eo, ws, specs)
        recs, counts = b.results()
        t_batch.append(time.perf_counter() - t0)
    captures, launches = b.graph_stats()
    same = True
    pos = 0
    for i, p in enumerate(plain):
        mine = recs[pos:pos + int(counts[i])]
        pos += int(counts[i])
        if (N.record_values(mine) != N.record_values(p) or
                mine["run"].tolist() != p["run"].tolist()):
            same = False
    tp, tb = statistics.median(t_plain), statistics.median(t_batch)
    return {"windows": len(ws), "engines": len(engines),
            "candidates": cand, "survivors": int(recs.size),
            "plain_s": tp, "batched_s": tb,
            "plain_s_runs": t_plain, "batched_s_runs": t_batch,
            "plain_rate": cand / tp, "batched_rate": cand / tb,
            "speedup": tp / tb, "streams_equal": same,
            "batched_graph": {"captures": captures, "launches": launches},
            "note": f"medians of {repeats} passes; the batch replays its "
                    f"captured launch sequence"}


# ------------------------------------ serving: Poisson arrivals

def run_serving(count, rate, seed=4, repeats=3):
    """Poisson arrivals of range-sweep requests of all three families,
    spans log-uniform in [2^18, 2^32] line values, depths 8..14, Dickson k
    up to about 1e22 (192-bit values); continuous batching over 4 slots.
    The figures are medians over `repeats` runs of the same arrivals."""
    import fk_serve as S
    reqs = S.poisson_requests(seed, count, rate)
    srv = S.Server(slots=4, budget=64 << 30)
    # engines are built (resident tables) before the clock starts
    for r in reqs:
        srv.engine(r.family, r.n, r.q2)
    runs = []
    for _ in range(repeats):
        srv.submissions = []
        st = srv.run(reqs)
        st["submissions"] = len(srv.submissions)
        st["mean_requests_per_submission"] = (
            sum(srv.submissions) / len(srv.submissions)
            if srv.submissions else 0)
        runs.append(st)
    out = dict(runs[0])
    for k in ("p50_ms", "p99_ms", "wall_s", "candidates_per_s",
              "mean_requests_per_submission"):
        out[k] = statistics.median(r[k] for r in runs)
    out["p99_ms_runs"] = [r["p99_ms"] for r in runs]
    out["rate_per_s"] = rate
    out["slots"] = 4
    out["note"] = f"medians of {repeats} runs of the same arrivals"
    return out


# ------------------------------------ the Euler n = 17 shape, two wheels

EULER_WIDE = (611 * 10**18, 2 * 10**16)     # euler-prime-runs/score.py:57-77


def run_euler_wide(golden_xor=None):
    """The Euler SCORE_WIDE shape (n = 17, p in [6.11e20, +2e16), every
    prime below 2^16 sieved) on the 31# wheel (29,937,600 residues) and on
    the 37# wheel (598,752,000, built factored on the device): p/s of each
    sweep after one warm-up sweep, and each stream's count and XOR."""
    import euler_gpu
    lo, span = EULER_WIDE
    out = {}
    for name, wp in (("wheel_31", [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31]),
                     ("wheel_37", [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31,
                                   37])):
        eng = euler_gpu.GpuEngine(17, wheel_primes=wp)
        eng.survivors_pre_mr(lo - 10**15, lo)                # warm-up
        t0 = time.perf_counter()
        got = eng.survivors_pre_mr(lo, lo + span)
        wall = time.perf_counter() - t0
        x = 0
        for p in got:
            x ^= p
        out[name] = {"value": span / wall, "unit": "p-candidates/s",
                     "seconds": wall, "residues": eng.n_offs,
                     "count": len(got), "xor": x}
        eng._eng.close()
    out["speedup_37_over_31"] = (out["wheel_37"]["value"] /
                                 out["wheel_31"]["value"])
    return out


# ------------------------------------ the reference's frozen benchmark shapes

# The reference's own published rates at these shapes, all on one RTX 4090
# (host survivor tests excluded): dickson-ladders/BENCHMARKS.md:38-40,
# primorial-ap/BENCHMARKS.md:35-38 (best of four captures),
# euler-prime-runs/BENCHMARKS.md:107 and :420-424 (best capture).
REF_SHAPE_RATES = {
    ("dickson", "SCORE"): 2.06e12, ("dickson", "SCORE12"): 2.49e12,
    ("dickson", "SCORE13"): 5.43e12, ("ap", "SCORE"): 4.45e10,
    ("euler", "SCORE"): 6.23e15, ("euler", "SCORE128"): 6.34e15,
    ("euler", "SCORE_WIDE"): 3.73e16,
}


def run_reference_shapes():
    """Each frozen benchmark shape of the reference (dickson-ladders/
    score.py:54-59, primorial-ap/score.py:42-60, euler-prime-runs/
    score.py:57-77), swept by the drop-in engines at its own parameters:
    line values per second, the median of five sweeps after two warm-up
    sweeps of the same shape (the first sweeps build the engine's batch
    buffers and capture its launch sequence; survivors found, not
    classified, as the reference's score), the stream's count and XOR
    against the reference's recorded fingerprint, and the reference's
    published rate."""
    import ap_gpu
    import euler_gpu
    import ladder_gpu
    g = json.loads((ROOT / "tests" / "golden" / "reference_vectors.json")
                   .read_text())
    out = []

    def big(v):
        return int(v)

    def timed(fam, label, make, sweep, span, want):
        eng = make()
        for _ in range(2):                         # warm-up: buffers, graphs
            sweep(eng)
        walls = []
        for _ in range(5):
            t0 = time.perf_counter()
            got = sweep(eng)
            walls.append(time.perf_counter() - t0)
        wall = statistics.median(walls)
        x = 0
        for v in got:
            x ^= int(v)
        ref = REF_SHAPE_RATES[(fam, label)]
        out.append({"family": fam, "shape": label, "span": span,
                    "seconds": wall, "rate": span / wall,
                    "fingerprint_equal": (len(got), x) == want,
                    "reference_rate_rtx4090": ref,
                    "ratio": span / wall / ref})

    for s in g["dickson_fingerprints"]:
        j0, span = big(s["j0"]), big(s["span"])
        timed("dickson", s["label"],
              lambda s=s: ladder_gpu.GpuEngine(s["n"], q2=s["q2"]),
              lambda e, j0=j0, span=span: e.survivors_j(j0, j0 + span).tolist(),
              span, (s["count"], big(s["xor"])))
    for s in g["ap_fingerprints"]:
        base, span = big(s["base"]), big(s["span"])

        def ap_sweep(e, base=base, span=span):
            vals = []
            for c in e.survivors(base, span):
                vals.extend(c.tolist())
            return vals
        timed("ap", s["label"],
              lambda s=s: ap_gpu.GpuEngine(s["n"], s["q2"], launch_u=s["launch_u"]),
              ap_sweep, span, (s["count"], big(s["xor"])))
    for s in g["euler_fingerprints"]:
        lo, span = big(s["lo"]), big(s["span"])
        timed("euler", s["label"], lambda s=s: euler_gpu.GpuEngine(s["n"]),
              lambda e, lo=lo, span=span: e.survivors_pre_mr(lo, lo + span),
              span, (s["count"], big(s["xor"])))
    return out


# ------------------------------------ config 2: the AP sweep

AP_N, AP_Q2 = 17, 1 << 24
AP_SPAN = 30 * ((1 << 32) // 30)


def ap_window(seed):
    base = 10**16 + splitmix64(seed) % 10**15
    base -= base % 30
    return base, base + AP_SPAN


def run_ap(steps=16, warmup=4):
    """The plain engine on config 2: n = 17, d = 17#, primes 61..2^24
    sieved (7..59 kill residue 0), windows of 30 floor(2^32 / 30)
    consecutive p near 1e16, survivors classified two-pass (base 2, every
    base from depth 6, cap 17); windows in order, four in flight."""
    import fk_native as N
    import fk_families as F
    eng = N.Engine(N.FAMILY_AP, AP_N, AP_Q2, **F.tuning("ap"))
    spec = (N.CLASSIFY_TWO_PASS, AP_N, 6)
    for _ in eng.sweep_many([ap_window(100 + s)
                             for s in range(max(warmup, N.PIPELINE_DEPTH))],
                            *spec):
        pass
    wins = [ap_window(s) for s in range(steps)]
    t0 = time.perf_counter()
    out = list(eng.sweep_many(wins, *spec))
    wall = time.perf_counter() - t0
    eng.stage_timing(True)
    eng.sweep(*wins[0], *spec)
    ms = eng.last_timing()[0]
    counts = eng.last_counts()
    eng.stage_timing(False)
    surv = sum(int(r.size) for r in out)
    eng.close()
    return {"value": steps * AP_SPAN / wall, "unit": "p-candidates/s",
            "ms_per_window": 1e3 * wall / steps, "windows": steps,
            "survivors": surv,
            "stages_ms": dict(zip(("prep", "sieve", "rounds", "classify",
                                   "whole"), ms)),
            "queue_after_each_stage": counts,
            "config": {"n": AP_N, "q2": AP_Q2, "window_p": AP_SPAN,
                       "classify": "two-pass, exact from 6, cap 17"}}


VALU_PEAK = 256 * 64 * 2.4e9           # lane-ops/s: 256 CUs x 64 lanes/clk


def rocprof_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from its committed PMC summary,
    profiles/<kernel>_dickson_sweep_traffic.json (None when there is none)."""
    path = ROOT / "profiles" / f"{kernel}_dickson_sweep_traffic.json"
    if path.exists():
        try:
            js = json.loads(path.read_text())
        except (OSError, ValueError):
            return None
        if kernel in js.get("kernel", ""):
            return js.get("bytes_per_dispatch")
    return None


def legendre_is_one(a, q):
    """(a | q) == 1 for arrays of a and odd primes q (q < 2^26), by
    a^((q-1)/2) mod q with numpy uint64 (products below 2^52)."""
    import numpy as np
    e = (q - 1) // 2
    base = a % q
    acc = np.ones_like(q)
    while e.any():
        odd = (e & 1).astype(bool)
        acc = np.where(odd, acc * base % q, acc)
        base = base * base % q
        e >>= 1
    return acc == 1


def round_work(info, counts, growth):
    """Algorithmic work of the Dickson sweep's compaction rounds, per
    window: round r tests every candidate left after the stage before it
    (counts[r]) against every prime of [q_med g^r, q_med g^(r+1)) (the last
    round ends at q2).  A test is the folded residue of k (below 2^53, one
    exact double) mod q -- multiply by 1/q, round, one FMA, convert, then
    min(u, q - u): 6 VALU ops -- and one compare per member of the prime's
    folded kill set, whose size is #{m <= n : (-m | q) = 1}; a prime with
    an empty set needs no test.  The kernels compare in groups of four, so
    this is a lower bound on their VALU work.  Returns (tests, VALU
    lane-ops, per-round [(first prime, primes, candidates)])."""
    import numpy as np
    import fk_families as F
    ps = np.array([q for q in F.primes_below(int(info.q2))
                   if q >= int(info.q_med)], dtype=np.uint64)
    n = int(info.n)
    cnt = np.zeros(ps.size, dtype=np.int64)
    for m in range(1, n + 1):
        cnt += legendre_is_one((ps - np.uint64(m)) % ps, ps)
    ops = np.where(cnt > 0, 6 + cnt, 0)
    # the engine's boundaries: q_med g^r (g an integer, or hundredths from
    # 101 up), each at the first prime at or past it, repeats dropped
    g = growth / 100.0 if growth >= 100 else float(growth)
    q2 = int(info.q2)
    idx = [0]
    v = int(info.q_med) * g
    while True:
        i = int(np.searchsorted(ps, int(min(v, q2))))
        if i > idx[-1]:
            idx.append(i)
        if v >= q2:
            break
        v *= g
    if idx[-1] < ps.size:
        idx.append(int(ps.size))
    tests = vops = 0
    rounds = []
    for r in range(len(idx) - 1):
        lo, hi = idx[r], idx[r + 1]
        c = int(counts[r]) if r < len(counts) else 0
        tests += c * int(hi - lo)
        vops += c * int(ops[lo:hi].sum())
        rounds.append((int(ps[lo]), int(hi - lo), c))
    return int(tests), int(vops), rounds


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
    ap.add_argument("--no-ap", action="store_true")
    ap.add_argument("--no-euler", action="store_true")
    ap.add_argument("--cpu-one-core-log2", type=int, default=27)
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=400.0)
    # a second arrival rate, past what one request at a time could serve,
    # where requests that arrive while the slots are busy share submissions
    ap.add_argument("--serve-requests-high", type=int, default=1200)
    ap.add_argument("--serve-rate-high", type=float, default=4000.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device copy of MIB MiB (counter "
                         "calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the engine has no multi-GPU path "
              f"(windows would shard across GPUs by range of the search "
              f"line). --gpus {args.gpus} is not supported.", file=sys.stderr)
        return 2

    P = run_plain(args.steps, args.warmup)
    eng, wall, stage = P["eng"], P["wall"], P["stage"]
    j_total = args.steps * SWEEP_SPAN
    value = j_total / wall
    info = eng._eng.info
    # per-launch HIP-event durations of the eager windows
    def launch_ms(kinds):
        per = [sum(ms for k, ms in w if k in kinds) for w in P["launches"]]
        n = [sum(1 for k, _ in w if k in kinds) for w in P["launches"]]
        return sum(per) / len(per), sum(n) / len(n)
    round_ms, round_launches = launch_ms({N_STEP_ROUND_C})
    roundp_ms, roundp_launches = launch_ms({N_STEP_ROUND_P})
    sieve_ms, _ = launch_ms({N_STEP_SIEVE})
    compact_ms, _ = launch_ms({N_STEP_COMPACT})
    # the rounds: tests and VALU lane-ops per window from the queue sizes
    import fk_families as F
    growth = int(F.tuning("dickson").get("round_growth") or 2)
    tests, vops, rounds = round_work(info, P["counts"][0], growth)
    r_achieved = vops / (round_ms / 1e3) if round_ms > 0 else 0.0
    r_traffic = rocprof_traffic("k_round_c")
    # the sieve: algorithmic LDS bytes of the window's wheel candidates
    kernel, per_cand = sieve_kernel(info, SWEEP_SPAN)
    s_launch = sieve_ms / 1e3
    s_achieved = P["cand"] * per_cand / s_launch if s_launch > 0 else 0.0
    s_traffic = rocprof_traffic(kernel)
    # the denominators measured on this device (fk_probe_peaks), beside the
    # published ones: VALU issue (the faster of the two probe loops) and LDS
    # read bandwidth
    import fk_native as N
    pf64, pint, plds = N.probe_peaks(0)
    valu_meas = max(pf64, pint)
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
                               "in order, four in flight",
                   "n": SWEEP_N, "q2": SWEEP_Q2, "window_j": SWEEP_SPAN,
                   "classify_cap": SWEEP_CAP,
                   "wheel_lines": int(info.nlines),
                   "wheel_period": int(info.line_step),
                   "q_pat": int(info.q_pat), "q_med": int(info.q_med),
                   "tile_bits": None},
        "survivors": P["surv"],
        "fingerprint": {"count": P["surv"], "xor": P["fp"],
                        "of": "the j of every survivor of the timed windows"},
        "deepest_ladder": {"run": P["deepest"][0], "j": P["deepest"][1]},
        "device_ms_per_window_alone": P["device_ms"],
        "stages_ms_per_step": dict(zip(("prep", "sieve", "rounds",
                                        "classify", "whole"), stage)),
        "launch_ms_per_step": {"k_round_c": round_ms,
                               "k_round_c_launches": round_launches,
                               "k_round_p": roundp_ms,
                               "k_round_p_launches": roundp_launches,
                               "k_round_compact": compact_ms,
                               "sieve": sieve_ms},
        "warmup_windows": P["warm_windows"],
        "stages_note": "per-stage and per-launch HIP events on three windows "
                       "swept eagerly after the timed loop; the timed sweeps "
                       "keep four windows in flight, each replaying its "
                       "slot's captured launch sequence (the warm-up sweeps "
                       "warmup_windows = max(warmup, 4) windows, one at "
                       "least per slot); "
                       "device_ms_per_window_alone is one window swept by "
                       "itself (graph replay, event-timed)",
        "measured_peaks": {"valu_f64_Tlane_op_s": pf64 / 1e12,
                           "valu_int_Tlane_op_s": pint / 1e12,
                           "lds_GBps": plds / 1e9,
                           "note": "fk_probe_peaks: eight dependent chains "
                                   "per lane of double FMAs / 32-bit funnel "
                                   "shifts over every CU, and conflict-free "
                                   "16-byte LDS reads; HIP-event timed"},
        "roofline": {"bound": "valu", "kernel": "k_round_c",
                     "achieved": r_achieved / 1e12, "peak": VALU_PEAK / 1e12,
                     "unit": "Tlane-op/s", "frac": r_achieved / VALU_PEAK,
                     "peak_measured": valu_meas / 1e12,
                     "frac_measured": (r_achieved / valu_meas
                                       if valu_meas > 0 else None),
                     "traffic": r_traffic,
                     "tests_per_window": tests,
                     "valu_lane_ops_per_window": vops,
                     "valu_ops_per_test": vops / tests if tests else None,
                     "launches_per_window": round_launches,
                     "rounds": [{"first_prime": a, "primes": b,
                                 "candidates": c} for a, b, c in rounds],
                     "hbm_GBps": (r_traffic * round_launches / (round_ms / 1e3)
                                  / 1e9 if r_traffic and round_ms > 0
                                  else None)},
        "roofline_sieve": {"bound": "lds", "kernel": kernel,
                           "achieved": s_achieved / 1e9,
                           "peak": LDS_PEAK / 1e9, "unit": "GB/s",
                           "frac": s_achieved / LDS_PEAK,
                           "peak_measured": plds / 1e9,
                           "frac_measured": (s_achieved / plds
                                             if plds > 0 else None),
                           "traffic": s_traffic,
                           "algorithmic_bytes_per_candidate": per_cand,
                           "candidates_per_launch": P["cand"],
                           "hbm_GBps": (s_traffic / s_launch / 1e9
                                        if s_traffic and s_launch > 0
                                        else None)},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(eng, args.cpu_sample_log2,
                                           threads, args.cpu_one_core_log2)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        out["gpu_over_cpu_one_core"] = (value /
                                        out["cpu_baseline"]["value_one_core"])
    if not args.no_batched:
        out["batched"] = run_batched()
    if not args.no_serving:
        out["serving"] = run_serving(args.serve_requests, args.serve_rate)
        out["serving_high_rate"] = run_serving(args.serve_requests_high,
                                               args.serve_rate_high)
    if not args.no_ap:
        out["config2_ap"] = run_ap()
    if not args.no_euler:
        out["euler_score_wide"] = run_euler_wide()
        out["reference_shapes"] = run_reference_shapes()
    print(json.dumps(out, default=_plain), flush=True)
    return 0


def _plain(x):
    """numpy scalars in the JSON line as plain numbers."""
    if hasattr(x, "item"):
        return x.item()
    raise TypeError(f"{type(x).__name__} is not JSON serializable")


if __name__ == "__main__":
    sys.exit(main())
