"""Autotuner for the sieve engine's free parameters.

The survivor stream does not depend on any of them (the GPU parity tests
check this across tile shapes, cutoffs, wheels and round growth), so the
tuner is free to pick whatever is fastest:

  threads, words_per_thread   workgroup size and tile length of k_sieve
  q_pat                       primes below it are OR-ed in as pattern words
  q_mark                      primes in [q_pat, q_mark) are stride-marked
                              in LDS
  q_bits                      primes in [q_mark, q_bits) OR in 64-bit
                              windows of periodic kill bitmaps
  q_med                       primes in [q_mark, q_med) are tested per
                              surviving position against LDS kill bitmaps;
                              above it per candidate in compaction rounds
  nwheel                      kill primes enumerated by the wheel
  round_growth                prime-range growth factor between rounds

A window whose lines hold at most short_words words (the engine's default,
1024) takes the short-line sieve instead of k_sieve: there every LDS prime
is a bitmap window and only q_med, nwheel and round_growth matter, so the
tuning window's length decides which parameters the tuner can see.

Coordinate descent from the current tuning (the saved one, else the family
defaults): each pass tries every value of one parameter with the others
held, keeps the fastest, and moves on; the result replaces the starting
point only if it is faster again when both are re-measured.
Timing is the wall time per window of a pipelined run over consecutive
representative windows (as the engine sweeps a range, PIPELINE_DEPTH
windows in flight), median of `reps`; --alone times one window's device
time instead (HIP events around the batch).  The result is written to
fk_tuning.json next to this file, which fk_families.tuning() reads.

python fk_autotune.py --family dickson --n 10 --q2 4194304 \
                      --lo 432900432900 --span 4294967296
"""

import argparse
import json
import pathlib
import statistics
import sys
import time

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

import fk_families as F   # noqa: E402
import fk_native as N     # noqa: E402

FAMILY = {"euler": N.FAMILY_EULER, "dickson": N.FAMILY_DICKSON,
          "ap": N.FAMILY_AP}

SPACE = {
    "threads": [256, 512, 1024],
    "words_per_thread": [1, 2, 3, 4, 6, 8],
    "q_pat": [64, 128, 192, 256, 320],
    "q_med": [128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096,
              8192, 16384],
    "nwheel": [0, 3, 4, 5, 6, 7],
    "round_growth": [2, 3, 4],
    "q_mark": [0, 192, 256, 512, 1024],
    "q_bits": [0, 1, 256, 512, 1024, 2048],
}

VALID_SHAPES = {(256, 1), (256, 2), (256, 3), (256, 4), (256, 6), (256, 8),
                (256, 16), (512, 1), (512, 2), (512, 3), (512, 4), (512, 8),
                (1024, 1), (1024, 2), (1024, 4)}


def measure(family, n, q2, lo, span, params, reps=3, classify=None,
            wheel_primes=None, pipelined=True):
    """Median milliseconds per window, or None if the configuration is not
    buildable (LDS, shape, wheel size).  pipelined: the wall time per
    window of 4 * PIPELINE_DEPTH consecutive windows swept as the engine
    sweeps a range (Engine.sweep_many, that many windows in flight), after
    one pass that captures the slots' launch sequences; else the device
    time of one window swept alone (HIP events around the batch)."""
    if (params["threads"], params["words_per_thread"]) not in VALID_SHAPES:
        return None, None
    kw = dict(params)
    if wheel_primes:
        kw.pop("nwheel", None)
        kw["wheel_primes"] = wheel_primes
    try:
        eng = N.Engine(FAMILY[family], n, q2, **kw)
    except (ValueError, RuntimeError):
        return None, None
    mode, cap, ef = classify or (N.CLASSIFY_NONE, 0, 0)
    try:
        eng.sweep(lo, lo + span, mode, cap, ef)       # warm
        times, stages = [], []
        if pipelined:
            wins = [(lo + k * span, lo + (k + 1) * span)
                    for k in range(4 * N.PIPELINE_DEPTH)]
            for _ in eng.sweep_many(wins[:N.PIPELINE_DEPTH], mode, cap, ef):
                pass
        for _ in range(reps):
            if pipelined:
                t0 = time.perf_counter()
                for _ in eng.sweep_many(wins, mode, cap, ef):
                    pass
                times.append(1e3 * (time.perf_counter() - t0) / len(wins))
                stages.append(None)
                continue
            eng.sweep(lo, lo + span, mode, cap, ef)
            ms, _, _ = eng.last_timing()
            times.append(ms[4])
            stages.append(ms)
    except (ValueError, RuntimeError):
        return None, None
    finally:
        eng.close()
    i = times.index(statistics.median(times))
    return statistics.median(times), stages[i]


def tune(family, n, q2, lo, span, passes=2, reps=3, classify=None,
         wheel_primes=None, log=print, pipelined=True):
    cur = dict(F.tuning(family))
    cur.setdefault("nwheel", 0)
    cur.setdefault("round_growth", 2)
    cur.setdefault("q_mark", 0)
    cur.setdefault("q_bits", 0)
    cur.pop("round_lane_prime", None)
    best, st = measure(family, n, q2, lo, span, cur, reps, classify,
                       wheel_primes, pipelined)
    log(f"start {cur} -> {best:.3f} ms {st}")
    start = dict(cur)
    keys = ["q_med", "q_bits", "q_mark", "q_pat", "threads",
            "words_per_thread", "round_growth"]
    if not wheel_primes:
        keys.append("nwheel")
    for _ in range(passes):
        improved = False
        for key in keys:
            for val in SPACE[key]:
                if val == cur.get(key):
                    continue
                trial = dict(cur)
                trial[key] = val
                if key == "threads":
                    # keep the tile length when the workgroup changes
                    bits = cur["threads"] * cur["words_per_thread"]
                    trial["words_per_thread"] = max(1, bits // val)
                t, st = measure(family, n, q2, lo, span, trial, reps,
                                classify, wheel_primes, pipelined)
                log(f"  {key}={val}: "
                    + ("unbuildable" if t is None else f"{t:.3f} ms {st}"))
                if t is not None and t < best * 0.98:
                    best, cur, improved = t, trial, True
        if not improved:
            break
    if cur != start:
        # single measurements are noisy (2-3% between near-equal shapes):
        # the winner must beat the starting point again, at three times the
        # repetitions, before it replaces it
        t_new, _ = measure(family, n, q2, lo, span, cur, 3 * reps, classify,
                           wheel_primes, pipelined)
        t_old, _ = measure(family, n, q2, lo, span, start, 3 * reps, classify,
                           wheel_primes, pipelined)
        log(f"confirm: start {t_old:.3f} ms, best {t_new:.3f} ms")
        if t_old is not None and (t_new is None or t_new > 0.99 * t_old):
            cur, best = start, t_old
        else:
            best = t_new
    log(f"best {cur} -> {best:.3f} ms")
    return cur, best


def save(family, params, path=F.TUNING_PATH):
    data = {}
    if path.exists():
        try:
            data = json.loads(path.read_text())
        except ValueError:
            data = {}
    data[family] = params
    path.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--family", required=True, choices=sorted(FAMILY))
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--q2", type=int, required=True)
    ap.add_argument("--lo", type=int, required=True)
    ap.add_argument("--span", type=int, required=True)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--cap", type=int, default=0)
    ap.add_argument("--wheel", type=str, default="")
    ap.add_argument("--save", action="store_true")
    ap.add_argument("--alone", action="store_true",
                    help="time one window swept alone, not a pipelined run")
    args = ap.parse_args()
    cls = (N.CLASSIFY_FULL, args.cap, 0) if args.cap else None
    wheel = [int(x) for x in args.wheel.split(",") if x] or None
    t0 = time.time()
    params, best = tune(args.family, args.n, args.q2, args.lo, args.span,
                        args.passes, args.reps, cls, wheel,
                        log=lambda s: print(s, flush=True),
                        pipelined=not args.alone)
    print(json.dumps({"family": args.family, "params": params,
                      "ms": best, "tuning_s": time.time() - t0}), flush=True)
    if args.save:
        save(args.family, params)
    return 0


if __name__ == "__main__":
    sys.exit(main())
