"""Drop-in for euler-prime-runs/euler_gpu.py's GpuEngine, on MI355X.

Same construction (GpuEngine(n, wheel_primes=None)), same range sweep
(survivors_pre_mr(lo, hi) -> sorted list of exact p in [lo, hi)), same
hunt (ascending [(p, exact_run)] with run >= n), same cursor unit (eng.M,
the wheel period, and eng.wheel_primes), same errors (ValueError below
P_FLOOR or above P_CEIL, RuntimeError on device trouble).  The sieve,
the compaction rounds and the strong-probable-prime chains all run in
libfk_amd.so on the GPU; nothing here computes a survivor.

The survivor set does not depend on the wheel (euler_search.py:1-9): a
survivor is a p with no prime q < 65536 dividing any of x^2+x+p, x < n.
"""

import fk_families as F
import fk_native as N

P_FLOOR = F.EULER_P_FLOOR
P_CEIL = F.EULER_P_CEIL
Q2 = F.EULER_Q2
SWEEP_CAP = 64          # chain values a sweep classifies per survivor


class GpuEngine:
    """euler-prime-runs/euler_gpu.py:897 GpuEngine, rebuilt for gfx950."""

    def __init__(self, n, wheel_primes=None, device=0, **tuning):
        self.n = int(n)
        self.wheel_primes = list(wheel_primes or F.euler_best_wheel(self.n))
        t = F.tuning("euler")
        t.update(tuning)
        self._eng = N.Engine(N.FAMILY_EULER, self.n, Q2, device=device,
                             wheel_primes=self.wheel_primes, **t)
        self.M = F.product(self.wheel_primes)
        self.n_offs = self._eng.nlines
        self.info = self._eng.info

    def _check(self, lo, hi):
        lo, hi = int(lo), int(hi)
        if lo < P_FLOOR:
            raise ValueError("GPU engine floor is %d" % P_FLOOR)
        if hi > P_CEIL:
            raise ValueError("GPU engine ceiling is %d" % P_CEIL)
        return lo, hi

    def _pieces(self, lo, hi, periods_per_launch):
        if not periods_per_launch:
            return [(lo, hi)]
        step = int(periods_per_launch) * self.M
        return [(a, min(a + step, hi)) for a in range(lo, hi, step)]

    def survivors_pre_mr(self, lo, hi, periods_per_launch=None):
        """Sorted list of exact survivor p (Python ints) in [lo, hi)."""
        lo, hi = self._check(lo, hi)
        out = []
        for a, b in self._pieces(lo, hi, periods_per_launch):
            if b > a:
                out.extend(N.record_values(self._eng.sweep(a, b)))
        return out

    def hunt(self, lo, hi, cap=100, periods_per_launch=None):
        """Ascending [(p, exact_run)] with run >= n in [lo, hi)
        (euler_gpu.py:1458-1464); the chains run on the device, 64 values
        in the sweep and further, to cap, for the chains that reach 64."""
        lo, hi = self._check(lo, hi)
        cap = int(cap)
        look = min(max(cap, self.n), SWEEP_CAP)
        out = []
        for a, b in self._pieces(lo, hi, periods_per_launch):
            if b <= a:
                continue
            recs = self._eng.sweep(a, b, N.CLASSIFY_FULL, look)
            ps = N.record_values(recs)
            runs = self._eng.extend_runs(ps, recs["run"].tolist(), look, cap)
            for p, run in zip(ps, runs):
                if run >= self.n:
                    out.append((p, min(int(run), cap)))
        return out

    def mr_run_length(self, p, cap=100):
        """euler_search.py:258-262 on the device (cap up to 256)."""
        return self._eng.classify([int(p)], N.CLASSIFY_FULL, int(cap))[0]

    def run_lengths(self, ps, cap=100):
        return self._eng.classify(ps, N.CLASSIFY_FULL, int(cap))
