"""Drop-in for dickson-ladders/ladder_gpu.py's GpuEngine, on MI355X.

Same construction (GpuEngine(n, q2=65536, ns, T, ratio, launch, warp_pop,
vote, fold)), same sweeps (survivors_j -> sorted np.uint64 j in
[j_lo, j_hi); survivors_j_deep -> sorted Python ints to the folded reach;
survivors_pre_mr -> the k-line view with both k ends inclusive), same
classification (run_length, hunt) and the same cursor attributes (W, P,
offs, J_REACH, q2).  The reference's tuning knobs (ns, T, ratio, launch,
warp_pop, vote) shape its CUDA kernels; they are accepted and ignored here,
where the tile shape and prime cutoffs come from the tuning table.
"""

import numpy as np

import fk_families as F
import fk_native as N

Q2_DEFAULT = F.DICKSON_Q2_DEFAULT
Q2_MAX = F.Q2_MAX
J_CEIL = F.DICKSON_J_CEIL
K_FLOOR = F.DICKSON_K_FLOOR
SPRP_EXACT_FROM = F.SPRP_EXACT_FROM


class GpuEngine:
    """dickson-ladders/ladder_gpu.py:424 GpuEngine, rebuilt for gfx950."""

    def __init__(self, n, q2=Q2_DEFAULT, ns=None, T=None, ratio=2.0,
                 launch=1 << 34, warp_pop=1 << 16, vote=1, fold="auto",
                 device=0, **tuning):
        if not (2 < q2 <= Q2_MAX):
            raise ValueError(
                f"q2 {q2} is outside the range this engine's arithmetic is "
                f"exact for (2, {Q2_MAX}]")
        self.n = int(n)
        self.q2 = int(q2)
        self.W = F.dickson_wheel_modulus(self.n)
        allp = [q for q in F.primes_below(self.q2) if q >= self.n + 2]
        # the fold prime and its offsets: the unit the reference's cursor
        # and J_REACH are expressed in (ladder_gpu.py:446-455)
        if fold == "auto":
            fold = allp[0] if len(allp) > 1 else None
        if fold:
            fold = int(fold)
            if fold not in allp:
                raise ValueError(f"fold prime {fold} is not a sieve prime "
                                 f"at n={n}, q2={q2}")
        self.P = int(fold) if fold else 1
        self.offs = (F.dickson_fold_offsets(self.W, self.P, self.n)
                     if fold else [0])
        self.J_REACH = self.P * J_CEIL
        self.primes = np.array([q for q in allp if q != fold],
                               dtype=np.uint64)
        self.NP = int(self.primes.size)
        t = F.tuning("dickson")
        t.update(tuning)
        self._eng = N.Engine(N.FAMILY_DICKSON, self.n, self.q2,
                             device=device, **t)
        self.info = self._eng.info

    def _sweep(self, j_lo, j_hi, mode=N.CLASSIFY_NONE, cap=0, exact_from=0):
        j_lo, j_hi = int(j_lo), int(j_hi)
        if j_hi <= j_lo:
            return np.empty(0, dtype=N.RECORD_DTYPE)
        return self._eng.sweep(j_lo, j_hi, mode, cap, exact_from)

    def survivors_j(self, j_lo, j_hi, launch=None):
        """Sorted u64 array of surviving j in [j_lo, j_hi)
        (ladder_gpu.py:757-781, with the same refusals)."""
        if j_hi > min(self.J_REACH, 1 << 62):
            raise ValueError(
                f"j {j_hi} past this API's reach "
                f"{min(self.J_REACH, 1 << 62)}; the deep zone is "
                f"survivors_j_deep's (Python integers)")
        if j_lo * self.W < K_FLOOR:
            raise ValueError("engines refuse to run below K_FLOOR")
        recs = self._sweep(j_lo, j_hi)
        return recs["v0"].astype(np.uint64)

    def survivors_j_deep(self, j_lo, j_hi, launch=None):
        """Sorted list of surviving j as Python ints, to J_REACH
        (ladder_gpu.py:783-804)."""
        j_lo, j_hi = int(j_lo), int(j_hi)
        if j_hi > self.J_REACH:
            raise ValueError(f"j {j_hi} past the enforced reach "
                             f"{self.J_REACH} (= fold {self.P} x J_CEIL)")
        if j_lo * self.W < K_FLOOR:
            raise ValueError("engines refuse to run below K_FLOOR")
        return N.record_values(self._sweep(j_lo, j_hi))

    def survivors_pre_mr(self, k_lo, k_hi, launch=None):
        """The same stream on the k line, both ends inclusive."""
        j_lo = max(1, -(-int(k_lo) // self.W))
        j_hi = int(k_hi) // self.W + 1
        return self.survivors_j(j_lo, j_hi, launch)

    def run_length(self, k, cap=64):
        """Full-bases strong-probable-prime run of k (ladder_gpu.py:813-817),
        on the device."""
        return self._eng.classify([int(k)], N.CLASSIFY_FULL, int(cap))[0]

    def run_lengths(self, ks, cap=64):
        return self._eng.classify(ks, N.CLASSIFY_FULL, int(cap))

    def sprp_runs(self, ks, cap, exact_from=SPRP_EXACT_FROM):
        """dickson-ladders/launch.py:310-342 sprp_run for many k at once."""
        return self._eng.classify(ks, N.CLASSIFY_TWO_PASS, int(cap),
                                  int(exact_from))

    def hunt(self, k_lo, k_hi, cap=None):
        """[(k, run)] for every survivor whose run reaches n
        (ladder_gpu.py:819-829); runs computed on the device."""
        cap = int(cap or self.n + 8)
        k_lo, k_hi = int(k_lo), int(k_hi)
        j_lo = max(1, -(-k_lo // self.W))
        j_hi = k_hi // self.W + 1
        look = min(cap, 64)
        recs = self._sweep(j_lo, j_hi, N.CLASSIFY_FULL, look)
        js = N.record_values(recs)
        runs = self._eng.extend_runs([j * self.W for j in js],
                                     recs["run"].tolist(), look, cap)
        out = []
        for j, r in zip(js, runs):
            k = j * self.W
            if k < k_lo or k > k_hi:
                continue
            if r >= self.n:
                out.append((k, int(r)))
        return out

    def sweep_classified(self, j_lo, j_hi, cap, two_pass=True,
                         exact_from=SPRP_EXACT_FROM):
        """Survivors and their runs in one device pass: [(j, run)] sorted."""
        mode = N.CLASSIFY_TWO_PASS if two_pass else N.CLASSIFY_FULL
        recs = self._sweep(j_lo, j_hi, mode, int(cap), int(exact_from))
        return list(zip(N.record_values(recs),
                        [int(r) for r in recs["run"].tolist()]))
