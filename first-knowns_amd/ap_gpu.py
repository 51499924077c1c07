"""Drop-in for primorial-ap/ap_gpu.py's GpuEngine, on MI355X.

Same construction (GpuEngine(n, q2=65536, launch_u=2^27)), same sweep
(survivors(base, span) -> generator of ascending np.uint64 offsets
relative to base, base and span multiples of 30), same membership test
(survives), same classification (chain_depth: the two-pass chain capped at
n, primorial-ap/ap_search.py:160-194), the same cursor attributes (floor,
launch_u) and the same refusals.  Primes that divide P(n) still kill
residue 0 (ap_search.py:100-104): the engine treats them like any other
kill prime, and the mod-30 lanes are part of its wheel.
"""

import numpy as np

import fk_families as F
import fk_native as N

W0 = F.AP_W0
W0_RESIDUES = F.AP_W0_RESIDUES
P_FLOOR = F.AP_P_FLOOR
P_CEIL = F.AP_P_CEIL
Q2_DEFAULT = F.AP_Q2_DEFAULT
Q2_MAX = F.Q2_MAX
LAUNCH_U = F.AP_LAUNCH_U
DEPTH_EXACT_FROM = F.DEPTH_EXACT_FROM


def difference(n):
    return F.ap_difference(n)


class GpuEngine:
    """primorial-ap/ap_gpu.py:398 GpuEngine, rebuilt for gfx950."""

    def __init__(self, n, q2=Q2_DEFAULT, launch_u=LAUNCH_U, device=0,
                 **tuning):
        if not 7 <= q2 <= Q2_MAX:
            raise ValueError(f"q2 {q2} outside [7, {Q2_MAX}]")
        self.n = int(n)
        self.q2 = int(q2)
        self.d = difference(self.n)
        self.launch_u = int(launch_u)
        self.primes = [q for q in F.primes_below(self.q2) if q >= 7]
        self.floor = max(P_FLOOR, self.q2)
        t = F.tuning("ap")
        t.update(tuning)
        self._eng = N.Engine(N.FAMILY_AP, self.n, self.q2, device=device,
                             **t)
        self.info = self._eng.info

    def _check(self, base, span):
        base, span = int(base), int(span)
        if base % W0 or span % W0:
            raise ValueError(f"base and span must be multiples of {W0}")
        if base < self.floor:
            raise ValueError(f"the engine refuses to sieve below {self.floor}")
        if base + span > P_CEIL:
            raise ValueError(f"p {base + span} is past the enforced ceiling "
                             f"{P_CEIL}")
        return base, span

    def survivors(self, base, span, launch_u=None):
        """Yield uint64 arrays of offsets o with base + o surviving,
        ascending, one array per launch of launch_u wheel periods
        (ap_gpu.py:555-579)."""
        base, span = self._check(base, span)
        step_u = int(launch_u or self.launch_u)
        done = 0
        while done < span:
            u = min(step_u, (span - done) // W0)
            if u <= 0:
                break
            lo = base + done
            recs = self._eng.sweep(lo, lo + u * W0)
            if recs.size:
                vals = N.record_values(recs)
                yield np.array([v - base for v in vals], dtype=np.uint64)
            done += u * W0

    def sweep(self, base, span):
        """All survivor offsets of [base, base + span) in one device sweep."""
        base, span = self._check(base, span)
        recs = self._eng.sweep(base, base + span)
        return np.array([v - base for v in N.record_values(recs)],
                        dtype=np.uint64)

    def survives(self, p):
        """One-candidate membership: the sieve run on the window [p, p+1)."""
        p = int(p)
        if any(p % q == 0 for q in (2, 3, 5)):
            return False
        return self._eng.sweep(p, p + 1).size == 1

    def chain_depth(self, p, cap=None):
        """ap_search.py:160-194: base-2 chain, redone with all seven bases
        from DEPTH_EXACT_FROM, capped at n -- on the device."""
        cap = self.n if cap is None else int(cap)
        return self._eng.classify([int(p)], N.CLASSIFY_TWO_PASS, cap,
                                  DEPTH_EXACT_FROM)[0]

    def chain_depths(self, ps, cap=None):
        cap = self.n if cap is None else int(cap)
        return self._eng.classify(ps, N.CLASSIFY_TWO_PASS, cap,
                                  DEPTH_EXACT_FROM)

    def hunt(self, base, span, min_depth=None):
        """[(p, depth)] for every survivor whose chain reaches min_depth."""
        want = self.n if min_depth is None else int(min_depth)
        base, span = self._check(base, span)
        recs = self._eng.sweep(base, base + span, N.CLASSIFY_TWO_PASS,
                               self.n, DEPTH_EXACT_FROM)
        return [(p, int(r)) for p, r in zip(N.record_values(recs),
                                            recs["run"].tolist())
                if r >= want]
