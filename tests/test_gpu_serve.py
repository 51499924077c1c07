"""The serving loop on the device: a seeded Poisson mix of requests of all
three families (64-, 128- and 192-bit values), served with continuous
batching over several slots, gives every request exactly the records
(values and classified runs) of the same window swept alone."""
import sys

import pytest

import fk_native as N
import fk_serve as S

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    reqs = S.poisson_requests(5, 40, rate=2000.0, log2_span=(24, 31))
    print("serving", len(reqs), "requests", flush=True)
    sys.stdout.flush()
    srv = S.Server(slots=3, budget=8 << 30)
    st = srv.run(reqs)
    assert st["completed"] == len(reqs)
    assert max(srv.submissions) > 1          # requests shared submissions
    assert st["p99_ms"] >= st["p50_ms"] > 0
    total = 0
    for r in reqs:
        eng = srv.engines[(r.family, r.n, r.q2)]
        alone = eng.sweep(r.lo, r.hi, *r.classify)
        assert N.record_values(r.records) == N.record_values(alone), r.rid
        assert r.records["run"].tolist() == alone["run"].tolist(), r.rid
        total += alone.size
    assert total >= 20
