"""The campaign segment loop over the real engine: a campaign stopped after
some segments and resumed from its cursor writes exactly the records of one
sweep of the whole range (values and classified runs), which are the
oracle's survivors and runs, and a cursor under another configuration is
refused."""
import sys

import pytest

import fk_campaign as FC
import fk_checkpoint as C
import fk_families as F
import fk_native as N

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_dickson_campaign_resumes_to_one_sweep(tmp_path):
    n, q2 = 10, 1 << 16
    lo, seg = 10**12, 1 << 27
    hi = lo + 5 * seg + 12345
    spec = FC.default_classify(N.FAMILY_DICKSON, n)
    say("campaign dickson")
    eng = N.Engine(N.FAMILY_DICKSON, n, q2, **F.tuning("dickson"))
    whole = eng.sweep(lo, hi, *spec)
    want = list(zip(N.record_values(whole),
                    [int(r) for r in whole["run"].tolist()]))
    assert len(want) >= 100

    path = tmp_path / "dickson.json"
    stop_at = lo + 3 * seg

    class Stop(Exception):
        pass

    def stop(s0, s1, values, runs, state):
        if s1 == stop_at:
            raise Stop

    c = FC.Campaign(eng, lo, hi, seg, path, classify=spec)
    with pytest.raises(Stop):
        c.run(on_segment=stop)
    assert c.cursor == stop_at

    c2 = FC.Campaign(eng, lo, hi, seg, path, classify=spec)
    assert c2.cursor == stop_at
    st = c2.run()
    assert c2.records() == want
    # and both are the oracle's survivors and base-2-then-full runs
    import fk_oracle as O
    js = O.dickson_survivors(n, lo, hi, q2=q2)
    W = O.wheel_modulus(n)
    assert c2.records() == [(j, O.dickson_sprp_run(j * W, spec[1],
                                                   exact_from=spec[2]))
                            for j in js]
    assert int(st["next"]) == hi and st["survivors"] == len(want)
    assert st["best_run"] == max(r for _, r in want)

    other = N.Engine(N.FAMILY_DICKSON, n + 1, q2, **F.tuning("dickson"))
    with pytest.raises(C.CursorRefused):
        FC.Campaign(other, lo, hi, seg, path)


def test_ap_campaign_matches_one_sweep(tmp_path):
    n, q2 = 12, 1 << 16
    lo = 4 * 10**13 - 10
    seg = 30 * (1 << 22)
    hi = lo + 4 * seg
    spec = FC.default_classify(N.FAMILY_AP, n)
    say("campaign ap")
    eng = N.Engine(N.FAMILY_AP, n, q2, **F.tuning("ap"))
    whole = eng.sweep(lo, hi, *spec)
    want = list(zip(N.record_values(whole),
                    [int(r) for r in whole["run"].tolist()]))
    assert len(want) >= 20
    c = FC.Campaign(eng, lo, hi, seg, tmp_path / "ap.json", classify=spec)
    c.run()
    assert c.records() == want
    import fk_oracle as O
    ps = [lo + o for o in O.ap_survivors(n, lo, hi - lo, q2=q2)]
    assert c.records() == [(p, O.ap_chain_depth(p, n)) for p in ps]
