"""Operator output of the campaign loop, on the CPU: the log-line and
census formats of huntlib/hlog.py, the evidence ledger's guarantees of
huntlib/evidence.py (a redone segment rewrites its record, the ledger stays
sorted by the found value, a torn ledger is rebuilt from the evidence
files), the wall-clock heartbeat, and the campaign writing evidence before
its cursor."""
import io
import json
import os
import re
import time

import pytest

import fk_campaign as FC
import fk_checkpoint as C
import fk_evidence as EV
import fk_log as LG

from test_campaign_host import HI, LO, SEG, StubBatch, StubEngine, fake_run, whole


def test_line_and_census_formats():
    s = LG.line("STATUS", "at 12 census 7:1", now=0)
    assert re.fullmatch(r"\[\d\d:\d\d:\d\d\] \[STATUS\] at 12 census 7:1", s)
    assert LG.census_str({"7": 280, 8: 71, "10": 8}, 7, 10) == \
        "census 7:280 8:71 9:0 10:8"
    assert LG.census_str({}, 5, 4) == "census -"
    out = io.StringIO()
    LG.log("WARN", "x", stream=out)
    assert out.getvalue().endswith("[WARN] x\n")


def test_ledger_upserts_and_sorts(tmp_path):
    d = tmp_path / "evidence"
    ledger = d / "ledger.json"
    EV.record({"value": 12345, "run": 10, "note": "first"}, d,
              "hit_12345.json", ledger, label="DISCOVERY", now=1.0)
    EV.record({"value": 12345, "run": 10, "note": "redone"}, d,
              "hit_12345.json", ledger, label="DISCOVERY", now=2.0)
    EV.record({"value": 999, "run": 9}, d, "hit_999.json", ledger, now=3.0)
    rows = EV.load_ledger(ledger)
    assert [r["value"] for r in rows] == [999, 12345]
    assert rows[1]["note"] == "redone" and rows[1]["label"] == "DISCOVERY"
    assert json.loads((d / "hit_12345.json").read_text())["note"] == "redone"
    assert EV.load_ledger(tmp_path / "none.json") == []


def test_torn_ledger_is_rebuilt_from_evidence(tmp_path):
    d = tmp_path / "evidence"
    ledger = d / "ledger.json"
    for v in (30, 10, 20):
        EV.record({"value": v, "run": 11}, d, f"hit_{v}.json", ledger)
    ledger.write_bytes(b"\0" * 64)          # torn write
    EV.record({"value": 15, "run": 12}, d, "hit_15.json", ledger)
    assert [r["value"] for r in EV.load_ledger(ledger)] == [10, 15, 20, 30]


def test_heartbeat_rate_and_stall():
    t = [100.0]
    hb = LG.Heartbeat(interval=30.0, clock=lambda: t[0])
    hb.mark(0)
    t[0] = 110.0
    hb.mark(1000)
    t[0] = 130.0
    assert hb.rate() == pytest.approx(1000 / 30.0)   # stall time included
    lines = []
    hb.beat(lambda: "ok", emit=lambda tag, m: lines.append((tag, m)))
    hb.doing("verifying")
    t[0] = 145.0
    assert hb.stalled() == ("verifying", 15.0)
    hb.beat(lambda: 1 / 0, emit=lambda tag, m: lines.append((tag, m)))
    assert lines[0] == ("STATUS", "ok") and lines[1][0] == "WARN"


def test_heartbeat_thread_emits_on_the_clock():
    lines = []
    hb = LG.Heartbeat(interval=0.02)
    hb.start(lambda: "tick", emit=lambda tag, m: lines.append(tag))
    time.sleep(0.2)
    hb.stop()
    assert lines.count("STATUS") >= 3


def test_campaign_writes_evidence_before_its_cursor(tmp_path, monkeypatch):
    ev_dir = tmp_path / "evidence"
    logs = []
    kw = dict(evidence_dir=ev_dir, evidence_min_run=16,
              log=lambda tag, m: logs.append((tag, m)), heartbeat=0.01)
    c = FC.Campaign(StubEngine(), LO, HI, SEG, tmp_path / "cur.json",
                    batches=[StubBatch(), StubBatch()], **kw)
    c.run(max_segments=2)
    # a crash after a segment's evidence, before its cursor: the redone
    # segment rewrites the same records
    real = C.save_json
    state = {"cursor_saves": 0}

    def flaky(path, obj):
        if os.fspath(path).endswith("cur.json"):
            state["cursor_saves"] += 1
            if state["cursor_saves"] == 1:
                raise OSError("power lost")
        real(path, obj)

    monkeypatch.setattr(C, "save_json", flaky)
    with pytest.raises(OSError):
        FC.Campaign(StubEngine(), LO, HI, SEG, tmp_path / "cur.json",
                    batches=[StubBatch(), StubBatch()], **kw).run()
    monkeypatch.setattr(C, "save_json", real)
    c3 = FC.Campaign(StubEngine(), LO, HI, SEG, tmp_path / "cur.json",
                     batches=[StubBatch(), StubBatch()], **kw)
    c3.run()
    want = sorted(v for v, r in whole() if r >= 16)
    rows = EV.load_ledger(ev_dir / "ledger.json")
    assert [r["value"] for r in rows] == want
    assert all(r["run"] == fake_run(r["value"]) for r in rows)
    assert sorted(f for f in os.listdir(ev_dir) if f.startswith("hit_")
                  and f.endswith(".json")) == sorted(f"hit_{v}.json" for v in want)
    assert any(t == "DISCOVERY" for t, _ in logs)
    line = c3.status_line()
    assert "census" in line and f"at {HI}" in line


# ------------------------------------------- verification in the loop

with open(os.path.join(os.path.dirname(__file__), "golden",
                       "reference_vectors.json")) as _f:
    GOLD = json.load(_f)
D10 = GOLD["evidence"]["dickson_full"][0]           # run 10, n = 10


class OneRecordBatch:
    """Returns one record, (j, run), for the window that holds j."""

    def __init__(self, j, run):
        self.j, self.run = j, run

    def submit(self, engines, engine_of, windows, specs):
        (lo, hi), = windows
        self.emit = lo <= self.j < hi

    def results(self):
        import numpy as np
        import fk_native as N
        recs = np.zeros(1 if self.emit else 0, dtype=N.RECORD_DTYPE)
        if self.emit:
            recs["v0"] = self.j
            recs["run"] = self.run
        return recs, np.array([recs.size], dtype=np.uint64)


def _verified_campaign(tmp_path, run, **kw):
    import fk_verify as VF
    j = int(D10["k"]) // VF.dickson_wheel(10)
    logs = []
    c = FC.Campaign(StubEngine(), j, j + 10, 5, tmp_path / "cur.json",
                    batches=[OneRecordBatch(j, run), OneRecordBatch(j, run)],
                    log=lambda t, m: logs.append((t, m)),
                    verify=True, **kw)
    return c, j, logs


def test_campaign_verifies_before_writing_evidence(tmp_path):
    c, j, logs = _verified_campaign(tmp_path, D10["run"],
                                    evidence_dir=tmp_path / "ev",
                                    evidence_min_run=10)
    c.run()
    rows = EV.load_ledger(tmp_path / "ev" / "ledger.json")
    assert [r["value"] for r in rows] == [j]
    v = rows[0]["verification"]
    assert str(v["k"]) == D10["k"] and v["run"] == D10["run"]
    assert v["certificates_bls75"] == D10["certificates_bls75"]
    assert v["breaker_factor"] == D10["breaker_factor"]
    assert any(t == "DISCOVERY" and "verified" in m for t, m in logs)


def test_campaign_alarm_stops_before_the_cursor(tmp_path):
    c, j, logs = _verified_campaign(tmp_path, D10["run"] + 1,
                                    evidence_dir=tmp_path / "ev",
                                    evidence_min_run=10)
    with pytest.raises(FC.CorruptEngineError):
        c.run()
    assert c.cursor == j                              # nothing committed
    assert any(t == "ALARM" for t, _ in logs)
    assert EV.load_ledger(tmp_path / "ev" / "ledger.json") == []


def test_campaign_near_is_verified_not_evidenced(tmp_path):
    c, j, logs = _verified_campaign(tmp_path, D10["run"],
                                    evidence_dir=tmp_path / "ev",
                                    evidence_min_run=11, near_min_run=9)
    c.run()
    assert any(t == "NEAR" and f"value={j}" in m for t, m in logs)
    assert EV.load_ledger(tmp_path / "ev" / "ledger.json") == []
