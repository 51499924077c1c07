"""The campaign segment loop and its cursor rules, on the CPU.

fk_checkpoint restates huntlib/checkpoint.py's rules: write durably, keep
a .bak, treat corrupt as an error rather than as absent, and refuse a
cursor under another key. The campaign loop (fk_campaign) is driven here by
stand-in batches that return the oracle's survivors of each segment, with
runs given by a fixed function of the value. That exercises the loop's
cursor, ledger and resume logic without a device. The same loop over the
real engine is tested on the GPU in test_gpu_campaign.py."""
import json
import os

import numpy as np
import pytest

import fk_campaign as FC
import fk_checkpoint as C
import fk_native as N
import fk_oracle as O


# ------------------------------------------------------------ checkpoint

def test_save_keeps_backup_and_loads(tmp_path):
    p = tmp_path / "c.json"
    C.save_json(p, {"key": "k", "next": 1})
    C.save_json(p, {"key": "k", "next": 2})
    assert C.load(p, "k")["next"] == 2
    assert json.loads((tmp_path / "c.json.bak").read_text())["next"] == 1
    assert not (tmp_path / "c.json.tmp").exists()


def test_absent_is_none_and_key_mismatch_is_ignored(tmp_path):
    p = tmp_path / "c.json"
    assert C.load(p, "k") is None
    C.save_json(p, {"key": "k", "next": 1})
    msgs = []
    assert C.load(p, "other", warn=msgs.append) is None
    assert msgs


def test_corrupt_falls_back_to_backup(tmp_path):
    p = tmp_path / "c.json"
    C.save_json(p, {"key": "k", "next": 1})
    C.save_json(p, {"key": "k", "next": 2})
    p.write_bytes(b"\0" * 40)            # a torn write: right size, all NUL
    msgs = []
    assert C.load(p, "k", warn=msgs.append)["next"] == 1
    assert msgs


def test_corrupt_without_backup_is_an_error(tmp_path):
    p = tmp_path / "c.json"
    p.write_bytes(b"\0" * 40)
    with pytest.raises(C.CheckpointCorrupt):
        C.load(p, "k")
    with pytest.raises(C.CheckpointCorrupt):
        C.refuse_mismatch(p, "k")
    C.refuse_mismatch(p, "k", fresh=True)


def test_refuse_mismatch(tmp_path):
    p = tmp_path / "c.json"
    C.refuse_mismatch(p, "k")            # nothing there: fine
    C.save_json(p, {"key": "k", "next": 5})
    C.refuse_mismatch(p, "k")
    with pytest.raises(C.CursorRefused):
        C.refuse_mismatch(p, "other")
    C.refuse_mismatch(p, "other", fresh=True)


# ---------------------------------------------------------- the loop

N_D, Q2_D = 10, 4096


class StubEngine:
    family, n, q2 = N.FAMILY_DICKSON, N_D, Q2_D
    line_step, nlines = 96577, 3927


def fake_run(j):
    return 10 + (j * 2654435761) % 7


class StubBatch:
    """Stands in for fk_native.Batch: the oracle's survivors of the
    submitted window, with runs from fake_run."""

    def __init__(self, fail_on=None, log=None):
        self.fail_on = fail_on
        self.log = log if log is not None else []

    def submit(self, engines, engine_of, windows, specs):
        (self.lo, self.hi), = windows
        self.log.append((self.lo, self.hi))

    def results(self):
        if self.fail_on is not None and self.lo == self.fail_on:
            raise RuntimeError("device lost")
        js = O.dickson_survivors(N_D, self.lo, self.hi, q2=Q2_D)
        recs = np.zeros(len(js), dtype=N.RECORD_DTYPE)
        recs["v0"] = np.array(js, dtype=np.uint64)
        recs["run"] = [fake_run(j) for j in js]
        return recs, np.array([len(js)], dtype=np.uint64)


LO, SEG = 10**9, 1 << 18
HI = LO + 5 * SEG + 12345


def whole():
    js = O.dickson_survivors(N_D, LO, HI, q2=Q2_D)
    return [(j, fake_run(j)) for j in js]


def campaign(tmp_path, **kw):
    b = kw.pop("batches", None) or [StubBatch(), StubBatch()]
    return FC.Campaign(StubEngine(), LO, HI, SEG, tmp_path / "cur.json",
                       batches=b, **kw)


def test_campaign_matches_one_sweep(tmp_path):
    c = campaign(tmp_path)
    st = c.run()
    assert c.records() == whole()
    assert int(st["next"]) == HI and st["segments"] == 6
    assert st["survivors"] == len(whole())
    assert int(st["swept"]) == HI - LO
    assert sum(st["census"].values()) == st["survivors"]


def test_resume_after_device_error_redoes_nothing_twice(tmp_path):
    fail = LO + 3 * SEG
    c = campaign(tmp_path, batches=[StubBatch(fail), StubBatch(fail)])
    with pytest.raises(RuntimeError):
        c.run()
    assert c.cursor == fail               # last completed boundary
    c2 = campaign(tmp_path)
    assert c2.cursor == fail
    c2.run()
    assert c2.records() == whole()


def test_crash_between_ledger_and_cursor(tmp_path, monkeypatch):
    c = campaign(tmp_path)
    c.run(max_segments=2)
    real = C.save_json
    calls = {"n": 0}

    def flaky(path, obj):
        calls["n"] += 1
        if calls["n"] == 2:
            raise OSError("power lost")
        real(path, obj)

    monkeypatch.setattr(C, "save_json", flaky)
    with pytest.raises(OSError):
        campaign(tmp_path).run()
    monkeypatch.setattr(C, "save_json", real)
    # the ledger holds one segment past the cursor; resume truncates it
    c3 = campaign(tmp_path)
    assert c3.cursor == LO + 3 * SEG
    c3.run()
    assert c3.records() == whole()


def test_key_mismatch_is_refused(tmp_path):
    campaign(tmp_path).run(max_segments=1)

    class Other(StubEngine):
        n = 11
    with pytest.raises(C.CursorRefused):
        FC.Campaign(Other(), LO, HI, SEG, tmp_path / "cur.json",
                    batches=[StubBatch(), StubBatch()])
    c = FC.Campaign(Other(), LO, HI, SEG, tmp_path / "cur.json",
                    batches=[StubBatch(), StubBatch()], fresh=True)
    assert c.cursor == LO and c.records() == []


def test_lost_ledger_is_an_error(tmp_path):
    c = campaign(tmp_path)
    c.run(max_segments=2)
    os.remove(c.ledger)
    with pytest.raises(C.CheckpointCorrupt):
        campaign(tmp_path)


def test_segments_alternate_batches(tmp_path):
    log = []
    b = [StubBatch(log=log), StubBatch(log=log)]
    campaign(tmp_path, batches=b).run()
    assert log[0] == (LO, LO + SEG) and log[-1][1] == HI
    assert all(a[1] == b_[0] for a, b_ in zip(log, log[1:]))


def test_crash_inside_save_keeps_the_cursor(tmp_path, monkeypatch):
    """A crash after the backup is refreshed and before the new state is
    renamed into place leaves the old cursor readable under its own name:
    the resume redoes one segment and loses no record."""
    c = campaign(tmp_path)
    c.run(max_segments=2)
    real = os.replace

    def dies_on_main_rename(src, dst):
        if str(src).endswith("cur.json.tmp"):
            raise OSError("power lost")
        real(src, dst)

    monkeypatch.setattr(os, "replace", dies_on_main_rename)
    with pytest.raises(OSError):
        campaign(tmp_path).run()
    monkeypatch.setattr(os, "replace", real)
    assert (tmp_path / "cur.json").exists()
    c3 = campaign(tmp_path)
    assert c3.cursor == LO + 2 * SEG
    c3.run()
    assert c3.records() == whole()


def test_missing_cursor_resumes_from_backup(tmp_path):
    c = campaign(tmp_path)
    c.run(max_segments=3)
    os.remove(c.path)
    msgs = []
    c2 = campaign(tmp_path, warn=msgs.append)
    assert c2.cursor == LO + 2 * SEG      # the backup: one segment behind
    assert msgs and "redone" in msgs[0]
    c2.run()
    assert c2.records() == whole()


def test_records_without_cursor_are_not_discarded(tmp_path):
    c = campaign(tmp_path)
    c.run(max_segments=2)
    size = os.path.getsize(c.ledger)
    assert size > 0
    os.remove(c.path)
    os.remove(c.path + ".bak")
    with pytest.raises(C.CheckpointCorrupt):
        campaign(tmp_path)
    assert os.path.getsize(c.ledger) == size
    c2 = campaign(tmp_path, fresh=True)
    assert c2.cursor == LO and c2.records() == []


def test_backup_is_a_separate_file(tmp_path):
    """Rewriting the main file in place never changes the backup."""
    p = tmp_path / "c.json"
    C.save_json(p, {"key": "k", "next": 1})
    C.save_json(p, {"key": "k", "next": 2})
    C.save_json(p, {"key": "k", "next": 3})
    p.write_text("{")
    assert C.load(p, "k")["next"] == 2
    state, source, problems = C.read_state(p)
    assert source.endswith(".bak") and len(problems) == 1
