"""Campaign segment loop over the MI355X engine.

A campaign sweeps a range of one family's search line in segments. Every
survivor is classified on the device, inside the sweep. Each segment's
records are appended to a durable ledger, and the cursor is persisted at
every segment boundary.

The reference loops are euler-prime-runs/launch.py:648-781,
dickson-ladders/launch.py:1072-1117 and primorial-ap/launch.py:583-642
(`_finalize` :686-770). They keep one segment in flight on the GPU while a
host process pool classifies the previous one. Here the strong tests run on
the device, so the overlap is different: two batches alternate, and the
device sweeps segment i+1 while the host persists segment i.

Operator output follows huntlib's formats: [STATUS] lines with the census
of run lengths from a wall-clock heartbeat thread (fk_log), and an evidence
directory holding one durable file per find at or above a chosen run
length, with an idempotent ledger (fk_evidence).  Evidence for a segment is
written before the cursor that covers it, so a redone segment rewrites its
records rather than duplicating them.

Durability follows huntlib/checkpoint.py (see fk_checkpoint):
* A segment's records are appended to the ledger and fsynced before the
  cursor that covers them is saved.
* The cursor stores the ledger's byte length at its boundary, and a resume
  truncates the ledger to that length. A crash between the two writes
  therefore redoes that one segment and never duplicates its records.
* The cursor's key names what its position is counted in: family, depth,
  sieve bound, wheel and classification. A cursor under another key is
  refused, not reinterpreted.
"""

import json
import os

import fk_checkpoint as C
import fk_evidence as EV
import fk_log as LG
import fk_native as N
import fk_verify as VF

FAMILY_NAME = {N.FAMILY_EULER: "euler", N.FAMILY_DICKSON: "dickson",
               N.FAMILY_AP: "ap"}


class CorruptEngineError(RuntimeError):
    """A record the verification legs do not confirm.  The campaign stops
    before the segment's cursor is written, as the reference's launchers
    stop on an [ALARM] (dickson-ladders/launch.py:218, 923-925)."""


def default_classify(family, n):
    """(mode, cap, exact_from) the reference's launchers record. Dickson
    uses base-2 chains redone with every base from run 5
    (dickson-ladders/launch.py:307-342). AP uses two-pass depth from 6
    (primorial-ap/ap_search.py:79-86, 160-194). Euler uses full-base runs
    (euler-prime-runs/euler_search.py:258-262)."""
    if family == N.FAMILY_DICKSON:
        return (N.CLASSIFY_TWO_PASS, n + 8, 5)
    if family == N.FAMILY_AP:
        return (N.CLASSIFY_TWO_PASS, n, 6)
    return (N.CLASSIFY_FULL, n + 8, 0)


def engine_key(engine, classify):
    """The parameters that give a cursor position its meaning."""
    return {"family": FAMILY_NAME.get(engine.family, str(engine.family)),
            "n": int(engine.n), "q2": int(engine.q2),
            "line_step": int(engine.line_step),
            "nlines": int(engine.nlines),
            "classify": [int(x) for x in classify]}


class Campaign:
    """Sweep [lo, hi) of `engine`'s line in segments of `seg`, with the
    cursor at `path` and the records in `ledger` (JSON lines of
    {"value", "run"}; default `path` + ".records.jsonl").

    `batches` is for tests: two objects with the fk_native.Batch interface.
    `evidence_dir` with `evidence_min_run`: every record with run >= that
    gets an evidence file and a ledger row there.  With `verify`, each such
    record first passes the verification legs with certificates and a
    breaker witness (fk_verify), whose output goes into the evidence, and
    each record with near_min_run <= run < evidence_min_run passes the
    three legs alone and is logged as [NEAR]; a record the legs do not
    confirm raises CorruptEngineError.  `log` (tag, message) is where
    STATUS / DISCOVERY / NEAR / ALARM lines go (fk_log.log by default);
    `heartbeat` is the [STATUS] interval in seconds (None: no heartbeat
    thread).
    """

    def __init__(self, engine, lo, hi, seg, path, ledger=None,
                 classify=None, fresh=False, batches=None, warn=None,
                 evidence_dir=None, evidence_min_run=None, log=None,
                 heartbeat=None, verify=False, near_min_run=None):
        lo, hi, seg = int(lo), int(hi), int(seg)
        if not lo < hi:
            raise ValueError(f"empty range [{lo}, {hi})")
        if seg <= 0:
            raise ValueError(f"segment length {seg} must be positive")
        self.engine = engine
        self.lo, self.hi, self.seg = lo, hi, seg
        self.path = os.fspath(path)
        self.ledger = os.fspath(ledger) if ledger else self.path + ".records.jsonl"
        self.classify = tuple(classify or default_classify(engine.family,
                                                           engine.n))
        self.key = engine_key(engine, self.classify)
        self.warn = warn
        self.evidence_dir = os.fspath(evidence_dir) if evidence_dir else None
        self.evidence_min_run = evidence_min_run
        self.verify = bool(verify)
        self.near_min_run = near_min_run
        self.log = log or LG.log
        self.heartbeat = heartbeat
        self._hb = None
        C.refuse_mismatch(self.path, self.key, fresh=fresh,
                          describe=lambda s: f"cursor at {s.get('next')}")
        state = None if fresh else C.load(self.path, self.key, warn=warn)
        if state is None:
            # a ledger with records and no cursor covering them is a lost
            # cursor, not a new campaign: only `fresh` may empty it
            if (not fresh and os.path.exists(self.ledger) and
                    os.path.getsize(self.ledger) > 0):
                raise C.CheckpointCorrupt(
                    f"{self.ledger} holds records but no cursor under this "
                    f"key covers them ({self.path}); nothing was changed. "
                    f"Restore the cursor, or pass fresh=True to discard "
                    f"the records on purpose")
            state = {"key": self.key, "next": str(lo), "segments": 0,
                     "survivors": 0, "swept": "0", "census": {},
                     "best_run": -1, "best_value": None, "ledger_bytes": 0}
            self._truncate_ledger(0)
        else:
            self._truncate_ledger(int(state["ledger_bytes"]))
        self.state = state
        # segments of one length repeat one launch sequence: the batches
        # run it as a captured HIP graph after the first segment; the two
        # run side by side, each on a hardware queue of its own
        self.batches = batches if batches is not None else [
            N.Batch(getattr(engine, "device", 0), graph=True, dedicated=True,
                    throughput=True) for _ in range(2)]

    # ------------------------------------------------------------ ledger

    def _truncate_ledger(self, nbytes):
        if nbytes == 0:
            with open(self.ledger, "w"):
                pass
            return
        if not os.path.exists(self.ledger) or \
                os.path.getsize(self.ledger) < nbytes:
            raise C.CheckpointCorrupt(
                f"{self.ledger} is shorter than the {nbytes} bytes its cursor "
                f"covers: records below the cursor were lost")
        with open(self.ledger, "r+") as f:
            f.truncate(nbytes)
            f.flush()
            os.fsync(f.fileno())

    def _append(self, values, runs):
        with open(self.ledger, "a", newline="\n") as f:
            for v, r in zip(values, runs):
                f.write(json.dumps({"value": int(v), "run": int(r)}) + "\n")
            f.flush()
            os.fsync(f.fileno())
        return os.path.getsize(self.ledger)

    # -------------------------------------------------------------- loop

    @property
    def cursor(self):
        return int(self.state["next"])

    def _legs(self, v, r, full):
        """The verification legs on one record (certificates and witness
        when `full`); raises CorruptEngineError when they do not confirm."""
        if self._hb is not None:
            self._hb.doing(f"verifying run {r} at {v}")
        ok, res = VF.verify_record(self.engine.family, self.key["n"], v, r,
                                   self.key["q2"], certify=full, witness=full)
        if not ok:
            why = res if isinstance(res, str) else res.get(
                "error", f"legs {res.get('legs')}")
            msg = f"verify failed at value={v} run={r}: {why}"
            self.log("ALARM", msg)
            raise CorruptEngineError(msg)
        return res

    def _evidence(self, s0, s1, values, runs):
        ev_min = (self.evidence_min_run if self.evidence_dir else None)
        near = self.near_min_run if self.verify else None
        if ev_min is None and near is None:
            return
        ledger = (os.path.join(self.evidence_dir, "ledger.json")
                  if self.evidence_dir else None)
        for v, r in zip(values, runs):
            if ev_min is not None and r >= ev_min:
                ev = {"family": self.key["family"], "n": self.key["n"],
                      "value": int(v), "run": int(r),
                      "segment": [str(s0), str(s1)],
                      "classify": self.key["classify"]}
                if self.verify:
                    ev["verification"] = self._legs(int(v), int(r), True)
                EV.record(ev, self.evidence_dir, f"hit_{int(v)}.json", ledger,
                          key="value", label="DISCOVERY")
                self.log("DISCOVERY", f"{self.key['family']} n={self.key['n']} "
                                      f"value={int(v)} run={int(r)}"
                                      + (" (verified)" if self.verify else ""))
            elif near is not None and r >= near:
                self._legs(int(v), int(r), False)
                self.log("NEAR", f"{self.key['family']} n={self.key['n']} "
                                 f"value={int(v)} run={int(r)} (three legs "
                                 f"agree)")

    def status_line(self):
        """The [STATUS] text: cursor, swept, segments, survivors, best run,
        rate (position units per second, from the heartbeat) and census."""
        st = self.state
        census = st["census"]
        lo_r = min((int(k) for k in census), default=0)
        hi_r = max((int(k) for k in census), default=-1)
        msg = (f"{self.key['family']} n={self.key['n']} at {st['next']} "
               f"swept {st['swept']} segments {st['segments']} "
               f"survivors {st['survivors']} best {st['best_run']}")
        if self._hb is not None:
            msg += f" rate {self._hb.rate():.4g}/s"
            stall = self._hb.stalled()
            if stall:
                msg += f" (busy {stall[0]} for {stall[1]:.0f} s)"
        return msg + " " + LG.census_str(census, lo_r, hi_r)

    def _consume(self, s0, s1, batch, on_segment):
        recs, _ = batch.results()
        values = N.record_values(recs)
        runs = [int(r) for r in recs["run"].tolist()]
        self._evidence(s0, s1, values, runs)
        size = self._append(values, runs)
        st = dict(self.state)
        census = dict(st["census"])
        best_run, best_value = st["best_run"], st["best_value"]
        for v, r in zip(values, runs):
            census[str(r)] = census.get(str(r), 0) + 1
            if r > best_run:
                best_run, best_value = r, str(v)
        st.update(next=str(s1), segments=st["segments"] + 1,
                  survivors=st["survivors"] + len(values),
                  swept=str(int(st["swept"]) + (s1 - s0)), census=census,
                  best_run=best_run, best_value=best_value,
                  ledger_bytes=size)
        C.save_json(self.path, st)
        self.state = st
        if self._hb is not None:
            self._hb.mark(s1 - self.lo)
            self._hb.doing(f"segment at {s1}")
        if on_segment:
            on_segment(s0, s1, values, runs, st)

    def run(self, max_segments=None, on_segment=None):
        """Sweep from the cursor to hi (or max_segments more segments).
        Returns the final cursor state. An interrupt leaves the cursor at
        the last completed segment boundary; so does any other error, and
        then nothing more is written."""
        if self.heartbeat:
            self._hb = LG.Heartbeat(self.heartbeat)
            self._hb.mark(self.cursor - self.lo)
            self._hb.start(self.status_line, self.log)
        try:
            return self._run(max_segments, on_segment)
        finally:
            if self._hb is not None:
                self._hb.stop()
                self._hb = None

    def _run(self, max_segments, on_segment):
        eng = self.engine
        cur = self.cursor
        pending, slot, submitted = None, 0, 0
        while True:
            new = None
            if cur < self.hi and (max_segments is None or
                                  submitted < max_segments):
                s1 = min(cur + self.seg, self.hi)
                b = self.batches[slot]
                slot ^= 1
                b.submit([eng], [0], [(cur, s1)], [self.classify])
                new = (cur, s1, b)
                cur = s1
                submitted += 1
            if pending is not None:
                self._consume(*pending, on_segment)
            elif new is None:
                break
            pending = new
        return self.state

    def records(self):
        """Every (value, run) in the ledger, in sweep order."""
        out = []
        with open(self.ledger) as f:
            for line in f:
                if line.strip():
                    r = json.loads(line)
                    out.append((int(r["value"]), int(r["run"])))
        return out
