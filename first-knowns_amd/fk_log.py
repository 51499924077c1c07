"""Operator log lines, the census token string and the wall-clock status
heartbeat, in the formats the reference launchers print
(huntlib/hlog.py:44-71 and :74-175), so a campaign on this engine reads
the same to an operator:

* every line is ``[HH:MM:SS] [TAG] message``, plain ASCII, one writer at a
  time (tags: STAGE, STATUS, MILESTONE, NEAR, DISCOVERY, WARN, ALARM);
* the census of run lengths lo..hi is one token per length,
  ``census 7:280 8:71 9:28 10:8`` (absent lengths print as 0);
* [STATUS] comes from a timer thread on the wall clock, not from the
  segment loop, so a long segment never looks like a hang; its rate is
  position per second over the last interval *including* the time since the
  last segment boundary, and a line with no boundary since the previous one
  says what the loop is busy with and for how long.
"""

import collections
import sys
import threading
import time

_lock = threading.Lock()


def line(tag, msg, now=None):
    """The text of one log line (no newline)."""
    t = time.localtime(time.time() if now is None else now)
    return f"[{time.strftime('%H:%M:%S', t)}] [{tag}] {msg}"


def log(tag, msg, stream=None):
    """Write one line; lines from different threads never interleave."""
    s = line(tag, msg)
    with _lock:
        out = stream if stream is not None else sys.stdout
        out.write(s + "\n")
        out.flush()
    return s


def census_str(counts, lo, hi):
    """``census r:n ...`` for run lengths lo..hi; `counts` maps a length
    (int or its string) to a count."""
    counts = counts or {}
    toks = [f"{r}:{int(counts.get(str(r), counts.get(r, 0)))}"
            for r in range(int(lo), int(hi) + 1)]
    return "census " + (" ".join(toks) if toks else "-")


class Heartbeat:
    """A daemon thread that emits ``line_fn()`` as [STATUS] every `interval`
    seconds of wall clock.  The main loop calls mark(position) at segment
    boundaries and doing(what) when it starts something long; line_fn reads
    rate() and stalled().  An exception in line_fn is reported as [WARN]
    and the thread keeps running.  `clock` is for tests."""

    def __init__(self, interval=30.0, clock=time.time):
        self.interval = float(interval)
        self.clock = clock
        self._marks = collections.deque(maxlen=4096)      # (t, position)
        self._what, self._what_t = "", clock()
        self._marks_at_line = 0
        self._nmarks = 0
        self._stop = threading.Event()
        self._thread = None

    # ---- main loop side
    def mark(self, pos):
        self._marks.append((self.clock(), pos))
        self._nmarks += 1

    def doing(self, what):
        self._what, self._what_t = str(what), self.clock()

    # ---- line_fn side
    def position(self):
        return self._marks[-1][1] if self._marks else None

    def rate(self, now=None):
        """Position units per second: movement since the newest mark at
        least one interval old (else the oldest), divided by the time from
        that mark to now."""
        if len(self._marks) < 2:
            return 0.0
        now = self.clock() if now is None else now
        marks = list(self._marks)
        ref = marks[0]
        for m in reversed(marks[:-1]):
            if m[0] <= now - self.interval:
                ref = m
                break
        return (marks[-1][1] - ref[1]) / max(now - ref[0], 1e-9)

    def stalled(self, now=None):
        """(what, seconds) when no mark has landed since the last line."""
        if self._nmarks != self._marks_at_line:
            return None
        now = self.clock() if now is None else now
        return (self._what or "waiting", now - self._what_t)

    def beat(self, line_fn, emit=None):
        """One [STATUS] line now (the thread calls this every interval)."""
        emit = emit or log
        try:
            msg = line_fn()
        except Exception as exc:          # the heartbeat must not die
            emit("WARN", f"status line failed: {exc!r}")
        else:
            emit("STATUS", msg)
        self._marks_at_line = self._nmarks

    def start(self, line_fn, emit=None):
        if self._thread is not None:
            return
        self._stop.clear()

        def loop():
            while not self._stop.wait(self.interval):
                self.beat(line_fn, emit)

        self._thread = threading.Thread(target=loop, daemon=True,
                                        name="fk-heartbeat")
        self._thread.start()

    def stop(self):
        if self._thread is None:
            return
        self._stop.set()
        self._thread.join()
        self._thread = None
