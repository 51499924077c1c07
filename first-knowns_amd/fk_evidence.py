"""The evidence directory: one durable file per first occurrence and a
ledger of them, with the reference's two guarantees
(huntlib/evidence.py:1-57):

* idempotence -- the segment in flight at a crash is redone on resume, so
  the same find can be recorded twice; the ledger is keyed by the found
  value and upserted, so it stays a set sorted by that value;
* durability -- evidence files and the ledger are written through the
  checkpoint path (temporary file, fsync, replace, .bak kept).

A ledger that cannot be read is rebuilt from the evidence files beside it
(every file in the directory whose JSON carries the key), not trusted and
not silently emptied.
"""

import json
import os
import time

import fk_checkpoint as C


def _read_ledger(ledger):
    with open(ledger) as f:
        rows = json.load(f)
    if not isinstance(rows, list):
        raise ValueError("ledger is not a list")
    return rows


def _rebuild(dirname, ledger, key):
    rows = []
    for name in sorted(os.listdir(dirname)):
        path = os.path.join(dirname, name)
        if (not name.endswith(".json") or name.endswith(".bak")
                or os.path.abspath(path) == os.path.abspath(ledger)):
            continue
        try:
            with open(path) as f:
                ev = json.load(f)
        except (OSError, ValueError):
            continue
        if isinstance(ev, dict) and key in ev:
            rows.append(ev)
    return rows


def record(ev, dirname, filename, ledger, key="value", label=None, now=None):
    """Write `ev` (a dict of exact values; ev[key] identifies the find) to
    dirname/filename and upsert it into the ledger.  Returns the path."""
    os.makedirs(dirname, exist_ok=True)
    path = os.path.join(dirname, filename)
    C.save_json(path, ev)
    if os.path.exists(ledger):
        try:
            rows = _read_ledger(ledger)
        except (OSError, ValueError):
            rows = _rebuild(dirname, ledger, key)
    else:
        rows = []
    row = dict(ev)
    if label is not None:
        row["label"] = label
    row["t"] = time.time() if now is None else now
    kv = int(ev[key])
    rows = [r for r in rows if int(r.get(key, -1)) != kv]
    rows.append(row)
    rows.sort(key=lambda r: int(r.get(key, 0)))
    C.save_json(ledger, rows)
    return path


def load_ledger(ledger):
    """The ledger's rows, or [] when there is none yet."""
    if not os.path.exists(ledger):
        return []
    return _read_ledger(ledger)
