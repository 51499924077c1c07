"""Crash-durable, config-keyed JSON cursor files for campaign segment loops.

A cursor file is one JSON object with a `key` (the engine parameters its
position is counted in) beside the position itself.  The file format and
its rules are huntlib/checkpoint.py:1-27's, so that either engine reads
the other's cursors: a write is all or nothing, the previous file stays
behind as `<path>.bak`, an unreadable file is an error rather than an
absent one, and a cursor under another key is refused, not reinterpreted.

How this module meets them:

* `save_json` writes the new text to `<path>.tmp` and fsyncs it, makes the
  current file the backup by a hard link (or a durable copy where the file
  system has no links), and renames the temporary file over the main one,
  then fsyncs the directory.  The main file therefore exists at every
  instant from its first save on: a crash at any point leaves the old or
  the new state under `<path>`, and never a missing cursor beside a full
  records ledger.
* `read_state` looks at the main file and then the backup and returns the
  first that parses, with where it came from.  Files that exist but none
  of which parse raise CheckpointCorrupt; no file at all is "no cursor".
"""

import json
import os
import shutil


class CheckpointCorrupt(RuntimeError):
    """A cursor file exists and no readable state stands behind it."""


class CursorRefused(RuntimeError):
    """A cursor exists that this configuration must not reinterpret."""


def _sync_dir(path):
    """Make a rename in `path`'s directory durable (best effort where the
    platform cannot open directories)."""
    try:
        fd = os.open(os.path.dirname(os.path.abspath(path)) or ".",
                     os.O_RDONLY)
    except OSError:
        return
    try:
        os.fsync(fd)
    except OSError:
        pass
    finally:
        os.close(fd)


def _write_synced(path, text):
    with open(path, "w", newline="\n") as f:
        f.write(text)
        f.flush()
        os.fsync(f.fileno())


def _refresh_backup(path):
    """<path>.bak becomes the current content of <path>, with <path>
    itself left in place."""
    bak = path + ".bak"
    staged = bak + ".tmp"
    try:
        if os.path.exists(staged):
            os.unlink(staged)
        os.link(path, staged)
    except OSError:
        # no hard links here: a synced copy does the same job
        shutil.copyfile(path, staged)
        with open(staged, "r+b") as f:
            os.fsync(f.fileno())
    os.replace(staged, bak)


def save_json(path, obj):
    """Durably replace the JSON at `path` by `obj`, the old content kept as
    `<path>.bak`."""
    path = os.fspath(path)
    fresh = path + ".tmp"
    _write_synced(fresh, json.dumps(obj, indent=1) + "\n")
    if os.path.exists(path):
        try:
            _refresh_backup(path)
        except OSError:
            pass            # a missing backup costs a fallback, not the state
    os.replace(fresh, path)
    _sync_dir(path)


def _parse(path):
    with open(path) as f:
        obj = json.load(f)
    if not isinstance(obj, dict):
        raise ValueError("not a JSON object")
    return obj


def read_state(path):
    """(state, source, problems): the state from `path`, else from its
    backup (source names the file read), or (None, None, []) when neither
    file exists.  problems lists the files that exist but did not parse.
    Raises CheckpointCorrupt when files exist and none of them parses."""
    path = os.fspath(path)
    problems = []
    for cand in (path, path + ".bak"):
        if not os.path.exists(cand):
            continue
        try:
            return _parse(cand), cand, problems
        except (OSError, ValueError) as err:
            problems.append(f"{cand}: {err}")
    if problems:
        raise CheckpointCorrupt(
            "no readable cursor: " + "; ".join(problems) +
            ".  Nothing was changed; start again with fresh=True only to "
            "discard this frontier on purpose")
    return None, None, []


def load(path, expect_key, warn=None):
    """The cursor under `expect_key`, or None when there is none or it was
    counted under another key.  A state recovered from the backup is one
    segment behind the lost one; `warn` is told so."""
    path = os.fspath(path)
    state, source, problems = read_state(path)
    if state is None:
        return None
    if source != path and warn:
        lost = "; ".join(problems) if problems else f"{path} is missing"
        warn(f"resuming from {source} ({lost}): one segment is redone")
    if state.get("key") != expect_key:
        if warn:
            warn(f"cursor in {source} is keyed {state.get('key')!r}, not "
                 f"{expect_key!r}; it is not used")
        return None
    return state


def refuse_mismatch(path, expect_key, fresh=False, describe=None):
    """Return when starting under `expect_key` is safe: there is no cursor,
    or its key is this one, or `fresh` discards it on purpose.  Raise
    CursorRefused for a cursor under another key (resuming it would
    misplace its frontier, restarting would abandon it) and
    CheckpointCorrupt when no readable state exists."""
    if fresh:
        return
    path = os.fspath(path)
    state, source, _ = read_state(path)
    if state is None or state.get("key") == expect_key:
        return
    where = ""
    if describe is not None:
        try:
            where = f" ({describe(state)})"
        except Exception:           # describing is a courtesy only
            where = ""
    raise CursorRefused(
        f"{source} holds a cursor for {state.get('key')!r}{where}; this "
        f"run is {expect_key!r}.  Use another cursor path, or fresh=True to "
        f"give that frontier up")
