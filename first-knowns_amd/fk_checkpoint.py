"""Crash-durable, config-keyed JSON cursors for campaign segment loops.

The rules are huntlib/checkpoint.py:1-27's, kept because a campaign's
cursor files must mean the same thing on either engine:

* a file is written to a temporary name, flushed and fsynced, and only
  then renamed over the old one, so a crash leaves either the old or the
  new file, never a torn one (a rename can reach the disk before the data,
  which is why the fsync comes first);
* the file that was on disk before the rename is kept as `<path>.bak`, and
  `load` falls back to it when the main file cannot be read;
* a file that exists but cannot be read is an error (CheckpointCorrupt),
  never "no checkpoint": for a live frontier those need opposite answers;
* every cursor carries a `key` naming the engine parameters its position
  is counted in; `load` ignores a file whose key differs, and
  `refuse_mismatch` refuses to start over one unless told `fresh`.
"""

import json
import os


class CheckpointCorrupt(RuntimeError):
    """A cursor file exists and neither it nor its backup can be read."""


class CursorRefused(RuntimeError):
    """A cursor exists that this configuration must not reinterpret."""


def save_json(path, obj):
    """Write obj durably: temporary file, fsync, keep the old file as
    .bak, then rename into place (huntlib/checkpoint.py:40-66)."""
    path = os.fspath(path)
    tmp = path + ".tmp"
    with open(tmp, "w", newline="\n") as f:
        json.dump(obj, f, indent=1)
        f.flush()
        os.fsync(f.fileno())
    if os.path.exists(path):
        try:
            os.replace(path, path + ".bak")
        except OSError:
            pass                      # the backup is best effort
    os.replace(tmp, path)


def _read(path):
    with open(path) as f:
        obj = json.load(f)
    if not isinstance(obj, dict):
        raise ValueError("cursor file does not hold a JSON object")
    return obj


def load(path, expect_key, warn=None):
    """The cursor dict, or None when there is none or its key differs
    (huntlib/checkpoint.py:80-110).  Falls back to `.bak` when the file
    cannot be read, and raises CheckpointCorrupt when neither can."""
    path = os.fspath(path)
    if not os.path.exists(path):
        return None
    try:
        state = _read(path)
    except Exception as err:
        bak = path + ".bak"
        if not os.path.exists(bak):
            raise CheckpointCorrupt(
                f"{path} cannot be read ({err}) and there is no {bak}; "
                f"restart deliberately with fresh=True") from err
        try:
            state = _read(bak)
        except Exception as err2:
            raise CheckpointCorrupt(
                f"{path} cannot be read ({err}), nor {bak} ({err2})") from err2
        if warn:
            warn(f"{path} cannot be read ({err}); resuming from {bak}, "
                 f"one segment behind")
    if state.get("key") != expect_key:
        if warn:
            warn(f"cursor key {state.get('key')!r} differs; ignoring it")
        return None
    return state


def refuse_mismatch(path, expect_key, fresh=False, describe=None):
    """Raise unless starting is safe (huntlib/checkpoint.py:117-165): an
    existing cursor under another key is refused rather than ignored, since
    ignoring it would restart the sweep at its floor and reinterpreting it
    would misplace the frontier."""
    path = os.fspath(path)
    if fresh or not os.path.exists(path):
        return
    try:
        state = _read(path)
    except Exception as err:
        bak = path + ".bak"
        if os.path.exists(bak):
            try:
                _read(bak)
                return                # load() recovers it and says so
            except Exception:
                pass
        raise CheckpointCorrupt(
            f"{path} exists but cannot be read ({err}) and no usable backup "
            f"stands behind it; restart deliberately with fresh=True") from err
    if state.get("key") == expect_key:
        return
    extra = ""
    if describe:
        try:
            extra = f"; stored: {describe(state)}"
        except Exception:
            extra = ""
    raise CursorRefused(
        f"a cursor exists under another configuration: stored key "
        f"{state.get('key')!r}, wanted {expect_key!r}{extra}.  Starting would "
        f"abandon that frontier; pass fresh=True to discard it deliberately.")
