"""The C ABI library, on the CPU: it loads, its version matches, it exports
every entry point include/fk_engine.h declares, and its struct layouts
agree with the header.  No compute call is made without a GPU."""
import ctypes
import pathlib
import re

import pytest

import fk_native as N

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "fk_engine.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^int\s+(fk_\w+)\s*\(", text, re.M)))


def test_library_loads_and_version_matches():
    L = N.lib()
    assert L.fk_abi_version() == N.ABI_VERSION
    m = re.search(r"#define FK_ABI_VERSION (\d+)", HEADER.read_text())
    assert int(m.group(1)) == N.ABI_VERSION


def test_every_declared_symbol_is_exported():
    L = N.lib()
    names = declared_functions()
    assert len(names) == len(N.EXPORTS)
    assert set(names) == set(N.EXPORTS)
    for name in names:
        assert hasattr(L, name), name


def test_struct_sizes_match_header():
    # sizes implied by the header's field lists (LP64)
    assert ctypes.sizeof(N.EngineDesc) == 4 * 5 + 4 * 12 + 4 * 11
    assert ctypes.sizeof(N.Window) == 32
    assert ctypes.sizeof(N.ClassifySpec) == 16
    assert N.RECORD_DTYPE.itemsize == 24
    assert ctypes.sizeof(N.EngineInfo) == 16 + 64 + 8 * 7 + 40 + 8 + 8 + 8


def test_status_codes_match_header():
    text = HEADER.read_text()
    for name, val in (("FK_OK", N.FK_OK), ("FK_E_ARG", N.FK_E_ARG),
                      ("FK_E_RANGE", N.FK_E_RANGE),
                      ("FK_E_OVERFLOW", N.FK_E_OVERFLOW),
                      ("FK_E_DEVICE", N.FK_E_DEVICE),
                      ("FK_E_OOM", N.FK_E_OOM), ("FK_E_NODEV", N.FK_E_NODEV),
                      ("FK_E_BUSY", N.FK_E_BUSY)):
        assert re.search(rf"\b{name} = {val}\b", text), name


def test_errors_map_to_reference_exceptions():
    with pytest.raises(ValueError):
        N.check(N.FK_E_RANGE)
    with pytest.raises(RuntimeError):
        N.check(N.FK_E_DEVICE)


def test_drop_in_modules_refuse_bad_parameters_before_the_device():
    import ladder_gpu
    import ap_gpu
    with pytest.raises(ValueError):
        ladder_gpu.GpuEngine(10, q2=2)
    with pytest.raises(ValueError):
        ap_gpu.GpuEngine(16, q2=5)


def test_batch_flags_are_checked_before_the_device():
    """fk_batch_create_ex refuses flags it does not know (FK_E_ARG) without
    touching a device; the dedicated-queue flag matches the header."""
    text = HEADER.read_text()
    m = re.search(r"#define FK_BATCH_DEDICATED_QUEUE (\d+)u", text)
    assert int(m.group(1)) == N.BATCH_DEDICATED_QUEUE
    m = re.search(r"#define FK_BATCH_THROUGHPUT (\d+)u", text)
    assert int(m.group(1)) == N.BATCH_THROUGHPUT
    h = ctypes.c_void_p()
    rc = N.lib().fk_batch_create_ex(0, 4, ctypes.byref(h))
    assert rc == N.FK_E_ARG
    assert not h.value
