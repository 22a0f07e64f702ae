"""The library's launch-shape autotuner (aprl_autotune, aprl_autotune_report):
launch shapes that give bitwise the same results are timed per kernel and
grid shape on their first uses and the fastest is kept; APRL_TUNE_CACHE
names a file the decisions are read from and appended to."""
import ctypes as C

from ._capi import check, lib


def setAutotune(enabled: bool = True):
    """Tuning on (the default) or off (every kernel launches its default shape)."""
    check(lib().aprl_autotune(1 if enabled else 0))


def autotuneDecisions() -> dict:
    """{key: chosen value} of the decisions taken so far."""
    n = C.c_size_t()
    check(lib().aprl_autotune_report(None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(lib().aprl_autotune_report(buf, n.value + 1, C.byref(n)))
    out = {}
    for line in buf.value.decode().splitlines():
        key, _, value = line.rpartition(" ")
        out[key] = int(value)
    return out
