"""The autotuner's table of launch choices, and its JSON cache.

The library times the candidate configurations of each tunable launch on the first launch of
a case (a key naming the launch and the shapes that matter) and keeps the fastest for the
process.  The knobs are:

- ``paint_rows``: the row-block shape of the fused paint (rows per workgroup, LDS footprint);
- ``xpass_cols``: columns per tile of the binning x pass (FFT tile and LDS layout);
- ``ypass_cols``: columns per tile of the y pass;
- ``sort_chunk``: particles per workgroup of the row-sort scatter (LDS staging).

The table can be saved as JSON and loaded in a later process, so the measurements run once per
machine: :func:`save` / :func:`load`, or set ``JAXPOWER_AMD_TUNING_CACHE`` to a file path, which
is loaded when the package is imported and written back at exit when new cases were measured.
jax-power itself leaves launch configuration to XLA; this module has no counterpart there.
"""
import atexit
import contextlib
import ctypes
import json
import os

from . import _lib

KNOBS = {'paint_rows': 0, 'xpass_cols': 1, 'ypass_cols': 2, 'sort_chunk': 3}
ENV_CACHE = 'JAXPOWER_AMD_TUNING_CACHE'
_FORMAT = 2  # 2: y-pass candidates timed on a copy larger than the last-level cache (format-1 choices dropped)


def table():
    """The current table: {case key: {'knob': name, 'choice': q, 'ms': [...], 'measured': bool}}."""
    lib = _lib.load()
    size = lib.mp_tune_export(None, 0) + 1
    buf = ctypes.create_string_buffer(size)
    lib.mp_tune_export(buf, size)
    return json.loads(buf.value.decode())


def _device_name():
    import torch
    if not torch.cuda.is_available():
        return None
    return torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName


def save(path):
    """Writes the table to ``path`` as JSON (with the device it was measured on)."""
    doc = {'format': _FORMAT, 'device': _device_name(), 'entries': table()}
    directory = os.path.dirname(os.path.abspath(path))
    os.makedirs(directory, exist_ok=True)
    tmp = path + '.tmp'
    with open(tmp, 'w') as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
    return len(doc['entries'])


def load(path, device=None):
    """Imports the entries of a JSON file written by :func:`save`; entries measured on another
    device type are skipped.  Returns the number of entries imported."""
    with open(path) as f:
        doc = json.load(f)
    if doc.get('format') != _FORMAT:
        raise ValueError(f'{path}: tuning cache format {doc.get("format")!r}, expected {_FORMAT}')
    device = device if device is not None else _device_name()
    if doc.get('device') and device and doc['device'] != device:
        return 0
    lib = _lib.load()
    count = 0
    for key, entry in doc['entries'].items():
        knob = KNOBS.get(entry.get('knob'))
        choice = entry.get('choice')
        if knob is None or not isinstance(choice, int):
            raise ValueError(f'{path}: malformed entry {key!r}: {entry!r}')
        _lib.check(lib.mp_tune_set(key.encode(), knob, choice), 'mp_tune_set')
        count += 1
    return count


def clear():
    """Empties the table (cases are measured again on their next launch)."""
    _lib.check(_lib.load().mp_tune_clear(), 'mp_tune_clear')


def enable(on=True):
    """Turns measurement on or off (off: cases not in the table take the default candidate);
    returns the previous setting."""
    return bool(_lib.load().mp_tune_enable(1 if on else 0))


@contextlib.contextmanager
def force(knob, choice):
    """Within the block, every launch of ``knob`` uses candidate ``choice`` (where it applies;
    otherwise the default)."""
    lib = _lib.load()
    k = KNOBS[knob]
    old = lib.mp_tune_force(k, int(choice))
    try:
        yield
    finally:
        lib.mp_tune_force(k, old)


def _autoload():
    path = os.environ.get(ENV_CACHE)
    if not path:
        return
    start = 0
    if os.path.exists(path):
        try:
            start = load(path)
        except (OSError, ValueError) as exc:
            import warnings
            warnings.warn(f'tuning cache {path} not loaded: {exc}')

    def _autosave():
        try:
            entries = table()
        except Exception:  # the library may be gone at interpreter exit
            return
        if any(e.get('measured') for e in entries.values()) or len(entries) != start:
            save(path)

    atexit.register(_autosave)
