"""The process-global libc rand() stream the reference's sweeps consume.

qmc.QuantumAnneal and sa.Anneal call libc ``rand()`` (piqmc/qmc.pyx:22,132,
piqmc/sa.pyx:19,116).  The replay kernels run that generator (glibc TYPE_3,
31-word additive feedback) on the device, so the drop-in needs the current
libc state before a call and must leave libc in the state the reference would
leave it in afterwards.

With glibc the state is read and written through the documented
``initstate``/``setstate`` pair: ``initstate`` returns the previous state
array, whose first word encodes (rear index * 5 + generator type), followed by
the 31 state words.  Elsewhere a module-level emulated stream is used, seeded
like ``srand(1)``.
"""
import ctypes as C

import numpy as np

from . import _lib

_MAX_TYPES, _TYPE_3 = 5, 3
_glibc = None
_tmp = C.create_string_buffer(132)


def _load_glibc():
    global _glibc
    if _glibc is None:
        try:
            libc = C.CDLL("libc.so.6")
            libc.gnu_get_libc_version  # glibc only
            libc.initstate.argtypes = [C.c_uint, C.c_void_p, C.c_size_t]
            libc.initstate.restype = C.c_void_p
            libc.setstate.argtypes = [C.c_void_p]
            libc.setstate.restype = C.c_void_p
            _glibc = libc
        except (OSError, AttributeError):
            _glibc = False
    return _glibc


_emulated = np.zeros(32, np.uint32)
_lib.call("piqmc_rand_seed", 1, _emulated.ctypes.data)
sync_libc = True  # False: use the module's own emulated stream


def _glibc_table():
    libc = _load_glibc()
    prev = libc.initstate(1, C.addressof(_tmp), 128)
    words = (C.c_int32 * 32).from_address(prev)
    if words[0] % _MAX_TYPES != _TYPE_3:
        libc.setstate(prev)
        return None, None
    return prev, words


def get_state():
    """Current rand() state as uint32[32] (31 words + front index)."""
    if sync_libc and _load_glibc():
        prev, words = _glibc_table()
        if prev is not None:
            rear = words[0] // _MAX_TYPES
            st = np.array([w & 0xFFFFFFFF for w in words[1:32]] + [(rear + 3) % 31], np.uint32)
            _load_glibc().setstate(prev)
            return st
    return _emulated.copy()


def set_state(st):
    """Make the next rand() calls continue from state st (uint32[32])."""
    st = np.asarray(st, np.uint32)
    if sync_libc and _load_glibc():
        prev, words = _glibc_table()
        if prev is not None:
            front = int(st[31])
            rear = (front + 28) % 31
            words[0] = rear * _MAX_TYPES + _TYPE_3
            for t in range(31):
                words[1 + t] = int(st[t]) - (1 << 32) if st[t] >= (1 << 31) else int(st[t])
            _load_glibc().setstate(prev)
            return
    _emulated[:] = st


def srand(seed):
    """Seed the stream like libc srand(seed)."""
    st = np.zeros(32, np.uint32)
    _lib.call("piqmc_rand_seed", int(seed) & 0xFFFFFFFF, st.ctypes.data)
    set_state(st)
