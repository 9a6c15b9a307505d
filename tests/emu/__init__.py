"""ctypes binding of the CPU emulator of the kernel bodies (test infrastructure)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libemu.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        _lib = ctypes.CDLL(path)
    return _lib


def far_field(prec, theta, phi, lam, recs, nb, vrecs=None, fast_diff=False):
    """Kernel bodies on the CPU: complex [r, c, theta, phi]."""
    import oracle
    from abt_amd.lib import raw_to_complex
    th = np.ascontiguousarray(theta, np.float64)
    ph = np.ascontiguousarray(phi, np.float64)
    x, y, z = oracle.make_far_field_bins(th, ph)
    cp, sp = np.cos(ph), np.sin(ph)
    ti = int(np.argmax(th > oracle.PI_R4 / 180)) if np.any(th > oracle.PI_R4 / 180) else len(th) - 1
    out = np.empty(8 * len(th) * len(ph))
    recs = np.ascontiguousarray(recs, np.float64)
    vr = recs if vrecs is None else np.ascontiguousarray(vrecs, np.float64)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib().emu_far_field(ctypes.c_int(prec), ctypes.c_int(len(ph)), ctypes.c_int(len(th)),
                        ctypes.c_int(ti), ctypes.c_double(lam), P(x), P(y), P(z), P(cp), P(sp),
                        P(recs), P(vr), ctypes.c_int64(nb), ctypes.c_int(int(fast_diff)), P(out))
    return raw_to_complex(out, len(ph), len(th))
