"""qgw_diagnostics_device, the allocation-free entry point: its sums against
qgw_diagnostics on the same state (bit for bit: the same kernels and block
order), on two calls in a row, the second on a smaller run (the scratch keeps
the first call's contents), and its argument checks."""
import ctypes
import sys

import numpy as np
import pytest

import qgw_amd as Q
from qgw_amd import configs

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            pass
    raise RuntimeError("libamdhip64 not found")


def test_diagnostics_device_matches_diagnostics():
    hip = _hip()
    lib = Q.hip_lib()
    lib.qgw_diagnostics_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.qgw_diagnostics_device.restype = ctypes.c_int
    out = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(out), ctypes.c_size_t(8 * 2 * 64)) == 0
    try:
        for NX, nl, prec in ((256, 5, Q.FP32), (64, 2, Q.FP64)):
            print("[case] diagnostics_device", NX, nl, prec, flush=True)
            sys.stdout.flush()
            p = configs.c1(NX=NX, NY=NX, nl=nl, dh=list(np.linspace(300.0, 3000.0, nl)),
                           N2=[1e-5] * (nl - 1))
            modes = Q.eigmode(p)
            q0 = Q.make_ic(p, modes)[1]
            m = Q.Model(p, prec, modes)
            m.set_q(q0)
            m.run(3)
            m.invert_pv()
            El, Zl, _, _ = m.diagnostics()
            assert lib.qgw_diagnostics_device(m.ctx, out) == 0
            m.synchronize()
            h = np.zeros(2 * nl)
            assert hip.hipMemcpy(h.ctypes.data_as(ctypes.c_void_p), out, ctypes.c_size_t(8 * 2 * nl), 2) == 0
            w = (p.Lx / p.NX) ** 2 / p.Lx ** 2
            assert np.array_equal(-h[0::2] * w, El), (NX, nl)
            assert np.array_equal(h[1::2] * w, Zl), (NX, nl)
            assert lib.qgw_diagnostics_device(m.ctx, None) < 0  # a null output is refused
            m.close()
    finally:
        hip.hipFree(out)
