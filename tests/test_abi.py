"""The C-ABI libraries load and export every entry point the headers declare
(no device compute here)."""
import ctypes
import os
import re

import qgw_amd as Q

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qgwh?_[a-z0-9_]+)\s*\(", text)))


def test_host_library_exports_header():
    lib = Q.host_lib()
    for h in ("qgw_host.h", "qgw_nc.h"):
        names = _declared(h)
        assert names
        for n in names:
            assert hasattr(lib, n), (h, n)


def test_hip_library_exports_header():
    lib = Q.hip_lib()
    names = _declared("qgw_hip.h")
    assert "qgw_timestep" in names and "qgw_invert_pv" in names
    for n in names:
        assert hasattr(lib, n), n


def test_params_struct_layout():
    assert Q.host_lib().qgwh_params_size() == ctypes.sizeof(Q.Params)


def test_strerror_codes():
    lib = Q.hip_lib()
    for code in (0, -1, -2, -3, -4, -5, -6, -99):
        assert lib.qgw_strerror(code)


def test_hip_library_is_gfx950():
    so = os.path.join(ROOT, "qgw_amd", "libqgw_hip.so")
    data = open(so, "rb").read()
    assert b"gfx950" in data


def test_batch_rejects_adaptive_step():
    """A batch needs a fixed step in every member; the check precedes any device call."""
    p = Q.default_params(NX=64, NY=64, nl=2, dt_fixed=0.0)
    st = ctypes.c_int(0)
    b = Q.hip_lib().qgw_batch_create(1, ctypes.byref(p), Q.FP32, ctypes.byref(st))
    assert not b
    assert st.value == -1


def test_forcing_parameters_checked_before_device():
    """File forcing needs dt_forc_period a positive multiple of dt_forc > 0
    (src/forcing.h:238-251); stochastic forcing needs the square grid of the
    reference's N x N transform; the same checks hold for batch members.
    All rejected before any device call."""
    lib = Q.hip_lib()
    st = ctypes.c_int(0)
    for kw in (dict(dt_forc=0.0, dt_forc_period=100.0), dict(dt_forc=30.0, dt_forc_period=100.0),
               dict(dt_forc=-10.0, dt_forc_period=-100.0), dict(NY=48, k_f=10.0, sigma_f=1.0)):
        kw = dict(dict(NX=64, NY=64, nl=2, dt_fixed=600.0), **kw)
        p = Q.default_params(**kw)
        assert not lib.qgw_create(ctypes.byref(p), Q.FP64, ctypes.byref(st)), kw
        assert st.value == -1, kw
    for kw in (dict(dt_forc=30.0, dt_forc_period=100.0), dict(NY=48, k_f=10.0, sigma_f=1.0)):
        p = Q.default_params(**dict(dict(NX=64, NY=64, nl=2, dt_fixed=600.0), **kw))
        assert not lib.qgw_batch_create(1, ctypes.byref(p), Q.FP64, ctypes.byref(st))
        assert st.value == -1


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    so, before any device call."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
