"""Forceable failures: each one must end in a clean negative status, never a
hang, a fault or a silently wrong state.

- a NaN in q on the fixed-step path (single run and batch) and on the
  adaptive path: QGW_ERR_NUMERIC from the call that waits for the steps;
- the 1024^2 x 3 configuration with its biharmonic viscosity as written
  (nu4 = 1e10, AB3-unstable: |lambda| dt ~ 1.34 > 0.545), on a 256^2 grid of
  the same spacing, overflows within 25 steps (the oracle overflows there
  too); the stable nu4 = 2e9 runs the same steps;
- a device-memory ceiling below what the run needs: QGW_ERR_NOMEM at create;
- a corrupt restart.nc given to the C driver: a failed read, exit status 1.

Every case prints its name before the first device call.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import qgw_amd as Q
from qgw_amd import configs

pytestmark = pytest.mark.gpu


def _announce(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _small(**over):
    kw = dict(NX=64, NY=64, Lx=1000e3)
    kw.update(over)
    return configs.c1(**kw)


def _numeric(excinfo):
    return "numeric" in str(excinfo.value).lower() or "nan" in str(excinfo.value).lower()


@pytest.mark.parametrize("prec", [Q.FP64, Q.FP32])
def test_nan_in_q_fixed_step(prec):
    p = _small()
    _announce("NaN in q, fixed step", p.NX, prec)
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes)
    m = Q.Model(p, prec, modes)
    bad = q0.copy()
    bad[1, 20, 30] = np.nan
    m.set_q(bad)
    with pytest.raises(Q.QGWError) as e:
        m.run(3)
    assert _numeric(e)
    with pytest.raises(Q.QGWError):
        m.synchronize()      # the flag is sticky
    # a new state clears it: q, and the tendency history the NaN has reached
    zero = np.zeros_like(q0)
    for f in (Q.FIELD_F0, Q.FIELD_F1, Q.FIELD_F2):
        m.upload(f, zero)
    m.set_q(q0)
    m.run(3)
    assert np.isfinite(m.q).all()
    m.close()


def test_nan_in_q_adaptive_step():
    p = _small(dt_fixed=0.0, tend=1e9, dt_out=1e9)
    _announce("NaN in q, adaptive step", p.NX)
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes)
    m = Q.Model(p, Q.FP64, modes)
    bad = q0.copy()
    bad[0, 10, 10] = np.inf
    m.set_q(bad)
    with pytest.raises(Q.QGWError) as e:
        m.timestep()
    assert _numeric(e)
    m.close()


def test_nan_in_one_batch_member():
    ps = [_small(seed=configs.SEED0 + i) for i in range(2)]
    _announce("NaN in one member of a batch", ps[0].NX)
    modes = [Q.eigmode(p) for p in ps]
    b = Q.Batch(ps, Q.FP32, modes)
    for i, p in enumerate(ps):
        q = Q.make_ic(p, modes[i])[1]
        if i == 1:
            q[0, 5, 7] = np.nan
        b.set_q(i, q)
    assert b.groups()[0] == 1
    with pytest.raises(Q.QGWError) as e:
        b.step(12)           # two graph replays
    assert _numeric(e)
    b.close()


def test_as_written_viscosity_blows_up_cleanly():
    # 1024^2 x 3 physics with nu4 = 1e10 on 256^2 at the same grid spacing
    # (Lx = 1000 km): the hyperviscous AB3 factor is unchanged (1.34)
    unstable = configs.c2(as_written=True, NX=256, NY=256, Lx=1000e3)
    stable = configs.c2(NX=256, NY=256, Lx=1000e3)
    _announce("as-written nu4 blow-up", unstable.NX, unstable.nl)
    for p, blows in ((unstable, True), (stable, False)):
        modes = Q.eigmode(p)
        _, q0 = Q.make_ic(p, modes)
        m = Q.Model(p, Q.FP64, modes)
        m.set_q(q0)
        if blows:
            with pytest.raises(Q.QGWError) as e:
                for _ in range(8):
                    m.run(25)
            assert _numeric(e)
        else:
            m.run(200)
            assert np.isfinite(m.q).all()
        m.close()


def test_device_memory_limit_fails_create(monkeypatch):
    p = _small()
    _announce("device memory limit", p.NX)
    modes = Q.eigmode(p)
    monkeypatch.setenv("QGW_MAX_DEVICE_BYTES", "200000")
    with pytest.raises(Q.QGWError) as e:
        Q.Model(p, Q.FP64, modes)
    assert "alloc" in str(e.value).lower() or "memory" in str(e.value).lower()
    monkeypatch.delenv("QGW_MAX_DEVICE_BYTES")
    m = Q.Model(p, Q.FP64, modes)  # and nothing is left behind
    m.close()


def test_driver_rejects_corrupt_restart(tmp_path):
    drv = os.path.join(os.path.dirname(Q.__file__), "qg_hip")
    _announce("driver with a corrupt restart.nc")
    with open(tmp_path / "params.in", "w") as f:
        f.write("NX = 32\nNY = 32\nnl = 1\nLx = 1e6\ndh = [1000.0]\ndt_fixed = 600.0\nnsteps = 2\n")
    with open(tmp_path / "restart.nc", "wb") as f:
        f.write(b"CDF\x01" + os.urandom(64))
    r = subprocess.run([drv, "params.in"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "restart.nc" in r.stderr
