"""4-D file forcing on the GPU (init_4d_forcing/calc_4d_forcing,
src/forcing.h:226-289, src/timestep.h:152-163) against the oracle.

The records are written to a netCDF file with scipy and read by the library
(qgwh_nc_read_var); the oracle gets the same float values directly."""
import sys

import numpy as np
import pytest
from scipy.io import netcdf_file

import qgw_amd as Q
import oracle as O
from conftest import rel

pytestmark = pytest.mark.gpu


def _announce(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _params(**kw):
    base = dict(NX=64, NY=48, nl=3, Lx=1000e3, dh=[300.0, 700.0, 2000.0], N2=[2e-5, 6e-6], f0=1e-4, beta=2e-11,
                hEkb=3.0, n_hyper=4, nu_hyper=1e8, tau0=5e-5, forc_mode=2.0, bc_fac=0.0, dt_fixed=600.0,
                ic_urms=0.1, ic_lambda0=150e3, seed=77, dt_forc=1000.0, dt_forc_period=3000.0)
    base.update(kw)
    return Q.default_params(**base)


def _records(p, nrec, rng, amp=2e-10):
    """nrec float-valued storage fields of nl layers (a periodic run's ring stays zero)."""
    g = Q.grid(p)[0]
    r = amp * rng.standard_normal((nrec, p.nl, g.nyp2, g.nxp2))
    r = r.astype(np.float32).astype(np.float64)
    if p.bc_fac == -1.0:
        ring = np.ones((g.nyp2, g.nxp2), bool)
        ring[1:-1, 1:-1] = False
        r[:, :, ring] = 0.0
    return r


def _write(path, p, recs, name="q_forcing_3d"):
    g = Q.grid(p)[0]
    b = 1 if p.bc_fac == -1.0 else 0
    f = netcdf_file(str(path), "w")
    f.createDimension("t", None)
    f.createDimension("level", p.nl)
    f.createDimension("y", g.nyp2 - 2 * b)
    f.createDimension("x", g.nxp2 - 2 * b)
    v = f.createVariable(name, "f", ("t", "level", "y", "x"))
    for r in range(recs.shape[0]):
        v[r] = recs[r][:, b:g.nyp2 - b, b:g.nxp2 - b]
    f.close()


def _pair(p, recs, path, precision=Q.FP64):
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes)
    m = Q.Model(p, precision, modes)
    m.set_forcing_file(str(path))
    m.set_q(q0)
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.set_forcing4d(recs, p.dt_forc)
    return m, o


@pytest.mark.parametrize("case", ["bounded", "periodic", "adaptive", "one_record"])
def test_forcing_fp64_vs_oracle(tmp_path, case):
    kw = {"bounded": dict(),
          "periodic": dict(bc_fac=-1.0, NX=64, NY=64, Lx=2000e3, dt_forc=900.0, dt_forc_period=3600.0),
          "adaptive": dict(dt_fixed=0.0, cfl=0.3),
          "one_record": dict(dt_forc=5000.0, dt_forc_period=5000.0)}[case]
    p = _params(**kw)
    nrec = int(round(p.dt_forc_period / p.dt_forc))
    _announce("4-D forcing FP64", case, p.NX, p.NY, p.nl, "records", nrec)
    recs = _records(p, nrec, np.random.default_rng(5))
    _write(tmp_path / "forc.nc", p, recs)
    m, o = _pair(p, recs, tmp_path / "forc.nc")
    if case == "adaptive":
        m.set_next_output(20 * 3600.0)
        t = o.get_time()
        o.set_time(0.0, t["dt"], t["dt0"], t["dt1"], 0, 20 * 3600.0)
    # 14 steps of 600 s cross every record boundary and the wrap to record 0
    for _ in range(14):
        m.timestep()
        o.timestep()
    assert m.get_time()["t"] == pytest.approx(o.get_time()["t"], rel=1e-14)
    assert rel(m.q, o.field(O.Q)) < 1e-12
    # the forcing is visible: the same run without it ends elsewhere
    o2 = O.Oracle(p)
    o2.set_modes(*Q.eigmode(p))
    o2.set_q(Q.make_ic(p)[1])
    if case == "adaptive":
        t = o2.get_time()
        o2.set_time(0.0, t["dt"], t["dt0"], t["dt1"], 0, 20 * 3600.0)
    o2.run(14)
    assert rel(o2.field(O.Q), o.field(O.Q)) > 1e-6


def test_forcing_closed_form_on_device(tmp_path):
    """A forced sine mode with every other tendency off: q = dt * sum A(t_n) F
    with the hand-computed triangle wave 0, .5, 1, 1.5, 2, 1, 0, .5, 1 of
    records 0, F, 2F at dt_forc = 2 dt (psi is proportional to q, so the
    Jacobian vanishes)."""
    p = _params(NX=32, NY=32, nl=2, dh=[400.0, 1600.0], N2=[1e-5], beta=0.0, tau0=0.0, h_topo=0.0, nu=0.0,
                nu_hyper=0.0, hEkb=0.0, bc_fac=0.0, dt_fixed=100.0, dt_forc=200.0, dt_forc_period=600.0)
    _announce("4-D forcing closed form", p.NX, p.nl)
    g = Q.grid(p)[0]
    i, j = np.arange(g.nxp2), np.arange(g.nyp2)
    s = np.outer(np.sin(np.pi * j / (g.m + 1)), np.sin(2 * np.pi * i / (g.n + 1)))
    F = np.stack([1e-9 * s, -4e-10 * s]).astype(np.float32).astype(np.float64)
    F[:, 0, :] = F[:, -1, :] = F[:, :, 0] = F[:, :, -1] = 0
    recs = np.stack([0 * F, F, 2 * F])
    _write(tmp_path / "forc.nc", p, recs)
    m = Q.Model(p, Q.FP64)
    m.set_forcing_file(str(tmp_path / "forc.nc"))
    m.run(9)
    assert rel(m.q, 100.0 * 7.5 * F) < 1e-12


def test_forcing_fp32_vs_fp64(tmp_path):
    p = _params()
    _announce("4-D forcing FP32 vs FP64", p.NX, p.nl)
    recs = _records(p, 3, np.random.default_rng(6))
    _write(tmp_path / "forc.nc", p, recs)
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes)
    out = {}
    for prec in (Q.FP64, Q.FP32, Q.FP32_PLAIN):
        m = Q.Model(p, prec, modes)
        m.set_forcing_file(str(tmp_path / "forc.nc"))
        m.set_q(q0)
        m.run(20)
        out[prec] = m.q
    assert rel(out[Q.FP32], out[Q.FP64]) < 2e-5
    assert rel(out[Q.FP32_PLAIN], out[Q.FP64]) < 2e-5


def test_forcing_default_file_and_switch(tmp_path, monkeypatch):
    """Default input_vars.nc in the working directory (src/forcing.h:262);
    qgw_set_forcing_file reads the records again from the new file."""
    p = _params()
    _announce("4-D forcing default file and switch", p.NX, p.nl)
    rng = np.random.default_rng(8)
    ra, rb = _records(p, 3, rng), _records(p, 3, rng)
    _write(tmp_path / "input_vars.nc", p, ra)
    _write(tmp_path / "other.nc", p, rb)
    monkeypatch.chdir(tmp_path)
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes)
    m = Q.Model(p, Q.FP64, modes)
    m.set_q(q0)
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.set_forcing4d(ra, p.dt_forc)
    m.run(3)
    o.run(3)
    m.set_forcing_file("other.nc")
    o.set_forcing4d(rb, p.dt_forc)
    m.run(4)
    o.run(4)
    assert rel(m.q, o.field(O.Q)) < 1e-12


def test_forcing_missing_file_fails_before_the_step(tmp_path):
    p = _params()
    _announce("4-D forcing missing file", p.NX, p.nl)
    m = Q.Model(p, Q.FP64)
    m.set_forcing_file(str(tmp_path / "none.nc"))
    m.set_q(Q.make_ic(p)[1])
    q_before = m.q
    with pytest.raises(Q.QGWError):
        m.timestep()
    assert m.get_time()["t"] == 0.0 and m.get_time()["it"] == 0
    assert np.array_equal(m.q, q_before)
    # a file without the variable
    recs = _records(p, 3, np.random.default_rng(9))
    _write(tmp_path / "wrong.nc", p, recs, name="something_else")
    m.set_forcing_file(str(tmp_path / "wrong.nc"))
    with pytest.raises(Q.QGWError):
        m.timestep()
    # and then the right one
    _write(tmp_path / "ok.nc", p, recs)
    m.set_forcing_file(str(tmp_path / "ok.nc"))
    m.timestep()
    assert m.get_time()["it"] == 1
