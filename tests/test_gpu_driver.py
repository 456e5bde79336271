"""The C driver (qgw_amd/qg_hip, src/qg.c:103-226 restated) end to end on the
GPU: parameter file, output directory, restart.nc read, vars.nc records at
the output times.  The records are checked against the oracle stepped from
the same (float-rounded) restart state."""
import os
import subprocess
import sys

import numpy as np
import pytest
from scipy.io import netcdf_file

import qgw_amd as Q
import oracle as O
from conftest import rel

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(Q.__file__), "qg_hip")

KEYS = dict(NX=48, NY=48, nl=2, Lx=1000e3, dh=[500.0, 2500.0], N2=[2e-5], f0=1e-4, beta=2e-11, hEkb=3.0,
            n_hyper=4, nu_hyper=1e8, tau0=5e-5, forc_mode=2.0, bc_fac=0.5, h_topo=200.0, w_topo=100e3,
            dt_fixed=900.0, nsteps=6, dt_out=2700.0, tend=1e9)


def _params_file(path, keys):
    with open(path, "w") as f:
        for k, v in keys.items():
            if isinstance(v, list):
                f.write("%s = [%s]\n" % (k, ", ".join(repr(x) for x in v)))
            else:
                f.write("%s = %r\n" % (k, v))


@pytest.mark.parametrize("periodic", [False, True])
def test_driver_restart_to_vars_nc(tmp_path, periodic):
    keys = dict(KEYS)
    if periodic:
        keys.update(bc_fac=-1.0, NX=32, NY=32, Lx=2e6)
    print("[case] driver restart -> vars.nc", keys["NX"], keys["nl"], "periodic" if periodic else "bounded",
          flush=True)
    _params_file(str(tmp_path / "params.in"), keys)
    p = Q.read_params(str(tmp_path / "params.in"))
    g = Q.grid(p)[0]
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes, seed=11)
    rng = np.random.default_rng(12)
    topo = Q.topo(p) + 5.0 * rng.standard_normal((g.nyp2, g.nxp2))
    b = 1 if periodic else 0
    f = netcdf_file(str(tmp_path / "restart.nc"), "w")
    f.createDimension("t", None)
    f.createDimension("level", p.nl)
    f.createDimension("y", g.nyp2 - 2 * b)
    f.createDimension("x", g.nxp2 - 2 * b)
    f.createVariable("q", "f", ("t", "level", "y", "x"))[0] = q0[:, b:g.nyp2 - b, b:g.nxp2 - b]
    f.createVariable("topo", "f", ("t", "y", "x"))[0] = topo[b:g.nyp2 - b, b:g.nxp2 - b]
    f.close()

    r = subprocess.run([DRIVER, "params.in"], cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0
    assert "restart.nc: q topo" in r.stdout
    out = tmp_path / "outdir_0001"
    assert (out / "params.in").read_text() == (tmp_path / "params.in").read_text()

    # the oracle from what the driver read (restart values rounded to float;
    # a periodic run's ring keeps the computed values)
    q_in, topo_in, found = Q.read_restart(str(tmp_path / "restart.nc"), p, q=np.zeros_like(q0), topo=Q.topo(p))
    assert found == {"q", "topo"}
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.field(O.TOPO)[...] = topo_in
    o.set_q(q_in)
    v = netcdf_file(str(out / "vars.nc"), "r", mmap=False)
    try:
        T = v.variables["t"][:]
        assert np.array_equal(T, np.array([0.0, 2700.0], np.float32))
        for rec, steps in ((0, 0), (1, 3)):
            o.run(steps)
            o.invert_pv()
            qo = o.field(O.Q)[:, b:g.nyp2 - b, b:g.nxp2 - b]
            po = o.field(O.PSI)[:, b:g.nyp2 - b, b:g.nxp2 - b]
            qf, pf = v.variables["q"][rec].astype(np.float64), v.variables["psi"][rec].astype(np.float64)
            # FP64 parity is ~1e-13; the file holds floats (relative 2^-24 per value)
            assert rel(qf, qo) < 1e-7 and rel(pf, po) < 1e-7
            assert np.abs(qf - qo).max() <= 2 ** -23 * np.abs(qo).max()
    finally:
        v.close()
