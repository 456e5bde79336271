"""Stochastic ring forcing on the GPU (init_stoch_forc/calc_forc,
src/forcing.h:67-208, src/timestep.h:143-150) against the oracle.

The device draws its coefficients from the context's own stream, which starts
as the C library's rand() does; the oracle draws from rand() itself after
srand(1).  Both therefore force with the same realisation, and the states
agree to FP64 rounding."""
import sys

import numpy as np
import pytest

import qgw_amd as Q
import oracle as O
from conftest import rel

pytestmark = pytest.mark.gpu


def _announce(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _params(**kw):
    # test/params_06.in (periodic, one layer, k_f = 32, sigma_f = 1e-3) on 64^2
    base = dict(NX=64, NY=64, nl=1, Lx=1.0, dh=[1.0], N2=[0.0], f0=2.0, beta=5.0, tau0=0.0, nu=0.0,
                nu_hyper=1e-17, n_hyper=6, hEkb=5e-4, bc_fac=-1.0, k_f=32.0, sigma_f=1e-3,
                dt_fixed=2e-3, ic_urms=0.02, ic_lambda0=0.1, seed=5)
    base.update(kw)
    return Q.default_params(**base)


def _pair(p, precision=Q.FP64):
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes)
    m = Q.Model(p, precision, modes)
    m.set_q(q0)
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.set_stochastic(p.k_f, p.sigma_f)
    return m, o


CASES = {
    "periodic_params06": dict(),
    "nyquist_column": dict(NX=32, NY=32, k_f=2 * np.pi * 15.7),   # the ring reaches i = N/2
    "bounded_square": dict(bc_fac=0.0, NX=64, NY=64, nl=2, dh=[0.4, 0.6], N2=[40.0], f0=2.0, hEkb=0.0),
    "with_file_forcing": dict(bc_fac=0.0, dt_forc=2.0 ** -8, dt_forc_period=3 * 2.0 ** -8),
}


@pytest.mark.parametrize("case", list(CASES))
def test_stochastic_fp64_vs_oracle(case, tmp_path):
    p = _params(**CASES[case])
    ii, jj, env = Q.stoch_ring(p)
    _announce("stochastic FP64", case, p.NX, p.nl, "points", len(ii))
    m, o = _pair(p)
    if p.dt_forc_period:
        from scipy.io import netcdf_file
        g = Q.grid(p)[0]
        recs = (1e-3 * np.random.default_rng(3).standard_normal((3, p.nl, g.nyp2, g.nxp2))).astype(np.float32)
        f = netcdf_file(str(tmp_path / "forc.nc"), "w")
        f.createDimension("t", None)
        f.createDimension("level", p.nl)
        f.createDimension("y", g.nyp2)
        f.createDimension("x", g.nxp2)
        v = f.createVariable("q_forcing_3d", "f", ("t", "level", "y", "x"))
        for r in range(3):
            v[r] = recs[r]
        f.close()
        m.set_forcing_file(str(tmp_path / "forc.nc"))
        o.set_forcing4d(recs.astype(np.float64), p.dt_forc)
    q_start = m.q
    for _ in range(20):
        m.timestep()
        o.timestep()
    assert rel(m.q, o.field(O.Q)) < 1e-12
    # the forcing is there: the same run unforced ends elsewhere
    o2 = O.Oracle(p)
    o2.set_modes(*Q.eigmode(p))
    o2.set_q(q_start)
    o2.run(20)
    assert rel(o2.field(O.Q), o.field(O.Q)) > 1e-6


def test_stochastic_field_of_one_step():
    """From rest with every tendency off, one step leaves q = forcing * sqrt(dt)
    on layer 0: the device's synthesis against the oracle's c2r field."""
    p = _params(beta=0.0, hEkb=0.0, nu_hyper=0.0, NX=48, NY=48, k_f=2 * np.pi * 6.4)
    _announce("stochastic field", p.NX)
    m = Q.Model(p, Q.FP64)
    o = O.Oracle(p)
    o.set_modes(*Q.eigmode(p))
    o.set_stochastic(p.k_f, p.sigma_f)
    m.timestep()
    o.timestep()
    x = o.field(O.FORC)[1:-1, 1:-1]
    assert np.abs(x).max() > 0
    assert rel(m.q[0, 1:-1, 1:-1] / np.sqrt(p.dt_fixed), x) < 1e-13


def test_stochastic_fp32_vs_fp64():
    p = _params()
    _announce("stochastic FP32 vs FP64", p.NX)
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes)
    out = {}
    for prec in (Q.FP64, Q.FP32, Q.FP32_PLAIN):
        m = Q.Model(p, prec, modes)
        m.set_q(q0)
        m.run(30)
        out[prec] = m.q
    assert rel(out[Q.FP32], out[Q.FP64]) < 2e-5
    assert rel(out[Q.FP32_PLAIN], out[Q.FP64]) < 2e-5


def test_stochastic_streams_private_and_reseeded():
    p = _params()
    _announce("stochastic streams", p.NX)
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes)
    a, b, c = (Q.Model(p, Q.FP64, modes) for _ in range(3))
    for m in (a, b, c):
        m.set_q(q0)
    c.set_stoch_seed(2)
    for _ in range(6):  # interleaved: each context keeps its own stream
        a.timestep()
        b.timestep()
        c.timestep()
    assert np.array_equal(a.q, b.q)
    assert rel(c.q, a.q) > 1e-8
    # the stream restarted mid-run, and the oracle's rand() restarted (srand(1)) at the same step
    m, o = _pair(p)
    m.run(3)
    o.run(3)
    m.set_stoch_seed(1)
    o.set_stochastic(p.k_f, p.sigma_f)
    m.run(3)
    o.run(3)
    assert rel(m.q, o.field(O.Q)) < 1e-12


def test_stochastic_autotune_keeps_the_stream():
    p = _params()
    _announce("stochastic autotune", p.NX)
    modes = Q.eigmode(p)
    _, q0 = Q.make_ic(p, modes)
    a, b = Q.Model(p, Q.FP64, modes), Q.Model(p, Q.FP64, modes)
    a.set_q(q0)
    b.set_q(q0)
    a.run(3)
    b.run(3)
    b.autotune(reps=2, cache=False)
    a.run(5)
    b.run(5)
    assert rel(b.q, a.q) < 1e-12
