"""GPU parity of the periodic branch (bc_fac = -1) against the oracle:
invert_pv (src/elliptic.h:32-47,103-118), rhs on the torus and trajectories,
in single runs and ensembles.  Every case prints its name before the first
device call."""
import os
import sys

import numpy as np
import pytest

import qgw_amd as Q
import oracle as O
from conftest import rel

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _announce(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _per(NX=64, nl=1, **kw):
    dh = [4000.0] if nl == 1 else list(np.linspace(400.0, 1600.0, nl))
    N2 = [1e-5 * (1 + 0.5 * l) for l in range(max(nl - 1, 1))]
    base = dict(NX=NX, NY=NX, nl=nl, Lx=1.0e6, dh=dh, N2=N2, f0=1e-4, beta=2e-11, bc_fac=-1.0, n_hyper=4,
                nu_hyper=1e8, nu=0.0, hEkb=3.0, tau0=5e-5, forc_mode=2.0, dt_fixed=600.0, ic_urms=0.1,
                ic_lambda0=8 * 1.0e6 / NX, seed=31 + NX + nl)
    base.update(kw)
    return Q.default_params(**base)


def _pair(p, precision=Q.FP64):
    modes = Q.eigmode(p)
    psi0, q0 = Q.make_ic(p, modes)
    m = Q.Model(p, precision, modes)
    m.set_q(q0)
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    return m, o, psi0, q0


@pytest.mark.parametrize("shape", [(32, 1), (64, 3), (96, 2), (256, 5)])
def test_periodic_invert_pv(shape):
    NX, nl = shape
    p = _per(NX, nl)
    _announce("periodic invert_pv", shape)
    m, o, psi0, q0 = _pair(p)
    m.invert_pv()
    o.invert_pv()
    assert rel(m.psi, o.field(O.PSI)) < 1e-12
    assert rel(m.psi, psi0) < 1e-12
    assert rel(m.q, o.field(O.Q)) < 1e-15  # the ring is a copy of the interior
    m.close()
    o.close()


@pytest.mark.parametrize("case", [
    dict(n_hyper=0, nu_hyper=0.0), dict(n_hyper=2, nu_hyper=1e3), dict(n_hyper=4, nu_hyper=1e8),
    dict(n_hyper=6, nu_hyper=1e13, nu=5.0), dict(n_hyper=8, nu_hyper=1e18, h_topo=200.0, w_topo=1e5)])
def test_periodic_rhs(case):
    p = _per(64, 3, **case)
    _announce("periodic rhs", case)
    m, o, _, _ = _pair(p)
    f = m.rhs()
    fo = o.rhs()
    assert rel(f[:, 1:-1, 1:-1], fo[:, 1:-1, 1:-1]) < 1e-12
    m.close()
    o.close()


@pytest.mark.parametrize("shape", [(64, 2), (96, 3)])
def test_periodic_trajectory(shape):
    NX, nl = shape
    p = _per(NX, nl)
    _announce("periodic trajectory FP64", shape)
    m, o, _, _ = _pair(p)
    m.run(50)
    o.run(50)
    m.invert_pv()
    o.invert_pv()
    assert rel(m.q, o.field(O.Q)) < 1e-12
    assert rel(m.psi, o.field(O.PSI)) < 1e-12
    m.close()
    o.close()


def test_periodic_fp32_forms():
    p = _per(64, 3)
    _announce("periodic FP32 forms")
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    out = {}
    for prec in (Q.FP64, Q.FP32, Q.FP32_PLAIN):
        m = Q.Model(p, prec, modes)
        m.set_q(q0)
        m.run(30)
        m.invert_pv()
        out[prec] = (m.q, m.diagnostics()[2:])
        m.close()
    assert rel(out[Q.FP32][0], out[Q.FP64][0]) < 1e-5
    for prec in (Q.FP32, Q.FP32_PLAIN):
        E, Z = out[prec][1]
        E64, Z64 = out[Q.FP64][1]
        assert abs(E - E64) <= 1e-5 * abs(E64) and abs(Z - Z64) <= 1e-5 * abs(Z64)


def test_periodic_jacobian_conservation_on_device():
    """Only the Jacobian: sum psi*f and sum q*f vanish to round-off on the torus."""
    p = _per(128, 1, beta=0.0, nu_hyper=0.0, n_hyper=0, hEkb=0.0, tau0=0.0)
    _announce("periodic conservation")
    m, o, _, _ = _pair(p)
    f = m.rhs()[0, 1:-1, 1:-1]
    psi = m.psi[0, 1:-1, 1:-1]
    q = m.q[0, 1:-1, 1:-1]
    assert abs((psi * f).sum()) <= 1e-12 * np.abs(psi * f).sum()
    assert abs((q * f).sum()) <= 1e-12 * np.abs(q * f).sum()
    m.close()
    o.close()


def test_reference_params_06():
    """The reference's periodic configuration (test/params_06.in: 256^2, one
    layer, beta, hyperviscosity n_hyper = 6, drag) without its stochastic
    forcing, 20 FP64 steps against the oracle."""
    p = Q.read_params(os.path.join(HERE, "golden", "params", "params_06.in"))
    p.k_f = 0.0
    p.dt_fixed = 1e-3
    p.ic_kind = 1
    p.ic_urms = 1.0
    p.ic_lambda0 = 0.1
    p.seed = 606
    _announce("params_06 periodic", p.NX, p.n_hyper)
    m, o, _, _ = _pair(p)
    m.run(20)
    o.run(20)
    m.invert_pv()
    o.invert_pv()
    assert rel(m.q, o.field(O.Q)) < 1e-12
    assert rel(m.psi, o.field(O.PSI)) < 1e-12
    m.close()
    o.close()


def test_periodic_batch():
    """Periodic members in one ensemble against the oracle (14 steps: two
    graph replays and two single steps)."""
    params = [_per(64, 2, tau0=3e-5 * (1 + i), seed=70 + i) for i in range(3)] + [_per(96, 3, seed=80)]
    _announce("periodic batch", len(params))
    modes = [Q.eigmode(p) for p in params]
    b = Q.Batch(params, Q.FP64, modes)
    q0 = []
    for i, p in enumerate(params):
        q0.append(Q.make_ic(p, modes[i])[1])
        b.set_q(i, q0[-1])
    assert b.groups()[0] == 2
    b.step(14)
    b.invert_pv()
    for i, p in enumerate(params):
        o = O.Oracle(p)
        o.set_modes(*modes[i])
        o.set_q(q0[i])
        o.run(14)
        o.invert_pv()
        assert rel(b.q(i), o.field(O.Q)) < 1e-12, i
        assert rel(b.psi(i), o.field(O.PSI)) < 1e-12, i
        o.close()
    b.close()
