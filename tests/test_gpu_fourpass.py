"""The inversion in four passes (QGW_TUNE_FUSE = 1: projections inside the row
transforms, column transforms on column tiles of the mode-major spectral
array) against the six-pass form and against the oracle: FP64 to the
oracle's 1e-12, FP32 within the FP32 bar (2e-5), on square, non-square and
3*2^k grids, odd and even layer counts, and ensemble groups whose members
are stacked along the layer axis (each with its own mode matrices).

Every case prints its name and shape before the first device call.
"""
import sys

import numpy as np
import pytest

import qgw_amd as Q
from qgw_amd import configs
import oracle as O
from conftest import rel

pytestmark = pytest.mark.gpu


def _announce(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _params(NX, nl, **kw):
    dh = list(np.linspace(300.0, 1500.0, nl)) if nl > 1 else [4000.0]
    N2 = [1e-5 * (1.0 + 0.3 * l) for l in range(max(nl - 1, 1))]
    base = dict(NX=NX, NY=NX, nl=nl, Lx=1000e3 * NX / 256, dh=dh, N2=N2, f0=1e-4, beta=2e-11, hEkb=3.0,
                n_hyper=4, nu_hyper=1e8, tau0=5e-5, forc_mode=2.0, bc_fac=0.0, dt_fixed=600.0, ic_urms=0.1,
                ic_lambda0=80e3, seed=configs.SEED0 + NX + nl)
    base.update(kw)
    return Q.default_params(**base)


def _run(p, prec, modes, q0, fuse, steps):
    m = Q.Model(p, prec, modes)
    m.tune_set(Q.TUNE_FUSE, fuse)
    assert m.tune_get(Q.TUNE_FUSE) == fuse
    m.set_q(q0)
    m.invert_pv()
    psi = m.psi
    m.run(steps)
    m.invert_pv()
    out = (psi, m.q, m.psi)
    m.close()
    return out


@pytest.mark.parametrize("NX,NY,nl,kw", [
    (64, 64, 1, dict(Ld=40e3)),
    (128, 128, 2, dict()),
    (128, 96, 3, dict(bc_fac=0.5)),
    (256, 256, 5, dict(h_topo=100.0, w_topo=60e3)),
    (192, 192, 4, dict()),
    (1024, 1024, 10, dict()),
])
def test_four_pass_vs_six_and_oracle(NX, NY, nl, kw):
    p = _params(NX, nl, NY=NY, **kw)
    _announce("four-pass inversion", NX, NY, nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    steps = 3
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.invert_pv()
    psi_o = o.field(O.PSI).copy()
    o.run(steps)
    o.invert_pv()
    q_o, p_o = o.field(O.Q).copy(), o.field(O.PSI).copy()
    o.close()
    f4 = _run(p, Q.FP64, modes, q0, 1, steps)
    f6 = _run(p, Q.FP64, modes, q0, 0, steps)
    for a, b in zip(f4, (psi_o, q_o, p_o)):
        assert rel(a, b) < 1e-12
    assert rel(f4[0], f6[0]) < 1e-13
    s4 = _run(p, Q.FP32, modes, q0, 1, steps)
    for a, b in zip(s4, (psi_o, q_o, p_o)):
        assert rel(a, b) < 2e-5


def test_four_pass_batch_members():
    # three members of one group (stacked layers, each its own modes and wind),
    # 12 steps: two graph replays of the four-pass step, against the oracle
    ps = [_params(128, 3, seed=configs.SEED0 + 40 + i, tau0=4e-5 + 1e-5 * i,
                  dh=[300.0 + 50 * i, 700.0 - 50 * i, 2000.0]) for i in range(3)]
    _announce("four-pass batch", 128, len(ps))
    modes = [Q.eigmode(p) for p in ps]
    ics = [Q.make_ic(p, mo)[1] for p, mo in zip(ps, modes)]
    for prec, tol in ((Q.FP64, 1e-12), (Q.FP32, 2e-5)):
        b = Q.Batch(ps, prec, modes)
        assert b.groups()[0] == 1
        b.tune_set(0, Q.TUNE_FUSE, 1)
        for i, q in enumerate(ics):
            b.set_q(i, q)
        b.step(12)
        got = [b.q(i) for i in range(len(ps))]
        b.close()
        for i, p in enumerate(ps):
            o = O.Oracle(p)
            o.set_modes(*modes[i])
            o.set_q(ics[i])
            o.run(12)
            assert rel(got[i], o.field(O.Q)) < tol, (prec, i)
            o.close()


@pytest.mark.parametrize("fft_e,col", [(8, 8), (16, 16), (8, 16), (16, 8)])
def test_four_pass_radix_schedules(fft_e, col):
    p = _params(1024, 2, Lx=4000e3)
    _announce("four-pass radix schedules", fft_e, col)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    m = Q.Model(p, Q.FP64, modes)
    m.tune_set(Q.TUNE_FUSE, 1)
    m.tune_set(Q.TUNE_FFT_E, fft_e)
    m.tune_set(Q.TUNE_FFT_E_COL, col)
    m.set_q(q0)
    m.run(3)
    m.invert_pv()
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.run(3)
    o.invert_pv()
    assert rel(m.q, o.field(O.Q)) < 1e-12
    assert rel(m.psi, o.field(O.PSI)) < 1e-12
    m.close()
    o.close()
