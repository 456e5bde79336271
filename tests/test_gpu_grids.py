"""Grids beyond the power-of-two and 3*2^k lengths (the reference's FFTW
plans take any NX, NY: src/elliptic.h:63-66), against the oracle, whose own
transform is a mixed-radix FFT of any length.

FFT sine transforms with radix-5 and radix-7 stages:
- NX = 100, 200, 500, 1000 and 2000 (common ocean set-ups; transform
  lengths 2 NX = 2^a 5^b) and NX = 224 (2^6 7), square and mixed: FP64 against
  the oracle at 1e-12 in every form (odd-extension, half-length one block
  per row pair, half-length pipelined, four passes);
- the FP32 forms against the oracle at the FP32 bar (2e-5);
- periodic NX = 100 and 200 (lengths 100 and 200) at 1e-12.

Dense sine transforms (matrix products on the matrix cores) for lengths the
FFT kernels do not cover (NX = 150, 1200, 2100 and a 100 x 60 grid):
- FP64 inversion and steps against the oracle at 1e-12;
- the FP32 form at 2e-5;
- a batch of such members: the graph replay against launch by launch (bit
  for bit) and against the oracle.

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


def _grid_params(NX, NY, nl, **kw):
    dh = list(np.linspace(400.0, 2000.0, nl))
    N2 = [1e-5 * (1.0 + 0.5 * l) for l in range(max(nl - 1, 1))]
    Lx = 10e3 * NX  # 10 km spacing
    base = dict(NX=NX, NY=NY, nl=nl, Lx=Lx, dh=dh, N2=N2, f0=1e-4, beta=2e-11, hEkb=3.0, n_hyper=4,
                nu_hyper=2e8, tau0=5e-5, forc_mode=2.0, bc_fac=0.0, dt_fixed=900.0, ic_urms=0.05,
                ic_lambda0=100e3, seed=configs.SEED0 + NX + NY + nl, h_topo=100.0, w_topo=50e3)
    base.update(kw)
    return Q.default_params(**base)


def _oracle_run(p, modes, q0, steps):
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.invert_pv()
    psi = o.field(O.PSI).copy()
    o.run(steps)
    o.invert_pv()
    out = (psi, o.field(O.Q).copy(), o.field(O.PSI).copy())
    o.close()
    return out


_FORMS = [("odd-extension", 0, 0), ("half-length", 1, 0), ("half-length pipelined", 2, 0), ("four passes", 2, 1)]


def _device_run(p, prec, modes, q0, steps, dst=None, fuse=None):
    m = Q.Model(p, prec, modes)
    if dst is not None:
        m.tune_set(Q.TUNE_DST, dst)
        assert m.tune_get(Q.TUNE_DST) == dst
    if fuse is not None:
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


@pytest.mark.timeout(300)
@pytest.mark.parametrize("NX,NY,nl,steps", [(100, 100, 3, 6), (200, 200, 2, 6), (200, 100, 2, 6), (224, 224, 2, 6),
                                            (500, 500, 2, 4), (1000, 1000, 2, 3), (2000, 2000, 2, 2)])
def test_radix57_fft_fp64_vs_oracle(NX, NY, nl, steps):
    p = _grid_params(NX, NY, nl, bc_fac=0.5 if NX == 100 else 0.0)
    modes = Q.eigmode(p)
    psi0, q0 = Q.make_ic(p, modes)
    psi_o, q_o, p_o = _oracle_run(p, modes, q0, steps)
    for name, dst, fuse in _FORMS:
        _announce("radix-5/7 FFT FP64", name, NX, NY, nl)
        psi_d, q_d, p_d = _device_run(p, Q.FP64, modes, q0, steps, dst, fuse)
        assert rel(psi_d, psi_o) < 1e-12, name
        assert rel(q_d, q_o) < 1e-12, name
        assert rel(p_d, p_o) < 1e-12, name
        if p.bc_fac == 0.0:
            assert rel(psi_d, psi0) < 1e-12, name  # the seeded q is the exact inverse of psi0


@pytest.mark.parametrize("NX,NY,nl", [(100, 100, 3), (1000, 1000, 2), (224, 200, 2)])
def test_radix57_fft_fp32_vs_oracle(NX, NY, nl):
    p = _grid_params(NX, NY, nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    _, q_o, p_o = _oracle_run(p, modes, q0, 4)
    for name, dst, fuse in _FORMS[1:]:
        _announce("radix-5/7 FFT FP32", name, NX, NY, nl)
        _, q_d, p_d = _device_run(p, Q.FP32, modes, q0, 4, dst, fuse)
        assert rel(q_d, q_o) < 2e-5, name
        assert rel(p_d, p_o) < 2e-5, name


def test_radix57_default_form():
    """Grids of radix-5/7 lengths take the FFT transforms by default (the
    half-length forms), not the dense products."""
    for NX, NY in ((100, 100), (2000, 2000), (224, 200)):
        p = _grid_params(NX, NY, 2)
        _announce("radix-5/7 default form", NX, NY)
        m = Q.Model(p, Q.FP64, Q.eigmode(p))
        assert m.tuning()["dst"].startswith("half-length"), (NX, NY)
        m.close()


@pytest.mark.parametrize("NX,nl", [(100, 2), (200, 3)])
def test_radix57_periodic_vs_oracle(NX, nl):
    p = _grid_params(NX, NX, nl, bc_fac=-1.0, Lx=1.0e6, ic_lambda0=8 * 1.0e6 / NX, h_topo=0.0)
    _announce("radix-5 periodic FP64", NX, nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    psi_d, q_d, p_d = _device_run(p, Q.FP64, modes, q0, 4)
    psi_o, q_o, p_o = _oracle_run(p, modes, q0, 4)
    assert rel(psi_d, psi_o) < 1e-12
    assert rel(q_d, q_o) < 1e-12
    assert rel(p_d, p_o) < 1e-12


@pytest.mark.timeout(300)
@pytest.mark.parametrize("NX,NY,nl,steps", [(150, 150, 3, 6), (100, 60, 2, 6), (1200, 1200, 2, 3),
                                            (2100, 2100, 2, 2)])
def test_dense_transforms_fp64_vs_oracle(NX, NY, nl, steps):
    p = _grid_params(NX, NY, nl, bc_fac=0.5 if NX == 150 else 0.0)
    _announce("dense sine transforms FP64", NX, NY, nl)
    modes = Q.eigmode(p)
    psi0, q0 = Q.make_ic(p, modes)
    m = Q.Model(p, Q.FP64, modes)
    assert m.tuning()["dst"].startswith("dense")
    m.set_q(q0)
    m.invert_pv()
    psi_d = m.psi
    m.run(steps)
    m.invert_pv()
    q_d, p_d = m.q, m.psi
    m.close()
    psi_o, q_o, p_o = _oracle_run(p, modes, q0, steps)
    assert rel(psi_d, psi_o) < 1e-12
    assert rel(q_d, q_o) < 1e-12
    assert rel(p_d, p_o) < 1e-12
    if p.bc_fac == 0.0:
        assert rel(psi_d, psi0) < 1e-12  # the seeded q is the exact inverse of psi0


@pytest.mark.parametrize("NX,NY,nl", [(150, 150, 3), (240, 120, 2)])
def test_dense_transforms_fp32_vs_oracle(NX, NY, nl):
    p = _grid_params(NX, NY, nl)
    _announce("dense sine transforms FP32", NX, NY, nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    m = Q.Model(p, Q.FP32, modes)
    assert m.tuning()["dst"].startswith("dense")
    m.set_q(q0)
    m.run(4)
    m.invert_pv()
    _, q_o, p_o = _oracle_run(p, modes, q0, 4)
    assert rel(m.q, q_o) < 2e-5
    assert rel(m.psi, p_o) < 2e-5
    m.close()


def test_dense_transforms_batch_graph():
    ps = [_grid_params(150, 150, 2, seed=configs.SEED0 + i, tau0=4e-5 + 1e-5 * i) for i in range(3)]
    _announce("dense sine transforms, batch graph", 150, len(ps))
    modes = [Q.eigmode(p) for p in ps]
    ics = [Q.make_ic(p, mo)[1] for p, mo in zip(ps, modes)]
    out = {}
    for graph in (True, False):
        b = Q.Batch(ps, Q.FP64, modes)
        b.use_graph(graph)
        for i, q in enumerate(ics):
            b.set_q(i, q)
        b.step(12)
        out[graph] = [b.q(i) for i in range(len(ps))]
        if graph:
            assert b.graph_launches() == 2
        b.close()
    for i in range(len(ps)):
        assert np.array_equal(out[True][i], out[False][i]), i
    _, q_o, _ = _oracle_run(ps[1], modes[1], ics[1], 12)
    assert rel(out[True][1], q_o) < 1e-12
