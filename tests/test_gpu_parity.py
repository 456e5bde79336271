"""GPU parity of the hot path (libqgw_hip.so) against the oracle.

Every case prints its name, shape and seed before the first device call so
that a failing run names its case.  The oracle and the device get the same
vertical modes (the host library's) and the same seeded FP64 inputs.
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


def _pair(p, precision=Q.FP64, seed=None):
    modes = Q.eigmode(p)
    psi0, q0 = Q.make_ic(p, modes, seed)
    m = Q.Model(p, precision, modes)
    m.set_q(q0)
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    return m, o, psi0, q0


def _small(NX, nl, **kw):
    dh = list(np.linspace(300.0, 1500.0, nl)) if nl > 1 else [4000.0]
    N2 = [1e-5 * (1.0 + 0.3 * l) for l in range(max(nl - 1, 1))]
    base = dict(NX=NX, NY=NX, nl=nl, Lx=1000e3, dh=dh, N2=N2, f0=1e-4, beta=2e-11, hEkb=3.0, n_hyper=4,
                nu_hyper=1e8, tau0=5e-5, forc_mode=2.0, bc_fac=0.0, dt_fixed=600.0, ic_urms=0.1,
                ic_lambda0=8 * 1000e3 / NX, seed=configs.SEED0 + NX + nl)
    base.update(kw)
    return Q.default_params(**base)


def test_invert_pv_c1():
    p = configs.c1()
    _announce("invert_pv c1", p.NX, p.nl)
    m, o, psi0, q0 = _pair(p)
    m.invert_pv()
    o.invert_pv()
    assert rel(m.psi, o.field(O.PSI)) < 1e-13
    assert rel(m.psi, psi0) < 1e-13
    assert np.array_equal(m.q[:, 0, :], o.field(O.Q)[:, 0, :])


def test_rhs_c1():
    p = configs.c1()
    _announce("rhs c1", p.NX, p.nl)
    m, o, _, _ = _pair(p)
    f = m.rhs()
    fo = o.rhs()
    assert rel(f, fo) < 1e-12


@pytest.mark.parametrize("NX,nl,kw", [
    (8, 1, dict()),
    (16, 2, dict(n_hyper=0, nu_hyper=1e-7)),
    (32, 3, dict(bc_fac=1.0)),
    (48, 2, dict(bc_fac=0.5, n_hyper=2, nu_hyper=1e4)),
    (64, 4, dict(h_topo=200.0, w_topo=100e3)),
    (96, 5, dict(n_hyper=6, nu_hyper=1e12, nu=20.0)),
    (128, 1, dict(Ld=30e3, n_hyper=8, nu_hyper=1e14)),
    (64, 16, dict()),
])
def test_rhs_small(NX, nl, kw):
    p = _small(NX, nl, **kw)
    _announce("rhs small", NX, nl, kw)
    m, o, _, _ = _pair(p)
    f = m.rhs()
    fo = o.rhs()
    assert rel(f, fo) < 1e-12
    assert rel(m.psi, o.field(O.PSI)) < 1e-13
    assert np.allclose(m.q, o.field(O.Q), rtol=0, atol=1e-12 * np.abs(o.field(O.Q)).max())


def test_nonsquare():
    p = _small(64, 3, NY=48)
    _announce("rhs nonsquare", p.NX, p.NY, p.nl)
    m, o, _, _ = _pair(p)
    assert rel(m.rhs(), o.rhs()) < 1e-12


def test_trajectory_c1_100():
    p = configs.c1()
    _announce("trajectory c1 100 steps", p.NX, p.nl)
    m, o, _, _ = _pair(p)
    m.run(100)
    o.run(100)
    m.invert_pv()
    o.invert_pv()
    assert rel(m.q, o.field(O.Q)) < 1e-12
    assert rel(m.psi, o.field(O.PSI)) < 1e-12
    tm, to = m.get_time(), o.get_time()
    assert tm["t"] == to["t"] and tm["it"] == to["it"] == 100


@pytest.mark.parametrize("NX,nl,kw", [
    # no-slip walls with hyperviscosity are unstable in the reference's scheme
    # (the ring of lap^n psi is Dirichlet), so the no-slip cases use nu only
    (32, 3, dict(bc_fac=1.0, nu_hyper=0.0, nu=50.0, h_topo=100.0, w_topo=50e3)),
    (48, 2, dict(bc_fac=1.0, n_hyper=0, nu_hyper=1e-7, nu=20.0)),
    (96, 2, dict(nu=5.0)),
])
def test_trajectory_small(NX, nl, kw):
    p = _small(NX, nl, **kw)
    _announce("trajectory small", NX, nl, kw)
    m, o, _, _ = _pair(p)
    m.run(30)
    o.run(30)
    assert rel(m.q, o.field(O.Q)) < 1e-12


def test_adaptive_dt_matches_oracle():
    p = _small(64, 2, dt_fixed=0.0, cfl=0.2)
    _announce("adaptive dt", p.NX, p.nl)
    m, o, _, _ = _pair(p)
    m.set_next_output(3 * 86400.0)
    o.set_time(0.0, o.get_time()["dt"], o.get_time()["dt0"], o.get_time()["dt1"], 0, 3 * 86400.0)
    for _ in range(12):
        m.timestep()
        o.timestep()
        assert m.get_time()["dt"] == pytest.approx(o.get_time()["dt"], rel=1e-12)
    assert rel(m.q, o.field(O.Q)) < 1e-12


def test_diagnostics_match_oracle():
    p = configs.c1()
    _announce("diagnostics c1", p.NX, p.nl)
    m, o, _, _ = _pair(p)
    m.invert_pv()
    o.invert_pv()
    Em, Zm, Etm, Ztm = m.diagnostics()
    Eo, Zo, Eto, Zto = o.diagnostics()
    assert np.allclose(Em, Eo, rtol=1e-12) and np.allclose(Zm, Zo, rtol=1e-12)
    assert Etm == pytest.approx(Eto, rel=1e-12) and Ztm == pytest.approx(Zto, rel=1e-12)


@pytest.mark.parametrize("col", [8, 16])
@pytest.mark.parametrize("fft_e", [8, 16])
def test_fp32_radix_schedules(fft_e, col):
    """The FP32 form with each radix schedule of the 1024-point row and column
    transforms against the FP64 oracle: 6 steps within the FP32 budget of the
    other FP32 checks (2e-5)."""
    p = _small(1024, 2, NY=1024, Lx=4000e3)
    _announce("fp32 radix schedule rows", fft_e, "columns", col)
    m, o, _, _ = _pair(p, precision=Q.FP32)
    m.tune_set(Q.TUNE_FFT_E, fft_e)
    m.tune_set(Q.TUNE_FFT_E_COL, col)
    m.run(6)
    o.run(6)
    assert rel(m.q, o.field(O.Q)) < 2e-5
    m.close()
    o.close()


@pytest.mark.parametrize("prec,tol", [(Q.FP32, 2e-5), (Q.FP32_PLAIN, 2e-5)])
def test_fp32_forms_c1(prec, tol):
    p = configs.c1()
    _announce("fp32 form c1", prec)
    m, o, _, _ = _pair(p, precision=prec)
    m.run(50)
    o.run(50)
    m.invert_pv()
    o.invert_pv()
    assert rel(m.q, o.field(O.Q)) < tol
    Em, Zm, Etm, Ztm = m.diagnostics()
    Eo, Zo, Eto, Zto = o.diagnostics()
    assert abs(Etm - Eto) / abs(Eto) < 1e-4
    assert abs(Ztm - Zto) / abs(Zto) < 1e-4


def test_upload_download_roundtrip_fp32_keeps_residual():
    """The FP32 form keeps q's FP64 rounding residual in its compensation (16
    bits below q's last place): the download returns q to ~2^-39 relative,
    where FP32 alone gives ~2^-24 (3e-8), and a dropped or mis-scaled residual
    gives 1e-8 or worse."""
    p = configs.c1()
    _announce("fp32 upload residual")
    m = Q.Model(p, Q.FP32)
    _, q0 = Q.make_ic(p)
    m.set_q(q0)
    assert rel(m.q, q0) < 2e-12
    # without the compensation: plain FP32 rounding
    assert rel(q0.astype(np.float32).astype(np.float64), q0) > 1e-9


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("prec", [Q.FP64, Q.FP32])
def test_p4_tiles_agree(tile, prec):
    # every tile shape of the tendency kernel, on a grid with partial edge
    # tiles, walls with slip, topography and the q ring
    p = _small(96, 3, bc_fac=0.5, nu=10.0, nu_hyper=1e8, h_topo=150.0, w_topo=80e3, NY=64)
    _announce("p4 tile", Q.P4_TILES[tile], "prec", prec)
    m, o, _, _ = _pair(p, precision=prec)
    m.tune_set(Q.TUNE_P4_TILE, tile)
    assert m.tune_get(Q.TUNE_P4_TILE) == tile
    tol = 1e-12 if prec == Q.FP64 else 2e-5
    assert rel(m.rhs(), o.rhs()) < tol
    m.run(5)
    o.run(5)
    assert rel(m.q, o.field(O.Q)) < tol


@pytest.mark.parametrize("nl", [20, 32])
def test_large_nl_fp64_vs_oracle(nl):
    p = _small(64, nl)
    _announce("large nl FP64 vs oracle", p.NX, nl)
    m, o, _, _ = _pair(p)
    m.run(8)
    o.run(8)
    m.invert_pv()
    o.invert_pv()
    assert rel(m.q, o.field(O.Q)) < 1e-12
    assert rel(m.psi, o.field(O.PSI)) < 1e-12


@pytest.mark.parametrize("nl", [3, 20, 32])
def test_projection_matrix_cores_vs_valu(nl):
    """FP32 projections on the matrix cores agree with the VALU projections and
    track the FP64 run."""
    p = _small(64, nl)
    _announce("projection MFMA vs VALU", p.NX, nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    runs = {}
    for name, prec, tk in (("valu", Q.FP32, 1), ("mfma", Q.FP32, 2), ("mfma16", Q.FP32, 3), ("fp64", Q.FP64, 0)):
        m = Q.Model(p, prec, modes)
        m.tune_set(Q.TUNE_T_KERNEL, tk)
        m.set_q(q0)
        m.run(8)
        m.invert_pv()
        runs[name] = (m.q, m.psi)
        m.close()
    assert rel(runs["mfma"][1], runs["valu"][1]) < 1e-5
    assert rel(runs["mfma"][0], runs["valu"][0]) < 1e-5
    assert rel(runs["mfma"][1], runs["fp64"][1]) < 1e-5
    assert rel(runs["mfma16"][1], runs["mfma"][1]) < 1e-5


@pytest.mark.parametrize("nl", [33, 48, 70])
def test_many_layers_vs_oracle(nl):
    """More than 32 layers (the reference takes up to 1000, src/qg.c:45): the
    projections run over 32 x 32 blocks of the mode matrices, a partial block
    at 33 and 70.  FP64 inversion and 4 steps against the oracle (1e-12), and
    the FP32 form within the FP32 bar (2e-5)."""
    p = _small(48, nl)
    _announce("many layers vs oracle", p.NX, nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.invert_pv()
    psi_o = o.field(O.PSI).copy()
    o.run(4)
    q_o = o.field(O.Q).copy()
    o.close()
    for prec, tol in ((Q.FP64, 1e-12), (Q.FP32, 2e-5)):
        m = Q.Model(p, prec, modes)
        m.set_q(q0)
        m.invert_pv()
        assert rel(m.psi, psi_o) < tol, prec
        m.run(4)
        assert rel(m.q, q_o) < tol, prec
        m.close()


@pytest.mark.parametrize("n_hyper,periodic", [(10, False), (12, False), (18, False), (10, True)])
def test_deep_hyperviscosity_vs_oracle(n_hyper, periodic):
    """n_hyper above 8 (calc_diff nests n_hyper/2 + 1 Laplacians,
    src/dynamics.h:25-43): the streaming update on bounded grids and the
    wrapped tile on a periodic one, 4 steps against the oracle (1e-12)."""
    # |lambda| dt = nu (8/Delta^2)^(n_hyper/2) dt = 0.1 at the grid scale (AB3-stable);
    # the model sets the sign of nu_hyper itself (src/domain.h:210)
    kw = dict(n_hyper=n_hyper, nu_hyper=0.1 * (1000e3 / 96) ** n_hyper / (8.0 ** (n_hyper // 2) * 600.0))
    # free-slip walls: a partial-slip ring re-injected at every nested level
    # (2 bc_fac/Delta^2 psi) diverges within a step at these depths, in the
    # reference scheme as here
    kw.update(bc_fac=-1.0 if periodic else 0.0)
    p = _small(96, 2, **kw)
    _announce("deep hyperviscosity vs oracle", n_hyper, "periodic" if periodic else "bounded")
    m, o, _, _ = _pair(p)
    m.run(4)
    o.run(4)
    assert rel(m.q, o.field(O.Q)) < 1e-12
    m.invert_pv()
    o.invert_pv()
    assert rel(m.psi, o.field(O.PSI)) < 1e-12
    m.close()
    o.close()


@pytest.mark.parametrize("prec", [Q.FP64, Q.FP32])
@pytest.mark.parametrize("hyper", [(4, 0.0), (0, 1e3), (2, -1e4), (4, 1e8), (6, -1e9), (8, 1e10)])
def test_p4_stream_paths(prec, hyper):
    """The streaming P4 (interior walks with buffer loads, edge walks with the
    wall rules) against the LDS-tile P4 on a grid with interior strips and
    chunks, for every nesting depth of the Laplacians (n_hyper 0..8, and no
    hyperviscosity)."""
    n_hyper, nu_h = hyper
    p = _small(256, 3, n_hyper=n_hyper, nu_hyper=nu_h, NY=192, Lx=1000e3)
    _announce("p4 stream paths", prec, n_hyper, nu_h)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    out = {}
    for tile in (1, 4, 5):
        m = Q.Model(p, prec, modes)
        m.tune_set(Q.TUNE_P4_TILE, tile)
        m.set_q(q0)
        m.run(3)
        out[tile] = (m.q, m.download(Q.FIELD_F0))
        m.close()
    tol = 1e-12 if prec == Q.FP64 else 1e-5
    for tile in (4, 5):
        assert rel(out[tile][1], out[1][1]) < tol
        assert rel(out[tile][0], out[1][0]) < tol


@pytest.mark.parametrize("prec", [Q.FP64, Q.FP32, Q.FP32_PLAIN])
@pytest.mark.parametrize("case", [(600, 320, 4, 1e8), (500, 200, 12, None), (1024, 96, 2, -1e4)])
def test_p4_aligned_blocks(prec, case):
    """The streaming P4 over aligned column blocks (4 wavefronts over 224
    output columns, 192 at n_hyper 12; a partial last block whose trailing
    wavefronts have no columns; interior and edge walks in one block) against
    the streaming P4 and the oracle, on grids whose width is not a multiple of
    the block's."""
    NX, NY, n_hyper, nu_h = case
    bc = 0.5
    if n_hyper > 8:
        # AB3-stable at the grid scale, free-slip walls (a partial-slip ring
        # re-injected at every nested level diverges at these depths)
        nu_h, bc = 0.1 * (1000e3 / NX) ** n_hyper / (8.0 ** (n_hyper // 2) * 600.0), 0.0
    p = _small(NX, 2, NY=NY, n_hyper=n_hyper, nu_hyper=nu_h, bc_fac=bc, h_topo=150.0, w_topo=80e3)
    _announce("p4 aligned blocks", prec, case)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    out = {}
    for tile in (4, 5):
        m = Q.Model(p, prec, modes)
        m.tune_set(Q.TUNE_P4_TILE, tile)
        m.set_q(q0)
        m.run(3)
        out[tile] = (m.q, m.download(Q.FIELD_F0))
        m.close()
    # FP32: the two forms may contract multiply-adds differently where their
    # interior and edge walks differ, and six nested Laplacians amplify that
    # rounding in the tendency (FP64 agreement to 1e-12 shows the arithmetic is
    # the same)
    tol = 1e-12 if prec == Q.FP64 else (1e-5 if n_hyper <= 8 else 1e-4)
    assert rel(out[5][1], out[4][1]) < tol
    assert rel(out[5][0], out[4][0]) < tol
    if prec == Q.FP64:
        o = O.Oracle(p)
        o.set_modes(*modes)
        o.set_q(q0)
        o.run(3)
        assert rel(out[5][0], o.field(O.Q)) < 1e-12
        o.close()


@pytest.mark.parametrize("prec", [Q.FP64, Q.FP32, Q.FP32_PLAIN])
@pytest.mark.parametrize("case", [(600, 320, 4, 1e8, {}), (500, 200, 12, None, {}),
                                  (1024, 96, 2, -1e4, {}), (448, 256, 0, 0.0, dict(nu=20.0, nl=3))])
def test_p4_reversed_walks(prec, case):
    """The interior walks of odd row chunks run up their chunk
    (QGW_TUNE_P4_REV 1, the default) or down (0), and the walks are dealt
    layers or chunks first (QGW_TUNE_P4_ORDER), in the streaming P4 and over
    aligned column blocks: the same bits in all four combinations (q and the
    tendency after three steps: every stencil -- both Jacobians, the nested
    Laplacians, nu lap q, the drag -- is evaluated in the same order in both
    directions; a y difference taken the wrong way round, or a walk given the
    wrong layer or chunk, changes the bits), and the oracle's results (FP64).
    Row chunks of 16 rows on these grids, so most walks are interior and half
    of them run up."""
    NX, NY, n_hyper, nu_h, kw = case
    kw = dict(kw)
    nl = kw.pop("nl", 2)
    bc = 0.5
    if n_hyper > 8:
        nu_h, bc = 0.1 * (1000e3 / NX) ** n_hyper / (8.0 ** (n_hyper // 2) * 600.0), 0.0
    p = _small(NX, nl, NY=NY, n_hyper=n_hyper, nu_hyper=nu_h, bc_fac=bc, h_topo=150.0, w_topo=80e3, **kw)
    _announce("p4 reversed walks", prec, case)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    for tile in (4, 5):
        out = {}
        for rev, order in ((0, 0), (1, 0), (0, 1), (1, 1)):
            m = Q.Model(p, prec, modes)
            m.tune_set(Q.TUNE_P4_TILE, tile)
            m.tune_set(Q.TUNE_P4_REV, rev)
            m.tune_set(Q.TUNE_P4_ORDER, order)
            assert m.tune_get(Q.TUNE_P4_REV) == rev and m.tune_get(Q.TUNE_P4_ORDER) == order
            m.set_q(q0)
            m.run(3)
            out[rev, order] = (m.q, m.download(Q.FIELD_F0))
            m.close()
        for key in ((1, 0), (0, 1), (1, 1)):
            assert np.array_equal(out[key][0], out[0, 0][0]), (tile, key)
            assert np.array_equal(out[key][1], out[0, 0][1]), (tile, key)
        out[1] = out[1, 1]
        if prec == Q.FP64:
            o = O.Oracle(p)
            o.set_modes(*modes)
            o.set_q(q0)
            o.run(3)
            assert rel(out[1][0], o.field(O.Q)) < 1e-12
            o.close()


@pytest.mark.parametrize("prec", [Q.FP64, Q.FP32])
@pytest.mark.parametrize("fuse,dst", [(0, 2), (0, 1), (1, 2)])
def test_psi_store_policy(prec, fuse, dst):
    """Plain and non-temporal psi stores of the inverse row passes
    (QGW_TUNE_PSI_NT; pipelined and one-block rows, six and four passes) give
    the same results bit for bit, and the oracle's (FP64)."""
    p = _small(256, 3, NY=192, bc_fac=0.5)
    _announce("psi store policy", prec, fuse, dst)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    out = {}
    for nt in (0, 1):
        m = Q.Model(p, prec, modes)
        m.tune_set(Q.TUNE_FUSE, fuse)
        m.tune_set(Q.TUNE_DST, dst)
        m.tune_set(Q.TUNE_PSI_NT, nt)
        assert m.tune_get(Q.TUNE_PSI_NT) == nt
        m.set_q(q0)
        m.run(3)
        q = m.q
        m.invert_pv()  # psi of the new q (the inversion also rewrites q's wall ring)
        out[nt] = (q, m.psi)
        m.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    if prec == Q.FP64:
        o = O.Oracle(p)
        o.set_modes(*modes)
        o.set_q(q0)
        o.run(3)
        assert rel(out[1][0], o.field(O.Q)) < 1e-12
        o.close()


@pytest.mark.parametrize("shape", [(64, 64, 3), (96, 128, 2), (256, 192, 5), (384, 256, 2)])
def test_dst_half_vs_full(shape):
    """Sine transforms through half-length FFTs (one block per row pair, and
    pipelined persistent blocks) against the odd-extension FFTs (FP64, 6
    steps, and the oracle), on power-of-two and 3*2^k grids: the row passes
    of a 3*2^k length run the pipelined blocks with the LDS Stockham, and the
    mixed grids pair power-of-two and 3*2^k transforms in one step."""
    NX, NY, nl = shape
    p = _small(NX, nl, NY=NY)
    _announce("dst half vs full", shape)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    out = {}
    for form in (0, 1, 2):
        m = Q.Model(p, Q.FP64, modes)
        m.tune_set(Q.TUNE_DST, form)
        assert m.tune_get(Q.TUNE_DST) == form
        m.set_q(q0)
        m.run(6)
        m.invert_pv()
        out[form] = (m.q, m.psi)
        m.close()
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.run(6)
    o.invert_pv()
    for form in (0, 1, 2):
        assert rel(out[form][0], o.field(O.Q)) < 1e-12, form
        assert rel(out[form][1], o.field(O.PSI)) < 1e-12, form
    assert rel(out[1][1], out[0][1]) < 1e-13
    assert rel(out[2][1], out[0][1]) < 1e-13


_BIG = {}


def _big_case():
    """2048^2 x 4 layers: inputs and the oracle's inversion and 3-step state
    (computed once, about 10 s on the host)."""
    if not _BIG:
        p = _small(2048, 4, Lx=4000e3, nu_hyper=1e4, ic_lambda0=200e3)
        modes = Q.eigmode(p)
        q0 = Q.make_ic(p, modes)[1]
        o = O.Oracle(p)
        o.set_modes(*modes)
        o.set_q(q0)
        o.invert_pv()
        psi_o = o.field(O.PSI).copy()
        o.run(3)
        q_o = o.field(O.Q).copy()
        o.close()
        _BIG.update(p=p, modes=modes, q0=q0, psi_o=psi_o, q_o=q_o)
    return _BIG


@pytest.mark.parametrize("NX,nl", [(1536, 3), (3072, 2)])
def test_pipelined_rows_3x2k_many_items(NX, nl):
    """Pipelined row transforms of a 3*2^k length (half lengths 1536 and 3072,
    the LDS Stockham with its twiddle table in LDS) with several items per
    persistent block: 1536^2 x 3 layers (3070 items) and 3072^2 x 2 (3071),
    FP64 inversion and 3 steps against the oracle, and the one-block form
    against the pipelined one."""
    p = _small(NX, nl, Lx=2e6 * NX / 1024, nu_hyper=1e4, ic_lambda0=150e3)
    _announce("pipelined 3*2^k rows, many items per block", p.NX, p.nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.invert_pv()
    psi_o = o.field(O.PSI).copy()
    o.run(3)
    q_o = o.field(O.Q).copy()
    o.close()
    out = {}
    for form in (1, 2):
        m = Q.Model(p, Q.FP64, modes)
        m.tune_set(Q.TUNE_DST, form)
        m.set_q(q0)
        m.invert_pv()
        psi = m.psi
        m.run(3)
        out[form] = (psi, m.q)
        m.close()
        assert rel(psi, psi_o) < 1e-12, form
        assert rel(out[form][1], q_o) < 1e-12, form
    assert rel(out[2][0], out[1][0]) < 1e-13


@pytest.mark.parametrize("fft_e", [8, 16])
def test_pipelined_rows_many_items(fft_e):
    """Pipelined row transforms with several items per persistent block:
    2048^2 x 4 layers has 4094 (row, layer pair) items against at most a few
    per compute unit resident, so each block stages, prefetches and transforms
    a run of items.  FP64 inversion and 3 steps against the oracle, and the
    one-block form against the pipelined one."""
    c = _big_case()
    p, modes, q0, psi_o, q_o = c["p"], c["modes"], c["q0"], c["psi_o"], c["q_o"]
    _announce("pipelined rows, many items per block", p.NX, p.nl, "fft_e", fft_e)
    out = {}
    for form in (1, 2):
        m = Q.Model(p, Q.FP64, modes)
        m.tune_set(Q.TUNE_DST, form)
        m.tune_set(Q.TUNE_FFT_E, fft_e)
        m.tune_set(Q.TUNE_FFT_E_COL, fft_e)
        m.set_q(q0)
        m.invert_pv()
        psi = m.psi
        m.run(3)
        out[form] = (psi, m.q)
        m.close()
        assert rel(psi, psi_o) < 1e-12, form
        assert rel(out[form][1], q_o) < 1e-12, form
    assert rel(out[2][0], out[1][0]) < 1e-13


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("fft_e,col", [(8, 8), (16, 16), (8, 16), (16, 8)])
def test_dst_half_radix_schedules(fft_e, col, form):
    """Both radix schedules of the half-length FFTs (1024-point rows and
    columns, each pass with either), one block per row pair and pipelined,
    against the oracle."""
    p = _small(1024, 2, NY=1024, Lx=4000e3)
    _announce("dst half radix schedule rows", fft_e, "columns", col, "form", form)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    m = Q.Model(p, Q.FP64, modes)
    m.tune_set(Q.TUNE_DST, form)
    m.tune_set(Q.TUNE_FFT_E, fft_e)
    m.tune_set(Q.TUNE_FFT_E_COL, col)
    m.set_q(q0)
    m.run(4)
    m.invert_pv()
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.run(4)
    o.invert_pv()
    assert rel(m.q, o.field(O.Q)) < 1e-12
    assert rel(m.psi, o.field(O.PSI)) < 1e-12
    m.close()
    o.close()


@pytest.mark.parametrize("nl", [16, 20, 32])
def test_projection_matrix_cores_vs_oracle(nl):
    """The FP32 matrix-core projections (32- and 16-column tiles) against the
    FP64 oracle directly: the inversion of the seeded state and 8 steps, at
    the FP32 bar of the other FP32 checks (2e-5); a transposed or misplaced
    coefficient of the 32 x 32 contraction gives O(1)."""
    p = _small(64, nl)
    _announce("projection MFMA vs oracle", p.NX, nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.invert_pv()
    psi_o = o.field(O.PSI).copy()
    o.run(8)
    q_o = o.field(O.Q).copy()
    o.close()
    for tk in (2, 3):
        m = Q.Model(p, Q.FP32, modes)
        m.tune_set(Q.TUNE_T_KERNEL, tk)
        m.set_q(q0)
        m.invert_pv()
        assert rel(m.psi, psi_o) < 2e-5, tk
        m.run(8)
        assert rel(m.q, q_o) < 2e-5, tk
        m.close()


@pytest.mark.parametrize("NX,nl", [(96, 2), (384, 3), (1536, 3)])
def test_pipelined_rows_3x2k_fp32(NX, nl):
    """FP32 row transforms of 3*2^k half lengths (96, 384, 1536): the
    pipelined persistent blocks against the one-block form (the same Stockham
    code and twiddle values: 1e-6), and both against the FP64 oracle within
    the FP32 bar (2e-5)."""
    p = _small(NX, nl, Lx=2e6 * NX / 1024, nu_hyper=1e4, ic_lambda0=max(150e3, 8 * 2e6 / 1024))
    _announce("pipelined 3*2^k rows FP32", p.NX, p.nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    o.run(3)
    q_o = o.field(O.Q).copy()
    o.close()
    out = {}
    for form in (1, 2):
        m = Q.Model(p, Q.FP32, modes)
        m.tune_set(Q.TUNE_DST, form)
        m.set_q(q0)
        m.run(3)
        out[form] = m.q
        m.close()
        assert rel(out[form], q_o) < 2e-5, form
    assert rel(out[2], out[1]) < 1e-6


@pytest.mark.parametrize("NX,prec,form", [(1536, Q.FP64, 1), (1536, Q.FP32, 2), (2048, Q.FP64, 2), (2048, Q.FP32, 2),
                                          (384, Q.FP64, 1), (96, Q.FP64, 1), (64, Q.FP64, 2)])
def test_default_sine_transform_form(NX, prec, form):
    """Default row-transform form: the pipelined persistent blocks, except
    FP64 rows of a 3*2^k half length, which take one block per row pair (the
    faster of the two there); both forms stay selectable."""
    p = _small(NX, 2)
    _announce("default sine-transform form", NX, prec)
    m = Q.Model(p, prec, Q.eigmode(p))
    assert m.tune_get(Q.TUNE_DST) == form
    for f in (1, 2):
        m.tune_set(Q.TUNE_DST, f)
        assert m.tune_get(Q.TUNE_DST) == f
    m.close()
