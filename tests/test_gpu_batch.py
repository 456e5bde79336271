"""GPU parity of the ensemble path (qgw_batch_*): members against the oracle,
graph replays against step-by-step launches, and the state-preserving
autotuner.  Every case prints its name before the first device call."""
import sys

import numpy as np
import pytest

import qgw_amd as Q
from qgw_amd import configs
import oracle as O
from conftest import rel

pytestmark = pytest.mark.gpu

# every knob the autotuner sets
KNOBS = (Q.TUNE_P4_TILE, Q.TUNE_P4_WAVES, Q.TUNE_T_KERNEL, Q.TUNE_DST, Q.TUNE_FFT_E, Q.TUNE_FFT_E_COL,
         Q.TUNE_FUSE, Q.TUNE_P2M_PAIRS, Q.TUNE_PSI_NT, Q.TUNE_P4_ORDER)


def _announce(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _member(NX, nl, member, tau, **kw):
    """A small member: L = 1000 km, stable hyperviscosity, dt 600 s."""
    dh = list(np.linspace(300.0, 1500.0, nl)) if nl > 1 else [4000.0]
    N2 = [1e-5 * (1.0 + 0.3 * l) for l in range(max(nl - 1, 1))]
    base = dict(NX=NX, NY=NX, nl=nl, Lx=1000e3, dh=dh, N2=N2, f0=1e-4, beta=2e-11, hEkb=3.0, n_hyper=4,
                nu_hyper=1e8, tau0=tau, forc_mode=2.0, bc_fac=0.0, dt_fixed=600.0, ic_urms=0.1,
                ic_lambda0=8 * 1000e3 / NX, seed=configs.SEED0 + member)
    base.update(kw)
    return Q.default_params(**base)


def _members():
    """Three groups: 3 x (64^2, 2 layers), 2 x (96^2, 3), 2 x (64^2, 5); winds differ per member."""
    out = []
    for NX, nl, cnt in ((64, 2, 3), (96, 3, 2), (64, 5, 2)):
        for _ in range(cnt):
            out.append(_member(NX, nl, len(out), 3e-5 * (1 + 0.25 * len(out))))
    return out


def _batch(params, prec, graph=True):
    modes = [Q.eigmode(p) for p in params]
    b = Q.Batch(params, prec, modes)
    b.use_graph(graph)
    q0 = []
    for i, p in enumerate(params):
        q0.append(Q.make_ic(p, modes[i])[1])
        b.set_q(i, q0[-1])
    return b, modes, q0


def test_batch_groups():
    params = _members()
    _announce("batch groups", len(params))
    b, _, _ = _batch(params, Q.FP64)
    ng, gm = b.groups()
    assert ng == 3
    assert gm == [0, 0, 0, 1, 1, 2, 2]
    b.close()


@pytest.mark.parametrize("nsteps", [14])
def test_batch_members_vs_oracle(nsteps):
    params = _members()
    _announce("batch FP64 vs oracle", len(params), "members", nsteps, "steps")
    b, modes, q0 = _batch(params, Q.FP64)
    b.step(nsteps)
    b.invert_pv()
    assert b.graph_launches() == nsteps // 6
    for i, p in enumerate(params):
        _announce("  member", i, p.NX, p.nl)
        o = O.Oracle(p)
        o.set_modes(*modes[i])
        o.set_q(q0[i])
        o.run(nsteps)
        o.invert_pv()
        assert rel(b.q(i), o.field(O.Q)) < 1e-12, i
        assert rel(b.psi(i), o.field(O.PSI)) < 1e-12, i
        o.close()
    assert b.get_time()["it"] == nsteps
    b.close()


@pytest.mark.parametrize("prec", [Q.FP64, Q.FP32])
def test_batch_graph_replays_match_eager(prec):
    """Each replay of the captured six-step graph, with eager launches and
    synchronisations between replays, equals the same steps launched one by one."""
    params = _members()
    _announce("batch graph vs eager", prec)
    bg, _, _ = _batch(params, prec, graph=True)
    be, _, _ = _batch(params, prec, graph=False)
    for chunk in (6, 6, 7, 12):
        bg.step(chunk)
        be.step(chunk)
        for i in range(len(params)):
            assert np.array_equal(bg.q(i), be.q(i)), (chunk, i)
        # eager work between replays: an inversion and a reduction on the same stream
        bg.invert_pv()
        be.invert_pv()
        Eg = bg.diagnostics(0)[2]
        Ee = be.diagnostics(0)[2]
        assert Eg == Ee
    assert bg.graph_launches() == 1 + 1 + 1 + 2
    assert be.graph_launches() == 0
    bg.close()
    be.close()


def test_batch_fp32_vs_fp64():
    """FP32 + Kahan members track the FP64 members (relative error after 30 steps)."""
    params = _members()
    _announce("batch FP32 vs FP64")
    b64, _, _ = _batch(params, Q.FP64)
    b32, _, _ = _batch(params, Q.FP32)
    b64.step(30)
    b32.step(30)
    for i in range(len(params)):
        assert rel(b32.q(i), b64.q(i)) < 2e-5, i
        E32, E64 = b32.diagnostics(i)[2], b64.diagnostics(i)[2]
        assert abs(E32 - E64) <= 1e-4 * abs(E64)
    b64.close()
    b32.close()


def test_batch_autotune_preserves_state():
    params = _members()
    _announce("batch autotune")
    b1, _, _ = _batch(params, Q.FP32)
    b2, _, _ = _batch(params, Q.FP32)
    b1.step(4)
    b2.step(4)
    choice = b1.autotune(reps=2, cache=False)
    assert len(choice) == 3
    for g in range(3):
        for knob in KNOBS:
            b2.tune_set(g, knob, b1.tune_get(g, knob))
    assert b1.get_time()["it"] == 4
    b1.step(8)
    b2.step(8)
    for i in range(len(params)):
        assert np.array_equal(b1.q(i), b2.q(i)), i
    b1.close()
    b2.close()


def test_model_autotune_preserves_state():
    p = _member(96, 3, 0, 4e-5)
    _announce("model autotune", p.NX, p.nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    m1, m2 = Q.Model(p, Q.FP64, modes), Q.Model(p, Q.FP64, modes)
    m1.set_q(q0)
    m2.set_q(q0)
    m1.run(3)
    m2.run(3)
    t, ms = Q.ctypes.c_int(), Q.ctypes.c_double()
    m1._chk(m1.lib.qgw_autotune(m1.ctx, 2, Q.ctypes.byref(t), Q.ctypes.byref(ms)), "autotune")
    assert m1.tune_get(Q.TUNE_P4_TILE) == t.value
    for knob in KNOBS:
        m2.tune_set(knob, m1.tune_get(knob))
    assert m1.get_time()["it"] == 3
    m1.run(5)
    m2.run(5)
    assert np.array_equal(m1.q, m2.q)
    m1.close()
    m2.close()


def test_autotune_choices_persist_across_objects():
    # a second model of the same shape takes the first one's choices from the
    # JSON cache instead of retuning
    p = _member(96, 3, 0, 4e-5)
    _announce("autotune cache", p.NX, p.nl)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]
    m1 = Q.Model(p, Q.FP32, modes)
    m1.set_q(q0)
    first = m1.autotune(reps=1, cache=False)
    key = Q.tune_key(Q.FP32, m1.n, m1.m, m1.nl, p.n_hyper, p.nu_hyper != 0)
    assert key in Q.tune_cache_load()
    m2 = Q.Model(p, Q.FP32, modes)
    m2.set_q(q0)
    again = m2.autotune(reps=1)
    assert again["cached"]
    for knob in KNOBS:
        assert m2.tune_get(knob) == m1.tune_get(knob) == int(first[str(knob)])
    m1.close()
    m2.close()


def test_batch_with_forced_members(tmp_path):
    """A batch whose members include a stochastically forced run and a 4-D
    file-forced run beside two plain ones: the forced members get groups of
    their own, stepped launch by launch beside the replayed graph of the
    plain group.  Every member matches the oracle after 12 steps (1e-12),
    the forced ones with the same realisation and records."""
    from scipy.io import netcdf_file
    base = dict(NX=64, NY=64, nl=2, Lx=1000e3, dh=[500.0, 2500.0], N2=[2e-5], f0=1e-4, beta=2e-11, hEkb=3.0,
                n_hyper=4, nu_hyper=1e8, tau0=5e-5, forc_mode=2.0, bc_fac=0.0, dt_fixed=600.0, ic_urms=0.1,
                ic_lambda0=150e3)
    ps = [Q.default_params(**dict(base, seed=11)), Q.default_params(**dict(base, seed=12)),
          Q.default_params(**dict(base, seed=13, k_f=2 * np.pi * 6 / 1000e3, sigma_f=2e-12)),
          Q.default_params(**dict(base, seed=14, dt_forc=1000.0, dt_forc_period=3000.0))]
    _announce("batch with forced members", len(ps))
    g = Q.grid(ps[3])[0]
    recs = (2e-10 * np.random.default_rng(3).standard_normal((3, 2, g.nyp2, g.nxp2))).astype(np.float32)
    recs = recs.astype(np.float64)
    path = tmp_path / "forc.nc"
    f = netcdf_file(str(path), "w")
    f.createDimension("t", None)
    f.createDimension("level", 2)
    f.createDimension("y", g.nyp2)
    f.createDimension("x", g.nxp2)
    v = f.createVariable("q_forcing_3d", "f", ("t", "level", "y", "x"))
    for r in range(3):
        v[r] = recs[r]
    f.close()
    modes = [Q.eigmode(p) for p in ps]
    ics = [Q.make_ic(p, mo)[1] for p, mo in zip(ps, modes)]
    b = Q.Batch(ps, Q.FP64, modes)
    assert b.groups()[0] == 3
    b.set_forcing_file(3, str(path))
    for i, q in enumerate(ics):
        b.set_q(i, q)
    b.step(12)
    assert b.graph_launches() == 2
    got = [b.q(i) for i in range(len(ps))]
    b.close()
    for i, p in enumerate(ps):
        o = O.Oracle(p)
        o.set_modes(*modes[i])
        o.set_q(ics[i])
        if i == 2:
            o.set_stochastic(p.k_f, p.sigma_f)
        if i == 3:
            o.set_forcing4d(recs, p.dt_forc)
        o.run(12)
        assert rel(got[i], o.field(O.Q)) < 1e-12, i
        o.close()
