"""The oracle's 4-D forcing (calc_4d_forcing, src/forcing.h:246-289, applied
in timestep, src/timestep.h:152-163), pinned by hand-computed cases.

On a single sine mode with every other tendency switched off (beta, wind,
topography, drag and viscosity zero) the inversion gives psi proportional to
q in every layer, so the Jacobian vanishes and q is the sum of the forcing
increments alone: q(N dt) = dt * sum_n A(n dt) * F, with A the interpolated
record weight at the time before step n.  For records 0, 1, 2 with
dt_forc = 2 dt the weights run 0, .5, 1, 1.5, 2, 1, 0, .5, 1, ... (a triangle
wave that wraps from the last record to the first), worked out by hand below.
"""
import numpy as np
import pytest

import qgw_amd as Q
import oracle as O
from conftest import rel


def _quiet(**kw):
    base = dict(NX=16, NY=12, nl=2, Lx=1e6, dh=[400.0, 1600.0], N2=[1e-5], f0=1e-4, beta=0.0, tau0=0.0,
                h_topo=0.0, nu=0.0, nu_hyper=0.0, n_hyper=4, hEkb=0.0, bc_fac=0.0, dt_fixed=100.0)
    base.update(kw)
    return Q.default_params(**base)


def _mode(o, layer_w):
    n, m = o.n, o.m
    i = np.arange(n + 2)
    j = np.arange(m + 2)
    s = np.outer(np.sin(np.pi * j / (m + 1)), np.sin(2 * np.pi * i / (n + 1)))
    s[0, :] = s[-1, :] = 0
    s[:, 0] = s[:, -1] = 0
    return np.stack([w * s for w in layer_w])


def _run(p, records, dt_forc, nsteps):
    o = O.Oracle(p)
    o.set_modes(*Q.eigmode(p))
    o.set_forcing4d(records, dt_forc)
    o.run(nsteps)
    return o


def test_triangle_wave_of_three_records():
    p = _quiet()
    o0 = O.Oracle(p)
    F = _mode(o0, [1e-9, -4e-10])
    recs = np.stack([0 * F, 1 * F, 2 * F])
    dt = p.dt_fixed
    # A(t_n) for t_n = n dt, dt_forc = 2 dt: r = floor(n/2) mod 3, w2 = n/2 - floor(n/2)
    by_hand = [0.0, 0.5, 1.0, 1.5, 2.0, 1.0, 0.0, 0.5, 1.0]
    for N in (1, 4, 6, 9):
        o = _run(p, recs, 2 * dt, N)
        assert rel(o.field(O.Q), dt * sum(by_hand[:N]) * F) < 1e-12, N
        assert o.get_time()["t"] == N * dt


def test_constant_forcing_single_record():
    """dt_forc = dt_forc_period: one record, the stationary forcing of the
    reference's documentation (doc/forcing.org:264-265)."""
    p = _quiet(NX=24, NY=24, nl=3, dh=[300.0, 700.0, 2000.0], N2=[2e-5, 5e-6])
    o0 = O.Oracle(p)
    F = _mode(o0, [2e-9, 1e-9, -1e-9])
    o = _run(p, F[None], 500.0, 7)
    assert rel(o.field(O.Q), 7 * p.dt_fixed * F) < 1e-12


def test_forcing_between_record_times():
    """dt_forc not a multiple of dt: weights change every step."""
    p = _quiet(dt_fixed=30.0)
    o0 = O.Oracle(p)
    F = _mode(o0, [1e-9, 2e-9])
    a = [3.0, -1.0]
    recs = np.stack([a[0] * F, a[1] * F])
    dt_f = 70.0
    o = _run(p, recs, dt_f, 10)
    # t = 0, 30, ..., 270: A = a[r](1-w) + a[r+1 mod 2] w with r = floor(t/70) mod 2
    tot = 0.0
    for t in [0, 30, 60, 90, 120, 150, 180, 210, 240, 270]:
        x = t / 70.0
        r = int(np.floor(x)) % 2
        w = x - np.floor(x)
        tot += a[r] * (1 - w) + a[(r + 1) % 2] * w
    # the same sum worked by hand: t/70 = 0, 3/7, 6/7, 9/7, 12/7, 15/7, 18/7, 3, 24/7, 27/7
    hand = (3.0 + (3 - 4 * 3 / 7) + (3 - 4 * 6 / 7) + (-1 + 4 * 2 / 7) + (-1 + 4 * 5 / 7) + (3 - 4 * 1 / 7)
            + (3 - 4 * 4 / 7) + (-1.0) + (-1 + 4 * 3 / 7) + (-1 + 4 * 6 / 7))
    assert tot == pytest.approx(hand, rel=1e-14)
    assert rel(o.field(O.Q), p.dt_fixed * hand * F) < 1e-12


def test_forcing_only_on_the_interior_and_off_by_default():
    p = _quiet()
    o0 = O.Oracle(p)
    F = np.ones(o0.shape)  # ring values must not be used
    o = _run(p, F[None], 100.0, 1)
    q = o.field(O.Q)
    assert (q[:, 0, :] == 0).all() and (q[:, -1, :] == 0).all() and (q[:, :, 0] == 0).all()
    assert np.allclose(q[:, 1:-1, 1:-1], 100.0)
    o.set_forcing4d(None, 0.0)
    o.set_q(np.zeros(o.shape))
    o.run(2)
    assert not o.field(O.Q).any()
