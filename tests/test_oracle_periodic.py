"""The periodic branch of the oracle (bc_fac = -1): invert_pv
(src/elliptic.h:32-47,103-118), the periodic ring (src/domain.h:225-256) and
the rhs on the torus, pinned by closed forms, numpy's FFT and the Arakawa
conservation identities."""
import os

import numpy as np
import pytest

import qgw_amd as Q
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def _per(NX=32, nl=1, **kw):
    dh = [4000.0] if nl == 1 else list(np.linspace(400.0, 1600.0, nl))
    N2 = [1e-5 * (1 + 0.5 * l) for l in range(max(nl - 1, 1))]
    base = dict(NX=NX, NY=NX, nl=nl, Lx=1.0e6, dh=dh, N2=N2, f0=1e-4, beta=0.0, bc_fac=-1.0, n_hyper=0,
                nu_hyper=0.0, nu=0.0, hEkb=0.0, tau0=0.0, forc_mode=0.0, dt_fixed=600.0, ic_urms=0.1,
                ic_lambda0=2.0e5, seed=11)
    base.update(kw)
    return Q.default_params(**base)


def _fourier(p, a, b, phase=0.3):
    g, X, Y, _, _ = Q.grid(p)
    kx, ky = 2 * np.pi * a / p.Lx, 2 * np.pi * b / p.Lx
    return np.cos(kx * X[None, :] + phase) * np.cos(ky * Y[:, None]), kx * kx + ky * ky


@pytest.mark.parametrize("nl", [1, 3])
@pytest.mark.parametrize("ab", [(1, 0), (3, 5), (7, 2), (16, 16)])
def test_single_fourier_mode_closed_form(nl, ab):
    """q = v_m cos(k x + phi) cos(l y) in vertical mode m gives
    psi = -q / (k^2 + l^2 + 1/Rd_m^2) (the Nyquist cosine included)."""
    p = _per(nl=nl)
    o = O.Oracle(p)
    cl2m, cm2l, iRd2 = [np.asarray(x) for x in o.modes()]
    for mode in range(nl):
        f, K2 = _fourier(p, *ab)
        v = cm2l.reshape(nl, nl)[:, mode]
        q = v[:, None, None] * f[None]
        o.set_q(q)
        o.invert_pv()
        psi = o.field(O.PSI)
        expect = -q / (K2 + iRd2[mode])
        assert np.abs(psi - expect).max() <= 1e-12 * np.abs(expect).max(), (mode, ab)
    o.close()


def test_mean_is_removed_and_constant_gives_zero():
    """The (0,0) coefficient is zeroed for every mode (src/elliptic.h:115-118)."""
    p = _per(nl=2)
    o = O.Oracle(p)
    f, K2 = _fourier(p, 2, 1)
    q = np.stack([f + 3.0, 0.5 * f - 1.0])
    o.set_q(q)
    o.invert_pv()
    psi = o.field(O.PSI)[:, 1:-1, 1:-1]
    assert np.abs(psi.mean(axis=(1, 2))).max() < 1e-15 * np.abs(psi).max()
    o.set_q(np.full_like(q, 2.5))
    o.invert_pv()
    k2min = (2 * np.pi / p.Lx) ** 2  # a constant leaves only transform round-off
    assert np.abs(o.field(O.PSI)).max() <= 1e-14 * 2.5 / k2min
    o.close()


@pytest.mark.parametrize("NX", [24, 64])
def test_invert_matches_numpy_fft(NX):
    """A random multi-layer q against numpy's 2-D FFT of the same inversion."""
    p = _per(NX=NX, nl=3)
    o = O.Oracle(p)
    cl2m, cm2l, iRd2 = [np.asarray(x) for x in o.modes()]
    nl = 3
    rng = np.random.default_rng(4)
    q = rng.standard_normal((nl, NX + 2, NX + 2))
    o.set_q(q)
    o.invert_pv()
    qi = q[:, 1:-1, 1:-1]
    k = 2 * np.pi / p.Lx * np.fft.fftfreq(NX, 1.0 / NX)
    k[NX // 2] = -k[NX // 2]  # the reference's Nyquist sign (only its square enters)
    KY, KX = np.meshgrid(k, k, indexing="ij")
    qm = np.tensordot(cl2m.reshape(nl, nl), qi, axes=(1, 0))
    Qh = np.fft.fft2(qm)
    fact = -(KX ** 2 + KY ** 2 + iRd2[:, None, None])
    Ph = np.where(fact != 0, Qh / np.where(fact != 0, fact, 1.0), 0)
    Ph[:, 0, 0] = 0
    pm = np.real(np.fft.ifft2(Ph))
    expect = np.tensordot(cm2l.reshape(nl, nl), pm, axes=(1, 0))
    psi = o.field(O.PSI)[:, 1:-1, 1:-1]
    assert np.abs(psi - expect).max() <= 1e-12 * np.abs(expect).max()
    o.close()


def test_periodic_ring():
    """After invert_pv both rings are periodic copies, corners included."""
    p = _per(NX=32, nl=2)
    o = O.Oracle(p)
    rng = np.random.default_rng(5)
    o.set_q(rng.standard_normal((2, 34, 34)))
    o.invert_pv()
    for which in (O.PSI, O.Q):
        f = o.field(which)
        wrapped = f[:, 1:-1, 1:-1]
        full = np.pad(wrapped, ((0, 0), (1, 1), (1, 1)), mode="wrap")
        assert np.array_equal(f, full), which
    o.close()


def test_jacobian_conserves_energy_and_enstrophy_on_the_torus():
    """With only the Jacobian, sum psi*f and sum q*f vanish to round-off on a
    periodic domain (Arakawa 1966)."""
    p = _per(NX=48, nl=1)
    o = O.Oracle(p)
    modes = o.modes()
    psi0, q0 = Q.make_ic(p, modes)
    o.set_q(q0)
    f = o.rhs()[0, 1:-1, 1:-1]
    psi = o.field(O.PSI)[0, 1:-1, 1:-1]
    q = o.field(O.Q)[0, 1:-1, 1:-1]
    assert abs((psi * f).sum()) <= 1e-13 * np.abs(psi * f).sum()
    assert abs((q * f).sum()) <= 1e-13 * np.abs(q * f).sum()
    o.close()


@pytest.mark.parametrize("nl", [1, 3])
def test_periodic_ic_round_trip(nl):
    """The periodic initial condition is the exact inverse of the inversion:
    invert_pv(q) returns its psi (ring included), with zero mean."""
    p = _per(NX=64, nl=nl)
    o = O.Oracle(p)
    modes = o.modes()
    psi0, q0 = Q.make_ic(p, modes)
    assert np.abs(psi0[:, 1:-1, 1:-1].mean(axis=(1, 2))).max() < 1e-14 * np.abs(psi0).max()
    o.set_q(q0)
    o.invert_pv()
    assert np.abs(o.field(O.PSI) - psi0).max() <= 1e-12 * np.abs(psi0).max()
    o.close()


@pytest.mark.parametrize("name", ["params_05.in", "params_06.in"])
def test_reference_periodic_params(name):
    """The reference's periodic test configurations read unchanged and give a
    periodic grid (their stochastic forcing, k_f, is out of scope)."""
    p = Q.read_params(os.path.join(HERE, "golden", "params", name))
    assert p.bc_fac == -1.0 and p.NX == 256 and p.nl == 1
    g = Q.grid(p)[0]
    assert (g.n, g.m, g.Delta) == (256, 256, p.Lx / 256)
