"""Stochastic ring forcing (src/forcing.h:67-208, applied in timestep,
src/timestep.h:143-150): the host pieces of the product (forced points,
random stream, coefficients) and the oracle's restatement, pinned by

- a ring counted by hand and the integer geometry of the selection rule;
- the C library's own rand() stream, which the reference draws from;
- numpy.fft.irfft2, an independent real inverse FFT, for the c2r;
- the energy injection rate sigma^2 the forcing is built for
  (doc/forcing.org), as a statistical check over many steps.
"""
import ctypes
import subprocess
import sys

import numpy as np
import pytest

import qgw_amd as Q
import oracle as O
from conftest import rel


def _periodic(N=64, Lx=1.0, k_f=2 * np.pi * 8.3, sigma_f=1e-2, **kw):
    base = dict(NX=N, NY=N, nl=1, Lx=Lx, dh=[1.0], N2=[0.0], f0=1.0, beta=0.0, tau0=0.0, nu=0.0, nu_hyper=0.0,
                n_hyper=4, hEkb=0.0, bc_fac=-1.0, k_f=k_f, sigma_f=sigma_f, dt_fixed=1e-3)
    base.update(kw)
    return Q.default_params(**base)


# ---------------------------------------------------------------- ring

def test_ring_counted_by_hand():
    """test/params_05.in: NX = 256, Lx = 1, k_f = 5: a ring of radius
    5/(2 pi) = 0.796 grid wavenumbers.  Along y (dist_l) the points
    l = +-1 at k = 0 are 0.204 away; along x (dist_k) the point k = 1 at
    l = 0 is 0.204 away; nothing else is within 1/2.  Order: j outer."""
    p = _periodic(N=256, k_f=5.0, sigma_f=3.0)
    ii, jj, env = Q.stoch_ring(p)
    assert list(zip(ii, jj)) == [(1, 0), (0, 1), (0, 255)]
    # sqrt(8/((3-1)*2)) * pi * 5/(2 pi) = 2.5 sqrt(2)
    assert env == pytest.approx(2.5 * np.sqrt(2), rel=1e-15)
    o = O.Oracle(p)
    assert o.set_stochastic(p.k_f, p.sigma_f) == 3
    oi, oj = o.stoch_points()
    assert list(zip(oi, oj)) == [(1, 0), (0, 1), (0, 255)]


def _ring_by_squares(N, r):
    """The selection rule in integer geometry: the circle of radius r (grid
    wavenumbers) passes within 1/2 of point (i, l) along x iff
    (i - 1/2)^2 < r^2 - l^2 < (i + 1/2)^2 (0 <= r^2 - l^2 < 1/4 at i = 0),
    and along y likewise with |l| and i exchanged."""
    out = []
    r2 = r * r
    for j in range(N):
        l = ((j + N // 2) % N) - N // 2
        for i in range(N // 2 + 1):
            def near(a, b):  # circle crosses the segment around a at height b
                d = r2 - b * b
                if d < 0:
                    return False
                return d < 0.25 if a == 0 else (a - 0.5) ** 2 < d < (a + 0.5) ** 2
            if near(i, l) or near(abs(l), i):
                out.append((i, j))
    return out


@pytest.mark.parametrize("N,Lx,k_f", [(64, 1.0, 32.0), (128, 6.235, 10.0), (64, 1.0, 2 * np.pi * 8.3),
                                      (96, 2.0, 2 * np.pi * 9.1), (48, 1.0, 2 * np.pi * 23.7)])
def test_ring_matches_integer_geometry(N, Lx, k_f):
    p = _periodic(N=N, Lx=Lx, k_f=k_f)
    ii, jj, env = Q.stoch_ring(p)
    want = _ring_by_squares(N, k_f / (2 * np.pi) * Lx)
    assert list(zip(ii, jj)) == want
    o = O.Oracle(p)
    assert o.set_stochastic(k_f, 1.0) == len(want)
    assert list(zip(*o.stoch_points())) == want
    npts = len(want)
    assert env == pytest.approx(np.sqrt(8.0 / ((npts - 1) * 2)) * np.pi * k_f / (2 * np.pi), rel=1e-15)


def test_ring_bounded_uses_NX_points():
    """Bounded grids transform NXp1 = NX points (src/forcing.h:84-88, with NX
    already decremented by init_domain)."""
    p = _periodic(N=64, bc_fac=0.0)
    q = _periodic(N=64)
    assert [a.tolist() for a in Q.stoch_ring(p)[:2]] == [a.tolist() for a in Q.stoch_ring(q)[:2]]


# ---------------------------------------------------------------- random stream

def test_rng_is_the_c_library_rand_stream():
    code = ("import ctypes; c = ctypes.CDLL(None); "
            "a = [c.rand() for _ in range(20000)]; c.srand(1); b = [c.rand() for _ in range(20000)]; "
            "print(a == b); print(a[:3], a[-1])")
    out = subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE, text=True, check=True).stdout
    first, rest = out.splitlines()[0], out.splitlines()[1]
    assert first == "True"  # a fresh program's stream is srand(1)'s
    code2 = "import ctypes; c = ctypes.CDLL(None); print([c.rand() for _ in range(20000)])"
    ref = eval(subprocess.run([sys.executable, "-c", code2], stdout=subprocess.PIPE, text=True,
                              check=True).stdout)
    assert Q.Rng(1).next(20000).tolist() == ref
    assert Q.Rng(2).next(5).tolist() != ref[:5]


def test_coefficients_match_the_oracle_draws():
    """The product draws in an explicit order; the oracle evaluates the
    reference's Box-Muller expression with the C library's rand()."""
    p = _periodic(N=64)
    ii, jj, env = Q.stoch_ring(p)
    o = O.Oracle(p)
    o.set_stochastic(p.k_f, p.sigma_f)
    r = Q.Rng(1)
    for _ in range(3):
        a = r.coeffs(p.sigma_f, env, len(ii)).ravel()
        b = o.stoch_draw()
        assert np.array_equal(a, b)


def test_coefficient_distribution():
    """|coef|^2 = envelope^2 sigma^2 N(0,1)^2 with a uniform phase."""
    r = Q.Rng(7)
    c = r.coeffs(2.0, 0.5, 200000)
    z = c[:, 0] + 1j * c[:, 1]
    assert np.mean(np.abs(z) ** 2) == pytest.approx(1.0, rel=0.012)  # (0.5 * 2)^2, 4 sigma of chi^2_1 mean
    assert abs(np.mean(z / np.abs(z))) < 0.01


# ---------------------------------------------------------------- c2r

@pytest.mark.parametrize("N", [16, 15, 32, 12])
def test_c2r_matches_numpy_irfft2(N):
    """FFTW's unnormalised c2r = N^2 irfft2 (numpy: inverse DFT along the first
    axis, then the real inverse along the last, imaginary parts of the
    zero and Nyquist columns discarded), also for non-Hermitian input."""
    rng = np.random.default_rng(N)
    NK = N // 2 + 1
    spec = rng.standard_normal((N, NK)) + 1j * rng.standard_normal((N, NK))
    out = np.zeros(N * N)
    packed = np.ascontiguousarray(np.stack([spec.real, spec.imag], -1).ravel())
    O.lib().orc_c2r2d(packed.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), N,
                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    want = N * N * np.fft.irfft2(spec, s=(N, N))
    assert rel(out.reshape(N, N), want) < 1e-13


# ---------------------------------------------------------------- energy injection

def _ke_of_forcing(x, Lx):
    """0.5 <|grad psi|^2> with lap psi = x on a periodic square (numpy FFT)."""
    N = x.shape[0]
    k = 2 * np.pi / Lx * np.fft.fftfreq(N, 1.0 / N)
    K2 = k[None, :] ** 2 + k[:, None] ** 2
    xh = np.fft.fft2(x)
    K2[0, 0] = 1.0
    ph = -xh / K2
    ph[0, 0] = 0
    return 0.5 * np.sum(K2 * np.abs(ph) ** 2) / N ** 4


def test_energy_injection_rate():
    """Energy injection of the forcing (doc/forcing.org: the rate sigma^2).
    Exactly, per unit time, E[dKE] = 0.5 env^2 sigma^2 sum_m c_m / K_m^2 with
    c = 2 for i > 0 and 1/2 for i = 0 (Hermitian pairs), which is 1 - O(1/N_P)
    of sigma^2 on a thin ring.  Checked on the oracle's forcing fields over
    many steps and on the model's kinetic energy from rest."""
    p = _periodic(N=64, k_f=2 * np.pi * 8.3, sigma_f=1e-2)
    ii, jj, env = Q.stoch_ring(p)
    N, NP = p.NX, len(ii)
    l = ((jj + N // 2) % N) - N // 2
    c = np.where(ii == 0, 0.5, 2.0)
    rho = 0.5 * (8.3 ** 2 / (NP - 1)) * np.sum(c / (ii ** 2 + l ** 2))
    assert abs(rho - 1) < 0.02  # the sigma^2 rate of the design

    o = O.Oracle(p)
    o.set_modes(*Q.eigmode(p))
    o.set_stochastic(p.k_f, p.sigma_f)
    dt = p.dt_fixed
    inj = []
    for _ in range(400):
        o.timestep()
        x = o.field(O.FORC)[1:-1, 1:-1]
        inj.append(_ke_of_forcing(x, p.Lx) * dt)
    want = rho * p.sigma_f ** 2 * dt
    tol = 4 * np.sqrt(2.0 / (2 * NP * len(inj)))
    assert abs(np.mean(inj) / want - 1) < tol

    # the model's kinetic energy from rest, 48 segments of 40 steps (each a
    # chi^2 sample with 2 N_P degrees of freedom)
    ke = []
    for _ in range(48):
        o.set_q(np.zeros(o.shape))
        o.set_time(0.0, dt, dt, dt, 0, 0.0)
        o.run(40)
        o.invert_pv()
        psi, q = o.field(O.PSI)[0, 1:-1, 1:-1], o.field(O.Q)[0, 1:-1, 1:-1]
        ke.append(-0.5 * np.mean(psi * q))
    want = rho * p.sigma_f ** 2 * 40 * dt
    assert abs(np.mean(ke) / want - 1) < 4 * np.sqrt(2.0 / (2 * NP)) / np.sqrt(48)
