"""Pins of the oracle's right-hand-side terms, wall ring and topography.

Each term of `rhs` (src/dynamics.h:52-71, src/domain.h:12-33) is switched on
alone and compared with a closed form worked from the physics the reference
documents (doc/usage.org: f = f0 + beta*y, the wind forcing formula, the
Ekman drag formula), not from the oracle's code: on a single sine mode the
inversion is exact (continuous wavenumbers, src/domain.h:170-173) and the
centred differences reduce to trigonometric identities.  The wall ring
(src/domain.h:308-348) is pinned by Thom's no-slip wall vorticity and the
corner ownership, the bathtub topography (src/domain.h:200-207) at worked
points.  Every pin also runs against the oracle with a planted mistake
(oracle.MUT) and must fail there: the name of each case says what it catches.
"""
import numpy as np
import pytest

import oracle as O
import qgw_amd as Q
from conftest import rel

LX = 1.0e6


@pytest.fixture(autouse=True)
def _faithful_oracle():
    O.mutate("none")
    yield
    O.mutate("none")


def _grid(NX, NY=None):
    NY = NX if NY is None else NY
    D = LX / NX
    jj, ii = np.meshgrid(np.arange(NY + 1), np.arange(NX + 1), indexing="ij")
    return D, ii, jj


def _sine(NX, a, b):
    """sin(a pi x/Lx) sin(b pi y/Lx) on the storage grid of a square basin."""
    n = NX - 1
    _, ii, jj = _grid(NX)
    return np.sin(np.pi * a * ii / (n + 1)) * np.sin(np.pi * b * jj / (n + 1))


def _mode_state(o, NX, a, b, k, A):
    """q = cm2l[:, k] (-kappa^2) A s, so that psi_l = cm2l[l, k] A s exactly."""
    cl2m, cm2l, iRd2 = o.modes()
    nl = o.nl
    v = np.asarray(cm2l).reshape(nl, nl)[:, k]
    kap2 = (np.pi * a / LX) ** 2 + (np.pi * b / LX) ** 2 + iRd2[k]
    s = _sine(NX, a, b)
    o.set_q(-kap2 * A * v[:, None, None] * s[None])
    return v * A, s


def _params(nl=1, **kw):
    base = dict(NX=32, NY=32, nl=nl, Lx=LX, dh=[400.0, 3600.0][:nl] if nl <= 2 else [300.0, 700.0, 3000.0],
                N2=[1e-5, 4e-6][:max(nl - 1, 1)], f0=1e-4)
    base.update(kw)
    return Q.default_params(**base)


def _interior(f):
    return f[..., 1:-1, 1:-1]


def _check(err, tol, mutation):
    if mutation == "none":
        assert err < tol, err
    else:
        assert err > 1e3 * tol, (mutation, err)


# ------------------------------------------------------------------- beta

@pytest.mark.parametrize("mutation", ["none", "beta_sign", "beta_axis"])
def test_beta_term_is_minus_beta_v(mutation):
    # f = f0 + beta*y gives dq/dt = -beta * dpsi/dx; centred in x on
    # psi = A sin(a pi i/(n+1)) sin(b pi j/(n+1)):
    # (psi[i+1] - psi[i-1])/(2 Delta) = A sin(a pi/(n+1))/Delta cos(a pi i/(n+1)) sin(b pi j/(n+1))
    O.mutate(mutation)
    NX, a, b, A, beta = 32, 3, 5, 2.0e3, 2e-11
    o = O.Oracle(_params(beta=beta))
    amp, s = _mode_state(o, NX, a, b, 0, A)
    f = o.rhs()
    D, ii, jj = _grid(NX)
    n = NX - 1
    expect = -beta * amp[0] * np.sin(a * np.pi / (n + 1)) / D * np.cos(a * np.pi * ii / (n + 1)) * \
        np.sin(b * np.pi * jj / (n + 1))
    _check(rel(_interior(f[0]), _interior(expect)), 1e-12, mutation)


# ------------------------------------------------------------------- nu lap q

@pytest.mark.parametrize("mutation", ["none", "nu_psi"])
def test_harmonic_viscosity_acts_on_q(mutation):
    # nu lap(q) with the 5-point eigenvalue mu = -(4/Delta^2)(sin^2(a pi/2(n+1)) + sin^2(b pi/2(n+1)))
    O.mutate(mutation)
    NX, a, b, A, nu = 32, 4, 2, 1.0e3, 500.0
    o = O.Oracle(_params(nu=nu))
    _mode_state(o, NX, a, b, 0, A)
    q = o.field(O.Q).copy()
    f = o.rhs()
    D, _, _ = _grid(NX)
    n = NX - 1
    mu = -(4 / D ** 2) * (np.sin(a * np.pi / (2 * (n + 1))) ** 2 + np.sin(b * np.pi / (2 * (n + 1))) ** 2)
    _check(rel(_interior(f), _interior(nu * mu * q)), 1e-12, mutation)


# ------------------------------------------------------------------- wind

@pytest.mark.parametrize("mutation", ["none", "wind_layer", "wind_dh", "wind_lx"])
def test_wind_forcing_worked_points(mutation):
    # doc/usage.org: dq[0]/dt = -tau0/dh[0] m pi/Ly sin(m pi y/Ly), top layer
    # only, uniform in x.  A 40 x 24 basin (Ly = 0.6 Lx), dh = [400, 3600]:
    #   m = 2 at y = Ly/4 (j = 6): sin(pi/2) = 1 -> -tau0/400 * 2 pi/Ly
    #   m = 2 at y = Ly/2 (j = 12): sin(pi) = 0
    #   m = 2 at y = 3Ly/4 (j = 18): +tau0/400 * 2 pi/Ly
    #   m = 1 at y = Ly/2: -tau0/400 * pi/Ly;  at y = Ly/6 (j = 4): half of that
    O.mutate(mutation)
    tau0 = 8e-5
    Ly = 24 * LX / 40
    worst = 0.0
    for m, pts in ((2.0, {6: -1.0, 12: 0.0, 18: 1.0}), (1.0, {12: -1.0, 4: -0.5})):
        o = O.Oracle(_params(nl=2, NX=40, NY=24, tau0=tau0, forc_mode=m))
        f = o.rhs()
        unit = tau0 / 400.0 * m * np.pi / Ly
        for j, c in pts.items():
            worst = max(worst, np.abs(f[0, j, 1:-1] - c * unit).max() / unit)
        worst = max(worst, np.abs(f[1]).max() / unit)                    # bottom layer untouched
        worst = max(worst, np.abs(f[0, 1:-1, 1:-1] - f[0, 1:-1, 1:2]).max() / unit)  # uniform in x
    _check(worst, 1e-13, mutation)


# ------------------------------------------------------------------- drag

@pytest.mark.parametrize("mutation", ["none", "drag_dh", "drag_half", "drag_q"])
def test_bottom_drag_rate(mutation):
    # hEkb = 2 r dh[nl-1]/f0 makes the drag -r * zeta on the bottom layer
    # (doc/usage.org: -hEkb f0/(2 dh[nl-1]) lap(psi)), zeta = mu psi on a sine mode
    O.mutate(mutation)
    NX, a, b, A, r = 32, 2, 3, 1.0e3, 1e-7
    hEkb = 2 * r * 3600.0 / 1e-4
    o = O.Oracle(_params(nl=2, hEkb=hEkb))
    amp, s = _mode_state(o, NX, a, b, 1, A)  # baroclinic: the layers differ in sign
    f = o.rhs()
    D, _, _ = _grid(NX)
    n = NX - 1
    mu = -(4 / D ** 2) * (np.sin(a * np.pi / (2 * (n + 1))) ** 2 + np.sin(b * np.pi / (2 * (n + 1))) ** 2)
    expect = np.zeros_like(f)
    expect[1] = -r * mu * amp[1] * s
    _check(rel(_interior(f), _interior(expect)), 1e-12, mutation)


# ------------------------------------------------------------------- topography

def _arakawa_linear_factor(c, n):
    # Arakawa's J(psi, y) on a sine mode: the three Jacobians give
    # (2 + cos(c pi/(n+1)))/3 times the centred x-difference (c: the y wavenumber);
    # J(psi, x) likewise, minus the centred y-difference with the x wavenumber
    return (2 + np.cos(c * np.pi / (n + 1))) / 3


@pytest.mark.parametrize("axis", ["y", "x"])
@pytest.mark.parametrize("mutation", ["none", "topo_swap", "topo_layer"])
def test_topography_ramp_is_topographic_beta(mutation, axis):
    # -f0/dh[nl-1] J(psi_b, h) with h = g*y is a topographic beta f0 g/dh[nl-1]
    # acting on the bottom layer only (sign as for beta: a bottom rising to the north)
    O.mutate(mutation)
    NX, a, b, A, g = 32, 3, 2, 1.0e3, 1e-3
    o = O.Oracle(_params(nl=2))
    amp, s = _mode_state(o, NX, a, b, 1, A)
    D, ii, jj = _grid(NX)
    n = NX - 1
    o.field(O.TOPO)[...] = g * (jj if axis == "y" else ii) * D
    f = o.rhs()
    bt = 1e-4 * g / 3600.0
    expect = np.zeros_like(f)
    if axis == "y":  # J(psi, y) = psi_x
        dpsi = amp[1] * np.sin(a * np.pi / (n + 1)) / D * np.cos(a * np.pi * ii / (n + 1)) * \
            np.sin(b * np.pi * jj / (n + 1))
        expect[1] = -bt * _arakawa_linear_factor(b, n) * dpsi
    else:            # J(psi, x) = -psi_y
        dpsi = amp[1] * np.sin(b * np.pi / (n + 1)) / D * np.sin(a * np.pi * ii / (n + 1)) * \
            np.cos(b * np.pi * jj / (n + 1))
        expect[1] = bt * _arakawa_linear_factor(a, n) * dpsi
    _check(rel(_interior(f), _interior(expect)), 1e-12, mutation)


def test_arakawa_linear_factor_against_continuum():
    # the factor tends to 1 (the continuous psi_x) as the mode is resolved
    assert _arakawa_linear_factor(1, 4095) == pytest.approx(1.0, abs=1e-6)
    assert _arakawa_linear_factor(16, 16 - 1) == pytest.approx(1 / 3)  # 2 + cos(pi) = 1


# ------------------------------------------------------------------- superposition

def test_rhs_is_the_sum_of_its_terms():
    # with every term on, the tendency equals the Jacobian-only tendency plus
    # each term's own increment (no term dropped or counted twice together)
    base = dict(nl=2, NX=32, NY=32, n_hyper=4)
    terms = dict(beta=2e-11, nu=300.0, tau0=8e-5, hEkb=7.0, h_topo=200.0, w_topo=80e3, nu_hyper=1e9)
    rng = np.random.default_rng(11)
    q = np.zeros((2, 33, 33))
    q[:, 1:-1, 1:-1] = rng.standard_normal((2, 31, 31)) * 1e-5

    def f_of(**kw):
        o = O.Oracle(_params(**base, **kw))
        o.set_q(q)
        return o.rhs()

    f_j = f_of()
    total = f_j.copy()
    for k, v in terms.items():
        kw = {k: v}
        if k == "h_topo":
            kw["w_topo"] = terms["w_topo"]
        if k == "w_topo":
            continue
        total += f_of(**kw) - f_j
    f_all = f_of(**terms)
    assert rel(_interior(f_all), _interior(total)) < 1e-12
    for k in ("beta", "nu", "tau0", "hEkb", "h_topo", "nu_hyper"):   # each term is active
        kw = {k: terms[k]}
        if k == "h_topo":
            kw["w_topo"] = terms["w_topo"]
        assert rel(_interior(f_of(**kw)), _interior(f_j)) > 1e-6, k


# ------------------------------------------------------------------- wall ring

@pytest.mark.parametrize("bc_fac", [1.0, 0.5])
@pytest.mark.parametrize("mutation", ["none", "bc_factor"])
def test_wall_ring_is_thom_no_slip_vorticity(mutation, bc_fac):
    # no slip (bc_fac = 1): psi = 0 and dpsi/dn = 0 at the wall; for psi = d^2
    # (d: distance to the wall) the wall vorticity d2psi/dn2 = 2, and Thom's
    # formula 2 psi_1/Delta^2 gives it exactly; partial slip scales it by bc_fac
    O.mutate(mutation)
    NX, NY = 24, 16
    o = O.Oracle(_params(NX=NX, NY=NY, bc_fac=bc_fac))
    D, ii, jj = _grid(NX, NY)
    X, Y, Ly = ii * D, jj * D, NY * D
    worst = 0.0
    for wall, dist in (("S", Y), ("N", Ly - Y), ("W", X), ("E", LX - X)):
        psi = np.zeros(o.shape)
        psi[0, 1:-1, 1:-1] = dist[1:-1, 1:-1] ** 2
        o.field(O.PSI)[...] = psi
        ring = o.adjust_bc(np.zeros(o.shape), 1)[0]
        side = {"S": ring[0, 1:-1], "N": ring[-1, 1:-1], "W": ring[1:-1, 0], "E": ring[1:-1, -1]}[wall]
        worst = max(worst, np.abs(side - 2 * bc_fac).max())
        assert np.abs(o.adjust_bc(np.ones(o.shape), 0)[0][[0, -1], :]).max() == 0.0  # psi ring: 0
    _check(worst, 1e-12, mutation)


@pytest.mark.parametrize("mutation", ["none", "bc_order"])
def test_wall_ring_corners_belong_to_west_and_east(mutation):
    # src/domain.h:316-346 writes S, N, then W and E over the full column,
    # so a corner takes 2 bc_fac/Delta^2 psi of its x-neighbour (on the S/N ring)
    O.mutate(mutation)
    NX = 16
    o = O.Oracle(_params(NX=NX, bc_fac=1.0))
    D = LX / NX
    psi = np.random.default_rng(3).standard_normal(o.shape)  # nonzero ring to tell the owners apart
    o.field(O.PSI)[...] = psi
    ring = o.adjust_bc(np.zeros(o.shape), 1)[0]
    c = 2.0 / D ** 2
    p = psi[0]
    want = np.array([c * p[0, 1], c * p[0, -2], c * p[-1, 1], c * p[-1, -2]])
    got = np.array([ring[0, 0], ring[0, -1], ring[-1, 0], ring[-1, -1]])
    _check(np.abs(got - want).max() / np.abs(want).max(), 1e-14, mutation)


def test_calc_diff_ring_uses_psi_at_every_level():
    # calc_diff (src/dynamics.h:26-43) resets the ring of each nested
    # Laplacian to 2 bc_fac/Delta^2 * psi_adjacent: on a sine mode psi = s the
    # first level is mu s inside, and a point next to one wall gains
    # (2 bc_fac/Delta^4) psi there in the second level (the sine's own odd
    # extension would put 0 on the ring)
    NX, a, b = 32, 3, 4
    bc = 0.5
    o = O.Oracle(_params(NX=NX, bc_fac=bc, n_hyper=2))
    D, _, _ = _grid(NX)
    n = NX - 1
    s = _sine(NX, a, b)
    o.field(O.PSI)[...] = s[None]
    o.calc_diff()
    lapn = o.field(O.LAPN)[0]
    mu = -(4 / D ** 2) * (np.sin(a * np.pi / (2 * (n + 1))) ** 2 + np.sin(b * np.pi / (2 * (n + 1))) ** 2)
    c = 2 * bc / D ** 2
    expect = mu * mu * s
    expect[1:-1, 1] += c * s[1:-1, 1] / D ** 2
    expect[1:-1, -2] += c * s[1:-1, -2] / D ** 2
    expect[1, 1:-1] += c * s[1, 1:-1] / D ** 2
    expect[-2, 1:-1] += c * s[-2, 1:-1] / D ** 2
    assert rel(lapn[1:-1, 1:-1], expect[1:-1, 1:-1]) < 1e-12
    ring = np.concatenate([lapn[0, 1:-1] - c * s[1, 1:-1], lapn[1:-1, 0] - c * s[1:-1, 1]])
    assert np.abs(ring).max() < 1e-12 * c


# ------------------------------------------------------------------- bathtub topography

def _topo_worked(h, w, NX, NY):
    # topo = h (1 - shelf(x) shelf(Lx-x) shelf(y) shelf(Ly-y)), shelf(d) = 1 - exp(-d^2/(2 w^2));
    # Lx = 1e6 on 40 x 30 cells (Delta = 25 km, Ly = 750 km), w = 50 km (2 cells):
    #   every wall point:                      h            (a shelf of 0 distance is 0)
    #   x = w, y = Ly/2 (i = 2, j = 15):       h e^{-1/2}   (the y shelves are 1 - e^{-28.1}: 1.2e-12 h off)
    #   x = Lx/2, y = Ly - w (i = 20, j = 28): h e^{-1/2}
    #   x = y = w (i = j = 2):                 h (1 - (1 - e^{-1/2})^2)
    #   x = Lx/2, y = Ly/2:                    0 to 1e-16 h
    e = np.exp(-0.5)
    return {(2, 15): h * e, (20, 28): h * e, (2, 2): h * (1 - (1 - e) ** 2), (20, 15): 0.0}


@pytest.mark.parametrize("mutation", ["none", "topo_lx", "shelf_width"])
def test_bathtub_topography_worked_points(mutation):
    O.mutate(mutation)
    h, w, NX, NY = 300.0, 50e3, 40, 30
    p = _params(NX=NX, NY=NY, h_topo=h, w_topo=w)
    o = O.Oracle(p)
    t = o.field(O.TOPO)
    worst = max(np.abs(t[0, :] - h).max(), np.abs(t[-1, :] - h).max(),
                np.abs(t[:, 0] - h).max(), np.abs(t[:, -1] - h).max()) / h
    for (i, j), v in _topo_worked(h, w, NX, NY).items():
        worst = max(worst, abs(t[j, i] - v) / h)
    _check(worst, 1e-11, mutation)
    if mutation == "none":
        # the product's host topography (libqgw_host.so) at the same points
        th = Q.topo(p)
        for (i, j), v in _topo_worked(h, w, NX, NY).items():
            assert abs(th[j, i] - v) < 1e-11 * h
        assert np.abs(th[0, :] - h).max() < 1e-12 * h
