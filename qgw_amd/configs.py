"""Named workloads of the QG timestep hot path, as qgw parameter sets.

The physical descriptions (grid, layers, reduced gravity, drag, viscosity,
wind, step) are mapped onto the reference's parameter keys (src/qg.c:113-137,
meanings in doc/usage.org:10-66):

  * N2 at interface l = g'_l / dhc_l, dhc_l = (dh_l + dh_{l+1})/2 (src/eigmode.h:19-20)
  * bottom drag r -> hEkb = 2 r dh_bottom / f0 (the drag term of src/dynamics.h:68)
  * wind tau0 (1 - cos 2 pi y/L) [N/m^2] -> forc_mode = 2, tau0 = tau/rho0, rho0 = 1000
  * biharmonic nu4 -> n_hyper = 4, nu_hyper = nu4 (src/domain.h:210 gives it the damping sign)
  * bounded free-slip walls, bc_fac = 0
  * a fixed step dt_fixed (extension; the reference adapts dt every step)

Every workload starts from the seeded k^-3 initial condition of
qgwh_make_ic, seed 20261203 + member id.
"""
import math

import numpy as np

from . import default_params

SEED0 = 20261203
RHO0 = 1000.0


def _n2_from_gprime(dh, gprime):
    dh = list(dh)
    return [g / (0.5 * (dh[l] + dh[l + 1])) for l, g in enumerate(gprime)]


def _logdepth_layers(nl, H=4000.0, z1=40.0):
    """Layer bottoms z_l = z1 (H/z1)^(l/(nl-1)), l = 0..nl-1, thicknesses dh_l = z_l - z_{l-1}."""
    zb = [z1 * (H / z1) ** (l / (nl - 1)) for l in range(nl)]
    zb[-1] = H
    dh = [zb[0]] + [zb[l] - zb[l - 1] for l in range(1, nl)]
    return dh, zb


def _exp_stratification(zb, seed, efold=800.0):
    """N^2(z) = N0^2 exp(-z/efold) at the interfaces, N0^2 ~ U[1e-5, 5e-5] from the seed."""
    rng = np.random.default_rng(seed)
    n0 = rng.uniform(1e-5, 5e-5)
    return [n0 * math.exp(-z / efold) for z in zb[:-1]], n0


def c1(**over):
    """256^2 x 2 layers, L = 2000 km, H = [500, 3500], g' = 0.02, drag 1e-7, nu4 = 1e9, dt 1800 s."""
    dh = [500.0, 3500.0]
    p = default_params(NX=256, NY=256, nl=2, Lx=2000e3, dh=dh, N2=_n2_from_gprime(dh, [0.02]),
                       f0=1e-4, beta=2e-11, hEkb=2 * 1e-7 * dh[-1] / 1e-4, n_hyper=4, nu_hyper=1e9,
                       tau0=0.0, forc_mode=2.0, bc_fac=0.0, dt_fixed=1800.0, nsteps=100, ic_kind=1,
                       seed=SEED0, ic_urms=0.1, ic_lambda0=200e3)
    return p.set(**over)


def c2(as_written=False, **over):
    """1024^2 x 3 layers, L = 4000 km, H = [350, 750, 2900], g' = [0.025, 0.0125], double-gyre
    wind 0.08 N/m^2, drag 1e-7, dt 1200 s.  nu4 = 1e10 as written is AB3-unstable at this
    grid and step (|lambda| dt ~ 1.34 > 0.545); the default uses nu4 = 2e9."""
    dh = [350.0, 750.0, 2900.0]
    p = default_params(NX=1024, NY=1024, nl=3, Lx=4000e3, dh=dh, N2=_n2_from_gprime(dh, [0.025, 0.0125]),
                       f0=1e-4, beta=2e-11, hEkb=2 * 1e-7 * dh[-1] / 1e-4, n_hyper=4,
                       nu_hyper=1e10 if as_written else 2e9, tau0=0.08 / RHO0, forc_mode=2.0, bc_fac=0.0,
                       dt_fixed=1200.0, nsteps=2000, ic_kind=1, seed=SEED0, ic_urms=0.05, ic_lambda0=200e3)
    return p.set(**over)


def c3(**over):
    """2048^2 x 10 layers, L = 4000 km, log-depth layers summing to 4000 m, seeded exponential
    stratification (e-folding 800 m), double-gyre wind 0.08 N/m^2, drag 1e-7, nu4 = 2e8
    (stable limit ~3.4e8), dt 900 s.  Initial RMS velocity 0.05 m/s (as c2): at 0.1 m/s the
    eddy peaks in the two thin top layers (40 m and 27 m) exceed the AB3 advective limit of
    dt = 900 s and the run diverges near step 185, whereas 0.05 m/s runs the full 1000 steps."""
    nl = 10
    dh, zb = _logdepth_layers(nl)
    N2, _ = _exp_stratification(zb, SEED0)
    p = default_params(NX=2048, NY=2048, nl=nl, Lx=4000e3, dh=dh, N2=N2, f0=1e-4, beta=2e-11,
                       hEkb=2 * 1e-7 * dh[-1] / 1e-4, n_hyper=4, nu_hyper=2e8, tau0=0.08 / RHO0,
                       forc_mode=2.0, bc_fac=0.0, dt_fixed=900.0, nsteps=1000, ic_kind=1, seed=SEED0,
                       ic_urms=0.05, ic_lambda0=200e3)
    return p.set(**over)


def scaled(p, NX, **over):
    """The same physics on an NX x NX grid (Lx unchanged)."""
    q = type(p)()
    import ctypes
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(p))
    q.NX = NX
    q.NY = NX
    return q.set(**over)


# Ensemble size classes: (members, NX, nl, nu4).  nu4 keeps the hyperviscous
# AB3 factor 26 nu4 dt / Delta^4 at or below ~0.27 for dt = 1200 s (the AB3
# stability limit is 0.545).
C4_CLASSES = [(32, 512, 2, 1e10), (32, 1024, 3, 2e9), (24, 1536, 5, 4e8), (8, 2048, 4, 1.25e8)]


def c4_member(NX, nl, nu4, member, tau, dt=1200.0):
    """One ensemble member: the physics of the 1024^2 x 3 workload (L = 4000 km, f0 = 1e-4,
    beta = 2e-11, drag 1e-7, double-gyre wind) at NX^2 x nl.  nl = 3 uses that workload's layers;
    other layer counts use log-depth layers summing to 4000 m with the seeded exponential
    stratification of the 2048^2 x 10 workload (c3).  Seed 20261203 + member."""
    if nl == 3:
        dh = [350.0, 750.0, 2900.0]
        N2 = _n2_from_gprime(dh, [0.025, 0.0125])
    else:
        dh, zb = _logdepth_layers(nl)
        N2, _ = _exp_stratification(zb, SEED0)
    return default_params(NX=NX, NY=NX, nl=nl, Lx=4000e3, dh=dh, N2=N2, f0=1e-4, beta=2e-11,
                          hEkb=2 * 1e-7 * dh[-1] / 1e-4, n_hyper=4, nu_hyper=nu4, tau0=tau / RHO0,
                          forc_mode=2.0, bc_fac=0.0, dt_fixed=dt, nsteps=500, ic_kind=1,
                          seed=SEED0 + member, ic_urms=0.05, ic_lambda0=200e3)


def c4(classes=None, dt=1200.0):
    """The 96-member ensemble: 32 x (512^2, 2), 32 x (1024^2, 3), 24 x (1536^2, 5), 8 x (2048^2, 4);
    wind amplitude tau ~ U[0.04, 0.12] N/m^2 per member (seeded), FP32 + Kahan, dt 1200 s."""
    classes = C4_CLASSES if classes is None else classes
    total = sum(c[0] for c in classes)
    taus = np.random.default_rng(SEED0).uniform(0.04, 0.12, total)
    out = []
    for cnt, NX, nl, nu4 in classes:
        for _ in range(cnt):
            out.append(c4_member(NX, nl, nu4, len(out), float(taus[len(out)]), dt))
    return out


def c5(**over):
    """4096^2 x 32 layers, L = 8000 km, log-depth layers summing to 4000 m with the seeded
    exponential stratification of c3, double-gyre wind 0.08 N/m^2, drag 1e-7, dt 600 s,
    nu4 = 2e8 (26 nu4 dt / Delta^4 = 0.21)."""
    nl = 32
    dh, zb = _logdepth_layers(nl)
    N2, _ = _exp_stratification(zb, SEED0)
    p = default_params(NX=4096, NY=4096, nl=nl, Lx=8000e3, dh=dh, N2=N2, f0=1e-4, beta=2e-11,
                       hEkb=2 * 1e-7 * dh[-1] / 1e-4, n_hyper=4, nu_hyper=2e8, tau0=0.08 / RHO0,
                       forc_mode=2.0, bc_fac=0.0, dt_fixed=600.0, nsteps=200, ic_kind=1, seed=SEED0,
                       ic_urms=0.1, ic_lambda0=200e3)
    return p.set(**over)


WORKLOADS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c5": c5}
