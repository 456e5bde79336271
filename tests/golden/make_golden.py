"""Write the golden fixtures of tests/golden/ (run in a container with scipy/sympy).

Library vectors that pin the oracle and the host library where the reference
itself cannot run (FFTW, LAPACKE and netCDF are absent, and the reference's
own tests hold no vectors):

  dst_golden.npz     scipy.fft.dst(type=1) -- FFTW's RODFT00 definition
                     (src/elliptic.h:63-66) -- of seeded inputs, n = 3, 7,
                     63, 255, 1535, and of [1, 2, 3]; scipy.fft.dstn for a
                     31 x 23 array (the 2-D transform, both axes RODFT00)
  eig_golden.npz     scipy.linalg.eigh of the symmetrised stretching matrix
                     D^1/2 Gamma D^-1/2 for the 256^2 x 2, 1024^2 x 3 and
                     2048^2 x 10 layer sets; Gamma built independently here
                     from the definition in src/eigmode.h:99-122
  ab3_golden.npz     variable-step AB3 weights from an independent sympy
                     integration of the Lagrange interpolant through
                     (0, f0), (-dt0, f1), (-dt0-dt1, f2) over [0, dt]
                     (the construction described in doc/timestep.org:39-97)

Usage: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.fft as sf
import scipy.linalg as sl
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def dst_vectors():
    rng = np.random.default_rng(20261203)
    out = {}
    for n in (3, 7, 63, 255, 1535):
        x = rng.standard_normal(n)
        out["x_%d" % n] = x
        out["y_%d" % n] = sf.dst(x, type=1)
    out["x_123"] = np.array([1.0, 2.0, 3.0])
    out["y_123"] = sf.dst(out["x_123"], type=1)
    a = rng.standard_normal((23, 31))
    out["a2"] = a
    out["A2"] = sf.dstn(a, type=1)
    np.savez(os.path.join(HERE, "dst_golden.npz"), **out)


def gamma(dh, N2, f0):
    nl = len(dh)
    dhc = [0.5 * (dh[l] + dh[l + 1]) for l in range(nl - 1)]
    G = np.zeros((nl, nl))
    for l in range(nl):
        if l > 0:
            G[l, l - 1] = -f0 ** 2 / N2[l - 1] / (dhc[l - 1] * dh[l])
        if l < nl - 1:
            G[l, l + 1] = -f0 ** 2 / N2[l] / (dhc[l] * dh[l])
        G[l, l] = -G[l].sum()
    return G


def layer_sets():
    sets = {}
    dh = [500.0, 3500.0]
    sets["c1"] = (dh, [0.02 / 2000.0], 1e-4)
    dh = [350.0, 750.0, 2900.0]
    sets["c2"] = (dh, [0.025 / 550.0, 0.0125 / 1825.0], 1e-4)
    nl, H, z1 = 10, 4000.0, 40.0
    zb = [z1 * (H / z1) ** (l / (nl - 1)) for l in range(nl)]
    zb[-1] = H
    dh = [zb[0]] + [zb[l] - zb[l - 1] for l in range(1, nl)]
    n0 = np.random.default_rng(20261203).uniform(1e-5, 5e-5)
    sets["c3"] = (dh, [n0 * np.exp(-z / 800.0) for z in zb[:-1]], 1e-4)
    return sets


def eig_vectors():
    out = {}
    for name, (dh, N2, f0) in layer_sets().items():
        G = gamma(dh, N2, f0)
        d = np.sqrt(np.array(dh))
        T = (d[:, None] * G) / d[None, :]
        T = 0.5 * (T + T.T)
        w, U = sl.eigh(T)
        out[name + "_dh"] = np.array(dh)
        out[name + "_N2"] = np.array(N2)
        out[name + "_gamma"] = G
        out[name + "_w"] = w
        out[name + "_U"] = U
    np.savez(os.path.join(HERE, "eig_golden.npz"), **out)


def ab3_vectors():
    t, dt, dt0, dt1 = sp.symbols("t dt dt0 dt1", positive=True)
    f0, f1, f2 = sp.symbols("f0 f1 f2")
    nodes = [(0, f0), (-dt0, f1), (-dt0 - dt1, f2)]
    poly = 0
    for i, (ti, fi) in enumerate(nodes):
        term = fi
        for j, (tj, _) in enumerate(nodes):
            if j != i:
                term = term * (t - tj) / (ti - tj)
        poly += term
    integral = sp.expand(sp.integrate(poly, (t, 0, dt)))
    coeffs = [sp.simplify(integral.coeff(f)) for f in (f0, f1, f2)]
    cases = np.array([[1800.0, 1800.0, 1800.0], [2.0, 1.0, 1.0], [1.0, 2.0, 3.0], [900.0, 450.0, 1800.0],
                      [0.37, 0.81, 0.05]])
    vals = np.array([[float(c.subs({dt: a, dt0: b, dt1: cc})) for c in coeffs] for a, b, cc in cases])
    np.savez(os.path.join(HERE, "ab3_golden.npz"), cases=cases, coeffs=vals)


if __name__ == "__main__":
    dst_vectors()
    eig_vectors()
    ab3_vectors()
    print("fixtures written to", HERE)
