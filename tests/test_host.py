"""Host C library (libqgw_host.so): reader, grid, modes, initial condition."""
import os
import re

import numpy as np
import pytest

import qgw_amd as Q
from qgw_amd import configs
import oracle as O
from conftest import rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PARAMS = os.path.join(GOLD, "params")


# ---------------------------------------------------------------- reader

def _expected_from_text(path):
    """Reader semantics of src/extra.h:98-153 restated in Python: spaces
    removed, split at '=', registered keys only, int via leading digits,
    arrays: first field split on '[,]', later ones on ','."""
    keys_int = {"NX", "NY", "nl", "n_hyper"}
    keys_arr = {"dh", "N2"}
    keys_dbl = {"Lx", "tau0", "forc_mode", "f0", "beta", "nu", "nu_hyper", "hEkb", "Ld", "h_topo", "w_topo",
                "bc_fac", "tend", "dt_out", "dt_forc", "dt_forc_period", "sigma_f", "k_f", "cfl"}
    out = {}
    for line in open(path):
        line = line.replace(" ", "")
        parts = [x for x in line.split("=") if x != ""]
        if len(parts) < 2:
            continue
        k, v = parts[0], parts[1]
        if k in keys_int:
            out[k] = int(re.match(r"[+-]?\d*", v.strip()).group() or 0)
        elif k in keys_dbl:
            out[k] = float(re.match(r"[+-]?[\d.eE+-]*", v.strip()).group())
        elif k in keys_arr:
            v = v.split("\n")[0]
            toks = [t for t in re.split(r"[\[,\]]", v) if t != ""]
            out[k] = [float(re.match(r"[+-]?[\d.eE+-]*", t).group()) for t in toks]
    return out


@pytest.mark.parametrize("name", sorted(os.listdir(PARAMS)))
def test_reader_reference_params_files(name):
    path = os.path.join(PARAMS, name)
    p = Q.read_params(path)
    exp = _expected_from_text(path)
    d = p.as_dict()
    for k, v in exp.items():
        if isinstance(v, list):
            got = [getattr(p, k)[i] for i in range(len(v))]
            assert got == pytest.approx(v, rel=1e-15), k
        else:
            assert d[k] == pytest.approx(v, rel=1e-15), k
    # keys the reference does not register are ignored (e.g. dt in params_01)
    assert p.dt_fixed == 0.0


def test_reader_params03_values():
    p = Q.read_params(os.path.join(PARAMS, "params_03.in"))
    assert (p.NX, p.NY, p.nl) == (64, 64, 3)
    assert [p.dh[i] for i in range(3)] == [350.0, 750.0, 2900.0]
    assert [p.N2[i] for i in range(2)] == [4.54545455e-05, 6.84931507e-06]
    assert p.forc_mode == 2.0 and p.hEkb == 2.0 and p.nu == 100.0
    assert p.f0 == 9.4e-05 and p.beta == 1.7e-11 and p.tau0 == 1.3e-5


def test_reader_defaults_and_missing_file(tmp_path):
    p = Q.default_params()
    assert (p.nl, p.f0, p.cfl, p.sigma_f, p.forc_mode, p.w_topo) == (1, 1e-5, 0.2, 1.0, 1.0, 1.0)
    with pytest.raises(Q.QGWError):
        Q.read_params(str(tmp_path / "nope.in"))


def test_reader_extension_keys(tmp_path):
    f = tmp_path / "x.in"
    f.write_text("NX = 128\n# comment = 5\nnl = 2\ndt_fixed = 900\nseed = 42\nunknown = 7\n"
                 "dh=[ 100 , 200]\n")
    p = Q.read_params(str(f))
    assert (p.NX, p.nl, p.dt_fixed, p.seed) == (128, 2, 900.0, 42)
    assert (p.dh[0], p.dh[1]) == (100.0, 200.0)


# ---------------------------------------------------------------- grid and forcing

def test_grid_and_wavenumbers():
    p = Q.default_params(NX=64, NY=48, Lx=1000e3)
    g, X, Y, K, L = Q.grid(p)
    assert (g.n, g.m, g.nxp2, g.nyp2) == (63, 47, 65, 49)
    assert g.Delta == 1000e3 / 64 and g.Ly == 48 * g.Delta
    assert L[1] == np.pi / 1000e3 and L[63] == np.pi * 63 / 1000e3
    assert K[47] == np.pi * 47 / 1000e3  # K uses Lx (src/domain.h:170-171)

    # periodic (src/domain.h:71-75,144-147,163-167): NX points, zero of X at
    # the second grid point, K[i] = 2 pi i/Lx (i <= NX/2), L[j] signed
    g, X, Y, K, L = Q.grid(Q.default_params(NX=64, NY=64, Lx=2.0, bc_fac=-1.0))
    assert (g.n, g.m, g.nxp2, g.nyp2) == (64, 64, 66, 66)
    assert X[0] == -2.0 / 64 and X[1] == 0.0 and Y[65] == 64 * 2.0 / 64
    assert K[0] == 0.0 and K[32] == 2 * np.pi * 32 / 2.0
    assert L[1] == 2 * np.pi / 2.0 and L[31] == 2 * np.pi * 31 / 2.0
    assert L[32] == -2 * np.pi * 32 / 2.0 and L[63] == -2 * np.pi / 2.0
    for bad in (dict(NX=64, NY=48), dict(NX=63, NY=63)):  # square, even NX only
        with pytest.raises(Q.QGWError):
            Q.grid(Q.default_params(Lx=1.0, bc_fac=-1.0, **bad))
    with pytest.raises(Q.QGWError):
        Q.grid(Q.default_params(NX=64, NY=64, Lx=1.0, bc_fac=-0.5))


def test_topo_and_wind_match_oracle():
    p = Q.default_params(NX=32, NY=32, nl=2, Lx=1e6, dh=[500.0, 1500.0], N2=[1e-5], h_topo=300.0, w_topo=80e3,
                         tau0=1e-4, forc_mode=2.0)
    o = O.Oracle(p)
    assert np.array_equal(Q.topo(p), o.field(O.TOPO))
    w = Q.wind_row(p)
    g = Q.grid(p)[0]
    y = np.arange(g.nyp2) * g.Delta
    assert np.allclose(w, -1e-4 / 500.0 * 2 * np.pi / g.Ly * np.sin(2 * np.pi * y / g.Ly), rtol=1e-14, atol=1e-30)


def test_check_timestep_host_matches_oracle():
    for kw in [dict(tau0=1e-5, beta=1e-11), dict(nu=10.0), dict(n_hyper=4, nu_hyper=1e9, beta=2e-11),
               dict(tau0=1e-5, sigma_f=0.0, beta=1e-11), dict(tau0=0.0, sigma_f=2.0)]:
        p = Q.default_params(NX=64, NY=64, nl=1, Lx=1e6, dh=[1000.0], **kw)
        a, b = Q.check_timestep(p), O.check_timestep(p)
        assert a == b or (np.isnan(a[0]) and np.isnan(b[0])), kw


# ---------------------------------------------------------------- modes

@pytest.mark.parametrize("mk", [configs.c1, configs.c2, configs.c3])
def test_host_modes_match_oracle_and_golden(mk):
    p = mk()
    cl2m, cm2l, iRd2 = Q.eigmode(p)
    o = O.Oracle(Q.default_params(NX=8, NY=8, nl=p.nl, Lx=p.Lx, dh=list(p.dh)[:p.nl],
                                  N2=list(p.N2)[:p.nl - 1], f0=p.f0))
    a, b, c = o.modes()
    assert np.abs(cl2m - a).max() < 1e-10 * np.abs(a).max()
    assert np.abs(cm2l - b).max() < 1e-10 * np.abs(b).max()
    assert np.allclose(iRd2, c, rtol=1e-10, atol=0)
    nl = p.nl
    assert np.abs(cl2m.reshape(nl, nl) @ cm2l.reshape(nl, nl) - np.eye(nl)).max() < 1e-12


@pytest.mark.parametrize("nl", [32, 200])
def test_host_modes_many_layers_and_single_layer(nl):
    rng = np.random.default_rng(9)
    dh = list(rng.uniform(50, 400, nl))
    N2 = list(rng.uniform(1e-6, 1e-4, nl - 1))
    p = Q.default_params(NX=8, NY=8, nl=nl, Lx=1e6, dh=dh, N2=N2, f0=1e-4)
    cl2m, cm2l, iRd2 = Q.eigmode(p)
    A, B = cl2m.reshape(nl, nl), cm2l.reshape(nl, nl)
    assert np.abs(A @ B - np.eye(nl)).max() < 1e-10
    G = Q.gamma(p)
    lam = np.concatenate([[0.0], iRd2[1:]])
    assert np.abs(G.reshape(nl, nl) @ B - B * lam).max() < 1e-8 * np.abs(G).max() * np.abs(B).max()
    big = np.abs(B[0]) > 1e-8  # bottom-trapped modes can have a vanishing surface value
    assert (np.diff(iRd2) > 0).all() and (B[0][big] > 0).all()
    # eigenvalues against LAPACK's symmetric solver (numpy) on D^1/2 Gamma D^-1/2
    sq = np.sqrt(np.asarray(dh))
    S = (sq[:, None] * G.reshape(nl, nl)) / sq[None, :]
    w = np.linalg.eigvalsh(0.5 * (S + S.T))
    assert np.allclose(iRd2[1:], w[1:], rtol=1e-9, atol=1e-12 * np.abs(w).max())
    p1 = Q.default_params(NX=8, NY=8, nl=1, Lx=1e6, dh=[1000.0], Ld=30e3)
    c1, m1, i1 = Q.eigmode(p1)
    assert (c1[0], m1[0]) == (1.0, 1.0) and i1[0] == pytest.approx(1 / 30e3 ** 2)


# ---------------------------------------------------------------- transforms and IC

def test_host_dst2d_golden():
    g = np.load(os.path.join(GOLD, "dst_golden.npz"))
    assert rel(Q.dst2d(g["a2"]), g["A2"]) < 1e-14
    rng = np.random.default_rng(11)
    a = rng.standard_normal((95, 47))
    assert rel(Q.dst2d(a), O.dst2d(a)) < 1e-14
    # radix-5 and radix-7 stages of the host FFT: lengths 2(n+1) = 448 (2^6 7),
    # 280 (2^3 5 7) and 200 (2^3 5^2), against the direct sum
    for m, n in ((223, 139), (99, 223)):
        a = rng.standard_normal((m, n))
        Sx = 2.0 * np.sin(np.pi * np.outer(np.arange(1, n + 1), np.arange(1, n + 1)) / (n + 1))
        Sy = 2.0 * np.sin(np.pi * np.outer(np.arange(1, m + 1), np.arange(1, m + 1)) / (m + 1))
        assert rel(Q.dst2d(a), Sy @ a @ Sx) < 1e-13, (m, n)


def test_initial_condition_properties():
    p = configs.c1()
    modes = Q.eigmode(p)
    psi, q = Q.make_ic(p, modes)
    psi2, q2 = Q.make_ic(p, modes)
    assert np.array_equal(psi, psi2) and np.array_equal(q, q2)
    assert np.abs(psi[:, 0, :]).max() == 0 and np.abs(psi[:, :, -1]).max() == 0
    g = Q.grid(p)[0]
    u = -(psi[:, 1:-1, 2:-1] - psi[:, 1:-1, 1:-2]) / g.Delta
    v = (psi[:, 2:-1, 1:-1] - psi[:, 1:-2, 1:-1]) / g.Delta
    # forward differences over the interior with a zero ring at i = n+1, j = m+1
    uu = np.zeros((p.nl, g.m, g.n))
    vv = np.zeros_like(uu)
    uu[:, :, :-1] = u
    uu[:, :, -1] = psi[:, 1:-1, -2] / g.Delta
    vv[:, :-1, :] = v
    vv[:, -1, :] = -psi[:, -2, 1:-1] / g.Delta
    assert np.sqrt(np.mean(uu ** 2 + vv ** 2)) == pytest.approx(0.1, rel=1e-10)
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q)
    o.invert_pv()
    assert rel(o.field(O.PSI), psi) < 1e-13
    psi3, _ = Q.make_ic(p, modes, seed=p.seed + 1)
    assert rel(psi3, psi) > 0.5


def test_diagnostics_host_matches_oracle():
    p = configs.c1()
    psi, q = Q.make_ic(p)
    o = O.Oracle(p)
    o.set_q(q)
    o.field(O.PSI)[...] = psi
    a = Q.diagnostics_host(p, psi, q)
    b = o.diagnostics()
    assert np.allclose(a[0], b[0], rtol=1e-13) and np.allclose(a[1], b[1], rtol=1e-13)
    assert a[2] == pytest.approx(b[2], rel=1e-13) and a[3] == pytest.approx(b[3], rel=1e-13)
    assert a[2] > 0 and a[3] > 0


def test_tune_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setenv("QGW_TUNE_CACHE", str(tmp_path / "sub" / "tune.json"))
    assert Q.tune_cache_load() == {}
    k1 = Q.tune_key(Q.FP32, 2047, 2047, 10, 4, True)
    k2 = Q.tune_key(Q.FP64, 2047, 2047, 10, 4, True)
    assert k1 != k2 and "2047x2047" in k1 and "nl10" in k1
    Q.tune_cache_store({k1: {"1": 4, "3": 40}})
    Q.tune_cache_store({k2: {"1": 2}})
    d = Q.tune_cache_load()
    assert d[k1] == {"1": 4, "3": 40} and d[k2] == {"1": 2}
    (tmp_path / "sub" / "tune.json").write_text("{not json")
    assert Q.tune_cache_load() == {}  # a damaged cache only costs a retune
