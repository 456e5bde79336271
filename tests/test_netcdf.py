"""netCDF classic I/O of the driver (include/qgw_nc.h): the vars.nc writer and
the restart.nc reader, checked against scipy.io.netcdf_file, an independent
implementation of the same file format, in both directions.

The reference's files (src/netcdf_io.h) hold float variables; the layouts
checked here are the ones create_nc / write_nc / read_nc produce and expect
(src/netcdf_io.h:38-160, 171-277, 286-386)."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

import qgw_amd as Q


def _bounded():
    return Q.default_params(NX=32, NY=24, nl=3, Lx=1e6, dh=[300.0, 700.0, 2000.0], N2=[1e-5, 5e-6])


def _periodic():
    return Q.default_params(NX=16, NY=16, nl=2, Lx=2e6, dh=[500.0, 1500.0], N2=[1e-5], bc_fac=-1.0)


def _fields(p, rng):
    g = Q.grid(p)[0]
    return rng.standard_normal((p.nl, g.nyp2, g.nxp2)), rng.standard_normal((p.nl, g.nyp2, g.nxp2))


def _ibc(p):
    return 1 if p.bc_fac == -1.0 else 0  # src/netcdf_io.h:42-44


def _out(a, p):
    """The part of a storage array a run writes (src/netcdf_io.h:236-240)."""
    b = _ibc(p)
    return a[..., b:a.shape[-2] - b, b:a.shape[-1] - b]


# ---------------------------------------------------------------- writer

@pytest.mark.parametrize("mk", [_bounded, _periodic])
def test_writer_read_by_scipy(tmp_path, mk):
    p = mk()
    g = Q.grid(p)[0]
    rng = np.random.default_rng(1)
    path = str(tmp_path / "vars.nc")
    recs = []
    with Q.NcWriter(path, p) as w:
        assert w.records == 0
        for r in range(3):
            psi, q = _fields(p, rng)
            t = 1234.5 * r + 0.1
            w.write(t, psi, q)
            recs.append((t, psi, q))
        assert w.records == 3
    f = netcdf_file(path, "r", mmap=False)
    try:
        assert f.version_byte == 1  # classic, as nc_create(..., NC_CLOBBER) writes
        nyo, nxo = g.nyp2 - 2 * _ibc(p), g.nxp2 - 2 * _ibc(p)
        assert f.dimensions == {"t": None, "level": p.nl, "y": nyo, "x": nxo}
        V = f.variables
        assert list(V["psi"].dimensions) == ["t", "level", "y", "x"] and V["q"].typecode() == "f"
        assert V["t"].dimensions == ("t",) and V["level"].dimensions == ("level",)
        # coordinates (src/netcdf_io.h:134-152)
        assert np.array_equal(V["x"][:], (np.arange(nxo) * g.Delta).astype(np.float32))
        assert np.array_equal(V["y"][:], (np.arange(nyo) * g.Delta).astype(np.float32))
        assert np.array_equal(V["level"][:], np.arange(p.nl, dtype=np.float32))
        assert np.array_equal(V["t"][:], np.array([r[0] for r in recs], dtype=np.float32))
        for r, (_, psi, q) in enumerate(recs):
            assert np.array_equal(V["psi"][r], _out(psi, p).astype(np.float32))
            assert np.array_equal(V["q"][r], _out(q, p).astype(np.float32))
    finally:
        f.close()
    # fixed part, then records of t, psi, q (4 + 2 * vsize bytes each)
    slab = 4 * p.nl * (g.nyp2 - 2 * _ibc(p)) * (g.nxp2 - 2 * _ibc(p))
    hdr_and_coords = os.path.getsize(path) - 3 * (4 + 2 * slab)
    raw = open(path, "rb").read()
    assert raw[:4] == b"CDF\x01" and int.from_bytes(raw[4:8], "big") == 3
    assert hdr_and_coords % 4 == 0 and hdr_and_coords > 0


def test_writer_round_trips_through_reader(tmp_path):
    p = _bounded()
    rng = np.random.default_rng(2)
    path = str(tmp_path / "vars.nc")
    with Q.NcWriter(path, p) as w:
        fields = [_fields(p, rng) for _ in range(2)]
        for r, (psi, q) in enumerate(fields):
            w.write(r, psi, q)
    for r, (_, q) in enumerate(fields):
        got, topo, found = Q.read_restart(path, p, rec=r)
        assert found == {"q"}  # no topo in a vars.nc
        assert np.array_equal(got, q.astype(np.float32).astype(np.float64))
        assert not topo.any()


def test_writer_rejects_bad_shapes_and_paths(tmp_path):
    p = _bounded()
    with pytest.raises(Q.QGWError):
        Q.NcWriter(str(tmp_path / "no_such_dir" / "vars.nc"), p)
    with Q.NcWriter(str(tmp_path / "v.nc"), p) as w:
        psi, q = _fields(p, np.random.default_rng(0))
        with pytest.raises(Q.QGWError):
            w.write(0.0, psi[:, :-1], q)
        assert w.records == 0


# ---------------------------------------------------------------- reader

def _restart(path, p, q_recs, topo_recs=None, version=1, dtype="f", fixed_topo=None, swap_xy=False):
    """A restart file as a user would make it with any netCDF library."""
    g = Q.grid(p)[0]
    nyo, nxo = g.nyp2 - 2 * _ibc(p), g.nxp2 - 2 * _ibc(p)
    f = netcdf_file(path, "w", version=version)
    f.createDimension("time", None)
    f.createDimension("level", p.nl)
    ya, xa = ("x", "y") if swap_xy else ("y", "x")
    f.createDimension(ya, nyo)
    f.createDimension(xa, nxo)
    tv = f.createVariable("time", "f", ("time",))
    qv = f.createVariable("q", dtype, ("time", "level", ya, xa))
    for r, a in enumerate(q_recs):
        qv[r] = _out(a, p)
        tv[r] = r
    if topo_recs is not None:
        tp = f.createVariable("topo", dtype, ("time", ya, xa))
        for r, a in enumerate(topo_recs):
            tp[r] = _out(a, p)
    if fixed_topo is not None:
        tp = f.createVariable("topo", dtype, (ya, xa))
        tp[:] = _out(fixed_topo, p)
    f.close()


@pytest.mark.parametrize("mk", [_bounded, _periodic])
def test_reader_scipy_restart(tmp_path, mk):
    p = mk()
    g = Q.grid(p)[0]
    rng = np.random.default_rng(3)
    qs = [_fields(p, rng)[0] for _ in range(3)]
    tops = [rng.standard_normal((g.nyp2, g.nxp2)) for _ in range(3)]
    path = str(tmp_path / "restart.nc")
    _restart(path, p, qs, tops)
    q0 = np.full((p.nl, g.nyp2, g.nxp2), 7.0)
    t0 = np.full((g.nyp2, g.nxp2), -3.0)
    q, topo, found = Q.read_restart(path, p, rec=2, q=q0.copy(), topo=t0.copy())
    assert found == {"q", "topo"}
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    b = _ibc(p)
    inner = (slice(None), slice(b, g.nyp2 - b), slice(b, g.nxp2 - b))
    assert np.array_equal(q[inner], f32(qs[2])[inner])
    assert np.array_equal(topo[inner[1:]], f32(tops[2])[inner[1:]])
    if b:  # a periodic run's ring is not in the file and stays as it was
        ring = np.ones(q.shape[1:], bool)
        ring[1:-1, 1:-1] = False
        assert (q[:, ring] == 7.0).all() and (topo[ring] == -3.0).all()


def test_reader_double_variables_rounded_to_float(tmp_path):
    """nc_get_vara_float converts double data to float (src/netcdf_io.h:339,354)."""
    p = _bounded()
    rng = np.random.default_rng(4)
    q = _fields(p, rng)[0] * (1 + 1e-9)
    path = str(tmp_path / "restart.nc")
    _restart(path, p, [q], dtype="d")
    got, _, found = Q.read_restart(path, p)
    assert found == {"q"}
    assert np.array_equal(got, q.astype(np.float32).astype(np.float64))


def test_reader_fixed_topo_cdf2_and_single_record_variable(tmp_path):
    p = _bounded()
    g = Q.grid(p)[0]
    rng = np.random.default_rng(5)
    q = _fields(p, rng)[0]
    top = rng.standard_normal((g.nyp2, g.nxp2))
    path = str(tmp_path / "r2.nc")
    f = netcdf_file(path, "w", version=2)  # 64-bit offsets
    f.createDimension("t", None)
    f.createDimension("level", p.nl)
    f.createDimension("y", g.nyp2)
    f.createDimension("x", g.nxp2)
    f.attr = "global attribute to skip"
    v = f.createVariable("q", "f", ("t", "level", "y", "x"))  # the only record variable
    v.units = "s-1"
    v[0] = q
    v[1] = 2 * q
    tp = f.createVariable("topo", "f", ("y", "x"))
    tp[:] = top
    f.close()
    got, topo, found = Q.read_restart(path, p, rec=1)
    assert found == {"q", "topo"}
    assert np.array_equal(got, (2 * q).astype(np.float32).astype(np.float64))
    assert np.array_equal(topo, top.astype(np.float32).astype(np.float64))


def test_reader_transposed_dimensions_read_as_stored(tmp_path):
    """read_nc asks for (1, nl, NX_Out, NY_Out) and maps the values in storage
    order (src/netcdf_io.h:349-366): a square file with dimensions named
    (x, y) is read element for element, like (y, x)."""
    p = Q.default_params(NX=16, NY=16, nl=1, Lx=1e6, dh=[1000.0])
    rng = np.random.default_rng(6)
    q = _fields(p, rng)[0]
    path = str(tmp_path / "restart.nc")
    _restart(path, p, [q], swap_xy=True)
    got, _, _ = Q.read_restart(path, p)
    assert np.array_equal(got, q.astype(np.float32).astype(np.float64))


def test_reader_errors(tmp_path):
    p = _bounded()
    rng = np.random.default_rng(7)
    q = _fields(p, rng)[0]
    path = str(tmp_path / "restart.nc")
    _restart(path, p, [q])
    with pytest.raises(Q.QGWError):  # record beyond the file
        Q.read_restart(path, p, rec=1)
    with pytest.raises(Q.QGWError):  # grid of another size
        Q.read_restart(path, Q.default_params(NX=16, NY=24, nl=3, Lx=1e6, dh=[300.0, 700.0, 2000.0],
                                              N2=[1e-5, 5e-6]))
    with pytest.raises(Q.QGWError):  # missing file
        Q.read_restart(str(tmp_path / "none.nc"), p)
    bad = tmp_path / "bad.nc"
    bad.write_bytes(b"HDF\x01" + bytes(60))
    with pytest.raises(Q.QGWError):
        Q.read_restart(str(bad), p)
    trunc = tmp_path / "trunc.nc"
    trunc.write_bytes(open(path, "rb").read()[:200])
    with pytest.raises(Q.QGWError):
        Q.read_restart(str(trunc), p)
    other = tmp_path / "other.nc"
    f = netcdf_file(str(other), "w")
    f.createDimension("n", 4)
    f.createVariable("u", "f", ("n",))[:] = 1.0
    f.close()
    q0 = np.full_like(q, 5.0)
    got, _, found = Q.read_restart(str(other), p, q=q0)
    assert found == set() and (got == 5.0).all()  # nothing to read leaves the state as it is
