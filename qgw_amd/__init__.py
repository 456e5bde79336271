"""qgw_amd -- the MI355X timestep hot path of the multi-layer QG model qgw.

Python view of the two native libraries built in this directory:

  libqgw_host.so  host C: parameter reader, grid, vertical modes, seeded
                  initial condition, FP64 diagnostics (include/qgw_host.h)
  libqgw_hip.so   HIP for gfx950: the per-step hot path behind a C-ABI
                  (include/qgw_hip.h)

The names follow the reference's own call boundaries (src/qg.c:153-200):
``Model.invert_pv`` (src/elliptic.h:76), ``Model.rhs`` (src/dynamics.h:47),
``Model.adjust_timestep`` (src/timestep.h:69) and ``Model.timestep``
(src/timestep.h:114).  Arrays are numpy float64 in the reference storage
layout ``idx(i,j,k) = k*Nxp2*Nyp2 + j*Nxp2 + i`` (src/domain.h:62), i.e. shape
``(nl, m+2, n+2)``.

Every device call goes through libqgw_hip.so; there is no CPU fallback.  If
the library is missing or no GPU is present the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NL_MAX = 1000

FP64, FP32, FP32_PLAIN = 0, 1, 2
FIELD_Q, FIELD_PSI, FIELD_TOPO, FIELD_F0, FIELD_F1, FIELD_F2 = 1, 2, 3, 4, 5, 6
PASS_NAMES = ["P1", "T1", "P2", "T2", "P3", "P4", "CFL", "DIAG", "FORC"]
TUNE_P4_TILE = 1
TUNE_T_KERNEL = 2
TUNE_P4_WAVES = 3
TUNE_DST = 4
TUNE_FFT_E = 5
TUNE_FFT_E_COL = 6
TUNE_FUSE = 7
TUNE_P2M_PAIRS = 8
TUNE_PSI_NT = 9
TUNE_P4_REV = 10
TUNE_P4_ORDER = 11
P4_TILES = ["64x8", "64x16", "64x32", "128x16", "stream", "stream-aligned"]
TUNE_KNOBS = (TUNE_P4_TILE, TUNE_P4_WAVES, TUNE_T_KERNEL, TUNE_DST, TUNE_FFT_E, TUNE_FFT_E_COL, TUNE_FUSE,
              TUNE_P2M_PAIRS, TUNE_PSI_NT, TUNE_P4_REV, TUNE_P4_ORDER)
_TUNE_CACHE = {}

QGW_OK = 0


class QGWError(RuntimeError):
    pass


class Params(ctypes.Structure):
    """Mirror of ``qgw_params`` (include/qgw_params.h)."""

    _fields_ = [
        ("NX", ctypes.c_int), ("NY", ctypes.c_int), ("nl", ctypes.c_int), ("n_hyper", ctypes.c_int),
        ("Lx", ctypes.c_double), ("dh", ctypes.c_double * NL_MAX),
        ("tau0", ctypes.c_double), ("forc_mode", ctypes.c_double), ("f0", ctypes.c_double),
        ("beta", ctypes.c_double), ("nu", ctypes.c_double), ("nu_hyper", ctypes.c_double),
        ("hEkb", ctypes.c_double), ("N2", ctypes.c_double * NL_MAX),
        ("Ld", ctypes.c_double), ("h_topo", ctypes.c_double), ("w_topo", ctypes.c_double),
        ("bc_fac", ctypes.c_double), ("tend", ctypes.c_double), ("dt_out", ctypes.c_double),
        ("dt_forc", ctypes.c_double), ("dt_forc_period", ctypes.c_double), ("sigma_f", ctypes.c_double),
        ("k_f", ctypes.c_double), ("cfl", ctypes.c_double),
        ("dt_fixed", ctypes.c_double), ("nsteps", ctypes.c_int), ("ic_kind", ctypes.c_int),
        ("seed", ctypes.c_uint64), ("ic_urms", ctypes.c_double), ("ic_lambda0", ctypes.c_double),
        ("precision", ctypes.c_int), ("verbose", ctypes.c_int),
    ]

    def set(self, **kw):
        for k, v in kw.items():
            if k in ("dh", "N2"):
                arr = getattr(self, k)
                for i, x in enumerate(v):
                    arr[i] = float(x)
            else:
                setattr(self, k, v)
        return self

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if name in ("dh", "N2"):
                v = [v[i] for i in range(max(self.nl, 1))]
            d[name] = v
        return d


class Grid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("m", ctypes.c_int), ("nxp2", ctypes.c_int), ("nyp2", ctypes.c_int),
                ("Delta", ctypes.c_double), ("Lx", ctypes.c_double), ("Ly", ctypes.c_double)]


_host = None
_hip = None
_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def host_lib():
    """libqgw_host.so (pure C, usable without a GPU)."""
    global _host
    if _host is None:
        path = os.path.join(HERE, "libqgw_host.so")
        if not os.path.exists(path):
            raise QGWError("libqgw_host.so not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        P = ctypes.POINTER(Params)
        lib.qgwh_default_params.argtypes = [P]
        lib.qgwh_read_params.argtypes = [ctypes.c_char_p, P]
        lib.qgwh_grid.argtypes = [P, ctypes.POINTER(Grid), _dp, _dp, _dp, _dp]
        lib.qgwh_topo.argtypes = [P, _dp]
        lib.qgwh_eigmode.argtypes = [P, _dp, _dp, _dp]
        lib.qgwh_gamma.argtypes = [P, _dp]
        lib.qgwh_wind_row.argtypes = [P, _dp]
        lib.qgwh_make_ic.argtypes = [P, _dp, _dp, _dp, ctypes.c_uint64, _dp, _dp]
        lib.qgwh_dst2d.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        lib.qgwh_diagnostics.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.qgwh_check_timestep.argtypes = [P, ctypes.c_double, _dp, _dp]
        lib.qgwh_nu_hyper_signed.argtypes = [P]
        lib.qgwh_nu_hyper_signed.restype = ctypes.c_double
        lib.qgwh_strerror.argtypes = [ctypes.c_int]
        lib.qgwh_strerror.restype = ctypes.c_char_p
        lib.qgwh_params_size.restype = ctypes.c_size_t
        vp = ctypes.c_void_p
        lib.qgwh_nc_create.argtypes = [ctypes.c_char_p, P, ctypes.POINTER(ctypes.c_int)]
        lib.qgwh_nc_create.restype = vp
        lib.qgwh_nc_write.argtypes = [vp, ctypes.c_double, _dp, _dp]
        lib.qgwh_nc_records.argtypes = [vp]
        lib.qgwh_nc_close.argtypes = [vp]
        lib.qgwh_nc_read.argtypes = [ctypes.c_char_p, P, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
        lib.qgwh_nc_read_var.argtypes = [ctypes.c_char_p, P, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _dp,
                                         ctypes.POINTER(ctypes.c_int)]
        ip_ = ctypes.POINTER(ctypes.c_int)
        lib.qgwh_stoch_ring.argtypes = [P, ctypes.c_int, ip_, ip_, ip_, _dp]
        lib.qgwh_rng_create.argtypes = [ctypes.c_uint]
        lib.qgwh_rng_create.restype = vp
        lib.qgwh_rng_destroy.argtypes = [vp]
        lib.qgwh_rng_next.argtypes = [vp]
        lib.qgwh_stoch_coeffs.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        if lib.qgwh_params_size() != ctypes.sizeof(Params):
            raise QGWError("qgw_params layout mismatch between C and Python")
        _host = lib
    return _host


def hip_lib():
    """libqgw_hip.so (needs the HIP runtime; compute calls need a GPU)."""
    global _hip
    if _hip is None:
        host_lib()
        path = os.path.join(HERE, "libqgw_hip.so")
        if not os.path.exists(path):
            raise QGWError("libqgw_hip.so not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        vp = ctypes.c_void_p
        P = ctypes.POINTER(Params)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.qgw_create.argtypes = [P, ctypes.c_int, ip]
        lib.qgw_create.restype = vp
        for name, args in [
            ("qgw_destroy", [vp]), ("qgw_set_modes", [vp, _dp, _dp, _dp]),
            ("qgw_upload", [vp, ctypes.c_int, _dp]), ("qgw_download", [vp, ctypes.c_int, _dp]),
            ("qgw_invert_pv", [vp]), ("qgw_rhs", [vp]), ("qgw_adjust_timestep", [vp, _dp]),
            ("qgw_timestep", [vp]), ("qgw_run", [vp, ctypes.c_int]),
            ("qgw_diagnostics", [vp, _dp, _dp, _dp, _dp]),
            ("qgw_get_time", [vp, _dp, _dp, _dp, _dp, ip]),
            ("qgw_set_time", [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              ctypes.c_int]),
            ("qgw_set_next_output", [vp, ctypes.c_double]),
            ("qgw_set_forcing_file", [vp, ctypes.c_char_p]), ("qgw_set_stoch_seed", [vp, ctypes.c_uint]),
            ("qgw_synchronize", [vp]), ("qgw_time_run", [vp, ctypes.c_int, _dp]),
            ("qgw_profile_enable", [vp, ctypes.c_int]), ("qgw_profile_read", [vp, _dp, ip, ctypes.c_int]),
            ("qgw_device_bytes", [vp, ctypes.POINTER(ctypes.c_size_t)]),
            ("qgw_tune_set", [vp, ctypes.c_int, ctypes.c_int]), ("qgw_tune_get", [vp, ctypes.c_int, ip]),
            ("qgw_autotune", [vp, ctypes.c_int, ip, _dp]),
        ]:
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        lib.qgw_strerror.argtypes = [ctypes.c_int]
        lib.qgw_strerror.restype = ctypes.c_char_p
        lib.qgw_last_error.argtypes = [vp]
        lib.qgw_last_error.restype = ctypes.c_char_p
        lib.qgw_batch_create.argtypes = [ctypes.c_int, P, ctypes.c_int, ip]
        lib.qgw_batch_create.restype = vp
        ci = ctypes.c_int
        for name, args in [
            ("qgw_batch_destroy", [vp]), ("qgw_batch_groups", [vp, ip, ip]),
            ("qgw_batch_set_modes", [vp, ci, _dp, _dp, _dp]),
            ("qgw_batch_upload", [vp, ci, ci, _dp]), ("qgw_batch_download", [vp, ci, ci, _dp]),
            ("qgw_batch_invert_pv", [vp]), ("qgw_batch_step", [vp, ci]),
            ("qgw_batch_time_step", [vp, ci, _dp]), ("qgw_batch_synchronize", [vp]), ("qgw_batch_prepare", [vp]),
            ("qgw_batch_set_stoch_seed", [vp, ci, ctypes.c_uint]),
            ("qgw_batch_set_forcing_file", [vp, ci, ctypes.c_char_p]),
            ("qgw_batch_diagnostics", [vp, ci, _dp, _dp, _dp, _dp]),
            ("qgw_batch_get_time", [vp, _dp, ip]), ("qgw_batch_use_graph", [vp, ci]),
            ("qgw_batch_graph_launches", [vp, ip]), ("qgw_batch_autotune", [vp, ci, ip]),
            ("qgw_batch_tune_set", [vp, ci, ci, ci]), ("qgw_batch_tune_get", [vp, ci, ci, ip]),
            ("qgw_batch_profile_enable", [vp, ci]), ("qgw_batch_profile_read", [vp, _dp, ip, ci]),
            ("qgw_batch_device_bytes", [vp, ctypes.POINTER(ctypes.c_size_t)]),
        ]:
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        lib.qgw_batch_last_error.argtypes = [vp]
        lib.qgw_batch_last_error.restype = ctypes.c_char_p
        _hip = lib
    return _hip


# ---------------------------------------------------------------------------
# autotuner choices, persisted
# ---------------------------------------------------------------------------

def tune_cache_path():
    """JSON file of the autotuner's choices: $QGW_TUNE_CACHE, else
    ~/.cache/qgw_amd/tune.json."""
    p = os.environ.get("QGW_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "qgw_amd", "tune.json")


def device_name():
    """Name of HIP device 0 (hipDeviceGetName), part of every tuning key."""
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = ctypes.CDLL(name)
            break
        except OSError:
            hip = None
    if hip is None:
        return "unknown"
    buf = ctypes.create_string_buffer(256)
    if hip.hipDeviceGetName(buf, 256, 0) != 0:
        return "unknown"
    return buf.value.decode(errors="replace").strip() or "unknown"


def tune_key(precision, n, m, layers, n_hyper, hyper_on):
    """Key of one group's tuning: the kernels it launches depend only on the
    device, the precision form, the grid, the stacked layer count and the
    depth of the nested Laplacians."""
    return "%s|prec%d|%dx%d|nl%d|nh%d|%s" % (device_name(), precision, n, m, layers, n_hyper,
                                             "hv" if hyper_on else "nohv")


def tune_cache_load():
    """The persisted choices {key: {knob: value, ...}} (empty when absent or unreadable)."""
    import json
    try:
        with open(tune_cache_path()) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def tune_cache_store(entries):
    """Merge {key: choice} into the JSON cache (written to a temporary file and
    renamed, so a reader never sees a partial file)."""
    import json
    import tempfile
    path = tune_cache_path()
    d = tune_cache_load()
    d.update(entries)
    try:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(path)), suffix=".tmp")
        with os.fdopen(fd, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a retune


# ---------------------------------------------------------------------------
# host helpers
# ---------------------------------------------------------------------------

def default_params(**kw):
    p = Params()
    host_lib().qgwh_default_params(ctypes.byref(p))
    return p.set(**kw)


def read_params(path, p=None):
    """Parse a reference-style params.in (src/extra.h:98-153 semantics)."""
    if p is None:
        p = default_params()
    st = host_lib().qgwh_read_params(os.fsencode(path), ctypes.byref(p))
    if st != QGW_OK:
        raise QGWError("read_params(%s): %s" % (path, host_lib().qgwh_strerror(st).decode()))
    return p


def grid(p):
    """Grid geometry and the X, Y, K, L arrays of src/domain.h (storage extents
    n+2 and m+2: NX+1 bounded, NX+2 periodic)."""
    g = Grid()
    n2, m2 = p.NX + 2, p.NY + 2  # room for either layout
    X, Y, K, L = np.zeros(n2), np.zeros(m2), np.zeros(m2), np.zeros(n2)
    st = host_lib().qgwh_grid(ctypes.byref(p), ctypes.byref(g), _ptr(X), _ptr(Y), _ptr(K), _ptr(L))
    if st != QGW_OK:
        raise QGWError("grid: unsupported parameters")
    return g, X[:g.nxp2], Y[:g.nyp2], K[:g.nyp2], L[:g.nxp2]


def eigmode(p):
    """(cl2m, cm2l, iRd2) as flat float64 arrays (src/eigmode.h:87-271)."""
    nl = p.nl
    cl2m, cm2l, iRd2 = np.zeros(nl * nl), np.zeros(nl * nl), np.zeros(nl)
    st = host_lib().qgwh_eigmode(ctypes.byref(p), _ptr(cl2m), _ptr(cm2l), _ptr(iRd2))
    if st != QGW_OK:
        raise QGWError("eigmode failed: %s" % host_lib().qgwh_strerror(st).decode())
    return cl2m, cm2l, iRd2


def gamma(p):
    G = np.zeros(p.nl * p.nl)
    host_lib().qgwh_gamma(ctypes.byref(p), _ptr(G))
    return G.reshape(p.nl, p.nl)


def topo(p):
    g = grid(p)[0]
    t = np.zeros((g.nyp2, g.nxp2))
    host_lib().qgwh_topo(ctypes.byref(p), _ptr(t))
    return t


def wind_row(p):
    g = grid(p)[0]
    w = np.zeros(g.nyp2)
    host_lib().qgwh_wind_row(ctypes.byref(p), _ptr(w))
    return w


def make_ic(p, modes=None, seed=None):
    """Seeded initial condition (psi, q), each shape (nl, m+2, n+2)."""
    g = grid(p)[0]
    if modes is None:
        modes = eigmode(p)
    cl2m, cm2l, iRd2 = [np.ascontiguousarray(a, dtype=np.float64) for a in modes]
    psi = np.zeros((p.nl, g.nyp2, g.nxp2))
    q = np.zeros_like(psi)
    s = p.seed if seed is None else seed
    st = host_lib().qgwh_make_ic(ctypes.byref(p), _ptr(cl2m), _ptr(cm2l), _ptr(iRd2), ctypes.c_uint64(s),
                                 _ptr(psi), _ptr(q))
    if st != QGW_OK:
        raise QGWError("make_ic failed: %s" % host_lib().qgwh_strerror(st).decode())
    return psi, q


def dst2d(a):
    """Unnormalised 2-D RODFT00 of a (m, n) array (host implementation)."""
    b = np.ascontiguousarray(a, dtype=np.float64).copy()
    m, n = b.shape
    host_lib().qgwh_dst2d(_ptr(b), n, m)
    return b


def diagnostics_host(p, psi, q):
    nl = p.nl
    El, Zl = np.zeros(nl), np.zeros(nl)
    Et, Zt = ctypes.c_double(), ctypes.c_double()
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    host_lib().qgwh_diagnostics(ctypes.byref(p), _ptr(psi), _ptr(q), _ptr(El), _ptr(Zl),
                                ctypes.byref(Et), ctypes.byref(Zt))
    return El, Zl, Et.value, Zt.value


def check_timestep(p):
    dt, DT = ctypes.c_double(), ctypes.c_double()
    h = host_lib()
    h.qgwh_check_timestep(ctypes.byref(p), h.qgwh_nu_hyper_signed(ctypes.byref(p)), ctypes.byref(dt),
                          ctypes.byref(DT))
    return dt.value, DT.value


def stoch_ring(p):
    """Forced half-spectrum points (i, j) of the stochastic ring forcing in the
    reference's order, and the coefficient amplitude (init_stoch_forc,
    src/forcing.h:131-149,180)."""
    h = host_lib()
    n, env = ctypes.c_int(0), ctypes.c_double(0.0)
    st = h.qgwh_stoch_ring(ctypes.byref(p), 0, None, None, ctypes.byref(n), ctypes.byref(env))
    if st != QGW_OK:
        raise QGWError("stoch_ring: %s" % h.qgwh_strerror(st).decode())
    ii, jj = (ctypes.c_int * max(n.value, 1))(), (ctypes.c_int * max(n.value, 1))()
    h.qgwh_stoch_ring(ctypes.byref(p), n.value, ii, jj, ctypes.byref(n), ctypes.byref(env))
    return np.array(ii[:n.value]), np.array(jj[:n.value]), env.value


class Rng:
    """The forcing's random stream (seed 1: the C library's rand() stream)."""

    def __init__(self, seed=1):
        self._h = host_lib().qgwh_rng_create(seed)
        if not self._h:
            raise QGWError("rng_create failed")

    def next(self, n=1):
        return np.array([host_lib().qgwh_rng_next(self._h) for _ in range(n)])

    def coeffs(self, sigma_f, envelope, count):
        """One step of calc_forc's coefficients: (count, 2) re/im."""
        c = np.zeros(2 * max(count, 1))
        host_lib().qgwh_stoch_coeffs(self._h, sigma_f, envelope, count, _ptr(c))
        return c[:2 * count].reshape(count, 2)

    def __del__(self):
        if getattr(self, "_h", None):
            host_lib().qgwh_rng_destroy(self._h)
            self._h = None


class NcWriter:
    """netCDF classic output file of a run (create_nc / write_nc,
    src/netcdf_io.h:38-277): float record variables psi and q over
    (t, level, y, x); y and x cover the storage grid with its ring (bounded)
    or the NX x NY interior (periodic)."""

    def __init__(self, path, p):
        self.p = p
        self.g = grid(p)[0]
        st = ctypes.c_int(0)
        self._h = host_lib().qgwh_nc_create(os.fsencode(path), ctypes.byref(p), ctypes.byref(st))
        if not self._h:
            raise QGWError("nc_create(%s): %s" % (path, host_lib().qgwh_strerror(st.value).decode()))

    def write(self, t, psi, q):
        """Append one record: time t and the storage arrays psi, q
        (shape (nl, m+2, n+2), float64)."""
        shape = (self.p.nl, self.g.nyp2, self.g.nxp2)
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        if psi.shape != shape or q.shape != shape:
            raise QGWError("nc_write: fields must have shape %s" % (shape,))
        st = host_lib().qgwh_nc_write(self._h, float(t), _ptr(psi), _ptr(q))
        if st != QGW_OK:
            raise QGWError("nc_write: %s" % host_lib().qgwh_strerror(st).decode())

    @property
    def records(self):
        return host_lib().qgwh_nc_records(self._h)

    def close(self):
        if self._h:
            h, self._h = self._h, None
            st = host_lib().qgwh_nc_close(h)
            if st != QGW_OK:
                raise QGWError("nc_close: %s" % host_lib().qgwh_strerror(st).decode())

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "_h", None):
            host_lib().qgwh_nc_close(self._h)


def read_restart(path, p, rec=0, q=None, topo=None):
    """read_nc of a restart file (src/netcdf_io.h:286-386): variables "q"
    (nl levels) and "topo" (one level) at record rec, rounded to float as the
    reference reads them, into storage arrays.  q and topo are updated in
    place when given (a variable absent from the file leaves them untouched)
    and allocated as zeros otherwise.  Returns (q, topo, found) with found a
    set of the variable names read."""
    g = grid(p)[0]
    if q is None:
        q = np.zeros((p.nl, g.nyp2, g.nxp2))
    if topo is None:
        topo = np.zeros((g.nyp2, g.nxp2))
    for a, shape in ((q, (p.nl, g.nyp2, g.nxp2)), (topo, (g.nyp2, g.nxp2))):
        if a.shape != shape or a.dtype != np.float64 or not a.flags.c_contiguous:
            raise QGWError("read_restart: arrays must be contiguous float64 of shape %s" % (shape,))
    found = ctypes.c_int(0)
    st = host_lib().qgwh_nc_read(os.fsencode(path), ctypes.byref(p), int(rec), _ptr(q), _ptr(topo),
                                 ctypes.byref(found))
    if st != QGW_OK:
        raise QGWError("read_nc(%s): %s" % (path, host_lib().qgwh_strerror(st).decode()))
    names = {n for b, n in ((1, "q"), (2, "topo")) if found.value & b}
    return q, topo, names


# ---------------------------------------------------------------------------
# device model
# ---------------------------------------------------------------------------

class Model:
    """One run on the GPU: a ``qgw_ctx`` of libqgw_hip.so."""

    def __init__(self, p, precision=FP64, modes=None):
        self.lib = hip_lib()
        self.p = p
        self.precision = precision
        st = ctypes.c_int(0)
        self.ctx = self.lib.qgw_create(ctypes.byref(p), precision, ctypes.byref(st))
        if not self.ctx:
            raise QGWError("qgw_create failed: %s" % self.lib.qgw_strerror(st.value).decode())
        g = grid(p)[0]
        self.n, self.m, self.nl = g.n, g.m, p.nl
        self.shape = (p.nl, g.nyp2, g.nxp2)
        if modes is None:
            modes = eigmode(p)
        self.set_modes(*modes)

    def _chk(self, st, what):
        if st != QGW_OK:
            raise QGWError("%s: %s (%s)" % (what, self.lib.qgw_strerror(st).decode(),
                                            self.lib.qgw_last_error(self.ctx).decode()))

    def close(self):
        if self.ctx:
            self.lib.qgw_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_modes(self, cl2m, cm2l, iRd2):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (cl2m, cm2l, iRd2)]
        self._chk(self.lib.qgw_set_modes(self.ctx, _ptr(a[0]), _ptr(a[1]), _ptr(a[2])), "set_modes")

    def upload(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        need = self.shape[1] * self.shape[2] * (1 if field == FIELD_TOPO else self.nl)
        if a.size != need:
            raise QGWError("upload: expected %d values, got %d" % (need, a.size))
        self._chk(self.lib.qgw_upload(self.ctx, field, _ptr(a)), "upload")

    def download(self, field):
        shape = (self.shape[1], self.shape[2]) if field == FIELD_TOPO else self.shape
        a = np.zeros(shape)
        self._chk(self.lib.qgw_download(self.ctx, field, _ptr(a)), "download")
        return a

    def set_q(self, q):
        self.upload(FIELD_Q, q)

    @property
    def q(self):
        return self.download(FIELD_Q)

    @property
    def psi(self):
        return self.download(FIELD_PSI)

    def invert_pv(self):
        self._chk(self.lib.qgw_invert_pv(self.ctx), "invert_pv")

    def rhs(self):
        self._chk(self.lib.qgw_rhs(self.ctx), "rhs")
        return self.download(FIELD_F0)

    def adjust_timestep(self):
        dt = ctypes.c_double()
        self._chk(self.lib.qgw_adjust_timestep(self.ctx, ctypes.byref(dt)), "adjust_timestep")
        return dt.value

    def timestep(self):
        self._chk(self.lib.qgw_timestep(self.ctx), "timestep")

    def run(self, nsteps):
        self._chk(self.lib.qgw_run(self.ctx, int(nsteps)), "run")

    def time_run(self, nsteps):
        ms = ctypes.c_double()
        self._chk(self.lib.qgw_time_run(self.ctx, int(nsteps), ctypes.byref(ms)), "time_run")
        return ms.value

    def synchronize(self):
        self._chk(self.lib.qgw_synchronize(self.ctx), "synchronize")

    def diagnostics(self):
        El, Zl = np.zeros(self.nl), np.zeros(self.nl)
        Et, Zt = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.qgw_diagnostics(self.ctx, _ptr(El), _ptr(Zl), ctypes.byref(Et), ctypes.byref(Zt)),
                  "diagnostics")
        return El, Zl, Et.value, Zt.value

    def get_time(self):
        t, dt, dt0, dt1 = (ctypes.c_double() for _ in range(4))
        it = ctypes.c_int()
        self._chk(self.lib.qgw_get_time(self.ctx, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(dt0),
                                        ctypes.byref(dt1), ctypes.byref(it)), "get_time")
        return dict(t=t.value, dt=dt.value, dt0=dt0.value, dt1=dt1.value, it=it.value)

    def set_time(self, t, dt, dt0, dt1, it):
        self._chk(self.lib.qgw_set_time(self.ctx, t, dt, dt0, dt1, it), "set_time")

    def set_next_output(self, t_out):
        self._chk(self.lib.qgw_set_next_output(self.ctx, t_out), "set_next_output")

    def set_stoch_seed(self, seed):
        """Restart the stochastic forcing's random stream (seed 1: the C
        library's rand() stream, the reference's)."""
        self._chk(self.lib.qgw_set_stoch_seed(self.ctx, seed), "set_stoch_seed")

    def set_forcing_file(self, path):
        """File of the 4-D forcing records (variable q_forcing_3d); default
        input_vars.nc in the working directory."""
        self._chk(self.lib.qgw_set_forcing_file(self.ctx, os.fsencode(path)), "set_forcing_file")

    def profile(self, on=True):
        self._chk(self.lib.qgw_profile_enable(self.ctx, 1 if on else 0), "profile_enable")

    def profile_read(self):
        ms = np.zeros(len(PASS_NAMES))
        cnt = (ctypes.c_int * len(PASS_NAMES))()
        self._chk(self.lib.qgw_profile_read(self.ctx, _ptr(ms), cnt, len(PASS_NAMES)), "profile_read")
        return {PASS_NAMES[i]: (ms[i], cnt[i]) for i in range(len(PASS_NAMES))}

    def tune_set(self, knob, value):
        self._chk(self.lib.qgw_tune_set(self.ctx, knob, value), "tune_set")

    def tune_get(self, knob):
        v = ctypes.c_int()
        self._chk(self.lib.qgw_tune_get(self.ctx, knob, ctypes.byref(v)), "tune_get")
        return v.value

    def tuning(self):
        """The tuning choices, by name."""
        return {"p4": P4_TILES[self.tune_get(TUNE_P4_TILE)], "p4_waves": self.tune_get(TUNE_P4_WAVES),
                "t_kernel": ["auto", "valu", "mfma32", "mfma16"][self.tune_get(TUNE_T_KERNEL)],
                "dst": ["odd-extension", "half-length", "half-length pipelined", "dense (matrix cores)"][self.tune_get(TUNE_DST)],
                "fft_e": self.tune_get(TUNE_FFT_E), "fft_e_col": self.tune_get(TUNE_FFT_E_COL),
                "passes": 4 if self.tune_get(TUNE_FUSE) else 6, "p2m_pairs": self.tune_get(TUNE_P2M_PAIRS),
                "psi_nt": self.tune_get(TUNE_PSI_NT), "p4_rev": self.tune_get(TUNE_P4_REV),
                "p4_order": ["chunks first", "layers first"][self.tune_get(TUNE_P4_ORDER)]}

    def autotune(self, reps=5, cache=True):
        """Pick the tuning of this model by timing each variant on its own
        state (qgw_autotune copies the state first and restores it after).
        Choices are kept per device, form and shape in the JSON file of
        tune_cache_path(), so later processes reuse them; cache=False retunes."""
        key = tune_key(self.precision, self.n, self.m, self.nl, self.p.n_hyper, self.p.nu_hyper != 0)
        choice = tune_cache_load().get(key) if cache else None
        if choice is None:
            t, ms = ctypes.c_int(), ctypes.c_double()
            self._chk(self.lib.qgw_autotune(self.ctx, reps, ctypes.byref(t), ctypes.byref(ms)), "autotune")
            choice = {str(k): self.tune_get(k) for k in TUNE_KNOBS}
            choice["ms_per_step"] = ms.value
            tune_cache_store({key: choice})
        else:
            for k in TUNE_KNOBS:
                if str(k) in choice:
                    self.tune_set(k, int(choice[str(k)]))
        return dict(choice, cached=choice is not None and cache)

    def device_bytes(self):
        b = ctypes.c_size_t()
        self._chk(self.lib.qgw_device_bytes(self.ctx, ctypes.byref(b)), "device_bytes")
        return b.value


class Batch:
    """An ensemble on the GPU: a ``qgw_batch`` of libqgw_hip.so.

    ``params`` is a list of Params, one per member (each needs dt_fixed > 0).
    Members of equal grid and step scalars share every launch of a step; a
    step of all members is replayed from a captured hipGraph."""

    def __init__(self, params, precision=FP32, modes=None):
        self.lib = hip_lib()
        self.params = list(params)
        self.precision = precision
        n = len(self.params)
        arr = (Params * n)(*self.params)
        st = ctypes.c_int(0)
        self.b = self.lib.qgw_batch_create(n, arr, precision, ctypes.byref(st))
        if not self.b:
            raise QGWError("qgw_batch_create failed: %s" % self.lib.qgw_strerror(st.value).decode())
        self.shapes = []
        for p in self.params:
            g = grid(p)[0]
            self.shapes.append((p.nl, g.nyp2, g.nxp2))
        for i, p in enumerate(self.params):
            self.set_modes(i, *(modes[i] if modes is not None else eigmode(p)))

    def __len__(self):
        return len(self.params)

    def _chk(self, st, what):
        if st != QGW_OK:
            raise QGWError("%s: %s (%s)" % (what, self.lib.qgw_strerror(st).decode(),
                                            self.lib.qgw_batch_last_error(self.b).decode()))

    def close(self):
        if self.b:
            self.lib.qgw_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def groups(self):
        ng = ctypes.c_int()
        gm = (ctypes.c_int * len(self.params))()
        self._chk(self.lib.qgw_batch_groups(self.b, ctypes.byref(ng), gm), "groups")
        return ng.value, list(gm)

    def set_modes(self, member, cl2m, cm2l, iRd2):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (cl2m, cm2l, iRd2)]
        self._chk(self.lib.qgw_batch_set_modes(self.b, member, _ptr(a[0]), _ptr(a[1]), _ptr(a[2])), "set_modes")

    def upload(self, member, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        nl, ny, nx = self.shapes[member]
        need = ny * nx * (1 if field == FIELD_TOPO else nl)
        if a.size != need:
            raise QGWError("upload: expected %d values, got %d" % (need, a.size))
        self._chk(self.lib.qgw_batch_upload(self.b, member, field, _ptr(a)), "upload")

    def download(self, member, field):
        nl, ny, nx = self.shapes[member]
        a = np.zeros((ny, nx) if field == FIELD_TOPO else (nl, ny, nx))
        self._chk(self.lib.qgw_batch_download(self.b, member, field, _ptr(a)), "download")
        return a

    def set_q(self, member, q):
        self.upload(member, FIELD_Q, q)

    def q(self, member):
        return self.download(member, FIELD_Q)

    def psi(self, member):
        return self.download(member, FIELD_PSI)

    def invert_pv(self):
        self._chk(self.lib.qgw_batch_invert_pv(self.b), "invert_pv")

    def step(self, nsteps):
        self._chk(self.lib.qgw_batch_step(self.b, int(nsteps)), "step")

    def time_step(self, nsteps):
        ms = ctypes.c_double()
        self._chk(self.lib.qgw_batch_time_step(self.b, int(nsteps), ctypes.byref(ms)), "time_step")
        return ms.value

    def synchronize(self):
        self._chk(self.lib.qgw_batch_synchronize(self.b), "synchronize")

    def prepare(self):
        """Capture the graph of the current state now (no step is run)."""
        self._chk(self.lib.qgw_batch_prepare(self.b), "prepare")

    def set_stoch_seed(self, member, seed):
        self._chk(self.lib.qgw_batch_set_stoch_seed(self.b, member, int(seed)), "set_stoch_seed")

    def set_forcing_file(self, member, path):
        self._chk(self.lib.qgw_batch_set_forcing_file(self.b, member, os.fsencode(path)), "set_forcing_file")

    def diagnostics(self, member):
        nl = self.shapes[member][0]
        El, Zl = np.zeros(nl), np.zeros(nl)
        Et, Zt = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.qgw_batch_diagnostics(self.b, member, _ptr(El), _ptr(Zl), ctypes.byref(Et),
                                                 ctypes.byref(Zt)), "diagnostics")
        return El, Zl, Et.value, Zt.value

    def get_time(self):
        t, it = ctypes.c_double(), ctypes.c_int()
        self._chk(self.lib.qgw_batch_get_time(self.b, ctypes.byref(t), ctypes.byref(it)), "get_time")
        return dict(t=t.value, it=it.value)

    def use_graph(self, on=True):
        self._chk(self.lib.qgw_batch_use_graph(self.b, 1 if on else 0), "use_graph")

    def graph_launches(self):
        c = ctypes.c_int()
        self._chk(self.lib.qgw_batch_graph_launches(self.b, ctypes.byref(c)), "graph_launches")
        return c.value

    def _group_keys(self):
        ng, gof = self.groups()
        keys = []
        for g in range(ng):
            mem = [i for i in range(len(gof)) if gof[i] == g]
            p = self.params[mem[0]]
            nl, m2, n2 = self.shapes[mem[0]]
            keys.append(tune_key(self.precision, n2 - 2, m2 - 2, nl * len(mem), p.n_hyper, p.nu_hyper != 0))
        return keys

    def autotune(self, reps=3, cache=True):
        """Tune every group (as Model.autotune); choices come from and go to
        the JSON cache of tune_cache_path() unless cache=False."""
        keys = self._group_keys()
        known = tune_cache_load() if cache else {}
        if all(k in known for k in keys):
            for g, k in enumerate(keys):
                for knob in TUNE_KNOBS:
                    if str(knob) in known[k]:
                        self.tune_set(g, knob, int(known[k][str(knob)]))
        else:
            v = (ctypes.c_int * len(keys))()
            self._chk(self.lib.qgw_batch_autotune(self.b, reps, v), "autotune")
            tune_cache_store({k: {str(knob): self.tune_get(g, knob) for knob in TUNE_KNOBS}
                              for g, k in enumerate(keys)})
        return [self.tuning(g)["p4"] for g in range(len(keys))]

    def tune_set(self, group, knob, value):
        self._chk(self.lib.qgw_batch_tune_set(self.b, group, knob, value), "tune_set")

    def tune_get(self, group, knob):
        v = ctypes.c_int()
        self._chk(self.lib.qgw_batch_tune_get(self.b, group, knob, ctypes.byref(v)), "tune_get")
        return v.value

    def tuning(self, group=0):
        """The tuning choices of one group, by name."""
        return {"p4": P4_TILES[self.tune_get(group, TUNE_P4_TILE)],
                "p4_waves": self.tune_get(group, TUNE_P4_WAVES),
                "t_kernel": ["auto", "valu", "mfma32", "mfma16"][self.tune_get(group, TUNE_T_KERNEL)],
                "dst": ["odd-extension", "half-length", "half-length pipelined", "dense (matrix cores)"][self.tune_get(group, TUNE_DST)],
                "fft_e": self.tune_get(group, TUNE_FFT_E), "fft_e_col": self.tune_get(group, TUNE_FFT_E_COL),
                "passes": 4 if self.tune_get(group, TUNE_FUSE) else 6,
                "p2m_pairs": self.tune_get(group, TUNE_P2M_PAIRS),
                "psi_nt": self.tune_get(group, TUNE_PSI_NT), "p4_rev": self.tune_get(group, TUNE_P4_REV),
                "p4_order": ["chunks first", "layers first"][self.tune_get(group, TUNE_P4_ORDER)]}

    def profile(self, on=True):
        self._chk(self.lib.qgw_batch_profile_enable(self.b, 1 if on else 0), "profile_enable")

    def profile_read(self):
        ms = np.zeros(len(PASS_NAMES))
        cnt = (ctypes.c_int * len(PASS_NAMES))()
        self._chk(self.lib.qgw_batch_profile_read(self.b, _ptr(ms), cnt, len(PASS_NAMES)), "profile_read")
        return {PASS_NAMES[i]: (ms[i], cnt[i]) for i in range(len(PASS_NAMES))}

    def device_bytes(self):
        b = ctypes.c_size_t()
        self._chk(self.lib.qgw_batch_device_bytes(self.b, ctypes.byref(b)), "device_bytes")
        return b.value
