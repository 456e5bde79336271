"""TEST INFRASTRUCTURE ONLY -- ctypes view of the oracle (liboracle.so).

The oracle is a CPU FP64 restatement of the reference's timestep hot path
(see qgw_oracle.h for the function -> reference file:line map and the pins).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
it, and only as the checker; the product never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NL_MAX = 1000
Q, PSI, TOPO, F0, F1, F2, LAPN, FORC = 1, 2, 3, 4, 5, 6, 7, 8

# planted mistakes of orc_mutate (qgw_oracle.h ORC_MUT_*), for tests that show
# which mistake each pin catches
MUT = dict(none=0, beta_sign=1, beta_axis=2, nu_psi=3, wind_layer=4, wind_dh=5, wind_lx=6, drag_dh=7,
           drag_half=8, drag_q=9, topo_swap=10, topo_layer=11, bc_factor=12, bc_order=13, topo_lx=14,
           shelf_width=15)

_dp = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("NX", ctypes.c_int), ("NY", ctypes.c_int), ("nl", ctypes.c_int), ("n_hyper", ctypes.c_int),
        ("Lx", ctypes.c_double), ("dh", ctypes.c_double * NL_MAX), ("N2", ctypes.c_double * NL_MAX),
        ("tau0", ctypes.c_double), ("forc_mode", ctypes.c_double), ("f0", ctypes.c_double),
        ("beta", ctypes.c_double), ("nu", ctypes.c_double), ("nu_hyper", ctypes.c_double),
        ("hEkb", ctypes.c_double), ("Ld", ctypes.c_double), ("h_topo", ctypes.c_double),
        ("w_topo", ctypes.c_double), ("bc_fac", ctypes.c_double), ("cfl", ctypes.c_double),
        ("sigma_f", ctypes.c_double), ("tend", ctypes.c_double), ("dt_out", ctypes.c_double),
        ("dt_fixed", ctypes.c_double),
    ]


_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams), ip]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_modes.argtypes = [vp, _dp, _dp, _dp]
        L.orc_set_modes.argtypes = [vp, _dp, _dp, _dp]
        L.orc_field.argtypes = [vp, ctypes.c_int]
        L.orc_field.restype = _dp
        L.orc_sizes.argtypes = [vp, ip, ip, ip]
        for f in ("orc_invert_pv", "orc_calc_diff", "orc_timestep"):
            getattr(L, f).argtypes = [vp]
        L.orc_adjust_bc.argtypes = [vp, _dp, ctypes.c_int]
        L.orc_rhs.argtypes = [vp, ctypes.c_int]
        L.orc_adjust_timestep.argtypes = [vp]
        L.orc_adjust_timestep.restype = ctypes.c_double
        L.orc_run.argtypes = [vp, ctypes.c_int]
        L.orc_set_forcing4d.argtypes = [vp, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_set_stochastic.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.orc_stoch_points.argtypes = [vp, ip, ip]
        L.orc_stoch_draw.argtypes = [vp, _dp]
        L.orc_c2r2d.argtypes = [_dp, ctypes.c_int, _dp]
        L.orc_get_time.argtypes = [vp, _dp, _dp, _dp, _dp, ip, _dp]
        L.orc_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_double]
        L.orc_diagnostics.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.orc_ab3_coeffs.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_nu_hyper.argtypes = [vp]
        L.orc_nu_hyper.restype = ctypes.c_double
        L.orc_dst1d.argtypes = [_dp, ctypes.c_int]
        L.orc_dst2d.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        L.orc_fft.argtypes = [_dp, _dp, ctypes.c_int]
        L.orc_eig_sym.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_check_timestep.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp]
        L.orc_jacobian_field.argtypes = [vp, _dp, _dp, _dp]
        L.orc_laplacian_field.argtypes = [vp, _dp, _dp]
        L.orc_mutate.argtypes = [ctypes.c_int]
        L.orc_threads.argtypes = [ctypes.c_int]
        L.orc_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def params_from(p):
    """OrcParams from a qgw_amd.Params (same meanings) or a dict."""
    d = p.as_dict() if hasattr(p, "as_dict") else dict(p)
    o = OrcParams()
    for name, _ in OrcParams._fields_:
        if name in ("dh", "N2"):
            arr = getattr(o, name)
            for i, v in enumerate(d.get(name, [])):
                arr[i] = v
        elif name in d:
            setattr(o, name, d[name])
    return o


class Oracle:
    def __init__(self, p):
        self.L = lib()
        self.op = params_from(p)
        st = ctypes.c_int(0)
        self.s = self.L.orc_create(ctypes.byref(self.op), ctypes.byref(st))
        if not self.s:
            raise RuntimeError("orc_create failed (%d)" % st.value)
        n, m, nl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.L.orc_sizes(self.s, ctypes.byref(n), ctypes.byref(m), ctypes.byref(nl))
        self.n, self.m, self.nl = n.value, m.value, nl.value
        self.shape = (self.nl, self.m + 2, self.n + 2)

    def close(self):
        if self.s:
            self.L.orc_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, which):
        """A numpy view of an oracle field (shares memory; copy to keep)."""
        ptr = self.L.orc_field(self.s, which)
        one = which in (TOPO, FORC)
        size = self.shape[1] * self.shape[2] * (1 if one else self.nl)
        a = np.ctypeslib.as_array(ptr, shape=(size,))
        return a.reshape((self.shape[1], self.shape[2]) if one else self.shape)

    def modes(self):
        nl = self.nl
        a, b, c = np.zeros(nl * nl), np.zeros(nl * nl), np.zeros(nl)
        self.L.orc_modes(self.s, _p(a), _p(b), _p(c))
        return a, b, c

    def set_modes(self, cl2m, cm2l, iRd2):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (cl2m, cm2l, iRd2)]
        self.L.orc_set_modes(self.s, _p(a[0]), _p(a[1]), _p(a[2]))

    def set_q(self, q):
        self.field(Q)[...] = q

    def invert_pv(self):
        self.L.orc_invert_pv(self.s)

    def set_stochastic(self, k_f, sigma_f):
        """Stochastic ring forcing; resets the C library's rand() stream.
        Returns the number of forced points."""
        n = self.L.orc_set_stochastic(self.s, float(k_f), float(sigma_f))
        if n < 0:
            raise ValueError("stochastic forcing needs an NX x NX field inside the grid")
        return n

    def stoch_points(self):
        n = self.L.orc_stoch_points(self.s, None, None)
        ii, jj = (ctypes.c_int * max(n, 1))(), (ctypes.c_int * max(n, 1))()
        self.L.orc_stoch_points(self.s, ii, jj)
        return np.array(ii[:n]), np.array(jj[:n])

    def stoch_draw(self):
        n = self.L.orc_stoch_points(self.s, None, None)
        c = np.zeros(2 * max(n, 1))
        self.L.orc_stoch_draw(self.s, _p(c))
        return c[:2 * n]

    def set_forcing4d(self, records, dt_forc):
        """4-D forcing records, shape (n_rec, nl, m+2, n+2); None switches it off."""
        if records is None:
            self.L.orc_set_forcing4d(self.s, 0, 0.0, None)
            return
        r = np.ascontiguousarray(records, dtype=np.float64)
        if r.shape[1:] != self.shape:
            raise ValueError("forcing records must have shape (n_rec,) + %s" % (self.shape,))
        if self.L.orc_set_forcing4d(self.s, r.shape[0], float(dt_forc), _p(r)) != 0:
            raise ValueError("bad forcing records")

    def adjust_bc(self, field, field_type):
        """orc_adjust_bc on a copy of `field` (storage shape), reading the
        oracle's current psi; returns the copy."""
        f = np.ascontiguousarray(field, dtype=np.float64).copy()
        if f.shape != self.shape:
            raise ValueError("field must have shape %s" % (self.shape,))
        self.L.orc_adjust_bc(self.s, _p(f), int(field_type))
        return f

    def calc_diff(self):
        self.L.orc_calc_diff(self.s)

    def rhs(self, which=F0):
        self.L.orc_rhs(self.s, which)
        return self.field(which).copy()

    def adjust_timestep(self):
        return self.L.orc_adjust_timestep(self.s)

    def timestep(self):
        self.L.orc_timestep(self.s)

    def run(self, nsteps):
        self.L.orc_run(self.s, int(nsteps))

    def get_time(self):
        t, dt, dt0, dt1, DT = (ctypes.c_double() for _ in range(5))
        it = ctypes.c_int()
        self.L.orc_get_time(self.s, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(dt0), ctypes.byref(dt1),
                            ctypes.byref(it), ctypes.byref(DT))
        return dict(t=t.value, dt=dt.value, dt0=dt0.value, dt1=dt1.value, it=it.value, DT_max=DT.value)

    def set_time(self, t, dt, dt0, dt1, it, t_out):
        self.L.orc_set_time(self.s, t, dt, dt0, dt1, it, t_out)

    def diagnostics(self):
        El, Zl = np.zeros(self.nl), np.zeros(self.nl)
        Et, Zt = ctypes.c_double(), ctypes.c_double()
        self.L.orc_diagnostics(self.s, _p(El), _p(Zl), ctypes.byref(Et), ctypes.byref(Zt))
        return El, Zl, Et.value, Zt.value

    def jacobian(self, p, q):
        p = np.ascontiguousarray(p, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros(self.shape)
        self.L.orc_jacobian_field(self.s, _p(p), _p(q), _p(out))
        return out

    def laplacian(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        out = np.zeros(self.shape)
        self.L.orc_laplacian_field(self.s, _p(p), _p(out))
        return out

    def nu_hyper(self):
        return self.L.orc_nu_hyper(self.s)


def dst1d(x):
    a = np.ascontiguousarray(x, dtype=np.float64).copy()
    lib().orc_dst1d(_p(a), a.size)
    return a


def dst2d(a):
    b = np.ascontiguousarray(a, dtype=np.float64).copy()
    m, n = b.shape
    lib().orc_dst2d(_p(b), n, m)
    return b


def fft(z):
    re = np.ascontiguousarray(z.real, dtype=np.float64).copy()
    im = np.ascontiguousarray(z.imag, dtype=np.float64).copy()
    lib().orc_fft(_p(re), _p(im), re.size)
    return re + 1j * im


def eig_sym(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    w, V = np.zeros(n), np.zeros((n, n))
    lib().orc_eig_sym(n, _p(A), _p(w), _p(V))
    return w, V


def ab3_coeffs(dt, dt0, dt1):
    c = np.zeros(3)
    lib().orc_ab3_coeffs(dt, dt0, dt1, _p(c))
    return c


def check_timestep(p):
    o = params_from(p)
    dt, DT = ctypes.c_double(), ctypes.c_double()
    lib().orc_check_timestep(ctypes.byref(o), ctypes.byref(dt), ctypes.byref(DT))
    return dt.value, DT.value


def mutate(name):
    """Plant one of the mistakes of MUT in the oracle ("none" restores it)."""
    lib().orc_mutate(MUT[name])


def threads(n=0):
    return lib().orc_threads(n)
