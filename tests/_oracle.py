"""ctypes wrapper for the CPU oracle (oracle/_build/libmlegs_oracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product package
(mlegs_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_ROOT, "oracle", "_build", "libmlegs_oracle.so")

PPP, FFF, PFP, PFF = 0, 1, 2, 3


class OgParams(C.Structure):
    _fields_ = [
        ("nr", C.c_int), ("nth", C.c_int), ("nx", C.c_int),
        ("nrchop", C.c_int), ("ntchop", C.c_int), ("nxchop", C.c_int),
        ("minc", C.c_int), ("mklink", C.c_int),
        ("ell", C.c_double), ("zlen", C.c_double),
        ("nu", C.c_double), ("kappa", C.c_double), ("nup", C.c_double), ("kappap", C.c_double),
        ("visc_sw", C.c_int), ("visc_p", C.c_int),
        ("omega", C.c_double), ("bv0", C.c_double),
        ("adams", C.c_int), ("hyper_cn", C.c_int),
        ("dt", C.c_double),
    ]


class OgField(C.Structure):
    _fields_ = [("e", C.c_void_p), ("ln", C.c_double), ("space", C.c_int), ("len", C.c_long)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(OgField)
        g = C.c_void_p
        sigs = {
            "og_init": ([C.POINTER(OgParams)], g),
            "og_free": ([g], None),
            "og_dims": ([g, C.POINTER(C.c_int)], None),
            "og_table": ([g, C.c_char_p, C.POINTER(C.c_int)], C.POINTER(C.c_double)),
            "og_chopset": ([g, C.c_int], None),
            "og_nrchop": ([g], C.c_int),
            "og_field_new": ([g, C.c_int], P),
            "og_field_free": ([P], None),
            "og_pool_drain": ([], None),
            "og_field_size": ([g, C.c_int], C.c_long),
            "og_chopdo": ([g, P], None),
            "og_copy": ([g, P, P], None),
            "og_horfft": ([g, P, C.c_int], C.c_int),
            "og_verfft": ([g, P, C.c_int], C.c_int),
            "og_rtran": ([g, P, C.c_int], C.c_int),
            "og_toff": ([g, P], C.c_int),
            "og_tofp": ([g, P], C.c_int),
            "og_delsqh": ([g, P, P], C.c_int),
            "og_idelsqh": ([g, P, P], C.c_int),
            "og_mulxm": ([g, P, P], C.c_int),
            "og_xxdx": ([g, P, P], C.c_int),
            "og_del2": ([g, P, P], C.c_int),
            "og_idel2": ([g, P, P, C.c_int, C.c_double], C.c_int),
            "og_idel2ln": ([g, P, P], C.c_int),
            "og_ihelm": ([g, P, P, C.c_double], C.c_int),
            "og_ihelmp": ([g, C.c_int, P, P, C.c_double, C.c_double], C.c_int),
            "og_helmp": ([g, C.c_int, P, P, C.c_double, C.c_double], C.c_int),
            "og_grad": ([g, P, P, P, P], C.c_int),
            "og_pc2vel": ([g, P, P, P, P, P], C.c_int),
            "og_pc2vor": ([g, P, P, P, P, P], C.c_int),
            "og_vprodsub": ([g, P, P, P, P, P, P], C.c_int),
            "og_vdotpsub": ([g, P, P, P, P, P, P, P], C.c_int),
            "og_vel2pln": ([g, P, P, C.c_int, C.c_double], C.c_int),
            "og_project": ([g, P, P, P, P, P, C.c_int, C.c_double], C.c_int),
            "og_force": ([g, P, P, P, P, P, P], C.c_int),
            "og_nonlin": ([g, P, P, P, P, P, P, P, P], C.c_int),
            "og_etd_op": ([g, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)], None),
            "og_energy": ([g, P, P, P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
            "og_solver_new": ([g], C.c_void_p),
            "og_solver_free": ([C.c_void_p], None),
            "og_initialize": ([C.c_void_p, P, P, P], C.c_int),
            "og_time_step": ([C.c_void_p, P, P, P], C.c_int),
            "og_solver_prepare": ([C.c_void_p], C.c_int),
            "og_diagnost": ([g, P, P], C.c_int),
            "og_time": ([C.c_void_p], C.c_double),
            "og_nstep": ([C.c_void_p], C.c_int),
            "og_qvortex": ([g, C.c_double, P, P], C.c_int),
            "og_diagnost_it1": ([g, P, P, C.POINTER(C.c_double)], C.c_int),
            "og_set_sponge": ([g, C.c_int], None),
            "og_sponge": ([g, P, P, P, C.c_double], C.c_int),
            "og_ringcut": ([g, P, C.c_int, C.c_int, C.c_int], C.c_int),
            "og_smooth": ([C.POINTER(C.c_double), C.c_int, C.c_int], C.c_int),
            "og_remove": ([g, P, P, P], C.c_int),
            "og_enemon": ([g, P, P, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
            "og_specurv": ([C.POINTER(C.c_double), C.c_int, C.c_double], C.c_double),
            "og_speslop": ([C.POINTER(C.c_double), C.c_int, C.c_double], C.c_double),
            "og_pdajust": ([C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                            C.c_int, C.c_double], None),
            "og_hyp3adj": ([g, P, P, P], C.c_int),
            "og_hyperv_coeffs": ([g, C.POINTER(C.c_double)], None),
            "og_set_hyperv_coeffs": ([g, C.POINTER(C.c_double)], None),
            "og_set_march_controls": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int], None),
            "og_set_threads": ([C.c_int], None),
            "og_set_sum_order": ([C.c_int], None),
            "og_set_mutation": ([C.c_int], None),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def make_params(cfg):
    p = OgParams()
    for k, _ in OgParams._fields_:
        if k in cfg:
            setattr(p, k, cfg[k])
    if p.minc == 0:
        p.minc = 1
    return p


class OField:
    """An oracle field; `.arr()` is a live Fortran-ordered numpy view."""

    def __init__(self, grid, space=FFF, ptr=None):
        self.grid = grid
        self.ptr = ptr if ptr is not None else lib().og_field_new(grid.g, space)

    def __del__(self):
        try:
            if self.ptr and self.grid.g:
                lib().og_field_free(self.ptr)
        except Exception:
            pass

    @property
    def space(self):
        return self.ptr.contents.space

    @space.setter
    def space(self, v):
        self.ptr.contents.space = v

    @property
    def ln(self):
        return self.ptr.contents.ln

    @ln.setter
    def ln(self, v):
        self.ptr.contents.ln = float(v)

    def arr(self, space=None):
        sp = self.space if space is None else space
        shape = self.grid.shape(sp)
        n = int(np.prod(shape))
        buf = (C.c_double * (2 * n)).from_address(self.ptr.contents.e)
        return np.frombuffer(buf, dtype=np.complex128, count=n).reshape(shape, order="F")

    def set(self, a, space, ln=0.0):
        self.space = space
        v = self.arr(space)
        v[...] = a
        self.ln = ln
        return self

    def get(self):
        return self.arr().copy(order="F")


class OGrid:
    def __init__(self, cfg, threads=0):
        self.cfg = dict(cfg)
        self.params = make_params(cfg)
        if threads:
            lib().og_set_threads(threads)
        self.g = lib().og_init(C.byref(self.params))
        if not self.g:
            raise ValueError("og_init rejected the configuration")
        d = (C.c_int * 12)()
        lib().og_dims(self.g, d)
        (self.ndimr, self.ndimth, self.ndimx, self.nrchopdim, self.ntchopdim, self.nxchopdim,
         self.nrh, self.nxchoph, _, self.npf, self.nnorm, self.nr) = list(d)

    def close(self):
        if self.g:
            lib().og_pool_drain()
            lib().og_free(self.g)
            self.g = None

    def shape(self, space):
        if space == FFF:
            return (self.nrchopdim, self.ntchopdim, self.nxchopdim)
        if space == PFF:
            return (self.ndimr, self.ntchopdim, self.nxchopdim)
        return (self.ndimr, self.ndimth, self.ndimx)

    def table(self, name):
        n = C.c_int(0)
        p = lib().og_table(self.g, name.encode(), C.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def pf(self):
        return self.table("pf").reshape((self.nrh, self.npf, self.ntchopdim), order="F")

    @property
    def nrchop(self):
        return lib().og_nrchop(self.g)

    def chopset(self, iof):
        lib().og_chopset(self.g, iof)

    def chopdo(self, f):
        lib().og_chopdo(self.g, f.ptr)

    def field(self, space=FFF, data=None, ln=0.0):
        f = OField(self, space)
        if data is not None:
            f.set(data, space, ln)
        return f

    def call(self, name, *args):
        conv = [a.ptr if isinstance(a, OField) else a for a in args]
        rc = getattr(lib(), "og_" + name)(self.g, *conv)
        if rc:
            raise RuntimeError(f"og_{name} returned {rc}")
        return rc

    def etd_op(self, dt):
        e = (C.c_double * 16)()
        nl = (C.c_double * 16)()
        lib().og_etd_op(self.g, dt, e, nl)
        return (np.array(e).reshape(4, 4, order="F"), np.array(nl).reshape(4, 4, order="F"))

    def energy(self, psi, chi, b):
        nt, nxc = self.cfg["ntchop"], self.cfg["nxchop"]
        ke = (C.c_double * (nt * nxc))()
        pe = (C.c_double * (nt * nxc))()
        dke = (C.c_double * (nt * nxc))()
        rc = lib().og_energy(self.g, psi.ptr, chi.ptr, b.ptr, ke, pe, dke)
        if rc:
            raise RuntimeError(f"og_energy returned {rc}")
        f = lambda a: np.array(a).reshape(nt, nxc, order="F")
        return f(ke), f(pe), f(dke)


    def enemon(self, psi, chi):
        """ENEMON at the current truncation: (EM[NTCHOP], EK[NXCHOP])."""
        nt, nxc = self.cfg["ntchop"], self.cfg["nxchop"]
        em, ek = (C.c_double * nt)(), (C.c_double * nxc)()
        rc = lib().og_enemon(self.g, psi.ptr, chi.ptr, em, ek)
        if rc:
            raise RuntimeError(f"og_enemon returned {rc}")
        return np.array(em), np.array(ek)

    def diagnost_it1(self, psi, chi):
        """DIAGNOST's every-IT1 quantities, in mlegs_amd.hip.DIAGNOST_IT1_KEYS order."""
        o = (C.c_double * 12)()
        rc = lib().og_diagnost_it1(self.g, psi.ptr, chi.ptr, o)
        if rc:
            raise RuntimeError(f"og_diagnost_it1 returned {rc}")
        return np.array(o)

    def hyperv_coeffs(self):
        o = (C.c_double * 6)()
        lib().og_hyperv_coeffs(self.g, o)
        return np.array(o)

    def set_hyperv_coeffs(self, c):
        lib().og_set_hyperv_coeffs(self.g, (C.c_double * 6)(*[float(v) for v in c]))

    def hyperv_tables(self):
        """(H_R [NRCHOP x NTCHOPDIM], H_T, H_X, D_R, D_T, D_X) of the filters."""
        nrc, ntd, nxd = self.cfg["nrchop"], self.ntchopdim, self.nxchopdim
        out = []
        for name, n in (("hv_r", nrc * ntd), ("hv_t", ntd), ("hv_x", nxd), ("hd_r", nrc * ntd), ("hd_t", ntd),
                        ("hd_x", nxd)):
            out.append(self.table(name)[:n].copy())
        return out


def smooth(a, passes=1):
    """SMOOTH applied `passes` times to a complex vector."""
    v = np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64).copy()
    rc = lib().og_smooth(v.ctypes.data_as(C.POINTER(C.c_double)), len(v) // 2, passes)
    if rc:
        raise RuntimeError(f"og_smooth returned {rc}")
    return v.view(np.complex128)


def specurv(espec, ratio):
    a = np.ascontiguousarray(espec, dtype=np.float64)
    return lib().og_specurv(a.ctypes.data_as(C.POINTER(C.c_double)), len(a), float(ratio))


def speslop(espec, ratio=1.0 / 3.0):
    a = np.ascontiguousarray(espec, dtype=np.float64)
    return lib().og_speslop(a.ctypes.data_as(C.POINTER(C.c_double)), len(a), float(ratio))


def pdajust(coeff, cur, prev, base, params, thr=None):
    c = C.c_double(coeff)
    lib().og_pdajust(C.byref(c), cur, prev, base, (C.c_double * 4)(*params), 0 if thr is None else 1,
                     0.0 if thr is None else float(thr))
    return c.value


class OSolver:
    def __init__(self, grid):
        self.grid = grid
        self.s = lib().og_solver_new(grid.g)

    def initialize(self, psi, chi, b):
        rc = lib().og_initialize(self.s, psi.ptr, chi.ptr, b.ptr)
        if rc:
            raise RuntimeError(f"og_initialize returned {rc}")

    def prepare(self):
        """ETD operators only (no Richardson start): for timing steps."""
        rc = lib().og_solver_prepare(self.s)
        if rc:
            raise RuntimeError(f"og_solver_prepare returned {rc}")

    def step(self, psi, chi, b):
        return lib().og_time_step(self.s, psi.ptr, chi.ptr, b.ptr)

    def set_march_controls(self, rmv_sw=0, rmv_int=0, adj_sw=0, adj_int=0):
        lib().og_set_march_controls(self.s, rmv_sw, rmv_int, adj_sw, adj_int)

    @property
    def t(self):
        return lib().og_time(self.s)

    def close(self):
        if self.s:
            lib().og_solver_free(self.s)
            self.s = None
