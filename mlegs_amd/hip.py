"""Python interface to libmlegs_hip.so (ctypes over the C ABI in include/mlegs_hip.h).

The names mirror the reference's Fortran module procedures so that a caller
of MLegS finds the same operations with the same argument meaning:

* ``Grid``   -- LEGINIT / FINALIZE (src/mod_fft.f90:165-367), the CHOPSET
  truncation state (src/mod_scalar3.f90:174-195).
* ``Scalar`` -- TYPE(SCALAR) (src/mod_scalar3.f90:24-74): a device field with
  its space tag and log-term coefficient LN.
* transforms, operators and Boussinesq terms as free functions
  (src/mod_fft.f90, src/mod_legops.f90, src/mod_boussinesq.f90).
* ``Solver`` -- SOLVER_T INITIALIZE / TIME_STEPPING (src/mod_march.f90:171-307).

Errors: every call checks the returned status and raises ``MlegsError``
with the reference's error code (src/mod_scalar3.f90:88-106); the reference
aborts at the same points.  There is no CPU fallback: if the HIP library is
missing this module fails at import of the first call.
"""
import ctypes as C
import os

import numpy as np

PPP, FFF, PFP, PFF = 0, 1, 2, 3
SPACE_NAMES = {PPP: "PPP", FFF: "FFF", PFP: "PFP", PFF: "PFF"}

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libmlegs_hip.so")


class MlegsError(RuntimeError):
    """A non-zero status from the library (reference error codes, 100+ for runtime errors)."""

    def __init__(self, code, where):
        self.code = int(code)
        msg = _lib().mlg_error_string(self.code)
        name = msg.decode() if msg else "unknown"
        super().__init__(f"{where}: error {self.code} ({name})")


class Params(C.Structure):
    """mlg_params: read.input's DIMENSION/CHOPS/ELL/ZLEN/DT/VISCOSITY/BOUSSINESQ blocks."""

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

    @classmethod
    def from_dict(cls, cfg):
        p = cls()
        for k, _ in cls._fields_:
            if k in cfg:
                setattr(p, k, cfg[k])
        if p.minc == 0:
            p.minc = 1
        return p


NAME_LEN, MAX_OUTPUTS, MAX_MONITOR = 72, 512, 32


class RunSettings(C.Structure):
    """mlg_run_settings: read.input's file names, outputs and cadences."""
    _fields_ = [("savedir", C.c_char * NAME_LEN), ("psi0", C.c_char * NAME_LEN), ("chi0", C.c_char * NAME_LEN),
                ("b0", C.c_char * NAME_LEN), ("psii", C.c_char * NAME_LEN), ("chii", C.c_char * NAME_LEN),
                ("bi", C.c_char * NAME_LEN), ("it1", C.c_int), ("it2", C.c_int), ("rmv_sw", C.c_int),
                ("rmv_int", C.c_int), ("adj_sw", C.c_int), ("adj_int", C.c_int), ("noutputs", C.c_int),
                ("out_t", C.c_double * MAX_OUTPUTS), ("out_psi", (C.c_char * NAME_LEN) * MAX_OUTPUTS),
                ("out_chi", (C.c_char * NAME_LEN) * MAX_OUTPUTS), ("out_b", (C.c_char * NAME_LEN) * MAX_OUTPUTS),
                ("nmonitor", C.c_int), ("monitor_mk", (C.c_int * 2) * MAX_MONITOR), ("monitor_start", C.c_int),
                ("monitor_interval", C.c_int)]


_LIB = None
_V = C.c_void_p
_D = C.c_double
_I = C.c_int
_PD = C.POINTER(C.c_double)

# name -> (argtypes, restype); must list every symbol of include/mlegs_hip.h
SIGNATURES = {
    "mlg_init": ([C.POINTER(Params), C.POINTER(_V)], _I),
    "mlg_finalize": ([_V], _I),
    "mlg_dims": ([_V, C.POINTER(_I)], _I),
    "mlg_get_table": ([_V, C.c_char_p, _PD, C.c_long, C.POINTER(C.c_long)], _I),
    "mlg_set_stream": ([_V, _V], _I),
    "mlg_synchronize": ([_V], _I),
    "mlg_error_string": ([_I], C.c_char_p),
    "mlg_chopset": ([_V, _I], _I),
    "mlg_nrchop": ([_V], _I),
    "mlg_field_create": ([_V, _I, C.POINTER(_V)], _I),
    "mlg_field_destroy": ([_V], _I),
    "mlg_field_space": ([_V], _I),
    "mlg_field_upload": ([_V, _PD, _I, _D], _I),
    "mlg_field_download": ([_V, _PD, C.POINTER(_I), _PD], _I),
    "mlg_field_device_ptr": ([_V, C.POINTER(_V), C.POINTER(_V)], _I),
    "mlg_field_zero": ([_V, _I], _I),
    "mlg_copy": ([_V, _V], _I),
    "mlg_chopdo": ([_V], _I),
    "mlg_horfft": ([_V, _I], _I),
    "mlg_verfft": ([_V, _I], _I),
    "mlg_rtran": ([_V, _I], _I),
    "mlg_toff": ([_V], _I),
    "mlg_tofp": ([_V], _I),
    "mlg_plnfft": ([_V, _I], _I),
    "mlg_delsqh": ([_V, _V], _I),
    "mlg_idelsqh": ([_V, _V], _I),
    "mlg_mulxm": ([_V, _V], _I),
    "mlg_xxdx": ([_V, _V], _I),
    "mlg_del2": ([_V, _V], _I),
    "mlg_idel2": ([_V, _V, _I, _D], _I),
    "mlg_idel2ln": ([_V, _V], _I),
    "mlg_ihelm": ([_V, _V, _D], _I),
    "mlg_ihelmp": ([_V, _V, _I, _D, _D], _I),
    "mlg_grad": ([_V, _V, _V, _V], _I),
    "mlg_pc2vel": ([_V, _V, _V, _V, _V], _I),
    "mlg_pc2vor": ([_V, _V, _V, _V, _V], _I),
    "mlg_vprodsub": ([_V, _V, _V, _V, _V, _V], _I),
    "mlg_vdotpsub": ([_V, _V, _V, _V, _V, _V, _V], _I),
    "mlg_vel2pln": ([_V, _V, _I, _D], _I),
    "mlg_vel2pln_noalloc": ([_V, _V, _I, _D], _I),
    "mlg_project": ([_V, _V, _V, _V, _V, _I, _D], _I),
    "mlg_force": ([_V, _V, _V, _V, _V, _V], _I),
    "mlg_nonlin": ([_V, _V, _V, _V, _V, _V, _V, _V], _I),
    "mlg_energy": ([_V, _V, _V, _PD, _PD, _PD], _I),
    "mlg_etd_op": ([_V, _D, _PD, _PD], _I),
    "mlg_solver_create": ([_V, C.POINTER(_V)], _I),
    "mlg_solver_destroy": ([_V], _I),
    "mlg_initialize": ([_V, _V, _V, _V], _I),
    "mlg_time_step": ([_V, _V, _V, _V], _I),
    "mlg_time_steps": ([_V, _V, _V, _V, _I], _I),
    "mlg_time_steps_graph": ([_V, _V, _V, _V, _I], _I),
    "mlg_diagnost": ([_V, _V], _I),
    "mlg_solver_time": ([_V], _D),
    "mlg_solver_nstep": ([_V], _I),
    "mlg_qvortex": ([_V, _D, _V, _V], _I),
    "mlg_set_legendre_precision": ([_V, _I], _I),
    "mlg_legendre_precision": ([_V], _I),
    "mlg_set_split_fp64_modes": ([_V, _I], _I),
    "mlg_read_input": ([C.c_char_p, C.POINTER(Params), _PD], _I),
    "mlg_read_run_settings": ([C.c_char_p, C.POINTER(RunSettings)], _I),
    "mlg_field_save": ([_V, C.c_char_p], _I),
    "mlg_field_load": ([_V, C.c_char_p, _I, _I], _I),
    "mlg_field_m0_columns": ([_V, _I, _PD, C.POINTER(_I)], _I),
    "mlg_force_adjoint": ([_V] * 9, _I),
    "mlg_step_adjoint": ([_V] * 10, _I),
    "mlg_start_adjoint": ([_V] * 10, _I),
    "mlg_ke_gradient": ([_V, _V, _I, _V, _V], _I),
    "mlg_solver_history": ([_V, C.POINTER(_V), C.POINTER(_V), C.POINTER(_V)], _I),
    "mlg_adjoint_op": ([_I, C.POINTER(_V), _I, _D], _I),
    "mlg_diagnost_it1": ([_V, _V, _PD], _I),
    "mlg_set_march_controls": ([_V, _I, _I, _I, _I], _I),
    "mlg_remove": ([_V, _V, _V], _I),
    "mlg_ringcut": ([_V, _I, _I, _I], _I),
    "mlg_sponge": ([_V, _V, _V, _D], _I),
    "mlg_enemon": ([_V, _V, _PD, _PD], _I),
    "mlg_hyp3adj": ([_V, _V, _V, _V], _I),
    "mlg_hyperv_coeffs": ([_V, _PD], _I),
    "mlg_set_hyperv_coeffs": ([_V, _PD], _I),
    "mlg_sponge_table": ([_V, _PD, _I], _I),
    "mlg_sweep_create": ([_I, C.POINTER(_V), C.POINTER(_V), C.POINTER(_V), C.POINTER(_V), _I, C.POINTER(_V)], _I),
    "mlg_sweep_steps": ([_V, _I], _I),
    "mlg_sweep_info": ([_V, C.POINTER(_I)], _I),
    "mlg_sweep_codes": ([_V, C.POINTER(_I)], _I),
    "mlg_sweep_work": ([_V, _PD], _I),
    "mlg_sweep_destroy": ([_V], _I),
    "mlg_tune_rtran_bwd": ([_V, _I, C.POINTER(_I), _PD], _I),
    "mlg_timer_enable": ([_V, _I], _I),
    "mlg_timer_read": ([_V, C.c_char_p, _PD, C.POINTER(C.c_long), _PD, C.POINTER(_I)], _I),
}

TIMED_KERNELS = ("rtran_bwd", "rtran_fwd", "project", "verfft_fwd", "verfft_bwd", "horfft_fwd", "horfft_bwd",
                 "products", "band_solve", "etd")


def _lib():
    """Load the HIP library (no fallback: a missing library is an error)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmlegs_hip.so not built ({LIB_PATH}); run `make -C mlegs_amd`")
        # PyTorch ships its own HIP runtime, and a process keeps one copy per
        # soname: with torch's loaded first the library runs on it, and
        # torch's device work (tensors over the library's memory, its streams)
        # stays usable after the library is loaded.  Loaded the other way
        # round, torch finds no HIP device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def library_path():
    return LIB_PATH


def read_input(path):
    """read.input -> (parameter dict for Grid, end time TIM%LIMIT) via the library's parser."""
    p = Params()
    lim = C.c_double(0.0)
    _call("mlg_read_input", os.fsencode(path), C.byref(p), C.byref(lim))
    return {k: getattr(p, k) for k, _ in Params._fields_}, lim.value


def read_run_settings(path):
    """read.input's run settings (names, OUTPUTS, DIAGNOST, REMOVE, NUP_ADJUST, MONITOR_MK) as a dict."""
    st = RunSettings()
    _call("mlg_read_run_settings", os.fsencode(path), C.byref(st))

    def name(b):
        return bytes(b).decode(errors="replace").rstrip()

    out = {k: name(getattr(st, k)) for k in ("savedir", "psi0", "chi0", "b0", "psii", "chii", "bi")}
    out.update({k: getattr(st, k) for k in ("it1", "it2", "rmv_sw", "rmv_int", "adj_sw", "adj_int",
                                            "monitor_start", "monitor_interval")})
    out["outputs"] = [(st.out_t[i], name(st.out_psi[i]), name(st.out_chi[i]), name(st.out_b[i]))
                      for i in range(st.noutputs)]
    out["monitor_mk"] = [(st.monitor_mk[i][0], st.monitor_mk[i][1]) for i in range(st.nmonitor)]
    return out


def _check(rc, where):
    if rc:
        raise MlegsError(rc, where)


def _call(name, *args):
    _check(getattr(_lib(), name)(*args), name)


# ---------------------------------------------------------------------------
class Grid:
    """Grid, tables and truncation state on the current HIP device (LEGINIT)."""

    def __init__(self, cfg):
        self.cfg = dict(cfg)
        self.params = Params.from_dict(cfg)
        h = _V()
        _call("mlg_init", C.byref(self.params), C.byref(h))
        self.h = h
        d = (_I * 12)()
        _call("mlg_dims", self.h, d)
        (self.ndimr, self.ndimth, self.ndimx, self.nrchopdim, self.ntchopdim, self.nxchopdim,
         self.nrh, self.nxchoph, _, self.npf, self.nnorm, self.nr) = list(d)
        if cfg.get("split"):
            self.legendre_precision = "split32"

    def close(self):
        if getattr(self, "h", None):
            _lib().mlg_finalize(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, space):
        if space == FFF:
            return (self.nrchopdim, self.ntchopdim, self.nxchopdim)
        if space == PFF:
            return (self.ndimr, self.ntchopdim, self.nxchopdim)
        return (self.ndimr, self.ndimth, self.ndimx)

    def size(self, space):
        return int(np.prod(self.shape(space)))

    def table(self, name):
        n = C.c_long(0)
        _call("mlg_get_table", self.h, name.encode(), None, 0, C.byref(n))
        out = np.zeros(n.value, dtype=np.float64)
        _call("mlg_get_table", self.h, name.encode(), out.ctypes.data_as(_PD), n.value, C.byref(n))
        return out

    @property
    def nrchop(self):
        return _lib().mlg_nrchop(self.h)

    def chopset(self, iof):
        """CHOPSET(IOF): NRCHOP += IOF (src/mod_scalar3.f90:174-195)."""
        _call("mlg_chopset", self.h, int(iof))

    def synchronize(self):
        """Wait for the stream; raises the first sticky device-side error."""
        _call("mlg_synchronize", self.h)

    def set_stream(self, stream_handle):
        """Run on an external HIP stream (an int handle, e.g. torch.cuda.Stream.cuda_stream); None = own stream."""
        _call("mlg_set_stream", self.h, _V(stream_handle) if stream_handle else None)

    def etd_op(self, dt):
        """CALC_BOUSSI_ETD_OP: real parts of S e^{J dt} S^-1 and S J^-1 (e^{J dt} - I) S^-1."""
        e, nl = (_D * 16)(), (_D * 16)()
        _call("mlg_etd_op", self.h, float(dt), e, nl)
        return (np.array(e).reshape(4, 4, order="F"), np.array(nl).reshape(4, 4, order="F"))

    @property
    def legendre_precision(self):
        """'fp64' or 'split32': the arithmetic of the Legendre products (RTRAN, PROJECT)."""
        return "split32" if _lib().mlg_legendre_precision(self.h) == 1 else "fp64"

    @legendre_precision.setter
    def legendre_precision(self, mode):
        if mode not in ("fp64", "split32"):
            raise ValueError("legendre_precision is 'fp64' or 'split32'")
        _call("mlg_set_legendre_precision", self.h, 1 if mode == "split32" else 0)

    def set_split_fp64_modes(self, nmodes):
        """Split precision: the modes m < nmodes (>= 2; default 5) keep FP64 products."""
        _call("mlg_set_split_fp64_modes", self.h, int(nmodes))

    @property
    def hyperv_coeffs(self):
        """(NUP_R, NUP_T, NUP_X, KAPPAP_R, KAPPAP_T, KAPPAP_X) of the filters."""
        o = (_D * 6)()
        _call("mlg_hyperv_coeffs", self.h, o)
        return np.array(o)

    @hyperv_coeffs.setter
    def hyperv_coeffs(self, c):
        _call("mlg_set_hyperv_coeffs", self.h, (_D * 6)(*[float(v) for v in c]))

    def sponge_table(self):
        """TOPHAT_X of the axial sponge over the z points (zeros when off)."""
        o = (_D * self.ndimx)()
        _call("mlg_sponge_table", self.h, o, self.ndimx)
        return np.array(o)

    def timer(self, on):
        """Start (clearing earlier records) or stop the per-kernel HIP event timer."""
        _call("mlg_timer_enable", self.h, 1 if on else 0)

    def timer_read(self, kernel):
        """(ms_total, launches, work_total, unit) for one timed kernel class."""
        ms, n, w, ub = C.c_double(0), C.c_long(0), C.c_double(0), C.c_int(0)
        _call("mlg_timer_read", self.h, kernel.encode(), C.byref(ms), C.byref(n), C.byref(w), C.byref(ub))
        return ms.value, n.value, w.value, ("byte" if ub.value else "flop")

    def scalar(self, space=FFF, data=None, ln=0.0):
        s = Scalar(self, space)
        if data is not None:
            s.upload(data, space, ln)
        return s


class Scalar:
    """TYPE(SCALAR): device coefficients E, space tag and log coefficient LN."""

    def __init__(self, grid, space=FFF):
        self.grid = grid
        h = _V()
        _call("mlg_field_create", grid.h, int(space), C.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.grid.h:
            _lib().mlg_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def space(self):
        return _lib().mlg_field_space(self.h)

    def upload(self, arr, space, ln=0.0):
        shape = self.grid.shape(space)
        a = np.asarray(arr, dtype=np.complex128)
        if a.shape != shape:
            raise ValueError(f"array shape {a.shape} does not match {SPACE_NAMES[space]} shape {shape}")
        a = np.asfortranarray(a)
        _call("mlg_field_upload", self.h, a.ctypes.data_as(_PD), int(space), float(ln))
        return self

    def download(self):
        """Returns (E as a Fortran-ordered complex array, space, LN)."""
        sp = C.c_int(0)
        _check(_lib().mlg_field_download(self.h, None, C.byref(sp), None), "mlg_field_download")
        out = np.zeros(self.grid.shape(sp.value), dtype=np.complex128, order="F")
        ln = C.c_double(0.0)
        _call("mlg_field_download", self.h, out.ctypes.data_as(_PD), C.byref(sp), C.byref(ln))
        return out, sp.value, ln.value

    @property
    def e(self):
        return self.download()[0]

    @property
    def ln(self):
        return self.download()[2]

    def save(self, fn):
        """MSAVE / MPISAVEX (src/mod_fft.f90:2304-2377): FN.info and FN, written by the library."""
        _call("mlg_field_save", self.h, os.fsencode(fn))

    def load(self, fn, procs=(1, 1)):
        """MLOAD / MPILOADX (src/mod_fft.f90:2510-2642) of a file written on procs=(N1, N2) processes."""
        _call("mlg_field_load", self.h, os.fsencode(fn), int(procs[0]), int(procs[1]))
        return self

    def m0_columns(self, nk):
        """E(:, m=0, k) for k < nk in the current space (what INSPECT prints)."""
        nrad = C.c_int(0)
        _call("mlg_field_m0_columns", self.h, int(nk), None, C.byref(nrad))
        out = np.zeros((nrad.value, nk), dtype=np.complex128, order="F")
        _call("mlg_field_m0_columns", self.h, int(nk), out.ctypes.data_as(_PD), C.byref(nrad))
        return out

    def device_ptrs(self):
        e, ln = _V(), _V()
        _call("mlg_field_device_ptr", self.h, C.byref(e), C.byref(ln))
        return e.value, ln.value

    def zero(self, space=FFF):
        _call("mlg_field_zero", self.h, int(space))
        return self

    def as_tensor(self):
        """A torch view (complex128, device memory of this field, current space shape)."""
        import torch

        ptr, _ = self.device_ptrs()
        shape = self.grid.shape(self.space)
        n = int(np.prod(shape))

        class _Iface:
            __cuda_array_interface__ = {"shape": (n,), "typestr": "<c16", "data": (ptr, False), "version": 2,
                                        "strides": None}

        t = torch.as_tensor(_Iface(), device="cuda")
        return t.view(shape[::-1]).permute(2, 1, 0)


class _Borrowed(Scalar):
    """A field owned by the library (a solver's history field): not destroyed here."""

    def __init__(self, grid, handle):
        self.grid = grid
        self.h = handle

    def close(self):
        self.h = None


# ---------------------------------------------------------------------------
# transforms (src/mod_fft.f90)
def chopdo(a):
    _call("mlg_chopdo", a.h)


def copy(b, a):
    """B = A (with LN; FFF copies chopped)."""
    _call("mlg_copy", b.h, a.h)


def horfft(a, is_):
    _call("mlg_horfft", a.h, int(is_))


def verfft(a, is_):
    _call("mlg_verfft", a.h, int(is_))


def rtran(a, is_):
    _call("mlg_rtran", a.h, int(is_))


def plnfft(a, is_):
    _call("mlg_plnfft", a.h, int(is_))


def toff(a):
    _call("mlg_toff", a.h)


def tofp(a):
    _call("mlg_tofp", a.h)


# operators (src/mod_legops.f90)
def delsqh(a, b):
    _call("mlg_delsqh", a.h, b.h)


def idelsqh(b, a):
    _call("mlg_idelsqh", b.h, a.h)


def mulxm(a, b):
    _call("mlg_mulxm", a.h, b.h)


def xxdx(a, b):
    _call("mlg_xxdx", a.h, b.h)


def del2(a, b):
    _call("mlg_del2", a.h, b.h)


def idel2(b, a, ln=None):
    _call("mlg_idel2", b.h, a.h, 0 if ln is None else 1, 0.0 if ln is None else float(ln))


def idel2ln(b, a):
    _call("mlg_idel2ln", b.h, a.h)


def ihelm(b, a, alp):
    _call("mlg_ihelm", b.h, a.h, float(alp))


def ihelmp(p, b, a, alp, nu=0.0):
    """IHELMP(P, B, A, ALP, NUIN): A = (del^P + nu del^2 - alp)^-1 B (src/mod_legops.f90:1409-1592)."""
    _call("mlg_ihelmp", b.h, a.h, int(p), float(alp), float(nu))


def grad(a, rr, rp, z):
    _call("mlg_grad", a.h, rr.h, rp.h, z.h)


def pc2vel(psi, chi, rur, rup, uz):
    _call("mlg_pc2vel", psi.h, chi.h, rur.h, rup.h, uz.h)


def pc2vor(psi, chi, ror, rop, oz):
    _call("mlg_pc2vor", psi.h, chi.h, ror.h, rop.h, oz.h)


def vprodsub(rur, rup, uz, ror, rop, oz):
    _call("mlg_vprodsub", rur.h, rup.h, uz.h, ror.h, rop.h, oz.h)


def vdotpsub(rar, rap, az, rbr, rbp, bz, out):
    _call("mlg_vdotpsub", rar.h, rap.h, az.h, rbr.h, rbp.h, bz.h, out.h)


def vel2pln(rup, psi, ln=None):
    _call("mlg_vel2pln", rup.h, psi.h, 0 if ln is None else 1, 0.0 if ln is None else float(ln))


def vel2pln_noalloc(rup, psi, ln=None):
    """VEL2PLN with its scratch in a fixed per-GPU array (no device allocation; NRCHOP <= 1024)."""
    _call("mlg_vel2pln_noalloc", rup.h, psi.h, 0 if ln is None else 1, 0.0 if ln is None else float(ln))


def project(rur, rup, uz, psi, chi, ln=None):
    _call("mlg_project", rur.h, rup.h, uz.h, psi.h, chi.h, 0 if ln is None else 1, 0.0 if ln is None else float(ln))


# Boussinesq terms (src/mod_boussinesq.f90)
def calc_boussi_force(psi, chi, b, psin, chin, bn):
    _call("mlg_force", psi.h, chi.h, b.h, psin.h, chin.h, bn.h)


def calc_boussi_nonlin(b, rur, rup, uz, bn, ror, rop, oz):
    _call("mlg_nonlin", b.h, rur.h, rup.h, uz.h, bn.h, ror.h, rop.h, oz.h)


def calc_boussi_energy(psi, chi, b):
    """Returns (KE, PE, dKE) as (NTCHOP, NXCHOP) arrays over m >= 0, |k|."""
    cfg = psi.grid.cfg
    nt, nxc = cfg["ntchop"], cfg["nxchop"]
    ke, pe, dke = (_D * (nt * nxc))(), (_D * (nt * nxc))(), (_D * (nt * nxc))()
    _call("mlg_energy", psi.h, chi.h, b.h, ke, pe, dke)
    f = lambda a: np.array(a).reshape(nt, nxc, order="F")
    return f(ke), f(pe), f(dke)


DIAGNOST_IT1_KEYS = ("valvor", "valzmo", "valene", "valhel", "valang", "psiat1", "psiat1_spread", "chiat1",
                     "chiat1_spread", "x", "y", "r")


def diagnost_it1(psi, chi):
    """DIAGNOST's every-IT1 quantities (vorticity, z-momentum, energy,
    helicity, angular momentum, psi/chi at r = infinity, the vortex centre
    <X>, <Y> and radius <R>) as a dict."""
    o = (_D * 12)()
    _call("mlg_diagnost_it1", psi.h, chi.h, o)
    return dict(zip(DIAGNOST_IT1_KEYS, list(o)))


def remove(psi, chi, b=None):
    """REMOVE (mod_march.f90:2503-2551): far-field vorticity removal of psi,
    chi (and b), in place."""
    _call("mlg_remove", psi.h, chi.h, b.h if b is not None else None)


def ringcut(wk, ns0, ns, divxm2):
    """RINGCUT on a PFF field (1-based ns0, ns as in the reference)."""
    _call("mlg_ringcut", wk.h, int(ns0), int(ns), 1 if divxm2 else 0)


def sponge(psi, chi, b, dt):
    """SPONGE (mod_march.f90:2380-2420) over dt; b may be None."""
    _call("mlg_sponge", psi.h, chi.h, b.h if b is not None else None, float(dt))


def enemon(psi, chi):
    """ENEMON at the current truncation: (EM[NTCHOP], EK[NXCHOP])."""
    cfg = psi.grid.cfg
    em, ek = (_D * cfg["ntchop"])(), (_D * cfg["nxchop"])()
    _call("mlg_enemon", psi.h, chi.h, em, ek)
    return np.array(em), np.array(ek)


def hyp3adj(grid, psi, chi, b=None):
    """HYP3ADJ then HYP3SET: returns the new (NUP_R, NUP_T, NUP_X, KAPPAP_R,
    KAPPAP_T, KAPPAP_X)."""
    _call("mlg_hyp3adj", grid.h, psi.h, chi.h, b.h if b is not None else None)
    return grid.hyperv_coeffs


def force_adjoint(psi, chi, b, nbar, ubar):
    """ubar (3 fields) += CALC_BOUSSI_FORCE'(psi, chi, b)^T nbar (3 fields)."""
    _call("mlg_force_adjoint", psi.h, chi.h, b.h, *[f.h for f in nbar], *[f.h for f in ubar])


ADJ_TOFP, ADJ_TOFF, ADJ_PROJECT, ADJ_IDEL2LN, ADJ_IHELM, ADJ_DEL2, ADJ_VELVORGRAD, ADJ_PRODUCTS = range(8)


def adjoint_op(op, fields, alp=0.0):
    """Transpose of one operator (mlg_adjoint_op) on the given fields."""
    arr = (_V * len(fields))(*[f.h for f in fields])
    _call("mlg_adjoint_op", int(op), arr, len(fields), float(alp))


def ke_gradient(psi, chi, psi_bar, chi_bar, mmin=1):
    """d/d(psi, chi) of the sum over m >= mmin of CALC_BOUSSI_ENERGY's KE."""
    _call("mlg_ke_gradient", psi.h, chi.h, int(mmin), psi_bar.h, chi_bar.h)


def diagnost(psi, chi):
    """DIAGNOST's per-step checks (src/mod_march.f90:1726-1791): raises
    MlegsError(11) (dealiasing) or MlegsError(12) (divergence)."""
    _call("mlg_diagnost", psi.h, chi.h)


def qvortex(grid, q, psi, chi):
    """QVORTEX (src/mod_layout.f90:449-610): Gaussian vortex of strength q in psi, chi = 0."""
    _call("mlg_qvortex", grid.h, float(q), psi.h, chi.h)


class Solver:
    """SOLVER_T: AB2-CN or ETD2 time march with a Richardson first step."""

    def __init__(self, grid):
        self.grid = grid
        h = _V()
        _call("mlg_solver_create", grid.h, C.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.grid.h:
            _lib().mlg_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self, psi, chi, b):
        _call("mlg_initialize", self.h, psi.h, chi.h, b.h)

    def time_step(self, psi, chi, b):
        _call("mlg_time_step", self.h, psi.h, chi.h, b.h)

    def time_stepping(self, psi, chi, b, nsteps, graph=False):
        """NSTEPS steps, one synchronisation at the end (device checks are sticky).
        graph=True replays a captured step (same results, fewer launches)."""
        _call("mlg_time_steps_graph" if graph else "mlg_time_steps", self.h, psi.h, chi.h, b.h, int(nsteps))

    def set_march_controls(self, rmv_sw=0, rmv_int=0, adj_sw=0, adj_int=0):
        """TIME_STEPPING's REMOVE (RMV%SW, RMV%INT) and HYP3ADJ (VISC%ADJSW,
        VISC%ADJINT) cadence; rmv_sw=2 also turns on the ETD axial sponge
        (call before initialize)."""
        _call("mlg_set_march_controls", self.h, int(rmv_sw), int(rmv_int), int(adj_sw), int(adj_int))

    def history(self):
        """The solver's own AB2 history fields (N_psi, N_chi, N_b), borrowed."""
        a, b, c = _V(), _V(), _V()
        _call("mlg_solver_history", self.h, C.byref(a), C.byref(b), C.byref(c))
        return _Borrowed(self.grid, a), _Borrowed(self.grid, b), _Borrowed(self.grid, c)

    def step_adjoint(self, psi, chi, b, ub, hb):
        """Transpose of one AB2-CN step from the state (psi, chi, b): ub (3
        fields) in: adjoint of the state after the step, out: before it; hb
        the same for the history."""
        _call("mlg_step_adjoint", self.h, psi.h, chi.h, b.h, *[f.h for f in ub], *[f.h for f in hb])

    def start_adjoint(self, psi, chi, b, ub, hb):
        """Transpose of the Richardson start (initialize) from the initial state."""
        _call("mlg_start_adjoint", self.h, psi.h, chi.h, b.h, *[f.h for f in ub], *[f.h for f in hb])

    @property
    def t(self):
        return _lib().mlg_solver_time(self.h)

    @property
    def nstep(self):
        return _lib().mlg_solver_nstep(self.h)


# ---------------------------------------------------------------------------
SWEEP_GROUPED = 1
SWEEP_TUNE = 2
SWEEP_STREAMS = 4
SWEEP_BY_GRID = 8
BWD_TILE_SHAPES = ("64x64", "32x64", "64x32")   # radial nodes x real columns


class Sweep:
    """Independent runs of different resolution stepped together: one graph
    launch per step for all runs (mlg_sweep_*).  ``runs`` is a sequence of
    (solver, psi, chi, b), each solver on its own Grid and already started.
    grouped: the RTRAN backward products of all runs at each place of the
    step run as one grouped persistent launch; tune: tile shapes from the
    tuner; shape: force one tile shape (index into BWD_TILE_SHAPES);
    streams: each run's segments launched on the run's own stream instead of
    one graph launch per step; by_grid: one grouped launch per class of runs
    of one grid."""

    def __init__(self, runs, grouped=True, tune=True, shape=None, streams=False, by_grid=False):
        runs = list(runs)
        n = len(runs)
        self.runs = runs
        arr = [(_V * n)(*[r[i].h for r in runs]) for i in range(4)]
        flags = (SWEEP_GROUPED if grouped else 0) | (SWEEP_TUNE if tune and shape is None else 0)
        if shape is not None:
            flags |= (int(shape) + 1) << 4
        if streams:
            flags |= SWEEP_STREAMS
        if by_grid:
            flags |= SWEEP_BY_GRID
        h = _V()
        _call("mlg_sweep_create", n, arr[0], arr[1], arr[2], arr[3], flags, C.byref(h))
        self.h = h

    def steps(self, nsteps, check=True):
        """Advance every run by nsteps.  A run's DIAGNOST failure raises
        MlegsError (check=True) after all runs have stepped; with
        check=False the runs' codes are returned (codes())."""
        rc = _lib().mlg_sweep_steps(self.h, int(nsteps))
        if rc >= 100 or (rc and check):
            raise MlegsError(rc, "mlg_sweep_steps")
        return self.codes()

    def codes(self):
        """Each run's first DIAGNOST code since the sweep was created (0: none)."""
        o = (_I * len(self.runs))()
        _call("mlg_sweep_codes", self.h, o)
        return list(o)

    def info(self):
        o = (_I * 8)()
        _call("mlg_sweep_info", self.h, o)
        o = list(o)
        gf = C.c_double(0.0)
        _call("mlg_sweep_work", self.h, C.byref(gf))
        return dict(grouped_positions=o[0], graph_nodes=o[1], grouped_jobs=o[2],
                    jobs_per_shape=dict(zip(BWD_TILE_SHAPES, o[3:6])), grouped_launches=o[6], runs=o[7],
                    grouped_gflop_per_step=gf.value)

    def close(self):
        if getattr(self, "h", None):
            _lib().mlg_sweep_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_rtran_bwd(grid, nfields):
    """Tile tuner of the backward products at the grid's current truncation:
    {bucket: (chosen shape name, {shape name: microseconds})} per mode bucket
    (0: NRCHOPS(m) > 128, 1: 65..128, 2: 33..64, 3: <= 32) that has modes."""
    shape = (_I * 4)()
    us = (_D * 16)()
    _call("mlg_tune_rtran_bwd", grid.h, int(nfields), shape, us)
    out = {}
    for b in range(4):
        if shape[b] >= 0:
            out[b] = (BWD_TILE_SHAPES[shape[b]], {BWD_TILE_SHAPES[s]: us[4 * b + s] for s in range(3)})
    return out
