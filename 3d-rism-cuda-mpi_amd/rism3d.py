"""Python mirror of the reference's RISM3D interface over librism3d.so (ctypes).

The reference drives one solve as ``RISM3D::initialize`` -> ``iterate`` ->
``output`` (main.cu:63-65); :class:`RISM3D` offers the same three calls with
the same meaning, plus the in-process hot-path calls that
``RISM3D::iterate`` makes (``calculate``, ``cal_rms``, ``AN2::calculate``,
iterate.cu:48-72).  Everything runs in ``librism3d.so`` (HIP kernels for
gfx950 behind the C ABI of ``include/rism3d_c.h``); this module only
marshals arguments.  There is no CPU fallback: if the library is missing,
:func:`lib` raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librism3d.so")

RISM_OK = 0
RISM_ERR_ARG, RISM_ERR_IO, RISM_ERR_VERSION, RISM_ERR_CLOSURE = 1, 2, 3, 4
RISM_ERR_NATOM, RISM_ERR_KRANGE, RISM_ERR_HIP, RISM_ERR_NOMEM, RISM_ERR_STATE = 5, 6, 7, 8, 9
STATUS_CONVERGED, STATUS_DIVERGED, STATUS_MAXSTEP = 0, 1, 2
KH, HNC = 0, 1
ANALYSIS_GRAD, ANALYSIS_EUV, ANALYSIS_AD = 0, 1, 2


def PSE(n):
    """Closure code of PSE-n (Kast-Kloss partial series, 1 <= n <= 20)."""
    return 10 + int(n)

FIELD_T, FIELD_G, FIELD_H, FIELD_R, FIELD_U, FIELD_FR, FIELD_C = range(7)

# every symbol include/rism3d_c.h declares
EXPORTS = [
    "rism_version", "rism_last_error", "rism_create", "rism_destroy",
    "rism_setup_potentials", "rism_calculate", "rism_anderson", "rism_solve",
    "rism_thermo", "rism_get_field", "rism_set_field", "rism_get_timing", "rism_run",
    "rism_read_input", "rism_free_input", "rism_read_xvk", "rism_write_xmu",
    "rism_write_guv", "rism_write_tuv", "rism_read_tuv", "rism_k_fft_r2c",
    "rism_k_fft_c2r", "rism_k_get_fk", "rism_k_an2_state", "rism_set_kernel_timing",
    "rism_get_kernel_stats", "rism_set_precision", "rism_get_precision",
    "rism_solve_batch", "rism_run_batch", "rism_write_huv", "rism_write_cuv",
    "rism_analysis", "rism_k_copy_bench", "rism_tune_report", "rism_k_fft_noalloc",
]
KERNELS = ["zfwd_closure", "ypass_fwd", "xpass_oz", "ypass_inv", "zinv_resid", "reduce",
           "an2_update", "potential_r", "fk_half", "thermo_partial"]


class RismError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rism error {code}: {msg}")
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("box", ctypes.c_double * 3)]


class SolventC(ctypes.Structure):
    _fields_ = [("natv", ctypes.c_int), ("ntab", ctypes.c_int), ("dr", ctypes.c_double),
                ("temper", ctypes.c_double), ("xt", ctypes.c_double), ("sig", _dp),
                ("eps", _dp), ("q", _dp), ("rho", _dp), ("xvv", _dp)]


class SoluteC(ctypes.Structure):
    _fields_ = [("natu", ctypes.c_int), ("sig", _dp), ("eps", _dp), ("q", _dp), ("xyz", _dp)]


class Opts(ctypes.Structure):
    _fields_ = [("closure", ctypes.c_int), ("an2_count", ctypes.c_int),
                ("an2_m", ctypes.c_double), ("an2_mp", ctypes.c_double),
                ("ad_switch", ctypes.c_int), ("lambda_", ctypes.c_double),
                ("verbose", ctypes.c_int), ("chi_spline", ctypes.c_int),
                ("anderson_m", ctypes.c_int), ("precision", ctypes.c_int),
                ("mixed_switch", ctypes.c_double), ("mixed_block", ctypes.c_int)]


class Thermo(ctypes.Structure):
    _fields_ = [("natv", ctypes.c_int), ("sfe_sc", ctypes.c_double),
                ("sfe_sc_hnc", ctypes.c_double), ("sfe_gf", ctypes.c_double),
                ("pmv", ctypes.c_double), ("pressure", ctypes.c_double),
                ("correction", ctypes.c_double), ("sfec_sc", ctypes.c_double * 16),
                ("sfec_sc_hnc", ctypes.c_double * 16), ("sfec_gf", ctypes.c_double * 16)]

    def as_dict(self, closure=KH):
        n = self.natv
        return {
            "SFE_SC": self.sfe_sc, "SFEC_SC": np.array(self.sfec_sc[:n]),
            "SFE_SC_HNC": self.sfe_sc_hnc if closure == KH else None,
            "SFEC_SC_HNC": np.array(self.sfec_sc_hnc[:n]) if closure == KH else None,
            "SFE_GF": self.sfe_gf, "SFEC_GF": np.array(self.sfec_gf[:n]),
            "PMV": self.pmv, "Pressure": self.pressure, "Correction_Term": self.correction,
        }


class Timing(ctypes.Structure):
    _fields_ = [("setup_ms", ctypes.c_double), ("iterate_ms", ctypes.c_double),
                ("thermo_ms", ctypes.c_double), ("steps", ctypes.c_int),
                ("status", ctypes.c_int), ("rms", ctypes.c_double),
                ("steps_fp32", ctypes.c_int)]


class KStat(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_long), ("total_ms", ctypes.c_double)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("pass_", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int), ("natv", ctypes.c_int), ("fp32", ctypes.c_int),
                ("choice", ctypes.c_int), ("ms", ctypes.c_float * 3)]


class InputC(ctypes.Structure):
    _fields_ = [("outlist", ctypes.c_char * 64), ("ksave", ctypes.c_int),
                ("closure", ctypes.c_int), ("fsolvent", ctypes.c_char * 1024),
                ("convergence", ctypes.c_double), ("maxstep", ctypes.c_int),
                ("an2_count", ctypes.c_int), ("an2_m", ctypes.c_double),
                ("an2_mp", ctypes.c_double), ("box", ctypes.c_double * 3),
                ("grid", ctypes.c_int * 3), ("natu", ctypes.c_int), ("usig", _dp),
                ("ueps", _dp), ("uq", _dp), ("uxyz", _dp), ("shift", ctypes.c_double * 3),
                ("natv", ctypes.c_int), ("ntab", ctypes.c_int), ("dr", ctypes.c_double),
                ("temper", ctypes.c_double), ("xt", ctypes.c_double), ("vsig", _dp),
                ("veps", _dp), ("vq", _dp), ("vrho", _dp), ("xvv", _dp)]


_lib = None


def lib():
    """Loads librism3d.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C 3d-rism-cuda-mpi_amd); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    pi = ctypes.POINTER(ctypes.c_int)
    L.rism_version.restype = ctypes.c_char_p
    L.rism_last_error.restype = ctypes.c_char_p
    L.rism_create.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(SolventC),
                              ctypes.POINTER(SoluteC), ctypes.POINTER(Opts), pi]
    L.rism_create.restype = vp
    L.rism_destroy.argtypes = [vp]
    L.rism_destroy.restype = None
    L.rism_setup_potentials.argtypes = [vp]
    L.rism_calculate.argtypes = [vp, _dp]
    L.rism_anderson.argtypes = [vp]
    L.rism_set_precision.argtypes = [vp, ci]
    L.rism_solve_batch.argtypes = [ctypes.POINTER(vp), ci, ci, cd, ci, pi, pi, _dp]
    L.rism_run_batch.argtypes = [ctypes.POINTER(vp), ci, ci, cd, ci, ctypes.POINTER(Thermo),
                                 pi, pi]
    L.rism_get_precision.argtypes = [vp]
    L.rism_solve.argtypes = [vp, ci, cd, pi, pi, _dp]
    L.rism_thermo.argtypes = [vp, ctypes.POINTER(Thermo)]
    L.rism_get_field.argtypes = [vp, ci, ci, _dp]
    L.rism_set_field.argtypes = [vp, ci, ci, _dp]
    L.rism_get_timing.argtypes = [vp, ctypes.POINTER(Timing)]
    L.rism_run.argtypes = [vp, ci, cd, ctypes.POINTER(Thermo), ctypes.POINTER(Timing)]
    L.rism_read_input.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ci, ci, ctypes.c_char_p,
                                  ctypes.POINTER(ctypes.POINTER(InputC))]
    L.rism_free_input.argtypes = [ctypes.POINTER(InputC)]
    L.rism_free_input.restype = None
    L.rism_read_xvk.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(InputC))]
    L.rism_write_xmu.argtypes = [ctypes.c_char_p, ci, ctypes.POINTER(Thermo)]
    L.rism_write_guv.argtypes = [vp, ctypes.c_char_p, _dp]
    L.rism_write_tuv.argtypes = [vp, ctypes.c_char_p]
    L.rism_write_huv.argtypes = [vp, ctypes.c_char_p]
    L.rism_analysis.argtypes = [vp, ci, _dp]
    L.rism_write_cuv.argtypes = [vp, ctypes.c_char_p]
    L.rism_read_tuv.argtypes = [vp, ctypes.c_char_p]
    L.rism_k_fft_r2c.argtypes = [ci, ci, ci, _dp, _dp, _dp]
    L.rism_k_fft_c2r.argtypes = [ci, ci, ci, _dp, _dp, _dp]
    L.rism_k_get_fk.argtypes = [vp, _dp]
    L.rism_k_an2_state.argtypes = [vp, _dp, _dp, pi, pi]
    L.rism_set_kernel_timing.argtypes = [vp, ci]
    L.rism_get_kernel_stats.argtypes = [vp, ctypes.POINTER(KStat)]
    L.rism_k_copy_bench.argtypes = [ctypes.c_size_t, ci, _dp]
    L.rism_tune_report.argtypes = [ctypes.c_void_p, ci]
    L.rism_tune_report.restype = ci
    L.rism_k_fft_noalloc.argtypes = [ci, ci, ci, ci, _dp, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]
    _lib = L
    return L


def _check(rc):
    if rc != RISM_OK:
        raise RismError(rc, lib().rism_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------
# input files
# ---------------------------------------------------------------------------

def _arr(p, n):
    return np.ctypeslib.as_array(p, shape=(n,)).copy() if n > 0 else np.zeros(0)


def read_input(control, structure=None, center=True, q0=False, solvent_dir=None):
    """read_input.cu:21-97 + solvent_read.cc:10-94 through the library's parser."""
    out = ctypes.POINTER(InputC)()
    _check(lib().rism_read_input(str(control).encode(),
                                 str(structure).encode() if structure else None,
                                 int(bool(center)), int(bool(q0)),
                                 str(solvent_dir).encode() if solvent_dir else None,
                                 ctypes.byref(out)))
    try:
        c = out.contents
        n, nv, nt = c.natu, c.natv, c.ntab
        d = {
            "outlist": c.outlist.decode(), "ksave": c.ksave, "closure": c.closure,
            "fsolvent": c.fsolvent.decode(), "convergence": c.convergence,
            "maxstep": c.maxstep, "count": c.an2_count, "m": c.an2_m, "mp": c.an2_mp,
            "box": np.array(c.box[:]), "grid": np.array(c.grid[:]),
            "usig": _arr(c.usig, n), "ueps": _arr(c.ueps, n), "uq": _arr(c.uq, n),
            "uxyz": _arr(c.uxyz, 3 * n).reshape(n, 3), "shift": np.array(c.shift[:]),
            "natv": nv, "ntab": nt, "dr": c.dr, "temper": c.temper, "xt": c.xt,
            "vsig": _arr(c.vsig, nv), "veps": _arr(c.veps, nv), "vq": _arr(c.vq, nv),
            "vrho": _arr(c.vrho, nv), "xvv": _arr(c.xvv, nt * nv * nv).reshape(nt, nv, nv),
        }
    finally:
        lib().rism_free_input(out)
    return d


def read_xvk(path):
    out = ctypes.POINTER(InputC)()
    _check(lib().rism_read_xvk(str(path).encode(), ctypes.byref(out)))
    try:
        c = out.contents
        nv, nt = c.natv, c.ntab
        d = {"natv": nv, "ntab": nt, "dr": c.dr, "temper": c.temper, "xt": c.xt,
             "vsig": _arr(c.vsig, nv), "veps": _arr(c.veps, nv), "vq": _arr(c.vq, nv),
             "vrho": _arr(c.vrho, nv), "xvv": _arr(c.xvv, nt * nv * nv).reshape(nt, nv, nv)}
    finally:
        lib().rism_free_input(out)
    return d


def write_xmu(path, closure, thermo: Thermo):
    _check(lib().rism_write_xmu(str(path).encode(), int(closure), ctypes.byref(thermo)))


# ---------------------------------------------------------------------------
# kernel-level transforms (FFT3D::execute)
# ---------------------------------------------------------------------------

def fft_r2c(f, box):
    """Forward half-shifted transform of a real [nz][ny][nx] field on the GPU;
    returns the k_z > 0 half spectrum [nz/2][ny][nx] complex (reference
    convention; entry (m, jy, jx) is k-index (jx, jy, m + nz/2))."""
    f = _f64(f)
    nz, ny, nx = f.shape
    out = np.zeros((nz // 2, ny, nx), dtype=np.complex128)
    b = _f64(box)
    _check(lib().rism_k_fft_r2c(nx, ny, nz, _ptr(b), _ptr(f), out.ctypes.data_as(_dp)))
    return out


def fft_c2r(F, box):
    """Inverse of :func:`fft_r2c` (1/V normalisation)."""
    F = np.ascontiguousarray(F, dtype=np.complex128)
    m, ny, nx = F.shape
    nz = 2 * m
    out = np.zeros((nz, ny, nx))
    b = _f64(box)
    _check(lib().rism_k_fft_c2r(nx, ny, nz, _ptr(b), F.ctypes.data_as(_dp), _ptr(out)))
    return out


def fft_noalloc(direction, grid, box, real_ptr, spec_ptr, stream=0):
    """rism_k_fft_noalloc on device buffers given by address (e.g. a torch
    tensor's data_ptr()): direction -1 real -> half spectrum, +1 half
    spectrum (overwritten) -> real; enqueued on `stream`, nothing allocated."""
    nx, ny, nz = (int(g) for g in grid)
    b = _f64(box)
    _check(lib().rism_k_fft_noalloc(int(direction), nx, ny, nz, _ptr(b), ctypes.c_void_p(real_ptr),
                                    ctypes.c_void_p(spec_ptr), ctypes.c_void_p(stream)))


def copy_bench(nbytes, reps=5):
    """Average ms of one plain device copy of nbytes (rism_k_copy_bench)."""
    ms = ctypes.c_double(0.0)
    _check(lib().rism_k_copy_bench(int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


TUNED_PASSES = ("zfwd_closure", "zinv_resid", "ypass")


def tune_report():
    """The autotuner's choices so far in this process (rism_tune_report):
    one dict per tuned pass and grid shape with the chosen plan and the mean
    launch time of each plan (None: not built for that length)."""
    n = lib().rism_tune_report(None, 0)
    arr = (TuneEntry * max(n, 1))()
    n = lib().rism_tune_report(arr, n)
    out = []
    for e in arr[:n]:
        out.append({"pass": TUNED_PASSES[e.pass_], "grid": [e.nx, e.ny, e.nz], "natv": e.natv,
                    "dtype": "f32" if e.fp32 else "f64", "choice": e.choice,
                    "ms": [m if m >= 0 else None for m in e.ms]})
    return out


# ---------------------------------------------------------------------------
# the solver
# ---------------------------------------------------------------------------

def solve_batch(solvers, maxstep=10000, tol=1e-6, threads=8):
    """rism_solve on several RISM3D objects at once (after initialize());
    returns [(status, steps, rms)]."""
    n = len(solvers)
    arr = (ctypes.c_void_p * n)(*[g._ctx for g in solvers])
    st, sp = (ctypes.c_int * n)(), (ctypes.c_int * n)()
    rms = np.zeros(n)
    _check(lib().rism_solve_batch(arr, n, int(maxstep), float(tol), int(threads), st, sp,
                                  _ptr(rms)))
    return [(st[i], sp[i], rms[i]) for i in range(n)]


def run_batch(solvers, maxstep=10000, tol=1e-6, threads=8):
    """Potentials + batched solve + thermodynamics; returns [(thermo dict,
    status, steps)]."""
    n = len(solvers)
    arr = (ctypes.c_void_p * n)(*[g._ctx for g in solvers])
    th = (Thermo * n)()
    st, sp = (ctypes.c_int * n)(), (ctypes.c_int * n)()
    _check(lib().rism_run_batch(arr, n, int(maxstep), float(tol), int(threads), th, st, sp))
    out = []
    for i, g in enumerate(solvers):
        g.last_thermo = th[i]
        out.append((th[i].as_dict(g.closure), st[i], sp[i]))
    return out


class RISM3D:
    """One 3D-RISM solve on the GPU (RISM3D of rism3d.h:21-104).

    Construct from arrays (grid/box, merged solvent, centred solute) or use
    :meth:`from_inp` for the reference's .inp/.xvk files.
    """

    def __init__(self, grid, box, solvent, solute, closure=KH, count=2, m=0.6, mp=0.2,
                 ad_switch=0, lam=1.0, verbose=False, chi_spline=False, anderson_m=0,
                 precision=0, mixed_switch=0.0, mixed_block=0):
        self.grid = [int(g) for g in grid]
        self.box = [float(b) for b in box]
        self.natv = int(solvent["natv"])
        self.closure = int(closure)
        self.N = self.grid[0] * self.grid[1] * self.grid[2]
        # keep the arrays alive for the duration of rism_create
        self._keep = [_f64(solvent[k]) for k in ("vsig", "veps", "vq", "vrho")]
        self._keep.append(_f64(np.asarray(solvent["xvv"]).reshape(-1)))
        sv = SolventC(self.natv, int(solvent["ntab"]), float(solvent["dr"]),
                      float(solvent["temper"]), float(solvent["xt"]),
                      *[_ptr(a) for a in self._keep])
        natu = len(solute["usig"])
        us = [_f64(solute[k]) for k in ("usig", "ueps", "uq")]
        us.append(_f64(np.asarray(solute["uxyz"]).reshape(-1)))
        self._keep += us
        su = SoluteC(natu, *[_ptr(a) for a in us])
        self._natu = natu
        g = Grid(self.grid[0], self.grid[1], self.grid[2], (ctypes.c_double * 3)(*self.box))
        op = Opts(int(closure), int(count), float(m), float(mp), int(ad_switch), float(lam),
                  int(bool(verbose)), int(bool(chi_spline)), int(anderson_m), int(precision),
                  float(mixed_switch), int(mixed_block))
        err = ctypes.c_int(0)
        self._ctx = lib().rism_create(ctypes.byref(g), ctypes.byref(sv), ctypes.byref(su),
                                      ctypes.byref(op), ctypes.byref(err))
        if not self._ctx:
            raise RismError(err.value, lib().rism_last_error().decode())
        self.shift = np.zeros(3)

    @classmethod
    def from_inp(cls, control, structure=None, center=True, q0=False, solvent_dir=None,
                 ad_switch=0, lam=1.0, verbose=False, chi_spline=False, anderson_m=0,
                 precision=0, mixed_switch=0.0, mixed_block=0):
        d = read_input(control, structure, center, q0, solvent_dir)
        obj = cls(d["grid"], d["box"], d, d, d["closure"], d["count"], d["m"], d["mp"],
                  ad_switch, lam, verbose, chi_spline, anderson_m, precision, mixed_switch,
                  mixed_block)
        obj.input = d
        obj.shift = d["shift"]
        return obj

    def close(self):
        if getattr(self, "_ctx", None):
            lib().rism_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # initialize -> iterate -> output (main.cu:63-65)
    def initialize(self):
        """cal_potential + initialize_tuv (initialize.cu:43, iterate.cu:34)."""
        _check(lib().rism_setup_potentials(self._ctx))
        return self

    def calculate(self):
        """RISM3D::calculate + cal_rms: one step; returns the rms."""
        r = ctypes.c_double(0.0)
        _check(lib().rism_calculate(self._ctx, ctypes.byref(r)))
        return r.value

    def anderson(self):
        """AN2::calculate."""
        _check(lib().rism_anderson(self._ctx))

    def iterate(self, maxstep=10000, tol=1e-6):
        """RISM3D::iterate loop; returns (status, steps, rms)."""
        st, n, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().rism_solve(self._ctx, int(maxstep), float(tol), ctypes.byref(st),
                                ctypes.byref(n), ctypes.byref(r)))
        return st.value, n.value, r.value

    def thermo(self):
        th = Thermo()
        _check(lib().rism_thermo(self._ctx, ctypes.byref(th)))
        self.last_thermo = th
        return th.as_dict(self.closure)

    def output(self, fname, q0=False):
        """The 'm' and 'g' outputs of RISM3D::output."""
        self.thermo()
        write_xmu(fname + (".xmu0" if q0 else ".xmu"), self.closure, self.last_thermo)
        s = _f64(self.shift)
        _check(lib().rism_write_guv(self._ctx, (fname + ".guv").encode(), _ptr(s)))

    def run(self, maxstep=10000, tol=1e-6):
        """potentials + loop + thermodynamics, timed on the device."""
        th, tm = Thermo(), Timing()
        _check(lib().rism_run(self._ctx, int(maxstep), float(tol), ctypes.byref(th),
                              ctypes.byref(tm)))
        self.last_thermo = th
        return th.as_dict(self.closure), {
            "setup_ms": tm.setup_ms, "iterate_ms": tm.iterate_ms, "thermo_ms": tm.thermo_ms,
            "steps": tm.steps, "status": tm.status, "rms": tm.rms,
            "steps_fp32": tm.steps_fp32}

    def set_precision(self, prec):
        """Moves the loop state to FP64 (0) or FP32 (1) (mixed contexts)."""
        _check(lib().rism_set_precision(self._ctx, int(prec)))

    def precision(self):
        return lib().rism_get_precision(self._ctx)

    def get(self, which, site=0):
        out = np.zeros(self.N)
        _check(lib().rism_get_field(self._ctx, int(which), int(site), _ptr(out)))
        return out

    def set(self, which, site, values):
        a = _f64(values).reshape(-1)
        assert a.size == self.N
        _check(lib().rism_set_field(self._ctx, int(which), int(site), _ptr(a)))

    def fields(self, which):
        return np.stack([self.get(which, v) for v in range(self.natv)])

    def fk_half(self):
        out = np.zeros((self.grid[2] // 2, self.grid[1], self.grid[0]), dtype=np.complex128)
        _check(lib().rism_k_get_fk(self._ctx, out.ctypes.data_as(_dp)))
        return out

    def an2_state(self):
        x, s = np.zeros(2), np.zeros(2)
        cnt, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().rism_k_an2_state(self._ctx, _ptr(x), _ptr(s), ctypes.byref(cnt),
                                      ctypes.byref(b)))
        return x, s, cnt.value, b.value

    def kernel_timing(self, on=True):
        """Bracket every hot-path launch with hipEvents (clears the counters)."""
        _check(lib().rism_set_kernel_timing(self._ctx, int(bool(on))))

    def kernel_stats(self):
        """{kernel name: (launches, total ms)} since the last kernel_timing()."""
        arr = (KStat * len(KERNELS))()
        _check(lib().rism_get_kernel_stats(self._ctx, arr))
        return {KERNELS[i]: (arr[i].launches, arr[i].total_ms) for i in range(len(KERNELS))}

    def write_tuv(self, path):
        _check(lib().rism_write_tuv(self._ctx, str(path).encode()))

    def analysis(self, kind):
        """Per-atom integrals of g (cal_grad / cal_euv / cal_ad): kind 0 grad
        (natu x 3, kcal/mol/A), 1 euv (natu x (1 + natv), J/mol), 2 ad
        (natu, kcal/mol), as written to .gra / .euv / <base>-<s>-<lambda>."""
        n = self._natu
        width = {0: 3, 1: 1 + self.natv, 2: 1}[int(kind)]
        out = np.zeros(n * width)
        _check(lib().rism_analysis(self._ctx, int(kind), _ptr(out)))
        return out.reshape(n, width) if width > 1 else out

    def write_huv(self, path):
        """output_huv: text dump of h (last step) with the solute header."""
        _check(lib().rism_write_huv(self._ctx, str(path).encode()))

    def write_cuv(self, path):
        """output_cuv: text dump of c = h - t as "(re,im)" pairs."""
        _check(lib().rism_write_cuv(self._ctx, str(path).encode()))

    def read_tuv(self, path):
        _check(lib().rism_read_tuv(self._ctx, str(path).encode()))
