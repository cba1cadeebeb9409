"""Python mirror of the MI355X NFFT C ABI (``libnfft3_hip.so``).

The library exports NFFT3's plan API (``nfft_init_guru``,
``nfft_precompute_one_psi``, ``nfft_trafo``, ``nfft_adjoint``, ...; reference
include/nfft3.h:163-187) with plan structs laid out field for field as the
reference's.  This module declares those structs with ctypes and wraps them in
a small ``Plan`` class whose ``x``, ``f`` and ``f_hat`` are numpy views of the
plan's own host arrays, so tests read like the reference's (set x, set f_hat,
precompute, trafo, compare f).  Everything runs through the C entry points;
there is no Python or CPU fallback: the library must be built
(``python -m nfft_amd.build``) and a HIP device present.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnfft3_hip.so")

# plan flags (reference include/nfft3.h:195-208)
PRE_PHI_HUT = 1 << 0
FG_PSI = 1 << 1
PRE_LIN_PSI = 1 << 2
PRE_FG_PSI = 1 << 3
PRE_PSI = 1 << 4
PRE_FULL_PSI = 1 << 5
MALLOC_X = 1 << 6
MALLOC_F_HAT = 1 << 7
MALLOC_F = 1 << 8
FFT_OUT_OF_PLACE = 1 << 9
FFTW_INIT = 1 << 10
NFFT_SORT_NODES = 1 << 11
NFFT_OMP_BLOCKWISE_ADJOINT = 1 << 12
PRE_ONE_PSI = PRE_LIN_PSI | PRE_FG_PSI | PRE_PSI | PRE_FULL_PSI
FFTW_MEASURE = 0
FFTW_DESTROY_INPUT = 1 << 0
FFTW_ESTIMATE = 1 << 6

# solver flags (reference include/nfft3.h:823-829)
LANDWEBER = 1 << 0
STEEPEST_DESCENT = 1 << 1
CGNR = 1 << 2
CGNE = 1 << 3
NORMS_FOR_LANDWEBER = 1 << 4
PRECOMPUTE_WEIGHT = 1 << 5
PRECOMPUTE_DAMP = 1 << 6

_INT = ctypes.c_ssize_t  # NFFT_INT = ptrdiff_t
_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


def _plan_fields(R):
    P = ctypes.POINTER
    return [
        ("N_total", _INT), ("M_total", _INT),
        ("f_hat", P(R)), ("f", P(R)),
        ("mv_trafo", ctypes.c_void_p), ("mv_adjoint", ctypes.c_void_p),
        ("d", _INT), ("N", P(_INT)), ("sigma", P(R)), ("n", P(_INT)),
        ("n_total", _INT), ("m", _INT), ("b", P(R)), ("K", _INT),
        ("flags", ctypes.c_uint), ("fftw_flags", ctypes.c_uint),
        ("x", P(R)), ("MEASURE_TIME_t", R * 3),
        ("my_fftw_plan1", ctypes.c_void_p), ("my_fftw_plan2", ctypes.c_void_p),
        ("c_phi_inv", P(P(R))), ("psi", P(R)),
        ("psi_index_g", P(_INT)), ("psi_index_f", P(_INT)),
        ("g", P(R)), ("g_hat", P(R)), ("g1", P(R)), ("g2", P(R)),
        ("spline_coeffs", P(R)), ("index_x", P(_INT)),
    ]


def _solver_fields(R):
    P = ctypes.POINTER
    return [
        ("mv", ctypes.c_void_p), ("flags", ctypes.c_uint),
        ("w", P(R)), ("w_hat", P(R)), ("y", P(R)),
        ("f_hat_iter", P(R)), ("r_iter", P(R)), ("z_hat_iter", P(R)),
        ("p_hat_iter", P(R)), ("v_iter", P(R)),
        ("alpha_iter", R), ("beta_iter", R), ("dot_r_iter", R), ("dot_r_iter_old", R),
        ("dot_z_hat_iter", R), ("dot_z_hat_iter_old", R), ("dot_p_hat_iter", R),
        ("dot_v_iter", R),
    ]


class nfftf_plan(ctypes.Structure):
    _fields_ = _plan_fields(ctypes.c_float)


class nfft_plan(ctypes.Structure):
    _fields_ = _plan_fields(ctypes.c_double)


class solverf_plan_complex(ctypes.Structure):
    _fields_ = _solver_fields(ctypes.c_float)


class solver_plan_complex(ctypes.Structure):
    _fields_ = _solver_fields(ctypes.c_double)


# real-valued solver (solver[f]_plan_double) and its mv plan prefix
# (reference include/nfft3.h:54-60, 786-819): one real per vector entry
class solverf_plan_double(ctypes.Structure):
    _fields_ = _solver_fields(ctypes.c_float)


class solver_plan_double(ctypes.Structure):
    _fields_ = _solver_fields(ctypes.c_double)


def _mv_fields(R):
    P = ctypes.POINTER
    return [("N_total", _INT), ("M_total", _INT), ("f_hat", P(R)), ("f", P(R)),
            ("mv_trafo", _FN), ("mv_adjoint", _FN)]


class nfftf_mv_plan_double(ctypes.Structure):
    _fields_ = _mv_fields(ctypes.c_float)


class nfft_mv_plan_double(ctypes.Structure):
    _fields_ = _mv_fields(ctypes.c_double)


_LIB = None


def lib():
    """Load libnfft3_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libnfft3_hip.so is not built: run `python -m nfft_amd.build`")
        # torch (device buffers of the harness and the density helpers) must
        # create its GPU context before this library's first HIP call: when
        # the library initialised HIP first, torch later reported no GPU
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for pre, P, S, R in (("nfftf_", nfftf_plan, solverf_plan_complex, ctypes.c_float),
                             ("nfft_", nfft_plan, solver_plan_complex, ctypes.c_double)):
            pp = ctypes.POINTER(P)
            ip = ctypes.POINTER(ctypes.c_int)
            sig = {
                "init_guru": [pp, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, ctypes.c_uint, ctypes.c_uint],
                "init_lin": [pp, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, ctypes.c_int, ctypes.c_uint, ctypes.c_uint],
                "init": [pp, ctypes.c_int, ip, ctypes.c_int],
                "init_1d": [pp, ctypes.c_int, ctypes.c_int],
                "init_2d": [pp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
                "init_3d": [pp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
                "trafo_device": [pp, ctypes.c_void_p, ctypes.c_void_p],
                "adjoint_device": [pp, ctypes.c_void_p, ctypes.c_void_p],
                "set_stream": [pp, ctypes.c_void_p],
                "set_timing": [pp, ctypes.c_int],
                "timing_begin": [pp],
                "timing_end": [pp, ctypes.POINTER(ctypes.c_double)],
            }
            for name in ("precompute_one_psi", "precompute_psi", "precompute_full_psi",
                         "precompute_fg_psi", "precompute_lin_psi", "trafo", "trafo_1d",
                         "trafo_2d", "trafo_3d", "adjoint", "adjoint_1d", "adjoint_2d",
                         "adjoint_3d", "trafo_direct", "adjoint_direct", "finalize",
                         "synchronize"):
                sig[name] = [pp]
            for name, args in sig.items():
                fn = getattr(L, pre + name)
                fn.argtypes = args
                fn.restype = None
            getattr(L, pre + "check").argtypes = [pp]
            getattr(L, pre + "check").restype = ctypes.c_char_p
            getattr(L, pre + "get_stream").argtypes = [pp]
            getattr(L, pre + "get_stream").restype = ctypes.c_void_p
            getattr(L, pre + "num_subproblems").argtypes = [pp]
            getattr(L, pre + "num_subproblems").restype = ctypes.c_long
            spre = "solverf_" if pre == "nfftf_" else "solver_"
            sp = ctypes.POINTER(S)
            if hasattr(L, spre + "init_advanced_complex"):
                getattr(L, spre + "init_advanced_complex").argtypes = [sp, ctypes.c_void_p, ctypes.c_uint]
                getattr(L, spre + "init_complex").argtypes = [sp, ctypes.c_void_p]
                for name in ("before_loop_complex", "loop_one_step_complex", "finalize_complex"):
                    getattr(L, spre + name).argtypes = [sp]
                for name in ("init_advanced_complex", "init_complex", "before_loop_complex",
                             "loop_one_step_complex", "finalize_complex"):
                    getattr(L, spre + name).restype = None
        _LIB = L
    return _LIB


def _view(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    addr = ctypes.cast(ptr, ctypes.c_void_p).value
    if not addr:
        raise RuntimeError("array not allocated")
    buf = (ctypes.c_byte * (count * np.dtype(dtype).itemsize)).from_address(addr)
    return np.frombuffer(buf, dtype=dtype)


class Plan:
    """An nfft_plan (precision 'double') or nfftf_plan ('single') driven
    through the C ABI.  Construction calls ``init_guru`` (or ``init`` when
    ``n``, ``m`` and ``flags`` are all left out, ``init_1d/2d/3d`` with
    ``how='init_Nd'``, ``init_lin`` with ``how='init_lin'`` and ``lin_K``)."""

    def __init__(self, N, M, n=None, m=None, flags=None, precision="double",
                 fftw_flags=FFTW_ESTIMATE | FFTW_DESTROY_INPUT, how=None, lin_K=1 << 20):
        self.L = lib()
        self.precision = precision
        self.pre = "nfft_" if precision == "double" else "nfftf_"
        self.rdtype = np.float64 if precision == "double" else np.float32
        self.cdtype = np.complex128 if precision == "double" else np.complex64
        self.p = (nfft_plan if precision == "double" else nfftf_plan)()
        self.pp = ctypes.pointer(self.p)
        N = [int(t) for t in N]
        d = len(N)
        Na = (ctypes.c_int * d)(*N)
        if how == "init_Nd":
            getattr(self.L, self.pre + "init_%dd" % d)(self.pp, *N, int(M))
        elif how == "init_lin":
            # init_lin(plan, d, N, M, n, m, K, flags, fftw_flags): PRE_LIN_PSI
            # with a table of K window samples per unit
            n = n or [2 * _next_pow2(t) for t in N]
            m = m or (8 if precision == "double" else 4)
            if flags is None:
                flags = PRE_PHI_HUT | PRE_LIN_PSI | MALLOC_X | MALLOC_F_HAT | MALLOC_F | FFTW_INIT | FFT_OUT_OF_PLACE
            na = (ctypes.c_int * d)(*[int(t) for t in n])
            self._fn("init_lin")(self.pp, d, Na, int(M), na, int(m), int(lin_K), int(flags),
                                 int(fftw_flags))
        elif how == "init" or (n is None and m is None and flags is None):
            self._fn("init")(self.pp, d, Na, int(M))
        else:
            if n is None:
                n = [2 * _next_pow2(t) for t in N]
            if m is None:
                m = 8 if precision == "double" else 4
            if flags is None:
                flags = PRE_PHI_HUT | PRE_PSI | MALLOC_X | MALLOC_F_HAT | MALLOC_F | FFTW_INIT | FFT_OUT_OF_PLACE
            na = (ctypes.c_int * d)(*[int(t) for t in n])
            self._fn("init_guru")(self.pp, d, Na, int(M), na, int(m), int(flags), int(fftw_flags))
        self.alive = True

    def _fn(self, name):
        return getattr(self.L, self.pre + name)

    # numpy views of the plan's host arrays
    @property
    def d(self):
        return self.p.d

    @property
    def M(self):
        return self.p.M_total

    @property
    def N(self):
        return [self.p.N[t] for t in range(self.p.d)]

    @property
    def n(self):
        return [self.p.n[t] for t in range(self.p.d)]

    @property
    def m(self):
        return self.p.m

    @property
    def x(self):
        return _view(self.p.x, self.p.M_total * self.p.d, self.rdtype).reshape(self.p.M_total, self.p.d)

    @property
    def f(self):
        return _view(self.p.f, self.p.M_total, self.cdtype)

    @property
    def f_hat(self):
        return _view(self.p.f_hat, self.p.N_total, self.cdtype)

    @property
    def measure_time(self):
        return list(self.p.MEASURE_TIME_t)

    def precompute(self):
        self._fn("precompute_one_psi")(self.pp)

    def trafo(self):
        self._fn("trafo")(self.pp)

    def adjoint(self):
        self._fn("adjoint")(self.pp)

    def trafo_direct(self):
        self._fn("trafo_direct")(self.pp)

    def adjoint_direct(self):
        self._fn("adjoint_direct")(self.pp)

    def call(self, name):
        self._fn(name)(self.pp)

    def check(self):
        r = self._fn("check")(self.pp)
        return None if r is None else r.decode()

    def trafo_device(self, f_hat_ptr, f_ptr):
        self._fn("trafo_device")(self.pp, ctypes.c_void_p(f_hat_ptr), ctypes.c_void_p(f_ptr))

    def adjoint_device(self, f_ptr, f_hat_ptr):
        self._fn("adjoint_device")(self.pp, ctypes.c_void_p(f_ptr), ctypes.c_void_p(f_hat_ptr))

    def set_stream(self, stream_ptr):
        self._fn("set_stream")(self.pp, ctypes.c_void_p(stream_ptr))

    def set_timing(self, on):
        self._fn("set_timing")(self.pp, 1 if on else 0)

    def synchronize(self):
        self._fn("synchronize")(self.pp)

    def timing_begin(self):
        self._fn("timing_begin")(self.pp)

    def timing_end(self):
        """Summed stage seconds since timing_begin: dict with call counts and
        trafo D / FFT / B, adjoint B / FFT / D."""
        out = (ctypes.c_double * 9)()
        self._fn("timing_end")(self.pp, out)
        keys = ["n_trafo", "n_adjoint", "trafo_D", "trafo_FFT", "trafo_B",
                "adjoint_B", "adjoint_Bred", "adjoint_FFT", "adjoint_D"]
        return dict(zip(keys, list(out)))

    def gridding_device(self, f_ptr, w_ptr, f_hat_ptr):
        """f_hat = A^H (w .* f) on device buffers; w_ptr 0: plain adjoint."""
        self._fn("gridding_device")(self.pp, ctypes.c_void_p(f_ptr), ctypes.c_void_p(w_ptr or None),
                                    ctypes.c_void_p(f_hat_ptr))

    def gridding(self, w=None):
        """f_hat = A^H (w .* f) with the plan's host f and f_hat (f unchanged)."""
        if w is None:
            self._fn("gridding")(self.pp, None)
            return
        w = np.ascontiguousarray(w, dtype=self.rdtype)
        assert w.size == self.M
        self._fn("gridding")(self.pp, w.ctypes.data_as(ctypes.c_void_p))

    def trafo_batch_device(self, f_hat_ptr, f_ptr, nrhs):
        self._fn("trafo_batch_device")(self.pp, ctypes.c_void_p(f_hat_ptr), ctypes.c_void_p(f_ptr),
                                       int(nrhs))

    def adjoint_batch_device(self, f_ptr, f_hat_ptr, nrhs):
        self._fn("adjoint_batch_device")(self.pp, ctypes.c_void_p(f_ptr), ctypes.c_void_p(f_hat_ptr),
                                         int(nrhs))

    def num_subproblems(self):
        return self._fn("num_subproblems")(self.pp)

    def finalize(self):
        if self.alive:
            self._fn("finalize")(self.pp)
            self.alive = False

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass


class Solver:
    """A solver_plan_complex / solverf_plan_complex on a Plan, driven through
    the C ABI (solver_init_advanced_complex, solver_before_loop_complex,
    solver_loop_one_step_complex, solver_finalize_complex; reference
    include/nfft3.h:782-815).  y, w, w_hat, f_hat_iter, r_iter, ... are
    numpy views of the solver's host arrays."""

    def __init__(self, plan, flags=CGNR):
        self.plan = plan
        self.L = plan.L
        self.pre = "solver_" if plan.precision == "double" else "solverf_"
        self.s = (solver_plan_complex if plan.precision == "double" else solverf_plan_complex)()
        self.sp = ctypes.pointer(self.s)
        mv = ctypes.cast(plan.pp, ctypes.c_void_p)
        getattr(self.L, self.pre + "init_advanced_complex")(self.sp, mv, int(flags))
        self.alive = True

    def _c(self, name, n):
        return _view(getattr(self.s, name), n, self.plan.cdtype)

    def _r(self, name, n):
        return _view(getattr(self.s, name), n, self.plan.rdtype)

    @property
    def y(self):
        return self._c("y", self.plan.M)

    @property
    def f_hat_iter(self):
        return self._c("f_hat_iter", self.plan.p.N_total)

    @property
    def r_iter(self):
        return self._c("r_iter", self.plan.M)

    @property
    def z_hat_iter(self):
        return self._c("z_hat_iter", self.plan.p.N_total)

    @property
    def p_hat_iter(self):
        return self._c("p_hat_iter", self.plan.p.N_total)

    @property
    def w(self):
        return self._r("w", self.plan.M)

    @property
    def w_hat(self):
        return self._r("w_hat", self.plan.p.N_total)

    def before_loop(self):
        getattr(self.L, self.pre + "before_loop_complex")(self.sp)

    def loop_one_step(self):
        getattr(self.L, self.pre + "loop_one_step_complex")(self.sp)

    def set_sync(self, level):
        fn = getattr(self.L, self.pre + "set_sync_complex")
        fn.argtypes = [type(self.sp), ctypes.c_int]
        fn(self.sp, int(level))

    def sync(self):
        getattr(self.L, self.pre + "sync_complex")(self.sp)

    def finalize(self):
        if self.alive:
            getattr(self.L, self.pre + "finalize_complex")(self.sp)
            self.alive = False

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass


def _next_pow2(x):
    if x < 2:
        return x + 1
    p = 1
    while p < x:
        p <<= 1
    return p


class Graph:
    """hipGraph of the batched transforms of several plans (nfftf_graph_trafo /
    nfftf_graph_adjoint and nfft_hip_graph_launch / _synchronize / _destroy,
    include/nfft3_hip.h).  ``inputs`` / ``outputs`` are device pointers, one
    per plan, each holding ``nrhs`` vectors."""

    def __init__(self, plans, inputs, outputs, nrhs=1, adjoint=False):
        self.L = lib()
        prec = plans[0].precision
        assert all(p.precision == prec for p in plans)
        pre = "nfft_" if prec == "double" else "nfftf_"
        n = len(plans)
        PP = ctypes.c_void_p * n
        pl = PP(*[ctypes.cast(p.pp, ctypes.c_void_p) for p in plans])
        ins = PP(*[ctypes.c_void_p(int(a)) for a in inputs])
        outs = PP(*[ctypes.c_void_p(int(a)) for a in outputs])
        fn = getattr(self.L, pre + ("graph_adjoint" if adjoint else "graph_trafo"))
        fn.restype = ctypes.c_void_p
        self.h = fn(pl, n, ins, outs, int(nrhs))
        self.plans = plans  # keep the plans alive while the graph exists

    def launch(self):
        self.L.nfft_hip_graph_launch(ctypes.c_void_p(self.h))

    def synchronize(self):
        self.L.nfft_hip_graph_synchronize(ctypes.c_void_p(self.h))

    def destroy(self):
        if self.h:
            self.L.nfft_hip_graph_destroy(ctypes.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
