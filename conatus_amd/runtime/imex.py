"""Batched IMEX integrator on the MI355X: M_k dX/dt + L_k X = F for every mode k.

The device twin of the reference's ``cuda.Imex`` (python/conatus/runtime/cuda.py:267-308,
core/src/cuda/imex.cu:56-244) and of the host ``ImexIntegrator`` step arithmetic
(core/include/conatus/imex.hpp:73-219): SBDF1..4 with the lower-order startup, RK111 / RK222 /
RK443 and RKSMR / RKGFY (python/conatus/timestep/schemes.py:122-137), the Crank-Nicolson
multistep family CNAB1 / CNAB2 / MCNAB2 / CNLF2 (schemes.py:91-103, step of
python/conatus/timestep/integrator.py:576-602, CNAB1 until the history is full), factorizations
cached by their LHS coefficients (RK443: one), the continuation rule of multistep histories, and
F = g * X + h (pointwise) or an external F writer.

One SBDF step is two launches on the current stream: the fused stage kernel (M X, F, history
level and right-hand side in one pass) and the out-of-place batched solve into X.  History
levels live in a ring of fixed device buffers, so a full-order step captured as a HIP graph per
ring phase replays without copies (``run_graph``).
"""
from __future__ import annotations

import ctypes
from collections.abc import Sequence

import numpy as np
import scipy.sparse as sp

from .. import _lib
from ..errors import OptionError
from .banded import BatchHandle, _factor_dtype, csr_args, solve_operand

__all__ = ["Imex", "DeviceCsr", "SCHEMES", "CN_SCHEMES", "sbdf_coefficients", "rk_tableau", "cn_coefficients"]

CN_SCHEMES = ("CNAB1", "CNAB2", "MCNAB2", "CNLF2")
SCHEMES = ("SBDF1", "SBDF2", "SBDF3", "SBDF4", "RK111", "RK222", "RK443", "RKSMR", "RKGFY") + CN_SCHEMES


def sbdf_coefficients(order: int):
    """(a, c), a scaled by dt: the constant-step SBDF1..4 constants (imex.hpp:63-71)."""
    table = {1: ((1.0, -1.0), (0.0, 1.0)),
             2: ((1.5, -2.0, 0.5), (0.0, 2.0, -1.0)),
             3: ((11.0 / 6, -3.0, 1.5, -1.0 / 3), (0.0, 3.0, -3.0, 1.0)),
             4: ((25.0 / 12, -4.0, 3.0, -4.0 / 3, 0.25), (0.0, 4.0, -6.0, 4.0, -1.0))}
    if order not in table:
        raise OptionError("SBDF order must be 1..4")
    return table[order]


def cn_coefficients(name: str):
    """(a, b, c), a in units of 1/dt, of a CN-family multistep scheme (schemes.py:91-103):
    sum_j (a_j/dt) M X^{n+1-j} + sum_j b_j L X^{n+1-j} = sum_{j>=1} c_j F^{n+1-j}."""
    table = {"CNAB1": ((1.0, -1.0), (0.5, 0.5), (0.0, 1.0)),
             "CNAB2": ((1.0, -1.0, 0.0), (0.5, 0.5, 0.0), (0.0, 1.5, -0.5)),
             "MCNAB2": ((1.0, -1.0, 0.0), (9 / 16, 6 / 16, 1 / 16), (0.0, 1.5, -0.5)),
             "CNLF2": ((0.5, 0.0, -0.5), (0.5, 0.0, 0.5), (0.0, 1.0, 0.0))}
    if name not in table:
        raise OptionError(f"unknown CN scheme: {name}", fix=f"one of {CN_SCHEMES}")
    return table[name]


def rk_tableau(name: str):
    """(A, H, c) of the IMEX tableaux RK111, RK222, RK443 (imex.hpp:31-51) and RKSMR, RKGFY
    (schemes.py:122-137), in the cumulative form M X_i = M X_0 + dt sum (A_ij F_j - H_ij L X_j)."""
    if name == "RK111":
        return ((0.0, 0.0), (1.0, 0.0)), ((0.0, 0.0), (0.0, 1.0)), (0.0, 1.0)
    if name == "RK222":
        g = (2.0 - np.sqrt(2.0)) / 2.0
        d = 1.0 - 1.0 / g / 2.0
        return (((0.0, 0.0, 0.0), (g, 0.0, 0.0), (d, 1 - d, 0.0)),
                ((0.0, 0.0, 0.0), (0.0, g, 0.0), (0.0, 1 - g, g)), (0.0, g, 1.0))
    if name == "RK443":
        return (((0, 0, 0, 0, 0), (1 / 2, 0, 0, 0, 0), (11 / 18, 1 / 18, 0, 0, 0),
                 (5 / 6, -5 / 6, 1 / 2, 0, 0), (1 / 4, 7 / 4, 3 / 4, -7 / 4, 0)),
                ((0, 0, 0, 0, 0), (0, 1 / 2, 0, 0, 0), (0, 1 / 6, 1 / 2, 0, 0),
                 (0, -1 / 2, 1 / 2, 1 / 2, 0), (0, 3 / 2, -3 / 2, 1 / 2, 1 / 2)),
                (0, 1 / 2, 2 / 3, 1 / 2, 1))
    if name == "RKSMR":
        # Spalart-Moser-Rogers substeps X_i = X_{i-1} + dt [g_i F_{i-1} + z_i F_{i-2} - a_i L X_{i-1}
        # - b_i L X_i] summed into cumulative rows
        al, be = (29 / 96, -3 / 40, 1 / 6), (37 / 160, 5 / 24, 1 / 6)
        ga, ze = (8 / 15, 5 / 12, 3 / 4), (-17 / 60, -5 / 12)
        return (((0, 0, 0, 0), (ga[0], 0, 0, 0), (ga[0] + ze[0], ga[1], 0, 0),
                 (ga[0] + ze[0], ga[1] + ze[1], ga[2], 0)),
                ((0, 0, 0, 0), (al[0], be[0], 0, 0), (al[0], be[0] + al[1], be[1], 0),
                 (al[0], be[0] + al[1], be[1] + al[2], be[2])), (0, 8 / 15, 2 / 3, 1))
    if name == "RKGFY":                  # Heun on F, Crank-Nicolson on L
        return (((0, 0, 0), (1, 0, 0), (1 / 2, 1 / 2, 0)), ((0, 0, 0), (1 / 2, 1 / 2, 0), (1 / 2, 0, 1 / 2)),
                (0, 1, 1))
    raise OptionError(f"unknown RK scheme: {name}", fix=f"one of {SCHEMES}")


def _torch():
    import torch
    return torch


def _stream() -> int:
    return _torch().cuda.current_stream().cuda_stream


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


class DeviceCsr:
    """A batch of CSR matrices (same n) as one block-diagonal device operator (``cn_csr``)."""

    def __init__(self, mats: Sequence, dtype):
        self.dtype = np.dtype(dtype)
        csrs = [csr_args(A, self.dtype) for A in mats]
        self.batch, self.n = len(csrs), int(csrs[0][0])
        keep = [(c[1], c[2], np.ascontiguousarray(c[3], dtype=self.dtype)) for c in csrs]
        P = ctypes.c_void_p * self.batch
        h = ctypes.c_void_p()
        code = _lib.CN_C128 if self.dtype == np.complex128 else _lib.CN_F64
        st = _lib.lib().cn_csr_create(ctypes.byref(h), code, self.batch, self.n,
                                      P(*[k[0].ctypes.data for k in keep]), P(*[k[1].ctypes.data for k in keep]),
                                      P(*[k[2].ctypes.data for k in keep]),
                                      (ctypes.c_int64 * self.batch)(*[int(k[1].size) for k in keep]))
        _lib.check(st, "DeviceCsr")
        self._h = h

    def matvec(self, x, y, alpha: float = 1.0, accumulate: bool = False) -> None:
        _lib.check(_lib.lib().cn_csr_matvec(self._h, ctypes.c_void_p(_ptr(x)), ctypes.c_void_p(_ptr(y)),
                                            float(alpha), 1 if accumulate else 0, ctypes.c_void_p(_stream())),
                   "DeviceCsr.matvec")

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().cn_csr_destroy(h)
            except Exception:
                pass


def factor_operator(Ms, Ls, alpha: float, beta: float, border_rows, dtype, pivot: bool = True,
                    mode: str | None = None) -> BatchHandle:
    """The factored batch of alpha M_k + beta L_k (the implicit operator of a stage), built the
    same way wherever it is needed, so that equal operators give bit-identical factors."""
    dtype = np.dtype(dtype)
    lhs = [csr_args((dtype.type(alpha) * M + dtype.type(beta) * L).tocsr(), dtype) for M, L in zip(Ms, Ls)]
    return BatchHandle(lhs, border_rows, dtype, pivot, mode=mode)


def capture_graph(fn):
    """Capture the device work that ``fn`` enqueues into a HIP graph (torch's CUDAGraph on a side
    stream ordered after the current one)."""
    torch = _torch()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=s):
        fn()
    torch.cuda.current_stream().wait_stream(s)
    return g


class Imex:
    """Batched IMEX integrator on the GPU (same constructor as python/conatus/runtime/cuda.py:267-286).

    ``g``, ``h``: (batch, n) host arrays of the pointwise explicit term F = g * X + h (None: 0).
    Scheme and dt are required: nothing is chosen for the user."""

    def __init__(self, Ms: Sequence, Ls: Sequence, border_rows: Sequence[int], scheme: str, dt: float,
                 g=None, h=None, pivot: bool = True, dtype=None, explicit=None, mode: str | None = None):
        torch = _torch()
        if scheme not in SCHEMES:
            raise OptionError(f"Imex: unknown scheme {scheme!r}", fix=f"one of {SCHEMES}")
        if not float(dt) > 0:
            raise OptionError("dt must be > 0")
        Ms = [sp.csr_matrix(A) for A in Ms]
        Ls = [sp.csr_matrix(A) for A in Ls]
        if not Ms or len(Ms) != len(Ls):
            raise OptionError("Imex: M and L batches differ")
        self.dtype = np.dtype(dtype) if dtype is not None else np.dtype(_factor_dtype(Ms + Ls))
        self.n = Ms[0].shape[0]
        self.batch = len(Ms)
        for A in Ms + Ls:
            if A.shape != (self.n, self.n):
                raise OptionError("Imex: all modes need the same n")
        self.scheme, self.dt = scheme, float(dt)
        self.mode = mode              # kernel mode of the factorizations (None: automatic)
        self.border_rows = [int(i) for i in border_rows]
        self.pivot = bool(pivot)
        self._Mh = [A.astype(self.dtype) for A in Ms]
        self._Lh = [A.astype(self.dtype) for A in Ls]
        self._M = DeviceCsr(self._Mh, self.dtype)
        self._L = DeviceCsr(self._Lh, self.dtype)
        self._tdt = torch.complex128 if self.dtype == np.complex128 else torch.float64
        N = self.batch * self.n

        def conv(a):
            if a is None:
                return None
            a = solve_operand(a, self.dtype, "Imex (g, h)").ravel()
            if a.size != N:
                raise OptionError("pointwise rhs: expected batch*n values")
            return torch.from_numpy(a).to("cuda")
        self._g, self._hh = conv(g), conv(h)
        # explicit: an object with base() -> (batch*n) device tensor of the constant rows of F and
        # write(x, out) launching F(x) into the rows it owns (other rows of out keep base())
        self._explicit = explicit
        if explicit is not None and (g is not None or h is not None):
            raise OptionError("Imex: pass either the pointwise g, h or an explicit term, not both")
        self.cn = cn_coefficients(scheme) if scheme in CN_SCHEMES else None
        if scheme.startswith("SBDF"):
            self.order = int(scheme[4])
            self.tab = None
        elif self.cn is not None:
            self.order = 0
            self.tab = None
        else:
            self.order = 0
            self.tab = rk_tableau(scheme)
        self._factors: dict = {}
        z = lambda: torch.zeros(N, dtype=self._tdt, device="cuda")   # noqa: E731
        self._rhs = z()
        fz = (lambda: explicit.base().clone()) if explicit is not None else z
        if self.order:
            self._ring = [(z(), z()) for _ in range(self.order)]     # (MX, F) levels
            self._fnew = fz() if explicit is not None else None
        elif self.cn is not None:
            s = len(self.cn[0]) - 1
            self._ring = [(z(), fz(), z()) for _ in range(s)]     # (MX, F, LX) levels
        else:
            s = len(self.tab[2]) - 1
            self._X = [z() for _ in range(s)]
            self._F = [fz() for _ in range(s)]
            self._LX = [z() for _ in range(s)]
            self._MX0 = z()
        self._graphs = {}
        self._t_step = 0.0            # time of the state being stepped (for a time-aware F)
        self.reset()
        # the continuation check compares states byte for byte (as imex.cu:145-151 does); run it
        # once here so that its device code is loaded before any timed step
        zz = torch.zeros((self.batch, self.n), dtype=self._tdt, device="cuda")
        torch.equal(zz.view(torch.uint8), zz.clone().view(torch.uint8))

    def set_pointwise(self, g=None, h=None) -> None:
        """F = g * X + h from (batch, n) host arrays (None: absent), copied to the device once (the
        reference's set_pointwise_rhs, core/src/cuda/imex.cu:113-119); factors and history are kept."""
        if self._explicit is not None:
            raise OptionError("Imex.set_pointwise: this integrator has an explicit term writer")
        torch = _torch()
        N = self.batch * self.n

        def conv(a):
            if a is None:
                return None
            a = solve_operand(a, self.dtype, "Imex (g, h)").ravel()
            if a.size != N:
                raise OptionError("pointwise rhs: expected batch*n values")
            return torch.from_numpy(a).to("cuda")
        self._g, self._hh = conv(g), conv(h)

    # -- state ---------------------------------------------------------------------------------
    def reset(self) -> None:
        """Forget the SBDF history and the iteration count (factorizations kept)."""
        self._levels: list[int] = []          # ring slots, newest first
        self._last = None
        self.iteration = 0

    @property
    def history_length(self) -> int:
        return len(self._levels) if (self.order or self.cn is not None) else 0

    @property
    def n_factorizations(self) -> int:
        return len(self._factors)

    def _factor(self, alpha: float, beta: float) -> BatchHandle:
        key = (alpha, beta)
        f = self._factors.get(key)
        if f is None:
            f = factor_operator(self._Mh, self._Lh, alpha, beta, self.border_rows, self.dtype, self.pivot, self.mode)
            self._factors[key] = f
        return f

    # -- steps ---------------------------------------------------------------------------------
    def _write_F(self, x, out, t: float) -> None:
        """F(x) into out: the explicit writer of the integrator; one that declares ``wants_time``
        (a host callback F(x, t)) is given the time of the state, t + c_j dt for RK stage j
        (imex.hpp:186)."""
        if getattr(self._explicit, "wants_time", False):
            self._explicit.write(x, out, t)
        else:
            self._explicit.write(x, out)

    def _sbdf_step(self, x) -> None:
        # the slot of the new level: a free slot, or the oldest level's
        if len(self._levels) < self.order:
            used = set(self._levels)
            slot = next(i for i in range(self.order) if i not in used)
            levels = [slot] + self._levels
        else:
            slot = self._levels[-1]
            levels = [slot] + self._levels[:-1]
        self._sbdf_launch(x, levels)
        self._levels = levels

    def _sbdf_launch(self, x, levels) -> None:
        q = len(levels)
        a, c = sbdf_coefficients(q)
        mx0, f0 = self._ring[levels[0]]
        prev = levels[1:]
        P = ctypes.c_void_p * 4
        mxh = P(*([_ptr(self._ring[i][0]) for i in prev] + [0] * (4 - len(prev))))
        fh = P(*([_ptr(self._ring[i][1]) for i in prev] + [0] * (4 - len(prev))))
        cc = (ctypes.c_double * 5)(*([0.0] + [float(c[j]) for j in range(1, q + 1)] + [0.0] * (4 - q)))
        adt = (ctypes.c_double * 5)(*([0.0] + [float(a[j] / self.dt) for j in range(1, q + 1)] + [0.0] * (4 - q)))
        v = ctypes.c_void_p
        fext = None
        if self._explicit is not None:
            self._write_F(x, self._fnew, self._t_step)
            fext = v(_ptr(self._fnew))
        st = _lib.lib().cn_sbdf_stage(self._M._h, v(_ptr(x)), v(_ptr(self._g)), v(_ptr(self._hh)), fext,
                                      v(_ptr(mx0)), v(_ptr(f0)), mxh, fh, q, cc, adt, v(_ptr(self._rhs)),
                                      v(_stream()))
        _lib.check(st, "Imex (SBDF stage)")
        fac = self._factor(a[0] / self.dt, 1.0)
        fac.solve_mapped_ptr(_ptr(self._rhs), 0, _ptr(x), 0, _stream())

    def _cn_step(self, x) -> None:
        """One CN-family step: the new level (M X, F(X), L X) of the current state, then
        (M + dt b_0/a_0 L) X^{n+1} = M X_j* + dt/a_0 sum_{j>=1} (c_j F_j - b_j L X_j), the
        reference's (a_0/dt M + b_0 L) X^{n+1} = sum_j (c_j F_j - (a_j/dt) M X_j - b_j L X_j)
        divided by a_0/dt: every scheme of the family has one M X history term, a_j* = -a_0
        (CNAB1, CNAB2, MCNAB2: j* = 1; CNLF2: j* = 2).  The right-hand side is the RK stage
        kernel's M X_0 + sum dA F - sum dH L X."""
        s = len(self.cn[0]) - 1
        if len(self._levels) < s:
            used = set(self._levels)
            slot = next(i for i in range(s) if i not in used)
            levels = [slot] + self._levels
        else:
            levels = [self._levels[-1]] + self._levels[:-1]
        mx, f, lx = self._ring[levels[0]]
        N = self.batch * self.n
        code = _lib.CN_C128 if self.dtype == np.complex128 else _lib.CN_F64
        v = ctypes.c_void_p
        self._M.matvec(x, mx)
        self._L.matvec(x, lx)
        if self._explicit is not None:
            self._write_F(x, f, self._t_step)
        else:
            _lib.check(_lib.lib().cn_pointwise(code, N, v(_ptr(f)), v(_ptr(x)), v(_ptr(self._g)),
                                               v(_ptr(self._hh)), v(_stream())), "Imex (F)")
        self._levels = levels
        a, b, c = self.cn if len(levels) >= s else cn_coefficients("CNAB1")
        q = len(a) - 1
        jstar = [j for j in range(1, q + 1) if a[j] != 0]
        if len(jstar) != 1 or a[jstar[0]] != -a[0]:
            raise OptionError(f"Imex: {self.scheme} is not of the one-M-history CN form")
        P = ctypes.c_void_p * 4
        dA = [self.dt * c[j] / a[0] for j in range(1, q + 1)]
        dH = [self.dt * b[j] / a[0] for j in range(1, q + 1)]
        st = _lib.lib().cn_rk_rhs(code, N, v(_ptr(self._ring[levels[jstar[0] - 1]][0])),
                                  P(*([_ptr(self._ring[levels[k]][1]) if dA[k] != 0 else 0 for k in range(q)]
                                      + [0] * (4 - q))),
                                  P(*([_ptr(self._ring[levels[k]][2]) if dH[k] != 0 else 0 for k in range(q)]
                                      + [0] * (4 - q))),
                                  (ctypes.c_double * 4)(*(dA + [0.0] * (4 - q))),
                                  (ctypes.c_double * 4)(*(dH + [0.0] * (4 - q))), q, v(_ptr(self._rhs)),
                                  v(_stream()))
        _lib.check(st, "Imex (CN right-hand side)")
        fac = self._factor(1.0, self.dt * b[0] / a[0])
        fac.solve_mapped_ptr(_ptr(self._rhs), 0, _ptr(x), 0, _stream())

    def _rk_step(self, x) -> None:
        A, H, cc = self.tab
        s = len(cc) - 1
        N = self.batch * self.n
        code = _lib.CN_C128 if self.dtype == np.complex128 else _lib.CN_F64
        X = [x] + self._X[: s - 1] + [x]      # X_0 = x; the last stage is written into x
        self._M.matvec(x, self._MX0)
        needF = lambda j: any(A[i][j] != 0 for i in range(j + 1, s + 1))    # noqa: E731
        needLX = lambda j: any(H[i][j] != 0 for i in range(j + 1, s + 1))   # noqa: E731
        v = ctypes.c_void_p
        for i in range(1, s + 1):
            j = i - 1
            if needF(j):
                if self._explicit is not None:
                    self._write_F(X[j], self._F[j], self._t_step + float(cc[j]) * self.dt)
                else:
                    _lib.check(_lib.lib().cn_pointwise(code, N, v(_ptr(self._F[j])), v(_ptr(X[j])),
                                                       v(_ptr(self._g)), v(_ptr(self._hh)), v(_stream())),
                               "Imex (F)")
            if needLX(j):
                self._L.matvec(X[j], self._LX[j])
            P = ctypes.c_void_p * 4
            dA = [self.dt * A[i][k] for k in range(i)]
            dH = [self.dt * H[i][k] for k in range(i)]
            st = _lib.lib().cn_rk_rhs(code, N, v(_ptr(self._MX0)),
                                      P(*[_ptr(self._F[k]) if dA[k] != 0 else 0 for k in range(i)]),
                                      P(*[_ptr(self._LX[k]) if dH[k] != 0 else 0 for k in range(i)]),
                                      (ctypes.c_double * 4)(*dA), (ctypes.c_double * 4)(*dH), i,
                                      v(_ptr(self._rhs)), v(_stream()))
            _lib.check(st, "Imex (RK stage)")
            fac = self._factor(1.0, self.dt * H[i][i])
            fac.solve_mapped_ptr(_ptr(self._rhs), 0, _ptr(X[i]), 0, _stream())

    def _check_state(self, d) -> None:
        torch = _torch()
        if not isinstance(d, torch.Tensor) or d.device.type != "cuda" or d.dtype != self._tdt \
                or tuple(d.shape) != (self.batch, self.n) or not d.is_contiguous():
            raise OptionError(f"Imex.run: expected a contiguous ({self.batch}, {self.n}) "
                              f"{np.dtype(self.dtype).name} device tensor")

    def _continuation(self, d, t0: float) -> None:
        torch = _torch()
        # bitwise comparison (a state holding NaN still continues itself)
        if self.history_length and not (self._last is not None and self._last[1] == float(t0)
                                         and torch.equal(self._last[0].view(torch.uint8), d.view(torch.uint8))):
            raise OptionError(f"Imex.run: {self.scheme} holds the history of a previous run, and (x0, t0) "
                              "is not the state and time that run returned",
                              fix="call reset() to start a new trajectory (lower-order startup again), "
                                  "or pass the returned (x, t) to continue")

    def _remember(self, d, t) -> None:
        """Keep the (state, time) a call returned, which the next call must continue from while
        a multistep history (SBDF levels, or the CN family's M X, F, L X levels) is held."""
        if self.order or self.cn is not None:
            self._last = (d.clone(), float(t))
        else:
            self._last = None

    def _next_t(self, t: float, clock) -> float:
        """The time after the step just committed: ``clock(iteration)`` when the owner keeps a
        time origin (t = t_origin + (iteration - it_origin) dt, no accumulated round-off), else
        one addition of dt per step, as imex.hpp:112-116 does."""
        return float(clock(self.iteration)) if clock is not None else t + self.dt

    def adopt_history(self, sources: Sequence, d, t: float) -> None:
        """Continue the trajectories of other integrators' systems: ``sources`` is a list of
        (imex, rows) with rows a device index tensor of that integrator's systems; their states
        are now in ``d`` one source after another, at time ``t``.  Every source has the same
        scheme, dt and n and a full SBDF history (or none, RK); history level i (newest first)
        of each source is copied into level i here, so that the next step of every system is the
        one its source would have made."""
        torch = _torch()
        for src, _ in sources:
            if (self.scheme, self.dt, self.n) != (src.scheme, src.dt, src.n):
                raise OptionError("Imex.adopt_history: the integrators differ in scheme, dt or n")
            if self.order and src.history_length != len(sources[0][0]._levels):
                raise OptionError("Imex.adopt_history: the sources are at different startup orders")
        if self.order:
            q = sources[0][0].history_length
            self._levels = list(range(q))
            shape = (self.batch, self.n)
            for i in range(q):
                mx, f = self._ring[i]
                mx.view(shape).copy_(torch.cat([src._ring[src._levels[i]][0].view(src.batch, src.n).index_select(0, rows)
                                                for src, rows in sources]))
                f.view(shape).copy_(torch.cat([src._ring[src._levels[i]][1].view(src.batch, src.n).index_select(0, rows)
                                               for src, rows in sources]))
        self.iteration = max(src.iteration for src, _ in sources)
        self._remember(d, t)

    def run(self, d, t0: float, nsteps: int, clock=None) -> float:
        """Advance the (batch, n) device tensor ``d`` in place by ``nsteps``; returns the time
        after the last step: t0 plus one addition of dt per step, as the reference's integrator
        computes it (imex.hpp:112-116), or ``clock(iteration)`` when given (the IVP solver's
        t = t_origin + (iteration - it_origin) dt, integrator.py:680)."""
        self._check_state(d)
        if isinstance(nsteps, bool) or int(nsteps) != nsteps or nsteps < 0:
            raise OptionError("Imex.run: nsteps must be a whole number >= 0")
        self._continuation(d, t0)
        t = float(t0)
        for _ in range(int(nsteps)):
            self._t_step = t
            if self.order:
                self._sbdf_step(d)
            elif self.cn is not None:
                self._cn_step(d)
            else:
                self._rk_step(d)
            self.iteration += 1
            t = self._next_t(t, clock)
        self._remember(d, t)
        return t

    def run_graph(self, d, t0: float, nsteps: int, chunk: int = 8, clock=None) -> float:
        """As ``run``, with every full-order step replayed from captured HIP graphs.  Whole chunks
        of steps (``chunk`` rounded up to a multiple of the SBDF ring period) are one graph each,
        so that the host launches one graph per chunk; the remaining steps use one graph per ring
        phase (one graph for RK).  Startup steps run eagerly; CN-family schemes step eagerly."""
        if self.cn is not None:
            return self.run(d, t0, nsteps, clock)
        if getattr(self._explicit, "wants_time", False):
            raise OptionError("Imex.run_graph: the explicit term depends on t (a host callback), which a "
                              "captured graph would freeze", fix="use run()")
        self._check_state(d)
        self._continuation(d, t0)
        t, done = self._graph_startup(d, float(t0), nsteps, clock)
        period = max(self.order, 1)
        chunk = period * max(1, -(-int(chunk) // period))
        while done < nsteps:
            if chunk > 1 and nsteps - done >= chunk and (not self.order or self.history_length == self.order):
                gk = (d.data_ptr(), tuple(self._levels) if self.order else "rk", "chunk", chunk)
                g = self._graphs.get(gk)
                if g is None:
                    g = self._capture_chunk(d, chunk)
                    self._graphs[gk] = g
                g.replay()
                for _ in range(chunk):
                    self._graph_step(d)[2]()
                    t = self._next_t(t, clock)
                done += chunk
                continue
            gk, fn, commit = self._graph_step(d)
            g = self._graphs.get(gk)
            if g is None:
                g = capture_graph(fn)
                self._graphs[gk] = g
            g.replay()
            commit()
            t = self._next_t(t, clock)
            done += 1
        self._remember(d, t)
        return t

    def _capture_chunk(self, d, chunk: int):
        """One graph of ``chunk`` consecutive full-order steps from the current ring phase (the
        bookkeeping is restored: the caller commits the steps after each replay)."""
        saved = (list(self._levels), self.iteration)
        fns = []
        for _ in range(chunk):
            _, fn, commit = self._graph_step(d)
            fns.append(fn)
            commit()
        self._levels, self.iteration = saved
        return capture_graph(lambda: [f() for f in fns])

    def startup_step(self, d, t: float) -> None:
        """One eager SBDF step of the state ``d`` at time ``t``, at the order the history allows,
        for an owner that keeps the trajectory itself (no continuation check, no host
        synchronisation; every factorization the step uses must exist already)."""
        if not self.order:
            raise OptionError("Imex.startup_step: SBDF schemes only")
        self._t_step = float(t)
        self._sbdf_step(d)
        self.iteration += 1

    def replay_step(self, d) -> None:
        """One full-order step of ``d`` from the captured graphs, for an owner that keeps the
        trajectory itself (no continuation check, no host synchronisation): the history must be
        complete (startup done, or levels adopted from other integrators)."""
        if self.cn is not None:
            raise OptionError("Imex.replay_step: CN-family schemes step eagerly (run)")
        if self.order and self.history_length < self.order - 1:
            raise OptionError("Imex.replay_step: the SBDF history is not complete")
        if self.order:
            a, _ = sbdf_coefficients(self.order)
            self._factor(a[0] / self.dt, 1.0)
        gk, fn, commit = self._graph_step(d)
        g = self._graphs.get(gk)
        if g is None:
            g = capture_graph(fn)
            self._graphs[gk] = g
        g.replay()
        commit()

    # -- graph steps (shared with ensemble sets that capture several integrators in one graph) --
    def _graph_startup(self, d, t: float, nsteps: int, clock=None):
        """Eager lower-order startup steps (SBDF), and every factorization the full-order steps use
        built before any capture; returns (t, steps done)."""
        done = 0
        if self.order:
            while self.history_length < self.order - 1 and done < nsteps:
                self._t_step = t
                self._sbdf_step(d)
                self.iteration += 1
                t = self._next_t(t, clock)
                done += 1
            if done < nsteps:
                a, _ = sbdf_coefficients(self.order)
                self._factor(a[0] / self.dt, 1.0)
        else:
            A, H, cc = self.tab
            for i in range(1, len(cc)):
                self._factor(1.0, self.dt * H[i][i])
        return t, done

    def _graph_step(self, d):
        """(key, launch, commit) of the next full-order step: the key names the ring phase, launch
        enqueues the step's kernels, commit advances the history bookkeeping once it ran."""
        key = d.data_ptr()
        if self.order:
            if len(self._levels) < self.order:
                used = set(self._levels)
                slot = next(i for i in range(self.order) if i not in used)
                levels = [slot] + self._levels
            else:
                levels = [self._levels[-1]] + self._levels[:-1]
            gk = (key, tuple(levels))
            fn = lambda lv=levels: self._sbdf_launch(d, lv)     # noqa: E731
        else:
            levels = None
            gk = (key, "rk")
            fn = lambda: self._rk_step(d)                       # noqa: E731

        def commit(lv=levels):
            if lv is not None:
                self._levels = list(lv)
            self.iteration += 1
        return gk, fn, commit

    def explain(self) -> str:
        return (f"hip.Imex(scheme={self.scheme}, dt={self.dt:g}, batch={self.batch}, n={self.n}, "
                f"F={'g*X+h' if (self._g is not None or self._hh is not None) else type(self._explicit).__name__ if self._explicit is not None else 'none'}, "
                f"factorizations={self.n_factorizations}); state resident on the device, fused "
                "stage kernel + batched bordered banded solve per stage")
