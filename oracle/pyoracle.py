"""ctypes wrapper around the oracle libraries built from oracle/rf_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the rigidfibers_amd package.

``Oracle(64)`` is the FP64 restatement of the reference Fortran timestep
(fortran/fibers.f95:182-349); ``Oracle(32)`` is the same code in FP32 with
OpenMP, the CPU baseline.  Arrays are numpy, fiber vectors shaped [3*M], the
system matrix column-major (Fortran order) n x n.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class Params(C.Structure):
    _fields_ = [
        ("M", C.c_int),
        ("N", C.c_int),
        ("nqi", C.c_int),
        ("threads", C.c_int),
        ("slenderness", C.c_double),
        ("dt", C.c_double),
        ("force", C.c_double * 3),
    ]


def make_params(M, N=5, nqi=8, slenderness=0.01, dt=0.1, force=(0.0, 0.0, -1.0), threads=0):
    p = Params()
    p.M, p.N, p.nqi, p.threads = int(M), int(N), int(nqi), int(threads)
    p.slenderness, p.dt = float(slenderness), float(dt)
    for a in range(3):
        p.force[a] = float(force[a])
    return p


class Oracle:
    def __init__(self, bits=64):
        if bits not in (32, 64):
            raise ValueError("bits must be 32 or 64")
        path = os.path.join(_HERE, f"librf_oracle{bits}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = C.CDLL(path)
        self.dtype = np.float64 if bits == 64 else np.float32
        self.bits = bits
        real = C.c_double if bits == 64 else C.c_float
        P = C.POINTER(real)
        self._P = P
        lib = self.lib
        lib.rfo_tables.argtypes = [C.c_int, C.c_int, C.c_double] + [C.POINTER(C.c_double)] * 5
        lib.rfo_assemble.argtypes = [C.POINTER(Params), C.c_int, P, P, P, P]
        lib.rfo_solve_direct.argtypes = [C.c_int, P, P]
        lib.rfo_solve_gmres.argtypes = [C.c_int, P, P, P, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.POINTER(C.c_int), C.POINTER(C.c_double)]
        lib.rfo_velocities.argtypes = [C.POINTER(Params), C.c_int, P, P, P, P, P]
        lib.rfo_update_first.argtypes = [C.POINTER(Params)] + [P] * 6
        lib.rfo_update.argtypes = [C.POINTER(Params)] + [P] * 10
        lib.rfo_update.restype = None
        lib.rfo_update_first.restype = None
        lib.rfo_moments.argtypes = [real, real, C.c_int, C.c_int, P, P, P]
        lib.rfo_moments.restype = None
        lib.rfo_stokes_kernel.argtypes = [P, C.c_double, P]
        lib.rfo_stokes_kernel.restype = None
        lib.rfo_assemble_rows.argtypes = [C.POINTER(Params), C.c_int, P, P, C.c_int, C.c_int, P, P]
        lib.rfo_velocities_rows.argtypes = [C.POINTER(Params), C.c_int, P, P, P, C.c_int, C.c_int, P, P]
        lib.rfo_matvec_rows.argtypes = [C.c_int, C.c_int, P, P, P]
        lib.rfo_matvec_rows.restype = None
        assert lib.rfo_real_size() == (8 if bits == 64 else 4)

    # -- helpers -------------------------------------------------------------
    def _arr(self, a, copy=True):
        a = np.array(a, dtype=self.dtype, copy=copy, order="F")
        return a

    def _ptr(self, a):
        assert a.dtype == self.dtype and (a.flags.f_contiguous or a.flags.c_contiguous)
        return a.ctypes.data_as(self._P)

    # -- tables --------------------------------------------------------------
    @staticmethod
    def tables(nqi=8, N=5, eps=0.01):
        """(points[Q], weights[Q], legendre[N,Q] = P_{k+1}(s_q), lambda[N], eigen[N]) in FP64."""
        lib = C.CDLL(os.path.join(_HERE, "librf_oracle64.so"))
        lib.rfo_tables.argtypes = [C.c_int, C.c_int, C.c_double] + [C.POINTER(C.c_double)] * 5
        Q = 3 * nqi
        qp, qw = np.zeros(Q), np.zeros(Q)
        leg, lam, eig = np.zeros(N * Q), np.zeros(N), np.zeros(N)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        rc = lib.rfo_tables(nqi, N, eps, dp(qp), dp(qw), dp(leg), dp(lam), dp(eig))
        if rc:
            raise ValueError("invalid table parameters")
        return qp, qw, leg.reshape(N, Q), lam, eig

    # -- phases ----------------------------------------------------------------
    def assemble(self, params, pos, ori, analytical=False):
        M, N = params.M, params.N
        n = 3 * N * M
        pos, ori = self._arr(pos), self._arr(ori)
        A = np.zeros((n, n), dtype=self.dtype, order="F")
        b = np.zeros(n, dtype=self.dtype)
        rc = self.lib.rfo_assemble(C.byref(params), int(bool(analytical)), self._ptr(pos), self._ptr(ori),
                                   self._ptr(A), self._ptr(b))
        if rc:
            raise ValueError("rfo_assemble failed")
        return A, b

    def solve_direct(self, A, b):
        A = np.array(A, dtype=self.dtype, order="F", copy=True)
        x = self._arr(b)
        info = self.lib.rfo_solve_direct(A.shape[0], self._ptr(A), self._ptr(x))
        if info:
            raise np.linalg.LinAlgError(f"singular pivot {info}")
        return x

    def solve_gmres(self, A, b, restart=10, maxit=1000, tol=1e-5, x0=None):
        """x0=None starts from zero (shipped Fortran); otherwise from x0."""
        A = np.asfortranarray(A, dtype=self.dtype)
        b = self._arr(b)
        x = np.zeros_like(b) if x0 is None else self._arr(x0)
        it, rr = C.c_int(0), C.c_double(0.0)
        status = self.lib.rfo_solve_gmres(A.shape[0], self._ptr(A), self._ptr(b), self._ptr(x),
                                          0 if x0 is None else 1, restart, maxit, tol, C.byref(it), C.byref(rr))
        return x, {"status": status, "iterations": it.value, "relres": rr.value}

    def velocities(self, params, pos, ori, x, analytical=False):
        pos, ori, x = self._arr(pos), self._arr(ori), self._arr(x)
        v = np.zeros(3 * params.M, dtype=self.dtype)
        w = np.zeros(3 * params.M, dtype=self.dtype)
        if self.lib.rfo_velocities(C.byref(params), int(bool(analytical)), self._ptr(pos), self._ptr(ori), self._ptr(x), self._ptr(v),
                                   self._ptr(w)):
            raise ValueError("rfo_velocities failed")
        return v, w

    def update_first(self, params, pos, ori, v, w):
        pos, ori, v, w = (self._arr(a) for a in (pos, ori, v, w))
        pn, on = np.zeros_like(pos), np.zeros_like(ori)
        self.lib.rfo_update_first(C.byref(params), *(self._ptr(a) for a in (pos, ori, v, w, pn, on)))
        return pn, on

    def update(self, params, pos_prev, pos, ori_prev, ori, v_prev, v, w_prev, w):
        arrs = [self._arr(a) for a in (pos_prev, pos, ori_prev, ori, v_prev, v, w_prev, w)]
        pn, on = np.zeros_like(arrs[1]), np.zeros_like(arrs[3])
        self.lib.rfo_update(C.byref(params), *(self._ptr(a) for a in arrs + [pn, on]))
        return pn, on

    # -- row samples (CPU-baseline timing) ---------------------------------------
    def assemble_rows(self, params, pos, ori, i0, i1, analytical=False):
        n = 3 * params.N * params.M
        rows = 3 * params.N * (i1 - i0)
        pos, ori = self._arr(pos), self._arr(ori)
        A = np.zeros((rows, n), dtype=self.dtype, order="F")
        b = np.zeros(rows, dtype=self.dtype)
        if self.lib.rfo_assemble_rows(C.byref(params), int(bool(analytical)), self._ptr(pos), self._ptr(ori), i0, i1,
                                      self._ptr(A), self._ptr(b)):
            raise ValueError("rfo_assemble_rows failed")
        return A, b

    def velocities_rows(self, params, pos, ori, x, i0, i1, analytical=False):
        pos, ori, x = self._arr(pos), self._arr(ori), self._arr(x)
        v = np.zeros(3 * (i1 - i0), dtype=self.dtype)
        w = np.zeros_like(v)
        if self.lib.rfo_velocities_rows(C.byref(params), int(bool(analytical)), self._ptr(pos), self._ptr(ori),
                                        self._ptr(x), i0, i1, self._ptr(v), self._ptr(w)):
            raise ValueError("rfo_velocities_rows failed")
        return v, w

    def matvec_rows(self, A_rows, x):
        A_rows = np.asfortranarray(A_rows, dtype=self.dtype)
        x = self._arr(x)
        y = np.zeros(A_rows.shape[0], dtype=self.dtype)
        self.lib.rfo_matvec_rows(A_rows.shape[0], A_rows.shape[1], self._ptr(A_rows), self._ptr(x), self._ptr(y))
        return y

    def moments(self, b, l, nfwd, nkeep):
        I1 = np.zeros(nkeep + 4, dtype=self.dtype)
        I3, I5 = np.zeros_like(I1), np.zeros_like(I1)
        self.lib.rfo_moments(b, l, nfwd, nkeep, self._ptr(I1), self._ptr(I3), self._ptr(I5))
        return I1, I3, I5

    def stokes_kernel(self, R, eps):
        R = self._arr(R)
        K = np.zeros(6, dtype=self.dtype)
        self.lib.rfo_stokes_kernel(self._ptr(R), eps, self._ptr(K))
        return K

    # -- whole timesteps -----------------------------------------------------------
    def simulate(self, params, pos, ori, steps, solver="direct", analytical=False, gmres=(10, 1000, 1e-5),
                 record=None, analytical_velocities=None, warm_start=True):
        """Runs ``steps`` timesteps as fortran/fibers.f95:182-349 does.

        Returns the final (pos, ori).  If ``record`` is a list, a dict per step
        with b, x, v, w, and the updated pos/ori is appended (the quantities the
        reference dumps under VALIDATE, fortran/fibers.f95:205-303).
        ``analytical_velocities`` defaults to ``analytical``; ``warm_start``
        starts GMRES from the previous step's solution."""
        if analytical_velocities is None:
            analytical_velocities = analytical
        cur_p, cur_o = self._arr(pos), self._arr(ori)
        prev_p = prev_o = prev_v = prev_w = x = None
        for step in range(steps):
            A, b = self.assemble(params, cur_p, cur_o, analytical)
            if solver == "direct":
                x = self.solve_direct(A, b)
                info = {}
            else:
                x, info = self.solve_gmres(A, b, *gmres, x0=x if warm_start else None)
            v, w = self.velocities(params, cur_p, cur_o, x, analytical_velocities)
            if step == 0:
                nxt_p, nxt_o = self.update_first(params, cur_p, cur_o, v, w)
            else:
                nxt_p, nxt_o = self.update(params, prev_p, cur_p, prev_o, cur_o, prev_v, v, prev_w, w)
            if record is not None:
                record.append({"b": b, "x": x, "v": v, "w": w, "pos": nxt_p, "ori": nxt_o, "solver": info})
            prev_p, prev_o, prev_v, prev_w = cur_p, cur_o, v, w
            cur_p, cur_o = nxt_p, nxt_o
        return cur_p, cur_o
