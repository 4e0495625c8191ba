"""ctypes wrapper over oracle/liboracle.so — the CPU oracle.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the measured or
shipped path. See wenofd_oracle.cpp for what it restates and how it is pinned.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

TESTCASES = {"SMOOTH": 0, "RIEMANN1": 1, "RIEMANN2": 2, "RTI": 3, "VORTEX": 4, "QUADRANT": 5}
SCHEMES = {"JS": 0, "Z": 1, "ZP": 2, "ZPI": 3}

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.oracle_weno5.restype = d
        L.oracle_weno5.argtypes = [i, d, d, d, d, d, d]
        L.oracle_weno5_batch.argtypes = [i, i, _dp, d, _dp]
        L.oracle_eigen.argtypes = [i, _dp, d, _dp, _dp]
        L.oracle_split.argtypes = [i, _dp, d, _dp, _dp]
        L.oracle_phys_flux.argtypes = [i, _dp, d, _dp]
        L.oracle_max_eigen.restype = d
        L.oracle_max_eigen.argtypes = [_dp, d, d, d]
        L.oracle_face_flux.argtypes = [i, i, d, d, _dp, _dp]
        L.oracle_seeded_quadrant.argtypes = [ctypes.c_ulonglong, _dp]
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [i, i, i, d, i, i, d, _dp, _dp]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_grid.argtypes = [vp, _dp, _dp, _dp]
        for n in ("oracle_get_full", "oracle_set_full", "oracle_get_interior", "oracle_set_interior",
                  "oracle_assemble_rhs"):
            getattr(L, n).argtypes = [vp, _dp]
        L.oracle_fill_ghosts.argtypes = [vp]
        L.oracle_delta_t.restype = d
        L.oracle_delta_t.argtypes = [vp]
        L.oracle_step.restype = i
        L.oracle_step.argtypes = [vp, i, d]
        L.oracle_run.restype = i
        L.oracle_run.argtypes = [vp, i, d]
        L.oracle_now.restype = d
        L.oracle_now.argtypes = [vp]
        L.oracle_accuracy.argtypes = [vp, _dp]
        L.oracle_count_ops.argtypes = [i, i, i, i, i, ctypes.POINTER(ctypes.c_ulonglong)]
        fp = ctypes.POINTER(ctypes.c_float)
        u8 = ctypes.POINTER(ctypes.c_ubyte)
        L.oracle_linear5.restype = d
        L.oracle_linear5.argtypes = [d, d, d, d, d]
        L.oracle_hybrid_face.argtypes = [i, i, d, d, i, _dp, _dp]
        L.oracle_linear_face.restype = d
        L.oracle_linear_face.argtypes = [_dp, _dp]
        L.oracle_stencil_features.argtypes = [_dp, fp]
        L.oracle_weno_z_departure.restype = d
        L.oracle_weno_z_departure.argtypes = [_dp]
        L.oracle_samples.argtypes = [vp, i, fp, u8]
        L.oracle_set_flags.argtypes = [vp, u8, u8]
        L.oracle_mlp_loss_grad.restype = d
        L.oracle_mlp_loss_grad.argtypes = [fp, fp, u8, i, d, _dp]
        L.oracle_mlp_loss_grad_d.restype = d
        L.oracle_mlp_loss_grad_d.argtypes = [_dp, fp, u8, i, d, _dp]
        L.oracle_mlp_forward.argtypes = [fp, fp, i, _dp]
        L.oracle_mlp_infer_bf16.argtypes = [fp, fp, i, fp]
        _LIB = L
    return _LIB


def weno5(scheme, v5, zp_lambda=0.0):
    v5 = np.ascontiguousarray(v5, dtype=np.float64).reshape(-1, 5)
    out = np.empty(v5.shape[0])
    lib().oracle_weno5_batch(int(scheme), v5.shape[0], _ptr(v5), float(zp_lambda), _ptr(out))
    return out


def eigen(direction, U, gamma=1.4):
    U = np.ascontiguousarray(U, dtype=np.float64)
    L = np.empty(16)
    R = np.empty(16)
    lib().oracle_eigen(int(direction), _ptr(U), gamma, _ptr(L), _ptr(R))
    return L.reshape(4, 4), R.reshape(4, 4)


def split(direction, U, gamma=1.4):
    U = np.ascontiguousarray(U, dtype=np.float64)
    Fm, Fp = np.empty(4), np.empty(4)
    lib().oracle_split(int(direction), _ptr(U), gamma, _ptr(Fm), _ptr(Fp))
    return Fm, Fp


def phys_flux(direction, U, gamma=1.4):
    U = np.ascontiguousarray(U, dtype=np.float64)
    F = np.empty(4)
    lib().oracle_phys_flux(int(direction), _ptr(U), gamma, _ptr(F))
    return F


def max_eigen(U, nx, ny, gamma=1.4):
    U = np.ascontiguousarray(U, dtype=np.float64)
    return lib().oracle_max_eigen(_ptr(U), float(nx), float(ny), gamma)


def face_flux(direction, scheme, U6, gamma=1.4, zp_lambda=0.0):
    U6 = np.ascontiguousarray(U6, dtype=np.float64).reshape(6, 4)
    out = np.empty(4)
    lib().oracle_face_flux(int(direction), int(scheme), gamma, zp_lambda, _ptr(U6), _ptr(out))
    return out


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u8ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))


N_PARAMS = 1441  # W1[32][10], b1[32], W2[32][32], b2[32], w3[32], b3


def linear5(a, b, c, d, e):
    return lib().oracle_linear5(float(a), float(b), float(c), float(d), float(e))


def hybrid_face(direction, scheme, mask, U6, gamma=1.4, zp_lambda=0.0):
    """Hybrid face flux from the conserved states of points f-3..f+2 ([6][4])."""
    U = np.ascontiguousarray(U6, dtype=np.float64).reshape(24)
    out = np.empty(4)
    lib().oracle_hybrid_face(int(direction), int(scheme), float(gamma), float(zp_lambda), int(mask), _ptr(U),
                             _ptr(out))
    return out


def linear_face(fp5, fm5):
    """The cheap flux of one component: F+ at points f-3..f+1, F- at f-2..f+2."""
    a = np.ascontiguousarray(fp5, dtype=np.float64)
    b = np.ascontiguousarray(fm5, dtype=np.float64)
    return lib().oracle_linear_face(_ptr(a), _ptr(b))


def stencil_features(w5):
    w = np.ascontiguousarray(w5, dtype=np.float64)
    out = np.empty(5, dtype=np.float32)
    lib().oracle_stencil_features(_ptr(w), _fptr(out))
    return out


def weno_z_departure(w5):
    w = np.ascontiguousarray(w5, dtype=np.float64)
    return lib().oracle_weno_z_departure(_ptr(w))


def mlp_loss_grad(params, X, Y, pos_weight=1.0):
    """Mean weighted BCE loss and its gradient (FP64) at float or double params."""
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 10)
    Y = np.ascontiguousarray(Y, dtype=np.uint8).reshape(-1)
    G = np.empty(N_PARAMS)
    if np.asarray(params).dtype == np.float64:
        P = np.ascontiguousarray(params, dtype=np.float64)
        loss = lib().oracle_mlp_loss_grad_d(_ptr(P), _fptr(X), _u8ptr(Y), X.shape[0], float(pos_weight), _ptr(G))
    else:
        P = np.ascontiguousarray(params, dtype=np.float32)
        loss = lib().oracle_mlp_loss_grad(_fptr(P), _fptr(X), _u8ptr(Y), X.shape[0], float(pos_weight), _ptr(G))
    return loss, G


def mlp_forward(params, X):
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 10)
    z = np.empty(X.shape[0])
    lib().oracle_mlp_forward(_fptr(P), _fptr(X), X.shape[0], _ptr(z))
    return z


def mlp_infer_bf16(params, X):
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 10)
    z = np.empty(X.shape[0], dtype=np.float32)
    lib().oracle_mlp_infer_bf16(_fptr(P), _fptr(X), X.shape[0], _fptr(z))
    return z


def seeded_quadrant(seed):
    out = np.empty(16)
    lib().oracle_seeded_quadrant(ctypes.c_ulonglong(seed), _ptr(out))
    return out.reshape(4, 4)


def count_ops(testcase, scheme, rk, nx, ny):
    out = (ctypes.c_ulonglong * 5)()
    lib().oracle_count_ops(int(testcase), int(scheme), int(rk), int(nx), int(ny), out)
    return dict(zip(("add", "mul", "div", "sqrt", "cmp"), list(out)))


class Solver:
    """One oracle solve; mirrors the reference CWENOFD loop on a whole domain."""

    def __init__(self, testcase, scheme, rk, cfl, nx, ny, t_end=-1.0, quad=None, vortex=None):
        self.nx, self.ny = nx, ny
        tc = TESTCASES[testcase] if isinstance(testcase, str) else int(testcase)
        sc = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        q = None if quad is None else np.ascontiguousarray(quad, dtype=np.float64).reshape(16)
        v = None if vortex is None else np.ascontiguousarray(vortex, dtype=np.float64).reshape(3)
        self._q, self._v = q, v
        self.h = lib().oracle_create(tc, sc, int(rk), float(cfl), nx, ny, float(t_end),
                                     None if q is None else _ptr(q), None if v is None else _ptr(v))
        if not self.h:
            raise ValueError("oracle_create rejected the arguments")
        self.rk = int(rk)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def grid(self):
        dxdy = np.empty(3)
        xc = np.empty(self.nx + 8)
        yc = np.empty(self.ny + 8)
        lib().oracle_grid(self.h, _ptr(dxdy), _ptr(xc), _ptr(yc))
        return dxdy, xc, yc

    def get_state(self):
        out = np.empty((self.nx, self.ny, 4))
        lib().oracle_get_interior(self.h, _ptr(out))
        return out

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.nx, self.ny, 4)
        lib().oracle_set_interior(self.h, _ptr(U))

    def get_full(self):
        out = np.empty((self.nx + 8, self.ny + 8, 4))
        lib().oracle_get_full(self.h, _ptr(out))
        return out

    def set_full(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.nx + 8, self.ny + 8, 4)
        lib().oracle_set_full(self.h, _ptr(U))

    def fill_ghosts(self):
        lib().oracle_fill_ghosts(self.h)

    def assemble_rhs(self):
        out = np.empty((self.nx, self.ny, 4))
        lib().oracle_assemble_rhs(self.h, _ptr(out))
        return out

    def delta_t(self):
        return lib().oracle_delta_t(self.h)

    def step(self, dt, rk=None):
        rc = lib().oracle_step(self.h, self.rk if rk is None else int(rk), float(dt))
        if rc != 0:
            raise ValueError("bad RK method")

    def run(self, max_steps=0, dt_fixed=0.0):
        return lib().oracle_run(self.h, int(max_steps), float(dt_fixed))

    def now(self):
        return lib().oracle_now(self.h)

    def samples(self, direction):
        """Ghost fill, then switch features [faces, 4, 10] and labels [faces, 4]
        of every face of one direction (x: NY x (NX+1) faces, y: (NY+1) x NX)."""
        nf = self.ny * (self.nx + 1) if direction == 0 else (self.ny + 1) * self.nx
        X = np.empty((nf, 4, 10), dtype=np.float32)
        Y = np.empty((nf, 4), dtype=np.uint8)
        lib().oracle_samples(self.h, int(direction), _fptr(X), _u8ptr(Y))
        return X, Y

    def set_flags(self, flags_x, flags_y):
        """Hybrid flux (hybrid_face): per face a mask of flagged characteristic
        fields (bit k = field k); zero takes the cheap flux alone. None restores
        the plain solver."""
        if flags_x is None:
            self._flags = None
            lib().oracle_set_flags(self.h, None, None)
            return
        fx = np.ascontiguousarray(flags_x, dtype=np.uint8).reshape(self.ny, self.nx + 1)
        fy = np.ascontiguousarray(flags_y, dtype=np.uint8).reshape(self.ny + 1, self.nx)
        self._flags = (fx, fy)  # kept alive while the oracle holds the pointers
        lib().oracle_set_flags(self.h, _u8ptr(fx), _u8ptr(fy))

    def accuracy(self):
        out = np.empty(3)
        lib().oracle_accuracy(self.h, _ptr(out))
        return out  # Linf, L1, L2 of rho
