"""ctypes binding of libwenofd_amd.so (include/wenofd_amd.h) for tests and bench.

The library is the product: HIP kernels for gfx950 behind a C ABI. This module
only marshals arguments; it has no fallback — if the shared library is missing
or no GPU is present, calls fail loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# WENOFD_LIBRARY selects another build of the same library (A/B timing).
LIB_PATH = os.environ.get("WENOFD_LIBRARY") or os.path.join(HERE, "lib", "libwenofd_amd.so")

OK, E_CONFIG, E_BAD_ENUM, E_NAN_DT, E_HIP, E_ARG, E_IO = range(7)
LAYOUT_REF, LAYOUT_SOA, LAYOUT_REF_GHOSTED = 0, 1, 2
K_GHOST, K_FLUX_X, K_FLUX_Y, K_RK, K_DT, K_COUNT = range(6)
KERNEL_NAMES = ("ghost_fill", "flux_x_rk", "flux_y", "unused", "delta_t")

# Exported symbols declared in include/wenofd_amd.h.
EXPORTS = (
    "wenofd_last_error", "wenofd_version", "wenofd_create", "wenofd_create_from_file", "wenofd_destroy",
    "wenofd_parse_config", "wenofd_initialize_ave", "wenofd_calculate_delta_t", "wenofd_run_rk",
    "wenofd_set_boundary", "wenofd_get_flux", "wenofd_get_rhs", "wenofd_step", "wenofd_advance", "wenofd_run",
    "wenofd_get_state", "wenofd_set_state", "wenofd_get_info", "wenofd_set_time", "wenofd_output_dir",
    "wenofd_output_ave", "wenofd_output_error", "wenofd_output_accuracy", "wenofd_accuracy",
    "wenofd_set_timing", "wenofd_get_timing", "wenofd_synchronize", "wenofd_probe_weno5",
    "wenofd_probe_physics", "wenofd_set_tiles",
    "wenofd_batch_create", "wenofd_batch_destroy", "wenofd_batch_size", "wenofd_batch_initialize_ave",
    "wenofd_batch_advance", "wenofd_batch_set_graph_steps", "wenofd_batch_set_tiles", "wenofd_batch_get_tiles",
    "wenofd_batch_autotune", "wenofd_batch_get_state", "wenofd_batch_get_info", "wenofd_batch_accuracy",
    "wenofd_batch_synchronize",
)

_dp = ctypes.POINTER(ctypes.c_double)


class Info(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("cfl", "dx", "dy", "xL", "xR", "yL", "yR", "t_end", "now", "gamma", "zp_lambda")] + \
               [("steps", ctypes.c_longlong)] + \
               [(n, ctypes.c_int) for n in ("nx", "ny", "testcase", "scheme", "rkmeth", "device")]


class WenofdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"wenofd error {code}: {msg}")
        self.code = code


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                               "(make -C fd-weno-eulerequation_amd)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.wenofd_last_error.restype = ctypes.c_char_p
        L.wenofd_create.argtypes = [ctypes.c_char_p, i, ctypes.POINTER(vp)]
        L.wenofd_create_from_file.argtypes = [ctypes.c_char_p, i, ctypes.POINTER(vp)]
        L.wenofd_destroy.argtypes = [vp]
        L.wenofd_parse_config.argtypes = [ctypes.c_char_p, ctypes.POINTER(Info)]
        for n in ("wenofd_initialize_ave", "wenofd_set_boundary", "wenofd_get_flux", "wenofd_synchronize"):
            getattr(L, n).argtypes = [vp]
        L.wenofd_calculate_delta_t.argtypes = [vp, _dp]
        L.wenofd_run_rk.argtypes = [vp, i, d]
        L.wenofd_get_rhs.argtypes = [vp, _dp, i]
        L.wenofd_step.argtypes = [vp, d, _dp]
        L.wenofd_advance.argtypes = [vp, i, d, ctypes.POINTER(i)]
        L.wenofd_run.argtypes = [vp, i]
        L.wenofd_get_state.argtypes = [vp, _dp, i]
        L.wenofd_set_state.argtypes = [vp, _dp, i]
        L.wenofd_get_info.argtypes = [vp, ctypes.POINTER(Info)]
        L.wenofd_set_time.argtypes = [vp, d]
        L.wenofd_output_dir.argtypes = [vp, ctypes.c_char_p, i]
        L.wenofd_output_ave.argtypes = [vp, ctypes.c_char_p]
        L.wenofd_output_error.argtypes = [vp, ctypes.c_char_p]
        L.wenofd_output_accuracy.argtypes = [vp, ctypes.c_char_p, _dp]
        L.wenofd_accuracy.argtypes = [vp, _dp]
        L.wenofd_set_timing.argtypes = [vp, i]
        L.wenofd_get_timing.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_longlong), i]
        L.wenofd_probe_weno5.argtypes = [i, i, d, i, _dp, _dp]
        L.wenofd_probe_physics.argtypes = [i, d, i, _dp, _dp]
        L.wenofd_set_tiles.argtypes = [vp, i, i]
        ip = ctypes.POINTER(i)
        L.wenofd_batch_create.argtypes = [ctypes.POINTER(ctypes.c_char_p), i, i, ctypes.POINTER(vp)]
        for n in ("wenofd_batch_destroy", "wenofd_batch_initialize_ave", "wenofd_batch_synchronize"):
            getattr(L, n).argtypes = [vp]
        L.wenofd_batch_size.argtypes = [vp, ip]
        L.wenofd_batch_advance.argtypes = [vp, i, ip, ip]
        L.wenofd_batch_set_graph_steps.argtypes = [vp, i]
        L.wenofd_batch_set_tiles.argtypes = [vp, i, i]
        L.wenofd_batch_get_tiles.argtypes = [vp, ip, ip]
        L.wenofd_batch_autotune.argtypes = [vp, i, ip, ip, _dp]
        L.wenofd_batch_get_state.argtypes = [vp, i, _dp, i]
        L.wenofd_batch_get_info.argtypes = [vp, i, ctypes.POINTER(Info)]
        L.wenofd_batch_accuracy.argtypes = [vp, i, _dp]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != OK:
        raise WenofdError(rc, lib().wenofd_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_dp)


def config_text(**kw):
    return "".join(f"{k} = {v}\n" for k, v in kw.items())


def parse_config(text):
    info = Info()
    _check(lib().wenofd_parse_config(text.encode(), ctypes.byref(info)))
    return info


def probe_weno5(scheme, v5, zp_lambda=0.0, device=0):
    v5 = np.ascontiguousarray(v5, dtype=np.float64).reshape(-1, 5)
    out = np.empty(v5.shape[0])
    _check(lib().wenofd_probe_weno5(device, int(scheme), float(zp_lambda), v5.shape[0], _ptr(v5), _ptr(out)))
    return out


def probe_physics(states, gamma=1.4, device=0):
    s = np.ascontiguousarray(states, dtype=np.float64).reshape(-1, 4)
    out = np.empty((s.shape[0], 82))
    _check(lib().wenofd_probe_physics(device, float(gamma), s.shape[0], _ptr(s), _ptr(out)))
    return out


def vortex_batch_configs(n=64, seed=0, scheme=1, rkmeth=2, cfl=0.05, sizes=range(100, 401, 20), t_end=None):
    """Config texts of n seeded isentropic-vortex solves: N drawn uniformly
    from `sizes` (square grids), strength ~ U[3, 6], centre ~ U[3, 7]^2."""
    rng = np.random.default_rng(seed)
    sizes = list(sizes)
    out = []
    for _ in range(n):
        N = int(sizes[rng.integers(len(sizes))])
        keys = dict(TESTCASE=4, SCHEME=scheme, RKMETH=rkmeth, CFL="%.17g" % cfl, ELEMNUMX=N, ELEMNUMY=N,
                    VORTEX_EPS="%.17g" % rng.uniform(3.0, 6.0), VORTEX_XC="%.17g" % rng.uniform(3.0, 7.0),
                    VORTEX_YC="%.17g" % rng.uniform(3.0, 7.0))
        if t_end is not None:
            keys["TEND"] = "%.17g" % t_end
        out.append(config_text(**keys))
    return out


class Solver:
    """One solve on one GPU (wenofd_solver handle)."""

    def __init__(self, config=None, device=0, path=None, **keys):
        h = ctypes.c_void_p()
        if path is not None:
            _check(lib().wenofd_create_from_file(str(path).encode(), device, ctypes.byref(h)))
        else:
            text = config if config is not None else config_text(**keys)
            _check(lib().wenofd_create(text.encode(), device, ctypes.byref(h)))
        self.h = h
        inf = self.info()
        self.nx, self.ny = inf.nx, inf.ny

    def close(self):
        if getattr(self, "h", None):
            lib().wenofd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        inf = Info()
        _check(lib().wenofd_get_info(self.h, ctypes.byref(inf)))
        return inf

    def shape(self, layout):
        if layout == LAYOUT_REF:
            return (self.nx, self.ny, 4)
        if layout == LAYOUT_SOA:
            return (4, self.ny, self.nx)
        return (self.nx + 8, self.ny + 8, 4)

    def get_state(self, layout=LAYOUT_REF):
        out = np.empty(self.shape(layout))
        _check(lib().wenofd_get_state(self.h, _ptr(out), layout))
        return out

    def set_state(self, U, layout=LAYOUT_REF):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.shape(layout))
        _check(lib().wenofd_set_state(self.h, _ptr(U), layout))

    def initialize_ave(self):
        _check(lib().wenofd_initialize_ave(self.h))

    def set_boundary(self):
        _check(lib().wenofd_set_boundary(self.h))

    def get_flux(self):
        _check(lib().wenofd_get_flux(self.h))

    def get_rhs(self, layout=LAYOUT_REF):
        out = np.empty(self.shape(layout))
        _check(lib().wenofd_get_rhs(self.h, _ptr(out), layout))
        return out

    def calculate_delta_t(self):
        dt = ctypes.c_double()
        _check(lib().wenofd_calculate_delta_t(self.h, ctypes.byref(dt)))
        return dt.value

    def run_rk(self, rkmeth, dt):
        _check(lib().wenofd_run_rk(self.h, int(rkmeth), float(dt)))

    def step(self, dt_fixed=0.0):
        dt = ctypes.c_double()
        _check(lib().wenofd_step(self.h, float(dt_fixed), ctypes.byref(dt)))
        return dt.value

    def advance(self, max_steps=0, dt_fixed=0.0):
        n = ctypes.c_int()
        _check(lib().wenofd_advance(self.h, int(max_steps), float(dt_fixed), ctypes.byref(n)))
        return n.value

    def run(self, verbose=False):
        _check(lib().wenofd_run(self.h, 1 if verbose else 0))

    def set_time(self, now):
        _check(lib().wenofd_set_time(self.h, float(now)))

    def output_dir(self):
        buf = ctypes.create_string_buffer(4096)
        _check(lib().wenofd_output_dir(self.h, buf, 4096))
        return buf.value.decode()

    def output_ave(self, prefix):
        _check(lib().wenofd_output_ave(self.h, prefix.encode()))

    def output_error(self, prefix):
        _check(lib().wenofd_output_error(self.h, prefix.encode()))

    def output_accuracy(self, prefix):
        out = np.empty(3)
        _check(lib().wenofd_output_accuracy(self.h, prefix.encode(), _ptr(out)))
        return out

    def accuracy(self):
        out = np.empty(3)
        _check(lib().wenofd_accuracy(self.h, _ptr(out)))
        return out  # Linf, L1, L2 of rho

    def set_timing(self, enable=True):
        _check(lib().wenofd_set_timing(self.h, 1 if enable else 0))

    def get_timing(self):
        ms = np.zeros(K_COUNT)
        cnt = (ctypes.c_longlong * K_COUNT)()
        _check(lib().wenofd_get_timing(self.h, _ptr(ms), cnt, K_COUNT))
        return {KERNEL_NAMES[k]: (float(ms[k]), int(cnt[k])) for k in range(K_COUNT)}

    def set_tiles(self, x_faces, y_faces):
        _check(lib().wenofd_set_tiles(self.h, int(x_faces), int(y_faces)))

    def synchronize(self):
        _check(lib().wenofd_synchronize(self.h))


class Batch:
    """Many independent solves advanced together (wenofd_batch handle)."""

    def __init__(self, configs, device=0):
        texts = [c.encode() if isinstance(c, str) else config_text(**c).encode() for c in configs]
        arr = (ctypes.c_char_p * len(texts))(*texts)
        h = ctypes.c_void_p()
        _check(lib().wenofd_batch_create(arr, len(texts), device, ctypes.byref(h)))
        self.h = h
        self.n = len(texts)
        self.shapes = [(inf.nx, inf.ny) for inf in (self.info(s) for s in range(self.n))]

    def close(self):
        if getattr(self, "h", None):
            lib().wenofd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self, solve):
        inf = Info()
        _check(lib().wenofd_batch_get_info(self.h, int(solve), ctypes.byref(inf)))
        return inf

    def initialize_ave(self):
        _check(lib().wenofd_batch_initialize_ave(self.h))

    def advance(self, max_steps=0):
        """Returns (iterations launched, solves still running)."""
        done, active = ctypes.c_int(), ctypes.c_int()
        _check(lib().wenofd_batch_advance(self.h, int(max_steps), ctypes.byref(done), ctypes.byref(active)))
        return done.value, active.value

    def set_graph_steps(self, steps):
        _check(lib().wenofd_batch_set_graph_steps(self.h, int(steps)))

    def set_tiles(self, x_faces, y_faces):
        _check(lib().wenofd_batch_set_tiles(self.h, int(x_faces), int(y_faces)))

    def get_tiles(self):
        x, y = ctypes.c_int(), ctypes.c_int()
        _check(lib().wenofd_batch_get_tiles(self.h, ctypes.byref(x), ctypes.byref(y)))
        return x.value, y.value

    def autotune(self, steps=4):
        """Returns (x_faces, y_faces, ms per step of the chosen shape)."""
        x, y, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().wenofd_batch_autotune(self.h, int(steps), ctypes.byref(x), ctypes.byref(y), ctypes.byref(ms)))
        return x.value, y.value, ms.value

    def get_state(self, solve, layout=LAYOUT_REF):
        nx, ny = self.shapes[solve]
        shape = {LAYOUT_REF: (nx, ny, 4), LAYOUT_SOA: (4, ny, nx)}.get(layout, (nx + 8, ny + 8, 4))
        out = np.empty(shape)
        _check(lib().wenofd_batch_get_state(self.h, int(solve), _ptr(out), layout))
        return out

    def accuracy(self, solve):
        out = np.empty(3)
        _check(lib().wenofd_batch_accuracy(self.h, int(solve), _ptr(out)))
        return out

    def synchronize(self):
        _check(lib().wenofd_batch_synchronize(self.h))
