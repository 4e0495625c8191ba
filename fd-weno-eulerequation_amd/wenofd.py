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
K_GHOST, K_FLUX_X, K_FLUX_Y, K_SWITCH, K_DT, K_HYB, K_PATCH, K_COUNT = range(8)
KERNEL_NAMES = ("ghost_fill", "flux_x_rk", "flux_y", "switch", "delta_t", "hyb_stage", "hyb_patch")
SWITCH_OFF, SWITCH_MLP, SWITCH_LINEAR, SWITCH_WENO = range(4)
N_PARAMS = 1441
SWITCH_WEIGHTS = os.path.join(HERE, "switch", "switch_mlp.npz")
# Steps per graph replay given to every new Solver / Batch (None: the
# library's default of 16; 0 launches every step from the host).
DEFAULT_GRAPH_STEPS = None

# Exported symbols declared in include/wenofd_amd.h.
EXPORTS = (
    "wenofd_last_error", "wenofd_version", "wenofd_create", "wenofd_create_from_file", "wenofd_destroy",
    "wenofd_parse_config", "wenofd_initialize_ave", "wenofd_calculate_delta_t", "wenofd_run_rk",
    "wenofd_set_boundary", "wenofd_get_flux", "wenofd_get_rhs", "wenofd_assemble_boundary_term",
    "wenofd_assemble_source_term", "wenofd_step", "wenofd_advance", "wenofd_run",
    "wenofd_get_state", "wenofd_set_state", "wenofd_get_info", "wenofd_set_time", "wenofd_output_dir",
    "wenofd_output_ave", "wenofd_output_error", "wenofd_output_accuracy", "wenofd_accuracy",
    "wenofd_set_timing", "wenofd_get_timing", "wenofd_synchronize", "wenofd_probe_weno5",
    "wenofd_probe_physics", "wenofd_set_tiles", "wenofd_set_device_loop", "wenofd_delta_t_noalloc",
    "wenofd_batch_create", "wenofd_batch_destroy", "wenofd_batch_size", "wenofd_batch_initialize_ave",
    "wenofd_batch_advance", "wenofd_batch_set_graph_steps", "wenofd_batch_set_tiles", "wenofd_batch_get_tiles",
    "wenofd_batch_autotune", "wenofd_batch_get_state", "wenofd_batch_get_info", "wenofd_batch_accuracy",
    "wenofd_batch_synchronize",
    "wenofd_switch_set", "wenofd_switch_interval", "wenofd_switch_dilation", "wenofd_switch_classify", "wenofd_switch_flags", "wenofd_switch_stats", "wenofd_harvest",
    "wenofd_mlp_create", "wenofd_mlp_destroy", "wenofd_mlp_set_params", "wenofd_mlp_get_params",
    "wenofd_mlp_dataset_add", "wenofd_mlp_harvest", "wenofd_mlp_dataset_size", "wenofd_mlp_dataset_get",
    "wenofd_mlp_dataset_clear", "wenofd_mlp_grad", "wenofd_mlp_train", "wenofd_mlp_infer", "wenofd_mlp_evaluate",
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
        for n in ("wenofd_initialize_ave", "wenofd_set_boundary", "wenofd_get_flux", "wenofd_synchronize",
                  "wenofd_assemble_boundary_term", "wenofd_assemble_source_term"):
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
        L.wenofd_set_device_loop.argtypes = [vp, i]
        L.wenofd_delta_t_noalloc.argtypes = [vp, i, i, i, d, d, d, d, d, i, vp, vp]
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
        fp, u8, ll, ull = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_ubyte), ctypes.c_longlong, \
            ctypes.c_ulonglong
        llp = ctypes.POINTER(ll)
        L.wenofd_switch_set.argtypes = [vp, i, fp, ctypes.c_float]
        L.wenofd_switch_classify.argtypes = [vp, i, u8, u8, fp, fp]
        L.wenofd_switch_interval.argtypes = [vp, i]
        L.wenofd_switch_dilation.argtypes = [vp, i]
        L.wenofd_switch_flags.argtypes = [vp, u8, u8]
        L.wenofd_switch_stats.argtypes = [vp, llp]
        L.wenofd_harvest.argtypes = [vp, i, fp, u8]
        L.wenofd_mlp_create.argtypes = [i, ull, ctypes.POINTER(vp)]
        L.wenofd_mlp_destroy.argtypes = [vp]
        L.wenofd_mlp_set_params.argtypes = [vp, fp]
        L.wenofd_mlp_get_params.argtypes = [vp, fp]
        L.wenofd_mlp_dataset_add.argtypes = [vp, fp, u8, ll]
        L.wenofd_mlp_harvest.argtypes = [vp, vp, d, d, ull, llp]
        L.wenofd_mlp_dataset_size.argtypes = [vp, llp, llp]
        L.wenofd_mlp_dataset_get.argtypes = [vp, ll, ll, fp, u8]
        L.wenofd_mlp_dataset_clear.argtypes = [vp]
        L.wenofd_mlp_grad.argtypes = [vp, ctypes.POINTER(i), i, ctypes.c_float, fp, fp]
        L.wenofd_mlp_train.argtypes = [vp, i, i, ctypes.c_float, ctypes.c_float, ull, fp]
        L.wenofd_mlp_infer.argtypes = [vp, fp, ll, fp]
        L.wenofd_mlp_evaluate.argtypes = [vp, ll, ll, ctypes.c_float, llp]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != OK:
        raise WenofdError(rc, lib().wenofd_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u8ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))


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


def delta_t_noalloc(U_ptr, nx, ny, pitch, gamma, dx, dy, cfl, t_left, fp_contract, out_ptr, stream=None):
    """wenofd_delta_t_noalloc on device pointers (ints): {dt, nan} into out_ptr."""
    _check(lib().wenofd_delta_t_noalloc(ctypes.c_void_p(U_ptr), int(nx), int(ny), int(pitch), float(gamma),
                                        float(dx), float(dy), float(cfl), float(t_left), int(fp_contract),
                                        ctypes.c_void_p(out_ptr), ctypes.c_void_p(stream)))


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
        if DEFAULT_GRAPH_STEPS is not None:
            self.set_device_loop(DEFAULT_GRAPH_STEPS)

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

    def assemble_boundary_term(self):
        _check(lib().wenofd_assemble_boundary_term(self.h))

    def assemble_source_term(self):
        _check(lib().wenofd_assemble_source_term(self.h))

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

    def set_device_loop(self, graph_steps):
        """Steps per graph replay of advance()'s device loop; 0: one host-driven step at a time."""
        _check(lib().wenofd_set_device_loop(self.h, int(graph_steps)))

    def synchronize(self):
        _check(lib().wenofd_synchronize(self.h))

    # ---- learned switch
    def faces(self, direction):
        return self.ny * (self.nx + 1) if direction == 0 else (self.ny + 1) * self.nx

    def switch_set(self, mode, params=None, threshold=0.0):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32).reshape(N_PARAMS)
        _check(lib().wenofd_switch_set(self.h, int(mode), None if p is None else _fptr(p), float(threshold)))

    def switch_interval(self, steps):
        _check(lib().wenofd_switch_interval(self.h, int(steps)))

    def switch_dilation(self, radius):
        _check(lib().wenofd_switch_dilation(self.h, int(radius)))

    def switch_classify(self, dilate=True, logits=False):
        """(flags_x [NY][NX+1], flags_y [NY+1][NX], logits_x, logits_y) of the current state."""
        fx = np.empty((self.ny, self.nx + 1), np.uint8)
        fy = np.empty((self.ny + 1, self.nx), np.uint8)
        lx = np.empty((self.ny, self.nx + 1, 4), np.float32) if logits else None
        ly = np.empty((self.ny + 1, self.nx, 4), np.float32) if logits else None
        _check(lib().wenofd_switch_classify(self.h, 1 if dilate else 0, _u8ptr(fx), _u8ptr(fy),
                                            None if lx is None else _fptr(lx), None if ly is None else _fptr(ly)))
        return fx, fy, lx, ly

    def switch_flags(self):
        fx = np.empty((self.ny, self.nx + 1), np.uint8)
        fy = np.empty((self.ny + 1, self.nx), np.uint8)
        _check(lib().wenofd_switch_flags(self.h, _u8ptr(fx), _u8ptr(fy)))
        return fx, fy

    def switch_fallbacks(self):
        n = ctypes.c_longlong()
        _check(lib().wenofd_switch_stats(self.h, ctypes.byref(n)))
        return n.value

    def harvest(self, direction):
        nf = self.faces(direction)
        X = np.empty((nf, 4, 10), np.float32)
        Y = np.empty((nf, 4), np.uint8)
        _check(lib().wenofd_harvest(self.h, int(direction), _fptr(X), _u8ptr(Y)))
        return X, Y


def load_switch(path=SWITCH_WEIGHTS):
    """(params float32[1441], threshold) of a trained switch (switch_train.py)."""
    with np.load(path) as z:
        return np.asarray(z["params"], np.float32), float(z["threshold"])


class Mlp:
    """Switch MLP trainer (wenofd_mlp handle): parameters, Adam state, samples on the GPU."""

    def __init__(self, seed=0, device=0):
        h = ctypes.c_void_p()
        _check(lib().wenofd_mlp_create(device, ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().wenofd_mlp_destroy(self.h)
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

    @property
    def params(self):
        P = np.empty(N_PARAMS, np.float32)
        _check(lib().wenofd_mlp_get_params(self.h, _fptr(P)))
        return P

    @params.setter
    def params(self, P):
        P = np.ascontiguousarray(P, dtype=np.float32).reshape(N_PARAMS)
        _check(lib().wenofd_mlp_set_params(self.h, _fptr(P)))

    def add(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 10)
        Y = np.ascontiguousarray(Y, dtype=np.uint8).reshape(-1)
        _check(lib().wenofd_mlp_dataset_add(self.h, _fptr(X), _u8ptr(Y), X.shape[0]))

    def harvest(self, solver, keep_pos=1.0, keep_neg=1.0, seed=0):
        n = ctypes.c_longlong()
        _check(lib().wenofd_mlp_harvest(self.h, solver.h, float(keep_pos), float(keep_neg),
                                        ctypes.c_ulonglong(seed), ctypes.byref(n)))
        return n.value

    def size(self):
        n, p = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().wenofd_mlp_dataset_size(self.h, ctypes.byref(n), ctypes.byref(p)))
        return n.value, p.value

    def samples(self, start=0, n=None):
        if n is None:
            n = self.size()[0] - start
        X = np.empty((n, 10), np.float32)
        Y = np.empty(n, np.uint8)
        _check(lib().wenofd_mlp_dataset_get(self.h, int(start), int(n), _fptr(X), _u8ptr(Y)))
        return X, Y

    def clear(self):
        _check(lib().wenofd_mlp_dataset_clear(self.h))

    def grad(self, idx, pos_weight=1.0):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        G = np.empty(N_PARAMS, np.float32)
        loss = np.empty(1, np.float32)
        _check(lib().wenofd_mlp_grad(self.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), idx.size,
                                     float(pos_weight), _fptr(G), _fptr(loss)))
        return float(loss[0]), G

    def train(self, epochs, batch=16384, lr=3e-3, pos_weight=1.0, seed=0):
        out = np.empty(epochs, np.float32)
        _check(lib().wenofd_mlp_train(self.h, int(epochs), int(batch), float(lr), float(pos_weight),
                                      ctypes.c_ulonglong(seed), _fptr(out)))
        return out

    def infer(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 10)
        z = np.empty(X.shape[0], np.float32)
        _check(lib().wenofd_mlp_infer(self.h, _fptr(X), X.shape[0], _fptr(z)))
        return z

    def evaluate(self, start=0, n=None, threshold=0.0):
        """{TP, FP, FN, TN} of the deployed (bf16 MFMA) network."""
        if n is None:
            n = self.size()[0] - start
        c = (ctypes.c_longlong * 4)()
        _check(lib().wenofd_mlp_evaluate(self.h, int(start), int(n), float(threshold), c))
        return dict(zip(("tp", "fp", "fn", "tn"), (int(v) for v in c)))


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
        if DEFAULT_GRAPH_STEPS is not None:
            self.set_graph_steps(DEFAULT_GRAPH_STEPS)

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
