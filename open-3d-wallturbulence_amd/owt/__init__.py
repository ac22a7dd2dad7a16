"""Python binding of libowt (the MI355X channel-flow time step) over its C ABI.

The call surface mirrors the reference solver's free functions
(reference src/setup.hpp:84-114, call order src/main.cpp:63-87):

    ch = Channel("case.toml")        # load_params, allocate_var, init_mat, init_fft
    ch.init_fields(mode=1)            # init_fields
    ch.step(n)                        # time_Integration x n
    ch.field("u")                     # the wave-space field, [NZ][NXH][NY] complex
    ch.wall_shear()                   # fx_fric, bottom, upper (src/solver.cpp:314)
    ch.save("DATA")                   # save_instant_field (src/IO.cpp:198-283), HDF5

Every call goes to the HIP implementation in lib/libowt.so; there is no CPU
fallback, and a missing library raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libowt.so")

FIELDS = {"u": 0, "v": 1, "w": 2, "omx": 3, "omy": 4, "omz": 5,
          "lmbx": 6, "lmby": 7, "lmbz": 8, "Hv": 9, "Hg": 10, "p": 11}
PLAIN, BATCHED = 0, 1
FP64, MIXED = 0, 1
KERNELS = {"solve": 0, "horizontal": 1, "mean": 2, "z_inverse": 3, "x_lamb": 4, "z_forward": 5, "nonlinear": 6,
           "factor": 7, "helmholtz": 8}

# every symbol include/owt.h declares
EXPORTS = [
    "owt_create", "owt_create_params", "owt_destroy", "owt_last_error", "owt_get_sizes",
    "owt_init_fields", "owt_set_velocity", "owt_step", "owt_wall_shear", "owt_diag",
    "owt_steps_done", "owt_get_field", "owt_set_method", "owt_set_precision",
    "owt_checkpoint", "owt_restore", "owt_f2p", "owt_p2f", "owt_lamb", "owt_get_matrix",
    "owt_time_steps", "owt_time_kernel", "owt_get_factor", "owt_save",
    "owt_ensemble_create", "owt_ensemble_destroy", "owt_ensemble_last_error", "owt_ensemble_member",
    "owt_ensemble_init_fields", "owt_ensemble_step", "owt_ensemble_time_steps", "owt_ensemble_launches",
    "owt_ensemble_autotune",
    "owt_tangent_init", "owt_tangent_handle", "owt_tangent_step", "owt_adjoint_run", "owt_adjoint_apply",
    "owt_autotune", "owt_get_tune", "owt_set_tune", "owt_save_fields", "owt_statistics",
    "owt_f2p_noalloc", "owt_get_run_params",
]


class OwtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[owt status {code}] {msg}")
        self.code = code


class _Opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("precision", ctypes.c_int),
                ("method", ctypes.c_int), ("verbose", ctypes.c_int)]


class _RunParams(ctypes.Structure):
    _fields_ = [("NT", ctypes.c_long), ("save_interval", ctypes.c_long), ("init_mode", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("amp", ctypes.c_double), ("restart", ctypes.c_char * 1024)]


class _Params(ctypes.Structure):
    _fields_ = [("nu", ctypes.c_double), ("dt", ctypes.c_double), ("dkx", ctypes.c_double),
                ("dkz", ctypes.c_double), ("alpha_mesh", ctypes.c_double),
                ("NX", ctypes.c_int), ("NY", ctypes.c_int), ("NZ", ctypes.c_int),
                ("precision", ctypes.c_int), ("method", ctypes.c_int)]


_lib = None


def lib():
    """Load libowt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libowt.so not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        PP = ctypes.POINTER(ctypes.c_void_p)
        L.owt_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Opts), PP]
        L.owt_create_params.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(_Opts), PP]
        L.owt_destroy.argtypes = [P]
        L.owt_destroy.restype = None
        L.owt_last_error.argtypes = [P]
        L.owt_last_error.restype = ctypes.c_char_p
        L.owt_get_sizes.argtypes = [P, P]
        L.owt_init_fields.argtypes = [P, I, ctypes.c_uint64, D]
        L.owt_set_velocity.argtypes = [P, P, P, P]
        L.owt_step.argtypes = [P, I]
        L.owt_wall_shear.argtypes = [P, P, P, P]
        L.owt_diag.argtypes = [P, P, I]
        L.owt_steps_done.argtypes = [P]
        L.owt_steps_done.restype = ctypes.c_long
        L.owt_get_field.argtypes = [P, I, P]
        L.owt_set_method.argtypes = [P, I]
        L.owt_set_precision.argtypes = [P, I]
        L.owt_save.argtypes = [P, ctypes.c_char_p]
        L.owt_save_fields.argtypes = [P, ctypes.c_char_p, I]
        L.owt_statistics.argtypes = [P, P, P]
        L.owt_checkpoint.argtypes = [P, ctypes.c_char_p]
        L.owt_restore.argtypes = [P, ctypes.c_char_p]
        L.owt_f2p.argtypes = [P, P, I, P]
        L.owt_p2f.argtypes = [P, P, I, P]
        L.owt_f2p_noalloc.argtypes = [P, P, I, P]
        L.owt_get_run_params.argtypes = [P, ctypes.POINTER(_RunParams)]
        L.owt_lamb.argtypes = [P, P, P]
        L.owt_get_matrix.argtypes = [P, I, P]
        L.owt_time_steps.argtypes = [P, I, P]
        L.owt_time_kernel.argtypes = [P, I, I, P]
        L.owt_get_factor.argtypes = [P, I, I, P, P]
        L.owt_autotune.argtypes = [P, I, ctypes.c_char_p, I]
        L.owt_get_tune.argtypes = [P, P]
        L.owt_set_tune.argtypes = [P, P]
        L.owt_tangent_init.argtypes = [P, P, P, P]
        L.owt_tangent_handle.argtypes = [P]
        L.owt_tangent_handle.restype = P
        L.owt_tangent_step.argtypes = [P, I]
        L.owt_adjoint_run.argtypes = [P, I, I, I, I, I, P, P, P, P]
        L.owt_adjoint_apply.argtypes = [P, I, I, P, P, P, P, P, P]
        L.owt_ensemble_create.argtypes = [I, P, ctypes.POINTER(_Opts), PP]
        L.owt_ensemble_destroy.argtypes = [P]
        L.owt_ensemble_destroy.restype = None
        L.owt_ensemble_last_error.argtypes = [P]
        L.owt_ensemble_last_error.restype = ctypes.c_char_p
        L.owt_ensemble_member.argtypes = [P, I]
        L.owt_ensemble_member.restype = P
        L.owt_ensemble_init_fields.argtypes = [P, I, P, D]
        L.owt_ensemble_step.argtypes = [P, I]
        L.owt_ensemble_time_steps.argtypes = [P, I, P]
        L.owt_ensemble_launches.argtypes = [P, P]
        L.owt_ensemble_autotune.argtypes = [P, I]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Channel:
    """One channel on one GPU (a handle of the C ABI)."""

    def __init__(self, case=None, *, params=None, device=0, precision=-1, method=-1, verbose=False, _member=None):
        L = lib()
        self._owner = None
        if _member is not None:
            # a member of an Ensemble: borrowed handle, kept alive by the ensemble
            self._h, self._owner = _member
        else:
            opts = _Opts(device, precision, method, 1 if verbose else 0)
            h = ctypes.c_void_p()
            if case is not None:
                rc = L.owt_create(os.fsencode(case), ctypes.byref(opts), ctypes.byref(h))
            else:
                prm = _params_struct(params)
                rc = L.owt_create_params(ctypes.byref(prm), ctypes.byref(opts), ctypes.byref(h))
            if rc:
                raise OwtError(rc, L.owt_last_error(None).decode())
            self._h = h
        sz = np.zeros(7, dtype=np.int32)
        L.owt_get_sizes(self._h, _p(sz))
        self.NX, self.NY, self.NZ, self.NXH, self.NZH, self.NX2, self.NZ2 = (int(v) for v in sz)

    def close(self):
        if getattr(self, "_h", None):
            if self._owner is None:
                lib().owt_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc:
            raise OwtError(rc, lib().owt_last_error(self._h).decode())

    @property
    def wave_shape(self):
        return (self.NZ, self.NXH, self.NY)

    @property
    def phys_shape(self):
        return (self.NZ2, self.NX2, self.NY)

    @property
    def gridpoints(self):
        return self.NX * self.NY * self.NZ

    def run_params(self):
        """NT, save_interval and the initial-field choice of the case file (init_mode
        1 reference field, 2 seeded, 3 laminar, 4 restart from `restart`)."""
        r = _RunParams()
        self._check(lib().owt_get_run_params(self._h, ctypes.byref(r)))
        return {"NT": r.NT, "save_interval": r.save_interval, "init_mode": r.init_mode, "seed": r.seed,
                "amp": r.amp, "restart": r.restart.decode()}

    def init_fields(self, mode=1, seed=0, amp=0.15):
        self._check(lib().owt_init_fields(self._h, int(mode), int(seed), float(amp)))

    def set_velocity(self, u, v, w):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128).reshape(self.wave_shape) for a in (u, v, w)]
        self._check(lib().owt_set_velocity(self._h, *[_p(a) for a in arrs]))

    def step(self, n=1):
        self._check(lib().owt_step(self._h, int(n)))

    def field(self, name):
        out = np.empty(self.wave_shape, dtype=np.complex128)
        self._check(lib().owt_get_field(self._h, FIELDS[name], _p(out)))
        return out

    def wall_shear(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().owt_wall_shear(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def diag(self, max_rows=1 << 16):
        out = np.empty((max_rows, 3), dtype=np.float64)
        n = lib().owt_diag(self._h, _p(out), int(max_rows))
        if n < 0:
            self._check(-n)
        return out[:n].copy()

    def steps_done(self):
        n = lib().owt_steps_done(self._h)
        if n < 0:
            self._check(int(-n))
        return n

    def set_method(self, method):
        self._check(lib().owt_set_method(self._h, int(method)))

    def set_precision(self, precision):
        self._check(lib().owt_set_precision(self._h, int(precision)))

    def save(self, directory="DATA"):
        """save_instant_field: HDF5 snapshot <directory>/F-<step, 8 digits>.H5 of u, v, w;
        returns its path."""
        self._check(lib().owt_save(self._h, os.fsencode(directory)))
        return os.path.join(directory, "F-%08d.H5" % self.steps_done())

    def save_fields(self, directory="DATA", what=1):
        """Snapshot with a dataset mask (1 velocity, 2 vorticity, 4 Lamb vector in
        wave space; 8, 16, 32 the same in physical space); returns its path."""
        self._check(lib().owt_save_fields(self._h, os.fsencode(directory), int(what)))
        return os.path.join(directory, "F-%08d.H5" % self.steps_done())

    def statistics(self):
        """Plane statistics: dict of NY-profiles U, V, W, uu, vv, ww, uv, uw, vw, and Re_tau."""
        prof = np.empty((9, self.NY), dtype=np.float64)
        rt = ctypes.c_double()
        self._check(lib().owt_statistics(self._h, _p(prof), ctypes.byref(rt)))
        out = dict(zip(("U", "V", "W", "uu", "vv", "ww", "uv", "uw", "vw"), prof))
        out["re_tau"] = rt.value
        return out

    def checkpoint(self, path):
        self._check(lib().owt_checkpoint(self._h, os.fsencode(path)))

    def restore(self, path):
        self._check(lib().owt_restore(self._h, os.fsencode(path)))

    def f2p(self, wave):
        wave = np.ascontiguousarray(wave, dtype=np.complex128)
        nf = wave.size // (self.NZ * self.NXH * self.NY)
        out = np.empty((nf,) + self.phys_shape, dtype=np.float64)
        self._check(lib().owt_f2p(self._h, _p(wave), nf, _p(out)))
        return out if nf > 1 else out[0]

    def f2p_noalloc(self, wave_dev, nf, phys_dev):
        """owt_f2p on device buffers (addresses as ints), allocating nothing."""
        self._check(lib().owt_f2p_noalloc(self._h, ctypes.c_void_p(wave_dev), int(nf), ctypes.c_void_p(phys_dev)))

    def p2f(self, phys):
        phys = np.ascontiguousarray(phys, dtype=np.float64)
        nf = phys.size // (self.NZ2 * self.NX2 * self.NY)
        out = np.empty((nf,) + self.wave_shape, dtype=np.complex128)
        self._check(lib().owt_p2f(self._h, _p(phys), nf, _p(out)))
        return out if nf > 1 else out[0]

    def lamb(self, six):
        six = np.ascontiguousarray(six, dtype=np.complex128).reshape((6,) + self.wave_shape)
        out = np.empty((3,) + self.wave_shape, dtype=np.complex128)
        self._check(lib().owt_lamb(self._h, _p(six), _p(out)))
        return out

    def matrix(self, which):
        idx = {"DY1": 0, "DY2": 1, "DY3": 2, "DY4": 3}[which]
        out = np.empty((self.NY, self.NY), dtype=np.float64, order="F")
        self._check(lib().owt_get_matrix(self._h, idx, _p(out)))
        return out

    def factor(self, sys, op):
        """LU factor (dgetrf layout) and pivot rows of operator op (0 Av, 1 Ag)
        of the +-kz system sys = mz*NXH + i."""
        lu = np.empty((self.NY, self.NY), dtype=np.float64, order="F")
        perm = np.empty(self.NY, dtype=np.int32)
        self._check(lib().owt_get_factor(self._h, int(sys), int(op), _p(lu), _p(perm)))
        return lu, perm

    def time_steps(self, n):
        ms = ctypes.c_double()
        self._check(lib().owt_time_steps(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def autotune(self, reps=5):
        """Time every candidate launch shape on this channel and keep the fastest;
        returns {knob: {"choice": i, "ms": {shape: ms}}}."""
        import json
        buf = ctypes.create_string_buffer(4096)
        self._check(lib().owt_autotune(self._h, int(reps), buf, len(buf)))
        return json.loads(buf.value.decode())

    @property
    def tune(self):
        k = np.zeros(6, dtype=np.int32)
        self._check(lib().owt_get_tune(self._h, _p(k)))
        return [int(v) for v in k]

    @tune.setter
    def tune(self, knobs):
        k = np.ascontiguousarray(knobs, dtype=np.int32)
        self._check(lib().owt_set_tune(self._h, _p(k)))

    def tangent_init(self, du, dv, dw):
        """Attach the tangent (forward-mode linearization) started from the
        direction (du, dv, dw) of the initial velocity; returns it as a Channel."""
        arrs = [np.ascontiguousarray(a, dtype=np.complex128).reshape(self.wave_shape) for a in (du, dv, dw)]
        self._check(lib().owt_tangent_init(self._h, *[_p(a) for a in arrs]))
        return Channel(_member=(ctypes.c_void_p(lib().owt_tangent_handle(self._h)), self))

    def adjoint_run(self, K, ckpt_every=5, first=1, last=None):
        """J = mean friction force over steps first..last of K steps from the
        initial field, and dJ/d(u, v, w) of that field by the discrete adjoint."""
        last = K if last is None else last
        J = ctypes.c_double()
        g = [np.empty(self.wave_shape, dtype=np.complex128) for _ in range(3)]
        self._check(lib().owt_adjoint_run(self._h, int(K), int(ckpt_every), 0, int(first), int(last), ctypes.byref(J),
                                          *[_p(a) for a in g]))
        return J.value, g

    def adjoint_apply(self, K, wu, wv, ww, ckpt_every=5):
        """L^T w for the linear map L from the initial velocity to the velocity
        after K steps (about this channel's trajectory from its initial field),
        w = (wu, wv, ww): the gradients of <L d, w> with respect to d."""
        ws = [np.ascontiguousarray(a, dtype=np.complex128).reshape(self.wave_shape) for a in (wu, wv, ww)]
        g = [np.empty(self.wave_shape, dtype=np.complex128) for _ in range(3)]
        self._check(lib().owt_adjoint_apply(self._h, int(K), int(ckpt_every), *[_p(a) for a in ws], *[_p(a) for a in g]))
        return g

    def time_kernel(self, name, reps):
        ms = ctypes.c_double()
        self._check(lib().owt_time_kernel(self._h, KERNELS[name], int(reps), ctypes.byref(ms)))
        return ms.value


def _params_struct(p):
    return _Params(p["nu"], p["dt"], p["dkx"], p["dkz"], p["alpha_mesh"], p["NX"], p["NY"], p["NZ"],
                   p.get("precision", 0), p.get("method", 1))


class Ensemble:
    """Independent channels on one GPU, stepped together (one captured graph per
    step with the batched method). ens[m] is member m as a Channel."""

    def __init__(self, params_list, *, device=0, precision=-1, method=-1):
        L = lib()
        arr = (_Params * len(params_list))(*[_params_struct(p) for p in params_list])
        opts = _Opts(device, precision, method, 0)
        e = ctypes.c_void_p()
        rc = L.owt_ensemble_create(len(params_list), ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(opts),
                                   ctypes.byref(e))
        if rc:
            raise OwtError(rc, L.owt_ensemble_last_error(None).decode())
        self._e = e
        self.members = [Channel(_member=(ctypes.c_void_p(L.owt_ensemble_member(e, m)), self))
                        for m in range(len(params_list))]

    def __len__(self):
        return len(self.members)

    def __getitem__(self, m):
        return self.members[m]

    def _check(self, rc):
        if rc:
            raise OwtError(rc, lib().owt_ensemble_last_error(self._e).decode())

    def init_fields(self, mode=2, seeds=None, amp=0.15):
        if seeds is None:
            seeds = range(len(self))
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        self._check(lib().owt_ensemble_init_fields(self._e, int(mode), _p(s), float(amp)))

    def step(self, n=1):
        self._check(lib().owt_ensemble_step(self._e, int(n)))

    def time_steps(self, n):
        ms = ctypes.c_double()
        self._check(lib().owt_ensemble_time_steps(self._e, int(n), ctypes.byref(ms)))
        return ms.value

    def autotune(self, reps=5):
        """Launch-shape autotuner: one member per distinct parameter set is tuned,
        the members with the same parameters take its choices."""
        self._check(lib().owt_ensemble_autotune(self._e, int(reps)))

    def launches_per_step(self):
        """Kernel launches of one grouped ensemble step (0: members stepped one by one)."""
        n = ctypes.c_int()
        self._check(lib().owt_ensemble_launches(self._e, ctypes.byref(n)))
        return n.value

    @property
    def gridpoints(self):
        return sum(m.gridpoints for m in self.members)

    def close(self):
        if getattr(self, "_e", None):
            for m in self.members:
                m._h = None
            lib().owt_ensemble_destroy(self._e)
            self._e = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def case_params(nu, dt, dkx, dkz, alpha_mesh, NX, NY, NZ, precision=0, method=1):
    return dict(nu=nu, dt=dt, dkx=dkx, dkz=dkz, alpha_mesh=alpha_mesh, NX=NX, NY=NY, NZ=NZ,
                precision=precision, method=method)


# The workloads of the benchmark and parity suite (box, grid, stretch, nu, dt).
CONFIGS = {
    # Re_tau 180, box 4pi x 2 x 2pi, 64x65x64
    "re180_64": case_params(1 / 2800, 1e-3, 0.5, 1.0, 1.5, 64, 65, 64),
    # Re_tau 180, box 4pi x 2 x 4pi/3, 128x129x128
    "re180_128": case_params(1 / 2800, 1e-3, 0.5, 1.5, 1.5, 128, 129, 128),
    # Re_tau 550, box 2pi x 2 x pi, 384x193x384
    "re550_384": case_params(1e-4, 1e-3, 1.0, 2.0, 2.0, 384, 193, 384),
}

# Ensemble workloads: lists of member parameters.
ENSEMBLES = {
    # 32 Re_tau 180 channels of uneven size in the 4pi x 2 x 2pi box:
    # 16 at 64x65x64, 8 at 64x97x64, 8 at 96x49x96
    "re180_32": [case_params(1 / 2800, 1e-3, 0.5, 1.0, 1.5, *shape)
                 for shape in [(64, 65, 64)] * 16 + [(64, 97, 64)] * 8 + [(96, 49, 96)] * 8],
}
