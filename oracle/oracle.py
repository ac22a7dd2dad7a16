"""ctypes wrapper of the CPU parity oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does. See owt_oracle.cpp for what it restates
(reference src/solver.cpp, src/FFT.cpp, src/compdiff.hpp, src/linalg.hpp).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")

FIELDS = {"u": 0, "v": 1, "w": 2, "omx": 3, "omy": 4, "omz": 5,
          "lmbx": 6, "lmby": 7, "lmbz": 8, "Hv": 9, "Hg": 10, "p": 11}
MATRICES = {"DY1": 0, "DY2": 1, "DY3": 2, "DY4": 3, "A1": 4, "A2": 5, "A3": 6, "A4": 7,
            "B1": 8, "B2": 9, "B3": 10, "B4": 11, "D1": 12, "D2": 13, "D3": 14, "D4": 15}
VECTORS = {"y": 0, "s": 1, "dyds1": 2, "dyds2": 3, "dyds3": 4, "dyds4": 5, "c11": 6, "c21": 7,
           "c22": 8, "c41": 9, "c42": 10, "c43": 11, "c44": 12, "kx": 13, "kz": 14}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        D, I, P, U64 = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64
        L.orc_create.restype = P
        L.orc_create.argtypes = [D, D, D, D, D, I, I, I]
        L.orc_destroy.argtypes = [P]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_set_threads.argtypes = [I]
        L.orc_set_threads.restype = I
        L.orc_init.argtypes = [P, I, U64, D]
        L.orc_set_velocity.argtypes = [P, P, P, P]
        L.orc_step.argtypes = [P, I]
        L.orc_get_field.argtypes = [P, I, P]
        L.orc_get_matrix.argtypes = [P, I, P]
        L.orc_get_vector.argtypes = [P, I, P]
        L.orc_wall_shear.argtypes = [P, P, P, P]
        L.orc_diag_count.argtypes = [P]
        L.orc_diag.argtypes = [P, P]
        L.orc_f2p.argtypes = [P, P, P]
        L.orc_p2f.argtypes = [P, P, P]
        L.orc_lamb.argtypes = [P, P, P]
        L.orc_pressure_of.argtypes = [P, P, P, P, P, P]
        L.orc_peak_velocity.argtypes = [P, P]
        L.orc_fft.argtypes = [I, I, P, P]
        L.orc_lu_solve.argtypes = [I, P, I, P]
        L.orc_lu_factor.argtypes = [I, P, P]
        L.orc_seeded_gauss.argtypes = [U64, I, I, I, I, P]
        L.orc_set_noise.argtypes = [P, U64, D, I]
        L.orc_set_mutation.argtypes = [P, I]
        L.orc_set_factor_cache.argtypes = [P, I]
        L.orc_set_fft_variant.argtypes = [P, I]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    return lib().orc_set_threads(int(n))


class Oracle:
    """One channel: params are the reference's toml keys (src/IO.cpp:14-25) with PX = PZ = 1."""

    def __init__(self, nu, dt, dkx, dkz, alpha_mesh, NX, NY, NZ):
        self.NX, self.NY, self.NZ = int(NX), int(NY), int(NZ)
        self.NXH, self.NZH = self.NX // 2, self.NZ // 2
        self.NX2, self.NZ2 = self.NX // 2 * 3, self.NZ // 2 * 3
        self.nu, self.dt = nu, dt
        h = lib().orc_create(nu, dt, dkx, dkz, alpha_mesh, self.NX, self.NY, self.NZ)
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        self._h = h

    @classmethod
    def from_config(cls, cfg):
        return cls(cfg["nu"], cfg["dt"], cfg["dkx"], cfg["dkz"], cfg["alpha_mesh"], cfg["NX"], cfg["NY"], cfg["NZ"])

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    @property
    def wave_shape(self):
        return (self.NZ, self.NXH, self.NY)

    @property
    def phys_shape(self):
        return (self.NZ2, self.NX2, self.NY)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def init(self, mode=2, seed=0, amp=0.15):
        self._check(lib().orc_init(self._h, int(mode), int(seed), float(amp)))

    def set_velocity(self, u, v, w):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128).reshape(self.wave_shape) for a in (u, v, w)]
        self._check(lib().orc_set_velocity(self._h, *[_p(a) for a in arrs]))

    def step(self, n=1):
        self._check(lib().orc_step(self._h, int(n)))

    # test instruments (owt_oracle.cpp, Oracle::perturb and Oracle::mutate)
    NOISE_LAMB, NOISE_SOLVE = 1, 2
    MUT_HG_SIGN, MUT_FRICTION_SIGN, MUT_U_OMY_SIGN, MUT_HG_KX_SIGN = 1, 2, 4, 8
    MUT_LAMB_HISTORY, MUT_FRICTION_HISTORY = 16, 32

    def set_noise(self, seed, rel, where):
        """Multiply the fields named by `where` (NOISE_LAMB | NOISE_SOLVE) by (1 + rel g),
        g ~ N(0, 1), at every step: the spread that rounding alone causes."""
        self._check(lib().orc_set_noise(self._h, int(seed), float(rel), int(where)))

    def set_fft_variant(self, variant):
        """1: transforms by a second, equally valid FFT factorization (radix 2 first)."""
        self._check(lib().orc_set_fft_variant(self._h, int(variant)))

    def set_mutation(self, bits):
        """Planted mistakes (MUT_*) that the oracle's pins must detect."""
        self._check(lib().orc_set_mutation(self._h, int(bits)))

    def set_factor_cache(self, on=True):
        """Factor Av and Ag once per distinct k^2 and reuse the factors: the same
        fields bit for bit as factoring every pencil at every step, in a fraction
        of the time at large sizes."""
        self._check(lib().orc_set_factor_cache(self._h, int(bool(on))))

    def field(self, name):
        out = np.empty(self.wave_shape, dtype=np.complex128)
        self._check(lib().orc_get_field(self._h, FIELDS[name], _p(out)))
        return out

    def pressure_of(self, lx, ly, lz, v):
        """Dynamic pressure from a given Lamb vector and v (field "p" uses the state's)."""
        arrs = [np.ascontiguousarray(a, dtype=np.complex128).reshape(self.wave_shape) for a in (lx, ly, lz, v)]
        out = np.empty(self.wave_shape, dtype=np.complex128)
        self._check(lib().orc_pressure_of(self._h, *[_p(a) for a in arrs], _p(out)))
        return out

    def matrix(self, name):
        out = np.empty((self.NY, self.NY), dtype=np.float64, order="F")
        self._check(lib().orc_get_matrix(self._h, MATRICES[name], _p(out)))
        return out

    def vector(self, name):
        n = {"kx": self.NXH, "kz": self.NZ}.get(name, self.NY)
        out = np.empty(n, dtype=np.float64)
        self._check(lib().orc_get_vector(self._h, VECTORS[name], _p(out)))
        return out

    def wall_shear(self):
        a, b, c = (ctypes.c_double(), ctypes.c_double(), ctypes.c_double())
        lib().orc_wall_shear(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def diag(self):
        n = lib().orc_diag_count(self._h)
        out = np.empty((n, 3), dtype=np.float64)
        lib().orc_diag(self._h, _p(out))
        return out

    def f2p(self, wave):
        wave = np.ascontiguousarray(wave, dtype=np.complex128).reshape(self.wave_shape)
        out = np.empty(self.phys_shape, dtype=np.float64)
        lib().orc_f2p(self._h, _p(wave), _p(out))
        return out

    def p2f(self, phys):
        phys = np.ascontiguousarray(phys, dtype=np.float64).reshape(self.phys_shape)
        out = np.empty(self.wave_shape, dtype=np.complex128)
        lib().orc_p2f(self._h, _p(phys), _p(out))
        return out

    def lamb(self, six):
        six = np.ascontiguousarray(six, dtype=np.complex128).reshape((6,) + self.wave_shape)
        out = np.empty((3,) + self.wave_shape, dtype=np.complex128)
        lib().orc_lamb(self._h, _p(six), _p(out))
        return out

    def peak_velocity(self):
        v = ctypes.c_double()
        lib().orc_peak_velocity(self._h, ctypes.byref(v))
        return v.value


def fft(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_fft(x.size, int(sign), _p(x), _p(out))
    return out


def lu_solve(A, b):
    """Solve with the oracle's dgetrf/dgetrs restatement; b is (n,) or (n, nrhs)."""
    A = np.asfortranarray(A, dtype=np.float64)
    b2 = np.array(b, dtype=np.float64, order="F", copy=True)
    nrhs = 1 if b2.ndim == 1 else b2.shape[1]
    info = lib().orc_lu_solve(A.shape[0], _p(A), nrhs, _p(b2))
    if info:
        raise np.linalg.LinAlgError(f"singular (info={info})")
    return b2


def lu_factor(A):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    ipiv = np.empty(A.shape[0], dtype=np.int32)
    info = lib().orc_lu_factor(A.shape[0], _p(A), _p(ipiv))
    return A, ipiv, info


def seeded_gauss(seed, field, i, m, p):
    out = np.empty(2, dtype=np.float64)
    lib().orc_seeded_gauss(int(seed), field, i, m, p, _p(out))
    return complex(out[0], out[1])
