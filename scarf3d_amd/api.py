"""ctypes binding of include/scarf3d.h and include/scarf3d_srm.h.

`Library("f")` wraps libscarf3df.so (real = float), `Library("d")` wraps
libscarf3d.so (real = double). Loading fails loudly when the library has not
been built: there is no Python or CPU fallback for the field computation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

KERNELS = {
    "auto": 0,
    "direct_f32": 1,
    "direct_f64": 2,
    "mfma_bf16x3": 3,
    "mfma_f16x2": 4,
    "mfma_bf16x2": 5,
}

# every symbol the two public headers declare
EXPORTS = (
    "scarf_opt_init", "scarf_cart_initialize", "scarf_nocart_initialize", "scarf_execute",
    "scarf_finalize", "scarf_io_one", "scarf_io_slice",
    "scarf_set_nharm", "scarf_get_nharm", "scarf_set_kernel", "scarf_get_kernel", "scarf_last_error",
    "scarf_execute_device", "scarf_srm_draw", "scarf_srm_run", "scarf_last_timings", "scarf_stream",
    "scarf_setup_values", "scarf_draw_harmonics", "scarf_get_config", "scarf_set_option", "scarf_srm_run_fp64",
    "scarf_ensemble_execute", "scarf_ensemble_last_ms", "scarf_last_tile", "scarf_srm_run_noalloc",
)


class LibraryMissing(RuntimeError):
    pass


def _opt_struct(real):
    class ScarfOpt(C.Structure):
        # struct scarf_opt, src/scarf3d.h:54-70
        _fields_ = [
            ("method", C.c_int), ("hurst", real), ("dh", real), ("ds", real), ("poi", C.POINTER(real)),
            ("npoi", C.c_int), ("mute", real), ("taper", real), ("rescale", C.c_int), ("pad", C.c_int),
            ("nc", C.POINTER(real)), ("fc", C.POINTER(real)), ("alpha", real), ("beta", real), ("gamma", real),
        ]
    return ScarfOpt


def _member_struct(real):
    class ScarfMember(C.Structure):
        # struct scarf_member (include/scarf3d_srm.h)
        _fields_ = [
            ("nd", C.c_int), ("fs", C.c_int * 3), ("fe", C.c_int * 3), ("ds", real), ("dh", real), ("acf", C.c_int),
            ("cl", real * 3), ("sigma", real), ("hurst", real), ("alpha", real), ("beta", real), ("gamma", real),
            ("rescale", C.c_int), ("nharm", C.c_int), ("seed", C.c_int),
        ]
    return ScarfMember


class Library:
    """One flavour of the library; flavour 'f' (float real) or 'd' (double real)."""

    def __init__(self, flavour: str = "f", path: Optional[Path] = None):
        if flavour not in ("f", "d"):
            raise ValueError("flavour must be 'f' or 'd'")
        self.flavour = flavour
        self.real = C.c_float if flavour == "f" else C.c_double
        self.dtype = np.float32 if flavour == "f" else np.float64
        name = "libscarf3df.so" if flavour == "f" else "libscarf3d.so"
        self.path = Path(path) if path else LIB_DIR / name
        if not self.path.exists():
            raise LibraryMissing(f"{self.path} is not built (run: python -m scarf3d_amd.build)")
        self.lib = C.CDLL(str(self.path))
        self.Opt = _opt_struct(self.real)
        self.Member = _member_struct(self.real)
        self._keep = []  # arrays whose pointers the library borrows
        self._declare()

    # ------------------------------------------------------------------
    def _declare(self):
        L, R, P = self.lib, self.real, C.POINTER
        ip, rp = P(C.c_int), P(self.real)
        L.scarf_opt_init.argtypes = [P(self.Opt)]
        L.scarf_opt_init.restype = None
        L.scarf_cart_initialize.argtypes = [C.c_int, ip, ip, R, C.c_int, rp, R, P(self.Opt)]
        L.scarf_cart_initialize.restype = None
        L.scarf_nocart_initialize.argtypes = [C.c_int, C.c_int, rp, rp, rp, R, C.c_int, rp, R, P(self.Opt)]
        L.scarf_nocart_initialize.restype = None
        L.scarf_execute.argtypes = [C.c_int, C.c_void_p, rp]
        L.scarf_execute.restype = None
        L.scarf_finalize.argtypes = []
        L.scarf_finalize.restype = None
        L.scarf_io_one.argtypes = [C.c_int, ip, rp, C.c_char_p, ip]
        L.scarf_io_one.restype = None
        L.scarf_io_slice.argtypes = [ip, C.c_char_p, C.c_int, rp, C.c_char_p]
        L.scarf_io_slice.restype = None
        L.scarf_set_nharm.argtypes = [C.c_int]
        L.scarf_set_nharm.restype = C.c_int
        L.scarf_get_nharm.argtypes = []
        L.scarf_get_nharm.restype = C.c_int
        L.scarf_set_kernel.argtypes = [C.c_int]
        L.scarf_set_kernel.restype = C.c_int
        L.scarf_get_kernel.argtypes = []
        L.scarf_get_kernel.restype = C.c_int
        L.scarf_last_error.argtypes = []
        L.scarf_last_error.restype = C.c_char_p
        L.scarf_execute_device.argtypes = [C.c_int, C.c_void_p, rp]
        L.scarf_execute_device.restype = C.c_int
        L.scarf_srm_draw.argtypes = [C.c_int, P(C.c_double)]
        L.scarf_srm_draw.restype = C.c_int
        L.scarf_srm_run.argtypes = [C.c_void_p, rp]
        L.scarf_srm_run.restype = C.c_int
        L.scarf_srm_run_noalloc.argtypes = [C.c_void_p, rp]
        L.scarf_srm_run_noalloc.restype = C.c_int
        L.scarf_last_timings.argtypes = [P(C.c_double), C.c_int]
        L.scarf_last_timings.restype = C.c_int
        L.scarf_stream.argtypes = []
        L.scarf_stream.restype = C.c_void_p
        L.scarf_setup_values.argtypes = [P(C.c_double), C.c_int]
        L.scarf_setup_values.restype = C.c_int
        L.scarf_draw_harmonics.argtypes = [C.c_int, P(C.c_double), P(C.c_ubyte)]
        L.scarf_draw_harmonics.restype = C.c_long
        L.scarf_get_config.argtypes = [P(C.c_double), C.c_int]
        L.scarf_get_config.restype = C.c_int
        L.scarf_ensemble_execute.argtypes = [C.c_int, P(self.Member), P(C.c_void_p), rp, C.c_int]
        L.scarf_ensemble_execute.restype = C.c_int
        L.scarf_ensemble_last_ms.argtypes = []
        L.scarf_ensemble_last_ms.restype = C.c_double
        L.scarf_set_option.argtypes = [C.c_char_p, C.c_double]
        L.scarf_set_option.restype = C.c_int
        L.scarf_srm_run_fp64.argtypes = [C.c_void_p, P(C.c_double)]
        L.scarf_srm_run_fp64.restype = C.c_int
        L.scarf_last_tile.argtypes = [P(C.c_int)]
        L.scarf_last_tile.restype = C.c_int

    def _arr(self, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or self.dtype)
        self._keep.append(a)
        return a

    def _rp(self, a):
        return a.ctypes.data_as(C.POINTER(self.real))

    # ------------------------------------------------------------------
    # reference C API
    def opt(self, **kw):
        """struct scarf_opt initialised by scarf_opt_init, then the given members set."""
        o = self.Opt()
        self.lib.scarf_opt_init(C.byref(o))
        for k, v in kw.items():
            if v is None:
                continue
            if k == "poi":
                a = self._arr(np.asarray(v, dtype=self.dtype).reshape(-1))
                o.poi = self._rp(a)
            elif k in ("nc", "fc"):
                a = self._arr(v)
                setattr(o, k, self._rp(a))
            else:
                setattr(o, k, v)
        self._opt = o
        return o

    def cart_initialize(self, nd, fs, fe, ds, acf, cl, sigma, opt=None):
        self._keep = [] if opt is None else self._keep
        fs_a = self._arr(fs, np.int32)
        fe_a = self._arr(fe, np.int32)
        cl_a = self._arr(cl)
        self.lib.scarf_cart_initialize(
            nd, fs_a.ctypes.data_as(C.POINTER(C.c_int)), fe_a.ctypes.data_as(C.POINTER(C.c_int)), ds, acf,
            self._rp(cl_a), sigma, C.byref(opt) if opt is not None else None)
        self.grid = tuple(int(e) - int(s) + 1 for s, e in zip(fs, fe))
        self.npts = int(np.prod(self.grid))

    def nocart_initialize(self, nd, x, y, z, dh, acf, cl, sigma, opt=None):
        self._keep = [] if opt is None else self._keep
        xa, ya = self._arr(x), self._arr(y)
        za = self._arr(z) if z is not None else None
        cl_a = self._arr(cl)
        self.lib.scarf_nocart_initialize(
            nd, len(xa), self._rp(xa), self._rp(ya), self._rp(za) if za is not None else None, dh, acf,
            self._rp(cl_a), sigma, C.byref(opt) if opt is not None else None)
        self.grid = (len(xa),)
        self.npts = len(xa)

    def execute(self, seed: int, out: Optional[np.ndarray] = None):
        """scarf_execute: returns (field in Fortran order of the local block,
        stats[8]). `out`, a contiguous array of npts `real` values, receives
        the field instead of a new array (a caller's field array reused
        across realizations)."""
        if out is None:
            field = np.zeros(self.npts, dtype=self.dtype)
        else:
            field = out.reshape(-1, order="A")
            if field.dtype != self.dtype or field.size != self.npts or not (out.flags.c_contiguous or
                                                                             out.flags.f_contiguous):
                raise ValueError("out must be a contiguous array of npts real values")
        stats = np.zeros(8, dtype=self.dtype)
        self.lib.scarf_execute(seed, field.ctypes.data_as(C.c_void_p), self._rp(stats))
        return field.reshape(self.grid, order="F"), stats

    def finalize(self):
        self.lib.scarf_finalize()
        self._keep = []

    def io_one(self, npts, field, fname, nwriters=None):
        n = self._arr(npts, np.int32)
        f = self._arr(np.asarray(field, dtype=self.dtype).reshape(-1, order="F"))
        nw = None
        if nwriters is not None:
            nw = C.byref(C.c_int(nwriters))
        self.lib.scarf_io_one(len(n), n.ctypes.data_as(C.POINTER(C.c_int)), self._rp(f), str(fname).encode(), nw)

    def io_slice(self, npts, axis, plane, field, fname):
        n = self._arr(npts, np.int32)
        f = self._arr(np.asarray(field, dtype=self.dtype).reshape(-1, order="F"))
        self.lib.scarf_io_slice(n.ctypes.data_as(C.POINTER(C.c_int)), axis.encode(), plane, self._rp(f),
                                str(fname).encode())

    # ------------------------------------------------------------------
    # extensions
    def last_error(self) -> str:
        return self.lib.scarf_last_error().decode()

    def set_nharm(self, n: int) -> int:
        return self.lib.scarf_set_nharm(n)

    def set_kernel(self, k) -> int:
        return self.lib.scarf_set_kernel(KERNELS[k] if isinstance(k, str) else int(k))

    def kernel(self) -> str:
        inv = {v: k for k, v in KERNELS.items()}
        return inv.get(self.lib.scarf_get_kernel(), "unknown")

    def setup_values(self) -> dict:
        out = np.zeros(21)
        if self.lib.scarf_setup_values(out.ctypes.data_as(C.POINTER(C.c_double)), 21) != 0:
            raise RuntimeError(self.last_error())
        return {"ndim": int(out[0]), "kmin": out[1], "kmax": out[2], "kc": out[3], "info1": out[4],
                "info2": out[5], "scaling": out[6], "band": out[7], "all": out[8],
                "matrix": out[9:18].copy(), "bar": out[18:21].copy()}

    def config_values(self) -> dict:
        out = np.zeros(32)
        if self.lib.scarf_get_config(out.ctypes.data_as(C.POINTER(C.c_double)), 32) < 0:
            raise RuntimeError(self.last_error())
        keys = ["nd", "structured", "fs1", "fs2", "fs3", "fe1", "fe2", "fe3", "ds", "dh", "acf", "cl1", "cl2",
                "cl3", "sigma", "hurst", "nc1", "nc2", "nc3", "fc1", "fc2", "fc3", "alpha", "beta", "gamma",
                "mute", "taper", "rescale", "pad", "method", "npoi", "npts"]
        return dict(zip(keys, out.tolist()))

    def draw_harmonics(self, seed: int, nharm: int):
        harm = np.zeros((nharm, 5))
        used = np.zeros(nharm, dtype=np.uint8)
        nu = self.lib.scarf_draw_harmonics(seed, harm.ctypes.data_as(C.POINTER(C.c_double)),
                                           used.ctypes.data_as(C.POINTER(C.c_ubyte)))
        if nu < 0:
            raise RuntimeError(self.last_error())
        return harm, used, nu

    def srm_draw(self, seed: int):
        hs = C.c_double(0.0)
        n = self.lib.scarf_srm_draw(seed, C.byref(hs))
        if n < 0:
            raise RuntimeError(self.last_error())
        return n, hs.value

    def srm_run(self, d_field_ptr: int):
        stats = np.zeros(8, dtype=self.dtype)
        if self.lib.scarf_srm_run(C.c_void_p(d_field_ptr), self._rp(stats)) != 0:
            raise RuntimeError(self.last_error())
        return stats

    def srm_run_noalloc(self, d_field_ptr: int):
        """scarf_srm_run_noalloc: returns (return code, stats[8]); no device allocation."""
        stats = np.zeros(8, dtype=self.dtype)
        rc = self.lib.scarf_srm_run_noalloc(C.c_void_p(d_field_ptr), self._rp(stats))
        return rc, stats

    def srm_run_fp64(self, d_field_ptr: int):
        st = np.zeros(4)
        if self.lib.scarf_srm_run_fp64(C.c_void_p(d_field_ptr), st.ctypes.data_as(C.POINTER(C.c_double))) != 0:
            raise RuntimeError(self.last_error())
        return st

    def ensemble_execute(self, members, d_field_ptrs, kernel="auto"):
        """members: list of dicts with the scarf_member fields (nd, fs, fe, ds,
        dh, acf, cl, sigma, hurst, alpha, beta, gamma, rescale, nharm, seed).
        Returns stats (n x 8)."""
        n = len(members)
        arr = (self.Member * n)()
        for m, d in zip(arr, members):
            m.nd = d.get("nd", 3)
            for i in range(3):
                m.fs[i] = d.get("fs", (1, 1, 1))[i]
                m.fe[i] = d["fe"][i]
                m.cl[i] = d["cl"][i]
            m.ds, m.dh, m.acf = d["ds"], d.get("dh", 0.0), d["acf"]
            m.sigma, m.hurst = d["sigma"], d.get("hurst", 0.0)
            m.alpha, m.beta, m.gamma = d.get("alpha", 0.0), d.get("beta", 0.0), d.get("gamma", 0.0)
            m.rescale, m.nharm, m.seed = d.get("rescale", 0), d["nharm"], d["seed"]
        ptrs = (C.c_void_p * n)(*[C.c_void_p(p) for p in d_field_ptrs])
        stats = np.zeros(8 * n, dtype=self.dtype)
        k = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        if self.lib.scarf_ensemble_execute(n, arr, ptrs, self._rp(stats), k) != 0:
            raise RuntimeError(self.last_error())
        return stats.reshape(n, 8)

    def ensemble_ms(self) -> float:
        return self.lib.scarf_ensemble_last_ms()

    def last_tile(self):
        """(x points, j values, k values) of the last MFMA run's block tile, or None."""
        out = (C.c_int * 3)()
        return tuple(out) if self.lib.scarf_last_tile(out) == 3 else None

    def set_option(self, name: str, value: float) -> int:
        return self.lib.scarf_set_option(name.encode(), float(value))

    def execute_device(self, seed: int, d_field_ptr: int):
        stats = np.zeros(8, dtype=self.dtype)
        if self.lib.scarf_execute_device(seed, C.c_void_p(d_field_ptr), self._rp(stats)) != 0:
            raise RuntimeError(self.last_error())
        return stats

    def timings(self) -> dict:
        out = np.zeros(5)
        self.lib.scarf_last_timings(out.ctypes.data_as(C.POINTER(C.c_double)), 5)
        return dict(zip(["tables", "field", "stats", "epilogue", "total"], out.tolist()))

    def stream(self) -> int:
        return self.lib.scarf_stream() or 0

    # ------------------------------------------------------------------
    # Fortran-style generic interface (src/scarflib.f90:97-296)
    def scarf_initialize(self, *args, **kw):
        """Structured: (fs, fe, ds, acf, cl, sigma, **opts); unstructured:
        (dh, acf, cl, sigma, x, y, z=None, **opts). Optional arguments take the
        names of the Fortran module; through the C ABI a value counts as given
        by the rules of src/c_binding.c:89-103."""
        names = ("method", "hurst", "dh", "ds", "poi", "mute", "taper", "rescale", "pad", "nc", "fc",
                 "alpha", "beta", "gamma")
        opts = {k: kw.pop(k) for k in list(kw) if k in names}
        poi = opts.pop("poi", None)
        if poi is not None:
            poi = np.asarray(poi, dtype=self.dtype)
            opts["poi"] = poi.reshape(-1, order="F")
            opts["npoi"] = poi.shape[1] if poi.ndim == 2 else 1
        if len(args) >= 6 and np.ndim(args[0]) == 1:
            fs, fe, ds, acf, cl, sigma = args[:6]
            o = self.opt(**opts)
            self.cart_initialize(len(fs), fs, fe, ds, acf, cl, sigma, o)
        else:
            dh, acf, cl, sigma, x, y = args[:6]
            z = args[6] if len(args) > 6 else kw.pop("z", None)
            o = self.opt(**opts)
            self.nocart_initialize(len(cl), x, y, z, dh, acf, cl, sigma, o)

    def scarf_execute(self, seed: int):
        return self.execute(seed)

    def scarf_finalize(self):
        self.finalize()


_loaded = {}


def load(flavour: str = "f") -> Library:
    if flavour not in _loaded:
        _loaded[flavour] = Library(flavour)
    return _loaded[flavour]
