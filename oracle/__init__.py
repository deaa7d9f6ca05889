"""ctypes bindings of the test-side checkers.

TEST INFRASTRUCTURE ONLY. Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this package, and only to check or to
time the CPU reference beside the GPU. The product (``zest_amd``) never imports
it.

Two libraries are bound:

* ``liboracle.so`` -- the C restatement of zest's GLQ transforms
  (``oracle/zgl_oracle.c``), always built by ``oracle/Makefile``.
* ``_ref/libzest_ref.so`` -- the upstream zest library compiled from its own
  sources behind ``oracle/zest_ref_shim.cpp``. Present when it was built in a
  container holding the zest sources; ``REF`` is ``None`` otherwise.

Convention arguments use zest's enum values: ``zernike_norm`` 0 = normed,
1 = unnormed; ``sh_norm`` 0 = geo, 1 = qm; ``sh_phase`` -1 = none, +1 = cs;
SH ``layout`` 0 = LonLat (zest's default), 1 = LatLon.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_c_size = ctypes.c_size_t
_c_int = ctypes.c_int
_ptr = ctypes.c_void_p


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _load_oracle():
    path = os.path.join(_HERE, "liboracle.so")
    lib = ctypes.CDLL(path)
    lib.zo_gl_nodes_weights.argtypes = [_c_size, _ptr, _ptr]
    lib.zo_gl_packed.argtypes = [_c_size, _ptr, _ptr]
    lib.zo_plm_table.argtypes = [_c_int, _c_int, _c_size, _ptr, _c_size, _ptr]
    lib.zo_radial_table.argtypes = [_c_int, _c_size, _ptr, _c_size, _ptr]
    lib.zo_rdft_forward.argtypes = [_c_size, _c_size, _ptr, _c_size, ctypes.c_double, _ptr]
    lib.zo_rdft_backward.argtypes = [_c_size, _c_size, _ptr, _c_size, _ptr]
    lib.zo_zt_forward.argtypes = [_c_int, _c_int, _c_int, _c_size, _ptr, _c_size, _ptr]
    lib.zo_zt_backward.argtypes = [_c_int, _c_int, _c_int, _c_size, _ptr, _c_size, _ptr]
    lib.zo_st_forward.argtypes = [_c_int, _c_int, _c_int, _c_size, _ptr, _c_size, _ptr]
    lib.zo_st_backward.argtypes = [_c_int, _c_int, _c_int, _c_size, _ptr, _c_size, _ptr]
    lib.zo_zt_evaluate.argtypes = [_c_int, _c_int, _c_int, _c_size, _ptr, _ptr, _c_size, _ptr,
                                   _c_size, _ptr, _c_size, _ptr]
    lib.zo_st_evaluate.argtypes = [_c_int, _c_int, _c_size, _ptr, _ptr, _c_size, _ptr, _c_size, _ptr]
    lib.zo_st_power_spectrum.argtypes = [_c_int, _c_size, _ptr, _ptr, _ptr]
    lib.zo_zt_power_spectrum.argtypes = [_c_int, _c_size, _ptr, _ptr, _ptr]
    lib.zo_wignerd_table.argtypes = [_c_size, _ptr]
    _d = ctypes.c_double
    lib.zo_rotate.argtypes = [_c_int, _c_size, _ptr, _d, _d, _d, _c_int]
    lib.zo_polar_rotate.argtypes = [_c_int, _c_size, _ptr, _d, _c_int]
    return lib


def _host_has_avx512() -> bool:
    """x86-64-v4 (AVX-512 F, BW, CD, DQ, VL) on this host, from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    return {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    except OSError:
        pass
    return False


# Which build of compiled zest REF is: "x86-64-v4" (AVX-512, when the host has
# it and that build exists) or "x86-64-v3" (AVX2 + FMA); None without zest.
REF_ISA = None


def _load_ref():
    global REF_ISA
    v3 = os.path.join(_HERE, "_ref", "libzest_ref.so")
    v4 = os.path.join(_HERE, "_ref", "libzest_ref_v4.so")
    if os.path.exists(v4) and _host_has_avx512():
        path, REF_ISA = v4, "x86-64-v4"
    elif os.path.exists(v3):
        path, REF_ISA = v3, "x86-64-v3"
    else:
        return None
    lib = ctypes.CDLL(path)
    lib.zref_zt_create.restype = _ptr
    lib.zref_zt_create.argtypes = [_c_int, _c_int, _c_int, _c_size]
    lib.zref_zt_destroy.argtypes = [_ptr]
    lib.zref_zt_forward.argtypes = [_ptr, _ptr, _c_size, _ptr, _c_size]
    lib.zref_zt_backward.argtypes = [_ptr, _ptr, _c_size, _ptr, _c_size]
    lib.zref_st_create.restype = _ptr
    lib.zref_st_create.argtypes = [_c_int, _c_int, _c_int, _c_size]
    lib.zref_st_destroy.argtypes = [_ptr]
    lib.zref_st_forward.argtypes = [_ptr, _ptr, _c_size, _ptr, _c_size]
    lib.zref_st_backward.argtypes = [_ptr, _ptr, _c_size, _ptr, _c_size]
    lib.zref_st_backward_zernike_row.argtypes = [_ptr, _ptr, _c_size, _c_size, _ptr, _c_size]
    lib.zref_rotate_l.argtypes = [_c_size, _ptr, _ptr, _ptr, _ptr]
    lib.zref_polar_rotate_l.argtypes = [_c_size, _ptr, _ptr]
    lib.zref_gl_nodes_weights.argtypes = [_c_size, _c_int, _ptr, _ptr]
    lib.zref_gl_angles.argtypes = [_c_size, _ptr]
    lib.zref_plm_table.argtypes = [_c_int, _c_int, _c_size, _ptr, _c_size, _ptr]
    lib.zref_radial_table.argtypes = [_c_int, _c_size, _ptr, _c_size, _ptr]
    lib.zref_zt_evaluate.argtypes = [_c_int, _c_int, _c_int, _c_size, _ptr, _ptr, _c_size, _ptr,
                                     _c_size, _ptr, _c_size, _ptr]
    lib.zref_st_evaluate.argtypes = [_c_int, _c_int, _c_size, _ptr, _ptr, _c_size, _ptr, _c_size, _ptr]
    lib.zref_zt_power_spectrum.argtypes = [_c_int, _c_int, _c_size, _ptr, _ptr]
    _d = ctypes.c_double
    lib.zref_wignerd.argtypes = [_c_size, _ptr]
    lib.zref_euler_angles.argtypes = [_ptr, _ptr]
    lib.zref_rotate_many.argtypes = [_c_int, _c_size, _ptr, _c_size, _c_size, _ptr, _c_int]
    lib.zref_rotate.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_size, _ptr, _d, _d, _d, _c_int]
    lib.zref_polar_rotate.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_size, _ptr, _d, _c_int]
    return lib


ORACLE = _load_oracle()
REF = _load_ref()


# ---- shapes (include/zernike_glq_transformer.hpp:97-126,
#      include/sh_glq_transformer.hpp:48-195) ----
def zt_shape(order: int) -> tuple[int, int, int]:
    lon = 2 * order - min(1, order)
    return (lon, order + 1, (order + 1 + 7) // 8 * 8)


def zt_num_coeffs(order: int) -> int:
    return (order + 1) * (order + 3) * (2 * order + 1) // 24


def st_shape(order: int, layout: int = 0) -> tuple[int, int]:
    lon = 2 * order - min(1, order)
    if layout == 1:
        return (order, (lon + 7) // 8 * 8)
    return (lon, (order + 7) // 8 * 8)


def st_num_coeffs(order: int) -> int:
    return order * (order + 1) // 2


# ---- oracle primitives ----
def gl_nodes_weights(n: int):
    z = np.zeros(n)
    w = np.zeros(n)
    ORACLE.zo_gl_nodes_weights(n, _p(z), _p(w))
    return z, w


def gl_packed(n: int):
    nu = (n + 1) // 2
    z = np.zeros(nu)
    w = np.zeros(nu)
    ORACLE.zo_gl_packed(n, _p(z), _p(w))
    return z, w


def plm_table(sh_norm: int, sh_phase: int, order: int, z: np.ndarray) -> np.ndarray:
    z = np.ascontiguousarray(z, dtype=np.float64)
    out = np.zeros((order * (order + 1) // 2, z.size))
    ORACLE.zo_plm_table(sh_norm, sh_phase, order, _p(z), z.size, _p(out))
    return out


def radial_table(zernike_norm: int, order: int, r: np.ndarray) -> np.ndarray:
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(((order + 1) ** 2 // 4, r.size))
    ORACLE.zo_radial_table(zernike_norm, order, _p(r), r.size, _p(out))
    return out


def rdft_forward(x: np.ndarray, nfreq: int, scale: float = 1.0) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    cols = x.size // n
    out = np.zeros((nfreq, cols, 2))
    ORACLE.zo_rdft_forward(n, cols, _p(x), nfreq, scale, _p(out))
    return out


def rdft_backward(x: np.ndarray, n: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    nfreq, cols = x.shape[0], x.shape[1]
    out = np.zeros((n, cols))
    ORACLE.zo_rdft_backward(n, cols, _p(x), nfreq, _p(out))
    return out


def zt_forward(grid: np.ndarray, order: int, exp_order: int | None = None,
               zernike_norm: int = 1, sh_norm: int = 0, sh_phase: int = -1,
               out: np.ndarray | None = None) -> np.ndarray:
    exp_order = order if exp_order is None else exp_order
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    assert grid.size == np.prod(zt_shape(order))
    if out is None:
        out = np.zeros((zt_num_coeffs(exp_order), 2))
    ORACLE.zo_zt_forward(zernike_norm, sh_norm, sh_phase, order, _p(grid), exp_order, _p(out))
    return out


def zt_backward(coeffs: np.ndarray, order: int, exp_order: int | None = None,
                zernike_norm: int = 1, sh_norm: int = 0, sh_phase: int = -1) -> np.ndarray:
    exp_order = order if exp_order is None else exp_order
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    assert coeffs.size >= 2 * zt_num_coeffs(min(exp_order, order))
    grid = np.zeros(zt_shape(order))
    ORACLE.zo_zt_backward(zernike_norm, sh_norm, sh_phase, exp_order, _p(coeffs), order, _p(grid))
    return grid


def st_forward(grid: np.ndarray, order: int, exp_order: int | None = None,
               sh_norm: int = 0, sh_phase: int = -1, layout: int = 0,
               out: np.ndarray | None = None) -> np.ndarray:
    exp_order = order if exp_order is None else exp_order
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    assert grid.size == np.prod(st_shape(order, layout))
    if out is None:
        out = np.zeros((st_num_coeffs(exp_order), 2))
    ORACLE.zo_st_forward(sh_norm, sh_phase, layout, order, _p(grid), exp_order, _p(out))
    return out


def st_backward(coeffs: np.ndarray, order: int, exp_order: int | None = None,
                sh_norm: int = 0, sh_phase: int = -1, layout: int = 0) -> np.ndarray:
    exp_order = order if exp_order is None else exp_order
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    grid = np.zeros(st_shape(order, layout))
    ORACLE.zo_st_backward(sh_norm, sh_phase, layout, exp_order, _p(coeffs), order, _p(grid))
    return grid


def st_row_skipping_expansion(zcoeffs: np.ndarray, zorder: int, n: int, order: int) -> np.ndarray:
    """The full SH expansion that zest's row-skipping backward overload
    transforms when given row n of a Zernike expansion of order zorder and a
    grid of the given order (include/sh_glq_transformer.hpp:690-699).

    zest views row n's data (n + 1 rows of one parity, row l at (l*l)//4 in
    RowSkippingTriangleLayout, include/layout.hpp:334-370) truncated to
    o = min(n + 1, order); the truncated view's rows are l = (o+1) % 2, ...,
    o - 1 in steps of 2 (ParityIndexRange, include/indexing.hpp:226-227), so a
    truncation that changes the parity of the order reads the same data as rows
    of the other parity. Returns the (st_num_coeffs(o), 2) array with those rows
    filled and the others zero."""
    if not 0 <= n < zorder:
        raise ValueError("row n out of range")
    zcoeffs = np.asarray(zcoeffs, dtype=np.float64)
    o = min(n + 1, order)
    row = zcoeffs[zt_num_coeffs(n):zt_num_coeffs(n + 1)]
    out = np.zeros((st_num_coeffs(o), 2))
    for l in range((o + 1) % 2, o, 2):
        start = l * l // 4
        out[l * (l + 1) // 2:l * (l + 1) // 2 + l + 1] = row[start:start + l + 1]
    return out


# ---- evaluation on arbitrary grids (zt/st::GridEvaluator,
#      include/grid_evaluator.hpp:104-133, 227-265) ----
def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def zt_evaluate(coeffs: np.ndarray, order: int, lon, colat, r, zernike_norm: int = 1,
                sh_norm: int = 0, sh_phase: int = -1) -> np.ndarray:
    """Values [len(lon)][len(colat)][len(r)] of one Zernike expansion."""
    coeffs, lon, colat, r = _f64(coeffs), _f64(lon), _f64(colat), _f64(r)
    assert coeffs.size >= 2 * zt_num_coeffs(order)
    out = np.zeros((lon.size, colat.size, r.size))
    ORACLE.zo_zt_evaluate(zernike_norm, sh_norm, sh_phase, order, _p(coeffs), _p(lon), lon.size,
                          _p(colat), colat.size, _p(r), r.size, _p(out))
    return out


def st_evaluate(coeffs: np.ndarray, order: int, lon, colat, sh_norm: int = 0,
                sh_phase: int = -1) -> np.ndarray:
    """Values [len(lon)][len(colat)] of one spherical-harmonic expansion."""
    coeffs, lon, colat = _f64(coeffs), _f64(lon), _f64(colat)
    assert coeffs.size >= 2 * st_num_coeffs(order)
    out = np.zeros((lon.size, colat.size))
    ORACLE.zo_st_evaluate(sh_norm, sh_phase, order, _p(coeffs), _p(lon), lon.size, _p(colat),
                          colat.size, _p(out))
    return out


# ---- power spectra (include/power_spectra.hpp:43-141) ----
def st_power_spectrum(a: np.ndarray, order: int, sh_norm: int = 0, b: np.ndarray | None = None):
    """[order] (cross) power spectrum of SH coefficients a (and b)."""
    a = _f64(a)
    b = a if b is None else _f64(b)
    out = np.zeros(order)
    ORACLE.zo_st_power_spectrum(sh_norm, order, _p(a), _p(b), _p(out))
    return out


def zt_power_spectrum(a: np.ndarray, order: int, sh_norm: int = 0, b: np.ndarray | None = None):
    """[((order+1)^2)//4] (cross) power spectrum per (n, l) of Zernike coefficients."""
    a = _f64(a)
    b = a if b is None else _f64(b)
    out = np.zeros((order + 1) ** 2 // 4)
    ORACLE.zo_zt_power_spectrum(sh_norm, order, _p(a), _p(b), _p(out))
    return out


# ---- rotations (include/rotor.hpp, include/wignerd_collection.hpp) ----
ZERNIKE_KIND, SH_KIND = 0, 1


def wignerd_size(order: int) -> int:
    return order * (order + 1) * (2 * order + 1) // 6


def wignerd_table(order: int) -> np.ndarray:
    """Wigner d^l_{m1,m2}(pi/2), zest's layout [l(l+1)(2l+1)/6 + (l+1) m1 + m2]."""
    out = np.zeros(wignerd_size(order))
    ORACLE.zo_wignerd_table(order, _p(out))
    return out


def rotate(coeffs: np.ndarray, order: int, kind: int, angles, rtype: int = 0) -> np.ndarray:
    """A rotated copy (Euler angles alpha, beta, gamma; rtype 0 object, 1 coordinate)."""
    c = np.array(coeffs, dtype=np.float64, order="C", copy=True)
    ORACLE.zo_rotate(kind, order, _p(c), float(angles[0]), float(angles[1]), float(angles[2]), rtype)
    return c


def polar_rotate(coeffs: np.ndarray, order: int, kind: int, angle: float, rtype: int = 0) -> np.ndarray:
    c = np.array(coeffs, dtype=np.float64, order="C", copy=True)
    ORACLE.zo_polar_rotate(kind, order, _p(c), float(angle), rtype)
    return c


# ---- the upstream zest library (compiled reference) ----
class ZestZernike:
    """One zest ``zt::GLQTransformer`` instance (keeps its tables between calls)."""

    def __init__(self, order: int, zernike_norm: int = 1, sh_norm: int = 0, sh_phase: int = -1):
        if REF is None:
            raise RuntimeError("oracle/_ref/libzest_ref.so is not built")
        self.order = order
        self._h = REF.zref_zt_create(zernike_norm, sh_norm, sh_phase, order)
        if not self._h:
            raise MemoryError("zest transformer allocation failed")

    def __del__(self):
        if getattr(self, "_h", None):
            REF.zref_zt_destroy(self._h)
            self._h = None

    def forward(self, grid: np.ndarray, order: int | None = None, exp_order: int | None = None,
                out: np.ndarray | None = None) -> np.ndarray:
        order = self.order if order is None else order
        exp_order = order if exp_order is None else exp_order
        grid = np.ascontiguousarray(grid, dtype=np.float64)
        if out is None:
            out = np.zeros((zt_num_coeffs(exp_order), 2))
        REF.zref_zt_forward(self._h, _p(grid), order, _p(out), exp_order)
        return out

    def backward(self, coeffs: np.ndarray, order: int | None = None,
                 exp_order: int | None = None, out: np.ndarray | None = None) -> np.ndarray:
        order = self.order if order is None else order
        exp_order = order if exp_order is None else exp_order
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        if out is None:
            out = np.zeros(zt_shape(order))
        REF.zref_zt_backward(self._h, _p(coeffs), exp_order, _p(out), order)
        return out


class ZestSH:
    """One zest ``st::GLQTransformer`` instance."""

    def __init__(self, order: int, sh_norm: int = 0, sh_phase: int = -1, layout: int = 0):
        if REF is None:
            raise RuntimeError("oracle/_ref/libzest_ref.so is not built")
        self.order = order
        self.layout = layout
        self._h = REF.zref_st_create(sh_norm, sh_phase, layout, order)
        if not self._h:
            raise MemoryError("zest transformer allocation failed")

    def __del__(self):
        if getattr(self, "_h", None):
            REF.zref_st_destroy(self._h)
            self._h = None

    def forward(self, grid: np.ndarray, order: int | None = None, exp_order: int | None = None,
                out: np.ndarray | None = None) -> np.ndarray:
        order = self.order if order is None else order
        exp_order = order if exp_order is None else exp_order
        grid = np.ascontiguousarray(grid, dtype=np.float64)
        if out is None:
            out = np.zeros((st_num_coeffs(exp_order), 2))
        REF.zref_st_forward(self._h, _p(grid), order, _p(out), exp_order)
        return out

    def backward(self, coeffs: np.ndarray, order: int | None = None,
                 exp_order: int | None = None, out: np.ndarray | None = None) -> np.ndarray:
        order = self.order if order is None else order
        exp_order = order if exp_order is None else exp_order
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        if out is None:
            out = np.zeros(st_shape(order, self.layout))
        REF.zref_st_backward(self._h, _p(coeffs), exp_order, _p(out), order)
        return out

    def backward_zernike_row(self, zcoeffs: np.ndarray, zorder: int, n: int,
                             order: int | None = None) -> np.ndarray:
        """zest's row-skipping backward overload on row n of a Zernike
        expansion of order zorder (only l of n's parity present)."""
        order = self.order if order is None else order
        if not 0 <= n < zorder:
            raise ValueError("row n out of range")
        zcoeffs = np.ascontiguousarray(zcoeffs, dtype=np.float64)
        if zcoeffs.shape != (zt_num_coeffs(zorder), 2):
            raise ValueError("zcoeffs must have shape (zt_num_coeffs(zorder), 2)")
        out = np.zeros(st_shape(order, self.layout))
        REF.zref_st_backward_zernike_row(self._h, _p(zcoeffs), zorder, n, _p(out), order)
        return out


def ref_gl_nodes_weights(n: int, packed: bool = False):
    size = (n + 1) // 2 if packed else n
    z = np.zeros(size)
    w = np.zeros(size)
    REF.zref_gl_nodes_weights(n, int(packed), _p(z), _p(w))
    return z, w


def ref_gl_angles(n: int) -> np.ndarray:
    a = np.zeros(n)
    REF.zref_gl_angles(n, _p(a))
    return a


def ref_plm_table(sh_norm: int, sh_phase: int, order: int, z: np.ndarray) -> np.ndarray:
    z = np.ascontiguousarray(z, dtype=np.float64)
    out = np.zeros((order * (order + 1) // 2, z.size))
    REF.zref_plm_table(sh_norm, sh_phase, order, _p(z), z.size, _p(out))
    return out


def ref_radial_table(zernike_norm: int, order: int, r: np.ndarray) -> np.ndarray:
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(((order + 1) ** 2 // 4, r.size))
    REF.zref_radial_table(zernike_norm, order, _p(r), r.size, _p(out))
    return out


def ref_zt_evaluate(coeffs: np.ndarray, order: int, lon, colat, r, zernike_norm: int = 1,
                    sh_norm: int = 0, sh_phase: int = -1) -> np.ndarray:
    coeffs, lon, colat, r = _f64(coeffs), _f64(lon), _f64(colat), _f64(r)
    out = np.zeros((lon.size, colat.size, r.size))
    if out.size:
        REF.zref_zt_evaluate(zernike_norm, sh_norm, sh_phase, order, _p(coeffs), _p(lon), lon.size,
                             _p(colat), colat.size, _p(r), r.size, _p(out))
    return out


def ref_st_evaluate(coeffs: np.ndarray, order: int, lon, colat, sh_norm: int = 0,
                    sh_phase: int = -1) -> np.ndarray:
    coeffs, lon, colat = _f64(coeffs), _f64(lon), _f64(colat)
    out = np.zeros((lon.size, colat.size))
    if out.size:
        REF.zref_st_evaluate(sh_norm, sh_phase, order, _p(coeffs), _p(lon), lon.size, _p(colat),
                             colat.size, _p(out))
    return out


def ref_zt_power_spectrum(coeffs: np.ndarray, order: int, zernike_norm: int = 1, sh_norm: int = 0,
                          fill: float = np.nan) -> np.ndarray:
    """zest's zt::power_spectrum; the l = n entries it skips keep `fill`."""
    coeffs = _f64(coeffs)
    out = np.full((order + 1) ** 2 // 4, fill)
    REF.zref_zt_power_spectrum(zernike_norm, sh_norm, order, _p(coeffs), _p(out))
    return out


def ref_wignerd_table(order: int) -> np.ndarray:
    out = np.zeros(wignerd_size(order))
    REF.zref_wignerd(order, _p(out))
    return out


def ref_rotate(coeffs, order, kind, angles, rtype=0, zernike_norm=1, sh_norm=0, sh_phase=-1):
    c = np.array(coeffs, dtype=np.float64, order="C", copy=True)
    REF.zref_rotate(kind, zernike_norm, sh_norm, sh_phase, order, _p(c), float(angles[0]),
                    float(angles[1]), float(angles[2]), rtype)
    return c


def ref_polar_rotate(coeffs, order, kind, angle, rtype=0, zernike_norm=1, sh_norm=0, sh_phase=-1):
    c = np.array(coeffs, dtype=np.float64, order="C", copy=True)
    REF.zref_polar_rotate(kind, zernike_norm, sh_norm, sh_phase, order, _p(c), float(angle), rtype)
    return c


def _cvec(a, n: int) -> np.ndarray:
    a = np.array(a, dtype=np.complex128, order="C", copy=True)
    if a.shape != (n,):
        raise ValueError(f"expected {n} complex values")
    return a


def ref_rotate_l(vector, exp_gamma, exp_beta, exp_alpha) -> np.ndarray:
    """zest::rotate_l on one degree-l vector of l + 1 complex values."""
    n = len(vector)
    v = _cvec(vector, n)
    g, b, a = _cvec(exp_gamma, n), _cvec(exp_beta, n), _cvec(exp_alpha, n)
    REF.zref_rotate_l(n - 1, _p(v), _p(g), _p(b), _p(a))
    return v


def ref_polar_rotate_l(vector, exp_alpha) -> np.ndarray:
    """zest::polar_rotate_l on one degree-l vector of l + 1 complex values."""
    n = len(vector)
    v = _cvec(vector, n)
    a = _cvec(exp_alpha, n)
    REF.zref_polar_rotate_l(n - 1, _p(v), _p(a))
    return v


def ref_euler_angles(matrix) -> np.ndarray:
    m = np.ascontiguousarray(matrix, dtype=np.float64).reshape(3, 3)
    out = np.zeros(3)
    REF.zref_euler_angles(_p(m), _p(out))
    return out
