"""zest_amd: MI355X-native Gauss-Legendre-quadrature Zernike and spherical-
harmonic transforms.

The product is ``libzgl.so`` -- hand-written HIP kernels for gfx950 behind the
C ABI declared in ``include/zgl.h`` -- together with the C++20 drop-in headers
``include/zest_amd/zernike_glq_transformer.hpp`` and
``include/zest_amd/sh_glq_transformer.hpp``. This module is a thin ctypes binding
of that C ABI used by the tests and the benchmark; it moves no data itself and
has no CPU fallback: importing it without the built library raises.

Convention arguments take zest's enum values: ``zernike_norm`` 0 = normed,
1 = unnormed; ``sh_norm`` 0 = geo, 1 = qm; ``sh_phase`` -1 = none, +1 = cs; SH
``layout`` 0 = LonLat, 1 = LatLon.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzgl.so")

ZERNIKE, SH = 0, 1
FP64, FP32, FP32_REFINE, BF16X3 = 0, 1, 2, 3
MEM_DEVICE, MEM_HOST = 0, 1

_STATUS = {0: "success", 1: "invalid value", 2: "out of memory", 3: "HIP error", 4: "not supported"}


class ZglError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"zgl {_STATUS.get(status, status)}: {message}")
        self.status = status


def _load():
    # torch (ROCm build) bundles its own HIP runtime with the same soname as
    # the system one libzgl.so links. Importing torch first makes libzgl bind
    # to that already-loaded runtime; loading libzgl first would pull in a
    # second HIP runtime, which then finds no device in this process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, pd = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_ssize_t
    lib.zgl_plan_create.argtypes = [ctypes.POINTER(vp), i32, i32, i32, i32, i32, sz, i32]
    lib.zgl_plan_create.restype = i32
    lib.zgl_plan_destroy.argtypes = [vp]
    lib.zgl_plan_destroy.restype = i32
    lib.zgl_plan_resize.argtypes = [vp, sz]
    lib.zgl_plan_resize.restype = i32
    lib.zgl_plan_order.argtypes = [vp]
    lib.zgl_plan_order.restype = sz
    lib.zgl_grid_size.argtypes = [vp, sz]
    lib.zgl_grid_size.restype = sz
    lib.zgl_coeff_size.argtypes = [vp, sz]
    lib.zgl_coeff_size.restype = sz
    lib.zgl_forward.argtypes = [vp, vp, sz, vp, sz, sz, pd, pd, i32, vp]
    lib.zgl_forward.restype = i32
    lib.zgl_backward.argtypes = [vp, vp, sz, vp, sz, sz, pd, pd, i32, vp]
    lib.zgl_backward.restype = i32
    lib.zgl_forward_noalloc.argtypes = [vp, vp, sz, vp, sz, sz, pd, pd, vp]
    lib.zgl_forward_noalloc.restype = i32
    lib.zgl_forward_ragged.argtypes = [vp, ctypes.POINTER(Job), sz, vp]
    lib.zgl_forward_ragged.restype = i32
    lib.zgl_backward_ragged.argtypes = [vp, ctypes.POINTER(Job), sz, vp]
    lib.zgl_backward_ragged.restype = i32
    lib.zgl_ragged_lane_costs.argtypes = [vp, ctypes.POINTER(ctypes.c_double), i32, ctypes.POINTER(i32)]
    lib.zgl_ragged_lane_costs.restype = i32
    lib.zgl_backward_adjoint.argtypes = [vp, vp, sz, vp, sz, sz, pd, pd, i32, vp]
    lib.zgl_backward_adjoint.restype = i32
    lib.zgl_forward_adjoint.argtypes = [vp, vp, sz, vp, sz, sz, pd, pd, i32, vp]
    lib.zgl_forward_adjoint.restype = i32
    lib.zgl_forward_grid32.argtypes = [vp, vp, sz, vp, sz, sz, pd, pd, i32, vp]
    lib.zgl_forward_grid32.restype = i32
    lib.zgl_backward_grid32.argtypes = [vp, vp, sz, vp, sz, sz, pd, pd, i32, vp]
    lib.zgl_evaluate.argtypes = [vp, vp, sz, vp, sz, vp, sz, vp, sz, vp, sz, pd, pd, i32, vp]
    lib.zgl_evaluate.restype = i32
    lib.zgl_power_spectrum.argtypes = [vp, vp, sz, vp, sz, pd, pd, i32, vp]
    lib.zgl_power_spectrum.restype = i32
    lib.zgl_cross_power_spectrum.argtypes = [vp, vp, vp, sz, vp, sz, pd, pd, i32, vp]
    lib.zgl_cross_power_spectrum.restype = i32
    lib.zgl_rotate.argtypes = [vp, vp, sz, vp, i32, sz, pd, i32, vp]
    lib.zgl_rotate.restype = i32
    lib.zgl_polar_rotate.argtypes = [vp, vp, sz, vp, i32, sz, pd, i32, vp]
    lib.zgl_polar_rotate.restype = i32
    lib.zgl_wignerd_pi2.argtypes = [vp, sz, vp, i32, vp]
    lib.zgl_wignerd_pi2.restype = i32
    lib.zgl_euler_angles_from_rotation_matrix.argtypes = [vp, vp]
    lib.zgl_euler_angles_from_rotation_matrix.restype = None
    lib.zgl_backward_grid32.restype = i32
    lib.zgl_last_error_string.argtypes = []
    lib.zgl_last_error_string.restype = ctypes.c_char_p
    lib.zgl_plan_autotune.argtypes = [vp, i32]
    lib.zgl_plan_autotune.restype = i32
    lib.zgl_plan_tuning.argtypes = [vp, ctypes.POINTER(i32), i32, ctypes.POINTER(i32)]
    lib.zgl_plan_tuning.restype = i32
    lib.zgl_profile_enable.argtypes = [vp, i32]
    lib.zgl_profile_enable.restype = i32
    lib.zgl_profile_stats.argtypes = [vp, ctypes.POINTER(StageStats), i32, ctypes.POINTER(i32)]
    lib.zgl_profile_stats.restype = i32
    lib.zgl_set_device_allocator.argtypes = [DEVICE_ALLOC_FN, DEVICE_FREE_FN, vp]
    lib.zgl_set_device_allocator.restype = i32
    return lib


# zgl_device_alloc_fn / zgl_device_free_fn of include/zgl.h
DEVICE_ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)
DEVICE_FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)


class Job(ctypes.Structure):
    """zgl_job of include/zgl.h."""
    _fields_ = [("order", ctypes.c_size_t), ("grid", ctypes.c_void_p), ("coeffs", ctypes.c_void_p)]


class StageStats(ctypes.Structure):
    """zgl_stage_stats of include/zgl.h."""
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_longlong),
                ("instances", ctypes.c_longlong), ("total_ms", ctypes.c_double)]


LIB = _load()

# The C ABI entry points of include/zgl.h.
EXPORTS = (
    "zgl_plan_create", "zgl_plan_destroy", "zgl_plan_resize", "zgl_plan_order",
    "zgl_grid_size", "zgl_coeff_size", "zgl_forward", "zgl_backward",
    "zgl_forward_grid32", "zgl_backward_grid32", "zgl_backward_adjoint", "zgl_forward_adjoint",
    "zgl_forward_noalloc",
    "zgl_forward_ragged", "zgl_backward_ragged", "zgl_ragged_lane_costs",
    "zgl_plan_autotune", "zgl_plan_tuning", "zgl_evaluate",
    "zgl_power_spectrum", "zgl_cross_power_spectrum", "zgl_rotate", "zgl_polar_rotate",
    "zgl_wignerd_pi2", "zgl_euler_angles_from_rotation_matrix",
    "zgl_profile_enable", "zgl_profile_stats", "zgl_last_error_string",
    "zgl_set_device_allocator",
)


def _check(status: int):
    if status != 0:
        raise ZglError(status, LIB.zgl_last_error_string().decode())


def zt_shape(order: int) -> tuple[int, int, int]:
    """BallGLQGrid shape [lon][lat][rad] (include/zernike_glq_transformer.hpp:70-126)."""
    return (2 * order - min(1, order), order + 1, (order + 1 + 7) // 8 * 8)


def zt_num_coeffs(order: int) -> int:
    """RealZernikeExpansion size (include/layout.hpp:400-406)."""
    return (order + 1) * (order + 3) * (2 * order + 1) // 24


def st_shape(order: int, layout: int = 0) -> tuple[int, int]:
    """SphereGLQGrid shape (include/sh_glq_transformer.hpp:44-195)."""
    lon = 2 * order - min(1, order)
    if layout == 1:
        return (order, (lon + 7) // 8 * 8)
    return (lon, (order + 7) // 8 * 8)


def st_num_coeffs(order: int) -> int:
    return order * (order + 1) // 2


def _numel(x) -> int:
    return int(x.size) if isinstance(x, np.ndarray) else int(x.numel())


def _need(x, name: str, batch: int, stride: int, per: int):
    """Raise ValueError unless `x` holds `batch` instances of `per` elements
    `stride` elements apart (the last one needs only `per`)."""
    if batch <= 0 or per <= 0:
        return
    if stride < 0:
        raise ValueError(f"{name}: negative instance stride")
    need = (batch - 1) * stride + per
    if _numel(x) < need:
        raise ValueError(f"{name} holds {_numel(x)} elements, {need} needed "
                         f"({batch} instances of {per}, {stride} apart)")


def _ptr_and_mem(x, allow_f32: bool = False):
    """(pointer, memory kind, is_float32) of a torch tensor or numpy array."""
    if isinstance(x, np.ndarray):
        if not x.flags.c_contiguous:
            raise ValueError("arrays must be C-contiguous")
        f32 = x.dtype == np.float32
        if not (x.dtype == np.float64 or (allow_f32 and f32)):
            raise ValueError("arrays must be float64" + (" or float32" if allow_f32 else ""))
        return x.ctypes.data, MEM_HOST, f32
    import torch
    if not isinstance(x, torch.Tensor):
        raise TypeError("expected a numpy array or a torch tensor")
    if not x.is_contiguous():
        raise ValueError("tensors must be contiguous")
    f32 = x.dtype == torch.float32
    if not (x.dtype == torch.float64 or (allow_f32 and f32)):
        raise ValueError("tensors must be float64" + (" or float32" if allow_f32 else ""))
    return x.data_ptr(), (MEM_DEVICE if x.is_cuda else MEM_HOST), f32


class Plan:
    """One transform plan (the counterpart of one zest GLQTransformer)."""

    def __init__(self, kind: int, order: int, zernike_norm: int = 1, sh_norm: int = 0,
                 sh_phase: int = -1, layout: int = 0, precision: int = FP64):
        self.kind = kind
        self.layout = layout if kind == SH else 0
        self._h = ctypes.c_void_p()
        # held by the plan: module globals may already be cleared when a plan
        # is collected at interpreter exit
        self._lib = LIB
        _check(LIB.zgl_plan_create(ctypes.byref(self._h), kind, zernike_norm, sh_norm,
                                   sh_phase, layout, order, precision))

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.zgl_plan_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def order(self) -> int:
        return LIB.zgl_plan_order(self._h)

    def resize(self, order: int):
        _check(LIB.zgl_plan_resize(self._h, order))

    def grid_size(self, order: int) -> int:
        return LIB.zgl_grid_size(self._h, order)

    def coeff_size(self, order: int) -> int:
        return LIB.zgl_coeff_size(self._h, order)

    def autotune(self, enable: bool = True):
        """Time the candidate tile shapes on first use of each order (zgl_plan_autotune)."""
        _check(LIB.zgl_plan_autotune(self._h, int(enable)))

    def tuning(self) -> list:
        """[(order, bytes per real, bytes per grid value, stage name, choice), ...]
        (zgl_plan_tuning)."""
        n = ctypes.c_int()
        _check(LIB.zgl_plan_tuning(self._h, None, 0, ctypes.byref(n)))
        arr = (ctypes.c_int * (5 * max(1, n.value)))()
        _check(LIB.zgl_plan_tuning(self._h, arr, n.value, ctypes.byref(n)))
        # the library's stage order (kStageNames, zgl_api.hip)
        names = ("fft_forward", "legendre_forward", "radial_forward", "radial_backward",
                 "legendre_backward", "fft_backward", "eval_tables", "eval_radial", "eval_legendre",
                 "eval_longitude", "rotate", "rotate_wide", "polar_rotate", "power_spectrum",
                 "legendre_radial_forward", "radial_legendre_backward")
        return [(arr[5 * i], arr[5 * i + 1], arr[5 * i + 2], names[arr[5 * i + 3]], arr[5 * i + 4])
                for i in range(n.value)]

    def profile(self, enable: bool = True):
        """Bracket every kernel launch with HIP events (zgl_profile_enable)."""
        _check(LIB.zgl_profile_enable(self._h, int(enable)))

    def stage_stats(self) -> dict:
        """{stage: (launches, instances, total_ms)} since the last call."""
        arr = (StageStats * 16)()
        n = ctypes.c_int()
        _check(LIB.zgl_profile_stats(self._h, arr, 16, ctypes.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].instances, arr[i].total_ms)
                for i in range(min(n.value, 16))}

    def _check_sizes(self, grid, coeffs, grid_order, exp_order, batch, grid_stride, coeff_stride):
        """The buffers hold `batch` instances at the strides the C ABI will use."""
        nt = min(exp_order, grid_order)
        gsz = self.grid_size(grid_order)
        _need(grid, "grid", batch, grid_stride or gsz, gsz)
        _need(coeffs, "coefficients", batch, coeff_stride or 2 * self.coeff_size(exp_order),
              2 * self.coeff_size(nt))

    def grid_shape(self, order: int):
        return zt_shape(order) if self.kind == ZERNIKE else st_shape(order, self.layout)

    def forward(self, grid, coeffs, grid_order: int, exp_order: int | None = None,
                batch: int = 1, grid_stride: int = 0, coeff_stride: int = 0, stream=None):
        """Enqueue forward transforms of `batch` instances (zgl_forward, or
        zgl_forward_grid32 for a float32 grid)."""
        exp_order = grid_order if exp_order is None else exp_order
        gp, gm, g32 = _ptr_and_mem(grid, allow_f32=True)
        cp, cm, _ = _ptr_and_mem(coeffs)
        if gm != cm:
            raise ValueError("grid and coefficients must live in the same memory")
        self._check_sizes(grid, coeffs, grid_order, exp_order, batch, grid_stride, coeff_stride)
        fn = LIB.zgl_forward_grid32 if g32 else LIB.zgl_forward
        _check(fn(self._h, gp, grid_order, cp, exp_order, batch, grid_stride,
                  coeff_stride, gm, _stream_ptr(stream, gm)))

    def backward(self, coeffs, grid, grid_order: int, exp_order: int | None = None,
                 batch: int = 1, coeff_stride: int = 0, grid_stride: int = 0, stream=None):
        """Enqueue backward transforms of `batch` instances (zgl_backward, or
        zgl_backward_grid32 for a float32 grid)."""
        exp_order = grid_order if exp_order is None else exp_order
        cp, cm, _ = _ptr_and_mem(coeffs)
        gp, gm, g32 = _ptr_and_mem(grid, allow_f32=True)
        if gm != cm:
            raise ValueError("grid and coefficients must live in the same memory")
        self._check_sizes(grid, coeffs, grid_order, exp_order, batch, grid_stride, coeff_stride)
        fn = LIB.zgl_backward_grid32 if g32 else LIB.zgl_backward
        _check(fn(self._h, cp, exp_order, gp, grid_order, batch, coeff_stride,
                  grid_stride, gm, _stream_ptr(stream, gm)))


    def forward_noalloc(self, grid, coeffs, grid_order: int, exp_order: int | None = None,
                        batch: int = 1, grid_stride: int = 0, coeff_stride: int = 0, stream=None):
        """zgl_forward_noalloc: forward() with no allocation, scratch from the
        library's static device array (device tensors, FP64 grids; the plan
        must hold the tables of grid_order)."""
        exp_order = grid_order if exp_order is None else exp_order
        gp, gm, _ = _ptr_and_mem(grid)
        cp, cm, _ = _ptr_and_mem(coeffs)
        if gm != MEM_DEVICE or cm != MEM_DEVICE:
            raise ValueError("the no-allocation forward takes device tensors")
        self._check_sizes(grid, coeffs, grid_order, exp_order, batch, grid_stride, coeff_stride)
        _check(LIB.zgl_forward_noalloc(self._h, gp, grid_order, cp, exp_order, batch, grid_stride,
                                       coeff_stride, _stream_ptr(stream, MEM_DEVICE)))

    def evaluate(self, coeffs, order: int, longitudes, colatitudes, radii=None, values=None,
                 batch: int = 1, coeff_stride: int = 0, value_stride: int = 0, stream=None):
        """Values of `batch` expansions of `order` on the grid longitudes x
        colatitudes (x radii for a Zernike plan), zgl_evaluate. Every array
        lives in the same memory (torch tensors on the GPU, or host arrays);
        returns `values`, allocated like `coeffs` when not given, shaped
        (batch, nlon, nlat[, nrad])."""
        zt = self.kind == ZERNIKE
        if zt and radii is None:
            raise ValueError("a Zernike plan needs radii")
        cp, cm, _ = _ptr_and_mem(coeffs)
        lp, lm, _ = _ptr_and_mem(longitudes)
        tp, tm, _ = _ptr_and_mem(colatitudes)
        nlon, nlat = int(np.prod(longitudes.shape)), int(np.prod(colatitudes.shape))
        rp, nrad, mems = 0, 1, {cm, lm, tm}
        if zt:
            rp, rm, _ = _ptr_and_mem(radii)
            nrad = int(np.prod(radii.shape))
            mems.add(rm)
        shape = (batch, nlon, nlat) + ((nrad,) if zt else ())
        if values is None:
            if isinstance(coeffs, np.ndarray):
                values = np.empty(shape)
            else:
                import torch
                values = torch.empty(shape, dtype=torch.float64, device=coeffs.device)
        vp, vm, _ = _ptr_and_mem(values)
        mems.add(vm)
        if len(mems) != 1:
            raise ValueError("coefficients, points and values must live in the same memory")
        csz = 2 * self.coeff_size(order)
        _need(coeffs, "coefficients", batch, coeff_stride or csz, csz)
        npts = nlon * nlat * nrad
        _need(values, "values", batch, value_stride or npts, npts)
        _check(LIB.zgl_evaluate(self._h, cp, order, lp, nlon, tp, nlat, rp or None, nrad, vp, batch,
                                coeff_stride, value_stride, vm, _stream_ptr(stream, vm)))
        return values

    def power_spectrum(self, coeffs, order: int, out=None, batch: int = 1, other=None,
                       coeff_stride: int = 0, out_stride: int = 0, stream=None):
        """(Cross) power spectra of `batch` expansions (zgl_power_spectrum, or
        zgl_cross_power_spectrum with `other`); returns `out`, allocated like
        `coeffs` when not given, shaped (batch, entries)."""
        nent = (order + 1) ** 2 // 4 if self.kind == ZERNIKE else order
        if out is None:
            if isinstance(coeffs, np.ndarray):
                out = np.empty((batch, nent))
            else:
                import torch
                out = torch.empty((batch, nent), dtype=torch.float64, device=coeffs.device)
        ap, am, _ = _ptr_and_mem(coeffs)
        op, om, _ = _ptr_and_mem(out)
        csz = 2 * self.coeff_size(order)
        _need(coeffs, "coefficients", batch, coeff_stride or csz, csz)
        if other is not None:
            _need(other, "coefficients", batch, coeff_stride or csz, csz)
        _need(out, "output", batch, out_stride or nent, nent)
        if other is None:
            if am != om:
                raise ValueError("coefficients and output must live in the same memory")
            _check(LIB.zgl_power_spectrum(self._h, ap, order, op, batch, coeff_stride, out_stride, om,
                                          _stream_ptr(stream, om)))
        else:
            bp, bm, _ = _ptr_and_mem(other)
            if len({am, bm, om}) != 1:
                raise ValueError("coefficients and output must live in the same memory")
            _check(LIB.zgl_cross_power_spectrum(self._h, ap, bp, order, op, batch, coeff_stride,
                                                out_stride, om, _stream_ptr(stream, om)))
        return out

    def rotate(self, coeffs, order: int, angles, rotation_type: int = 0, batch: int = 1,
               coeff_stride: int = 0, stream=None):
        """In-place ZYZ rotation of `batch` expansions (zgl_rotate); `angles`
        holds (alpha, beta, gamma) per instance, in the coefficients' memory."""
        cp, cm, _ = _ptr_and_mem(coeffs)
        ap, am, _ = _ptr_and_mem(angles)
        if am != cm:
            raise ValueError("angles and coefficients must live in the same memory")
        csz = 2 * self.coeff_size(order)
        _need(coeffs, "coefficients", batch, coeff_stride or csz, csz)
        _need(angles, "angles", 1, 0, 3 * batch)
        _check(LIB.zgl_rotate(self._h, cp, order, ap, rotation_type, batch, coeff_stride, cm,
                              _stream_ptr(stream, cm)))
        return coeffs

    def polar_rotate(self, coeffs, order: int, angles, rotation_type: int = 0, batch: int = 1,
                     coeff_stride: int = 0, stream=None):
        """In-place rotation about z by angles[b] (zgl_polar_rotate)."""
        cp, cm, _ = _ptr_and_mem(coeffs)
        ap, am, _ = _ptr_and_mem(angles)
        if am != cm:
            raise ValueError("angles and coefficients must live in the same memory")
        csz = 2 * self.coeff_size(order)
        _need(coeffs, "coefficients", batch, coeff_stride or csz, csz)
        _need(angles, "angles", 1, 0, batch)
        _check(LIB.zgl_polar_rotate(self._h, cp, order, ap, rotation_type, batch, coeff_stride, cm,
                                    _stream_ptr(stream, cm)))
        return coeffs

    def wignerd_pi2(self, order: int) -> np.ndarray:
        """The plan's Wigner d(pi/2) table of degrees < order (zgl_wignerd_pi2), on the host."""
        out = np.zeros(order * (order + 1) * (2 * order + 1) // 6)
        _check(LIB.zgl_wignerd_pi2(self._h, order, out.ctypes.data, MEM_HOST, None))
        return out

    def _jobs(self, jobs):
        if isinstance(jobs, ctypes.Array):
            return jobs
        for order, grid, coeffs in jobs:
            if order:
                if grid is None or coeffs is None:
                    raise ValueError("a job of order > 0 needs a grid and coefficients")
                _need(grid, "job grid", 1, 0, self.grid_size(order))
                _need(coeffs, "job coefficients", 1, 0, 2 * self.coeff_size(order))
        return make_jobs(jobs)

    def forward_ragged(self, jobs, stream=None):
        """Forward transforms of [(order, grid, coeffs), ...] of uneven orders, or
        of an array from make_jobs (zgl_forward_ragged)."""
        arr = self._jobs(jobs)
        _check(LIB.zgl_forward_ragged(self._h, arr, len(jobs), _stream_ptr(stream, MEM_DEVICE)))

    def backward_ragged(self, jobs, stream=None):
        """Backward transforms of [(order, grid, coeffs), ...] or of an array from
        make_jobs (zgl_backward_ragged)."""
        arr = self._jobs(jobs)
        _check(LIB.zgl_backward_ragged(self._h, arr, len(jobs), _stream_ptr(stream, MEM_DEVICE)))

    def ragged_lane_costs(self) -> list:
        """Modelled work per unit of execution of the last ragged call: per
        resident workgroup of the costliest persistent launch, or per
        internal stream of the lanes pipeline."""
        n = ctypes.c_int()
        _check(LIB.zgl_ragged_lane_costs(self._h, None, 0, ctypes.byref(n)))
        arr = (ctypes.c_double * max(n.value, 1))()
        _check(LIB.zgl_ragged_lane_costs(self._h, arr, n.value, ctypes.byref(n)))
        return [arr[i] for i in range(n.value)]

    def backward_adjoint(self, grid, coeffs, grid_order: int, exp_order: int | None = None,
                         batch: int = 1, grid_stride: int = 0, coeff_stride: int = 0, stream=None):
        """coeffs = B^T grid, the transpose of the backward transform
        (zgl_backward_adjoint); arguments as forward()."""
        exp_order = grid_order if exp_order is None else exp_order
        gp, gm, _ = _ptr_and_mem(grid)
        cp, cm, _ = _ptr_and_mem(coeffs)
        if gm != cm:
            raise ValueError("grid and coefficients must live in the same memory")
        self._check_sizes(grid, coeffs, grid_order, exp_order, batch, grid_stride, coeff_stride)
        _check(LIB.zgl_backward_adjoint(self._h, gp, grid_order, cp, exp_order, batch, grid_stride,
                                        coeff_stride, gm, _stream_ptr(stream, gm)))

    def forward_adjoint(self, coeffs, grid, grid_order: int, exp_order: int | None = None,
                        batch: int = 1, coeff_stride: int = 0, grid_stride: int = 0, stream=None):
        """grid = F^T coeffs, the transpose of the forward transform
        (zgl_forward_adjoint); arguments as backward()."""
        exp_order = grid_order if exp_order is None else exp_order
        cp, cm, _ = _ptr_and_mem(coeffs)
        gp, gm, _ = _ptr_and_mem(grid)
        if gm != cm:
            raise ValueError("grid and coefficients must live in the same memory")
        self._check_sizes(grid, coeffs, grid_order, exp_order, batch, grid_stride, coeff_stride)
        _check(LIB.zgl_forward_adjoint(self._h, cp, exp_order, gp, grid_order, batch, coeff_stride,
                                       grid_stride, gm, _stream_ptr(stream, gm)))


def make_jobs(jobs):
    """zgl_job array of [(order, grid, coeffs), ...] with contiguous float64
    device tensors (None for order 0)."""
    arr = (Job * max(1, len(jobs)))()
    for i, (order, grid, coeffs) in enumerate(jobs):
        import torch
        for t in (grid, coeffs):
            if t is not None and (not isinstance(t, torch.Tensor) or not t.is_cuda
                                  or not t.is_contiguous() or t.dtype != torch.float64):
                raise ValueError("ragged jobs take contiguous float64 device tensors")
        arr[i].order = order
        arr[i].grid = grid.data_ptr() if grid is not None else None
        arr[i].coeffs = coeffs.data_ptr() if coeffs is not None else None
    return arr


def euler_angles_from_rotation_matrix(matrix) -> np.ndarray:
    """(alpha, beta, gamma) of a row-major 3x3 rotation matrix, as zest computes them."""
    m = np.ascontiguousarray(matrix, dtype=np.float64).reshape(3, 3)
    out = np.zeros(3)
    LIB.zgl_euler_angles_from_rotation_matrix(m.ctypes.data, out.ctypes.data)
    return out


class DeviceAllocator:
    """Routes the library's device allocations through two Python callables
    (zgl_set_device_allocator): ``alloc(nbytes) -> int address or 0`` and
    ``free(address, nbytes)``. A context manager: installed on entry, hipMalloc
    restored on exit. Memory still held by plans then goes back through
    ``free`` when the plans release it, so the callables must outlive it (the
    object keeps the C callbacks alive until it is garbage-collected).
    ``live`` counts the allocations made through it and not yet freed."""

    def __init__(self, alloc, free):
        self._alloc, self._free = alloc, free
        self.live = 0
        self.calls = 0

        def c_alloc(nbytes, _user):
            self.calls += 1
            try:
                p = int(self._alloc(int(nbytes)) or 0)
            except Exception:  # e.g. torch's OutOfMemoryError: report NULL
                p = 0
            if p:
                self.live += 1
            return p or None

        def c_free(ptr, nbytes, _user):
            self.live -= 1
            self._free(int(ptr), int(nbytes))

        self._c_alloc = DEVICE_ALLOC_FN(c_alloc)
        self._c_free = DEVICE_FREE_FN(c_free)

    def install(self):
        _check(LIB.zgl_set_device_allocator(self._c_alloc, self._c_free, None))

    def __enter__(self):
        self.install()
        return self

    def __exit__(self, *exc):
        restore_default_allocator()
        return False


def restore_default_allocator():
    """hipMalloc / hipFree for the library's device memory again."""
    _check(LIB.zgl_set_device_allocator(DEVICE_ALLOC_FN(), DEVICE_FREE_FN(), None))


def torch_device_allocator(fail_after: int | None = None) -> DeviceAllocator:
    """A DeviceAllocator over torch's caching allocator on the current device
    (torch.cuda.caching_allocator_alloc / _delete), so the library's tables
    and scratch count in torch's memory statistics and reuse its cache.
    ``fail_after=k`` makes every request after the first k return NULL (out of
    memory), for testing."""
    import torch

    state = {"n": 0}

    def alloc(nbytes):
        state["n"] += 1
        if fail_after is not None and state["n"] > fail_after:
            return 0
        return torch.cuda.caching_allocator_alloc(max(int(nbytes), 1))

    def free(ptr, _nbytes):
        torch.cuda.caching_allocator_delete(ptr)

    return DeviceAllocator(alloc, free)


def _stream_ptr(stream, mem):
    if mem == MEM_HOST:
        return None
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)
