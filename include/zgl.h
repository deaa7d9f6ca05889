/* zgl -- MI355X-native Gauss-Legendre-quadrature Zernike and spherical-
 * harmonic transforms: the C ABI.
 *
 * This is the thin boundary between host code and the HIP kernels for gfx950.
 * Only plain pointers, sizes and enums cross it. It replaces the two class
 * templates of zest (sebsassi/zest) that form the transform hot path:
 *
 *   zest::zt::GLQTransformer<ZernikeNorm, SHNorm, SHPhase, LonLatRadLayout<>>
 *       include/zernike_glq_transformer.hpp:511-965
 *   zest::st::GLQTransformer<SHNorm, SHPhase, LonLatLayout<> | LatLonLayout<>>
 *       include/sh_glq_transformer.hpp:513-1213
 *
 * Memory layouts are zest's, unchanged:
 *   Zernike grid   row-major [lon = 2N-1][lat = N+1][rad = round8(N+1)] doubles
 *                  (include/zernike_glq_transformer.hpp:50-131)
 *   SH grid        LonLat [lon = 2L-1][lat = round8(L)] or
 *                  LatLon [lat = L][lon = round8(2L-1)] doubles
 *                  (include/sh_glq_transformer.hpp:44-195)
 *   Zernike coeffs {cos, sin} double pairs at (n+1)(n+3)(2n+1)/24 + l*l/4 + m
 *                  (include/layout.hpp:384-421)
 *   SH coeffs      {cos, sin} double pairs at l(l+1)/2 + m
 *                  (include/layout.hpp:202-239)
 *
 * Convention values are zest's enum values (include/zernike_conventions.hpp:35,
 * include/sh_conventions.hpp:35-48).
 *
 * Semantics follow zest: the plan adapts to the grid order of each call
 * (zest's resize(), include/zernike_glq_transformer.hpp:580-632); the transform
 * works at order min(exp_order, grid_order); forward writes only that many
 * leading coefficients of each expansion and leaves the rest untouched
 * (include/zernike_glq_transformer.hpp:649-654); backward writes the whole grid.
 * With ZGL_MEM_DEVICE the work is enqueued on `stream` and the call returns
 * without waiting; with ZGL_MEM_HOST the call is synchronous, as zest is.
 * A plan must not be used from two threads at once (zest's instances are not
 * thread-safe either); distinct plans are independent. A plan's device calls
 * share its scratch buffers, so they must be ordered: issue them on one stream,
 * or synchronize between calls on different streams.
 */
#ifndef ZGL_H
#define ZGL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Each enum type has a last enumerator 0x7fffffff, as C APIs commonly do: the
 * type is int-sized and any int the caller passes is a value of it (C++ code
 * reading an out-of-range value of a narrower enum would be undefined), so an
 * invalid value reaches the library's argument checks. */
typedef enum zgl_status {
    ZGL_SUCCESS = 0,
    ZGL_ERROR_INVALID_VALUE = 1,
    ZGL_ERROR_OUT_OF_MEMORY = 2,
    ZGL_ERROR_HIP = 3,
    ZGL_ERROR_NOT_SUPPORTED = 4,
    ZGL_STATUS_MAX_ENUM = 0x7fffffff /* every int is a value of the type */
} zgl_status;

typedef enum zgl_kind {
    ZGL_ZERNIKE = 0, /* zt::GLQTransformer: BallGLQGrid <-> RealZernikeExpansion */
    ZGL_SH = 1,      /* st::GLQTransformer: SphereGLQGrid <-> RealSHExpansion */
    ZGL_KIND_MAX_ENUM = 0x7fffffff
} zgl_kind;

/* zest::zt::ZernikeNorm (include/zernike_conventions.hpp:35) */
enum { ZGL_ZNORM_NORMED = 0, ZGL_ZNORM_UNNORMED = 1 };
/* zest::st::SHNorm (include/sh_conventions.hpp:40-46) */
enum { ZGL_SHNORM_GEO = 0, ZGL_SHNORM_QM = 1 };
/* zest::st::SHPhase (include/sh_conventions.hpp:35) */
enum { ZGL_PHASE_NONE = -1, ZGL_PHASE_CS = 1 };
/* SH grid layouts (include/sh_glq_transformer.hpp:44-195) */
enum { ZGL_LAYOUT_LONLAT = 0, ZGL_LAYOUT_LATLON = 1 };

/* Arithmetic of the transforms. Coefficients are FP64 in memory in every
 * mode; grids are FP64 except through the *_grid32 entry points of the FP32
 * modes; the tables are always generated in FP64 (the FP32 modes use their
 * FP32 roundings). */
typedef enum zgl_precision {
    ZGL_FP64 = 0,        /* FP64 throughout: the drop-in form, matches zest */
    ZGL_FP32 = 1,        /* FP32 FFT and FP32-MFMA contractions, FP32 accumulation */
    ZGL_FP32_REFINE = 2, /* ZGL_FP32, plus one step of iterative refinement on the
                            forward transform: c += F32(g - B64(c)), the residual
                            formed by an FP64 backward transform */
    ZGL_BF16X3 = 3,      /* ZGL_FP32 with the Zernike Legendre products formed on the
                            BF16 MFMA from operands split into three BF16 values
                            (six partial products, FP32 accumulation) */
    ZGL_PRECISION_MAX_ENUM = 0x7fffffff
} zgl_precision;

typedef enum zgl_mem {
    ZGL_MEM_DEVICE = 0, /* grid and coefficient pointers are device memory */
    ZGL_MEM_HOST = 1,   /* host memory; staged through device buffers, synchronous */
    ZGL_MEM_MAX_ENUM = 0x7fffffff
} zgl_mem;

typedef struct zgl_plan zgl_plan;

/* Replaces GLQTransformer(size_t order) and GLQTransformer()
 * (include/zernike_glq_transformer.hpp:523-570,
 *  include/sh_glq_transformer.hpp:525-570). order 0 is allowed.
 * `layout` is ignored for ZGL_ZERNIKE (zest's only Zernike grid layout is
 * LonLatRadLayout). Tables are built on the current HIP device. */
zgl_status zgl_plan_create(zgl_plan **plan, zgl_kind kind, int zernike_norm,
                           int sh_norm, int sh_phase, int layout, size_t order,
                           zgl_precision precision);

zgl_status zgl_plan_destroy(zgl_plan *plan);

/* Replaces resize(order) (include/zernike_glq_transformer.hpp:580-632,
 * include/sh_glq_transformer.hpp:580-624). Tables of earlier orders stay
 * cached, so alternating orders does not rebuild them. */
zgl_status zgl_plan_resize(zgl_plan *plan, size_t order);

/* Replaces order() (include/zernike_glq_transformer.hpp:575). */
size_t zgl_plan_order(const zgl_plan *plan);

/* Number of grid doubles and coefficient pairs of one instance at `order`
 * for this plan's kind and layout (BallGLQGrid::size,
 * RealZernikeExpansion::size, SphereGLQGrid::size, RealSHExpansion::size). */
size_t zgl_grid_size(const zgl_plan *plan, size_t order);
size_t zgl_coeff_size(const zgl_plan *plan, size_t order);

/* Replaces forward_transform(grid_span, expansion_span)
 * (include/zernike_glq_transformer.hpp:640-655,
 *  include/sh_glq_transformer.hpp:632-647), batched over `batch`
 * independent instances. grid_stride and coeff_stride are the distances in
 * doubles between consecutive instances (0 means packed:
 * zgl_grid_size(order) and 2 * zgl_coeff_size(exp_order)). `stream` is a
 * hipStream_t (NULL = the null stream). */
zgl_status zgl_forward(zgl_plan *plan, const double *grid, size_t grid_order,
                       double *coeffs, size_t exp_order, size_t batch,
                       ptrdiff_t grid_stride, ptrdiff_t coeff_stride,
                       zgl_mem mem, void *stream);

/* Replaces backward_transform(expansion_span, grid_span)
 * (include/zernike_glq_transformer.hpp:663-679,
 *  include/sh_glq_transformer.hpp:655-668). */
zgl_status zgl_backward(zgl_plan *plan, const double *coeffs, size_t exp_order,
                        double *grid, size_t grid_order, size_t batch,
                        ptrdiff_t coeff_stride, ptrdiff_t grid_stride,
                        zgl_mem mem, void *stream);

/* zgl_forward without any allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. Arguments as
 * zgl_forward, device memory only (FP64 grids), any precision of the plan.
 * The call allocates nothing (no hipMalloc, hipFree, event or host-heap
 * allocation): its scratch is a static device array of 128 MiB that every GPU
 * the library runs on holds until the process ends, and the instances are
 * processed in chunks that fit it. Requirements, each reported as
 * ZGL_ERROR_INVALID_VALUE: the plan already holds the tables of grid_order
 * (created at that order or resized to it), and one instance of that order
 * fits the array (Zernike: orders up to 175 in FP64, 145 in FP32 with
 * refinement, 222 in FP32; SH in FP64: up to 2362). The
 * plan's tuned tile shapes are used; no tuning or stage
 * timing happens during the call. All calls on one GPU share the array:
 * calls that overlap on a GPU (other streams, other threads, other plans)
 * race, so order them. */
zgl_status zgl_forward_noalloc(zgl_plan *plan, const double *grid, size_t grid_order,
                               double *coeffs, size_t exp_order, size_t batch,
                               ptrdiff_t grid_stride, ptrdiff_t coeff_stride, void *stream);

/* Ragged batches: many instances of uneven orders in one call.
 * Replaces a loop of forward_transform / backward_transform calls over
 * transformers of different orders (include/zernike_glq_transformer.hpp:640-679,
 * include/sh_glq_transformer.hpp:632-668); each job's expansion order equals
 * its grid order. Device memory only; the work is enqueued on `stream`. The
 * jobs are grouped by order (largest first) and the groups spread over the
 * plan's internal streams; on a non-null stream the pipeline is captured into
 * a HIP graph on the first call and replayed while the job list (orders and
 * pointers) is unchanged. Jobs of order 0 are skipped. */
typedef struct zgl_job {
    size_t order;    /* grid order = expansion order of this instance */
    double *grid;    /* device pointer: grid of that order in zest layout */
    double *coeffs;  /* device pointer: {cos, sin} coefficients of that order */
} zgl_job;

zgl_status zgl_forward_ragged(zgl_plan *plan, const zgl_job *jobs, size_t njobs, void *stream);
zgl_status zgl_backward_ragged(zgl_plan *plan, const zgl_job *jobs, size_t njobs, void *stream);
/* The load balance of the last ragged call, as modelled work per unit of
 * execution; no counterpart in zest. Persistent pipeline: per resident
 * workgroup of its costliest launch; lanes pipeline: per internal stream.
 * num_lanes receives the count (costs may hold fewer, max_lanes). */
zgl_status zgl_ragged_lane_costs(const zgl_plan *plan, double *costs, int max_lanes,
                                 int *num_lanes);

/* Transposes of the two transforms, for gradients. No counterpart in zest;
 * they are the adjoints of the members replaced above
 * (include/zernike_glq_transformer.hpp:640-679,
 *  include/sh_glq_transformer.hpp:632-668) under the plain Euclidean inner
 * products on grids and on {cos, sin} coefficient arrays.
 *
 * zgl_backward_adjoint: coefficients = B^T grid, the transpose of the
 *   backward transform B (sum over grid points of the grid value times each
 *   basis function, without quadrature weights). Arguments as zgl_forward.
 * zgl_forward_adjoint:  grid = F^T coefficients, the transpose of the forward
 *   transform F. Since F = diag(nu) B^T W (W the quadrature weights of the
 *   grid points, nu = prefactor * kappa_n), F^T c = W * B(nu * c). Arguments
 *   as zgl_backward.
 * The sin part at m = 0 is ignored on input and written as zero on output,
 * as by the transforms themselves. Arithmetic follows the plan (FP64, or FP32
 * for both FP32 modes: refinement does not apply); grids are FP64. */
zgl_status zgl_backward_adjoint(zgl_plan *plan, const double *grid, size_t grid_order,
                                double *coeffs, size_t exp_order, size_t batch,
                                ptrdiff_t grid_stride, ptrdiff_t coeff_stride,
                                zgl_mem mem, void *stream);
zgl_status zgl_forward_adjoint(zgl_plan *plan, const double *coeffs, size_t exp_order,
                               double *grid, size_t grid_order, size_t batch,
                               ptrdiff_t coeff_stride, ptrdiff_t grid_stride,
                               zgl_mem mem, void *stream);

/* zgl_forward / zgl_backward with the grids stored in FP32 (same layouts,
 * floats instead of doubles; strides in floats). Only for plans of precision
 * ZGL_FP32 or ZGL_FP32_REFINE (ZGL_ERROR_NOT_SUPPORTED otherwise); the
 * coefficients stay FP64. With ZGL_FP32_REFINE the forward residual
 * g - B64(c) is formed in FP64 against the FP32 grid values, so the result
 * is the FP64-accurate transform of the grid as stored. Replaces the same
 * zest members as zgl_forward / zgl_backward
 * (include/zernike_glq_transformer.hpp:640-679,
 *  include/sh_glq_transformer.hpp:632-668) for callers that hold their
 * grids in single precision. */
zgl_status zgl_forward_grid32(zgl_plan *plan, const float *grid, size_t grid_order,
                              double *coeffs, size_t exp_order, size_t batch,
                              ptrdiff_t grid_stride, ptrdiff_t coeff_stride,
                              zgl_mem mem, void *stream);
zgl_status zgl_backward_grid32(zgl_plan *plan, const double *coeffs, size_t exp_order,
                               float *grid, size_t grid_order, size_t batch,
                               ptrdiff_t coeff_stride, ptrdiff_t grid_stride,
                               zgl_mem mem, void *stream);

/* Evaluation on arbitrary grids. Replaces zt::GridEvaluator::evaluate
 * (include/grid_evaluator.hpp:227-265) for a ZGL_ZERNIKE plan and
 * st::GridEvaluator::evaluate (include/grid_evaluator.hpp:104-133) for a
 * ZGL_SH plan, batched over `batch` expansions of order exp_order in the
 * plan's conventions. The grid is the tensor product of the given longitudes,
 * colatitudes and (Zernike) radii, in radians and ball radius units; any
 * values are accepted. Values of instance b start at values + b *
 * value_stride, row-major [nlon][nlat][nrad] (Zernike) or [nlon][nlat] (SH,
 * whatever the plan's grid layout; radii and nrad are ignored), as zest
 * returns them. Strides 0 mean packed (2 * zgl_coeff_size(exp_order) and the
 * point count). An empty axis writes nothing; order 0 writes zeros. FP64
 * throughout, whatever the plan's precision. `mem` applies to every pointer,
 * the point arrays included. Independent of the plan's transform order and
 * tables: the evaluator keeps its own buffers in the plan. */
zgl_status zgl_evaluate(zgl_plan *plan, const double *coeffs, size_t exp_order,
                        const double *longitudes, size_t nlon,
                        const double *colatitudes, size_t nlat,
                        const double *radii, size_t nrad,
                        double *values, size_t batch,
                        ptrdiff_t coeff_stride, ptrdiff_t value_stride,
                        zgl_mem mem, void *stream);

/* Power spectra, batched. zgl_power_spectrum replaces st::power_spectrum
 * (include/power_spectra.hpp:80-96) for an SH plan and zt::power_spectrum
 * (include/power_spectra.hpp:118-141) for a Zernike plan;
 * zgl_cross_power_spectrum replaces st::cross_power_spectrum
 * (include/power_spectra.hpp:43-58) and serves Zernike plans too.
 * Per instance: SH out[l], l < order; Zernike out[((n+1)^2)/4 + l/2] for
 * l <= n < order (RadialZernikeLayout), every l <= n written (zest's loop
 * stops at l < n and leaves the l = n entries unwritten). Each entry is
 * norm * (a_l0 b_l0 + sum_{m>=1} a_lm . b_lm) over the {cos, sin} pairs, norm =
 * 1 (qm) or 1/(4 pi) (geo) from the plan's SH convention. Strides 0 mean
 * packed (2 * zgl_coeff_size(order) and the entry count); batch <= 65535. */
zgl_status zgl_power_spectrum(zgl_plan *plan, const double *coeffs, size_t order,
                              double *out, size_t batch, ptrdiff_t coeff_stride,
                              ptrdiff_t out_stride, zgl_mem mem, void *stream);
zgl_status zgl_cross_power_spectrum(zgl_plan *plan, const double *a, const double *b,
                                    size_t order, double *out, size_t batch,
                                    ptrdiff_t coeff_stride, ptrdiff_t out_stride,
                                    zgl_mem mem, void *stream);

/* Rotations, batched and in place. zgl_rotate replaces Rotor::rotate
 * (include/rotor.hpp:124-288) with the WignerdPiHalfCollection it is given
 * (here built on the device and kept by the plan): intrinsic ZYZ Euler
 * angles, euler_angles[3 b .. 3 b + 2] = (alpha, beta, gamma) of instance b
 * (in the same memory as the coefficients); rotation_type 0 rotates the
 * object, 1 the coordinate system (zest's RotationType). zgl_polar_rotate
 * replaces Rotor::polar_rotate (include/rotor.hpp:321-380), angles[b] per
 * instance. Both take the plan's kind; the plan's conventions do not enter
 * (zest converts with the expansion's own conventions, which cancel). Orders
 * below 2 leave the coefficients unchanged; batch <= 65535. */
zgl_status zgl_rotate(zgl_plan *plan, double *coeffs, size_t order,
                      const double *euler_angles, int rotation_type, size_t batch,
                      ptrdiff_t coeff_stride, zgl_mem mem, void *stream);
zgl_status zgl_polar_rotate(zgl_plan *plan, double *coeffs, size_t order,
                            const double *angles, int rotation_type, size_t batch,
                            ptrdiff_t coeff_stride, zgl_mem mem, void *stream);
/* The Wigner d-matrices at pi/2 of degrees l < order, as
 * WignerdPiHalfCollection(order).matrices() (src/wignerd_collection.cpp:29-104):
 * order (order + 1)(2 order + 1)/6 doubles, d^l_{m1,m2} at
 * l(l+1)(2l+1)/6 + (l+1) m1 + m2. */
zgl_status zgl_wignerd_pi2(zgl_plan *plan, size_t order, double *out, zgl_mem mem,
                           void *stream);
/* euler_angles_from_rotation_matrix (include/rotor.hpp:49-60), on the host:
 * matrix row-major 3x3, angles = (alpha, beta, gamma). Computes beta from
 * sqrt((1 + r22^2)(1 - r22^2)) as zest does. */
void zgl_euler_angles_from_rotation_matrix(const double *matrix, double *angles);

/* Autotuning. No counterpart in the reference (zest has nothing to tune).
 * While enabled (also: environment ZGL_AUTOTUNE=1 at plan creation), the
 * first call of each stage at a new order and arithmetic times the candidate
 * tile shapes on that call's own operands -- the FFT's column pairs per
 * workgroup, the streaming contractions' instances per workgroup and, for the
 * FP32 Legendre stages, FP32 MFMA products against split-BF16x3 ones -- and
 * keeps the fastest for later calls. Results depend on the choice only
 * within the FP32 rounding of the mixed-precision transforms.
 * zgl_plan_tuning lists the choices made: 5 ints per entry (order, bytes per
 * real of the arithmetic, bytes per grid value, stage index as in
 * zgl_profile_stats, choice); a Legendre choice of 16 + g is the BF16x3
 * variant with g instances per workgroup. */
zgl_status zgl_plan_autotune(zgl_plan *plan, int enable);
zgl_status zgl_plan_tuning(const zgl_plan *plan, int *entries, int max_entries,
                           int *num_entries);

/* Stage timing. No counterpart in the reference: it serves benchmarks and
 * profilers. While enabled, every kernel launch of the plan is bracketed by
 * HIP events on the launch stream; zgl_profile_stats() waits for them and
 * returns, per stage, the launch count and the summed kernel time. */
typedef struct zgl_stage_stats {
    char name[32];
    long long launches;
    long long instances; /* instances processed by those launches */
    double total_ms;
} zgl_stage_stats;

zgl_status zgl_profile_enable(zgl_plan *plan, int enable);
zgl_status zgl_profile_stats(zgl_plan *plan, zgl_stage_stats *out, int max_stages,
                             int *num_stages);

/* Message of the last error on the calling thread (zest reports errors by
 * exceptions; the C ABI returns a status and keeps the text here). */
const char *zgl_last_error_string(void);

/* Device allocator. zest allocates its buffers through its AlignedAllocator
 * (include/alignment.hpp:157-187), whose failure is std::bad_alloc or
 * std::bad_array_new_length (include/alignment.hpp:169-176); this is the
 * counterpart for the device
 * memory of the library, with no counterpart as a hook in zest: it is for
 * callers that keep device memory in their own pool or caching allocator
 * (a framework's, an arena sized up front) and for tests that make
 * allocations fail.
 *
 * Every device allocation the library makes after the call -- tables, scratch,
 * work lists, staging buffers -- is requested from alloc(bytes, user), which
 * returns device memory of the current device, aligned to 256 bytes, or NULL
 * when it cannot. A NULL makes the call that needed the memory return
 * ZGL_ERROR_OUT_OF_MEMORY (the message names the request); the plan stays
 * valid, holds nothing of the failed request and can be used again. Memory is
 * handed back with free(ptr, bytes, user) by the allocator that made it, after
 * the library has waited for all device work (hipDeviceSynchronize), so the
 * allocator may be changed at any time, also while plans hold memory. alloc ==
 * NULL (free ignored) restores hipMalloc / hipFree. Process-wide; may be
 * called from any thread; alloc and free may be called from any thread that
 * calls into the library. The static scratch of zgl_forward_noalloc is not
 * allocated and does not pass through here. */
typedef void *(*zgl_device_alloc_fn)(size_t bytes, void *user);
typedef void (*zgl_device_free_fn)(void *ptr, size_t bytes, void *user);
zgl_status zgl_set_device_allocator(zgl_device_alloc_fn alloc, zgl_device_free_fn free, void *user);

#ifdef __cplusplus
}
#endif
#endif /* ZGL_H */
