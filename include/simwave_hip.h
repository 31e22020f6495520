/*
 * simwave_hip.h - C ABI of the HIP (gfx950) acoustic forward operator.
 *
 * One shared library per reference variant (dimension x density x
 * precision): simwave_amd/lib/libsimwave_hip_<dim>d_<cd|vd>_<f32|f64>.so.
 * Each library exports
 *
 *   - `forward`: the reference's FFI symbol with the reference's signature
 *     for that variant (the one spelled out below as
 *     simwave_hip_forward_<variant>), so the reference's ctypes loader
 *     (simwave/kernel/backend/middleware.py:26-51,107-158) binds it
 *     unchanged;
 *   - simwave_hip_forward_<variant>: the same function under a unique name;
 *   - simwave_hip_forward_device_<variant>: the same contract with device
 *     pointers for the arrays that live on the GPU across calls;
 *   - simwave_hip_plan_{create,run,destroy}_<variant>: the device entry
 *     split into chunks for serving loops, and
 *     simwave_hip_plans_run_<variant>, which advances several plans in
 *     batched launches;
 *   - simwave_hip_forward_device_noalloc_<variant>: the device entry with
 *     every library buffer in a static device array (no allocation);
 *   - simwave_hip_gather_<variant>: the device gather that builds the
 *     edge-extended models and damping mask of a device-resident shot;
 *   - simwave_hip_fixed_div_<variant>, simwave_hip_last_error,
 *     simwave_hip_last_timings, simwave_hip_variant, simwave_hip_trim_pool;
 *   - simwave_hip_{device,host}_{alloc,free}, simwave_hip_copy,
 *     simwave_hip_zero, simwave_hip_stream_*, simwave_hip_event_*: memory,
 *     streams and events for device-resident callers;
 *   - through libsimwave_hip_common.so, which every variant library links:
 *     simwave_hip_overlap_{acquire,release,holders}.
 *
 * Array layout (reference contract, middleware.py:160-285): C-contiguous,
 * last axis fastest; u is [num_snapshots][nz][nx][ny] (2D: no ny);
 * receivers is [end_timestep - begin_timestep + 1 rows used][num_receivers],
 * row n-1 written at step n.  Index arrays are size_t (numpy uint64).
 *
 * Errors: the reference has no error return (its C never checks).  These
 * entry points return a negative value on failure (bad sizes, unsupported
 * order, no device, HIP error) and leave a message for
 * simwave_hip_last_error() on the calling thread; they never exit().
 * Return value on success: wall seconds of the call, like the reference.
 *
 * Threading: synchronous; re-entrant on disjoint buffers.
 */
#ifndef SIMWAVE_HIP_H
#define SIMWAVE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 2d_cd_f32 */
/* Host arrays, as the reference's forward(): reference constant_density/2d/wave.c:23-36
 * (`forward` in libsimwave_hip_2d_cd_f32.so has exactly this signature). */
double simwave_hip_forward_2d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots);
/* Device arrays for u, velocity, damp and receivers (the
 * rest stay on the host); runs on the HIP stream `stream` (NULL: the null
 * stream) and returns after the stream has drained.  The models are only
 * read.  flags: 1 time the step kernel with HIP events, 2 force the
 * generic kernels, 4 run the receiver gather of large grids on a second
 * stream of the library's (event-ordered; one run or plan per process may
 * hold it, see simwave_hip_overlap_acquire; an error while another does).  Same arguments as
 * reference constant_density/2d/wave.c:23-36, plus stream and flags. */
double simwave_hip_forward_device_2d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Allocation-free device entry (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results as simwave_hip_forward_device_2d_cd_f32, flags 0
 * or 2 only; the library's buffers come from a static 64 MiB device
 * array per library and device (per-point arrays of float: 8 bytes a grid
 * point, so at most about 8.3 M points), and larger grids are rejected
 * with an error.  No hipMalloc, hipFree or allocating library call.
 * Synchronous; calls that overlap on one device (several threads)
 * race on the shared array. */
double simwave_hip_forward_device_noalloc_2d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Bytes of the static array of simwave_hip_forward_device_noalloc_2d_cd_f32
 * (host-only helper). */
size_t simwave_hip_scratch_bytes_2d_cd_f32(void);
/* Persistent plan (no counterpart in the reference, whose forward runs a
 * whole step range per call): the device entry split into create / run in
 * chunks / destroy, for callers that interleave shots (the serving loop) or
 * re-run step ranges (checkpointed recomputation).  Arguments as
 * simwave_hip_forward_device_2d_cd_f32; its host arrays are read only here.
 * flags as the device entry, plus 8: for stride 0, capture 48 steps once in a
 * hipGraph and replay it for every aligned 48-step stretch of a run (the
 * step number then comes from a device counter; results are identical).
 * Returns NULL on failure. */
void *simwave_hip_plan_create_2d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Enqueues the next `steps` steps of the plan on `stream`, ordered after the
 * plan's earlier work on any stream; does not wait for the device.  Returns
 * the steps left, or -1. */
long long simwave_hip_plan_run_2d_cd_f32(void *plan, size_t steps, void *stream);
/* Frees the plan once its enqueued work is done.  Returns 0, or -1. */
int simwave_hip_plan_destroy_2d_cd_f32(void *plan);
/* Batched runs (no counterpart in the reference; for the serving loop):
 * enqueues up to `steps` further steps of each of the `count` plans on
 * `stream`, after each plan's earlier work.  Plans that can be batched
 * (simwave_hip_plan_batchable_2d_cd_f32) advance together, radius by radius:
 * one step kernel and one receiver kernel per time step for all of them
 * (up to 48 per launch); the others run as simwave_hip_plan_run_2d_cd_f32.
 * left[i] receives plan i's remaining steps.  Returns 0, or -1. */
int simwave_hip_plans_run_2d_cd_f32(void **plans, int count, size_t steps,
                            void *stream, long long *left);
/* 1 when the plan can join batched runs (stride 0, fused z-marching step
 * kernel, no source box in the halo, flags without 1 and 4), else 0. */
int simwave_hip_plan_batchable_2d_cd_f32(void *plan);
/* Device gather (no counterpart in the reference; for the device-resident
 * shot, which builds its edge-extended models and damping mask on the GPU
 * from their distinct values): out[i][j][k] = small[idx0[i]][idx1[j]][idx2[k]]
 * for out of out_shape and small of small_shape (host arrays of ndim
 * entries, ndim the library's dimension; 2D: no k, idx2 ignored); out,
 * small and the int64 index arrays are device pointers.  Enqueued on
 * `stream` without waiting.  Returns 0, or -1. */
int simwave_hip_gather_2d_cd_f32(float *out, const float *small, int ndim,
                        const long long *out_shape, const long long *small_shape,
                        const long long *idx0, const long long *idx1,
                        const long long *idx2, void *stream);
/* Adjoint-state gradient (no counterpart in the reference; constant density).
 * Arguments of simwave_hip_forward_device_2d_cd_f32 (device arrays; stride 0,
 * steps 1..T, source boxes inside the interior), then: rec_adjoint, the
 * adjoint of the shot record [T][num_receivers] (device); grad_velocity,
 * float64 [extended grid] (device), set to the gradient with respect to
 * the velocity; grad_wavelet, [T][num_sources] (device) or NULL, set to
 * the gradient with respect to the wavelet samples; checkpoint_every, the
 * forward states saved every that many steps and recomputed per segment
 * (>= the step count: every level kept in device memory, no recomputation).
 * flags: 1 sets grad_velocity to the zero-lag cross-correlation image
 * sum_k u^k lambda^k of the forward and adjoint wavefields instead; 2 takes
 * rec_adjoint as observed data and drives the adjoint with the residual
 * (shot record - rec_adjoint) formed on the device after the forward (the
 * L2 misfit's gradient in one call); 4 runs the forward's receiver gather
 * on a second stream (large grids); 8 times every reverse update kernel
 * with HIP events (simwave_hip_last_timings: [4] their summed seconds,
 * [5] their count).
 * The shot starts from rest: u's three slots are zeroed first, whatever
 * they hold.  Also writes the shot record and leaves u as after a plain
 * forward.  Returns wall seconds, or -1. */
double simwave_hip_gradient_2d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    const float *rec_adjoint, double *grad_velocity, float *grad_wavelet,
    size_t checkpoint_every, void *stream, unsigned flags);
/* Host-only helper (no device work, no reference counterpart): the
 * fixed-divisor parameters the kernels use for the divisor b (a squared
 * spacing, cd-3d wave.c:186).  Returns 1 and y = RN(1/b) and the safe
 * numerator range [lo, hi], or 0 when IEEE division is used. */
int simwave_hip_fixed_div_2d_cd_f32(float b, float *y, float *lo, float *hi);
/* Host-only helper (no device work, no reference counterpart): 1 when the
 * 3D tiled kernel forms the quotients sum/b of the divisor b as
 * RN(sum * RN64(1/b)), the float64 product rounded to float (float32
 * libraries, b = B 2^t with B odd, t <= 0, b in [2^-126, 2^24): equal to
 * the IEEE quotient for every float sum); 0 otherwise. */
int simwave_hip_f64_quotient_2d_cd_f32(float b);

/* ---- 2d_cd_f64 */
/* Host arrays, as the reference's forward(): reference constant_density/2d/wave.c:23-36
 * (`forward` in libsimwave_hip_2d_cd_f64.so has exactly this signature). */
double simwave_hip_forward_2d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots);
/* Device arrays for u, velocity, damp and receivers (the
 * rest stay on the host); runs on the HIP stream `stream` (NULL: the null
 * stream) and returns after the stream has drained.  The models are only
 * read.  flags: 1 time the step kernel with HIP events, 2 force the
 * generic kernels, 4 run the receiver gather of large grids on a second
 * stream of the library's (event-ordered; one run or plan per process may
 * hold it, see simwave_hip_overlap_acquire; an error while another does).  Same arguments as
 * reference constant_density/2d/wave.c:23-36, plus stream and flags. */
double simwave_hip_forward_device_2d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Allocation-free device entry (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results as simwave_hip_forward_device_2d_cd_f64, flags 0
 * or 2 only; the library's buffers come from a static 64 MiB device
 * array per library and device (per-point arrays of double: 16 bytes a grid
 * point, so at most about 4.1 M points), and larger grids are rejected
 * with an error.  No hipMalloc, hipFree or allocating library call.
 * Synchronous; calls that overlap on one device (several threads)
 * race on the shared array. */
double simwave_hip_forward_device_noalloc_2d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Bytes of the static array of simwave_hip_forward_device_noalloc_2d_cd_f64
 * (host-only helper). */
size_t simwave_hip_scratch_bytes_2d_cd_f64(void);
/* Persistent plan (no counterpart in the reference, whose forward runs a
 * whole step range per call): the device entry split into create / run in
 * chunks / destroy, for callers that interleave shots (the serving loop) or
 * re-run step ranges (checkpointed recomputation).  Arguments as
 * simwave_hip_forward_device_2d_cd_f64; its host arrays are read only here.
 * flags as the device entry, plus 8: for stride 0, capture 48 steps once in a
 * hipGraph and replay it for every aligned 48-step stretch of a run (the
 * step number then comes from a device counter; results are identical).
 * Returns NULL on failure. */
void *simwave_hip_plan_create_2d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Enqueues the next `steps` steps of the plan on `stream`, ordered after the
 * plan's earlier work on any stream; does not wait for the device.  Returns
 * the steps left, or -1. */
long long simwave_hip_plan_run_2d_cd_f64(void *plan, size_t steps, void *stream);
/* Frees the plan once its enqueued work is done.  Returns 0, or -1. */
int simwave_hip_plan_destroy_2d_cd_f64(void *plan);
/* Batched runs (no counterpart in the reference; for the serving loop):
 * enqueues up to `steps` further steps of each of the `count` plans on
 * `stream`, after each plan's earlier work.  Plans that can be batched
 * (simwave_hip_plan_batchable_2d_cd_f64) advance together, radius by radius:
 * one step kernel and one receiver kernel per time step for all of them
 * (up to 48 per launch); the others run as simwave_hip_plan_run_2d_cd_f64.
 * left[i] receives plan i's remaining steps.  Returns 0, or -1. */
int simwave_hip_plans_run_2d_cd_f64(void **plans, int count, size_t steps,
                            void *stream, long long *left);
/* 1 when the plan can join batched runs (stride 0, fused z-marching step
 * kernel, no source box in the halo, flags without 1 and 4), else 0. */
int simwave_hip_plan_batchable_2d_cd_f64(void *plan);
/* Device gather (no counterpart in the reference; for the device-resident
 * shot, which builds its edge-extended models and damping mask on the GPU
 * from their distinct values): out[i][j][k] = small[idx0[i]][idx1[j]][idx2[k]]
 * for out of out_shape and small of small_shape (host arrays of ndim
 * entries, ndim the library's dimension; 2D: no k, idx2 ignored); out,
 * small and the int64 index arrays are device pointers.  Enqueued on
 * `stream` without waiting.  Returns 0, or -1. */
int simwave_hip_gather_2d_cd_f64(double *out, const double *small, int ndim,
                        const long long *out_shape, const long long *small_shape,
                        const long long *idx0, const long long *idx1,
                        const long long *idx2, void *stream);
/* Adjoint-state gradient (no counterpart in the reference; constant density).
 * Arguments of simwave_hip_forward_device_2d_cd_f64 (device arrays; stride 0,
 * steps 1..T, source boxes inside the interior), then: rec_adjoint, the
 * adjoint of the shot record [T][num_receivers] (device); grad_velocity,
 * float64 [extended grid] (device), set to the gradient with respect to
 * the velocity; grad_wavelet, [T][num_sources] (device) or NULL, set to
 * the gradient with respect to the wavelet samples; checkpoint_every, the
 * forward states saved every that many steps and recomputed per segment
 * (>= the step count: every level kept in device memory, no recomputation).
 * flags: 1 sets grad_velocity to the zero-lag cross-correlation image
 * sum_k u^k lambda^k of the forward and adjoint wavefields instead; 2 takes
 * rec_adjoint as observed data and drives the adjoint with the residual
 * (shot record - rec_adjoint) formed on the device after the forward (the
 * L2 misfit's gradient in one call); 4 runs the forward's receiver gather
 * on a second stream (large grids); 8 times every reverse update kernel
 * with HIP events (simwave_hip_last_timings: [4] their summed seconds,
 * [5] their count).
 * The shot starts from rest: u's three slots are zeroed first, whatever
 * they hold.  Also writes the shot record and leaves u as after a plain
 * forward.  Returns wall seconds, or -1. */
double simwave_hip_gradient_2d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    const double *rec_adjoint, double *grad_velocity, double *grad_wavelet,
    size_t checkpoint_every, void *stream, unsigned flags);
/* Host-only helper (no device work, no reference counterpart): the
 * fixed-divisor parameters the kernels use for the divisor b (a squared
 * spacing, cd-3d wave.c:186).  Returns 1 and y = RN(1/b) and the safe
 * numerator range [lo, hi], or 0 when IEEE division is used. */
int simwave_hip_fixed_div_2d_cd_f64(double b, double *y, double *lo, double *hi);
/* Host-only helper (no device work, no reference counterpart): 1 when the
 * 3D tiled kernel forms the quotients sum/b of the divisor b as
 * RN(sum * RN64(1/b)), the float64 product rounded to float (float32
 * libraries, b = B 2^t with B odd, t <= 0, b in [2^-126, 2^24): equal to
 * the IEEE quotient for every float sum); 0 otherwise. */
int simwave_hip_f64_quotient_2d_cd_f64(double b);

/* ---- 2d_vd_f32 */
/* Host arrays, as the reference's forward(): reference variable_density/2d/wave.c:23-36
 * (`forward` in libsimwave_hip_2d_vd_f32.so has exactly this signature). */
double simwave_hip_forward_2d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots);
/* Device arrays for u, velocity, density, damp and receivers (the
 * rest stay on the host); runs on the HIP stream `stream` (NULL: the null
 * stream) and returns after the stream has drained.  The models are only
 * read.  flags: 1 time the step kernel with HIP events, 2 force the
 * generic kernels, 4 run the receiver gather of large grids on a second
 * stream of the library's (event-ordered; one run or plan per process may
 * hold it, see simwave_hip_overlap_acquire; an error while another does).  Same arguments as
 * reference variable_density/2d/wave.c:23-36, plus stream and flags. */
double simwave_hip_forward_device_2d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Allocation-free device entry (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results as simwave_hip_forward_device_2d_vd_f32, flags 0
 * or 2 only; the library's buffers come from a static 64 MiB device
 * array per library and device (per-point arrays of float: 8 bytes a grid
 * point, so at most about 8.3 M points; from space order 12 on, the density
 * sums add 8 bytes a point: about 4.1 M points), and larger grids are
 * rejected with an error.  No hipMalloc, hipFree or allocating library call.
 * Synchronous; calls that overlap on one device (several threads)
 * race on the shared array. */
double simwave_hip_forward_device_noalloc_2d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Bytes of the static array of simwave_hip_forward_device_noalloc_2d_vd_f32
 * (host-only helper). */
size_t simwave_hip_scratch_bytes_2d_vd_f32(void);
/* Persistent plan (no counterpart in the reference, whose forward runs a
 * whole step range per call): the device entry split into create / run in
 * chunks / destroy, for callers that interleave shots (the serving loop) or
 * re-run step ranges (checkpointed recomputation).  Arguments as
 * simwave_hip_forward_device_2d_vd_f32; its host arrays are read only here.
 * flags as the device entry, plus 8: for stride 0, capture 48 steps once in a
 * hipGraph and replay it for every aligned 48-step stretch of a run (the
 * step number then comes from a device counter; results are identical).
 * Returns NULL on failure. */
void *simwave_hip_plan_create_2d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Enqueues the next `steps` steps of the plan on `stream`, ordered after the
 * plan's earlier work on any stream; does not wait for the device.  Returns
 * the steps left, or -1. */
long long simwave_hip_plan_run_2d_vd_f32(void *plan, size_t steps, void *stream);
/* Frees the plan once its enqueued work is done.  Returns 0, or -1. */
int simwave_hip_plan_destroy_2d_vd_f32(void *plan);
/* Batched runs (no counterpart in the reference; for the serving loop):
 * enqueues up to `steps` further steps of each of the `count` plans on
 * `stream`, after each plan's earlier work.  Plans that can be batched
 * (simwave_hip_plan_batchable_2d_vd_f32) advance together, radius by radius:
 * one step kernel and one receiver kernel per time step for all of them
 * (up to 48 per launch); the others run as simwave_hip_plan_run_2d_vd_f32.
 * left[i] receives plan i's remaining steps.  Returns 0, or -1. */
int simwave_hip_plans_run_2d_vd_f32(void **plans, int count, size_t steps,
                            void *stream, long long *left);
/* 1 when the plan can join batched runs (stride 0, fused z-marching step
 * kernel, no source box in the halo, flags without 1 and 4), else 0. */
int simwave_hip_plan_batchable_2d_vd_f32(void *plan);
/* Device gather (no counterpart in the reference; for the device-resident
 * shot, which builds its edge-extended models and damping mask on the GPU
 * from their distinct values): out[i][j][k] = small[idx0[i]][idx1[j]][idx2[k]]
 * for out of out_shape and small of small_shape (host arrays of ndim
 * entries, ndim the library's dimension; 2D: no k, idx2 ignored); out,
 * small and the int64 index arrays are device pointers.  Enqueued on
 * `stream` without waiting.  Returns 0, or -1. */
int simwave_hip_gather_2d_vd_f32(float *out, const float *small, int ndim,
                        const long long *out_shape, const long long *small_shape,
                        const long long *idx0, const long long *idx1,
                        const long long *idx2, void *stream);
/* Adjoint-state gradient (no counterpart in the reference; constant density only: this variable-density library returns -1).
 * Arguments of simwave_hip_forward_device_2d_vd_f32 (device arrays; stride 0,
 * steps 1..T, source boxes inside the interior), then: rec_adjoint, the
 * adjoint of the shot record [T][num_receivers] (device); grad_velocity,
 * float64 [extended grid] (device), set to the gradient with respect to
 * the velocity; grad_wavelet, [T][num_sources] (device) or NULL, set to
 * the gradient with respect to the wavelet samples; checkpoint_every, the
 * forward states saved every that many steps and recomputed per segment
 * (>= the step count: every level kept in device memory, no recomputation).
 * flags: 1 sets grad_velocity to the zero-lag cross-correlation image
 * sum_k u^k lambda^k of the forward and adjoint wavefields instead; 2 takes
 * rec_adjoint as observed data and drives the adjoint with the residual
 * (shot record - rec_adjoint) formed on the device after the forward (the
 * L2 misfit's gradient in one call); 4 runs the forward's receiver gather
 * on a second stream (large grids); 8 times every reverse update kernel
 * with HIP events (simwave_hip_last_timings: [4] their summed seconds,
 * [5] their count).
 * The shot starts from rest: u's three slots are zeroed first, whatever
 * they hold.  Also writes the shot record and leaves u as after a plain
 * forward.  Returns wall seconds, or -1. */
double simwave_hip_gradient_2d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, float dz, float dx,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    const float *rec_adjoint, double *grad_velocity, float *grad_wavelet,
    size_t checkpoint_every, void *stream, unsigned flags);
/* Host-only helper (no device work, no reference counterpart): the
 * fixed-divisor parameters the kernels use for the divisor b (a squared
 * spacing, cd-3d wave.c:186).  Returns 1 and y = RN(1/b) and the safe
 * numerator range [lo, hi], or 0 when IEEE division is used. */
int simwave_hip_fixed_div_2d_vd_f32(float b, float *y, float *lo, float *hi);
/* Host-only helper (no device work, no reference counterpart): 1 when the
 * 3D tiled kernel forms the quotients sum/b of the divisor b as
 * RN(sum * RN64(1/b)), the float64 product rounded to float (float32
 * libraries, b = B 2^t with B odd, t <= 0, b in [2^-126, 2^24): equal to
 * the IEEE quotient for every float sum); 0 otherwise. */
int simwave_hip_f64_quotient_2d_vd_f32(float b);

/* ---- 2d_vd_f64 */
/* Host arrays, as the reference's forward(): reference variable_density/2d/wave.c:23-36
 * (`forward` in libsimwave_hip_2d_vd_f64.so has exactly this signature). */
double simwave_hip_forward_2d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots);
/* Device arrays for u, velocity, density, damp and receivers (the
 * rest stay on the host); runs on the HIP stream `stream` (NULL: the null
 * stream) and returns after the stream has drained.  The models are only
 * read.  flags: 1 time the step kernel with HIP events, 2 force the
 * generic kernels, 4 run the receiver gather of large grids on a second
 * stream of the library's (event-ordered; one run or plan per process may
 * hold it, see simwave_hip_overlap_acquire; an error while another does).  Same arguments as
 * reference variable_density/2d/wave.c:23-36, plus stream and flags. */
double simwave_hip_forward_device_2d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Allocation-free device entry (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results as simwave_hip_forward_device_2d_vd_f64, flags 0
 * or 2 only; the library's buffers come from a static 64 MiB device
 * array per library and device (per-point arrays of double: 16 bytes a grid
 * point, so at most about 4.1 M points; from space order 12 on, the density
 * sums add 16 bytes a point: about 2.0 M points), and larger grids are
 * rejected with an error.  No hipMalloc, hipFree or allocating library call.
 * Synchronous; calls that overlap on one device (several threads)
 * race on the shared array. */
double simwave_hip_forward_device_noalloc_2d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Bytes of the static array of simwave_hip_forward_device_noalloc_2d_vd_f64
 * (host-only helper). */
size_t simwave_hip_scratch_bytes_2d_vd_f64(void);
/* Persistent plan (no counterpart in the reference, whose forward runs a
 * whole step range per call): the device entry split into create / run in
 * chunks / destroy, for callers that interleave shots (the serving loop) or
 * re-run step ranges (checkpointed recomputation).  Arguments as
 * simwave_hip_forward_device_2d_vd_f64; its host arrays are read only here.
 * flags as the device entry, plus 8: for stride 0, capture 48 steps once in a
 * hipGraph and replay it for every aligned 48-step stretch of a run (the
 * step number then comes from a device counter; results are identical).
 * Returns NULL on failure. */
void *simwave_hip_plan_create_2d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Enqueues the next `steps` steps of the plan on `stream`, ordered after the
 * plan's earlier work on any stream; does not wait for the device.  Returns
 * the steps left, or -1. */
long long simwave_hip_plan_run_2d_vd_f64(void *plan, size_t steps, void *stream);
/* Frees the plan once its enqueued work is done.  Returns 0, or -1. */
int simwave_hip_plan_destroy_2d_vd_f64(void *plan);
/* Batched runs (no counterpart in the reference; for the serving loop):
 * enqueues up to `steps` further steps of each of the `count` plans on
 * `stream`, after each plan's earlier work.  Plans that can be batched
 * (simwave_hip_plan_batchable_2d_vd_f64) advance together, radius by radius:
 * one step kernel and one receiver kernel per time step for all of them
 * (up to 48 per launch); the others run as simwave_hip_plan_run_2d_vd_f64.
 * left[i] receives plan i's remaining steps.  Returns 0, or -1. */
int simwave_hip_plans_run_2d_vd_f64(void **plans, int count, size_t steps,
                            void *stream, long long *left);
/* 1 when the plan can join batched runs (stride 0, fused z-marching step
 * kernel, no source box in the halo, flags without 1 and 4), else 0. */
int simwave_hip_plan_batchable_2d_vd_f64(void *plan);
/* Device gather (no counterpart in the reference; for the device-resident
 * shot, which builds its edge-extended models and damping mask on the GPU
 * from their distinct values): out[i][j][k] = small[idx0[i]][idx1[j]][idx2[k]]
 * for out of out_shape and small of small_shape (host arrays of ndim
 * entries, ndim the library's dimension; 2D: no k, idx2 ignored); out,
 * small and the int64 index arrays are device pointers.  Enqueued on
 * `stream` without waiting.  Returns 0, or -1. */
int simwave_hip_gather_2d_vd_f64(double *out, const double *small, int ndim,
                        const long long *out_shape, const long long *small_shape,
                        const long long *idx0, const long long *idx1,
                        const long long *idx2, void *stream);
/* Adjoint-state gradient (no counterpart in the reference; constant density only: this variable-density library returns -1).
 * Arguments of simwave_hip_forward_device_2d_vd_f64 (device arrays; stride 0,
 * steps 1..T, source boxes inside the interior), then: rec_adjoint, the
 * adjoint of the shot record [T][num_receivers] (device); grad_velocity,
 * float64 [extended grid] (device), set to the gradient with respect to
 * the velocity; grad_wavelet, [T][num_sources] (device) or NULL, set to
 * the gradient with respect to the wavelet samples; checkpoint_every, the
 * forward states saved every that many steps and recomputed per segment
 * (>= the step count: every level kept in device memory, no recomputation).
 * flags: 1 sets grad_velocity to the zero-lag cross-correlation image
 * sum_k u^k lambda^k of the forward and adjoint wavefields instead; 2 takes
 * rec_adjoint as observed data and drives the adjoint with the residual
 * (shot record - rec_adjoint) formed on the device after the forward (the
 * L2 misfit's gradient in one call); 4 runs the forward's receiver gather
 * on a second stream (large grids); 8 times every reverse update kernel
 * with HIP events (simwave_hip_last_timings: [4] their summed seconds,
 * [5] their count).
 * The shot starts from rest: u's three slots are zeroed first, whatever
 * they hold.  Also writes the shot record and leaves u as after a plain
 * forward.  Returns wall seconds, or -1. */
double simwave_hip_gradient_2d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, double dz, double dx,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    const double *rec_adjoint, double *grad_velocity, double *grad_wavelet,
    size_t checkpoint_every, void *stream, unsigned flags);
/* Host-only helper (no device work, no reference counterpart): the
 * fixed-divisor parameters the kernels use for the divisor b (a squared
 * spacing, cd-3d wave.c:186).  Returns 1 and y = RN(1/b) and the safe
 * numerator range [lo, hi], or 0 when IEEE division is used. */
int simwave_hip_fixed_div_2d_vd_f64(double b, double *y, double *lo, double *hi);
/* Host-only helper (no device work, no reference counterpart): 1 when the
 * 3D tiled kernel forms the quotients sum/b of the divisor b as
 * RN(sum * RN64(1/b)), the float64 product rounded to float (float32
 * libraries, b = B 2^t with B odd, t <= 0, b in [2^-126, 2^24): equal to
 * the IEEE quotient for every float sum); 0 otherwise. */
int simwave_hip_f64_quotient_2d_vd_f64(double b);

/* ---- 3d_cd_f32 */
/* Host arrays, as the reference's forward(): reference constant_density/3d/wave.c:23-36
 * (`forward` in libsimwave_hip_3d_cd_f32.so has exactly this signature). */
double simwave_hip_forward_3d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots);
/* Device arrays for u, velocity, damp and receivers (the
 * rest stay on the host); runs on the HIP stream `stream` (NULL: the null
 * stream) and returns after the stream has drained.  The models are only
 * read.  flags: 1 time the step kernel with HIP events, 2 force the
 * generic kernels, 4 run the receiver gather of large grids on a second
 * stream of the library's (event-ordered; one run or plan per process may
 * hold it, see simwave_hip_overlap_acquire; an error while another does).  Same arguments as
 * reference constant_density/3d/wave.c:23-36, plus stream and flags. */
double simwave_hip_forward_device_3d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Allocation-free device entry (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results as simwave_hip_forward_device_3d_cd_f32, flags 0
 * or 2 only; the library's buffers come from a static 64 MiB device
 * array per library and device (per-point arrays of float: 12 bytes a grid
 * point, so at most about 5.5 M points), and larger grids are rejected
 * with an error.  No hipMalloc, hipFree or allocating library call.
 * Synchronous; calls that overlap on one device (several threads)
 * race on the shared array. */
double simwave_hip_forward_device_noalloc_3d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Bytes of the static array of simwave_hip_forward_device_noalloc_3d_cd_f32
 * (host-only helper). */
size_t simwave_hip_scratch_bytes_3d_cd_f32(void);
/* Persistent plan (no counterpart in the reference, whose forward runs a
 * whole step range per call): the device entry split into create / run in
 * chunks / destroy, for callers that interleave shots (the serving loop) or
 * re-run step ranges (checkpointed recomputation).  Arguments as
 * simwave_hip_forward_device_3d_cd_f32; its host arrays are read only here.
 * flags as the device entry, plus 8: for stride 0, capture 48 steps once in a
 * hipGraph and replay it for every aligned 48-step stretch of a run (the
 * step number then comes from a device counter; results are identical).
 * Returns NULL on failure. */
void *simwave_hip_plan_create_3d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Enqueues the next `steps` steps of the plan on `stream`, ordered after the
 * plan's earlier work on any stream; does not wait for the device.  Returns
 * the steps left, or -1. */
long long simwave_hip_plan_run_3d_cd_f32(void *plan, size_t steps, void *stream);
/* Frees the plan once its enqueued work is done.  Returns 0, or -1. */
int simwave_hip_plan_destroy_3d_cd_f32(void *plan);
/* Batched runs (no counterpart in the reference; for the serving loop):
 * enqueues up to `steps` further steps of each of the `count` plans on
 * `stream`, after each plan's earlier work.  Plans that can be batched
 * (simwave_hip_plan_batchable_3d_cd_f32) advance together, radius by radius:
 * one step kernel and one receiver kernel per time step for all of them
 * (up to 48 per launch); the others run as simwave_hip_plan_run_3d_cd_f32.
 * left[i] receives plan i's remaining steps.  Returns 0, or -1. */
int simwave_hip_plans_run_3d_cd_f32(void **plans, int count, size_t steps,
                            void *stream, long long *left);
/* 1 when the plan can join batched runs (stride 0, fused z-marching step
 * kernel, no source box in the halo, flags without 1 and 4), else 0. */
int simwave_hip_plan_batchable_3d_cd_f32(void *plan);
/* Device gather (no counterpart in the reference; for the device-resident
 * shot, which builds its edge-extended models and damping mask on the GPU
 * from their distinct values): out[i][j][k] = small[idx0[i]][idx1[j]][idx2[k]]
 * for out of out_shape and small of small_shape (host arrays of ndim
 * entries, ndim the library's dimension; 2D: no k, idx2 ignored); out,
 * small and the int64 index arrays are device pointers.  Enqueued on
 * `stream` without waiting.  Returns 0, or -1. */
int simwave_hip_gather_3d_cd_f32(float *out, const float *small, int ndim,
                        const long long *out_shape, const long long *small_shape,
                        const long long *idx0, const long long *idx1,
                        const long long *idx2, void *stream);
/* Adjoint-state gradient (no counterpart in the reference; constant density).
 * Arguments of simwave_hip_forward_device_3d_cd_f32 (device arrays; stride 0,
 * steps 1..T, source boxes inside the interior), then: rec_adjoint, the
 * adjoint of the shot record [T][num_receivers] (device); grad_velocity,
 * float64 [extended grid] (device), set to the gradient with respect to
 * the velocity; grad_wavelet, [T][num_sources] (device) or NULL, set to
 * the gradient with respect to the wavelet samples; checkpoint_every, the
 * forward states saved every that many steps and recomputed per segment
 * (>= the step count: every level kept in device memory, no recomputation).
 * flags: 1 sets grad_velocity to the zero-lag cross-correlation image
 * sum_k u^k lambda^k of the forward and adjoint wavefields instead; 2 takes
 * rec_adjoint as observed data and drives the adjoint with the residual
 * (shot record - rec_adjoint) formed on the device after the forward (the
 * L2 misfit's gradient in one call); 4 runs the forward's receiver gather
 * on a second stream (large grids); 8 times every reverse update kernel
 * with HIP events (simwave_hip_last_timings: [4] their summed seconds,
 * [5] their count).
 * The shot starts from rest: u's three slots are zeroed first, whatever
 * they hold.  Also writes the shot record and leaves u as after a plain
 * forward.  Returns wall seconds, or -1. */
double simwave_hip_gradient_3d_cd_f32(
    float *u, float *velocity, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    const float *rec_adjoint, double *grad_velocity, float *grad_wavelet,
    size_t checkpoint_every, void *stream, unsigned flags);
/* Host-only helper (no device work, no reference counterpart): the
 * fixed-divisor parameters the kernels use for the divisor b (a squared
 * spacing, cd-3d wave.c:186).  Returns 1 and y = RN(1/b) and the safe
 * numerator range [lo, hi], or 0 when IEEE division is used. */
int simwave_hip_fixed_div_3d_cd_f32(float b, float *y, float *lo, float *hi);
/* Host-only helper (no device work, no reference counterpart): 1 when the
 * 3D tiled kernel forms the quotients sum/b of the divisor b as
 * RN(sum * RN64(1/b)), the float64 product rounded to float (float32
 * libraries, b = B 2^t with B odd, t <= 0, b in [2^-126, 2^24): equal to
 * the IEEE quotient for every float sum); 0 otherwise. */
int simwave_hip_f64_quotient_3d_cd_f32(float b);

/* ---- 3d_cd_f64 */
/* Host arrays, as the reference's forward(): reference constant_density/3d/wave.c:23-36
 * (`forward` in libsimwave_hip_3d_cd_f64.so has exactly this signature). */
double simwave_hip_forward_3d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots);
/* Device arrays for u, velocity, damp and receivers (the
 * rest stay on the host); runs on the HIP stream `stream` (NULL: the null
 * stream) and returns after the stream has drained.  The models are only
 * read.  flags: 1 time the step kernel with HIP events, 2 force the
 * generic kernels, 4 run the receiver gather of large grids on a second
 * stream of the library's (event-ordered; one run or plan per process may
 * hold it, see simwave_hip_overlap_acquire; an error while another does).  Same arguments as
 * reference constant_density/3d/wave.c:23-36, plus stream and flags. */
double simwave_hip_forward_device_3d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Allocation-free device entry (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results as simwave_hip_forward_device_3d_cd_f64, flags 0
 * or 2 only; the library's buffers come from a static 64 MiB device
 * array per library and device (per-point arrays of double: 24 bytes a grid
 * point, so at most about 2.7 M points), and larger grids are rejected
 * with an error.  No hipMalloc, hipFree or allocating library call.
 * Synchronous; calls that overlap on one device (several threads)
 * race on the shared array. */
double simwave_hip_forward_device_noalloc_3d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Bytes of the static array of simwave_hip_forward_device_noalloc_3d_cd_f64
 * (host-only helper). */
size_t simwave_hip_scratch_bytes_3d_cd_f64(void);
/* Persistent plan (no counterpart in the reference, whose forward runs a
 * whole step range per call): the device entry split into create / run in
 * chunks / destroy, for callers that interleave shots (the serving loop) or
 * re-run step ranges (checkpointed recomputation).  Arguments as
 * simwave_hip_forward_device_3d_cd_f64; its host arrays are read only here.
 * flags as the device entry, plus 8: for stride 0, capture 48 steps once in a
 * hipGraph and replay it for every aligned 48-step stretch of a run (the
 * step number then comes from a device counter; results are identical).
 * Returns NULL on failure. */
void *simwave_hip_plan_create_3d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Enqueues the next `steps` steps of the plan on `stream`, ordered after the
 * plan's earlier work on any stream; does not wait for the device.  Returns
 * the steps left, or -1. */
long long simwave_hip_plan_run_3d_cd_f64(void *plan, size_t steps, void *stream);
/* Frees the plan once its enqueued work is done.  Returns 0, or -1. */
int simwave_hip_plan_destroy_3d_cd_f64(void *plan);
/* Batched runs (no counterpart in the reference; for the serving loop):
 * enqueues up to `steps` further steps of each of the `count` plans on
 * `stream`, after each plan's earlier work.  Plans that can be batched
 * (simwave_hip_plan_batchable_3d_cd_f64) advance together, radius by radius:
 * one step kernel and one receiver kernel per time step for all of them
 * (up to 48 per launch); the others run as simwave_hip_plan_run_3d_cd_f64.
 * left[i] receives plan i's remaining steps.  Returns 0, or -1. */
int simwave_hip_plans_run_3d_cd_f64(void **plans, int count, size_t steps,
                            void *stream, long long *left);
/* 1 when the plan can join batched runs (stride 0, fused z-marching step
 * kernel, no source box in the halo, flags without 1 and 4), else 0. */
int simwave_hip_plan_batchable_3d_cd_f64(void *plan);
/* Device gather (no counterpart in the reference; for the device-resident
 * shot, which builds its edge-extended models and damping mask on the GPU
 * from their distinct values): out[i][j][k] = small[idx0[i]][idx1[j]][idx2[k]]
 * for out of out_shape and small of small_shape (host arrays of ndim
 * entries, ndim the library's dimension; 2D: no k, idx2 ignored); out,
 * small and the int64 index arrays are device pointers.  Enqueued on
 * `stream` without waiting.  Returns 0, or -1. */
int simwave_hip_gather_3d_cd_f64(double *out, const double *small, int ndim,
                        const long long *out_shape, const long long *small_shape,
                        const long long *idx0, const long long *idx1,
                        const long long *idx2, void *stream);
/* Adjoint-state gradient (no counterpart in the reference; constant density).
 * Arguments of simwave_hip_forward_device_3d_cd_f64 (device arrays; stride 0,
 * steps 1..T, source boxes inside the interior), then: rec_adjoint, the
 * adjoint of the shot record [T][num_receivers] (device); grad_velocity,
 * float64 [extended grid] (device), set to the gradient with respect to
 * the velocity; grad_wavelet, [T][num_sources] (device) or NULL, set to
 * the gradient with respect to the wavelet samples; checkpoint_every, the
 * forward states saved every that many steps and recomputed per segment
 * (>= the step count: every level kept in device memory, no recomputation).
 * flags: 1 sets grad_velocity to the zero-lag cross-correlation image
 * sum_k u^k lambda^k of the forward and adjoint wavefields instead; 2 takes
 * rec_adjoint as observed data and drives the adjoint with the residual
 * (shot record - rec_adjoint) formed on the device after the forward (the
 * L2 misfit's gradient in one call); 4 runs the forward's receiver gather
 * on a second stream (large grids); 8 times every reverse update kernel
 * with HIP events (simwave_hip_last_timings: [4] their summed seconds,
 * [5] their count).
 * The shot starts from rest: u's three slots are zeroed first, whatever
 * they hold.  Also writes the shot record and leaves u as after a plain
 * forward.  Returns wall seconds, or -1. */
double simwave_hip_gradient_3d_cd_f64(
    double *u, double *velocity, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    const double *rec_adjoint, double *grad_velocity, double *grad_wavelet,
    size_t checkpoint_every, void *stream, unsigned flags);
/* Host-only helper (no device work, no reference counterpart): the
 * fixed-divisor parameters the kernels use for the divisor b (a squared
 * spacing, cd-3d wave.c:186).  Returns 1 and y = RN(1/b) and the safe
 * numerator range [lo, hi], or 0 when IEEE division is used. */
int simwave_hip_fixed_div_3d_cd_f64(double b, double *y, double *lo, double *hi);
/* Host-only helper (no device work, no reference counterpart): 1 when the
 * 3D tiled kernel forms the quotients sum/b of the divisor b as
 * RN(sum * RN64(1/b)), the float64 product rounded to float (float32
 * libraries, b = B 2^t with B odd, t <= 0, b in [2^-126, 2^24): equal to
 * the IEEE quotient for every float sum); 0 otherwise. */
int simwave_hip_f64_quotient_3d_cd_f64(double b);

/* ---- 3d_vd_f32 */
/* Host arrays, as the reference's forward(): reference variable_density/3d/wave.c:23-36
 * (`forward` in libsimwave_hip_3d_vd_f32.so has exactly this signature). */
double simwave_hip_forward_3d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots);
/* Device arrays for u, velocity, density, damp and receivers (the
 * rest stay on the host); runs on the HIP stream `stream` (NULL: the null
 * stream) and returns after the stream has drained.  The models are only
 * read.  flags: 1 time the step kernel with HIP events, 2 force the
 * generic kernels, 4 run the receiver gather of large grids on a second
 * stream of the library's (event-ordered; one run or plan per process may
 * hold it, see simwave_hip_overlap_acquire; an error while another does).  Same arguments as
 * reference variable_density/3d/wave.c:23-36, plus stream and flags. */
double simwave_hip_forward_device_3d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Allocation-free device entry (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results as simwave_hip_forward_device_3d_vd_f32, flags 0
 * or 2 only; the library's buffers come from a static 64 MiB device
 * array per library and device (per-point arrays of float: 12 bytes a grid
 * point, so at most about 5.5 M points; from space order 12 on, the density
 * sums add 12 bytes a point: about 2.7 M points), and larger grids are
 * rejected with an error.  No hipMalloc, hipFree or allocating library call.
 * Synchronous; calls that overlap on one device (several threads)
 * race on the shared array. */
double simwave_hip_forward_device_noalloc_3d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Bytes of the static array of simwave_hip_forward_device_noalloc_3d_vd_f32
 * (host-only helper). */
size_t simwave_hip_scratch_bytes_3d_vd_f32(void);
/* Persistent plan (no counterpart in the reference, whose forward runs a
 * whole step range per call): the device entry split into create / run in
 * chunks / destroy, for callers that interleave shots (the serving loop) or
 * re-run step ranges (checkpointed recomputation).  Arguments as
 * simwave_hip_forward_device_3d_vd_f32; its host arrays are read only here.
 * flags as the device entry, plus 8: for stride 0, capture 48 steps once in a
 * hipGraph and replay it for every aligned 48-step stretch of a run (the
 * step number then comes from a device counter; results are identical).
 * Returns NULL on failure. */
void *simwave_hip_plan_create_3d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Enqueues the next `steps` steps of the plan on `stream`, ordered after the
 * plan's earlier work on any stream; does not wait for the device.  Returns
 * the steps left, or -1. */
long long simwave_hip_plan_run_3d_vd_f32(void *plan, size_t steps, void *stream);
/* Frees the plan once its enqueued work is done.  Returns 0, or -1. */
int simwave_hip_plan_destroy_3d_vd_f32(void *plan);
/* Batched runs (no counterpart in the reference; for the serving loop):
 * enqueues up to `steps` further steps of each of the `count` plans on
 * `stream`, after each plan's earlier work.  Plans that can be batched
 * (simwave_hip_plan_batchable_3d_vd_f32) advance together, radius by radius:
 * one step kernel and one receiver kernel per time step for all of them
 * (up to 48 per launch); the others run as simwave_hip_plan_run_3d_vd_f32.
 * left[i] receives plan i's remaining steps.  Returns 0, or -1. */
int simwave_hip_plans_run_3d_vd_f32(void **plans, int count, size_t steps,
                            void *stream, long long *left);
/* 1 when the plan can join batched runs (stride 0, fused z-marching step
 * kernel, no source box in the halo, flags without 1 and 4), else 0. */
int simwave_hip_plan_batchable_3d_vd_f32(void *plan);
/* Device gather (no counterpart in the reference; for the device-resident
 * shot, which builds its edge-extended models and damping mask on the GPU
 * from their distinct values): out[i][j][k] = small[idx0[i]][idx1[j]][idx2[k]]
 * for out of out_shape and small of small_shape (host arrays of ndim
 * entries, ndim the library's dimension; 2D: no k, idx2 ignored); out,
 * small and the int64 index arrays are device pointers.  Enqueued on
 * `stream` without waiting.  Returns 0, or -1. */
int simwave_hip_gather_3d_vd_f32(float *out, const float *small, int ndim,
                        const long long *out_shape, const long long *small_shape,
                        const long long *idx0, const long long *idx1,
                        const long long *idx2, void *stream);
/* Adjoint-state gradient (no counterpart in the reference; constant density only: this variable-density library returns -1).
 * Arguments of simwave_hip_forward_device_3d_vd_f32 (device arrays; stride 0,
 * steps 1..T, source boxes inside the interior), then: rec_adjoint, the
 * adjoint of the shot record [T][num_receivers] (device); grad_velocity,
 * float64 [extended grid] (device), set to the gradient with respect to
 * the velocity; grad_wavelet, [T][num_sources] (device) or NULL, set to
 * the gradient with respect to the wavelet samples; checkpoint_every, the
 * forward states saved every that many steps and recomputed per segment
 * (>= the step count: every level kept in device memory, no recomputation).
 * flags: 1 sets grad_velocity to the zero-lag cross-correlation image
 * sum_k u^k lambda^k of the forward and adjoint wavefields instead; 2 takes
 * rec_adjoint as observed data and drives the adjoint with the residual
 * (shot record - rec_adjoint) formed on the device after the forward (the
 * L2 misfit's gradient in one call); 4 runs the forward's receiver gather
 * on a second stream (large grids); 8 times every reverse update kernel
 * with HIP events (simwave_hip_last_timings: [4] their summed seconds,
 * [5] their count).
 * The shot starts from rest: u's three slots are zeroed first, whatever
 * they hold.  Also writes the shot record and leaves u as after a plain
 * forward.  Returns wall seconds, or -1. */
double simwave_hip_gradient_3d_vd_f32(
    float *u, float *velocity, float *density, float *damp,
    float *wavelet, size_t wavelet_size, size_t wavelet_count,
    float *coeff_order2, float *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    float *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    float *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    float *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, float dz, float dx, float dy,
    size_t saving_stride, float dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    const float *rec_adjoint, double *grad_velocity, float *grad_wavelet,
    size_t checkpoint_every, void *stream, unsigned flags);
/* Host-only helper (no device work, no reference counterpart): the
 * fixed-divisor parameters the kernels use for the divisor b (a squared
 * spacing, cd-3d wave.c:186).  Returns 1 and y = RN(1/b) and the safe
 * numerator range [lo, hi], or 0 when IEEE division is used. */
int simwave_hip_fixed_div_3d_vd_f32(float b, float *y, float *lo, float *hi);
/* Host-only helper (no device work, no reference counterpart): 1 when the
 * 3D tiled kernel forms the quotients sum/b of the divisor b as
 * RN(sum * RN64(1/b)), the float64 product rounded to float (float32
 * libraries, b = B 2^t with B odd, t <= 0, b in [2^-126, 2^24): equal to
 * the IEEE quotient for every float sum); 0 otherwise. */
int simwave_hip_f64_quotient_3d_vd_f32(float b);

/* ---- 3d_vd_f64 */
/* Host arrays, as the reference's forward(): reference variable_density/3d/wave.c:23-36
 * (`forward` in libsimwave_hip_3d_vd_f64.so has exactly this signature). */
double simwave_hip_forward_3d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots);
/* Device arrays for u, velocity, density, damp and receivers (the
 * rest stay on the host); runs on the HIP stream `stream` (NULL: the null
 * stream) and returns after the stream has drained.  The models are only
 * read.  flags: 1 time the step kernel with HIP events, 2 force the
 * generic kernels, 4 run the receiver gather of large grids on a second
 * stream of the library's (event-ordered; one run or plan per process may
 * hold it, see simwave_hip_overlap_acquire; an error while another does).  Same arguments as
 * reference variable_density/3d/wave.c:23-36, plus stream and flags. */
double simwave_hip_forward_device_3d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Allocation-free device entry (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results as simwave_hip_forward_device_3d_vd_f64, flags 0
 * or 2 only; the library's buffers come from a static 64 MiB device
 * array per library and device (per-point arrays of double: 24 bytes a grid
 * point, so at most about 2.7 M points; from space order 12 on, the density
 * sums add 24 bytes a point: about 1.3 M points), and larger grids are
 * rejected with an error.  No hipMalloc, hipFree or allocating library call.
 * Synchronous; calls that overlap on one device (several threads)
 * race on the shared array. */
double simwave_hip_forward_device_noalloc_3d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Bytes of the static array of simwave_hip_forward_device_noalloc_3d_vd_f64
 * (host-only helper). */
size_t simwave_hip_scratch_bytes_3d_vd_f64(void);
/* Persistent plan (no counterpart in the reference, whose forward runs a
 * whole step range per call): the device entry split into create / run in
 * chunks / destroy, for callers that interleave shots (the serving loop) or
 * re-run step ranges (checkpointed recomputation).  Arguments as
 * simwave_hip_forward_device_3d_vd_f64; its host arrays are read only here.
 * flags as the device entry, plus 8: for stride 0, capture 48 steps once in a
 * hipGraph and replay it for every aligned 48-step stretch of a run (the
 * step number then comes from a device counter; results are identical).
 * Returns NULL on failure. */
void *simwave_hip_plan_create_3d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    void *stream, unsigned flags);
/* Enqueues the next `steps` steps of the plan on `stream`, ordered after the
 * plan's earlier work on any stream; does not wait for the device.  Returns
 * the steps left, or -1. */
long long simwave_hip_plan_run_3d_vd_f64(void *plan, size_t steps, void *stream);
/* Frees the plan once its enqueued work is done.  Returns 0, or -1. */
int simwave_hip_plan_destroy_3d_vd_f64(void *plan);
/* Batched runs (no counterpart in the reference; for the serving loop):
 * enqueues up to `steps` further steps of each of the `count` plans on
 * `stream`, after each plan's earlier work.  Plans that can be batched
 * (simwave_hip_plan_batchable_3d_vd_f64) advance together, radius by radius:
 * one step kernel and one receiver kernel per time step for all of them
 * (up to 48 per launch); the others run as simwave_hip_plan_run_3d_vd_f64.
 * left[i] receives plan i's remaining steps.  Returns 0, or -1. */
int simwave_hip_plans_run_3d_vd_f64(void **plans, int count, size_t steps,
                            void *stream, long long *left);
/* 1 when the plan can join batched runs (stride 0, fused z-marching step
 * kernel, no source box in the halo, flags without 1 and 4), else 0. */
int simwave_hip_plan_batchable_3d_vd_f64(void *plan);
/* Device gather (no counterpart in the reference; for the device-resident
 * shot, which builds its edge-extended models and damping mask on the GPU
 * from their distinct values): out[i][j][k] = small[idx0[i]][idx1[j]][idx2[k]]
 * for out of out_shape and small of small_shape (host arrays of ndim
 * entries, ndim the library's dimension; 2D: no k, idx2 ignored); out,
 * small and the int64 index arrays are device pointers.  Enqueued on
 * `stream` without waiting.  Returns 0, or -1. */
int simwave_hip_gather_3d_vd_f64(double *out, const double *small, int ndim,
                        const long long *out_shape, const long long *small_shape,
                        const long long *idx0, const long long *idx1,
                        const long long *idx2, void *stream);
/* Adjoint-state gradient (no counterpart in the reference; constant density only: this variable-density library returns -1).
 * Arguments of simwave_hip_forward_device_3d_vd_f64 (device arrays; stride 0,
 * steps 1..T, source boxes inside the interior), then: rec_adjoint, the
 * adjoint of the shot record [T][num_receivers] (device); grad_velocity,
 * float64 [extended grid] (device), set to the gradient with respect to
 * the velocity; grad_wavelet, [T][num_sources] (device) or NULL, set to
 * the gradient with respect to the wavelet samples; checkpoint_every, the
 * forward states saved every that many steps and recomputed per segment
 * (>= the step count: every level kept in device memory, no recomputation).
 * flags: 1 sets grad_velocity to the zero-lag cross-correlation image
 * sum_k u^k lambda^k of the forward and adjoint wavefields instead; 2 takes
 * rec_adjoint as observed data and drives the adjoint with the residual
 * (shot record - rec_adjoint) formed on the device after the forward (the
 * L2 misfit's gradient in one call); 4 runs the forward's receiver gather
 * on a second stream (large grids); 8 times every reverse update kernel
 * with HIP events (simwave_hip_last_timings: [4] their summed seconds,
 * [5] their count).
 * The shot starts from rest: u's three slots are zeroed first, whatever
 * they hold.  Also writes the shot record and leaves u as after a plain
 * forward.  Returns wall seconds, or -1. */
double simwave_hip_gradient_3d_vd_f64(
    double *u, double *velocity, double *density, double *damp,
    double *wavelet, size_t wavelet_size, size_t wavelet_count,
    double *coeff_order2, double *coeff_order1,
    size_t *boundary_conditions,
    size_t *src_points_interval, size_t src_points_interval_size,
    double *src_points_values, size_t src_points_values_size,
    size_t *src_points_values_offset,
    size_t *rec_points_interval, size_t rec_points_interval_size,
    double *rec_points_values, size_t rec_points_values_size,
    size_t *rec_points_values_offset,
    double *receivers, size_t num_sources, size_t num_receivers,
    size_t nz, size_t nx, size_t ny, double dz, double dx, double dy,
    size_t saving_stride, double dt, size_t begin_timestep,
    size_t end_timestep, size_t space_order, size_t num_snapshots,
    const double *rec_adjoint, double *grad_velocity, double *grad_wavelet,
    size_t checkpoint_every, void *stream, unsigned flags);
/* Host-only helper (no device work, no reference counterpart): the
 * fixed-divisor parameters the kernels use for the divisor b (a squared
 * spacing, cd-3d wave.c:186).  Returns 1 and y = RN(1/b) and the safe
 * numerator range [lo, hi], or 0 when IEEE division is used. */
int simwave_hip_fixed_div_3d_vd_f64(double b, double *y, double *lo, double *hi);
/* Host-only helper (no device work, no reference counterpart): 1 when the
 * 3D tiled kernel forms the quotients sum/b of the divisor b as
 * RN(sum * RN64(1/b)), the float64 product rounded to float (float32
 * libraries, b = B 2^t with B odd, t <= 0, b in [2^-126, 2^24): equal to
 * the IEEE quotient for every float sum); 0 otherwise. */
int simwave_hip_f64_quotient_3d_vd_f64(double b);

/* ---- every library */
/* Message of the last failed call on this thread ("" after a success). */
const char *simwave_hip_last_error(void);
/* Phase timings of the last call on this thread: total s, prepare s, time
 * loop s, download s, step-kernel s (flags & 1), step launches, 1 when the
 * fused kernels ran, step kernel (1 3D tiled, 2 z-marching, 0 other).
 * Writes min(n, 8) values; returns that count. */
int simwave_hip_last_timings(double *out, int n);
/* The library's variant, e.g. "3d_cd_f32". */
const char *simwave_hip_variant(void);
/* No reference counterpart.  Hands the freed memory that the current
 * device's default pool keeps for reuse (the library's stream-ordered
 * allocations; up to 8 GiB is kept between runs) back to the device, after
 * waiting for the device, so that a free-memory query sees it.  0, or -1. */
int simwave_hip_trim_pool(void);

/* No reference counterpart (the reference's forward takes host arrays
 * only, wave.c:23-36).  Device memory, pinned host memory, copies, zero
 * fills, streams and events on the current device, for callers that keep
 * a shot resident through simwave_hip_forward_device_<variant> or the plan
 * entries and hold no other GPU runtime (the Python device path and serving
 * loop use these).  Streams are non-blocking with respect to the null
 * stream, at the default priority.  Each returns 0, or -1 with
 * simwave_hip_last_error() set. */
/* Stream-ordered device memory from the current device's default pool:
 * allocated for work enqueued on `stream` after the call; a free returns it
 * once the work enqueued on `stream` before the call is done (work on other
 * streams that reads it must be complete).  Neither waits on the host. */
int simwave_hip_device_alloc(unsigned long long bytes, void *stream, void **ptr);
int simwave_hip_device_free(void *ptr, void *stream);
int simwave_hip_host_alloc(unsigned long long bytes, void **ptr);
int simwave_hip_host_free(void *ptr);
/* Stream-ordered copy of `bytes` between device, pinned or pageable host
 * memory (the direction follows from the pointers). */
int simwave_hip_copy(void *dst, const void *src, unsigned long long bytes, void *stream);
/* Zeroes `bytes` (a multiple of 4) of device memory with a fill kernel on
 * `stream`. */
int simwave_hip_zero(void *dst, unsigned long long bytes, void *stream);
int simwave_hip_stream_create(void **stream);
int simwave_hip_stream_destroy(void *stream);
int simwave_hip_stream_synchronize(void *stream);
int simwave_hip_event_create(void **event);
int simwave_hip_event_record(void *event, void *stream);
int simwave_hip_event_synchronize(void *event);
int simwave_hip_event_destroy(void *event);

/* ---- libsimwave_hip_common.so (linked by every variant library; one copy
 * per process).  No reference counterpart.  The receiver-overlap lease:
 * flag 4 joins a second stream of the library's with event waits in both
 * directions, and such waits enqueued from several host threads at once
 * can order into a cycle on the device's shared hardware queues, so one
 * run or plan per process holds it.  The entry points take it themselves
 * (flag 4 returns an error while another holds it; the host entry `forward`
 * then runs its receivers on its one stream instead). */
/* 1 if the caller now holds the lease, 0 if another holder has it. */
int simwave_hip_overlap_acquire(void);
/* Releases the lease. */
void simwave_hip_overlap_release(void);
/* Current holders (0 or 1). */
int simwave_hip_overlap_holders(void);

#ifdef __cplusplus
}
#endif

#endif /* SIMWAVE_HIP_H */
