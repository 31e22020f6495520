/*
 * ORACLE — test infrastructure only. Not part of the product path.
 *
 * CPU restatement (C99, optional OpenMP) of simwave's acoustic `forward()`
 * operator.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load the libraries built from this file, and only as the checker or
 * the timed CPU baseline.
 *
 * One source file covers the eight reference variants through macros:
 *   SW_DIM      2 | 3
 *   SW_VD       0 (constant density) | 1 (variable density)
 *   SW_DOUBLE   defined -> f_type is double, else float
 *   SW_OPENMP   defined -> outer loops run under OpenMP (the cpu_openmp twin)
 *
 * Every build exports `forward` with the exact signature of the reference
 * variant it restates:
 *   cd-3d  simwave/kernel/backend/c_code/forward/constant_density/3d/wave.c:23-36
 *   cd-2d  simwave/kernel/backend/c_code/forward/constant_density/2d/wave.c:23-36
 *   vd-3d  simwave/kernel/backend/c_code/forward/variable_density/3d/wave.c:23-36
 *   vd-2d  simwave/kernel/backend/c_code/forward/variable_density/2d/wave.c:23-36
 *
 * The floating-point expressions keep the reference's C types and
 * left-to-right evaluation (float products, the double-promoted
 * `1.0/(v*v)`, `1.0 +/- ...`, `2.0/den * u`) so that a build with the same
 * flags (-O3 -std=c99, no contraction) is bit-identical to the reference.
 *
 * Parity pinning (tests/test_oracle.py): bit identity with the reference C
 * compiled from its own sources (oracle/Makefile -> oracle/_ref/), both as
 * recorded digests (tests/golden/ref_outputs.json, all eight variants'
 * families incl. the 3D golden grid) and live where oracle/_ref is built;
 * the golden 2D wavefields the reference's tests ship (atol 1e-5, the
 * reference's own bar); the reference's invariants (saving stride, OpenMP
 * vs sequential, unit density) and the analytic 2D Hankel response.  No
 * function of this file is left unpinned.
 */
#include <stddef.h>
#include <sys/time.h>
#ifdef SW_OPENMP
#include <omp.h>
#endif

#ifdef SW_DOUBLE
typedef double f_type;
#else
typedef float f_type;
#endif

#ifndef SW_DIM
#error "SW_DIM must be 2 or 3"
#endif
#ifndef SW_VD
#define SW_VD 0
#endif

#ifdef SW_OPENMP
#define SW_PRAGMA_FOR _Pragma("omp parallel for")
#else
#define SW_PRAGMA_FOR
#endif

/* Grid description shared by the section helpers.  Axis order is (z, x[, y]);
 * the last axis is contiguous.  In 2D `ny` is 1 and the contiguous axis is x. */
typedef struct {
    size_t nz, nx, ny;     /* extended shape (ny == 1 in 2D)             */
    size_t npts;           /* points per time slot                        */
    size_t r;              /* stencil radius = space_order / 2            */
    size_t s_z, s_x, s_c;  /* strides of z, middle and contiguous axis    */
} grid_t;

/* ---------------------------------------------------------------- Section 1
 * Leap-frog update with damping over the interior [r, n-r) of every axis.
 * Reference: cd-3d wave.c:142-188, cd-2d wave.c:140-180,
 *            vd-3d wave.c:145-214, vd-2d wave.c:128-201.
 */
static void update_interior(const grid_t *g, f_type *uprev, const f_type *ucur,
                            f_type *unext, const f_type *vel,
                            const f_type *rho, const f_type *damp,
                            const f_type *c2, const f_type *c1,
                            f_type dz2, f_type dx2, f_type dy2,
                            f_type dt, f_type dt2)
{
    const size_t r = g->r;
    (void)rho; (void)c1; (void)dy2;
    SW_PRAGMA_FOR
    for (size_t i = r; i < g->nz - r; i++) {
#if SW_DIM == 3
        for (size_t j = r; j < g->nx - r; j++) {
            for (size_t k = r; k < g->ny - r; k++) {
                const size_t p = (i * g->nx + j) * g->ny + k;
#else
        {
            for (size_t k = r; k < g->nx - r; k++) {
                const size_t p = i * g->nx + k;
#endif
                const f_type *c = ucur + p;
                /* per-axis second-derivative sums, centre term first */
                f_type acc_c = c2[0] * c[0];   /* contiguous axis (y / 2D x) */
#if SW_DIM == 3
                f_type acc_m = c2[0] * c[0];   /* middle axis x (3D only)    */
#endif
                f_type acc_z = c2[0] * c[0];
#if SW_VD
                f_type dpu_c = 0.0, dpr_c = 0.0;
#if SW_DIM == 3
                f_type dpu_m = 0.0, dpr_m = 0.0;
#endif
                f_type dpu_z = 0.0, dpr_z = 0.0;
                const f_type *d = rho + p;
#endif
                for (size_t m = 1; m <= r; m++) {
                    acc_c += c2[m] * (c[m] + c[-(ptrdiff_t)m]);
#if SW_VD
                    dpu_c += c1[m] * (c[m] - c[-(ptrdiff_t)m]);
                    dpr_c += c1[m] * (d[m] - d[-(ptrdiff_t)m]);
#endif
#if SW_DIM == 3
                    {
                        const size_t o = m * g->ny;
                        acc_m += c2[m] * (c[o] + c[-(ptrdiff_t)o]);
#if SW_VD
                        /* the reference steps the middle-axis first
                         * derivatives by nx, not ny (vd-3d wave.c:185-186) */
                        const size_t oq = m * g->nx;
                        dpu_m += c1[m] * (c[oq] - c[-(ptrdiff_t)oq]);
                        dpr_m += c1[m] * (d[oq] - d[-(ptrdiff_t)oq]);
#endif
                    }
#endif
                    {
                        const size_t o = m * g->s_z;
                        acc_z += c2[m] * (c[o] + c[-(ptrdiff_t)o]);
#if SW_VD
                        dpu_z += c1[m] * (c[o] - c[-(ptrdiff_t)o]);
                        dpr_z += c1[m] * (d[o] - d[-(ptrdiff_t)o]);
#endif
                    }
                }
                f_type value = 0.0;
#if SW_DIM == 3
                value += acc_c / dy2 + acc_m / dx2 + acc_z / dz2;
#else
                value += acc_c / dx2 + acc_z / dz2;
#endif
#if SW_VD
#if SW_DIM == 3
                {
                    f_type t_c = (dpu_c * dpr_c) / (4 * dy2);
                    f_type t_m = (dpu_m * dpr_m) / (4 * dx2);
                    f_type t_z = (dpu_z * dpr_z) / (4 * dz2);
                    value -= (t_c + t_m + t_z) / d[0];
                }
#else
                {
                    f_type t_c = (dpu_c * dpr_c) / (4 * dx2);
                    f_type t_z = (dpu_z * dpr_z) / (4 * dz2);
                    value -= (t_c + t_z) / d[0];
                }
#endif
#endif
                {
                    const f_type v = vel[p];
                    const f_type slow = 1.0 / (v * v);
                    const f_type den = (1.0 + damp[p] * dt / (2 * slow));
                    const f_type num = (1.0 - damp[p] * dt / (2 * slow));
                    value *= (dt2 / slow) / den;
                    unext[p] = 2.0 / den * c[0] - (num / den) * uprev[p] + value;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------- Section 2
 * Windowed-sinc source injection into level n+1, sources in index order.
 * Reference: cd-3d wave.c:208-295, cd-2d wave.c:199-271.
 */
static void inject_sources(const grid_t *g, f_type *unext, const f_type *vel,
                           const f_type *damp, const f_type *wavelet,
                           size_t wavelet_count, size_t n,
                           const size_t *box, const f_type *w,
                           const size_t *woff, size_t nsrc,
                           f_type dt, f_type dt2)
{
    const size_t D = SW_DIM;
    for (size_t s = 0; s < nsrc; s++) {
        const size_t widx = (wavelet_count > 1) ? (n - 1) * nsrc + s : n - 1;
        const f_type amp = wavelet[widx];
        if (amp == 0.0)
            continue;
        const size_t *b = box + s * 2 * D;
        const size_t cz = b[1] - b[0] + 1;
#if SW_DIM == 3
        const size_t cx = b[3] - b[2] + 1;
        const f_type *wz = w + woff[s], *wx = wz + cz, *wy = wx + cx;
#else
        const f_type *wz = w + woff[s], *wx = wz + cz;
#endif
        for (size_t i = b[0]; i <= b[1]; i++) {
            for (size_t j = b[2]; j <= b[3]; j++) {
#if SW_DIM == 3
                for (size_t k = b[4]; k <= b[5]; k++) {
                    const f_type kws = wz[i - b[0]] * wx[j - b[2]] * wy[k - b[4]];
                    const size_t p = (i * g->nx + j) * g->ny + k;
#else
                {
                    const f_type kws = wz[i - b[0]] * wx[j - b[2]];
                    const size_t p = i * g->nx + j;
#endif
                    const f_type v = vel[p];
                    const f_type slow = 1.0 / (v * v);
                    const f_type den = (1.0 + damp[p] * dt / (2 * slow));
                    const f_type add = dt2 / slow * kws * amp / den;
                    unext[p] += add;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------- Section 3
 * Boundary conditions on level n+1.  Codes: 0 none, 1 null Dirichlet (zero
 * the first interior plane), 2 null Neumann (mirror into the halo).  The
 * contiguous axis goes first, then the middle axis (3D), then z; each face
 * line spans only the interior of the other axes.
 * Reference: cd-3d wave.c:304-480, cd-2d wave.c:280-393.
 *
 * One line is handled by `face_line`: `base` is the offset of the line's
 * (interior) origin, `st` the stride along the line's own axis, `n` the axis
 * length.  The four operations run in the reference order.
 */
static void face_line(f_type *u, size_t base, size_t st, size_t n, size_t r,
                      size_t before, size_t after)
{
    const size_t lo = base + r * st;            /* first interior plane */
    const size_t hi = base + (n - r - 1) * st;  /* last interior plane  */
    if (before == 1)
        u[lo] = 0.0;
    if (before == 2)
        for (size_t m = 1; m <= r; m++)
            u[lo - m * st] = u[lo + m * st];
    if (after == 1)
        u[hi] = 0.0;
    if (after == 2)
        for (size_t m = 1; m <= r; m++)
            u[hi + m * st] = u[hi - m * st];
}

static void apply_bc(const grid_t *g, f_type *u, const size_t *bc)
{
    const size_t r = g->r;
#if SW_DIM == 3
    /* y faces: lines along the contiguous axis */
    if (bc[4] || bc[5]) {
        SW_PRAGMA_FOR
        for (size_t i = r; i < g->nz - r; i++)
            for (size_t j = r; j < g->nx - r; j++)
                face_line(u, (i * g->nx + j) * g->ny, 1, g->ny, r, bc[4], bc[5]);
    }
    /* x faces */
    if (bc[2] || bc[3]) {
        SW_PRAGMA_FOR
        for (size_t i = r; i < g->nz - r; i++)
            for (size_t k = r; k < g->ny - r; k++)
                face_line(u, i * g->nx * g->ny + k, g->ny, g->nx, r, bc[2], bc[3]);
    }
    /* z faces */
    if (bc[0] || bc[1]) {
        SW_PRAGMA_FOR
        for (size_t j = r; j < g->nx - r; j++)
            for (size_t k = r; k < g->ny - r; k++)
                face_line(u, j * g->ny + k, g->nx * g->ny, g->nz, r, bc[0], bc[1]);
    }
#else
    /* x faces: lines along the contiguous axis */
    if (bc[2] || bc[3]) {
        SW_PRAGMA_FOR
        for (size_t i = r; i < g->nz - r; i++)
            face_line(u, i * g->nx, 1, g->nx, r, bc[2], bc[3]);
    }
    /* z faces */
    if (bc[0] || bc[1]) {
        SW_PRAGMA_FOR
        for (size_t j = r; j < g->nx - r; j++)
            face_line(u, j, g->nx, g->nz, r, bc[0], bc[1]);
    }
#endif
}

/* ---------------------------------------------------------------- Section 4
 * Receivers sample level n (the input of this step) into row n-1.
 * Reference: cd-3d wave.c:499-566, cd-2d wave.c:412-464.
 */
static void sample_receivers(const grid_t *g, const f_type *ucur,
                             f_type *rec_out, const size_t *box,
                             const f_type *w, const size_t *woff, size_t nrec)
{
    const size_t D = SW_DIM;
    SW_PRAGMA_FOR
    for (size_t q = 0; q < nrec; q++) {
        const size_t *b = box + q * 2 * D;
        const size_t cz = b[1] - b[0] + 1;
#if SW_DIM == 3
        const size_t cx = b[3] - b[2] + 1;
        const f_type *wz = w + woff[q], *wx = wz + cz, *wy = wx + cx;
#else
        const f_type *wz = w + woff[q], *wx = wz + cz;
#endif
        f_type acc = 0.0;
        for (size_t i = b[0]; i <= b[1]; i++) {
            for (size_t j = b[2]; j <= b[3]; j++) {
#if SW_DIM == 3
                for (size_t k = b[4]; k <= b[5]; k++) {
                    const f_type kws = wz[i - b[0]] * wx[j - b[2]] * wy[k - b[4]];
                    acc += ucur[(i * g->nx + j) * g->ny + k] * kws;
                }
#else
                const f_type kws = wz[i - b[0]] * wx[j - b[2]];
                acc += ucur[i * g->nx + j] * kws;
#endif
            }
        }
        rec_out[q] = acc;
    }
}

/* ---------------------------------------------------------------- driver
 * Time loop and snapshot-slot bookkeeping.
 * Reference: cd-3d wave.c:110-124 (slot choice), :569-618 (stride logic).
 */
static double run(f_type *u, const f_type *vel, const f_type *rho,
                  const f_type *damp, const f_type *wavelet,
                  size_t wavelet_count, const f_type *c2, const f_type *c1,
                  const size_t *bc,
                  const size_t *sbox, const f_type *sw, const size_t *soff,
                  const size_t *rbox, const f_type *rw, const size_t *roff,
                  f_type *receivers, size_t nsrc, size_t nrec,
                  size_t nz, size_t nx, size_t ny,
                  f_type dz, f_type dx, f_type dy,
                  size_t stride, f_type dt, size_t n0, size_t n1,
                  size_t space_order)
{
    struct timeval t_start, t_end;
    gettimeofday(&t_start, NULL);

    grid_t g;
    g.nz = nz; g.nx = nx; g.ny = ny;
    g.npts = nz * nx * ny;
    g.r = space_order / 2;
#if SW_DIM == 3
    g.s_z = nx * ny; g.s_x = ny; g.s_c = 1;
#else
    g.s_z = nx; g.s_x = 1; g.s_c = 1;
#endif
    const f_type dz2 = dz * dz, dx2 = dx * dx, dy2 = dy * dy, dt2 = dt * dt;

    size_t prev = 0, cur = 1, next = 2;
    for (size_t n = n0; n <= n1; n++) {
        if (stride == 0) {
            prev = (n - 1) % 3; cur = n % 3; next = (n + 1) % 3;
        } else if (stride == 1) {
            prev = n - 1; cur = n; next = n + 1;
        }
        f_type *up = u + prev * g.npts;
        f_type *uc = u + cur * g.npts;
        f_type *un = u + next * g.npts;

        update_interior(&g, up, uc, un, vel, rho, damp, c2, c1,
                        dz2, dx2, dy2, dt, dt2);
        inject_sources(&g, un, vel, damp, wavelet, wavelet_count, n,
                       sbox, sw, soff, nsrc, dt, dt2);
        apply_bc(&g, un, bc);
        sample_receivers(&g, uc, receivers + (n - 1) * nrec,
                         rbox, rw, roff, nrec);

        if (stride > 1) {
            if (n % stride == 1) {
                prev = cur;
                cur += 1;
                next += 1;
                if (stride % 2 == 0 && n < n1) {
                    /* even stride: exchange the contents of the two slots */
                    const size_t t = cur; cur = next; next = t;
                    f_type *a = u + cur * g.npts, *b = u + next * g.npts;
                    SW_PRAGMA_FOR
                    for (size_t p = 0; p < g.npts; p++) {
                        const f_type x = a[p]; a[p] = b[p]; b[p] = x;
                    }
                }
            } else {
                prev = cur;
                cur = next;
                next = prev;
            }
        }
    }

    gettimeofday(&t_end, NULL);
    return (double)(t_end.tv_sec - t_start.tv_sec)
         + (double)(t_end.tv_usec - t_start.tv_usec) / 1000000.0;
}

/* ------------------------------------------------------------- entry points
 * Same parameter lists and order as the reference `forward` of each variant.
 * The *_size arguments and num_snapshots are accepted for ABI parity; the
 * reference's CPU path does not read them either.
 */
#if SW_DIM == 3 && !SW_VD
double forward(f_type *u, f_type *velocity, f_type *damp,
               f_type *wavelet, size_t wavelet_size, size_t wavelet_count,
               f_type *coeff, size_t *boundary_conditions,
               size_t *src_points_interval, size_t src_points_interval_size,
               f_type *src_points_values, size_t src_points_values_size,
               size_t *src_points_values_offset,
               size_t *rec_points_interval, size_t rec_points_interval_size,
               f_type *rec_points_values, size_t rec_points_values_size,
               size_t *rec_points_values_offset,
               f_type *receivers, size_t num_sources, size_t num_receivers,
               size_t nz, size_t nx, size_t ny, f_type dz, f_type dx, f_type dy,
               size_t saving_stride, f_type dt,
               size_t begin_timestep, size_t end_timestep,
               size_t space_order, size_t num_snapshots)
{
    (void)wavelet_size; (void)src_points_interval_size;
    (void)src_points_values_size; (void)rec_points_interval_size;
    (void)rec_points_values_size; (void)num_snapshots;
    return run(u, velocity, NULL, damp, wavelet, wavelet_count, coeff, NULL,
               boundary_conditions, src_points_interval, src_points_values,
               src_points_values_offset, rec_points_interval,
               rec_points_values, rec_points_values_offset, receivers,
               num_sources, num_receivers, nz, nx, ny, dz, dx, dy,
               saving_stride, dt, begin_timestep, end_timestep, space_order);
}
#elif SW_DIM == 3 && SW_VD
double forward(f_type *u, f_type *velocity, f_type *density, f_type *damp,
               f_type *wavelet, size_t wavelet_size, size_t wavelet_count,
               f_type *coeff_order2, f_type *coeff_order1,
               size_t *boundary_conditions,
               size_t *src_points_interval, size_t src_points_interval_size,
               f_type *src_points_values, size_t src_points_values_size,
               size_t *src_points_values_offset,
               size_t *rec_points_interval, size_t rec_points_interval_size,
               f_type *rec_points_values, size_t rec_points_values_size,
               size_t *rec_points_values_offset,
               f_type *receivers, size_t num_sources, size_t num_receivers,
               size_t nz, size_t nx, size_t ny, f_type dz, f_type dx, f_type dy,
               size_t saving_stride, f_type dt,
               size_t begin_timestep, size_t end_timestep,
               size_t space_order, size_t num_snapshots)
{
    (void)wavelet_size; (void)src_points_interval_size;
    (void)src_points_values_size; (void)rec_points_interval_size;
    (void)rec_points_values_size; (void)num_snapshots;
    return run(u, velocity, density, damp, wavelet, wavelet_count,
               coeff_order2, coeff_order1, boundary_conditions,
               src_points_interval, src_points_values,
               src_points_values_offset, rec_points_interval,
               rec_points_values, rec_points_values_offset, receivers,
               num_sources, num_receivers, nz, nx, ny, dz, dx, dy,
               saving_stride, dt, begin_timestep, end_timestep, space_order);
}
#elif SW_DIM == 2 && !SW_VD
double forward(f_type *u, f_type *velocity, f_type *damp,
               f_type *wavelet, size_t wavelet_size, size_t wavelet_count,
               f_type *coeff, size_t *boundary_conditions,
               size_t *src_points_interval, size_t src_points_interval_size,
               f_type *src_points_values, size_t src_points_values_size,
               size_t *src_points_values_offset,
               size_t *rec_points_interval, size_t rec_points_interval_size,
               f_type *rec_points_values, size_t rec_points_values_size,
               size_t *rec_points_values_offset,
               f_type *receivers, size_t num_sources, size_t num_receivers,
               size_t nz, size_t nx, f_type dz, f_type dx,
               size_t saving_stride, f_type dt,
               size_t begin_timestep, size_t end_timestep,
               size_t space_order, size_t num_snapshots)
{
    (void)wavelet_size; (void)src_points_interval_size;
    (void)src_points_values_size; (void)rec_points_interval_size;
    (void)rec_points_values_size; (void)num_snapshots;
    return run(u, velocity, NULL, damp, wavelet, wavelet_count, coeff, NULL,
               boundary_conditions, src_points_interval, src_points_values,
               src_points_values_offset, rec_points_interval,
               rec_points_values, rec_points_values_offset, receivers,
               num_sources, num_receivers, nz, nx, 1, dz, dx, dx,
               saving_stride, dt, begin_timestep, end_timestep, space_order);
}
#else
double forward(f_type *u, f_type *velocity, f_type *density, f_type *damp,
               f_type *wavelet, size_t wavelet_size, size_t wavelet_count,
               f_type *coeff_order2, f_type *coeff_order1,
               size_t *boundary_conditions,
               size_t *src_points_interval, size_t src_points_interval_size,
               f_type *src_points_values, size_t src_points_values_size,
               size_t *src_points_values_offset,
               size_t *rec_points_interval, size_t rec_points_interval_size,
               f_type *rec_points_values, size_t rec_points_values_size,
               size_t *rec_points_values_offset,
               f_type *receivers, size_t num_sources, size_t num_receivers,
               size_t nz, size_t nx, f_type dz, f_type dx,
               size_t saving_stride, f_type dt,
               size_t begin_timestep, size_t end_timestep,
               size_t space_order, size_t num_snapshots)
{
    (void)wavelet_size; (void)src_points_interval_size;
    (void)src_points_values_size; (void)rec_points_interval_size;
    (void)rec_points_values_size; (void)num_snapshots;
    return run(u, velocity, density, damp, wavelet, wavelet_count,
               coeff_order2, coeff_order1, boundary_conditions,
               src_points_interval, src_points_values,
               src_points_values_offset, rec_points_interval,
               rec_points_values, rec_points_values_offset, receivers,
               num_sources, num_receivers, nz, nx, 1, dz, dx, dx,
               saving_stride, dt, begin_timestep, end_timestep, space_order);
}
#endif
