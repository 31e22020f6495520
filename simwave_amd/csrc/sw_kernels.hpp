// Device kernels of the acoustic forward operator for gfx950 (MI355X).
//
// Reference semantics: simwave/kernel/backend/c_code/forward/
//   {constant,variable}_density/{2d,3d}/wave.c, sections 1-4 of the time loop.
//
// Floating-point contract: every kernel reproduces the reference C's types and
// evaluation order (centre term first, pairs added in radius order, division by
// h^2, the double-promoted `2.0/den * u` update), and the library is compiled
// with -ffp-contract=off, so results are bit-identical to the reference C on
// x86 (IEEE binary32/64, round-to-nearest, denormals kept).
//
// Per-point model coefficients are precomputed once per run by k_prepare:
//   q = dt^2 / s              (the reference's `dtSquared / slowness`)
//   e = damp*dt / (2 s)       (the reference's damping term)
// with s = 1/(v*v) rounded to the working type exactly as the C does.  Where
// e == 0 (the whole interior outside the damping layer) the reference's
// den/num are exactly 1 and the update reduces to 2u - u_prev + q*L with the
// same roundings, which the kernels exploit without changing a bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#if defined(SW_DOUBLE)
typedef double real_t;
#else
typedef float real_t;
#endif

#ifndef SW_DIM
#error "SW_DIM must be 2 or 3"
#endif
#ifndef SW_VD
#define SW_VD 0
#endif

namespace sw {

constexpr int kMaxR = 10;

// Inclusive index box of one source or receiver plus the offset of its
// per-axis weights (z weights, then x, then y) in the weight array.
struct Box {
    int lo[3];
    int hi[3];
    int woff;
    int pad;
};

// Everything a time step needs that does not change between steps.
struct Params {
    int nz, nx, ny;          // extended shape; ny == 1 in 2D
    int r;                   // stencil radius
    long long sz, sx;        // element strides of z and the middle axis (3D)
    int bc[6];               // z_b z_a x_b x_a y_b y_a : 0 none, 1 Dirichlet, 2 Neumann
    real_t dz2, dx2, dy2;    // squared spacings (2D: dx2 is the contiguous axis)
    real_t c2[kMaxR + 1];    // second-derivative half coefficients
    real_t c1[kMaxR + 1];    // first-derivative half coefficients (VD)
    const real_t *q;         // dt^2/s per point
    const real_t *e;         // damping term per point
    const real_t *rho;       // density (VD)
    // sources
    int nsrc;
    int wav_multi;           // wavelet_count > 1
    const Box *sbox;
    const real_t *sw;        // source weights
    const real_t *wavelet;
    // graph replay: when set, the step number a kernel receives is relative
    // and *nbase is added to it on the device (sources, receiver rows)
    const int *nbase;
    // variable density, radius >= kVdPre, z-marching kernels: the density's
    // first-derivative sums along z, the middle axis and the contiguous axis
    // (k_rho_diff), which do not change between steps
    const real_t *drho[3];
};

// Radius from which the z-marching variable-density kernels read the
// density's derivative sums from arrays made once per run instead of
// forming them from 2R density neighbours per axis at every step.
constexpr int kVdPre = 6;

// ---------------------------------------------------------------- helpers

// Leap-frog update with damping; `value` is the spatial operator result
// (the reference's `value` before `value *= (dt2/s)/den`).
// Reference: cd-3d wave.c:177-185.
__device__ __forceinline__ real_t leapfrog(real_t value, real_t q, real_t e,
                                           real_t uc, real_t up)
{
    if (e == (real_t)0) {
        // den == num == 1: (dt2/s)/den == q, 2.0/den == 2.0, (num/den)*up == up
        value = value * q;
        double t = 2.0 * (double)uc - (double)up;
        t = t + (double)value;
        return (real_t)t;
    }
    const real_t den = (real_t)(1.0 + (double)e);
    const real_t num = (real_t)(1.0 - (double)e);
    value = value * (q / den);
    const double a = 2.0 / (double)den;
    const real_t b = num / den;
    double t = a * (double)uc - (double)(b * up);
    t = t + (double)value;
    return (real_t)t;
}

// Correctly rounded a / b for a fixed divisor b (Markstein's theorem):
// with y = RN(1/b), q0 = RN(a*y), r = a - b*q0 is exact and RN(q0 + r*y) is
// RN(a/b), provided nothing under- or overflows.  `lo`/`hi` bound |a| so that
// this holds (set on the host per divisor); zero is also safe (a -0 result
// may come out +0, which no later operation distinguishes).
struct FixedDiv {
    real_t b, y, lo, hi;
};

__device__ __forceinline__ bool fixed_div_ok(real_t a, const FixedDiv &d)
{
    const real_t m = a < (real_t)0 ? -a : a;
    return (m >= d.lo && m <= d.hi) || a == (real_t)0;
}

__device__ __forceinline__ real_t fixed_div(real_t a, const FixedDiv &d)
{
    const real_t q0 = a * d.y;
#if defined(SW_DOUBLE)
    const real_t r = __builtin_fma(-q0, d.b, a);
    return __builtin_fma(r, d.y, q0);
#else
    const real_t r = __builtin_fmaf(-q0, d.b, a);
    return __builtin_fmaf(r, d.y, q0);
#endif
}

// The divisors of the Laplacian's three axis sums (dy2, dx2, dz2) for
// fixed_div, and the numerator range safe for all three; ok == 0 keeps the
// IEEE quotients.
struct LapDiv {
    FixedDiv dy, dx, dz;
    real_t lo, hi;
    int ok;
};

// acc_y / dy2 + acc_x / dx2 + acc_z / dz2 (cd-3d wave.c:186), every quotient
// correctly rounded: fixed-divisor quotients when all active lanes of the
// wavefront have their three numerators in the safe range (or all zero),
// IEEE quotients otherwise.  Same bits either way.
__device__ __forceinline__ real_t lap_quotients(const Params &P, const LapDiv &D,
                                                real_t ay, real_t ax, real_t az)
{
    const real_t mx = fmax(fmax(fabs(ay), fabs(ax)), fabs(az));
    const real_t mn = fmin(fmin(fabs(ay), fabs(ax)), fabs(az));
    const bool ok = (D.ok != 0) & (((mn >= D.lo) & (mx <= D.hi)) | (mx == (real_t)0));
    if (__all(ok))
        return fixed_div(ay, D.dy) + fixed_div(ax, D.dx) + fixed_div(az, D.dz);
    return ay / P.dy2 + ax / P.dx2 + az / P.dz2;
}

// Source contribution at one point: dt2/s * kws * w / den.
// Reference: cd-3d wave.c:272-277.
__device__ __forceinline__ real_t source_term(real_t q, real_t e, real_t kws,
                                              real_t amp)
{
    const real_t den = (real_t)(1.0 + (double)e);
    return q * kws * amp / den;
}

__device__ __forceinline__ real_t wavelet_amp(const Params &P, int n, int s)
{
    if (P.nbase)
        n += *P.nbase;
    const long long idx = P.wav_multi ? (long long)(n - 1) * P.nsrc + s
                                      : (long long)(n - 1);
    return P.wavelet[idx];
}

// kws weight of box point (i, j[, k]) of box b.
__device__ __forceinline__ real_t box_weight(const Box &b, const real_t *w,
                                             int i, int j, int k)
{
    const int cz = b.hi[0] - b.lo[0] + 1;
    const real_t *wz = w + b.woff;
    const real_t *wx = wz + cz;
#if SW_DIM == 3
    const int cx = b.hi[1] - b.lo[1] + 1;
    const real_t *wy = wx + cx;
    return wz[i - b.lo[0]] * wx[j - b.lo[1]] * wy[k - b.lo[2]];
#else
    (void)k;
    return wz[i - b.lo[0]] * wx[j - b.lo[1]];
#endif
}

__device__ __forceinline__ bool box_has(const Box &b, int i, int j, int k)
{
#if SW_DIM == 3
    return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] &&
           k >= b.lo[2] && k <= b.hi[2];
#else
    (void)k;
    return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1];
#endif
}

// Adds, in source order, every source whose box holds interior point p to
// the freshly computed value v.  Reference: cd-3d wave.c:208-295 (the
// sequential C order of the `+=`).
__device__ __forceinline__ real_t add_sources(const Params &P, int n, int i,
                                              int j, int k, long long p,
                                              real_t v)
{
    for (int s = 0; s < P.nsrc; s++) {
        const Box b = P.sbox[s];
        if (!box_has(b, i, j, k))
            continue;
        const real_t amp = wavelet_amp(P, n, s);
        if (amp == (real_t)0)
            continue;
        const real_t kws = box_weight(b, P.sw, i, j, k);
        v += source_term(P.q[p], P.e[p], kws, amp);
    }
    return v;
}

// ------------------------------------------------------------ k_prepare
// q = dt2/s and e = damp*dt/(2s) with s = 1/(v*v).  Reference: cd-3d
// wave.c:177-181 (per-point terms that do not depend on the time step).
// Called in place (q aliases vel, e aliases damp): each point is read
// before it is written, by the same thread.
__global__ void k_prepare(const real_t *vel, const real_t *damp,
                          real_t *q, real_t *e,
                          long long npts, real_t dt, real_t dt2)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         p < npts; p += (long long)gridDim.x * blockDim.x) {
        const real_t v = vel[p];
        const real_t s = (real_t)(1.0 / (double)(v * v));
        q[p] = dt2 / s;
        e[p] = damp[p] * dt / (2 * s);
    }
}

#if SW_VD
// Density derivative sums of the variable-density stencil at every interior
// point: d[0] along z, d[1] along the middle axis (3D: x, stride sx), d[2]
// along the contiguous axis (3D: y; 2D: x), each
// sum_{m=1..R} c1[m] (rho[p + m s] - rho[p - m s]) in the reference's order
// (vd-3d wave.c:175-186, vd-2d wave.c:160-170), starting from zero, so the
// kernels that read them form the same bits as from the neighbours.
template <int R>
__global__ void k_rho_diff(Params P, real_t *__restrict__ dz, real_t *__restrict__ dm,
                           real_t *__restrict__ dc)
{
    const long long ny = SW_DIM == 3 ? P.ny : 1;
    const long long iy = SW_DIM == 3 ? P.ny - 2 * R : 1;
    const long long ix = P.nx - 2 * R, iz = P.nz - 2 * R;
    const long long cnt = iz * ix * iy;
    const real_t *__restrict__ rho = P.rho;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cnt;
         t += (long long)gridDim.x * blockDim.x) {
        long long rem = t;
        const long long k = SW_DIM == 3 ? R + rem % iy : 0;
        rem = SW_DIM == 3 ? rem / iy : rem;
        const long long j = R + rem % ix;
        const long long i = R + rem / ix;
        const long long p = i * P.sz + j * (SW_DIM == 3 ? P.sx : 1) + (SW_DIM == 3 ? k : 0);
        (void)ny;
        real_t az = 0, am = 0, ac = 0;
#pragma unroll
        for (int m = 1; m <= R; m++) {
            az += P.c1[m] * (rho[p + m * P.sz] - rho[p - m * P.sz]);
#if SW_DIM == 3
            am += P.c1[m] * (rho[p + m * P.sx] - rho[p - m * P.sx]);
            ac += P.c1[m] * (rho[p + m] - rho[p - m]);
#else
            ac += P.c1[m] * (rho[p + m] - rho[p - m]);
#endif
        }
        dz[p] = az;
        if (dm)
            dm[p] = am;
        dc[p] = ac;
    }
}
#endif

// C = (dt2/s)/den with den = (float)(1.0 + e): the reference's per-point
// factor `(dtSquared / slowness) / denominator` (cd-3d wave.c:184), which
// does not depend on the step.  C == q where e == 0.
__global__ void k_scaled_q(const real_t *__restrict__ q, const real_t *__restrict__ e,
                           real_t *__restrict__ C, long long npts)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         p < npts; p += (long long)gridDim.x * blockDim.x) {
        const real_t den = (real_t)(1.0 + (double)e[p]);
        C[p] = q[p] / den;
    }
}

// -------------------------------------------------- boundary predicates
// For grids whose interior spans at least r+1 points on every axis, the
// Neumann mirrors only read interior points and only write halo points, so
// the three boundary phases of one step can be applied pointwise by the
// thread that owns the mirrored interior point.  A value read by a later
// phase is zero iff an earlier phase's Dirichlet face zeroed it.
// Reference order: contiguous axis, middle axis (3D), z; each face line
// applies before-Dirichlet, before-Neumann, after-Dirichlet, after-Neumann.
// Reference: cd-3d wave.c:304-480, cd-2d wave.c:280-393.
//
// Writes the final value of interior point (i,j,k) and every halo point the
// Neumann faces copy from it.  v is the post-source value.
__device__ __forceinline__ void store_with_bc(const Params &P, real_t *un,
                                              long long p, int i, int j, int k,
                                              real_t v)
{
    const int r = P.r;
#if SW_DIM == 3
    const int cb = P.bc[4], ca = P.bc[5], cn = P.ny;  // contiguous axis: y
    const int mb = P.bc[2], ma = P.bc[3], mn = P.nx;  // middle axis: x
    const int kc = k, km = j;
    const long long sc = 1, sm = P.sx;
#else
    const int cb = P.bc[2], ca = P.bc[3], cn = P.nx;  // contiguous axis: x
    const int kc = j;
    const long long sc = 1;
#endif
    const int zb = P.bc[0], za = P.bc[1];
    const long long sz = P.sz;

    const bool zc = (cb == 1 && kc == r) || (ca == 1 && kc == cn - r - 1);
#if SW_DIM == 3
    const bool zm = (mb == 1 && km == r) || (ma == 1 && km == mn - r - 1);
#else
    const bool zm = false;
#endif
    const bool zz = (zb == 1 && i == r) || (za == 1 && i == P.nz - r - 1);
    const real_t zero = (real_t)0;
    un[p] = (zc || zm || zz) ? zero : v;

    // contiguous-axis faces (first phase)
    if (cb == 2) {
        const int m = kc - r;
        if (m >= 1 && m <= r)
            un[p - 2 * m * sc] = v;
    }
    if (ca == 2) {
        const int m = (cn - r - 1) - kc;
        if (m >= 1 && m <= r)
            un[p + 2 * m * sc] = (cb == 1 && kc == r) ? zero : v;
    }
#if SW_DIM == 3
    // middle-axis faces (second phase): reads see contiguous-axis zeroing
    if (mb == 2) {
        const int m = km - r;
        if (m >= 1 && m <= r)
            un[p - 2 * m * sm] = zc ? zero : v;
    }
    if (ma == 2) {
        const int m = (mn - r - 1) - km;
        if (m >= 1 && m <= r)
            un[p + 2 * m * sm] = (zc || (mb == 1 && km == r)) ? zero : v;
    }
#endif
    // z faces (last phase)
    if (zb == 2) {
        const int m = i - r;
        if (m >= 1 && m <= r)
            un[p - 2 * m * sz] = (zc || zm) ? zero : v;
    }
    if (za == 2) {
        const int m = (P.nz - r - 1) - i;
        if (m >= 1 && m <= r)
            un[p + 2 * m * sz] = (zc || zm || (zb == 1 && i == r)) ? zero : v;
    }
}

// ------------------------------------------------ fused step, 3D
// One workgroup = TY (=64, one wavefront row) points along the contiguous
// axis y x TX rows along x, marching along z over [z0, z0+zlen).  The
// current plane of u (and rho for VD) lives in LDS with an R-wide halo; the
// z neighbours live in a per-thread register queue of 2R+1 values.
#if SW_DIM == 3
template <int R, int TX>
__global__ void __launch_bounds__(64 * TX)
k_step3d(Params P, const real_t *__restrict__ uc, const real_t *up,
         real_t *un, int n, int zlen)
{
    constexpr int TY = 64;
    constexpr int W = TY + 2 * R;
    __shared__ real_t tile[TX + 2 * R][W];
#if SW_VD
    __shared__ real_t rtile[TX + 2 * R][W];
#endif
    const int tx = threadIdx.x;          // along y
    const int ty = threadIdx.y;          // along x
    const int k = R + blockIdx.x * TY + tx;
    const int j = R + blockIdx.y * TX + ty;
    const int z0 = R + blockIdx.z * zlen;
    const int z1 = min(z0 + zlen, P.nz - R);
    const int ny = P.ny, nx = P.nx;
    const long long sz = P.sz, sx = P.sx;
    const bool in_arr = (k < ny) && (j < nx);
    const bool active = (k < ny - R) && (j < nx - R);
    const long long col = (long long)j * sx + k;

    real_t qz[2 * R + 1];
#if SW_VD
    real_t rz[2 * R + 1];
#endif
#pragma unroll
    for (int m = 0; m < 2 * R; m++) {
        const long long o = (long long)(z0 - R + m) * sz + col;
        qz[m + 1] = in_arr ? uc[o] : (real_t)0;
#if SW_VD
        rz[m + 1] = in_arr ? P.rho[o] : (real_t)0;
#endif
    }

    for (int i = z0; i < z1; i++) {
#pragma unroll
        for (int m = 0; m < 2 * R; m++) {
            qz[m] = qz[m + 1];
#if SW_VD
            rz[m] = rz[m + 1];
#endif
        }
        {
            const long long o = (long long)(i + R) * sz + col;
            qz[2 * R] = in_arr ? uc[o] : (real_t)0;
#if SW_VD
            rz[2 * R] = in_arr ? P.rho[o] : (real_t)0;
#endif
        }
        const long long plane = (long long)i * sz;
        __syncthreads();
        tile[ty + R][tx + R] = qz[R];
#if SW_VD
        rtile[ty + R][tx + R] = rz[R];
#endif
        // x halo rows: R rows below and above the tile
        if (k < ny) {
            const int j0 = R + (int)blockIdx.y * TX;
            for (int h = ty; h < R; h += TX) {
                const int jl = j0 - R + h;          // >= 0
                const int jh = j0 + TX + h;
                tile[h][tx + R] = uc[plane + (long long)jl * sx + k];
#if SW_VD
                rtile[h][tx + R] = P.rho[plane + (long long)jl * sx + k];
#endif
                if (jh < nx) {
                    tile[TX + R + h][tx + R] = uc[plane + (long long)jh * sx + k];
#if SW_VD
                    rtile[TX + R + h][tx + R] = P.rho[plane + (long long)jh * sx + k];
#endif
                }
            }
        }
        // y halo columns
        if (tx < R && j < nx) {
            const int kl = k - R;                   // >= 0
            const int kh = blockIdx.x * TY + R + TY + tx;
            tile[ty + R][tx] = uc[plane + (long long)j * sx + kl];
#if SW_VD
            rtile[ty + R][tx] = P.rho[plane + (long long)j * sx + kl];
#endif
            if (kh < ny) {
                tile[ty + R][TY + R + tx] = uc[plane + (long long)j * sx + kh];
#if SW_VD
                rtile[ty + R][TY + R + tx] = P.rho[plane + (long long)j * sx + kh];
#endif
            }
        }
        __syncthreads();
        if (!active)
            continue;

        const long long p = plane + col;
        const real_t c = qz[R];
        real_t acc_c = P.c2[0] * c;
        real_t acc_m = P.c2[0] * c;
        real_t acc_z = P.c2[0] * c;
#if SW_VD
        real_t du_c = 0, dr_c = 0, du_m = 0, dr_m = 0, du_z = 0, dr_z = 0;
#endif
#pragma unroll
        for (int m = 1; m <= R; m++) {
            const real_t cp = tile[ty + R][tx + R + m], cm = tile[ty + R][tx + R - m];
            const real_t mp = tile[ty + R + m][tx + R], mm = tile[ty + R - m][tx + R];
            acc_c += P.c2[m] * (cp + cm);
            acc_m += P.c2[m] * (mp + mm);
            acc_z += P.c2[m] * (qz[R + m] + qz[R - m]);
#if SW_VD
            du_c += P.c1[m] * (cp - cm);
            dr_c += P.c1[m] * (rtile[ty + R][tx + R + m] - rtile[ty + R][tx + R - m]);
            du_m += P.c1[m] * (mp - mm);
            dr_m += P.c1[m] * (rtile[ty + R + m][tx + R] - rtile[ty + R - m][tx + R]);
            du_z += P.c1[m] * (qz[R + m] - qz[R - m]);
            dr_z += P.c1[m] * (rz[R + m] - rz[R - m]);
#endif
        }
        real_t value = 0;
        value += acc_c / P.dy2 + acc_m / P.dx2 + acc_z / P.dz2;
#if SW_VD
        {
            const real_t t_c = (du_c * dr_c) / (4 * P.dy2);
            const real_t t_m = (du_m * dr_m) / (4 * P.dx2);
            const real_t t_z = (du_z * dr_z) / (4 * P.dz2);
            value -= (t_c + t_m + t_z) / rz[R];
        }
#endif
        real_t v = leapfrog(value, P.q[p], P.e[p], c, up[p]);
        if (P.nsrc)
            v = add_sources(P, n, i, j, k, p, v);
        store_with_bc(P, un, p, i, j, k, v);
    }
}
#endif

#if SW_DIM == 3
// Fused 3D step, one point per thread: a thread owns a (j, k) column of a z
// chunk, keeps u's (and rho's) z neighbours in a register queue and loads
// every operand of plane i+1 (queue heads, x/y neighbours, u_prev, q,
// damping term) while plane i is computed; x/y neighbours come through the
// caches.  Same operations in the same order as k_step3d.  Plain 64-bit
// addressing, so it also serves grids whose planes are too large for the
// tiled kernel's 32-bit buffer offsets.  Its small per-thread state (72
// VGPRs at radius 4, constant density) keeps 7 wavefronts per SIMD, which
// hides latency better than the tiled kernel wherever the grid holds too
// few tiles to keep that kernel's 2 per SIMD busy.  Variable density: the
// fused path requires nx == ny (wave.c's x-axis first derivatives step by
// nx), so the x neighbours at m * sx are the reference's.
// rows (of 64 threads) per block; 4 and 16 rows measured: 16 slower
// everywhere, 4 faster on shallow grids (228 x 324 x 324 so8: 9%) and
// slower on deep so8 cubes (176^3-352^3: 1-2.5%)
constexpr int kZLines = 8;

// One block of the step: block (bx, by, bz) of the (y, x, z-chunk) grid.
template <int R>
__device__ __forceinline__ void step3d_z_block(const Params &P, const LapDiv &D,
                                               const real_t *__restrict__ uc,
                                               const real_t *__restrict__ up,
                                               real_t *__restrict__ un, int n, int zlen,
                                               int bx, int by, int bz)
{
    const int k = R + bx * 64 + threadIdx.x;
    const int j = R + by * kZLines + threadIdx.y;
    if (k >= P.ny - R || j >= P.nx - R)
        return;
    const int z0 = R + bz * zlen;
    const int z1 = min(z0 + zlen, P.nz - R);
    const long long sz = P.sz, sx = P.sx;
    const long long col = (long long)j * sx + k;
    real_t qz[2 * R + 1];
#if SW_VD
    // PRE: the density derivative sums come from P.drho (kVdPre)
    constexpr bool PRE = R >= kVdPre;
    const real_t *__restrict__ rho = P.rho;
    real_t rz[2 * R + 1];
#endif
#pragma unroll
    for (int m = 0; m < 2 * R; m++) {
        qz[m + 1] = uc[(long long)(z0 - R + m) * sz + col];
#if SW_VD
        if constexpr (!PRE)
            rz[m + 1] = rho[(long long)(z0 - R + m) * sz + col];
#endif
    }
    struct Ops {
        real_t head, up, q, e;
        real_t yn[2 * R], xn[2 * R];
#if SW_VD
        real_t rhead;
        real_t ryn[2 * R], rxn[2 * R];
        real_t r0, dz, dm, dc;
#endif
    };
    auto load = [&](int i, Ops &o) {
        const long long p = (long long)i * sz + col;
        o.head = uc[p + R * sz];
#pragma unroll
        for (int m = 1; m <= R; m++) {
            o.yn[2 * m - 2] = uc[p + m];
            o.yn[2 * m - 1] = uc[p - m];
            o.xn[2 * m - 2] = uc[p + m * sx];
            o.xn[2 * m - 1] = uc[p - m * sx];
        }
#if SW_VD
        if constexpr (PRE) {
            o.r0 = rho[p];
            o.dz = P.drho[0][p];
            o.dm = P.drho[1][p];
            o.dc = P.drho[2][p];
        } else {
            o.rhead = rho[p + R * sz];
#pragma unroll
            for (int m = 1; m <= R; m++) {
                o.ryn[2 * m - 2] = rho[p + m];
                o.ryn[2 * m - 1] = rho[p - m];
                o.rxn[2 * m - 2] = rho[p + m * sx];
                o.rxn[2 * m - 1] = rho[p - m * sx];
            }
        }
#endif
        o.up = up[p];
        o.q = P.q[p];
        o.e = P.e[p];
    };
    Ops cur, nxt;
    load(z0, cur);
    for (int i = z0; i < z1; i++) {
        if (i + 1 < z1)
            load(i + 1, nxt);
#pragma unroll
        for (int m = 0; m < 2 * R; m++) {
            qz[m] = qz[m + 1];
#if SW_VD
            if constexpr (!PRE)
                rz[m] = rz[m + 1];
#endif
        }
        qz[2 * R] = cur.head;
#if SW_VD
        if constexpr (!PRE)
            rz[2 * R] = cur.rhead;
#endif
        const long long p = (long long)i * sz + col;
        const real_t c = qz[R];
        real_t acc_c = P.c2[0] * c;
        real_t acc_m = P.c2[0] * c;
        real_t acc_z = P.c2[0] * c;
#if SW_VD
        real_t du_c = 0, dr_c = 0, du_m = 0, dr_m = 0, du_z = 0, dr_z = 0;
#endif
#pragma unroll
        for (int m = 1; m <= R; m++) {
            const real_t cp = cur.yn[2 * m - 2], cm = cur.yn[2 * m - 1];
            const real_t mp = cur.xn[2 * m - 2], mm = cur.xn[2 * m - 1];
            acc_c += P.c2[m] * (cp + cm);
            acc_m += P.c2[m] * (mp + mm);
            acc_z += P.c2[m] * (qz[R + m] + qz[R - m]);
#if SW_VD
            du_c += P.c1[m] * (cp - cm);
            du_m += P.c1[m] * (mp - mm);
            du_z += P.c1[m] * (qz[R + m] - qz[R - m]);
            if constexpr (!PRE) {
                dr_c += P.c1[m] * (cur.ryn[2 * m - 2] - cur.ryn[2 * m - 1]);
                dr_m += P.c1[m] * (cur.rxn[2 * m - 2] - cur.rxn[2 * m - 1]);
                dr_z += P.c1[m] * (rz[R + m] - rz[R - m]);
            }
#endif
        }
        real_t value = 0;
        value += lap_quotients(P, D, acc_c, acc_m, acc_z);
#if SW_VD
        {
            if constexpr (PRE) {
                dr_c = cur.dc;
                dr_m = cur.dm;
                dr_z = cur.dz;
            }
            const real_t t_c = (du_c * dr_c) / (4 * P.dy2);
            const real_t t_m = (du_m * dr_m) / (4 * P.dx2);
            const real_t t_z = (du_z * dr_z) / (4 * P.dz2);
            value -= (t_c + t_m + t_z) / (PRE ? cur.r0 : rz[R]);
        }
#endif
        real_t v = leapfrog(value, cur.q, cur.e, c, cur.up);
        if (P.nsrc)
            v = add_sources(P, n, i, j, k, p, v);
        store_with_bc(P, un, p, i, j, k, v);
        cur = nxt;
    }
}

template <int R>
__global__ void __launch_bounds__(64 * kZLines)
k_step3d_z(Params P, LapDiv D, const real_t *__restrict__ uc, const real_t *__restrict__ up,
           real_t *__restrict__ un, int n, int zlen)
{
    step3d_z_block<R>(P, D, uc, up, un, n, zlen, blockIdx.x, blockIdx.y, blockIdx.z);
}
#endif

// ------------------------------------------------ fused step, 2D
// One workgroup = TB points along the contiguous axis x, marching along z.
#if SW_DIM == 2
template <int R, int TB>
__global__ void __launch_bounds__(TB)
k_step2d(Params P, const real_t *__restrict__ uc, const real_t *up,
         real_t *un, int n, int zlen)
{
    constexpr int W = TB + 2 * R;
    __shared__ real_t row[W];
#if SW_VD
    __shared__ real_t rrow[W];
#endif
    const int tx = threadIdx.x;
    const int j = R + blockIdx.x * TB + tx;
    const int z0 = R + blockIdx.y * zlen;
    const int z1 = min(z0 + zlen, P.nz - R);
    const int nx = P.nx;
    const long long sz = P.sz;
    const bool in_arr = j < nx;
    const bool active = j < nx - R;

    real_t qz[2 * R + 1];
#if SW_VD
    real_t rz[2 * R + 1];
#endif
#pragma unroll
    for (int m = 0; m < 2 * R; m++) {
        const long long o = (long long)(z0 - R + m) * sz + j;
        qz[m + 1] = in_arr ? uc[o] : (real_t)0;
#if SW_VD
        rz[m + 1] = in_arr ? P.rho[o] : (real_t)0;
#endif
    }
    for (int i = z0; i < z1; i++) {
#pragma unroll
        for (int m = 0; m < 2 * R; m++) {
            qz[m] = qz[m + 1];
#if SW_VD
            rz[m] = rz[m + 1];
#endif
        }
        {
            const long long o = (long long)(i + R) * sz + j;
            qz[2 * R] = in_arr ? uc[o] : (real_t)0;
#if SW_VD
            rz[2 * R] = in_arr ? P.rho[o] : (real_t)0;
#endif
        }
        const long long plane = (long long)i * sz;
        __syncthreads();
        row[tx + R] = qz[R];
#if SW_VD
        rrow[tx + R] = rz[R];
#endif
        if (tx < R) {
            const int jl = j - R;
            const int jh = blockIdx.x * TB + R + TB + tx;
            row[tx] = uc[plane + jl];
#if SW_VD
            rrow[tx] = P.rho[plane + jl];
#endif
            if (jh < nx) {
                row[TB + R + tx] = uc[plane + jh];
#if SW_VD
                rrow[TB + R + tx] = P.rho[plane + jh];
#endif
            }
        }
        __syncthreads();
        if (!active)
            continue;
        const long long p = plane + j;
        const real_t c = qz[R];
        real_t acc_c = P.c2[0] * c;
        real_t acc_z = P.c2[0] * c;
#if SW_VD
        real_t du_c = 0, dr_c = 0, du_z = 0, dr_z = 0;
#endif
#pragma unroll
        for (int m = 1; m <= R; m++) {
            const real_t cp = row[tx + R + m], cm = row[tx + R - m];
            acc_c += P.c2[m] * (cp + cm);
            acc_z += P.c2[m] * (qz[R + m] + qz[R - m]);
#if SW_VD
            du_c += P.c1[m] * (cp - cm);
            dr_c += P.c1[m] * (rrow[tx + R + m] - rrow[tx + R - m]);
            du_z += P.c1[m] * (qz[R + m] - qz[R - m]);
            dr_z += P.c1[m] * (rz[R + m] - rz[R - m]);
#endif
        }
        real_t value = 0;
        value += acc_c / P.dx2 + acc_z / P.dz2;
#if SW_VD
        {
            const real_t t_c = (du_c * dr_c) / (4 * P.dx2);
            const real_t t_z = (du_z * dr_z) / (4 * P.dz2);
            value -= (t_c + t_z) / rz[R];
        }
#endif
        real_t v = leapfrog(value, P.q[p], P.e[p], c, up[p]);
        if (P.nsrc)
            v = add_sources(P, n, i, j, 0, p, v);
        store_with_bc(P, un, p, i, j, 0, v);
    }
}
#endif

#if SW_DIM == 2
// Fused 2D step, one point per thread, the 2D counterpart of k_step3d_z: a
// thread owns column j of a z chunk, keeps u's (and rho's) z neighbours in
// a register queue and loads every operand of row i+1 (queue heads, x
// neighbours, u_prev, q, damping term) while row i is computed; x
// neighbours come through the caches, so no LDS and no barrier.  Same
// operations in the same order as k_step2d.
constexpr int kZ2Block = 256;

// One block of the step: block (bx, bz) of the (x, z-chunk) grid.
template <int R>
__device__ __forceinline__ void step2d_z_block(const Params &P, const real_t *__restrict__ uc,
                                               const real_t *__restrict__ up,
                                               real_t *__restrict__ un, int n, int zlen,
                                               int bx, int bz)
{
    const int j = R + bx * kZ2Block + threadIdx.x;
    if (j >= P.nx - R)
        return;
    const int z0 = R + bz * zlen;
    const int z1 = min(z0 + zlen, P.nz - R);
    const long long sz = P.sz;
    real_t qz[2 * R + 1];
#if SW_VD
    // PRE: the density derivative sums come from P.drho (kVdPre)
    constexpr bool PRE = R >= kVdPre;
    const real_t *__restrict__ rho = P.rho;
    real_t rz[2 * R + 1];
#endif
#pragma unroll
    for (int m = 0; m < 2 * R; m++) {
        qz[m + 1] = uc[(long long)(z0 - R + m) * sz + j];
#if SW_VD
        if constexpr (!PRE)
            rz[m + 1] = rho[(long long)(z0 - R + m) * sz + j];
#endif
    }
    struct Ops {
        real_t head, up, q, e;
        real_t xn[2 * R];
#if SW_VD
        real_t rhead;
        real_t rxn[2 * R];
        real_t r0, dz, dc;
#endif
    };
    auto load = [&](int i, Ops &o) {
        const long long p = (long long)i * sz + j;
        o.head = uc[p + R * sz];
#pragma unroll
        for (int m = 1; m <= R; m++) {
            o.xn[2 * m - 2] = uc[p + m];
            o.xn[2 * m - 1] = uc[p - m];
        }
#if SW_VD
        if constexpr (PRE) {
            o.r0 = rho[p];
            o.dz = P.drho[0][p];
            o.dc = P.drho[2][p];
        } else {
            o.rhead = rho[p + R * sz];
#pragma unroll
            for (int m = 1; m <= R; m++) {
                o.rxn[2 * m - 2] = rho[p + m];
                o.rxn[2 * m - 1] = rho[p - m];
            }
        }
#endif
        o.up = up[p];
        o.q = P.q[p];
        o.e = P.e[p];
    };
    Ops cur, nxt;
    load(z0, cur);
    for (int i = z0; i < z1; i++) {
        if (i + 1 < z1)
            load(i + 1, nxt);
#pragma unroll
        for (int m = 0; m < 2 * R; m++) {
            qz[m] = qz[m + 1];
#if SW_VD
            if constexpr (!PRE)
                rz[m] = rz[m + 1];
#endif
        }
        qz[2 * R] = cur.head;
#if SW_VD
        if constexpr (!PRE)
            rz[2 * R] = cur.rhead;
#endif
        const long long p = (long long)i * sz + j;
        const real_t c = qz[R];
        real_t acc_c = P.c2[0] * c;
        real_t acc_z = P.c2[0] * c;
#if SW_VD
        real_t du_c = 0, dr_c = 0, du_z = 0, dr_z = 0;
#endif
#pragma unroll
        for (int m = 1; m <= R; m++) {
            const real_t cp = cur.xn[2 * m - 2], cm = cur.xn[2 * m - 1];
            acc_c += P.c2[m] * (cp + cm);
            acc_z += P.c2[m] * (qz[R + m] + qz[R - m]);
#if SW_VD
            du_c += P.c1[m] * (cp - cm);
            du_z += P.c1[m] * (qz[R + m] - qz[R - m]);
            if constexpr (!PRE) {
                dr_c += P.c1[m] * (cur.rxn[2 * m - 2] - cur.rxn[2 * m - 1]);
                dr_z += P.c1[m] * (rz[R + m] - rz[R - m]);
            }
#endif
        }
        real_t value = 0;
        value += acc_c / P.dx2 + acc_z / P.dz2;
#if SW_VD
        {
            if constexpr (PRE) {
                dr_c = cur.dc;
                dr_z = cur.dz;
            }
            const real_t t_c = (du_c * dr_c) / (4 * P.dx2);
            const real_t t_z = (du_z * dr_z) / (4 * P.dz2);
            value -= (t_c + t_z) / (PRE ? cur.r0 : rz[R]);
        }
#endif
        real_t v = leapfrog(value, cur.q, cur.e, c, cur.up);
        if (P.nsrc)
            v = add_sources(P, n, i, j, 0, p, v);
        store_with_bc(P, un, p, i, j, 0, v);
        cur = nxt;
    }
}

template <int R>
__global__ void __launch_bounds__(kZ2Block)
k_step2d_z(Params P, const real_t *__restrict__ uc, const real_t *__restrict__ up,
           real_t *__restrict__ un, int n, int zlen)
{
    step2d_z_block<R>(P, uc, up, un, n, zlen, blockIdx.x, blockIdx.y);
}
#endif

// ------------------------------------------------ generic (exact) path
// Used where the fused kernel's preconditions do not hold (interior thinner
// than r+1 points, too many sources, or the 3D variable-density x stride
// quirk with nx != ny).  Plain global loads, one thread per interior point.
// Reference: vd-3d wave.c:145-214 including the nx stride of the middle-axis
// first derivatives (:184-186).
template <int R>
__global__ void __launch_bounds__(256)
k_stencil_generic(Params P, const real_t *__restrict__ uc, const real_t *up,
                  real_t *un)
{
    const long long li = (long long)blockIdx.x * blockDim.x + threadIdx.x;
#if SW_DIM == 3
    const long long Iy = P.ny - 2 * R, Ix = P.nx - 2 * R, Iz = P.nz - 2 * R;
    if (li >= Iy * Ix * Iz)
        return;
    const int k = R + (int)(li % Iy);
    const int j = R + (int)((li / Iy) % Ix);
    const int i = R + (int)(li / (Iy * Ix));
    const long long p = (long long)i * P.sz + (long long)j * P.sx + k;
#else
    const long long Ix = P.nx - 2 * R, Iz = P.nz - 2 * R;
    if (li >= Ix * Iz)
        return;
    const int j = R + (int)(li % Ix);
    const int i = R + (int)(li / Ix);
    const long long p = (long long)i * P.sz + j;
#endif
    const real_t *c = uc + p;
    real_t acc_c = P.c2[0] * c[0];
#if SW_DIM == 3
    real_t acc_m = P.c2[0] * c[0];
#endif
    real_t acc_z = P.c2[0] * c[0];
#if SW_VD
    const real_t *d = P.rho + p;
    real_t du_c = 0, dr_c = 0, du_z = 0, dr_z = 0;
#if SW_DIM == 3
    real_t du_m = 0, dr_m = 0;
#endif
#endif
#pragma unroll
    for (int m = 1; m <= R; m++) {
        acc_c += P.c2[m] * (c[m] + c[-m]);
#if SW_VD
        du_c += P.c1[m] * (c[m] - c[-m]);
        dr_c += P.c1[m] * (d[m] - d[-m]);
#endif
#if SW_DIM == 3
        {
            const long long o = (long long)m * P.sx;
            acc_m += P.c2[m] * (c[o] + c[-o]);
#if SW_VD
            const long long oq = (long long)m * P.nx;   // reference quirk
            du_m += P.c1[m] * (c[oq] - c[-oq]);
            dr_m += P.c1[m] * (d[oq] - d[-oq]);
#endif
        }
#endif
        {
            const long long o = (long long)m * P.sz;
            acc_z += P.c2[m] * (c[o] + c[-o]);
#if SW_VD
            du_z += P.c1[m] * (c[o] - c[-o]);
            dr_z += P.c1[m] * (d[o] - d[-o]);
#endif
        }
    }
    real_t value = 0;
#if SW_DIM == 3
    value += acc_c / P.dy2 + acc_m / P.dx2 + acc_z / P.dz2;
#if SW_VD
    {
        const real_t t_c = (du_c * dr_c) / (4 * P.dy2);
        const real_t t_m = (du_m * dr_m) / (4 * P.dx2);
        const real_t t_z = (du_z * dr_z) / (4 * P.dz2);
        value -= (t_c + t_m + t_z) / d[0];
    }
#endif
#else
    value += acc_c / P.dx2 + acc_z / P.dz2;
#if SW_VD
    {
        const real_t t_c = (du_c * dr_c) / (4 * P.dx2);
        const real_t t_z = (du_z * dr_z) / (4 * P.dz2);
        value -= (t_c + t_z) / d[0];
    }
#endif
#endif
    un[p] = leapfrog(value, P.q[p], P.e[p], c[0], up[p]);
}

// Source injection in the reference's sequential order: one workgroup walks
// the sources in index order; the points of one box are distinct, so its
// threads update them in parallel, and a barrier separates sources.
// halo_only: skip interior points (the fused kernel adds those itself).
// Reference: cd-3d wave.c:208-295.
__global__ void __launch_bounds__(256)
k_inject(Params P, real_t *un, int n, int halo_only)
{
    const int r = P.r;
    for (int s = 0; s < P.nsrc; s++) {
        const real_t amp = wavelet_amp(P, n, s);
        if (amp != (real_t)0) {
            const Box b = P.sbox[s];
            const int cz = b.hi[0] - b.lo[0] + 1;
            const int cx = b.hi[1] - b.lo[1] + 1;
#if SW_DIM == 3
            const int cy = b.hi[2] - b.lo[2] + 1;
#else
            const int cy = 1;
#endif
            const int cnt = cz * cx * cy;
            for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
                const int kk = t % cy;
                const int jj = (t / cy) % cx;
                const int ii = t / (cy * cx);
                const int i = b.lo[0] + ii, j = b.lo[1] + jj;
#if SW_DIM == 3
                const int k = b.lo[2] + kk;
                const bool interior = i >= r && i < P.nz - r && j >= r &&
                                      j < P.nx - r && k >= r && k < P.ny - r;
                const long long p = (long long)i * P.sz + (long long)j * P.sx + k;
#else
                const int k = 0;
                const bool interior = i >= r && i < P.nz - r && j >= r && j < P.nx - r;
                const long long p = (long long)i * P.sz + j;
#endif
                if (halo_only && interior)
                    continue;
                const real_t kws = box_weight(b, P.sw, i, j, k);
                un[p] += source_term(P.q[p], P.e[p], kws, amp);
            }
        }
        __syncthreads();
    }
}

// One boundary phase: each thread owns one face line and applies the four
// reference operations in order.  axis: 0 z, 1 middle x (3D), 2 contiguous.
// Reference: cd-3d wave.c:324-480 (face_line semantics).
__global__ void __launch_bounds__(256)
k_bc_phase(Params P, real_t *un, int axis)
{
    const int r = P.r;
    const long long li = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long base, st;
    int n, before, after;
#if SW_DIM == 3
    const long long Iy = P.ny - 2 * r, Ix = P.nx - 2 * r, Iz = P.nz - 2 * r;
    if (axis == 2) {          // lines along y over interior (i, j)
        if (li >= Iz * Ix) return;
        const int j = r + (int)(li % Ix), i = r + (int)(li / Ix);
        base = (long long)i * P.sz + (long long)j * P.sx; st = 1; n = P.ny;
        before = P.bc[4]; after = P.bc[5];
    } else if (axis == 1) {   // lines along x over interior (i, k)
        if (li >= Iz * Iy) return;
        const int k = r + (int)(li % Iy), i = r + (int)(li / Iy);
        base = (long long)i * P.sz + k; st = P.sx; n = P.nx;
        before = P.bc[2]; after = P.bc[3];
    } else {                  // lines along z over interior (j, k)
        if (li >= Ix * Iy) return;
        const int k = r + (int)(li % Iy), j = r + (int)(li / Iy);
        base = (long long)j * P.sx + k; st = P.sz; n = P.nz;
        before = P.bc[0]; after = P.bc[1];
    }
#else
    const long long Ix = P.nx - 2 * r, Iz = P.nz - 2 * r;
    if (axis == 2) {          // lines along x over interior i
        if (li >= Iz) return;
        base = (long long)(r + li) * P.sz; st = 1; n = P.nx;
        before = P.bc[2]; after = P.bc[3];
    } else {                  // lines along z over interior j
        if (li >= Ix) return;
        base = r + li; st = P.sz; n = P.nz;
        before = P.bc[0]; after = P.bc[1];
    }
#endif
    const long long lo = base + (long long)r * st;
    const long long hi = base + (long long)(n - r - 1) * st;
    if (before == 1)
        un[lo] = 0;
    if (before == 2)
        for (int m = 1; m <= r; m++)
            un[lo - m * st] = un[lo + m * st];
    if (after == 1)
        un[hi] = 0;
    if (after == 2)
        for (int m = 1; m <= r; m++)
            un[hi + m * st] = un[hi - m * st];
}

// Receivers: one thread per receiver, sequential sum over its box in the
// reference order (z, then x, then y).  Reads level n (the step's input).
// Reference: cd-3d wave.c:499-566.
__global__ void __launch_bounds__(256)
k_receivers(Params P, const real_t *__restrict__ uc, const Box *__restrict__ rbox,
            const real_t *__restrict__ rw, int nrec, real_t *__restrict__ out)
{
    if (P.nbase)
        out += (long long)(*P.nbase) * nrec;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nrec)
        return;
    const Box b = rbox[q];
    const int cz = b.hi[0] - b.lo[0] + 1;
    const real_t *wz = rw + b.woff;
    const real_t *wx = wz + cz;
#if SW_DIM == 3
    const int cx = b.hi[1] - b.lo[1] + 1;
    const real_t *wy = wx + cx;
#endif
    real_t acc = 0;
    for (int i = b.lo[0]; i <= b.hi[0]; i++) {
        const real_t fz = wz[i - b.lo[0]];
        for (int j = b.lo[1]; j <= b.hi[1]; j++) {
            const real_t fzx = fz * wx[j - b.lo[1]];
#if SW_DIM == 3
            const real_t *row = uc + (long long)i * P.sz + (long long)j * P.sx;
            for (int k = b.lo[2]; k <= b.hi[2]; k++)
                acc += row[k] * (fzx * wy[k - b.lo[2]]);
#else
            acc += uc[(long long)i * P.sz + j] * fzx;
#endif
        }
    }
    out[q] = acc;
}

// Receivers, cooperative form: a block serves RPB receivers.  All 256
// threads gather box values and multiply them by their weights into LDS (CH
// points per receiver per round), then lane q of the first wave adds its
// receiver's products in box order.  The 256/RPB threads of one receiver
// are consecutive, so a wavefront's loads walk along box rows (contiguous
// y) of a few receivers; the LDS rows are padded by one word against bank
// conflicts.  Every product and every partial sum is
// rounded exactly as in the reference's sequential loop (wave.c:533-565); only
// the loads and multiplies run in parallel.
// Receivers bx*RPB .. bx*RPB+RPB-1 of one shot.
template <int RPB, int CH>
__device__ __forceinline__ void receivers_coop_block(const Params &P,
                                                     const real_t *__restrict__ uc,
                                                     const Box *__restrict__ rbox,
                                                     const real_t *__restrict__ rw, int nrec,
                                                     real_t *__restrict__ out, int bx)
{
    __shared__ real_t prod[CH][RPB + 1];
    __shared__ Box boxes[RPB];
    __shared__ int counts[RPB];
    __shared__ int maxcount;
    const int tid = threadIdx.x;
    const int r0 = bx * RPB;
    if (tid == 0)
        maxcount = 0;
    __syncthreads();
    if (tid < RPB) {
        const int q = r0 + tid;
        Box b{};
        int cnt = 0;
        if (q < nrec) {
            b = rbox[q];
            cnt = (b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1);
#if SW_DIM == 3
            cnt *= b.hi[2] - b.lo[2] + 1;
#endif
        }
        boxes[tid] = b;
        counts[tid] = cnt;
        atomicMax(&maxcount, cnt);
    }
    __syncthreads();
    const int total_max = maxcount;
    constexpr int STEP = 256 / RPB;       // points advanced per thread per pass
    const int l = tid / STEP;             // receiver served by this thread
    const Box b = boxes[l];
    const int cnt = counts[l];
    const int cz = b.hi[0] - b.lo[0] + 1, cx = b.hi[1] - b.lo[1] + 1;
#if SW_DIM == 3
    const int cy = b.hi[2] - b.lo[2] + 1;
#else
    const int cy = 1;
#endif
    const real_t *wz = rw + b.woff;
    const real_t *wx = wz + cz;
#if SW_DIM == 3
    const real_t *wy = wx + cx;
#endif
    real_t acc = 0;
    for (int base = 0; base < total_max; base += CH) {
        // this thread's first point of the round, decoded once, then walked
        int t = base + tid % STEP;
        int ii = 0, jj = 0, kk = 0;
        if (t < cnt) {
            ii = t / (cx * cy);
            const int rem = t - ii * cx * cy;
            jj = rem / cy;
            kk = rem - jj * cy;
        }
        for (int c = tid % STEP; c < CH && t < cnt; c += STEP, t += STEP) {
#if SW_DIM == 3
            const real_t kws = wz[ii] * wx[jj] * wy[kk];
            const long long p = (long long)(b.lo[0] + ii) * P.sz +
                                (long long)(b.lo[1] + jj) * P.sx + (b.lo[2] + kk);
#else
            const real_t kws = wz[ii] * wx[jj];
            const long long p = (long long)(b.lo[0] + ii) * P.sz + (b.lo[1] + jj);
#endif
            prod[c][l] = uc[p] * kws;
            kk += STEP;
            while (kk >= cy) {
                kk -= cy;
                if (++jj >= cx) {
                    jj = 0;
                    ii++;
                }
            }
        }
        __syncthreads();
        if (tid < RPB) {
            // lane tid sums receiver tid (its loads above served receiver l)
            const int nround = min(CH, counts[tid] - base);
            for (int c = 0; c < nround; c++)
                acc += prod[c][tid];
        }
        __syncthreads();
    }
    if (tid < RPB && r0 + tid < nrec)
        out[r0 + tid] = acc;
}

template <int RPB, int CH>
__global__ void __launch_bounds__(256)
k_receivers_coop(Params P, const real_t *__restrict__ uc, const Box *__restrict__ rbox,
                 const real_t *__restrict__ rw, int nrec, real_t *__restrict__ out)
{
    if (P.nbase)
        out += (long long)(*P.nbase) * nrec;
    receivers_coop_block<RPB, CH>(P, uc, rbox, rw, nrec, out, blockIdx.x);
}

#if SW_DIM == 2
// One 2D step and its receiver gather in one launch: blocks 0 .. nrb-1
// gather the receivers of level n (they only read u^n, which the step does
// not write), the rest are k_step2d_z's blocks (gx along x).  Saves the
// second launch per step of small 2D grids, whose steps take a few
// microseconds.  The gather's chunk CH only sets how many products a round
// stages in LDS (small here, so the step blocks keep their occupancy); the
// sum per receiver runs in the same order for any CH.
template <int R, int RPB, int CH>
__global__ void __launch_bounds__(kZ2Block)
k_step2d_z_rec(Params P, const real_t *__restrict__ uc, const real_t *__restrict__ up,
               real_t *__restrict__ un, int n, int zlen, int gx, int nrb,
               const Box *__restrict__ rbox, const real_t *__restrict__ rw, int nrec,
               real_t *__restrict__ out)
{
    static_assert(kZ2Block == 256, "the gather's block shape");
    const int b = blockIdx.x;
    if (b < nrb) {
        if (P.nbase)
            out += (long long)(*P.nbase) * nrec;
        receivers_coop_block<RPB, CH>(P, uc, rbox, rw, nrec, out, b);
        return;
    }
    const int s = b - nrb;
    step2d_z_block<R>(P, uc, up, un, n, zlen, s % gx, s / gx);
}

// Grid-wide barrier of a cooperative launch (every block resident).  Each
// thread makes its stores visible device-wide (agent-scope fence: the
// stores complete and the XCD's L2 writes them back), the block's first
// thread adds one to *count and waits until it reaches `target`, and every
// thread then fences again (its later loads see the other XCDs' stores).
// The wait is bounded: after `spin_limit` polls, or once another block has
// given up, a block stores `code` in *abort (unless a code is there); every
// block returns false then and leaves the kernel, so no launch can hang on
// a missing arrival.
__device__ __forceinline__ bool grid_sync(unsigned *count, unsigned target, int *abort,
                                          int spin_limit, int code)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(count, 1u);
        int spins = 0;
        while (__hip_atomic_load(count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                ++spins > spin_limit) {
                int none = 0;
                __hip_atomic_compare_exchange_strong(abort, &none, code, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    __threadfence();
    return __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
}

// Steps n0 .. n1 of a stride-0 2D shot in one cooperative launch: per step
// the blocks take the receiver gather of level n (nrb blocks) and the
// k_step2d_z blocks (nsb, gx along x) in turn, then meet at a grid barrier
// before the next step reads what this one wrote.  A roll call opens the
// launch: if not every block is resident within `roll_spin` polls, all
// blocks leave before writing anything (*abort = 1) and the caller runs the
// steps launch by launch instead; a barrier given up later leaves
// *abort = 2.  Slots follow the
// stride-0 cycle of the reference's time loop (wave.c:112-124): step n
// reads levels n-1 and n (slots (n-1) mod 3, n mod 3) and writes level n+1
// (slot (n+1) mod 3); its receivers go to row n - n0 of rec0.  Same block
// bodies as the per-step launches, so the same bits.  For small grids,
// whose steps take a few microseconds, where launching every step costs
// more than the step.
template <int R, int RPB, int CH>
__global__ void __launch_bounds__(kZ2Block)
k_step2d_persist(Params P, real_t *__restrict__ u, long long np, int n0, int n1, int zlen,
                 int gx, int nsb, int nrb, const Box *__restrict__ rbox,
                 const real_t *__restrict__ rw, int nrec, real_t *__restrict__ rec0,
                 unsigned *count, int *abort, int spin_limit, int roll_spin, int roll_extra)
{
    static_assert(kZ2Block == 256, "the gather's block shape");
    // roll_extra > 0 (tests of the fallback) asks for arrivals that never come
    if (!grid_sync(count, gridDim.x + roll_extra, abort, roll_spin, 1))
        return;
    for (int n = n0; n <= n1; n++) {
        const real_t *uc = u + (long long)(n % 3) * np;
        const real_t *up = u + (long long)((n - 1) % 3) * np;
        real_t *un = u + (long long)((n + 1) % 3) * np;
        real_t *row = rec0 + (long long)(n - n0) * nrec;
        for (int b = blockIdx.x; b < nrb + nsb; b += gridDim.x) {
            if (b < nrb) {
                receivers_coop_block<RPB, CH>(P, uc, rbox, rw, nrec, row, b);
            } else {
                const int s = b - nrb;
                step2d_z_block<R>(P, uc, up, un, n, zlen, s % gx, s / gx);
            }
        }
        if (!grid_sync(count, (unsigned)(n - n0 + 2) * gridDim.x + roll_extra, abort,
                       spin_limit, 2))
            return;
    }
}
#endif

// Zero-fills a word array (a kernel rather than a memset, so it behaves the
// same inside captured graphs).
// The step base of a graph replay (a kernel, so that no memset node is
// needed around graphs).
__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *p = v;
}

// out[i][j][k] = small[i0[i]][i1[j]][i2[k]] (2D: no k, i2 null, m2 = 1),
// out contiguous n0 x n1 x n2 and small m0 x m1 x m2: the edge-extended
// models and the damping mask built on the device from their distinct
// values (copies only, so the bits are the host's).
__global__ void k_gather_axes(real_t *__restrict__ out, const real_t *__restrict__ small,
                              long long n0, long long n1, long long n2, long long m1,
                              long long m2, const long long *__restrict__ i0,
                              const long long *__restrict__ i1,
                              const long long *__restrict__ i2)
{
    const long long total = n0 * n1 * n2;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long k = t % n2, r = t / n2, j = r % n1, i = r / n1;
        const long long c = i2 ? i2[k] : 0;
        out[t] = small[(i0[i] * m1 + i1[j]) * m2 + c];
    }
}

__global__ void k_zero_u32(unsigned *p, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// Exchange two slots (the reference's even-stride swap, wave.c:578-611).
// dst[idx[i]] = val[i]: places the halo frame of a slot uploaded packed.
__global__ void k_scatter(real_t *__restrict__ dst, const unsigned long long *__restrict__ idx,
                          const real_t *__restrict__ val, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[idx[i]] = val[i];
}

__global__ void k_swap(real_t *a, real_t *b, long long npts)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         p < npts; p += (long long)gridDim.x * blockDim.x) {
        const real_t t = a[p];
        a[p] = b[p];
        b[p] = t;
    }
}

}  // namespace sw
