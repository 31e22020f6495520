// Adjoint (reverse-time) kernels of the constant-density forward operator.
//
// The forward step n maps (u^{n-1}, u^n) to u^{n+1} (cd-3d wave.c:142-480):
//   interior p:  u^{n+1}_p = a_p u^n_p - b_p u^{n-1}_p + c_p (L u^n)_p
//   sources:     u^{n+1}_p += c_p kws_{s,p} w_s^{n-1}
//   boundaries:  per face line, Dirichlet zeroes and Neumann copies
// with den = 1 + e, a = 2/den, b = (1 - e)/den, c = q/den, q = dt^2 v^2,
// e = damp dt v^2 / 2, and receivers d^n = K u^n (wave.c:499-566).  These
// kernels apply the transpose of that linear map, step by step backwards,
// and accumulate the derivative of the data with respect to the velocity:
//   da/dv = db/dv = -4 e / (v den^2),  dc/dv = 2 q / (v den^2).
// L is the stencil sum over the axes of c2-weighted pairs divided by h^2;
// it is symmetric, so its transpose is the same stencil evaluated on a
// field that is zero outside the interior.
#pragma once
#include "sw_kernels.hpp"

namespace sw {

// Stencil of `f` at point (i, j, k), treating points outside the grid as
// zero (the transpose of L applied to a field supported on the interior).
__device__ __forceinline__ real_t stencil_clipped(const Params &P, const real_t *f, int i,
                                                  int j, int k)
{
    const long long p = (long long)i * P.sz + (long long)j * P.sx + k;
    const int r = P.r;
    const real_t c0 = f[p] * P.c2[0];
    real_t sz = c0, sx = c0;
#if SW_DIM == 3
    real_t sy = c0;
#endif
    for (int m = 1; m <= r; m++) {
        const real_t cm = P.c2[m];
        real_t t = 0;
        if (i + m < P.nz) t += f[p + m * P.sz];
        if (i - m >= 0) t += f[p - m * P.sz];
        sz += cm * t;
        t = 0;
        if (j + m < P.nx) t += f[p + m * P.sx];
        if (j - m >= 0) t += f[p - m * P.sx];
        sx += cm * t;
#if SW_DIM == 3
        t = 0;
        if (k + m < P.ny) t += f[p + m];
        if (k - m >= 0) t += f[p - m];
        sy += cm * t;
#endif
    }
#if SW_DIM == 3
    return sy / P.dy2 + sx / P.dx2 + sz / P.dz2;
#else
    return sx / P.dx2 + sz / P.dz2;
#endif
}

// Point of this thread in a whole-grid launch (adj_grid): 64 threads along
// the contiguous axis, 4 lines per block.
constexpr int kAdjLines = 4;

__device__ __forceinline__ bool thread_point(const Params &P, int &i, int &j, int &k)
{
#if SW_DIM == 3
    i = blockIdx.z;
    j = blockIdx.y * kAdjLines + threadIdx.y;
    k = blockIdx.x * 64 + threadIdx.x;
    return j < P.nx && k < P.ny;
#else
    i = blockIdx.y * kAdjLines + threadIdx.y;
    j = blockIdx.x * 64 + threadIdx.x;
    k = 0;
    return i < P.nz && j < P.nx;
#endif
}

inline dim3 adj_grid(const Params &P)
{
#if SW_DIM == 3
    return dim3((unsigned)((P.ny + 63) / 64), (unsigned)((P.nx + kAdjLines - 1) / kAdjLines),
                (unsigned)P.nz);
#else
    return dim3((unsigned)((P.nx + 63) / 64), (unsigned)((P.nz + kAdjLines - 1) / kAdjLines), 1u);
#endif
}

inline dim3 adj_block() { return dim3(64u, (unsigned)kAdjLines, 1u); }

__device__ __forceinline__ bool interior(const Params &P, int i, int j, int k)
{
    const int r = P.r;
    bool in = i >= r && i < P.nz - r && j >= r && j < P.nx - r;
#if SW_DIM == 3
    in = in && k >= r && k < P.ny - r;
#else
    (void)k;
#endif
    return in;
}

// ubar += K^T res: each receiver spreads its residual over its box.
__global__ void __launch_bounds__(256)
k_adj_receivers(Params P, const Box *__restrict__ rbox, const real_t *__restrict__ rw,
                int nrec, const real_t *__restrict__ res, real_t *ubar)
{
    const int q = blockIdx.x;
    if (q >= nrec)
        return;
    const Box b = rbox[q];
    const real_t rv = res[q];
    if (rv == (real_t)0)
        return;
    const int cz = b.hi[0] - b.lo[0] + 1, cx = b.hi[1] - b.lo[1] + 1;
#if SW_DIM == 3
    const int cy = b.hi[2] - b.lo[2] + 1;
#else
    const int cy = 1;
#endif
    for (int t = threadIdx.x; t < cz * cx * cy; t += blockDim.x) {
        const int ii = t / (cx * cy), rem = t - ii * cx * cy;
        const int jj = rem / cy, kk = rem - jj * cy;
        const int i = b.lo[0] + ii, j = b.lo[1] + jj;
#if SW_DIM == 3
        const int k = b.lo[2] + kk;
#else
        const int k = 0;
        (void)kk;
#endif
        const long long p = (long long)i * P.sz + (long long)j * P.sx + k;
        atomicAdd(&ubar[p], box_weight(b, rw, i, j, k) * rv);
    }
}

// grad += gacc and gacc = 0 where gacc holds anything: the float64
// gradient takes the working-precision sums of the last few reverse steps
// (the update kernels add each step's term to gacc, which halves their
// gradient traffic in float32).
__global__ void __launch_bounds__(256)
k_grad_flush(real_t *__restrict__ gacc, double *__restrict__ grad, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const real_t g = gacc[i];
        if (g != (real_t)0) {
            grad[i] += (double)g;
            gacc[i] = (real_t)0;
        }
    }
}

// The same spread as a gather, with no atomics and a fixed order.  Built
// once per gradient: every (point, receiver) pair of the receiver boxes as
// an entry, sorted by point (a stable sort of entries made in receiver
// order), so each touched point sums its receivers' terms in ascending
// receiver order, as the receiver loop of the transposed step does.

// Entry off[q] + t is point t of receiver q's box: its linear point index
// (sort key) and its own position (sort value).
__global__ void __launch_bounds__(256)
k_adj_rec_entries(Params P, const Box *__restrict__ rbox, int nrec,
                  const unsigned *__restrict__ off, unsigned *__restrict__ key,
                  unsigned *__restrict__ idx)
{
    const int q = blockIdx.x;
    if (q >= nrec)
        return;
    const Box b = rbox[q];
    const int cz = b.hi[0] - b.lo[0] + 1, cx = b.hi[1] - b.lo[1] + 1;
#if SW_DIM == 3
    const int cy = b.hi[2] - b.lo[2] + 1;
#else
    const int cy = 1;
#endif
    const unsigned base = off[q];
    for (int t = threadIdx.x; t < cz * cx * cy; t += blockDim.x) {
        const int ii = t / (cx * cy), rem = t - ii * cx * cy;
        const int jj = rem / cy, kk = rem - jj * cy;
#if SW_DIM == 3
        const long long p = (long long)(b.lo[0] + ii) * P.sz +
                            (long long)(b.lo[1] + jj) * P.sx + (b.lo[2] + kk);
#else
        (void)kk;
        const long long p = (long long)(b.lo[0] + ii) * P.sz + (b.lo[1] + jj);
#endif
        key[base + t] = (unsigned)p;
        idx[base + t] = base + t;
    }
}

// Sorted entry s (point key[s], original entry idx[s]): its receiver (the q
// with off[q] <= idx[s] < off[q+1]) and weight.
__global__ void __launch_bounds__(256)
k_adj_rec_permute(Params P, const Box *__restrict__ rbox, const real_t *__restrict__ rw,
                  int nrec, const unsigned *__restrict__ off,
                  const unsigned *__restrict__ key, const unsigned *__restrict__ idx,
                  unsigned n, unsigned *__restrict__ qs, real_t *__restrict__ ws)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    const unsigned e = idx[s];
    int lo = 0, hi = nrec;                 // off[lo] <= e < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    const long long p = key[s];
#if SW_DIM == 3
    const int i = (int)(p / P.sz);
    const long long rem = p - (long long)i * P.sz;
    const int j = (int)(rem / P.sx), k = (int)(rem - (long long)j * P.sx);
#else
    const int i = (int)(p / P.sz), j = (int)(p - (long long)i * P.sz), k = 0;
#endif
    qs[s] = (unsigned)lo;
    ws[s] = box_weight(rbox[lo], rw, i, j, k);
}

// Per reverse step: each touched point adds its receivers' terms in
// receiver order (receivers whose residual is zero add nothing, as in
// k_adj_receivers).
__global__ void __launch_bounds__(256)
k_adj_rec_gather(const unsigned *__restrict__ pts, const unsigned *__restrict__ start,
                 const unsigned *__restrict__ cnt, int nu, const unsigned *__restrict__ qs,
                 const real_t *__restrict__ ws, const real_t *__restrict__ res,
                 real_t *__restrict__ ubar)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu)
        return;
    const unsigned p = pts[u];
    real_t acc = ubar[p];
    unsigned s = start[u];
    const unsigned s1 = s + cnt[u];
    // four entries' loads in flight, then their terms in order
    for (; s + 4 <= s1; s += 4) {
        real_t w[4], rv[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            w[t] = ws[s + t];
            rv[t] = res[qs[s + t]];
        }
#pragma unroll
        for (int t = 0; t < 4; t++)
            if (rv[t] != (real_t)0)
                acc += w[t] * rv[t];
    }
    for (; s < s1; s++) {
        const real_t rv = res[qs[s]];
        if (rv != (real_t)0)
            acc += ws[s] * rv;
    }
    ubar[p] = acc;
}

// Transpose of one boundary phase (k_bc_phase): per face line, the after
// face's Neumann copies, then its Dirichlet zero, then the same for the
// before face.
__global__ void k_adj_bc_phase(Params P, real_t *ub, int axis)
{
    const int r = P.r;
    const long long li = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long base, st;
    int n, before, after;
#if SW_DIM == 3
    const long long Iy = P.ny - 2 * r, Ix = P.nx - 2 * r, Iz = P.nz - 2 * r;
    if (axis == 2) {
        if (li >= Iz * Ix) return;
        const int j = r + (int)(li % Ix), i = r + (int)(li / Ix);
        base = (long long)i * P.sz + (long long)j * P.sx; st = 1; n = P.ny;
        before = P.bc[4]; after = P.bc[5];
    } else if (axis == 1) {
        if (li >= Iz * Iy) return;
        const int k = r + (int)(li % Iy), i = r + (int)(li / Iy);
        base = (long long)i * P.sz + k; st = P.sx; n = P.nx;
        before = P.bc[2]; after = P.bc[3];
    } else {
        if (li >= Ix * Iy) return;
        const int k = r + (int)(li % Iy), j = r + (int)(li / Iy);
        base = (long long)j * P.sx + k; st = P.sz; n = P.nz;
        before = P.bc[0]; after = P.bc[1];
    }
#else
    const long long Ix = P.nx - 2 * r, Iz = P.nz - 2 * r;
    if (axis == 2) {
        if (li >= Iz) return;
        base = (long long)(r + li) * P.sz; st = 1; n = P.nx;
        before = P.bc[2]; after = P.bc[3];
    } else {
        if (li >= Ix) return;
        base = r + li; st = P.sz; n = P.nz;
        before = P.bc[0]; after = P.bc[1];
    }
#endif
    const long long lo = base + (long long)r * st;
    const long long hi = base + (long long)(n - r - 1) * st;
    if (after == 2)
        for (int m = 1; m <= r; m++) {
            ub[hi - m * st] += ub[hi + m * st];
            ub[hi + m * st] = 0;
        }
    if (after == 1)
        ub[hi] = 0;
    if (before == 2)
        for (int m = 1; m <= r; m++) {
            ub[lo + m * st] += ub[lo - m * st];
            ub[lo - m * st] = 0;
        }
    if (before == 1)
        ub[lo] = 0;
}

// W = c * ubar on the interior, 0 elsewhere, with c = C = q / den the
// per-point coefficient the forward step multiplies the Laplacian by
// (k_scaled_q; wave.c:177-185).
__global__ void __launch_bounds__(64 * kAdjLines)
k_adj_cw(Params P, const real_t *__restrict__ C, const real_t *__restrict__ ub,
         real_t *__restrict__ W)
{
    int i, j, k;
    if (thread_point(P, i, j, k)) {
        const long long p = (long long)i * P.sz + (long long)j * P.sx + k;
        W[p] = interior(P, i, j, k) ? C[p] * ub[p] : (real_t)0;
    }
}

// The transpose of the interior update of step n (k = n+1) and the
// velocity derivative at every point (w = ub, the adjoint of u^{n+1};
// u0, u1, u2 = u^{n+1}, u^n, u^{n-1}):
//   ub1 += a w + L^T W,  ub2 -= b w,
//   grad += w (da (u1 - u2) + dc (L u1 + s))
// where s is the source term of the step.  c (L u1 + s) is what the step
// added to a u1 - b u2, so (L u1 + s) = (u0 - a u1 + b u2) / c and
// dc / c = 2 / (v den): no second stencil is needed.  Points the boundary
// phases zeroed (Dirichlet lines) have w = 0.  In image mode grad collects
// the zero-lag cross-correlation w u0 instead.  ub is cleared on the way
// (its buffer becomes the adjoint of u^{n-2}).
__global__ void __launch_bounds__(64 * kAdjLines)
k_adj_update(Params P, const real_t *__restrict__ vel, real_t *__restrict__ ub,
             const real_t *__restrict__ W, real_t *__restrict__ ub1, real_t *__restrict__ ub2,
             const real_t *__restrict__ u0, const real_t *__restrict__ u1,
             const real_t *__restrict__ u2, real_t *__restrict__ gacc, int image)
{
    int i, j, k;
    if (thread_point(P, i, j, k)) {
        const long long p = (long long)i * P.sz + (long long)j * P.sx + k;
        const real_t lt = stencil_clipped(P, W, i, j, k);
        if (!interior(P, i, j, k)) {
            ub1[p] += lt;
            ub[p] = 0;
            return;
        }
        const real_t w = ub[p];
        ub[p] = 0;
        const double e = (double)P.e[p];
        const double rden = 1.0 / (1.0 + e);
        const double a = 2.0 * rden, b = (1.0 - e) * rden;
        ub1[p] = (real_t)((double)ub1[p] + a * (double)w + (double)lt);
        ub2[p] = (real_t)((double)ub2[p] - b * (double)w);
        if (w != (real_t)0) {
            const double x0 = (double)u0[p], x1 = (double)u1[p], x2 = (double)u2[p];
            if (image) {
                gacc[p] += (real_t)((double)w * x0);
            } else {
                const double rv = 1.0 / (double)vel[p];
                const double da = -4.0 * e * rden * rden * rv;
                const double step = x0 - a * x1 + b * x2;     // c (L u1 + s)
                gacc[p] += (real_t)((double)w * (da * (x1 - x2) + 2.0 * rden * rv * step));
            }
        }
    }
}

#if SW_DIM == 3
// The update of k_adj_update marching z: each thread owns a (j, k) column
// of a z chunk and keeps W's z neighbours in a register queue, so every W
// value is read from memory once for its z terms instead of 2R+1 times.
// The stencil and update take k_adj_update's operations in its order, with
// one difference: ub2 (the adjoint of u^{n-1}) is written, not accumulated.
// That buffer's last role was as ub of step n+1, and between its
// consumption there and this write nothing adds to its interior; so ub
// needs clearing only on the halo (its interior is overwritten as the next
// step's ub2 before anything reads it).
constexpr int kAdjZChunk = 8;

template <int R>
__global__ void __launch_bounds__(64 * kAdjLines)
k_adj_update_z(Params P, const real_t *__restrict__ vel, real_t *__restrict__ ub,
               const real_t *__restrict__ W, real_t *__restrict__ ub1,
               real_t *__restrict__ ub2, const real_t *__restrict__ u0,
               const real_t *__restrict__ u1, const real_t *__restrict__ u2,
               real_t *__restrict__ gacc, int image)
{
    const int j = blockIdx.y * kAdjLines + threadIdx.y;
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (j >= P.nx || k >= P.ny)
        return;
    const int z0 = blockIdx.z * kAdjZChunk;
    const int z1 = min(z0 + kAdjZChunk, P.nz);
    const long long col = (long long)j * P.sx + k;
    // wq[R + d] = W at plane i + d (zero outside the grid)
    real_t wq[2 * R + 1];
#pragma unroll
    for (int d = -R; d < R; d++) {
        const int z = z0 + d;
        wq[R + d + 1] = (z >= 0 && z < P.nz) ? W[(long long)z * P.sz + col] : (real_t)0;
    }
    const bool in_xy = j >= R && j < P.nx - R && k >= R && k < P.ny - R;
    // Operands of one point: loaded a plane ahead (all in bounds), so the
    // loads of plane i+1 overlap the arithmetic of plane i.
    struct Ops {
        real_t head, w, e, v, b1, x0, x1, x2;
        real_t xn[2 * R], yn[2 * R];
    };
    auto load = [&](int i, Ops &o) {
        const long long p = (long long)i * P.sz + col;
        o.head = (i + R < P.nz) ? W[p + R * P.sz] : (real_t)0;
#pragma unroll
        for (int m = 1; m <= R; m++) {
            o.xn[2 * m - 2] = (j + m < P.nx) ? W[p + m * P.sx] : (real_t)0;
            o.xn[2 * m - 1] = (j - m >= 0) ? W[p - m * P.sx] : (real_t)0;
            o.yn[2 * m - 2] = (k + m < P.ny) ? W[p + m] : (real_t)0;
            o.yn[2 * m - 1] = (k - m >= 0) ? W[p - m] : (real_t)0;
        }
        o.w = ub[p];
        o.e = P.e[p];
        o.v = vel[p];
        o.b1 = ub1[p];
        o.x0 = u0[p];
        o.x1 = u1[p];
        o.x2 = u2[p];
    };
    Ops cur, nxt;
    load(z0, cur);
    for (int i = z0; i < z1; i++) {
        if (i + 1 < z1)
            load(i + 1, nxt);
#pragma unroll
        for (int d = 0; d < 2 * R; d++)
            wq[d] = wq[d + 1];
        wq[2 * R] = cur.head;
        const long long p = (long long)i * P.sz + col;
        // stencil_clipped's operations in its order, with the z terms from
        // the queue and the x/y terms preloaded
        const real_t c0 = wq[R] * P.c2[0];
        real_t sz = c0, sx = c0, sy = c0;
#pragma unroll
        for (int m = 1; m <= R; m++) {
            const real_t cm = P.c2[m];
            real_t t = 0;
            if (i + m < P.nz) t += wq[R + m];
            if (i - m >= 0) t += wq[R - m];
            sz += cm * t;
            t = 0;
            if (j + m < P.nx) t += cur.xn[2 * m - 2];
            if (j - m >= 0) t += cur.xn[2 * m - 1];
            sx += cm * t;
            t = 0;
            if (k + m < P.ny) t += cur.yn[2 * m - 2];
            if (k - m >= 0) t += cur.yn[2 * m - 1];
            sy += cm * t;
        }
        const real_t lt = sy / P.dy2 + sx / P.dx2 + sz / P.dz2;
        const real_t w = cur.w;
        if (!(in_xy && i >= R && i < P.nz - R)) {
            ub1[p] = cur.b1 + lt;
            ub[p] = 0;
        } else {
            const double e = (double)cur.e;
            const double rden = 1.0 / (1.0 + e);
            const double a = 2.0 * rden, b = (1.0 - e) * rden;
            ub1[p] = (real_t)((double)cur.b1 + a * (double)w + (double)lt);
            ub2[p] = (real_t)(-b * (double)w);
            if (w != (real_t)0) {
                const double x0 = (double)cur.x0, x1 = (double)cur.x1, x2 = (double)cur.x2;
                if (image) {
                    gacc[p] += (real_t)((double)w * x0);
                } else {
                    const double rv = 1.0 / (double)cur.v;
                    const double da = -4.0 * e * rden * rden * rv;
                    const double step = x0 - a * x1 + b * x2;     // c (L u1 + s)
                    gacc[p] += (real_t)((double)w * (da * (x1 - x2) + 2.0 * rden * rv * step));
                }
            }
        }
        cur = nxt;
    }
}
#endif

#if SW_DIM == 2
// 2D form of k_adj_update_z: each thread owns an x column of a z chunk
// (x is the contiguous axis), W's z neighbours in a register queue, the
// next row's operands loaded while the current one is computed, and ub2
// written rather than accumulated (see k_adj_update_z).
constexpr int kAdjZChunk = 8;

template <int R>
__global__ void __launch_bounds__(256)
k_adj_update_z2d(Params P, const real_t *__restrict__ vel, real_t *__restrict__ ub,
                 const real_t *__restrict__ W, real_t *__restrict__ ub1,
                 real_t *__restrict__ ub2, const real_t *__restrict__ u0,
                 const real_t *__restrict__ u1, const real_t *__restrict__ u2,
                 real_t *__restrict__ gacc, int image)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.nx)
        return;
    const int z0 = blockIdx.y * kAdjZChunk;
    const int z1 = min(z0 + kAdjZChunk, P.nz);
    real_t wq[2 * R + 1];
#pragma unroll
    for (int d = -R; d < R; d++) {
        const int z = z0 + d;
        wq[R + d + 1] = (z >= 0 && z < P.nz) ? W[(long long)z * P.sz + j] : (real_t)0;
    }
    const bool in_x = j >= R && j < P.nx - R;
    struct Ops {
        real_t head, w, e, v, b1, x0, x1, x2;
        real_t xn[2 * R];
    };
    auto load = [&](int i, Ops &o) {
        const long long p = (long long)i * P.sz + j;
        o.head = (i + R < P.nz) ? W[p + R * P.sz] : (real_t)0;
#pragma unroll
        for (int m = 1; m <= R; m++) {
            o.xn[2 * m - 2] = (j + m < P.nx) ? W[p + m] : (real_t)0;
            o.xn[2 * m - 1] = (j - m >= 0) ? W[p - m] : (real_t)0;
        }
        o.w = ub[p];
        o.e = P.e[p];
        o.v = vel[p];
        o.b1 = ub1[p];
        o.x0 = u0[p];
        o.x1 = u1[p];
        o.x2 = u2[p];
    };
    Ops cur, nxt;
    load(z0, cur);
    for (int i = z0; i < z1; i++) {
        if (i + 1 < z1)
            load(i + 1, nxt);
#pragma unroll
        for (int d = 0; d < 2 * R; d++)
            wq[d] = wq[d + 1];
        wq[2 * R] = cur.head;
        const long long p = (long long)i * P.sz + j;
        // stencil_clipped's operations in its order (2D: x then z sums)
        const real_t c0 = wq[R] * P.c2[0];
        real_t sz = c0, sx = c0;
#pragma unroll
        for (int m = 1; m <= R; m++) {
            const real_t cm = P.c2[m];
            real_t t = 0;
            if (i + m < P.nz) t += wq[R + m];
            if (i - m >= 0) t += wq[R - m];
            sz += cm * t;
            t = 0;
            if (j + m < P.nx) t += cur.xn[2 * m - 2];
            if (j - m >= 0) t += cur.xn[2 * m - 1];
            sx += cm * t;
        }
        const real_t lt = sx / P.dx2 + sz / P.dz2;
        const real_t w = cur.w;
        if (!(in_x && i >= R && i < P.nz - R)) {
            ub1[p] = cur.b1 + lt;
            ub[p] = 0;
        } else {
            const double e = (double)cur.e;
            const double rden = 1.0 / (1.0 + e);
            const double a = 2.0 * rden, b = (1.0 - e) * rden;
            ub1[p] = (real_t)((double)cur.b1 + a * (double)w + (double)lt);
            ub2[p] = (real_t)(-b * (double)w);
            if (w != (real_t)0) {
                const double x0 = (double)cur.x0, x1 = (double)cur.x1, x2 = (double)cur.x2;
                if (image) {
                    gacc[p] += (real_t)((double)w * x0);
                } else {
                    const double rv = 1.0 / (double)cur.v;
                    const double da = -4.0 * e * rden * rden * rv;
                    const double step = x0 - a * x1 + b * x2;     // c (L u1 + s)
                    gacc[p] += (real_t)((double)w * (da * (x1 - x2) + 2.0 * rden * rv * step));
                }
            }
        }
        cur = nxt;
    }
}
#endif

// Wavelet gradient of step n (k = n+1): gw[s] = sum_p kws_{s,p} c_p ub_p
// (c = q / den) over the source box's interior points, in float64, one
// block per source.  (The velocity derivative of the source term is part
// of k_adj_update's step difference; `grad` is unused and may be null.)
__global__ void __launch_bounds__(256)
k_adj_sources(Params P, const real_t *__restrict__ ub, real_t *__restrict__ gw)
{
    __shared__ double part[256];
    const int s = blockIdx.x;
    if (s >= P.nsrc)
        return;
    const Box b = P.sbox[s];
    const int cz = b.hi[0] - b.lo[0] + 1, cx = b.hi[1] - b.lo[1] + 1;
#if SW_DIM == 3
    const int cy = b.hi[2] - b.lo[2] + 1;
#else
    const int cy = 1;
#endif
    double acc = 0;
    for (int t = threadIdx.x; t < cz * cx * cy; t += blockDim.x) {
        const int ii = t / (cx * cy), rem = t - ii * cx * cy;
        const int jj = rem / cy, kk = rem - jj * cy;
        const int i = b.lo[0] + ii, j = b.lo[1] + jj;
#if SW_DIM == 3
        const int k = b.lo[2] + kk;
#else
        const int k = 0;
        (void)kk;
#endif
        if (!interior(P, i, j, k))
            continue;
        const long long p = (long long)i * P.sz + (long long)j * P.sx + k;
        const double kws = (double)box_weight(b, P.sw, i, j, k);
        const double q = (double)P.q[p], den = 1.0 + (double)P.e[p];
        acc += kws * (q / den) * (double)ub[p];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        gw[s] = (real_t)part[0];
}

__global__ void k_fill(real_t *p, long long n, real_t v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// r = d - o (the L2 misfit's adjoint source from the shot record d and the
// observed data o)
__global__ void k_residual(const real_t *__restrict__ d, const real_t *__restrict__ o,
                           real_t *__restrict__ r, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        r[i] = d[i] - o[i];
}

__global__ void k_fill_f64(double *p, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

}  // namespace sw
