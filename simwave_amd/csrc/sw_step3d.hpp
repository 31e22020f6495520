// Tiled, software-pipelined 3D step kernel for gfx950.
//
// One step of the reference time loop (stencil, sources, boundaries; wave.c
// sections 1-3) over the interior, fused in one launch.  Work unit = tile:
// 64 points along the contiguous axis y (one wavefront row) x TXR = 4*RPT rows
// along x x zlen planes along z.  A persistent grid walks the tiles; the
// blocks resident on one XCD (blockIdx % 8) take a contiguous range of tiles
// so neighbouring tiles share that XCD's L2.  Inside a tile each thread owns
// RPT x-rows of one y column and marches along z:
//   * z neighbours: a register queue of 2R+1 values per row;
//   * x/y neighbours: the current plane in LDS with an R-wide halo, double
//     buffered so one barrier per plane suffices;
//   * the next plane's centre values, halo, u_prev, q and the damping term are
//     loaded while the current plane is computed;
//   * every global access is a buffer instruction: an SGPR descriptor rebased
//     at the tile's first plane, a per-plane SGPR offset and a per-thread
//     VGPR offset fixed for the whole tile, so the z march needs no vector
//     address arithmetic.  Offsets of lanes outside the grid are clamped to a
//     valid element, so no load is predicated.
// Per-(tile, plane) flags mark planes without damping (their damping terms
// are not read and the update needs no division), and tile-uniform tests
// keep the boundary and source code out of interior tiles.
#pragma once
#include "sw_kernels.hpp"

namespace sw {

#if SW_DIM == 3

struct Tiling {
    int gk, gj, gz;     // tiles along y (64 wide), x (TXR rows), z (zlen planes)
    int zlen;
    int ntiles;
    int ncols;          // gk * gj
    int nwords;         // 32-bit mask words per tile column (nz/32 + 3)
    const unsigned *dmask;  // [ncols][nwords]: bit z set where plane z of the
                            // tile column holds a non-zero damping term
};

// Sources restricted to the tile's mask, in index order (the reference's
// sequential += order, wave.c:208-295).
__device__ __forceinline__ real_t add_sources_masked(const Params &P, int n, int i,
                                                     int j, int k, real_t qv,
                                                     real_t ev, real_t v,
                                                     unsigned long long mask)
{
    while (mask) {
        const int s = __builtin_ctzll(mask);
        mask &= mask - 1;
        const Box b = P.sbox[s];
        if (i < b.lo[0] || i > b.hi[0])
            continue;
        if (j < b.lo[1] || j > b.hi[1] || k < b.lo[2] || k > b.hi[2])
            continue;
        const real_t amp = wavelet_amp(P, n, s);
        if (amp == (real_t)0)
            continue;
        v += source_term(qv, ev, box_weight(b, P.sw, i, j, k), amp);
    }
    return v;
}

// Marks tile planes holding a non-zero damping term.  grid (ncols, nz); the
// mask words must be zeroed beforehand.
template <int TXR>
__global__ void __launch_bounds__(256)
k_damp_flags(const real_t *__restrict__ e, int nz, int nx, int ny, int R, int gk,
             int nwords, unsigned *__restrict__ mask)
{
    const int col = blockIdx.x, z = blockIdx.y;
    const int kt = col % gk, jt = col / gk;
    const int k0 = R + kt * 64, j0 = R + jt * TXR;
    int any = 0;
    for (int t = threadIdx.x; t < 64 * TXR; t += blockDim.x) {
        const int k = k0 + (t & 63), j = j0 + (t >> 6);
        if (k < ny && j < nx)
            any |= e[((long long)z * nx + j) * ny + k] != (real_t)0;
    }
    any = __syncthreads_or(any);
    if (threadIdx.x == 0 && any)
        atomicOr(&mask[(long long)col * nwords + (z >> 5)], 1u << (z & 31));
}

// Bits z0 .. z0+63 of a tile column's damping mask.
__device__ __forceinline__ unsigned long long damp_bits(const Tiling &T, int col, int z0)
{
    const unsigned *w = T.dmask + (long long)col * T.nwords + (z0 >> 5);
    const unsigned sh = z0 & 31;
    const unsigned long long lo = (unsigned long long)w[0] | ((unsigned long long)w[1] << 32);
    const unsigned long long hi = w[2];
    return sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
}

typedef __amdgpu_buffer_rsrc_t rsrc_t;

__device__ __forceinline__ rsrc_t make_rsrc(const void *base, unsigned long long bytes)
{
    const int nrec = bytes >= 0xffffffffull ? (int)0xffffffff : (int)bytes;
    // dword3 0x00020000: 32-bit data format, the usual raw-buffer setting
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, nrec,
                                             0x00020000);
}

// Buffer loads/stores move raw bits; the values are bit-cast, not converted.
__device__ __forceinline__ real_t bload(rsrc_t r, unsigned voff, unsigned soff)
{
#if defined(SW_DOUBLE)
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
#else
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
#endif
}

__device__ __forceinline__ double bload_f64(rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

__device__ __forceinline__ void bstore(real_t v, rsrc_t r, unsigned voff, unsigned soff)
{
#if defined(SW_DOUBLE)
    typedef unsigned v2u32 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u32, v), r, voff, soff, 0);
#else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
#endif
}

// The kernel's run constants, kept small so they stay in SGPRs.
struct TiledArgs {
    const real_t *uc, *up;
    const real_t *C, *B;   // update coefficients (working type)
    const double *A;       // 2.0/den
    real_t *un;
    int nz, nx, ny;
    unsigned pbytes;    // bytes per z plane (working type)
    int bc;             // 6 faces x 2 bits: z_b z_a x_b x_a y_b y_a
    int fastdiv;        // FixedDiv usable for all three divisors
    real_t dz2, dx2, dy2;
    FixedDiv dvz, dvx, dvy;
    real_t div_lo, div_hi;  // numerator range safe for all three divisors
    real_t c2[kMaxR + 1];
    int nsrc;
    const Box *sbox;
    Tiling T;
};

template <int R, int RPT, int NW>
__global__ void __launch_bounds__(64 * NW)
k_step3d_tiled(TiledArgs A, const Params *__restrict__ Pd, int n)
{
    constexpr int TY = 64;
    constexpr int NT = 64 * NW;
    constexpr int TXR = NW * RPT;
    constexpr int LS = 128;                 // LDS row stride (elements)
    constexpr int H = TXR + 2 * R + 1;      // + one junk row
    constexpr int JUNK = TXR + 2 * R;
    constexpr int HX = (2 * R + NW - 1) / NW;
    constexpr int NYH = TXR * 2 * R;
    constexpr int HY = (NYH + NT - 1) / NT;
    constexpr unsigned ascale = sizeof(double) / sizeof(real_t);
    __shared__ real_t lds[2][H][LS];

    const int tx = threadIdx.x, w = threadIdx.y;
    const int tid = w * TY + tx;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const unsigned PB = A.pbytes;
    const unsigned esz = sizeof(real_t);
    const Tiling &T = A.T;

    const int nb8 = gridDim.x >> 3;
    const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3;
    const int t_begin = (int)(((long long)xcd * T.ntiles) >> 3);
    const int t_end = (int)(((long long)(xcd + 1) * T.ntiles) >> 3);

    for (int t = t_begin + lb; t < t_end; t += nb8) {
        const int zc = t % T.gz, col = t / T.gz;
        const int kt = col % T.gk, jt = col / T.gk;
        const int k0 = R + kt * TY, j0 = R + jt * TXR;
        const int z0 = R + zc * T.zlen;
        const int z1 = min(z0 + T.zlen, nz - R);
        const int zb = z0 - R;              // first plane the tile touches
        const int k = k0 + tx;

        // descriptors rebased at plane zb
        const unsigned long long off0 = (unsigned long long)zb * PB;
        const unsigned long long rem = (unsigned long long)(nz - zb) * PB;
        const rsrc_t r_uc = make_rsrc(A.uc + off0 / esz, rem);
        const rsrc_t r_up = make_rsrc(A.up + off0 / esz, rem);
        const rsrc_t r_C = make_rsrc(A.C + off0 / esz, rem);
        const rsrc_t r_B = make_rsrc(A.B + off0 / esz, rem);
        const rsrc_t r_A = make_rsrc(A.A + off0 / esz, rem * ascale);
        const rsrc_t r_un = make_rsrc(A.un + off0 / esz, rem);

        unsigned cb[RPT];
        bool act[RPT];
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int j = j0 + w + NW * q;
            const bool in = (k < ny) && (j < nx);
            act[q] = (k < ny - R) && (j < nx - R);
            cb[q] = in ? ((unsigned)j * (unsigned)ny + (unsigned)k) * esz : 0u;
        }
        unsigned xh_off[HX];
        int xh_lds[HX];
#pragma unroll
        for (int h = 0; h < HX; h++) {
            const int hr = w + NW * h;
            const int jr = hr < R ? j0 - R + hr : j0 + TXR + (hr - R);
            const bool ok = hr < 2 * R && k < ny && jr < nx;
            xh_off[h] = ok ? ((unsigned)jr * (unsigned)ny + (unsigned)k) * esz : 0u;
            xh_lds[h] = (ok ? (hr < R ? hr : hr + TXR) : JUNK) * LS + R + tx;
        }
        unsigned yh_off[HY];
        int yh_lds[HY];
#pragma unroll
        for (int h = 0; h < HY; h++) {
            const int idx = tid + NT * h;
            const int row = idx / (2 * R), c = idx % (2 * R);
            const int jr = j0 + row;
            const int kr = c < R ? k0 - R + c : k0 + TY + (c - R);
            const bool ok = idx < NYH && jr < nx && kr < ny;
            yh_off[h] = ok ? ((unsigned)jr * (unsigned)ny + (unsigned)kr) * esz : 0u;
            yh_lds[h] = ok ? (R + row) * LS + (c < R ? c : c + TY) : JUNK * LS + tx;
        }

        // tile-uniform boundary and source tests
        const int bc = A.bc;
        const int r2 = 2 * R;
        bool bc_xy = false;
        if ((bc >> 8 & 3) && k0 <= r2)
            bc_xy = true;
        if ((bc >> 10 & 3) && k0 + TY - 1 >= ny - r2 - 1)
            bc_xy = true;
        if ((bc >> 4 & 3) && j0 <= r2)
            bc_xy = true;
        if ((bc >> 6 & 3) && j0 + TXR - 1 >= nx - r2 - 1)
            bc_xy = true;
        const bool bz_b = (bc & 3) != 0, bz_a = (bc >> 2 & 3) != 0;
        unsigned long long smask = 0;
        for (int s = 0; s < A.nsrc; s++) {
            const Box b = A.sbox[s];
            if (b.hi[1] >= j0 && b.lo[1] < j0 + TXR && b.hi[2] >= k0 &&
                b.lo[2] < k0 + TY)
                smask |= 1ull << s;
        }

        // prologue: z queue planes zb .. zb+2R-1 and the first plane
        real_t qz[RPT][2 * R + 1];
#pragma unroll
        for (int m = 0; m < 2 * R; m++)
#pragma unroll
            for (int q = 0; q < RPT; q++)
                qz[q][m + 1] = bload(r_uc, cb[q], (unsigned)m * PB);
        real_t nc[RPT], np[RPT], nC[RPT], nB[RPT], nhx[HX], nhy[HY];
        double nA[RPT];
        const unsigned long long dbits = damp_bits(T, col, z0);
        auto load_plane = [&](int I) {
            const unsigned sp = (unsigned)(I - zb) * PB;   // plane I
            const unsigned sc = sp + (unsigned)R * PB;     // plane I+R
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                nc[q] = bload(r_uc, cb[q], sc);
                np[q] = bload(r_up, cb[q], sp);
                nC[q] = bload(r_C, cb[q], sp);
            }
            if ((dbits >> (I - z0)) & 1) {
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    nB[q] = bload(r_B, cb[q], sp);
                    nA[q] = bload_f64(r_A, cb[q] * ascale, sp * ascale);
                }
            } else {
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    nB[q] = (real_t)1;
                    nA[q] = 2.0;
                }
            }
#pragma unroll
            for (int h = 0; h < HX; h++)
                nhx[h] = bload(r_uc, xh_off[h], sp);
#pragma unroll
            for (int h = 0; h < HY; h++)
                nhy[h] = bload(r_uc, yh_off[h], sp);
        };
        load_plane(z0);

        // Stores of plane i are issued in iteration i+1, ahead of that
        // iteration's prefetch: vmcnt counts stores too, so a store issued
        // after the prefetch would make the next loop-top wait cover it.
        real_t pend[RPT];
        bool have_pend = false;
        unsigned pend_so = 0;

        for (int i = z0; i < z1; i++) {
            real_t *lb_ = &lds[(i - z0) & 1][0][0];
            real_t cp[RPT], cC[RPT], cB[RPT];
            double cA[RPT];
#pragma unroll
            for (int q = 0; q < RPT; q++) {
#pragma unroll
                for (int m = 0; m < 2 * R; m++)
                    qz[q][m] = qz[q][m + 1];
                qz[q][2 * R] = nc[q];
                cp[q] = np[q];
                cC[q] = nC[q];
                cB[q] = nB[q];
                cA[q] = nA[q];
                lb_[(R + w + NW * q) * LS + R + tx] = qz[q][R];
            }
#pragma unroll
            for (int h = 0; h < HX; h++)
                lb_[xh_lds[h]] = nhx[h];
#pragma unroll
            for (int h = 0; h < HY; h++)
                lb_[yh_lds[h]] = nhy[h];
            if (have_pend) {
#pragma unroll
                for (int q = 0; q < RPT; q++)
                    if (act[q])
                        bstore(pend[q], r_un, cb[q], pend_so);
            }
            if (i + 1 < z1)
                load_plane(i + 1);
            __syncthreads();

            const bool bc_here = bc_xy || (bz_b && i <= r2) || (bz_a && i >= nz - r2 - 1);
            const bool special = (smask != 0) | bc_here;
            const unsigned so = (unsigned)(i - zb) * PB;
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                // lowest address this point reads: (row - R, column - R)
                const real_t *lo = lb_ + (w + NW * q) * LS + tx;
                const real_t c = qz[q][R];
                real_t acc_c = A.c2[0] * c;
                real_t acc_m = A.c2[0] * c;
                real_t acc_z = A.c2[0] * c;
#pragma unroll
                for (int m = 1; m <= R; m++) {
                    acc_c += A.c2[m] * (lo[R * LS + R + m] + lo[R * LS + R - m]);
                    acc_m += A.c2[m] * (lo[(R + m) * LS + R] + lo[(R - m) * LS + R]);
                    acc_z += A.c2[m] * (qz[q][R + m] + qz[q][R - m]);
                }
                // fixed-divisor path when every numerator is in the safe
                // range or all are zero; IEEE division otherwise
                const real_t mx = fmax(fmax(fabs(acc_c), fabs(acc_m)), fabs(acc_z));
                const real_t mn = fmin(fmin(fabs(acc_c), fabs(acc_m)), fabs(acc_z));
                const bool fast = A.fastdiv & (((mn >= A.div_lo) & (mx <= A.div_hi)) |
                                               (mx == (real_t)0));
                real_t lap;
                if (fast)
                    lap = fixed_div(acc_c, A.dvy) + fixed_div(acc_m, A.dvx) +
                          fixed_div(acc_z, A.dvz);
                else
                    lap = acc_c / A.dy2 + acc_m / A.dx2 + acc_z / A.dz2;
                real_t value = 0;
                value += lap;
                real_t v = leapfrog_coef(value, cC[q], cB[q], cA[q], c, cp[q]);
                if (special) {
                    // tiles touching a source box or a boundary face: store now
                    if (!act[q])
                        continue;
                    const int j = j0 + w + NW * q;
                    const long long p = (long long)i * nx * ny + (long long)j * ny + k;
                    if (smask)
                        v = add_sources_masked(*Pd, n, i, j, k, Pd->q[p], Pd->e[p], v, smask);
                    if (bc_here)
                        store_with_bc(*Pd, A.un, p, i, j, k, v);
                    else
                        bstore(v, r_un, cb[q], so);
                } else {
                    pend[q] = v;
                }
            }
            have_pend = !special;
            pend_so = so;
        }
        if (have_pend) {
#pragma unroll
            for (int q = 0; q < RPT; q++)
                if (act[q])
                    bstore(pend[q], r_un, cb[q], pend_so);
        }
        __syncthreads();
    }
}

#endif  // SW_DIM == 3

}  // namespace sw
