// Tiled 3D step kernel for gfx950 with 16- or 8-byte lanes.
//
// One step of the reference time loop (stencil, sources, boundaries; wave.c
// sections 1-3) over the interior, fused in one launch.  Every lane owns V
// consecutive points of the contiguous axis y, moved by one buffer access
// (16 bytes: V = 4 floats or 2 doubles; 8 bytes: 2 floats or 1 double); a
// wavefront covers WR rows of 64/WR lanes, so a tile row is TY = 64/WR * V
// points wide.  A work-group of NW wavefronts owns a tile
// of TXR = NW*WR*RPT rows along x and marches it along z over zlen planes:
//   * z neighbours: a per-lane register queue of 2R+1 vectors per row;
//   * x/y neighbours: the current plane in LDS with an R-wide halo (rounded
//     up to whole vectors), double buffered so one barrier per plane suffices;
//   * the next plane's centre values, halos, u_prev, C = (dt2/s)/den and
//     (on planes that hold damping) the damping term are loaded while the
//     current plane is computed;
//   * every global access is a buffer instruction on a descriptor rebased at
//     the tile's first plane.  The z-queue loads carry the plane offset in the
//     VGPR offset, so the descriptor's range check (which covers the VGPR
//     offset) turns the only possible overrun, the tail vector of the array's
//     last row, into zeros; the other accesses stay inside the array by
//     construction and use an SGPR plane offset.
// Per-(tile column, plane) bits mark planes without damping: there the update
// needs no division and the damping term is not read.  Tile-uniform tests
// keep the boundary and source code out of interior tiles.  Tiles are visited
// column-fastest (all columns of one z range before the next range); with
// order 1 the columns are cut into 8 contiguous chunks, one per XCD (blocks
// are dispatched to XCDs round-robin), so tiles sharing halo rows share an L2.
#pragma once
#include <type_traits>

#include "sw_kernels.hpp"

namespace sw {

#if SW_DIM == 3

// A lane's V points as one vector register group.
template <int V>
struct LaneVec {
    typedef real_t type __attribute__((ext_vector_type(V)));
};
constexpr int kVec16 = 16 / (int)sizeof(real_t);   // points per 16-byte lane
constexpr int kVec8 = 8 / (int)sizeof(real_t);     // points per 8-byte lane
typedef unsigned v4u32 __attribute__((ext_vector_type(4)));
typedef unsigned v2u32 __attribute__((ext_vector_type(2)));

struct Tiling {
    int gk, gj, gz;     // tiles along y (ty wide), x (txr rows), z (zlen planes)
    int ty, txr;        // tile width along y and rows along x
    int zlen;
    int ntiles;
    int ncols;          // gk * gj
    int order;          // 0 column-fastest, 1 column-fastest in 8 XCD chunks
    int chunk;          // columns per XCD chunk (order 1)
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
__global__ void __launch_bounds__(256)
k_damp_flags(const real_t *__restrict__ e, int nz, int nx, int ny, int R, int gk,
             int ty, int txr, int nwords, unsigned *__restrict__ mask)
{
    const int col = blockIdx.x, z = blockIdx.y;
    const int kt = col % gk, jt = col / gk;
    const int k0 = R + kt * ty, j0 = R + jt * txr;
    int any = 0;
    for (int t = threadIdx.x; t < ty * txr; t += blockDim.x) {
        const int k = k0 + t % ty, j = j0 + t / ty;
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

// A buffer offset past the end of every descriptor the tiled kernel builds
// (their ranges stay below 2^31 bytes), also with a plane offset (below
// 2^31) added: loads return zeros, stores are dropped.
constexpr unsigned kOOB = 0x80000000u;

__device__ __forceinline__ rsrc_t make_rsrc(const void *base, unsigned long long bytes)
{
    const int nrec = bytes >= 0xffffffffull ? (int)0xffffffff : (int)bytes;
    // dword3 0x00020000: 32-bit data format, the usual raw-buffer setting
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, nrec,
                                             0x00020000);
}

// 16- or 8-byte buffer accesses; the values are bit-cast, not converted.
template <int V>
__device__ __forceinline__ typename LaneVec<V>::type vload(rsrc_t r, unsigned voff,
                                                           unsigned soff)
{
    typedef typename LaneVec<V>::type vt;
    if constexpr (sizeof(vt) == 16)
        return __builtin_bit_cast(vt, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
    else
        return __builtin_bit_cast(vt, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// The store adds the plane offset into the VGPR offset and passes a zero
// SGPR offset.  A 128-bit buffer store reads its data VGPRs after issue: an
// instruction that overwrites them right behind it needs wait states, which
// the compiler inserts only for stores without an SGPR offset (the ISA's
// wait-state table exempts those).  On gfx950 an SGPR-offset store followed
// at once by a write of its data registers stored the new values (seen as
// byte offsets in the wavefield); with the zero SGPR offset the wait states
// are there.  8-byte stores take the same form.
template <typename VT>
__device__ __forceinline__ void vstore(VT v, rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(VT) == 16)
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32, v), r, voff + soff,
                                               0u, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u32, v), r, voff + soff, 0u,
                                              0);
}

__device__ __forceinline__ void sstore(real_t v, rsrc_t r, unsigned voff, unsigned soff)
{
#if defined(SW_DOUBLE)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u32, v), r, voff, soff, 0);
#else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
#endif
}

// Leap-frog update on a plane holding damping: the reference's full
// expression (wave.c:177-185) with its step-independent factor
// C = (dt2/s)/den precomputed (k_scaled_q).  With e == 0 it gives the same
// bits as leapfrog_plain, so one form serves every point of such a plane.
__device__ __forceinline__ real_t leapfrog_damped(real_t value, real_t C, real_t e,
                                                  real_t uc, real_t up)
{
    const real_t den = (real_t)(1.0 + (double)e);
    const real_t num = (real_t)(1.0 - (double)e);
    value = value * C;
    const double a = 2.0 / (double)den;
    const real_t b = num / den;
    double t = a * (double)uc - (double)(b * up);
    t = t + (double)value;
    return (real_t)t;
}

// The same update in float32 with one correctly rounded float64 reciprocal
// y = RN(1/den) for the two quotients: RN(2/den) = 2y exactly, and for float
// operands num/den (a normal quotient), rounding num*y first to double and
// then to float gives RN(num/den): the product is within 2^-52 of the
// quotient, and a quotient of two floats that is not a float lies at least
// 2^-49 (relative) from every midpoint between floats.
__device__ __forceinline__ float leapfrog_damped_rcp(float value, float C, float e,
                                                     float uc, float up)
{
    const float den = (float)(1.0 + (double)e);
    const float num = (float)(1.0 - (double)e);
    value = value * C;
    const double y = 1.0 / (double)den;
    const double a = 2.0 * y;
    const float b = (float)((double)num * y);
    double t = a * (double)uc - (double)(b * up);
    t = t + (double)value;
    return (float)t;
}

// Leap-frog update where den == num == 1 (no damping at the point; there
// C == q).
__device__ __forceinline__ real_t leapfrog_plain(real_t value, real_t C, real_t uc,
                                                 real_t up)
{
    value = value * C;
    double t = 2.0 * (double)uc - (double)up;
    t = t + (double)value;
    return (real_t)t;
}

// The kernel's run constants, kept small so they stay in SGPRs.
struct StepArgs {
    const real_t *uc, *up;
    const real_t *qc;   // (dt2/s)/den per point
    const real_t *e;
    real_t *un;
    int nz, nx, ny;
    unsigned pbytes;    // bytes per z plane (working type)
    int bc;             // 6 faces x 2 bits: z_b z_a x_b x_a y_b y_a
    int fastdiv;        // FixedDiv usable for all three divisors
    real_t dz2, dx2, dy2;
    FixedDiv dvz, dvx, dvy;
    real_t div_lo, div_hi;  // numerator range safe for all three divisors
    double ry, rx, rz;      // RN(1/dy2) ... in float64 (quotient mode D64)
    real_t c2[kMaxR + 1];
    int nsrc;
    const Box *sbox;
    Tiling T;
};

template <int R, int WR, int RPT, int NW, int VP>
struct TileShape {
    static constexpr int V = VP;
    static constexpr int LW = 64 / WR;                 // lanes per row
    static constexpr int TY = LW * V;                  // tile width along y
    static constexpr int TXR = NW * WR * RPT;          // tile rows along x
    static constexpr int PADL = (R + V - 1) / V * V;   // y halo, whole vectors
    static constexpr int HV = PADL / V;                // halo vectors per side
    static constexpr int LS = TY + 2 * PADL;           // LDS row stride
    static constexpr int H = TXR + 2 * R;              // LDS rows
    static constexpr int SLOTS = NW * WR;              // rows one instruction covers
    static constexpr int HX = (2 * R + SLOTS - 1) / SLOTS;
    static constexpr int NYV = TXR * 2 * HV;           // y-halo vectors per plane
    static constexpr int HY = (NYV + 64 * NW - 1) / (64 * NW);
};

// D64 (float32 only): the Laplacian quotients sum/h2 as RN(sum * RN(1/h2)),
// the product in float64; equal to the IEEE float quotient for every float
// sum when h2 = B 2^t (B odd) has t <= 0 and lies in [2^-126, 2^24), which
// the host checks (the product is within 2^-52 of the quotient, every
// non-float quotient at least 2^-49 from a rounding midpoint, and with t <= 0
// no quotient in the subnormal range lands on a midpoint).  Otherwise the
// fixed-divisor float quotients with their range vote.
template <int R, int WR, int RPT, int NW, int VP, bool D64 = false>
__global__ void __launch_bounds__(64 * NW)
k_step3d_vec(StepArgs A, const Params *__restrict__ Pd, int n)
{
    using S = TileShape<R, WR, RPT, NW, VP>;
    typedef typename LaneVec<VP>::type vreal;
    constexpr int V = S::V, LW = S::LW, TY = S::TY, TXR = S::TXR, PADL = S::PADL;
    constexpr int HV = S::HV, LS = S::LS, H = S::H, SLOTS = S::SLOTS, HX = S::HX;
    constexpr int NYV = S::NYV, HY = S::HY;
    constexpr int NT = 64 * NW;
    constexpr unsigned esz = sizeof(real_t);
    // two plane buffers, each followed by one dump slot per thread that
    // takes the halo writes of lanes without a halo vector, so that every
    // lane issues the same LDS stores (no branches around them)
    constexpr int DUMP = H * LS;
    __shared__ __align__(16) real_t lds[2][H * LS + NT * V];

    const int lane = threadIdx.x, w = threadIdx.y;
    const int tid = w * 64 + lane;
    const int slot = w * WR + lane / LW;     // row slot of this lane
    const int lc = lane % LW;                // vector column of this lane
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const unsigned PB = A.pbytes;
    const Tiling &T = A.T;

    int col, zc;
    if (T.order == 1) {
        const int xcd = blockIdx.x & 7, s = blockIdx.x >> 3;
        zc = s / T.chunk;
        col = xcd * T.chunk + s % T.chunk;
        if (col >= T.ncols)
            return;
    } else {
        col = blockIdx.x % T.ncols;
        zc = blockIdx.x / T.ncols;
    }
    const int kt = col % T.gk, jt = col / T.gk;
    const int k0 = R + kt * TY, j0 = R + jt * TXR;
    const int z0 = R + zc * T.zlen;
    const int z1 = min(z0 + T.zlen, nz - R);
    const int zb = z0 - R;                  // first plane the tile touches

    // descriptors rebased at plane zb, covering the planes the tile reads,
    // zb .. z1+R-1 (at most 64+2R planes: below 2^31 bytes, checked by the
    // host), so that kOOB is out of range for every descriptor
    const unsigned long long off0 = (unsigned long long)zb * PB;
    const unsigned long long rem = (unsigned long long)(z1 + R - zb) * PB;
    const rsrc_t r_uc = make_rsrc(A.uc + off0 / esz, rem);
    const rsrc_t r_up = make_rsrc(A.up + off0 / esz, rem);
    const rsrc_t r_q = make_rsrc(A.qc + off0 / esz, rem);
    const rsrc_t r_e = make_rsrc(A.e + off0 / esz, rem);
    const rsrc_t r_un = make_rsrc(A.un + off0 / esz, rem);

    // This lane's points: rows j0 + slot*RPT + q (adjacent rows share their
    // x neighbours), columns k .. k+V-1.
    // Lanes past the row end read the row start (values unused); a lane
    // straddling the row end reads at most V-1 elements of the next row.
    const int k = k0 + lc * V;
    const int kl = k < ny ? k : 0;
    unsigned cb[RPT];
    bool act[RPT];      // row computed
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int j = j0 + slot * RPT + q;
        act[q] = j < nx - R;
        const int jl = j < nx ? j : nx - 1;
        cb[q] = ((unsigned)jl * (unsigned)ny + (unsigned)kl) * esz;
    }
    const bool full = k + V <= ny - R;      // every point of the vector is interior
    // x halo: rows j0-R .. j0-1 and j0+TXR .. j0+TXR+R-1
    unsigned xh_off[HX];
    int xh_lds[HX];
#pragma unroll
    for (int h = 0; h < HX; h++) {
        const int hr = slot + SLOTS * h;
        const int jr = hr < R ? j0 - R + hr : j0 + TXR + (hr - R);
        const int jl = jr < nx ? jr : nx - 1;
        xh_off[h] = ((unsigned)jl * (unsigned)ny + (unsigned)kl) * esz;
        const int xr = hr < 2 * R ? (hr < R ? hr : hr + TXR) : -1;
        xh_lds[h] = xr >= 0 ? xr * LS + PADL + lc * V : DUMP + tid * V;
    }
    // y halo: HV vectors left of k0 and right of k0+TY on every tile row
    unsigned yh_off[HY];
    int yh_lds[HY];
#pragma unroll
    for (int h = 0; h < HY; h++) {
        const int idx = tid + NT * h;
        const int row = idx / (2 * HV), c = idx % (2 * HV);
        const int jr = j0 + row;
        const int jl = jr < nx ? jr : nx - 1;
        int kr = c < HV ? k0 - PADL + V * c : k0 + TY + V * (c - HV);
        kr = kr < ny ? kr : 0;
        // kr > -V and jl >= R >= 1 with ny >= V, so the offset is not negative
        yh_off[h] = (unsigned)((int)((unsigned)jl * (unsigned)ny) + kr) * esz;
        yh_lds[h] = idx < NYV ? (R + row) * LS + (c < HV ? V * c : PADL + TY + V * (c - HV))
                              : DUMP + tid * V;
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
        if (b.hi[1] >= j0 && b.lo[1] < j0 + TXR && b.hi[2] >= k0 && b.lo[2] < k0 + TY)
            smask |= 1ull << s;
    }

    // A clean tile touches no Neumann face and no source box, and each of
    // its lanes holds a vector that is wholly interior or wholly outside.
    // Its march issues the same memory instructions on every plane (loads
    // that are not needed and stores of lanes outside go to kOOB, which the
    // descriptor range check turns into zeros / dropped writes), so the
    // compiler's wait counts only ever wait for the data a plane consumes.
    // A Dirichlet face only zeroes the interior plane next to it (its
    // phase mirrors nothing), so a clean tile applies the Dirichlet faces
    // it touches itself: zeros stored where i, j or k is that plane (the
    // reference's zeroing, wave.c:304-480).  The z planes are tile-uniform
    // (zpb, zpa); the x rows and y columns (dxy) take extra stores in a
    // branch: per-element selects of the stored vectors took the kernel
    // above 256 VGPRs, one wavefront per SIMD.
    const bool lanes_whole = ((ny - 2 * R) % V == 0) || (k0 + TY <= ny - R);
    bool nm_xy = false;     // a Neumann x or y face within 2R
    if ((bc >> 8 & 3) == 2 && k0 <= r2)
        nm_xy = true;
    if ((bc >> 10 & 3) == 2 && k0 + TY - 1 >= ny - r2 - 1)
        nm_xy = true;
    if ((bc >> 4 & 3) == 2 && j0 <= r2)
        nm_xy = true;
    if ((bc >> 6 & 3) == 2 && j0 + TXR - 1 >= nx - r2 - 1)
        nm_xy = true;
    const bool dxy = ((bc >> 4 & 3) == 1 && j0 <= r2) ||
                     ((bc >> 6 & 3) == 1 && j0 + TXR - 1 >= nx - r2 - 1) ||
                     ((bc >> 8 & 3) == 1 && k0 <= r2) ||
                     ((bc >> 10 & 3) == 1 && k0 + TY - 1 >= ny - r2 - 1);
    const int zpb = (bc & 3) == 1 ? R : -1;
    const int zpa = (bc >> 2 & 3) == 1 ? nz - R - 1 : -1;
    const bool clean = !nm_xy && smask == 0 && !((bc & 3) == 2 && z0 <= r2) &&
                       !((bc >> 2 & 3) == 2 && z1 - 1 >= nz - r2 - 1) && lanes_whole;

    // The march is instantiated twice, for clean and other tiles, with
    // everything it keeps in registers declared inside, so the two share
    // no live values.
    auto march = [&](auto clean_tag) {
    constexpr bool CLEAN = decltype(clean_tag)::value;
    // prologue: z queue planes zb .. zb+2R-1 (plane offset in the VGPR
    // offset: range checked against the descriptor end)
    vreal qz[RPT][2 * R + 1];
#pragma unroll
    for (int m = 0; m < 2 * R; m++)
#pragma unroll
        for (int q = 0; q < RPT; q++)
            qz[q][m + 1] = vload<V>(r_uc, cb[q] + (unsigned)m * PB, 0u);
    // One register set per prefetched quantity: the centre plane and its
    // halos for plane i+1 are loaded right after plane i is written to LDS;
    // u_prev, C and the damping term for plane i+1 right after plane i's
    // update has consumed them.
    vreal nc[RPT], np[RPT], nq[RPT], ne[RPT], nhx[HX], nhy[HY];
    const unsigned long long dbits = damp_bits(T, col, z0);
    auto load_centre = [&](int I) {
        const unsigned sp = (unsigned)(I - zb) * PB;   // plane I
        const unsigned vc = (unsigned)(I + R - zb) * PB;
#pragma unroll
        for (int q = 0; q < RPT; q++)
            nc[q] = vload<V>(r_uc, cb[q] + vc, 0u);
#pragma unroll
        for (int h = 0; h < HX; h++)
            nhx[h] = vload<V>(r_uc, xh_off[h], sp);
#pragma unroll
        for (int h = 0; h < HY; h++)
            nhy[h] = vload<V>(r_uc, yh_off[h], sp);
    };
    auto load_update = [&](int I) {
        const unsigned sp = (unsigned)(I - zb) * PB;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            np[q] = vload<V>(r_up, cb[q], sp);
            nq[q] = vload<V>(r_q, cb[q], sp);
        }
        if ((dbits >> (I - z0)) & 1) {
#pragma unroll
            for (int q = 0; q < RPT; q++)
                ne[q] = vload<V>(r_e, cb[q], sp);
        }
    };
    // Clean-tile loads: always issued; the prefetch past the last plane
    // and the damping term of undamped planes read kOOB (no traffic).
    auto load_centre_clean = [&](int I) {
        const bool live = I < z1;
        const unsigned sp = live ? (unsigned)(I - zb) * PB : 0u;
        const unsigned vc = (unsigned)(I + R - zb) * PB;
#pragma unroll
        for (int q = 0; q < RPT; q++)
            nc[q] = vload<V>(r_uc, live ? cb[q] + vc : kOOB, 0u);
#pragma unroll
        for (int h = 0; h < HX; h++)
            nhx[h] = vload<V>(r_uc, live ? xh_off[h] : kOOB, sp);
#pragma unroll
        for (int h = 0; h < HY; h++)
            nhy[h] = vload<V>(r_uc, live ? yh_off[h] : kOOB, sp);
    };
    auto load_update_clean = [&](int I) {
        const bool live = I < z1;
        const bool damp = live && ((dbits >> (I - z0)) & 1);
        const unsigned sp = live ? (unsigned)(I - zb) * PB : 0u;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            np[q] = vload<V>(r_up, live ? cb[q] : kOOB, sp);
            nq[q] = vload<V>(r_q, live ? cb[q] : kOOB, sp);
            ne[q] = vload<V>(r_e, damp ? cb[q] : kOOB, sp);
        }
    };
    // The scheduler barriers keep the prologue's loads in the loop's order
    // (centre and halos before the update operands), so the wait counts
    // merged at the loop header stay exact.
    if constexpr (CLEAN) {
        load_centre_clean(z0);
        __builtin_amdgcn_sched_barrier(0);
        load_update_clean(z0);
    } else {
        load_centre(z0);
        __builtin_amdgcn_sched_barrier(0);
        load_update(z0);
    }
    __builtin_amdgcn_sched_barrier(0);

    // store offsets of clean tiles: rows past the interior and lanes
    // outside it write nothing
    unsigned st_off[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++)
        st_off[q] = (act[q] && full) ? cb[q] : kOOB;
    for (int i = z0; i < z1; i++) {
        real_t *lb = &lds[(i - z0) & 1][0];
        const bool damped = (dbits >> (i - z0)) & 1;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
#pragma unroll
            for (int m = 0; m < 2 * R; m++)
                qz[q][m] = qz[q][m + 1];
            qz[q][2 * R] = nc[q];
            *(vreal *)&lb[(R + slot * RPT + q) * LS + PADL + lc * V] = qz[q][R];
        }
#pragma unroll
        for (int h = 0; h < HX; h++)
            *(vreal *)&lb[xh_lds[h]] = nhx[h];
#pragma unroll
        for (int h = 0; h < HY; h++)
            *(vreal *)&lb[yh_lds[h]] = nhy[h];
        const bool more = i + 1 < z1;
        if constexpr (CLEAN)
            load_centre_clean(i + 1);
        else if (more)
            load_centre(i + 1);
        __syncthreads();

        const bool bc_here =
            !CLEAN && (bc_xy || (bz_b && i <= r2) || (bz_a && i >= nz - r2 - 1));
        const bool special = !CLEAN && ((smask != 0) | bc_here);
        const unsigned so = (unsigned)(i - zb) * PB;
        vreal out[RPT];
        // x neighbours of the lane's RPT adjacent rows: rows -R .. RPT-1+R
        // around its first row, the own rows from registers
        // (one row per lane: read inside the stencil loop instead, which
        // keeps fewer vectors live)
        const real_t *lrow0 = lb + (R + slot * RPT) * LS + lc * V + PADL;
        vreal X[RPT + 2 * R];
        if constexpr (RPT > 1) {
#pragma unroll
            for (int d = -R; d < RPT + R; d++)
                X[d + R] = (d >= 0 && d < RPT) ? qz[d][R] : *(const vreal *)&lrow0[d * LS];
        }
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int row = R + slot * RPT + q;
            const real_t *lrow = lb + row * LS + lc * V;   // y window start
            // y window: points k-PADL .. k+V+PADL-1
            real_t win[V + 2 * PADL];
#pragma unroll
            for (int h = 0; h < HV; h++) {
                const vreal a = *(const vreal *)&lrow[V * h];
                const vreal b = *(const vreal *)&lrow[PADL + V + V * h];
#pragma unroll
                for (int e = 0; e < V; e++) {
                    win[V * h + e] = a[e];
                    win[PADL + V + V * h + e] = b[e];
                }
            }
#pragma unroll
            for (int e = 0; e < V; e++)
                win[PADL + e] = qz[q][R][e];
            real_t acc_y[V], acc_x[V], acc_z[V];
#pragma unroll
            for (int e = 0; e < V; e++) {
                const real_t c0 = A.c2[0] * qz[q][R][e];
                acc_y[e] = c0;
                acc_x[e] = c0;
                acc_z[e] = c0;
            }
#pragma unroll
            for (int m = 1; m <= R; m++) {
                vreal xp, xm;
                if constexpr (RPT > 1) {
                    xp = X[q + m + R];
                    xm = X[q - m + R];
                } else {
                    xp = *(const vreal *)&lrow0[m * LS];
                    xm = *(const vreal *)&lrow0[-m * LS];
                }
#pragma unroll
                for (int e = 0; e < V; e++) {
                    acc_y[e] += A.c2[m] * (win[PADL + e + m] + win[PADL + e - m]);
                    acc_x[e] += A.c2[m] * (xp[e] + xm[e]);
                    acc_z[e] += A.c2[m] * (qz[q][R + m][e] + qz[q][R - m][e]);
                }
            }
            real_t lap[V];
            if constexpr (D64) {
#pragma unroll
                for (int e = 0; e < V; e++)
                    lap[e] = (real_t)((double)acc_y[e] * A.ry) +
                             (real_t)((double)acc_x[e] * A.rx) +
                             (real_t)((double)acc_z[e] * A.rz);
            } else {
            // Fixed-divisor quotients for every point; the IEEE quotients
            // replace them where a numerator is outside the safe range (or
            // zero next to a non-zero one), decided once per wavefront.
            bool ok = true;
#pragma unroll
            for (int e = 0; e < V; e++) {
                const real_t ay = acc_y[e], ax = acc_x[e], az = acc_z[e];
                const real_t mx = fmax(fmax(fabs(ay), fabs(ax)), fabs(az));
                const real_t mn = fmin(fmin(fabs(ay), fabs(ax)), fabs(az));
                ok &= ((mn >= A.div_lo) & (mx <= A.div_hi)) | (mx == (real_t)0);
                lap[e] = fixed_div(ay, A.dvy) + fixed_div(ax, A.dvx) + fixed_div(az, A.dvz);
            }
            if (!__all(ok & (A.fastdiv != 0))) {
#pragma unroll
                for (int e = 0; e < V; e++) {
                    const real_t ay = acc_y[e], ax = acc_x[e], az = acc_z[e];
                    const real_t mx = fmax(fmax(fabs(ay), fabs(ax)), fabs(az));
                    const real_t mn = fmin(fmin(fabs(ay), fabs(ax)), fabs(az));
                    const bool fe = A.fastdiv &
                                    (((mn >= A.div_lo) & (mx <= A.div_hi)) | (mx == (real_t)0));
                    if (!fe)
                        lap[e] = ay / A.dy2 + ax / A.dx2 + az / A.dz2;
                }
            }
            }
#pragma unroll
            for (int e = 0; e < V; e++) {
                real_t value = 0;
                value += lap[e];
#if defined(SW_DOUBLE)
                out[q][e] = damped ? leapfrog_damped(value, nq[q][e], ne[q][e], qz[q][R][e],
                                                     np[q][e])
                                   : leapfrog_plain(value, nq[q][e], qz[q][R][e], np[q][e]);
#else
                out[q][e] = damped ? leapfrog_damped_rcp(value, nq[q][e], ne[q][e],
                                                         qz[q][R][e], np[q][e])
                                   : leapfrog_plain(value, nq[q][e], qz[q][R][e], np[q][e]);
#endif
            }
            if (!CLEAN && special && act[q]) {
                // tiles touching a source box or a boundary face
                const int j = j0 + slot * RPT + q;
#pragma unroll
                for (int e = 0; e < V; e++) {
                    const int kk = k + e;
                    if (kk >= ny - R)
                        continue;
                    const long long p = (long long)i * nx * ny + (long long)j * ny + kk;
                    real_t v = out[q][e];
                    if (smask)
                        v = add_sources_masked(*Pd, n, i, j, kk, Pd->q[p], Pd->e[p], v,
                                               smask);
                    if (bc_here)
                        store_with_bc(*Pd, A.un, p, i, j, kk, v);
                    else
                        sstore(v, r_un, cb[q] + e * esz, so);
                }
            }
        }
        // stores first, then the next plane's update operands, so that
        // waiting for those never waits for the stores
        if constexpr (CLEAN) {
            if (i == zpb || i == zpa) {
#pragma unroll
                for (int q = 0; q < RPT; q++)
                    out[q] = (vreal)(real_t)0;
            }
#pragma unroll
            for (int q = 0; q < RPT; q++)
                vstore(out[q], r_un, st_off[q], so);
            if (dxy) {
                // the Dirichlet x rows and y columns the tile holds: zeros
                // over the values just stored (stores of one lane to one
                // address complete in program order)
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    const int j = j0 + slot * RPT + q;
                    if (((bc >> 4 & 3) == 1 && j == R) || ((bc >> 6 & 3) == 1 && j == nx - R - 1))
                        vstore((vreal)(real_t)0, r_un, st_off[q], so);
                    if (act[q] && (bc >> 8 & 3) == 1 && k <= R && R < k + V)
                        sstore((real_t)0, r_un, cb[q] + (unsigned)(R - k) * esz, so);
                    if (act[q] && (bc >> 10 & 3) == 1 && k <= ny - R - 1 && ny - R - 1 < k + V)
                        sstore((real_t)0, r_un, cb[q] + (unsigned)(ny - R - 1 - k) * esz, so);
                }
            }
            load_update_clean(i + 1);
        } else {
            if (!special) {
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    if (!act[q])
                        continue;
                    if (full) {
                        vstore(out[q], r_un, cb[q], so);
                    } else {
#pragma unroll
                        for (int e = 0; e < V; e++)
                            if (k + e < ny - R)
                                sstore(out[q][e], r_un, cb[q] + e * esz, so);
                    }
                }
            }
            if (more)
                load_update(i + 1);
        }
    }
    };
    if (clean)
        march(std::true_type{});
    else
        march(std::false_type{});
}

#endif  // SW_DIM == 3

}  // namespace sw
