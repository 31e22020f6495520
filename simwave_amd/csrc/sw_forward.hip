// Host orchestration and C-ABI of the HIP forward operator for gfx950.
//
// This translation unit is compiled once per reference variant:
//   -DSW_DIM=2|3  -DSW_VD=0|1  [-DSW_DOUBLE]
// and each build exports the reference FFI symbol `forward` with that
// variant's exact signature (simwave/kernel/backend/middleware.py:149-154
// binds it by that name), plus suffixed names and a device-pointer entry.
//
// Host-side emulation of the reference time loop
// (simwave/kernel/backend/c_code/forward/constant_density/3d/wave.c:110-618):
// the same prev/cur/next slot indices are computed on the host for every step,
// and each logical slot of the caller's `u` array is backed by a device buffer
// while it is in use.  A slot is uploaded before its first use and downloaded
// after its last use, so a saving_stride run of any length needs only a few
// slots of HBM, and the even-stride content exchange becomes a swap of the two
// device pointers.
#include "sw_kernels.hpp"
#include "sw_step3d.hpp"
#include "sw_adjoint.hpp"
#include "sw_batch.hpp"

#include <hip/hip_ext.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#define SW_STR2(x) #x
#define SW_STR(x) SW_STR2(x)
#if SW_VD
#define SW_DENS vd
#else
#define SW_DENS cd
#endif
#if defined(SW_DOUBLE)
#define SW_PREC f64
#else
#define SW_PREC f32
#endif
#define SW_CAT5(a, b, c, d, e) a##b##c##d##e
#define SW_NAME5(a, b, c, d, e) SW_CAT5(a, b, c, d, e)
#define SW_VARIANT_STR SW_STR(SW_NAME5(SW_DIM, d_, SW_DENS, _, SW_PREC))
#define SW_SUFFIX(base) SW_NAME5(base, SW_DIM, d_, SW_DENS, SW_NAME5(_, SW_PREC, , , ))

#define SW_EXPORT extern "C" __attribute__((visibility("default")))

// Device memory of the allocation-free device entry
// (simwave_hip_forward_device_noalloc_<variant>): the per-point coefficient
// arrays, damping mask words, parameter copy and source/receiver tables of
// one run are carved from this array instead of allocated.  Each device
// holds its own copy for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Timing of one step-kernel launch: events that the runtime sets at the
// kernel's own start and end (hipExtLaunchKernelGGL), so a timed run puts
// no marker packets between the kernels and reads what a profiler reads
// (marker pairs around a 30 us launch added about 6 us).  Both null: an
// ordinary launch.
struct LaunchTimer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

template <typename F, typename... Args>
void launch_timed(F kernel, dim3 grid, dim3 block, hipStream_t st, LaunchTimer tm,
                  Args... args)
{
    if (tm.e0)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, st, tm.e0, tm.e1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
}

using sw::Box;
using sw::Params;

thread_local std::string g_error;
thread_local double g_timings[8];

struct Failure {
    std::string msg;
};

[[noreturn]] void fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Failure{buf};
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        fail("%s: %s", what, hipGetErrorString(e));
}

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

constexpr unsigned kFlagTimeSteps = 1u;    // accumulate per-launch step time
constexpr unsigned kFlagGeneric = 2u;      // force the generic kernels
// receiver gather on a second stream of the library's own, joined by event
// waits (large grids only).  Opt-in: a caller that drives many streams from
// many threads keeps each shot on its one stream instead, since event waits
// between streams that share a hardware queue can then order into a cycle.
constexpr unsigned kFlagOverlapReceivers = 4u;
// persistent plans of stride-0 shots: capture kGraphSteps steps once in a
// hipGraph and replay it (the step number comes from a device counter)
constexpr unsigned kFlagGraph = 8u;
constexpr size_t kGraphSteps = 48;   // a multiple of the 3-slot cycle
// adjoint entry: accumulate the zero-lag cross-correlation image of the
// forward and adjoint wavefields instead of the velocity gradient
constexpr unsigned kAdjImage = 1u;
// adjoint entry: rec_adjoint holds observed data; the adjoint source is the
// residual (shot record - observed), formed on the device after the forward
constexpr unsigned kAdjResidual = 2u;
// adjoint entry: time every reverse update kernel with HIP events
// (simwave_hip_last_timings: [4] their summed seconds, [5] their count)
constexpr unsigned kAdjTimeUpdate = 8u;
// reverse steps whose gradient terms are summed in the working precision
// before they are added to the float64 gradient
constexpr size_t kGradFlush = 16;
constexpr int kMaxFusedSources = 64;
// cooperative receiver kernel: receivers per block and points per round
// (LDS = kRecChunk * (kRecPerBlock + 1) values)
constexpr int kRecPerBlock = sizeof(real_t) == 4 ? 16 : 8;
constexpr int kRecChunk = 768;
// grids from this many points on run the receiver gather on a second stream
constexpr long long kOverlapMinPoints = 1ll << 22;

// The receiver-overlap lease (sw_common.cpp): one run or plan per process may
// put its receiver gather on a second stream.  `required`: a caller that asked
// for the overlap gets an error while another holds it; otherwise the run
// quietly keeps its receivers on its one stream.
extern "C" int simwave_hip_overlap_acquire(void);
extern "C" void simwave_hip_overlap_release(void);

struct OverlapLease {
    bool held = false;
    OverlapLease() = default;
    OverlapLease(const OverlapLease &) = delete;
    OverlapLease &operator=(const OverlapLease &) = delete;
    bool acquire(bool required)
    {
        if (held)
            return true;
        held = simwave_hip_overlap_acquire() != 0;
        if (!held && required)
            fail("receiver overlap (flag 4) is already in use by another run or plan in "
                 "this process; run without it, or after that run or plan has ended");
        return held;
    }
    ~OverlapLease()
    {
        if (held)
            simwave_hip_overlap_release();
    }
};

// All arguments of one forward call, normalised across variants.
struct Args {
    real_t *u;
    const real_t *vel, *rho, *damp, *wavelet;
    size_t wavelet_size, wavelet_count;
    const real_t *c2, *c1;
    const size_t *bc;
    const size_t *sint;
    size_t sint_size;
    const real_t *svals;
    size_t svals_size;
    const size_t *soff;
    const size_t *rint;
    size_t rint_size;
    const real_t *rvals;
    size_t rvals_size;
    const size_t *roff;
    real_t *receivers;
    size_t nsrc, nrec;
    size_t nz, nx, ny;
    real_t dz, dx, dy;
    size_t stride;
    real_t dt;
    size_t n0, n1;
    size_t space_order, num_snapshots;
};

// Slot indices of one step and whether the even-stride exchange follows it.
struct StepSlots {
    size_t prev, cur, next;
    bool swap_after;
    size_t swap_a, swap_b;
};

// Replays the reference's slot bookkeeping (wave.c:112-124, 569-618).
std::vector<StepSlots> plan_slots(const Args &a)
{
    std::vector<StepSlots> out;
    if (a.n1 < a.n0)
        return out;
    out.reserve(a.n1 - a.n0 + 1);
    size_t prev = 0, cur = 1, next = 2;
    for (size_t n = a.n0; n <= a.n1; n++) {
        if (a.stride == 0) {
            prev = (n - 1) % 3; cur = n % 3; next = (n + 1) % 3;
        } else if (a.stride == 1) {
            prev = n - 1; cur = n; next = n + 1;
        }
        StepSlots s{prev, cur, next, false, 0, 0};
        if (a.stride > 1) {
            if (n % a.stride == 1) {
                prev = cur;
                cur += 1;
                next += 1;
                if (a.stride % 2 == 0 && n < a.n1) {
                    std::swap(cur, next);
                    s.swap_after = true;
                    s.swap_a = cur;
                    s.swap_b = next;
                }
            } else {
                prev = cur;
                cur = next;
                next = prev;
            }
        }
        out.push_back(s);
    }
    return out;
}

// Converts one interval/value/offset description (source.py:123-159 layout)
// into device boxes, validating every index against the grid.
std::vector<Box> make_boxes(const char *what, size_t count, const size_t *iv,
                            size_t iv_size, const size_t *off, size_t vsize,
                            const size_t dims[3])
{
    const int D = SW_DIM;
    std::vector<Box> boxes(count);
    if (count == 0)
        return boxes;
    if (!iv || !off)
        fail("%s: interval/offset arrays are NULL", what);
    if (iv_size < count * 2 * D)
        fail("%s: interval array holds %zu entries, %zu needed", what, iv_size,
             count * 2 * D);
    for (size_t s = 0; s < count; s++) {
        Box b{};
        size_t nvals = 0;
        for (int a = 0; a < D; a++) {
            const size_t lo = iv[s * 2 * D + 2 * a], hi = iv[s * 2 * D + 2 * a + 1];
            if (lo > hi || hi >= dims[a])
                fail("%s %zu: axis %d interval [%zu, %zu] outside [0, %zu)",
                     what, s, a, lo, hi, dims[a]);
            b.lo[a] = (int)lo;
            b.hi[a] = (int)hi;
            nvals += hi - lo + 1;
        }
        if (off[s] + nvals > vsize || off[s] > (size_t)0x7fffffff)
            fail("%s %zu: weights [%zu, %zu) exceed the %zu values", what, s,
                 off[s], off[s] + nvals, vsize);
        b.woff = (int)off[s];
        boxes[s] = b;
    }
    return boxes;
}

// Stream-ordered device allocations released when the object dies.  With a
// base pointer it instead hands out consecutive 256-byte aligned pieces of
// that caller-owned region (no allocation, nothing to free) and fails once
// `cap` bytes are used.
struct DeviceArena {
    hipStream_t stream;
    std::vector<void *> ptrs;
    char *base = nullptr;
    size_t cap = 0, used = 0;
    explicit DeviceArena(hipStream_t s) : stream(s) {}
    DeviceArena(hipStream_t s, void *region, size_t bytes)
        : stream(s), base(static_cast<char *>(region)), cap(bytes)
    {
    }
    template <typename T>
    T *alloc(size_t count)
    {
        if (count == 0)
            return nullptr;
        if (base) {
            const size_t bytes = (count * sizeof(T) + 255) / 256 * 256;
            if (bytes > cap - used)
                fail("the run needs more than the %zu bytes of the library's static "
                     "scratch (%zu used before a request of %zu bytes); use the "
                     "allocating entry point for this grid", cap, used, count * sizeof(T));
            void *p = base + used;
            used += bytes;
            return static_cast<T *>(p);
        }
        void *p = nullptr;
        hip_check(hipMallocAsync(&p, count * sizeof(T), stream), "hipMallocAsync");
        ptrs.push_back(p);
        return static_cast<T *>(p);
    }
    void release(void *p)
    {
        auto it = std::find(ptrs.begin(), ptrs.end(), p);
        if (it != ptrs.end()) {
            (void)hipFreeAsync(p, stream);
            ptrs.erase(it);
        }
    }
    ~DeviceArena()
    {
        for (void *p : ptrs)
            (void)hipFreeAsync(p, stream);
        (void)hipStreamSynchronize(stream);
    }
};

// Freed stream-ordered allocations stay in the device's default pool up to
// this many bytes, so repeated runs of one grid reuse their memory without
// new mappings; beyond it the pool returns memory to the device at the next
// synchronisation, where other allocators (PyTorch's) can use it.
constexpr uint64_t kPoolKeepBytes = 8ull << 30;

bool default_pool(hipMemPool_t &pool)
{
    int dev = 0;
    return hipGetDevice(&dev) == hipSuccess &&
           hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess;
}

void keep_pool_memory()
{
    static std::once_flag once;
    std::call_once(once, [] {
        hipMemPool_t pool;
        if (default_pool(pool)) {
            uint64_t thr = kPoolKeepBytes;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
    });
}

// Everything the time loop needs on the device, independent of the slots.
struct Prepared {
    Params P{};
    const Box *rbox = nullptr;
    const real_t *rw = nullptr;
    bool fused = false;
    bool halo_sources = false;   // some source box reaches the halo
    int r = 0;
    long long npts = 0;
#if SW_DIM == 3
    bool zmarch = false;         // use k_step3d_z
    int zmarch_zlen = 16;
    sw::LapDiv lapdiv{};         // its fixed-divisor quotients
    bool tiled = false;          // use k_step3d_vec
    void (*tiled_launch)(const struct Prepared &, const real_t *, const real_t *,
                         real_t *, int, hipStream_t, LaunchTimer) = nullptr;
    const Params *Pd = nullptr;   // device copy of P for the tiled kernel
    const real_t *qc = nullptr;   // (dt2/s)/den per point for the tiled kernel
    bool fastdiv = false;
    bool div64 = false;          // float64-reciprocal quotients in the tiled kernel
    sw::FixedDiv dvz{}, dvx{}, dvy{};
    sw::Tiling tiling{};
    int tiled_blocks = 0;
#endif
};

#if SW_DIM == 3
void setup_tiling(Prepared &pr, DeviceArena &arena, hipStream_t st);
bool prefer_zmarch(const Prepared &pr);
void set_lap_div(Prepared &pr);
#else
bool zmarch_2d(int R);
#endif

bool env_flag(const char *name)
{
    const char *v = getenv(name);
    return v && *v && strcmp(v, "0") != 0;
}

int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return (v && *v) ? atoi(v) : dflt;
}

// Argument checks that need no device; run before any HIP call so a bad
// call reports its own error on any host.
void validate(const Args &a)
{
    const size_t so = a.space_order;
    if (so < 2 || so > 20 || so % 2)
        fail("space_order %zu: even orders 2..20 are supported", so);
    const int r = (int)(so / 2);
    const size_t dims[3] = {a.nz, a.nx, SW_DIM == 3 ? a.ny : 1};
    for (int d = 0; d < SW_DIM; d++)
        if (dims[d] < (size_t)(2 * r) || dims[d] > (size_t)0x7fffffff)
            fail("axis %d has %zu points; the stencil radius %d needs >= %d",
                 d, dims[d], r, 2 * r);
    if (a.n0 < 1 && a.n1 >= a.n0)
        fail("begin_timestep must be >= 1 (got %zu)", a.n0);
    if (a.n1 >= a.n0) {
        if (a.wavelet_count > 1 && a.wavelet_count != a.nsrc)
            fail("wavelet_count %zu != num_sources %zu", a.wavelet_count, a.nsrc);
        if (a.n1 > a.wavelet_size)
            fail("end_timestep %zu exceeds wavelet_size %zu (the shot record "
                 "has wavelet_size rows)", a.n1, a.wavelet_size);
    }
    if (!a.u || !a.vel || !a.damp || !a.c2 || !a.bc)
        fail("NULL array argument");
#if SW_VD
    if (!a.rho || !a.c1)
        fail("NULL density or first-order coefficients");
#endif
    if (a.nrec && !a.receivers)
        fail("NULL receivers array");
    if (a.nsrc && !a.wavelet)
        fail("NULL wavelet array");
}

#if SW_VD
template <int R>
void rho_diff_launch(const Params &P, real_t *dz, real_t *dm, real_t *dc, long long cnt,
                     hipStream_t st)
{
    const unsigned blocks = (unsigned)std::min<long long>((cnt + 255) / 256, 16384);
    hipLaunchKernelGGL(sw::k_rho_diff<R>, dim3(blocks), dim3(256), 0, st, P, dz, dm, dc);
}

// The density derivative sums the z-marching kernels read from radius
// sw::kVdPre on (k_rho_diff), once per run.
void rho_diff(Prepared &pr, DeviceArena &arena, hipStream_t st)
{
    Params &P = pr.P;
    const int r = pr.r;
    real_t *dz = arena.alloc<real_t>(pr.npts);
    real_t *dm = SW_DIM == 3 ? arena.alloc<real_t>(pr.npts) : nullptr;
    real_t *dc = arena.alloc<real_t>(pr.npts);
    long long cnt = (long long)(P.nz - 2 * r) * (P.nx - 2 * r);
    if (SW_DIM == 3)
        cnt *= P.ny - 2 * r;
    switch (r) {
    case 6: rho_diff_launch<6>(P, dz, dm, dc, cnt, st); break;
    case 7: rho_diff_launch<7>(P, dz, dm, dc, cnt, st); break;
    case 8: rho_diff_launch<8>(P, dz, dm, dc, cnt, st); break;
    case 9: rho_diff_launch<9>(P, dz, dm, dc, cnt, st); break;
    case 10: rho_diff_launch<10>(P, dz, dm, dc, cnt, st); break;
    default: fail("density sums for radius %d", r);
    }
    hip_check(hipGetLastError(), "k_rho_diff launch");
    P.drho[0] = dz;
    P.drho[1] = dm;
    P.drho[2] = dc;
}
#endif

// Uploads the run constants (the arguments passed validate()).
// `dev_models` tells whether vel/rho/damp are device pointers (device
// entry) or host pointers.
Prepared prepare(const Args &a, bool dev_models, unsigned flags,
                 DeviceArena &arena, hipStream_t st)
{
    Prepared pr;
    const int r = (int)(a.space_order / 2);
    pr.r = r;
    const size_t dims[3] = {a.nz, a.nx, SW_DIM == 3 ? a.ny : 1};

    const long long npts = (long long)dims[0] * (long long)dims[1] * (long long)dims[2];
    pr.npts = npts;
    Params &P = pr.P;
    P.nz = (int)a.nz;
    P.nx = (int)a.nx;
    P.ny = SW_DIM == 3 ? (int)a.ny : 1;
    P.r = r;
#if SW_DIM == 3
    P.sz = (long long)a.nx * (long long)a.ny;
    P.sx = (long long)a.ny;
#else
    P.sz = (long long)a.nx;
    P.sx = 1;
#endif
    for (int f = 0; f < 2 * SW_DIM; f++) {
        const size_t c = a.bc[f];
        P.bc[f] = (c == 1 || c == 2) ? (int)c : 0;
    }
    P.dz2 = a.dz * a.dz;
    P.dx2 = a.dx * a.dx;
    P.dy2 = a.dy * a.dy;
    for (int m = 0; m <= r; m++) {
        P.c2[m] = a.c2[m];
        P.c1[m] = SW_VD ? a.c1[m] : (real_t)0;
    }

    // per-point coefficients: upload v and damp into the q/e buffers and
    // convert in place (k_prepare reads a point before writing it)
    real_t *q = arena.alloc<real_t>(npts);
    real_t *e = arena.alloc<real_t>(npts);
    const hipMemcpyKind mk = dev_models ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    hip_check(hipMemcpyAsync(q, a.vel, npts * sizeof(real_t), mk, st), "copy velocity");
    hip_check(hipMemcpyAsync(e, a.damp, npts * sizeof(real_t), mk, st), "copy damp");
    {
        const real_t dt = a.dt, dt2 = a.dt * a.dt;
        const int blocks = (int)std::min<long long>((npts + 255) / 256, 8192);
        hipLaunchKernelGGL(sw::k_prepare, dim3(blocks), dim3(256), 0, st, q, e, q, e,
                           npts, dt, dt2);
        hip_check(hipGetLastError(), "k_prepare launch");
    }
    P.q = q;
    P.e = e;
#if SW_VD
    if (dev_models) {
        P.rho = a.rho;
    } else {
        real_t *rho = arena.alloc<real_t>(npts);
        hip_check(hipMemcpyAsync(rho, a.rho, npts * sizeof(real_t),
                                 hipMemcpyHostToDevice, st), "copy density");
        P.rho = rho;
    }
#else
    P.rho = nullptr;
#endif

    // sources
    std::vector<Box> sb = make_boxes("source", a.nsrc, a.sint, a.sint_size,
                                     a.soff, a.svals_size, dims);
    std::vector<Box> rb = make_boxes("receiver", a.nrec, a.rint, a.rint_size,
                                     a.roff, a.rvals_size, dims);
    P.nsrc = (int)a.nsrc;
    P.wav_multi = a.wavelet_count > 1;
    if (a.nsrc) {
        Box *d_sb = arena.alloc<Box>(a.nsrc);
        real_t *d_sw = arena.alloc<real_t>(std::max<size_t>(a.svals_size, 1));
        const size_t wlen = a.wavelet_size * std::max<size_t>(a.wavelet_count, 1);
        real_t *d_wav = arena.alloc<real_t>(std::max<size_t>(wlen, 1));
        hip_check(hipMemcpyAsync(d_sb, sb.data(), a.nsrc * sizeof(Box),
                                 hipMemcpyHostToDevice, st), "copy source boxes");
        if (a.svals_size)
            hip_check(hipMemcpyAsync(d_sw, a.svals, a.svals_size * sizeof(real_t),
                                     hipMemcpyHostToDevice, st), "copy source weights");
        if (wlen)
            hip_check(hipMemcpyAsync(d_wav, a.wavelet, wlen * sizeof(real_t),
                                     hipMemcpyHostToDevice, st), "copy wavelet");
        P.sbox = d_sb;
        P.sw = d_sw;
        P.wavelet = d_wav;
        for (const Box &b : sb) {
            for (int d = 0; d < SW_DIM; d++)
                if (b.lo[d] < r || b.hi[d] >= (int)dims[d] - r)
                    pr.halo_sources = true;
        }
    }
    if (a.nrec) {
        Box *d_rb = arena.alloc<Box>(a.nrec);
        real_t *d_rw = arena.alloc<real_t>(std::max<size_t>(a.rvals_size, 1));
        hip_check(hipMemcpyAsync(d_rb, rb.data(), a.nrec * sizeof(Box),
                                 hipMemcpyHostToDevice, st), "copy receiver boxes");
        if (a.rvals_size)
            hip_check(hipMemcpyAsync(d_rw, a.rvals, a.rvals_size * sizeof(real_t),
                                     hipMemcpyHostToDevice, st), "copy receiver weights");
        pr.rbox = d_rb;
        pr.rw = d_rw;
    }

    // fused-kernel preconditions: interior of every axis at least r+1 thick
    // (Neumann mirrors read interior points only), few sources, and for 3D
    // variable density the middle-axis quirk stride nx equal to ny.
    bool fused = !(flags & kFlagGeneric) && !env_flag("SIMWAVE_HIP_GENERIC");
    for (int d = 0; d < SW_DIM; d++)
        if (dims[d] < (size_t)(3 * r + 1))
            fused = false;
    if (a.nsrc > (size_t)kMaxFusedSources)
        fused = false;
#if SW_DIM == 3 && SW_VD
    if (a.nx != a.ny)
        fused = false;
#endif
    pr.fused = fused;
#if SW_DIM == 3 && !SW_VD
    // the z-marching one-point-per-thread kernel where it is the faster one
    // (prefer_zmarch), when asked for, or where the tiled kernel's 32-bit
    // buffer offsets do not fit; the tiled kernel otherwise or when a tile
    // configuration is asked for (SIMWAVE_HIP_TILE)
    if (fused && !env_flag("SIMWAVE_HIP_UNTILED")) {
        const char *tile = getenv("SIMWAVE_HIP_TILE");
        const bool want_tile = tile && *tile;
        if (!env_flag("SIMWAVE_HIP_ZMARCH") && (want_tile || !prefer_zmarch(pr)))
            setup_tiling(pr, arena, st);
        if (!pr.tiled) {
            pr.zmarch = true;
            pr.zmarch_zlen = std::max(1, std::min(env_int("SIMWAVE_HIP_ZLEN", 8), 256));
            set_lap_div(pr);
        }
    }
#elif SW_DIM == 3
    // variable density: the z-marching kernel where it is the faster one
    // (prefer_zmarch) or when asked for; the LDS-tiled k_step3d otherwise.
    // Its z chunks: 16 planes, halved (down to 4) while the launch holds
    // fewer than 2048 blocks (448^3 so8: 854 us at 16 planes, 894 at 8;
    // 176^3 so8: 65 us at 4 planes, 71 at 16)
    if (fused && !env_flag("SIMWAVE_HIP_UNTILED") &&
        (env_flag("SIMWAVE_HIP_ZMARCH") || prefer_zmarch(pr))) {
        pr.zmarch = true;
        int zl = env_int("SIMWAVE_HIP_ZLEN", 0);
        if (zl <= 0) {
            const long long cols = (long long)((P.ny - 2 * r + 63) / 64) *
                                   ((P.nx - 2 * r + sw::kZLines - 1) / sw::kZLines);
            const int iz = P.nz - 2 * r;
            zl = 16;
            while (zl > 4 && cols * ((iz + zl - 1) / zl) < 2048)
                zl /= 2;
        }
        pr.zmarch_zlen = std::max(1, std::min(zl, 256));
        set_lap_div(pr);
    }
#endif
#if SW_VD
#if SW_DIM == 3
    const bool zm = pr.zmarch;
#else
    const bool zm = fused && zmarch_2d(r);
#endif
    if (zm && r >= sw::kVdPre)
        rho_diff(pr, arena, st);
#endif
    return pr;
}

// ------------------------------------------------------------ launches

// Divisor data for sw::fixed_div: y = RN(1/b), chosen among the neighbours of
// an extended-precision estimate by the smallest exact residual |1 - b*y|,
// and the numerator magnitudes [lo, hi] for which q0 = RN(a*y) is normal,
// r = a - b*q0 is exact and nothing overflows.  false: use IEEE division.
bool make_fixed_div(real_t b, sw::FixedDiv &d)
{
    if (!(b > (real_t)0) || !std::isfinite(b))
        return false;
    int eb = 0;
    (void)std::frexp(b, &eb);
    eb -= 1;                                   // b in [2^eb, 2^(eb+1))
    if (eb < -60 || eb > 60)
        return false;
    constexpr int p = std::numeric_limits<real_t>::digits;
    constexpr int emin = std::numeric_limits<real_t>::min_exponent - 1;
    constexpr int emax = std::numeric_limits<real_t>::max_exponent - 1;
    const real_t y0 = (real_t)(1.0L / (long double)b);
    const real_t cand[3] = {std::nextafter(y0, (real_t)0), y0,
                            std::nextafter(y0, std::numeric_limits<real_t>::infinity())};
    real_t best = y0;
    double best_res = 1e300;
    for (real_t y : cand) {
        // float: b*y is exact in double; double: fma rounds the tiny residual once
        const double res = std::fabs(std::fma(-(double)b, (double)y, 1.0));
        if (res < best_res) {
            best_res = res;
            best = y;
        }
    }
    d.b = b;
    d.y = best;
    d.lo = (real_t)std::ldexp(1.0, std::max(emin + eb + 2, emin + p));
    d.hi = (real_t)std::ldexp(1.0, emax - 2 + std::min(0, eb));
    return true;
}

// Whether RN(a * RN64(1/b)), the product rounded to float64 and then to
// float32, equals the IEEE float quotient a/b for every float a: b a
// positive float in [2^-126, 2^24) whose odd significand B (b = B 2^t) has
// t <= 0.  The float64 product is within 2^-52 (relative) of a/b.  A normal
// quotient that is not a float lies at least 2^-49 from every midpoint
// between floats (a - m b is a non-zero multiple of a unit at that scale,
// as m b has more significant bits than a float holds).  A subnormal one
// lies at least 2^-48 from the midpoints (2j+1) 2^-150: with t <= 0,
// (2j+1) B 2^(t-150) is an odd multiple of 2^(t-150) and a is a multiple
// of 2^-149, so they differ.  Both separations exceed the product's error.
bool f64_quotient_exact(real_t b)
{
    if (sizeof(real_t) != 4)
        return false;
    const double bd = (double)b;
    if (!(bd >= std::ldexp(1.0, -126) && bd < std::ldexp(1.0, 24)))
        return false;
    int e = 0;
    const double m = std::frexp(bd, &e);              // bd = m 2^e, m in [0.5, 1)
    unsigned long long S = (unsigned long long)std::ldexp(m, 24);   // 24-bit significand
    int t = e - 24;
    while (S && !(S & 1ull)) {
        S >>= 1;
        t++;
    }
    return t <= 0;
}

#if SW_DIM == 3

// Fixed divisors for k_step3d_z's Laplacian quotients (IEEE ones when a
// divisor has none or SIMWAVE_HIP_IEEE_DIV is set).
void set_lap_div(Prepared &pr)
{
    sw::LapDiv &D = pr.lapdiv;
    D.ok = !env_flag("SIMWAVE_HIP_IEEE_DIV") && make_fixed_div(pr.P.dy2, D.dy) &&
           make_fixed_div(pr.P.dx2, D.dx) && make_fixed_div(pr.P.dz2, D.dz);
    D.lo = std::max({D.dy.lo, D.dx.lo, D.dz.lo});
    D.hi = std::min({D.dy.hi, D.dx.hi, D.dz.hi});
}
#endif

#if SW_DIM == 3
// Tile configurations of k_step3d_vec as (rows per wavefront WR, rows per
// thread RPT, wavefronts per block NW, and 16- or 8-byte lanes).  Radius 4
// (space order 8) carries every configuration; other radii carry 0, 4, 8,
// 17 and 18.  SIMWAVE_HIP_TILE picks one; otherwise default_tile().
// Measured per launch on one MI355X, float32, radius 4, 6-plane z chunks:
//  * spacings whose Laplacian quotients are float64 products (div64):
//    8 rows x 64 points of 8-byte lanes with 4 rows per lane are fastest,
//    1-3% ahead of the same tile of 16-byte lanes with 2 rows per lane
//    (472^3: 515 against 532 us; interior 340^3 245 against 248, 380^3 318
//    against 322, 420^3 411 against 420; 300^3 even).  The 8-byte lanes halve the registers of a z
//    queue entry, so a lane holds 4 adjacent rows and takes more of its x
//    neighbours from registers.  The 4 x 128 shape of 8-byte lanes was
//    fastest at 472^3 alone (511 us) and 2-8% behind elsewhere; shapes with
//    3 wavefronts per SIMD (133-137 VGPRs) came to 541-578 us at 472^3.
//  * other spacings (fixed-divisor quotients): the 16-byte 8 x 64 tile, the
//    8-byte one within 2% either way (300^3-380^3 interior), 5% behind at
//    260^3.
// Radii 5-10: 8 rows x 32 points of 16-byte lanes and 4-plane chunks
// (10-13% ahead of 4-wavefront tiles at 312^3-376^3; the 8-byte shapes
// 1-10% behind at 320^3 so12 and so16).  Float64: 8 rows x 32 points of
// 16-byte lanes, 8-plane chunks at radius 4 (472^3: 881 us; the 8-byte
// shapes 894-1097 us), 12-plane chunks elsewhere.
constexpr int kTileShapes = 24;
constexpr int default_tile(int R, bool div64)
{
    return R == 4 ? (sizeof(real_t) == 4 && div64 ? 17 : 10)
         : R < 4 ? 0 : (sizeof(real_t) == 4 ? 8 : 4);
}
constexpr int default_tile_zlen(int R)
{
    return sizeof(real_t) == 4 ? (R == 4 ? 6 : R > 4 ? 4 : 12) : (R == 4 ? 8 : 12);
}

// Whether the z-marching one-point-per-thread kernel beats the tiled one
// for this grid.  Its 7 wavefronts per SIMD hide latency better while the
// grid holds too few tiles to keep the tiled kernel's 2 per SIMD busy; its
// neighbour reads through the caches lose once there are plenty at radius 4.
// Measured per launch (float32, cubes): radius 1-3 at every size up to
// 466^3 (1.2-1.9x faster than any tile configuration); radius 4 up to
// ~30 M interior points (interior 260^3: 126 against 144 us; 300^3, 27 M:
// 172 against 187 us, 198 us with float64-product quotients; 320^3, 33 M:
// tiled 192 against 209 us; 384^3, 57 M: 327 against 347 us; with the
// tiled kernel's 6-plane chunks and Dirichlet faces in its clean march);
// radius 5+ never (its register queue is 2R+1 deep).
// Float64: the tiled kernel (z-marching 26% slower at 22 M, radius 4).
// Variable density: float32 at every radius and size measured, 1.2-1.8x
// faster than the LDS-tiled k_step3d (176^3-448^3, orders 2-16); float64 at
// radius <= 4 (272^3 so8: 527 vs 690 us; 268^3 so4: 346 vs 413 us), above
// which its double-precision register queues would spill.
bool prefer_zmarch(const Prepared &pr)
{
    if (SW_VD)
        return sizeof(real_t) == 4 || pr.r <= 4;
    if (sizeof(real_t) != 4)
        return false;
    const Params &P = pr.P;
    const int R = pr.r;
    const long long interior = (long long)(P.nz - 2 * R) * (P.nx - 2 * R) * (P.ny - 2 * R);
    if (R <= 3)
        return true;
    return R == 4 && interior < 30ll * 1000 * 1000;
}

template <int R, int WR, int RPT, int NW, int V, bool D64>
void tiled_launch(const Prepared &pr, const real_t *uc, const real_t *up, real_t *un,
                  int n, hipStream_t st, LaunchTimer tm)
{
    const Params &P = pr.P;
    sw::StepArgs A{};
    A.uc = uc;
    A.up = up;
    A.qc = pr.qc;
    A.e = P.e;
    A.un = un;
    A.fastdiv = pr.fastdiv ? 1 : 0;
    A.dvz = pr.dvz;
    A.dvx = pr.dvx;
    A.dvy = pr.dvy;
    A.div_lo = std::max({pr.dvz.lo, pr.dvx.lo, pr.dvy.lo});
    A.div_hi = std::min({pr.dvz.hi, pr.dvx.hi, pr.dvy.hi});
    A.ry = 1.0 / (double)P.dy2;
    A.rx = 1.0 / (double)P.dx2;
    A.rz = 1.0 / (double)P.dz2;
    A.nz = P.nz;
    A.nx = P.nx;
    A.ny = P.ny;
    A.pbytes = (unsigned)((long long)P.nx * P.ny * (long long)sizeof(real_t));
    A.bc = 0;
    for (int f = 0; f < 6; f++)
        A.bc |= (P.bc[f] & 3) << (2 * f);
    A.dz2 = P.dz2;
    A.dx2 = P.dx2;
    A.dy2 = P.dy2;
    for (int m = 0; m <= sw::kMaxR; m++)
        A.c2[m] = P.c2[m];
    A.nsrc = P.nsrc;
    A.sbox = P.sbox;
    A.T = pr.tiling;
    launch_timed(sw::k_step3d_vec<R, WR, RPT, NW, V, D64>, dim3(pr.tiled_blocks), dim3(64, NW),
                 st, tm, A, pr.Pd, n);
}

using TiledFn = void (*)(const Prepared &, const real_t *, const real_t *, real_t *, int,
                         hipStream_t, LaunchTimer);
struct TiledKernel {
    TiledFn launch;
    TiledFn launch64;   // float64-reciprocal quotients (float32 builds)
    int ty, txr;
};

template <int R, int WR, int RPT, int NW, int V = sw::kVec16>
TiledKernel tiled_config()
{
    using S = sw::TileShape<R, WR, RPT, NW, V>;
    constexpr bool f32 = sizeof(real_t) == 4;
    return TiledKernel{&tiled_launch<R, WR, RPT, NW, V, false>,
                       &tiled_launch<R, WR, RPT, NW, V, f32>, S::TY, S::TXR};
}

template <int R>
TiledKernel tiled_entry(int v)
{
    if constexpr (R == 4) {
        switch (v) {
        case 1: return tiled_config<R, 1, 2, 4>();
        case 2: return tiled_config<R, 1, 1, 8>();
        case 3: return tiled_config<R, 2, 1, 4>();
        case 4: return tiled_config<R, 4, 1, 4>();
        case 5: return tiled_config<R, 2, 2, 4>();
        case 6: return tiled_config<R, 1, 2, 8>();
        case 7: return tiled_config<R, 4, 1, 2>();
        case 8: return tiled_config<R, 8, 1, 1>();
        case 9: return tiled_config<R, 8, 1, 2>();
        case 10: return tiled_config<R, 4, 2, 1>();
        case 11: return tiled_config<R, 4, 1, 1>();
        case 12: return tiled_config<R, 4, 1, 8>();
        case 13: return tiled_config<R, 2, 1, 8>();
        case 14: return tiled_config<R, 8, 2, 1>();
        case 15: return tiled_config<R, 4, 2, 2>();
        case 16: return tiled_config<R, 8, 2, 2>();
        // 8-byte lanes
        case 17: return tiled_config<R, 2, 4, 1, sw::kVec8>();
        case 18: return tiled_config<R, 2, 4, 2, sw::kVec8>();
        case 19: return tiled_config<R, 4, 4, 1, sw::kVec8>();
        case 20: return tiled_config<R, 1, 4, 1, sw::kVec8>();
        case 21: return tiled_config<R, 2, 2, 2, sw::kVec8>();
        case 22: return tiled_config<R, 1, 2, 4, sw::kVec8>();
        case 23: return tiled_config<R, 4, 2, 1, sw::kVec8>();
        default: return tiled_config<R, 1, 1, 4>();
        }
    } else {
        return v == 4 ? tiled_config<R, 4, 1, 4>()
             : v == 8 ? tiled_config<R, 8, 1, 1>()
             : v == 17 ? tiled_config<R, 4, 2, 1, sw::kVec8>()
             : v == 18 ? tiled_config<R, 2, 2, 1, sw::kVec8>()
                       : tiled_config<R, 1, 1, 4>();
    }
}

TiledKernel tiled_kernel(int R, int v)
{
    switch (R) {
    case 1: return tiled_entry<1>(v);
    case 2: return tiled_entry<2>(v);
    case 3: return tiled_entry<3>(v);
    case 4: return tiled_entry<4>(v);
    case 5: return tiled_entry<5>(v);
    case 6: return tiled_entry<6>(v);
    case 7: return tiled_entry<7>(v);
    case 8: return tiled_entry<8>(v);
    case 9: return tiled_entry<9>(v);
    default: return tiled_entry<10>(v);
    }
}

// Chooses the tile grid and launch order and marks the tile planes that
// hold damping.  Leaves pr.tiled false when the grid does not fit the
// kernel's 32-bit buffer offsets.
void setup_tiling(Prepared &pr, DeviceArena &arena, hipStream_t st)
{
    const Params &P = pr.P;
    const int R = pr.r;
    // float32 Laplacian quotients as float64 products with RN(1/h^2)
    // where that gives the IEEE quotient for every numerator
    const bool div64 = sizeof(real_t) == 4 && !env_flag("SIMWAVE_HIP_IEEE_DIV") &&
                       !env_flag("SIMWAVE_HIP_NO_DIV64") && f64_quotient_exact(P.dz2) &&
                       f64_quotient_exact(P.dx2) && f64_quotient_exact(P.dy2);
    int shape = env_int("SIMWAVE_HIP_TILE", default_tile(R, div64));
    shape = shape >= 0 && shape < kTileShapes ? shape : default_tile(R, div64);
    const TiledKernel K = tiled_kernel(R, shape);
    const int Iy = P.ny - 2 * R, Ix = P.nx - 2 * R, Iz = P.nz - 2 * R;
    const long long pbytes = (long long)P.nx * P.ny * (long long)sizeof(real_t);
    // buffer offsets are 32-bit: the planes one tile touches (at most
    // 64 + 2R) plus one row of overrun must stay below 2 GiB
    if (pbytes * (64 + 2 * R + 1) >= (1ll << 31) || P.ny < sw::kVec16)
        return;
    sw::Tiling T{};
    T.ty = K.ty;
    T.txr = K.txr;
    T.gk = (Iy + K.ty - 1) / K.ty;
    T.gj = (Ix + K.txr - 1) / K.txr;
    T.ncols = T.gk * T.gj;
    T.order = env_int("SIMWAVE_HIP_ORDER", 1) == 1 ? 1 : 0;
    T.chunk = (T.ncols + 7) / 8;
    // short z chunks (default_tile_zlen planes): short tiles keep many
    // independent blocks in every phase (prologue, march) on each CU, which
    // overlaps their loads and arithmetic better than long marches, at
    // 2R/zlen extra centre-plane reads
    int zlen = env_int("SIMWAVE_HIP_ZLEN", 0);
    if (zlen <= 0) {
        const int target = default_tile_zlen(R);
        const int chunks = std::max(1, (Iz + target - 1) / target);
        zlen = (Iz + chunks - 1) / chunks;
    }
    zlen = std::max(1, std::min({zlen, Iz, 64}));   // one 64-bit damping mask per tile
    T.zlen = zlen;
    T.gz = (Iz + zlen - 1) / zlen;
    T.ntiles = T.ncols * T.gz;
    const long long blocks =
        T.order == 1 ? 8ll * T.chunk * T.gz : (long long)T.ntiles;
    if (blocks >= (1ll << 31))
        return;
    T.nwords = P.nz / 32 + 3;
    unsigned *mask = arena.alloc<unsigned>((size_t)T.ncols * T.nwords);
    {
        const long long words = (long long)T.ncols * T.nwords;
        hipLaunchKernelGGL(sw::k_zero_u32,
                           dim3((unsigned)std::min<long long>((words + 255) / 256, 4096)),
                           dim3(256), 0, st, mask, words);
    }
    T.dmask = mask;
    hipLaunchKernelGGL(sw::k_damp_flags, dim3(T.ncols, P.nz), dim3(256), 0, st, P.e, P.nz,
                       P.nx, P.ny, R, T.gk, T.ty, T.txr, T.nwords, mask);
    hip_check(hipGetLastError(), "k_damp_flags launch");
    Params *Pd = arena.alloc<Params>(1);
    hip_check(hipMemcpyAsync(Pd, &P, sizeof(Params), hipMemcpyHostToDevice, st),
              "copy params");
    pr.Pd = Pd;
    {
        real_t *qc = arena.alloc<real_t>(pr.npts);
        const int nb = (int)std::min<long long>((pr.npts + 255) / 256, 8192);
        hipLaunchKernelGGL(sw::k_scaled_q, dim3(nb), dim3(256), 0, st, P.q, P.e, qc,
                           pr.npts);
        hip_check(hipGetLastError(), "k_scaled_q launch");
        pr.qc = qc;
    }
    pr.fastdiv = !env_flag("SIMWAVE_HIP_IEEE_DIV") && make_fixed_div(P.dz2, pr.dvz) &&
                 make_fixed_div(P.dx2, pr.dvx) && make_fixed_div(P.dy2, pr.dvy);
    pr.div64 = div64;
    pr.tiled_launch = pr.div64 ? K.launch64 : K.launch;
    pr.tiling = T;
    pr.tiled_blocks = (int)blocks;
    pr.tiled = true;
}
#endif

#if SW_DIM == 3
constexpr int kTX = 8;

template <int R>
void launch_fused(const Params &P, const real_t *uc, const real_t *up, real_t *un,
                  int n, hipStream_t st, LaunchTimer tm)
{
    const int Iy = P.ny - 2 * R, Ix = P.nx - 2 * R, Iz = P.nz - 2 * R;
    const int gx = (Iy + 63) / 64, gy = (Ix + kTX - 1) / kTX;
    int zlen = env_int("SIMWAVE_HIP_ZLEN", 0);
    if (zlen <= 0) {
        const long long cols = (long long)gx * gy;
        int nzc = (int)std::max<long long>(1, (2048 + cols - 1) / cols);
        nzc = std::min(nzc, std::max(1, Iz / 16));
        zlen = (Iz + nzc - 1) / nzc;
    }
    const int gz = (Iz + zlen - 1) / zlen;
    launch_timed(sw::k_step3d<R, kTX>, dim3(gx, gy, gz), dim3(64, kTX), st, tm, P, uc, up, un,
                 n, zlen);
}
#else
constexpr int kTB = 256;

// Whether 2D steps of radius R run k_step2d_z (see launch_fused).
bool zmarch_2d(int R)
{
    return env_flag("SIMWAVE_HIP_ZMARCH") ||
           ((sizeof(real_t) == 4 || R <= 4) && !env_flag("SIMWAVE_HIP_UNTILED"));
}

// Rows per block of the 2D step kernels: about 2048 blocks, down to one
// row each: a block's rows form one dependent load chain, so small grids run
// fastest with the most blocks (517^2 so4: 7.7 us per step at 1 row against
// 14.4 us at 9 rows; 1136 x 3136 so16: 57 us at 8 rows, within 3% of the
// best)
int zlen_2d(const Params &P, int R)
{
    const int Ix = P.nx - 2 * R, Iz = P.nz - 2 * R;
    const int gx = (Ix + kTB - 1) / kTB;
    int zlen = env_int("SIMWAVE_HIP_ZLEN", 0);
    if (zlen <= 0) {
        const int nzc = std::min(std::max(1, (2048 + gx - 1) / gx), std::max(1, Iz));
        zlen = (Iz + nzc - 1) / nzc;
    }
    return zlen;
}

template <int R>
void launch_fused(const Params &P, const real_t *uc, const real_t *up, real_t *un,
                  int n, hipStream_t st, LaunchTimer tm)
{
    const int Ix = P.nx - 2 * R, Iz = P.nz - 2 * R;
    const int gx = (Ix + kTB - 1) / kTB;
    const int zlen = zlen_2d(P, R);
    const int gz = (Iz + zlen - 1) / zlen;
    // the z-marching kernel for float32 (0-27% faster than the LDS row
    // kernel on 2D grids of 304^2 to 1136 x 3136, orders 2-20, CD and VD:
    // 41.7 vs 57.0 us per step at 1136 x 3136 so16 VD) and for float64 up
    // to radius 4 (517^2 so4: 8.2 vs 8.7 us); k_step2d for float64 above
    // (1136 x 3136 so16 VD: 66.5 vs 71.3 us).  SIMWAVE_HIP_UNTILED forces
    // k_step2d, SIMWAVE_HIP_ZMARCH the z-marching one.
    if (zmarch_2d(R)) {
        static_assert(sw::kZ2Block == kTB, "2D launch shapes differ");
        launch_timed(sw::k_step2d_z<R>, dim3(gx, gz), dim3(sw::kZ2Block), st, tm, P, uc, up,
                     un, n, zlen);
        return;
    }
    launch_timed(sw::k_step2d<R, kTB>, dim3(gx, gz), dim3(kTB), st, tm, P, uc, up, un, n,
                 zlen);
}

// k_step2d_z with the receiver gather of level n in the same launch (the
// 2D z-marching path with receivers on the step stream); 64-product rounds
// keep the gather's LDS at about 4 KB per block.
constexpr int kRecChunkFused = 64;

template <int R>
void launch_fused_rec(const Params &P, const real_t *uc, const real_t *up, real_t *un, int n,
                      hipStream_t st, LaunchTimer tm, const Box *rbox, const real_t *rw,
                      int nrec, real_t *row)
{
    const int Ix = P.nx - 2 * R, Iz = P.nz - 2 * R;
    const int gx = (Ix + kTB - 1) / kTB;
    const int zlen = zlen_2d(P, R);
    const int gz = (Iz + zlen - 1) / zlen;
    const int nrb = (nrec + kRecPerBlock - 1) / kRecPerBlock;
    launch_timed(sw::k_step2d_z_rec<R, kRecPerBlock, kRecChunkFused>,
                 dim3((unsigned)(nrb + gx * gz)), dim3(sw::kZ2Block), st, tm, P, uc, up, un, n,
                 zlen, gx, nrb, rbox, rw, nrec, row);
}

// Steps n0 .. n1 of a stride-0 2D shot in one cooperative launch
// (k_step2d_persist) of at most two blocks per compute unit (and at most
// the blocks of one step).  bar: two words, the barrier count and the
// abort word, zeroed here.  Polls per barrier before a block gives up: far
// beyond any step's time, so only a missing arrival ends a run early; the
// roll call that opens the launch gives up sooner, before anything is
// written.
constexpr int kPersistSpin = 1 << 22;
constexpr int kPersistRollSpin = 1 << 17;
constexpr int kPersistBlocksPerCU = 2;

template <int R>
void launch_persist(const Params &P, real_t *u, long long np, int n0, int n1, hipStream_t st,
                    const Box *rbox, const real_t *rw, int nrec, real_t *rec0, unsigned *bar)
{
    const auto kernel = sw::k_step2d_persist<R, kRecPerBlock, kRecChunkFused>;
    static int per_cu = -1, ncu = 0;
    static std::mutex mu;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (per_cu < 0) {
            int dev = 0;
            hip_check(hipGetDevice(&dev), "hipGetDevice");
            hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev),
                      "hipDeviceGetAttribute");
            hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, sw::kZ2Block,
                                                                   0),
                      "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        }
    }
    const int Ix = P.nx - 2 * R, Iz = P.nz - 2 * R;
    int gx = (Ix + kTB - 1) / kTB;
    int zlen = zlen_2d(P, R);
    int nsb = gx * ((Iz + zlen - 1) / zlen);
    int nrb = (nrec + kRecPerBlock - 1) / kRecPerBlock;
    const int bpc = std::max(1, env_int("SIMWAVE_HIP_PERSIST_BPC", kPersistBlocksPerCU));
    const long long resident = (long long)std::min(per_cu, bpc) * ncu;
    if (resident < 1)
        fail("no block of the whole-run 2D kernel fits on a compute unit");
    const unsigned grid = (unsigned)std::min<long long>(resident, (long long)nsb + nrb);
    // the barrier count reaches (steps + 1) x grid within the launch
    if ((long long)(n1 - n0 + 2) * grid >= (1ll << 31))
        fail("%d steps are too many for one whole-run launch", n1 - n0 + 1);
    Params Pk = P;
    int spin = kPersistSpin, roll = kPersistRollSpin;
    // test knob: a roll call that cannot complete, to exercise the fallback
    int roll_extra = env_flag("SIMWAVE_HIP_PERSIST_TEST_FALLBACK") ? 1 : 0;
    if (roll_extra)
        roll = 1 << 10;
    int a0 = n0, a1 = n1;
    unsigned *count = bar;
    int *abort_flag = reinterpret_cast<int *>(bar + 1);
    hipLaunchKernelGGL(sw::k_zero_u32, dim3(1), dim3(64), 0, st, bar, 2ll);
    void *args[] = {&Pk, &u, &np, &a0, &a1, &zlen, &gx, &nsb, &nrb, &rbox, &rw, &nrec, &rec0,
                    &count, &abort_flag, &spin, &roll, &roll_extra};
    hip_check(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(kernel), dim3(grid),
                                         dim3(sw::kZ2Block), args, 0, st),
              "cooperative launch");
}
#endif

template <int R>
void launch_generic(const Params &P, const real_t *uc, const real_t *up, real_t *un,
                    hipStream_t st)
{
#if SW_DIM == 3
    const long long cnt = (long long)(P.ny - 2 * R) * (P.nx - 2 * R) * (P.nz - 2 * R);
#else
    const long long cnt = (long long)(P.nx - 2 * R) * (P.nz - 2 * R);
#endif
    if (cnt <= 0)
        return;
    hipLaunchKernelGGL((sw::k_stencil_generic<R>), dim3((unsigned)((cnt + 255) / 256)),
                       dim3(256), 0, st, P, uc, up, un);
}

using StepFn = void (*)(const Params &, const real_t *, const real_t *, real_t *,
                        int, hipStream_t, LaunchTimer);
using GenFn = void (*)(const Params &, const real_t *, const real_t *, real_t *,
                       hipStream_t);

#if SW_DIM == 2
using StepRecFn = void (*)(const Params &, const real_t *, const real_t *, real_t *, int,
                           hipStream_t, LaunchTimer, const Box *, const real_t *, int,
                           real_t *);
using PersistFn = void (*)(const Params &, real_t *, long long, int, int, hipStream_t,
                           const Box *, const real_t *, int, real_t *, unsigned *);
#endif

template <int... Rs>
struct Tables {
    static constexpr StepFn fused[] = {nullptr, &launch_fused<Rs>...};
    static constexpr GenFn generic[] = {nullptr, &launch_generic<Rs>...};
#if SW_DIM == 2
    static constexpr StepRecFn fused_rec[] = {nullptr, &launch_fused_rec<Rs>...};
    static constexpr PersistFn persist[] = {nullptr, &launch_persist<Rs>...};
#endif
};
using RTable = Tables<1, 2, 3, 4, 5, 6, 7, 8, 9, 10>;

void launch_bc_phases(const Params &P, real_t *un, hipStream_t st)
{
    const long long r = P.r;
#if SW_DIM == 3
    const long long Iy = P.ny - 2 * r, Ix = P.nx - 2 * r, Iz = P.nz - 2 * r;
    struct { int axis; int b, a; long long lines; } ph[3] = {
        {2, P.bc[4], P.bc[5], Iz * Ix},
        {1, P.bc[2], P.bc[3], Iz * Iy},
        {0, P.bc[0], P.bc[1], Ix * Iy},
    };
    const int nph = 3;
#else
    const long long Ix = P.nx - 2 * r, Iz = P.nz - 2 * r;
    struct { int axis; int b, a; long long lines; } ph[2] = {
        {2, P.bc[2], P.bc[3], Iz},
        {0, P.bc[0], P.bc[1], Ix},
    };
    const int nph = 2;
#endif
    for (int k = 0; k < nph; k++) {
        if ((!ph[k].b && !ph[k].a) || ph[k].lines <= 0)
            continue;
        hipLaunchKernelGGL(sw::k_bc_phase, dim3((unsigned)((ph[k].lines + 255) / 256)),
                           dim3(256), 0, st, P, un, ph[k].axis);
    }
}

// Second stream and events for the receiver gather (see time_loop).  A
// plan owns one across its runs; a one-shot call makes its own.
struct RecStream {
    hipStream_t rs = nullptr;
    hipEvent_t cur = nullptr, rec[3] = {nullptr, nullptr, nullptr};
    void create()
    {
        if (rs)
            return;
        hip_check(hipStreamCreateWithFlags(&rs, hipStreamNonBlocking), "hipStreamCreate");
        for (hipEvent_t *e : {&cur, &rec[0], &rec[1], &rec[2]})
            hip_check(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
    }
    ~RecStream()
    {
        if (rs)
            (void)hipStreamSynchronize(rs);
        for (hipEvent_t e : {cur, rec[0], rec[1], rec[2]})
            if (e)
                (void)hipEventDestroy(e);
        if (rs)
            (void)hipStreamDestroy(rs);
    }
};

#if SW_DIM == 3
// One step of k_step3d_z over the interior.
void zmarch_launch(const Prepared &pr, const real_t *uc, const real_t *up, real_t *un, int n,
                   hipStream_t st, LaunchTimer tm)
{
    const Params &P = pr.P;
    const int R = pr.r, zl = pr.zmarch_zlen;
    const dim3 grid((unsigned)((P.ny - 2 * R + 63) / 64),
                    (unsigned)((P.nx - 2 * R + sw::kZLines - 1) / sw::kZLines),
                    (unsigned)((P.nz - 2 * R + zl - 1) / zl));
    const dim3 block(64, sw::kZLines);
    switch (R) {
#define SW_Z_CASE(RR)                                                                   \
    case RR:                                                                            \
        launch_timed(sw::k_step3d_z<RR>, grid, block, st, tm, P, pr.lapdiv, uc, up, un, n, zl); \
        return;
    SW_Z_CASE(1) SW_Z_CASE(2) SW_Z_CASE(3) SW_Z_CASE(4) SW_Z_CASE(5)
    SW_Z_CASE(6) SW_Z_CASE(7) SW_Z_CASE(8) SW_Z_CASE(9) SW_Z_CASE(10)
#undef SW_Z_CASE
    default:
        fail("stencil radius %d out of range", R);
    }
}
#endif

// Runs steps [a.n0, a.n1].  `slot_ptr(i)` returns the device buffer backing
// logical slot i; `before_step`/`after_step` manage residency (host entry;
// `uploads(idx)` tells whether before_step(idx) writes a device buffer) and
// `swap` exchanges two logical slots.
//
// The receiver gather of step n only reads u^n (slot cur(n)), which no step
// before n+2 writes in the reference's slot cycle, so it runs on a second
// stream concurrently with the following kernels.  Ordering: it starts after
// everything enqueued before step n on `st` (u^n complete); `st` waits for
// it before any later step whose output slot is cur(n), before step n+2 in
// any case, before uploads into recycled buffers and before slot exchanges.
//
// With kFlagTimeSteps every fused step kernel is launched with its own
// event pair, set by the runtime at the kernel's start and end (the generic
// path's stencil kernel is bracketed by event records instead), and the
// durations are summed after the loop, so the measurement adds no
// synchronisation between steps.
template <typename SlotPtr, typename Before, typename After, typename Swap,
          typename Uploads>
void time_loop(const Args &a, const Prepared &pr, const std::vector<StepSlots> &plan,
               size_t i0, size_t i1, real_t *rec_dev, size_t rec_row0, unsigned flags,
               hipStream_t st, SlotPtr slot_ptr, Before before_step, After after_step,
               Swap swap, Uploads uploads, RecStream *rec_stream = nullptr,
               size_t nshift = 0)
{
    const Params &P = pr.P;
    const int r = pr.r;
    StepFn fused = RTable::fused[r];
    GenFn generic = RTable::generic[r];
    const bool timed = flags & kFlagTimeSteps;
    // plan entries i0 .. i1-1 (step a.n0 + idx for entry idx)
    const size_t nsteps = i1 - i0;
    struct Timing {
        std::vector<hipEvent_t> t;
        ~Timing()
        {
            for (hipEvent_t e : t)
                if (e)
                    (void)hipEventDestroy(e);
        }
    } tev;
    RecStream local;
    RecStream &ev = rec_stream ? *rec_stream : local;
    if (timed) {
        tev.t.assign(2 * nsteps, nullptr);
        for (hipEvent_t &e : tev.t)
            hip_check(hipEventCreate(&e), "hipEventCreate");
    }
    // the overlap costs three extra stream operations per step: worth it
    // where a step kernel is long enough to hide the gather
    const bool overlap_rec = a.nrec && (flags & kFlagOverlapReceivers) &&
                             pr.npts >= kOverlapMinPoints &&
                             !env_flag("SIMWAVE_HIP_REC_SERIAL");
    if (overlap_rec)
        ev.create();
    // waited[j]: st already waits for the receivers of entry i0 + j
    std::vector<char> waited(nsteps, 0);
    auto wait_rec = [&](long long j) {
        if (!overlap_rec || j < 0 || waited[j])
            return;
        for (long long i = j; i >= 0 && !waited[i]; i--)   // events complete in order
            waited[i] = 1;
        hip_check(hipStreamWaitEvent(st, ev.rec[j % 3], 0), "hipStreamWaitEvent");
    };
    const unsigned rec_blocks = (unsigned)((a.nrec + 255) / 256);
    const bool coop_receivers = !env_flag("SIMWAVE_HIP_REC_THREAD");
#if SW_DIM == 2
    // 2D z-marching steps carry their receiver gather (one launch per step)
    const bool rec_in_step = pr.fused && a.nrec && !overlap_rec && coop_receivers &&
                             zmarch_2d(r) && !env_flag("SIMWAVE_HIP_REC_SEPARATE");
#else
    const bool rec_in_step = false;
#endif
    const unsigned coop_blocks = (unsigned)((a.nrec + kRecPerBlock - 1) / kRecPerBlock);
    auto receivers = [&](size_t n, const real_t *uc, hipStream_t s) {
        real_t *row = rec_dev + (n - 1 - rec_row0) * a.nrec;
        if (coop_receivers)
            hipLaunchKernelGGL((sw::k_receivers_coop<kRecPerBlock, kRecChunk>),
                               dim3(coop_blocks), dim3(256), 0, s, P, uc, pr.rbox, pr.rw,
                               (int)a.nrec, row);
        else
            hipLaunchKernelGGL(sw::k_receivers, dim3(rec_blocks), dim3(256), 0, s, P, uc,
                               pr.rbox, pr.rw, (int)a.nrec, row);
    };
    for (size_t idx = i0; idx < i1; idx++) {
        // nshift: capture of a graph replayed with a device step base
        // (Params::nbase); kernels then see the step relative to it
        const size_t n = a.n0 + idx - nshift;
        const StepSlots &s = plan[idx];
        const long long ii = (long long)(idx - i0);
        // the receivers of steps <= idx-2 are done before step idx; those of
        // idx-1 too where step idx overwrites their slot or buffers are
        // recycled by an upload
        wait_rec(ii - 2);
        if (ii > 0 && (s.next == plan[idx - 1].cur || uploads(idx)))
            wait_rec(ii - 1);
        before_step(idx);
        const real_t *up = slot_ptr(s.prev);
        const real_t *uc = slot_ptr(s.cur);
        real_t *un = slot_ptr(s.next);
        if (overlap_rec) {
            hip_check(hipEventRecord(ev.cur, st), "hipEventRecord");
            hip_check(hipStreamWaitEvent(ev.rs, ev.cur, 0), "hipStreamWaitEvent");
            receivers(n, uc, ev.rs);
            hip_check(hipEventRecord(ev.rec[ii % 3], ev.rs), "hipEventRecord");
        }
        hipEvent_t t0 = timed ? tev.t[2 * ii] : nullptr, t1 = timed ? tev.t[2 * ii + 1] : nullptr;
        if (pr.fused) {
            if (pr.halo_sources)
                hipLaunchKernelGGL(sw::k_inject, dim3(1), dim3(256), 0, st, P, un,
                                   (int)n, 1);
            const LaunchTimer tm{t0, t1};
#if SW_DIM == 3
            if (pr.zmarch)
                zmarch_launch(pr, uc, up, un, (int)n, st, tm);
            else if (pr.tiled)
                pr.tiled_launch(pr, uc, up, un, (int)n, st, tm);
            else
#else
            if (rec_in_step)
                RTable::fused_rec[r](P, uc, up, un, (int)n, st, tm, pr.rbox, pr.rw,
                                     (int)a.nrec, rec_dev + (n - 1 - rec_row0) * a.nrec);
            else
#endif
                fused(P, uc, up, un, (int)n, st, tm);
        } else {
            if (timed)
                hip_check(hipEventRecord(t0, st), "hipEventRecord");
            generic(P, uc, up, un, st);
            if (timed)
                hip_check(hipEventRecord(t1, st), "hipEventRecord");
            if (P.nsrc)
                hipLaunchKernelGGL(sw::k_inject, dim3(1), dim3(256), 0, st, P, un,
                                   (int)n, 0);
            launch_bc_phases(P, un, st);
        }
        if (a.nrec && !overlap_rec && !rec_in_step)
            receivers(n, uc, st);
        hip_check(hipGetLastError(), "step launch");
        if (s.swap_after) {
            wait_rec(ii);
            swap(s.swap_a, s.swap_b);
        }
        after_step(idx);
    }
    // st depends on every receiver gather of the range before it goes on
    if (nsteps)
        wait_rec((long long)nsteps - 1);
    if (timed) {
        hip_check(hipStreamSynchronize(st), "timed steps");
        double step_ms = 0.0;
        for (size_t i = 0; i < nsteps; i++) {
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, tev.t[2 * i], tev.t[2 * i + 1]),
                      "hipEventElapsedTime");
            step_ms += ms;
        }
        g_timings[4] = step_ms * 1e-3;
        g_timings[5] = (double)nsteps;
    }
}

// --------------------------------------------------------- host entry
// Copies between pageable and pinned host memory with a few threads: a
// single thread moves about 10 GB/s, short of what the DMA engines do.
class HostCopier {
public:
    explicit HostCopier(int nthreads)
    {
        for (int t = 0; t < nthreads; t++)
            threads_.emplace_back([this] { work(); });
    }
    ~HostCopier()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : threads_)
            t.join();
    }
    // Blocking parallel memcpy.
    void copy(void *dst, const void *src, size_t bytes)
    {
        const size_t parts = threads_.size() + 1;
        std::unique_lock<std::mutex> g(m_);
        dst_ = static_cast<char *>(dst);
        src_ = static_cast<const char *>(src);
        bytes_ = bytes;
        chunk_ = (bytes + parts - 1) / parts;
        next_ = 0;
        done_ = 0;
        parts_ = parts;
        gen_++;
        g.unlock();
        cv_.notify_all();
        run_chunks();
        g.lock();
        done_cv_.wait(g, [&] { return done_ == parts_; });
    }

private:
    void run_chunks()
    {
        for (;;) {
            size_t part;
            {
                std::lock_guard<std::mutex> g(m_);
                if (next_ >= parts_)
                    return;
                part = next_++;
            }
            const size_t lo = part * chunk_;
            const size_t hi = std::min(bytes_, lo + chunk_);
            if (lo < hi)
                std::memcpy(dst_ + lo, src_ + lo, hi - lo);
            std::lock_guard<std::mutex> g(m_);
            if (++done_ == parts_)
                done_cv_.notify_all();
        }
    }
    void work()
    {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_)
                    return;
                seen = gen_;
            }
            run_chunks();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    bool stop_ = false;
    unsigned long long gen_ = 0;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0, chunk_ = 0, next_ = 0, done_ = 0, parts_ = 0;
};

// Downloads finished slots without stalling the time loop: a device slot is
// copied on `ds` into one of a ring of pinned staging buffers, and a thread
// moves it from there into the caller's array once the copy has landed.
class SlotDownloader {
public:
    SlotDownloader(size_t bytes, int depth, hipStream_t ds)
        : bytes_(bytes), ds_(ds), copier_(4)
    {
        stage_.assign(depth, nullptr);
        done_.assign(depth, nullptr);
        busy_.assign(depth, false);
        for (int k = 0; k < depth; k++) {
            hip_check(hipHostMalloc((void **)&stage_[k], bytes, hipHostMallocDefault),
                      "hipHostMalloc (staging)");
            hip_check(hipEventCreateWithFlags(&done_[k], hipEventDisableTiming),
                      "hipEventCreate");
        }
        thread_ = std::thread([this] { work(); });
    }
    ~SlotDownloader()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        thread_.join();
        for (size_t k = 0; k < stage_.size(); k++) {
            if (done_[k])
                (void)hipEventDestroy(done_[k]);
            if (stage_[k])
                (void)hipHostFree(stage_[k]);
        }
    }
    // Enqueues the copy of device slot `src` (complete once `after` fires)
    // into `dst`; blocks while every staging buffer is in flight.
    void push(const real_t *src, real_t *dst, hipEvent_t after)
    {
        int k;
        {
            std::unique_lock<std::mutex> g(m_);
            cv_.wait(g, [&] { return !busy_[ring_] || !error_.empty(); });
            if (!error_.empty())
                fail("%s", error_.c_str());
            k = ring_;
            ring_ = (ring_ + 1) % (int)stage_.size();
            busy_[k] = true;
        }
        hip_check(hipStreamWaitEvent(ds_, after, 0), "hipStreamWaitEvent");
        hip_check(hipMemcpyAsync(stage_[k], src, bytes_, hipMemcpyDeviceToHost, ds_),
                  "download slot");
        hip_check(hipEventRecord(done_[k], ds_), "hipEventRecord");
        {
            std::lock_guard<std::mutex> g(m_);
            jobs_.push_back({k, dst});
        }
        cv_.notify_all();
    }
    // Waits until every pushed slot is in the caller's array.
    void drain()
    {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] {
            if (!error_.empty())
                return true;
            if (!jobs_.empty())
                return false;
            for (bool b : busy_)
                if (b)
                    return false;
            return true;
        });
        if (!error_.empty())
            fail("%s", error_.c_str());
    }

private:
    struct Job {
        int k;
        real_t *dst;
    };
    void work()
    {
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || !jobs_.empty(); });
                if (jobs_.empty())
                    return;
                j = jobs_.front();
                jobs_.pop_front();
            }
            const hipError_t e = hipEventSynchronize(done_[j.k]);
            if (e == hipSuccess)
                copier_.copy(j.dst, stage_[j.k], bytes_);
            {
                std::lock_guard<std::mutex> g(m_);
                if (e != hipSuccess && error_.empty())
                    error_ = std::string("slot download: ") + hipGetErrorString(e);
                busy_[j.k] = false;
            }
            cv_.notify_all();
        }
    }
    size_t bytes_;
    hipStream_t ds_;
    HostCopier copier_;
    std::vector<real_t *> stage_;
    std::vector<hipEvent_t> done_;
    std::vector<bool> busy_;
    int ring_ = 0;
    std::deque<Job> jobs_;
    std::mutex m_;
    std::condition_variable cv_;
    bool stop_ = false;
    std::string error_;
    std::thread thread_;
};

// Element indices of the halo frame (points outside [r, n-r) on some axis)
// of one slot, in ascending order.
std::vector<unsigned long long> frame_indices(const Params &P)
{
    std::vector<unsigned long long> out;
    const long long r = P.r, nz = P.nz, nx = P.nx, ny = P.ny;
    const bool d3 = SW_DIM == 3;
    for (long long i = 0; i < nz; i++)
        for (long long j = 0; j < nx; j++) {
            const unsigned long long row = (unsigned long long)(i * nx + j) * (d3 ? ny : 1);
            const bool whole = i < r || i >= nz - r || (d3 && (j < r || j >= nx - r));
            if (!d3) {
                if (whole || j < r || j >= nx - r)
                    out.push_back(row);
                continue;
            }
            if (whole) {
                for (long long k = 0; k < ny; k++)
                    out.push_back(row + k);
            } else {
                for (long long k = 0; k < r; k++)
                    out.push_back(row + k);
                for (long long k = ny - r; k < ny; k++)
                    out.push_back(row + k);
            }
        }
    return out;
}

// Arrays are host pointers (the reference contract).  Returns wall seconds.
//
// Slot residency: a logical slot lives on the device from its first to its
// last use.  A slot first used as a step's input is uploaded whole; one
// first used as the step's output only needs its halo frame (the step
// writes every interior point), which is uploaded packed and scattered.
// A slot is downloaded after its last use through SlotDownloader, so the
// transfers overlap the time loop; its device buffer is reused only after
// its download has fired.
double run_host(const Args &a)
{
    const double t0 = now_s();
    for (double &t : g_timings)
        t = 0.0;
    validate(a);
    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (ndev < 1)
        fail("no HIP device visible");
    keep_pool_memory();
    hipStream_t st = nullptr, ds = nullptr;
    hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard()
        {
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
        }
    } sg{st};
    hip_check(hipStreamCreateWithFlags(&ds, hipStreamNonBlocking), "hipStreamCreate");
    StreamGuard sgd{ds};

    std::vector<StepSlots> plan = plan_slots(a);
    // slot residency: first and last step index of every logical slot
    std::vector<long long> first(a.num_snapshots, -1), last(a.num_snapshots, -1);
    auto touch = [&](size_t slot, size_t idx) {
        if (slot >= a.num_snapshots)
            fail("slot %zu used by step %zu is outside the %zu snapshots", slot,
                 a.n0 + idx, a.num_snapshots);
        if (first[slot] < 0)
            first[slot] = (long long)idx;
        last[slot] = (long long)idx;
    };
    for (size_t idx = 0; idx < plan.size(); idx++) {
        touch(plan[idx].prev, idx);
        touch(plan[idx].cur, idx);
        touch(plan[idx].next, idx);
        if (plan[idx].swap_after) {
            touch(plan[idx].swap_a, idx);
            touch(plan[idx].swap_b, idx);
        }
    }

    double t_prep = 0, t_loop = 0, t_down = 0;
    {
        DeviceArena arena(st);
        Prepared pr = prepare(a, false, 0, arena, st);
        const size_t npts = (size_t)pr.npts;
        const size_t slot_bytes = npts * sizeof(real_t);

        // receivers rows n0-1 .. n1-1 are fully written by the run
        const size_t rows = plan.size();
        real_t *rec_dev = arena.alloc<real_t>(rows * a.nrec);

        std::vector<real_t *> phys(a.num_snapshots, nullptr);
        std::vector<real_t *> pool;
        std::map<real_t *, hipEvent_t> free_after;   // buffer -> download done
        auto acquire = [&]() -> real_t * {
            if (!pool.empty()) {
                real_t *p = pool.back();
                pool.pop_back();
                auto it = free_after.find(p);
                if (it != free_after.end()) {
                    hip_check(hipStreamWaitEvent(st, it->second, 0), "hipStreamWaitEvent");
                    free_after.erase(it);
                }
                return p;
            }
            return arena.alloc<real_t>(npts);
        };
        // bucket slots by the step that uploads / downloads them
        std::vector<std::vector<size_t>> up_at(plan.size()), down_at(plan.size());
        bool any_frame = false;
        for (size_t sl = 0; sl < a.num_snapshots; sl++) {
            if (first[sl] < 0)
                continue;
            up_at[first[sl]].push_back(sl);
            down_at[last[sl]].push_back(sl);
            const StepSlots &s = plan[first[sl]];
            any_frame |= s.next == sl && s.prev != sl && s.cur != sl;
        }
        // halo frame: indices on the device, packed values staged pinned
        std::vector<unsigned long long> fidx;
        unsigned long long *d_fidx = nullptr;
        real_t *d_fval = nullptr, *h_fval = nullptr;
        hipEvent_t fval_free = nullptr;
        struct PinnedGuard {
            real_t *p = nullptr;
            hipEvent_t e = nullptr;
            ~PinnedGuard()
            {
                if (e)
                    (void)hipEventDestroy(e);
                if (p)
                    (void)hipHostFree(p);
            }
        } pg;
        if (any_frame) {
            fidx = frame_indices(pr.P);
            d_fidx = arena.alloc<unsigned long long>(fidx.size());
            d_fval = arena.alloc<real_t>(fidx.size());
            hip_check(hipMemcpyAsync(d_fidx, fidx.data(), fidx.size() * sizeof(fidx[0]),
                                     hipMemcpyHostToDevice, st), "upload frame indices");
            hip_check(hipHostMalloc((void **)&h_fval, fidx.size() * sizeof(real_t),
                                    hipHostMallocDefault), "hipHostMalloc (frame)");
            pg.p = h_fval;
            hip_check(hipEventCreateWithFlags(&fval_free, hipEventDisableTiming),
                      "hipEventCreate");
            pg.e = fval_free;
            hip_check(hipEventRecord(fval_free, st), "hipEventRecord");
        }
        hipEvent_t slot_done = nullptr;
        hip_check(hipEventCreateWithFlags(&slot_done, hipEventDisableTiming), "hipEventCreate");
        struct EventGuard {
            hipEvent_t e;
            ~EventGuard() { (void)hipEventDestroy(e); }
        } eg{slot_done};
        std::vector<hipEvent_t> buf_events;    // one per recycled-buffer download
        struct EventsGuard {
            std::vector<hipEvent_t> &v;
            ~EventsGuard()
            {
                for (hipEvent_t e : v)
                    (void)hipEventDestroy(e);
            }
        } evg{buf_events};
        SlotDownloader down(slot_bytes, 3, ds);
        hip_check(hipStreamSynchronize(st), "prepare");
        t_prep = now_s() - t0;
        const double tl0 = now_s();

        auto before = [&](size_t idx) {
            const StepSlots &s = plan[idx];
            for (size_t sl : up_at[idx]) {
                real_t *d = acquire();
                const real_t *src = a.u + sl * npts;
                if (s.next == sl && s.prev != sl && s.cur != sl) {
                    hip_check(hipEventSynchronize(fval_free), "hipEventSynchronize");
                    const size_t nf = fidx.size();
                    for (size_t i = 0; i < nf; i++)
                        h_fval[i] = src[fidx[i]];
                    hip_check(hipMemcpyAsync(d_fval, h_fval, nf * sizeof(real_t),
                                             hipMemcpyHostToDevice, st), "upload frame");
                    hipLaunchKernelGGL(sw::k_scatter,
                                       dim3((unsigned)std::min<size_t>((nf + 255) / 256, 4096)),
                                       dim3(256), 0, st, d, d_fidx, d_fval, (long long)nf);
                    hip_check(hipEventRecord(fval_free, st), "hipEventRecord");
                } else {
                    hip_check(hipMemcpyAsync(d, src, slot_bytes, hipMemcpyHostToDevice, st),
                              "upload slot");
                }
                phys[sl] = d;
            }
        };
        auto after = [&](size_t idx) {
            if (down_at[idx].empty())
                return;
            hip_check(hipEventRecord(slot_done, st), "hipEventRecord");
            for (size_t sl : down_at[idx]) {
                down.push(phys[sl], a.u + sl * npts, slot_done);
                // the buffer may be rewritten once its download has fired
                hipEvent_t e = nullptr;
                hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
                buf_events.push_back(e);
                hip_check(hipEventRecord(e, ds), "hipEventRecord");
                free_after[phys[sl]] = e;
                pool.push_back(phys[sl]);
                phys[sl] = nullptr;
            }
        };
        auto slot_ptr = [&](size_t sl) { return phys[sl]; };
        auto swap = [&](size_t x, size_t y) { std::swap(phys[x], phys[y]); };

        auto uploads = [&](size_t idx) { return !up_at[idx].empty(); };
        // the receiver overlap where no other run or plan of the process
        // holds it (see OverlapLease)
        OverlapLease lease;
        const unsigned hflags = lease.acquire(false) ? kFlagOverlapReceivers : 0u;
        time_loop(a, pr, plan, 0, plan.size(), rec_dev, a.n0 - 1, hflags, st,
                  slot_ptr, before, after, swap, uploads);
        hip_check(hipStreamSynchronize(st), "time loop");
        const double tl1 = now_s();
        t_loop = tl1 - tl0;
        if (rows && a.nrec)
            hip_check(hipMemcpyAsync(a.receivers + (a.n0 - 1) * a.nrec, rec_dev,
                                     rows * a.nrec * sizeof(real_t),
                                     hipMemcpyDeviceToHost, st), "download receivers");
        hip_check(hipStreamSynchronize(st), "download");
        down.drain();
        t_down = now_s() - tl1;
        g_timings[6] = pr.fused ? 1.0 : 0.0;
#if SW_DIM == 3
        g_timings[7] = pr.tiled ? 1.0 : pr.zmarch ? 2.0 : 0.0;
#else
        g_timings[7] = pr.fused && zmarch_2d(pr.r) ? 2.0 : 0.0;
#endif
    }
    const double total = now_s() - t0;
    g_timings[0] = total;
    g_timings[1] = t_prep;
    g_timings[2] = t_loop;
    g_timings[3] = t_down;
    return total;
}

// -------------------------------------------------------- device entry
// u, velocity, density, damp and receivers are device pointers; u holds all
// num_snapshots slots, receivers wavelet_size rows.  Everything else is host
// memory, read only while the plan is created.  A plan keeps everything a
// shot needs on the device and runs its steps in order, in any number of
// chunks, on any stream.
struct Plan {
    OverlapLease lease;          // held with kFlagOverlapReceivers (released last)
    Args a;
    unsigned flags;
    hipStream_t st;              // stream of creation; the arena frees on it
    DeviceArena arena;
    Prepared pr;
    std::vector<StepSlots> slots;
    size_t done = 0;             // plan entries enqueued so far
    hipEvent_t last = nullptr;   // end of the latest enqueued work
    RecStream rec;
    // graph replay (kFlagGraph): kGraphSteps steps captured with the step
    // number relative to *nbase; prg is pr with P.nbase set
    Prepared prg;
    int *nbase = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // batched runs (run_plans_batched): the shot's constants on the device
    // and the blocks of its step and receiver grids; null when the plan
    // runs only on its own
    const sw::BatchShot *bshot = nullptr;
    int bblocks = 0, rblocks = 0;
    // whole-run 2D launches (k_step2d_persist): the barrier counter and the
    // abort word; null when the plan runs step by step
    unsigned *bar = nullptr;
    Plan(const Args &args, hipStream_t s, unsigned f) : a(args), flags(f), st(s), arena(s) {}
    ~Plan()
    {
        // the arena frees on `st`: order it after the latest run
        if (last) {
            if (gexec)
                (void)hipEventSynchronize(last);   // replays done before destroying
            (void)hipStreamWaitEvent(st, last, 0);
            (void)hipEventDestroy(last);
        }
        if (gexec)
            (void)hipGraphExecDestroy(gexec);
        if (graph)
            (void)hipGraphDestroy(graph);
    }
};

// Whether the plan's slot cycle repeats every 3 steps with no exchanges
// (stride 0), so that one captured chunk serves every aligned position.
bool periodic_slots(const std::vector<StepSlots> &slots)
{
    for (size_t i = 0; i < slots.size(); i++) {
        const StepSlots &s = slots[i], &b = slots[i % 3];
        if (s.swap_after || s.prev != b.prev || s.cur != b.cur || s.next != b.next)
            return false;
    }
    return true;
}

// Captures plan entries 0 .. kGraphSteps-1 on the plan's stream with the
// step numbers relative to the device counter p.nbase (1 .. kGraphSteps);
// a replay starting at entry e sets *nbase = n0 + e - 1 first.  The chunk's
// kernels take the same arguments as in a direct run, except the
// step-dependent ones (sources, receiver rows), which add *nbase.
void capture_graph(Plan &p)
{
    p.nbase = p.arena.alloc<int>(1);
    p.prg = p.pr;
    p.prg.P.nbase = p.nbase;
#if SW_DIM == 3
    if (p.pr.tiled) {
        Params *Pg = p.arena.alloc<Params>(1);
        hip_check(hipMemcpyAsync(Pg, &p.prg.P, sizeof(Params), hipMemcpyHostToDevice, p.st),
                  "copy params");
        p.prg.Pd = Pg;
    }
#endif
    // replays run after the plan's preparation (ordered by the plan's
    // event); the capture records without executing, on a stream of its
    // own (the default stream cannot be captured), at the default priority
    struct OwnStream {
        hipStream_t s = nullptr;
        ~OwnStream()
        {
            if (s)
                (void)hipStreamDestroy(s);
        }
    } cs;
    hip_check(hipStreamCreateWithFlags(&cs.s, hipStreamNonBlocking), "capture stream");
    hipStream_t st = cs.s;
    const size_t npts = (size_t)p.pr.npts;
    real_t *u = p.a.u;
    auto slot_ptr = [&](size_t sl) { return u + sl * npts; };
    auto nop = [](size_t) {};
    auto noswap = [](size_t, size_t) {};
    auto no_uploads = [](size_t) { return false; };
    hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
        time_loop(p.a, p.prg, p.slots, 0, kGraphSteps, p.a.receivers, 0, 0u, st, slot_ptr,
                  nop, nop, noswap, no_uploads, nullptr, p.a.n0 - 1);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g)
            (void)hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(st, &p.graph), "end capture");
    hip_check(hipGraphInstantiate(&p.gexec, p.graph, nullptr, nullptr, 0), "graph instantiate");
}

size_t run_plan(Plan &p, size_t count, hipStream_t st, bool whole_run = false);

// A plan can join batched runs when its steps use the fused z-marching
// kernel of a stride-0 run with no source in the halo and no per-step
// timing or receiver stream.
void setup_batch(Plan &p)
{
    const Prepared &pr = p.pr;
    const Params &P = pr.P;
    const int R = pr.r;
    if (!pr.fused || pr.halo_sources || p.a.stride != 0 || !periodic_slots(p.slots) ||
        (p.flags & (kFlagTimeSteps | kFlagOverlapReceivers)) || env_flag("SIMWAVE_HIP_NO_BATCH"))
        return;
    sw::BatchShot b{};
#if SW_DIM == 3
    if (!pr.zmarch)
        return;
    b.D = pr.lapdiv;
    b.zlen = pr.zmarch_zlen;
    b.gx = (P.ny - 2 * R + 63) / 64;
    b.gy = (P.nx - 2 * R + sw::kZLines - 1) / sw::kZLines;
    const long long blocks =
        (long long)b.gx * b.gy * ((P.nz - 2 * R + b.zlen - 1) / b.zlen);
#else
    if (!zmarch_2d(R))
        return;
    b.zlen = zlen_2d(P, R);
    b.gx = (P.nx - 2 * R + kTB - 1) / kTB;
    b.gy = 1;
    const long long blocks = (long long)b.gx * ((P.nz - 2 * R + b.zlen - 1) / b.zlen);
#endif
    if (blocks >= (1ll << 28))
        return;
    b.P = P;
    b.P.nbase = nullptr;
    b.u = p.a.u;
    b.rec = p.a.receivers;
    b.rbox = pr.rbox;
    b.rw = pr.rw;
    b.npts = pr.npts;
    b.nrec = (int)p.a.nrec;
    sw::BatchShot *d = p.arena.alloc<sw::BatchShot>(1);
    hip_check(hipMemcpyAsync(d, &b, sizeof(b), hipMemcpyHostToDevice, p.st),
              "copy batch constants");
    p.bshot = d;
    p.bblocks = (int)blocks;
    p.rblocks = (int)((p.a.nrec + kRecPerBlock - 1) / kRecPerBlock);
}

template <int R>
void launch_step_batch(const sw::BatchArgs &A, hipStream_t st)
{
#if SW_DIM == 3
    hipLaunchKernelGGL(sw::k_step3d_z_batch<R>, dim3((unsigned)A.first[A.count]),
                       dim3(64, sw::kZLines), 0, st, A);
#else
    hipLaunchKernelGGL(sw::k_step2d_z_batch<R>, dim3((unsigned)A.first[A.count]),
                       dim3(sw::kZ2Block), 0, st, A);
#endif
}

using BatchFn = void (*)(const sw::BatchArgs &, hipStream_t);
constexpr BatchFn kStepBatch[] = {nullptr,
                                  &launch_step_batch<1>, &launch_step_batch<2>,
                                  &launch_step_batch<3>, &launch_step_batch<4>,
                                  &launch_step_batch<5>, &launch_step_batch<6>,
                                  &launch_step_batch<7>, &launch_step_batch<8>,
                                  &launch_step_batch<9>, &launch_step_batch<10>};

// Enqueues up to `steps` further steps of every plan on `st`, after each
// plan's earlier work.  Plans that can be batched (setup_batch) advance
// together, radius by radius, up to kMaxBatch per launch: one step kernel
// and one receiver kernel per time step for all of them, in segments over
// which the set of unfinished plans is fixed.  The others run on their own
// (run_plan).  left[i]: steps plan i has left.
void run_plans_batched(Plan *const *plans, int count, size_t steps, hipStream_t st,
                       long long *left)
{
    for (int i = 0; i < count; i++) {
        if (!plans[i])
            fail("NULL plan %d", i);
        for (int j = 0; j < i; j++)
            if (plans[j] == plans[i])
                fail("plan %d listed twice", i);
    }
    std::vector<Plan *> group[sw::kMaxR + 1];
    for (int i = 0; i < count; i++) {
        Plan *p = plans[i];
        if (p->bshot) {
            hip_check(hipStreamWaitEvent(st, p->last, 0), "hipStreamWaitEvent");
            group[p->pr.r].push_back(p);
        } else {
            run_plan(*p, steps, st);
        }
    }
    for (int R = 1; R <= sw::kMaxR; R++) {
        std::vector<Plan *> &g = group[R];
        std::vector<size_t> want(g.size());
        for (size_t i = 0; i < g.size(); i++)
            want[i] = std::min(g[i]->slots.size(), g[i]->done + steps);
        for (;;) {
            std::vector<Plan *> act;
            size_t seg = std::numeric_limits<size_t>::max();
            for (size_t i = 0; i < g.size(); i++)
                if (g[i]->done < want[i]) {
                    act.push_back(g[i]);
                    seg = std::min(seg, want[i] - g[i]->done);
                }
            if (act.empty())
                break;
            for (size_t c0 = 0; c0 < act.size(); c0 += sw::kMaxBatch) {
                const int cnt = (int)std::min<size_t>(sw::kMaxBatch, act.size() - c0);
                sw::BatchArgs S{}, Rc{};
                S.count = cnt;
                S.first[0] = 0;
                int rc = 0;
                for (int i = 0; i < cnt; i++) {
                    const Plan *p = act[c0 + i];
                    S.shot[i] = p->bshot;
                    S.n[i] = (int)(p->a.n0 + p->done);
                    S.first[i + 1] = S.first[i] + p->bblocks;
                    if (p->rblocks) {
                        Rc.shot[rc] = p->bshot;
                        Rc.n[rc] = S.n[i];
                        Rc.first[rc + 1] = Rc.first[rc] + p->rblocks;
                        rc++;
                    }
                }
                Rc.count = rc;
                for (size_t k = 0; k < seg; k++) {
                    kStepBatch[R](S, st);
                    if (rc)
                        hipLaunchKernelGGL((sw::k_receivers_batch<kRecPerBlock, kRecChunk>),
                                           dim3((unsigned)Rc.first[rc]), dim3(256), 0, st, Rc);
                    hip_check(hipGetLastError(), "batched step");
                    for (int i = 0; i < cnt; i++)
                        S.n[i]++;
                    for (int i = 0; i < rc; i++)
                        Rc.n[i]++;
                }
            }
            for (Plan *p : act)
                p->done += seg;
        }
    }
    for (int i = 0; i < count; i++) {
        Plan *p = plans[i];
        if (p->bshot)
            hip_check(hipEventRecord(p->last, st), "hipEventRecord");
        left[i] = (long long)(p->slots.size() - p->done);
    }
}

#if SW_DIM == 2
// Whether a plan may run all its steps in one cooperative launch
// (k_step2d_persist, run_device only): stride 0 on the fused z-marching 2D
// kernel with the cooperative receiver gather, no halo-reaching sources, no
// graph and no second stream, on grids up to kPersistMaxPoints, where a
// step takes a few microseconds and launching it costs as much.
// SIMWAVE_HIP_PERSIST=0 switches it off, =1 takes it on any such grid.
constexpr long long kPersistMaxPoints = 1ll << 21;

void setup_persist(Plan &p)
{
    const char *env = getenv("SIMWAVE_HIP_PERSIST");
    const int want = env ? atoi(env) : -1;
    if (want == 0)
        return;
    const Prepared &pr = p.pr;
    const bool ok = p.a.stride == 0 && pr.fused && !pr.halo_sources && zmarch_2d(pr.r) &&
                    !p.gexec && !(p.flags & kFlagOverlapReceivers) &&
                    !env_flag("SIMWAVE_HIP_REC_THREAD") && p.a.n0 >= 1 &&
                    p.slots.size() < (1u << 21) && pr.P.nbase == nullptr;
    if (!ok || (want != 1 && pr.npts > kPersistMaxPoints))
        return;
    p.bar = p.arena.alloc<unsigned>(2);
}
#endif

Plan *make_plan(const Args &a, hipStream_t st, unsigned flags)
{
    validate(a);
    keep_pool_memory();
    std::unique_ptr<Plan> p(new Plan(a, st, flags));
    if (flags & kFlagOverlapReceivers)
        p->lease.acquire(true);
    p->slots = plan_slots(a);
    for (size_t idx = 0; idx < p->slots.size(); idx++) {
        const StepSlots &s = p->slots[idx];
        size_t m = std::max({s.prev, s.cur, s.next});
        if (s.swap_after)
            m = std::max({m, s.swap_a, s.swap_b});
        if (m >= a.num_snapshots)
            fail("slot %zu used by step %zu is outside the %zu snapshots", m,
                 a.n0 + idx, a.num_snapshots);
    }
    p->pr = prepare(a, true, flags, p->arena, st);
    if ((flags & kFlagGraph) && !(flags & (kFlagOverlapReceivers | kFlagTimeSteps)) &&
        p->slots.size() >= kGraphSteps && periodic_slots(p->slots))
        capture_graph(*p);
    setup_batch(*p);
#if SW_DIM == 2
    setup_persist(*p);
#endif
    hip_check(hipEventCreateWithFlags(&p->last, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventRecord(p->last, st), "hipEventRecord");
    return p.release();
}

// Enqueues the next `count` steps of the plan on `st` (after all earlier
// work of the plan, whatever stream it ran on); returns the steps left.
// whole_run: the caller waits for the stream afterwards and checks the
// plan's abort word (run_device), so a 2D plan set up for it may run the
// steps in one cooperative launch.
size_t run_plan(Plan &p, size_t count, hipStream_t st, bool whole_run)
{
    const size_t i0 = p.done, i1 = std::min(p.slots.size(), i0 + count);
    hip_check(hipStreamWaitEvent(st, p.last, 0), "hipStreamWaitEvent");
#if SW_DIM == 2
    if (whole_run && p.bar && i1 > i0) {
        const bool timed = p.flags & kFlagTimeSteps;
        hipEvent_t e[2] = {nullptr, nullptr};
        if (timed)
            for (hipEvent_t &x : e) {
                hip_check(hipEventCreate(&x), "hipEventCreate");
                hip_check(hipEventRecord(x, st), "hipEventRecord");
            }
        const long long n0 = (long long)(p.a.n0 + i0), n1 = (long long)(p.a.n0 + i1 - 1);
        RTable::persist[p.pr.r](p.pr.P, p.a.u, p.pr.npts, (int)n0, (int)n1, st, p.pr.rbox,
                                p.pr.rw, (int)p.a.nrec,
                                p.a.receivers + (n0 - 1) * (long long)p.a.nrec, p.bar);
        if (timed) {
            hip_check(hipEventRecord(e[1], st), "hipEventRecord");
            hip_check(hipEventSynchronize(e[1]), "timed steps");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, e[0], e[1]), "hipEventElapsedTime");
            g_timings[4] = ms * 1e-3;
            g_timings[5] = (double)(i1 - i0);
            for (hipEvent_t x : e)
                (void)hipEventDestroy(x);
        }
        hip_check(hipEventRecord(p.last, st), "hipEventRecord");
        p.done = i1;
        g_timings[6] = 1.0;
        g_timings[7] = 3.0;
        return p.slots.size() - p.done;
    }
#endif
    const size_t npts = (size_t)p.pr.npts;
    real_t *u = p.a.u;
    auto slot_ptr = [&](size_t sl) { return u + sl * npts; };
    auto nop = [](size_t) {};
    auto swap = [&](size_t x, size_t y) {
        const int blocks = (int)std::min<long long>((long long)(npts + 255) / 256, 16384);
        hipLaunchKernelGGL(sw::k_swap, dim3(blocks), dim3(256), 0, st, u + x * npts,
                           u + y * npts, (long long)npts);
    };
    auto no_uploads = [](size_t) { return false; };
    size_t i = i0;
    while (i < i1) {
        if (p.gexec && i % 3 == 0 && i + kGraphSteps <= i1) {
            // replay: entries i .. i+kGraphSteps-1 have the captured slots
            hipLaunchKernelGGL(sw::k_set_int, dim3(1), dim3(1), 0, st, p.nbase,
                               (int)(p.a.n0 + i - 1));
            hip_check(hipGetLastError(), "step base");
            hip_check(hipGraphLaunch(p.gexec, st), "graph launch");
            i += kGraphSteps;
            continue;
        }
        // direct launches up to the next replay position (or the end)
        size_t j = i1;
        if (p.gexec) {
            const size_t next = (i + 2) / 3 * 3;
            if (next > i && next + kGraphSteps <= i1)
                j = next;
        }
        time_loop(p.a, p.pr, p.slots, i, j, p.a.receivers, 0, p.flags, st, slot_ptr, nop,
                  nop, swap, no_uploads, &p.rec);
        i = j;
    }
    hip_check(hipEventRecord(p.last, st), "hipEventRecord");
    p.done = i1;
    g_timings[6] = p.pr.fused ? 1.0 : 0.0;
#if SW_DIM == 3
    g_timings[7] = p.pr.tiled ? 1.0 : p.pr.zmarch ? 2.0 : 0.0;
#else
    g_timings[7] = p.pr.fused && zmarch_2d(p.pr.r) ? 2.0 : 0.0;
#endif
    return p.slots.size() - p.done;
}

// One whole shot on `stream`; returns after the stream has drained (wall
// seconds of the call).
double run_device(const Args &a, hipStream_t st, unsigned flags)
{
    const double t0 = now_s();
    for (double &t : g_timings)
        t = 0.0;
    std::unique_ptr<Plan> p(make_plan(a, st, flags));
    const double tl0 = now_s();
    run_plan(*p, p->slots.size(), st, true);
    hip_check(hipStreamSynchronize(st), "time loop");
    if (p->bar) {
        int aborted = 0;
        hip_check(hipMemcpy(&aborted, p->bar + 1, sizeof(int), hipMemcpyDeviceToHost),
                  "abort word");
        if (aborted == 1) {
            // not every block was resident (the device was shared): nothing
            // was written, so the steps run launch by launch instead
            p->done = 0;
            run_plan(*p, p->slots.size(), st, false);
            hip_check(hipStreamSynchronize(st), "time loop");
        } else if (aborted) {
            fail("the whole-run 2D launch gave up waiting at a grid barrier; run with "
                 "SIMWAVE_HIP_PERSIST=0");
        }
    }
    g_timings[2] = now_s() - tl0;
    const double total = now_s() - t0;
    g_timings[0] = total;
    return total;
}

// Address of this library's scratch array on the current device, looked up
// once per device (with that device current) and kept.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];

void *scratch_region(hipStream_t st)
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (st) {
        int sdev = -1;
        hip_check(hipStreamGetDevice(st, &sdev), "hipStreamGetDevice");
        if (sdev != dev)
            fail("the stream belongs to device %d but device %d is current; make the "
                 "stream's device current", sdev, dev);
    }
    if (dev < 0 || dev >= kMaxDevices)
        fail("device %d is beyond the %d the library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p)
            fail("no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// One whole shot on `stream` like run_device, with every library buffer in
// the static scratch array: no device allocation or free.  Returns after the
// stream has drained, so that the next call may reuse the array; calls that
// overlap on one device (from several threads) race on it.
double run_device_noalloc(const Args &a, hipStream_t st, unsigned flags)
{
    if (flags & ~kFlagGeneric)
        fail("flags 0x%x: the allocation-free entry takes only flag %u (generic "
             "kernels); step timing and the receiver stream need allocations",
             flags, kFlagGeneric);
    const double t0 = now_s();
    for (double &t : g_timings)
        t = 0.0;
    validate(a);
    const std::vector<StepSlots> slots = plan_slots(a);
    for (size_t idx = 0; idx < slots.size(); idx++) {
        const StepSlots &s = slots[idx];
        size_t m = std::max({s.prev, s.cur, s.next});
        if (s.swap_after)
            m = std::max({m, s.swap_a, s.swap_b});
        if (m >= a.num_snapshots)
            fail("slot %zu used by step %zu is outside the %zu snapshots", m,
                 a.n0 + idx, a.num_snapshots);
    }
    // the per-point arrays (q, e and, for the tiled 3D kernel, C) decide
    // whether the grid fits; the small tables follow them
    const size_t npts = a.nz * a.nx * (SW_DIM == 3 ? a.ny : 1);
    // variable density from radius sw::kVdPre on adds the density sums
    const size_t per_point =
        ((SW_DIM == 3 ? 3 : 2) + (SW_VD && a.space_order / 2 >= (size_t)sw::kVdPre ? SW_DIM : 0)) *
        sizeof(real_t);
    if (npts > (kScratchBytes - (1u << 20)) / per_point)
        fail("grid of %zu points needs %zu bytes of per-point arrays, more than the "
             "%zu-byte static scratch holds (at most %zu points); use "
             "simwave_hip_forward_device for this grid",
             npts, npts * per_point, kScratchBytes, (kScratchBytes - (1u << 20)) / per_point);
    void *base = scratch_region(st);
    DeviceArena arena(st, base, kScratchBytes);
    const Prepared pr = prepare(a, true, flags, arena, st);
    real_t *u = a.u;
    auto slot_ptr = [&](size_t sl) { return u + sl * (size_t)pr.npts; };
    auto nop = [](size_t) {};
    auto swap = [&](size_t x, size_t y) {
        const long long n = pr.npts;
        const int blocks = (int)std::min<long long>((n + 255) / 256, 16384);
        hipLaunchKernelGGL(sw::k_swap, dim3(blocks), dim3(256), 0, st, u + x * n, u + y * n, n);
    };
    auto no_uploads = [](size_t) { return false; };
    const double tl0 = now_s();
    time_loop(a, pr, slots, 0, slots.size(), a.receivers, 0, flags, st, slot_ptr, nop, nop,
              swap, no_uploads);
    hip_check(hipStreamSynchronize(st), "time loop");
    g_timings[2] = now_s() - tl0;
    g_timings[0] = now_s() - t0;
    return g_timings[0];
}

// ------------------------------------------------------------- adjoint
// Reverse interior update of one step: the z-marching kernels (3D, 2D);
// the whole-grid k_adj_update for radii they do not carry.
#if SW_DIM == 3
template <int R>
void adj_update_z(const sw::Params &P, hipStream_t st, LaunchTimer tm, const real_t *vel,
                  real_t *ub, const real_t *W, real_t *ub1, real_t *ub2, const real_t *u0,
                  const real_t *u1, const real_t *u2, real_t *gacc, int image)
{
    const dim3 grid((unsigned)((P.ny + 63) / 64),
                    (unsigned)((P.nx + sw::kAdjLines - 1) / sw::kAdjLines),
                    (unsigned)((P.nz + sw::kAdjZChunk - 1) / sw::kAdjZChunk));
    launch_timed(sw::k_adj_update_z<R>, grid, sw::adj_block(), st, tm, P, vel, ub, W, ub1, ub2,
                 u0, u1, u2, gacc, image);
}
#endif

void adj_update(const sw::Params &P, hipStream_t st, LaunchTimer tm, const real_t *vel,
                real_t *ub, const real_t *W, real_t *ub1, real_t *ub2, const real_t *u0,
                const real_t *u1, const real_t *u2, real_t *gacc, int image)
{
#if SW_DIM == 3
    switch (P.r) {
#define SW_ADJ_CASE(R)                                                              \
    case R:                                                                         \
        adj_update_z<R>(P, st, tm, vel, ub, W, ub1, ub2, u0, u1, u2, gacc, image);  \
        return;
    SW_ADJ_CASE(1) SW_ADJ_CASE(2) SW_ADJ_CASE(3) SW_ADJ_CASE(4) SW_ADJ_CASE(5)
    SW_ADJ_CASE(6) SW_ADJ_CASE(7) SW_ADJ_CASE(8) SW_ADJ_CASE(9) SW_ADJ_CASE(10)
#undef SW_ADJ_CASE
    default:
        break;
    }
#else
    switch (P.r) {
#define SW_ADJ_CASE(R)                                                              \
    case R:                                                                         \
        launch_timed(sw::k_adj_update_z2d<R>,                                      \
                     dim3((unsigned)((P.nx + 255) / 256),                           \
                          (unsigned)((P.nz + sw::kAdjZChunk - 1) / sw::kAdjZChunk)), \
                     dim3(256), st, tm, P, vel, ub, W, ub1, ub2, u0, u1, u2, gacc, image); \
        return;
    SW_ADJ_CASE(1) SW_ADJ_CASE(2) SW_ADJ_CASE(3) SW_ADJ_CASE(4) SW_ADJ_CASE(5)
    SW_ADJ_CASE(6) SW_ADJ_CASE(7) SW_ADJ_CASE(8) SW_ADJ_CASE(9) SW_ADJ_CASE(10)
#undef SW_ADJ_CASE
    default:
        break;
    }
#endif
    hipLaunchKernelGGL(sw::k_adj_update, sw::adj_grid(P), sw::adj_block(), 0, st, P, vel, ub, W,
                       ub1, ub2, u0, u1, u2, gacc, image);
}

// Shot record and its adjoint: runs the forward of a constant-density shot
// (device arrays, as the device entry; stride 0, steps 1..T) saving the two
// time levels (u^m, u^{m+1}) every `K` steps, then walks the checkpoint
// segments backwards: recomputes each segment's states, then applies the
// transpose of every step (sw_adjoint.hpp) to the adjoint of the shot
// record `rec_adj` ([T][num_receivers], device).  Accumulates into
// `grad` (float64, one value per extended grid point) the velocity
// gradient and writes into `gw` ([T][num_sources], device, or null) the
// wavelet gradient.
// The receiver-adjoint spread as a gather (k_adj_rec_gather): the touched
// points, where each one's entries start, how many, and the sorted entries'
// receivers and weights.  nu == 0: the atomic spread (k_adj_receivers) is
// used instead (no receivers, more than 2^31 points or entries, or
// SIMWAVE_HIP_REC_ATOMIC set).
struct RecGather {
    const unsigned *pts = nullptr, *start = nullptr, *cnt = nullptr, *qs = nullptr;
    const real_t *ws = nullptr;
    int nu = 0;
};

RecGather build_rec_gather(const Args &a, const Prepared &pr, DeviceArena &arena,
                           hipStream_t st)
{
    RecGather g;
    if (!a.nrec || env_flag("SIMWAVE_HIP_REC_ATOMIC"))
        return g;
    const size_t dims[3] = {a.nz, a.nx, SW_DIM == 3 ? a.ny : 1};
    const std::vector<Box> rb = make_boxes("receiver", a.nrec, a.rint, a.rint_size, a.roff,
                                           a.rvals_size, dims);
    std::vector<unsigned> off(a.nrec + 1);
    unsigned long long tot = 0;
    for (size_t q = 0; q < a.nrec; q++) {
        off[q] = (unsigned)tot;
        unsigned long long cnt = 1;
        for (int d = 0; d < SW_DIM; d++)
            cnt *= (unsigned long long)(rb[q].hi[d] - rb[q].lo[d] + 1);
        tot += cnt;
        if (tot >= (1ull << 31))
            return g;
    }
    if (pr.npts >= (1ll << 31) || a.nrec >= (1u << 31))
        return g;
    off[a.nrec] = (unsigned)tot;
    const unsigned n = (unsigned)tot;
    const int nrec = (int)a.nrec;
    unsigned *d_off = arena.alloc<unsigned>(a.nrec + 1);
    hip_check(hipMemcpyAsync(d_off, off.data(), (a.nrec + 1) * sizeof(unsigned),
                             hipMemcpyHostToDevice, st), "copy receiver entry offsets");
    unsigned *key = arena.alloc<unsigned>(n), *idx = arena.alloc<unsigned>(n);
    unsigned *key_s = arena.alloc<unsigned>(n), *idx_s = arena.alloc<unsigned>(n);
    hipLaunchKernelGGL(sw::k_adj_rec_entries, dim3((unsigned)nrec), dim3(256), 0, st, pr.P,
                       pr.rbox, nrec, d_off, key, idx);
    hip_check(hipGetLastError(), "k_adj_rec_entries launch");
    int end_bit = 1;
    while (end_bit < 32 && (1ll << end_bit) < pr.npts)
        end_bit++;
    // a stable sort: entries of one point keep their receiver order
    size_t tb = 0;
    hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, key_s, idx, idx_s, (int)n,
                                                 0, end_bit, st), "radix sort size");
    void *tmp = arena.alloc<unsigned char>(std::max<size_t>(tb, 1));
    hip_check(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, key_s, idx, idx_s, (int)n, 0,
                                                 end_bit, st), "radix sort");
    unsigned *pts = arena.alloc<unsigned>(n), *cnt = arena.alloc<unsigned>(n);
    int *d_nu = arena.alloc<int>(1);
    size_t tb2 = 0;
    hip_check(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, key_s, pts, cnt, d_nu, (int)n,
                                                    st), "run-length size");
    void *tmp2 = arena.alloc<unsigned char>(std::max<size_t>(tb2, 1));
    hip_check(hipcub::DeviceRunLengthEncode::Encode(tmp2, tb2, key_s, pts, cnt, d_nu, (int)n,
                                                    st), "run-length encode");
    int nu = 0;
    hip_check(hipMemcpyAsync(&nu, d_nu, sizeof(int), hipMemcpyDeviceToHost, st),
              "copy point count");
    hip_check(hipStreamSynchronize(st), "receiver entries");
    unsigned *start = arena.alloc<unsigned>((size_t)std::max(nu, 1));
    size_t tb3 = 0;
    hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, cnt, start, nu, st), "scan size");
    void *tmp3 = arena.alloc<unsigned char>(std::max<size_t>(tb3, 1));
    hip_check(hipcub::DeviceScan::ExclusiveSum(tmp3, tb3, cnt, start, nu, st), "scan");
    unsigned *qs = arena.alloc<unsigned>(n);
    real_t *ws = arena.alloc<real_t>(n);
    hipLaunchKernelGGL(sw::k_adj_rec_permute, dim3((n + 255) / 256), dim3(256), 0, st, pr.P,
                       pr.rbox, pr.rw, nrec, d_off, key_s, idx_s, n, qs, ws);
    hip_check(hipGetLastError(), "k_adj_rec_permute launch");
    for (void *p : {(void *)key, (void *)idx, (void *)key_s, (void *)idx_s, tmp, tmp2, tmp3,
                    (void *)d_off, (void *)d_nu})
        arena.release(p);
    g.pts = pts;
    g.start = start;
    g.cnt = cnt;
    g.qs = qs;
    g.ws = ws;
    g.nu = nu;
    return g;
}

double run_adjoint(const Args &a, const real_t *rec_adj, double *grad, real_t *gw, size_t K,
                   hipStream_t st, unsigned flags)
{
    const double t0 = now_s();
    for (double &t : g_timings)
        t = 0.0;
    validate(a);
#if SW_VD
    (void)rec_adj; (void)grad; (void)gw; (void)K; (void)st; (void)flags;
    fail("the adjoint covers constant density only");
    return -1.0;
#else
    if (a.stride != 0)
        fail("the adjoint needs saving_stride 0 (got %zu)", a.stride);
    if (a.n0 != 1 || a.n1 < 1)
        fail("the adjoint runs steps 1..T (got %zu..%zu)", a.n0, a.n1);
    if (a.num_snapshots < 3)
        fail("the adjoint needs the three wavefield slots");
    if (K < 1)
        fail("checkpoint interval must be >= 1");
    if (!grad)
        fail("NULL gradient array");
    if (a.nrec && !rec_adj)
        fail("NULL shot-record adjoint");
    const unsigned known = kAdjImage | kAdjResidual | kFlagOverlapReceivers | kAdjTimeUpdate;
    if (flags & ~known)
        fail("unknown adjoint flags 0x%x", flags & ~known);
    const bool image = flags & kAdjImage;
    const int r = (int)(a.space_order / 2);
    const size_t dims[3] = {a.nz, a.nx, SW_DIM == 3 ? a.ny : (size_t)(3 * r + 1)};
    for (int d = 0; d < SW_DIM; d++)
        if (dims[d] < (size_t)(3 * r + 1))
            fail("the adjoint needs at least 3r+1 points per axis");
    OverlapLease lease;
    if (flags & kFlagOverlapReceivers)
        lease.acquire(true);
    keep_pool_memory();
    struct PoolTrim {
        // the sweep's levels and checkpoints are the largest allocations of
        // the library: hand them back to the device once it has finished
        ~PoolTrim()
        {
            hipMemPool_t pool;
            if (default_pool(pool))
                (void)hipMemPoolTrimTo(pool, kPoolKeepBytes);
        }
    } trim;
    DeviceArena arena(st);
    Prepared pr = prepare(a, true, 0u, arena, st);
    if (pr.halo_sources)
        fail("the adjoint needs every source box inside the interior");
    const RecGather rg = build_rec_gather(a, pr, arena, st);
    const Params &P = pr.P;
    const size_t npts = (size_t)pr.npts;
    const size_t T = a.n1;
    // K >= T: every level stays in HBM (no checkpoints, no recomputation)
    const bool store_all = K >= T;
    if (store_all)
        K = T;
    const size_t ncp = store_all ? 1 : (T + K - 1) / K;
    const int nb = (int)std::min<long long>(((long long)npts + 255) / 256, 16384);
    auto zero = [&](real_t *p, size_t count) {
        hipLaunchKernelGGL(sw::k_fill, dim3(nb), dim3(256), 0, st, p, (long long)count,
                           (real_t)0);
    };
    real_t *cp = store_all ? nullptr : arena.alloc<real_t>(2 * ncp * npts);
    real_t *seg = arena.alloc<real_t>((K + 2) * npts);
    real_t *ubar = arena.alloc<real_t>(3 * npts);
    real_t *W = arena.alloc<real_t>(npts);
    // the reverse steps add their gradient terms here, in the working
    // precision; every kGradFlush steps (counted over the whole sweep, so
    // the sums do not depend on K) and at the end they go into grad
    real_t *gacc = arena.alloc<real_t>(npts);
    size_t reverse_steps = 0;
    // the forward's coefficient C = q / den per point (the tiled kernel's
    // copy where it was built)
#if SW_DIM == 3
    const real_t *Cq = pr.qc;
#else
    const real_t *Cq = nullptr;
#endif
    if (!Cq) {
        real_t *c = arena.alloc<real_t>(npts);
        hipLaunchKernelGGL(sw::k_scaled_q, dim3(nb), dim3(256), 0, st, P.q, P.e, c,
                           (long long)npts);
        Cq = c;
    }
    auto flush_grad = [&]() {
        hipLaunchKernelGGL(sw::k_grad_flush, dim3(nb), dim3(256), 0, st, gacc, grad,
                           (long long)npts);
    };
    // the shot starts from rest whatever the caller's slots hold, so that
    // the all-levels and checkpointed sweeps see the same initial state
    zero(a.u, 3 * npts);
    zero(seg, (K + 2) * npts);
    zero(ubar, 3 * npts);
    zero(gacc, npts);
    hipLaunchKernelGGL(sw::k_fill_f64, dim3(nb), dim3(256), 0, st, grad, (long long)npts);
    if (gw && a.nsrc)
        hipLaunchKernelGGL(sw::k_fill, dim3(1), dim3(256), 0, st, gw,
                           (long long)(T * a.nsrc), (real_t)0);
    hip_check(hipGetLastError(), "adjoint set-up");

    // forward with checkpoints
    std::vector<StepSlots> plan = plan_slots(a);
    real_t *u = a.u;
    auto slot_ptr = [&](size_t sl) { return u + sl * npts; };
    auto save = [&](size_t idx) {
        if (idx % K)
            return;
        const StepSlots &s = plan[idx];
        const size_t j = idx / K;
        hip_check(hipMemcpyAsync(cp + (2 * j) * npts, slot_ptr(s.prev), npts * sizeof(real_t),
                                 hipMemcpyDeviceToDevice, st), "checkpoint");
        hip_check(hipMemcpyAsync(cp + (2 * j + 1) * npts, slot_ptr(s.cur),
                                 npts * sizeof(real_t), hipMemcpyDeviceToDevice, st),
                  "checkpoint");
    };
    auto nop = [](size_t) {};
    auto noswap = [](size_t, size_t) {};
    auto no_uploads = [](size_t) { return false; };
    auto seg_ptr = [&](size_t l) { return seg + l * npts; };
    if (store_all) {
        // level l of the run in seg + l*npts
        std::vector<StepSlots> sp;
        for (size_t n = 1; n <= T; n++)
            sp.push_back(StepSlots{n - 1, n, n + 1, false, 0, 0});
        time_loop(a, pr, sp, 0, sp.size(), a.receivers, 0, flags & kFlagOverlapReceivers, st,
                  seg_ptr, nop, nop, noswap, no_uploads);
        // the caller's slots end as after a plain forward: level l in l % 3
        for (size_t l = T - 1; l <= T + 1; l++)
            hip_check(hipMemcpyAsync(slot_ptr(l % 3), seg_ptr(l), npts * sizeof(real_t),
                                     hipMemcpyDeviceToDevice, st),
                      "final levels");
    } else {
        time_loop(a, pr, plan, 0, plan.size(), a.receivers, 0, flags & kFlagOverlapReceivers,
                  st, slot_ptr, save, nop, noswap, no_uploads);
    }
    if ((flags & kAdjResidual) && a.nrec) {
        // the shot record rows 0 .. T-1 are complete on `st`
        real_t *res = arena.alloc<real_t>(T * a.nrec);
        const long long n = (long long)(T * a.nrec);
        hipLaunchKernelGGL(sw::k_residual,
                           dim3((unsigned)std::min<long long>((n + 255) / 256, 16384)),
                           dim3(256), 0, st, a.receivers, rec_adj, res, n);
        hip_check(hipGetLastError(), "residual");
        rec_adj = res;
    }
    const double t_fwd = now_s();

    // event pairs around every reverse update kernel (kAdjTimeUpdate)
    struct UpdateTimer {
        std::vector<hipEvent_t> ev;
        size_t used = 0;
        ~UpdateTimer()
        {
            for (hipEvent_t e : ev)
                if (e)
                    (void)hipEventDestroy(e);
        }
    } utime;
    if (flags & kAdjTimeUpdate) {
        utime.ev.assign(2 * T, nullptr);
        for (hipEvent_t &e : utime.ev)
            hip_check(hipEventCreate(&e), "hipEventCreate");
    }

    // backward over the checkpoint segments
    Args seg_args = a;
    seg_args.nrec = 0;
    for (size_t j = ncp; j-- > 0;) {
        const size_t m = j * K;                   // levels m, m+1 saved
        const size_t n_lo = m + 1, n_hi = std::min(m + K, T);
        if (!store_all) {
            hip_check(hipMemcpyAsync(seg, cp + (2 * j) * npts, 2 * npts * sizeof(real_t),
                                     hipMemcpyDeviceToDevice, st),
                      "restore checkpoint");
            std::vector<StepSlots> sp;
            for (size_t n = n_lo; n <= n_hi; n++)
                sp.push_back(StepSlots{n - 1 - m, n - m, n + 1 - m, false, 0, 0});
            seg_args.n0 = n_lo;
            seg_args.n1 = n_hi;
            time_loop(seg_args, pr, sp, 0, sp.size(), nullptr, 0, 0u, st, seg_ptr, nop, nop,
                      noswap, no_uploads);
        }
        for (size_t n = n_hi; n >= n_lo; n--) {
            const size_t k = n + 1;
            real_t *ub = ubar + (k % 3) * npts;
            real_t *ub1 = ubar + ((k - 1) % 3) * npts;
            real_t *ub2 = ubar + ((k - 2) % 3) * npts;
            if (k <= T && a.nrec && rg.nu)
                hipLaunchKernelGGL(sw::k_adj_rec_gather, dim3((unsigned)((rg.nu + 255) / 256)),
                                   dim3(256), 0, st, rg.pts, rg.start, rg.cnt, rg.nu, rg.qs,
                                   rg.ws, rec_adj + (k - 1) * a.nrec, ub);
            else if (k <= T && a.nrec)
                hipLaunchKernelGGL(sw::k_adj_receivers, dim3((unsigned)a.nrec), dim3(256), 0,
                                   st, P, pr.rbox, pr.rw, (int)a.nrec,
                                   rec_adj + (k - 1) * a.nrec, ub);
#if SW_DIM == 3
            const int axes[3] = {0, 1, 2};
            const long long lines[3] = {
                (long long)(P.nx - 2 * r) * (P.ny - 2 * r),
                (long long)(P.nz - 2 * r) * (P.ny - 2 * r),
                (long long)(P.nz - 2 * r) * (P.nx - 2 * r)};
            const int nph = 3;
#else
            const int axes[2] = {0, 2};
            const long long lines[2] = {(long long)(P.nx - 2 * r), (long long)(P.nz - 2 * r)};
            const int nph = 2;
#endif
            for (int ph = 0; ph < nph; ph++)   // z first: the reverse of y, x, z
                hipLaunchKernelGGL(sw::k_adj_bc_phase,
                                   dim3((unsigned)((lines[ph] + 255) / 256)), dim3(256), 0, st,
                                   P, ub, axes[ph]);
            hipLaunchKernelGGL(sw::k_adj_cw, sw::adj_grid(P), sw::adj_block(), 0, st, P, Cq, ub,
                               W);
            if (a.nsrc && gw)
                hipLaunchKernelGGL(sw::k_adj_sources, dim3((unsigned)a.nsrc), dim3(256), 0, st,
                                   P, ub, gw + (n - 1) * a.nsrc);
            const bool timed = utime.used + 2 <= utime.ev.size();
            LaunchTimer tm;
            if (timed) {
                tm = LaunchTimer{utime.ev[utime.used], utime.ev[utime.used + 1]};
                utime.used += 2;
            }
            adj_update(P, st, tm, a.vel, ub, W, ub1, ub2, seg + (n + 1 - m) * npts,
                       seg + (n - m) * npts, seg + (n - 1 - m) * npts, gacc, image ? 1 : 0);
            if (++reverse_steps % kGradFlush == 0)
                flush_grad();
            hip_check(hipGetLastError(), "adjoint step");
            if (n == n_lo)
                break;
        }
    }
    flush_grad();
    hip_check(hipStreamSynchronize(st), "adjoint");
    if (utime.used) {
        double sec = 0.0;
        for (size_t i = 0; i < utime.used; i += 2) {
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, utime.ev[i], utime.ev[i + 1]),
                      "hipEventElapsedTime");
            sec += ms * 1e-3;
        }
        g_timings[4] = sec;
        g_timings[5] = (double)(utime.used / 2);
    }
    g_timings[1] = t_fwd - t0;
    g_timings[2] = now_s() - t_fwd;
    const double total = now_s() - t0;
    g_timings[0] = total;
    return total;
#endif
}

template <typename F>
double guarded(F &&f)
{
    try {
        g_error.clear();
        return f();
    } catch (const Failure &e) {
        g_error = e.msg;
    } catch (const std::exception &e) {
        g_error = std::string("internal error: ") + e.what();
    } catch (...) {
        g_error = "internal error";
    }
    (void)hipGetLastError();
    return -1.0;
}

}  // namespace

// ================================================================ C ABI
// The parameter lists below are the reference's, variant by variant
// (constant_density/3d/wave.c:23-36, constant_density/2d/wave.c:23-36,
//  variable_density/3d/wave.c:23-36, variable_density/2d/wave.c:23-36).

#if SW_DIM == 3
#define SW_GRID_PARAMS size_t nz, size_t nx, size_t ny, real_t dz, real_t dx, real_t dy
#define SW_GRID_FILL a.nz = nz; a.nx = nx; a.ny = ny; a.dz = dz; a.dx = dx; a.dy = dy;
#define SW_GRID_ARGS nz, nx, ny, dz, dx, dy
#else
#define SW_GRID_PARAMS size_t nz, size_t nx, real_t dz, real_t dx
#define SW_GRID_FILL a.nz = nz; a.nx = nx; a.ny = 1; a.dz = dz; a.dx = dx; a.dy = dx;
#define SW_GRID_ARGS nz, nx, dz, dx
#endif

#if SW_VD
#define SW_MODEL_PARAMS real_t *u, real_t *velocity, real_t *density, real_t *damp
#define SW_MODEL_FILL a.u = u; a.vel = velocity; a.rho = density; a.damp = damp;
#define SW_MODEL_ARGS u, velocity, density, damp
#define SW_COEFF_PARAMS real_t *coeff_order2, real_t *coeff_order1
#define SW_COEFF_FILL a.c2 = coeff_order2; a.c1 = coeff_order1;
#define SW_COEFF_ARGS coeff_order2, coeff_order1
#else
#define SW_MODEL_PARAMS real_t *u, real_t *velocity, real_t *damp
#define SW_MODEL_FILL a.u = u; a.vel = velocity; a.rho = nullptr; a.damp = damp;
#define SW_MODEL_ARGS u, velocity, damp
#define SW_COEFF_PARAMS real_t *coeff
#define SW_COEFF_FILL a.c2 = coeff; a.c1 = nullptr;
#define SW_COEFF_ARGS coeff
#endif

#define SW_FORWARD_PARAMS                                                        \
    SW_MODEL_PARAMS, real_t *wavelet, size_t wavelet_size, size_t wavelet_count, \
        SW_COEFF_PARAMS, size_t *boundary_conditions, size_t *src_points_interval, \
        size_t src_points_interval_size, real_t *src_points_values,              \
        size_t src_points_values_size, size_t *src_points_values_offset,         \
        size_t *rec_points_interval, size_t rec_points_interval_size,            \
        real_t *rec_points_values, size_t rec_points_values_size,                \
        size_t *rec_points_values_offset, real_t *receivers, size_t num_sources, \
        size_t num_receivers, SW_GRID_PARAMS, size_t saving_stride, real_t dt,   \
        size_t begin_timestep, size_t end_timestep, size_t space_order,          \
        size_t num_snapshots

#define SW_FORWARD_ARGS                                                          \
    SW_MODEL_ARGS, wavelet, wavelet_size, wavelet_count, SW_COEFF_ARGS,          \
        boundary_conditions, src_points_interval, src_points_interval_size,      \
        src_points_values, src_points_values_size, src_points_values_offset,     \
        rec_points_interval, rec_points_interval_size, rec_points_values,        \
        rec_points_values_size, rec_points_values_offset, receivers,             \
        num_sources, num_receivers, SW_GRID_ARGS, saving_stride, dt,             \
        begin_timestep, end_timestep, space_order, num_snapshots

static Args fill_args(SW_FORWARD_PARAMS)
{
    Args a{};
    SW_MODEL_FILL
    a.wavelet = wavelet;
    a.wavelet_size = wavelet_size;
    a.wavelet_count = wavelet_count;
    SW_COEFF_FILL
    a.bc = boundary_conditions;
    a.sint = src_points_interval;
    a.sint_size = src_points_interval_size;
    a.svals = src_points_values;
    a.svals_size = src_points_values_size;
    a.soff = src_points_values_offset;
    a.rint = rec_points_interval;
    a.rint_size = rec_points_interval_size;
    a.rvals = rec_points_values;
    a.rvals_size = rec_points_values_size;
    a.roff = rec_points_values_offset;
    a.receivers = receivers;
    a.nsrc = num_sources;
    a.nrec = num_receivers;
    SW_GRID_FILL
    a.stride = saving_stride;
    a.dt = dt;
    a.n0 = begin_timestep;
    a.n1 = end_timestep;
    a.space_order = space_order;
    a.num_snapshots = num_snapshots;
    return a;
}

// Reference FFI name: one variant per library, like the reference's
// runtime-compiled wave.c (middleware.py:26-51).
SW_EXPORT double forward(SW_FORWARD_PARAMS)
{
    return guarded([&] { return run_host(fill_args(SW_FORWARD_ARGS)); });
}

SW_EXPORT double SW_SUFFIX(simwave_hip_forward_)(SW_FORWARD_PARAMS)
{
    return guarded([&] { return run_host(fill_args(SW_FORWARD_ARGS)); });
}

SW_EXPORT double SW_SUFFIX(simwave_hip_forward_device_)(SW_FORWARD_PARAMS, void *stream,
                                                        unsigned flags)
{
    return guarded([&] {
        return run_device(fill_args(SW_FORWARD_ARGS), (hipStream_t)stream, flags);
    });
}

SW_EXPORT double SW_SUFFIX(simwave_hip_forward_device_noalloc_)(SW_FORWARD_PARAMS,
                                                                void *stream,
                                                                unsigned flags)
{
    return guarded([&] {
        return run_device_noalloc(fill_args(SW_FORWARD_ARGS), (hipStream_t)stream, flags);
    });
}

// Host-only helper (no device work): size in bytes of the static scratch
// array of the allocation-free entry.
SW_EXPORT size_t SW_SUFFIX(simwave_hip_scratch_bytes_)(void)
{
    return kScratchBytes;
}

SW_EXPORT void *SW_SUFFIX(simwave_hip_plan_create_)(SW_FORWARD_PARAMS, void *stream,
                                                    unsigned flags)
{
    Plan *out = nullptr;
    const double r = guarded([&] {
        for (double &t : g_timings)
            t = 0.0;
        out = make_plan(fill_args(SW_FORWARD_ARGS), (hipStream_t)stream, flags);
        return 0.0;
    });
    return r < 0 ? nullptr : out;
}

SW_EXPORT long long SW_SUFFIX(simwave_hip_plan_run_)(void *plan, size_t steps, void *stream)
{
    long long left = -1;
    const double r = guarded([&] {
        if (!plan)
            fail("NULL plan");
        left = (long long)run_plan(*static_cast<Plan *>(plan), steps, (hipStream_t)stream);
        return 0.0;
    });
    return r < 0 ? -1 : left;
}

SW_EXPORT int SW_SUFFIX(simwave_hip_plans_run_)(void **plans, int count, size_t steps,
                                                void *stream, long long *left)
{
    const double r = guarded([&] {
        if (count < 0 || (count && (!plans || !left)))
            fail("bad plan list");
        run_plans_batched(reinterpret_cast<Plan *const *>(plans), count, steps,
                          (hipStream_t)stream, left);
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int SW_SUFFIX(simwave_hip_plan_batchable_)(void *plan)
{
    return plan && static_cast<Plan *>(plan)->bshot ? 1 : 0;
}

SW_EXPORT int SW_SUFFIX(simwave_hip_gather_)(real_t *out, const real_t *small, int ndim,
                                              const long long *out_shape,
                                              const long long *small_shape,
                                              const long long *idx0, const long long *idx1,
                                              const long long *idx2, void *stream)
{
    const double r = guarded([&] {
        if (ndim != SW_DIM)
            fail("gather: %d axes for a %dD library", ndim, SW_DIM);
        if (!out || !small || !out_shape || !small_shape || !idx0 || !idx1 ||
            (SW_DIM == 3 && !idx2))
            fail("gather: null argument");
        long long n[3] = {1, 1, 1}, m[3] = {1, 1, 1};
        for (int a = 0; a < SW_DIM; a++) {
            n[a] = out_shape[a];
            m[a] = small_shape[a];
            if (n[a] < 1 || m[a] < 1)
                fail("gather: empty axis %d", a);
        }
        const long long total = n[0] * n[1] * n[2];
        const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(sw::k_gather_axes, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           out, small, n[0], n[1], n[2], m[1], m[2], idx0, idx1,
                           SW_DIM == 3 ? idx2 : nullptr);
        hip_check(hipGetLastError(), "gather launch");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int SW_SUFFIX(simwave_hip_plan_destroy_)(void *plan)
{
    const double r = guarded([&] {
        delete static_cast<Plan *>(plan);
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT double SW_SUFFIX(simwave_hip_gradient_)(SW_FORWARD_PARAMS, const real_t *rec_adjoint,
                                                 double *grad_velocity,
                                                 real_t *grad_wavelet,
                                                 size_t checkpoint_every, void *stream,
                                                 unsigned flags)
{
    return guarded([&] {
        return run_adjoint(fill_args(SW_FORWARD_ARGS), rec_adjoint, grad_velocity,
                           grad_wavelet, checkpoint_every, (hipStream_t)stream, flags);
    });
}

SW_EXPORT const char *simwave_hip_last_error(void)
{
    return g_error.c_str();
}

// Device memory, pinned host memory, copies, zero fills, streams and
// events on the current device, so that a caller holding no other GPU
// runtime can keep a shot resident and drive it on streams of its own.
// Device memory is stream-ordered (the default pool, which keeps freed
// memory for reuse up to kPoolKeepBytes): a free waits for nothing on the
// host, where hipFree would wait for the whole device, so shots that end
// while others run do not stall.  Streams are non-blocking with respect to
// the null stream, at the default priority.  Each returns 0, or -1 with
// simwave_hip_last_error() set.
SW_EXPORT int simwave_hip_device_alloc(unsigned long long bytes, void *stream, void **ptr)
{
    const double r = guarded([&] {
        if (!ptr)
            fail("device_alloc: NULL result pointer");
        *ptr = nullptr;
        keep_pool_memory();
        hip_check(hipMallocAsync(ptr, bytes ? bytes : 1, (hipStream_t)stream),
                  "hipMallocAsync");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

// The memory returns to the pool once the work enqueued on `stream` before
// the call is done; work on other streams that reads it must be complete.
SW_EXPORT int simwave_hip_device_free(void *ptr, void *stream)
{
    const double r = guarded([&] {
        if (ptr)
            hip_check(hipFreeAsync(ptr, (hipStream_t)stream), "hipFreeAsync");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_host_alloc(unsigned long long bytes, void **ptr)
{
    const double r = guarded([&] {
        if (!ptr)
            fail("host_alloc: NULL result pointer");
        *ptr = nullptr;
        hip_check(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault),
                  "hipHostMalloc");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_host_free(void *ptr)
{
    const double r = guarded([&] {
        if (ptr)
            hip_check(hipHostFree(ptr), "hipHostFree");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

// Stream-ordered copy of `bytes` between any two of device, pinned and
// pageable host memory (the direction follows from the pointers).
SW_EXPORT int simwave_hip_copy(void *dst, const void *src, unsigned long long bytes,
                               void *stream)
{
    const double r = guarded([&] {
        if (bytes && (!dst || !src))
            fail("copy: NULL pointer");
        if (bytes)
            hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream),
                      "hipMemcpyAsync");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

// Zeroes `bytes` (a multiple of 4) of device memory with a fill kernel on
// `stream` (a kernel, not a memset node, so that it also behaves under
// graph capture).
SW_EXPORT int simwave_hip_zero(void *dst, unsigned long long bytes, void *stream)
{
    const double r = guarded([&] {
        if (bytes % 4)
            fail("zero: %llu bytes is not a multiple of 4", bytes);
        if (bytes && !dst)
            fail("zero: NULL pointer");
        const long long n = (long long)(bytes / 4);
        if (n) {
            const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 8192);
            hipLaunchKernelGGL(sw::k_zero_u32, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                               (unsigned *)dst, n);
            hip_check(hipGetLastError(), "zero launch");
        }
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_stream_create(void **stream)
{
    const double r = guarded([&] {
        if (!stream)
            fail("stream_create: NULL result pointer");
        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        *stream = s;
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_stream_destroy(void *stream)
{
    const double r = guarded([&] {
        if (stream)
            hip_check(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_stream_synchronize(void *stream)
{
    const double r = guarded([&] {
        hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_event_create(void **event)
{
    const double r = guarded([&] {
        if (!event)
            fail("event_create: NULL result pointer");
        hipEvent_t e = nullptr;
        hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        *event = e;
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_event_record(void *event, void *stream)
{
    const double r = guarded([&] {
        hip_check(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_event_synchronize(void *event)
{
    const double r = guarded([&] {
        hip_check(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_event_destroy(void *event)
{
    const double r = guarded([&] {
        if (event)
            hip_check(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

// Returns the freed memory the current device's default pool keeps for
// reuse to the device (the library's stream-ordered allocations come from
// that pool), so that a free-memory query afterwards sees it.  0, or -1.
SW_EXPORT int simwave_hip_trim_pool(void)
{
    const double r = guarded([&] {
        hipMemPool_t pool;
        if (!default_pool(pool))
            fail("no default memory pool on the current device");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipMemPoolTrimTo(pool, 0), "hipMemPoolTrimTo");
        return 0.0;
    });
    return r < 0 ? -1 : 0;
}

SW_EXPORT int simwave_hip_last_timings(double *out, int n)
{
    const int m = std::min(n, 8);
    for (int i = 0; i < m; i++)
        out[i] = g_timings[i];
    return m;
}

// Host-only helper (no device work): the fixed-divisor parameters the
// kernels use for a divisor b.  Returns 0 when IEEE division is used instead.
SW_EXPORT int SW_SUFFIX(simwave_hip_fixed_div_)(real_t b, real_t *y, real_t *lo,
                                                 real_t *hi)
{
    sw::FixedDiv d{};
    if (!make_fixed_div(b, d))
        return 0;
    *y = d.y;
    *lo = d.lo;
    *hi = d.hi;
    return 1;
}

// Host-only helper (no device work): 1 when the 3D tiled kernel of a float32
// library forms the quotients sum/b as RN(sum * RN64(1/b)) for the divisor b
// (see f64_quotient_exact), 0 otherwise.
SW_EXPORT int SW_SUFFIX(simwave_hip_f64_quotient_)(real_t b)
{
    return f64_quotient_exact(b) ? 1 : 0;
}

SW_EXPORT const char *simwave_hip_variant(void)
{
    return SW_VARIANT_STR;
}
