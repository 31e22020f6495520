// Batched step kernels for the shot server on gfx950.
//
// One launch advances several stride-0 shots of the same radius by one time
// step each: the blocks of the launch are the shots' own step grids laid
// end to end, and every block finds its shot by a search over the prefix
// of block counts (kernel arguments), then runs the single-shot block body
// (step3d_z_block / step2d_z_block / receivers_coop_block) on that shot's
// arrays with its step number.  Same operations per point as the
// single-shot kernels, so a batched shot gives the same bits as it alone.
// Slots follow the stride-0 cycle of the reference's time loop
// (wave.c:112-124): step n reads levels n-1 (slot (n-1) mod 3) and n (slot
// n mod 3) and writes level n+1 (slot (n+1) mod 3); its receivers read
// level n into row n-1 of the shot record.
#pragma once
#include "sw_kernels.hpp"

namespace sw {

// One shot's constants for batched launches, kept on the device by its plan.
struct BatchShot {
    Params P;                  // run constants (no graph step base)
    LapDiv D;                  // Laplacian quotient divisors (3D)
    real_t *u;                 // the three wavefield slots
    real_t *rec;               // shot record [steps][nrec]
    const Box *rbox;
    const real_t *rw;
    long long npts;            // slot stride (points)
    int nrec;
    int zlen;                  // z chunk of the step grid
    int gx, gy;                // step grid: blocks along the contiguous axis, then x
};

constexpr int kMaxBatch = 48;

// Arguments of one batched launch (by value): the shots, their step numbers
// and the first block of each (first[count] = blocks of the launch).
struct BatchArgs {
    const BatchShot *shot[kMaxBatch];
    int n[kMaxBatch];
    int first[kMaxBatch + 1];
    int count;
};

__device__ __forceinline__ int batch_shot(const BatchArgs &A, int b)
{
    int lo = 0, hi = A.count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (A.first[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

#if SW_DIM == 3
template <int R>
__global__ void __launch_bounds__(64 * kZLines)
k_step3d_z_batch(BatchArgs A)
{
    const int s = batch_shot(A, blockIdx.x);
    const BatchShot *__restrict__ b = A.shot[s];
    const Params P = b->P;
    const LapDiv D = b->D;
    const int n = A.n[s];
    const int local = blockIdx.x - A.first[s];
    const int gx = b->gx, gy = b->gy;
    const int bx = local % gx, t = local / gx;
    const long long np = b->npts;
    real_t *u = b->u;
    step3d_z_block<R>(P, D, u + (long long)(n % 3) * np, u + (long long)((n - 1) % 3) * np,
                      u + (long long)((n + 1) % 3) * np, n, b->zlen, bx, t % gy, t / gy);
}
#else
template <int R>
__global__ void __launch_bounds__(kZ2Block)
k_step2d_z_batch(BatchArgs A)
{
    const int s = batch_shot(A, blockIdx.x);
    const BatchShot *__restrict__ b = A.shot[s];
    const Params P = b->P;
    const int n = A.n[s];
    const int local = blockIdx.x - A.first[s];
    const int gx = b->gx;
    const long long np = b->npts;
    real_t *u = b->u;
    step2d_z_block<R>(P, u + (long long)(n % 3) * np, u + (long long)((n - 1) % 3) * np,
                      u + (long long)((n + 1) % 3) * np, n, b->zlen, local % gx, local / gx);
}
#endif

template <int RPB, int CH>
__global__ void __launch_bounds__(256)
k_receivers_batch(BatchArgs A)
{
    const int s = batch_shot(A, blockIdx.x);
    const BatchShot *__restrict__ b = A.shot[s];
    const Params P = b->P;
    const int n = A.n[s];
    receivers_coop_block<RPB, CH>(P, b->u + (long long)(n % 3) * b->npts, b->rbox, b->rw,
                                  b->nrec, b->rec + (long long)(n - 1) * b->nrec,
                                  blockIdx.x - A.first[s]);
}

}  // namespace sw
