// Block-wide exclusive scan of one int per thread (64-wide wavefronts), shared by the runtime's
// scans and the binning kernels that scan inside another kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace mdg {

__device__ inline int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across the block (at most 32 wavefronts; sh holds 33
// ints); returns the thread's offset and the block total in *tot.
__device__ inline int block_excl_scan(int v, int* sh, int* tot)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    if (w == 0) {
        int s  = lane < nw ? sh[lane] : 0;
        int si = wave_incl_scan(s);
        if (lane < nw) sh[lane] = si - s;
        if (lane == nw - 1) sh[32] = si;
    }
    __syncthreads();
    int r = inc - v + sh[w];
    *tot  = sh[32];
    __syncthreads();
    return r;
}

}  // namespace mdg
