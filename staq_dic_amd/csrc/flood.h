// Bit-parallel region fill of a window held as 64-bit column words (row
// counts up to 64): the connected component of the nonzero pixels through a
// seed, as bwselect (funICGNQuadtree.m:163-176, funICGN_Subpb1.m:93,160-165)
// computes it.  Shared by the wavefront kernels and the workgroup kernel.
#pragma once

#include "common.h"

namespace staq {
namespace {

// Bits of bw connected to the seeds s (s a subset of bw) along the word: a
// carry entering a run of ones at a seed runs to the run's top, so
// (bw + s) ^ bw marks seed-to-top; the bit-reversed word gives seed-to-bottom.
__device__ __forceinline__ unsigned long long run_fill(unsigned long long bw, unsigned long long s)
{
    const unsigned long long up = ((bw + s) ^ bw) & bw;
    const unsigned long long rb = __builtin_bitreverse64(bw), rs = __builtin_bitreverse64(s);
    const unsigned long long dn = __builtin_bitreverse64(((rb + rs) ^ rb) & rb);
    return up | dn | s;
}

// Connected component (4- or 8-connected) through the seed of the window
// whose column j is the 64-bit word bw held by lane j (bit i = row i).
__device__ unsigned long long wave_flood(unsigned long long bw, int ny, int seed_r, int seed_c,
                                         int conn)
{
    const int lane = threadIdx.x & 63;
    unsigned long long sel = 0ull;
    if (lane >= ny) bw = 0ull;
    if (seed_c >= 1 && seed_c <= ny && seed_r >= 1 && seed_r <= 64) {
        // start from a cross: the seed row's horizontal run through the seed
        // (one ballot gives the row as a word) and the vertical runs of its
        // columns; every pixel of it is 4-connected to the seed
        const unsigned long long row = __ballot((bw >> (seed_r - 1)) & 1ull);
        const unsigned long long sb = 1ull << (seed_c - 1);
        if (row & sb) {
            const unsigned long long run = run_fill(row, sb);
            sel = ((run >> lane) & 1ull) ? (1ull << (seed_r - 1)) : 0ull;
            sel = run_fill(bw, sel);
        }
    }
    for (;;) {
        unsigned long long l = dpp_u64<0x138>(sel);  // wave_shr:1 - column j-1
        unsigned long long r = dpp_u64<0x130>(sel);  // wave_shl:1 - column j+1
        if (conn == 8) {
            l |= (l << 1) | (l >> 1);
            r |= (r << 1) | (r >> 1);
        }
        // across one column, then along whole columns at once
        const unsigned long long m = run_fill(bw, bw & (sel | l | r));
        const bool ch = m != sel;
        sel = m;
        if (__ballot(ch) == 0ull) break;
    }
    return sel;
}

}  // namespace
}  // namespace staq
