// common.hpp -- device-side building blocks shared by every kernel file:
// field layout, Philox-4x32-10, Box-Muller, compensated (Neumaier / TwoSum)
// block reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace askja {

// Sites are tiled by 64 (one wavefront) for the register kernels (N <= 3):
// element e of link l lives at ((l / 64) * msz + e) * 64 + l % 64, so one
// wavefront reading element e of 64 consecutive links touches 64 consecutive
// reals.  For N >= 4 the tile is 1 (each matrix contiguous).
constexpr int kSiteTile = 64;

__host__ __device__ inline size_t elem_off(long l, int e, int msz, int tile)
{
    return ((size_t)(l / tile) * (size_t)msz + (size_t)e) * (size_t)tile + (size_t)(l % tile);
}

// Compact Hermitian storage of a momentum: N diagonal reals, then the upper
// triangle (i < j, row-major) as (re, im) pairs.  N*N reals per link.
__host__ __device__ inline int herm_upper_index(int n, int i, int j)
{
    return n + 2 * (i * n - i * (i + 1) / 2 + (j - i - 1));
}

// ---------------------------------------------------------------- Philox
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        U4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
    }
    return c;
}

// Philox stream purposes (top byte of the last counter word).
enum : uint32_t { kPurposeMomentum = 1u, kPurposeHaar = 2u, kPurposeAccept = 3u };

__host__ __device__ inline U4 philox_block(uint64_t seed, uint64_t counter, uint32_t purpose,
                                           uint32_t link, uint32_t block)
{
    U4 c;
    c.x = block;
    c.y = link;
    c.z = (uint32_t)counter;
    c.w = (uint32_t)(counter >> 32) ^ (purpose << 24);
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__host__ __device__ inline double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

// 4 normals from one Philox block: two Box-Muller pairs.
__device__ inline void normals4(U4 r, double z[4])
{
    double rad = sqrt(-2.0 * log(u01(r.x)));
    double ang = 6.283185307179586 * u01(r.y);
    z[0] = rad * cos(ang);
    z[1] = rad * sin(ang);
    rad = sqrt(-2.0 * log(u01(r.z)));
    ang = 6.283185307179586 * u01(r.w);
    z[2] = rad * cos(ang);
    z[3] = rad * sin(ang);
}

// ------------------------------------------------------------ XCD mapping
// The hardware hands consecutive workgroups of a launch to the 8 XCDs in
// turn (workgroup i to XCD i % 8), each XCD with its own L2.  This maps the
// launch index to a logical block so that every XCD walks one contiguous
// range of logical blocks: spatial neighbours then share an L2.  Bijective
// for any grid size.
constexpr unsigned kNumXcd = 8;

__device__ inline unsigned xcd_block(unsigned i, unsigned nblocks)
{
    const unsigned q = nblocks / kNumXcd, r = nblocks % kNumXcd;
    const unsigned x = i % kNumXcd, k = i / kNumXcd;
    return x * q + (x < r ? x : r) + k;
}

// Site block sb of a launch over both parities, blocks of `bs` sites:
// blocks of even and odd sites alternate, so that a contiguous range of site
// blocks covers one region of the lattice in both parities (the neighbours
// of an even site are odd).  Returns the first device site of the block and
// the number of valid sites in it.
__device__ inline long parity_block(long sb, long V, long V0, int bs, int &count)
{
    const long V1 = V - V0;
    const long n0 = (V0 + bs - 1) / bs, n1 = (V1 + bs - 1) / bs;
    const long m = n0 < n1 ? n0 : n1;
    long p, r;
    if (sb < 2 * m) {
        p = sb & 1;
        r = sb >> 1;
    } else {
        p = n0 > n1 ? 0 : 1;
        r = m + (sb - 2 * m);
    }
    const long base = p ? V0 : 0, len = p ? V1 : V0;
    const long first = r * bs;
    const long left = len - first;
    count = left < bs ? (int)left : bs;
    return base + first;
}

// ---------------------------------------------------------- reductions
// Neumaier-compensated accumulation; a (s, c) pair represents s + c.
__device__ inline void kahan_add(double &s, double &c, double x)
{
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

// Combine two (s, c) pairs with an error-free TwoSum of the leading parts.
__device__ inline void pair_add(double &s, double &c, double s2, double c2)
{
    const double t = s + s2;
    const double bp = t - s;
    const double e = (s - (t - bp)) + (s2 - bp);
    s = t;
    c = c + c2 + e;
}

// Block-wide reduction of K compensated pairs held per thread; thread 0
// writes out[2*k] = s, out[2*k+1] = c.  blockDim.x must be a power of two
// <= 1024.  Fixed tree order: deterministic.
template <int K>
__device__ inline void block_reduce_store(double (&s)[K], double (&c)[K], double *out)
{
    __shared__ double sh_s[K][1024 / 64];
    __shared__ double sh_c[K][1024 / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // wavefront stage: fixed butterfly order
#pragma unroll
    for (int k = 0; k < K; k++) {
        for (int off = 32; off > 0; off >>= 1) {
            double s2 = __shfl_down(s[k], off, 64);
            double c2 = __shfl_down(c[k], off, 64);
            if (lane < off) pair_add(s[k], c[k], s2, c2);
        }
        if (lane == 0) {
            sh_s[k][wid] = s[k];
            sh_c[k][wid] = c[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int k = 0; k < K; k++) {
            double a = sh_s[k][0], b = sh_c[k][0];
            for (int w = 1; w < nw; w++) pair_add(a, b, sh_s[k][w], sh_c[k][w]);
            out[2 * k] = a;
            out[2 * k + 1] = b;
        }
    }
}

}  // namespace askja
