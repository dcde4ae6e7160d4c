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

// ---------------------------------------------- reproducible Box-Muller
// The logarithm and the sine/cosine of the Box-Muller transform are written
// here as fixed sequences of IEEE-754 double operations (exact scalings and
// subtractions, one correctly rounded division, correctly rounded sqrt, FMA
// Horner steps, no contraction), so that any IEEE machine that performs the
// same steps (the host C library's fma and sqrt are correctly rounded too)
// gets the same bits: the Philox momenta of the device equal a host
// computation bit for bit.  Accuracy: within 2 ulp of the true values.

// log x for normal x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s) = 2 s sum_k s^2k / (2k + 1), s = (m - 1) / (m + 1),
// |s| <= 0.1716 so 13 terms reach 2^-70; e log 2 with log 2 split in a
// 32-bit head (e ln2_hi exact) and a tail.
__device__ inline double repro_log(double x)
{
#pragma clang fp contract(off)
    const long long b = __double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((b & 0x000fffffffffffffLL) | 0x3ff0000000000000LL);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double f = m - 1.0;  // exact (Sterbenz)
    const double s = f / (m + 1.0);
    const double z = s * s;
    double p = 0.04;
    p = fma(p, z, 0.043478260869565216);
    p = fma(p, z, 0.047619047619047616);
    p = fma(p, z, 0.05263157894736842);
    p = fma(p, z, 0.058823529411764705);
    p = fma(p, z, 0.06666666666666667);
    p = fma(p, z, 0.07692307692307693);
    p = fma(p, z, 0.09090909090909091);
    p = fma(p, z, 0.1111111111111111);
    p = fma(p, z, 0.14285714285714285);
    p = fma(p, z, 0.2);
    p = fma(p, z, 0.3333333333333333);
    p = fma(p, z, 1.0);
    const double lm = (2.0 * s) * p;
    const double de = (double)e;
    return de * 0.6931471803691238 + fma(de, 1.9082149292705877e-10, lm);
}

// (sin, cos) of 2 pi u for u in [0, 1): t = 4u, quadrant q = floor(t + 1/2),
// theta = (t - q) pi / 2 in [-pi/4, pi/4] (t - q exact), Taylor polynomials
// to theta^19 / theta^20, then the exact quadrant rotation.
__device__ inline void repro_sincos2pi(double u, double *sn, double *cs)
{
#pragma clang fp contract(off)
    const double t = 4.0 * u;
    const double q = floor(t + 0.5);
    const double th = (t - q) * 1.5707963267948966;
    const double z = th * th;
    double ps = -8.22063524662433e-18;
    ps = fma(ps, z, 2.8114572543455206e-15);
    ps = fma(ps, z, -7.647163731819816e-13);
    ps = fma(ps, z, 1.6059043836821613e-10);
    ps = fma(ps, z, -2.505210838544172e-08);
    ps = fma(ps, z, 2.7557319223985893e-06);
    ps = fma(ps, z, -0.0001984126984126984);
    ps = fma(ps, z, 0.008333333333333333);
    ps = fma(ps, z, -0.16666666666666666);
    ps = fma(ps, z, 1.0);
    double pc = 4.110317623312165e-19;
    pc = fma(pc, z, -1.5619206968586225e-16);
    pc = fma(pc, z, 4.779477332387385e-14);
    pc = fma(pc, z, -1.1470745597729725e-11);
    pc = fma(pc, z, 2.08767569878681e-09);
    pc = fma(pc, z, -2.755731922398589e-07);
    pc = fma(pc, z, 2.48015873015873e-05);
    pc = fma(pc, z, -0.001388888888888889);
    pc = fma(pc, z, 0.041666666666666664);
    pc = fma(pc, z, -0.5);
    pc = fma(pc, z, 1.0);
    const double s0 = th * ps, c0 = pc;
    switch ((int)q & 3) {
    case 0: *sn = s0; *cs = c0; break;
    case 1: *sn = c0; *cs = -s0; break;
    case 2: *sn = -s0; *cs = -c0; break;
    default: *sn = -c0; *cs = s0; break;
    }
}

// Box-Muller pair of two uniforms words: r = sqrt(-2 log u0),
// (r cos 2 pi u1, r sin 2 pi u1).
__device__ inline void box_muller(uint32_t w0, uint32_t w1, double *zc, double *zs)
{
#pragma clang fp contract(off)
    const double rad = sqrt(-2.0 * repro_log(u01(w0)));
    double sn, cs;
    repro_sincos2pi(u01(w1), &sn, &cs);
    *zc = rad * cs;
    *zs = rad * sn;
}

// 4 normals from one Philox block: two Box-Muller pairs.
__device__ inline void normals4(U4 r, double z[4])
{
    box_muller(r.x, r.y, &z[0], &z[1]);
    box_muller(r.z, r.w, &z[2], &z[3]);
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
