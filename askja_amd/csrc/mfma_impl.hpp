// mfma_impl.hpp -- kick (+ drift) for 9 <= N <= 64 on the MFMA units.
// Every complex link product C = A B is computed as two real products of
// the stacked form
//   Re C = [Ar | Ai] [Br ; -Bi],   Im C = [Ar | Ai] [Bi ; Br]   (K = 2 NMAX)
// in one of three arithmetic policies:
//   F64    : FP64 links, v_mfma_f64_16x16x4_f64;
//   F32    : FP32 links, v_mfma_f32_16x16x4_f32;
//   BF16X3 : FP32 links split as x = hi + lo (hi = bf16(x), lo = bf16(x - hi)),
//            hi*hi + hi*lo + lo*hi accumulated in FP32 by
//            v_mfma_f32_16x16x32_bf16.
// One workgroup per link; the (zero padded) NMAX x NMAX matrices are cut in
// 16-row blocks, each handled by WPR wavefronts that split its 16-column
// tiles.  Lane l holds, for every tile of its wavefront, four outputs in
// column l%16 of the tile, at rows 4*(l/16)+q (f32 accumulators) or
// 4*q+l/16 (f64 accumulators), q < 4: the MFMA accumulator layouts, both
// checked on gfx950 with one-hot operands.  The A operand of lane l is
// A[l%16][k0 + l/16] and the B operand B[k0 + l/16][l%16] for both 16x16x4
// forms.
// Operands are staged in two LDS slots (planar re/im, row stride Cfg::LD;
// BF16x3 slots column-swizzled, see bsw).
#pragma once
#include "kernels.hpp"
#include "common.hpp"

namespace askja {
namespace mfma {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

enum Policy { kF64 = 0, kF32 = 1, kBF16X3 = 2 };

template <int P>
struct Pol;
template <>
struct Pol<kF64> {
    typedef double R;
    typedef f64x4 V4;
};
template <>
struct Pol<kF32> {
    typedef float R;
    typedef f32x4 V4;
};
template <>
struct Pol<kBF16X3> {
    typedef float R;
    typedef f32x4 V4;
};

template <int P, int NMAX>
struct Cfg {
    static constexpr int WPR = (P == kF64 && NMAX == 64) ? 2 : 1;  // wavefronts per 16-row block
    static constexpr int NT = 64 * (NMAX / 16) * WPR;
    static constexpr int NCT = NMAX / 16 / WPR;  // column tiles per wavefront
    // operand prefetch through registers, except at NMAX = 64 in FP32 and
    // BF16x3, where the extra registers would halve the occupancy
    static constexpr bool PF = !(P != kF64 && NMAX == 64);
    // LDS row stride in elements.  FP64 / FP32: NMAX + 2 spreads the 16 rows
    // a wavefront reads for one operand over the banks (an unpadded stride
    // maps them all to one bank group; + 2 measured a little faster than
    // + 4).  BF16x3: NMAX + 8 bf16 (16 bytes) keeps the 16-byte reads of
    // eight consecutive rows on distinct banks.
    static constexpr int LD = NMAX + (P == kBF16X3 ? 8 : 2);
    static constexpr int SLOT = NMAX * LD;
};

template <int P, int NMAX>
struct Acc {
    typename Pol<P>::V4 r[Cfg<P, NMAX>::NCT], i[Cfg<P, NMAX>::NCT];
};

template <int P, int NMAX>
__device__ inline void acc_zero(Acc<P, NMAX> &a)
{
#pragma unroll
    for (int t = 0; t < Cfg<P, NMAX>::NCT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) a.r[t][q] = a.i[t][q] = 0;
}

// wavefront geometry: 16-row block and first column tile of this wavefront
template <int P, int NMAX>
__device__ inline int wave_rowblock()
{
    return (threadIdx.x >> 6) / Cfg<P, NMAX>::WPR;
}
template <int P, int NMAX>
__device__ inline int wave_tile0()
{
    return ((threadIdx.x >> 6) % Cfg<P, NMAX>::WPR) * Cfg<P, NMAX>::NCT;
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// bf16x8 with the sign of every element flipped where m = 0x80008000
__device__ inline bf16x8 xsign(bf16x8 v, unsigned m)
{
    u32x4 x = __builtin_bit_cast(u32x4, v);
    x ^= m;
    return __builtin_bit_cast(bf16x8, x);
}

__device__ inline bf16x8 ld8(const __bf16 *p)
{
    return *(const bf16x8 *)p;
}

// Column of a BF16x3 slot element as stored: in rows 8-15 of every 16-row
// block the two 16-byte chunks of each 16 k are swapped (k ^ 8).  A
// wavefront writing one slot column over rows c = 0..15 (a transposed
// store: rows c and c + 8 are 12 c dwords apart at row stride 24, the same
// bank modulo 32) then hits distinct banks, while a product's read of eight
// consecutive k stays one aligned 16-byte chunk.
__device__ inline int bsw(int row, int k)
{
    return k ^ (row & 8);
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Two elements x = hi + lo at once: hi = bf16(x) (round to nearest even, one
// packed conversion), lo = bf16(x - hi).  Each difference x_j - hi_j is one
// bf16 dot product straight from the packed hi word, x_j + (-1) hi_j + 0 hi_k
// (v_dot2c_f32_bf16): the products are exact and x_j - hi_j is representable
// in FP32 (hi_j is x_j rounded to 8 bits, within a factor 2 of it), so the
// dot product returns it exactly, as the subtraction from the widened hi_j
// did.  A pair costs four VALU instructions (two conversions, two dot
// products) instead of six (a shift and a mask to widen hi, two
// subtractions).  The (-1, 0) and (0, -1) bf16 pairs are moved into SGPRs
// by hand: written as constants, the compiler encodes 0x0000bf80 as the
// inline constant -1.0, which the instruction reads as the FP32 pattern
// 0xbf800000, i.e. the pair (0, -1) (checked on gfx950: the first element's
// lo then came out as x0 + hi1).
__device__ inline void split2(float x0, float x1, bf16x2 &hi, bf16x2 &lo)
{
    const f32x2 x = {x0, x1};
    hi = __builtin_convertvector(x, bf16x2);
    unsigned b0, b1;
    asm("s_mov_b32 %0, 0xbf80" : "=s"(b0));
    asm("s_mov_b32 %0, 0xbf800000" : "=s"(b1));
    const bf16x2 m0 = __builtin_bit_cast(bf16x2, b0), m1 = __builtin_bit_cast(bf16x2, b1);
    const f32x2 d = {__builtin_amdgcn_fdot2_f32_bf16(hi, m0, x0, false),
                     __builtin_amdgcn_fdot2_f32_bf16(hi, m1, x1, false)};
    lo = __builtin_convertvector(d, bf16x2);
}

// Operand slots.  FP64 / FP32: planar re / im matrices, row-major, row
// stride LD.  BF16x3: every element is split once, when it is written, as
// x = hi + lo (hi = bf16(x), lo = bf16(x - hi)); each of the re and im
// arrays then holds a hi plane followed by a lo plane of bf16, row stride LD,
// stored so that the MFMA's k index runs along the row: as A for A and for a
// daggered B, transposed (TR) for a daggered A and for a B.  The products
// read eight consecutive k of one plane with one 16-byte load.
template <int P, int NMAX, bool TR, typename R>
__device__ inline void put(R *Sr, R *Si, int r, int c, R re, R im)
{
    constexpr int LD = Cfg<P, NMAX>::LD;
    if constexpr (P == kBF16X3) {
        const int o = TR ? c * LD + bsw(c, r) : r * LD + bsw(r, c);
        __bf16 *rh = (__bf16 *)Sr, *ih = (__bf16 *)Si;
        bf16x2 h, l;
        split2(re, im, h, l);
        rh[o] = h[0];
        rh[NMAX * LD + o] = l[0];
        ih[o] = h[1];
        ih[NMAX * LD + o] = l[1];
    } else {
        Sr[r * LD + c] = re;
        Si[r * LD + c] = im;
    }
}

// The stacked product on the 16x16x4 forms, both halves in one pass over k:
// lane l holds A[row][k] and B[k][col], k = 4 ks + l/16, and each pair
// (Br, Bi) it reads from LDS feeds four products,
//   Re += Ar Br + (-Ai) Bi,   Im += Ar Bi + Ai Br
// (Ai, Bi with the sign of the conjugate for a daggered operand).
template <int P, int NMAX, bool DA, bool DB, typename R>
__device__ inline void cmul_vec(const R *Ar, const R *Ai, const R *Br, const R *Bi, Acc<P, NMAX> &acc)
{
    constexpr int NCT = Cfg<P, NMAX>::NCT, LD = Cfg<P, NMAX>::LD;
    const int lane = threadIdx.x & 63;
    const int row = 16 * wave_rowblock<P, NMAX>() + (lane & 15);
    const int t0 = wave_tile0<P, NMAX>();
#pragma unroll 4
    for (int ks = 0; ks < NMAX / 4; ks++) {
        const int k = 4 * ks + (lane >> 4);
        const R ar = DA ? Ar[k * LD + row] : Ar[row * LD + k];
        const R ai = DA ? -Ai[k * LD + row] : Ai[row * LD + k];
        const R nai = -ai;
#pragma unroll
        for (int t = 0; t < NCT; t++) {
            const int col = 16 * (t0 + t) + (lane & 15);
            const R br = DB ? Br[col * LD + k] : Br[k * LD + col];
            const R bi = DB ? -Bi[col * LD + k] : Bi[k * LD + col];
            if constexpr (P == kF64) {
                acc.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, acc.r[t], 0, 0, 0);
                acc.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acc.i[t], 0, 0, 0);
                acc.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bi, acc.r[t], 0, 0, 0);
                acc.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, acc.i[t], 0, 0, 0);
            } else {
                acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, acc.r[t], 0, 0, 0);
                acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, acc.i[t], 0, 0, 0);
                acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(nai, bi, acc.r[t], 0, 0, 0);
                acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, acc.i[t], 0, 0, 0);
            }
        }
    }
}

// BF16x3 steps ks in [KS0, KS1) of the stacked product.  Lane l holds
// A[row][k0..k0+7] and B[k0..k0+7][col], k0 = 32 ks + 8 (l/16) in the
// stacked index; the slots hold A and B in the orientation that makes k
// contiguous.  Every sign (conjugation, the -Bi block of [Br ; -Bi]) is
// folded into the A operand, once per step: the real output takes
// (+Ar | -+Ai) against [Br ; Bi] rows, the imaginary output (+-Ar | +-Ai)
// against [Bi ; Br].  H = 0 / 1: all lanes in the real / imaginary half
// (NMAX >= 32); H = -1: the half is decided per lane and step.
template <int NMAX, bool DA, bool DB, int H>
__device__ inline void bf_steps(const __bf16 *arh, const __bf16 *aih, const __bf16 *brh, const __bf16 *bih,
                                Acc<kBF16X3, NMAX> &acc)
{
    constexpr int P = kBF16X3, NCT = Cfg<P, NMAX>::NCT, LD = Cfg<P, NMAX>::LD, PL = NMAX * LD;
    constexpr int KS0 = H == 1 ? NMAX / 32 : 0, KS1 = H == 0 ? NMAX / 32 : (2 * NMAX) / 32;
    constexpr unsigned SG = 0x80008000u;
    const int lane = threadIdx.x & 63;
    const int row = 16 * wave_rowblock<P, NMAX>() + (lane & 15);
    const int t0 = wave_tile0<P, NMAX>();
    const int kq = 8 * (lane >> 4);
#pragma unroll 1  // unrolled, the 16-byte operand loads of every step stay live
    for (int ks = KS0; ks < KS1; ks++) {
        const int k0 = 32 * ks + kq;
        const bool imag = H < 0 ? k0 >= NMAX : H == 1;
        const int kb = imag ? k0 - NMAX : k0;  // column of A / row of B
        // real output: Ar Br - Ai_eff Bi_eff; imaginary: Ar Bi_eff + Ai_eff Br,
        // with Ai_eff = -Ai for A^+ and Bi_eff = -Bi for B^+
        const unsigned mr = imag ? ((DA == DB) ? SG : 0u) : 0u;
        const unsigned mi = imag ? (DA ? SG : 0u) : (DB ? SG : 0u);
        const __bf16 *ap = (imag ? aih : arh) + row * LD + bsw(row, kb);
        const bf16x8 ah = ld8(ap), al = ld8(ap + PL);
        const bf16x8 rh = xsign(ah, mr), rl = xsign(al, mr), ih = xsign(ah, mi), il = xsign(al, mi);
        const __bf16 *bp = imag ? bih : brh, *cp = imag ? brh : bih;
#pragma unroll
        for (int t = 0; t < NCT; t++) {
            const int br = 16 * (t0 + t) + (lane & 15);
            const int o = br * LD + bsw(br, kb);
            const bf16x8 bh = ld8(bp + o), bl = ld8(bp + PL + o);
            const bf16x8 ch = ld8(cp + o), cl = ld8(cp + PL + o);
            acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh, bh, acc.r[t], 0, 0, 0);
            acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh, bl, acc.r[t], 0, 0, 0);
            acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rl, bh, acc.r[t], 0, 0, 0);
            acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ih, ch, acc.i[t], 0, 0, 0);
            acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ih, cl, acc.i[t], 0, 0, 0);
            acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(il, ch, acc.i[t], 0, 0, 0);
        }
    }
}

// acc (+)= op(A) op(B) on the wavefront's rows and column tiles; A, B in LDS.
template <int P, int NMAX, bool DA, bool DB, typename R>
__device__ inline void cmul(const R *Ar, const R *Ai, const R *Br, const R *Bi, Acc<P, NMAX> &acc)
{
    if constexpr (P == kBF16X3) {
        const __bf16 *arh = (const __bf16 *)Ar, *aih = (const __bf16 *)Ai;
        const __bf16 *brh = (const __bf16 *)Br, *bih = (const __bf16 *)Bi;
        if constexpr (NMAX >= 64) {  // NMAX = 32 measured faster with the per-lane form
            bf_steps<NMAX, DA, DB, 0>(arh, aih, brh, bih, acc);
            bf_steps<NMAX, DA, DB, 1>(arh, aih, brh, bih, acc);
        } else {
            bf_steps<NMAX, DA, DB, -1>(arh, aih, brh, bih, acc);
        }
    } else {
        cmul_vec<P, NMAX, DA, DB>(Ar, Ai, Br, Bi, acc);
    }
}

// row / column of accumulator entry (t, q) of this lane
template <int P, int NMAX>
__device__ inline int acc_row(int q)
{
    const int lr = (threadIdx.x & 63) >> 4;
    return 16 * wave_rowblock<P, NMAX>() + (P == kF64 ? 4 * q + lr : 4 * lr + q);
}
template <int P, int NMAX>
__device__ inline int acc_col(int t)
{
    return 16 * (wave_tile0<P, NMAX>() + t) + (threadIdx.x & 15);
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// four consecutive elements x -> hi / lo bf16 (see split2)
__device__ inline void split4(const f32x4 &x, bf16x4 &hi, bf16x4 &lo)
{
    bf16x2 h0, l0, h1, l1;
    split2(x[0], x[1], h0, l0);
    split2(x[2], x[3], h1, l1);
    hi = __builtin_shufflevector(h0, h1, 0, 1, 2, 3);
    lo = __builtin_shufflevector(l0, l1, 0, 1, 2, 3);
}

// accumulators -> operand slot (orientation TR, see put).  A BF16x3 slot
// stored transposed takes the lane's four rows q = 0..3 (consecutive in the
// f32 accumulator layout) with one 8-byte write per plane.
template <bool TR, int P, int NMAX, typename R>
__device__ inline void acc_store(const Acc<P, NMAX> &a, R *Sr, R *Si)
{
    if constexpr (P == kBF16X3 && TR) {
        constexpr int LD = Cfg<P, NMAX>::LD, PL = NMAX * LD;
        __bf16 *rp = (__bf16 *)Sr, *ip = (__bf16 *)Si;
#pragma unroll
        for (int t = 0; t < Cfg<P, NMAX>::NCT; t++) {
            const int o = acc_col<P, NMAX>(t) * LD + bsw(acc_col<P, NMAX>(t), acc_row<P, NMAX>(0));
            bf16x4 h, l;
            split4(a.r[t], h, l);
            *(bf16x4 *)(rp + o) = h;
            *(bf16x4 *)(rp + PL + o) = l;
            split4(a.i[t], h, l);
            *(bf16x4 *)(ip + o) = h;
            *(bf16x4 *)(ip + PL + o) = l;
        }
    } else {
#pragma unroll
        for (int t = 0; t < Cfg<P, NMAX>::NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                put<P, NMAX, TR, R>(Sr, Si, acc_row<P, NMAX>(q), acc_col<P, NMAX>(t), a.r[t][q], a.i[t][q]);
    }
}

// accumulators -> planar row-major matrix at full precision (read element
// by element, not as a product operand)
template <int P, int NMAX, typename R>
__device__ inline void acc_store_full(const Acc<P, NMAX> &a, R *Sr, R *Si)
{
#pragma unroll
    for (int t = 0; t < Cfg<P, NMAX>::NCT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Sr[acc_row<P, NMAX>(q) * Cfg<P, NMAX>::LD + acc_col<P, NMAX>(t)] = a.r[t][q];
            Si[acc_row<P, NMAX>(q) * Cfg<P, NMAX>::LD + acc_col<P, NMAX>(t)] = a.i[t][q];
        }
}

// Operand items: each thread moves E elements of a slot, in items of W
// elements.  FP64 / FP32: W = 1, item p is element (p / NMAX, p % NMAX).
// BF16x3: W = 4, item p is the 2 x 2 block at rows 2 br, 2 br + 1 and columns
// 2 bc, 2 bc + 1 (br = p / (NMAX / 2), bc = p % (NMAX / 2)), element w at
// (2 br + w / 2, 2 bc + w % 2): a block is two 16-byte runs of the link in
// memory, and either orientation of the slot takes it as two 4-byte writes
// per plane (rows 2 br, 2 br + 1 of the slot, or columns when TR).
template <int P, int NMAX>
struct Items {
    static constexpr int NT = Cfg<P, NMAX>::NT, E = NMAX * NMAX / NT;
    static constexpr int W = P == kBF16X3 ? 4 : 1, EI = E / W;
    static_assert(E * NT == NMAX * NMAX && EI * W == E, "slot not a multiple of the workgroup");
    // link element (r, c) of item p, element w
    static __device__ inline void rc(int p, int w, int &r, int &c)
    {
        if constexpr (W == 1) {
            r = p / NMAX;
            c = p % NMAX;
        } else {
            r = 2 * (p / (NMAX / 2)) + (w >> 1);
            c = 2 * (p % (NMAX / 2)) + (w & 1);
        }
    }
    // write the item's W complex values (link elements of rc) to the slot
    template <bool TR, typename R>
    static __device__ inline void write(R *Sr, R *Si, int p, const R (&re)[W], const R (&im)[W])
    {
        if constexpr (W == 1) {
            int r, c;
            rc(p, 0, r, c);
            put<P, NMAX, TR, R>(Sr, Si, r, c, re[0], im[0]);
        } else {
            constexpr int LD = Cfg<P, NMAX>::LD, PL = NMAX * LD;
            int r, c;
            rc(p, 0, r, c);
            // pairs along the stored row: (w0, w1), (w2, w3) as rows r, r + 1;
            // (w0, w2), (w1, w3) as columns c, c + 1 when TR
            constexpr int a0 = 0, a1 = TR ? 2 : 1, b0 = TR ? 1 : 2, b1 = 3;
            // rows r, r + 1 (or c, c + 1) lie in one 8-row half: one swizzle
            const int o0 = TR ? c * LD + bsw(c, r) : r * LD + bsw(r, c), o1 = o0 + LD;
            bf16x2 h, l;
            __bf16 *rp = (__bf16 *)Sr, *ip = (__bf16 *)Si;
            split2(re[a0], re[a1], h, l);
            *(bf16x2 *)(rp + o0) = h;
            *(bf16x2 *)(rp + PL + o0) = l;
            split2(re[b0], re[b1], h, l);
            *(bf16x2 *)(rp + o1) = h;
            *(bf16x2 *)(rp + PL + o1) = l;
            split2(im[a0], im[a1], h, l);
            *(bf16x2 *)(ip + o0) = h;
            *(bf16x2 *)(ip + PL + o0) = l;
            split2(im[b0], im[b1], h, l);
            *(bf16x2 *)(ip + o1) = h;
            *(bf16x2 *)(ip + PL + o1) = l;
        }
    }
    // global link (matrix contiguous, n x n complex) -> the item's values.
    // EN (n == NMAX): a BF16x3 block is read as two 16-byte loads with no
    // bounds test; otherwise element by element, zero in the padding.
    template <bool EN, typename S, typename V2>
    static __device__ inline void fetch(const S *__restrict__ b, int p, int n, V2 (&v)[W])
    {
        if constexpr (W == 4 && EN) {
            typedef S S4 __attribute__((ext_vector_type(4)));
            int r, c;
            rc(p, 0, r, c);
            const S4 x = *(const S4 *)(b + 2 * (r * NMAX + c));
            const S4 y = *(const S4 *)(b + 2 * ((r + 1) * NMAX + c));
            v[0] = V2{x[0], x[1]};
            v[1] = V2{x[2], x[3]};
            v[2] = V2{y[0], y[1]};
            v[3] = V2{y[2], y[3]};
        } else {
            const V2 *b2 = (const V2 *)b;
#pragma unroll
            for (int w = 0; w < W; w++) {
                int r, c;
                rc(p, w, r, c);
                v[w] = (EN || (r < n && c < n)) ? b2[r * n + c] : V2{S(0), S(0)};
            }
        }
    }
};

// global link -> zero-padded slot
template <int P, int NMAX, bool TR, bool EN = false, typename R, typename S>
__device__ inline void ld_link(R *Sr, R *Si, const S *__restrict__ f, long l, int n)
{
    typedef Items<P, NMAX> I;
    typedef S S2 __attribute__((ext_vector_type(2)));
    const S *b = f + (size_t)l * 2 * n * n;
    for (int p = threadIdx.x; p < I::EI * I::NT; p += I::NT) {
        S2 v[I::W];
        I::template fetch<EN>(b, p, n, v);
        R re[I::W], im[I::W];
#pragma unroll
        for (int w = 0; w < I::W; w++) {
            re[w] = (R)v[w][0];
            im[w] = (R)v[w][1];
        }
        I::template write<TR, R>(Sr, Si, p, re, im);
    }
}

// Register staging of one link operand: fetch() issues its global loads,
// stash() writes it to an LDS slot.  With PF the fetch of the next operands
// is issued before the current product and the stash follows it, so the
// loads run under the MFMA work; without PF the operand is loaded straight
// into its slot at the stash point (fewer live registers).  TR is the
// orientation of the slot the operand goes to (see put); EN: n == NMAX.
template <int P, int NMAX, typename S, bool EN>
struct Stage {
    typedef Items<P, NMAX> I;
    typedef S S2 __attribute__((ext_vector_type(2)));
    S2 v[I::EI][I::W];
    __device__ inline void fetch(const S *__restrict__ f, long l, int n)
    {
        const S *b = f + (size_t)l * 2 * n * n;
#pragma unroll
        for (int j = 0; j < I::EI; j++) I::template fetch<EN>(b, (int)threadIdx.x + j * I::NT, n, v[j]);
    }
    template <bool TR, typename R>
    __device__ inline void stash(R *Sr, R *Si) const
    {
#pragma unroll
        for (int j = 0; j < I::EI; j++) {
            R re[I::W], im[I::W];
#pragma unroll
            for (int w = 0; w < I::W; w++) {
                re[w] = (R)v[j][w][0];
                im[w] = (R)v[j][w][1];
            }
            I::template write<TR, R>(Sr, Si, threadIdx.x + j * I::NT, re, im);
        }
    }
    template <bool PF, bool TR>
    __device__ inline void early(const S *__restrict__ f, long l, int n)
    {
        if constexpr (PF) fetch(f, l, n);
    }
    template <bool PF, bool TR, typename R>
    __device__ inline void late(const S *__restrict__ f, long l, int n, R *Sr, R *Si)
    {
        if constexpr (PF)
            stash<TR>(Sr, Si);
        else
            ld_link<P, NMAX, TR, EN>(Sr, Si, f, l, n);
    }
};

template <typename R>
__device__ inline R mom_el(const R *__restrict__ P, long l, int n, int i, int j, R &im)
{
    const R *b = P + (size_t)l * n * n;
    if (i == j) {
        im = R(0);
        return b[i];
    }
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int h = hup(n, lo, hi);
    im = i < j ? b[h + 1] : -b[h + 1];
    return b[h];
}

// Row and column (i < j) of entry v of the upper triangle in row order:
// row i starts at s_i = i n - i (i + 1) / 2; i from the quadratic (the
// hardware square root, within an ulp), then corrected by one where the root
// rounds across a row start.
__device__ inline void upper_ij(int n, int v, int &i, int &j)
{
    const float m = (float)(2 * n - 1);
    int r = (int)(0.5f * (m - __builtin_amdgcn_sqrtf(m * m - 8.0f * (float)v)));
    r = r < 0 ? 0 : (r > n - 2 ? n - 2 : r);
    if (r * n - r * (r + 1) / 2 > v) r--;
    if ((r + 1) * n - (r + 1) * (r + 2) / 2 <= v) r++;
    i = r;
    j = v - (r * n - r * (r + 1) / 2) + r + 1;
}

// The kick on the compact momentum (n diagonal reals, then the upper
// triangle's (re, im) pairs in row order, see hup): thread-strided units, a
// unit being one diagonal element or one upper pair, so the momentum is read
// and written in contiguous runs and each unit by one thread.  W = U_mu(x) S
// at full precision in (Wr, Wi).  Drift: the new momentum also goes to
// (Fr, Fi) as a full Hermitian matrix.  sum: {ReTr W, Tr Pin^2, Tr Pout^2}.
// TAH: sunmat.h:355-371; kick: sun.cpp:716.
template <int P, int NMAX, bool Drift, typename R>
__device__ inline void kick_compact(const KickArgs &a, long l, int n, const R *Wr, const R *Wi, R *Fr, R *Fi,
                                    double (&sum)[3])
{
    constexpr int LD = Cfg<P, NMAX>::LD, NT = Cfg<P, NMAX>::NT;
    const R eps = (R)a.eps_kick, inv_n = (R)a.inv_n;
    R tri = R(0);
    for (int i = 0; i < n; i++) tri += Wi[i * LD + i] + Wi[i * LD + i];
    const R corr = tri * inv_n;
    R *b = (R *)a.P + (size_t)l * n * n;
    R *bp = (!Drift && a.Pprev) ? (R *)a.Pprev + (size_t)l * n * n : nullptr;
    const int units = n + n * (n - 1) / 2;
    const bool sums = a.part != nullptr;  // the energies are read from the partials
    for (int u = threadIdx.x; u < units; u += NT) {
        if (u < n) {
            const R ir = b[u];
            const R mi = (Wi[u * LD + u] + Wi[u * LD + u]) - corr;
            const R pr = ir - (R(0.5) * mi) * eps;
            b[u] = pr;
            if (bp) bp[u] = ir;
            if (sums) {
                sum[0] += (double)Wr[u * LD + u];
                sum[1] += (double)ir * ir;
                sum[2] += (double)pr * pr;
            }
            if constexpr (Drift) {
                Fr[u * LD + u] = pr;
                Fi[u * LD + u] = R(0);
            }
        } else {
            const int v = u - n, h = n + 2 * v;
            int i, j;
            upper_ij(n, v, i, j);
            const R ur = b[h], ui = b[h + 1];
            const R dr = Wr[i * LD + j] - Wr[j * LD + i];
            const R di = Wi[i * LD + j] + Wi[j * LD + i];
            const R vr = ur - (R(0.5) * di) * eps, vi = ui + (R(0.5) * dr) * eps;
            b[h] = vr;
            b[h + 1] = vi;
            if (bp) {
                bp[h] = ur;
                bp[h + 1] = ui;
            }
            if (sums) {
                sum[1] += 2.0 * ((double)ur * ur + (double)ui * ui);
                sum[2] += 2.0 * ((double)vr * vr + (double)vi * vi);
            }
            if constexpr (Drift) {
                Fr[i * LD + j] = vr;
                Fi[i * LD + j] = vi;
                Fr[j * LD + i] = vr;
                Fi[j * LD + i] = -vi;
            }
        }
    }
}

// LS: link storage type (FP32 links with FP64 arithmetic for the exact
// action of the reduced-precision modes).  ActionOnly: only sum ReTr W
// (= 4 sum of the plaquettes touching the link) into the partials.
// The drift's truncated exponential X = sum_{j <= L} R^j / j! of R = i eps P
// (the Taylor series of sunmat.h:200-248) in accumulator layout.  On entry
// slot A holds R (stored as a left operand) and (rr, ri) are R's lane
// entries; slots A and B are overwritten.  The caller has a barrier between
// its last read of slot B and this call.  L <= 3: Horner, X = 1 + (R / k) X,
// L - 1 products.  L >= 4: the even/odd split X = sum_i R^{2i} (a_{2i} + a_{2i+1} R),
// a_j = 1 / j!, by Horner in R^2: 1 + ceil((L + 1) / 2) - 1 products (5
// instead of 7 at L = 8, 4 instead of 5 at L = 6).  Same polynomial, other
// rounding: the non-exact modes only.
template <int P, int NMAX, typename R>
__device__ inline void taylor_poly(const R (&rr)[Cfg<P, NMAX>::NCT][4], const R (&ri)[Cfg<P, NMAX>::NCT][4], int level,
                                   const double *cinv, int n, R *Are, R *Aim, R *Bre, R *Bim, Acc<P, NMAX> &X)
{
    constexpr int NCT = Cfg<P, NMAX>::NCT;
    if (level < 4) {
        const R cl = level <= 1 ? R(1) : (R)cinv[level];
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                const R d = (i == j && i < n) ? R(1) : R(0);
                X.r[t][q] = level == 0 ? d : d + cl * rr[t][q];
                X.i[t][q] = level == 0 ? R(0) : cl * ri[t][q];
            }
        for (int k = level - 1; k >= 1; k--) {
            __syncthreads();
            acc_store<true>(X, Bre, Bim);
            __syncthreads();
            Acc<P, NMAX> Q;
            acc_zero(Q);
            cmul<P, NMAX, false, false>(Are, Aim, Bre, Bim, Q);
            const R ck = (R)cinv[k];
#pragma unroll
            for (int t = 0; t < NCT; t++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                    X.r[t][q] = ((i == j && i < n) ? R(1) : R(0)) + ck * Q.r[t][q];
                    X.i[t][q] = ck * Q.i[t][q];
                }
        }
        return;
    }
    // a_j = 1 / j! = cinv[1] ... cinv[j], zero beyond the level (uniform
    // scalars, recomputed per block rather than indexed from a local array)
    auto fact_inv = [&](int j) {
        double f = 1.0;
        for (int k = 1; k <= j; k++) f *= cinv[k];
        return j <= level ? f : 0.0;
    };
    // R^2 = R R: R also as the right operand, then R^2 as the left one
#pragma unroll
    for (int t = 0; t < NCT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            X.r[t][q] = rr[t][q];
            X.i[t][q] = ri[t][q];
        }
    acc_store<true>(X, Bre, Bim);
    __syncthreads();
    Acc<P, NMAX> Q;
    acc_zero(Q);
    cmul<P, NMAX, false, false>(Are, Aim, Bre, Bim, Q);
    __syncthreads();
    acc_store<false>(Q, Are, Aim);
    // Horner in R^2 over the blocks a_{2i} + a_{2i+1} R, from the last
    const int nb = (level + 2) / 2;
    // X <- a_{2b} + a_{2b+1} R (+ Q)
    auto block = [&](int b, bool addq) {
        const R c0 = (R)fact_inv(2 * b), c1 = (R)fact_inv(2 * b + 1);
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                X.r[t][q] = ((i == j && i < n) ? c0 : R(0)) + c1 * rr[t][q] + (addq ? Q.r[t][q] : R(0));
                X.i[t][q] = c1 * ri[t][q] + (addq ? Q.i[t][q] : R(0));
            }
    };
    block(nb - 1, false);
    for (int b = nb - 2; b >= 0; b--) {
        __syncthreads();
        acc_store<true>(X, Bre, Bim);
        __syncthreads();
        acc_zero(Q);
        cmul<P, NMAX, false, false>(Are, Aim, Bre, Bim, Q);
        block(b, true);
    }
}

// Register budgets (launch-bounds minimum of wavefronts per SIMD).  FP64 at
// NMAX = 32 needs 236 VGPRs plus 64 AGPRs, one wavefront per SIMD; compiled
// for two (a few registers spilled to scratch) the kick+drift runs 22-25 %
// faster (SU(32) 8^4 2.18 -> 1.67 ms, SU(24) 8^4 2.11 -> 1.64 ms, SU(32)
// 96^2 1.63 -> 1.22 ms).  FP64 at NMAX = 16: four instead of three
// (126 VGPRs + 32 AGPRs before, 16-24 bytes spilled now): SU(16) 6x24^3
// 4.61 -> 4.46 ms, its action pass over FP32 links 3.12 -> 2.92 ms.  FP32
// and BF16x3 at NMAX = 16: five (96 VGPRs, at most a few spilled): SU(16)
// 6x24^3 FP32 2.25 -> 2.19 ms.
template <int P, int NMAX>
constexpr int kMinWaves = (P == kF64 && NMAX == 32) ? 2 : (P == kF64 && NMAX == 16) ? 4 : (NMAX == 16) ? 5 : 1;

// EN: the link size N equals NMAX (no padding), a compile-time constant, so
// that every bounds test on the padding folds away.
template <int P, int NMAX, bool Drift, typename LS = typename Pol<P>::R, bool ActionOnly = false, bool EN = false>
__global__ __launch_bounds__((Cfg<P, NMAX>::NT), (kMinWaves<P, NMAX>)) void k_kick(KickArgs a)
{
    typedef typename Pol<P>::R R;
    constexpr int NCT = Cfg<P, NMAX>::NCT;
    // 16-byte aligned: the BF16x3 products read the slots 16 bytes at a time
    __shared__ __align__(16) R Are[Cfg<P, NMAX>::SLOT];
    __shared__ __align__(16) R Aim[Cfg<P, NMAX>::SLOT];
    __shared__ __align__(16) R Bre[Cfg<P, NMAX>::SLOT];
    __shared__ __align__(16) R Bim[Cfg<P, NMAX>::SLOT];
    const Geom &g = a.g;
    const int n = EN ? NMAX : g.N;
    const long V = g.V;
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(lb % (unsigned)g.D);
    int cnt;
    const long s = parity_block((long)(lb / (unsigned)g.D), V, g.V0, 1, cnt);
    const long l = (long)mu * V + s;
    const LS *U = (const LS *)a.U;
    const long xp = g.nbr[(2 * mu) * V + s];

    // operands in the order the products consume them, per direction nu:
    // U_nu(x+mu), U_mu(x+nu) | U_nu(x) | U_nu(x+mu-nu), U_mu(x-nu) | U_nu(x-nu)
    constexpr bool PF = Cfg<P, NMAX>::PF;
    Stage<P, NMAX, LS, EN> s0, s1;
    int nu = mu == 0 ? 1 : 0;
    s0.template early<PF, false>(U, nu * V + xp, n);
    s1.template early<PF, false>(U, mu * V + g.nbr[(2 * nu) * V + s], n);
    Acc<P, NMAX> S, T;
    acc_zero(S);
    while (nu < g.D) {
        int nx = nu + 1;
        if (nx == mu) nx++;
        const long xpn = g.nbr[(2 * nu) * V + s];
        const long xmn = g.nbr[(2 * nu + 1) * V + s];
        const long xpmn = g.nbr[(2 * nu + 1) * V + xp];
        // S += U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
        __syncthreads();
        s0.template late<PF, false>(U, nu * V + xp, n, Are, Aim);
        s1.template late<PF, false>(U, mu * V + xpn, n, Bre, Bim);
        __syncthreads();
        s0.template early<PF, false>(U, nu * V + s, n);
        acc_zero(T);
        if constexpr (P == kBF16X3) {
            // the transpose T^T = conj(U_mu(x+nu)) U_nu(x+mu)^T from the slots
            // as they stand: the transposed store then writes T row-major
            // with one 8-byte write per plane and tile
            cmul<P, NMAX, true, false>(Bre, Bim, Are, Aim, T);
            __syncthreads();
            acc_store<true>(T, Are, Aim);
        } else {
            cmul<P, NMAX, false, true>(Are, Aim, Bre, Bim, T);
            __syncthreads();
            acc_store<false>(T, Are, Aim);
        }
        s0.template late<PF, false>(U, nu * V + s, n, Bre, Bim);
        __syncthreads();
        s0.template early<PF, true>(U, nu * V + xpmn, n);
        s1.template early<PF, false>(U, mu * V + xmn, n);
        cmul<P, NMAX, false, true>(Are, Aim, Bre, Bim, S);
        // S += U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
        __syncthreads();
        s0.template late<PF, true>(U, nu * V + xpmn, n, Are, Aim);
        s1.template late<PF, false>(U, mu * V + xmn, n, Bre, Bim);
        __syncthreads();
        s0.template early<PF, true>(U, nu * V + xmn, n);
        acc_zero(T);
        if constexpr (P == kBF16X3) {
            // T^T = conj(U_mu(x-nu)) conj(U_nu(x+mu-nu)), as above
            cmul<P, NMAX, true, true>(Bre, Bim, Are, Aim, T);
            __syncthreads();
            acc_store<true>(T, Are, Aim);
        } else {
            cmul<P, NMAX, true, true>(Are, Aim, Bre, Bim, T);
            __syncthreads();
            acc_store<false>(T, Are, Aim);
        }
        s0.template late<PF, true>(U, nu * V + xmn, n, Bre, Bim);
        __syncthreads();
        if (nx < g.D) {
            s0.template early<PF, false>(U, nx * V + xp, n);
            s1.template early<PF, false>(U, mu * V + g.nbr[(2 * nx) * V + s], n);
        } else {
            s0.template early<PF, false>(U, mu * V + s, n);
        }
        cmul<P, NMAX, false, false>(Are, Aim, Bre, Bim, S);
        nu = nx;
    }
    // W = U_mu(x) S
    __syncthreads();
    acc_store<true>(S, Are, Aim);
    s0.template late<PF, false>(U, mu * V + s, n, Bre, Bim);
    __syncthreads();
    acc_zero(T);
    cmul<P, NMAX, false, false>(Bre, Bim, Are, Aim, T);
    __syncthreads();
    acc_store_full(T, Are, Aim);
    __syncthreads();
    if constexpr (ActionOnly) {
        double sum[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                if (i == j && i < n) sum[0] += (double)Are[i * Cfg<P, NMAX>::LD + i];
            }
        block_reduce_store<3>(sum, cmp, a.part + 6 * (size_t)lb);
        return;
    }
    // TAH and kick on the compact momentum (sunmat.h:355-371, sun.cpp:716)
    double sum[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    kick_compact<P, NMAX, Drift>(a, l, n, Are, Aim, Bre, Bim, sum);
    if constexpr (Drift) {
        // rhs = i eps_d P into slot A; X = 1 + c_L rhs in registers; Horner
        // steps X = 1 + c_k (rhs X) (sunmat.h:200-248).  The new momentum is
        // in slot B as a full matrix.
        constexpr int LD = Cfg<P, NMAX>::LD;
        const R ed = (R)a.eps_drift;
        int level = a.level;
        if (level < 0 || level > 8) level = 8;
        __syncthreads();
        R pr[NCT][4], pim[NCT][4];
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                const bool in = EN || (i < n && j < n);
                pr[t][q] = in ? Bre[i * LD + j] : R(0);
                pim[t][q] = in ? Bim[i * LD + j] : R(0);
            }
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                put<P, NMAX, false, R>(Are, Aim, i, j, -(pim[t][q] * ed), pr[t][q] * ed);  // rhs as A
            }
        s0.template early<PF, true>(U, l, n);  // U_mu(x) for the final product
        // R = i eps_d P in registers; every lane has read slot B (the momentum)
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const R re = -(pim[t][q] * ed), im = pr[t][q] * ed;
                pr[t][q] = re;
                pim[t][q] = im;
            }
        __syncthreads();
        Acc<P, NMAX> X;
        taylor_poly<P, NMAX, R>(pr, pim, level, a.cinv, n, Are, Aim, Bre, Bim, X);
        __syncthreads();
        acc_store<false>(X, Bre, Bim);
        s0.template late<PF, true>(U, l, n, Are, Aim);
        __syncthreads();
        Acc<P, NMAX> Un;
        acc_zero(Un);
        R *ob = (R *)a.Uout + (size_t)l * 2 * n * n;
        if constexpr (P == kBF16X3) {
            // the transpose Un^T = U_mu(x)^T X^T (slot A holds U_mu(x)^T
            // row-major, slot B X row-major): lane (t, q) holds Un[col][row +
            // q], four consecutive elements of one row of the link
            cmul<P, NMAX, false, false>(Are, Aim, Bre, Bim, Un);
#pragma unroll
            for (int t = 0; t < NCT; t++) {
                const int i = acc_col<P, NMAX>(t), j0 = acc_row<P, NMAX>(0);
                if constexpr (EN) {
                    f32x4 *o4 = (f32x4 *)(ob + 2 * (i * NMAX + j0));
                    o4[0] = f32x4{Un.r[t][0], Un.i[t][0], Un.r[t][1], Un.i[t][1]};
                    o4[1] = f32x4{Un.r[t][2], Un.i[t][2], Un.r[t][3], Un.i[t][3]};
                } else {
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (i < n && j0 + q < n) {
                            ob[2 * (i * n + j0 + q)] = Un.r[t][q];
                            ob[2 * (i * n + j0 + q) + 1] = Un.i[t][q];
                        }
                }
            }
        } else {
            cmul<P, NMAX, false, false>(Bre, Bim, Are, Aim, Un);
#pragma unroll
            for (int t = 0; t < NCT; t++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                    if (i < n && j < n) {
                        ob[2 * (i * n + j)] = Un.r[t][q];
                        ob[2 * (i * n + j) + 1] = Un.i[t][q];
                    }
                }
        }
    }
    if (a.part) block_reduce_store<3>(sum, cmp, a.part + 6 * (size_t)lb);
}

// Standalone drift U <- T_L(i eps P) U of every link in place (the Horner
// loop of the fused kernel on the MFMA units; one workgroup per link, which
// reads and writes its own link only).
template <int P, int NMAX>
__global__ __launch_bounds__((Cfg<P, NMAX>::NT), (kMinWaves<P, NMAX>)) void k_drift(Geom g, void *Uv, const void *Pv,
                                                                                     double eps, int level,
                                                                                     Cinv9 cinv)
{
    typedef typename Pol<P>::R R;
    constexpr int NCT = Cfg<P, NMAX>::NCT;
    __shared__ __align__(16) R Are[Cfg<P, NMAX>::SLOT];
    __shared__ __align__(16) R Aim[Cfg<P, NMAX>::SLOT];
    __shared__ __align__(16) R Bre[Cfg<P, NMAX>::SLOT];
    __shared__ __align__(16) R Bim[Cfg<P, NMAX>::SLOT];
    const int n = g.N;
    const long l = blockIdx.x;
    R *U = (R *)Uv;
    const R *Pm = (const R *)Pv;
    const R ed = (R)eps;
    if (level < 0 || level > 8) level = 8;
    R pr[NCT][4], pim[NCT][4];
#pragma unroll
    for (int t = 0; t < NCT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
            pr[t][q] = pim[t][q] = R(0);
            if (i < n && j < n) pr[t][q] = mom_el(Pm, l, n, i, j, pim[t][q]);
        }
    // R = i eps P into slot A (zero outside n x n) and in registers; the
    // truncated exponential of R (sunmat.h:200-248)
#pragma unroll
    for (int t = 0; t < NCT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const R re = -(pim[t][q] * ed), im = pr[t][q] * ed;
            put<P, NMAX, false, R>(Are, Aim, acc_row<P, NMAX>(q), acc_col<P, NMAX>(t), re, im);
            pr[t][q] = re;
            pim[t][q] = im;
        }
    Acc<P, NMAX> X;
    taylor_poly<P, NMAX, R>(pr, pim, level, cinv.c, n, Are, Aim, Bre, Bim, X);
    __syncthreads();
    acc_store<false>(X, Bre, Bim);
    ld_link<P, NMAX, true>(Are, Aim, U, l, n);
    __syncthreads();
    Acc<P, NMAX> Un;
    acc_zero(Un);
    cmul<P, NMAX, false, false>(Bre, Bim, Are, Aim, Un);
    R *ob = U + (size_t)l * 2 * n * n;
#pragma unroll
    for (int t = 0; t < NCT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
            if (i < n && j < n) {
                ob[2 * (i * n + j)] = Un.r[t][q];
                ob[2 * (i * n + j) + 1] = Un.i[t][q];
            }
        }
}

template <int P, int NMAX>
hipError_t drift(const Geom &g, void *U, const void *Pm, double eps, int level, const double *cinv, hipStream_t st)
{
    Cinv9 c;
    for (int k = 0; k < 9; k++) c.c[k] = cinv[k];
    hipLaunchKernelGGL((k_drift<P, NMAX>), dim3((unsigned)(g.V * g.D)), dim3(Cfg<P, NMAX>::NT), 0, st, g, U, Pm, eps,
                       level, c);
    return hipGetLastError();
}

// sum ReTr W in FP64 arithmetic over FP32 (f32_links) or FP64 links
// (partials as the kick's)
template <int NMAX>
hipError_t action_f64(const KickArgs &a, bool f32_links, hipStream_t st)
{
    const unsigned nb = (unsigned)(a.g.V * a.g.D);
    constexpr int NT = Cfg<kF64, NMAX>::NT;
    if (f32_links) hipLaunchKernelGGL((k_kick<kF64, NMAX, false, float, true>), dim3(nb), dim3(NT), 0, st, a);
    else hipLaunchKernelGGL((k_kick<kF64, NMAX, false, double, true>), dim3(nb), dim3(NT), 0, st, a);
    return hipGetLastError();
}

template <int P, int NMAX>
hipError_t kick(const KickArgs &a, bool drift, hipStream_t st)
{
    typedef typename Pol<P>::R R;
    const unsigned nb = (unsigned)(a.g.V * a.g.D);
    constexpr int NT = Cfg<P, NMAX>::NT;
    const bool en = a.g.N == NMAX;
    if (drift && en) hipLaunchKernelGGL((k_kick<P, NMAX, true, R, false, true>), dim3(nb), dim3(NT), 0, st, a);
    else if (drift) hipLaunchKernelGGL((k_kick<P, NMAX, true>), dim3(nb), dim3(NT), 0, st, a);
    else if (en) hipLaunchKernelGGL((k_kick<P, NMAX, false, R, false, true>), dim3(nb), dim3(NT), 0, st, a);
    else hipLaunchKernelGGL((k_kick<P, NMAX, false>), dim3(nb), dim3(NT), 0, st, a);
    return hipGetLastError();
}

}  // namespace mfma
}  // namespace askja
