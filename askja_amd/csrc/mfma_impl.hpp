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
// Operands are staged in two LDS slots (planar re/im, row stride NMAX + 4).
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
    // operand prefetch through registers, except at NMAX = 64 in FP32, where
    // the extra registers would halve the occupancy
    static constexpr bool PF = !(P == kF32 && NMAX == 64);
    // LDS row stride in elements.  FP64 / FP32: NMAX + 4 puts the 16 rows a
    // wavefront reads for one operand on distinct banks (an unpadded stride
    // maps them all to one bank group).  BF16x3: NMAX + 8 bf16 (16 bytes)
    // keeps the 16-byte reads of eight consecutive rows on distinct banks.
    static constexpr int LD = NMAX + (P == kBF16X3 ? 8 : 4);
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
        // no contraction: x - hi must not fuse with a product that formed x
#pragma clang fp contract(off)
        const int o = TR ? c * LD + r : r * LD + c;
        __bf16 *rh = (__bf16 *)Sr, *ih = (__bf16 *)Si;
        const __bf16 hr = (__bf16)re, hi = (__bf16)im;
        rh[o] = hr;
        rh[NMAX * LD + o] = (__bf16)(re - (float)hr);
        ih[o] = hi;
        ih[NMAX * LD + o] = (__bf16)(im - (float)hi);
    } else {
        Sr[r * LD + c] = re;
        Si[r * LD + c] = im;
    }
}

// One half of the stacked product on the 16x16x4 forms: IM = false takes
// k < NMAX ([Ar] against [Br ; Bi] rows), IM = true the rest ([Ai] against
// [-Bi ; Br]).  Lane l holds A[row][k] and B[k][col], k = 4 ks + l/16.
template <int P, int NMAX, bool DA, bool DB, bool IM, typename R>
__device__ inline void cmul_half(const R *A, const R *Br, const R *Bi, Acc<P, NMAX> &acc)
{
    constexpr int NCT = Cfg<P, NMAX>::NCT, LD = Cfg<P, NMAX>::LD;
    const int lane = threadIdx.x & 63;
    const int row = 16 * wave_rowblock<P, NMAX>() + (lane & 15);
    const int t0 = wave_tile0<P, NMAX>();
#pragma unroll 4
    for (int ks = 0; ks < NMAX / 4; ks++) {
        const int k = 4 * ks + (lane >> 4);
        const R a = DA ? A[k * LD + row] : A[row * LD + k];
        const R av = (IM && DA) ? -a : a;
#pragma unroll
        for (int t = 0; t < NCT; t++) {
            const int col = 16 * (t0 + t) + (lane & 15);
            const R bre = DB ? Br[col * LD + k] : Br[k * LD + col];
            const R bim = DB ? -Bi[col * LD + k] : Bi[k * LD + col];
            const R br = IM ? -bim : bre, bi = IM ? bre : bim;
            if constexpr (P == kF64) {
                acc.r[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, br, acc.r[t], 0, 0, 0);
                acc.i[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bi, acc.i[t], 0, 0, 0);
            } else {
                acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, br, acc.r[t], 0, 0, 0);
                acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bi, acc.i[t], 0, 0, 0);
            }
        }
    }
}

// acc (+)= op(A) op(B) on the wavefront's rows and column tiles; A, B in LDS.
template <int P, int NMAX, bool DA, bool DB, typename R>
__device__ inline void cmul(const R *Ar, const R *Ai, const R *Br, const R *Bi, Acc<P, NMAX> &acc)
{
    if constexpr (P == kBF16X3) {
        // lane l: A[row][k0..k0+7] and B[k0..k0+7][col], k0 = 32 ks + 8 (l/16)
        // in the stacked index; the slots hold A and B in the orientation
        // that makes k contiguous (conjugation = sign flip of the im planes)
        constexpr int NCT = Cfg<P, NMAX>::NCT, LD = Cfg<P, NMAX>::LD, PL = NMAX * LD;
        const __bf16 *arh = (const __bf16 *)Ar, *aih = (const __bf16 *)Ai;
        const __bf16 *brh = (const __bf16 *)Br, *bih = (const __bf16 *)Bi;
        const int lane = threadIdx.x & 63;
        const int row = 16 * wave_rowblock<P, NMAX>() + (lane & 15);
        const int t0 = wave_tile0<P, NMAX>();
        const int kq = 8 * (lane >> 4);
        constexpr unsigned SG = 0x80008000u;
#pragma unroll 1  // unrolled, the 16-byte operand loads of every step stay live
        for (int ks = 0; ks < (2 * NMAX) / 32; ks++) {
            const int k0 = 32 * ks + kq;           // first of this lane's 8 stacked k
            const bool imag = k0 >= NMAX;          // all 8 in the same half
            const int kb = imag ? k0 - NMAX : k0;  // column of A / row of B
            // [Ar | Ai] (Ai -> -Ai for A^+)
            const __bf16 *ap = (imag ? aih : arh) + row * LD + kb;
            const unsigned ma = (DA && imag) ? SG : 0u;
            const bf16x8 ah = xsign(ld8(ap), ma), al = xsign(ld8(ap + PL), ma);
            // real output: [Br ; -Bi], imaginary output: [Bi ; Br]
            // (Bi -> -Bi for B^+)
            const __bf16 *bp = imag ? bih : brh, *cp = imag ? brh : bih;
            const unsigned mb = (imag && !DB) ? SG : 0u, mc = (!imag && DB) ? SG : 0u;
#pragma unroll
            for (int t = 0; t < NCT; t++) {
                const int o = (16 * (t0 + t) + (lane & 15)) * LD + kb;
                const bf16x8 bh = xsign(ld8(bp + o), mb), bl = xsign(ld8(bp + PL + o), mb);
                const bf16x8 ch = xsign(ld8(cp + o), mc), cl = xsign(ld8(cp + PL + o), mc);
                acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc.r[t], 0, 0, 0);
                acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc.r[t], 0, 0, 0);
                acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc.r[t], 0, 0, 0);
                acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, ch, acc.i[t], 0, 0, 0);
                acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, cl, acc.i[t], 0, 0, 0);
                acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, ch, acc.i[t], 0, 0, 0);
            }
        }
    } else {
        cmul_half<P, NMAX, DA, DB, false>(Ar, Br, Bi, acc);
        cmul_half<P, NMAX, DA, DB, true>(Ai, Br, Bi, acc);
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

// accumulators -> operand slot (orientation TR, see put)
template <bool TR, int P, int NMAX, typename R>
__device__ inline void acc_store(const Acc<P, NMAX> &a, R *Sr, R *Si)
{
#pragma unroll
    for (int t = 0; t < Cfg<P, NMAX>::NCT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            put<P, NMAX, TR, R>(Sr, Si, acc_row<P, NMAX>(q), acc_col<P, NMAX>(t), a.r[t][q], a.i[t][q]);
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

// global link (matrix contiguous, n x n complex) -> zero-padded slot
template <int P, int NMAX, bool TR, typename R, typename S>
__device__ inline void ld_link(R *Sr, R *Si, const S *__restrict__ f, long l, int n)
{
    constexpr int NT = Cfg<P, NMAX>::NT;
    const S *b = f + (size_t)l * 2 * n * n;
    for (int e = threadIdx.x; e < NMAX * NMAX; e += NT) {
        const int r = e / NMAX, c = e % NMAX;
        const bool in = r < n && c < n;
        put<P, NMAX, TR, R>(Sr, Si, r, c, in ? (R)b[2 * (r * n + c)] : R(0), in ? (R)b[2 * (r * n + c) + 1] : R(0));
    }
}

// Register staging of one link operand: fetch() issues its global loads,
// stash() writes it to an LDS slot.  With PF the fetch of the next operands
// is issued before the current product and the stash follows it, so the
// loads run under the MFMA work; without PF the operand is loaded straight
// into its slot at the stash point (fewer live registers).
template <int P, int NMAX, typename S>
struct Stage {
    static constexpr int NT = Cfg<P, NMAX>::NT, E = NMAX * NMAX / NT;
    static_assert(E * NT == NMAX * NMAX, "slot not a multiple of the workgroup");
    typedef S S2 __attribute__((ext_vector_type(2)));
    S2 v[E];
    __device__ inline void fetch(const S *__restrict__ f, long l, int n)
    {
        const S2 *b = (const S2 *)(f + (size_t)l * 2 * n * n);
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * NT, r = e / NMAX, c = e % NMAX;
            v[j] = (r < n && c < n) ? b[r * n + c] : S2{S(0), S(0)};
        }
    }
    template <bool TR, typename R>
    __device__ inline void stash(R *Sr, R *Si) const
    {
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int e = threadIdx.x + j * NT, r = e / NMAX, c = e % NMAX;
            put<P, NMAX, TR, R>(Sr, Si, r, c, (R)v[j][0], (R)v[j][1]);
        }
    }
    template <bool PF>
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
            ld_link<P, NMAX, TR>(Sr, Si, f, l, n);
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

// LS: link storage type (FP32 links with FP64 arithmetic for the exact
// action of the reduced-precision modes).  ActionOnly: only sum ReTr W
// (= 4 sum of the plaquettes touching the link) into the partials.
template <int P, int NMAX, bool Drift, typename LS = typename Pol<P>::R, bool ActionOnly = false>
__global__ __launch_bounds__((Cfg<P, NMAX>::NT)) void k_kick(KickArgs a)
{
    typedef typename Pol<P>::R R;
    constexpr int NCT = Cfg<P, NMAX>::NCT;
    // 16-byte aligned: the BF16x3 products read the slots 16 bytes at a time
    __shared__ __align__(16) R Are[Cfg<P, NMAX>::SLOT];
    __shared__ __align__(16) R Aim[Cfg<P, NMAX>::SLOT];
    __shared__ __align__(16) R Bre[Cfg<P, NMAX>::SLOT];
    __shared__ __align__(16) R Bim[Cfg<P, NMAX>::SLOT];
    const Geom &g = a.g;
    const int n = g.N;
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
    Stage<P, NMAX, LS> s0, s1;
    int nu = mu == 0 ? 1 : 0;
    s0.template early<PF>(U, nu * V + xp, n);
    s1.template early<PF>(U, mu * V + g.nbr[(2 * nu) * V + s], n);
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
        s0.template early<PF>(U, nu * V + s, n);
        acc_zero(T);
        cmul<P, NMAX, false, true>(Are, Aim, Bre, Bim, T);
        __syncthreads();
        acc_store<false>(T, Are, Aim);
        s0.template late<PF, false>(U, nu * V + s, n, Bre, Bim);
        __syncthreads();
        s0.template early<PF>(U, nu * V + xpmn, n);
        s1.template early<PF>(U, mu * V + xmn, n);
        cmul<P, NMAX, false, true>(Are, Aim, Bre, Bim, S);
        // S += U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
        __syncthreads();
        s0.template late<PF, true>(U, nu * V + xpmn, n, Are, Aim);
        s1.template late<PF, false>(U, mu * V + xmn, n, Bre, Bim);
        __syncthreads();
        s0.template early<PF>(U, nu * V + xmn, n);
        acc_zero(T);
        cmul<P, NMAX, true, true>(Are, Aim, Bre, Bim, T);
        __syncthreads();
        acc_store<false>(T, Are, Aim);
        s0.template late<PF, true>(U, nu * V + xmn, n, Bre, Bim);
        __syncthreads();
        if (nx < g.D) {
            s0.template early<PF>(U, nx * V + xp, n);
            s1.template early<PF>(U, mu * V + g.nbr[(2 * nx) * V + s], n);
        } else {
            s0.template early<PF>(U, mu * V + s, n);
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
    // TAH and kick on the lane's entries (sunmat.h:355-371, sun.cpp:716)
    const R eps = (R)a.eps_kick, inv_n = (R)a.inv_n;
    R *Pm = (R *)a.P;
    R tri = R(0);
    for (int i = 0; i < n; i++) tri += Aim[i * Cfg<P, NMAX>::LD + i] + Aim[i * Cfg<P, NMAX>::LD + i];
    const R corr = tri * inv_n;
    double sum[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    R pr[NCT][4], pim[NCT][4], opr[NCT][4], opi[NCT][4];
#pragma unroll
    for (int t = 0; t < NCT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
            pr[t][q] = pim[t][q] = opr[t][q] = opi[t][q] = R(0);
            if (i >= n || j >= n) continue;
            R ii;
            const R ir = mom_el(Pm, l, n, i, j, ii);
            opr[t][q] = ir;
            opi[t][q] = ii;
            if (i == j) {
                const R mi = (Aim[i * Cfg<P, NMAX>::LD + i] + Aim[i * Cfg<P, NMAX>::LD + i]) - corr;
                pr[t][q] = ir - (R(0.5) * mi) * eps;
                sum[0] += (double)Are[i * Cfg<P, NMAX>::LD + i];
            } else {
                const int lo = i < j ? i : j, hi = i < j ? j : i;
                const R dr = Are[lo * Cfg<P, NMAX>::LD + hi] - Are[hi * Cfg<P, NMAX>::LD + lo];
                const R di = Aim[lo * Cfg<P, NMAX>::LD + hi] + Aim[hi * Cfg<P, NMAX>::LD + lo];
                R ui;
                const R ur = mom_el(Pm, l, n, lo, hi, ui);
                const R vr = ur - (R(0.5) * di) * eps, vi = ui + (R(0.5) * dr) * eps;
                pr[t][q] = vr;
                pim[t][q] = i < j ? vi : -vi;
            }
            sum[1] += (double)ir * ir + (double)ii * ii;
            sum[2] += (double)pr[t][q] * pr[t][q] + (double)pim[t][q] * pim[t][q];
        }
    __syncthreads();  // every lane has read the old momentum
    {
        R *b = Pm + (size_t)l * n * n;
        R *bp = (!Drift && a.Pprev) ? (R *)a.Pprev + (size_t)l * n * n : nullptr;
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                if (i >= n || j >= n || i > j) continue;
                if (i == j) {
                    b[i] = pr[t][q];
                    if (bp) bp[i] = opr[t][q];
                } else {
                    const int h = hup(n, i, j);
                    b[h] = pr[t][q];
                    b[h + 1] = pim[t][q];
                    if (bp) {
                        bp[h] = opr[t][q];
                        bp[h + 1] = opi[t][q];
                    }
                }
            }
    }
    if constexpr (Drift) {
        // rhs = i eps_d P into slot A; X = 1 + c_L rhs in registers; Horner
        // steps X = 1 + c_k (rhs X) (sunmat.h:200-248)
        const R ed = (R)a.eps_drift;
        int level = a.level;
        if (level < 0 || level > 8) level = 8;
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                put<P, NMAX, false, R>(Are, Aim, i, j, -(pim[t][q] * ed), pr[t][q] * ed);  // rhs as A
            }
        s0.template early<PF>(U, l, n);  // U_mu(x) for the final product
        Acc<P, NMAX> X;
        const R cl = level <= 1 ? R(1) : (R)a.cinv[level];
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                const R d = (i == j && i < n) ? R(1) : R(0);
                X.r[t][q] = level == 0 ? d : d - cl * (pim[t][q] * ed);
                X.i[t][q] = level == 0 ? R(0) : cl * (pr[t][q] * ed);
            }
        for (int k = level - 1; k >= 1; k--) {
            __syncthreads();
            acc_store<true>(X, Bre, Bim);
            __syncthreads();
            Acc<P, NMAX> Q;
            acc_zero(Q);
            cmul<P, NMAX, false, false>(Are, Aim, Bre, Bim, Q);
            const R ck = (R)a.cinv[k];
#pragma unroll
            for (int t = 0; t < NCT; t++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = acc_row<P, NMAX>(q), j = acc_col<P, NMAX>(t);
                    X.r[t][q] = ((i == j && i < n) ? R(1) : R(0)) + ck * Q.r[t][q];
                    X.i[t][q] = ck * Q.i[t][q];
                }
        }
        __syncthreads();
        acc_store<false>(X, Bre, Bim);
        s0.template late<PF, true>(U, l, n, Are, Aim);
        __syncthreads();
        Acc<P, NMAX> Un;
        acc_zero(Un);
        cmul<P, NMAX, false, false>(Bre, Bim, Are, Aim, Un);
        R *ob = (R *)a.Uout + (size_t)l * 2 * n * n;
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
    block_reduce_store<3>(sum, cmp, a.part + 6 * (size_t)lb);
}

// sum ReTr W in FP64 arithmetic over FP32 links (partials as the kick's)
template <int NMAX>
hipError_t action_f64(const KickArgs &a, hipStream_t st)
{
    const unsigned nb = (unsigned)(a.g.V * a.g.D);
    constexpr int NT = Cfg<kF64, NMAX>::NT;
    hipLaunchKernelGGL((k_kick<kF64, NMAX, false, float, true>), dim3(nb), dim3(NT), 0, st, a);
    return hipGetLastError();
}

template <int P, int NMAX>
hipError_t kick(const KickArgs &a, bool drift, hipStream_t st)
{
    const unsigned nb = (unsigned)(a.g.V * a.g.D);
    constexpr int NT = Cfg<P, NMAX>::NT;
    if (drift) hipLaunchKernelGGL((k_kick<P, NMAX, true>), dim3(nb), dim3(NT), 0, st, a);
    else hipLaunchKernelGGL((k_kick<P, NMAX, false>), dim3(nb), dim3(NT), 0, st, a);
    return hipGetLastError();
}

}  // namespace mfma
}  // namespace askja
