// mfma_impl.hpp -- kick (+ drift) for 9 <= N <= 64 in split-precision
// BF16x3 on the MFMA units: FP32 links and momenta; every complex link
// product C = A B is computed as two real products of the stacked form
//   Re C = [Ar | Ai] [Br ; -Bi],   Im C = [Ar | Ai] [Bi ; Br]   (K = 2 NMAX)
// with each FP32 operand split as x = hi + lo (hi = bf16(x), lo = bf16(x - hi))
// and hi*hi + hi*lo + lo*hi accumulated in FP32 by v_mfma_f32_16x16x32_bf16.
//
// One workgroup per link, one wavefront per 16-row block of the (zero
// padded) NMAX x NMAX matrices; lane l of a wavefront holds, for every
// 16-column tile, the outputs at rows 4*(l/16)+i (i < 4), column l%16 of the
// tile (the MFMA accumulator layout).  Operands are staged in two LDS slots
// (planar re/im, row stride NMAX).
#pragma once
#include "kernels.hpp"
#include "common.hpp"

namespace askja {
namespace mfma {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int NMAX>
struct Cfg {
    static constexpr int NT = 64 * (NMAX / 16);  // one wavefront per row block
    static constexpr int NCT = NMAX / 16;        // column tiles
    static constexpr int SLOT = NMAX * NMAX;
};

template <int NMAX>
struct Acc {
    f32x4 r[NMAX / 16], i[NMAX / 16];
};

template <int NMAX>
__device__ inline void acc_zero(Acc<NMAX> &a)
{
#pragma unroll
    for (int t = 0; t < NMAX / 16; t++) {
        a.r[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        a.i[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
}

__device__ inline void split8(const float (&x)[8], bf16x8 &hi, bf16x8 &lo)
{
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const __bf16 h = (__bf16)x[j];
        hi[j] = h;
        lo[j] = (__bf16)(x[j] - (float)h);
    }
}

// acc (+)= op(A) op(B) on the wavefront's 16-row block; A, B in LDS slots.
template <int NMAX, bool DA, bool DB>
__device__ inline void cmul(const float *Ar, const float *Ai, const float *Br, const float *Bi, Acc<NMAX> &acc)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = 16 * w + (lane & 15);
    const int kq = 8 * (lane >> 4);
#pragma unroll
    for (int ks = 0; ks < (2 * NMAX) / 32; ks++) {
        const int k0 = 32 * ks + kq;           // first of this lane's 8 stacked k
        const bool imag = k0 >= NMAX;          // all 8 in the same half
        const int kb = imag ? k0 - NMAX : k0;  // column of A / row of B
        float av[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = kb + j;
            // op(A)[row][k]
            const float re = DA ? Ar[k * NMAX + row] : Ar[row * NMAX + k];
            const float im = DA ? -Ai[k * NMAX + row] : Ai[row * NMAX + k];
            av[j] = imag ? im : re;
        }
        bf16x8 ah, al;
        split8(av, ah, al);
#pragma unroll
        for (int t = 0; t < NMAX / 16; t++) {
            const int col = 16 * t + (lane & 15);
            float brv[8], biv[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = kb + j;
                // op(B)[k][col]
                const float re = DB ? Br[col * NMAX + k] : Br[k * NMAX + col];
                const float im = DB ? -Bi[col * NMAX + k] : Bi[k * NMAX + col];
                brv[j] = imag ? -im : re;  // [Br ; -Bi]
                biv[j] = imag ? re : im;   // [Bi ; Br]
            }
            bf16x8 bh, bl, ch, cl;
            split8(brv, bh, bl);
            split8(biv, ch, cl);
            acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc.r[t], 0, 0, 0);
            acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc.r[t], 0, 0, 0);
            acc.r[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc.r[t], 0, 0, 0);
            acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, ch, acc.i[t], 0, 0, 0);
            acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, cl, acc.i[t], 0, 0, 0);
            acc.i[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, ch, acc.i[t], 0, 0, 0);
        }
    }
}

// row / column of accumulator entry (t, q) of this lane
__device__ inline int acc_row(int q) { return 16 * (threadIdx.x >> 6) + 4 * ((threadIdx.x & 63) >> 4) + q; }
__device__ inline int acc_col(int t) { return 16 * t + (threadIdx.x & 15); }

template <int NMAX>
__device__ inline void acc_store(const Acc<NMAX> &a, float *Sr, float *Si)
{
#pragma unroll
    for (int t = 0; t < NMAX / 16; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Sr[acc_row(q) * NMAX + acc_col(t)] = a.r[t][q];
            Si[acc_row(q) * NMAX + acc_col(t)] = a.i[t][q];
        }
}

// global link (matrix contiguous, n x n complex) -> zero-padded slot
template <int NMAX>
__device__ inline void ld_link(float *Sr, float *Si, const float *__restrict__ f, long l, int n)
{
    constexpr int NT = Cfg<NMAX>::NT;
    const float *b = f + (size_t)l * 2 * n * n;
    for (int e = threadIdx.x; e < NMAX * NMAX; e += NT) {
        const int r = e / NMAX, c = e % NMAX;
        const bool in = r < n && c < n;
        Sr[e] = in ? b[2 * (r * n + c)] : 0.f;
        Si[e] = in ? b[2 * (r * n + c) + 1] : 0.f;
    }
}

template <int NMAX>
__device__ inline float mom_el(const float *__restrict__ P, long l, int n, int i, int j, float &im)
{
    const float *b = P + (size_t)l * n * n;
    if (i == j) {
        im = 0.f;
        return b[i];
    }
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int h = hup(n, lo, hi);
    im = i < j ? b[h + 1] : -b[h + 1];
    return b[h];
}

template <int NMAX, bool Drift>
__global__ __launch_bounds__(Cfg<NMAX>::NT) void k_kick(KickArgs a)
{
    constexpr int NCT = Cfg<NMAX>::NCT;
    __shared__ float Are[Cfg<NMAX>::SLOT], Aim[Cfg<NMAX>::SLOT], Bre[Cfg<NMAX>::SLOT], Bim[Cfg<NMAX>::SLOT];
    const Geom &g = a.g;
    const int n = g.N;
    const long V = g.V;
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(lb % (unsigned)g.D);
    int cnt;
    const long s = parity_block((long)(lb / (unsigned)g.D), V, g.V0, 1, cnt);
    const long l = (long)mu * V + s;
    const float *U = (const float *)a.U;
    const long xp = g.nbr[(2 * mu) * V + s];

    Acc<NMAX> S, T;
    acc_zero(S);
    for (int nu = 0; nu < g.D; nu++) {
        if (nu == mu) continue;
        const long xpn = g.nbr[(2 * nu) * V + s];
        const long xmn = g.nbr[(2 * nu + 1) * V + s];
        const long xpmn = g.nbr[(2 * nu + 1) * V + xp];
        // S += U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
        __syncthreads();
        ld_link<NMAX>(Are, Aim, U, nu * V + xp, n);
        ld_link<NMAX>(Bre, Bim, U, mu * V + xpn, n);
        __syncthreads();
        acc_zero(T);
        cmul<NMAX, false, true>(Are, Aim, Bre, Bim, T);
        __syncthreads();
        acc_store(T, Are, Aim);
        ld_link<NMAX>(Bre, Bim, U, nu * V + s, n);
        __syncthreads();
        cmul<NMAX, false, true>(Are, Aim, Bre, Bim, S);
        // S += U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
        __syncthreads();
        ld_link<NMAX>(Are, Aim, U, nu * V + xpmn, n);
        ld_link<NMAX>(Bre, Bim, U, mu * V + xmn, n);
        __syncthreads();
        acc_zero(T);
        cmul<NMAX, true, true>(Are, Aim, Bre, Bim, T);
        __syncthreads();
        acc_store(T, Are, Aim);
        ld_link<NMAX>(Bre, Bim, U, nu * V + xmn, n);
        __syncthreads();
        cmul<NMAX, false, false>(Are, Aim, Bre, Bim, S);
    }
    // W = U_mu(x) S
    __syncthreads();
    acc_store(S, Are, Aim);
    ld_link<NMAX>(Bre, Bim, U, mu * V + s, n);
    __syncthreads();
    acc_zero(T);
    cmul<NMAX, false, false>(Bre, Bim, Are, Aim, T);
    __syncthreads();
    acc_store(T, Are, Aim);
    __syncthreads();
    // TAH and kick on the lane's entries
    const float eps = (float)a.eps_kick, inv_n = (float)a.inv_n;
    float *P = (float *)a.P;
    float tri = 0.f;
    for (int i = 0; i < n; i++) tri += Aim[i * NMAX + i] + Aim[i * NMAX + i];
    const float corr = tri * inv_n;
    double sum[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    float pr[NCT][4], pim[NCT][4], opr[NCT][4], opi[NCT][4];
#pragma unroll
    for (int t = 0; t < NCT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = acc_row(q), j = acc_col(t);
            pr[t][q] = pim[t][q] = opr[t][q] = opi[t][q] = 0.f;
            if (i >= n || j >= n) continue;
            float ii;
            const float ir = mom_el<NMAX>(P, l, n, i, j, ii);
            opr[t][q] = ir;
            opi[t][q] = ii;
            if (i == j) {
                const float mi = (Aim[i * NMAX + i] + Aim[i * NMAX + i]) - corr;
                pr[t][q] = ir - (0.5f * mi) * eps;
                sum[0] += (double)Are[i * NMAX + i];
            } else {
                const int lo = i < j ? i : j, hi = i < j ? j : i;
                const float dr = Are[lo * NMAX + hi] - Are[hi * NMAX + lo];
                const float di = Aim[lo * NMAX + hi] + Aim[hi * NMAX + lo];
                float ui;
                const float ur = mom_el<NMAX>(P, l, n, lo, hi, ui);
                const float vr = ur - (0.5f * di) * eps, vi = ui + (0.5f * dr) * eps;
                pr[t][q] = vr;
                pim[t][q] = i < j ? vi : -vi;
            }
            sum[1] += (double)ir * ir + (double)ii * ii;
            sum[2] += (double)pr[t][q] * pr[t][q] + (double)pim[t][q] * pim[t][q];
        }
    __syncthreads();  // every lane has read the old momentum
    {
        float *b = P + (size_t)l * n * n;
        float *bp = (!Drift && a.Pprev) ? (float *)a.Pprev + (size_t)l * n * n : nullptr;
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row(q), j = acc_col(t);
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
        // rhs = i eps_d P into slot A; X = 1 + rhs c_L in slot B (lane entries)
        const float ed = (float)a.eps_drift;
        int level = a.level;
        if (level < 0 || level > 8) level = 8;
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row(q), j = acc_col(t);
                Are[i * NMAX + j] = -(pim[t][q] * ed);
                Aim[i * NMAX + j] = pr[t][q] * ed;
            }
        Acc<NMAX> X;
        const float cl = level <= 1 ? 1.f : (float)a.cinv[level];
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row(q), j = acc_col(t);
                const bool d = i == j && i < n;
                X.r[t][q] = level == 0 ? (d ? 1.f : 0.f) : (d ? 1.f : 0.f) - cl * (pim[t][q] * ed);
                X.i[t][q] = level == 0 ? 0.f : cl * (pr[t][q] * ed);
            }
        for (int k = level - 1; k >= 1; k--) {
            __syncthreads();
            acc_store(X, Bre, Bim);
            __syncthreads();
            Acc<NMAX> Q;
            acc_zero(Q);
            cmul<NMAX, false, false>(Are, Aim, Bre, Bim, Q);
            const float ck = (float)a.cinv[k];
#pragma unroll
            for (int t = 0; t < NCT; t++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = acc_row(q), j = acc_col(t);
                    X.r[t][q] = (i == j && i < n ? 1.f : 0.f) + ck * Q.r[t][q];
                    X.i[t][q] = ck * Q.i[t][q];
                }
        }
        __syncthreads();
        acc_store(X, Bre, Bim);
        ld_link<NMAX>(Are, Aim, U, l, n);
        __syncthreads();
        Acc<NMAX> Un;
        acc_zero(Un);
        cmul<NMAX, false, false>(Bre, Bim, Are, Aim, Un);
        float *ob = (float *)a.Uout + (size_t)l * 2 * n * n;
#pragma unroll
        for (int t = 0; t < NCT; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = acc_row(q), j = acc_col(t);
                if (i < n && j < n) {
                    ob[2 * (i * n + j)] = Un.r[t][q];
                    ob[2 * (i * n + j) + 1] = Un.i[t][q];
                }
            }
    }
    block_reduce_store<3>(sum, cmp, a.part + 6 * (size_t)lb);
}

template <int NMAX>
hipError_t kick(const KickArgs &a, bool drift, hipStream_t st)
{
    const unsigned nb = (unsigned)(a.g.V * a.g.D);
    constexpr int NT = Cfg<NMAX>::NT;
    if (drift) hipLaunchKernelGGL((k_kick<NMAX, true>), dim3(nb), dim3(NT), 0, st, a);
    else hipLaunchKernelGGL((k_kick<NMAX, false>), dim3(nb), dim3(NT), 0, st, a);
    return hipGetLastError();
}

}  // namespace mfma
}  // namespace askja
