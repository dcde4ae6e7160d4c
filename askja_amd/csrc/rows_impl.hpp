// rows_impl.hpp -- kick (+ drift) for 4 <= N <= 8: one wavefront holds
// LPB = 64 / N links and each thread owns one row of one link.  The left
// operand of every product is the thread's row (registers), the right
// operand the whole matrix in a per-link LDS slot, so the chain products of
// the staples and of the Horner drift never send their left operand through
// LDS, and each workgroup is one wavefront (its barriers are cheap).
// Operation order per output element is that of the workgroup kernels
// (team_impl.hpp): k ascending from 0, the exact mode without contraction
// and with products rounded as (ar*br - ai*bi), as GSL's zgemm.
#pragma once
#include "kernels.hpp"
#include "common.hpp"

namespace askja {
namespace rows {

constexpr int kT = 64;

template <int N>
struct Cfg {
    static constexpr int LPB = kT / N;  // links per workgroup
    // LDS slot stride: N^2 + 1, so that the slots (read by broadcast, one
    // address per link) start on different banks
    static constexpr int SS = N * N + 1;
};

// c (+)= (ck a) op(B) for this thread's row a; B in LDS (planar, row stride N)
template <int N, bool DB, bool Acc, bool Exact, typename R>
__device__ inline void rowmul(const R (&ar)[N], const R (&ai)[N], const R *bre, const R *bim, R ck, R (&cr)[N],
                              R (&ci)[N])
{
#pragma unroll
    for (int j = 0; j < N; j++) {
        R sr = Acc ? cr[j] : R(0), si = Acc ? ci[j] : R(0);
        if constexpr (Exact) {
#pragma clang fp contract(off)
#pragma unroll
            for (int k = 0; k < N; k++) {
                const R a_r = ar[k] * ck, a_i = ai[k] * ck;
                const R b_r = DB ? bre[j * N + k] : bre[k * N + j];
                const R b_i = DB ? -bim[j * N + k] : bim[k * N + j];
                sr = sr + (a_r * b_r - a_i * b_i);
                si = si + (a_r * b_i + a_i * b_r);
            }
        } else {
#pragma unroll
            for (int k = 0; k < N; k++) {
                const R a_r = ar[k] * ck, a_i = ai[k] * ck;
                const R b_r = DB ? bre[j * N + k] : bre[k * N + j];
                const R b_i = DB ? -bim[j * N + k] : bim[k * N + j];
                sr = fma(a_r, b_r, sr);
                sr = fma(-a_i, b_i, sr);
                si = fma(a_r, b_i, si);
                si = fma(a_i, b_r, si);
            }
        }
        cr[j] = sr;
        ci[j] = si;
    }
}

// row i (or, with Dag, row i of the conjugate transpose) of link l
template <int N, bool Dag, typename R>
__device__ inline void ld_row(const R *__restrict__ U, long l, int i, R (&r)[N], R (&im)[N])
{
    const R *b = U + (size_t)l * 2 * N * N;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int e = Dag ? k * N + i : i * N + k;
        r[k] = b[2 * e];
        im[k] = Dag ? -b[2 * e + 1] : b[2 * e + 1];
    }
}

// this thread's row i of link l into the link's LDS slot
template <int N, typename R>
__device__ inline void st_row_from(const R *__restrict__ U, long l, int i, R *sre, R *sim)
{
    const R *b = U + (size_t)l * 2 * N * N + 2 * i * N;
#pragma unroll
    for (int k = 0; k < N; k++) {
        sre[i * N + k] = b[2 * k];
        sim[i * N + k] = b[2 * k + 1];
    }
}

template <int N, typename R>
__device__ inline void st_row(const R (&r)[N], const R (&im)[N], int i, R *sre, R *sim)
{
#pragma unroll
    for (int k = 0; k < N; k++) {
        sre[i * N + k] = r[k];
        sim[i * N + k] = im[k];
    }
}

template <typename R>
__device__ inline void mom_el(const R *__restrict__ b, int n, int i, int j, R &re, R &im)
{
    if (i == j) {
        re = b[i];
        im = R(0);
    } else if (i < j) {
        const int h = hup(n, i, j);
        re = b[h];
        im = b[h + 1];
    } else {
        const int h = hup(n, j, i);
        re = b[h];
        im = -b[h + 1];
    }
}

// Link of logical index t = (site block, direction): the D links of one site
// are consecutive, sites interleaved even / odd (parity_block with blocks of
// one site), workgroups XCD-contiguous.  MW = the wavefronts per SIMD the
// register allocation must allow (launch-bounds minimum).
template <int N, typename R, bool Exact, bool Drift, int MW>
__global__ __launch_bounds__(kT, MW) void k_kick(KickArgs a)
{
    constexpr int LPB = Cfg<N>::LPB;
    constexpr int SS = Cfg<N>::SS;
    __shared__ R Sre[LPB * SS], Sim[LPB * SS];
    const Geom &g = a.g;
    const long V = g.V;
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int slot = threadIdx.x / N, i = threadIdx.x % N;
    const long t = (long)lb * LPB + slot;
    bool ok = slot < LPB && t < V * g.D;
    int mu = 0;
    long s = 0;
    if (ok) {
        mu = (int)(t % g.D);
        int cnt;
        s = parity_block(t / g.D, V, g.V0, 1, cnt);
        ok = cnt > 0;
    }
    R *sre = Sre + (slot < LPB ? slot : 0) * SS, *sim = Sim + (slot < LPB ? slot : 0) * SS;
    const long l = (long)mu * V + s;
    const R *U = (const R *)a.U;
    const long xp = ok ? g.nbr[(2 * mu) * V + s] : 0;
    R wr[N], wi[N], tr[N], ti[N], ar_[N], ai_[N];
#pragma unroll
    for (int k = 0; k < N; k++) wr[k] = wi[k] = R(0);
    // every lane steps through every direction and every barrier; the links
    // of one workgroup differ in mu, so the work is guarded, not the loop
    for (int nu = 0; nu < g.D; nu++) {
        const bool act = ok && nu != mu;
        const long xpn = act ? g.nbr[(2 * nu) * V + s] : 0;
        const long xmn = act ? g.nbr[(2 * nu + 1) * V + s] : 0;
        const long xpmn = act ? g.nbr[(2 * nu + 1) * V + xp] : 0;
        if constexpr (Exact) {
            // ((U_mu(x) U_nu(x+mu)) U_mu(x+nu)^+) U_nu(x)^+
            __syncthreads();
            if (act) st_row_from<N>(U, nu * V + xp, i, sre, sim);
            __syncthreads();
            if (act) {
                ld_row<N, false>(U, mu * V + s, i, ar_, ai_);
                rowmul<N, false, false, true>(ar_, ai_, sre, sim, R(1), tr, ti);
            }
            __syncthreads();
            if (act) st_row_from<N>(U, mu * V + xpn, i, sre, sim);
            __syncthreads();
            if (act) rowmul<N, true, false, true>(tr, ti, sre, sim, R(1), ar_, ai_);
            __syncthreads();
            if (act) st_row_from<N>(U, nu * V + s, i, sre, sim);
            __syncthreads();
            if (act) {
                rowmul<N, true, false, true>(ar_, ai_, sre, sim, R(1), tr, ti);
                {
#pragma clang fp contract(off)
#pragma unroll
                    for (int k = 0; k < N; k++) {
                        wr[k] = wr[k] + tr[k];
                        wi[k] = wi[k] + ti[k];
                    }
                }
            }
            // ((U_mu(x) U_nu(x+mu-nu)^+) U_mu(x-nu)^+) U_nu(x-nu)
            __syncthreads();
            if (act) st_row_from<N>(U, nu * V + xpmn, i, sre, sim);
            __syncthreads();
            if (act) {
                ld_row<N, false>(U, mu * V + s, i, ar_, ai_);
                rowmul<N, true, false, true>(ar_, ai_, sre, sim, R(1), tr, ti);
            }
            __syncthreads();
            if (act) st_row_from<N>(U, mu * V + xmn, i, sre, sim);
            __syncthreads();
            if (act) rowmul<N, true, false, true>(tr, ti, sre, sim, R(1), ar_, ai_);
            __syncthreads();
            if (act) st_row_from<N>(U, nu * V + xmn, i, sre, sim);
            __syncthreads();
            if (act) {
                rowmul<N, false, false, true>(ar_, ai_, sre, sim, R(1), tr, ti);
                {
#pragma clang fp contract(off)
#pragma unroll
                    for (int k = 0; k < N; k++) {
                        wr[k] = wr[k] + tr[k];
                        wi[k] = wi[k] + ti[k];
                    }
                }
            }
        } else {
            // S += U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ and
            // S += U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu); the next right
            // operand's row is loaded while the current product runs
            R br[N], bi[N];
            if (act) {
                ld_row<N, false>(U, mu * V + xpn, i, br, bi);
                ld_row<N, false>(U, nu * V + xp, i, ar_, ai_);
            }
            __syncthreads();
            if (act) st_row<N>(br, bi, i, sre, sim);
            __syncthreads();
            if (act) {
                ld_row<N, false>(U, nu * V + s, i, br, bi);
                rowmul<N, true, false, false>(ar_, ai_, sre, sim, R(1), tr, ti);
            }
            __syncthreads();
            if (act) st_row<N>(br, bi, i, sre, sim);
            __syncthreads();
            if (act) {
                ld_row<N, false>(U, mu * V + xmn, i, br, bi);
                ld_row<N, true>(U, nu * V + xpmn, i, ar_, ai_);
                rowmul<N, true, true, false>(tr, ti, sre, sim, R(1), wr, wi);
            }
            __syncthreads();
            if (act) st_row<N>(br, bi, i, sre, sim);
            __syncthreads();
            if (act) {
                ld_row<N, false>(U, nu * V + xmn, i, br, bi);
                rowmul<N, true, false, false>(ar_, ai_, sre, sim, R(1), tr, ti);
            }
            __syncthreads();
            if (act) st_row<N>(br, bi, i, sre, sim);
            __syncthreads();
            if (act) rowmul<N, false, true, false>(tr, ti, sre, sim, R(1), wr, wi);
        }
    }
    if constexpr (!Exact) {
        // W = U_mu(x) S
        __syncthreads();
        if (ok) st_row<N>(wr, wi, i, sre, sim);
        __syncthreads();
        if (ok) {
            ld_row<N, false>(U, mu * V + s, i, ar_, ai_);
            rowmul<N, false, false, false>(ar_, ai_, sre, sim, R(1), wr, wi);
        }
    }
    // TAH and kick on this row (sunmat.h:355-371, sun.cpp:716)
    __syncthreads();
    if (ok) st_row<N>(wr, wi, i, sre, sim);
    __syncthreads();
    const R eps = (R)a.eps_kick, inv_n = (R)a.inv_n;
    R *P = (R *)a.P;
    double sum[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    R pr[N], pim[N], ppr[N], ppi[N];
    if (ok) {
#pragma clang fp contract(off)
        const R *pb = P + (size_t)l * N * N;
        R tri = R(0);
        for (int q = 0; q < N; q++) tri = tri + (sim[q * N + q] + sim[q * N + q]);
        const R corr = tri * inv_n;
#pragma unroll
        for (int j = 0; j < N; j++) {
            R inr, ini;
            mom_el(pb, N, i, j, inr, ini);
            ppr[j] = inr;
            ppi[j] = ini;
            if (i == j) {
                const R mi = (sim[i * N + i] + sim[i * N + i]) - corr;
                pr[j] = inr - (R(0.5) * mi) * eps;
                pim[j] = R(0);
            } else {
                const int lo = i < j ? i : j, hi = i < j ? j : i;
                const R dr = sre[lo * N + hi] - sre[hi * N + lo];
                const R di = sim[lo * N + hi] + sim[hi * N + lo];
                const R qr = (R(0.5) * di) * eps;
                const R qi = (R(0.5) * dr) * eps;
                R ur, ui;
                mom_el(pb, N, lo, hi, ur, ui);
                const R vr = ur - qr, vi = ui + qi;
                pr[j] = vr;
                pim[j] = i < j ? vi : -vi;
            }
        }
        sum[0] = (double)wr[i];
#pragma unroll
        for (int j = 0; j < N; j++) {
            sum[1] += (double)ppr[j] * (double)ppr[j] + (double)ppi[j] * (double)ppi[j];
            sum[2] += (double)pr[j] * (double)pr[j] + (double)pim[j] * (double)pim[j];
        }
    }
    // new momentum: the upper triangle of row i (every row of the link has
    // read the old one: one wavefront, and the stores follow all loads of it)
    __syncthreads();
    if (ok) {
        R *b = P + (size_t)l * N * N;
        R *bp = (!Drift && a.Pprev) ? (R *)a.Pprev + (size_t)l * N * N : nullptr;
        b[i] = pr[i];
        if (bp) bp[i] = ppr[i];
#pragma unroll
        for (int j = 0; j < N; j++) {
            if (j <= i) continue;
            const int h = hup(N, i, j);
            b[h] = pr[j];
            b[h + 1] = pim[j];
            if (bp) {
                bp[h] = ppr[j];
                bp[h + 1] = ppi[j];
            }
        }
    }
    if constexpr (Drift) {
        // rhs = i (eps_d P) as the left operand row; X in the slot; Horner
        // X = 1 + (c_k rhs) X, k = L-1 .. 1 (sunmat.h:200-248); then X U
        const R ed = (R)a.eps_drift;
        int level = a.level;
        if (level < 0 || level > 8) level = 8;
        R rr[N], ri[N], xr[N], xi[N];
        {
#pragma clang fp contract(off)
#pragma unroll
            for (int j = 0; j < N; j++) {
                rr[j] = -(pim[j] * ed);
                ri[j] = pr[j] * ed;
            }
            if (level == 0) {
#pragma unroll
                for (int j = 0; j < N; j++) {
                    xr[j] = i == j ? R(1) : R(0);
                    xi[j] = R(0);
                }
            } else {
                const R c = level == 1 ? R(1) : (R)a.cinv[level];
#pragma unroll
                for (int j = 0; j < N; j++) {
                    xr[j] = (i == j ? R(1) : R(0)) + rr[j] * c;
                    xi[j] = R(0) + ri[j] * c;
                }
            }
        }
        for (int k = level - 1; k >= 1; k--) {
            __syncthreads();
            if (ok) st_row<N>(xr, xi, i, sre, sim);
            __syncthreads();
            if (ok) {
                R qr[N], qi[N];
                rowmul<N, false, false, Exact>(rr, ri, sre, sim, (R)a.cinv[k], qr, qi);
#pragma unroll
                for (int j = 0; j < N; j++) {  // additions only: nothing to contract
                    xr[j] = (i == j ? R(1) : R(0)) + qr[j];
                    xi[j] = R(0) + qi[j];
                }
            }
        }
        __syncthreads();
        if (ok) st_row_from<N>(U, l, i, sre, sim);
        __syncthreads();
        if (ok) {
            R ur[N], ui[N];
            rowmul<N, false, false, Exact>(xr, xi, sre, sim, R(1), ur, ui);
            R *b = (R *)a.Uout + (size_t)l * 2 * N * N + 2 * i * N;
#pragma unroll
            for (int j = 0; j < N; j++) {
                b[2 * j] = ur[j];
                b[2 * j + 1] = ui[j];
            }
        }
    }
    if (a.part) block_reduce_store<3>(sum, cmp, a.part + 6 * (size_t)lb);
}

template <int N, typename R, bool Exact, int MW>
hipError_t kick_mw(const KickArgs &a, bool drift, unsigned nb, hipStream_t st)
{
    if (drift) hipLaunchKernelGGL((k_kick<N, R, Exact, true, MW>), dim3(nb), dim3(kT), 0, st, a);
    else hipLaunchKernelGGL((k_kick<N, R, Exact, false, MW>), dim3(nb), dim3(kT), 0, st, a);
    return hipGetLastError();
}

// The FP64 kick at N >= 7 keeps more values live than two wavefronts per
// SIMD allow (256 VGPRs plus spill AGPRs: one wavefront).  Where the grid
// gives each of the 1024 SIMDs more than about 1.5 workgroups, the variant
// compiled for two wavefronts (a few registers spilled to scratch) hides
// more latency: kick+drift SU(8) 10^4 245 -> 212 us, SU(7) 8^4 73.6 ->
// 58.8 us (1821 workgroups); with 1024 workgroups (SU(8) 8^3 x 4) the
// one-wavefront variant is faster (52.7 against 57.0 us).  The SU(8) kick
// without the drift already fits two wavefronts (242 VGPRs) and keeps it.
constexpr unsigned kTwoWaveMinBlocks = 1536;

template <int N, typename R, bool Exact>
hipError_t kick(const KickArgs &a, bool drift, hipStream_t st)
{
    const long nl = a.g.V * a.g.D;
    const unsigned nb = (unsigned)((nl + Cfg<N>::LPB - 1) / Cfg<N>::LPB);
    if constexpr (N >= 7 && !Exact && sizeof(R) == 8)
        if (nb >= kTwoWaveMinBlocks && (N == 7 || drift)) return kick_mw<N, R, Exact, 2>(a, drift, nb, st);
    return kick_mw<N, R, Exact, 1>(a, drift, nb, st);
}

}  // namespace rows
}  // namespace askja
