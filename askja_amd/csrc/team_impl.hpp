// team_impl.hpp -- MD kernels for 4 <= N <= 64: one workgroup of NT threads
// per link, matrices staged in LDS (two slots, planar re/im), each thread
// owning the output elements e = tid + q*NT (q < EPL) of every product.
// Included by team<NMAX>.hip, which instantiates one size class.
#pragma once
#include "kernels.hpp"
#include "common.hpp"

namespace askja {
namespace team {

template <int NMAX, int NT, typename R>
struct Cfg {
    static constexpr int EPL = (NMAX * NMAX + NT - 1) / NT;
    static constexpr int SLOT = NMAX * NMAX;
};

// global (matrix-contiguous link) -> LDS slot
template <int NT, typename R>
__device__ inline void ld_link(R *sre, R *sim, const R *__restrict__ f, long l, int n)
{
    const int nr = 2 * n * n;
    const R *b = f + (size_t)l * nr;
    for (int r = threadIdx.x; r < nr; r += NT) {
        const R v = b[r];
        if (r & 1) sim[r >> 1] = v;
        else sre[r >> 1] = v;
    }
}

// Register staging of one link operand: fetch() issues the global loads of
// link l (K = 2 NMAX^2 / NT values per thread), put() writes them to an LDS
// slot.  The kick fetches the next operand before the current product and
// puts it after, so that the loads run under the product instead of in
// series with it.  That pays where a link is one small matrix (NMAX = 8, two
// values per thread: the small lattices of 4 <= N <= 8, latency bound); for
// the larger size classes the staging registers would cost occupancy or
// spill, and there fetch() only records the link and put() loads it straight
// into the slot.
template <int NT, int K, typename R, bool Stage = (K <= 2)>
struct Pref {
    R v[K];
    __device__ inline void fetch(const R *__restrict__ f, long l, int n)
    {
        const int nr = 2 * n * n;
        const R *b = f + (size_t)l * nr;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int r = (int)threadIdx.x + k * NT;
            v[k] = r < nr ? b[r] : R(0);
        }
    }
    __device__ inline void put(R *sre, R *sim, int n) const
    {
        const int nr = 2 * n * n;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int r = (int)threadIdx.x + k * NT;
            if (r < nr) {
                if (r & 1) sim[r >> 1] = v[k];
                else sre[r >> 1] = v[k];
            }
        }
    }
};

template <int NT, int K, typename R>
struct Pref<NT, K, R, false> {
    const R *f = nullptr;
    long l = 0;
    __device__ inline void fetch(const R *__restrict__ f_, long l_, int) { f = f_, l = l_; }
    __device__ inline void put(R *sre, R *sim, int n) const { ld_link<NT>(sre, sim, f, l, n); }
};

// Owned outputs: element index, row, column, validity.
template <int EPL, int NT>
struct Own {
    int i[EPL], j[EPL];
    bool ok[EPL];
    __device__ inline void init(int n)
    {
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            const int e = threadIdx.x + q * NT;
            ok[q] = e < n * n;
            i[q] = ok[q] ? e / n : 0;
            j[q] = ok[q] ? e % n : 0;
        }
    }
};

// c (+)= op(ck*A) op(B) for the owned elements; k ascending from 0.
// Exact: no FMA and products rounded as (ar*br - ai*bi), as GSL's zgemm.
template <bool DA, bool DB, bool Acc, bool Exact, int EPL, int NT, typename R>
__device__ inline void mul(const Own<EPL, NT> &o, const R *ar_, const R *ai_, const R *br_, const R *bi_, int n,
                           R ck, R (&cr)[EPL], R (&ci)[EPL])
{
    if constexpr (!Acc) {
#pragma unroll
        for (int q = 0; q < EPL; q++) cr[q] = ci[q] = R(0);
    }
    if constexpr (Exact) {
#pragma clang fp contract(off)
        for (int k = 0; k < n; k++) {
#pragma unroll
            for (int q = 0; q < EPL; q++) {
                const int ia = DA ? k * n + o.i[q] : o.i[q] * n + k;
                const int ib = DB ? o.j[q] * n + k : k * n + o.j[q];
                const R ar = ar_[ia] * ck, ai = (DA ? -ai_[ia] : ai_[ia]) * ck;
                const R br = br_[ib], bi = DB ? -bi_[ib] : bi_[ib];
                cr[q] = cr[q] + (ar * br - ai * bi);
                ci[q] = ci[q] + (ar * bi + ai * br);
            }
        }
    } else {
        for (int k = 0; k < n; k++) {
#pragma unroll
            for (int q = 0; q < EPL; q++) {
                const int ia = DA ? k * n + o.i[q] : o.i[q] * n + k;
                const int ib = DB ? o.j[q] * n + k : k * n + o.j[q];
                const R ar = ar_[ia] * ck, ai = (DA ? -ai_[ia] : ai_[ia]) * ck;
                const R br = br_[ib], bi = DB ? -bi_[ib] : bi_[ib];
                cr[q] = fma(ar, br, cr[q]);
                cr[q] = fma(-ai, bi, cr[q]);
                ci[q] = fma(ar, bi, ci[q]);
                ci[q] = fma(ai, br, ci[q]);
            }
        }
    }
}

template <int EPL, int NT, typename R>
__device__ inline void st_regs(const Own<EPL, NT> &o, R *sre, R *sim, int n, const R (&cr)[EPL], const R (&ci)[EPL])
{
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q]) {
            sre[o.i[q] * n + o.j[q]] = cr[q];
            sim[o.i[q] * n + o.j[q]] = ci[q];
        }
}

// Compact Hermitian element (i, j) of momentum link l (tile 1).
template <typename R>
__device__ inline void mom_get(const R *__restrict__ P, long l, int n, int i, int j, R &re, R &im)
{
    const R *b = P + (size_t)l * n * n;
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

// X = 1 + c*rhs on the owned elements (Horner seed, sunmat.h:200-248)
template <bool Exact, int EPL, int NT, typename R>
__device__ inline void horner_seed(const Own<EPL, NT> &o, const R *rre, const R *rim, int n, R c, R (&xr)[EPL],
                                   R (&xi)[EPL])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        const int e = o.i[q] * n + o.j[q];
        xr[q] = (o.i[q] == o.j[q] ? R(1) : R(0)) + rre[e] * c;
        xi[q] = R(0) + rim[e] * c;
    }
}

// Drift of one link whose i*eps*P is in slot A: Uout = T_L(A) U_l.
// Uses both slots; on return slot contents are undefined.
template <bool Exact, int EPL, int NT, typename R>
__device__ inline void drift_core(const Own<EPL, NT> &o, R *Are, R *Aim, R *Bre, R *Bim, int n, int level,
                                  const double *cinv, const R *__restrict__ Uin, long l, R *__restrict__ Uout)
{
#pragma clang fp contract(off)
    if (level < 0 || level > 8) level = 8;
    R xr[EPL], xi[EPL];
    if (level == 0) {
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            xr[q] = o.i[q] == o.j[q] ? R(1) : R(0);
            xi[q] = R(0);
        }
    } else {
        horner_seed<Exact>(o, Are, Aim, n, level == 1 ? R(1) : (R)cinv[level], xr, xi);
        for (int k = level - 1; k >= 1; k--) {
            __syncthreads();
            st_regs(o, Bre, Bim, n, xr, xi);
            __syncthreads();
            R qr[EPL], qi[EPL];
            mul<false, false, false, Exact>(o, Are, Aim, Bre, Bim, n, (R)cinv[k], qr, qi);
#pragma unroll
            for (int q = 0; q < EPL; q++) {
                xr[q] = (o.i[q] == o.j[q] ? R(1) : R(0)) + qr[q];
                xi[q] = R(0) + qi[q];
            }
        }
    }
    __syncthreads();
    st_regs(o, Bre, Bim, n, xr, xi);
    ld_link<NT>(Are, Aim, Uin, l, n);
    __syncthreads();
    R ur[EPL], ui[EPL];
    mul<false, false, false, Exact>(o, Bre, Bim, Are, Aim, n, R(1), ur, ui);
    R *b = Uout + (size_t)l * 2 * n * n;
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q]) {
            const int e = o.i[q] * n + o.j[q];
            b[2 * e] = ur[q];
            b[2 * e + 1] = ui[q];
        }
}

template <int NMAX, int NT, typename R, bool Exact, bool Drift>
__global__ __launch_bounds__(NT) void k_kick(KickArgs a)
{
    using C = Cfg<NMAX, NT, R>;
    constexpr int EPL = C::EPL;
    __shared__ R Are[C::SLOT], Aim[C::SLOT], Bre[C::SLOT], Bim[C::SLOT];
    const Geom &g = a.g;
    const int n = g.N;
    const long V = g.V;
    // the D links of one site are consecutive logical blocks, and each XCD
    // takes one contiguous range of sites
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(lb % (unsigned)g.D);
    int cnt;
    const long s = parity_block((long)(lb / (unsigned)g.D), V, g.V0, 1, cnt);
    const long l = (long)mu * V + s;
    const R *U = (const R *)a.U;
    Own<EPL, NT> o;
    o.init(n);
    const long xp = g.nbr[(2 * mu) * V + s];

    R wr[EPL], wi[EPL], tr[EPL], ti[EPL];
#pragma unroll
    for (int q = 0; q < EPL; q++) wr[q] = wi[q] = R(0);
    // the staple operands in the order the products consume them; pa / pb
    // hold the next A / B operand (see Pref)
    constexpr int K = (2 * NMAX * NMAX + NT - 1) / NT;
    Pref<NT, K, R> pa, pb;
    int nu = mu == 0 ? 1 : 0;
    if (nu < g.D) {
        if constexpr (Exact) pa.fetch(U, mu * V + s, n);
        else pa.fetch(U, nu * V + xp, n);
        if constexpr (Exact) pb.fetch(U, nu * V + xp, n);
        else pb.fetch(U, mu * V + g.nbr[(2 * nu) * V + s], n);
    } else if constexpr (!Exact) {
        pb.fetch(U, mu * V + s, n);
    }
    while (nu < g.D) {
        int nx = nu + 1;
        if (nx == mu) nx++;
        const long xpn = g.nbr[(2 * nu) * V + s];
        const long xmn = g.nbr[(2 * nu + 1) * V + s];
        const long xpmn = g.nbr[(2 * nu + 1) * V + xp];
        if constexpr (Exact) {
            // ((U_mu(x) U_nu(x+mu)) U_mu(x+nu)^+) U_nu(x)^+
            __syncthreads();
            pa.put(Are, Aim, n);
            pb.put(Bre, Bim, n);
            __syncthreads();
            pb.fetch(U, mu * V + xpn, n);
            mul<false, false, false, true>(o, Are, Aim, Bre, Bim, n, R(1), tr, ti);
            __syncthreads();
            st_regs(o, Are, Aim, n, tr, ti);
            pb.put(Bre, Bim, n);
            __syncthreads();
            pb.fetch(U, nu * V + s, n);
            mul<false, true, false, true>(o, Are, Aim, Bre, Bim, n, R(1), tr, ti);
            __syncthreads();
            st_regs(o, Are, Aim, n, tr, ti);
            pb.put(Bre, Bim, n);
            __syncthreads();
            pa.fetch(U, mu * V + s, n);
            pb.fetch(U, nu * V + xpmn, n);
            mul<false, true, false, true>(o, Are, Aim, Bre, Bim, n, R(1), tr, ti);
            {
#pragma clang fp contract(off)
#pragma unroll
                for (int q = 0; q < EPL; q++) {
                    wr[q] = wr[q] + tr[q];
                    wi[q] = wi[q] + ti[q];
                }
            }
            // ((U_mu(x) U_nu(x+mu-nu)^+) U_mu(x-nu)^+) U_nu(x-nu)
            __syncthreads();
            pa.put(Are, Aim, n);
            pb.put(Bre, Bim, n);
            __syncthreads();
            pb.fetch(U, mu * V + xmn, n);
            mul<false, true, false, true>(o, Are, Aim, Bre, Bim, n, R(1), tr, ti);
            __syncthreads();
            st_regs(o, Are, Aim, n, tr, ti);
            pb.put(Bre, Bim, n);
            __syncthreads();
            pb.fetch(U, nu * V + xmn, n);
            mul<false, true, false, true>(o, Are, Aim, Bre, Bim, n, R(1), tr, ti);
            __syncthreads();
            st_regs(o, Are, Aim, n, tr, ti);
            pb.put(Bre, Bim, n);
            __syncthreads();
            if (nx < g.D) {
                pa.fetch(U, mu * V + s, n);
                pb.fetch(U, nx * V + xp, n);
            }
            mul<false, false, false, true>(o, Are, Aim, Bre, Bim, n, R(1), tr, ti);
            {
#pragma clang fp contract(off)
#pragma unroll
                for (int q = 0; q < EPL; q++) {
                    wr[q] = wr[q] + tr[q];
                    wi[q] = wi[q] + ti[q];
                }
            }
        } else {
            // S += U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
            __syncthreads();
            pa.put(Are, Aim, n);
            pb.put(Bre, Bim, n);
            __syncthreads();
            pb.fetch(U, nu * V + s, n);
            mul<false, true, false, false>(o, Are, Aim, Bre, Bim, n, R(1), tr, ti);
            __syncthreads();
            st_regs(o, Are, Aim, n, tr, ti);
            pb.put(Bre, Bim, n);
            __syncthreads();
            pa.fetch(U, nu * V + xpmn, n);
            pb.fetch(U, mu * V + xmn, n);
            mul<false, true, true, false>(o, Are, Aim, Bre, Bim, n, R(1), wr, wi);
            // S += U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
            __syncthreads();
            pa.put(Are, Aim, n);
            pb.put(Bre, Bim, n);
            __syncthreads();
            pb.fetch(U, nu * V + xmn, n);
            mul<true, true, false, false>(o, Are, Aim, Bre, Bim, n, R(1), tr, ti);
            __syncthreads();
            st_regs(o, Are, Aim, n, tr, ti);
            pb.put(Bre, Bim, n);
            __syncthreads();
            if (nx < g.D) {
                pa.fetch(U, nx * V + xp, n);
                pb.fetch(U, mu * V + g.nbr[(2 * nx) * V + s], n);
            } else {
                pb.fetch(U, mu * V + s, n);  // U_mu(x) for W = U_mu(x) S
            }
            mul<false, false, true, false>(o, Are, Aim, Bre, Bim, n, R(1), wr, wi);
        }
        nu = nx;
    }
    if constexpr (!Exact) {
        // W = U_mu(x) S
        __syncthreads();
        st_regs(o, Are, Aim, n, wr, wi);
        pb.put(Bre, Bim, n);
        __syncthreads();
        mul<false, false, false, false>(o, Bre, Bim, Are, Aim, n, R(1), tr, ti);
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            wr[q] = tr[q];
            wi[q] = ti[q];
        }
    }
    // TAH and kick on the owned elements (sunmat.h:355-371, sun.cpp:716)
    __syncthreads();
    st_regs(o, Are, Aim, n, wr, wi);
    __syncthreads();
    const R eps = (R)a.eps_kick, inv_n = (R)a.inv_n;
    R *P = (R *)a.P;
    double sum[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    R pr[EPL], pi[EPL], ppr[EPL], ppi[EPL];
    {
#pragma clang fp contract(off)
        R tri = R(0);
        for (int i = 0; i < n; i++) tri = tri + (Aim[i * n + i] + Aim[i * n + i]);
        const R corr = tri * inv_n;
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            const int i = o.i[q], j = o.j[q];
            R inr, ini;
            mom_get(P, l, n, i, j, inr, ini);
            ppr[q] = inr;
            ppi[q] = ini;
            if (i == j) {
                const R mi = (Aim[i * n + i] + Aim[i * n + i]) - corr;
                pr[q] = inr - (R(0.5) * mi) * eps;
                pi[q] = R(0);
            } else {
                const int lo = i < j ? i : j, hi = i < j ? j : i;
                const R dr = Are[lo * n + hi] - Are[hi * n + lo];
                const R di = Aim[lo * n + hi] + Aim[hi * n + lo];
                const R qr = (R(0.5) * di) * eps;
                const R qi = (R(0.5) * dr) * eps;
                // value of the (lo, hi) element; (hi, lo) is its conjugate
                R ur, ui;
                mom_get(P, l, n, lo, hi, ur, ui);
                const R vr = ur - qr, vi = ui + qi;
                pr[q] = vr;
                pi[q] = i < j ? vi : -vi;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q]) {
            if (o.i[q] == o.j[q]) sum[0] += (double)wr[q];
            sum[1] += (double)ppr[q] * (double)ppr[q] + (double)ppi[q] * (double)ppi[q];
            sum[2] += (double)pr[q] * (double)pr[q] + (double)pi[q] * (double)pi[q];
        }
    // store the new momentum (compact: owners of i <= j); every thread has
    // finished reading the old one
    __syncthreads();
    {
        R *b = P + (size_t)l * n * n;
        R *bp = (!Drift && a.Pprev) ? (R *)a.Pprev + (size_t)l * n * n : nullptr;
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            const int i = o.i[q], j = o.j[q];
            if (!o.ok[q] || i > j) continue;
            if (i == j) {
                b[i] = pr[q];
                if (bp) bp[i] = ppr[q];
            } else {
                const int h = hup(n, i, j);
                b[h] = pr[q];
                b[h + 1] = pi[q];
                if (bp) {
                    bp[h] = ppr[q];
                    bp[h + 1] = ppi[q];
                }
            }
        }
    }
    if constexpr (Drift) {
        // rhs = i * (eps_d * P) into slot A
        const R ed = (R)a.eps_drift;
        __syncthreads();
        {
#pragma clang fp contract(off)
#pragma unroll
            for (int q = 0; q < EPL; q++)
                if (o.ok[q]) {
                    const int e = o.i[q] * n + o.j[q];
                    Are[e] = -(pi[q] * ed);
                    Aim[e] = pr[q] * ed;
                }
        }
        __syncthreads();
        drift_core<Exact>(o, Are, Aim, Bre, Bim, n, a.level, a.cinv, U, l, (R *)a.Uout);
    }
    if (a.part) block_reduce_store<3>(sum, cmp, a.part + 6 * (size_t)lb);
}

template <int NMAX, int NT, typename R, bool Exact>
__global__ __launch_bounds__(NT) void k_drift(Geom g, R *U, const R *P, double eps, int level, Cinv9 cinv)
{
    using C = Cfg<NMAX, NT, R>;
    constexpr int EPL = C::EPL;
    __shared__ R Are[C::SLOT], Aim[C::SLOT], Bre[C::SLOT], Bim[C::SLOT];
    const int n = g.N;
    const long l = blockIdx.x;
    Own<EPL, NT> o;
    o.init(n);
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int q = 0; q < EPL; q++)
            if (o.ok[q]) {
                R re, im;
                mom_get(P, l, n, o.i[q], o.j[q], re, im);
                const int e = o.i[q] * n + o.j[q];
                Are[e] = -(im * (R)eps);
                Aim[e] = re * (R)eps;
            }
    }
    __syncthreads();
    drift_core<Exact>(o, Are, Aim, Bre, Bim, n, level, cinv.c, U, l, U);
}

// SU(N) projection of one link per workgroup (N >= 4), FP64 in LDS:
// classical Gram-Schmidt with one re-orthogonalisation pass on the rows,
// then the determinant of the orthonormal rows by LU with partial pivoting
// and multiplication by exp(-i arg(det)/N).  Each projection coefficient
// and each updated element is summed by a group of G = NT / NMAX adjacent
// lanes and reduced with shuffles; the pivot search is one wavefront's
// argmax.  The orthonormal rows are written to the link, the LU then
// overwrites them in LDS, and the phase is applied to the link in place
// (one LDS matrix, odd row stride).
// haar: the link is first filled with a Ginibre matrix from Philox (purpose
// 2, counter = stream, link index = dir * V + lexicographic site; normal 2e
// and 2e+1 are the real and imaginary parts of element e, scaled 1/sqrt 2).
template <int NMAX, int NT, typename R>
__global__ __launch_bounds__(NT) void k_project(Geom g, R *U, int haar, uint64_t seed, uint64_t stream)
{
    constexpr int LDA = NMAX + 1;  // odd row stride: rows on different banks
    constexpr int G = NT / NMAX;   // lanes per coefficient
    constexpr int NW = NT / 64;
    static_assert(G * NMAX == NT && (G & (G - 1)) == 0 && G <= 64 && NW * 64 == NT, "workgroup shape");
    __shared__ double ar[NMAX * LDA], ai[NMAX * LDA];
    __shared__ double dr[NMAX], di[NMAX], red[NW];
    __shared__ int piv;
    const int n = g.N, tid = threadIdx.x;
    const int grp = tid / G, gl = tid % G;  // row q / column k of this group, lane in the group
    const long l = blockIdx.x;
    R *b = U + (size_t)l * 2 * n * n;
    if (haar) {
        const int d = (int)(l / g.V);
        const uint32_t link = (uint32_t)(d * g.V + g.dev2lex[l - d * g.V]);
        for (int blk = tid; 4 * blk < 2 * n * n; blk += NT) {
            double z[4];
            normals4(philox_block(seed, stream, kPurposeHaar, link, (uint32_t)blk), z);
            for (int q = 0; q < 4 && 4 * blk + q < 2 * n * n; q++) {
                const int a = 4 * blk + q, e = a >> 1;
                double *dst = (a & 1) ? ai : ar;
                dst[(e / n) * LDA + e % n] = z[q] * 0.7071067811865476;
            }
        }
    } else {
        for (int e = tid; e < n * n; e += NT) {
            ar[(e / n) * LDA + e % n] = (double)b[2 * e];
            ai[(e / n) * LDA + e % n] = (double)b[2 * e + 1];
        }
    }
    __syncthreads();
    for (int r = 0; r < n; r++) {
        for (int pass = 0; pass < 2; pass++) {
            // d_q = <row q, row r> = sum_k conj(a_qk) a_rk for q < r
            {
                const int q = grp;
                double sr = 0.0, si = 0.0;
                if (q < r)
                    for (int k = gl; k < n; k += G) {
                        const double qr = ar[q * LDA + k], qi = ai[q * LDA + k];
                        const double xr = ar[r * LDA + k], xi = ai[r * LDA + k];
                        sr += qr * xr + qi * xi;
                        si += qr * xi - qi * xr;
                    }
#pragma unroll
                for (int off = G / 2; off > 0; off >>= 1) {
                    sr += __shfl_xor(sr, off, 64);
                    si += __shfl_xor(si, off, 64);
                }
                if (q < r && gl == 0) {
                    dr[q] = sr;
                    di[q] = si;
                }
            }
            __syncthreads();
            // a_rk -= sum_q d_q a_qk
            {
                const int k = grp;
                double xr = 0.0, xi = 0.0;
                if (k < n)
                    for (int q = gl; q < r; q += G) {
                        const double qr = ar[q * LDA + k], qi = ai[q * LDA + k];
                        xr += dr[q] * qr - di[q] * qi;
                        xi += dr[q] * qi + di[q] * qr;
                    }
#pragma unroll
                for (int off = G / 2; off > 0; off >>= 1) {
                    xr += __shfl_xor(xr, off, 64);
                    xi += __shfl_xor(xi, off, 64);
                }
                if (k < n && gl == 0) {
                    ar[r * LDA + k] -= xr;
                    ai[r * LDA + k] -= xi;
                }
            }
            __syncthreads();
        }
        // normalise row r
        double part = 0.0;
        for (int k = tid; k < n; k += NT) part += ar[r * LDA + k] * ar[r * LDA + k] + ai[r * LDA + k] * ai[r * LDA + k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        if ((tid & 63) == 0) red[tid >> 6] = part;
        __syncthreads();
        double nrm = 0.0;
#pragma unroll
        for (int w = 0; w < NW; w++) nrm += red[w];
        const double inv = 1.0 / sqrt(nrm);
        for (int k = tid; k < n; k += NT) {
            ar[r * LDA + k] *= inv;
            ai[r * LDA + k] *= inv;
        }
        __syncthreads();
    }
    // orthonormal rows -> link; determinant by LU in place
    for (int e = tid; e < n * n; e += NT) {
        b[2 * e] = (R)ar[(e / n) * LDA + e % n];
        b[2 * e + 1] = (R)ai[(e / n) * LDA + e % n];
    }
    double detr = 1.0, deti = 0.0;  // kept identical in every thread
    for (int c = 0; c < n; c++) {
        if (tid < 64) {
            // first row of largest |m_rc|^2, r >= c (n <= 64 rows, one per lane)
            const int rr = tid;
            double v = -1.0;
            if (rr >= c && rr < n) v = ar[rr * LDA + c] * ar[rr * LDA + c] + ai[rr * LDA + c] * ai[rr * LDA + c];
            int idx = rr;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double v2 = __shfl_xor(v, off, 64);
                const int i2 = __shfl_xor(idx, off, 64);
                if (v2 > v || (v2 == v && i2 < idx)) {
                    v = v2;
                    idx = i2;
                }
            }
            if (tid == 0) piv = idx;
        }
        __syncthreads();
        const int p = piv;
        if (p != c) {
            for (int k = tid; k < n; k += NT) {
                double t = ar[c * LDA + k];
                ar[c * LDA + k] = ar[p * LDA + k];
                ar[p * LDA + k] = t;
                t = ai[c * LDA + k];
                ai[c * LDA + k] = ai[p * LDA + k];
                ai[p * LDA + k] = t;
            }
            detr = -detr;
            deti = -deti;
        }
        __syncthreads();
        const double pr = ar[c * LDA + c], pim = ai[c * LDA + c];
        const double nr = detr * pr - deti * pim, ni = detr * pim + deti * pr;
        detr = nr;
        deti = ni;
        const double den = pr * pr + pim * pim;
        if (den == 0.0) break;
        // eliminate below the pivot: element (r, k), r > c, k > c
        const int m = n - c - 1;
        for (int e = tid; e < m * m; e += NT) {
            const int r = c + 1 + e / m, k = c + 1 + e % m;
            const double xr = ar[r * LDA + c], xi = ai[r * LDA + c];
            const double fr = (xr * pr + xi * pim) / den, fi = (xi * pr - xr * pim) / den;
            const double br = ar[c * LDA + k], bi = ai[c * LDA + k];
            ar[r * LDA + k] -= fr * br - fi * bi;
            ai[r * LDA + k] -= fr * bi + fi * br;
        }
        __syncthreads();
    }
    const double ph = -atan2(deti, detr) / n;
    const double cr = cos(ph), ci = sin(ph);
    for (int e = tid; e < n * n; e += NT) {
        const double qr = (double)b[2 * e], qi = (double)b[2 * e + 1];
        b[2 * e] = (R)(qr * cr - qi * ci);
        b[2 * e + 1] = (R)(qr * ci + qi * cr);
    }
}

// global link (matrix contiguous, stored as LS) -> LDS slot in FP64
template <int NT, typename LS>
__device__ inline void ld_link_f64(double *sre, double *sim, const LS *__restrict__ f, long l, int n)
{
    const int nr = 2 * n * n;
    const LS *b = f + (size_t)l * nr;
    for (int r = threadIdx.x; r < nr; r += NT) {
        const double v = (double)b[r];
        if (r & 1) sim[r >> 1] = v;
        else sre[r >> 1] = v;
    }
}

// Plaquette sums of one site per workgroup, FP64 arithmetic whatever the
// storage (EnergyMag sun.cpp:507-537, split as PlaqExp
// measurements4d.cpp:885-927): 1 - ReTr P_jk(x) / N over the temporal planes
// (0, k) and, separately, the spatial planes (j, k), j >= 1, with
// P_jk(x) = U_j(x) U_k(x+j) U_j(x+k)^+ U_k(x)^+.  Each thread adds
// 1/N - P_ii/N for the diagonal elements it owns; the workgroup sum of
// those is 1 - ReTr P / N.
template <int NMAX, int NT, typename LS>
__global__ __launch_bounds__(NT) void k_plaq(Geom g, const LS *U, double inv_n, double *part)
{
    using C = Cfg<NMAX, NT, double>;
    constexpr int EPL = C::EPL;
    __shared__ double Are[C::SLOT], Aim[C::SLOT], Bre[C::SLOT], Bim[C::SLOT];
    const int n = g.N;
    const long V = g.V;
    const long x = blockIdx.x;
    Own<EPL, NT> o;
    o.init(n);
    double s[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
    for (int j = 0; j < g.D; j++) {
        const long xj = g.nbr[(2 * j) * V + x];
        for (int k = j + 1; k < g.D; k++) {
            const long xk = g.nbr[(2 * k) * V + x];
            double tr[EPL], ti[EPL];
            __syncthreads();
            ld_link_f64<NT>(Are, Aim, U, j * V + x, n);
            ld_link_f64<NT>(Bre, Bim, U, k * V + xj, n);
            __syncthreads();
            mul<false, false, false, false>(o, Are, Aim, Bre, Bim, n, 1.0, tr, ti);
            __syncthreads();
            st_regs(o, Are, Aim, n, tr, ti);
            ld_link_f64<NT>(Bre, Bim, U, j * V + xk, n);
            __syncthreads();
            mul<false, true, false, false>(o, Are, Aim, Bre, Bim, n, 1.0, tr, ti);
            __syncthreads();
            st_regs(o, Are, Aim, n, tr, ti);
            ld_link_f64<NT>(Bre, Bim, U, k * V + x, n);
            __syncthreads();
            mul<false, true, false, false>(o, Are, Aim, Bre, Bim, n, 1.0, tr, ti);
#pragma unroll
            for (int q = 0; q < EPL; q++)
                if (o.ok[q] && o.i[q] == o.j[q]) kahan_add(s[j == 0 ? 0 : 1], c[j == 0 ? 0 : 1], inv_n - tr[q] * inv_n);
        }
    }
    block_reduce_store<2>(s, c, part + 4 * (size_t)blockIdx.x);
}

// Straight-line products W = U_dir(x) U_dir(x + dir) ... U_dir(x + (L-1) dir)
// (WilsonLoop sun.cpp:307-345, PolyakovLoop measurements4d.cpp:410-466), one
// workgroup per line, FP64 in the reference's product order (GSL zgemm: k
// ascending, no contraction).  Line q = blockIdx.x starts at the site with
// x_dir = 0 and the other coordinates of q in lexicographic order, or, with
// lex1 >= 0, at lexicographic site lex1.  traces (may be null): Tr W / N as
// (re, im) at 2q; wfull (may be null): W, row-major complex.
template <int NMAX, int NT, typename LS>
__global__ __launch_bounds__(NT) void k_line(Geom g, const LS *U, int dir, long lex1, double *traces, double *wfull)
{
    using C = Cfg<NMAX, NT, double>;
    constexpr int EPL = C::EPL;
    __shared__ double Are[C::SLOT], Aim[C::SLOT], Bre[C::SLOT], Bim[C::SLOT];
    const int n = g.N;
    const long V = g.V, q = blockIdx.x;
    long lex = lex1;
    if (lex < 0) {
        long rem = q;
        lex = 0;
        for (int d = g.D - 1; d >= 0; d--) {
            if (d == dir) continue;
            lex += (rem % g.ext[d]) * g.helparr[d];
            rem /= g.ext[d];
        }
    }
    long s = g.lex2dev[lex];
    Own<EPL, NT> o;
    o.init(n);
    ld_link_f64<NT>(Are, Aim, U, dir * V + s, n);
    double wr[EPL], wi[EPL];
    for (int k = 1; k < g.ext[dir]; k++) {
        s = g.nbr[(2 * dir) * V + s];
        __syncthreads();
        ld_link_f64<NT>(Bre, Bim, U, dir * V + s, n);
        __syncthreads();
        mul<false, false, false, true>(o, Are, Aim, Bre, Bim, n, 1.0, wr, wi);
        __syncthreads();
        st_regs(o, Are, Aim, n, wr, wi);
    }
    __syncthreads();
    if (wfull)
        for (int e = threadIdx.x; e < n * n; e += NT) {
            wfull[2 * e] = Are[e];
            wfull[2 * e + 1] = Aim[e];
        }
    if (traces && threadIdx.x == 0) {
        double re = 0.0, im = 0.0;
        for (int i = 0; i < n; i++) {
            re = re + Are[i * n + i];
            im = im + Aim[i * n + i];
        }
        traces[2 * q] = re / n;
        traces[2 * q + 1] = im / n;
    }
}

template <int NMAX, int NT>
hipError_t lines(const Geom &g, const void *U, int dir, long nlines, long lex1, int mode, double *traces,
                 double *wfull, hipStream_t st)
{
    if (f32_mode(mode))
        hipLaunchKernelGGL((k_line<NMAX, NT, float>), dim3((unsigned)nlines), dim3(NT), 0, st, g, (const float *)U, dir,
                           lex1, traces, wfull);
    else
        hipLaunchKernelGGL((k_line<NMAX, NT, double>), dim3((unsigned)nlines), dim3(NT), 0, st, g, (const double *)U,
                           dir, lex1, traces, wfull);
    return hipGetLastError();
}

// UViol (sun.cpp:441-471): sum over links of tr2(U U^+ - 1) = ReTr(T T) with
// T = U U^+ - 1, products in the reference's order (FP64).  Workgroups take
// the links l = blockIdx.x, blockIdx.x + gridDim.x, ...; each thread adds
// the diagonal elements of T T it owns; one (s, c) pair per workgroup.
template <int NMAX, int NT, typename LS>
__global__ __launch_bounds__(NT) void k_uviol(Geom g, const LS *U, double *part)
{
    using C = Cfg<NMAX, NT, double>;
    constexpr int EPL = C::EPL;
    __shared__ double Are[C::SLOT], Aim[C::SLOT], Bre[C::SLOT], Bim[C::SLOT];
    const int n = g.N;
    Own<EPL, NT> o;
    o.init(n);
    double s[1] = {0.0}, c[1] = {0.0};
    for (long l = blockIdx.x; l < g.V * g.D; l += gridDim.x) {
        double tr[EPL], ti[EPL];
        __syncthreads();
        ld_link_f64<NT>(Are, Aim, U, l, n);
        __syncthreads();
        mul<false, true, false, true>(o, Are, Aim, Are, Aim, n, 1.0, tr, ti);
#pragma unroll
        for (int q = 0; q < EPL; q++)
            if (o.i[q] == o.j[q]) tr[q] -= 1.0;
        st_regs(o, Bre, Bim, n, tr, ti);
        __syncthreads();
        mul<false, false, false, true>(o, Bre, Bim, Bre, Bim, n, 1.0, tr, ti);
#pragma unroll
        for (int q = 0; q < EPL; q++)
            if (o.ok[q] && o.i[q] == o.j[q]) kahan_add(s[0], c[0], tr[q]);
    }
    block_reduce_store<1>(s, c, part + 2 * (size_t)blockIdx.x);
}

// GaussLaw (sun.cpp:415-438): the matrix sum over links of P_l - U_l^+ P_l U_l
// (P_l the Hermitian momentum of link l; summed over all sites, the
// reference's U^+(x-mu) P(x-mu) U(x-mu) runs over every link once), products
// in the reference's order (FP64).  Workgroups take the links
// l = blockIdx.x, blockIdx.x + gridDim.x, ...; each thread keeps compensated
// sums of the elements it owns and writes them to row blockIdx.x of part (2N^2
// (s, c) pairs: element (i, j) at 2 (i N + j), re then im), for a column
// reduction over the rows.
template <int NMAX, int NT, typename LS>
__global__ __launch_bounds__(NT) void k_gauss(Geom g, const LS *U, const LS *P, double *part)
{
    using C = Cfg<NMAX, NT, double>;
    constexpr int EPL = C::EPL;
    __shared__ double Are[C::SLOT], Aim[C::SLOT], Bre[C::SLOT], Bim[C::SLOT];
    const int n = g.N;
    Own<EPL, NT> o;
    o.init(n);
    double sr[EPL], cr[EPL], si[EPL], ci[EPL];
#pragma unroll
    for (int q = 0; q < EPL; q++) sr[q] = cr[q] = si[q] = ci[q] = 0.0;
    for (long l = blockIdx.x; l < g.V * g.D; l += gridDim.x) {
        double pr[EPL], pi[EPL], tr[EPL], ti[EPL];
        __syncthreads();
        ld_link_f64<NT>(Are, Aim, U, l, n);
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            pr[q] = pi[q] = 0.0;
            if (!o.ok[q]) continue;
            LS re, im;
            mom_get(P, l, n, o.i[q], o.j[q], re, im);
            pr[q] = (double)re;
            pi[q] = (double)im;
        }
        st_regs(o, Bre, Bim, n, pr, pi);
        __syncthreads();
        mul<true, false, false, true>(o, Are, Aim, Bre, Bim, n, 1.0, tr, ti);  // U^+ P
        __syncthreads();
        st_regs(o, Bre, Bim, n, tr, ti);  // P itself stays in pr, pi
        __syncthreads();
        mul<false, false, false, true>(o, Bre, Bim, Are, Aim, n, 1.0, tr, ti);  // (U^+ P) U
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            kahan_add(sr[q], cr[q], pr[q] - tr[q]);
            kahan_add(si[q], ci[q], pi[q] - ti[q]);
        }
    }
    const int nr = 2 * n * n;
    double *row = part + (size_t)blockIdx.x * 2 * nr;
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q]) {
            const int e = 2 * (o.i[q] * n + o.j[q]);
            row[2 * e] = sr[q];
            row[2 * e + 1] = cr[q];
            row[2 * (e + 1)] = si[q];
            row[2 * (e + 1) + 1] = ci[q];
        }
}

template <int NMAX, int NT>
hipError_t uviol(const Geom &g, const void *U, int mode, long nwg, double *part, hipStream_t st)
{
    if (f32_mode(mode))
        hipLaunchKernelGGL((k_uviol<NMAX, NT, float>), dim3((unsigned)nwg), dim3(NT), 0, st, g, (const float *)U, part);
    else hipLaunchKernelGGL((k_uviol<NMAX, NT, double>), dim3((unsigned)nwg), dim3(NT), 0, st, g, (const double *)U, part);
    return hipGetLastError();
}

template <int NMAX, int NT>
hipError_t gauss(const Geom &g, const void *U, const void *P, int mode, long nwg, double *part, hipStream_t st)
{
    if (f32_mode(mode))
        hipLaunchKernelGGL((k_gauss<NMAX, NT, float>), dim3((unsigned)nwg), dim3(NT), 0, st, g, (const float *)U,
                           (const float *)P, part);
    else
        hipLaunchKernelGGL((k_gauss<NMAX, NT, double>), dim3((unsigned)nwg), dim3(NT), 0, st, g, (const double *)U,
                           (const double *)P, part);
    return hipGetLastError();
}

template <int NMAX, int NT>
hipError_t plaq(const Geom &g, const void *U, int mode, double inv_n, double *part, hipStream_t st)
{
    const unsigned nb = (unsigned)g.V;
    if (f32_mode(mode))
        hipLaunchKernelGGL((k_plaq<NMAX, NT, float>), dim3(nb), dim3(NT), 0, st, g, (const float *)U, inv_n, part);
    else
        hipLaunchKernelGGL((k_plaq<NMAX, NT, double>), dim3(nb), dim3(NT), 0, st, g, (const double *)U, inv_n, part);
    return hipGetLastError();
}

template <int NMAX, int NT>
hipError_t project(const Geom &g, void *U, int mode, int haar, uint64_t seed, uint64_t stream, hipStream_t st)
{
    const unsigned nb = (unsigned)(g.V * g.D);
    if (f32_mode(mode))
        hipLaunchKernelGGL((k_project<NMAX, NT, float>), dim3(nb), dim3(NT), 0, st, g, (float *)U, haar, seed, stream);
    else
        hipLaunchKernelGGL((k_project<NMAX, NT, double>), dim3(nb), dim3(NT), 0, st, g, (double *)U, haar, seed, stream);
    return hipGetLastError();
}

template <int NMAX, int NT>
hipError_t kick(const KickArgs &a, int mode, bool drift, hipStream_t st)
{
    const unsigned nb = (unsigned)(a.g.V * a.g.D);
    switch (mode) {
    case kModeF64:
        if (drift) hipLaunchKernelGGL((k_kick<NMAX, NT, double, false, true>), dim3(nb), dim3(NT), 0, st, a);
        else hipLaunchKernelGGL((k_kick<NMAX, NT, double, false, false>), dim3(nb), dim3(NT), 0, st, a);
        break;
    case kModeF64Exact:
        if (drift) hipLaunchKernelGGL((k_kick<NMAX, NT, double, true, true>), dim3(nb), dim3(NT), 0, st, a);
        else hipLaunchKernelGGL((k_kick<NMAX, NT, double, true, false>), dim3(nb), dim3(NT), 0, st, a);
        break;
    default:
        if (drift) hipLaunchKernelGGL((k_kick<NMAX, NT, float, false, true>), dim3(nb), dim3(NT), 0, st, a);
        else hipLaunchKernelGGL((k_kick<NMAX, NT, float, false, false>), dim3(nb), dim3(NT), 0, st, a);
    }
    return hipGetLastError();
}

template <int NMAX, int NT>
hipError_t drift(const Geom &g, void *U, const void *P, double eps, int level, const double *c, int mode,
                 hipStream_t st)
{
    const unsigned nb = (unsigned)(g.V * g.D);
    Cinv9 ci;
    for (int k = 0; k < 9; k++) ci.c[k] = c[k];
    switch (mode) {
    case kModeF64:
        hipLaunchKernelGGL((k_drift<NMAX, NT, double, false>), dim3(nb), dim3(NT), 0, st, g, (double *)U,
                           (const double *)P, eps, level, ci);
        break;
    case kModeF64Exact:
        hipLaunchKernelGGL((k_drift<NMAX, NT, double, true>), dim3(nb), dim3(NT), 0, st, g, (double *)U,
                           (const double *)P, eps, level, ci);
        break;
    default:
        hipLaunchKernelGGL((k_drift<NMAX, NT, float, false>), dim3(nb), dim3(NT), 0, st, g, (float *)U,
                           (const float *)P, eps, level, ci);
    }
    return hipGetLastError();
}

}  // namespace team
}  // namespace askja
