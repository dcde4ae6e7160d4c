// small.hip -- MD, action and diagnostic kernels for N <= 3: one lane per
// link (or site, or line), matrices held in registers, fields in the
// 64-link tiled structure-of-arrays layout so that a wavefront reading one
// matrix element of 64 consecutive links touches one contiguous segment.
#include "kernels.hpp"
#include "su_small.hpp"

namespace askja {

namespace {

constexpr int kBlock = 256;

template <typename R>
struct Cinv {
    R c[9];
};

template <typename R>
__device__ inline Cinv<R> cinv_of(const double *c)
{
    Cinv<R> r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.c[k] = (R)c[k];
    return r;
}

// W = U_mu(x) * sum_{nu != mu} [up + down staple]  (updateE, sun.cpp:694-716).
// Exact: every staple as ((U_mu(x) U_nu(x+mu)) U_mu(x+nu)^+) U_nu(x)^+ and
// ((U_mu(x) U_nu(x+mu-nu)^+) U_mu(x-nu)^+) U_nu(x-nu), nu ascending, up then
// down, summed from zero, as the reference does.  Fast: the sum of the
// 2(D-1) three-link staples first, then one product with U_mu(x).
template <int N, typename R, bool Exact>
__device__ inline void staple_w(const Geom &g, const R *__restrict__ U, int mu, long s, CM<N, R> &W)
{
    const long V = g.V;
    const long xp = g.nbr[(2 * mu) * V + s];
    if constexpr (Exact) {
        CM<N, R> Umu, B, T1, T2;
        load_link<N, R>(U, mu * V + s, Umu);
        cm_zero(W);
        for (int nu = 0; nu < g.D; nu++) {
            if (nu == mu) continue;
            const long xpn = g.nbr[(2 * nu) * V + s];
            const long xmn = g.nbr[(2 * nu + 1) * V + s];
            const long xpmn = g.nbr[(2 * nu + 1) * V + xp];
            load_link<N, R>(U, nu * V + xp, B);
            mul_exact<false, false>(Umu, B, T1);
            load_link<N, R>(U, mu * V + xpn, B);
            mul_exact<false, true>(T1, B, T2);
            load_link<N, R>(U, nu * V + s, B);
            mul_exact<false, true>(T2, B, T1);
            cm_add_exact(W, T1);
            load_link<N, R>(U, nu * V + xpmn, B);
            mul_exact<false, true>(Umu, B, T1);
            load_link<N, R>(U, mu * V + xmn, B);
            mul_exact<false, true>(T1, B, T2);
            load_link<N, R>(U, nu * V + xmn, B);
            mul_exact<false, false>(T2, B, T1);
            cm_add_exact(W, T1);
        }
    } else {
        CM<N, R> S, A, B, T;
        cm_zero(S);
        for (int nu = 0; nu < g.D; nu++) {
            if (nu == mu) continue;
            const long xpn = g.nbr[(2 * nu) * V + s];
            const long xmn = g.nbr[(2 * nu + 1) * V + s];
            const long xpmn = g.nbr[(2 * nu + 1) * V + xp];
            load_link<N, R>(U, nu * V + xp, A);
            load_link<N, R>(U, mu * V + xpn, B);
            mul_fast<false, true, false>(A, B, T);
            load_link<N, R>(U, nu * V + s, B);
            mul_fast<false, true, true>(T, B, S);
            load_link<N, R>(U, nu * V + xpmn, A);
            load_link<N, R>(U, mu * V + xmn, B);
            mul_fast<true, true, false>(A, B, T);
            load_link<N, R>(U, nu * V + xmn, B);
            mul_fast<false, false, true>(T, B, S);
        }
        load_link<N, R>(U, mu * V + s, A);
        mul_fast<false, false, false>(A, S, W);
    }
}

// One kick of every link, optionally fused with the drift of the same link:
//   P <- P - eps_k TAH(W)                       (updateE, sun.cpp:672-723)
//   Uout <- T_L(i eps_d P) U                    (updateUandcopyDown, 774-790)
// The fused drift writes a second link buffer, so every kick reads the links
// of one time slice only.  Per block: compensated sums of ReTr W, Tr Pin^2,
// Tr Pout^2.
template <int N, typename R, bool Exact, bool Drift>
__global__ __launch_bounds__(kBlock) void k_kick(KickArgs a)
{
    const Geom &g = a.g;
    // blocks of the D directions of one run of sites are consecutive, and
    // each XCD receives one contiguous range of them
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(lb % (unsigned)g.D);
    int cnt;
    const long s = parity_block((long)(lb / (unsigned)g.D), g.V, g.V0, kBlock, cnt) + threadIdx.x;
    double sum[3] = {0.0, 0.0, 0.0}, cmp[3] = {0.0, 0.0, 0.0};
    if ((int)threadIdx.x < cnt) {
        const R *U = (const R *)a.U;
        R *P = (R *)a.P;
        const long l = mu * g.V + s;
        CM<N, R> W, Pin, Pout;
        staple_w<N, R, Exact>(g, U, mu, s, W);
        load_mom<N, R>(P, l, Pin);
        tah_kick<Exact>(W, Pin, (R)a.eps_kick, (R)a.inv_n, Pout);
        store_mom<N, R>(P, l, Pout);
        if constexpr (!Drift) {
            if (a.Pprev) store_mom<N, R>((R *)a.Pprev, l, Pin);
        } else {
            const Cinv<R> ci = cinv_of<R>(a.cinv);
            CM<N, R> Ul, Un;
            load_link<N, R>(U, l, Ul);
            if constexpr (Exact) taylor_drift<true>(Pout, (R)a.eps_drift, a.level, ci.c, Ul, Un);
            else taylor_drift_fast(Pout, (R)a.eps_drift, a.level, ci.c, Ul, Un);
            store_link<N, R>((R *)a.Uout, l, Un);
        }
        sum[0] = (double)retr(W);
        sum[1] = herm_norm2(Pin);
        sum[2] = herm_norm2(Pout);
    }
    if (a.part) block_reduce_store<3>(sum, cmp, a.part + 6 * (size_t)lb);
}

// Standalone drift, in place: U <- T_L(i eps P) U.
template <int N, typename R, bool Exact>
__global__ __launch_bounds__(kBlock) void k_drift(Geom g, R *U, const R *P, double eps, int level, Cinv<R> ci)
{
    const long l = (long)blockIdx.x * kBlock + threadIdx.x;
    if (l >= g.V * g.D) return;
    CM<N, R> Pl, Ul, Un;
    load_mom<N, R>(P, l, Pl);
    load_link<N, R>(U, l, Ul);
    if constexpr (Exact) taylor_drift<true>(Pl, (R)eps, level, ci.c, Ul, Un);
    else taylor_drift_fast(Pl, (R)eps, level, ci.c, Ul, Un);
    store_link<N, R>(U, l, Un);
}

// Plaquette sums per site: temporal planes (0, k) and spatial planes (j, k),
// 1 <= j < k, of 1 - ReTr P_jk(x)/N (EnergyMag sun.cpp:507-537, PlaqExp
// measurements4d.cpp:885-927).
template <int N, typename R, bool Exact, typename C = R>
__global__ __launch_bounds__(kBlock) void k_plaq(Geom g, const R *U, double inv_n, double *part)
{
    const long s = (long)blockIdx.x * kBlock + threadIdx.x;
    double sum[2] = {0.0, 0.0}, cmp[2] = {0.0, 0.0};
    if (s < g.V) {
        const long V = g.V;
        for (int j = 0; j < g.D; j++) {
            const long xj = g.nbr[(2 * j) * V + s];
            for (int k = j + 1; k < g.D; k++) {
                const long xk = g.nbr[(2 * k) * V + s];
                double tr;
                CM<N, C> A, B, T1, T2;
                if constexpr (Exact) {
                    load_link<N, R>(U, j * V + s, A);
                    load_link<N, R>(U, k * V + xj, B);
                    mul_exact<false, false>(A, B, T1);
                    load_link<N, R>(U, j * V + xk, B);
                    mul_exact<false, true>(T1, B, T2);
                    load_link<N, R>(U, k * V + s, B);
                    mul_exact<false, true>(T2, B, T1);
                    tr = (double)retr(T1);
                } else {
                    load_link<N, R>(U, j * V + s, A);
                    load_link<N, R>(U, k * V + xj, B);
                    mul_fast<false, false, false>(A, B, T1);
                    load_link<N, R>(U, k * V + s, A);
                    load_link<N, R>(U, j * V + xk, B);
                    mul_fast<false, false, false>(A, B, T2);
                    tr = (double)retr_abh(T1, T2);
                }
                kahan_add(sum[j == 0 ? 0 : 1], cmp[j == 0 ? 0 : 1], 1.0 - tr * inv_n);
            }
        }
    }
    block_reduce_store<2>(sum, cmp, part + 4 * (size_t)blockIdx.x);
}

// UViol (sun.cpp:441-471): tr2(U U^+ - 1) per link.
template <int N, typename R>
__global__ __launch_bounds__(kBlock) void k_uviol(Geom g, const R *U, double *part)
{
    const long l = (long)blockIdx.x * kBlock + threadIdx.x;
    double sum[1] = {0.0}, cmp[1] = {0.0};
    if (l < g.V * g.D) {
        CM<N, R> A, T;
        load_link<N, R>(U, l, A);
        mul_exact<false, true>(A, A, T);
#pragma unroll
        for (int i = 0; i < N; i++) T.re[i][i] -= R(1);
        CM<N, R> Q;
        mul_exact<false, false>(T, T, Q);
        sum[0] = (double)retr(Q);
    }
    block_reduce_store<1>(sum, cmp, part + 2 * (size_t)blockIdx.x);
}

// GaussLaw summand per link: P - U^+ P U (sun.cpp:415-438, 345-352).
template <int N, typename R>
__global__ __launch_bounds__(kBlock) void k_gauss(Geom g, const R *U, const R *P, double *part)
{
    const long l = (long)blockIdx.x * kBlock + threadIdx.x;
    double sum[2 * N * N], cmp[2 * N * N];
#pragma unroll
    for (int q = 0; q < 2 * N * N; q++) sum[q] = cmp[q] = 0.0;
    if (l < g.V * g.D) {
        CM<N, R> A, Pl, T, Q;
        load_link<N, R>(U, l, A);
        load_mom<N, R>(P, l, Pl);
        mul_exact<true, false>(A, Pl, T);
        mul_exact<false, false>(T, A, Q);
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) {
                sum[2 * (i * N + j)] = (double)Pl.re[i][j] - (double)Q.re[i][j];
                sum[2 * (i * N + j) + 1] = (double)Pl.im[i][j] - (double)Q.im[i][j];
            }
    }
    block_reduce_store<2 * N * N>(sum, cmp, part + 4 * N * N * (size_t)blockIdx.x);
}

// Straight Wilson lines: line q starts at the site whose coordinates other
// than dir enumerate q lexicographically and x_dir = 0; product of ext[dir]
// links in direction dir (WilsonLoop sun.cpp:307-345, measurements4d.cpp
// 311-355, 410-466).
template <int N, typename R>
__global__ __launch_bounds__(kBlock) void k_wlines(Geom g, const R *U, int dir, long nlines, double *out)
{
    const long q = (long)blockIdx.x * kBlock + threadIdx.x;
    if (q >= nlines) return;
    long rem = q, lex = 0;
    for (int d = g.D - 1; d >= 0; d--) {
        if (d == dir) continue;
        const long x = rem % g.ext[d];
        rem /= g.ext[d];
        lex += x * g.helparr[d];
    }
    long s = g.lex2dev[lex];
    CM<N, R> W, B, T;
    load_link<N, R>(U, dir * g.V + s, W);
    for (int t = 1; t < g.ext[dir]; t++) {
        s = g.nbr[(2 * dir) * g.V + s];
        load_link<N, R>(U, dir * g.V + s, B);
        mul_exact<false, false>(W, B, T);
        W = T;
    }
    out[2 * q] = (double)retr(W) / N;
    out[2 * q + 1] = (double)imtr(W) / N;
}

// SU(N) projection of one link held in registers (N <= 3): modified
// Gram-Schmidt on the rows, then exp(-i arg(det)/N); FP64 working copy.
// haar: the link is first replaced by a Ginibre matrix from Philox (purpose
// 2, counter = stream, link index = dir * V + lexicographic site).
template <int N, typename R>
__global__ __launch_bounds__(kBlock) void k_project(Geom g, R *U, int haar, uint64_t seed, uint64_t stream)
{
#pragma clang fp contract(off)
    const long l = (long)blockIdx.x * kBlock + threadIdx.x;
    if (l >= g.V * g.D) return;
    double ar[N][N], ai[N][N];
    if (haar) {
        const int d = (int)(l / g.V);
        const uint32_t link = (uint32_t)(d * g.V + g.dev2lex[l - d * g.V]);
        double z[4 * ((2 * N * N + 3) / 4)];
#pragma unroll
        for (int b = 0; b < (2 * N * N + 3) / 4; b++) normals4(philox_block(seed, stream, kPurposeHaar, link, (uint32_t)b), z + 4 * b);
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) {
                ar[i][j] = z[2 * (i * N + j)] * 0.7071067811865476;
                ai[i][j] = z[2 * (i * N + j) + 1] * 0.7071067811865476;
            }
    } else {
        CM<N, R> m;
        load_link<N, R>(U, l, m);
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) {
                ar[i][j] = (double)m.re[i][j];
                ai[i][j] = (double)m.im[i][j];
            }
    }
#pragma unroll
    for (int r = 0; r < N; r++) {
#pragma unroll
        for (int q = 0; q < r; q++) {
            double dr = 0.0, di = 0.0;
#pragma unroll
            for (int k = 0; k < N; k++) {
                dr += ar[q][k] * ar[r][k] + ai[q][k] * ai[r][k];
                di += ar[q][k] * ai[r][k] - ai[q][k] * ar[r][k];
            }
#pragma unroll
            for (int k = 0; k < N; k++) {
                ar[r][k] -= dr * ar[q][k] - di * ai[q][k];
                ai[r][k] -= dr * ai[q][k] + di * ar[q][k];
            }
        }
        double nrm = 0.0;
#pragma unroll
        for (int k = 0; k < N; k++) nrm += ar[r][k] * ar[r][k] + ai[r][k] * ai[r][k];
        nrm = 1.0 / sqrt(nrm);
#pragma unroll
        for (int k = 0; k < N; k++) {
            ar[r][k] *= nrm;
            ai[r][k] *= nrm;
        }
    }
    // determinant by cofactors
    double dr, di;
    if constexpr (N == 1) {
        dr = ar[0][0];
        di = ai[0][0];
    } else if constexpr (N == 2) {
        dr = (ar[0][0] * ar[1][1] - ai[0][0] * ai[1][1]) - (ar[0][1] * ar[1][0] - ai[0][1] * ai[1][0]);
        di = (ar[0][0] * ai[1][1] + ai[0][0] * ar[1][1]) - (ar[0][1] * ai[1][0] + ai[0][1] * ar[1][0]);
    } else {
        // row 0 times the cofactors of rows 1, 2
        dr = 0.0;
        di = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int a = (j + 1) % 3, b = (j + 2) % 3;
            const double cr = (ar[1][a] * ar[2][b] - ai[1][a] * ai[2][b]) - (ar[1][b] * ar[2][a] - ai[1][b] * ai[2][a]);
            const double ci = (ar[1][a] * ai[2][b] + ai[1][a] * ar[2][b]) - (ar[1][b] * ai[2][a] + ai[1][b] * ar[2][a]);
            dr += ar[0][j] * cr - ai[0][j] * ci;
            di += ar[0][j] * ci + ai[0][j] * cr;
        }
    }
    const double ph = -atan2(di, dr) / N;
    const double c = cos(ph), sn = sin(ph);
    CM<N, R> out;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            out.re[i][j] = (R)(ar[i][j] * c - ai[i][j] * sn);
            out.im[i][j] = (R)(ar[i][j] * sn + ai[i][j] * c);
        }
    store_link<N, R>(U, l, out);
}

template <int N>
hipError_t project_mode(const Geom &g, void *U, int mode, int haar, uint64_t seed, uint64_t stream, hipStream_t st)
{
    const unsigned nb = (unsigned)((g.V * g.D + kBlock - 1) / kBlock);
    if (f32_mode(mode))
        hipLaunchKernelGGL((k_project<N, float>), dim3(nb), dim3(kBlock), 0, st, g, (float *)U, haar, seed, stream);
    else hipLaunchKernelGGL((k_project<N, double>), dim3(nb), dim3(kBlock), 0, st, g, (double *)U, haar, seed, stream);
    return hipGetLastError();
}

template <int N, typename R, bool Exact>
hipError_t kick_n(const KickArgs &a, bool drift, hipStream_t st)
{
    dim3 grid((unsigned)kick_site_blocks(a.g) * (unsigned)a.g.D);
    if (drift) hipLaunchKernelGGL((k_kick<N, R, Exact, true>), grid, dim3(kBlock), 0, st, a);
    else hipLaunchKernelGGL((k_kick<N, R, Exact, false>), grid, dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int N>
hipError_t kick_mode(const KickArgs &a, int mode, bool drift, hipStream_t st)
{
    switch (mode) {
    case kModeF64: return kick_n<N, double, false>(a, drift, st);
    case kModeF64Exact: return kick_n<N, double, true>(a, drift, st);
    default: return kick_n<N, float, false>(a, drift, st);
    }
}

template <int N, typename R, bool Exact>
hipError_t drift_n(const Geom &g, void *U, const void *P, double eps, int level, const double *c, hipStream_t st)
{
    Cinv<R> ci;
    for (int k = 0; k < 9; k++) ci.c[k] = (R)c[k];
    const long nl = g.V * g.D;
    hipLaunchKernelGGL((k_drift<N, R, Exact>), dim3((unsigned)((nl + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                       g, (R *)U, (const R *)P, eps, level, ci);
    return hipGetLastError();
}

template <int N>
hipError_t drift_mode(const Geom &g, void *U, const void *P, double eps, int level, const double *c, int mode,
                      hipStream_t st)
{
    switch (mode) {
    case kModeF64: return drift_n<N, double, false>(g, U, P, eps, level, c, st);
    case kModeF64Exact: return drift_n<N, double, true>(g, U, P, eps, level, c, st);
    default: return drift_n<N, float, false>(g, U, P, eps, level, c, st);
    }
}

template <int N>
hipError_t plaq_mode(const Geom &g, const void *U, int mode, double inv_n, double *part, long *np, hipStream_t st)
{
    const unsigned nb = (unsigned)((g.V + kBlock - 1) / kBlock);
    *np = nb;
    switch (mode) {
    case kModeF64:
        hipLaunchKernelGGL((k_plaq<N, double, false>), dim3(nb), dim3(kBlock), 0, st, g, (const double *)U, inv_n, part);
        break;
    case kModeF64Exact:
        hipLaunchKernelGGL((k_plaq<N, double, true>), dim3(nb), dim3(kBlock), 0, st, g, (const double *)U, inv_n, part);
        break;
    default:  // FP32 links, FP64 arithmetic: the exact action of the stored links
        hipLaunchKernelGGL((k_plaq<N, float, false, double>), dim3(nb), dim3(kBlock), 0, st, g, (const float *)U, inv_n,
                           part);
    }
    return hipGetLastError();
}

template <int N>
hipError_t uviol_mode(const Geom &g, const void *U, int mode, double *part, long *np, hipStream_t st)
{
    const unsigned nb = (unsigned)((g.V * g.D + kBlock - 1) / kBlock);
    *np = nb;
    if (f32_mode(mode)) hipLaunchKernelGGL((k_uviol<N, float>), dim3(nb), dim3(kBlock), 0, st, g, (const float *)U, part);
    else hipLaunchKernelGGL((k_uviol<N, double>), dim3(nb), dim3(kBlock), 0, st, g, (const double *)U, part);
    return hipGetLastError();
}

template <int N>
hipError_t gauss_mode(const Geom &g, const void *U, const void *P, int mode, double *part, long *np, double *out,
                      hipStream_t st)
{
    const unsigned nb = (unsigned)((g.V * g.D + kBlock - 1) / kBlock);
    *np = nb;
    if (f32_mode(mode))
        hipLaunchKernelGGL((k_gauss<N, float>), dim3(nb), dim3(kBlock), 0, st, g, (const float *)U, (const float *)P, part);
    else
        hipLaunchKernelGGL((k_gauss<N, double>), dim3(nb), dim3(kBlock), 0, st, g, (const double *)U, (const double *)P,
                           part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_finalize(part, nb, 2 * N * N, out, st);
}

template <int N>
hipError_t wlines_mode(const Geom &g, const void *U, int dir, int mode, double *out, hipStream_t st)
{
    const long nlines = g.V / g.ext[dir];
    const unsigned nb = (unsigned)((nlines + kBlock - 1) / kBlock);
    if (f32_mode(mode))
        hipLaunchKernelGGL((k_wlines<N, float>), dim3(nb), dim3(kBlock), 0, st, g, (const float *)U, dir, nlines, out);
    else
        hipLaunchKernelGGL((k_wlines<N, double>), dim3(nb), dim3(kBlock), 0, st, g, (const double *)U, dir, nlines, out);
    return hipGetLastError();
}

}  // namespace

// ---- dispatch for N <= 3 (called from the generic launchers) ---------------
hipError_t small_kick(const KickArgs &a, int mode, bool drift, hipStream_t st)
{
    switch (a.g.N) {
    case 1: return kick_mode<1>(a, mode, drift, st);
    case 2: return kick_mode<2>(a, mode, drift, st);
    default: return kick_mode<3>(a, mode, drift, st);
    }
}

hipError_t small_drift(const Geom &g, void *U, const void *P, double eps, int level, const double *c, int mode,
                       hipStream_t st)
{
    switch (g.N) {
    case 1: return drift_mode<1>(g, U, P, eps, level, c, mode, st);
    case 2: return drift_mode<2>(g, U, P, eps, level, c, mode, st);
    default: return drift_mode<3>(g, U, P, eps, level, c, mode, st);
    }
}

hipError_t small_plaq(const Geom &g, const void *U, int mode, double inv_n, double *part, long *np, hipStream_t st)
{
    switch (g.N) {
    case 1: return plaq_mode<1>(g, U, mode, inv_n, part, np, st);
    case 2: return plaq_mode<2>(g, U, mode, inv_n, part, np, st);
    default: return plaq_mode<3>(g, U, mode, inv_n, part, np, st);
    }
}

hipError_t small_uviol(const Geom &g, const void *U, int mode, double *part, long *np, hipStream_t st)
{
    switch (g.N) {
    case 1: return uviol_mode<1>(g, U, mode, part, np, st);
    case 2: return uviol_mode<2>(g, U, mode, part, np, st);
    default: return uviol_mode<3>(g, U, mode, part, np, st);
    }
}

hipError_t small_gauss(const Geom &g, const void *U, const void *P, int mode, double *part, long *np, double *out,
                       hipStream_t st)
{
    switch (g.N) {
    case 1: return gauss_mode<1>(g, U, P, mode, part, np, out, st);
    case 2: return gauss_mode<2>(g, U, P, mode, part, np, out, st);
    default: return gauss_mode<3>(g, U, P, mode, part, np, out, st);
    }
}

hipError_t small_project(const Geom &g, void *U, int mode, int haar, uint64_t seed, uint64_t stream, hipStream_t st)
{
    switch (g.N) {
    case 1: return project_mode<1>(g, U, mode, haar, seed, stream, st);
    case 2: return project_mode<2>(g, U, mode, haar, seed, stream, st);
    default: return project_mode<3>(g, U, mode, haar, seed, stream, st);
    }
}

hipError_t small_wlines(const Geom &g, const void *U, int dir, int mode, double *out, hipStream_t st)
{
    switch (g.N) {
    case 1: return wlines_mode<1>(g, U, dir, mode, out, st);
    case 2: return wlines_mode<2>(g, U, dir, mode, out, st);
    default: return wlines_mode<3>(g, U, dir, mode, out, st);
    }
}

}  // namespace askja
