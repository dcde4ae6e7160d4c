// generic.hip -- kernels that run for every N: layout conversion, cold
// start, momentum refresh, kinetic sums, the final stage of every reduction,
// and the line matrix of N <= 3.  Also the top-level launchers that pick the
// register (N <= 3), row (4 <= N <= 8), workgroup (team_impl.hpp) or MFMA
// kernels.
#include "kernels.hpp"
#include "common.hpp"

namespace askja {

// exported by small.hip and team<NMAX>.hip
hipError_t small_kick(const KickArgs &a, int mode, bool drift, hipStream_t st);
hipError_t small_drift(const Geom &g, void *U, const void *P, double eps, int level, const double *c, int mode,
                       hipStream_t st);
hipError_t small_plaq(const Geom &g, const void *U, int mode, double inv_n, double *part, long *np, hipStream_t st);
hipError_t small_uviol(const Geom &g, const void *U, int mode, double *part, long *np, hipStream_t st);
hipError_t small_gauss(const Geom &g, const void *U, const void *P, int mode, double *part, long *np, double *out,
                       hipStream_t st);
hipError_t small_wlines(const Geom &g, const void *U, int dir, int mode, double *out, hipStream_t st);
hipError_t small_project(const Geom &g, void *U, int mode, int haar, uint64_t seed, uint64_t stream, hipStream_t st);
#define ASKJA_TEAM_DECL(NM)                                                                                   \
    hipError_t team_kick_##NM(const KickArgs &a, int mode, bool drift, hipStream_t st);                       \
    hipError_t team_drift_##NM(const Geom &g, void *U, const void *P, double eps, int level, const double *c, \
                               int mode, hipStream_t st);                                                     \
    hipError_t team_project_##NM(const Geom &g, void *U, int mode, int haar, uint64_t seed, uint64_t stream,    \
                                 hipStream_t st);                                                             \
    hipError_t team_plaq_##NM(const Geom &g, const void *U, int mode, double inv_n, double *part, hipStream_t st); \
    hipError_t team_lines_##NM(const Geom &g, const void *U, int dir, long nlines, long lex1, int mode,         \
                               double *traces, double *wfull, hipStream_t st);                                \
    hipError_t team_uviol_##NM(const Geom &g, const void *U, int mode, long nwg, double *part, hipStream_t st); \
    hipError_t team_gauss_##NM(const Geom &g, const void *U, const void *P, int mode, long nwg, double *part,   \
                               hipStream_t st);
ASKJA_TEAM_DECL(8)
ASKJA_TEAM_DECL(16)
ASKJA_TEAM_DECL(32)
ASKJA_TEAM_DECL(64)
#undef ASKJA_TEAM_DECL
hipError_t mfma_kick_16(const KickArgs &a, int mode, bool drift, hipStream_t st);
hipError_t mfma_kick_32(const KickArgs &a, int mode, bool drift, hipStream_t st);
hipError_t mfma_kick_64(const KickArgs &a, int mode, bool drift, hipStream_t st);
hipError_t mfma_drift_16(const Geom &g, void *U, const void *P, double eps, int level, const double *cinv, int mode,
                         hipStream_t st);
hipError_t mfma_drift_32(const Geom &g, void *U, const void *P, double eps, int level, const double *cinv, int mode,
                         hipStream_t st);
hipError_t mfma_drift_64(const Geom &g, void *U, const void *P, double eps, int level, const double *cinv, int mode,
                         hipStream_t st);
hipError_t rows_kick(const KickArgs &a, int mode, bool drift, hipStream_t st);
long rows_kick_blocks(const Geom &g);
hipError_t mfma_action_16(const KickArgs &a, bool f32_links, hipStream_t st);
hipError_t mfma_action_32(const KickArgs &a, bool f32_links, hipStream_t st);
hipError_t mfma_action_64(const KickArgs &a, bool f32_links, hipStream_t st);

namespace {

constexpr int kB = 256;

inline unsigned nblk(long n, int b = kB) { return (unsigned)((n + b - 1) / b); }

template <typename R>
__device__ inline R &lk(R *f, const Geom &g, long l, int e)
{
    return f[foff(l, e, 2 * g.N * g.N, g.tile)];
}
template <typename R>
__device__ inline R lkc(const R *f, const Geom &g, long l, int e)
{
    return f[foff(l, e, 2 * g.N * g.N, g.tile)];
}
template <typename R>
__device__ inline R &mk(R *f, const Geom &g, long l, int c)
{
    return f[foff(l, c, g.N * g.N, g.tile)];
}

// (i, j), i < j, of compact pair index p (0-based over the strict upper
// triangle, row-major).
__device__ inline void pair_ij(int n, int p, int &i, int &j)
{
    i = 0;
    while (p >= n - 1 - i) {
        p -= n - 1 - i;
        i++;
    }
    j = i + 1 + p;
}

// ------------------------------------------------------------- reductions
template <int K>
__global__ __launch_bounds__(1024) void k_finalize(const double *part, long np, double *out)
{
    double s[K], c[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = c[k] = 0.0;
    for (long p = threadIdx.x; p < np; p += 1024) {
#pragma unroll
        for (int k = 0; k < K; k++) pair_add(s[k], c[k], part[(size_t)p * 2 * K + 2 * k], part[(size_t)p * 2 * K + 2 * k + 1]);
    }
    block_reduce_store<K>(s, c, out);
}

// Column sums of np rows of K (s, c) pairs, for large K: thread per column.
__global__ void k_finalize_cols(const double *part, long np, int K, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    double s = 0.0, c = 0.0;
    for (long p = 0; p < np; p++) pair_add(s, c, part[(size_t)p * 2 * K + 2 * k], part[(size_t)p * 2 * K + 2 * k + 1]);
    out[2 * k] = s;
    out[2 * k + 1] = c;
}

// ------------------------------------------------------ layout conversion
template <typename R>
__global__ void k_links_from_aos(Geom g, const double *aos, R *U)
{
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    const int nr = 2 * g.N * g.N;
    const long tot = g.V * g.D * nr;
    if (t >= tot) return;
    const int e = (int)(t % nr);
    const long sd = t / nr;
    const int d = (int)(sd % g.D);
    const long i = sd / g.D;
    const long l = d * g.V + g.lex2dev[i];
    lk(U, g, l, e) = (R)aos[t];
}

template <typename R>
__global__ void k_links_to_aos(Geom g, const R *U, double *aos)
{
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    const int nr = 2 * g.N * g.N;
    const long tot = g.V * g.D * nr;
    if (t >= tot) return;
    const int e = (int)(t % nr);
    const long sd = t / nr;
    const int d = (int)(sd % g.D);
    const long i = sd / g.D;
    const long l = d * g.V + g.lex2dev[i];
    aos[t] = (double)lkc(U, g, l, e);
}

// compact element c of link l <- full Hermitian AoS matrix
template <typename R>
__global__ void k_mom_from_aos(Geom g, const double *aos, R *P)
{
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    const int n = g.N, nc = n * n;
    const long tot = g.V * g.D * nc;
    if (t >= tot) return;
    const int c = (int)(t % nc);
    const long sd = t / nc;
    const int d = (int)(sd % g.D);
    const long i = sd / g.D;
    const long l = d * g.V + g.lex2dev[i];
    const double *m = aos + sd * 2 * nc;
    double v;
    if (c < n) {
        v = m[2 * (c * n + c)];
    } else {
        int a, b;
        pair_ij(n, (c - n) >> 1, a, b);
        v = m[2 * (a * n + b) + ((c - n) & 1)];
    }
    mk(P, g, l, c) = (R)v;
}

template <typename R>
__global__ void k_mom_to_aos(Geom g, const R *P, double *aos)
{
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    const int n = g.N, nn = n * n;
    const long tot = g.V * g.D * nn;
    if (t >= tot) return;
    const int e = (int)(t % nn);
    const long sd = t / nn;
    const int d = (int)(sd % g.D);
    const long i = sd / g.D;
    const long l = d * g.V + g.lex2dev[i];
    const int a = e / n, b = e % n;
    double re, im;
    if (a == b) {
        re = (double)mk(const_cast<R *>(P), g, l, a);
        im = 0.0;
    } else if (a < b) {
        const int h = hup(n, a, b);
        re = (double)mk(const_cast<R *>(P), g, l, h);
        im = (double)mk(const_cast<R *>(P), g, l, h + 1);
    } else {
        const int h = hup(n, b, a);
        re = (double)mk(const_cast<R *>(P), g, l, h);
        im = -(double)mk(const_cast<R *>(P), g, l, h + 1);
    }
    aos[2 * t] = re;
    aos[2 * t + 1] = im;
}

template <typename R>
__global__ void k_cold(Geom g, R *U)
{
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    const int nr = 2 * g.N * g.N;
    if (t >= g.V * g.D * nr) return;
    const int e = (int)(t % nr);
    const long l = t / nr;
    const int ij = e >> 1;
    lk(U, g, l, e) = ((e & 1) == 0 && ij / g.N == ij % g.N) ? R(1) : R(0);
}

// ---------------------------------------------------------- random fields
// Normal deviate number a of (purpose, link) at `counter`: Philox block a/4,
// Box-Muller pair (a%4)/2, cos for even a, sin for odd a.
__device__ inline double normal_at(uint64_t seed, uint64_t counter, uint32_t purpose, uint32_t link, int a)
{
    const U4 r = philox_block(seed, counter, purpose, link, (uint32_t)(a >> 2));
    const int h = (a >> 1) & 1;
    double zc, zs;
    box_muller(h ? r.z : r.x, h ? r.w : r.y, &zc, &zs);
    return (a & 1) ? zs : zc;
}

// pi = c sum_a xi_a T^a with the reference's generator order (sun.cpp:142-189,
// 209-252): symmetric/antisymmetric pairs for i < j, then the Cartan
// generators diag(1,..,1,-(m+1),0,..)/sqrt(2(m+1)(m+2)).  RNG link index =
// dir * V + lexicographic site.
template <typename R>
__global__ void k_mom_refresh(Geom g, R *P, uint64_t seed, uint64_t counter, double c)
{
#pragma clang fp contract(off)
    const long l = (long)blockIdx.x * kB + threadIdx.x;
    if (l >= g.V * g.D) return;
    const int n = g.N;
    const int d = (int)(l / g.V);
    const uint32_t link = (uint32_t)(d * g.V + g.dev2lex[l - d * g.V]);
    // off-diagonal pairs, normals 0 .. n(n-1)-1 in order
    int a = 0;
    U4 blk = {0u, 0u, 0u, 0u};
    double z[4];
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            if ((a & 3) == 0) {
                blk = philox_block(seed, counter, kPurposeMomentum, link, (uint32_t)(a >> 2));
                normals4(blk, z);
            }
            const double xs = z[a & 3], xa = z[(a & 3) + 1];
            const int h = hup(n, i, j);
            mk(P, g, l, h) = (R)(0.5 * c * xs);
            mk(P, g, l, h + 1) = (R)(-(0.5 * c * xa));
            a += 2;
        }
    // diagonal: P_rr = sum_{m >= r} h_m - r h_{r-1}, h_m = c xi_m / sqrt(2 (m+1)(m+2))
    const int a0 = n * (n - 1);
    double suffix = 0.0;
    for (int r = n - 1; r >= 0; r--) {
        double v = suffix;
        if (r >= 1) {
            const int m = r - 1;
            const double hm = c * normal_at(seed, counter, kPurposeMomentum, link, a0 + m) *
                              sqrt(0.5 / ((double)(m + 1) * (m + 2)));
            v -= (double)r * hm;
            suffix += hm;
        }
        mk(P, g, l, r) = (R)v;
    }
}

template <typename R>
__global__ void k_mom_scale(Geom g, R *P, double f)
{
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    const int nc = g.N * g.N;
    if (t >= g.V * g.D * nc) return;
    const long l = t / nc;
    const int c = (int)(t % nc);
    mk(P, g, l, c) = (R)((double)mk(P, g, l, c) * f);
}


// ------------------------------------------------------------- kinetic
// {Tr P^2, 1/4 Tr (P + Q)^2} per link from the compact Hermitian layout.
template <typename R>
__global__ __launch_bounds__(kB) void k_kinetic(Geom g, const R *P, const R *Q, double *part)
{
    const long l = (long)blockIdx.x * kB + threadIdx.x;
    double s[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
    if (l < g.V * g.D) {
        const int n = g.N, nc = n * n;
        double a = 0.0, b = 0.0;
        for (int e = 0; e < nc; e++) {
            const double w = e < n ? 1.0 : 2.0;
            const double p = (double)mk(const_cast<R *>(P), g, l, e);
            a += w * p * p;
            if (Q) {
                const double pq = p + (double)mk(const_cast<R *>(Q), g, l, e);
                b += w * pq * pq;
            }
        }
        s[0] = a;
        s[1] = 0.25 * b;
    }
    block_reduce_store<2>(s, c, part + 4 * (size_t)blockIdx.x);
}

// ------------------------------------------- N >= 4 diagnostics (FP64 work)
// C = op(A) op(B), n x n complex, interleaved, k ascending, no contraction.
__device__ inline void gmul(const double *A, bool dA, const double *B, bool dB, double *C, int n)
{
#pragma clang fp contract(off)
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double cr = 0.0, ci = 0.0;
            for (int k = 0; k < n; k++) {
                const int ia = dA ? k * n + i : i * n + k;
                const int ib = dB ? j * n + k : k * n + j;
                const double ar = A[2 * ia], ai = dA ? -A[2 * ia + 1] : A[2 * ia + 1];
                const double br = B[2 * ib], bi = dB ? -B[2 * ib + 1] : B[2 * ib + 1];
                cr = cr + (ar * br - ai * bi);
                ci = ci + (ar * bi + ai * br);
            }
            C[2 * (i * n + j)] = cr;
            C[2 * (i * n + j) + 1] = ci;
        }
}

template <typename R>
__device__ inline void gload(const R *U, const Geom &g, long l, double *A)
{
    const int nr = 2 * g.N * g.N;
    for (int e = 0; e < nr; e++) A[e] = (double)lkc(U, g, l, e);
}

// Full product matrix of one straight line for N <= 3 (the matrix W of
// WilsonLoop, sun.cpp:307-345), one thread, FP64 in the reference's product
// order.
template <typename R>
__global__ void k_line_matrix(Geom g, const R *U, int dir, long lex, double *scratch, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int n = g.N, nr = 2 * n * n;
    double *B = scratch, *T = B + nr;
    long s = g.lex2dev[lex];
    gload(U, g, dir * g.V + s, out);
    for (int k = 1; k < g.ext[dir]; k++) {
        s = g.nbr[(2 * dir) * g.V + s];
        gload(U, g, dir * g.V + s, B);
        gmul(out, false, B, false, T, n);
        for (int e = 0; e < nr; e++) out[e] = T[e];
    }
}


}  // namespace


// ---------------------------------------------------------------- dispatch
// 4 <= N <= 8: the row kernel puts 64/N links in one single-wavefront
// workgroup and walks each link's staples and Horner products serially, so a
// small lattice leaves most SIMDs idle and the launch is latency bound (the
// reference's example parameters, N 8, D 10, 100 sites: 1000 links in 125
// workgroups).  Below 2560 links (2.5 per SIMD) the workgroup-per-link
// kernel, one matrix element per lane, runs instead.  Measured crossovers
// (kick+drift, MI355X): N 8 about 3600 links in FP64 and 6000 in the exact
// mode, N 4 about 2200 and 4300.
constexpr long kRowsMinLinks = 2560;

static bool small_lattice_team(const Geom &g)
{
    return g.N >= 4 && g.N <= 8 && g.V * g.D < kRowsMinLinks;
}

long plaq_partials(const Geom &g)
{
    if (g.N <= 3) return (g.V + 255) / 256;
    return g.V;
}

long kick_partials(const Geom &g)
{
    if (g.N <= 3) return kick_site_blocks(g) * g.D;
    if (g.N <= 8 && !small_lattice_team(g)) return rows_kick_blocks(g);
    return g.V * g.D;
}

hipError_t launch_kick(const KickArgs &a, int mode, bool drift, hipStream_t st)
{
    const int n = a.g.N;
    // N >= 9: complex link products on MFMA, except the exact mode, which
    // keeps the reference's operation order on the VALU
    if (n > 8 && mode != kModeF64Exact) {
        if (n <= 16) return mfma_kick_16(a, mode, drift, st);
        if (n <= 32) return mfma_kick_32(a, mode, drift, st);
        return mfma_kick_64(a, mode, drift, st);
    }
    if (mode == kModeBF16x3) mode = kModeF32;
    if (n <= 3) return small_kick(a, mode, drift, st);
    if (n <= 8) return small_lattice_team(a.g) ? team_kick_8(a, mode, drift, st) : rows_kick(a, mode, drift, st);
    if (n <= 16) return team_kick_16(a, mode, drift, st);
    if (n <= 32) return team_kick_32(a, mode, drift, st);
    return team_kick_64(a, mode, drift, st);
}

hipError_t launch_action_f64(const KickArgs &a, int mode, hipStream_t st)
{
    const int n = a.g.N;
    if (n <= 8) return hipErrorInvalidValue;
    const bool f32 = f32_mode(mode);
    if (n <= 16) return mfma_action_16(a, f32, st);
    if (n <= 32) return mfma_action_32(a, f32, st);
    return mfma_action_64(a, f32, st);
}

// The action pass's partials (one per link workgroup lb, direction lb % D,
// first (s, c) pair = ReTr W of the link) -> out = {sum over every link, sum
// over the links of direction 0}, compensated.
__global__ __launch_bounds__(1024) void k_action_split(const double *part, long np, int D, double *out)
{
    double s[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
    for (long p = threadIdx.x; p < np; p += 1024) {
        const double v = part[(size_t)p * 6], w = part[(size_t)p * 6 + 1];
        pair_add(s[0], c[0], v, w);
        if (p % D == 0) pair_add(s[1], c[1], v, w);
    }
    block_reduce_store<2>(s, c, out);
}

hipError_t launch_action_split(const double *part, long np, int D, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_action_split, dim3(1), dim3(1024), 0, st, part, np, D, out);
    return hipGetLastError();
}

hipError_t launch_drift(const Geom &g, void *U, const void *P, double eps, int level, const double *cinv, int mode,
                        hipStream_t st)
{
    const int n = g.N;
    if (mode == kModeBF16x3) mode = kModeF32;
    if (n <= 3) return small_drift(g, U, P, eps, level, cinv, mode, st);
    if (n <= 8) return team_drift_8(g, U, P, eps, level, cinv, mode, st);
    // N >= 9: the Horner products on MFMA, except the exact mode (the
    // reference's operation order on the VALU)
    if (mode != kModeF64Exact) {
        if (n <= 16) return mfma_drift_16(g, U, P, eps, level, cinv, mode, st);
        if (n <= 32) return mfma_drift_32(g, U, P, eps, level, cinv, mode, st);
        return mfma_drift_64(g, U, P, eps, level, cinv, mode, st);
    }
    if (n <= 16) return team_drift_16(g, U, P, eps, level, cinv, mode, st);
    if (n <= 32) return team_drift_32(g, U, P, eps, level, cinv, mode, st);
    return team_drift_64(g, U, P, eps, level, cinv, mode, st);
}

hipError_t launch_finalize(const double *part, long np, int K, double *out, hipStream_t st)
{
    switch (K) {
    case 1: hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(1024), 0, st, part, np, out); break;
    case 2: hipLaunchKernelGGL((k_finalize<2>), dim3(1), dim3(1024), 0, st, part, np, out); break;
    case 3: hipLaunchKernelGGL((k_finalize<3>), dim3(1), dim3(1024), 0, st, part, np, out); break;
    default: hipLaunchKernelGGL(k_finalize_cols, dim3(nblk(K)), dim3(kB), 0, st, part, np, K, out);
    }
    return hipGetLastError();
}

hipError_t launch_plaq(const Geom &g, const void *U, int mode, double inv_n, double *part, long *np, hipStream_t st)
{
    if (g.N <= 3) return small_plaq(g, U, mode, inv_n, part, np, st);
    // one workgroup per site, products through LDS
    *np = g.V;
    if (g.N <= 8) return team_plaq_8(g, U, mode, inv_n, part, st);
    if (g.N <= 16) return team_plaq_16(g, U, mode, inv_n, part, st);
    if (g.N <= 32) return team_plaq_32(g, U, mode, inv_n, part, st);
    return team_plaq_64(g, U, mode, inv_n, part, st);
}

long gauss_rows(const Geom &g)
{
    const long nl = g.V * g.D;
    return nl < 256 ? nl : 256;
}

hipError_t launch_uviol(const Geom &g, const void *U, int mode, double *part, long *np, hipStream_t st)
{
    if (g.N <= 3) return small_uviol(g, U, mode, part, np, st);
    // workgroups over the links (grid-stride), one (s, c) pair each
    const long nl = g.V * g.D, nwg = nl < 1024 ? nl : 1024;
    *np = nwg;
    if (g.N <= 8) return team_uviol_8(g, U, mode, nwg, part, st);
    if (g.N <= 16) return team_uviol_16(g, U, mode, nwg, part, st);
    if (g.N <= 32) return team_uviol_32(g, U, mode, nwg, part, st);
    return team_uviol_64(g, U, mode, nwg, part, st);
}

hipError_t launch_gauss(const Geom &g, const void *U, const void *P, int mode, double *part, long *np, double *out,
                        hipStream_t st)
{
    if (g.N <= 3) return small_gauss(g, U, P, mode, part, np, out, st);
    // gauss_rows(g) workgroups over the links, one row of 2N^2 (s, c) pairs
    // each, then the column sums
    const int nr = 2 * g.N * g.N;
    const long nwg = gauss_rows(g);
    *np = nwg;
    hipError_t e;
    if (g.N <= 8) e = team_gauss_8(g, U, P, mode, nwg, part, st);
    else if (g.N <= 16) e = team_gauss_16(g, U, P, mode, nwg, part, st);
    else if (g.N <= 32) e = team_gauss_32(g, U, P, mode, nwg, part, st);
    else e = team_gauss_64(g, U, P, mode, nwg, part, st);
    if (e != hipSuccess) return e;
    return launch_finalize(part, nwg, nr, out, st);
}

hipError_t launch_kinetic(const Geom &g, const void *P, const void *Q, int mode, double *part, long *np,
                          hipStream_t st)
{
    const long nl = g.V * g.D;
    *np = nblk(nl);
    if (f32_mode(mode))
        hipLaunchKernelGGL(k_kinetic<float>, dim3(nblk(nl)), dim3(kB), 0, st, g, (const float *)P, (const float *)Q, part);
    else
        hipLaunchKernelGGL(k_kinetic<double>, dim3(nblk(nl)), dim3(kB), 0, st, g, (const double *)P, (const double *)Q,
                           part);
    return hipGetLastError();
}

// straight-line products of 4 <= N <= 64: one workgroup per line
static hipError_t team_lines(const Geom &g, const void *U, int dir, long nlines, long lex1, int mode, double *traces,
                             double *wfull, hipStream_t st)
{
    if (g.N <= 8) return team_lines_8(g, U, dir, nlines, lex1, mode, traces, wfull, st);
    if (g.N <= 16) return team_lines_16(g, U, dir, nlines, lex1, mode, traces, wfull, st);
    if (g.N <= 32) return team_lines_32(g, U, dir, nlines, lex1, mode, traces, wfull, st);
    return team_lines_64(g, U, dir, nlines, lex1, mode, traces, wfull, st);
}

hipError_t launch_wlines(const Geom &g, const void *U, int dir, int mode, double *out, hipStream_t st)
{
    if (g.N <= 3) return small_wlines(g, U, dir, mode, out, st);
    return team_lines(g, U, dir, g.V / g.ext[dir], -1, mode, out, nullptr, st);
}

hipError_t launch_line_matrix(const Geom &g, const void *U, int dir, long lex, int mode, double *scratch, double *out,
                              hipStream_t st)
{
    if (g.N >= 4) return team_lines(g, U, dir, 1, lex, mode, nullptr, out, st);
    if (f32_mode(mode))
        hipLaunchKernelGGL(k_line_matrix<float>, dim3(1), dim3(64), 0, st, g, (const float *)U, dir, lex, scratch, out);
    else hipLaunchKernelGGL(k_line_matrix<double>, dim3(1), dim3(64), 0, st, g, (const double *)U, dir, lex, scratch, out);
    return hipGetLastError();
}

hipError_t launch_links_from_aos(const Geom &g, const double *aos, void *U, int mode, hipStream_t st)
{
    const long tot = g.V * g.D * 2L * g.N * g.N;
    if (f32_mode(mode)) hipLaunchKernelGGL(k_links_from_aos<float>, dim3(nblk(tot)), dim3(kB), 0, st, g, aos, (float *)U);
    else hipLaunchKernelGGL(k_links_from_aos<double>, dim3(nblk(tot)), dim3(kB), 0, st, g, aos, (double *)U);
    return hipGetLastError();
}

hipError_t launch_links_to_aos(const Geom &g, const void *U, double *aos, int mode, hipStream_t st)
{
    const long tot = g.V * g.D * 2L * g.N * g.N;
    if (f32_mode(mode)) hipLaunchKernelGGL(k_links_to_aos<float>, dim3(nblk(tot)), dim3(kB), 0, st, g, (const float *)U, aos);
    else hipLaunchKernelGGL(k_links_to_aos<double>, dim3(nblk(tot)), dim3(kB), 0, st, g, (const double *)U, aos);
    return hipGetLastError();
}

hipError_t launch_mom_from_aos(const Geom &g, const double *aos, void *P, int mode, hipStream_t st)
{
    const long tot = g.V * g.D * (long)g.N * g.N;
    if (f32_mode(mode)) hipLaunchKernelGGL(k_mom_from_aos<float>, dim3(nblk(tot)), dim3(kB), 0, st, g, aos, (float *)P);
    else hipLaunchKernelGGL(k_mom_from_aos<double>, dim3(nblk(tot)), dim3(kB), 0, st, g, aos, (double *)P);
    return hipGetLastError();
}

hipError_t launch_mom_to_aos(const Geom &g, const void *P, double *aos, int mode, hipStream_t st)
{
    const long tot = g.V * g.D * (long)g.N * g.N;
    if (f32_mode(mode)) hipLaunchKernelGGL(k_mom_to_aos<float>, dim3(nblk(tot)), dim3(kB), 0, st, g, (const float *)P, aos);
    else hipLaunchKernelGGL(k_mom_to_aos<double>, dim3(nblk(tot)), dim3(kB), 0, st, g, (const double *)P, aos);
    return hipGetLastError();
}

hipError_t launch_links_cold(const Geom &g, void *U, int mode, hipStream_t st)
{
    const long tot = g.V * g.D * 2L * g.N * g.N;
    if (f32_mode(mode)) hipLaunchKernelGGL(k_cold<float>, dim3(nblk(tot)), dim3(kB), 0, st, g, (float *)U);
    else hipLaunchKernelGGL(k_cold<double>, dim3(nblk(tot)), dim3(kB), 0, st, g, (double *)U);
    return hipGetLastError();
}

hipError_t launch_mom_refresh(const Geom &g, void *P, uint64_t seed, uint64_t counter, double c, int mode,
                              hipStream_t st)
{
    const long nl = g.V * g.D;
    if (f32_mode(mode))
        hipLaunchKernelGGL(k_mom_refresh<float>, dim3(nblk(nl)), dim3(kB), 0, st, g, (float *)P, seed, counter, c);
    else hipLaunchKernelGGL(k_mom_refresh<double>, dim3(nblk(nl)), dim3(kB), 0, st, g, (double *)P, seed, counter, c);
    return hipGetLastError();
}

hipError_t launch_mom_scale(const Geom &g, void *P, double f, int mode, hipStream_t st)
{
    const long tot = g.V * g.D * (long)g.N * g.N;
    if (f32_mode(mode)) hipLaunchKernelGGL(k_mom_scale<float>, dim3(nblk(tot)), dim3(kB), 0, st, g, (float *)P, f);
    else hipLaunchKernelGGL(k_mom_scale<double>, dim3(nblk(tot)), dim3(kB), 0, st, g, (double *)P, f);
    return hipGetLastError();
}

hipError_t launch_project(const Geom &g, void *U, int mode, int haar, uint64_t seed, uint64_t stream, hipStream_t st)
{
    if (g.N <= 3) return small_project(g, U, mode, haar, seed, stream, st);
    if (g.N <= 8) return team_project_8(g, U, mode, haar, seed, stream, st);
    if (g.N <= 16) return team_project_16(g, U, mode, haar, seed, stream, st);
    if (g.N <= 32) return team_project_32(g, U, mode, haar, seed, stream, st);
    if (g.N <= 64) return team_project_64(g, U, mode, haar, seed, stream, st);
    return hipErrorInvalidValue;  // askja_create admits N <= 64
}

}  // namespace askja
