// measure.hip -- device side of the measurement program (the counterpart of
// measurements4d.cpp): per-time-slice link traces (scalarLoopcorr,
// measurements4d.cpp:357-408) and the eigen-phases of straight Wilson and
// Polyakov lines (getevs, measurements4d.cpp:185-309).
#include "kernels.hpp"
#include "common.hpp"

namespace askja {
namespace {

constexpr int kT = 256;

template <typename R>
__device__ inline double lget(const R *U, const Geom &g, long l, int e)
{
    return (double)U[foff(l, e, 2 * g.N * g.N, g.tile)];
}

// Compensated partial sums of ReTr U_dir(x) / N over the sites x of time
// slice x0 = blockIdx.y (lexicographic sites x0 * slice .. x0 * slice +
// slice - 1); block (b, x0) writes its (s, c) pair to part[x0 * nb + b].
template <typename R>
__global__ __launch_bounds__(kT) void k_slice_trace(Geom g, const R *U, int dir, double *part)
{
    const long slice = g.V / g.ext[0];
    const long x0 = blockIdx.y;
    double s[1] = {0.0}, c[1] = {0.0};
    for (long k = (long)blockIdx.x * kT + threadIdx.x; k < slice; k += (long)gridDim.x * kT) {
        const long l = (long)dir * g.V + g.lex2dev[x0 * slice + k];
        double tr = 0.0;
        for (int i = 0; i < g.N; i++) tr += lget(U, g, l, 2 * (i * g.N + i));
        kahan_add(s[0], c[0], tr / g.N);
    }
    block_reduce_store<1>(s, c, part + 2 * ((size_t)x0 * gridDim.x + blockIdx.x));
}

// One workgroup per line.  The line product W = prod_t U_dir(x + t dir) is
// formed in LDS and kept in wbuf (2 N^2 doubles per line).  Its eigen-phases:
// W is normal, so it shares its eigenvectors with the Hermitian
// H = (e W + (e W)^+)/2, e = exp(i phi); H is diagonalised by parallel cyclic
// Jacobi (round-robin pairing: NP/2 disjoint rotations per step, NP - 1
// steps per sweep) and each phase is arg(v^+ W v).  If some Rayleigh
// quotient has |v^+ W v| < 1 - 1e-9 (two eigenvalues of H that coincide for
// this phi mixed their vectors), the diagonalisation is redone with the next
// phi.  Then, as getevs does: winding w = sum theta / 2 pi, theta -= 2 pi w / N
// when |w| > 0.1, and the N phases are sorted ascending.
template <typename R, int NMAX>
__global__ __launch_bounds__(kT) void k_line_phases(Geom g, const R *U, int dir, const long *starts, double *wbuf,
                                                    double *phases)
{
    __shared__ double Ar[NMAX * NMAX], Ai[NMAX * NMAX], Br[NMAX * NMAX], Bi[NMAX * NMAX];
    __shared__ int rp[NMAX / 2], rq[NMAX / 2];
    __shared__ double rcs[NMAX / 2], rsn[NMAX / 2], rphr[NMAX / 2], rphi[NMAX / 2];
    __shared__ double th[NMAX], red[kT / 64];
    __shared__ int bad;
    const int n = g.N, nn = n * n, tid = threadIdx.x;
    const int line = blockIdx.x;
    double *W = wbuf + (size_t)line * 2 * nn;

    // ---- line product, A <- A B
    long s = g.lex2dev[starts[line]];
    for (int e = tid; e < nn; e += kT) {
        Ar[e] = lget(U, g, (long)dir * g.V + s, 2 * e);
        Ai[e] = lget(U, g, (long)dir * g.V + s, 2 * e + 1);
    }
    for (int t = 1; t < g.ext[dir]; t++) {
        s = g.nbr[(2 * dir) * g.V + s];
        for (int e = tid; e < nn; e += kT) {
            Br[e] = lget(U, g, (long)dir * g.V + s, 2 * e);
            Bi[e] = lget(U, g, (long)dir * g.V + s, 2 * e + 1);
        }
        __syncthreads();
        double cr[(NMAX * NMAX + kT - 1) / kT], ci[(NMAX * NMAX + kT - 1) / kT];
#pragma unroll
        for (int q = 0; q < (NMAX * NMAX + kT - 1) / kT; q++) {
            const int e = tid + q * kT;
            cr[q] = ci[q] = 0.0;
            if (e < nn) {
                const int i = e / n, j = e % n;
                for (int k = 0; k < n; k++) {
                    cr[q] = fma(Ar[i * n + k], Br[k * n + j], cr[q]);
                    cr[q] = fma(-Ai[i * n + k], Bi[k * n + j], cr[q]);
                    ci[q] = fma(Ar[i * n + k], Bi[k * n + j], ci[q]);
                    ci[q] = fma(Ai[i * n + k], Br[k * n + j], ci[q]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < (NMAX * NMAX + kT - 1) / kT; q++) {
            const int e = tid + q * kT;
            if (e < nn) {
                Ar[e] = cr[q];
                Ai[e] = ci[q];
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < nn; e += kT) {
        W[2 * e] = Ar[e];
        W[2 * e + 1] = Ai[e];
    }
    __syncthreads();

    const int NP = n + (n & 1);
    const double phis[3] = {0.7390851332151607, 1.9140095476, 2.6532925713};
    for (int attempt = 0; attempt < 3; attempt++) {
        const double er = cos(phis[attempt]), ei = sin(phis[attempt]);
        // H = (e W + (e W)^+)/2 into A, V = 1 into B
        for (int e = tid; e < nn; e += kT) {
            const int i = e / n, j = e % n;
            const double wr = W[2 * (i * n + j)], wi = W[2 * (i * n + j) + 1];
            const double vr = W[2 * (j * n + i)], vi = W[2 * (j * n + i) + 1];
            // e w_ij + conj(e w_ji)
            Ar[e] = 0.5 * ((er * wr - ei * wi) + (er * vr - ei * vi));
            Ai[e] = 0.5 * ((er * wi + ei * wr) - (er * vi + ei * vr));
            Br[e] = i == j ? 1.0 : 0.0;
            Bi[e] = 0.0;
        }
        __syncthreads();
        for (int sweep = 0; sweep < 40; sweep++) {
            // off-diagonal norm
            double off = 0.0;
            for (int e = tid; e < nn; e += kT)
                if (e / n != e % n) off += Ar[e] * Ar[e] + Ai[e] * Ai[e];
            for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o);
            if ((tid & 63) == 0) red[tid >> 6] = off;
            __syncthreads();
            double tot = 0.0;
            for (int w = 0; w < kT / 64; w++) tot += red[w];
            __syncthreads();
            if (tot < 1e-28) break;
            for (int step = 0; step < NP - 1; step++) {
                if (tid < NP / 2) {
                    // circle method: position 0 fixed, the others rotate
                    auto at = [&](int pos) { return pos == 0 ? 0 : (pos - 1 + step) % (NP - 1) + 1; };
                    int p = at(tid), q = at(NP - 1 - tid);
                    if (p > q) {
                        const int tmp = p;
                        p = q;
                        q = tmp;
                    }
                    double c = 1.0, sn = 0.0, phr = 1.0, phi = 0.0;
                    if (q < n) {
                        const double hr = Ar[p * n + q], hi = Ai[p * n + q];
                        const double a = sqrt(hr * hr + hi * hi);
                        if (a > 1e-300) {
                            const double tau = (Ar[q * n + q] - Ar[p * n + p]) / (2.0 * a);
                            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                            c = 1.0 / sqrt(1.0 + t * t);
                            sn = t * c;
                            phr = hr / a;
                            phi = hi / a;
                        }
                    }
                    rp[tid] = p;
                    rq[tid] = q < n ? q : -1;
                    rcs[tid] = c;
                    rsn[tid] = sn;
                    rphr[tid] = phr;
                    rphi[tid] = phi;
                }
                __syncthreads();
                // H <- H J and V <- V J: columns p, q of every row k;
                // J = [c, s ph; -s conj(ph), c] in the (p, q) plane
                for (int it = tid; it < n * (NP / 2); it += kT) {
                    const int j = it / n, k = it % n;
                    const int q = rq[j];
                    if (q < 0) continue;
                    const int p = rp[j];
                    const double c = rcs[j], sn = rsn[j], pr = rphr[j], pi = rphi[j];
                    {
                        const double xr = Ar[k * n + p], xi = Ai[k * n + p], yr = Ar[k * n + q], yi = Ai[k * n + q];
                        // x' = c x - s conj(ph) y,  y' = s ph x + c y
                        Ar[k * n + p] = c * xr - sn * (pr * yr + pi * yi);
                        Ai[k * n + p] = c * xi - sn * (pr * yi - pi * yr);
                        Ar[k * n + q] = sn * (pr * xr - pi * xi) + c * yr;
                        Ai[k * n + q] = sn * (pr * xi + pi * xr) + c * yi;
                    }
                    {
                        const double xr = Br[k * n + p], xi = Bi[k * n + p], yr = Br[k * n + q], yi = Bi[k * n + q];
                        Br[k * n + p] = c * xr - sn * (pr * yr + pi * yi);
                        Bi[k * n + p] = c * xi - sn * (pr * yi - pi * yr);
                        Br[k * n + q] = sn * (pr * xr - pi * xi) + c * yr;
                        Bi[k * n + q] = sn * (pr * xi + pi * xr) + c * yi;
                    }
                }
                __syncthreads();
                // H <- J^+ H: rows p, q of every column k
                for (int it = tid; it < n * (NP / 2); it += kT) {
                    const int j = it / n, k = it % n;
                    const int q = rq[j];
                    if (q < 0) continue;
                    const int p = rp[j];
                    const double c = rcs[j], sn = rsn[j], pr = rphr[j], pi = rphi[j];
                    const double xr = Ar[p * n + k], xi = Ai[p * n + k], yr = Ar[q * n + k], yi = Ai[q * n + k];
                    // x' = c x - s ph y,  y' = s conj(ph) x + c y
                    Ar[p * n + k] = c * xr - sn * (pr * yr - pi * yi);
                    Ai[p * n + k] = c * xi - sn * (pr * yi + pi * yr);
                    Ar[q * n + k] = sn * (pr * xr + pi * xi) + c * yr;
                    Ai[q * n + k] = sn * (pr * xi - pi * xr) + c * yi;
                }
                __syncthreads();
            }
        }
        // Rayleigh quotients r_k = v_k^+ W v_k
        if (tid == 0) bad = 0;
        __syncthreads();
        for (int k = tid; k < n; k += kT) {
            double rr = 0.0, ri = 0.0;
            for (int i = 0; i < n; i++) {
                double wr = 0.0, wi = 0.0;
                for (int j = 0; j < n; j++) {
                    const double ar = W[2 * (i * n + j)], ai = W[2 * (i * n + j) + 1];
                    const double vr = Br[j * n + k], vi = Bi[j * n + k];
                    wr += ar * vr - ai * vi;
                    wi += ar * vi + ai * vr;
                }
                const double ur = Br[i * n + k], ui = -Bi[i * n + k];
                rr += ur * wr - ui * wi;
                ri += ur * wi + ui * wr;
            }
            th[k] = atan2(ri, rr);
            if (sqrt(rr * rr + ri * ri) < 1.0 - 1e-9) bad = 1;
        }
        __syncthreads();
        if (!bad) break;
        __syncthreads();
    }
    if (tid == 0) {
        double wind = 0.0;
        for (int k = 0; k < n; k++) wind += th[k] / (2.0 * M_PI);
        if (wind > 0.1 || wind < -0.1)
            for (int k = 0; k < n; k++) th[k] -= wind * 2.0 * M_PI / n;
        for (int k = 1; k < n; k++) {  // insertion sort, ascending
            const double v = th[k];
            int j = k - 1;
            while (j >= 0 && th[j] > v) {
                th[j + 1] = th[j];
                j--;
            }
            th[j + 1] = v;
        }
        for (int k = 0; k < n; k++) phases[(size_t)line * n + k] = th[k];
    }
}

template <typename R>
hipError_t phases_dispatch(const Geom &g, const R *U, int dir, int nlines, const long *starts, double *wbuf,
                           double *out, hipStream_t st)
{
    const int n = g.N;
    if (n <= 8) hipLaunchKernelGGL((k_line_phases<R, 8>), dim3(nlines), dim3(kT), 0, st, g, U, dir, starts, wbuf, out);
    else if (n <= 16)
        hipLaunchKernelGGL((k_line_phases<R, 16>), dim3(nlines), dim3(kT), 0, st, g, U, dir, starts, wbuf, out);
    else if (n <= 32)
        hipLaunchKernelGGL((k_line_phases<R, 32>), dim3(nlines), dim3(kT), 0, st, g, U, dir, starts, wbuf, out);
    else hipLaunchKernelGGL((k_line_phases<R, 64>), dim3(nlines), dim3(kT), 0, st, g, U, dir, starts, wbuf, out);
    return hipGetLastError();
}

}  // namespace

long slice_trace_blocks(const Geom &g)
{
    const long slice = g.V / g.ext[0];
    long nb = (slice + kT * 8 - 1) / (kT * 8);
    return nb < 1 ? 1 : nb;
}

hipError_t launch_slice_trace(const Geom &g, const void *U, int dir, int mode, double *part, hipStream_t st)
{
    const dim3 grid((unsigned)slice_trace_blocks(g), (unsigned)g.ext[0]);
    if (f32_mode(mode)) hipLaunchKernelGGL(k_slice_trace<float>, grid, dim3(kT), 0, st, g, (const float *)U, dir, part);
    else hipLaunchKernelGGL(k_slice_trace<double>, grid, dim3(kT), 0, st, g, (const double *)U, dir, part);
    return hipGetLastError();
}

hipError_t launch_line_phases(const Geom &g, const void *U, int dir, int mode, int nlines, const long *starts,
                              double *wbuf, double *out, hipStream_t st)
{
    if (f32_mode(mode)) return phases_dispatch(g, (const float *)U, dir, nlines, starts, wbuf, out, st);
    return phases_dispatch(g, (const double *)U, dir, nlines, starts, wbuf, out, st);
}

}  // namespace askja
