// flow2d.hip -- the spectral field transformation U = F(V) of
// two-dimensional SU(N) lattices (2 <= N <= 64), used as the FTHMC proposal
// map: forward with its log-Jacobian, and the hand-written backward of
// Phi = beta S_W(F(V)) - log J(V).
//
// Sub-layer (mu, off), nu = 1 - mu, updates U_mu(x) for x_nu = off (mod 4);
// every other link is frozen.  For an updated link
//   S  = U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+   (frozen: x+nu lies in row off+1)
//   W  = U_mu(x) S                            (the plaquette through the link)
//   f  = ReTr P(x+nu+k mu) / N, k = -1, 0, 1  (frozen plaquettes of row off+1)
//   t  = t_max tanh(b2 + w2 . tanh(W1 f + b1))
//   B  = 1 - t W,   h = B^{-1} (W - t)
//   U' = e^{2 i phi / N} h S^+,   phi = Im Tr log B (continuous branch),
// i.e. every eigenvalue of W goes through the Moebius map
// lambda -> (lambda - t) / (1 - t lambda), and a common phase restores
// det = 1.  Weyl's integration formula gives the Jacobian against the Haar
// measure without eigenvalues:
//   log J = N^2 log(1 - t^2) - 2N log|det B|
//           + log((1 + t^2 - (2t/N) ReTr W) / (1 - t^2)).
// B is factorised by LU without pivoting: its Hermitian part is at least
// (1 - |t|) > 0, so every leading block is nonsingular and the growth is
// bounded.  The pivots give log|det B| and arg det B mod 2 pi; phi is the
// value of that class nearest the truncated series
// -sum_{k<=4} t^k Im Tr W^k / k, whose error is below
// N |t|^5 / (5 (1 - |t|)) <= 0.8 rad for |t| <= 1/2, N <= 64.
//
// Backward (G' = Lie gradient of Phi at the output link; AH(M) =
// (M - M^+) / (2i), TL = traceless part, c = -1 for the -log J of Phi,
// m = 1 + t^2 - (2t/N) ReTr W):
//   own link:    g_U = TL[(1 - t^2) B^{-+} G' B^{-1} + c g_L],
//                g_L = -N t AH(h) / (1 - t^2) + t / (N m) AH(W);
//   staples, with g_s = g_U - E^+ G' E, E = U' U^+:
//                U_nu(x+mu) += U^+ g_s U,  U_mu(x+nu) += -A^+ g_s A
//                (A = W U_nu(x)),  U_nu(x) += -W^+ g_s W;
//   dPhi/dt = 4 Tr(G' AH(h)) / (1 - t^2) + c dlogJ/dt, then the MLP backward;
//   the links of every feature plaquette += dPhi/df (-AH(P_l) / (2N)), P_l
//   the plaquette rotated to start at the link (k_featgrad).
// Gradients live in the momentum layout (compact Hermitian, N^2 reals per
// link), so that the FTHMC kick is P <- P - eps (2N / beta) g elementwise.
// Contributions to frozen links are FP64 atomic adds (unsafeAtomicAdd: the
// buffers are ordinary device memory, where the hardware add is exact).
#include "kernels.hpp"
#include "common.hpp"

namespace askja {
namespace sflow {

constexpr int kHid = 32, kNF = 3;

struct Args {
    Geom g;
    double *U;          // field (FP64); forward: updated links written in place;
                        // backward: their inputs restored from `log`
    int mu, off;
    const float *w;     // MLP block
    double tmax;
    double *ptr;        // ReTr P / N per device site (rows off+1)
    double *log;        // the updated links' inputs, 2 N^2 reals each (nullable in the forward)
    double *part;       // forward: log J per updated link as (s, c) pairs
    double *gr;         // backward: Lie gradient, momentum layout
    double c_logj;
    double *gw;         // backward: weight gradient (nullable)
    double *dF;         // backward: dPhi/df per feature plaquette (device site)
};

template <int NMAX, int NT>
struct Cfg {
    static constexpr int LD = NMAX + 1;  // odd leading dimension: transposed reads spread over banks
    static constexpr int SZ = NMAX * LD;
    static constexpr int EPL = (NMAX * NMAX + NT - 1) / NT;
    static constexpr int NW = NT / 64;
};

__device__ inline size_t loff(const Geom &g, long l, int r, int n) { return elem_off(l, r, 2 * n * n, g.tile); }
__device__ inline size_t moff(const Geom &g, long l, int e, int n) { return elem_off(l, e, n * n, g.tile); }

// device site of active index b: x_mu = b % L_mu, x_nu = row + 4 (b / L_mu)
__device__ inline long act_site(const Geom &g, int mu, int row, long b)
{
    const long Lm = g.ext[mu];
    const long xm = b % Lm, xn = row + 4 * (b / Lm);
    const long x0 = mu == 0 ? xm : xn, x1 = mu == 0 ? xn : xm;
    return g.lex2dev[x0 * g.ext[1] + x1];
}

template <int NMAX, int NT>
struct Own {
    static constexpr int EPL = Cfg<NMAX, NT>::EPL;
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

template <int NMAX, int NT>
struct Reg {
    double r[Cfg<NMAX, NT>::EPL], i[Cfg<NMAX, NT>::EPL];
};

// global link -> LDS matrix
template <int NMAX, int NT>
__device__ inline void ld(const Geom &g, double *re, double *im, const double *f, long l, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
    for (int r = threadIdx.x; r < 2 * n * n; r += NT) {
        const double v = f[loff(g, l, r, n)];
        const int e = r >> 1;
        (r & 1 ? im : re)[(e / n) * LD + e % n] = v;
    }
}

// contiguous interleaved link (the input log) -> LDS matrix
template <int NMAX, int NT>
__device__ inline void ld_log(double *re, double *im, const double *b, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
    for (int r = threadIdx.x; r < 2 * n * n; r += NT) {
        const int e = r >> 1;
        (r & 1 ? im : re)[(e / n) * LD + e % n] = b[r];
    }
}

template <int NMAX, int NT>
__device__ inline void st(const Own<NMAX, NT> &o, double *re, double *im, const Reg<NMAX, NT> &c)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
#pragma unroll
    for (int q = 0; q < Own<NMAX, NT>::EPL; q++)
        if (o.ok[q]) {
            re[o.i[q] * LD + o.j[q]] = c.r[q];
            im[o.i[q] * LD + o.j[q]] = c.i[q];
        }
}

// c = op(A) op(B) on the owned elements (op = ^+ where DA / DB)
template <bool DA, bool DB, int NMAX, int NT>
__device__ inline void mm(const Own<NMAX, NT> &o, const double *ar_, const double *ai_, const double *br_,
                          const double *bi_, int n, Reg<NMAX, NT> &c)
{
    constexpr int LD = Cfg<NMAX, NT>::LD, EPL = Cfg<NMAX, NT>::EPL;
#pragma unroll
    for (int q = 0; q < EPL; q++) c.r[q] = c.i[q] = 0.0;
    for (int k = 0; k < n; k++) {
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            const int ia = DA ? k * LD + o.i[q] : o.i[q] * LD + k;
            const int ib = DB ? o.j[q] * LD + k : k * LD + o.j[q];
            const double ar = ar_[ia], ai = DA ? -ai_[ia] : ai_[ia];
            const double br = br_[ib], bi = DB ? -bi_[ib] : bi_[ib];
            c.r[q] = fma(ar, br, c.r[q]);
            c.r[q] = fma(-ai, bi, c.r[q]);
            c.i[q] = fma(ar, bi, c.i[q]);
            c.i[q] = fma(ai, br, c.i[q]);
        }
    }
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum of v over the workgroup, returned to every thread (red: NW * K doubles)
template <int NT, int K>
__device__ inline void bsum(double (&v)[K], double *red)
{
    constexpr int NW = NT / 64;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) red[w * K + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int u = 0; u < NW; u++) s += red[u * K + k];
        v[k] = s;
    }
    __syncthreads();
}

// 1 / (pr + i pi)
__device__ inline void crecip(double pr, double pi, double &ir, double &ii)
{
    const double d = 1.0 / (pr * pr + pi * pi);
    ir = pr * d;
    ii = -pi * d;
}

// In-place LU of the matrix in (Ar, Ai) without pivoting, the multipliers
// left unscaled: L_ik = A_ik / A_kk (i > k), Ub_kj = A_kj (j >= k).
template <int NMAX, int NT>
__device__ inline void lu(double *Ar, double *Ai, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
    for (int k = 0; k < n - 1; k++) {
        double ipr, ipi;
        crecip(Ar[k * LD + k], Ai[k * LD + k], ipr, ipi);
        const int m = n - 1 - k;
        for (int e = threadIdx.x; e < m * m; e += NT) {
            const int i = k + 1 + e / m, j = k + 1 + e % m;
            const double ar = Ar[i * LD + k], ai = Ai[i * LD + k];
            const double lr = ar * ipr - ai * ipi, li = ar * ipi + ai * ipr;
            const double ur = Ar[k * LD + j], ui = Ai[k * LD + j];
            Ar[i * LD + j] -= lr * ur - li * ui;
            Ai[i * LD + j] -= lr * ui + li * ur;
        }
        __syncthreads();
    }
}

// (Br, Bi) <- A^{-1} (Br, Bi) with A = L Ub from lu()
template <int NMAX, int NT>
__device__ inline void solve_left(const double *Ar, const double *Ai, double *Br, double *Bi, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
    for (int k = 0; k < n - 1; k++) {  // unit lower L
        double ipr, ipi;
        crecip(Ar[k * LD + k], Ai[k * LD + k], ipr, ipi);
        const int m = n - 1 - k;
        for (int e = threadIdx.x; e < m * n; e += NT) {
            const int i = k + 1 + e / n, j = e % n;
            const double ar = Ar[i * LD + k], ai = Ai[i * LD + k];
            const double lr = ar * ipr - ai * ipi, li = ar * ipi + ai * ipr;
            const double br = Br[k * LD + j], bi = Bi[k * LD + j];
            Br[i * LD + j] -= lr * br - li * bi;
            Bi[i * LD + j] -= lr * bi + li * br;
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 1; k--) {  // upper Ub, row k divided at the end
        double ipr, ipi;
        crecip(Ar[k * LD + k], Ai[k * LD + k], ipr, ipi);
        for (int e = threadIdx.x; e < k * n; e += NT) {
            const int i = e / n, j = e % n;
            const double ur = Ar[i * LD + k], ui = Ai[i * LD + k];
            const double xr0 = Br[k * LD + j], xi0 = Bi[k * LD + j];
            const double xr = xr0 * ipr - xi0 * ipi, xi = xr0 * ipi + xi0 * ipr;
            Br[i * LD + j] -= ur * xr - ui * xi;
            Bi[i * LD + j] -= ur * xi + ui * xr;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < n * n; e += NT) {
        const int i = e / n, j = e % n;
        double ipr, ipi;
        crecip(Ar[i * LD + i], Ai[i * LD + i], ipr, ipi);
        const double xr = Br[i * LD + j], xi = Bi[i * LD + j];
        Br[i * LD + j] = xr * ipr - xi * ipi;
        Bi[i * LD + j] = xr * ipi + xi * ipr;
    }
    __syncthreads();
}

// (Rr, Ri) <- (Rr, Ri) A^{-1} with A = L Ub from lu(): Z Ub = R, then Y L = Z
template <int NMAX, int NT>
__device__ inline void solve_right(const double *Ar, const double *Ai, double *Rr, double *Ri, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
    for (int k = 0; k < n - 1; k++) {  // columns j > k -= Z_:k Ub_kj, Z_:k = R_:k / Ub_kk
        double ipr, ipi;
        crecip(Ar[k * LD + k], Ai[k * LD + k], ipr, ipi);
        const int m = n - 1 - k;
        for (int e = threadIdx.x; e < n * m; e += NT) {
            const int i = e / m, j = k + 1 + e % m;
            const double zr0 = Rr[i * LD + k], zi0 = Ri[i * LD + k];
            const double zr = zr0 * ipr - zi0 * ipi, zi = zr0 * ipi + zi0 * ipr;
            const double ur = Ar[k * LD + j], ui = Ai[k * LD + j];
            Rr[i * LD + j] -= zr * ur - zi * ui;
            Ri[i * LD + j] -= zr * ui + zi * ur;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < n * n; e += NT) {
        const int i = e / n, j = e % n;
        double ipr, ipi;
        crecip(Ar[j * LD + j], Ai[j * LD + j], ipr, ipi);
        const double xr = Rr[i * LD + j], xi = Ri[i * LD + j];
        Rr[i * LD + j] = xr * ipr - xi * ipi;
        Ri[i * LD + j] = xr * ipi + xi * ipr;
    }
    __syncthreads();
    for (int k = n - 1; k >= 1; k--) {  // columns j < k -= Y_:k L_kj, L_kj = A_kj / A_jj
        const int m = k;
        for (int e = threadIdx.x; e < n * m; e += NT) {
            const int i = e / m, j = e % m;
            double ipr, ipi;
            crecip(Ar[j * LD + j], Ai[j * LD + j], ipr, ipi);
            const double ar = Ar[k * LD + j], ai = Ai[k * LD + j];
            const double lr = ar * ipr - ai * ipi, li = ar * ipi + ai * ipr;
            const double yr = Rr[i * LD + k], yi = Ri[i * LD + k];
            Rr[i * LD + j] -= yr * lr - yi * li;
            Ri[i * LD + j] -= yr * li + yi * lr;
        }
        __syncthreads();
    }
}

// MLP of the sub-layer's block: z = b2 + w2 . tanh(W1 f + b1)
struct Mlp {
    double f[kNF], hid[kHid], z;
};

__device__ inline void mlp_fwd(const float *w, Mlp &m)
{
    double z = (double)w[kHid * kNF + 2 * kHid];
#pragma unroll
    for (int k = 0; k < kHid; k++) {
        double a = (double)w[kHid * kNF + k];
#pragma unroll
        for (int q = 0; q < kNF; q++) a += (double)w[k * kNF + q] * m.f[q];
        m.hid[k] = tanh(a);
        z += (double)w[kHid * kNF + kHid + k] * m.hid[k];
    }
    m.z = z;
}

// the three features of the updated link at device site s
__device__ inline void features(const Geom &g, const double *ptr, int mu, int nu, long s, double (&f)[kNF])
{
    const long V = g.V;
    const long y = g.nbr[(2 * nu) * V + s];
    f[0] = ptr[g.nbr[(2 * mu + 1) * V + y]];
    f[1] = ptr[y];
    f[2] = ptr[g.nbr[(2 * mu) * V + y]];
}

// Hermitian matrix of link l from the compact momentum layout -> LDS
template <int NMAX, int NT>
__device__ inline void ld_herm(const Geom &g, double *re, double *im, const double *gr, long l, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
    for (int e = threadIdx.x; e < n * n; e += NT) {
        const int i = e / n, j = e % n;
        double vr, vi;
        if (i == j) {
            vr = gr[moff(g, l, i, n)];
            vi = 0.0;
        } else {
            const int lo = i < j ? i : j, hi = i < j ? j : i;
            const int h = hup(n, lo, hi);
            vr = gr[moff(g, l, h, n)];
            vi = gr[moff(g, l, h + 1, n)];
            if (i > j) vi = -vi;
        }
        re[i * LD + j] = vr;
        im[i * LD + j] = vi;
    }
}

// owned upper-triangle elements of a Hermitian matrix: store / atomic add
// into the compact layout (scaled by s)
template <bool Add, int NMAX, int NT>
__device__ inline void put_herm(const Geom &g, const Own<NMAX, NT> &o, double *gr, long l, int n,
                                const Reg<NMAX, NT> &c, double s)
{
#pragma unroll
    for (int q = 0; q < Own<NMAX, NT>::EPL; q++) {
        if (!o.ok[q] || o.i[q] > o.j[q]) continue;
        const int i = o.i[q], j = o.j[q];
        if (i == j) {
            double *p = gr + moff(g, l, i, n);
            if constexpr (Add) unsafeAtomicAdd(p, s * c.r[q]);
            else *p = s * c.r[q];
        } else {
            const int h = hup(n, i, j);
            double *pr = gr + moff(g, l, h, n), *pi = gr + moff(g, l, h + 1, n);
            if constexpr (Add) {
                unsafeAtomicAdd(pr, s * c.r[q]);
                unsafeAtomicAdd(pi, s * c.i[q]);
            } else {
                *pr = s * c.r[q];
                *pi = s * c.i[q];
            }
        }
    }
}

// AH(M)_ij = (M_ij - conj(M_ji)) / (2i) for the owned elements of the
// matrix held in (re, im)
template <int NMAX, int NT>
__device__ inline void ah_of(const Own<NMAX, NT> &o, const double *re, const double *im, Reg<NMAX, NT> &c)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
#pragma unroll
    for (int q = 0; q < Own<NMAX, NT>::EPL; q++) {
        const int a = o.i[q] * LD + o.j[q], b = o.j[q] * LD + o.i[q];
        c.r[q] = 0.5 * (im[a] + im[b]);
        c.i[q] = -0.5 * (re[a] - re[b]);
    }
}

// ReTr P(y) / N of the plaquettes of row `row` (every L_mu-th of them per b)
template <int NMAX, int NT>
__global__ __launch_bounds__(NT) void k_ptr(Args a, int row)
{
    __shared__ double Ar[Cfg<NMAX, NT>::SZ], Ai[Cfg<NMAX, NT>::SZ], Br[Cfg<NMAX, NT>::SZ],
        Bi[Cfg<NMAX, NT>::SZ];
    __shared__ double red[Cfg<NMAX, NT>::NW * 2];
    const Geom &g = a.g;
    const int n = g.N;
    const long V = g.V;
    const long y = act_site(g, a.mu, row, blockIdx.x);
    Own<NMAX, NT> o;
    o.init(n);
    // X = U_0(y) U_1(y+0), Y = U_1(y) U_0(y+1): ReTr P = Re sum X_ij conj(Y_ij)
    ld<NMAX, NT>(g, Ar, Ai, a.U, 0 * V + y, n);
    ld<NMAX, NT>(g, Br, Bi, a.U, 1 * V + g.nbr[0 * V + y], n);
    __syncthreads();
    Reg<NMAX, NT> X, Y;
    mm<false, false>(o, Ar, Ai, Br, Bi, n, X);
    __syncthreads();
    ld<NMAX, NT>(g, Ar, Ai, a.U, 1 * V + y, n);
    ld<NMAX, NT>(g, Br, Bi, a.U, 0 * V + g.nbr[2 * V + y], n);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, Y);
    double v[1] = {0.0};
#pragma unroll
    for (int q = 0; q < Own<NMAX, NT>::EPL; q++)
        if (o.ok[q]) v[0] += X.r[q] * Y.r[q] + X.i[q] * Y.i[q];
    bsum<NT, 1>(v, red);
    if (threadIdx.x == 0) a.ptr[y] = v[0] / n;
}

template <int NMAX, int NT>
__global__ __launch_bounds__(NT) void k_forward(Args a)
{
    constexpr int LD = Cfg<NMAX, NT>::LD, EPL = Cfg<NMAX, NT>::EPL;
    __shared__ double Ar[Cfg<NMAX, NT>::SZ], Ai[Cfg<NMAX, NT>::SZ], Br[Cfg<NMAX, NT>::SZ],
        Bi[Cfg<NMAX, NT>::SZ];
    __shared__ double red[Cfg<NMAX, NT>::NW * 5];
    const Geom &g = a.g;
    const int n = g.N, mu = a.mu, nu = 1 - mu;
    const long V = g.V, b = blockIdx.x;
    const long s = act_site(g, mu, a.off, b);
    const long l = (long)mu * V + s;
    const long xp = g.nbr[(2 * mu) * V + s], xpn = g.nbr[(2 * nu) * V + s];
    Mlp m;
    features(g, a.ptr, mu, nu, s, m.f);
    mlp_fwd(a.w, m);
    const double t = a.tmax * tanh(m.z);
    Own<NMAX, NT> o;
    o.init(n);
    Reg<NMAX, NT> T, S, W;
    // S = U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
    ld<NMAX, NT>(g, Ar, Ai, a.U, (long)nu * V + xp, n);
    ld<NMAX, NT>(g, Br, Bi, a.U, (long)mu * V + xpn, n);
    __syncthreads();
    mm<false, true>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Ar, Ai, T);
    ld<NMAX, NT>(g, Br, Bi, a.U, (long)nu * V + s, n);
    __syncthreads();
    mm<false, true>(o, Ar, Ai, Br, Bi, n, S);
    __syncthreads();
    st(o, Br, Bi, S);
    ld<NMAX, NT>(g, Ar, Ai, a.U, l, n);
    if (a.log) {
        double *lg = a.log + (size_t)b * 2 * n * n;
        for (int r = threadIdx.x; r < 2 * n * n; r += NT) lg[r] = a.U[loff(g, l, r, n)];
    }
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, W);  // W = U S
    __syncthreads();
    st(o, Ar, Ai, W);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Ar, Ai, n, T);  // W^2
    st(o, Br, Bi, T);
    __syncthreads();
    // Re Tr W, Im Tr W^k (k = 1..4) for the branch of phi
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        if (!o.ok[q]) continue;
        const int tr = o.j[q] * LD + o.i[q];
        if (o.i[q] == o.j[q]) {
            v[0] += W.r[q];
            v[1] += W.i[q];
        }
        v[2] += W.r[q] * Ai[tr] + W.i[q] * Ar[tr];
        v[3] += T.r[q] * Ai[tr] + T.i[q] * Ar[tr];
        v[4] += T.r[q] * Bi[tr] + T.i[q] * Br[tr];
    }
    bsum<NT, 5>(v, red);
    // B = 1 - t W into A, W - t into B, then h = B^{-1} (W - t) into B
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q]) {
            const int e = o.i[q] * LD + o.j[q];
            const double d = o.i[q] == o.j[q] ? 1.0 : 0.0;
            Ar[e] = d - t * W.r[q];
            Ai[e] = -t * W.i[q];
            Br[e] = W.r[q] - t * d;
            Bi[e] = W.i[q];
        }
    __syncthreads();
    lu<NMAX, NT>(Ar, Ai, n);
    solve_left<NMAX, NT>(Ar, Ai, Br, Bi, n);
    double pv[2] = {0.0, 0.0};
    if ((int)threadIdx.x < n) {
        const double pr = Ar[threadIdx.x * LD + threadIdx.x], pi = Ai[threadIdx.x * LD + threadIdx.x];
        pv[0] = 0.5 * log(pr * pr + pi * pi);
        pv[1] = atan2(pi, pr);
    }
    bsum<NT, 2>(pv, red);
    const double t2 = t * t;
    const double est = -(t * v[1] + t2 * v[2] / 2 + t2 * t * v[3] / 3 + t2 * t2 * v[4] / 4);
    const double twopi = 6.283185307179586;
    const double phi = pv[1] + twopi * rint((est - pv[1]) / twopi);
    double cr, ci;
    sincos(2.0 * phi / n, &ci, &cr);
    // U' = e^{2 i phi / N} h S^+
    st(o, Ar, Ai, S);
    __syncthreads();
    mm<false, true>(o, Br, Bi, Ar, Ai, n, T);
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q]) {
            const int e = o.i[q] * n + o.j[q];
            a.U[loff(g, l, 2 * e, n)] = cr * T.r[q] - ci * T.i[q];
            a.U[loff(g, l, 2 * e + 1, n)] = cr * T.i[q] + ci * T.r[q];
        }
    if (threadIdx.x == 0) {
        const double mm_ = 1.0 + t2 - 2.0 * t / n * v[0];
        a.part[2 * b] = (double)n * n * log1p(-t2) - 2.0 * n * pv[0] + log(mm_ / (1.0 - t2));
        a.part[2 * b + 1] = 0.0;
    }
}

template <int NMAX, int NT>
__global__ __launch_bounds__(NT) void k_backward(Args a)
{
    constexpr int LD = Cfg<NMAX, NT>::LD, EPL = Cfg<NMAX, NT>::EPL;
    __shared__ double Ar[Cfg<NMAX, NT>::SZ], Ai[Cfg<NMAX, NT>::SZ], Br[Cfg<NMAX, NT>::SZ],
        Bi[Cfg<NMAX, NT>::SZ];
    __shared__ double red[Cfg<NMAX, NT>::NW * 5];
    const Geom &g = a.g;
    const int n = g.N, mu = a.mu, nu = 1 - mu;
    const long V = g.V, b = blockIdx.x;
    const long s = act_site(g, mu, a.off, b);
    const long l = (long)mu * V + s;
    const long xp = g.nbr[(2 * mu) * V + s], xpn = g.nbr[(2 * nu) * V + s];
    const double *uin = a.log + (size_t)b * 2 * n * n;
    Mlp m;
    features(g, a.ptr, mu, nu, s, m.f);
    mlp_fwd(a.w, m);
    const double th = tanh(m.z), t = a.tmax * th, t2 = t * t;
    Own<NMAX, NT> o;
    o.init(n);
    Reg<NMAX, NT> T, S, W, Gp, EGE, GL;
    // E^+ G' E = U U'^+ G' U' U^+ (U' = the current link, U = its input)
    ld<NMAX, NT>(g, Ar, Ai, a.U, l, n);
    ld_herm<NMAX, NT>(g, Br, Bi, a.gr, l, n);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        Gp.r[q] = Br[o.i[q] * LD + o.j[q]];
        Gp.i[q] = Bi[o.i[q] * LD + o.j[q]];
    }
    mm<true, false>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
    __syncthreads();
    mm<false, false>(o, Br, Bi, Ar, Ai, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
    ld_log<NMAX, NT>(Ar, Ai, uin, n);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
    __syncthreads();
    mm<false, true>(o, Br, Bi, Ar, Ai, n, EGE);
    __syncthreads();
    // S, W of the input configuration
    ld<NMAX, NT>(g, Ar, Ai, a.U, (long)nu * V + xp, n);
    ld<NMAX, NT>(g, Br, Bi, a.U, (long)mu * V + xpn, n);
    __syncthreads();
    mm<false, true>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Ar, Ai, T);
    ld<NMAX, NT>(g, Br, Bi, a.U, (long)nu * V + s, n);
    __syncthreads();
    mm<false, true>(o, Ar, Ai, Br, Bi, n, S);
    __syncthreads();
    st(o, Br, Bi, S);
    ld_log<NMAX, NT>(Ar, Ai, uin, n);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, W);
    __syncthreads();
    st(o, Ar, Ai, W);
    __syncthreads();
    ah_of(o, Ar, Ai, GL);  // AH(W) for now
    double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q] && o.i[q] == o.j[q]) {
            v[0] += W.r[q];
            v[1] += W.i[q];
        }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q]) {
            const int e = o.i[q] * LD + o.j[q];
            const double d = o.i[q] == o.j[q] ? 1.0 : 0.0;
            Ar[e] = d - t * W.r[q];
            Ai[e] = -t * W.i[q];
            Br[e] = W.r[q] - t * d;
            Bi[e] = W.i[q];
        }
    __syncthreads();
    lu<NMAX, NT>(Ar, Ai, n);
    solve_left<NMAX, NT>(Ar, Ai, Br, Bi, n);  // h in B
    // Re Tr h, Im Tr h, Tr(G' AH(h)); g_L
    double tga = 0.0;
    Reg<NMAX, NT> AHh;
    ah_of(o, Br, Bi, AHh);
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q]) {
            tga += Gp.r[q] * AHh.r[q] + Gp.i[q] * AHh.i[q];
            if (o.i[q] == o.j[q]) {
                v[2] += Br[o.i[q] * LD + o.i[q]];
                v[3] += Bi[o.i[q] * LD + o.i[q]];
            }
        }
    {
        double w5[5] = {v[0], v[1], v[2], v[3], tga};
        bsum<NT, 5>(w5, red);
        v[0] = w5[0], v[1] = w5[1], v[2] = w5[2], v[3] = w5[3], tga = w5[4];
    }
    const double mq = 1.0 + t2 - 2.0 * t / n * v[0];
    const double ca = -n * t / (1.0 - t2), cb = t / (n * mq);
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        GL.r[q] = ca * AHh.r[q] + cb * GL.r[q];
        GL.i[q] = ca * AHh.i[q] + cb * GL.i[q];
    }
    const double trgl = ca * v[3] + cb * v[1];
    // g1 = B^{-+} G' B^{-1}: R = G' B^{-1}, then (R^+) B^{-1}
    st(o, Br, Bi, Gp);
    __syncthreads();
    solve_right<NMAX, NT>(Ar, Ai, Br, Bi, n);
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        T.r[q] = Br[o.j[q] * LD + o.i[q]];
        T.i[q] = -Bi[o.j[q] * LD + o.i[q]];
    }
    __syncthreads();
    st(o, Br, Bi, T);
    __syncthreads();
    solve_right<NMAX, NT>(Ar, Ai, Br, Bi, n);
    double tr1[1] = {0.0};
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        T.r[q] = Br[o.i[q] * LD + o.j[q]];
        T.i[q] = Bi[o.i[q] * LD + o.j[q]];
        if (o.ok[q] && o.i[q] == o.j[q]) tr1[0] += T.r[q];
    }
    bsum<NT, 1>(tr1, red);
    const double c = a.c_logj, omt = 1.0 - t2;
    const double trU = (omt * tr1[0] + c * trgl) / n;
    // g_U (T), g_s (EGE <- g_U - E^+ G' E)
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        T.r[q] = omt * T.r[q] + c * GL.r[q] - (o.i[q] == o.j[q] ? trU : 0.0);
        T.i[q] = omt * T.i[q] + c * GL.i[q];
        EGE.r[q] = T.r[q] - EGE.r[q];
        EGE.i[q] = T.i[q] - EGE.i[q];
    }
    put_herm<false>(g, o, a.gr, l, n, T, 1.0);
    // U_nu(x+mu) += U^+ g_s U
    st(o, Br, Bi, EGE);
    ld_log<NMAX, NT>(Ar, Ai, uin, n);
    __syncthreads();
    mm<true, false>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
    __syncthreads();
    mm<false, false>(o, Br, Bi, Ar, Ai, n, T);
    put_herm<true>(g, o, a.gr, (long)nu * V + xp, n, T, 1.0);
    __syncthreads();
    // U_nu(x) += -W^+ g_s W
    st(o, Ar, Ai, W);
    st(o, Br, Bi, EGE);
    __syncthreads();
    mm<true, false>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
    __syncthreads();
    mm<false, false>(o, Br, Bi, Ar, Ai, n, T);
    put_herm<true>(g, o, a.gr, (long)nu * V + s, n, T, -1.0);
    __syncthreads();
    // U_mu(x+nu) += -A^+ g_s A, A = W U_nu(x)
    ld<NMAX, NT>(g, Br, Bi, a.U, (long)nu * V + s, n);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Ar, Ai, T);
    st(o, Br, Bi, EGE);
    __syncthreads();
    mm<true, false>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
    __syncthreads();
    mm<false, false>(o, Br, Bi, Ar, Ai, n, T);
    put_herm<true>(g, o, a.gr, (long)mu * V + xpn, n, T, -1.0);
    // restore the input link
    for (int r = threadIdx.x; r < 2 * n * n; r += NT) a.U[loff(g, l, r, n)] = uin[r];
    // dPhi/dt, MLP backward: weight gradient and dPhi/df of the features
    const double dl = -2.0 * t * n * n / omt + 2.0 * n * (v[2] + n * t) / omt + (2.0 * t - 2.0 / n * v[0]) / mq +
                      2.0 * t / omt;
    const double dt = 4.0 * tga / omt + c * dl;
    const double dz = dt * a.tmax * (1.0 - th * th);
    if (threadIdx.x == 0) {
        double df[kNF] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < kHid; k++) {
            const double dp = dz * (double)a.w[kHid * kNF + kHid + k] * (1.0 - m.hid[k] * m.hid[k]);
#pragma unroll
            for (int q = 0; q < kNF; q++) df[q] += dp * (double)a.w[k * kNF + q];
        }
        const long y = g.nbr[(2 * nu) * V + s];
        unsafeAtomicAdd(a.dF + g.nbr[(2 * mu + 1) * V + y], df[0]);
        unsafeAtomicAdd(a.dF + y, df[1]);
        unsafeAtomicAdd(a.dF + g.nbr[(2 * mu) * V + y], df[2]);
    }
    if (a.gw) {
        const int k = threadIdx.x;
        if (k < kHid) {  // hidden unit k of this thread, recomputed (no dynamic index into m.hid)
            double pre = (double)a.w[kHid * kNF + k];
#pragma unroll
            for (int q = 0; q < kNF; q++) pre += (double)a.w[k * kNF + q] * m.f[q];
            const double hk = tanh(pre);
            const double dp = dz * (double)a.w[kHid * kNF + kHid + k] * (1.0 - hk * hk);
#pragma unroll
            for (int q = 0; q < kNF; q++) unsafeAtomicAdd(a.gw + k * kNF + q, dp * m.f[q]);
            unsafeAtomicAdd(a.gw + kHid * kNF + k, dp);
            unsafeAtomicAdd(a.gw + kHid * kNF + kHid + k, dz * hk);
        }
        if (k == 0) unsafeAtomicAdd(a.gw + kHid * kNF + 2 * kHid, dz);
    }
}

// Feature plaquettes P(y) = a b c^+ d^+ of row `row` (a = U_0(y),
// b = U_1(y+0), c = U_0(y+1), d = U_1(y)): each link += co TA(P_l),
// co = -dPhi/df / (2N), P_l the plaquette rotated to start at the link:
// a b c^+ d^+, b c^+ d^+ a, c b^+ a^+ d, d c b^+ a^+.
// put_ta stores P in (re, im), which nobody may be reading on entry.
template <int NMAX, int NT>
__device__ inline void put_ta(const Geom &g, const Own<NMAX, NT> &o, double *re, double *im, double *red,
                              const Reg<NMAX, NT> &P, double *gr, long l, int n, double co)
{
    constexpr int EPL = Cfg<NMAX, NT>::EPL;
    st(o, re, im, P);
    __syncthreads();
    Reg<NMAX, NT> H;
    ah_of(o, re, im, H);
    double tr[1] = {0.0};
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok[q] && o.i[q] == o.j[q]) tr[0] += P.i[q];
    bsum<NT, 1>(tr, red);  // ends with a barrier: (re, im) free again
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.i[q] == o.j[q]) H.r[q] -= tr[0] / n;
    put_herm<true>(g, o, gr, l, n, H, co);
}

template <int NMAX, int NT>
__global__ __launch_bounds__(NT) void k_featgrad(Args a, int row)
{
    __shared__ double Ar[Cfg<NMAX, NT>::SZ], Ai[Cfg<NMAX, NT>::SZ], Br[Cfg<NMAX, NT>::SZ],
        Bi[Cfg<NMAX, NT>::SZ];
    __shared__ double red[Cfg<NMAX, NT>::NW];
    const Geom &g = a.g;
    const int n = g.N;
    const long V = g.V;
    const long y = act_site(g, a.mu, row, blockIdx.x);
    const double co = -a.dF[y] / (2.0 * n);
    const long la = y, lb = V + g.nbr[y], lc = g.nbr[2 * V + y], ld_ = V + y;
    Own<NMAX, NT> o;
    o.init(n);
    Reg<NMAX, NT> AB, DC, P, Q;
    ld<NMAX, NT>(g, Ar, Ai, a.U, la, n);
    ld<NMAX, NT>(g, Br, Bi, a.U, lb, n);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, AB);
    __syncthreads();
    ld<NMAX, NT>(g, Ar, Ai, a.U, ld_, n);
    ld<NMAX, NT>(g, Br, Bi, a.U, lc, n);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, DC);
    __syncthreads();
    st(o, Ar, Ai, AB);
    st(o, Br, Bi, DC);
    __syncthreads();
    mm<false, true>(o, Ar, Ai, Br, Bi, n, P);  // (ab)(dc)^+
    mm<false, true>(o, Br, Bi, Ar, Ai, n, Q);  // (dc)(ab)^+
    __syncthreads();
    put_ta(g, o, Ar, Ai, red, P, a.gr, la, n, co);
    put_ta(g, o, Ar, Ai, red, Q, a.gr, ld_, n, co);
    // b (dc)^+ a  (B still holds dc)
    ld<NMAX, NT>(g, Ar, Ai, a.U, la, n);
    __syncthreads();
    mm<true, false>(o, Br, Bi, Ar, Ai, n, P);
    __syncthreads();
    st(o, Ar, Ai, P);
    ld<NMAX, NT>(g, Br, Bi, a.U, lb, n);
    __syncthreads();
    mm<false, false>(o, Br, Bi, Ar, Ai, n, P);
    __syncthreads();
    put_ta(g, o, Ar, Ai, red, P, a.gr, lb, n, co);
    // c (ab)^+ d
    st(o, Ar, Ai, AB);
    ld<NMAX, NT>(g, Br, Bi, a.U, ld_, n);
    __syncthreads();
    mm<true, false>(o, Ar, Ai, Br, Bi, n, P);
    __syncthreads();
    st(o, Br, Bi, P);
    ld<NMAX, NT>(g, Ar, Ai, a.U, lc, n);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, P);
    __syncthreads();
    put_ta(g, o, Ar, Ai, red, P, a.gr, lc, n, co);
}

__global__ void k_axpy(double *y, const double *x, double s, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += s * x[i];
}

template <typename F>
hipError_t dispatch(int n, F f)
{
    if (n <= 4) return f.template run<4, 64>();
    if (n <= 8) return f.template run<8, 64>();
    if (n <= 16) return f.template run<16, 64>();
    if (n <= 32) return f.template run<32, 256>();
    return f.template run<64, 512>();
}

struct LaunchPtr {
    const Args &a;
    int row;
    hipStream_t st;
    template <int NMAX, int NT>
    hipError_t run() const
    {
        hipLaunchKernelGGL((k_ptr<NMAX, NT>), dim3((unsigned)(a.g.V / 4)), dim3(NT), 0, st, a, row);
        return hipGetLastError();
    }
};
struct LaunchFwd {
    const Args &a;
    hipStream_t st;
    template <int NMAX, int NT>
    hipError_t run() const
    {
        hipLaunchKernelGGL((k_forward<NMAX, NT>), dim3((unsigned)(a.g.V / 4)), dim3(NT), 0, st, a);
        return hipGetLastError();
    }
};
struct LaunchBwd {
    const Args &a;
    hipStream_t st;
    template <int NMAX, int NT>
    hipError_t run() const
    {
        hipLaunchKernelGGL((k_backward<NMAX, NT>), dim3((unsigned)(a.g.V / 4)), dim3(NT), 0, st, a);
        return hipGetLastError();
    }
};
struct LaunchFeat {
    const Args &a;
    int row;
    hipStream_t st;
    template <int NMAX, int NT>
    hipError_t run() const
    {
        hipLaunchKernelGGL((k_featgrad<NMAX, NT>), dim3((unsigned)(a.g.V / 4)), dim3(NT), 0, st, a, row);
        return hipGetLastError();
    }
};

}  // namespace sflow

long sflow_active(const Geom &g) { return g.V / 4; }

hipError_t launch_sflow_forward(const Geom &g, void *U, int mu, int off, const float *w, double tmax, double *ptr,
                                double *log, double *part, hipStream_t st)
{
    sflow::Args a{};
    a.g = g;
    a.U = (double *)U;
    a.mu = mu;
    a.off = off;
    a.w = w;
    a.tmax = tmax;
    a.ptr = ptr;
    a.log = log;
    a.part = part;
    hipError_t e = sflow::dispatch(g.N, sflow::LaunchPtr{a, (off + 1) % 4, st});
    if (e != hipSuccess) return e;
    return sflow::dispatch(g.N, sflow::LaunchFwd{a, st});
}

hipError_t launch_sflow_backward(const Geom &g, void *U, int mu, int off, const float *w, double tmax, double *ptr,
                                 double *log, double *gr, double c_logj, double *gw, double *dF, hipStream_t st)
{
    sflow::Args a{};
    a.g = g;
    a.U = (double *)U;
    a.mu = mu;
    a.off = off;
    a.w = w;
    a.tmax = tmax;
    a.ptr = ptr;
    a.log = log;
    a.gr = gr;
    a.c_logj = c_logj;
    a.gw = gw;
    a.dF = dF;
    const int row = (off + 1) % 4;
    hipError_t e = hipMemsetAsync(dF, 0, (size_t)g.V * sizeof(double), st);
    if (e != hipSuccess) return e;
    if ((e = sflow::dispatch(g.N, sflow::LaunchPtr{a, row, st})) != hipSuccess) return e;
    if ((e = sflow::dispatch(g.N, sflow::LaunchBwd{a, st})) != hipSuccess) return e;
    return sflow::dispatch(g.N, sflow::LaunchFeat{a, row, st});
}

hipError_t launch_axpy(double *y, const double *x, double s, long n, hipStream_t st)
{
    hipLaunchKernelGGL(sflow::k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, x, s, n);
    return hipGetLastError();
}

}  // namespace askja
