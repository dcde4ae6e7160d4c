// flow_spectral.hip -- the spectral field transformation U = F(V) of SU(N)
// lattices (2 <= N <= 64, any D >= 2), used as the FTHMC proposal map:
// forward with its log-Jacobian, and the hand-written backward of
// Phi = beta S_W(F(V)) - log J(V).
//
// Sub-layer (mu, off), nu = (mu + 1) mod D (nu = 1 - mu in D = 2), updates
// U_mu(x) for x_nu = off (mod 4); every other link is frozen.  Everything the
// map reads lies in the (mu, nu) plane.  For an updated link
//   S  = U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+   (frozen: x+nu lies in row off+1)
//   W  = U_mu(x) S                            (the plaquette through the link)
//   f  = ReTr P_mu,nu(x+nu+k mu) / N, k = -1, 0, 1  (frozen plaquettes of row off+1,
//        P_mu,nu(y) = U_mu(y) U_nu(y+mu) U_mu(y+nu)^+ U_nu(y)^+)
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
//   own link:    g_U = TL[(1 - t^2) B^{-+} G' B^{-1} + c g_L]  (B^{-1} kept
//                from the forward, which forms it by LU and substitution),
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
#include "mfma_impl.hpp"

namespace askja {
namespace sflow {

constexpr int kHid = 32, kNF = 3;

struct Args {
    Geom g;
    double *U;          // field (FP64); forward: updated links written in place;
                        // backward: their inputs restored from `log`
    int mu, nu, off;
    const float *w;     // MLP block
    double tmax;
    double *ptr;        // ReTr P / N per device site (rows off+1)
    double *log;        // the updated links' inputs, 2 N^2 reals each (nullable in the forward)
    double *inv;        // B^{-1} of every updated link, 2 N^2 reals each (nullable in the forward)
    double *part;       // forward: log J per updated link as (s, c) pairs
    double *gr;         // backward: Lie gradient, momentum layout
    double c_logj;
    double *gw;         // backward: weight gradient (nullable)
    double *dF;         // backward: dPhi/df per feature plaquette (device site)
};

// Size classes NMAX >= 16 multiply on the FP64 MFMA units (mfma::cmul, the
// kick's product: planar slots with row stride NMAX + 2, its workgroup
// shape NT); the smaller ones on the VALU.
template <int NMAX, int NT>
struct Cfg {
    static constexpr bool MF = NMAX >= 16;
    static_assert(!MF || NT == mfma::Cfg<mfma::kF64, NMAX>::NT, "MFMA size class: workgroup shape");
    static constexpr int LD = NMAX + 2;
    static constexpr int SZ = NMAX * LD;
    static constexpr int EPL = (NMAX * NMAX + NT - 1) / NT;
    static constexpr int NW = NT / 64;
};

__device__ inline size_t loff(const Geom &g, long l, int r, int n) { return elem_off(l, r, 2 * n * n, g.tile); }
__device__ inline size_t moff(const Geom &g, long l, int e, int n) { return elem_off(l, e, n * n, g.tile); }

// device site of active index b: the V / L_nu sites of one nu-slice in
// lexicographic order of the other coordinates (j = b mod (V / L_nu)), the
// slices x_nu = row + 4 (b / (V / L_nu))
__device__ inline long act_site(const Geom &g, int nu, int row, long b)
{
    const long m = g.V / g.ext[nu];
    long j = b % m, lex = (row + 4 * (b / m)) * g.helparr[nu];
    for (int d = g.D - 1; d >= 0; d--) {
        if (d == nu) continue;
        lex += (j % g.ext[d]) * g.helparr[d];
        j /= g.ext[d];
    }
    return g.lex2dev[lex];
}

template <int NMAX, int NT>
struct Own {
    static constexpr bool MF = Cfg<NMAX, NT>::MF;
    static constexpr int RS = NT / NMAX;
    static constexpr int EPL = Cfg<NMAX, NT>::EPL;
    static_assert(MF ? EPL == 4 * mfma::Cfg<mfma::kF64, NMAX>::NCT : (NT % NMAX == 0 && EPL == (NMAX + RS - 1) / RS),
                  "size class");
    int r0, c, n;
    __device__ inline void init(int n_)
    {
        r0 = threadIdx.x / NMAX;
        c = threadIdx.x % NMAX;
        n = n_;
    }
    __device__ inline int ri(int q) const { return MF ? mfma::acc_row<mfma::kF64, NMAX>(q & 3) : r0 + RS * q; }
    __device__ inline int rj(int q) const { return MF ? mfma::acc_col<mfma::kF64, NMAX>(q >> 2) : c; }
    __device__ inline bool ok(int q) const { return ri(q) < n && rj(q) < n; }
    __device__ inline int i(int q) const { return ok(q) ? ri(q) : 0; }
    __device__ inline int j(int q) const { return ok(q) ? rj(q) : 0; }
};

template <int NMAX, int NT>
struct Reg {
    double r[Cfg<NMAX, NT>::EPL], i[Cfg<NMAX, NT>::EPL];
};

// global link -> LDS matrix.  Every load of the thread is issued before its
// first LDS store (a compile-time count per thread), so that the loads
// overlap instead of each waiting for the previous store.
template <int NMAX, int NT>
__device__ inline void ld(const Geom &g, double *re, double *im, const double *f, long l, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD, K = (2 * NMAX * NMAX + NT - 1) / NT;
    const int nr = 2 * n * n;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = threadIdx.x + k * NT;
        v[k] = r < nr ? f[loff(g, l, r, n)] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = threadIdx.x + k * NT;
        if (r < nr) {
            const int e = r >> 1;
            (r & 1 ? im : re)[(e / n) * LD + e % n] = v[k];
        }
    }
}

// contiguous interleaved link (the input log) -> LDS matrix
template <int NMAX, int NT>
__device__ inline void ld_log(double *re, double *im, const double *b, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD, K = (2 * NMAX * NMAX + NT - 1) / NT;
    const int nr = 2 * n * n;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = threadIdx.x + k * NT;
        v[k] = r < nr ? b[r] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = threadIdx.x + k * NT;
        if (r < nr) {
            const int e = r >> 1;
            (r & 1 ? im : re)[(e / n) * LD + e % n] = v[k];
        }
    }
}

// global -> global copy of one link's 2 N^2 reals (src in the field layout
// when FL, dst contiguous; or the reverse)
template <int NMAX, int NT, bool FromField>
__device__ inline void copy_link(const Geom &g, double *dst, const double *src, long l, int n)
{
    constexpr int K = (2 * NMAX * NMAX + NT - 1) / NT;
    const int nr = 2 * n * n;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = threadIdx.x + k * NT;
        v[k] = r < nr ? src[FromField ? loff(g, l, r, n) : (size_t)r] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = threadIdx.x + k * NT;
        if (r < nr) dst[FromField ? (size_t)r : loff(g, l, r, n)] = v[k];
    }
}

template <int NMAX, int NT>
__device__ inline void st(const Own<NMAX, NT> &o, double *re, double *im, const Reg<NMAX, NT> &c)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
#pragma unroll
    for (int q = 0; q < Own<NMAX, NT>::EPL; q++)
        if (o.ok(q)) {
            re[o.i(q) * LD + o.j(q)] = c.r[q];
            im[o.i(q) * LD + o.j(q)] = c.i[q];
        }
}

// c = op(A) op(B) on the owned elements (op = ^+ where DA / DB).  MFMA
// classes need the slots zero outside n x n (zero_slots at kernel start:
// nothing writes there afterwards).
template <bool DA, bool DB, int NMAX, int NT>
__device__ inline void mm(const Own<NMAX, NT> &o, const double *ar_, const double *ai_, const double *br_,
                          const double *bi_, int n, Reg<NMAX, NT> &c)
{
    constexpr int LD = Cfg<NMAX, NT>::LD, EPL = Cfg<NMAX, NT>::EPL;
    if constexpr (Cfg<NMAX, NT>::MF) {
        // the kick's stacked product (mfma::cmul_vec) with a shorter unroll:
        // these kernels hold more matrices in registers than the kick
        typedef mfma::f64x4 V4;
        constexpr int NCT = mfma::Cfg<mfma::kF64, NMAX>::NCT;
        V4 accr[NCT], acci[NCT];
#pragma unroll
        for (int t = 0; t < NCT; t++) accr[t] = acci[t] = V4{0.0, 0.0, 0.0, 0.0};
        const int lane = threadIdx.x & 63;
        const int row = 16 * mfma::wave_rowblock<mfma::kF64, NMAX>() + (lane & 15);
        const int t0 = mfma::wave_tile0<mfma::kF64, NMAX>();
#pragma unroll 1
        for (int ks = 0; ks < NMAX / 4; ks++) {
            const int k = 4 * ks + (lane >> 4);
            const double ar = DA ? ar_[k * LD + row] : ar_[row * LD + k];
            const double ai = DA ? -ai_[k * LD + row] : ai_[row * LD + k];
            const double nai = -ai;
#pragma unroll
            for (int t = 0; t < NCT; t++) {
                const int col = 16 * (t0 + t) + (lane & 15);
                const double br = DB ? br_[col * LD + k] : br_[k * LD + col];
                const double bi = DB ? -bi_[col * LD + k] : bi_[k * LD + col];
                accr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, accr[t], 0, 0, 0);
                acci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acci[t], 0, 0, 0);
                accr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bi, accr[t], 0, 0, 0);
                acci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, acci[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int q = 0; q < EPL; q++) {
            c.r[q] = accr[q >> 2][q & 3];
            c.i[q] = acci[q >> 2][q & 3];
        }
    } else {
#pragma unroll
        for (int q = 0; q < EPL; q++) c.r[q] = c.i[q] = 0.0;
        const int jc = o.c < n ? o.c : 0;
        for (int k = 0; k < n; k++) {
            const int ib = DB ? jc * LD + k : k * LD + jc;
            const double br = br_[ib], bi = DB ? -bi_[ib] : bi_[ib];
#pragma unroll
            for (int q = 0; q < EPL; q++) {
                const int ia = DA ? k * LD + o.i(q) : o.i(q) * LD + k;
                const double ar = ar_[ia], ai = DA ? -ai_[ia] : ai_[ia];
                c.r[q] = fma(ar, br, c.r[q]);
                c.r[q] = fma(-ai, bi, c.r[q]);
                c.i[q] = fma(ar, bi, c.i[q]);
                c.i[q] = fma(ai, br, c.i[q]);
            }
        }
    }
}

// zero both operand slots (MFMA classes: the padding outside n x n must be 0)
template <int NMAX, int NT>
__device__ inline void zero_slots(double *Ar, double *Ai, double *Br, double *Bi)
{
    if constexpr (Cfg<NMAX, NT>::MF) {
        for (int e = threadIdx.x; e < Cfg<NMAX, NT>::SZ; e += NT) Ar[e] = Ai[e] = Br[e] = Bi[e] = 0.0;
        __syncthreads();
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

// Thread map of the factorisation and the triangular solves: thread tid
// works on column c = tid % NMAX and rows r0 + RS q (r0 = tid / NMAX,
// q < RP), so that a wavefront covers whole rows (row-k and column-k reads are
// broadcasts or contiguous) and no index needs a division by a runtime value.
template <int NMAX, int NT>
struct Tri {
    static constexpr int RS = NT / NMAX;                 // rows per pass (NT >= NMAX in every size class)
    static constexpr int RP = (NMAX + RS - 1) / RS;      // passes
    static_assert(NT % NMAX == 0, "workgroup not a multiple of the matrix size class");
    static __device__ inline int col() { return threadIdx.x % NMAX; }
    static __device__ inline int row(int q) { return threadIdx.x / NMAX + RS * q; }
};

// In-place LU of the matrix in (Ar, Ai) without pivoting, the multipliers
// left unscaled: L_ik = A_ik / A_kk (i > k), Ub_kj = A_kj (j >= k); then the
// pivots' reciprocals into (ipr, ipi) (LDS, NMAX each).
template <int NMAX, int NT>
__device__ inline void lu(double *Ar, double *Ai, double *ipr, double *ipi, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
    typedef Tri<NMAX, NT> M;
    const int c = M::col();
    for (int k = 0; k < n - 1; k++) {
        if (c > k && c < n) {
            double pr, pi;
            crecip(Ar[k * LD + k], Ai[k * LD + k], pr, pi);
            const double ur = Ar[k * LD + c], ui = Ai[k * LD + c];
#pragma unroll
            for (int q = 0; q < M::RP; q++) {
                const int r = M::row(q);
                if (r > k && r < n) {
                    const double ar = Ar[r * LD + k], ai = Ai[r * LD + k];
                    const double lr = ar * pr - ai * pi, li = ar * pi + ai * pr;
                    Ar[r * LD + c] -= lr * ur - li * ui;
                    Ai[r * LD + c] -= lr * ui + li * ur;
                }
            }
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < n) crecip(Ar[threadIdx.x * LD + threadIdx.x], Ai[threadIdx.x * LD + threadIdx.x],
                                     ipr[threadIdx.x], ipi[threadIdx.x]);
    __syncthreads();
}

// (Br, Bi) <- A^{-1} (Br, Bi) with A = L Ub from lu()
template <int NMAX, int NT>
__device__ inline void solve_left(const double *Ar, const double *Ai, const double *ipr, const double *ipi,
                                  double *Br, double *Bi, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD;
    typedef Tri<NMAX, NT> M;
    const int c = M::col();
    for (int k = 0; k < n - 1; k++) {  // unit lower L: rows r > k -= L_rk B_k
        if (c < n) {
            const double br = Br[k * LD + c], bi = Bi[k * LD + c];
            const double pr = ipr[k], pi = ipi[k];
#pragma unroll
            for (int q = 0; q < M::RP; q++) {
                const int r = M::row(q);
                if (r > k && r < n) {
                    const double ar = Ar[r * LD + k], ai = Ai[r * LD + k];
                    const double lr = ar * pr - ai * pi, li = ar * pi + ai * pr;
                    Br[r * LD + c] -= lr * br - li * bi;
                    Bi[r * LD + c] -= lr * bi + li * br;
                }
            }
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 1; k--) {  // upper Ub: rows r < k -= Ub_rk x_k, x_k = B_k / Ub_kk
        if (c < n) {
            const double b0 = Br[k * LD + c], b1 = Bi[k * LD + c];
            const double xr = b0 * ipr[k] - b1 * ipi[k], xi = b0 * ipi[k] + b1 * ipr[k];
#pragma unroll
            for (int q = 0; q < M::RP; q++) {
                const int r = M::row(q);
                if (r < k) {
                    const double ur = Ar[r * LD + k], ui = Ai[r * LD + k];
                    Br[r * LD + c] -= ur * xr - ui * xi;
                    Bi[r * LD + c] -= ur * xi + ui * xr;
                }
            }
        }
        __syncthreads();
    }
    if (c < n) {
#pragma unroll
        for (int q = 0; q < M::RP; q++) {
            const int r = M::row(q);
            if (r < n) {
                const double b0 = Br[r * LD + c], b1 = Bi[r * LD + c];
                Br[r * LD + c] = b0 * ipr[r] - b1 * ipi[r];
                Bi[r * LD + c] = b0 * ipi[r] + b1 * ipr[r];
            }
        }
    }
    __syncthreads();
}

// MLP of the sub-layer's block: z = b2 + w2 . tanh(W1 f + b1)
struct Mlp {
    double f[kNF], z;
};

__device__ inline void mlp_fwd(const float *w, Mlp &m)
{
    double z = (double)w[kHid * kNF + 2 * kHid];
#pragma unroll
    for (int k = 0; k < kHid; k++) {
        double a = (double)w[kHid * kNF + k];
#pragma unroll
        for (int q = 0; q < kNF; q++) a += (double)w[k * kNF + q] * m.f[q];
        z += (double)w[kHid * kNF + kHid + k] * tanh(a);
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

// Hermitian matrix of link l from the compact momentum layout -> LDS (the
// loads of the thread issued first, as in ld)
template <int NMAX, int NT>
__device__ inline void ld_herm(const Geom &g, double *re, double *im, const double *gr, long l, int n)
{
    constexpr int LD = Cfg<NMAX, NT>::LD, K = (NMAX * NMAX + NT - 1) / NT;
    double vr[K], vi[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = threadIdx.x + k * NT;
        vr[k] = vi[k] = 0.0;
        if (e < n * n) {
            const int i = e / n, j = e % n;
            if (i == j) {
                vr[k] = gr[moff(g, l, i, n)];
            } else {
                const int lo = i < j ? i : j, hi = i < j ? j : i;
                const int h = hup(n, lo, hi);
                vr[k] = gr[moff(g, l, h, n)];
                vi[k] = i > j ? -gr[moff(g, l, h + 1, n)] : gr[moff(g, l, h + 1, n)];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = threadIdx.x + k * NT;
        if (e < n * n) {
            re[(e / n) * LD + e % n] = vr[k];
            im[(e / n) * LD + e % n] = vi[k];
        }
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
        if (!o.ok(q) || o.i(q) > o.j(q)) continue;
        const int i = o.i(q), j = o.j(q);
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
        const int a = o.i(q) * LD + o.j(q), b = o.j(q) * LD + o.i(q);
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
    const int mu = a.mu, nu = a.nu;
    const long y = act_site(g, nu, row, blockIdx.x);
    Own<NMAX, NT> o;
    o.init(n);
    zero_slots<NMAX, NT>(Ar, Ai, Br, Bi);
    // X = U_mu(y) U_nu(y+mu), Y = U_nu(y) U_mu(y+nu): ReTr P = Re sum X_ij conj(Y_ij)
    ld<NMAX, NT>(g, Ar, Ai, a.U, (long)mu * V + y, n);
    ld<NMAX, NT>(g, Br, Bi, a.U, (long)nu * V + g.nbr[(2 * mu) * V + y], n);
    __syncthreads();
    Reg<NMAX, NT> X, Y;
    mm<false, false>(o, Ar, Ai, Br, Bi, n, X);
    __syncthreads();
    ld<NMAX, NT>(g, Ar, Ai, a.U, (long)nu * V + y, n);
    ld<NMAX, NT>(g, Br, Bi, a.U, (long)mu * V + g.nbr[(2 * nu) * V + y], n);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, Y);
    double v[1] = {0.0};
#pragma unroll
    for (int q = 0; q < Own<NMAX, NT>::EPL; q++)
        if (o.ok(q)) v[0] += X.r[q] * Y.r[q] + X.i[q] * Y.i[q];
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
    __shared__ double Pr[NMAX], Pi[NMAX];  // pivot reciprocals of the LU
    const Geom &g = a.g;
    const int n = g.N, mu = a.mu, nu = a.nu;
    const long V = g.V, b = blockIdx.x;
    const long s = act_site(g, nu, a.off, b);
    const long l = (long)mu * V + s;
    const long xp = g.nbr[(2 * mu) * V + s], xpn = g.nbr[(2 * nu) * V + s];
    Mlp m;
    features(g, a.ptr, mu, nu, s, m.f);
    mlp_fwd(a.w, m);
    const double t = a.tmax * tanh(m.z);
    Own<NMAX, NT> o;
    o.init(n);
    zero_slots<NMAX, NT>(Ar, Ai, Br, Bi);
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
    if (a.log) copy_link<NMAX, NT, true>(g, a.log + (size_t)b * 2 * n * n, a.U, l, n);
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
        if (!o.ok(q)) continue;
        const int tr = o.j(q) * LD + o.i(q);
        if (o.i(q) == o.j(q)) {
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
        if (o.ok(q)) {
            const int e = o.i(q) * LD + o.j(q);
            const double d = o.i(q) == o.j(q) ? 1.0 : 0.0;
            Ar[e] = d - t * W.r[q];
            Ai[e] = -t * W.i[q];
            Br[e] = d;
            Bi[e] = 0.0;
        }
    __syncthreads();
    lu<NMAX, NT>(Ar, Ai, Pr, Pi, n);
    solve_left<NMAX, NT>(Ar, Ai, Pr, Pi, Br, Bi, n);  // B^{-1} in slot B
    if (a.inv) {  // kept for the backward
        double *iv = a.inv + (size_t)b * 2 * n * n;
        for (int e = threadIdx.x; e < n * n; e += NT) {
            iv[2 * e] = Br[(e / n) * LD + e % n];
            iv[2 * e + 1] = Bi[(e / n) * LD + e % n];
        }
    }
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
    // h = B^{-1} (W - t), then U' = e^{2 i phi / N} h S^+
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok(q)) {
            const int e = o.i(q) * LD + o.j(q);
            Ar[e] = W.r[q] - (o.i(q) == o.j(q) ? t : 0.0);
            Ai[e] = W.i[q];
        }
    __syncthreads();
    mm<false, false>(o, Br, Bi, Ar, Ai, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
    st(o, Ar, Ai, S);
    __syncthreads();
    mm<false, true>(o, Br, Bi, Ar, Ai, n, T);
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok(q)) {
            const int e = o.i(q) * n + o.j(q);
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
    __shared__ double Pr[NMAX], Pi[NMAX];  // pivot reciprocals of the LU
    const Geom &g = a.g;
    const int n = g.N, mu = a.mu, nu = a.nu;
    const long V = g.V, b = blockIdx.x;
    const long s = act_site(g, nu, a.off, b);
    const long l = (long)mu * V + s;
    const long xp = g.nbr[(2 * mu) * V + s], xpn = g.nbr[(2 * nu) * V + s];
    const double *uin = a.log + (size_t)b * 2 * n * n;
    Mlp m;
    features(g, a.ptr, mu, nu, s, m.f);
    mlp_fwd(a.w, m);
    const double th = tanh(m.z), t = a.tmax * th, t2 = t * t;
    Own<NMAX, NT> o;
    o.init(n);
    zero_slots<NMAX, NT>(Ar, Ai, Br, Bi);
    Reg<NMAX, NT> T, W, X, GL;
    // S = U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ and W = U S of the input configuration
    ld<NMAX, NT>(g, Ar, Ai, a.U, (long)nu * V + xp, n);
    ld<NMAX, NT>(g, Br, Bi, a.U, (long)mu * V + xpn, n);
    __syncthreads();
    mm<false, true>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Ar, Ai, T);
    ld<NMAX, NT>(g, Br, Bi, a.U, (long)nu * V + s, n);
    __syncthreads();
    mm<false, true>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
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
        if (o.ok(q) && o.i(q) == o.j(q)) {
            v[0] += W.r[q];
            v[1] += W.i[q];
        }
    __syncthreads();
    // h = B^{-1} (W - t) with B^{-1} kept by the forward (slot A)
    ld_log<NMAX, NT>(Ar, Ai, a.inv + (size_t)b * 2 * n * n, n);
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok(q)) {
            const int e = o.i(q) * LD + o.j(q);
            Br[e] = W.r[q] - (o.i(q) == o.j(q) ? t : 0.0);
            Bi[e] = W.i[q];
        }
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, X);
    __syncthreads();
    st(o, Br, Bi, X);
    __syncthreads();
    // AH(h), Re Tr h, Im Tr h; then G' into B and Tr(G' AH(h))
    ah_of(o, Br, Bi, X);
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok(q) && o.i(q) == o.j(q)) {
            v[2] += Br[o.i(q) * LD + o.i(q)];
            v[3] += Bi[o.i(q) * LD + o.i(q)];
        }
    __syncthreads();
    ld_herm<NMAX, NT>(g, Br, Bi, a.gr, l, n);
    __syncthreads();
    double tga = 0.0;
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok(q)) tga += Br[o.i(q) * LD + o.j(q)] * X.r[q] + Bi[o.i(q) * LD + o.j(q)] * X.i[q];
    {
        double w5[5] = {v[0], v[1], v[2], v[3], tga};
        bsum<NT, 5>(w5, red);
        v[0] = w5[0], v[1] = w5[1], v[2] = w5[2], v[3] = w5[3], tga = w5[4];
    }
    const double mq = 1.0 + t2 - 2.0 * t / n * v[0];
    const double ca = -n * t / (1.0 - t2), cb = t / (n * mq);
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        GL.r[q] = ca * X.r[q] + cb * GL.r[q];
        GL.i[q] = ca * X.i[q] + cb * GL.i[q];
    }
    const double trgl = ca * v[3] + cb * v[1];
    // g1 = B^{-+} (G' B^{-1})
    mm<false, false>(o, Br, Bi, Ar, Ai, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
    __syncthreads();
    mm<true, false>(o, Ar, Ai, Br, Bi, n, T);
    double tr1[1] = {0.0};
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.ok(q) && o.i(q) == o.j(q)) tr1[0] += T.r[q];
    bsum<NT, 1>(tr1, red);
    const double c = a.c_logj, omt = 1.0 - t2;
    const double trU = (omt * tr1[0] + c * trgl) / n;
    // g_U in T
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        T.r[q] = omt * T.r[q] + c * GL.r[q] - (o.i(q) == o.j(q) ? trU : 0.0);
        T.i[q] = omt * T.i[q] + c * GL.i[q];
    }
    // E^+ G' E = U U'^+ G' U' U^+ (U' = the current link, U = its input), then
    // g_s = g_U - E^+ G' E into X
    ld<NMAX, NT>(g, Ar, Ai, a.U, l, n);
    ld_herm<NMAX, NT>(g, Br, Bi, a.gr, l, n);
    __syncthreads();
    mm<true, false>(o, Ar, Ai, Br, Bi, n, X);
    __syncthreads();
    st(o, Br, Bi, X);
    __syncthreads();
    mm<false, false>(o, Br, Bi, Ar, Ai, n, X);
    __syncthreads();
    st(o, Br, Bi, X);
    ld_log<NMAX, NT>(Ar, Ai, uin, n);
    __syncthreads();
    mm<false, false>(o, Ar, Ai, Br, Bi, n, X);
    __syncthreads();
    st(o, Br, Bi, X);
    __syncthreads();
    mm<false, true>(o, Br, Bi, Ar, Ai, n, X);
#pragma unroll
    for (int q = 0; q < EPL; q++) {
        X.r[q] = T.r[q] - X.r[q];
        X.i[q] = T.i[q] - X.i[q];
    }
    put_herm<false>(g, o, a.gr, l, n, T, 1.0);
    __syncthreads();
    // U_nu(x+mu) += U^+ g_s U
    st(o, Br, Bi, X);
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
    st(o, Br, Bi, X);
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
    st(o, Br, Bi, X);
    __syncthreads();
    mm<true, false>(o, Ar, Ai, Br, Bi, n, T);
    __syncthreads();
    st(o, Br, Bi, T);
    __syncthreads();
    mm<false, false>(o, Br, Bi, Ar, Ai, n, T);
    put_herm<true>(g, o, a.gr, (long)mu * V + xpn, n, T, -1.0);
    // restore the input link
    copy_link<NMAX, NT, false>(g, a.U, uin, l, n);
    // dPhi/dt, MLP backward: weight gradient and dPhi/df of the features
    const double dl = -2.0 * t * n * n / omt + 2.0 * n * (v[2] + n * t) / omt + (2.0 * t - 2.0 / n * v[0]) / mq +
                      2.0 * t / omt;
    const double dt = 4.0 * tga / omt + c * dl;
    const double dz = dt * a.tmax * (1.0 - th * th);
    if (threadIdx.x == 0) {
        double df[kNF] = {0.0, 0.0, 0.0};
#pragma unroll 1
        for (int k = 0; k < kHid; k++) {  // hidden units recomputed (no live array of them)
            double pre = (double)a.w[kHid * kNF + k];
#pragma unroll
            for (int q = 0; q < kNF; q++) pre += (double)a.w[k * kNF + q] * m.f[q];
            const double hk = tanh(pre);
            const double dp = dz * (double)a.w[kHid * kNF + kHid + k] * (1.0 - hk * hk);
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

// Feature plaquettes P(y) = a b c^+ d^+ of row `row` (a = U_mu(y),
// b = U_nu(y+mu), c = U_mu(y+nu), d = U_nu(y)): each link += co TA(P_l),
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
        if (o.ok(q) && o.i(q) == o.j(q)) tr[0] += P.i[q];
    bsum<NT, 1>(tr, red);  // ends with a barrier: (re, im) free again
#pragma unroll
    for (int q = 0; q < EPL; q++)
        if (o.i(q) == o.j(q)) H.r[q] -= tr[0] / n;
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
    const int mu = a.mu, nu = a.nu;
    const long y = act_site(g, nu, row, blockIdx.x);
    const double co = -a.dF[y] / (2.0 * n);
    const long la = (long)mu * V + y, lb = (long)nu * V + g.nbr[(2 * mu) * V + y];
    const long lc = (long)mu * V + g.nbr[(2 * nu) * V + y], ld_ = (long)nu * V + y;
    Own<NMAX, NT> o;
    o.init(n);
    zero_slots<NMAX, NT>(Ar, Ai, Br, Bi);
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
    if (n <= 32) return f.template run<32, 128>();
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
                                double *log, double *inv, double *part, hipStream_t st)
{
    sflow::Args a{};
    a.g = g;
    a.U = (double *)U;
    a.mu = mu;
    a.nu = (mu + 1) % g.D;
    a.off = off;
    a.w = w;
    a.tmax = tmax;
    a.ptr = ptr;
    a.log = log;
    a.inv = inv;
    a.part = part;
    hipError_t e = sflow::dispatch(g.N, sflow::LaunchPtr{a, (off + 1) % 4, st});
    if (e != hipSuccess) return e;
    return sflow::dispatch(g.N, sflow::LaunchFwd{a, st});
}

hipError_t launch_sflow_backward(const Geom &g, void *U, int mu, int off, const float *w, double tmax, double *ptr,
                                 double *log, double *inv, double *gr, double c_logj, double *gw, double *dF,
                                 hipStream_t st)
{
    sflow::Args a{};
    a.g = g;
    a.U = (double *)U;
    a.mu = mu;
    a.nu = (mu + 1) % g.D;
    a.off = off;
    a.w = w;
    a.tmax = tmax;
    a.ptr = ptr;
    a.log = log;
    a.inv = inv;
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
