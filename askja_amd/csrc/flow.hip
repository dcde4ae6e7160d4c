// flow.hip -- the learned field transformation U = F(V) used as an HMC
// proposal map (field-transformation HMC), for SU(3).
//
// F is a stack of masked sub-layers.  Sub-layer k updates the links of one
// direction mu_k on the sites of one parity p_k; every other link is frozen.
// For an updated link l = (x, mu):
//   S      = sum_{nu != mu} [U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
//                            + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)]  (frozen staples)
//   W      = U_l S                                  (the staple loop through l)
//   P_j    = plaquettes at x in the planes not containing mu (frozen)
//   R_j    = the 2x1 and 1x2 rectangles at x in those planes (frozen)
//   Lam    = TAH(sum_j P_j),  Lhat = Lam / sqrt(2 Tr Lam^2 + delta^2)
//   rho    = rho_max tanh(MLP(ReTr P_j / 3, ReTr R_j / 3))   (2 layers, hidden 32)
//   theta  = rho Im Tr(Lhat W)
//   U'_l   = exp(i theta Lhat) U_l
// Lhat, rho and the staples depend on frozen links only, so the Jacobian of
// the sub-layer is the rank-one update of a left translation along Lhat:
//   det J = 1 + d theta / dt (e^{i t Lhat} U_l) = 1 + rho ReTr(Lhat^2 W),
// positive for |rho| < 1/3 (Tr Lhat^2 <= 1/2, ||W|| <= 2(D-1) = 6 at D=4), and
// the sub-layer is a diffeomorphism (on every orbit t -> t + theta(t) is
// monotone).  Gauge equivariance: Lhat and W transform covariantly at x,
// theta and the features are invariant.
//
// The backward kernel propagates Lie-algebra gradients g_l = sum_a D^a Phi
// T^a (Hermitian traceless, D^a = d/dt at e^{i t T^a} U_l) of
// Phi = beta S_W(F(V)) - log J(V) from the output links to the input links,
// recomputing every sub-layer's forward quantities from its stored input.
#include "kernels.hpp"
#include "common.hpp"

namespace askja {
namespace flow {

constexpr int kB = 256;
constexpr int kHid = 32;

struct C3 {
    double r[3][3], i[3][3];
};

__device__ inline void zero(C3 &a)
{
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) a.r[p][q] = a.i[p][q] = 0.0;
}

__device__ inline void eye(C3 &a)
{
    zero(a);
#pragma unroll
    for (int p = 0; p < 3; p++) a.r[p][p] = 1.0;
}

// C = op(A) op(B)
template <bool DA, bool DB>
__device__ inline C3 mm(const C3 &A, const C3 &B)
{
    C3 C;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            double cr = 0.0, ci = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double ar = DA ? A.r[k][p] : A.r[p][k], ai = DA ? -A.i[k][p] : A.i[p][k];
                const double br = DB ? B.r[q][k] : B.r[k][q], bi = DB ? -B.i[q][k] : B.i[k][q];
                cr = fma(ar, br, cr);
                cr = fma(-ai, bi, cr);
                ci = fma(ar, bi, ci);
                ci = fma(ai, br, ci);
            }
            C.r[p][q] = cr;
            C.i[p][q] = ci;
        }
    return C;
}

__device__ inline void addto(C3 &a, const C3 &b, double s = 1.0)
{
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            a.r[p][q] += s * b.r[p][q];
            a.i[p][q] += s * b.i[p][q];
        }
}

// multiply by a complex scalar (cr + i ci)
__device__ inline C3 cscale(const C3 &a, double cr, double ci)
{
    C3 b;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            b.r[p][q] = a.r[p][q] * cr - a.i[p][q] * ci;
            b.i[p][q] = a.r[p][q] * ci + a.i[p][q] * cr;
        }
    return b;
}

__device__ inline double retr(const C3 &a) { return a.r[0][0] + a.r[1][1] + a.r[2][2]; }
__device__ inline double imtr(const C3 &a) { return a.i[0][0] + a.i[1][1] + a.i[2][2]; }

// traceless Hermitian part of (Z + Z^+)/2
__device__ inline C3 proj(const C3 &z)
{
    C3 h;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            h.r[p][q] = 0.5 * (z.r[p][q] + z.r[q][p]);
            h.i[p][q] = 0.5 * (z.i[p][q] - z.i[q][p]);
        }
    const double t = (h.r[0][0] + h.r[1][1] + h.r[2][2]) / 3.0;
#pragma unroll
    for (int p = 0; p < 3; p++) h.r[p][p] -= t;
    return h;
}

// TAH(C) = -i/2 [(C - C^+) - Tr(C - C^+)/3]
__device__ inline C3 tah(const C3 &c)
{
    C3 d;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            d.r[p][q] = c.r[p][q] - c.r[q][p];
            d.i[p][q] = c.i[p][q] + c.i[q][p];
        }
    const double ti = (d.i[0][0] + d.i[1][1] + d.i[2][2]) / 3.0;
#pragma unroll
    for (int p = 0; p < 3; p++) d.i[p][p] -= ti;
    return cscale(d, 0.0, -0.5);
}

__device__ inline double fnorm2(const C3 &a)
{
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) s += a.r[p][q] * a.r[p][q] + a.i[p][q] * a.i[p][q];
    return s;
}

// exp(i X) for Hermitian X by Taylor series to convergence (||X|| < ~3)
__device__ inline C3 expi(const C3 &X)
{
    C3 E, T;
    eye(E);
    eye(T);
    for (int k = 1; k < 48; k++) {
        T = mm<false, false>(T, X);
        T = cscale(T, 0.0, 1.0 / k);  // T <- T (iX) / k
        addto(E, T);
        if (fnorm2(T) < 1e-36) break;
    }
    return E;
}

// sum_k (-i)^k/(k+1)! ad_X^k(G) = int_0^1 e^{-isX} G e^{isX} ds
__device__ inline C3 avg_ad(const C3 &X, const C3 &G)
{
    C3 S = G, T = G;
    double fac = 1.0;
    for (int k = 1; k < 48; k++) {
        C3 c = mm<false, false>(X, T);
        addto(c, mm<false, false>(T, X), -1.0);
        T = cscale(c, 0.0, -1.0);  // T <- -i [X, T]
        fac /= (k + 1);
        addto(S, T, fac);
        if (fnorm2(T) * fac * fac < 1e-36) break;
    }
    return S;
}

// ---- field access (N = 3, links tiled by 64) -----------------------------
template <typename R>
__device__ inline C3 ld(const R *U, long l)
{
    const R *b = U + (size_t)(l >> 6) * (18 * 64) + (l & 63);
    C3 m;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            m.r[p][q] = (double)b[((p * 3 + q) * 2) * 64];
            m.i[p][q] = (double)b[((p * 3 + q) * 2 + 1) * 64];
        }
    return m;
}

template <typename R>
__device__ inline void st(R *U, long l, const C3 &m)
{
    R *b = U + (size_t)(l >> 6) * (18 * 64) + (l & 63);
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            b[((p * 3 + q) * 2) * 64] = (R)m.r[p][q];
            b[((p * 3 + q) * 2 + 1) * 64] = (R)m.i[p][q];
        }
}

// ---- stout sub-layer: the su(3) Jacobian -------------------------------------
// The stout update of link l is U' = exp(iQ) U with Q = -rho TAH(W), W = U S
// (S the frozen staple sum, rho from the MLP of frozen features).  In
// left-invariant coordinates x_a = 2 Tr(T^a X) (Tr T^a T^b = delta_ab / 2)
//   J = Ad_{exp(iQ)} (1 + g(ad_{iQ}) M),  g(z) = (1 - e^{-z}) / z,
//   M X = -rho TAH(i X W)  (the change of Q when U -> e^{iX} U),
// and det Ad = 1, so log det J = log det(1 + g(ad_{iQ}) M).  Everything is
// evaluated in the eigenbasis of Q = V diag(lambda) V^+ (the hat frame,
// Y^ = V^+ Y V): there g(ad_{iQ}) multiplies element (a, b) by
// g(i (lambda_a - lambda_b)), and M^ X^ = -rho TAH(i X^ W^).  The 8 x 8
// matrix 1 + g M in the basis T^a (taken in the hat frame; the determinant
// does not depend on the basis) is inverted in place by Gauss-Jordan
// elimination, whose pivots give the log det (|g M| < 1 for
// rho < 1/(2(D-1)), so no pivoting is needed).
//
// Gradient of L = log det(1 + g M) with G = (1 + g M)^{-1}, U_b = sum_a G_ba T^a:
//   through M:  dL = -2 drho ReTr(Lam W) - 2 rho ReTr(Lam dW),
//               Lam = sum_b g(ad_{-iQ})(U_b) T^b;
//   through g:  dL = 2 Tr(Z dQ), with (Daleckii-Krein form of the derivative
//               of g(ad_{iQ}) in the hat frame)
//               Z^_ab = i sum_{b', l} [D_abl M^_b',al U^_b',lb
//                                      - conj(D_abl) U^_b',al M^_b',lb],
//               M_b' = -rho TAH(i T^b' W), D_abl = sum_m c_{m+1} i^m
//               h_m(lambda_a - lambda_l, lambda_b - lambda_l) with
//               c_k = (-1)^k / (k + 1)! and h_m the complete homogeneous
//               polynomial of degree m (the divided difference of g).

// Gell-Mann matrix b / 2, element (p, q), real and imaginary parts
__device__ inline double gm_re(int b, int p, int q)
{
    constexpr double h = 0.5, c8 = 0.28867513459481287, d8 = -0.5773502691896258;
    constexpr double R[8][3][3] = {
        {{0, h, 0}, {h, 0, 0}, {0, 0, 0}}, {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
        {{h, 0, 0}, {0, -h, 0}, {0, 0, 0}}, {{0, 0, h}, {0, 0, 0}, {h, 0, 0}},
        {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, {{0, 0, 0}, {0, 0, h}, {0, h, 0}},
        {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, {{c8, 0, 0}, {0, c8, 0}, {0, 0, d8}}};
    return R[b][p][q];
}
__device__ inline double gm_im(int b, int p, int q)
{
    constexpr double h = 0.5;
    constexpr double I[8][3][3] = {
        {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, {{0, -h, 0}, {h, 0, 0}, {0, 0, 0}},
        {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
        {{0, 0, -h}, {0, 0, 0}, {h, 0, 0}}, {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
        {{0, 0, 0}, {0, 0, -h}, {0, h, 0}}, {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}};
    return I[b][p][q];
}

// the same matrices for a runtime generator index (uniform across a
// wavefront: scalar loads): element (p, q) at [b][3 p + q]
__constant__ double kGmRe[8][9] = {
    {0, 0.5, 0, 0.5, 0, 0, 0, 0, 0},  {0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0.5, 0, 0, 0, -0.5, 0, 0, 0, 0}, {0, 0, 0.5, 0, 0, 0, 0.5, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 0},      {0, 0, 0, 0, 0, 0.5, 0, 0.5, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 0},      {0.28867513459481287, 0, 0, 0, 0.28867513459481287, 0, 0, 0, -0.5773502691896258}};
__constant__ double kGmIm[8][9] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0},  {0, -0.5, 0, 0.5, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 0},  {0, 0, -0.5, 0, 0, 0, 0.5, 0, 0}, {0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, -0.5, 0, 0.5, 0}, {0, 0, 0, 0, 0, 0, 0, 0, 0}};

__device__ inline C3 gm_mat(int b)
{
    C3 T;
#pragma unroll
    for (int e = 0; e < 9; e++) {
        T.r[e / 3][e % 3] = kGmRe[b][e];
        T.i[e / 3][e % 3] = kGmIm[b][e];
    }
    return T;
}

// i T^b Y (b a compile-time constant after unrolling)
__device__ inline C3 i_gm_mul(int b, const C3 &Y)
{
    C3 Z;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            double zr = 0.0, zi = 0.0;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const double tr = gm_re(b, p, r), ti = gm_im(b, p, r);
                // i (tr + i ti) (yr + i yi) = (-ti yr - tr yi) + i (tr yr - ti yi)
                zr += -ti * Y.r[r][q] - tr * Y.i[r][q];
                zi += tr * Y.r[r][q] - ti * Y.i[r][q];
            }
            Z.r[p][q] = zr;
            Z.i[p][q] = zi;
        }
    return Z;
}

// 2 Re Tr(T^a Y)
__device__ inline double gm_coord(int a, const C3 &Y)
{
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int r = 0; r < 3; r++) s += gm_re(a, p, r) * Y.r[r][p] - gm_im(a, p, r) * Y.i[r][p];
    return 2.0 * s;
}

// Eigen-decomposition of a Hermitian 3 x 3 matrix by cyclic complex Jacobi
// rotations: A = V diag(lam) V^+ (columns of V the eigenvectors).
__device__ inline void herm_eig3(C3 A, double lam[3], C3 &V)
{
    eye(V);
#pragma unroll 1
    for (int sweep = 0; sweep < 10; sweep++) {
        const double off = A.r[0][1] * A.r[0][1] + A.i[0][1] * A.i[0][1] + A.r[0][2] * A.r[0][2] +
                           A.i[0][2] * A.i[0][2] + A.r[1][2] * A.r[1][2] + A.i[1][2] * A.i[1][2];
        const double dia = A.r[0][0] * A.r[0][0] + A.r[1][1] * A.r[1][1] + A.r[2][2] * A.r[2][2];
        if (off <= 1e-36 * dia || off < 1e-300) break;
#pragma unroll
        for (int pr = 0; pr < 3; pr++) {
            const int p = pr == 2 ? 1 : 0, q = pr == 0 ? 1 : 2;
            const double ar = A.r[p][q], ai = A.i[p][q];
            const double a = sqrt(ar * ar + ai * ai);
            if (a < 1e-300) continue;
            const double phr = ar / a, phi = ai / a;  // ph = a_pq / |a_pq|
            const double tau = (A.r[q][q] - A.r[p][p]) / (2.0 * a);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
            // G = 1 but G_pp = G_qq = c, G_pq = s ph, G_qp = -s conj(ph);
            // A <- G^+ A G (columns, then rows), V <- V G
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double xr = A.r[k][p], xi = A.i[k][p], yr = A.r[k][q], yi = A.i[k][q];
                // col p: c x - s conj(ph) y ; col q: s ph x + c y
                A.r[k][p] = c * xr - sn * (phr * yr + phi * yi);
                A.i[k][p] = c * xi - sn * (phr * yi - phi * yr);
                A.r[k][q] = sn * (phr * xr - phi * xi) + c * yr;
                A.i[k][q] = sn * (phr * xi + phi * xr) + c * yi;
                const double vxr = V.r[k][p], vxi = V.i[k][p], vyr = V.r[k][q], vyi = V.i[k][q];
                V.r[k][p] = c * vxr - sn * (phr * vyr + phi * vyi);
                V.i[k][p] = c * vxi - sn * (phr * vyi - phi * vyr);
                V.r[k][q] = sn * (phr * vxr - phi * vxi) + c * vyr;
                V.i[k][q] = sn * (phr * vxi + phi * vxr) + c * vyi;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double xr = A.r[p][k], xi = A.i[p][k], yr = A.r[q][k], yi = A.i[q][k];
                // row p: c x - s ph y ; row q: s conj(ph) x + c y
                A.r[p][k] = c * xr - sn * (phr * yr - phi * yi);
                A.i[p][k] = c * xi - sn * (phr * yi + phi * yr);
                A.r[q][k] = sn * (phr * xr + phi * xi) + c * yr;
                A.i[q][k] = sn * (phr * xi - phi * xr) + c * yi;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) lam[k] = A.r[k][k];
}

// g(i d) = (1 - e^{-id}) / (id) = sin(d)/d - i (1 - cos d)/d, by its series
// (|d| <= 2 here: 14 terms reach 2^-60)
__device__ inline void g_of(double d, double &gr, double &gi)
{
    const double z = d * d;
    double fs = 1.0, fc = 0.5, sg = 1.0, sc = 0.0, oc = 0.0;
#pragma unroll
    for (int k = 0; k < 14; k++) {
        sc += sg * fs;
        oc += sg * fc;
        sg *= -z;
        fs /= (double)((2 * k + 2) * (2 * k + 3));
        fc /= (double)((2 * k + 3) * (2 * k + 4));
    }
    gr = sc;
    gi = -d * oc;
}

// D(iu, iv) = sum_m c_{m+1} i^m h_m(u, v): the divided difference
// (g(iu) - g(iv)) / (iu - iv), summed without division (|u|, |v| <= 2)
__device__ inline void dd_of(double u, double v, double &re, double &im)
{
    // c_{m+1} i^m = (-1)^(m+1) i^m / (m+2)!: real -(-1)^(m/2) / (m+2)! for
    // even m, imaginary (-1)^((m-1)/2) / (m+2)! for odd m
    double h = 1.0, vm = 1.0, fact = 2.0;
    re = 0.0;
    im = 0.0;
#pragma unroll
    for (int m = 0; m < 28; m++) {
        const double s2 = (m & 2) ? -1.0 : 1.0;
        if (m & 1) im += s2 * h / fact;
        else re -= s2 * h / fact;
        vm *= v;
        h = u * h + vm;
        fact *= (double)(m + 3);
    }
}

// 8 x 8 matrix 1 + g(ad_{iQ}) M in the hat frame (column b = image of T^b)
__device__ inline void stout_jac(const C3 &Wh, double rho, const double (&gr)[3][3], const double (&gi)[3][3],
                                 double (&J)[8][8])
{
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const C3 Mb = cscale(tah(i_gm_mul(b, Wh)), -rho, 0.0);
        C3 Y;
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                Y.r[p][q] = gr[p][q] * Mb.r[p][q] - gi[p][q] * Mb.i[p][q];
                Y.i[p][q] = gr[p][q] * Mb.i[p][q] + gi[p][q] * Mb.r[p][q];
            }
#pragma unroll
        for (int a = 0; a < 8; a++) J[a][b] = (a == b ? 1.0 : 0.0) + gm_coord(a, Y);
    }
}

// In-place inverse by Gauss-Jordan elimination without pivoting; returns
// the log of the determinant (sum of the log pivots)
__device__ inline double gauss_jordan8(double (&A)[8][8])
{
    double ld = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const double piv = A[k][k];
        ld += log(piv);
        const double inv = 1.0 / piv;
        A[k][k] = 1.0;
#pragma unroll
        for (int j = 0; j < 8; j++) A[k][j] *= inv;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            if (i == k) continue;
            const double f = A[i][k];
            A[i][k] = 0.0;
#pragma unroll
            for (int j = 0; j < 8; j++) A[i][j] -= f * A[k][j];
        }
    }
    return ld;
}

// the hat frame of one link: Q = -rho TAH(W), its eigenbasis, W^ and the
// factors g(i (lambda_a - lambda_b))
__device__ inline void stout_frame(const C3 &W, double rho, double (&lam)[3], C3 &V, C3 &Wh, double (&gr)[3][3],
                                   double (&gi)[3][3])
{
    herm_eig3(cscale(tah(W), -rho, 0.0), lam, V);
    Wh = mm<true, false>(V, mm<false, false>(W, V));
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) g_of(lam[p] - lam[q], gr[p][q], gi[p][q]);
}

// gradient buffer: 9 doubles per link (compact Hermitian: diag, then
// (0,1), (0,2), (1,2) as re, im)
__device__ inline C3 gget(const double *g, long l)
{
    const double *b = g + 9 * (size_t)l;
    C3 h;
    h.r[0][0] = b[0];
    h.r[1][1] = b[1];
    h.r[2][2] = b[2];
    h.i[0][0] = h.i[1][1] = h.i[2][2] = 0.0;
    h.r[0][1] = h.r[1][0] = b[3];
    h.i[0][1] = b[4];
    h.i[1][0] = -b[4];
    h.r[0][2] = h.r[2][0] = b[5];
    h.i[0][2] = b[6];
    h.i[2][0] = -b[6];
    h.r[1][2] = h.r[2][1] = b[7];
    h.i[1][2] = b[8];
    h.i[2][1] = -b[8];
    return h;
}

__device__ inline void gput(double *g, long l, const C3 &h)
{
    double *b = g + 9 * (size_t)l;
    b[0] = h.r[0][0];
    b[1] = h.r[1][1];
    b[2] = h.r[2][2];
    b[3] = h.r[0][1];
    b[4] = h.i[0][1];
    b[5] = h.r[0][2];
    b[6] = h.i[0][2];
    b[7] = h.r[1][2];
    b[8] = h.i[1][2];
}

// Contribution slots of the backward (no atomics): the backward thread of
// updated link t writes what it adds to each neighbouring link into its own
// slot, slot-major planes of 9 doubles: c[(slot * 9 + e) * T + t]; the
// gather kernel then sums, for every link, the slots that name it.
// Slots of link (x, MU): staple direction i (nu != MU, i = nu - (nu > MU))
// roles 0..5 = targets (x+MU, nu), (x+nu, MU), (x, nu), (x+MU-nu, nu),
// (x-nu, MU), (x-nu, nu) -> slot 6 i + role; plaquette plane j (nu < la,
// neither MU, in loop order) roles 0..3 = targets (x, nu), (x+nu, la),
// (x+la, nu), (x, la) -> slot 6 (D-1) + 4 j + role.
__device__ inline void cput(double *c, int slot, long t, long T, const C3 &h)
{
    double *b = c + (size_t)slot * 9 * T + t;
    b[0 * T] = h.r[0][0];
    b[1 * T] = h.r[1][1];
    b[2 * T] = h.r[2][2];
    b[3 * T] = h.r[0][1];
    b[4 * T] = h.i[0][1];
    b[5 * T] = h.r[0][2];
    b[6 * T] = h.i[0][2];
    b[7 * T] = h.r[1][2];
    b[8 * T] = h.i[1][2];
}

__device__ inline void cacc(double (&a)[9], const double *c, int slot, long t, long T)
{
    const double *b = c + (size_t)slot * 9 * T + t;
#pragma unroll
    for (int e = 0; e < 9; e++) a[e] += b[e * T];
}

struct Sub {
    int mu, par;
    const float *w;  // MLP block: W1[kHid][nf], b1[kHid], w2[kHid], b2
    int nf;          // number of features = planes not containing mu
    double rho_max, delta;
};

// planes (nu < la) without a given direction, and MLP features: the
// plaquette of every plane, then the 2x1 and 1x2 rectangles plane by plane
constexpr int nplanes(int D) { return (D - 1) * (D - 2) / 2; }
constexpr int nfeat(int D) { return 3 * nplanes(D); }

// The two rectangles at x in plane (nu, la), as six links A_1 ... A_6 with
// A_1..3 = X_1..3 and A_4..6 = X_4..6^+:
//   R1: X = U_nu(x), U_nu(x+nu), U_la(x+2nu), U_nu(x+nu+la), U_nu(x+la), U_la(x)
//   R2: X = U_nu(x), U_la(x+nu), U_la(x+nu+la), U_nu(x+2la), U_la(x+la), U_la(x)
struct Rect {
    long l[6];
};
__device__ inline void rectangles(const Geom &g, long s, int nu, int la, Rect &r1, Rect &r2)
{
    const long V = g.V;
    const long xpn = g.nbr[(2 * nu) * V + s], xla = g.nbr[(2 * la) * V + s];
    const long x2n = g.nbr[(2 * nu) * V + xpn], xnl = g.nbr[(2 * la) * V + xpn], x2l = g.nbr[(2 * la) * V + xla];
    r1 = Rect{{nu * V + s, nu * V + xpn, la * V + x2n, nu * V + xnl, nu * V + xla, la * V + s}};
    r2 = Rect{{nu * V + s, la * V + xpn, la * V + xnl, nu * V + x2l, la * V + xla, la * V + s}};
}

template <typename R>
__device__ inline C3 rect_loop(const R *U, const Rect &r)
{
    C3 M = mm<false, false>(mm<false, false>(ld(U, r.l[0]), ld(U, r.l[1])), ld(U, r.l[2]));
    M = mm<false, true>(mm<false, true>(mm<false, true>(M, ld(U, r.l[3])), ld(U, r.l[4])), ld(U, r.l[5]));
    return M;
}

// the forward quantities of one updated link that the backward reuses
template <int D>
struct Fwd {
    C3 U, W, Lh;  // own link (sub-layer input), U S, normalised TAH(C)
    double n, rho, theta, K;
    double f[nfeat(D)];
};

// Two wavefronts share the work of 64 updated links: the forward and
// backward kernels run workgroups of 2 x kLinks threads, and wavefront `role`
// (0 or 1) holds lane t of the same kLinks links.  Role 0 computes the
// plaquettes at x (their sum C and their features) and the 2x1 rectangles,
// role 1 the 1x2 rectangles and the staple sum S; the MLP is split by hidden
// unit.  The parts meet in LDS (one column per link); both wavefronts then
// hold C, S and every feature and compute the rest alike.  A sub-layer
// updates few links (32768 at 16^4) with 256-VGPR threads, so the split is
// what gives every SIMD a wavefront.
constexpr int kLinks = 64;

template <int D>
struct Xch {
    double c[18][kLinks], sm[18][kLinks];  // C (role 0), S (role 1): re then im, row-major
    double f[nfeat(D)][kLinks];            // features, each written by the role that computes it
    double z[2][kLinks];                   // MLP output partials of the two roles
    double df[2][nfeat(D)][kLinks];        // backward: dPhi/df partials of the two roles
};

__device__ inline void xput(double (*d)[kLinks], int lane, const C3 &m)
{
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            d[p * 3 + q][lane] = m.r[p][q];
            d[9 + p * 3 + q][lane] = m.i[p][q];
        }
}

__device__ inline C3 xget(const double (*d)[kLinks], int lane)
{
    C3 m;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            m.r[p][q] = d[p * 3 + q][lane];
            m.i[p][q] = d[9 + p * 3 + q][lane];
        }
    return m;
}

// Plaquette and rectangle features, MLP, staples and rotation angle of link
// (x = s, MU), split between the two roles as above.  F.U must hold the own
// link on entry (ok lanes).  Every thread of the workgroup calls it (it has
// two barriers).  The loops stay rolled (a fully unrolled stencil hoists
// every neighbour load and spills).
template <typename R, int D, int MU, bool ST>
__device__ inline void forward_link(const Geom &g, const R *U, const Sub &sb, long s, bool ok, int role, int lane,
                                    Xch<D> &x, Fwd<D> &F)
{
    constexpr int nf = nfeat(D), np = nplanes(D);
    const long V = g.V;
    if (ok && role == 0) {
        C3 C;
        zero(C);
        int j = 0;
#pragma unroll 1
        for (int nu = 0; nu < D; nu++) {
            if (nu == MU) continue;
            const long xpn = g.nbr[(2 * nu) * V + s];
#pragma unroll 1
            for (int la = nu + 1; la < D; la++) {
                if (la == MU) continue;
                const long xla = g.nbr[(2 * la) * V + s];
                const C3 P = mm<false, true>(
                    mm<false, true>(mm<false, false>(ld(U, nu * V + s), ld(U, la * V + xpn)), ld(U, nu * V + xla)),
                    ld(U, la * V + s));
                x.f[j][lane] = retr(P) / 3.0;
                addto(C, P);
                Rect r1, r2;
                rectangles(g, s, nu, la, r1, r2);
                x.f[np + 2 * j][lane] = retr(rect_loop(U, r1)) / 3.0;
                j++;
            }
        }
        xput(x.c, lane, C);
    } else if (ok) {
        int j = 0;
#pragma unroll 1
        for (int nu = 0; nu < D; nu++) {
            if (nu == MU) continue;
#pragma unroll 1
            for (int la = nu + 1; la < D; la++) {
                if (la == MU) continue;
                Rect r1, r2;
                rectangles(g, s, nu, la, r1, r2);
                x.f[np + 2 * j + 1][lane] = retr(rect_loop(U, r2)) / 3.0;
                j++;
            }
        }
        const long xp = g.nbr[(2 * MU) * V + s];
        C3 S;
        zero(S);
#pragma unroll 1
        for (int nu = 0; nu < D; nu++) {
            if (nu == MU) continue;
            const long xpn = g.nbr[(2 * nu) * V + s];
            const long xmn = g.nbr[(2 * nu + 1) * V + s];
            const long xpmn = g.nbr[(2 * nu + 1) * V + xp];
            addto(S, mm<false, true>(mm<false, true>(ld(U, nu * V + xp), ld(U, MU * V + xpn)), ld(U, nu * V + s)));
            addto(S, mm<false, false>(mm<true, true>(ld(U, nu * V + xpmn), ld(U, MU * V + xmn)), ld(U, nu * V + xmn)));
        }
        xput(x.sm, lane, S);
    }
    __syncthreads();
    const float *w = sb.w;
    if (ok) {
#pragma unroll
        for (int q = 0; q < nf; q++) F.f[q] = x.f[q][lane];
        if constexpr (!ST) {
            const C3 Lam = tah(xget(x.c, lane));
            F.n = sqrt(2.0 * retr(mm<false, false>(Lam, Lam)) + sb.delta * sb.delta);
            F.Lh = cscale(Lam, 1.0 / F.n, 0.0);
        }
        // hidden units [role kHid/2, (role + 1) kHid/2); role 0 adds b2
        double zp = role == 0 ? (double)w[kHid * nf + 2 * kHid] : 0.0;
#pragma unroll 1
        for (int k = role * (kHid / 2); k < (role + 1) * (kHid / 2); k++) {
            double a = (double)w[kHid * nf + k];
#pragma unroll
            for (int q = 0; q < nf; q++) a += (double)w[k * nf + q] * F.f[q];
            zp += (double)w[kHid * nf + kHid + k] * tanh(a);
        }
        x.z[role][lane] = zp;
    }
    __syncthreads();
    if (ok) {
        F.rho = sb.rho_max * tanh(x.z[0][lane] + x.z[1][lane]);
        F.W = mm<false, false>(F.U, xget(x.sm, lane));
        if constexpr (!ST) {
            F.theta = F.rho * imtr(mm<false, false>(F.Lh, F.W));
            F.K = retr(mm<false, false>(mm<false, false>(F.Lh, F.Lh), F.W));
        }
    }
}

// sites of parity p: device sites [0, V0) or [V0, V)
__device__ inline bool site_of(const Geom &g, int par, long t, long &s)
{
    const long base = par ? g.V0 : 0, len = par ? g.V - g.V0 : g.V0;
    s = base + t;
    return t < len;
}

// One sub-layer in place (Uin may equal Uout: no thread reads another
// thread's updated link).  With log, the sub-layer's input links are kept,
// link of parity-site t at index t (64-link tiles), for the backward.
// ST: the stout update U' = exp(-i rho TAH(W)) U instead, with W, rho and the
// features of every updated link written to wr (stride T: rows 0-17 W, 18
// rho, 19 on the features) for k_stout_logdet and the backward; no partials.
template <typename R, int D, int MU, bool ST>
__global__ __launch_bounds__(2 * kLinks) void k_forward(Geom g, const R *Uin, R *Uout, Sub sb, double *part, R *log,
                                                        double *wr, long T)
{
    __shared__ Xch<D> x;
    const int lane = (int)(threadIdx.x % kLinks), role = (int)(threadIdx.x / kLinks);
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);  // XCD-contiguous link blocks
    const long t = (long)lb * kLinks + lane;
    double sum[1] = {0.0}, cmp[1] = {0.0};
    long s;
    const bool ok = site_of(g, sb.par, t, s);
    Fwd<D> F;
    if (ok) {
        F.U = ld(Uin, MU * g.V + s);
        if (log && role == 0) st(log, t, F.U);
    }
    forward_link<R, D, MU, ST>(g, Uin, sb, s, ok, role, lane, x, F);
    if constexpr (ST) {
        if (ok && role == 0) {
            const C3 E = expi(cscale(tah(F.W), -F.rho, 0.0));
            st(Uout, MU * g.V + s, mm<false, false>(E, F.U));
        } else if (ok) {
#pragma unroll
            for (int e = 0; e < 9; e++) {
                wr[(size_t)e * T + t] = F.W.r[e / 3][e % 3];
                wr[(size_t)(9 + e) * T + t] = F.W.i[e / 3][e % 3];
            }
            wr[(size_t)18 * T + t] = F.rho;
#pragma unroll
            for (int q = 0; q < nfeat(D); q++) wr[(size_t)(19 + q) * T + t] = F.f[q];
        }
    } else {
        if (ok && role == 0) {
            const C3 E = expi(cscale(F.Lh, F.theta, 0.0));
            st(Uout, MU * g.V + s, mm<false, false>(E, F.U));
            sum[0] = log1p(F.rho * F.K);
        }
        block_reduce_store<1>(sum, cmp, part + 2 * (size_t)lb);
    }
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Backward of one sub-layer.  U holds the sub-layer's output field on entry;
// the kernel restores its input (the updated links from log) in place.  gr
// holds the gradient with respect to the sub-layer's output on entry and
// with respect to its input on exit.  c_logj = dPhi/dlogJ (-1 for
// Phi = beta S - log J).  gw (may be null) accumulates dPhi/dweights of the
// sub-layer's MLP block.  Every forward quantity is recomputed from the
// sub-layer input (layer recomputation; nothing but the input is stored),
// except that the stout kind reads W, rho and the features from wr, where
// its forward wrote them (the same values: the same code on the same links).
// Lie-gradient contributions of c ReTr R (c = dPhi/df / 6 with
// f = ReTr R / 3) to the six links of rectangle R = A_1 ... A_6 (A_1..3 = X,
// A_4..6 = X^+): with the cyclic rotations M_k = A_k ... A_6 A_1 ... A_{k-1},
// link k gets proj(+i c M_k) for k <= 3 and proj(-i c M_{k+1}) for k >= 4
// (M_7 = M_1).  M_{k+1} = A_k^{-1} M_k A_k, with A^{-1} = A^+ (SU(3) links).
template <typename R>
__device__ inline void rect_grad(const R *U, const Rect &r, double c, double *contrib, int sl, long t, long T)
{
    const C3 M1 = rect_loop(U, r);
    cput(contrib, sl + 0, t, T, proj(cscale(M1, 0.0, c)));
    cput(contrib, sl + 5, t, T, proj(cscale(M1, 0.0, -c)));
    C3 M = M1;
#pragma unroll
    for (int k = 0; k < 2; k++) {  // M_2, M_3
        const C3 X = ld(U, r.l[k]);
        M = mm<true, false>(X, mm<false, false>(M, X));
        cput(contrib, sl + 1 + k, t, T, proj(cscale(M, 0.0, c)));
    }
    {  // M_4 = X_3^+ M_3 X_3
        const C3 X = ld(U, r.l[2]);
        M = mm<true, false>(X, mm<false, false>(M, X));
    }
#pragma unroll
    for (int k = 3; k < 5; k++) {  // M_5, M_6 (A_k = X_k^+)
        const C3 X = ld(U, r.l[k]);
        M = mm<false, true>(mm<false, false>(X, M), X);
        cput(contrib, sl + k, t, T, proj(cscale(M, 0.0, -c)));
    }
}

// ST: the stout sub-layer's backward; the log det part of the gradient
// comes from k_stout_logdet_grad in lg (stride T), c_logj already applied.
template <typename R, int D, int MU, bool ST>
__global__ __launch_bounds__(2 * kLinks) void k_backward(Geom g, R *U, Sub sb, double *gr, double c_logj, double *gw,
                                                         const R *log, double *contrib, long T, const double *lg,
                                                         const double *wr)
{
    constexpr int nf = nfeat(D), np = nplanes(D);
    __shared__ Xch<D> x;
    __shared__ double sga[18][kLinks];  // stout: Gamma = R^T(G') from role 0
    __shared__ double sdr[kLinks];      //        and its d / d rho part
    const int lane = (int)(threadIdx.x % kLinks), role = (int)(threadIdx.x / kLinks);
    const long t = (long)xcd_block(blockIdx.x, gridDim.x) * kLinks + lane;
    long s;
    const bool ok = site_of(g, sb.par, t, s);
    const long V = g.V;
    const long l = MU * V + s;
    Fwd<D> F;
#pragma unroll
    for (int q = 0; q < nf; q++) F.f[q] = 0.0;
    if (ok) {
        F.U = ld(log, t);
        if (role == 0) st(U, l, F.U);
    }
    if constexpr (ST) {
        if (ok) {
#pragma unroll
            for (int e = 0; e < 9; e++) {
                F.W.r[e / 3][e % 3] = wr[(size_t)e * T + t];
                F.W.i[e / 3][e % 3] = wr[(size_t)(9 + e) * T + t];
            }
            F.rho = wr[(size_t)18 * T + t];
#pragma unroll
            for (int q = 0; q < nf; q++) F.f[q] = wr[(size_t)(19 + q) * T + t];
        }
    } else {
        forward_link<R, D, MU, ST>(g, U, sb, s, ok, role, lane, x, F);
    }
    // Both roles need dPhi/dz; role 0 then needs dPhi/dC (through Lhat, the
    // dexp adjoint series), role 1 the own link's gradient (through E) and
    // dPhi/dS.
    double dz = 0.0;
    C3 gU, GC, GS;
    if constexpr (ST) {
        // U' = E U, E = exp(iQ), Q = -rho TAH(W):
        //   dPhi = 2 Tr(G' dE E^+) + 2 Tr(E^+ G' E X)     (dU = iXU)
        //   2 Tr(G' dE E^+) = 2 Tr(Gamma dQ), Gamma = sum_k ad_{-iQ}^k(G') / (k+1)!
        //   dQ = -drho TAH(W) - rho TAH(dW): dPhi/drho = -2 Im Tr(Gamma W),
        //   G_W = i rho Gamma (dPhi = 2 ReTr(G_W dW)), then the log det's part
        if (ok) {
            const C3 gp = gget(gr, l);
            const C3 Q = cscale(tah(F.W), -F.rho, 0.0);
            if (role == 0) {
                const C3 Ga = avg_ad(Q, gp);
                xput(sga, lane, Ga);
                sdr[lane] = -2.0 * imtr(mm<false, false>(Ga, F.W));
                zero(GC);
            } else {
                const C3 E = expi(Q);
                gU = proj(mm<true, false>(E, mm<false, false>(gp, E)));
            }
        }
        __syncthreads();
        if (ok) {
            const double drho = sdr[lane] + lg[(size_t)18 * T + t];
            const double th = F.rho / sb.rho_max;
            dz = drho * sb.rho_max * (1.0 - th * th);
            if (role == 1) {
                C3 GW = cscale(xget(sga, lane), 0.0, F.rho);
#pragma unroll
                for (int e = 0; e < 9; e++) {
                    GW.r[e / 3][e % 3] += lg[(size_t)e * T + t];
                    GW.i[e / 3][e % 3] += lg[(size_t)(9 + e) * T + t];
                }
                addto(gU, proj(cscale(mm<false, false>(F.W, GW), 0.0, 1.0)));
                GS = mm<false, false>(GW, F.U);
            }
        }
    } else if (ok) {
        const C3 gp = gget(gr, l);
        const double dtheta = 2.0 * retr(mm<false, false>(gp, F.Lh));
        // theta = rho Im Tr(Lhat W)
        double drho = dtheta * imtr(mm<false, false>(F.Lh, F.W));
        // log J = log(1 + rho K), K = ReTr(Lhat^2 W)
        const double a = c_logj / (1.0 + F.rho * F.K);
        drho += a * F.K;
        // rho = rho_max tanh(z)
        const double th = F.rho / sb.rho_max;
        dz = drho * sb.rho_max * (1.0 - th * th);
        if (role == 0) {
            // U' = E U with E = exp(i theta Lhat): through Lhat
            C3 gLh = cscale(avg_ad(cscale(F.Lh, F.theta, 0.0), gp), F.theta, 0.0);
            {
                C3 WmWd;
#pragma unroll
                for (int p = 0; p < 3; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        WmWd.r[p][q] = F.W.r[p][q] - F.W.r[q][p];
                        WmWd.i[p][q] = F.W.i[p][q] + F.W.i[q][p];
                    }
                addto(gLh, cscale(WmWd, 0.0, -0.25 * dtheta * F.rho));  // dtheta rho (W - W^+)/(4i)
            }
            {
                C3 M = mm<false, false>(F.Lh, F.W);
                addto(M, mm<false, false>(F.W, F.Lh));
                C3 Md;
#pragma unroll
                for (int p = 0; p < 3; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        Md.r[p][q] = M.r[p][q] + M.r[q][p];
                        Md.i[p][q] = M.i[p][q] - M.i[q][p];
                    }
                addto(gLh, Md, 0.25 * a * F.rho);
            }
            // Lhat = Lam / n, Lam = n Lhat
            const double tgl = 2.0 * F.n * retr(mm<false, false>(gLh, F.Lh));
            C3 gLam = cscale(gLh, 1.0 / F.n, 0.0);
            addto(gLam, F.Lh, -tgl / (F.n * F.n));
            GC = cscale(proj(gLam), 0.0, -1.0);  // Lam = TAH(C)
        } else {
            C3 GW = cscale(F.Lh, 0.0, -0.5 * dtheta * F.rho);  // dtheta rho (-i Lhat)/2
            addto(GW, mm<false, false>(F.Lh, F.Lh), 0.5 * a * F.rho);
            // own link: through E, then through W = U S
            {
                const C3 E = expi(cscale(F.Lh, F.theta, 0.0));
                gU = proj(mm<true, false>(E, mm<false, false>(gp, E)));
            }
            addto(gU, proj(cscale(mm<false, false>(F.W, GW), 0.0, 1.0)));
            // staple links: G_S = G_W U
            GS = mm<false, false>(GW, F.U);
        }
    }
    // MLP backward over this role's hidden units, every lane (inactive lanes
    // carry dz = 0, f = 0) so that the weight gradient can be summed over
    // the wavefront first; dPhi/df partials meet in LDS
    const float *w = sb.w;
    double df[nf > 0 ? nf : 1];
#pragma unroll
    for (int q = 0; q < nf; q++) df[q] = 0.0;
    const bool lead = lane == 0;
#pragma unroll 1
    for (int k = role * (kHid / 2); k < (role + 1) * (kHid / 2); k++) {
        double a = (double)w[kHid * nf + k];
#pragma unroll
        for (int q = 0; q < nf; q++) a += (double)w[k * nf + q] * F.f[q];
        const double h = tanh(a);
        const double dh = dz * (double)w[kHid * nf + kHid + k];
        const double da = dh * (1.0 - h * h);
#pragma unroll
        for (int q = 0; q < nf; q++) df[q] += da * (double)w[k * nf + q];
        if (gw) {
            const double s_w2 = wave_sum(dz * h), s_b1 = wave_sum(da);
            if (lead) {
                atomicAdd(gw + kHid * nf + kHid + k, s_w2);
                atomicAdd(gw + kHid * nf + k, s_b1);
            }
#pragma unroll
            for (int q = 0; q < nf; q++) {
                const double s_w1 = wave_sum(da * F.f[q]);
                if (lead) atomicAdd(gw + k * nf + q, s_w1);
            }
        }
    }
    if (gw && role == 0) {
        const double s_b2 = wave_sum(dz);
        if (lead) atomicAdd(gw + kHid * nf + 2 * kHid, s_b2);
    }
#pragma unroll
    for (int q = 0; q < nf; q++) x.df[role][q][lane] = df[q];
    // also orders every read of gr[l] above before its write below
    __syncthreads();
    if (!ok) return;
#pragma unroll
    for (int q = 0; q < nf; q++) df[q] = x.df[0][q][lane] + x.df[1][q][lane];
    if (role == 0) {
        // plaquettes P = p1 p2 p3^+ p4^+ in the planes (nu, la) without MU:
        // dPhi/dLam through C = sum P, and dPhi/df through ReTr P / 3; then
        // the 2x1 rectangles: dPhi/df (f = ReTr R / 3) only
        int j = 0;
#pragma unroll 1
        for (int nu = 0; nu < D; nu++) {
            if (nu == MU) continue;
            const long xpn = g.nbr[(2 * nu) * V + s];
#pragma unroll 1
            for (int la = nu + 1; la < D; la++) {
                if (la == MU) continue;
                const long xla = g.nbr[(2 * la) * V + s];
                const C3 p1 = ld(U, nu * V + s), p2 = ld(U, la * V + xpn), p3 = ld(U, nu * V + xla),
                         p4 = ld(U, la * V + s);
                C3 GP = GC;
                double dfj = 0.0, df1 = 0.0;
#pragma unroll
                for (int q = 0; q < nf; q++) {
                    if (q == j) dfj = df[q];
                    if (q == np + 2 * j) df1 = df[q];
                }
                const double cf_ = dfj / 6.0;
#pragma unroll
                for (int p = 0; p < 3; p++) GP.r[p][p] += cf_;
                const C3 p12 = mm<false, false>(p1, p2);
                const C3 p123 = mm<false, true>(p12, p3);
                const C3 P = mm<false, true>(p123, p4);
                const int sl = 6 * (D - 1) + 16 * j;
                cput(contrib, sl + 0, t, T, proj(cscale(mm<false, false>(P, GP), 0.0, 1.0)));
                cput(contrib, sl + 1, t, T,
                     proj(cscale(mm<false, false>(mm<false, true>(mm<false, true>(p2, p3), p4), mm<false, false>(GP, p1)),
                                 0.0, 1.0)));
                cput(contrib, sl + 2, t, T, proj(cscale(mm<true, false>(p4, mm<false, false>(GP, p123)), 0.0, -1.0)));
                cput(contrib, sl + 3, t, T, proj(cscale(mm<false, false>(GP, P), 0.0, -1.0)));
                Rect r1, r2;
                rectangles(g, s, nu, la, r1, r2);
                rect_grad(U, r1, df1 / 6.0, contrib, sl + 4, t, T);
                j++;
            }
        }
    } else {
        gput(gr, l, gU);
        // staple links through S
        const long xp = g.nbr[(2 * MU) * V + s];
#pragma unroll 1
        for (int nu = 0; nu < D; nu++) {
            if (nu == MU) continue;
            const long xpn = g.nbr[(2 * nu) * V + s];
            const long xmn = g.nbr[(2 * nu + 1) * V + s];
            const long xpmn = g.nbr[(2 * nu + 1) * V + xp];
            const int sl = 6 * (nu - (nu > MU));
            {
                const C3 ca = ld(U, nu * V + xp), cb = ld(U, MU * V + xpn), cc = ld(U, nu * V + s);
                // t1 = a b^+ c^+
                const C3 abd = mm<false, true>(ca, cb);
                const C3 t1 = mm<false, true>(abd, cc);
                cput(contrib, sl + 0, t, T, proj(cscale(mm<false, false>(t1, GS), 0.0, 1.0)));
                cput(contrib, sl + 1, t, T, proj(cscale(mm<true, false>(cc, mm<false, false>(GS, abd)), 0.0, -1.0)));
                cput(contrib, sl + 2, t, T, proj(cscale(mm<false, false>(GS, t1), 0.0, -1.0)));
            }
            {
                const C3 cd = ld(U, nu * V + xpmn), ce = ld(U, MU * V + xmn), cf = ld(U, nu * V + xmn);
                // t2 = d^+ e^+ f
                const C3 de = mm<true, true>(cd, ce);
                const C3 fGS = mm<false, false>(cf, GS);
                cput(contrib, sl + 3, t, T, proj(cscale(mm<false, true>(mm<true, false>(ce, fGS), cd), 0.0, -1.0)));
                const C3 fd = mm<false, false>(fGS, de);
                cput(contrib, sl + 4, t, T, proj(cscale(fd, 0.0, -1.0)));
                cput(contrib, sl + 5, t, T, proj(cscale(fd, 0.0, 1.0)));
            }
        }
        // the 1x2 rectangles: dPhi/df only
        int j = 0;
#pragma unroll 1
        for (int nu = 0; nu < D; nu++) {
            if (nu == MU) continue;
#pragma unroll 1
            for (int la = nu + 1; la < D; la++) {
                if (la == MU) continue;
                double df2 = 0.0;
#pragma unroll
                for (int q = 0; q < nf; q++)
                    if (q == np + 2 * j + 1) df2 = df[q];
                Rect r1, r2;
                rectangles(g, s, nu, la, r1, r2);
                rect_grad(U, r2, df2 / 6.0, contrib, 6 * (D - 1) + 16 * j + 10, t, T);
                j++;
            }
        }
    }
}

// Plane-loop roles of the contribution slots (per plane (p0, p1), 16 slots):
// plaquette links p1..p4, then rectangle R1 A_1..A_6, then R2 A_1..A_6.  Role
// r names link (x + kRoleA[r] p0 + kRoleB[r] p1, p0 if kRoleDir[r] == 0 else p1).
__constant__ int kRoleDir[16] = {0, 1, 0, 1, 0, 0, 1, 0, 0, 1, 0, 1, 1, 0, 1, 1};
__constant__ int kRoleA[16] = {0, 1, 0, 0, 0, 1, 2, 1, 0, 0, 0, 1, 1, 0, 0, 0};
__constant__ int kRoleB[16] = {0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 2, 1, 0};

// gr[y, nu] += the contribution slots of the sub-layer's updated links (x,
// MU) of parity p that name link (y, nu) (the inverse of the slot map).
template <typename R, int D, int MU>
__global__ __launch_bounds__(kB) void k_gather(Geom g, double *gr, const double *contrib, int par, long T)
{
    const long L = (long)blockIdx.x * kB + threadIdx.x;
    const long V = g.V;
    if (L >= V * D) return;
    const int nu = (int)(L / V);
    const long y = L - nu * V;
    const int py = y >= g.V0;
    const long base = par ? g.V0 : 0;
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (nu == MU) {
        if (py == par) return;  // updated link: its gradient is final
        for (int n2 = 0; n2 < D; n2++) {
            if (n2 == MU) continue;
            const int sl = 6 * (n2 - (n2 > MU));
            cacc(a, contrib, sl + 1, g.nbr[(2 * n2 + 1) * V + y] - base, T);  // x = y - n2
            cacc(a, contrib, sl + 4, g.nbr[(2 * n2) * V + y] - base, T);      // x = y + n2
        }
    } else {
        const int sl = 6 * (nu - (nu > MU));
        if (py == par) {
            cacc(a, contrib, sl + 2, y - base, T);  // x = y
            const long x = g.nbr[(2 * nu) * V + g.nbr[(2 * MU + 1) * V + y]];
            cacc(a, contrib, sl + 3, x - base, T);  // x = y - MU + nu
        } else {
            cacc(a, contrib, sl + 0, g.nbr[(2 * MU + 1) * V + y] - base, T);  // x = y - MU
            cacc(a, contrib, sl + 5, g.nbr[(2 * nu) * V + y] - base, T);      // x = y + nu
        }
        // plane loops (plaquette and rectangles) through link (y, nu): role r
        // of plane (p0, p1) names link (x + a p0 + b p1, dir); x = y - a p0 - b p1
        int j = 0;
        for (int p0 = 0; p0 < D; p0++) {
            if (p0 == MU) continue;
            for (int p1 = p0 + 1; p1 < D; p1++) {
                if (p1 == MU) continue;
                if (nu == p0 || nu == p1) {
                    const int side = nu == p0 ? 0 : 1;
                    for (int r = 0; r < 16; r++) {
                        if (kRoleDir[r] != side) continue;
                        const int ra = kRoleA[r], rb = kRoleB[r];
                        if (((py + ra + rb) & 1) != par) continue;
                        long x = y;
                        for (int q = 0; q < ra; q++) x = g.nbr[(2 * p0 + 1) * V + x];
                        for (int q = 0; q < rb; q++) x = g.nbr[(2 * p1 + 1) * V + x];
                        cacc(a, contrib, 6 * (D - 1) + 16 * j + r, x - base, T);
                    }
                }
                j++;
            }
        }
    }
    double *b = gr + 9 * (size_t)L;
#pragma unroll
    for (int e = 0; e < 9; e++) b[e] += a[e];
}

// g_l = (beta / 2N) TAH(U_l sum staples): gradient of beta S_W
template <typename R>
__global__ __launch_bounds__(kB) void k_wilson_grad(Geom g, const R *U, double coef, double *gr, double *part)
{
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    double sum[1] = {0.0}, cmp[1] = {0.0};
    if (t < g.V * g.D) {
        const long V = g.V;
        const int mu = (int)(t / V);
        const long s = t - mu * V;
        const long xp = g.nbr[(2 * mu) * V + s];
        C3 S;
        zero(S);
        for (int nu = 0; nu < g.D; nu++) {
            if (nu == mu) continue;
            const long xpn = g.nbr[(2 * nu) * V + s];
            const long xmn = g.nbr[(2 * nu + 1) * V + s];
            const long xpmn = g.nbr[(2 * nu + 1) * V + xp];
            addto(S, mm<false, true>(mm<false, true>(ld(U, nu * V + xp), ld(U, mu * V + xpn)), ld(U, nu * V + s)));
            addto(S, mm<false, false>(mm<true, true>(ld(U, nu * V + xpmn), ld(U, mu * V + xmn)), ld(U, nu * V + xmn)));
        }
        const C3 W = mm<false, false>(ld(U, t), S);
        gput(gr, t, cscale(tah(W), coef, 0.0));
        sum[0] = retr(W);
    }
    block_reduce_store<1>(sum, cmp, part + 2 * (size_t)blockIdx.x);
}

// P <- P - eps c g (compact momenta, tile 64)
template <typename R>
__global__ void k_kick_g(Geom g, R *P, const double *gr, double eps_c)
{
    const long l = (long)blockIdx.x * kB + threadIdx.x;
    if (l >= g.V * g.D) return;
    const double *b = gr + 9 * (size_t)l;
    R *p = P + (size_t)(l >> 6) * (9 * 64) + (l & 63);
    // momentum compact order: diag 0..2, (0,1), (0,2), (1,2) as (re, im): same
#pragma unroll
    for (int c = 0; c < 9; c++) p[c * 64] = (R)((double)p[c * 64] - eps_c * b[c]);
}

// ---- stout sub-layer: log det J and its gradient, one thread per link ----
// wr holds W = U S (18 doubles, re then im, row-major) and rho of every
// updated link t (parity-site index), structure of arrays with stride T.

__device__ inline C3 wr_get(const double *wr, long t, long T, double &rho)
{
    C3 W;
#pragma unroll
    for (int e = 0; e < 9; e++) {
        W.r[e / 3][e % 3] = wr[(size_t)e * T + t];
        W.i[e / 3][e % 3] = wr[(size_t)(9 + e) * T + t];
    }
    rho = wr[(size_t)18 * T + t];
    return W;
}

// log det J per link into per-block partials; with aux also (1 + g M)^{-1}
// (64 doubles), V (18) and lambda (3) for the gradient kernel
__global__ __launch_bounds__(kLinks) void k_stout_logdet(Geom g, int par, const double *wr, long T, double *part,
                                                         double *aux)
{
    const long t = (long)blockIdx.x * kLinks + threadIdx.x;
    double sum[1] = {0.0}, cmp[1] = {0.0};
    long s;
    if (site_of(g, par, t, s)) {
        double rho;
        const C3 W = wr_get(wr, t, T, rho);
        double lam[3], gr[3][3], gi[3][3];
        C3 V, Wh;
        stout_frame(W, rho, lam, V, Wh, gr, gi);
        double J[8][8];
        stout_jac(Wh, rho, gr, gi, J);
        sum[0] = gauss_jordan8(J);
        if (aux) {
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int b = 0; b < 8; b++) aux[(size_t)(a * 8 + b) * T + t] = J[a][b];
#pragma unroll
            for (int e = 0; e < 9; e++) {
                aux[(size_t)(64 + e) * T + t] = V.r[e / 3][e % 3];
                aux[(size_t)(73 + e) * T + t] = V.i[e / 3][e % 3];
            }
#pragma unroll
            for (int k = 0; k < 3; k++) aux[(size_t)(82 + k) * T + t] = lam[k];
        }
    }
    block_reduce_store<1>(sum, cmp, part + 2 * (size_t)blockIdx.x);
}

// U_b = sum_a G_ba T^a from row b of the inverse (in aux); b runtime
__device__ inline C3 row_gen(const double *aux, int b, long t, long T)
{
    C3 U;
    zero(U);
#pragma unroll 1
    for (int a = 0; a < 8; a++) {
        const double gba = aux[(size_t)(b * 8 + a) * T + t];
#pragma unroll
        for (int e = 0; e < 9; e++) {
            U.r[e / 3][e % 3] += gba * kGmRe[a][e];
            U.i[e / 3][e % 3] += gba * kGmIm[a][e];
        }
    }
    return U;
}

// lg = c (i rho Z - rho Lam) (18 doubles: the G_W of c log det J) and
// c (-2 Im Tr(Z W) - 2 Re Tr(Lam W)) (d / d rho), per updated link.  The
// generator loops run over a runtime index (matrices from constant memory)
// to keep the register footprint small.
__global__ __launch_bounds__(kLinks) void k_stout_logdet_grad(Geom g, int par, const double *wr, const double *aux,
                                                              long T, double c, double *lg)
{
    __shared__ double sD[2][27][kLinks];  // divided differences D_abl, one column per thread
    const int lane = (int)threadIdx.x;
    const long t = (long)blockIdx.x * kLinks + threadIdx.x;
    long s;
    if (!site_of(g, par, t, s)) return;
    double rho;
    const C3 W = wr_get(wr, t, T, rho);
    C3 V;
    double lam[3];
#pragma unroll
    for (int e = 0; e < 9; e++) {
        V.r[e / 3][e % 3] = aux[(size_t)(64 + e) * T + t];
        V.i[e / 3][e % 3] = aux[(size_t)(73 + e) * T + t];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) lam[k] = aux[(size_t)(82 + k) * T + t];
    const C3 Wh = mm<true, false>(V, mm<false, false>(W, V));
    double gr[3][3], gi[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) g_of(lam[p] - lam[q], gr[p][q], gi[p][q]);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = a; b < 3; b++)
#pragma unroll
            for (int l = 0; l < 3; l++) {
                double dr, di;
                dd_of(lam[a] - lam[l], lam[b] - lam[l], dr, di);
                sD[0][(a * 3 + b) * 3 + l][lane] = sD[0][(b * 3 + a) * 3 + l][lane] = dr;
                sD[1][(a * 3 + b) * 3 + l][lane] = sD[1][(b * 3 + a) * 3 + l][lane] = di;
            }
    // Lam^ = sum_b g(ad_{-iQ})(U_b) T^b (g(-ix) = conj(g(ix))) and Z^, one
    // generator at a time
    C3 Lh, Zh;
    zero(Lh);
    zero(Zh);
#pragma unroll 1
    for (int b = 0; b < 8; b++) {
        const C3 U = row_gen(aux, b, t, T);
        const C3 Tb = gm_mat(b);
        C3 Y;
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                Y.r[p][q] = gr[p][q] * U.r[p][q] + gi[p][q] * U.i[p][q];
                Y.i[p][q] = gr[p][q] * U.i[p][q] - gi[p][q] * U.r[p][q];
            }
        addto(Lh, mm<false, false>(Y, Tb));
        // M_b = -rho TAH(i T^b W^)
        const C3 M = cscale(tah(cscale(mm<false, false>(Tb, Wh), 0.0, 1.0)), -rho, 0.0);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                double ar = 0.0, ai = 0.0;
#pragma unroll
                for (int l = 0; l < 3; l++) {
                    const double dr = sD[0][(a * 3 + q) * 3 + l][lane], di = sD[1][(a * 3 + q) * 3 + l][lane];
                    // M_al U_lq
                    const double p1r = M.r[a][l] * U.r[l][q] - M.i[a][l] * U.i[l][q];
                    const double p1i = M.r[a][l] * U.i[l][q] + M.i[a][l] * U.r[l][q];
                    // U_al M_lq
                    const double p2r = U.r[a][l] * M.r[l][q] - U.i[a][l] * M.i[l][q];
                    const double p2i = U.r[a][l] * M.i[l][q] + U.i[a][l] * M.r[l][q];
                    // D p1 - conj(D) p2
                    ar += dr * p1r - di * p1i - (dr * p2r + di * p2i);
                    ai += dr * p1i + di * p1r - (dr * p2i - di * p2r);
                }
                // Z^ += i (ar + i ai)
                Zh.r[a][q] -= ai;
                Zh.i[a][q] += ar;
            }
    }
    const C3 Lo = mm<false, true>(mm<false, false>(V, Lh), V);
    const C3 Zo = mm<false, true>(mm<false, false>(V, Zh), V);
    C3 G = cscale(Zo, 0.0, rho);  // i rho Z
    addto(G, Lo, -rho);
    G = cscale(G, c, 0.0);
    const double dr = c * (-2.0 * imtr(mm<false, false>(Zo, W)) - 2.0 * retr(mm<false, false>(Lo, W)));
#pragma unroll
    for (int e = 0; e < 9; e++) {
        lg[(size_t)e * T + t] = G.r[e / 3][e % 3];
        lg[(size_t)(9 + e) * T + t] = G.i[e / 3][e % 3];
    }
    lg[(size_t)18 * T + t] = dr;
}

}  // namespace flow

// ---------------------------------------------------------------- launchers
static flow::Sub to_sub(const FlowSub &f) { return flow::Sub{f.mu, f.par, f.w, f.nf, f.rho_max, f.delta}; }

static unsigned par_blocks(const Geom &g, int par)
{
    const long len = par ? g.V - g.V0 : g.V0;
    return (unsigned)((len + flow::kLinks - 1) / flow::kLinks);
}

long flow_forward_partials(const Geom &g, int par) { return par_blocks(g, par); }

size_t flow_log_reals(const Geom &g)
{
    const long n = g.V0 > g.V - g.V0 ? g.V0 : g.V - g.V0;
    return (size_t)((n + 63) / 64) * 64 * 18;
}

// (D, mu) dispatch for the flow kernels: the plane enumeration is unrolled
// at compile time
template <template <typename, int, int> class K, typename R, typename... Args>
static hipError_t flow_dispatch(int D, int mu, unsigned nb, hipStream_t st, Args... args)
{
#define ASKJA_FLOW_CASE(DD, MM)                                                              \
    if (D == DD && mu == MM) {                                                               \
        hipLaunchKernelGGL((K<R, DD, MM>::fn), dim3(nb), dim3(K<R, DD, MM>::block), 0, st, args...); \
        return hipGetLastError();                                                            \
    }
    ASKJA_FLOW_CASE(3, 0) ASKJA_FLOW_CASE(3, 1) ASKJA_FLOW_CASE(3, 2)
    ASKJA_FLOW_CASE(4, 0) ASKJA_FLOW_CASE(4, 1) ASKJA_FLOW_CASE(4, 2) ASKJA_FLOW_CASE(4, 3)
#undef ASKJA_FLOW_CASE
    return hipErrorInvalidValue;
}

template <typename R, int D, int MU>
struct FwdK {
    static constexpr auto fn = flow::k_forward<R, D, MU, false>;
    static constexpr int block = 2 * flow::kLinks;
};
template <typename R, int D, int MU>
struct FwdKS {
    static constexpr auto fn = flow::k_forward<R, D, MU, true>;
    static constexpr int block = 2 * flow::kLinks;
};
template <typename R, int D, int MU>
struct BwdK {
    static constexpr auto fn = flow::k_backward<R, D, MU, false>;
    static constexpr int block = 2 * flow::kLinks;
};
template <typename R, int D, int MU>
struct BwdKS {
    static constexpr auto fn = flow::k_backward<R, D, MU, true>;
    static constexpr int block = 2 * flow::kLinks;
};

// parity-site stride of the per-link structure-of-arrays buffers
static long sub_stride(const Geom &g) { return (long)(flow_log_reals(g) / 18); }

size_t stout_wr_doubles(const Geom &g) { return (size_t)(19 + flow::nfeat(g.D)) * sub_stride(g); }
size_t stout_aux_doubles(const Geom &g) { return (size_t)85 * sub_stride(g); }
size_t stout_lg_doubles(const Geom &g) { return (size_t)19 * sub_stride(g); }

hipError_t launch_stout_logdet(const Geom &g, const FlowSub &f, const double *st_wr, double *part, double *aux,
                               hipStream_t st)
{
    hipLaunchKernelGGL(flow::k_stout_logdet, dim3(par_blocks(g, f.par)), dim3(flow::kLinks), 0, st, g, f.par, st_wr,
                       sub_stride(g), part, aux);
    return hipGetLastError();
}

hipError_t launch_stout_logdet_grad(const Geom &g, const FlowSub &f, const double *st_wr, const double *aux,
                                    double c_logj, double *lg, hipStream_t st)
{
    hipLaunchKernelGGL(flow::k_stout_logdet_grad, dim3(par_blocks(g, f.par)), dim3(flow::kLinks), 0, st, g, f.par,
                       st_wr, aux, sub_stride(g), c_logj, lg);
    return hipGetLastError();
}
template <typename R, int D, int MU>
struct GatK {
    static constexpr auto fn = flow::k_gather<R, D, MU>;
    static constexpr int block = flow::kB;
};

hipError_t launch_flow_forward(const Geom &g, const void *Uin, void *Uout, const FlowSub &f, int mode, double *part,
                               void *log, double *st_wr, hipStream_t st)
{
    const unsigned nb = par_blocks(g, f.par);
    const long T = sub_stride(g);
    const bool stout = f.kind == kFlowStout;
    if (stout && !st_wr) return hipErrorInvalidValue;
    if (f32_mode(mode)) {
        if (stout)
            return flow_dispatch<FwdKS, float>(g.D, f.mu, nb, st, g, (const float *)Uin, (float *)Uout, to_sub(f), part,
                                               (float *)log, st_wr, T);
        return flow_dispatch<FwdK, float>(g.D, f.mu, nb, st, g, (const float *)Uin, (float *)Uout, to_sub(f), part,
                                          (float *)log, st_wr, T);
    }
    if (stout)
        return flow_dispatch<FwdKS, double>(g.D, f.mu, nb, st, g, (const double *)Uin, (double *)Uout, to_sub(f),
                                            part, (double *)log, st_wr, T);
    return flow_dispatch<FwdK, double>(g.D, f.mu, nb, st, g, (const double *)Uin, (double *)Uout, to_sub(f), part,
                                       (double *)log, st_wr, T);
}

hipError_t launch_flow_backward(const Geom &g, void *U, const FlowSub &f, int mode, double *gr, double c_logj,
                                double *gw, const void *log, double *contrib, const double *lg, const double *st_wr,
                                hipStream_t st)
{
    const unsigned nb = par_blocks(g, f.par);
    const long T = sub_stride(g);
    const unsigned nl = (unsigned)((g.V * g.D + flow::kB - 1) / flow::kB);
    const bool stout = f.kind == kFlowStout;
    if (stout && (!lg || !st_wr)) return hipErrorInvalidValue;
    hipError_t e;
    if (f32_mode(mode)) {
        e = stout ? flow_dispatch<BwdKS, float>(g.D, f.mu, nb, st, g, (float *)U, to_sub(f), gr, c_logj, gw,
                                                (const float *)log, contrib, T, lg, st_wr)
                  : flow_dispatch<BwdK, float>(g.D, f.mu, nb, st, g, (float *)U, to_sub(f), gr, c_logj, gw,
                                               (const float *)log, contrib, T, lg, st_wr);
        if (e == hipSuccess) e = flow_dispatch<GatK, float>(g.D, f.mu, nl, st, g, gr, (const double *)contrib, f.par, T);
    } else {
        e = stout ? flow_dispatch<BwdKS, double>(g.D, f.mu, nb, st, g, (double *)U, to_sub(f), gr, c_logj, gw,
                                                 (const double *)log, contrib, T, lg, st_wr)
                  : flow_dispatch<BwdK, double>(g.D, f.mu, nb, st, g, (double *)U, to_sub(f), gr, c_logj, gw,
                                                (const double *)log, contrib, T, lg, st_wr);
        if (e == hipSuccess)
            e = flow_dispatch<GatK, double>(g.D, f.mu, nl, st, g, gr, (const double *)contrib, f.par, T);
    }
    return e;
}

size_t flow_contrib_doubles(const Geom &g)
{
    return (size_t)(6 * (g.D - 1) + 16 * flow::nplanes(g.D)) * 9 * (flow_log_reals(g) / 18);
}

hipError_t launch_wilson_grad(const Geom &g, const void *U, double coef, int mode, double *gr, double *part,
                              long *np, hipStream_t st)
{
    const long nl = g.V * g.D;
    const unsigned nb = (unsigned)((nl + flow::kB - 1) / flow::kB);
    *np = nb;
    if (f32_mode(mode))
        hipLaunchKernelGGL(flow::k_wilson_grad<float>, dim3(nb), dim3(flow::kB), 0, st, g, (const float *)U, coef, gr,
                           part);
    else
        hipLaunchKernelGGL(flow::k_wilson_grad<double>, dim3(nb), dim3(flow::kB), 0, st, g, (const double *)U, coef, gr,
                           part);
    return hipGetLastError();
}

hipError_t launch_kick_g(const Geom &g, void *P, const double *gr, double eps_c, int mode, hipStream_t st)
{
    const long nl = g.V * g.D;
    const unsigned nb = (unsigned)((nl + flow::kB - 1) / flow::kB);
    if (f32_mode(mode))
        hipLaunchKernelGGL(flow::k_kick_g<float>, dim3(nb), dim3(flow::kB), 0, st, g, (float *)P, gr, eps_c);
    else hipLaunchKernelGGL(flow::k_kick_g<double>, dim3(nb), dim3(flow::kB), 0, st, g, (double *)P, gr, eps_c);
    return hipGetLastError();
}

}  // namespace askja
