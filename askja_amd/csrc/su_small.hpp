// su_small.hpp -- register-resident N x N complex matrices for N <= 3, one
// matrix per lane.  Two arithmetic flavours:
//   Exact = true : the reference's operation order (sunmat.h:81-114 with
//                  GSL's reference zgemm: C_ij = 0, then C_ij += A_ik B_kj for
//                  k ascending, complex product (ar br - ai bi, ar bi + ai br)),
//                  no FMA contraction -> bit-identical to the reference's
//                  double arithmetic;
//   Exact = false: FMA chains.
#pragma once
#include "common.hpp"

namespace askja {

template <int N, typename R>
struct CM {
    R re[N][N];
    R im[N][N];
};

template <int N, typename R>
__device__ inline void cm_zero(CM<N, R> &m)
{
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            m.re[i][j] = R(0);
            m.im[i][j] = R(0);
        }
}

// ---- loads/stores in the 64-site tiled SoA layout -------------------------
// storage type S, arithmetic type R (S = R except for FP64 arithmetic on FP32
// links)
template <int N, typename S, typename R>
__device__ inline void load_link(const S *__restrict__ f, long l, CM<N, R> &m)
{
    const S *b = f + (size_t)(l >> 6) * (2 * N * N * 64) + (l & 63);
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            m.re[i][j] = b[((i * N + j) * 2) * 64];
            m.im[i][j] = b[((i * N + j) * 2 + 1) * 64];
        }
}

template <int N, typename R>
__device__ inline void store_link(R *__restrict__ f, long l, const CM<N, R> &m)
{
    R *b = f + (size_t)(l >> 6) * (2 * N * N * 64) + (l & 63);
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            b[((i * N + j) * 2) * 64] = m.re[i][j];
            b[((i * N + j) * 2 + 1) * 64] = m.im[i][j];
        }
}

// momentum: compact Hermitian, N*N reals per link
template <int N, typename R>
__device__ inline void load_mom(const R *__restrict__ f, long l, CM<N, R> &m)
{
    const R *b = f + (size_t)(l >> 6) * (N * N * 64) + (l & 63);
#pragma unroll
    for (int i = 0; i < N; i++) {
        m.re[i][i] = b[i * 64];
        m.im[i][i] = R(0);
    }
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = i + 1; j < N; j++) {
            const int q = herm_upper_index(N, i, j);
            const R r = b[q * 64], s = b[(q + 1) * 64];
            m.re[i][j] = r;
            m.im[i][j] = s;
            m.re[j][i] = r;
            m.im[j][i] = -s;
        }
}

template <int N, typename R>
__device__ inline void store_mom(R *__restrict__ f, long l, const CM<N, R> &m)
{
    R *b = f + (size_t)(l >> 6) * (N * N * 64) + (l & 63);
#pragma unroll
    for (int i = 0; i < N; i++) b[i * 64] = m.re[i][i];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = i + 1; j < N; j++) {
            const int q = herm_upper_index(N, i, j);
            b[q * 64] = m.re[i][j];
            b[(q + 1) * 64] = m.im[i][j];
        }
}

// ---- products ---------------------------------------------------------------
// C = op(A) op(B), op = identity or conjugate transpose (DA / DB).
template <bool DA, bool DB, int N, typename R>
__device__ inline void mul_exact(const CM<N, R> &A, const CM<N, R> &B, CM<N, R> &C)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            R cr = R(0), ci = R(0);
#pragma unroll
            for (int k = 0; k < N; k++) {
                const R ar = DA ? A.re[k][i] : A.re[i][k];
                const R ai = DA ? -A.im[k][i] : A.im[i][k];
                const R br = DB ? B.re[j][k] : B.re[k][j];
                const R bi = DB ? -B.im[j][k] : B.im[k][j];
                cr = cr + (ar * br - ai * bi);
                ci = ci + (ar * bi + ai * br);
            }
            C.re[i][j] = cr;
            C.im[i][j] = ci;
        }
}

// C (+)= op(A) op(B) with FMA; Acc selects accumulate into C.
template <bool DA, bool DB, bool Acc, int N, typename R>
__device__ inline void mul_fast(const CM<N, R> &A, const CM<N, R> &B, CM<N, R> &C)
{
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            R cr = Acc ? C.re[i][j] : R(0), ci = Acc ? C.im[i][j] : R(0);
#pragma unroll
            for (int k = 0; k < N; k++) {
                const R ar = DA ? A.re[k][i] : A.re[i][k];
                const R ai = DA ? -A.im[k][i] : A.im[i][k];
                const R br = DB ? B.re[j][k] : B.re[k][j];
                const R bi = DB ? -B.im[j][k] : B.im[k][j];
                cr = fma(ar, br, cr);
                cr = fma(-ai, bi, cr);
                ci = fma(ar, bi, ci);
                ci = fma(ai, br, ci);
            }
            C.re[i][j] = cr;
            C.im[i][j] = ci;
        }
}

template <int N, typename R>
__device__ inline void cm_add_exact(CM<N, R> &S, const CM<N, R> &P)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            S.re[i][j] = S.re[i][j] + P.re[i][j];
            S.im[i][j] = S.im[i][j] + P.im[i][j];
        }
}

// Pout = pin - eps * TAH(W), TAH(W) = -i/2 [(W - W^+) - Tr(W - W^+)/N]
// (sunmat.h:355-371, sun.cpp:716).  In the exact flavour every rounding of the
// reference is reproduced: D_ii.im = W_ii.im + W_ii.im, tri = sum D_ii.im,
// M_ii.im = D_ii.im - tri*(1/N), F = ((0.5*M.im)*eps, -(0.5*M.re)*eps).
template <bool Exact, int N, typename R>
__device__ inline void tah_kick(const CM<N, R> &W, const CM<N, R> &pin, R eps, R inv_n, CM<N, R> &out)
{
#pragma clang fp contract(off)
    R tri = R(0);
#pragma unroll
    for (int i = 0; i < N; i++) tri = tri + (W.im[i][i] + W.im[i][i]);
    const R corr = tri * inv_n;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const R mi = (W.im[i][i] + W.im[i][i]) - corr;
        const R fr = (R(0.5) * mi) * eps;
        out.re[i][i] = pin.re[i][i] - fr;
        out.im[i][i] = R(0);
    }
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = i + 1; j < N; j++) {
            const R dr = W.re[i][j] - W.re[j][i];
            const R di = W.im[i][j] + W.im[j][i];
            const R fr = (R(0.5) * di) * eps;
            const R fi = (R(0.5) * dr) * eps;  // F_ij.im = -fi
            const R pr = pin.re[i][j] - fr;
            const R pi = pin.im[i][j] + fi;
            out.re[i][j] = pr;
            out.im[i][j] = pi;
            out.re[j][i] = pr;
            out.im[j][i] = -pi;
        }
}

// T_L(i*eps*P) * U by Horner (sunmat.h:200-248): rhs = i*(eps*P);
// X = 1 + rhs/L; X = 1 + (rhs/k) X for k = L-1..2; T = 1 + rhs X.
// cinv[k] = 1./k computed on the host.
template <bool Exact, int N, typename R>
__device__ inline void taylor_drift(const CM<N, R> &P, R eps, int level, const R *cinv,
                                    const CM<N, R> &U, CM<N, R> &out)
{
#pragma clang fp contract(off)
    CM<N, R> rhs, X, Q;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            rhs.re[i][j] = -(P.im[i][j] * eps);
            rhs.im[i][j] = P.re[i][j] * eps;
        }
    if (level < 0 || level > 8) level = 8;
    if (level == 0) {
        out = U;
        return;
    }
    if (level == 1) {
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) {
                X.re[i][j] = (i == j ? R(1) : R(0)) + rhs.re[i][j];
                X.im[i][j] = R(0) + rhs.im[i][j];
            }
    } else {
        const R cl = cinv[level];
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) {
                X.re[i][j] = (i == j ? R(1) : R(0)) + rhs.re[i][j] * cl;
                X.im[i][j] = R(0) + rhs.im[i][j] * cl;
            }
        for (int k = level - 1; k >= 1; k--) {
            const R ck = cinv[k];
            CM<N, R> S;
#pragma unroll
            for (int i = 0; i < N; i++)
#pragma unroll
                for (int j = 0; j < N; j++) {
                    S.re[i][j] = rhs.re[i][j] * ck;  // ck = 1 for k = 1: exact
                    S.im[i][j] = rhs.im[i][j] * ck;
                }
            if (Exact) mul_exact<false, false>(S, X, Q);
            else mul_fast<false, false, false>(S, X, Q);
#pragma unroll
            for (int i = 0; i < N; i++)
#pragma unroll
                for (int j = 0; j < N; j++) {
                    X.re[i][j] = (i == j ? R(1) : R(0)) + Q.re[i][j];
                    X.im[i][j] = R(0) + Q.im[i][j];
                }
        }
    }
    if (Exact) mul_exact<false, false>(X, U, out);
    else mul_fast<false, false, false>(X, U, out);
}

// Fast Horner form of T_L(i eps P) U: with rhs = i eps P, every step
// X <- 1 + (rhs/k) X is evaluated as Q = P X, X = 1 + i (eps/k) Q, so only
// P, X and Q are live (FMA flavour; the exact flavour is taylor_drift).
template <int N, typename R>
__device__ inline void taylor_drift_fast(const CM<N, R> &P, R eps, int level, const R *cinv, const CM<N, R> &U,
                                         CM<N, R> &out)
{
    if (level < 0 || level > 8) level = 8;
    if (level == 0) {
        out = U;
        return;
    }
    CM<N, R> X, Q;
    R f = eps * (level == 1 ? R(1) : cinv[level]);
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            X.re[i][j] = (i == j ? R(1) : R(0)) - f * P.im[i][j];
            X.im[i][j] = f * P.re[i][j];
        }
    for (int k = level - 1; k >= 1; k--) {
        mul_fast<false, false, false>(P, X, Q);
        f = eps * cinv[k];
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int j = 0; j < N; j++) {
                X.re[i][j] = (i == j ? R(1) : R(0)) - f * Q.im[i][j];
                X.im[i][j] = f * Q.re[i][j];
            }
    }
    mul_fast<false, false, false>(X, U, out);
}

// ReTr(A B^+) = sum_ij Re(A_ij conj(B_ij)) (FMA flavour).
template <int N, typename R>
__device__ inline R retr_abh(const CM<N, R> &A, const CM<N, R> &B)
{
    R t = R(0);
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            t = fma(A.re[i][j], B.re[i][j], t);
            t = fma(A.im[i][j], B.im[i][j], t);
        }
    return t;
}

// TrR: real part of the trace, i ascending from 0 (sunmat.h:280-292).
template <int N, typename R>
__device__ inline R retr(const CM<N, R> &A)
{
#pragma clang fp contract(off)
    R t = R(0);
#pragma unroll
    for (int i = 0; i < N; i++) t = t + A.re[i][i];
    return t;
}

template <int N, typename R>
__device__ inline R imtr(const CM<N, R> &A)
{
#pragma clang fp contract(off)
    R t = R(0);
#pragma unroll
    for (int i = 0; i < N; i++) t = t + A.im[i][i];
    return t;
}

// Squared Frobenius norm of a compact Hermitian momentum = Tr P^2.
template <int N, typename R>
__device__ inline double herm_norm2(const CM<N, R> &P)
{
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) t += (double)P.re[i][i] * (double)P.re[i][i];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = i + 1; j < N; j++)
            t += 2.0 * ((double)P.re[i][j] * (double)P.re[i][j] + (double)P.im[i][j] * (double)P.im[i][j]);
    return t;
}

}  // namespace askja
