/*
 * askja_oracle.c -- CPU FP64 oracle for the SU(N) HMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it to compare the HIP
 * product path against.  The product library never links or calls it.
 *
 * It is a plain-C restatement of the reference algorithm (paro8929/askja),
 * keeping the reference's operation order so that values agree to the last
 * bit where the arithmetic allows:
 *   - complex matrix product = GSL's reference CBLAS zgemm (C = 0; for i, for
 *     k, skip A_ik == 0, for j: C_ij += A_ik * B_kj with the textbook complex
 *     product), called from sunmat.h:81-114;
 *   - matrix scaling by a real d is a complex scale by (d, 0), division by d
 *     is a scale by (1./d, 0) (sunmat.h:45-59, 143-169, GSL
 *     gsl_matrix_complex_scale);
 *   - no fused multiply-add anywhere: build with -ffp-contract=off, as the
 *     reference's generic x86-64 build has no FMA.
 * Third-party arithmetic restated here: GSL (version unpinned by the
 * reference, INSTALL:5): reference CBLAS zgemm, gsl_matrix_complex_scale,
 * the mt19937 generator (gsl_rng_set seeding = Matsumoto-Nishimura
 * init_genrand, seed 0 -> 4357), gsl_rng_uniform = x / 2^32,
 * gsl_rng_uniform_pos (rejects 0) and gsl_ran_gaussian (Knuth's polar
 * method, TAOCP vol. 2 algorithm P, returning sigma*y*sqrt(-2 ln r2 / r2)).
 *
 * Functions without a reference counterpart (new rows of the build) are
 * defined here and pinned by mathematics in the tests: Philox-4x32-10
 * (Salmon et al., SC'11) momenta, Haar-random SU(N) starts (Mezzadri 2007 via
 * Gram-Schmidt), reunitarisation, compensated sums.
 *
 * Layout: AoS exactly as the reference's U-*.conf: for each site in
 * lexicographic order (x[D-1] fastest, structs.h:130-165), for each direction,
 * an N x N complex matrix, row-major, (re, im) interleaved.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXDIM 32

typedef struct {
    int nc, dim, nn;              /* nn = nc*nc */
    int ext[OR_MAXDIM];
    long helparr[OR_MAXDIM];
    long V;
    double beta;
    int level;
    size_t msz;                   /* doubles per matrix = 2*nn */
    double *u, *U, *pi, *Pi, *uconf;
    double *gens;                 /* (nn-1) matrices */
} oracle_t;

/* ------------------------------------------------------------------------ */
/* complex matrix primitives (GSL semantics)                                 */
/* ------------------------------------------------------------------------ */

/* C = A*B.  gslcblas zgemm, NoTrans/NoTrans, alpha = 1, beta = 1 on a zeroed
 * C (sunmat.h:85,95): temp = alpha*A_ik computed as (1*ar - 0*ai, 1*ai +
 * 0*ar); zero temp skipped; C_ij += temp*B_kj. */
static void mm(int n, const double *A, const double *B, double *C)
{
    for (int q = 0; q < 2 * n * n; q++) C[q] = 0.0;
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < n; k++) {
            const double ar = A[2 * (i * n + k)], ai = A[2 * (i * n + k) + 1];
            const double tr = 1.0 * ar - 0.0 * ai;
            const double ti = 1.0 * ai + 0.0 * ar;
            if (tr == 0.0 && ti == 0.0) continue;
            for (int j = 0; j < n; j++) {
                const double br = B[2 * (k * n + j)], bi = B[2 * (k * n + j) + 1];
                C[2 * (i * n + j)] += tr * br - ti * bi;
                C[2 * (i * n + j) + 1] += tr * bi + ti * br;
            }
        }
    }
}

/* B = A^dagger (sunmat.h:251-264). */
static void herm(int n, const double *A, double *B)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            B[2 * (j * n + i)] = A[2 * (i * n + j)];
            B[2 * (j * n + i) + 1] = -A[2 * (i * n + j) + 1];
        }
}

/* gsl_matrix_complex_scale by (xr, xi). */
static void cscale(int n, double *A, double xr, double xi)
{
    for (int q = 0; q < n * n; q++) {
        const double ar = A[2 * q], ai = A[2 * q + 1];
        A[2 * q] = ar * xr - ai * xi;
        A[2 * q + 1] = ar * xi + ai * xr;
    }
}

static void madd(int n, const double *A, const double *B, double *C)
{
    for (int q = 0; q < 2 * n * n; q++) C[q] = A[q] + B[q];
}

static void msub(int n, const double *A, const double *B, double *C)
{
    for (int q = 0; q < 2 * n * n; q++) C[q] = A[q] - B[q];
}

static void mident(int n, double *A)
{
    for (int q = 0; q < 2 * n * n; q++) A[q] = 0.0;
    for (int i = 0; i < n; i++) A[2 * (i * n + i)] = 1.0;
}

/* TrR / Tr: gsl_complex_add from (0,0), i ascending (sunmat.h:280-325). */
static void ctrace(int n, const double *A, double *re, double *im)
{
    double r = 0.0, m = 0.0;
    for (int i = 0; i < n; i++) {
        r = r + A[2 * (i * n + i)];
        m = m + A[2 * (i * n + i) + 1];
    }
    *re = r;
    *im = m;
}

static double trR(int n, const double *A)
{
    double r, m;
    ctrace(n, A, &r, &m);
    return r;
}

/* tr2(A) = TrR(A*A) through a full product (sunmat.h:331-338). */
static double tr2(int n, const double *A, double *work)
{
    mm(n, A, A, work);
    return trR(n, work);
}

/* Adjoint(W) = -i/2 [ (W - W^+) - 1 Tr(W - W^+)/n ] (sunmat.h:355-371),
 * operation by operation. work must hold 3 matrices. */
static void tah(int n, const double *W, double *out, double *work)
{
    const size_t ms = 2 * (size_t)n * n;
    double *H = work, *D = work + ms, *T = work + 2 * ms;
    herm(n, W, H);
    msub(n, W, H, D);                       /* dummy = dummy - Hermitian(abc) */
    double trr, tri;
    ctrace(n, D, &trr, &tri);               /* Tr(dummy) */
    mident(n, T);
    cscale(n, T, trr, tri);                 /* unity*Tr(dummy) */
    cscale(n, T, 1. / (double)n, 0.0);      /* ... / n */
    msub(n, D, T, out);                     /* dummy - ... */
    cscale(n, out, 1. / 2., 0.0);           /* (...)/2. */
    cscale(n, out, 0.0, -1.0);              /* temp*minusi */
}

/* mexpi(abc, L): Taylor exp of i*abc to order L by Horner
 * (sunmat.h:200-248). abc is already scaled by eps. work: 3 matrices. */
static void mexpi(int n, const double *abc, int level, double *out, double *work)
{
    const size_t ms = 2 * (size_t)n * n;
    double *rhs = work, *X = work + ms, *P = work + 2 * ms;
    if (level < 0 || level > 8) level = 8;
    memcpy(rhs, abc, ms * sizeof(double));
    cscale(n, rhs, 0.0, 1.0);               /* rhs *= i */
    if (level == 0) {
        mident(n, out);
        return;
    }
    if (level == 1) {
        mident(n, out);
        madd(n, out, rhs, out);
        return;
    }
    /* innermost: X = temp + rhs/L */
    memcpy(P, rhs, ms * sizeof(double));
    cscale(n, P, 1. / (double)level, 0.0);
    mident(n, X);
    madd(n, X, P, X);
    for (int k = level - 1; k >= 2; k--) {
        /* X = temp + (rhs/k) * X */
        memcpy(P, rhs, ms * sizeof(double));
        cscale(n, P, 1. / (double)k, 0.0);
        mm(n, P, X, out);
        mident(n, X);
        madd(n, X, out, X);
    }
    /* temp + rhs * X */
    mm(n, rhs, X, P);
    mident(n, out);
    madd(n, out, P, out);
}

/* ------------------------------------------------------------------------ */
/* lattice geometry (structs.h:109-233)                                      */
/* ------------------------------------------------------------------------ */

long oracle_getplus(const oracle_t *o, long ind, int dir)
{
    long temp;
    if (dir == 0) temp = ind / o->helparr[0];
    else {
        temp = ind % o->helparr[dir - 1];
        temp /= o->helparr[dir];
    }
    int nb = (int)temp;
    nb++;
    nb %= o->ext[dir];
    return ind + (nb - temp) * o->helparr[dir];
}

long oracle_getminus(const oracle_t *o, long ind, int dir)
{
    long temp;
    if (dir == 0) temp = ind / o->helparr[0];
    else {
        temp = ind % o->helparr[dir - 1];
        temp /= o->helparr[dir];
    }
    int nb = (int)temp;
    nb--;
    nb %= o->ext[dir];
    if (nb < 0) nb += o->ext[dir];
    return ind + (nb - temp) * o->helparr[dir];
}

long oracle_getind(const oracle_t *o, const int *x)
{
    long t = 0;
    for (int i = 0; i < o->dim; i++) t += x[i] * o->helparr[i];
    return t;
}

void oracle_getcoord(const oracle_t *o, long ind, int *x)
{
    long temp;
    x[0] = (int)(ind / o->helparr[0]);
    temp = ind % o->helparr[0];
    for (int i = 1; i < o->dim; i++) {
        x[i] = (int)(temp / o->helparr[i]);
        temp %= o->helparr[i];
    }
}

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */

static double *LNK(const oracle_t *o, double *f, long site, int dir)
{
    return f + ((size_t)site * o->dim + dir) * o->msz;
}

void oracle_generators(int nc, double *gen);

oracle_t *oracle_create(int nc, int dim, const int *ext, double beta, int level)
{
    if (nc < 1 || dim < 1 || dim > OR_MAXDIM) return NULL;
    oracle_t *o = (oracle_t *)calloc(1, sizeof(oracle_t));
    o->nc = nc;
    o->dim = dim;
    o->nn = nc * nc;
    o->msz = 2 * (size_t)nc * nc;
    o->V = 1;
    for (int i = 0; i < dim; i++) {
        o->ext[i] = ext[i];
        o->V *= ext[i];
    }
    for (int i = 0; i < dim; i++) {
        o->helparr[i] = 1;
        for (int j = i + 1; j < dim; j++) o->helparr[i] *= ext[j];
    }
    o->beta = beta;
    o->level = level;
    size_t fs = (size_t)o->V * dim * o->msz;
    o->u = (double *)calloc(fs, sizeof(double));
    o->U = (double *)calloc(fs, sizeof(double));
    o->pi = (double *)calloc(fs, sizeof(double));
    o->Pi = (double *)calloc(fs, sizeof(double));
    o->uconf = (double *)calloc(fs, sizeof(double));
    o->gens = (double *)calloc((size_t)(o->nn > 1 ? o->nn - 1 : 1) * o->msz, sizeof(double));
    if (!o->u || !o->U || !o->pi || !o->Pi || !o->uconf || !o->gens) return NULL;
    oracle_generators(nc, o->gens);
    return o;
}

void oracle_destroy(oracle_t *o)
{
    if (!o) return;
    free(o->u);
    free(o->U);
    free(o->pi);
    free(o->Pi);
    free(o->uconf);
    free(o->gens);
    free(o);
}

long oracle_volume(const oracle_t *o) { return o->V; }
void oracle_set_beta(oracle_t *o, double beta) { o->beta = beta; }
void oracle_set_level(oracle_t *o, int level) { o->level = level; }

/* 0=u 1=U 2=pi 3=Pi 4=uconf; the caller reads/writes V*D*2N^2 doubles. */
double *oracle_field(oracle_t *o, int which)
{
    switch (which) {
    case 0: return o->u;
    case 1: return o->U;
    case 2: return o->pi;
    case 3: return o->Pi;
    case 4: return o->uconf;
    }
    return NULL;
}

/* ------------------------------------------------------------------------ */
/* generators and momenta (sun.cpp:142-189, 209-252)                         */
/* ------------------------------------------------------------------------ */

void oracle_generators(int nc, double *gen)
{
    const size_t ms = 2 * (size_t)nc * nc;
    const int ng = nc * nc - 1;
    memset(gen, 0, (size_t)(ng > 0 ? ng : 0) * ms * sizeof(double));
    int num = 0;
    for (int i = 0; i < nc; i++)
        for (int j = i + 1; j < nc; j++) {
            double *g0 = gen + num * ms, *g1 = gen + (num + 1) * ms;
            g0[2 * (i * nc + j)] = 1.0;
            g0[2 * (j * nc + i)] = 1.0;
            g1[2 * (i * nc + j) + 1] = -1.0;
            g1[2 * (j * nc + i) + 1] = 1.0;
            num += 2;
        }
    for (int i = 0; i < nc - 1; i++) {
        double *g = gen + num * ms;
        for (int j = 0; j < i + 1; j++) g[2 * (j * nc + j)] = 1.0;
        g[2 * ((i + 1) * nc + i + 1)] = (double)(-i - 1);
        num++;
    }
    double *work = (double *)malloc(ms * sizeof(double));
    for (int k = 0; k < ng; k++) {
        double *g = gen + k * ms;
        mm(nc, g, g, work);
        double norm = sqrt(0.5 / trR(nc, work));
        cscale(nc, g, norm, 0.0);
    }
    free(work);
}

/* MT19937 (Matsumoto & Nishimura 1998), GSL seeding (seed 0 -> 4357). */
typedef struct {
    uint32_t mt[624];
    int mti;
} mt_t;

static void mt_seed(mt_t *s, unsigned long seed)
{
    if (seed == 0) seed = 4357;
    s->mt[0] = (uint32_t)(seed & 0xffffffffUL);
    for (int i = 1; i < 624; i++)
        s->mt[i] = (uint32_t)(1812433253UL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + i);
    s->mti = 624;
}

static uint32_t mt_next(mt_t *s)
{
    static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
    uint32_t y;
    if (s->mti >= 624) {
        int kk;
        for (kk = 0; kk < 624 - 397; kk++) {
            y = (s->mt[kk] & 0x80000000u) | (s->mt[kk + 1] & 0x7fffffffu);
            s->mt[kk] = s->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; kk < 623; kk++) {
            y = (s->mt[kk] & 0x80000000u) | (s->mt[kk + 1] & 0x7fffffffu);
            s->mt[kk] = s->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        y = (s->mt[623] & 0x80000000u) | (s->mt[0] & 0x7fffffffu);
        s->mt[623] = s->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
        s->mti = 0;
    }
    y = s->mt[s->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

static double mt_uniform(mt_t *s) { return mt_next(s) / 4294967296.0; }

static double mt_uniform_pos(mt_t *s)
{
    double x;
    do x = mt_uniform(s);
    while (x == 0);
    return x;
}

static double mt_gaussian(mt_t *s, double sigma)
{
    double x, y, r2;
    do {
        x = -1 + 2 * mt_uniform_pos(s);
        y = -1 + 2 * mt_uniform_pos(s);
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0);
    return sigma * y * sqrt(-2.0 * log(r2) / r2);
}

void oracle_mt_stream(unsigned long seed, int n, uint32_t *out)
{
    mt_t s;
    mt_seed(&s, seed);
    for (int i = 0; i < n; i++) out[i] = mt_next(&s);
}

void oracle_gauss_stream(unsigned long seed, int n, double *out)
{
    mt_t s;
    mt_seed(&s, seed);
    for (int i = 0; i < n; i++) out[i] = mt_gaussian(&s, 1.0);
}

/* first gsl_rng_uniform of a freshly seeded mt19937 (Metropolis draw,
 * sun.cpp:1277-1287) */
double oracle_mt_first_uniform(unsigned long seed)
{
    mt_t s;
    mt_seed(&s, seed);
    return mt_uniform(&s);
}

/* initializeE(seed): pi = sqrt(2N/beta) sum_a xi_a T^a, site -> dir -> a. */
void oracle_initializeE(oracle_t *o, unsigned long seed)
{
    const int nc = o->nc;
    mt_t s;
    mt_seed(&s, seed);
    double *test = (double *)malloc(o->msz * sizeof(double));
    double *tmp = (double *)malloc(o->msz * sizeof(double));
    const double fac = sqrt(2 * nc / o->beta);
    for (long i = 0; i < o->V; i++)
        for (int j = 0; j < o->dim; j++) {
            memset(test, 0, o->msz * sizeof(double));
            for (int n = 0; n < o->nn - 1; n++) {
                double nummi = mt_gaussian(&s, 1.0);
                memcpy(tmp, o->gens + n * o->msz, o->msz * sizeof(double));
                cscale(nc, tmp, nummi, 0.0);
                madd(nc, test, tmp, test);
            }
            double *p = LNK(o, o->pi, i, j);
            memcpy(p, test, o->msz * sizeof(double));
            cscale(nc, p, fac, 0.0);
        }
    free(test);
    free(tmp);
}

void oracle_initializeU(oracle_t *o)
{
    for (long i = 0; i < o->V; i++)
        for (int j = 0; j < o->dim; j++) {
            mident(o->nc, LNK(o, o->u, i, j));
            memcpy(LNK(o, o->U, i, j), LNK(o, o->u, i, j), o->msz * sizeof(double));
        }
}

/* ------------------------------------------------------------------------ */
/* kick and drift (sun.cpp:672-723, 774-801)                                 */
/* ------------------------------------------------------------------------ */

/* Pi = pi - Adjoint(W)*eps, W = sum of the 2(D-1) oriented staples
 * ((U_j(x) U_k(x+j)) U_j(x+k)^+) U_k(x)^+ and ((U_j(x) U_k(x+j-k)^+)
 * U_j(x-k)^+) U_k(x-k), k ascending, up then down. */
void oracle_updateE(oracle_t *o, double eps)
{
    const int nc = o->nc, D = o->dim;
    const size_t ms = o->msz;
#pragma omp parallel
    {
        double *w = (double *)malloc(9 * ms * sizeof(double));
        double *st = w, *t1 = w + ms, *t2 = w + 2 * ms, *t3 = w + 3 * ms, *h = w + 4 * ms;
        double *ad = w + 5 * ms, *wk = w + 6 * ms;
#pragma omp for schedule(dynamic, 64)
        for (long i = 0; i < o->V; i++) {
            for (int j = 0; j < D; j++) {
                long njplus = oracle_getplus(o, i, j);
                memset(st, 0, ms * sizeof(double));
                for (int k = 0; k < D; k++) {
                    long nkplus = oracle_getplus(o, i, k);
                    long nkminus = oracle_getminus(o, i, k);
                    long njkminus = oracle_getminus(o, njplus, k);
                    if (k == j) continue;
                    mm(nc, LNK(o, o->u, i, j), LNK(o, o->u, njplus, k), t1);
                    herm(nc, LNK(o, o->u, nkplus, j), h);
                    mm(nc, t1, h, t2);
                    herm(nc, LNK(o, o->u, i, k), h);
                    mm(nc, t2, h, t3);
                    madd(nc, st, t3, st);
                    herm(nc, LNK(o, o->u, njkminus, k), h);
                    mm(nc, LNK(o, o->u, i, j), h, t1);
                    herm(nc, LNK(o, o->u, nkminus, j), h);
                    mm(nc, t1, h, t2);
                    mm(nc, t2, LNK(o, o->u, nkminus, k), t3);
                    madd(nc, st, t3, st);
                }
                tah(nc, st, ad, wk);
                cscale(nc, ad, eps, 0.0);
                msub(nc, LNK(o, o->pi, i, j), ad, LNK(o, o->Pi, i, j));
            }
        }
        free(w);
    }
}

/* U = mexpi(Pi*eps, LEVEL) * u; u = U; pi = Pi */
void oracle_updateUandcopyDown(oracle_t *o, double eps)
{
    const int nc = o->nc, D = o->dim;
    const size_t ms = o->msz;
#pragma omp parallel
    {
        double *w = (double *)malloc(6 * ms * sizeof(double));
        double *a = w, *T = w + ms, *wk = w + 2 * ms;
#pragma omp for schedule(dynamic, 64)
        for (long i = 0; i < o->V; i++)
            for (int j = 0; j < D; j++) {
                memcpy(a, LNK(o, o->Pi, i, j), ms * sizeof(double));
                cscale(nc, a, eps, 0.0);
                mexpi(nc, a, o->level, T, wk);
                mm(nc, T, LNK(o, o->u, i, j), LNK(o, o->U, i, j));
                memcpy(LNK(o, o->u, i, j), LNK(o, o->U, i, j), ms * sizeof(double));
                memcpy(LNK(o, o->pi, i, j), LNK(o, o->Pi, i, j), ms * sizeof(double));
            }
        free(w);
    }
}

void oracle_copyDown(oracle_t *o)
{
    size_t fs = (size_t)o->V * o->dim * o->msz * sizeof(double);
    memcpy(o->u, o->U, fs);
    memcpy(o->pi, o->Pi, fs);
}

void oracle_takeconfiguration(oracle_t *o)
{
    memcpy(o->uconf, o->u, (size_t)o->V * o->dim * o->msz * sizeof(double));
}

void oracle_loadconfiguration(oracle_t *o)
{
    memcpy(o->u, o->uconf, (size_t)o->V * o->dim * o->msz * sizeof(double));
}

/* stand-alone kernels for unit tests */
void oracle_matmul(int n, const double *A, const double *B, double *C) { mm(n, A, B, C); }
void oracle_tah(int n, const double *W, double *out)
{
    double *wk = (double *)malloc(6 * (size_t)n * n * sizeof(double));
    tah(n, W, out, wk);
    free(wk);
}
void oracle_mexpi(int n, const double *A, int level, double *out)
{
    double *wk = (double *)malloc(6 * (size_t)n * n * sizeof(double));
    mexpi(n, A, level, out, wk);
    free(wk);
}

/* ------------------------------------------------------------------------ */
/* energies and diagnostics (sun.cpp:415-576)                                */
/* ------------------------------------------------------------------------ */

static void plaq(const oracle_t *o, long i, int j, int k, double *P, double *w)
{
    const int nc = o->nc;
    const size_t ms = o->msz;
    long njplus = oracle_getplus(o, i, j);
    long nkplus = oracle_getplus(o, i, k);
    double *t1 = w, *t2 = w + ms, *h = w + 2 * ms;
    mm(nc, LNK(o, o->u, i, j), LNK(o, o->u, njplus, k), t1);
    herm(nc, LNK(o, o->u, nkplus, j), h);
    mm(nc, t1, h, t2);
    herm(nc, LNK(o, o->u, i, k), h);
    mm(nc, t2, h, P);
}

/* EnergyMag: sum_x sum_{j<k} (1 - ReTr P_jk / N), serial (sun.cpp:507-537) */
double oracle_energy_mag(const oracle_t *o)
{
    double *w = (double *)malloc(4 * o->msz * sizeof(double));
    double res = 0;
    for (long i = 0; i < o->V; i++)
        for (int j = 0; j < o->dim; j++)
            for (int k = j + 1; k < o->dim; k++) {
                plaq(o, i, j, k, w + 3 * o->msz, w);
                res += 1 - trR(o->nc, w + 3 * o->msz) / o->nc;
            }
    free(w);
    return res;
}

/* PlaqExp (measurements4d.cpp:885-927): temporal = sum over (0,k), spatial =
 * sum over (j,k), 1 <= j < k. */
void oracle_plaq_exp(const oracle_t *o, double *temporal, double *spatial)
{
    double *w = (double *)malloc(4 * o->msz * sizeof(double));
    double t = 0, s = 0;
    for (long i = 0; i < o->V; i++) {
        for (int j = 1; j < o->dim; j++)
            for (int k = j + 1; k < o->dim; k++) {
                plaq(o, i, j, k, w + 3 * o->msz, w);
                s += 1 - trR(o->nc, w + 3 * o->msz) / o->nc;
            }
        for (int k = 1; k < o->dim; k++) {
            plaq(o, i, 0, k, w + 3 * o->msz, w);
            t += 1 - trR(o->nc, w + 3 * o->msz) / o->nc;
        }
    }
    *temporal = t;
    *spatial = s;
    free(w);
}

/* simpleenergyEl: sum tr2(p) over the chosen field (2 = pi, 3 = Pi) */
double oracle_simple_energy_el(const oracle_t *o, int which)
{
    double *f = which == 2 ? o->pi : o->Pi;
    double *w = (double *)malloc(o->msz * sizeof(double));
    double res = 0;
    for (long i = 0; i < o->V; i++)
        for (int j = 0; j < o->dim; j++) res += tr2(o->nc, LNK(o, f, i, j), w);
    free(w);
    return res;
}

/* EnergyEl: sum 0.25*tr2(pi + Pi) (sun.cpp:560-567) */
double oracle_energy_el(const oracle_t *o)
{
    double *w = (double *)malloc(2 * o->msz * sizeof(double));
    double res = 0;
    for (long i = 0; i < o->V; i++)
        for (int j = 0; j < o->dim; j++) {
            madd(o->nc, LNK(o, o->pi, i, j), LNK(o, o->Pi, i, j), w);
            res += 0.25 * tr2(o->nc, w, w + o->msz);
        }
    free(w);
    return res;
}

double oracle_total_energy(const oracle_t *o)
{
    return 2 * o->nc * oracle_energy_mag(o) + oracle_energy_el(o);
}

/* GaussLaw: tr2 of ONE global sum of Pi(x,j) - U^+ Pi U at x-j (sun.cpp:415-438) */
double oracle_gauss_law(const oracle_t *o)
{
    const int nc = o->nc;
    const size_t ms = o->msz;
    double *w = (double *)malloc(5 * ms * sizeof(double));
    double *temp = w, *h = w + ms, *t1 = w + 2 * ms, *tl = w + 3 * ms, *s = w + 4 * ms;
    memset(temp, 0, ms * sizeof(double));
    for (long i = 0; i < o->V; i++)
        for (int j = 0; j < o->dim; j++) {
            long nm = oracle_getminus(o, i, j);
            herm(nc, LNK(o, o->u, nm, j), h);
            mm(nc, h, LNK(o, o->Pi, nm, j), t1);
            mm(nc, t1, LNK(o, o->u, nm, j), tl);
            madd(nc, temp, LNK(o, o->Pi, i, j), s);
            msub(nc, s, tl, temp);
        }
    double r = tr2(nc, temp, h);
    free(w);
    return r;
}

/* UViol: sum tr2(u u^+ - 1) (sun.cpp:441-471) */
double oracle_uviol(const oracle_t *o)
{
    const int nc = o->nc;
    const size_t ms = o->msz;
    double *w = (double *)malloc(4 * ms * sizeof(double));
    double *h = w, *t = w + ms, *one = w + 2 * ms, *d = w + 3 * ms;
    mident(nc, one);
    double res = 0;
    for (long i = 0; i < o->V; i++)
        for (int j = 0; j < o->dim; j++) {
            herm(nc, LNK(o, o->u, i, j), h);
            mm(nc, LNK(o, o->u, i, j), h, t);
            msub(nc, t, one, d);
            res += tr2(nc, d, t);
        }
    free(w);
    return res;
}

/* ------------------------------------------------------------------------ */
/* Wilson lines                                                              */
/* ------------------------------------------------------------------------ */

/* product of u_dir along direction dir starting at site pos, ext[dir]
 * factors; returns Tr */
static void line_trace(const oracle_t *o, long pos, int dir, double *re, double *im)
{
    const size_t ms = o->msz;
    double *W = (double *)malloc(2 * ms * sizeof(double));
    double *T = W + ms;
    memcpy(W, LNK(o, o->u, pos, dir), ms * sizeof(double));
    for (int i = 0; i < o->ext[dir] - 1; i++) {
        pos = oracle_getplus(o, pos, dir);
        mm(o->nc, W, LNK(o, o->u, pos, dir), T);
        memcpy(W, T, ms * sizeof(double));
    }
    ctrace(o->nc, W, re, im);
    free(W);
}

/* sun.cpp:307-345: |Tr prod_t U_0(t, 0)| / N at the origin */
double oracle_wilson_origin(const oracle_t *o)
{
    double re, im;
    line_trace(o, 0, 0, &re, &im);
    return hypot(re, im) / o->nc;
}

/* measurements4d.cpp:410-466: mean over x1 of |Tr prod U_0|/N.  The loop
 * variable co.x[0] is left at N0-1 by the inner loop, so lines after the
 * first start at x0 = N0-1 (a cyclic shift of the same product). */
double oracle_polyakov_measure(const oracle_t *o)
{
    int co[OR_MAXDIM] = {0};
    double acc = 0;
    for (co[1] = 0; co[1] < o->ext[1]; co[1]++) {
        long pos = oracle_getind(o, co);
        double re, im;
        line_trace(o, pos, 0, &re, &im);
        co[0] = o->ext[0] - 1 > 0 ? o->ext[0] - 1 : 0;
        acc += hypot(re, im) / o->nc;
    }
    return acc / o->ext[1];
}

/* measurements4d.cpp:311-355: mean over x0 of |Tr prod U_dir|/N, with the
 * same loop-variable carry-over in co.x[dir]. */
double oracle_wilson_measure(const oracle_t *o, int dir)
{
    int co[OR_MAXDIM] = {0};
    double acc = 0;
    for (co[0] = 0; co[0] < o->ext[0]; co[0]++) {
        long pos = oracle_getind(o, co);
        double re, im;
        line_trace(o, pos, dir, &re, &im);
        co[dir] = o->ext[dir] - 1 > 0 ? o->ext[dir] - 1 : 0;
        acc += hypot(re, im) / o->nc;
    }
    return acc / o->ext[0];
}

/* Per-spatial-site complex Tr(prod_t U_0)/N for every site with x0 = 0, in
 * lexicographic order of those sites (used for the volume-averaged Polyakov
 * loop of the product path). */
void oracle_polyakov_field(const oracle_t *o, double *out)
{
    long ns = o->V / o->ext[0];
    for (long s = 0; s < ns; s++) {
        double re, im;
        line_trace(o, s, 0, &re, &im);
        out[2 * s] = re / o->nc;
        out[2 * s + 1] = im / o->nc;
    }
}

/* ------------------------------------------------------------------------ */
/* Philox-4x32-10 and counter-based momenta (no reference counterpart)       */
/* ------------------------------------------------------------------------ */

void oracle_philox4x32_10(const uint32_t *ctr_in, const uint32_t *key_in, uint32_t *out)
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* uniform in (0,1) from 32 bits: (x + 0.5) / 2^32 */
static double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

/* Box-Muller with the logarithm and the sine/cosine written as fixed
 * sequences of IEEE double operations (exact scalings, one correctly rounded
 * division, correctly rounded sqrt and fma from the C library, Horner steps),
 * so that a device performing the same steps produces the same bits.
 * log x: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh s =
 * 2 s sum_{k=0}^{12} s^2k/(2k+1), s = (m-1)/(m+1); e log 2 with log 2 split
 * as a 32-bit head and a tail.  (sin, cos)(2 pi u): t = 4u, q = floor(t+1/2),
 * theta = (t-q) pi/2, Taylor polynomials to theta^19 / theta^20, exact
 * quadrant rotation.  Both are checked against the C library to 2 ulp in the
 * tests. */
double oracle_repro_log(double x)
{
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    uint64_t mb = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m;
    memcpy(&m, &mb, sizeof m);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    double f = m - 1.0, s = f / (m + 1.0), z = s * s;
    static const double c[13] = {1.0, 1.0 / 3, 1.0 / 5, 1.0 / 7, 1.0 / 9, 1.0 / 11, 1.0 / 13,
                                 1.0 / 15, 1.0 / 17, 1.0 / 19, 1.0 / 21, 1.0 / 23, 1.0 / 25};
    double p = c[12];
    for (int k = 11; k >= 0; k--) p = fma(p, z, c[k]);
    double lm = (2.0 * s) * p, de = (double)e;
    double hi = de * 0.6931471803691238;
    return hi + fma(de, 1.9082149292705877e-10, lm);
}

void oracle_repro_sincos2pi(double u, double *sn, double *cs)
{
    double t = 4.0 * u, q = floor(t + 0.5), th = (t - q) * 1.5707963267948966, z = th * th;
    /* (-1)^k / (2k+1)! and (-1)^k / (2k)!, nearest doubles */
    static const double ks[10] = {1.0, -0.16666666666666666, 0.008333333333333333, -0.0001984126984126984,
                                  2.7557319223985893e-06, -2.505210838544172e-08, 1.6059043836821613e-10,
                                  -7.647163731819816e-13, 2.8114572543455206e-15, -8.22063524662433e-18};
    static const double kc[11] = {1.0, -0.5, 0.041666666666666664, -0.001388888888888889, 2.48015873015873e-05,
                                  -2.755731922398589e-07, 2.08767569878681e-09, -1.1470745597729725e-11,
                                  4.779477332387385e-14, -1.5619206968586225e-16, 4.110317623312165e-19};
    double ps = ks[9], pc = kc[10];
    for (int k = 8; k >= 0; k--) ps = fma(ps, z, ks[k]);
    for (int k = 9; k >= 0; k--) pc = fma(pc, z, kc[k]);
    double s0 = th * ps, c0 = pc;
    switch ((int)q & 3) {
    case 0: *sn = s0; *cs = c0; break;
    case 1: *sn = c0; *cs = -s0; break;
    case 2: *sn = -s0; *cs = -c0; break;
    default: *sn = -c0; *cs = s0; break;
    }
}

/* Normal deviates for stream `purpose` of link `link` at counter `counter`:
 * Philox(ctr = {block, link, counter_lo, counter_hi ^ purpose<<24}, key =
 * seed) -> 4 uniforms -> 2 Box-Muller pairs (r = sqrt(-2 ln u0),
 * (r cos 2 pi u1, r sin 2 pi u1), then the same for u2, u3). */
void oracle_philox_normals(uint64_t seed, uint64_t counter, uint32_t purpose, uint32_t link,
                           int n, double *out)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int b = 0; b * 4 < n; b++) {
        uint32_t ctr[4] = {(uint32_t)b, link, (uint32_t)counter,
                           (uint32_t)(counter >> 32) ^ (purpose << 24)};
        uint32_t r[4];
        oracle_philox4x32_10(ctr, key, r);
        for (int h = 0; h < 2; h++) {
            double rad = sqrt(-2.0 * oracle_repro_log(u01(r[2 * h]))), sn, cs;
            oracle_repro_sincos2pi(u01(r[2 * h + 1]), &sn, &cs);
            int q = 4 * b + 2 * h;
            if (q < n) out[q] = rad * cs;
            if (q + 1 < n) out[q + 1] = rad * sn;
        }
    }
}

/* Counter-based momentum refresh into pi: link index = dir*V + lexicographic
 * site; xi_a (a in the reference generator order) -> pi = c sum_a xi_a T^a
 * built directly from the closed form of the generators. */
void oracle_mom_philox(oracle_t *o, uint64_t seed, uint64_t counter)
{
    const int nc = o->nc, ng = o->nn - 1;
    const double c = sqrt(2.0 * nc / o->beta);
    double *xi = (double *)malloc((size_t)(ng > 0 ? ng : 1) * sizeof(double));
    for (long i = 0; i < o->V; i++)
        for (int j = 0; j < o->dim; j++) {
            uint32_t link = (uint32_t)(j * o->V + i);
            oracle_philox_normals(seed, counter, 1u, link, ng, xi);
            double *p = LNK(o, o->pi, i, j);
            memset(p, 0, o->msz * sizeof(double));
            int a = 0;
            for (int r = 0; r < nc; r++)
                for (int s = r + 1; s < nc; s++) {
                    double re = 0.5 * c * xi[a], im = 0.5 * c * xi[a + 1];
                    p[2 * (r * nc + s)] = re;
                    p[2 * (r * nc + s) + 1] = -im;
                    p[2 * (s * nc + r)] = re;
                    p[2 * (s * nc + r) + 1] = im;
                    a += 2;
                }
            /* Cartan part: h_m = c xi_m / sqrt(2 (m+1)(m+2)) multiplies
             * diag(1, .., 1, -(m+1), 0, ..), so pi_rr = sum_{m >= r} h_m -
             * r h_{r-1}; summed from the top (m descending), then the
             * subtraction */
            double suffix = 0.0;
            for (int r = nc - 1; r >= 0; r--) {
                double v = suffix;
                if (r >= 1) {
                    int m = r - 1;
                    double h = c * xi[a + m] * sqrt(0.5 / ((double)(m + 1) * (m + 2)));
                    v -= (double)r * h;
                    suffix += h;
                }
                p[2 * (r * nc + r)] = v;
            }
        }
    free(xi);
}

/* ------------------------------------------------------------------------ */
/* SU(N) projection (no reference counterpart)                               */
/* ------------------------------------------------------------------------ */

/* det of an n x n complex matrix by LU with partial pivoting */
static void cdet(int n, const double *A, double *dr, double *di)
{
    double *M = (double *)malloc(2 * (size_t)n * n * sizeof(double));
    memcpy(M, A, 2 * (size_t)n * n * sizeof(double));
    double rr = 1, ri = 0;
    for (int c = 0; c < n; c++) {
        int p = c;
        double best = -1;
        for (int r = c; r < n; r++) {
            double v = hypot(M[2 * (r * n + c)], M[2 * (r * n + c) + 1]);
            if (v > best) {
                best = v;
                p = r;
            }
        }
        if (p != c) {
            for (int k = 0; k < 2 * n; k++) {
                double t = M[2 * c * n + k];
                M[2 * c * n + k] = M[2 * p * n + k];
                M[2 * p * n + k] = t;
            }
            rr = -rr;
            ri = -ri;
        }
        double pr = M[2 * (c * n + c)], pi = M[2 * (c * n + c) + 1];
        double nr = rr * pr - ri * pi, ni = rr * pi + ri * pr;
        rr = nr;
        ri = ni;
        double den = pr * pr + pi * pi;
        if (den == 0) {
            rr = ri = 0;
            break;
        }
        for (int r = c + 1; r < n; r++) {
            double ar = M[2 * (r * n + c)], ai = M[2 * (r * n + c) + 1];
            double fr = (ar * pr + ai * pi) / den, fi = (ai * pr - ar * pi) / den;
            for (int k = c; k < n; k++) {
                double br = M[2 * (c * n + k)], bi = M[2 * (c * n + k) + 1];
                M[2 * (r * n + k)] -= fr * br - fi * bi;
                M[2 * (r * n + k) + 1] -= fr * bi + fi * br;
            }
        }
    }
    *dr = rr;
    *di = ri;
    free(M);
}

/* Modified Gram-Schmidt on the rows, then multiply by exp(-i arg(det)/n). */
void oracle_project_sun(int n, double *A)
{
    for (int r = 0; r < n; r++) {
        double *a = A + 2 * (size_t)r * n;
        for (int q = 0; q < r; q++) {
            const double *b = A + 2 * (size_t)q * n;
            double dr = 0, di = 0; /* <b, a> = sum conj(b) a */
            for (int k = 0; k < n; k++) {
                dr += b[2 * k] * a[2 * k] + b[2 * k + 1] * a[2 * k + 1];
                di += b[2 * k] * a[2 * k + 1] - b[2 * k + 1] * a[2 * k];
            }
            for (int k = 0; k < n; k++) {
                a[2 * k] -= dr * b[2 * k] - di * b[2 * k + 1];
                a[2 * k + 1] -= dr * b[2 * k + 1] + di * b[2 * k];
            }
        }
        double nrm = 0;
        for (int k = 0; k < 2 * n; k++) nrm += a[k] * a[k];
        nrm = 1.0 / sqrt(nrm);
        for (int k = 0; k < 2 * n; k++) a[k] *= nrm;
    }
    double dr, di;
    cdet(n, A, &dr, &di);
    double ph = -atan2(di, dr) / n;
    double cr = cos(ph), ci = sin(ph);
    for (int q = 0; q < n * n; q++) {
        double ar = A[2 * q], ai = A[2 * q + 1];
        A[2 * q] = ar * cr - ai * ci;
        A[2 * q + 1] = ar * ci + ai * cr;
    }
}

void oracle_det(int n, const double *A, double *out) { cdet(n, A, out, out + 1); }

void oracle_reunitarize(oracle_t *o)
{
    size_t nl = (size_t)o->V * o->dim;
    for (size_t l = 0; l < nl; l++) oracle_project_sun(o->nc, o->u + l * o->msz);
}

/* Haar start: Ginibre entries (xi + i eta)/sqrt(2) from the Philox stream
 * (purpose 2, counter = stream), then the SU(N) projection.  Sets u, U and
 * uconf. */
void oracle_haar(oracle_t *o, uint64_t seed, uint64_t stream)
{
    const int nc = o->nc;
    double *g = (double *)malloc(2 * (size_t)o->nn * sizeof(double));
    for (long i = 0; i < o->V; i++)
        for (int j = 0; j < o->dim; j++) {
            uint32_t link = (uint32_t)(j * o->V + i);
            oracle_philox_normals(seed, stream, 2u, link, 2 * o->nn, g);
            double *p = LNK(o, o->u, i, j);
            for (int q = 0; q < 2 * o->nn; q++) p[q] = g[q] * 0.7071067811865476;
            oracle_project_sun(nc, p);
        }
    size_t fs = (size_t)o->V * o->dim * o->msz * sizeof(double);
    memcpy(o->U, o->u, fs);
    memcpy(o->uconf, o->u, fs);
    free(g);
}

/* ------------------------------------------------------------------------ */
/* trajectories                                                              */
/* ------------------------------------------------------------------------ */

/* H = beta*(EnergyMag + sum tr2(p)/2N) on the chosen momentum field */
double oracle_hamiltonian(const oracle_t *o, int which)
{
    return (oracle_energy_mag(o) + oracle_simple_energy_el(o, which) / (2 * o->nc)) * o->beta;
}

/* One MD trajectory with the reference's kick-drift-kick schedule
 * (sun.cpp:1174-1179, 1212-1218, 1249-1251, 1268-1272) starting from
 * (u, pi): updateE(dt/2); ndrift x [updateUandcopyDown(dt); updateE(dt)
 * except after the last drift]; updateE(dt/2).  The final momentum is in Pi.
 * (The copyDown after the very first half kick, sun.cpp:1179, only copies
 * U = u and pi = Pi, which the first drift does anyway.)  The first
 * trajectory of a fresh reference run has ndrift = UPDATE2 - 1, later ones
 * UPDATE2. */
void oracle_md_trajectory(oracle_t *o, double dt, int ndrift)
{
    oracle_updateE(o, dt / 2);
    for (int t = 1; t <= ndrift; t++) {
        oracle_updateUandcopyDown(o, dt);
        if (t < ndrift) oracle_updateE(o, dt);
    }
    oracle_updateE(o, dt / 2);
}

/* Plain HMC chain with counter-based randomness (momenta: Philox purpose 1,
 * counter = trajectory number; Metropolis uniform: Philox purpose 3, link
 * 0, counter = trajectory number).  Per trajectory it records
 * rec[0..9] = {dH, accepted, S/beta-normalised plaquette <ReTr P/N>,
 * Re P, Im P (volume-averaged Polyakov loop), |P|, H_old, H_new,
 * measure4d's spatial Wilson line W_1 and Polyakov loop P-bar}.
 * Reunitarises the chain state before each trajectory when reunit != 0. */
void oracle_hmc(oracle_t *o, double dt, int ndrift, int ntraj, uint64_t seed,
                uint64_t first_counter, int reunit, double *rec)
{
    const long ns = o->V / o->ext[0];
    double *pf = (double *)malloc(2 * (size_t)ns * sizeof(double));
    const long nplaq = o->V * o->dim * (o->dim - 1) / 2;
    for (int t = 0; t < ntraj; t++) {
        uint64_t ctr = first_counter + (uint64_t)t;
        if (reunit) oracle_reunitarize(o);
        oracle_takeconfiguration(o);
        oracle_mom_philox(o, seed, ctr);
        double h0 = oracle_hamiltonian(o, 2);
        oracle_md_trajectory(o, dt, ndrift);
        double h1 = oracle_hamiltonian(o, 3);
        double dh = h1 - h0;
        uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        uint32_t cc[4] = {0u, 0u, (uint32_t)ctr, (uint32_t)(ctr >> 32) ^ (3u << 24)};
        uint32_t r[4];
        oracle_philox4x32_10(cc, key, r);
        double uni = u01(r[0]);
        int acc = (h1 < h0) || (uni < exp(h0 - h1));
        if (!acc) oracle_loadconfiguration(o);
        double *R = rec + 10 * (size_t)t;
        R[0] = dh;
        R[1] = acc;
        R[2] = 1.0 - oracle_energy_mag(o) / nplaq;
        oracle_polyakov_field(o, pf);
        double pr = 0, pim = 0;
        for (long s = 0; s < ns; s++) {
            pr += pf[2 * s];
            pim += pf[2 * s + 1];
        }
        pr /= ns;
        pim /= ns;
        R[3] = pr;
        R[4] = pim;
        R[5] = hypot(pr, pim);
        R[6] = h0;
        R[7] = h1;
        R[8] = o->dim > 1 ? oracle_wilson_measure(o, 1) : 0.0; /* measurements4d.cpp:311-355, dir 1 */
        R[9] = o->dim > 1 ? oracle_polyakov_measure(o) : 0.0;  /* measurements4d.cpp:410-466 */
    }
    free(pf);
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
