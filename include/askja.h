/*
 * askja.h -- C ABI of the MI355X SU(N) HMC engine (libaskja.so).
 *
 * The reference (paro8929/askja) has no library API: its hot path is a set of
 * void functions over global arrays in sun.cpp.  Each entry point below
 * replaces one of them (cited as file:line of the reference), operating on
 * device-resident fields owned by a context.  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every call returns ASKJA_OK (0) or a negative ASKJA_E* code; the message
 *    is kept per context (askja_last_error).  Errors are sticky: after a
 *    failure every further call on that context returns ASKJA_ESTICKY until
 *    askja_clear_error.
 *  - One host thread per context.  Kernels run on the context's HIP stream;
 *    only calls that return scalars or fill host buffers synchronise.
 *  - Host buffers are caller-owned and copied.  Links and momenta cross the
 *    boundary in the reference's U-*.conf layout: for each site in
 *    lexicographic order (x[D-1] fastest), for each direction mu = 0..D-1, an
 *    N x N complex matrix, row-major, (re, im) interleaved doubles
 *    (sun.cpp:289-302, structs.h:130-165).  Length askja_field_doubles().
 *  - Device layout is private: links grouped by direction, sites in
 *    checkerboard order (even-parity sites, then odd; within a parity sorted
 *    by slab of the longest direction, then lexicographically);
 *    for N <= 3 the links are tiled in groups of 64 (one wavefront) with each
 *    matrix element contiguous across the group, for N >= 4 every matrix is
 *    contiguous.  Momenta are stored as compact Hermitian matrices (N^2 reals).
 */
#ifndef ASKJA_H
#define ASKJA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct askja_ctx askja_ctx;

/* precision modes */
enum {
    ASKJA_FP64 = 0,       /* FP64 storage and arithmetic, FMA, factored staples */
    ASKJA_FP64_EXACT = 1, /* FP64 in the reference's operation order, no FMA:
                             links and momenta bit-identical to the reference's
                             double arithmetic (sunmat.h:81-371, sun.cpp:672-801) */
    ASKJA_FP32 = 2,       /* FP32 links/momenta/MD; FP64 compensated sums,
                             FP64 reunitarisation, and an FP64 Metropolis test
                             whose energies are the FP64-arithmetic action and
                             kinetic energy of the stored FP32 fields */
    ASKJA_BF16X3 = 3      /* as FP32, but for N >= 9 the kick/drift link products
                             run on MFMA as split BF16 (hi*hi + hi*lo + lo*hi)
                             with FP32 accumulation */
};

/* error codes */
enum {
    ASKJA_OK = 0,
    ASKJA_EINVAL = -1,   /* bad argument */
    ASKJA_EHIP = -2,     /* HIP runtime error */
    ASKJA_ENOMEM = -3,   /* device allocation failed */
    ASKJA_EUNSUP = -4,   /* unsupported N / mode combination */
    ASKJA_ESTICKY = -5,  /* an earlier call on this context failed */
    ASKJA_EIO = -6       /* file read/write failed */
};

/* ---- context ----------------------------------------------------------------- */
/* allocate_memory + generators + xygrid.init (sun.cpp:96-125, 142-189,
 * 1086-1097; structs.h:109-147).  extents[0..dim-1] = N0..N_{D-1}; extents of
 * 1 are allowed.  1 <= nc <= 64, 1 <= dim <= 32.  Links start cold. */
int askja_create(int nc, int dim, const int *extents, int precision, askja_ctx **out);
/* free_memory (sun.cpp:128-139) */
int askja_destroy(askja_ctx *ctx);
/* Message of the last failure on ctx ("" if none); valid until the next call. */
const char *askja_last_error(const askja_ctx *ctx);
int askja_clear_error(askja_ctx *ctx);
int askja_info(const askja_ctx *ctx, long *volume, int *nc, int *dim, int *precision);
/* Number of doubles in one host field buffer (V*D*N*N*2). */
size_t askja_field_doubles(const askja_ctx *ctx);

/* ---- links ------------------------------------------------------------------- */
/* loadconf (sun.cpp:744-765), its in-memory half: links <- host buffer */
int askja_links_upload(askja_ctx *ctx, const double *aos);
/* storeconf (sun.cpp:289-302), its in-memory half: host buffer <- links */
int askja_links_download(askja_ctx *ctx, double *aos);
/* Asynchronous askja_links_download for writing configurations while the
 * chain continues: the AoS conversion runs on the context's stream (ordered
 * after everything queued so far), the device-to-host copy into `pinned`
 * (from askja_host_alloc, askja_field_doubles long) on a second stream.
 * Complete when askja_download_wait returns; `pinned` must not be read
 * before.  A new download first waits for the previous copy. */
int askja_links_download_async(askja_ctx *ctx, double *pinned);
int askja_download_wait(askja_ctx *ctx);
/* Page-locked host buffers for askja_links_download_async. */
int askja_host_alloc(size_t ndoubles, double **out);
int askja_host_free(double *p);
/* initializeU (sun.cpp:194-207): cold start, all links = identity */
int askja_links_cold(askja_ctx *ctx);
/* Haar-random SU(N) start (no reference counterpart): Ginibre matrix from
 * Philox-4x32-10 (key = seed, counter = stream), then the projection of
 * askja_reunitarize. */
int askja_links_haar(askja_ctx *ctx, uint64_t seed, uint64_t stream);
/* Project every link onto SU(N) in FP64 (no reference counterpart; the
 * reference only monitors unitarity, sun.cpp:441-471): modified Gram-Schmidt
 * on the rows, then multiplication by exp(-i arg(det)/N). */
int askja_reunitarize(askja_ctx *ctx);
/* takeconfiguration / loadconfiguration (sun.cpp:725-742):
 * op 0 = save links to the snapshot, 1 = restore links from it. */
int askja_snapshot(askja_ctx *ctx, int op);

/* ---- momenta ------------------------------------------------------------------ */
/* momenta <- host buffer of Hermitian matrices (links layout); the momentum
 * before the last kick is set to the same field. */
int askja_mom_upload(askja_ctx *ctx, const double *aos);
/* host buffer <- current momentum field (the reference's Pi) */
int askja_mom_download(askja_ctx *ctx, double *aos);
/* host buffer <- momentum before the last askja_kick (the reference's pi) */
int askja_mom_download_prev(askja_ctx *ctx, double *aos);
/* initializeE(seed) (sun.cpp:209-252) exactly as the reference computes it:
 * GSL's mt19937 seeded with `seed`, Gaussian deviates by the polar method,
 * order site -> direction -> generator, pi = sqrt(2N/beta) sum_a xi_a T^a
 * with the reference's roundings.  Generated on the host, then uploaded. */
int askja_mom_gsl(askja_ctx *ctx, unsigned long seed, double beta);
/* initializeE, counter-based form (used for every trajectory after the
 * first): xi from Philox-4x32-10 (key = seed, counter = counter) and
 * Box-Muller, generated on the device.  The Box-Muller logarithm and
 * sine/cosine are fixed sequences of IEEE double operations, so the momenta
 * are reproducible bit for bit by any IEEE host performing the same steps. */
int askja_mom_refresh(askja_ctx *ctx, uint64_t seed, uint64_t counter, double beta);
/* Flip the sign of the momenta (reversibility checks). */
int askja_mom_negate(askja_ctx *ctx);

/* ---- molecular dynamics -------------------------------------------------------- */
/* updateE (sun.cpp:672-723): Pi = pi - eps * TAH(U_mu(x) sum_nu staples), with
 * the momentum before the kick kept for askja_mom_download_prev and the E_L
 * and Gauss-law diagnostics. */
int askja_kick(askja_ctx *ctx, double eps);
/* updateUandcopyDown (sun.cpp:774-790): U = T_L(i eps Pi) U with T_L the
 * order-`level` Taylor exp by Horner (sunmat.h:200-248; level outside 0..8
 * means 8). */
int askja_drift(askja_ctx *ctx, double eps, int level);
/* A kick followed by the drift of the same links in one pass over the
 * lattice (the pair updateE(eps_kick); updateUandcopyDown(eps_drift)). */
int askja_kick_drift(askja_ctx *ctx, double eps_kick, double eps_drift, int level);
/* The MD part of one trajectory (sun.cpp:1178-1218, 1249-1251):
 * kick(dt/2), ndrift x [drift(dt); kick(dt) except after the last drift],
 * kick(dt/2).  The reference's first trajectory has ndrift = UPDATE2-1.
 * energies (may be NULL) receives {S(start), sum Tr pi^2 (start), S(end),
 * sum Tr Pi^2 (end)} with S = EnergyMag (sun.cpp:507-537), in the FP32 and
 * BF16x3 modes the FP64-arithmetic action of the stored links. */
int askja_md_trajectory(askja_ctx *ctx, double dt, int ndrift, int level, double *energies);

/* ---- reductions (FP64, compensated) ------------------------------------------- */
/* EnergyMag (sun.cpp:507-537) split as PlaqExp (measurements4d.cpp:885-927):
 * s_t = sum_x sum_{k>=1} (1 - ReTr P_0k/N), s_s = sum over 1 <= j < k.
 * EnergyMag = s_t + s_s.  Either pointer may be NULL. */
int askja_action(askja_ctx *ctx, double *s_t, double *s_s);
/* simpleenergyEl (sun.cpp:540-548): sum Tr(Pi^2) of the current momenta */
int askja_kinetic(askja_ctx *ctx, double *out);
/* EnergyEl (sun.cpp:560-567): 1/4 sum Tr (pi + Pi)^2 */
int askja_energy_el(askja_ctx *ctx, double *out);
/* GaussLaw (sun.cpp:415-438): tr2 of sum_{x,mu} [Pi_mu(x) - U^+ Pi U at x-mu] */
int askja_gauss_law(askja_ctx *ctx, double *out);
/* UViol (sun.cpp:441-471): sum over links of tr2(U U^+ - 1) */
int askja_uviol(askja_ctx *ctx, double *out);

/* ---- Wilson lines ----------------------------------------------------------------- */
/* Traces Tr(prod_{t} U_dir(x + t dir))/N of every straight line in direction
 * dir that starts at x_dir = 0, ordered lexicographically by start site.
 * out holds 2*V/extent[dir] doubles (re, im).  dir = 0 gives the Polyakov
 * loops: WilsonLoop (sun.cpp:307-345) is |entry 0|, PolyakovLoop
 * (measurements4d.cpp:410-466) the mean of |.| over x1 at x2.. = 0, and
 * WilsonLoop(evs, dir) (measurements4d.cpp:311-355) the mean over x0. */
int askja_wilson_lines(askja_ctx *ctx, int dir, double *out);

/* The N x N product matrix (host buffer, 2 N^2 doubles) of the straight line
 * in direction dir starting at coordinates x[0..D-1]: the matrix W of
 * WilsonLoop (sun.cpp:307-345) for x = 0, dir = 0, whose eigen-phases fill
 * the Wilson-*.dat columns. */
int askja_line_matrix(askja_ctx *ctx, int dir, const int *x, double *out);

/* ---- measurement (the device side of measurements4d.cpp) ------------------------- */
/* scalarLoopcorr (measurements4d.cpp:357-408): out[x0] = mean over the sites
 * of time slice x0 of ReTr U_dir(x) / N (the reference uses dir = DIM-1).
 * out holds extent[0] doubles. */
int askja_slice_traces(askja_ctx *ctx, int dir, double *out);
/* getevs (measurements4d.cpp:185-309): for each of nlines straight lines in
 * direction dir starting at the lexicographic sites starts[k], the N
 * eigen-phases arg(lambda) of the line product, shifted by 2 pi w / N when the
 * winding w = sum(phases) / 2 pi has |w| > 0.1, sorted ascending.  out holds
 * nlines * N doubles.  (Batched on the device: one workgroup per line.) */
int askja_line_phases(askja_ctx *ctx, int dir, int nlines, const int64_t *starts, double *out);

/* ---- HMC -------------------------------------------------------------------------- */
typedef struct {
    double beta;
    double dt;
    int ndrift;      /* drifts per trajectory (UPDATE2) */
    int level;       /* Taylor order */
    int reunit;      /* project the chain state onto SU(N) before the trajectory */
    uint64_t seed;   /* Philox key for momenta and the Metropolis draw */
} askja_hmc_params;

typedef struct {
    double h_old, h_new, dh;
    double uniform;    /* Metropolis draw */
    double action;     /* EnergyMag of the chain state after the decision */
    int accepted;
    int accept_case;   /* 1 = H decreased, 2 = uniform < exp(-dH), 0 = rejected */
} askja_hmc_result;

/* One HMC trajectory of the reference's loop (sun.cpp:1249-1318) with a
 * counter-based stream in place of the time(NULL) seeds: momentum refresh
 * (counter = traj_counter), H_old, MD, H_new, accept if H_new < H_old or
 * u < exp(H_old - H_new) with u from Philox; on reject the links return to
 * the state before the trajectory.  NaN energies reject, as in the reference.
 * In the FP32 and BF16x3 modes H_old and H_new are the FP64-arithmetic
 * energies of the stored FP32 fields (an extra action pass at each end), so
 * the accept test itself carries no reduced-precision error; the chain is
 * still not exact there, because FP32/BF16x3 MD is reversible and
 * area-preserving only to its rounding (and reunit projects the FP32 links
 * outside the accept test). */
int askja_hmc_trajectory(askja_ctx *ctx, const askja_hmc_params *p, uint64_t traj_counter,
                         askja_hmc_result *res);

/* askja_hmc_trajectory without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate, such as code under HIP graph
 * capture or a tight serving loop.  The same kernels and the same results,
 * bit for bit; their block partial sums and finalised sums go to a static
 * device buffer of 16 MiB per GPU (2^21 doubles of partials), held until the
 * process ends, instead of the context's reduction buffer, which
 * askja_hmc_trajectory grows with hipMalloc on first use.  Lattices whose
 * partials exceed the buffer are rejected with ASKJA_EINVAL (kick partials:
 * 6 doubles per 256-site block and direction for N <= 3, per link for
 * N >= 9: SU(3) 8x32^3, SU(16) 6x24^3 and SU(64) 128^2 fit).  The context's device must be
 * current.  All calls on one GPU share that GPU's buffer: calls that overlap
 * on one GPU (two host threads, two contexts) race. */
int askja_hmc_trajectory_noalloc(askja_ctx *ctx, const askja_hmc_params *p, uint64_t traj_counter,
                                 askja_hmc_result *res);

/* The Metropolis uniform in (0, 1) that askja_hmc_trajectory draws for
 * trajectory `counter` (replaces gsl_rng_uniform of an mt19937 seeded with
 * time(NULL), sun.cpp:1277-1287). */
int askja_uniform(askja_ctx *ctx, uint64_t seed, uint64_t counter, double *out);

/* ---- autocorrelation (no reference counterpart) ----------------------------------- */
/* Integrated autocorrelation time of the series x[0..n-1] by automatic
 * windowing (U. Wolff, Comput. Phys. Commun. 156 (2004) 143; S is the window
 * factor, 1.5 customary): *tau = 1 + 2 sum_{t=1}^{W} rho(t), 1 for
 * uncorrelated data, with rho(t) the autocorrelation normalised by the n - t
 * pairs, and *window = W (may be NULL).  The reference's error bars assume
 * independent samples (measurements4d.cpp:1213-1257).  Host code, no
 * context: returns ASKJA_EINVAL for n < 8, S <= 0 or null pointers.  A
 * constant series gives tau = 1, window 0. */
int askja_tau_int(const double *x, long n, double S, double *tau, long *window);

/* ---- learned field transformation (no reference counterpart; SU(3), D = 3 or 4) ----
 * U = F(V): nsub masked sub-layers; sub-layer k updates the links of
 * direction mu[k] on the sites of parity parity[k] as
 *   U'_l = exp(i theta Lhat) U_l,  theta = rho Im Tr(Lhat U_l S_l),
 * with S_l the staple sum, Lhat the normalised traceless anti-Hermitian part of
 * the plaquettes at x in the planes without mu (both frozen in the sub-layer),
 * rho = rho_max tanh(MLP(f)) with features f = ReTr / 3 of those plaquettes,
 * then of the 2x1 and 1x2 rectangles at x in the same planes, plane by plane;
 * the MLP (hidden 32, FP32) of sub-layer k is weight block block[k].
 * log det J = sum over updated links of log(1 + rho ReTr(Lhat^2 U_l S_l)).
 * Requires even extents and 0 < rho_max < 1/(D-1) (invertibility).  Block
 * layout: W1[32][nf], b1[32], w2[32], b2 with nf = 3 (D-1)(D-2)/2. */
int askja_flow_define(askja_ctx *ctx, int nsub, const int *mu, const int *parity, const int *block, int nblocks,
                      double rho_max, double delta);
/* The stout form of the map (no reference counterpart; SU(3), D = 3 or 4):
 * the same masks, features, MLP and weight blocks, with sub-layer k updating
 * its links as
 *   U'_l = exp(i Q) U_l,  Q = -rho TAH(U_l S_l),  rho = rho_max tanh(MLP(f)),
 * (S_l the frozen staple sum: a smearing step along the force direction) and
 * log det J = sum over updated links of the log det of the 8 x 8 Jacobian
 * 1 + g(ad_{iQ}) M in su(3) coordinates (g(z) = (1 - e^{-z}) / z, M the
 * change of Q with the link), computed exactly per link.  Requires even
 * extents and 0 < rho_max < 1/(2(D-1)) (then |g M| < 1: invertible). */
int askja_flow_define_stout(askja_ctx *ctx, int nsub, const int *mu, const int *parity, const int *block, int nblocks,
                            double rho_max);
/* The spectral form of the map for any D >= 2 and any 2 <= N <= 64 (no
 * reference counterpart; the map of the large-N two-dimensional workload):
 * sub-layer k updates U_mu(x), mu = mu[k], for every x with
 * x_nu = offset[k] mod 4 (nu = (mu + 1) mod D, so nu = 1 - mu in D = 2),
 * through the plaquette W = U_mu(x) S (S = the frozen staple
 * U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+): every eigenvalue of W goes to
 * e^{2i phi/N} (lambda - t) / (1 - t lambda) (phi restores det = 1), i.e.
 *   U'_mu(x) = e^{2i phi/N} (1 - tW)^{-1} (W - t) S^+,
 *   t = t_max tanh(MLP(f)),  f = ReTr P_mu,nu(x+nu+k mu) / N, k = -1, 0, 1
 * (frozen plaquettes of the same plane), and per updated link
 *   log J = N^2 log(1-t^2) - 2N log|det(1-tW)| + log((1+t^2-(2t/N) ReTr W)/(1-t^2)).
 * Requires every extent a multiple of 4, an FP64 precision mode and
 * 0 < t_max <= 1/2.  Block layout: W1[32][3], b1[32], w2[32], b2 (161).
 * The gradient's contributions to frozen links are FP64 atomic adds (the
 * order of those sums is not fixed). */
int askja_flow_define_spectral(askja_ctx *ctx, int nsub, const int *mu, const int *offset, const int *block,
                               int nblocks, double t_max);
/* Floats per weight block. */
int askja_flow_block_size(askja_ctx *ctx, int *bsize);
/* All nblocks * bsize weights, from host memory or (on_device != 0) device memory. */
int askja_flow_set_weights(askja_ctx *ctx, const float *w, int on_device);
/* U = F(V) of the current links V (kept in the context), log det J. */
int askja_flow_apply(askja_ctx *ctx, double *logj);
/* host AoS buffer <- U = F(V) of the last apply / gradient call */
int askja_flow_download(askja_ctx *ctx, double *aos);
/* Phi = beta S_W(F(V)) - log J(V): its Lie gradient with respect to V (kept
 * for askja_flow_gradient_download) by the hand-written backward kernels, and
 * with wgrad != 0 the gradient with respect to the weights.  The forward pass
 * keeps only each sub-layer's input links (its masked links, 1/(2D) of a
 * field per sub-layer); the backward recomputes every sub-layer's
 * intermediates from that input (the stout kind: from W, rho and the
 * features its forward kept per updated link).  Deterministic (no atomics on the link
 * gradient).  action = S_W(F(V)) (EnergyMag normalisation), logj = log det J(V). */
int askja_flow_gradient(askja_ctx *ctx, double beta, int wgrad, double *action, double *logj);
/* host AoS buffer <- Hermitian traceless g_l = sum_a T^a d/dt Phi(e^{i t T^a} V_l) */
int askja_flow_gradient_download(askja_ctx *ctx, double *aos);
/* host buffer <- dPhi/dweights (nblocks * bsize doubles) of the last gradient call with wgrad */
int askja_flow_weight_gradient(askja_ctx *ctx, double *out);
/* One field-transformation HMC trajectory in V: H = beta S_W(F(V)) - log J(V)
 * + beta/(2N) sum Tr P^2, kick-drift-kick with the flow force, exact
 * Metropolis; on reject V returns to its value before the trajectory (kept
 * in a buffer of the flow's own: the askja_snapshot slot is not touched). */
int askja_flow_hmc_trajectory(askja_ctx *ctx, const askja_hmc_params *p, uint64_t traj_counter,
                              askja_hmc_result *res);
/* The same trajectory with two time scales (Sexton and Weingarten's nested
 * leapfrog): Phi = alpha beta S_W(V) + [Phi(V) - alpha beta S_W(V)].  The
 * scaled plain action moves on the fine step dt with the fused kick+drift of
 * askja_hmc_trajectory, `inner` steps per coarse step; the flow's correction
 * kicks on the coarse step inner * dt, so a trajectory of ndrift fine steps
 * (inner must divide ndrift) evaluates the flow gradient ndrift / inner + 1
 * times instead of ndrift + 1.  alpha > 0 is free (the split is exact for
 * any): the closer alpha beta S_W's force is to Phi's, the smaller the
 * correction and the longer the coarse step it allows.  The splitting is
 * symplectic and reversible and H is the same, so the accept test keeps the
 * chain exact; inner = 1, alpha = 1 is askja_flow_hmc_trajectory up to
 * rounding.  No counterpart in the reference (no learned proposal there). */
int askja_flow_hmc_trajectory_nested(askja_ctx *ctx, const askja_hmc_params *p, int inner, double alpha,
                                     uint64_t traj_counter, askja_hmc_result *res);
/* askja_wilson_lines of U = F(V). */
int askja_flow_wilson_lines(askja_ctx *ctx, int dir, double *out);

/* ---- plumbing ------------------------------------------------------------------------ */
/* Synchronise the context's stream. */
int askja_sync(askja_ctx *ctx);
/* Record a start event / a stop event on the context's stream; the stop call
 * synchronises and returns the elapsed milliseconds between the two. */
int askja_timer_start(askja_ctx *ctx);
int askja_timer_stop(askja_ctx *ctx, double *ms);
/* Device pointers of the current links and momenta (private layout). */
int askja_device_fields(askja_ctx *ctx, void **links, void **mom, void **mom_prev);

#ifdef __cplusplus
}
#endif

#endif
