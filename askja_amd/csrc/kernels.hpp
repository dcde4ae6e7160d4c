// kernels.hpp -- host-side view of the device layout and the launchers that
// each kernel file exports.  Plain C++ types only; every launcher returns a
// hipError_t and launches on `stream`.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace askja {

// Arithmetic flavours of the MD kernels.
enum Mode : int {
    kModeF64 = 0,       // FP64, FMA, factored staples
    kModeF64Exact = 1,  // FP64 in the reference's operation order, no FMA
    kModeF32 = 2,       // FP32 storage and arithmetic, FP64 sums
    kModeBF16x3 = 3,    // FP32 storage; N >= 9 link products as split BF16 on MFMA
};

// FP32 storage (every mode but the FP64 ones)
__host__ __device__ inline bool f32_mode(int mode) { return mode == kModeF32 || mode == kModeBF16x3; }

// Geometry shared by every kernel.  Device site index s: all sites of even
// parity (sum of coordinates mod 2), then all odd ones; within a parity the
// sites are ordered by slab of the longest direction (cut into at most 8
// slabs), then lexicographically.  Link l = mu * V + s.  nbr[(2*d + 0) * V + s] = s + e_d, nbr[(2*d + 1) * V +
// s] = s - e_d (periodic), dev2lex[s] = lexicographic index (x[D-1] fastest).
struct Geom {
    long V;
    int D;
    int N;
    int tile;       // 64 for N <= 3 (links tiled by wavefront), 1 otherwise
    const int *nbr;
    const int *dev2lex;
    const int *lex2dev;
    long V0;        // number of even-parity sites (device sites 0 .. V0-1)
    int ext[32];
    long helparr[32];  // lexicographic strides, structs.h:130-136
};

// Reals per link (2 N^2) and per momentum (N^2, compact Hermitian).
__host__ __device__ inline int link_reals(int n) { return 2 * n * n; }
__host__ __device__ inline int mom_reals(int n) { return n * n; }

// Offset of real number e of link l in a field tiled by `tile` links.
__host__ __device__ inline size_t foff(long l, int e, int per, int tile)
{
    return ((size_t)(l / tile) * (size_t)per + (size_t)e) * (size_t)tile + (size_t)(l % tile);
}

// Index of the (i < j) pair in the compact Hermitian layout: N diagonal
// reals, then the strict upper triangle row-major as (re, im).
__host__ __device__ inline int hup(int n, int i, int j)
{
    return n + 2 * (i * n - i * (i + 1) / 2 + (j - i - 1));
}

struct Cinv9 {
    double c[9];
};

struct KickArgs {
    Geom g;
    const void *U;     // links read (staples)
    void *Uout;        // links written by the fused drift (may equal nothing)
    void *P;           // momenta, updated in place
    void *Pprev;       // receives P before the kick (kick-only variant), may be null
    double eps_kick;
    double eps_drift;
    int level;
    double cinv[9];    // 1./k, k = 0..8, computed on the host
    double inv_n;      // 1./N computed on the host
    double *part;      // per-block partial sums: {ReTr W, Tr Pin^2, Tr Pout^2} x (s, c);
                       // null: not wanted (the kick skips its block reductions)
};

// Number of partial-sum blocks a kick launch writes (each 3 (s, c) pairs).
long kick_partials(const Geom &g);
// Site blocks of 256 of the register-path kick (even and odd sites blocked
// separately).
inline long kick_site_blocks(const Geom &g)
{
    return (g.V0 + 255) / 256 + (g.V - g.V0 + 255) / 256;
}

hipError_t launch_kick(const KickArgs &a, int mode, bool drift, hipStream_t st);
// N >= 9, FP32 links: sum ReTr W in FP64 arithmetic into the kick partials
// (the exact action of the stored links; no momentum is touched)
hipError_t launch_action_f64(const KickArgs &a, int mode, hipStream_t st);
// Sums of the action pass's partials: {every link, the links of direction 0}
// (2 (s, c) pairs into out).
hipError_t launch_action_split(const double *part, long np, int D, double *out, hipStream_t st);
hipError_t launch_drift(const Geom &g, void *U, const void *P, double eps, int level,
                        const double *cinv, int mode, hipStream_t st);

// ---- reductions -----------------------------------------------------------
// Per-block partials (K (s, c) pairs each) -> out[2K] on the device.
hipError_t launch_finalize(const double *part, long nparts, int K, double *out, hipStream_t st);

// Plaquette sums: {sum over (0,k) planes, sum over (j,k), 1 <= j < k} of
// (1 - ReTr P / N).  Returns the number of partial blocks written, at most
// plaq_partials(g) (2 (s, c) pairs each).
long plaq_partials(const Geom &g);
hipError_t launch_plaq(const Geom &g, const void *U, int mode, double inv_n, double *part,
                       long *nparts, hipStream_t st);
// Rows of the Gauss-law partials for N >= 4 (2N^2 (s, c) pairs each).
long gauss_rows(const Geom &g);
// sum over links of ||U U^+ - 1||_F^2 (tr2 of U U^+ - 1)
hipError_t launch_uviol(const Geom &g, const void *U, int mode, double *part, long *nparts, hipStream_t st);
// sum over links of the N x N matrix P - U^+ P U (2 N^2 reals), written to out
// (device, 2 N^2 (s, c) pairs).
hipError_t launch_gauss(const Geom &g, const void *U, const void *P, int mode, double *part,
                        long *nparts, double *out, hipStream_t st);
// {sum Tr P^2, sum 1/4 Tr (P + Q)^2} over links, Q may be null
hipError_t launch_kinetic(const Geom &g, const void *P, const void *Q, int mode, double *part,
                          long *nparts, hipStream_t st);
// Straight Wilson lines in direction dir starting at x_dir = 0: out[2q..2q+1]
// = Tr(prod U)/N of line q (lines in lexicographic order of their start).
hipError_t launch_wlines(const Geom &g, const void *U, int dir, int mode, double *out, hipStream_t st);

// Product matrix (2 N^2 doubles, device) of the straight line in direction
// dir starting at lexicographic site lex; scratch >= 4 N^2 doubles (used
// for N <= 3).
hipError_t launch_line_matrix(const Geom &g, const void *U, int dir, long lex, int mode, double *scratch, double *out,
                              hipStream_t st);

// ---- field set-up -----------------------------------------------------------
// Host AoS (reference U-*.conf layout, doubles) <-> device layout.
hipError_t launch_links_from_aos(const Geom &g, const double *aos, void *U, int mode, hipStream_t st);
hipError_t launch_links_to_aos(const Geom &g, const void *U, double *aos, int mode, hipStream_t st);
hipError_t launch_mom_from_aos(const Geom &g, const double *aos, void *P, int mode, hipStream_t st);
hipError_t launch_mom_to_aos(const Geom &g, const void *P, double *aos, int mode, hipStream_t st);
hipError_t launch_links_cold(const Geom &g, void *U, int mode, hipStream_t st);
hipError_t launch_mom_refresh(const Geom &g, void *P, uint64_t seed, uint64_t counter, double c,
                              int mode, hipStream_t st);
// Doubles of FP64 working memory each generic (N >= 4) diagnostic or
// projection thread needs.
hipError_t launch_mom_scale(const Geom &g, void *P, double f, int mode, hipStream_t st);
// SU(N) projection (Gram-Schmidt on rows + determinant phase) in FP64.
// haar != 0 first fills every link with a Ginibre matrix from Philox (seed,
// stream).
hipError_t launch_project(const Geom &g, void *U, int mode, int haar, uint64_t seed, uint64_t stream,
                          hipStream_t st);

// ---- measurement (measure.hip) ---------------------------------------------
// blocks per time slice of the slice-trace partials
long slice_trace_blocks(const Geom &g);
// part[x0 * nb + b] (s, c): sums of ReTr U_dir / N over the sites of slice x0
hipError_t launch_slice_trace(const Geom &g, const void *U, int dir, int mode, double *part, hipStream_t st);
// sorted, winding-recentred eigen-phases of the lines in direction dir that
// start at the lexicographic sites starts[0 .. nlines); wbuf: 2 N^2 per line
hipError_t launch_line_phases(const Geom &g, const void *U, int dir, int mode, int nlines, const long *starts,
                              double *wbuf, double *out, hipStream_t st);

// ---- learned field transformation (flow.hip, N = 3) -------------------------
enum { kFlowRankOne = 0, kFlowStout = 1, kFlowSpectral = 2 };
struct FlowSub {
    int mu, par;     // updated links: direction mu on sites of parity par
    const float *w;  // device MLP block of the sub-layer
    int nf;          // features = planes not containing mu
    double rho_max, delta;
    int kind;        // kFlowRankOne, kFlowStout or kFlowSpectral (then par = the row offset
                     // off: links U_mu(x) with x_nu = off mod 4, rho_max = t_max)
};
long flow_forward_partials(const Geom &g, int par);
// one sub-layer in place; log (may be null) receives its input links; for
// the stout kind, st_wr (stout_wr_doubles) receives W = U S, rho and the MLP
// features of every updated link: launch_stout_logdet turns W and rho into
// log det J, and the backward reads all three instead of recomputing them
hipError_t launch_flow_forward(const Geom &g, const void *Uin, void *Uout, const FlowSub &f, int mode, double *part,
                               void *log, double *st_wr, hipStream_t st);
// backward of one sub-layer; restores its input links from log into U
// (contrib: flow_contrib_doubles scratch for the neighbour contributions;
// lg: for the stout kind, the log-det gradient of launch_stout_logdet_grad;
// st_wr: for the stout kind, what the sub-layer's forward wrote there)
hipError_t launch_flow_backward(const Geom &g, void *U, const FlowSub &f, int mode, double *gr, double c_logj,
                                double *gw, const void *log, double *contrib, const double *lg, const double *st_wr,
                                hipStream_t st);
// stout kind: per updated link W (18 doubles), rho and the features,
// structure of arrays
size_t stout_wr_doubles(const Geom &g);
size_t stout_aux_doubles(const Geom &g);  // J^{-1}, eigenbasis (85 doubles per link)
size_t stout_lg_doubles(const Geom &g);   // the log-det gradient (19 per link)
// log det J of every updated link of a stout sub-layer into per-block
// partials (flow_forward_partials of its parity); with aux, also J^{-1} and
// the eigenbasis of Q for launch_stout_logdet_grad
hipError_t launch_stout_logdet(const Geom &g, const FlowSub &f, const double *st_wr, double *part, double *aux,
                               hipStream_t st);
// c_logj d log det J / dW (as G_W: d = 2 ReTr(G_W dW)) and / d rho per link
hipError_t launch_stout_logdet_grad(const Geom &g, const FlowSub &f, const double *st_wr, const double *aux,
                                    double c_logj, double *lg, hipStream_t st);
size_t flow_contrib_doubles(const Geom &g);
// link reals of one sub-layer's input log (sites of one parity, 64-link tiles)
size_t flow_log_reals(const Geom &g);
// gr = coef * TAH(U sum staples) (9 doubles per link), part: ReTr W per block
hipError_t launch_wilson_grad(const Geom &g, const void *U, double coef, int mode, double *gr, double *part,
                              long *np, hipStream_t st);
// P <- P - eps_c * gr
hipError_t launch_kick_g(const Geom &g, void *P, const double *gr, double eps_c, int mode, hipStream_t st);

// ---- spectral field transformation (flow_spectral.hip, any N and D) ----------
// updated links per sub-layer: V / 4 (rows x_nu = off mod 4)
long sflow_active(const Geom &g);
// one sub-layer (mu, off) in place on U (FP64): feature plaquettes into ptr
// (V doubles), the updated links' inputs into log and (1 - tW)^{-1} into inv
// (2 N^2 per updated link each, both may be null), log J per updated link
// into part ((s, c) pairs)
hipError_t launch_sflow_forward(const Geom &g, void *U, int mu, int off, const float *w, double tmax, double *ptr,
                                double *log, double *inv, double *part, hipStream_t st);
// backward of one sub-layer: gr (momentum layout) from the gradient at its
// output to the gradient at its input, c_logj d log J included; gw (may be
// null) += the MLP weight gradient; restores the input links from log into
// U; inv: what the forward kept.  ptr, dF: V doubles of scratch.
hipError_t launch_sflow_backward(const Geom &g, void *U, int mu, int off, const float *w, double tmax, double *ptr,
                                 double *log, double *inv, double *gr, double c_logj, double *gw, double *dF,
                                 hipStream_t st);
// y += s x (n doubles)
hipError_t launch_axpy(double *y, const double *x, double s, long n, hipStream_t st);

}  // namespace askja
