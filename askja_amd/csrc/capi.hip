// capi.hip -- the C ABI of include/askja.h: a context owning the
// device-resident lattice (geometry tables, three link buffers, momenta,
// reduction buffers) and the host side of every entry point.
#include "../../include/askja.h"
#include "kernels.hpp"
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <mutex>
#include <string>
#include <vector>

using namespace askja;

// Allocation-free partial sums (askja_hmc_trajectory_noalloc): a file-scope
// device array of a fixed size, one copy per GPU, held until the process
// ends.  Its device address is looked up once per device, with that device
// current, by askja_create; host code uses only that address (the array's
// name is a host-side stand-in).  Every call on a GPU shares that GPU's
// copy: calls that overlap on one GPU race.
namespace {
constexpr size_t kScratchResDoubles = 64;                 // finalised sums
constexpr size_t kScratchPartDoubles = (size_t)1 << 21;   // block partials (16 MiB)
__device__ __align__(256) double scratch[kScratchResDoubles + kScratchPartDoubles];
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;
}  // namespace

struct askja_ctx {
    int nc = 0, dim = 0, prec = 0, mode = 0;
    int dev = 0;  // the HIP device current at askja_create
    int ext[32] = {0};
    long V = 0, NL = 0;
    long helparr[32] = {0};
    int tile = 1;
    size_t link_reals = 0, mom_reals = 0;  // per field, padded to the tile
    std::vector<int> lex2dev, dev2lex;

    int *d_nbr = nullptr, *d_dev2lex = nullptr, *d_lex2dev = nullptr;
    void *links[3] = {nullptr, nullptr, nullptr};
    int cur = 0;
    void *snap = nullptr;
    void *mom = nullptr, *mom_prev = nullptr;
    double *part = nullptr;
    size_t part_doubles = 0;
    bool part_static = false;  // part/res point into the static buffer
    double *res = nullptr;      // device, 64 doubles
    double *matwork = nullptr;  // FP64 working memory: two N x N matrices
    size_t matwork_doubles = 0;
    double *staging = nullptr;  // host<->device AoS staging
    size_t staging_doubles = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // asynchronous configuration download: own staging buffer and copy stream
    double *astaging = nullptr;
    hipStream_t cst = nullptr;
    hipEvent_t dl_ready = nullptr, dl_done = nullptr;
    bool dl_pending = false;
    int sticky = 0;
    std::string err;
    Geom geom{};
    // learned field transformation
    int flow_nsub = 0, flow_nblocks = 0, flow_bsize = 0;
    std::vector<FlowSub> flow_subs;
    std::vector<int> flow_block;
    float *flow_w = nullptr;     // nblocks * bsize MLP weights
    double *flow_gw = nullptr;   // their gradient
    double *flow_g = nullptr;    // 9 doubles per link: Lie gradient
    double *flow_lj = nullptr;   // log J (s, c) of the whole map
    void *flow_out = nullptr;    // U = F(V)
    void *flow_work = nullptr;   // backward working field
    void *flow_log = nullptr;    // input links of every sub-layer
    double *flow_contrib = nullptr;  // backward neighbour contributions
    void *flow_save = nullptr;       // chain state at the start of a flow trajectory
    int flow_kind = kFlowRankOne;
    double *flow_st_wr = nullptr;    // stout: W, rho and the features of every sub-layer's updated links
    double *flow_st_aux = nullptr;   // stout: inverse Jacobian and eigenbasis of every sub-layer
    double *flow_st_lg = nullptr;    // stout: log-det gradient of every sub-layer
    // stout: the log-det gradients run on a second stream beside the rest of
    // the flow gradient; per sub-layer an event after its forward (main
    // stream) and one after its log-det gradient (second stream)
    hipStream_t fst = nullptr;
    std::vector<hipEvent_t> flow_ev;
    double *flow_sp_ptr = nullptr;   // spectral: feature plaquette traces (V doubles)
    double *flow_sp_dF = nullptr;    // spectral: their gradients (V doubles)
    double *flow_sp_inv = nullptr;   // spectral: (1 - tW)^{-1} of every sub-layer's updated links
    double *flow_gp = nullptr;       // the plain force's gradient at V (nested trajectory), flow_g's layout
};

namespace {

int fail(askja_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) {
        c->err = buf;
        c->sticky = code;
    }
    return code;
}

#define HIPCHK(c, expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(c, ASKJA_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define ENTER(c)                                                               \
    do {                                                                       \
        if (!(c)) return ASKJA_EINVAL;                                         \
        if ((c)->sticky) return ASKJA_ESTICKY;                                 \
    } while (0)

size_t elem_bytes(const askja_ctx *c) { return f32_mode(c->mode) ? 4 : 8; }

int ensure_part(askja_ctx *c, size_t nd)
{
    if (nd <= c->part_doubles) return ASKJA_OK;
    if (c->part_static)  // never grow (or free) the static buffer
        return fail(c, ASKJA_EINVAL, "partial sums of %zu doubles exceed the static buffer", nd);
    if (c->part) (void)hipFree(c->part);
    c->part = nullptr;
    c->part_doubles = 0;
    HIPCHK(c, hipMalloc(&c->part, nd * sizeof(double)));
    c->part_doubles = nd;
    return ASKJA_OK;
}

int ensure_matwork(askja_ctx *c)
{
    if (c->matwork) return ASKJA_OK;
    // two FP64 N x N work matrices (the line matrix of N <= 3)
    const size_t want = (size_t)4 * c->nc * c->nc;
    HIPCHK(c, hipMalloc(&c->matwork, want * sizeof(double)));
    c->matwork_doubles = want;
    return ASKJA_OK;
}

int ensure_staging(askja_ctx *c)
{
    if (c->staging) return ASKJA_OK;
    const size_t nd = (size_t)c->V * c->dim * c->nc * c->nc * 2;
    HIPCHK(c, hipMalloc(&c->staging, nd * sizeof(double)));
    c->staging_doubles = nd;
    return ASKJA_OK;
}

// sum of the (s, c) pair at res[2k], res[2k+1] after a D2H copy
double pair_val(const double *h, int k) { return h[2 * k] + h[2 * k + 1]; }

int read_res(askja_ctx *c, double *host, int n)
{
    HIPCHK(c, hipMemcpyAsync(host, c->res, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return ASKJA_OK;
}

void fill_cinv(double *ci)
{
    ci[0] = 1.0;
    for (int k = 1; k < 9; k++) ci[k] = 1. / (double)k;
}

KickArgs kick_args(askja_ctx *c, const void *U, void *Uout, void *Pprev, double ek, double ed, int level)
{
    KickArgs a{};
    a.g = c->geom;
    a.U = U;
    a.Uout = Uout;
    a.P = c->mom;
    a.Pprev = Pprev;
    a.eps_kick = ek;
    a.eps_drift = ed;
    a.level = level;
    fill_cinv(a.cinv);
    a.inv_n = 1. / (double)c->nc;
    a.part = c->part;
    return a;
}

// kick (optionally fused with a drift into `out`) + optional finalize of its
// three sums into res[off .. off+5]
int do_kick(askja_ctx *c, const void *U, void *Uout, void *Pprev, double ek, double ed, int level, bool drift,
            int res_off)
{
    const long np = kick_partials(c->geom);
    int rc = ensure_part(c, (size_t)np * 6);
    if (rc) return rc;
    KickArgs a = kick_args(c, U, Uout, Pprev, ek, ed, level);
    // the sums of a kick whose energies nobody reads (every one of a
    // trajectory but the first and the last) are not computed
    if (res_off < 0) a.part = nullptr;
    HIPCHK(c, launch_kick(a, c->mode, drift, c->st));
    if (res_off >= 0) HIPCHK(c, launch_finalize(c->part, np, 3, c->res + res_off, c->st));
    return ASKJA_OK;
}

// Metropolis uniform of trajectory `counter`: Philox block 0 of link 0,
// purpose 3, first word mapped to (0, 1).
double metropolis_uniform(uint64_t seed, uint64_t counter)
{
    const U4 r = philox_block(seed, counter, kPurposeAccept, 0u, 0u);
    return u01(r.x);
}

long n_plaquettes(const askja_ctx *c) { return c->V * (long)c->dim * (c->dim - 1) / 2; }

// S = sum_p (1 - ReTr P / N) of the links in U in FP64 arithmetic.  The
// FP32 and BF16x3 modes take the Metropolis energies from this rather than
// from the MD's reduced-precision products, so that H is the exact FP64
// energy of the stored (FP32) links.  Overwrites res.
int exact_action(askja_ctx *c, const void *U, double *s)
{
    int rc;
    if (c->nc > 8) {
        const long np = kick_partials(c->geom);
        if ((rc = ensure_part(c, (size_t)np * 6))) return rc;
        const KickArgs a = kick_args(c, U, nullptr, nullptr, 0.0, 0.0, 0);
        HIPCHK(c, launch_action_f64(a, c->mode, c->st));
        HIPCHK(c, launch_finalize(c->part, np, 3, c->res, c->st));
        double h[6];
        if ((rc = read_res(c, h, 6))) return rc;
        *s = (double)n_plaquettes(c) - pair_val(h, 0) / (4.0 * c->nc);
        return ASKJA_OK;
    }
    long np = 0;
    if ((rc = ensure_part(c, (size_t)plaq_partials(c->geom) * 4))) return rc;
    HIPCHK(c, launch_plaq(c->geom, U, c->mode, 1. / (double)c->nc, c->part, &np, c->st));
    HIPCHK(c, launch_finalize(c->part, np, 2, c->res, c->st));
    double h[4];
    if ((rc = read_res(c, h, 4))) return rc;
    *s = pair_val(h, 0) + pair_val(h, 1);
    return ASKJA_OK;
}

// MD of one trajectory with the chain-state buffer `in0` left untouched;
// returns the buffer index that holds the final links.  Energies of the
// first and last kick go to res[0..5] and res[6..11].
// The plain leapfrog of ndrift steps (kicks scaled by kscale: the force of
// kscale beta S_W)
int md_pingpong(askja_ctx *c, double dt, int ndrift, int level, int *final_buf, double kscale = 1.0)
{
    const int a = c->cur, b = (a + 1) % 3, d = (a + 2) % 3;
    const double kd = kscale * dt;
    int rc;
    if (ndrift <= 0) {
        // two half kicks on the same links
        if ((rc = do_kick(c, c->links[a], nullptr, nullptr, kd / 2, 0.0, level, false, 0))) return rc;
        if ((rc = do_kick(c, c->links[a], nullptr, nullptr, kd / 2, 0.0, level, false, 6))) return rc;
        *final_buf = a;
        return ASKJA_OK;
    }
    if ((rc = do_kick(c, c->links[a], c->links[b], nullptr, kd / 2, dt, level, true, 0))) return rc;
    int in = b, out = d;
    for (int t = 2; t <= ndrift; t++) {
        if ((rc = do_kick(c, c->links[in], c->links[out], nullptr, kd, dt, level, true, -1))) return rc;
        const int tmp = in;
        in = out;
        out = tmp;
    }
    if ((rc = do_kick(c, c->links[in], nullptr, nullptr, kd / 2, 0.0, level, false, 6))) return rc;
    *final_buf = in;
    return ASKJA_OK;
}

}  // namespace

extern "C" {

int askja_create(int nc, int dim, const int *extents, int precision, askja_ctx **out)
{
    if (!out || !extents) return ASKJA_EINVAL;
    *out = nullptr;
    if (nc < 1 || nc > 64 || dim < 1 || dim > 32) return ASKJA_EUNSUP;
    if (precision < ASKJA_FP64 || precision > ASKJA_BF16X3) return ASKJA_EUNSUP;
    long V = 1;
    for (int i = 0; i < dim; i++) {
        if (extents[i] < 1) return ASKJA_EINVAL;
        V *= extents[i];
        if (V * dim > (1L << 30)) return ASKJA_EUNSUP;
    }
    askja_ctx *c = new askja_ctx();
    c->nc = nc;
    c->dim = dim;
    c->prec = precision;
    c->mode = precision == ASKJA_FP64         ? kModeF64
              : precision == ASKJA_FP64_EXACT ? kModeF64Exact
              : precision == ASKJA_FP32       ? kModeF32
                                              : kModeBF16x3;
    c->V = V;
    c->NL = V * dim;
    for (int i = 0; i < dim; i++) c->ext[i] = extents[i];
    for (int i = 0; i < dim; i++) {
        c->helparr[i] = 1;
        for (int j = i + 1; j < dim; j++) c->helparr[i] *= extents[j];
    }
    c->tile = nc <= 3 ? 64 : 1;
    const size_t ntiles = (size_t)((c->NL + c->tile - 1) / c->tile);
    c->link_reals = ntiles * c->tile * 2 * nc * nc;
    c->mom_reals = ntiles * c->tile * nc * nc;

    // Device site order: parity-0 sites then parity-1; within a parity the
    // sites are sorted by (chunk of the longest direction, lexicographic
    // index).  The longest direction is cut into min(8, extent) chunks, so a
    // contiguous eighth of the sites (what one XCD processes) is a compact
    // slab of the lattice rather than a thin slice of the slowest direction.
    c->lex2dev.resize(V);
    c->dev2lex.resize(V);
    int split = 0;
    for (int d = 1; d < dim; d++)
        if (extents[d] > extents[split]) split = d;
    const int nchunk = extents[split] < 8 ? extents[split] : 8;
    const int chunk = (extents[split] + nchunk - 1) / nchunk;
    std::vector<long> key(V);
    std::vector<int> par(V);
    long n0 = 0;
    for (long i = 0; i < V; i++) {
        long r = i;
        int p = 0, xs = 0;
        for (int d = 0; d < dim; d++) {
            const int xd = (int)(r / c->helparr[d]);
            r %= c->helparr[d];
            p += xd;
            if (d == split) xs = xd;
        }
        par[i] = p & 1;
        key[i] = (long)(xs / chunk) * V + i;
        if (par[i] == 0) n0++;
    }
    {
        std::vector<long> order(V);
        for (long i = 0; i < V; i++) order[i] = i;
        std::sort(order.begin(), order.end(), [&](long a, long b) {
            if (par[a] != par[b]) return par[a] < par[b];
            return key[a] < key[b];
        });
        for (long s = 0; s < V; s++) {
            c->dev2lex[s] = (int)order[s];
            c->lex2dev[order[s]] = (int)s;
        }
    }
    // neighbour table in device order (periodic +-1 steps, structs.h:169-218)
    std::vector<int> nbr((size_t)2 * dim * V);
    for (long s = 0; s < V; s++) {
        const long i = c->dev2lex[s];
        for (int d = 0; d < dim; d++) {
            const long xd = (i / c->helparr[d]) % extents[d];
            const long up = (xd + 1) % extents[d], dn = (xd + extents[d] - 1) % extents[d];
            nbr[(size_t)(2 * d) * V + s] = c->lex2dev[i + (up - xd) * c->helparr[d]];
            nbr[(size_t)(2 * d + 1) * V + s] = c->lex2dev[i + (dn - xd) * c->helparr[d]];
        }
    }
#define CREATE_CHK(expr)                                                        \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            fail(c, ASKJA_EHIP, "%s: %s", #expr, hipGetErrorString(e_));        \
            askja_destroy(c);                                                   \
            return e_ == hipErrorOutOfMemory ? ASKJA_ENOMEM : ASKJA_EHIP;       \
        }                                                                       \
    } while (0)
    CREATE_CHK(hipGetDevice(&c->dev));
    if (c->dev >= 0 && c->dev < kMaxDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch[c->dev]) {
            void *p = nullptr;
            CREATE_CHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scratch[c->dev] = (double *)p;
        }
    }
    CREATE_CHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    CREATE_CHK(hipEventCreate(&c->ev0));
    CREATE_CHK(hipEventCreate(&c->ev1));
    CREATE_CHK(hipMalloc(&c->d_nbr, nbr.size() * sizeof(int)));
    CREATE_CHK(hipMalloc(&c->d_dev2lex, V * sizeof(int)));
    CREATE_CHK(hipMalloc(&c->d_lex2dev, V * sizeof(int)));
    CREATE_CHK(hipMemcpy(c->d_nbr, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice));
    CREATE_CHK(hipMemcpy(c->d_dev2lex, c->dev2lex.data(), V * sizeof(int), hipMemcpyHostToDevice));
    CREATE_CHK(hipMemcpy(c->d_lex2dev, c->lex2dev.data(), V * sizeof(int), hipMemcpyHostToDevice));
    const size_t eb = elem_bytes(c);
    for (int b = 0; b < 3; b++) {
        CREATE_CHK(hipMalloc(&c->links[b], c->link_reals * eb));
        CREATE_CHK(hipMemset(c->links[b], 0, c->link_reals * eb));
    }
    CREATE_CHK(hipMalloc(&c->mom, c->mom_reals * eb));
    CREATE_CHK(hipMalloc(&c->mom_prev, c->mom_reals * eb));
    CREATE_CHK(hipMemset(c->mom, 0, c->mom_reals * eb));
    CREATE_CHK(hipMemset(c->mom_prev, 0, c->mom_reals * eb));
    CREATE_CHK(hipMalloc(&c->res, 64 * sizeof(double)));
#undef CREATE_CHK
    Geom &g = c->geom;
    g.V = V;
    g.D = dim;
    g.N = nc;
    g.tile = c->tile;
    g.nbr = c->d_nbr;
    g.dev2lex = c->d_dev2lex;
    g.lex2dev = c->d_lex2dev;
    g.V0 = n0;
    for (int i = 0; i < dim; i++) {
        g.ext[i] = extents[i];
        g.helparr[i] = c->helparr[i];
    }
    if (askja_links_cold(c) != ASKJA_OK) {
        askja_destroy(c);
        return ASKJA_EHIP;
    }
    *out = c;
    return ASKJA_OK;
}

int askja_destroy(askja_ctx *c)
{
    if (!c) return ASKJA_EINVAL;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (int b = 0; b < 3; b++)
        if (c->links[b]) (void)hipFree(c->links[b]);
    void *bufs[] = {c->snap, c->mom, c->mom_prev, c->part, c->res, c->matwork, c->staging,
                    c->d_nbr, c->d_dev2lex, c->d_lex2dev, c->flow_w, c->flow_gw, c->flow_g, c->flow_lj,
                    c->flow_out, c->flow_work, c->flow_log, c->flow_contrib, c->flow_save,
                    c->flow_st_wr, c->flow_st_aux, c->flow_st_lg, c->flow_sp_ptr, c->flow_sp_dF, c->flow_sp_inv,
                    c->flow_gp};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (c->cst) (void)hipStreamSynchronize(c->cst);
    if (c->astaging) (void)hipFree(c->astaging);
    if (c->dl_ready) (void)hipEventDestroy(c->dl_ready);
    if (c->dl_done) (void)hipEventDestroy(c->dl_done);
    if (c->cst) (void)hipStreamDestroy(c->cst);
    if (c->fst) (void)hipStreamSynchronize(c->fst);
    for (hipEvent_t e : c->flow_ev) (void)hipEventDestroy(e);
    if (c->fst) (void)hipStreamDestroy(c->fst);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return ASKJA_OK;
}

const char *askja_last_error(const askja_ctx *c) { return c ? c->err.c_str() : "null context"; }

int askja_clear_error(askja_ctx *c)
{
    if (!c) return ASKJA_EINVAL;
    c->sticky = 0;
    c->err.clear();
    return ASKJA_OK;
}

int askja_info(const askja_ctx *c, long *volume, int *nc, int *dim, int *precision)
{
    if (!c) return ASKJA_EINVAL;
    if (volume) *volume = c->V;
    if (nc) *nc = c->nc;
    if (dim) *dim = c->dim;
    if (precision) *precision = c->prec;
    return ASKJA_OK;
}

size_t askja_field_doubles(const askja_ctx *c) { return c ? (size_t)c->V * c->dim * c->nc * c->nc * 2 : 0; }

int askja_links_upload(askja_ctx *c, const double *aos)
{
    ENTER(c);
    if (!aos) return fail(c, ASKJA_EINVAL, "links_upload: null buffer");
    int rc = ensure_staging(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->staging, aos, c->staging_doubles * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_links_from_aos(c->geom, c->staging, c->links[c->cur], c->mode, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return ASKJA_OK;
}

int askja_links_download(askja_ctx *c, double *aos)
{
    ENTER(c);
    if (!aos) return fail(c, ASKJA_EINVAL, "links_download: null buffer");
    int rc = ensure_staging(c);
    if (rc) return rc;
    HIPCHK(c, launch_links_to_aos(c->geom, c->links[c->cur], c->staging, c->mode, c->st));
    HIPCHK(c, hipMemcpyAsync(aos, c->staging, c->staging_doubles * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return ASKJA_OK;
}

int askja_links_download_async(askja_ctx *c, double *pinned)
{
    ENTER(c);
    if (!pinned) return fail(c, ASKJA_EINVAL, "links_download_async: null buffer");
    int rc = ensure_staging(c);  // sizes staging_doubles
    if (rc) return rc;
    if (!c->cst) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->cst, hipStreamNonBlocking));
        HIPCHK(c, hipEventCreateWithFlags(&c->dl_ready, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->dl_done, hipEventDisableTiming));
        HIPCHK(c, hipMalloc((void **)&c->astaging, c->staging_doubles * sizeof(double)));
    }
    // the previous copy must have left the staging buffer before it is refilled
    if (c->dl_pending) HIPCHK(c, hipStreamWaitEvent(c->st, c->dl_done, 0));
    HIPCHK(c, launch_links_to_aos(c->geom, c->links[c->cur], c->astaging, c->mode, c->st));
    HIPCHK(c, hipEventRecord(c->dl_ready, c->st));
    HIPCHK(c, hipStreamWaitEvent(c->cst, c->dl_ready, 0));
    HIPCHK(c, hipMemcpyAsync(pinned, c->astaging, c->staging_doubles * sizeof(double), hipMemcpyDeviceToHost, c->cst));
    HIPCHK(c, hipEventRecord(c->dl_done, c->cst));
    c->dl_pending = true;
    return ASKJA_OK;
}

int askja_download_wait(askja_ctx *c)
{
    ENTER(c);
    if (c->dl_pending) HIPCHK(c, hipEventSynchronize(c->dl_done));
    c->dl_pending = false;
    return ASKJA_OK;
}

int askja_host_alloc(size_t ndoubles, double **out)
{
    if (!out || ndoubles == 0) return ASKJA_EINVAL;
    *out = nullptr;
    if (hipHostMalloc((void **)out, ndoubles * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return ASKJA_ENOMEM;
    }
    return ASKJA_OK;
}

int askja_host_free(double *p)
{
    if (p && hipHostFree(p) != hipSuccess) return ASKJA_EHIP;
    return ASKJA_OK;
}

int askja_links_cold(askja_ctx *c)
{
    ENTER(c);
    HIPCHK(c, launch_links_cold(c->geom, c->links[c->cur], c->mode, c->st));
    return ASKJA_OK;
}

int askja_links_haar(askja_ctx *c, uint64_t seed, uint64_t stream)
{
    ENTER(c);
    HIPCHK(c, launch_project(c->geom, c->links[c->cur], c->mode, 1, seed, stream, c->st));
    return ASKJA_OK;
}

int askja_reunitarize(askja_ctx *c)
{
    ENTER(c);
    HIPCHK(c, launch_project(c->geom, c->links[c->cur], c->mode, 0, 0, 0, c->st));
    return ASKJA_OK;
}

int askja_snapshot(askja_ctx *c, int op)
{
    ENTER(c);
    const size_t bytes = c->link_reals * elem_bytes(c);
    if (op == 0) {
        if (!c->snap) HIPCHK(c, hipMalloc(&c->snap, bytes));
        HIPCHK(c, hipMemcpyAsync(c->snap, c->links[c->cur], bytes, hipMemcpyDeviceToDevice, c->st));
        return ASKJA_OK;
    }
    if (op == 1) {
        if (!c->snap) return fail(c, ASKJA_EINVAL, "snapshot restore without a saved snapshot");
        HIPCHK(c, hipMemcpyAsync(c->links[c->cur], c->snap, bytes, hipMemcpyDeviceToDevice, c->st));
        return ASKJA_OK;
    }
    return fail(c, ASKJA_EINVAL, "snapshot: op must be 0 (save) or 1 (restore)");
}

int askja_mom_upload(askja_ctx *c, const double *aos)
{
    ENTER(c);
    if (!aos) return fail(c, ASKJA_EINVAL, "mom_upload: null buffer");
    int rc = ensure_staging(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->staging, aos, c->staging_doubles * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_mom_from_aos(c->geom, c->staging, c->mom, c->mode, c->st));
    HIPCHK(c, hipMemcpyAsync(c->mom_prev, c->mom, c->mom_reals * elem_bytes(c), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return ASKJA_OK;
}

static int mom_down(askja_ctx *c, const void *src, double *aos)
{
    int rc = ensure_staging(c);
    if (rc) return rc;
    HIPCHK(c, launch_mom_to_aos(c->geom, src, c->staging, c->mode, c->st));
    HIPCHK(c, hipMemcpyAsync(aos, c->staging, c->staging_doubles * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return ASKJA_OK;
}

int askja_mom_download(askja_ctx *c, double *aos)
{
    ENTER(c);
    if (!aos) return fail(c, ASKJA_EINVAL, "mom_download: null buffer");
    return mom_down(c, c->mom, aos);
}

int askja_mom_download_prev(askja_ctx *c, double *aos)
{
    ENTER(c);
    if (!aos) return fail(c, ASKJA_EINVAL, "mom_download_prev: null buffer");
    return mom_down(c, c->mom_prev, aos);
}

int askja_mom_gsl(askja_ctx *c, unsigned long seed, double beta)
{
#pragma clang fp contract(off)
    ENTER(c);
    if (!(beta > 0)) return fail(c, ASKJA_EINVAL, "mom_gsl: beta must be positive");
    // GSL's mt19937: gsl_rng_set(r, 0) seeds with 4357; the state recursion
    // is the standard one, so std::mt19937 reproduces the stream.
    std::mt19937 mt((std::mt19937::result_type)(seed == 0 ? 4357UL : (seed & 0xffffffffUL)));
    auto uniform_pos = [&mt]() {
        double x;
        do x = mt() / 4294967296.0;
        while (x == 0);
        return x;
    };
    // gsl_ran_gaussian: polar (Marsaglia) method, returning y * sqrt(-2 ln r2 / r2)
    auto gaussian = [&]() {
        double x, y, r2;
        do {
            x = -1 + 2 * uniform_pos();
            y = -1 + 2 * uniform_pos();
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0);
        return 1.0 * y * std::sqrt(-2.0 * std::log(r2) / r2);
    };
    const int n = c->nc, ng = n * n - 1;
    // normalised Cartan generator entries (sun.cpp:170-189): 1*norm and
    // -(m+1)*norm with norm = sqrt(0.5 / TrR(g g)), TrR(g g) = (m+1) + (m+1)^2
    std::vector<double> cn(n > 1 ? n - 1 : 1), cd(n > 1 ? n - 1 : 1);
    for (int m = 0; m + 1 < n; m++) {
        const double tr = (double)(m + 1) + (double)(m + 1) * (double)(m + 1);
        const double norm = std::sqrt(0.5 / tr);
        cn[m] = 1.0 * norm;
        cd[m] = (double)(-m - 1) * norm;
    }
    const double fac = std::sqrt(2 * n / beta);
    const size_t nd = askja_field_doubles(c);
    std::vector<double> aos(nd, 0.0);
    std::vector<double> xi(ng > 0 ? ng : 1);
    const size_t ms = (size_t)2 * n * n;
    for (long i = 0; i < c->V; i++)
        for (int d = 0; d < c->dim; d++) {
            for (int a = 0; a < ng; a++) xi[a] = gaussian();
            double *p = aos.data() + ((size_t)i * c->dim + d) * ms;
            int a = 0;
            for (int r = 0; r < n; r++)
                for (int s = r + 1; s < n; s++) {
                    // sym generator: 0.5 at (r,s),(s,r); antisym: -0.5i at (r,s), +0.5i at (s,r)
                    const double vs = 0.5 * xi[a], va = -0.5 * xi[a + 1], vb = 0.5 * xi[a + 1];
                    p[2 * (r * n + s)] = vs * fac;
                    p[2 * (r * n + s) + 1] = va * fac;
                    p[2 * (s * n + r)] = vs * fac;
                    p[2 * (s * n + r) + 1] = vb * fac;
                    a += 2;
                }
            // diagonal: accumulate in generator order, then scale
            for (int r = 0; r < n; r++) {
                double acc = 0.0;
                bool any = false;
                for (int m = 0; m + 1 < n; m++) {
                    double v;
                    if (r <= m) v = cn[m] * xi[a + m];
                    else if (r == m + 1) v = cd[m] * xi[a + m];
                    else continue;
                    acc = any ? acc + v : v;
                    any = true;
                }
                p[2 * (r * n + r)] = acc * fac;
            }
        }
    int rc = askja_mom_upload(c, aos.data());
    return rc;
}

int askja_mom_refresh(askja_ctx *c, uint64_t seed, uint64_t counter, double beta)
{
    ENTER(c);
    if (!(beta > 0)) return fail(c, ASKJA_EINVAL, "mom_refresh: beta must be positive");
    HIPCHK(c, launch_mom_refresh(c->geom, c->mom, seed, counter, std::sqrt(2.0 * c->nc / beta), c->mode, c->st));
    return ASKJA_OK;
}

int askja_mom_negate(askja_ctx *c)
{
    ENTER(c);
    HIPCHK(c, launch_mom_scale(c->geom, c->mom, -1.0, c->mode, c->st));
    return ASKJA_OK;
}

int askja_kick(askja_ctx *c, double eps)
{
    ENTER(c);
    return do_kick(c, c->links[c->cur], nullptr, c->mom_prev, eps, 0.0, 0, false, -1);
}

int askja_drift(askja_ctx *c, double eps, int level)
{
    ENTER(c);
    double ci[9];
    fill_cinv(ci);
    HIPCHK(c, launch_drift(c->geom, c->links[c->cur], c->mom, eps, level, ci, c->mode, c->st));
    return ASKJA_OK;
}

int askja_kick_drift(askja_ctx *c, double eps_kick, double eps_drift, int level)
{
    ENTER(c);
    const int out = (c->cur + 1) % 3;
    int rc = do_kick(c, c->links[c->cur], c->links[out], nullptr, eps_kick, eps_drift, level, true, -1);
    if (rc) return rc;
    c->cur = out;
    return ASKJA_OK;
}

int askja_md_trajectory(askja_ctx *c, double dt, int ndrift, int level, double *energies)
{
    ENTER(c);
    const int start = c->cur;
    int fin = c->cur;
    int rc = md_pingpong(c, dt, ndrift, level, &fin);
    if (rc) return rc;
    c->cur = fin;
    if (energies) {
        double h[12];
        if ((rc = read_res(c, h, 12))) return rc;
        const double np = (double)n_plaquettes(c), n4 = 4.0 * c->nc;
        energies[0] = np - pair_val(h, 0) / n4;
        energies[1] = pair_val(h, 1);
        energies[2] = np - pair_val(h + 6, 0) / n4;
        energies[3] = pair_val(h + 6, 2);
        if (f32_mode(c->mode)) {
            if ((rc = exact_action(c, c->links[start], &energies[0]))) return rc;
            if ((rc = exact_action(c, c->links[fin], &energies[2]))) return rc;
        }
    }
    return ASKJA_OK;
}

int askja_action(askja_ctx *c, double *s_t, double *s_s)
{
    ENTER(c);
    long np = 0;
    int rc;
    if (c->nc > 8 && c->mode != kModeF64Exact) {
        // MFMA pass (FP64 arithmetic) of ReTr W = sum of the 2(D-1)
        // plaquettes through each link: all links give 4 sum_p ReTr P, the
        // links of direction 0 twice the sum over the temporal plaquettes
        // (each plane (0, k) plaquette holds two of them)
        np = kick_partials(c->geom);
        if ((rc = ensure_part(c, (size_t)np * 6))) return rc;
        const KickArgs a = kick_args(c, c->links[c->cur], nullptr, nullptr, 0.0, 0.0, 0);
        HIPCHK(c, launch_action_f64(a, c->mode, c->st));
        HIPCHK(c, launch_action_split(c->part, np, c->dim, c->res, c->st));
        double h[4];
        if ((rc = read_res(c, h, 4))) return rc;
        const double n = (double)c->nc;
        const double tot = (double)n_plaquettes(c) - pair_val(h, 0) / (4.0 * n);
        if (c->dim == 2) {  // one plane, (0, 1): every plaquette is temporal
            if (s_t) *s_t = tot;
            if (s_s) *s_s = 0.0;
            return ASKJA_OK;
        }
        const double st_ = (double)(c->V * (c->dim - 1)) - pair_val(h, 1) / (2.0 * n);
        if (s_t) *s_t = st_;
        if (s_s) *s_s = tot - st_;
        return ASKJA_OK;
    }
    if ((rc = ensure_part(c, (size_t)plaq_partials(c->geom) * 4))) return rc;
    HIPCHK(c, launch_plaq(c->geom, c->links[c->cur], c->mode, 1. / (double)c->nc, c->part, &np, c->st));
    HIPCHK(c, launch_finalize(c->part, np, 2, c->res, c->st));
    double h[4];
    if ((rc = read_res(c, h, 4))) return rc;
    if (s_t) *s_t = pair_val(h, 0);
    if (s_s) *s_s = pair_val(h, 1);
    return ASKJA_OK;
}

static int kinetic2(askja_ctx *c, const void *Q, double *out2)
{
    long np = 0;
    int rc = ensure_part(c, (size_t)((c->NL + 255) / 256) * 4);
    if (rc) return rc;
    HIPCHK(c, launch_kinetic(c->geom, c->mom, Q, c->mode, c->part, &np, c->st));
    HIPCHK(c, launch_finalize(c->part, np, 2, c->res, c->st));
    double h[4];
    if ((rc = read_res(c, h, 4))) return rc;
    out2[0] = pair_val(h, 0);
    out2[1] = pair_val(h, 1);
    return ASKJA_OK;
}

int askja_kinetic(askja_ctx *c, double *out)
{
    ENTER(c);
    if (!out) return fail(c, ASKJA_EINVAL, "kinetic: null output");
    double v[2];
    int rc = kinetic2(c, nullptr, v);
    if (rc) return rc;
    *out = v[0];
    return ASKJA_OK;
}

int askja_energy_el(askja_ctx *c, double *out)
{
    ENTER(c);
    if (!out) return fail(c, ASKJA_EINVAL, "energy_el: null output");
    double v[2];
    int rc = kinetic2(c, c->mom_prev, v);
    if (rc) return rc;
    *out = v[1];
    return ASKJA_OK;
}

int askja_gauss_law(askja_ctx *c, double *out)
{
    ENTER(c);
    if (!out) return fail(c, ASKJA_EINVAL, "gauss_law: null output");
    int rc;
    const int n = c->nc, nr = 2 * n * n;
    const size_t small_need = (size_t)((c->NL + 255) / 256) * 2 * nr;
    const size_t gen_need = (size_t)gauss_rows(c->geom) * 2 * nr;
    if ((rc = ensure_part(c, n <= 3 ? small_need : gen_need))) return rc;
    double *dout = nullptr;
    HIPCHK(c, hipMallocAsync((void **)&dout, (size_t)2 * nr * sizeof(double), c->st));
    long np = 0;
    hipError_t e = launch_gauss(c->geom, c->links[c->cur], c->mom, c->mode, c->part, &np, dout, c->st);
    std::vector<double> h((size_t)2 * nr);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dout, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipFreeAsync(dout, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return fail(c, ASKJA_EHIP, "gauss_law: %s", hipGetErrorString(e));
    // tr2(M) = ReTr(M M) of the summed matrix (sunmat.h:331-338)
    std::vector<double> m((size_t)nr);
    for (int q = 0; q < nr; q++) m[q] = h[2 * q] + h[2 * q + 1];
    double t = 0.0;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
            const double ar = m[2 * (i * n + k)], ai = m[2 * (i * n + k) + 1];
            const double br = m[2 * (k * n + i)], bi = m[2 * (k * n + i) + 1];
            t += ar * br - ai * bi;
        }
    *out = t;
    return ASKJA_OK;
}

int askja_uviol(askja_ctx *c, double *out)
{
    ENTER(c);
    if (!out) return fail(c, ASKJA_EINVAL, "uviol: null output");
    int rc;
    if ((rc = ensure_part(c, (size_t)((c->NL + 255) / 256 + 1024) * 2))) return rc;
    long np = 0;
    HIPCHK(c, launch_uviol(c->geom, c->links[c->cur], c->mode, c->part, &np, c->st));
    HIPCHK(c, launch_finalize(c->part, np, 1, c->res, c->st));
    double h[2];
    if ((rc = read_res(c, h, 2))) return rc;
    *out = pair_val(h, 0);
    return ASKJA_OK;
}

int askja_wilson_lines(askja_ctx *c, int dir, double *out)
{
    ENTER(c);
    if (!out) return fail(c, ASKJA_EINVAL, "wilson_lines: null output");
    if (dir < 0 || dir >= c->dim) return fail(c, ASKJA_EINVAL, "wilson_lines: dir %d out of range", dir);
    const long nl = c->V / c->ext[dir];
    double *d = nullptr;
    HIPCHK(c, hipMallocAsync((void **)&d, (size_t)2 * nl * sizeof(double), c->st));
    hipError_t e = launch_wlines(c->geom, c->links[c->cur], dir, c->mode, d, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)2 * nl * sizeof(double), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipFreeAsync(d, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return fail(c, ASKJA_EHIP, "wilson_lines: %s", hipGetErrorString(e));
    return ASKJA_OK;
}

int askja_slice_traces(askja_ctx *c, int dir, double *out)
{
    ENTER(c);
    if (!out) return fail(c, ASKJA_EINVAL, "slice_traces: null output");
    if (dir < 0 || dir >= c->dim) return fail(c, ASKJA_EINVAL, "slice_traces: dir %d out of range", dir);
    const long nb = slice_trace_blocks(c->geom), n0 = c->ext[0];
    int rc = ensure_part(c, (size_t)nb * n0 * 2);
    if (rc) return rc;
    double *d = nullptr;
    HIPCHK(c, hipMallocAsync((void **)&d, (size_t)2 * n0 * sizeof(double), c->st));
    hipError_t e = launch_slice_trace(c->geom, c->links[c->cur], dir, c->mode, c->part, c->st);
    if (e != hipSuccess) (void)hipFreeAsync(d, c->st);
    for (long x0 = 0; x0 < n0 && e == hipSuccess; x0++)
        e = launch_finalize(c->part + 2 * x0 * nb, nb, 1, d + 2 * x0, c->st);
    std::vector<double> h((size_t)2 * n0);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipFreeAsync(d, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return fail(c, ASKJA_EHIP, "slice_traces: %s", hipGetErrorString(e));
    const double slice = (double)(c->V / n0);
    for (long x0 = 0; x0 < n0; x0++) out[x0] = (h[2 * x0] + h[2 * x0 + 1]) / slice;
    return ASKJA_OK;
}

int askja_line_phases(askja_ctx *c, int dir, int nlines, const int64_t *starts, double *out)
{
    ENTER(c);
    if (!starts || !out || nlines < 1) return fail(c, ASKJA_EINVAL, "line_phases: bad arguments");
    if (dir < 0 || dir >= c->dim) return fail(c, ASKJA_EINVAL, "line_phases: dir %d out of range", dir);
    for (int k = 0; k < nlines; k++)
        if (starts[k] < 0 || starts[k] >= c->V) return fail(c, ASKJA_EINVAL, "line_phases: start %d out of range", k);
    const size_t nn = (size_t)c->nc * c->nc;
    long *ds = nullptr;
    double *wb = nullptr, *dp = nullptr;
    std::vector<long> hs(starts, starts + nlines);
    hipError_t e = hipMallocAsync((void **)&ds, nlines * sizeof(long), c->st);
    if (e == hipSuccess) e = hipMallocAsync((void **)&wb, (size_t)nlines * 2 * nn * sizeof(double), c->st);
    if (e == hipSuccess) e = hipMallocAsync((void **)&dp, (size_t)nlines * c->nc * sizeof(double), c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(ds, hs.data(), nlines * sizeof(long), hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = launch_line_phases(c->geom, c->links[c->cur], dir, c->mode, nlines, ds, wb, dp, c->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dp, (size_t)nlines * c->nc * sizeof(double), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    for (void *q : {(void *)ds, (void *)wb, (void *)dp})
        if (q) (void)hipFreeAsync(q, c->st);
    const hipError_t e2 = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail(c, ASKJA_EHIP, "line_phases: %s", hipGetErrorString(e));
    return ASKJA_OK;
}

int askja_line_matrix(askja_ctx *c, int dir, const int *x, double *out)
{
    ENTER(c);
    if (!x || !out) return fail(c, ASKJA_EINVAL, "line_matrix: null argument");
    if (dir < 0 || dir >= c->dim) return fail(c, ASKJA_EINVAL, "line_matrix: dir %d out of range", dir);
    long lex = 0;
    for (int d = 0; d < c->dim; d++) {
        if (x[d] < 0 || x[d] >= c->ext[d]) return fail(c, ASKJA_EINVAL, "line_matrix: coordinate %d out of range", d);
        lex += x[d] * c->helparr[d];
    }
    int rc = ensure_matwork(c);
    if (rc) return rc;
    const size_t nr = (size_t)2 * c->nc * c->nc;
    double *d = nullptr;
    HIPCHK(c, hipMallocAsync((void **)&d, nr * sizeof(double), c->st));
    hipError_t e = launch_line_matrix(c->geom, c->links[c->cur], dir, lex, c->mode, c->matwork, d, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, nr * sizeof(double), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipFreeAsync(d, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return fail(c, ASKJA_EHIP, "line_matrix: %s", hipGetErrorString(e));
    return ASKJA_OK;
}

}  // extern "C"

namespace {

// Doubles of block partials one trajectory's kernels write (the kicks' three
// sums per block; the FP64 action pass of the FP32 modes).
size_t trajectory_partials(const askja_ctx *c)
{
    size_t need = (size_t)kick_partials(c->geom) * 6;
    if (f32_mode(c->mode) && c->nc <= 8) need = std::max(need, (size_t)plaq_partials(c->geom) * 4);
    return need;
}

int hmc_trajectory(askja_ctx *c, const askja_hmc_params *p, uint64_t traj_counter, askja_hmc_result *r)
{
    int rc;
    if (p->reunit && (rc = askja_reunitarize(c))) return rc;
    if ((rc = askja_mom_refresh(c, p->seed, traj_counter, p->beta))) return rc;
    int fin = c->cur;
    if ((rc = md_pingpong(c, p->dt, p->ndrift, p->level, &fin))) return rc;
    double h[12];
    if ((rc = read_res(c, h, 12))) return rc;
    const double np = (double)n_plaquettes(c), n4 = 4.0 * c->nc, n2 = 2.0 * c->nc;
    double s0 = np - pair_val(h, 0) / n4, s1 = np - pair_val(h + 6, 0) / n4;
    const double k0 = pair_val(h, 1), k1 = pair_val(h + 6, 2);
    if (f32_mode(c->mode)) {
        if ((rc = exact_action(c, c->links[c->cur], &s0))) return rc;
        if ((rc = exact_action(c, c->links[fin], &s1))) return rc;
    }
    r->h_old = (s0 + k0 / n2) * p->beta;
    r->h_new = (s1 + k1 / n2) * p->beta;
    r->dh = r->h_new - r->h_old;
    r->uniform = metropolis_uniform(p->seed, traj_counter);
    if (r->h_new < r->h_old) r->accept_case = 1;
    else if (r->uniform < std::exp(r->h_old - r->h_new)) r->accept_case = 2;
    else r->accept_case = 0;
    r->accepted = r->accept_case != 0;
    if (r->accepted) c->cur = fin;
    r->action = r->accepted ? s1 : s0;
    return ASKJA_OK;
}

int hmc_check(askja_ctx *c, const askja_hmc_params *p, askja_hmc_result *r, const char *who)
{
    if (!p || !r) return fail(c, ASKJA_EINVAL, "%s: null argument", who);
    if (!(p->beta > 0) || !(p->dt > 0)) return fail(c, ASKJA_EINVAL, "%s: beta and dt must be positive", who);
    return ASKJA_OK;
}

}  // namespace

extern "C" {

int askja_hmc_trajectory(askja_ctx *c, const askja_hmc_params *p, uint64_t traj_counter, askja_hmc_result *r)
{
    ENTER(c);
    int rc = hmc_check(c, p, r, "hmc_trajectory");
    if (rc) return rc;
    return hmc_trajectory(c, p, traj_counter, r);
}

int askja_hmc_trajectory_noalloc(askja_ctx *c, const askja_hmc_params *p, uint64_t traj_counter,
                                 askja_hmc_result *r)
{
    ENTER(c);
    int rc = hmc_check(c, p, r, "hmc_trajectory_noalloc");
    if (rc) return rc;
    int cur = -1;
    HIPCHK(c, hipGetDevice(&cur));
    if (cur != c->dev)
        return fail(c, ASKJA_EINVAL, "hmc_trajectory_noalloc: device %d is current, the context lives on device %d",
                    cur, c->dev);
    if (c->dev < 0 || c->dev >= kMaxDevices || !g_scratch[c->dev])
        return fail(c, ASKJA_EUNSUP, "hmc_trajectory_noalloc: no static buffer on device %d", c->dev);
    const size_t need = trajectory_partials(c);
    if (need > kScratchPartDoubles)
        return fail(c, ASKJA_EINVAL,
                    "hmc_trajectory_noalloc: this lattice needs %zu doubles of partial sums, the static buffer "
                    "holds %zu (16 MiB); use askja_hmc_trajectory",
                    need, kScratchPartDoubles);
    // the trajectory's kernels write their block partials and finalised sums
    // into the static buffer instead of the context's (grown on demand)
    double *part = c->part, *res = c->res;
    const size_t part_doubles = c->part_doubles;
    c->res = g_scratch[c->dev];
    c->part = g_scratch[c->dev] + kScratchResDoubles;
    c->part_doubles = kScratchPartDoubles;
    c->part_static = true;
    rc = hmc_trajectory(c, p, traj_counter, r);
    c->part_static = false;
    c->part = part;
    c->res = res;
    c->part_doubles = part_doubles;
    return rc;
}

int askja_uniform(askja_ctx *c, uint64_t seed, uint64_t counter, double *out)
{
    ENTER(c);
    if (!out) return fail(c, ASKJA_EINVAL, "uniform: null output");
    *out = metropolis_uniform(seed, counter);
    return ASKJA_OK;
}

// --------------------------------------------------------- autocorrelation
int askja_tau_int(const double *x, long n, double S, double *tau, long *window)
{
    if (!x || !tau || n < 8 || !(S > 0)) return ASKJA_EINVAL;
    double mean = 0.0;
    for (long i = 0; i < n; i++) mean += x[i];
    mean /= (double)n;
    std::vector<double> d((size_t)n);
    double var = 0.0;
    for (long i = 0; i < n; i++) {
        d[i] = x[i] - mean;
        var += d[i] * d[i];
    }
    var /= (double)n;
    if (!(var > 0.0)) {
        *tau = 1.0;
        if (window) *window = 0;
        return ASKJA_OK;
    }
    // Wolff's window: the first W where exp(-W/tau_W) - tau_W/sqrt(W n) < 0,
    // tau_W = S / log((2 tint + 1) / (2 tint - 1)) from the running tint
    double tint = 0.5;
    for (long W = 1; W < n / 2; W++) {
        double c = 0.0;
        for (long i = 0; i + W < n; i++) c += d[i] * d[i + W];
        tint += c / (double)(n - W) / var;
        const double tw = tint > 0.5 ? S / std::log((2.0 * tint + 1.0) / (2.0 * tint - 1.0)) : 1e-9;
        if (std::exp(-(double)W / tw) - tw / std::sqrt((double)W * (double)n) < 0.0) {
            *tau = 2.0 * tint;
            if (window) *window = W;
            return ASKJA_OK;
        }
    }
    *tau = 2.0 * tint;
    if (window) *window = n / 2;
    return ASKJA_OK;
}

// ------------------------------------------------------------------ flow
namespace {

int flow_ready(askja_ctx *c)
{
    if (c->flow_nsub <= 0) return fail(c, ASKJA_EINVAL, "flow: askja_flow_define has not been called");
    return ASKJA_OK;
}

size_t link_bytes(const askja_ctx *c) { return c->link_reals * elem_bytes(c); }

// reals of one sub-layer's input log
size_t flow_log_stride(const askja_ctx *c)
{
    if (c->flow_kind == kFlowSpectral) return (size_t)sflow_active(c->geom) * 2 * c->nc * c->nc;
    return flow_log_reals(c->geom);
}

void *flow_log_of(askja_ctx *c, int k)
{
    return (char *)c->flow_log + (size_t)k * flow_log_stride(c) * elem_bytes(c);
}

double *flow_wr_of(askja_ctx *c, int k)
{
    return c->flow_st_wr ? c->flow_st_wr + (size_t)k * stout_wr_doubles(c->geom) : nullptr;
}

double *flow_aux_of(askja_ctx *c, int k)
{
    return c->flow_st_aux ? c->flow_st_aux + (size_t)k * stout_aux_doubles(c->geom) : nullptr;
}

double *flow_lg_of(askja_ctx *c, int k)
{
    return c->flow_st_lg ? c->flow_st_lg + (size_t)k * stout_lg_doubles(c->geom) : nullptr;
}

// U = F(V) into flow_out; with store, the input links of every sub-layer
// go to flow_log.  Every sub-layer writes its log J partials to its own rows
// of the partial buffer, and one finalize sums them all into flow_lj.  With
// store, a stout
// sub-layer's log-det gradient (which needs only that sub-layer's W, rho and
// inverse Jacobians) is launched on the second stream as soon as its forward
// is done, so that it runs beside the later sub-layers' forward and the
// backward; the backward of the sub-layer waits for it.
int flow_forward(askja_ctx *c, bool store)
{
    HIPCHK(c, hipMemcpyAsync(c->flow_out, c->links[c->cur], link_bytes(c), hipMemcpyDeviceToDevice, c->st));
    if (c->flow_kind == kFlowSpectral) {
        const long na = sflow_active(c->geom);
        int rc = ensure_part(c, (size_t)na * 2 * c->flow_nsub);
        if (rc) return rc;
        for (int k = 0; k < c->flow_nsub; k++) {
            const FlowSub &f = c->flow_subs[k];
            const size_t ko = (size_t)k * flow_log_stride(c);
            HIPCHK(c, launch_sflow_forward(c->geom, c->flow_out, f.mu, f.par, f.w, f.rho_max, c->flow_sp_ptr,
                                           store ? (double *)flow_log_of(c, k) : nullptr,
                                           store ? c->flow_sp_inv + ko : nullptr, c->part + (size_t)2 * na * k,
                                           c->st));
        }
        HIPCHK(c, launch_finalize(c->part, na * c->flow_nsub, 1, c->flow_lj, c->st));
        return ASKJA_OK;
    }
    long np_all = 0;
    for (int k = 0; k < c->flow_nsub; k++) np_all += flow_forward_partials(c->geom, c->flow_subs[k].par);
    int rc = ensure_part(c, (size_t)np_all * 2);
    if (rc) return rc;
    long off = 0;  // the sub-layer's first partial row
    for (int k = 0; k < c->flow_nsub; k++) {
        const FlowSub &f = c->flow_subs[k];
        double *pk = c->part + 2 * off;
        off += flow_forward_partials(c->geom, f.par);
        HIPCHK(c, launch_flow_forward(c->geom, c->flow_out, c->flow_out, f, c->mode, pk,
                                      store ? flow_log_of(c, k) : nullptr, flow_wr_of(c, k), c->st));
        // stout: the per-link 8 x 8 log det from the W and rho just written
        // (with store, the inverse Jacobians and eigenbases are kept for the
        // backward)
        if (f.kind == kFlowStout) {
            HIPCHK(c, launch_stout_logdet(c->geom, f, flow_wr_of(c, k), pk, store ? flow_aux_of(c, k) : nullptr,
                                          c->st));
            if (store) {  // c_logj d log det J of this sub-layer's links (c_logj = -1)
                HIPCHK(c, hipEventRecord(c->flow_ev[2 * k], c->st));
                HIPCHK(c, hipStreamWaitEvent(c->fst, c->flow_ev[2 * k], 0));
                HIPCHK(c, launch_stout_logdet_grad(c->geom, f, flow_wr_of(c, k), flow_aux_of(c, k), -1.0,
                                                   flow_lg_of(c, k), c->fst));
                HIPCHK(c, hipEventRecord(c->flow_ev[2 * k + 1], c->fst));
            }
        }
    }
    HIPCHK(c, launch_finalize(c->part, np_all, 1, c->flow_lj, c->st));
    return ASKJA_OK;
}

int flow_logj(askja_ctx *c, double *logj)
{
    double h[2];
    HIPCHK(c, hipMemcpyAsync(h, c->flow_lj, sizeof h, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    *logj = h[0] + h[1];
    return ASKJA_OK;
}

// The spectral kind after its forward: the Wilson gradient of U = F(V) by a
// kick from zero momenta (P_out = -eps TA(W) with eps = -beta / (2N) is
// beta TA(W) / (2N), the Lie gradient of beta S_W), then the sub-layers'
// backward in reverse.  flow_g is in the momentum layout.
int sflow_gradient(askja_ctx *c, double beta, bool wgrad, double *action)
{
    const long np = kick_partials(c->geom);
    int rc = ensure_part(c, (size_t)np * 6);
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(c->flow_g, 0, c->mom_reals * sizeof(double), c->st));
    KickArgs a = kick_args(c, c->flow_out, nullptr, nullptr, -beta / (2.0 * c->nc), 0.0, 0);
    a.P = c->flow_g;
    HIPCHK(c, launch_kick(a, c->mode, false, c->st));
    HIPCHK(c, launch_finalize(c->part, np, 3, c->res, c->st));
    if (wgrad)
        HIPCHK(c, hipMemsetAsync(c->flow_gw, 0, (size_t)c->flow_nblocks * c->flow_bsize * sizeof(double), c->st));
    HIPCHK(c, hipMemcpyAsync(c->flow_work, c->flow_out, link_bytes(c), hipMemcpyDeviceToDevice, c->st));
    for (int k = c->flow_nsub - 1; k >= 0; k--) {
        const FlowSub &f = c->flow_subs[k];
        double *gw = wgrad ? c->flow_gw + (size_t)c->flow_block[k] * c->flow_bsize : nullptr;
        HIPCHK(c, launch_sflow_backward(c->geom, c->flow_work, f.mu, f.par, f.w, f.rho_max, c->flow_sp_ptr,
                                        (double *)flow_log_of(c, k), c->flow_sp_inv + (size_t)k * flow_log_stride(c),
                                        c->flow_g, -1.0, gw, c->flow_sp_dF, c->st));
    }
    if (action) {
        double h[6];
        if ((rc = read_res(c, h, 6))) return rc;
        *action = (double)n_plaquettes(c) - pair_val(h, 0) / (4.0 * c->nc);
    }
    return ASKJA_OK;
}

// Gradient of Phi = beta S_W(F(V)) - log J(V) with respect to V into
// flow_g (and the MLP weights into flow_gw when wgrad); returns S_W(F(V)).
int flow_gradient(askja_ctx *c, double beta, bool wgrad, double *action)
{
    // action == nullptr: no host synchronisation (inside an MD trajectory)
    int rc = flow_forward(c, true);
    if (rc) return rc;
    if (c->flow_kind == kFlowSpectral) return sflow_gradient(c, beta, wgrad, action);
    const long nl = c->NL;
    if ((rc = ensure_part(c, (size_t)((nl + 255) / 256) * 2))) return rc;
    long np = 0;
    HIPCHK(c, launch_wilson_grad(c->geom, c->flow_out, beta / (2.0 * c->nc), c->mode, c->flow_g, c->part, &np, c->st));
    HIPCHK(c, launch_finalize(c->part, np, 1, c->res, c->st));
    if (wgrad)
        HIPCHK(c, hipMemsetAsync(c->flow_gw, 0, (size_t)c->flow_nblocks * c->flow_bsize * sizeof(double), c->st));
    // backward through the sub-layers in reverse on a copy of F(V); each
    // backward launch restores its sub-layer's input links from the log
    HIPCHK(c, hipMemcpyAsync(c->flow_work, c->flow_out, link_bytes(c), hipMemcpyDeviceToDevice, c->st));
    for (int k = c->flow_nsub - 1; k >= 0; k--) {
        const FlowSub &f = c->flow_subs[k];
        double *gw = wgrad ? c->flow_gw + (size_t)c->flow_block[k] * c->flow_bsize : nullptr;
        if (f.kind == kFlowStout)  // its log-det gradient, from the second stream
            HIPCHK(c, hipStreamWaitEvent(c->st, c->flow_ev[2 * k + 1], 0));
        HIPCHK(c, launch_flow_backward(c->geom, c->flow_work, f, c->mode, c->flow_g, -1.0, gw, flow_log_of(c, k),
                                       c->flow_contrib, flow_lg_of(c, k), flow_wr_of(c, k), c->st));
    }
    if (action) {
        double h[2];
        if ((rc = read_res(c, h, 2))) return rc;
        *action = (double)n_plaquettes(c) - pair_val(h, 0) / (4.0 * c->nc);
    }
    return ASKJA_OK;
}

// P <- P - eps_c g, g in flow_g's layout
hipError_t kick_with(askja_ctx *c, const double *g, double eps_c)
{
    if (c->flow_kind == kFlowSpectral) return launch_axpy((double *)c->mom, g, -eps_c, (long)c->mom_reals, c->st);
    return launch_kick_g(c->geom, c->mom, g, eps_c, c->mode, c->st);
}

// P <- P - eps_c flow_g
hipError_t flow_kick(askja_ctx *c, double eps_c) { return kick_with(c, c->flow_g, eps_c); }

// flow_gp <- the Lie gradient of beta S_W at the chain links V (the plain
// force, in flow_g's layout: 9 doubles per link, or the momentum layout for
// the spectral kind, whose gradient is formed by a kick from zero momenta)
int wilson_grad_v(askja_ctx *c, double beta)
{
    const bool sp = c->flow_kind == kFlowSpectral;
    const size_t n = sp ? c->mom_reals : (size_t)c->NL * 9;
    if (!c->flow_gp) HIPCHK(c, hipMalloc((void **)&c->flow_gp, n * sizeof(double)));
    if (sp) {
        HIPCHK(c, hipMemsetAsync(c->flow_gp, 0, n * sizeof(double), c->st));
        KickArgs a = kick_args(c, c->links[c->cur], nullptr, nullptr, -beta / (2.0 * c->nc), 0.0, 0);
        a.P = c->flow_gp;
        a.part = nullptr;
        HIPCHK(c, launch_kick(a, c->mode, false, c->st));
    } else {
        int rc = ensure_part(c, (size_t)((c->NL + 255) / 256) * 2);
        if (rc) return rc;
        long np = 0;
        HIPCHK(c, launch_wilson_grad(c->geom, c->links[c->cur], beta / (2.0 * c->nc), c->mode, c->flow_gp, c->part,
                                     &np, c->st));
    }
    return ASKJA_OK;
}

// the kick of the flow's correction Phi - alpha beta S_W:
// P <- P - eps_c (flow_g - alpha flow_gp)
int correction_kick(askja_ctx *c, double eps_c, double alpha)
{
    HIPCHK(c, kick_with(c, c->flow_g, eps_c));
    HIPCHK(c, kick_with(c, c->flow_gp, -alpha * eps_c));
    return ASKJA_OK;
}

int kinetic_now(askja_ctx *c, double *k);

// H, Metropolis and the reject path of a flow trajectory (links at the end
// in links[cur], the start in flow_save)
int flow_decide(askja_ctx *c, const askja_hmc_params *p, uint64_t traj_counter, double s0, double lj0, double k0,
                double s1, askja_hmc_result *r)
{
    double lj1, k1;
    int rc;
    if ((rc = flow_logj(c, &lj1))) return rc;
    if ((rc = kinetic_now(c, &k1))) return rc;
    r->h_old = p->beta * s0 - lj0 + k0 / (2.0 * c->nc) * p->beta;
    r->h_new = p->beta * s1 - lj1 + k1 / (2.0 * c->nc) * p->beta;
    r->dh = r->h_new - r->h_old;
    r->uniform = metropolis_uniform(p->seed, traj_counter);
    if (r->h_new < r->h_old) r->accept_case = 1;
    else if (r->uniform < std::exp(r->h_old - r->h_new)) r->accept_case = 2;
    else r->accept_case = 0;
    r->accepted = r->accept_case != 0;
    if (!r->accepted)
        HIPCHK(c, hipMemcpyAsync(c->links[c->cur], c->flow_save, link_bytes(c), hipMemcpyDeviceToDevice, c->st));
    r->action = r->accepted ? s1 : s0;
    return ASKJA_OK;
}

int kinetic_now(askja_ctx *c, double *k)
{
    long np = 0;
    int rc = ensure_part(c, (size_t)((c->NL + 255) / 256) * 4);
    if (rc) return rc;
    HIPCHK(c, launch_kinetic(c->geom, c->mom, nullptr, c->mode, c->part, &np, c->st));
    HIPCHK(c, launch_finalize(c->part, np, 2, c->res, c->st));
    double h[4];
    if ((rc = read_res(c, h, 4))) return rc;
    *k = pair_val(h, 0);
    return ASKJA_OK;
}

}  // namespace

namespace {

int flow_free(askja_ctx *c)
{
    void *old[] = {c->flow_w,   c->flow_gw,       c->flow_g,     c->flow_lj,     c->flow_out,    c->flow_work,
                   c->flow_log, c->flow_contrib,  c->flow_save,  c->flow_st_wr,  c->flow_st_aux, c->flow_st_lg,
                   c->flow_sp_ptr, c->flow_sp_dF, c->flow_sp_inv, c->flow_gp};
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (c->fst) HIPCHK(c, hipStreamSynchronize(c->fst));
    for (void *p : old)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c->flow_ev) (void)hipEventDestroy(e);
    c->flow_ev.clear();
    c->flow_w = nullptr;
    c->flow_gw = c->flow_g = c->flow_lj = nullptr;
    c->flow_out = c->flow_work = c->flow_log = c->flow_save = nullptr;
    c->flow_st_wr = c->flow_st_aux = c->flow_st_lg = nullptr;
    c->flow_contrib = c->flow_sp_ptr = c->flow_sp_dF = c->flow_sp_inv = c->flow_gp = nullptr;
    c->flow_nsub = 0;
    return ASKJA_OK;
}

// Spectral map, any N and D: sub-layer k updates U_mu[k](x) for
// x_nu = offset[k] mod 4 (nu = (mu + 1) mod D), t = t_max tanh(MLP(3 features))
int sflow_define(askja_ctx *c, int nsub, const int *mu, const int *offset, const int *block, int nblocks,
                 double t_max)
{
    if (c->dim < 2) return fail(c, ASKJA_EUNSUP, "flow_define_spectral: D >= 2 (plaquettes)");
    if (c->mode != kModeF64 && c->mode != kModeF64Exact)
        return fail(c, ASKJA_EUNSUP, "flow_define_spectral: FP64 precision modes only");
    for (int d = 0; d < c->dim; d++)
        if (c->ext[d] % 4)
            return fail(c, ASKJA_EUNSUP, "flow_define_spectral: every extent must be a multiple of 4 (row masks mod 4)");
    if (nsub < 1 || nsub > 4096 || nblocks < 1 || !mu || !offset || !block)
        return fail(c, ASKJA_EINVAL, "flow_define_spectral: bad sizes or null arrays");
    if (!(t_max > 0.0) || t_max > 0.5)
        return fail(c, ASKJA_EINVAL, "flow_define_spectral: t_max must be in (0, 1/2] (the phase branch's series bound)");
    for (int k = 0; k < nsub; k++)
        if (mu[k] < 0 || mu[k] >= c->dim || offset[k] < 0 || offset[k] > 3 || block[k] < 0 || block[k] >= nblocks)
            return fail(c, ASKJA_EINVAL, "flow_define_spectral: sub-layer %d out of range", k);
    int rc = flow_free(c);
    if (rc) return rc;
    c->flow_kind = kFlowSpectral;
    c->flow_bsize = 32 * 3 + 2 * 32 + 1;
    c->flow_nblocks = nblocks;
    const size_t nd = c->mom_reals;
    HIPCHK(c, hipMalloc(&c->flow_w, (size_t)nblocks * c->flow_bsize * sizeof(float)));
    HIPCHK(c, hipMemset(c->flow_w, 0, (size_t)nblocks * c->flow_bsize * sizeof(float)));
    HIPCHK(c, hipMalloc(&c->flow_gw, (size_t)nblocks * c->flow_bsize * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->flow_g, nd * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->flow_lj, (size_t)2 * nsub * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->flow_out, link_bytes(c)));
    HIPCHK(c, hipMalloc(&c->flow_work, link_bytes(c)));
    HIPCHK(c, hipMalloc(&c->flow_log, (size_t)nsub * flow_log_stride(c) * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->flow_save, link_bytes(c)));
    HIPCHK(c, hipMalloc((void **)&c->flow_sp_ptr, (size_t)c->V * sizeof(double)));
    HIPCHK(c, hipMalloc((void **)&c->flow_sp_dF, (size_t)c->V * sizeof(double)));
    HIPCHK(c, hipMalloc((void **)&c->flow_sp_inv, (size_t)nsub * flow_log_stride(c) * sizeof(double)));
    c->flow_subs.resize(nsub);
    c->flow_block.assign(block, block + nsub);
    for (int k = 0; k < nsub; k++)
        c->flow_subs[k] =
            FlowSub{mu[k], offset[k], c->flow_w + (size_t)block[k] * c->flow_bsize, 3, t_max, 0.0, kFlowSpectral};
    c->flow_nsub = nsub;
    return ASKJA_OK;
}

int flow_define(askja_ctx *c, int kind, int nsub, const int *mu, const int *parity, const int *block, int nblocks,
                double rho_max, double delta)
{
    if (c->nc != 3) return fail(c, ASKJA_EUNSUP, "flow: only SU(3) is supported");
    if (c->dim < 3 || c->dim > 4)
        return fail(c, ASKJA_EUNSUP, "flow: D = 3 or 4 (D >= 3: planes orthogonal to every direction)");
    if (nsub < 1 || nsub > 4096 || nblocks < 1 || !mu || !parity || !block)
        return fail(c, ASKJA_EINVAL, "flow_define: bad sizes or null arrays");
    if (kind == kFlowRankOne && (!(rho_max > 0.0) || rho_max * 2.0 * (c->dim - 1) >= 2.0))
        return fail(c, ASKJA_EINVAL, "flow_define: rho_max must be in (0, 1/(D-1)) for invertibility");
    if (kind == kFlowStout && (!(rho_max > 0.0) || rho_max * 2.0 * (c->dim - 1) >= 1.0))
        return fail(c, ASKJA_EINVAL, "flow_define_stout: rho_max must be in (0, 1/(2(D-1))) for invertibility");
    for (int d = 0; d < c->dim; d++)
        if (c->ext[d] % 2) return fail(c, ASKJA_EUNSUP, "flow: every extent must be even (checkerboard masks)");
    for (int k = 0; k < nsub; k++)
        if (mu[k] < 0 || mu[k] >= c->dim || (parity[k] != 0 && parity[k] != 1) || block[k] < 0 || block[k] >= nblocks)
            return fail(c, ASKJA_EINVAL, "flow_define: sub-layer %d out of range", k);
    const int nf = 3 * (c->dim - 1) * (c->dim - 2) / 2;  // plaquette and two rectangles per plane
    int rc = flow_free(c);
    if (rc) return rc;
    c->flow_kind = kind;
    c->flow_bsize = 32 * nf + 2 * 32 + 1;
    c->flow_nblocks = nblocks;
    HIPCHK(c, hipMalloc(&c->flow_w, (size_t)nblocks * c->flow_bsize * sizeof(float)));
    HIPCHK(c, hipMemset(c->flow_w, 0, (size_t)nblocks * c->flow_bsize * sizeof(float)));
    HIPCHK(c, hipMalloc(&c->flow_gw, (size_t)nblocks * c->flow_bsize * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->flow_g, (size_t)c->NL * 9 * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->flow_lj, (size_t)2 * nsub * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->flow_out, link_bytes(c)));
    HIPCHK(c, hipMalloc(&c->flow_work, link_bytes(c)));
    HIPCHK(c, hipMalloc(&c->flow_log, (size_t)nsub * flow_log_reals(c->geom) * elem_bytes(c)));
    HIPCHK(c, hipMalloc((void **)&c->flow_contrib, flow_contrib_doubles(c->geom) * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->flow_save, link_bytes(c)));
    if (kind == kFlowStout) {
        HIPCHK(c, hipMalloc((void **)&c->flow_st_wr, (size_t)nsub * stout_wr_doubles(c->geom) * sizeof(double)));
        HIPCHK(c, hipMalloc((void **)&c->flow_st_aux, (size_t)nsub * stout_aux_doubles(c->geom) * sizeof(double)));
        HIPCHK(c, hipMalloc((void **)&c->flow_st_lg, (size_t)nsub * stout_lg_doubles(c->geom) * sizeof(double)));
        if (!c->fst) HIPCHK(c, hipStreamCreateWithFlags(&c->fst, hipStreamNonBlocking));
        c->flow_ev.assign((size_t)2 * nsub, nullptr);
        for (hipEvent_t &e : c->flow_ev) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    c->flow_subs.resize(nsub);
    c->flow_block.assign(block, block + nsub);
    for (int k = 0; k < nsub; k++)
        c->flow_subs[k] =
            FlowSub{mu[k], parity[k], c->flow_w + (size_t)block[k] * c->flow_bsize, nf, rho_max, delta, kind};
    c->flow_nsub = nsub;
    return ASKJA_OK;
}

}  // namespace

int askja_flow_define(askja_ctx *c, int nsub, const int *mu, const int *parity, const int *block, int nblocks,
                      double rho_max, double delta)
{
    ENTER(c);
    return flow_define(c, kFlowRankOne, nsub, mu, parity, block, nblocks, rho_max, delta);
}

int askja_flow_define_stout(askja_ctx *c, int nsub, const int *mu, const int *parity, const int *block, int nblocks,
                            double rho_max)
{
    ENTER(c);
    return flow_define(c, kFlowStout, nsub, mu, parity, block, nblocks, rho_max, 0.0);
}

int askja_flow_define_spectral(askja_ctx *c, int nsub, const int *mu, const int *offset, const int *block,
                               int nblocks, double t_max)
{
    ENTER(c);
    return sflow_define(c, nsub, mu, offset, block, nblocks, t_max);
}

int askja_flow_block_size(askja_ctx *c, int *bsize)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if (bsize) *bsize = c->flow_bsize;
    return ASKJA_OK;
}

int askja_flow_set_weights(askja_ctx *c, const float *w, int on_device)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if (!w) return fail(c, ASKJA_EINVAL, "flow_set_weights: null buffer");
    HIPCHK(c, hipMemcpyAsync(c->flow_w, w, (size_t)c->flow_nblocks * c->flow_bsize * sizeof(float),
                             on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return ASKJA_OK;
}

int askja_flow_apply(askja_ctx *c, double *logj)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if ((rc = flow_forward(c, false))) return rc;
    double lj;
    if ((rc = flow_logj(c, &lj))) return rc;
    if (logj) *logj = lj;
    return ASKJA_OK;
}

int askja_flow_download(askja_ctx *c, double *aos)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if (!aos) return fail(c, ASKJA_EINVAL, "flow_download: null buffer");
    if ((rc = ensure_staging(c))) return rc;
    HIPCHK(c, launch_links_to_aos(c->geom, c->flow_out, c->staging, c->mode, c->st));
    HIPCHK(c, hipMemcpyAsync(aos, c->staging, c->staging_doubles * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return ASKJA_OK;
}

int askja_flow_gradient(askja_ctx *c, double beta, int wgrad, double *action, double *logj)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if (!(beta > 0)) return fail(c, ASKJA_EINVAL, "flow_gradient: beta must be positive");
    double s, lj;
    if ((rc = flow_gradient(c, beta, wgrad != 0, &s))) return rc;
    if ((rc = flow_logj(c, &lj))) return rc;
    if (action) *action = s;
    if (logj) *logj = lj;
    return ASKJA_OK;
}

int askja_flow_gradient_download(askja_ctx *c, double *aos)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if (!aos) return fail(c, ASKJA_EINVAL, "flow_gradient_download: null buffer");
    if (c->flow_kind == kFlowSpectral) return mom_down(c, c->flow_g, aos);
    std::vector<double> g((size_t)c->NL * 9);
    HIPCHK(c, hipMemcpyAsync(g.data(), c->flow_g, g.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const int n = 3;
    for (long i = 0; i < c->V; i++)
        for (int d = 0; d < c->dim; d++) {
            const double *b = g.data() + 9 * (size_t)(d * c->V + c->lex2dev[i]);
            double *m = aos + ((size_t)i * c->dim + d) * 2 * n * n;
            const double re[3][3] = {{b[0], b[3], b[5]}, {b[3], b[1], b[7]}, {b[5], b[7], b[2]}};
            const double im[3][3] = {{0, b[4], b[6]}, {-b[4], 0, b[8]}, {-b[6], -b[8], 0}};
            for (int p = 0; p < n; p++)
                for (int q = 0; q < n; q++) {
                    m[2 * (p * n + q)] = re[p][q];
                    m[2 * (p * n + q) + 1] = im[p][q];
                }
        }
    return ASKJA_OK;
}

int askja_flow_weight_gradient(askja_ctx *c, double *out)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if (!out) return fail(c, ASKJA_EINVAL, "flow_weight_gradient: null buffer");
    HIPCHK(c, hipMemcpyAsync(out, c->flow_gw, (size_t)c->flow_nblocks * c->flow_bsize * sizeof(double),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return ASKJA_OK;
}

int askja_flow_hmc_trajectory(askja_ctx *c, const askja_hmc_params *p, uint64_t traj_counter, askja_hmc_result *r)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if (!p || !r) return fail(c, ASKJA_EINVAL, "flow_hmc_trajectory: null argument");
    if (!(p->beta > 0) || !(p->dt > 0) || p->ndrift < 1)
        return fail(c, ASKJA_EINVAL, "flow_hmc_trajectory: beta, dt and ndrift must be positive");
    if (p->reunit && (rc = askja_reunitarize(c))) return rc;
    // the chain state goes to the flow's own buffer (askja_snapshot's slot
    // stays the caller's)
    HIPCHK(c, hipMemcpyAsync(c->flow_save, c->links[c->cur], link_bytes(c), hipMemcpyDeviceToDevice, c->st));
    if ((rc = askja_mom_refresh(c, p->seed, traj_counter, p->beta))) return rc;
    const double cf = 2.0 * c->nc / p->beta;
    double s0, lj0, k0, s1;
    if ((rc = kinetic_now(c, &k0))) return rc;
    if ((rc = flow_gradient(c, p->beta, false, &s0))) return rc;
    if ((rc = flow_logj(c, &lj0))) return rc;
    HIPCHK(c, flow_kick(c, 0.5 * p->dt * cf));
    double ci[9];
    fill_cinv(ci);
    s1 = s0;
    for (int t = 1; t <= p->ndrift; t++) {
        HIPCHK(c, launch_drift(c->geom, c->links[c->cur], c->mom, p->dt, p->level, ci, c->mode, c->st));
        if ((rc = flow_gradient(c, p->beta, false, t == p->ndrift ? &s1 : nullptr))) return rc;
        HIPCHK(c, flow_kick(c, (t < p->ndrift ? 1.0 : 0.5) * p->dt * cf));
    }
    return flow_decide(c, p, traj_counter, s0, lj0, k0, s1, r);
}

int askja_flow_hmc_trajectory_nested(askja_ctx *c, const askja_hmc_params *p, int inner, double alpha,
                                     uint64_t traj_counter, askja_hmc_result *r)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if (!p || !r) return fail(c, ASKJA_EINVAL, "flow_hmc_trajectory_nested: null argument");
    if (!(p->beta > 0) || !(p->dt > 0) || p->ndrift < 1)
        return fail(c, ASKJA_EINVAL, "flow_hmc_trajectory_nested: beta, dt and ndrift must be positive");
    if (inner < 1 || p->ndrift % inner != 0)
        return fail(c, ASKJA_EINVAL, "flow_hmc_trajectory_nested: inner (%d) must be positive and divide ndrift (%d)",
                    inner, p->ndrift);
    if (!(alpha > 0) || !std::isfinite(alpha))
        return fail(c, ASKJA_EINVAL, "flow_hmc_trajectory_nested: alpha must be positive and finite");
    if (p->reunit && (rc = askja_reunitarize(c))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->flow_save, c->links[c->cur], link_bytes(c), hipMemcpyDeviceToDevice, c->st));
    if ((rc = askja_mom_refresh(c, p->seed, traj_counter, p->beta))) return rc;
    const double cf = 2.0 * c->nc / p->beta, h = inner * p->dt;
    const int m = p->ndrift / inner;
    double s0, lj0, k0, s1;
    if ((rc = kinetic_now(c, &k0))) return rc;
    if ((rc = flow_gradient(c, p->beta, false, &s0))) return rc;
    if ((rc = flow_logj(c, &lj0))) return rc;
    if ((rc = wilson_grad_v(c, p->beta))) return rc;
    if ((rc = correction_kick(c, 0.5 * h * cf, alpha))) return rc;
    s1 = s0;
    for (int o = 1; o <= m; o++) {
        // inner leapfrog of alpha beta S_W, the fused kick+drift kernels
        int fin = c->cur;
        if ((rc = md_pingpong(c, p->dt, inner, p->level, &fin, alpha))) return rc;
        c->cur = fin;
        if ((rc = flow_gradient(c, p->beta, false, o == m ? &s1 : nullptr))) return rc;
        if ((rc = wilson_grad_v(c, p->beta))) return rc;
        if ((rc = correction_kick(c, (o < m ? 1.0 : 0.5) * h * cf, alpha))) return rc;
    }
    return flow_decide(c, p, traj_counter, s0, lj0, k0, s1, r);
}

int askja_flow_wilson_lines(askja_ctx *c, int dir, double *out)
{
    ENTER(c);
    int rc = flow_ready(c);
    if (rc) return rc;
    if ((rc = flow_forward(c, false))) return rc;
    void *save = c->links[c->cur];
    c->links[c->cur] = c->flow_out;
    rc = askja_wilson_lines(c, dir, out);
    c->links[c->cur] = save;
    return rc;
}

int askja_sync(askja_ctx *c)
{
    ENTER(c);
    HIPCHK(c, hipStreamSynchronize(c->st));
    return ASKJA_OK;
}

int askja_timer_start(askja_ctx *c)
{
    ENTER(c);
    HIPCHK(c, hipEventRecord(c->ev0, c->st));
    return ASKJA_OK;
}

int askja_timer_stop(askja_ctx *c, double *ms)
{
    ENTER(c);
    HIPCHK(c, hipEventRecord(c->ev1, c->st));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK(c, hipEventElapsedTime(&f, c->ev0, c->ev1));
    if (ms) *ms = f;
    return ASKJA_OK;
}

int askja_device_fields(askja_ctx *c, void **links, void **mom, void **mom_prev)
{
    ENTER(c);
    if (links) *links = c->links[c->cur];
    if (mom) *mom = c->mom;
    if (mom_prev) *mom_prev = c->mom_prev;
    return ASKJA_OK;
}

}  // extern "C"
