// askja_main.cpp -- the `askja <params>` program on the MI355X library.
//
// Drop-in for the reference's driver (sun.cpp:1059-1373): the same positional
// params file (load_parameters, sun.cpp:803-938), the same output directory
// and file names, the same text rows in constraints-*.dat and Wilson-*.dat,
// the same binary U-*.conf configurations and master-*.params checkpoints
// (store_parameters, sun.cpp:940-1057), so the reference's measure programs
// read the output unchanged.  The MD, energies and observables run on the GPU
// through the C ABI of include/askja.h.
//
// Deliberate differences (the reference's randomness is seeded by time(NULL)
// for every trajectory after the first, sun.cpp:1268-1311, which repeats
// streams within one second and cannot be reproduced):
//   - trajectory 1 of a fresh run with RANDOM 0 uses the reference's own
//     momenta (mt19937 seed 120351616, polar Gaussians: askja_mom_gsl);
//   - later momenta and the Metropolis uniforms come from Philox-4x32-10 with
//     key = run seed and counter = trajectory number;
//   - the run seed is appended to master-*.params after NTOT (both reference
//     readers stop before it), so a resumed run continues the same stream;
//   - the Wilson-*.dat phase columns, left uninitialised by the reference
//     (sun.cpp:1119, 346-360), hold the sorted eigen-phases of the loop.
// Extra options follow the params file: --precision fp64|fp64-exact|fp32|bf16x3,
// --reunit (project the chain state onto SU(N) before every trajectory).
#include "../../include/askja.h"
#include "params.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include <atomic>
#include <thread>

namespace {

// store_parameters (sun.cpp:940-1057): re-reads the original params file
// token by token and writes it with the run's counters appended.
bool store_parameters(const char *infile, const char *outfile, const Params &p, int myt, int mytaken)
{
    // the reference copies infile to infile-tmp first (sun.cpp:947-949), so
    // infile may be outfile itself (a resumed run rewrites its master file)
    std::stringstream myin;
    {
        std::ifstream f(infile);
        if (!f.is_open()) {
            printf("\nERROR: one of the params files %s,%s could not be opened\n", infile, outfile);
            return false;
        }
        myin << f.rdbuf();
    }
    std::ofstream myout(outfile);
    if (!myout.is_open()) {
        printf("\nERROR: one of the params files %s,%s could not be opened\n", infile, outfile);
        return false;
    }
    std::string tok;
    int iv;
    double dv;
    myin >> tok;
    myout << tok << "\n";
    myin >> tok;
    myout << tok << "\t\t";
    myin >> iv;
    myout << iv << "\n";
    myin >> tok;
    myout << tok << "\t\t";
    myin >> iv;
    myout << iv << "\n";
    for (int i = 0; i < p.DIM; i++) {
        myin >> tok;
        myout << tok << "\t\t";
        myin >> iv;
        myout << iv << "\n";
    }
    myin >> tok;
    myout << tok << "\t\t";
    myin >> dv;
    myout << dv << "\n";
    myin >> tok;
    myout << tok << "\t\t";
    myin >> dv;
    myout << dv << "\n";
    const int vals[] = {p.STEPS, p.UPDATE, p.UPDATE2, p.CONFTAKE, p.LEVEL, p.VERBOSE, 1, p.RANDOM};
    const char *seps[] = {"\t\t", "\t\t", "\t\t", "\t\t", "\t", "\t\t", "\t", "\t\t"};
    for (int k = 0; k < 8; k++) {
        myin >> tok;
        myout << tok << seps[k];
        myin >> iv;
        myout << vals[k] << "\n";
    }
    myout << "TIME-STEP\t" << myt << "\n";
    myout << "TAKEN\t\t" << mytaken << "\n";
    myout << "NACC\t\t" << p.NACC << "\n";
    myout << "NAC1\t\t" << p.NAC1 << "\n";
    myout << "NREC\t\t" << p.NREC << "\n";
    myout << "NTOT\t\t" << p.NTOT << "\n";
    myout << "SEED\t\t" << p.seed << "\n";
    return (bool)myout;
}

// Configuration output: the device-to-host copy of a stored configuration
// runs on the library's copy stream while the chain goes on; the copy is
// waited for at the end of the next trajectory (or before the next store, or
// at exit), and the file and master file are then written on a host thread.
// Two page-locked buffers, one writer at a time (joined before the next
// starts and before exit); only the main thread calls the library.
struct ConfWriter {
    double *pinned[2] = {nullptr, nullptr};
    int next = 0;
    std::thread t;
    std::atomic<bool> ok{true};
    // the configuration whose copy is in flight
    bool pending = false;
    std::string conf, master;
    Params snap;
    int tk = 0;
    double *data = nullptr;
    size_t nd = 0;
    bool join()
    {
        if (t.joinable()) t.join();
        return ok.load();
    }
    ~ConfWriter()
    {
        join();
        for (double *p : pinned) askja_host_free(p);
    }
};

int die(askja_ctx *c, int rc, const char *what)
{
    fprintf(stderr, "askja: %s failed (%d): %s\n", what, rc, c ? askja_last_error(c) : "");
    return 1;
}

#define CK(expr)                                 \
    do {                                         \
        int rc_ = (expr);                        \
        if (rc_ != ASKJA_OK) return die(ctx, rc_, #expr); \
    } while (0)

}  // namespace

int main(int argc, char *argv[])
{
    if (argc < 2) {
        printf("ERROR please specify a params file name");
        printf("\nFinished with errors\n");
        return 2;
    }
    const char *paramsfile = argv[1];
    int precision = ASKJA_FP64;
    bool reunit = false;
    for (int a = 2; a < argc; a++) {
        if (!strcmp(argv[a], "--precision") && a + 1 < argc) {
            const char *v = argv[++a];
            if (!strcmp(v, "fp64")) precision = ASKJA_FP64;
            else if (!strcmp(v, "fp64-exact")) precision = ASKJA_FP64_EXACT;
            else if (!strcmp(v, "fp32")) precision = ASKJA_FP32;
            else if (!strcmp(v, "bf16x3")) precision = ASKJA_BF16X3;
            else {
                fprintf(stderr, "askja: unknown precision %s\n", v);
                return 2;
            }
        } else if (!strcmp(argv[a], "--reunit")) {
            reunit = true;
        } else {
            fprintf(stderr, "askja: unknown option %s\n", argv[a]);
            return 2;
        }
    }
    Params P;
    const bool params_ok = load_parameters(paramsfile, P);
    if (params_ok && precision == ASKJA_BF16X3 && !P.RUNCONTINUE && P.NC > 8)
        fprintf(stderr,
                "askja: warning: bf16x3 from the cold start: next to the ordered start the force is small and "
                "BF16x3's error in the staple sum is not, so most trajectories are rejected until the chain is "
                "disordered; thermalise with --precision fp32 or fp64 first, then continue (RUNCONTINUE 1) in "
                "bf16x3\n");
    if (!params_ok) {
        printf("===>Info: Statistics: acc=%i (%i) rec=%i, rate=%f\n", 0, 0, 0, 0.0 / 0.0);
        printf("===>Took %i configurations\n", 0);
        printf("Finished with errors\n");
        return 1;
    }
    const int NC = P.NC, DIM = P.DIM;
    const double BETA = P.BETA, dt = P.dt;
    printf("SU(%i) simulation in %i space dimensions\n", NC, DIM);
    for (int i = 0; i < DIM; i++) printf("Dimension %i: %i points\n", i, P.x[i]);
    askja_ctx *ctx = nullptr;
    CK(askja_create(NC, DIM, P.x.data(), precision, &ctx));

    char outdir[255], buffer[1024];
    snprintf(outdir, sizeof outdir, "output-DIM%i", DIM);
    struct stat st;
    if (stat(outdir, &st) != 0) {
        printf("===> Info: directory %s not present, creating it\n", outdir);
        mkdir(outdir, S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    }
    const int Nt = P.x[0], Ns = DIM > 1 ? P.x[1] : 1;
    char tag[512];
    snprintf(tag, sizeof tag, "N%i-Nt%i-Ns%i-BETA%.2f-DT%.4f-U%.4f-CT%i", NC, Nt, Ns, BETA, dt, P.UPDATE2 * dt,
             P.CONFTAKE);
    snprintf(buffer, sizeof buffer, "%s/constraints-%s.dat", outdir, tag);
    FILE *constraints = fopen(buffer, P.RUNCONTINUE ? "a" : "w");
    snprintf(buffer, sizeof buffer, "%s/Wilson-%s.dat", outdir, tag);
    FILE *wilson = fopen(buffer, P.RUNCONTINUE ? "a" : "w");
    if (!constraints || !wilson) {
        fprintf(stderr, "askja: cannot open output files in %s\n", outdir);
        return 1;
    }
    if (!P.RUNCONTINUE) {
        fputs("#Time\t Gauss-law\t Total Energy\t E_L\t E_M\t Unitary-Violation\n", constraints);
        fputs("#Time\t <W>\t Wilson-Line-Phases\n", wilson);
    }

    // ---- initial state (sun.cpp:1146-1179)
    bool gsl_first = false;
    if (!P.RUNCONTINUE) {
        P.seed = P.RANDOM ? (unsigned long long)time(NULL) : 120351616ULL;
        gsl_first = !P.RANDOM;
        CK(askja_links_cold(ctx));
    } else {
        if (!P.has_seed) P.seed = (unsigned long long)time(NULL);
        snprintf(buffer, sizeof buffer, "%s/U-%s-%i.conf", outdir, tag, P.TAKEN - 1);
        printf("===> Info: Attempting to load configuration %s\n", buffer);
        FILE *f = fopen(buffer, "rb");
        if (!f) {
            printf("Error loading fields file %s ABORTED!\n", buffer);
            printf("Finished with errors\n");
            return 1;
        }
        std::vector<double> u(askja_field_doubles(ctx));
        const size_t got = fread(u.data(), sizeof(double), u.size(), f);
        fclose(f);
        if (got != u.size()) {
            printf("Error loading fields file %s: %zu of %zu values ABORTED!\n", buffer, got, u.size());
            printf("Finished with errors\n");
            return 1;
        }
        CK(askja_links_upload(ctx, u.data()));
    }
    CK(askja_snapshot(ctx, 0));  // takeconfiguration

    auto energy_mag = [&](double &em) -> int {
        double t, s;
        int rc = askja_action(ctx, &t, &s);
        em = t + s;
        return rc;
    };
    auto refresh = [&](uint64_t counter) -> int {
        if (gsl_first) {
            gsl_first = false;
            return askja_mom_gsl(ctx, (unsigned long)P.seed, BETA);
        }
        return askja_mom_refresh(ctx, P.seed, counter, BETA);
    };
    // H = (EnergyMag + sum Tr p^2 / 2N) beta on the current momenta
    auto hamiltonian = [&](double &h) -> int {
        double em, k;
        int rc = energy_mag(em);
        if (rc) return rc;
        if ((rc = askja_kinetic(ctx, &k))) return rc;
        h = (em + k / (2 * NC)) * BETA;
        return ASKJA_OK;
    };
    std::vector<double> wmat((size_t)2 * NC * NC);
    std::vector<int> origin(DIM, 0);
    // WilsonLoop (sun.cpp:307-362): |Tr W|/N at the origin and its phases
    auto wilson_loop = [&](std::vector<double> &evs) -> int {
        int rc = askja_line_matrix(ctx, 0, origin.data(), wmat.data());
        if (rc) return rc;
        std::vector<std::complex<double>> W((size_t)NC * NC);
        std::complex<double> tr = 0;
        for (int q = 0; q < NC * NC; q++) W[q] = std::complex<double>(wmat[2 * q], wmat[2 * q + 1]);
        for (int i = 0; i < NC; i++) tr += W[i * NC + i];
        evs.assign(NC + 1, 0.0);
        evs[0] = std::abs(tr) / NC;
        std::vector<double> ph = unitary_phases(W, NC);
        for (int i = 0; i < NC; i++) evs[i + 1] = ph[i];
        return ASKJA_OK;
    };
    auto write_wilson = [&](double t, const std::vector<double> &evs) {
        fprintf(wilson, "%f\t", t);
        for (int i = 0; i < NC + 1; i++) fprintf(wilson, "%g\t", evs[i]);
        fprintf(wilson, "\n");
        fflush(wilson);
    };

    // |volume average of Tr P / N| of the chain state after every trajectory,
    // for the integrated autocorrelation time reported at the end
    long nlines = 1;
    for (int i = 1; i < DIM; i++) nlines *= P.x[i];
    std::vector<double> lines((size_t)2 * nlines), poly;
    auto polyakov_abs = [&](double &out) -> int {
        int rc = askja_wilson_lines(ctx, 0, lines.data());
        if (rc) return rc;
        double re = 0.0, im = 0.0;
        for (long k = 0; k < nlines; k++) {
            re += lines[2 * k];
            im += lines[2 * k + 1];
        }
        out = std::hypot(re, im) / (double)nlines;
        return ASKJA_OK;
    };

    uint64_t traj_counter = (uint64_t)P.NTOT;
    CK(refresh(traj_counter));
    double CRITENERGY;
    CK(hamiltonian(CRITENERGY));
    printf("\t INFO: Initial energy =%f\n", CRITENERGY);
    // a kick is held back so that it can run fused with the next drift; it
    // is flushed (standalone kick) whenever the state after it is observed
    double pending = dt / 2;
    auto flush = [&]() -> int {
        if (pending == 0) return ASKJA_OK;
        const double e = pending;
        pending = 0;
        return askja_kick(ctx, e);
    };
    auto drift = [&]() -> int {
        if (pending != 0) {
            const double e = pending;
            pending = 0;
            return askja_kick_drift(ctx, e, dt, P.LEVEL);
        }
        return askja_drift(ctx, dt, P.LEVEL);
    };
    CK(flush());
    {
        // t = 0 row (copyDown has made pi = Pi, so E_L = sum tr2(Pi))
        double em, ee, gauss, uv;
        std::vector<double> evs;
        CK(energy_mag(em));
        CK(askja_kinetic(ctx, &ee));
        CK(askja_gauss_law(ctx, &gauss));
        const double energy = 2 * NC * em + ee;
        CK(wilson_loop(evs));
        if (!P.RUNCONTINUE) {
            if (P.VERBOSE)
                printf("t=%f Gauss Law =%g Energy=%g Wx=%g EL=%g EM=%g\n", 0., gauss, energy / (2 * NC) * BETA, evs[0],
                       ee / (2 * NC) * BETA, em * BETA);
            CK(askja_uviol(ctx, &uv));
            fprintf(constraints, "%f\t %g\t %g\t %g\t %f\t %g\n", 0., gauss, energy / (2 * NC) * BETA,
                    ee / (2 * NC) * BETA, em * BETA, uv);
            fflush(constraints);
            write_wilson(0., evs);
        }
    }

    const clock_t c0 = clock();
    const time_t w0 = time(NULL);
    const auto t0 = std::chrono::steady_clock::now();
    ConfWriter out;
    // wait for the in-flight copy, then hand the buffer to the writer thread
    auto start_writer = [&]() -> int {
        if (!out.pending) return 0;
        out.pending = false;
        int rc = askja_download_wait(ctx);
        if (rc != ASKJA_OK) return die(ctx, rc, "askja_download_wait(ctx)");
        if (!out.join()) return 1;
        out.t = std::thread([&out = out, conf = out.conf, master = out.master, snap = out.snap, tk = out.tk,
                             data = out.data, nd = out.nd, paramsfile]() {
            FILE *f = fopen(conf.c_str(), "wb");
            bool ok = f && fwrite(data, sizeof(double), nd, f) == nd;
            if (f) fclose(f);
            if (!ok) fprintf(stderr, "askja: writing %s failed\n", conf.c_str());
            else ok = store_parameters(paramsfile, master.c_str(), snap, tk, snap.TAKEN);
            if (!ok) out.ok = false;
        });
        return 0;
    };
#define CKW()                          \
    do {                               \
        if (int w_ = start_writer()) return w_; \
    } while (0)
    int timekeep;
    for (timekeep = P.START + 1; timekeep < P.STEPS + P.START + 1; timekeep++) {
        CK(drift());
        const bool traj_end = timekeep % P.UPDATE2 == 0;
        if (!traj_end) pending = dt;
        if (timekeep % P.UPDATE == 0) {
            CK(flush());
            double gauss, ee, em, uv;
            std::vector<double> evs;
            CK(askja_gauss_law(ctx, &gauss));
            if (traj_end) CK(askja_kinetic(ctx, &ee));  // no kick this step: pi = Pi
            else CK(askja_energy_el(ctx, &ee));
            CK(energy_mag(em));
            const double energy = 2 * NC * em + ee;
            CK(askja_uviol(ctx, &uv));
            CK(wilson_loop(evs));
            if (P.VERBOSE)
                printf("t=%f Gauss Law =%g Energy=%g Wx=%g EL=%g EM=%g U-Viol=%g \n", timekeep * dt, gauss,
                       energy * BETA / (2 * NC), evs[0], ee * BETA / (2 * NC), em * BETA, uv);
            fprintf(constraints, "%f\t %g\t %g\t %g\t %f\t %g\n", timekeep * dt, gauss, energy * BETA / (2 * NC),
                    ee * BETA / (2 * NC), em * BETA, uv);
            fflush(constraints);
            write_wilson(timekeep * dt, evs);
        }
        if (traj_end) {
            pending = dt / 2;
            CK(flush());
            double myenergy;
            CK(hamiltonian(myenergy));
            printf("\t INFO: Final energy =%f\n", myenergy);
            CKW();  // the last stored configuration's copy overlapped this trajectory
            P.NTOT++;
            // Metropolis uniform: Philox (key = seed, purpose 3, counter)
            bool accept;
            if (myenergy < CRITENERGY) {
                printf("\t ACCEPT case 1: %f<%f\n", myenergy, CRITENERGY);
                P.NACC++;
                P.NAC1++;
                accept = true;
            } else {
                double nummi;
                CK(askja_uniform(ctx, P.seed, traj_counter, &nummi));
                const double bound = std::exp(-myenergy + CRITENERGY);
                if (nummi < bound) {
                    printf("\t ACCEPT case 2: %f<%f\n", nummi, bound);
                    P.NACC++;
                    accept = true;
                } else {
                    printf("\t REJECT %f>%f\n", nummi, bound);
                    P.NREC++;
                    accept = false;
                }
            }
            if (accept) {
                if (reunit) CK(askja_reunitarize(ctx));
                CK(askja_snapshot(ctx, 0));
            } else {
                CK(askja_snapshot(ctx, 1));
            }
            {
                double pa;
                CK(polyakov_abs(pa));
                poly.push_back(pa);
            }
            if (P.VERBOSE) {
                double em;
                CK(energy_mag(em));
                printf(accept ? "===> Info: accepted new configuration, em=%g\n"
                              : "===> Info: loaded old configuration, em=%g\n",
                       em * BETA);
            }
            traj_counter++;
            CK(refresh(traj_counter));
            CK(hamiltonian(CRITENERGY));
            printf("\t INFO: initial energy =%f\n", CRITENERGY);
            pending = dt / 2;
            if (P.NTOT % P.CONFTAKE == 0) {
                CK(flush());
                snprintf(buffer, sizeof buffer, "%s/U-%s-%i.conf", outdir, tag, P.TAKEN);
                printf("==>Info: storing configuration %s\n", buffer);
                // device -> pinned buffer on the copy stream; the file and
                // the master file are written once the copy has landed (at
                // the end of the next trajectory), in order, one writer at a
                // time.  The buffer refilled here was last written by the
                // writer before the one just started, which start_writer
                // has joined.
                CKW();
                const size_t nd = askja_field_doubles(ctx);
                if (!out.pinned[out.next]) CK(askja_host_alloc(nd, &out.pinned[out.next]));
                CK(askja_links_download_async(ctx, out.pinned[out.next]));
                P.TAKEN++;
                out.conf = buffer;
                snprintf(buffer, sizeof buffer, "%s/master-%s.params", outdir, tag);
                out.master = buffer;
                out.snap = P;
                out.tk = timekeep;
                out.data = out.pinned[out.next];
                out.nd = nd;
                out.pending = true;
                out.next ^= 1;
            }
        }
    }
    CK(flush());
    CK(askja_sync(ctx));
    CKW();
#undef CKW
    if (!out.join()) return 1;
    fclose(constraints);
    fclose(wilson);
    const double wall = difftime(time(NULL), w0);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    (void)c0;
    askja_destroy(ctx);
    printf("===>Info: Statistics: acc=%i (%i) rec=%i, rate=%f\n", P.NACC, P.NAC1, P.NREC, (P.NACC * 1.0) / (P.NTOT));
    printf("===>Took %i configurations\n", P.TAKEN);
    printf("===>Info: %d MD steps in %.0f s\n", P.STEPS, wall);
    {
        // integrated autocorrelation time of |P| (Wolff's window); the
        // reference's statistics assume independent samples
        double tau;
        long W;
        if (askja_tau_int(poly.data(), (long)poly.size(), 1.5, &tau, &W) == ASKJA_OK)
            printf("===>Info: tau_int(|P|) = %f (window %ld) over %zu trajectories, %f independent configurations/s\n",
                   tau, W, poly.size(), secs > 0 ? poly.size() / secs / tau : 0.0);
        else printf("===>Info: tau_int(|P|) needs at least 8 trajectories (%zu)\n", poly.size());
    }
    printf("Finished successfully\n");
    return 0;
}
