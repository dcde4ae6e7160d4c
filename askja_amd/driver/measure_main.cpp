// askja-measure: the measurement program of the reference (measure4d,
// measurements4d.cpp:930-1306) on the MI355X engine.  Same command line
//   askja-measure <master params> [SKIP] [VERBOSE] [GET_EV_DIST]
// same inputs (output-DIM<D>/U-*.conf of configurations SKIP .. TAKEN-1, TAKEN
// from the master params file) and the same output files in rav-DIM<D>/:
//   Nrav-*.dat   running means (first pass) and the final means with errors
//                of the Polyakov loop, the x1 Wilson line and the temporal /
//                spatial plaquette sums (measurements4d.cpp:1148-1236)
//   W24-*.dat    mean and error of P^2 + W_1^2 + sum_{d >= 2} W_d^2 (1244-1245)
//   P5corr-*.dat the connected slice correlator of ReTr U_{D-1} (1247-1253)
//   Hist-*.dat   histograms of the per-configuration mean sorted eigen-phases
//                of the Polyakov and x1 Wilson lines (1255-1261)
// The line traces, slice traces, plaquette sums and line eigen-phases run on
// the device (askja_wilson_lines, askja_slice_traces, askja_action,
// askja_line_phases); the two passes of the reference over the files read
// every configuration twice, here the per-configuration values of the first
// read are kept (the reference recomputes identical values).
// Deliberate differences: the histograms are zero when GET_EV_DIST is 0 (the
// reference leaves them uninitialised), and an eigen-phase that falls outside
// [-pi, pi) after the winding shift is reported and not counted (the
// reference writes one element past its histogram for bin == bins).
#include "../../include/askja.h"
#include "params.hpp"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

namespace {

int die(askja_ctx *c, int rc, const char *what)
{
    fprintf(stderr, "askja-measure: %s failed (%d): %s\n", what, rc, c ? askja_last_error(c) : "");
    printf("Finished with errors\n");
    return 1;
}

#define CK(expr)                                          \
    do {                                                  \
        int rc_ = (expr);                                 \
        if (rc_ != ASKJA_OK) return die(ctx, rc_, #expr); \
    } while (0)

struct Conf {
    double P, Wx, Ut, Us, W2;
    std::vector<double> P5, Wdist, Pdist;
};

// mean over x0 of |line trace| of the lines in direction d through
// (x0, 0, ..., 0) (measure4d's WilsonLoop(evs, d); d = 0: PolyakovLoop,
// mean over x1 of the lines through (0, x1, 0, ..., 0))
double line_mean(const std::vector<double> &lines, const Params &p, int d)
{
    const long V = [&] {
        long v = 1;
        for (int x : p.x) v *= x;
        return v;
    }();
    const int along = d == 0 ? 1 : 0;  // the coordinate that varies
    const long stride = V / ((long)p.x[d] * p.x[along]);
    double acc = 0;
    for (int k = 0; k < p.x[along]; k++) acc += std::hypot(lines[2 * k * stride], lines[2 * k * stride + 1]);
    return acc / p.x[along];
}

}  // namespace

int main(int argc, char *argv[])
{
    if (argc < 2) {
        printf("ERROR please specify a params file name");
        return 1;
    }
    const char *paramsfile = argv[1];
    int SKIP = argc > 2 ? atoi(argv[2]) : 0;
    int VERBOSE = argc > 3 ? atoi(argv[3]) : 0;
    int GET_EV_DIST = argc > 4 ? atoi(argv[4]) : 0;
    Params p;
    if (!load_parameters(paramsfile, p)) {
        printf("Finished with errors\n");
        return 1;
    }
    if (p.DIM < 2) {
        printf("ERROR: the measurement needs DIM >= 2\n");
        printf("Finished with errors\n");
        return 1;
    }
    const int TAKEN = p.RUNCONTINUE ? p.TAKEN : 0;
    const int nconf = TAKEN - SKIP;
    if (nconf < 1) {
        printf("ERROR: no configurations to measure (TAKEN=%i, SKIP=%i)\n", TAKEN, SKIP);
        printf("Finished with errors\n");
        return 1;
    }
    printf("SU(%i) simulation in %i space dimensions\n", p.NC, p.DIM);

    askja_ctx *ctx = nullptr;
    CK(askja_create(p.NC, p.DIM, p.x.data(), ASKJA_FP64, &ctx));
    long V = 1;
    for (int x : p.x) V *= x;

    char buffer[512], outdir[64], indir[64];
    snprintf(outdir, sizeof outdir, "rav-DIM%i", p.DIM);
    struct stat st;
    if (stat(outdir, &st) != 0) {
        printf("===> Info: directory %s not present, creating it\n", outdir);
        mkdir(outdir, S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    }
    snprintf(indir, sizeof indir, "output-DIM%i", p.DIM);
    const int vol = p.x[0];
    const double width = M_PI / (2 * p.NC);
    const int bins = (int)std::floor(2 * M_PI / width);
    auto fname = [&](const char *kind) {
        snprintf(buffer, sizeof buffer, "%s/%s-N%i-Nt%i-Ns%i-BETA%.2f-DT%.4f-U%.4f-CT%i.dat", outdir, kind, p.NC,
                 p.x[0], p.x[1], p.BETA, p.dt, p.UPDATE2 * p.dt, p.CONFTAKE);
        return std::string(buffer);
    };
    std::ofstream rav(fname("Nrav")), W2(fname("W24")), P5corr(fname("P5corr")), Hist(fname("Hist"));
    printf("===> Info: opening files finished\n");

    // line starts (lexicographic): Wilson lines along x1 through (x0, 0, ..),
    // Polyakov lines along x0 through (0, x1, 0, ..)
    long h1 = 1;
    for (int d = 2; d < p.DIM; d++) h1 *= p.x[d];
    std::vector<int64_t> wstart(p.x[0]), pstart(p.x[1]);
    for (int k = 0; k < p.x[0]; k++) wstart[k] = (int64_t)k * p.x[1] * h1;
    for (int k = 0; k < p.x[1]; k++) pstart[k] = (int64_t)k * h1;

    const size_t nd = (size_t)V * p.DIM * 2 * p.NC * p.NC;
    std::vector<double> u(nd);
    std::vector<Conf> confs;
    std::vector<int> Whist(bins, 0), Phist(bins, 0);
    double Psum = 0, Wssum = 0, W2sum = 0, Utsum = 0, Ussum = 0;
    std::vector<double> P5sum(p.x[0], 0.0);
    for (int conf = SKIP; conf < TAKEN; conf++) {  // first pass
        snprintf(buffer, sizeof buffer, "%s/U-N%i-Nt%i-Ns%i-BETA%.2f-DT%.4f-U%.4f-CT%i-%i.conf", indir, p.NC, p.x[0],
                 p.x[1], p.BETA, p.dt, p.UPDATE2 * p.dt, p.CONFTAKE, conf);
        FILE *f = fopen(buffer, "rb");  // loadconf (measurements4d.cpp:161-182)
        if (!f || fread(u.data(), sizeof(double), nd, f) != nd) {
            printf("\nERROR: configuration file %s could not be read\n", buffer);
            if (f) fclose(f);
            askja_destroy(ctx);
            printf("Finished with errors\n");
            return 1;
        }
        fclose(f);
        CK(askja_links_upload(ctx, u.data()));
        Conf c;
        std::vector<double> lines;
        lines.assign(2 * (V / p.x[0]), 0.0);
        CK(askja_wilson_lines(ctx, 0, lines.data()));
        c.P = line_mean(lines, p, 0);
        c.W2 = c.P * c.P;
        for (int d = 1; d < p.DIM; d++) {
            if (d >= 2 && p.x[d] <= 1) continue;
            lines.assign(2 * (V / p.x[d]), 0.0);
            CK(askja_wilson_lines(ctx, d, lines.data()));
            const double w = line_mean(lines, p, d);
            if (d == 1) c.Wx = w;
            c.W2 += w * w;
        }
        c.P5.assign(p.x[0], 0.0);
        CK(askja_slice_traces(ctx, p.DIM - 1, c.P5.data()));
        double Ut, Us;
        CK(askja_action(ctx, &Ut, &Us));
        c.Ut = Ut * p.BETA / vol;
        c.Us = Us * p.BETA / vol;
        if (GET_EV_DIST) {  // getevs (measurements4d.cpp:185-309)
            std::vector<double> ph((size_t)p.x[0] * p.NC);
            CK(askja_line_phases(ctx, 1, p.x[0], wstart.data(), ph.data()));
            c.Wdist.assign(p.NC, 0.0);
            for (int k = 0; k < p.x[0]; k++)
                for (int i = 0; i < p.NC; i++) c.Wdist[i] += ph[(size_t)k * p.NC + i] / p.x[0];
            ph.assign((size_t)p.x[1] * p.NC, 0.0);
            CK(askja_line_phases(ctx, 0, p.x[1], pstart.data(), ph.data()));
            c.Pdist.assign(p.NC, 0.0);
            for (int k = 0; k < p.x[1]; k++)
                for (int i = 0; i < p.NC; i++) c.Pdist[i] += ph[(size_t)k * p.NC + i] / p.x[1];
            for (int i = 0; i < p.NC; i++) {
                const int gw = (int)std::floor((c.Wdist[i] + M_PI) / width);
                const int gp = (int)std::floor((c.Pdist[i] + M_PI) / width);
                if (gw < 0 || gw >= bins) printf("ERRROR in histogram binning!\n");
                else Whist[gw]++;
                if (gp < 0 || gp >= bins) printf("ERRROR in histogram binning!\n");
                else Phist[gp]++;
            }
        }
        if (VERBOSE) printf("conf=%i Pt=%f Wx=%f Ut=%f Us=%f\n", conf, c.P, c.Wx, c.Ut, c.Us);
        Wssum += c.Wx;
        Psum += c.P;
        W2sum += c.W2;
        Utsum += c.Ut;
        Ussum += c.Us;
        for (int x0 = 0; x0 < p.x[0]; x0++) P5sum[x0] += c.P5[x0];
        printf("Conf %i first pass done!\n", conf);
        if (conf > SKIP)
            rav << conf << "\t" << Psum / (conf - SKIP + 1) << "\t 0.\t" << Wssum / (conf - SKIP + 1) << "\t 0.\t"
                << Utsum / (conf - SKIP + 1) << "\t 0.\t" << Ussum / (conf - SKIP + 1) << "\t 0.\n";
        confs.push_back(c);
    }
    Psum /= nconf;
    Wssum /= nconf;
    W2sum /= nconf;
    Utsum /= nconf;
    Ussum /= nconf;
    for (double &v : P5sum) v /= nconf;

    double dW = 0, dPt = 0, dUt = 0, dUs = 0, dW2 = 0;
    std::vector<double> dP5sum(p.x[0] / 2, 0.0);
    for (int k = 0; k < nconf; k++) {  // second pass
        const Conf &c = confs[k];
        dW += (c.Wx - Wssum) * (c.Wx - Wssum);
        dPt += (c.P - Psum) * (c.P - Psum);
        const double temp = c.W2 - W2sum;
        dW2 += temp * temp;
        dUt += (c.Ut - Utsum) * (c.Ut - Utsum);
        dUs += (c.Us - Ussum) * (c.Us - Ussum);
        const int n0 = p.x[0];
        for (int x0 = 0; x0 < n0 / 2; x0++)
            dP5sum[x0] += 0.5 * ((c.P5[x0] - P5sum[x0]) * (c.P5[0] - P5sum[0]) +
                                 (c.P5[n0 - x0 - 1] - P5sum[n0 - x0 - 1]) * (c.P5[0] - P5sum[0]));
        printf("Conf %i second pass done!\n", SKIP + k);
    }
    const double nn = (double)nconf * nconf;
    dW /= nn;
    dW2 /= nn;
    dPt /= nn;
    dUt /= nn;
    dUs /= nn;
    for (double &v : dP5sum) v /= nn;

    rav << TAKEN << "\t" << Psum << "\t" << std::sqrt(dPt) << "\t" << Wssum << "\t" << std::sqrt(dW) << "\t" << Utsum
        << "\t" << std::sqrt(dUt) << "\t" << Ussum << "\t" << std::sqrt(dUs) << "\t" << p.x[0] << "\t" << p.x[1]
        << "\t" << p.BETA << "\t" << std::endl;
    W2 << "#NC" << "\t" << "BETA" << "\t" << "Ns" << "\t" << "W2" << "\t" << "dW2" << "\t" << "conf" << "\t" << "SKIP"
       << std::endl;
    W2 << p.NC << "\t" << p.BETA << "\t" << p.x[1] << "\t" << W2sum << "\t" << std::sqrt(dW2) << "\t" << nconf << "\t"
       << SKIP << std::endl;
    for (int x0 = 0; x0 < p.x[0] / 2; x0++) P5corr << x0 << "\t" << dP5sum[x0] << "\t" << std::endl;
    for (int i = 0; i < bins; i++) {
        const double center = -M_PI + (i + 0.5) * width;
        Hist << center << "\t" << Phist[i] / (1.0 * nconf) << "\t" << Whist[i] / (1.0 * nconf) << std::endl;
    }
    printf("Nt=%i Ns=%i BETA=%.2f <Wt>=%f+-%f <Wx>=%f+-%f Ut=%f+-%f Us=%f+-%f\n", p.x[0], p.x[1], p.BETA, Psum,
           std::sqrt(dPt), Wssum, std::sqrt(dW), Utsum, std::sqrt(dUt), Ussum, std::sqrt(dUs));
    askja_destroy(ctx);
    printf("Finished successfully\n");
    return 0;
}
