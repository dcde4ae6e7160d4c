// params.hpp -- host pieces shared by the askja and askja-measure programs:
// the positional params-file reader (sun.cpp:803-938; the measure programs'
// reader, measurements4d.cpp:752-882, reads the same leading fields) and the
// eigen-phases of a unitary matrix.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace {

struct Params {
    int NC = 2, DIM = 3;
    std::vector<int> x;
    double BETA = 1, dt = 0.1;
    int STEPS = 100, UPDATE = 10, UPDATE2 = 10, CONFTAKE = 10, LEVEL = 5, VERBOSE = 1, RUNCONTINUE = 0, RANDOM = 0;
    int START = 1, TAKEN = 0, NACC = 0, NAC1 = 0, NREC = 0, NTOT = 0;
    bool has_seed = false;
    unsigned long long seed = 0;
};

// load_parameters (sun.cpp:803-938): whitespace tokens, a header token, then
// (name, value) pairs in a fixed order; names are not checked.
bool load_parameters(const char *path, Params &p)
{
    std::ifstream in(path);
    if (!in.is_open()) {
        printf("\nERROR: params file %s could not be opened\n", path);
        return false;
    }
    std::string name;
    auto rint = [&](int &v) { in >> name >> v; };
    auto rdbl = [&](double &v) { in >> name >> v; };
    in >> name;  // header
    rint(p.NC);
    rint(p.DIM);
    printf("DIM =%i\n", p.DIM);
    if (p.DIM < 1 || p.DIM > 32) {
        printf("\nERROR: DIM = %i not supported\n", p.DIM);
        return false;
    }
    p.x.assign(p.DIM, 1);
    for (int i = 0; i < p.DIM; i++) rint(p.x[i]);
    rdbl(p.BETA);
    printf("BETA =%f\n", p.BETA);
    rdbl(p.dt);
    printf("dt =%f\n", p.dt);
    rint(p.STEPS);
    printf("STEPS =%i\n", p.STEPS);
    rint(p.UPDATE);
    rint(p.UPDATE2);
    rint(p.CONFTAKE);
    printf("CONFTAKE =%i\n", p.CONFTAKE);
    rint(p.LEVEL);
    rint(p.VERBOSE);
    rint(p.RUNCONTINUE);
    rint(p.RANDOM);
    if (!in) {
        printf("\nERROR: params file %s is truncated or malformed\n", path);
        return false;
    }
    if (p.RUNCONTINUE) {
        rint(p.START);
        printf("===> Info: Continuing run at t=%f\n", p.START * p.dt);
        rint(p.TAKEN);
        rint(p.NACC);
        p.NACC = (p.NACC / p.CONFTAKE) * p.CONFTAKE;  // correct for non-stored parts
        rint(p.NAC1);
        rint(p.NREC);
        rint(p.NTOT);
        if (!in) {
            printf("\nERROR: continuation fields missing in %s\n", path);
            return false;
        }
        printf("INFO: Taken=%i Nacc=%i Nac1=%i Nrec=%i Ntot=%i\n", p.TAKEN, p.NACC, p.NAC1, p.NREC, p.NTOT);
        unsigned long long s;
        if (in >> name >> s && name == "SEED") {
            p.has_seed = true;
            p.seed = s;
        }
    }
    if (p.NC < 1 || p.UPDATE < 1 || p.UPDATE2 < 1 || p.CONFTAKE < 1) {
        printf("\nERROR: NC, UPDATE, UPDATE2 and CONFTAKE must be positive\n");
        return false;
    }
    return true;
}

// Eigen-phases of a unitary matrix W, ascending.  W is normal, so it shares
// eigenvectors with the Hermitian H = (e^{i phi} W + e^{-i phi} W^+)/2 for any
// phi; H is diagonalised by cyclic complex Jacobi rotations and each phase is
// the argument of the Rayleigh quotient v^+ W v.  A fixed irrational phi
// separates eigenvalues of H that would coincide for pairs theta, -theta.
inline std::vector<double> unitary_phases(const std::vector<std::complex<double>> &W, int n)
{
    typedef std::complex<double> cd;
    const cd e(std::cos(0.7390851332), std::sin(0.7390851332));
    std::vector<cd> H(n * n), Vm(n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) H[i * n + j] = 0.5 * (e * W[i * n + j] + std::conj(e * W[j * n + i]));
    for (int i = 0; i < n; i++) Vm[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) off += std::norm(H[i * n + j]);
        if (off < 1e-30) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const cd hpq = H[p * n + q];
                const double a = std::abs(hpq);
                if (a < 1e-300) continue;
                const double app = H[p * n + p].real(), aqq = H[q * n + q].real();
                const double tau = (aqq - app) / (2 * a);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
                const double c = 1 / std::sqrt(1 + t * t), s = t * c;
                const cd ph = hpq / a;  // rotation in the (p, q) plane: [c, s ph; -s conj(ph), c]
                for (int k = 0; k < n; k++) {  // H <- H J
                    const cd hkp = H[k * n + p], hkq = H[k * n + q];
                    H[k * n + p] = c * hkp - s * std::conj(ph) * hkq;
                    H[k * n + q] = s * ph * hkp + c * hkq;
                }
                for (int k = 0; k < n; k++) {  // H <- J^+ H
                    const cd hpk = H[p * n + k], hqk = H[q * n + k];
                    H[p * n + k] = c * hpk - s * ph * hqk;
                    H[q * n + k] = s * std::conj(ph) * hpk + c * hqk;
                }
                for (int k = 0; k < n; k++) {  // V <- V J
                    const cd vkp = Vm[k * n + p], vkq = Vm[k * n + q];
                    Vm[k * n + p] = c * vkp - s * std::conj(ph) * vkq;
                    Vm[k * n + q] = s * ph * vkp + c * vkq;
                }
            }
    }
    std::vector<double> out(n);
    for (int k = 0; k < n; k++) {
        cd r = 0;
        for (int i = 0; i < n; i++) {
            cd wv = 0;
            for (int j = 0; j < n; j++) wv += W[i * n + j] * Vm[j * n + k];
            r += std::conj(Vm[i * n + k]) * wv;
        }
        out[k] = std::arg(r);
    }
    // NaN phases (links overflowed by a diverging trajectory) sort last: a
    // strict weak order whatever the values
    std::sort(out.begin(), out.end(),
              [](double a, double b) { return a < b || (!std::isnan(a) && std::isnan(b)); });
    return out;
}

}  // namespace
