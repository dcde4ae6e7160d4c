// team32.hip -- instantiates the workgroup-per-link MD kernels for
// N <= 32 (LDS slots of 32x32, 256 threads per link).
#include "team_impl.hpp"

namespace askja {

hipError_t team_kick_32(const KickArgs &a, int mode, bool drift, hipStream_t st)
{
    return team::kick<32, 256>(a, mode, drift, st);
}

hipError_t team_drift_32(const Geom &g, void *U, const void *P, double eps, int level, const double *c, int mode,
                       hipStream_t st)
{
    return team::drift<32, 256>(g, U, P, eps, level, c, mode, st);
}

hipError_t team_project_32(const Geom &g, void *U, int mode, int haar, uint64_t seed, uint64_t stream, hipStream_t st)
{
    return team::project<32, 256>(g, U, mode, haar, seed, stream, st);
}

hipError_t team_plaq_32(const Geom &g, const void *U, int mode, double inv_n, double *part, hipStream_t st)
{
    return team::plaq<32, 256>(g, U, mode, inv_n, part, st);
}

hipError_t team_lines_32(const Geom &g, const void *U, int dir, long nlines, long lex1, int mode, double *traces,
                      double *wfull, hipStream_t st)
{
    return team::lines<32, 256>(g, U, dir, nlines, lex1, mode, traces, wfull, st);
}

hipError_t team_uviol_32(const Geom &g, const void *U, int mode, long nwg, double *part, hipStream_t st)
{
    return team::uviol<32, 256>(g, U, mode, nwg, part, st);
}

hipError_t team_gauss_32(const Geom &g, const void *U, const void *P, int mode, long nwg, double *part, hipStream_t st)
{
    return team::gauss<32, 256>(g, U, P, mode, nwg, part, st);
}

}  // namespace askja
