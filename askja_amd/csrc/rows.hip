// rows.hip -- instantiates the row-per-thread kick for 4 <= N <= 8.
#include "rows_impl.hpp"

namespace askja {

namespace {
template <int N>
hipError_t rows_n(const KickArgs &a, int mode, bool drift, hipStream_t st)
{
    switch (mode) {
    case kModeF64: return rows::kick<N, double, false>(a, drift, st);
    case kModeF64Exact: return rows::kick<N, double, true>(a, drift, st);
    default: return rows::kick<N, float, false>(a, drift, st);
    }
}
}  // namespace

hipError_t rows_kick(const KickArgs &a, int mode, bool drift, hipStream_t st)
{
    switch (a.g.N) {
    case 4: return rows_n<4>(a, mode, drift, st);
    case 5: return rows_n<5>(a, mode, drift, st);
    case 6: return rows_n<6>(a, mode, drift, st);
    case 7: return rows_n<7>(a, mode, drift, st);
    case 8: return rows_n<8>(a, mode, drift, st);
    default: return hipErrorInvalidValue;
    }
}

long rows_kick_blocks(const Geom &g)
{
    const long lpb = rows::kT / g.N;
    return (g.V * g.D + lpb - 1) / lpb;
}

}  // namespace askja
