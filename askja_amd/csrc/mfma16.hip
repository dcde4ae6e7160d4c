// mfma16.hip -- instantiates the BF16x3 MFMA kick for N <= 16.
#include "mfma_impl.hpp"

namespace askja {

hipError_t mfma_kick_16(const KickArgs &a, bool drift, hipStream_t st) { return mfma::kick<16>(a, drift, st); }

}  // namespace askja
