// mfma64.hip -- instantiates the BF16x3 MFMA kick for N <= 64.
#include "mfma_impl.hpp"

namespace askja {

hipError_t mfma_kick_64(const KickArgs &a, bool drift, hipStream_t st) { return mfma::kick<64>(a, drift, st); }

}  // namespace askja
