// mfma32.hip -- instantiates the BF16x3 MFMA kick for N <= 32.
#include "mfma_impl.hpp"

namespace askja {

hipError_t mfma_kick_32(const KickArgs &a, bool drift, hipStream_t st) { return mfma::kick<32>(a, drift, st); }

}  // namespace askja
