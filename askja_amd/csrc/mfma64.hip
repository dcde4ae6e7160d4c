// mfma64.hip -- instantiates the MFMA kick for N <= 64 (FP64, FP32, BF16x3).
#include "mfma_impl.hpp"

namespace askja {

hipError_t mfma_kick_64(const KickArgs &a, int mode, bool drift, hipStream_t st)
{
    switch (mode) {
    case kModeF64: return mfma::kick<mfma::kF64, 64>(a, drift, st);
    case kModeF32: return mfma::kick<mfma::kF32, 64>(a, drift, st);
    default: return mfma::kick<mfma::kBF16X3, 64>(a, drift, st);
    }
}

hipError_t mfma_drift_64(const Geom &g, void *U, const void *P, double eps, int level, const double *cinv, int mode,
                         hipStream_t st)
{
    switch (mode) {
    case kModeF64: return mfma::drift<mfma::kF64, 64>(g, U, P, eps, level, cinv, st);
    case kModeF32: return mfma::drift<mfma::kF32, 64>(g, U, P, eps, level, cinv, st);
    default: return mfma::drift<mfma::kBF16X3, 64>(g, U, P, eps, level, cinv, st);
    }
}

hipError_t mfma_action_64(const KickArgs &a, bool f32_links, hipStream_t st)
{
    return mfma::action_f64<64>(a, f32_links, st);
}

}  // namespace askja
