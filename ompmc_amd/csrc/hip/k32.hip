// FP32 transport kernels: particle state, stepping and sampling in FP32;
// tallies and CSDA range / energy-loss inversion stay FP64.
#include "kernels_impl.h"

namespace omc {

template struct Launch<float>;

}  // namespace omc
