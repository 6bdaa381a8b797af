// TT entry points with the reference's names and C++ signatures
// (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:656, :799); see include/ttx_shim.h.
#include "shim_common.h"

float ttdec_half_lu_device_6(float* h_A, const int* ttRanks, const int* ttDemns, const int& calculateTimes,
                             const int& p, bool calError) {
  (void)p;
  return ttx_shim::run(h_A, ttRanks, ttDemns, calculateTimes, calError, false);
}

// The reference streams h_A in p slices because its tensors outgrow 32 GB devices
// (ttdec6_singleGPU.cu:540-796); the whole tensor fits in 288 GB of HBM, so this runs the
// device-resident sweep.
float ttdec_half_lu_host_6(float* h_A, int* ttRanks, const int* ttDemns, const int& calculateTimes, const int& p,
                           const bool calError) {
  (void)p;
  return ttx_shim::run(h_A, ttRanks, ttDemns, calculateTimes, calError, false);
}
