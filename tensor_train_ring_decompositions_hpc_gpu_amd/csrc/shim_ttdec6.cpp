// TT entry points with the reference's names and C++ signatures
// (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:656, :799); see include/ttx_shim.h.
#include "shim_common.h"

float ttdec_half_lu_device_6(float* h_A, const int* ttRanks, const int* ttDemns, const int& calculateTimes,
                             const int& p, bool calError) {
  (void)p;
  return ttx_shim::run(h_A, ttRanks, ttDemns, calculateTimes, calError ? 1 : 0, false);
}

// The reference streams h_A in p column slices (ttdec6_singleGPU.cu:540-796); so does this, through
// ttx_tt_svd_host (N_1 <= 64; a wider first mode runs the device-resident sweep)
float ttdec_half_lu_host_6(float* h_A, int* ttRanks, const int* ttDemns, const int& calculateTimes, const int& p,
                           const bool calError) {
  return ttx_shim::run(h_A, ttRanks, ttDemns, calculateTimes, calError ? 1 : 0, false, 6, p > 0 ? p : 1);
}

// third-order TT driver with the reference's name and C++ signature
// (ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:778, func.h:16)
float ttdec_gpu_float(float* h_tensor, int* h_n_array, int* h_r_array, double* eps, float* singleError,
                      float* singleComRatio) {
  (void)eps;
  return ttx_shim::run3(h_tensor, h_n_array, h_r_array, singleError, singleComRatio, false);
}

// the harness's input generator with the reference's name and C++ signature (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:185)
void generalTTTensor(float* h_A, const int* ttRanks, const int* ttDemns, const int& p) {
  (void)p;
  ttx_shim::gen_ref(h_A, ttRanks, ttDemns, false);
}
