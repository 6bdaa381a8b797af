// TR entry points with the reference's names and C++ signatures
// (trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:668, :827); see include/ttx_shim.h.
#include "shim_common.h"

float ttdec_half_lu_device_6(float* h_A, const int* ttRanks, const int* ttDemns, const int& calculateTimes,
                             const int& p, bool calError) {
  (void)p;
  return ttx_shim::run(h_A, ttRanks, ttDemns, calculateTimes, calError ? 1 : 0, true);
}

// The reference streams h_A in p column slices (trdec6_singleGPU.cu:668-826); so does this, through
// ttx_tr_svd_host (N_1 <= 64 and r_0 r_1 <= 64; otherwise the device-resident sweep)
float ttdec_half_lu_host_6(float* h_A, int* ttRanks, const int* ttDemns, const int& calculateTimes, const int& p,
                           const bool calError) {
  return ttx_shim::run(h_A, ttRanks, ttDemns, calculateTimes, calError ? 1 : 0, true, 6, p > 0 ? p : 1);
}

// third-order TR driver with the reference's name and C++ signature
// (trdecomposition/third_order_GPU/opt_GPU/trdec_gpu_float_rsvd_lu_half.cu:734, func.h:15)
float trdec_gpu_float(float* h_tensor, int* h_n_array, int* h_r_array, double* eps, float* singleError,
                      float* singleComRatio) {
  (void)eps;
  return ttx_shim::run3(h_tensor, h_n_array, h_r_array, singleError, singleComRatio, true);
}

// the harness's input generator with the reference's name and C++ signature (trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:231)
void generalTRTensor(float* h_A, const int* ttRanks, const int* ttDemns, const int& p) {
  (void)p;
  ttx_shim::gen_ref(h_A, ttRanks, ttDemns, true);
}
