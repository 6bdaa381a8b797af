/* Reference-name drop-in shims.
 *
 * Two libraries export the reference's six-mode entry points with their exact C++ signatures; the
 * TT and TR versions carry the same names in the reference, so each lives in its own library:
 *
 *   libttx_ttdec6.so  <- ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:656 (_host_6) and :799 (_device_6)
 *   libttx_trdec6.so  <- trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:668 (_host_6) and :827 (_device_6)
 *
 *   float ttdec_half_lu_device_6(float* h_A, const int* ttRanks, const int* ttDemns,
 *                                const int& calculateTimes, const int& p, bool calError);
 *   float ttdec_half_lu_host_6(float* h_A, int* ttRanks, const int* ttDemns,
 *                              const int& calculateTimes, const int& p, const bool calError);
 *
 * Semantics kept from the reference: h_A is the caller's host tensor (column-major, 6 modes,
 * ttDemns[0..5]); ttRanks[0..6] are the TT ranks (1, r1..r5, 1) or the TR ranks (r0, r1..r5, r0,
 * first step of rank r0 r1); the timed loop runs calculateTimes decompositions, each including the
 * host-to-device copy of h_A, and the function returns the mean seconds per decomposition; with
 * calError the relative reconstruction error ||A - A_hat||_F / ||A||_F is printed per run, as the
 * reference's calMSE line.  p is the number of column slices in which
 * ttdec_half_lu_host_6 streams h_A from host memory (ttx_tt_svd_host / ttx_tr_svd_host; N_1 <= 64,
 * TR also r_0 r_1 <= 64, else the device path), as the reference's host variants do; the device
 * variants ignore it (the reference slices its GEMMs with it).  Differences: h_A is never overwritten; errors print a message and return -1 instead of
 * aborting through assert; the cores are not freed but kept for ttx_last_cores.
 *
 * The third-order drivers, one decomposition per call, returning its seconds (host copy included):
 *
 *   libttx_ttdec6.so  <- ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:778 (func.h:16)
 *   libttx_trdec6.so  <- trdecomposition/third_order_GPU/opt_GPU/trdec_gpu_float_rsvd_lu_half.cu:734 (func.h:15)
 *
 *   float ttdec_gpu_float(float* h_tensor, int* h_n_array, int* h_r_array, double* eps,
 *                         float* singleError, float* singleComRatio);
 *   float trdec_gpu_float(float* h_tensor, int* h_n_array, int* h_r_array, double* eps,
 *                         float* singleError, float* singleComRatio);
 *
 * h_n_array[0..2] are N1..N3; h_r_array[0..3] the TT ranks (1, r1, r2, 1) or the TR ranks (r0, r1,
 * r2, r0).  The reference declares singleError and singleComRatio and never writes them; here, when
 * non-null, singleError[0] receives ||A - A_hat||_F / ||A||_F (computed after the timer stops) and
 * singleComRatio[0] the tensor's entries divided by the cores' entries.  eps is not read, as in
 * the reference.
 *
 * Engine settings (environment, read at each call): TTX_PRECISION = fp32 | bf16 (default) | fp16,
 * TTX_OVERSAMPLING (default 10), TTX_POWER_ITERS (default 2, the ceiling of the adaptive range
 * check), TTX_ADAPTIVE = 0 | 1 (default 1).
 */
#pragma once
#include <stdint.h>

float ttdec_half_lu_device_6(float* h_A, const int* ttRanks, const int* ttDemns, const int& calculateTimes,
                             const int& p, bool calError);
float ttdec_half_lu_host_6(float* h_A, int* ttRanks, const int* ttDemns, const int& calculateTimes, const int& p,
                           const bool calError);

float ttdec_gpu_float(float* h_tensor, int* h_n_array, int* h_r_array, double* eps, float* singleError,
                      float* singleComRatio);
float trdec_gpu_float(float* h_tensor, int* h_n_array, int* h_r_array, double* eps, float* singleError,
                      float* singleComRatio);

/* The harness's input generators (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:185-271 in
 * libttx_ttdec6.so, trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:231-329 in libttx_trdec6.so):
 * fill the caller's host h_A (six modes ttDemns, ranks ttRanks) with the reference's tensor, cores from
 * glibc's rand() after srand(1134) in the order G5, G6, G4, G3, G2, G1 (ttx_tt_generate_ref /
 * ttx_tr_generate_ref); p is the reference's GEMM slice count and is not needed. */
void generalTTTensor(float* h_A, const int* ttRanks, const int* ttDemns, const int& p);
void generalTRTensor(float* h_A, const int* ttRanks, const int* ttDemns, const int& p);

extern "C" {
/* Side channel for the cores the reference discards (ttdec6_singleGPU.cu:947-952): device pointers
 * of the last call's d cores (TT or TR layouts, ttx.h) and its ranks (d + 1 entries); valid until
 * the next shim call.  Returns the order d, 0 when no call has succeeded. */
int32_t ttx_last_cores(float* const** d_cores, int32_t* ranks);
/* relative reconstruction error of the last call (computed only with calError; -1 otherwise) */
double ttx_last_rel_error(void);
}
