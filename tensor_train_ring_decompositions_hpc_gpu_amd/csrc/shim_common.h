// Shared body of the reference-name shims (ttx_shim.h).  Each shim library includes this file once
// and defines its two entry points with TR = false (TT) or true (TR).
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ttx.h"
#include "../../include/ttx_shim.h"

namespace ttx_shim {

struct State {
  ttx_handle h = nullptr;
  std::vector<float*> cores;
  int32_t ranks[TTX_MAX_ORDER + 1] = {0};
  int32_t order = 0;
  double err = -1.0;
};
static State g;

static int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return (v && *v) ? std::atoi(v) : dflt;
}

static void release_cores() {
  for (float* c : g.cores)
    if (c) (void)hipFree(c);
  g.cores.clear();
  g.order = 0;
}

// d-mode sweep over the host tensor h_A, timed from the host copy to the cores; err_mode 0: no
// error, 1: the relative error of every run, printed inside the timed loop (the six-mode drivers),
// 2: the error of the last run, computed after the timer stops (the third-order drivers)
// host_slices > 0 (N_1 <= TTX_MAX_ELL; TR also r_0 r_1 <= TTX_MAX_ELL): the host-streamed sweep, the
// tensor read from h_A in that many column slices; the device copy of the tensor is then made only
// for the error
static float run(const float* h_A, const int* ttRanks, const int* ttDemns, int times, int err_mode, bool tr,
                 int d = 6, int host_slices = 0) {
  if (!h_A || !ttRanks || !ttDemns || times <= 0 || d < 3 || d > TTX_MAX_ORDER) {
    std::fprintf(stderr, "ttx shim: bad arguments\n");
    return -1.0f;
  }
  if (!g.h && ttx_create(&g.h, nullptr) != TTX_OK) {
    std::fprintf(stderr, "ttx shim: ttx_create failed\n");
    return -1.0f;
  }
  ttx_params p;
  int64_t dims[TTX_MAX_ORDER];
  int32_t ranks[TTX_MAX_ORDER + 1];
  for (int k = 0; k < d; ++k) dims[k] = ttDemns[k];
  for (int k = 0; k <= d; ++k) ranks[k] = ttRanks[k];
  ttx_params_init(&p, d, dims, ranks);
  const char* prec = std::getenv("TTX_PRECISION");
  p.precision = (prec && !std::strcmp(prec, "fp32")) ? TTX_PREC_FP32
              : (prec && !std::strcmp(prec, "fp16")) ? TTX_PREC_FP16 : TTX_PREC_BF16;
  p.oversampling = env_int("TTX_OVERSAMPLING", 10);
  p.power_iters = env_int("TTX_POWER_ITERS", 2);
  if (env_int("TTX_ADAPTIVE", 1)) p.flags |= TTX_FLAG_ADAPTIVE;
  int64_t total = 1;
  for (int k = 0; k < d; ++k) total *= dims[k];

  release_cores();
  const bool host = host_slices > 0 && dims[0] <= TTX_MAX_ELL && (!tr || (int64_t)ranks[0] * ranks[1] <= TTX_MAX_ELL);
  const int64_t slice_cols = host ? (total / dims[0] + host_slices - 1) / host_slices : 0;
  float* d_A = nullptr;
  if ((!host || err_mode != 0) && hipMalloc(&d_A, (size_t)total * sizeof(float)) != hipSuccess) {
    std::fprintf(stderr, "ttx shim: out of device memory for the tensor\n");
    return -1.0f;
  }
  g.cores.assign(d, nullptr);
  for (int k = 0; k < d; ++k) {
    const int64_t sz = ttx_core_size(&p, tr ? 1 : 0, k);
    if (hipMalloc(&g.cores[k], (size_t)sz * sizeof(float)) != hipSuccess) {
      std::fprintf(stderr, "ttx shim: out of device memory for core %d\n", k);
      (void)hipFree(d_A);
      release_cores();
      return -1.0f;
    }
  }
  ttx_report rep;
  ttx_status st = TTX_OK;
  g.err = -1.0;
  (void)hipDeviceSynchronize();
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < times && st == TTX_OK; ++i) {
    if (host) {
      st = tr ? ttx_tr_svd_host(g.h, h_A, &p, slice_cols, g.cores.data(), nullptr, &rep)
              : ttx_tt_svd_host(g.h, h_A, &p, slice_cols, g.cores.data(), nullptr, &rep);
    } else {
      // the reference copies the host tensor inside its timed loop (ttdec6_singleGPU.cu:824)
      if (hipMemcpy(d_A, h_A, (size_t)total * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        st = TTX_ERR_HIP;
        break;
      }
      st = tr ? ttx_tr_svd(g.h, d_A, &p, g.cores.data(), nullptr, &rep)
              : ttx_tt_svd(g.h, d_A, &p, g.cores.data(), nullptr, &rep);
    }
    if (st == TTX_RESIDUAL_FALLBACK_USED) st = TTX_OK;
    if (st == TTX_OK && err_mode == 1 && host &&
        hipMemcpy(d_A, h_A, (size_t)total * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      st = TTX_ERR_HIP;
    if (st == TTX_OK && err_mode == 1) {
      double e = -1.0;
      st = tr ? ttx_tr_rel_error(g.h, d, dims, rep.ranks, g.cores.data(), d_A, &e)
              : ttx_tt_rel_error(g.h, d, dims, rep.ranks, g.cores.data(), d_A, &e);
      g.err = e;
      std::printf("\n*********************** relative error: %e ******************\n", e);
    }
  }
  (void)hipDeviceSynchronize();
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st == TTX_OK && err_mode == 2 && host &&
      hipMemcpy(d_A, h_A, (size_t)total * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    st = TTX_ERR_HIP;
  if (st == TTX_OK && err_mode == 2) {
    double e = -1.0;
    st = tr ? ttx_tr_rel_error(g.h, d, dims, rep.ranks, g.cores.data(), d_A, &e)
            : ttx_tt_rel_error(g.h, d, dims, rep.ranks, g.cores.data(), d_A, &e);
    g.err = e;
  }
  if (d_A) (void)hipFree(d_A);
  if (st != TTX_OK) {
    std::fprintf(stderr, "ttx shim: %s: %s\n", ttx_status_string(st), ttx_last_error(g.h));
    release_cores();
    return -1.0f;
  }
  g.order = d;
  for (int k = 0; k <= d; ++k) g.ranks[k] = rep.ranks[k];
  return (float)(sec / times);
}

// the third-order drivers: one decomposition of the N1 x N2 x N3 tensor with ranks h_r_array
// (TT: 1, r1, r2, 1; TR: r0, r1, r2, r0, first step of rank r0 r1); the relative error and the
// compression ratio (tensor entries / core entries, the ratio the reference's tntorch baseline
// prints, ttdecomposition/third_order_GPUs/python/tntorch.py:11) go to the output arguments the
// reference declares but leaves unset; eps is not read (neither does the reference read it)
static float run3(const float* h_tensor, const int* h_n_array, const int* h_r_array, float* singleError,
                  float* singleComRatio, bool tr) {
  const float sec = run(h_tensor, h_r_array, h_n_array, 1, singleError ? 2 : 0, tr, 3);
  if (sec < 0.0f) return sec;
  if (singleError) *singleError = (float)g.err;
  if (singleComRatio) {
    double entries = 1.0, core_entries = 0.0;
    for (int k = 0; k < 3; ++k) {
      entries *= (double)h_n_array[k];
      core_entries += (double)g.ranks[k] * h_n_array[k] * g.ranks[k + 1];
    }
    *singleComRatio = (float)(entries / core_entries);
  }
  return sec;
}

// the harness's input generators (generalTTTensor / generalTRTensor): the six-mode tensor of cores
// drawn from glibc's rand() after srand(1134) (the reference seeds inside the generator), contracted
// on the device, copied into the caller's host buffer h_A; p (the reference's GEMM slice count) is
// not needed
static void gen_ref(float* h_A, const int* ttRanks, const int* ttDemns, bool tr) {
  if (!h_A || !ttRanks || !ttDemns) {
    std::fprintf(stderr, "ttx shim: bad arguments\n");
    return;
  }
  if (!g.h && ttx_create(&g.h, nullptr) != TTX_OK) {
    std::fprintf(stderr, "ttx shim: ttx_create failed\n");
    return;
  }
  int64_t dims[6];
  int32_t ranks[7];
  int64_t total = 1;
  for (int k = 0; k < 6; ++k) total *= (dims[k] = ttDemns[k]);
  for (int k = 0; k <= 6; ++k) ranks[k] = ttRanks[k];
  float* d_A = nullptr;
  if (hipMalloc(&d_A, (size_t)total * sizeof(float)) != hipSuccess) {
    std::fprintf(stderr, "ttx shim: out of device memory for the tensor\n");
    return;
  }
  const ttx_status st = tr ? ttx_tr_generate_ref(g.h, 6, dims, ranks, 1134u, d_A, nullptr)
                           : ttx_tt_generate_ref(g.h, 6, dims, ranks, 1134u, d_A, nullptr);
  if (st != TTX_OK)
    std::fprintf(stderr, "ttx shim: %s: %s\n", ttx_status_string(st), ttx_last_error(g.h));
  else if (hipMemcpy(h_A, d_A, (size_t)total * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    std::fprintf(stderr, "ttx shim: copy of the generated tensor failed\n");
  (void)hipFree(d_A);
}

}  // namespace ttx_shim

extern "C" int32_t ttx_last_cores(float* const** d_cores, int32_t* ranks) {
  if (ttx_shim::g.order == 0) return 0;
  if (d_cores) *d_cores = ttx_shim::g.cores.data();
  if (ranks)
    for (int k = 0; k <= ttx_shim::g.order; ++k) ranks[k] = ttx_shim::g.ranks[k];
  return ttx_shim::g.order;
}

extern "C" double ttx_last_rel_error(void) { return ttx_shim::g.err; }
