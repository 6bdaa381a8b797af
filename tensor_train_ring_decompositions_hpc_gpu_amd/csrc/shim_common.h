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

static float run(const float* h_A, const int* ttRanks, const int* ttDemns, int times, bool calError, bool tr) {
  const int d = 6;
  if (!h_A || !ttRanks || !ttDemns || times <= 0) {
    std::fprintf(stderr, "ttx shim: bad arguments\n");
    return -1.0f;
  }
  if (!g.h && ttx_create(&g.h, nullptr) != TTX_OK) {
    std::fprintf(stderr, "ttx shim: ttx_create failed\n");
    return -1.0f;
  }
  ttx_params p;
  int64_t dims[6];
  int32_t ranks[7];
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
  float* d_A = nullptr;
  if (hipMalloc(&d_A, (size_t)total * sizeof(float)) != hipSuccess) {
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
    // the reference copies the host tensor inside its timed loop (ttdec6_singleGPU.cu:824)
    if (hipMemcpy(d_A, h_A, (size_t)total * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
      st = TTX_ERR_HIP;
      break;
    }
    st = tr ? ttx_tr_svd(g.h, d_A, &p, g.cores.data(), nullptr, &rep)
            : ttx_tt_svd(g.h, d_A, &p, g.cores.data(), nullptr, &rep);
    if (st == TTX_RESIDUAL_FALLBACK_USED) st = TTX_OK;
    if (st == TTX_OK && calError) {
      double e = -1.0;
      st = tr ? ttx_tr_rel_error(g.h, d, dims, rep.ranks, g.cores.data(), d_A, &e)
              : ttx_tt_rel_error(g.h, d, dims, rep.ranks, g.cores.data(), d_A, &e);
      g.err = e;
      std::printf("\n*********************** relative error: %e ******************\n", e);
    }
  }
  (void)hipDeviceSynchronize();
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  (void)hipFree(d_A);
  if (st != TTX_OK) {
    std::fprintf(stderr, "ttx shim: %s: %s\n", ttx_status_string(st), ttx_last_error(g.h));
    release_cores();
    return -1.0f;
  }
  g.order = d;
  for (int k = 0; k <= d; ++k) g.ranks[k] = rep.ranks[k];
  return (float)(sec / times);
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
