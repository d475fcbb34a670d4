// Vector-FMA throughput probes for the units the descriptor kernels are bound
// by: FP32 FMA without packed instructions (as every kernel here is compiled,
// so FP32 runs at the FP64 vector rate) and FP64 FMA.  Each probe keeps eight
// independent accumulator chains per lane in registers, with no memory
// traffic but one store per lane, on a grid of eight waves per SIMD, timed
// with HIP events.  The rates sit beside the published peaks in the
// benchmark's record as measurements of this part under this clock.
#include "../../include/gapx.h"
#include "gapx_internal.h"

namespace gapx {

namespace {

constexpr int kChains = 8;

template <typename T>
__global__ __launch_bounds__(256) void k_probe_fma(int iters, T m, double *out) {
  T acc[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) acc[c] = (T)(threadIdx.x + c) * (T)1e-3;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) acc[c] = acc[c] * m + (T)1e-7;
  }
  T s = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = (double)s;
}

}  // namespace

}  // namespace gapx

using namespace gapx;

extern "C" int gapx_probe_fma_rates(double *tflops) {
  if (!tflops) return fail(GAPX_ERR_ARG, "gapx_probe_fma_rates: null argument");
  int cus = 0, dev = 0;
  GAPX_HIP(hipGetDevice(&dev));
  GAPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks_valu = 8 * cus;  // 256-thread blocks: 8 waves per SIMD
  double *out = nullptr;
  if (hipMalloc(&out, sizeof(double) * 256 * (size_t)blocks_valu) != hipSuccess)
    return fail(GAPX_ERR_NOMEM, "gapx_probe_fma_rates: hipMalloc failed");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GAPX_OK;
  auto timed = [&](auto launch, double flops, double *res) -> int {
    launch();  // warm-up
    GAPX_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < 5; ++r) launch();
    GAPX_HIP(hipEventRecord(e1, 0));
    GAPX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GAPX_HIP(hipEventElapsedTime(&ms, e0, e1));
    *res = 5.0 * flops / (ms * 1e-3) / 1e12;
    return GAPX_OK;
  };
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail(GAPX_ERR_HIP, "gapx_probe_fma_rates: hipEventCreate failed");
  } else {
    const int iv = 4000;
    const double waves_v = blocks_valu * 4.0;
    rc = timed([&] { k_probe_fma<float><<<blocks_valu, 256>>>(iv, 0.999f, out); }, waves_v * 64 * iv * kChains * 2.0,
               &tflops[0]);
    if (!rc)
      rc = timed([&] { k_probe_fma<double><<<blocks_valu, 256>>>(iv, 0.999, out); },
                 waves_v * 64 * iv * kChains * 2.0, &tflops[1]);
    if (!rc && hipGetLastError() != hipSuccess) rc = fail(GAPX_ERR_HIP, "gapx_probe_fma_rates: launch failed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return rc;
}
