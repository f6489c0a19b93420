// tdhf_probe.h -- FP64 FMA throughput probe: the rate the device sustains for
// a stream of independent FP64 fused multiply-adds (2 flops each), next to
// which the kernels' flop rates can be read.  Included at the end of tdhf.hip.
#pragma once

namespace tdhf {

// Every thread runs kProbeChains independent FMA chains for iters rounds;
// the chains' sum is stored so that nothing is dead code.
constexpr int kProbeChains = 8;

__global__ void __launch_bounds__(256) fma_probe_kernel(double* out, double b, double c,
                                                        int iters) {
  double x[kProbeChains];
#pragma unroll
  for (int k = 0; k < kProbeChains; ++k) x[k] = 1e-3 * (double)(threadIdx.x + k);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) x[k] = fma(x[k], b, c);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kProbeChains; ++k) s += x[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace tdhf

extern "C" int tdhf_fp64_fma_probe(int iters, double* tflops, double* ms) {
  if (iters < 1 || !tflops) return fail(nullptr, TDHF_EINVAL, "invalid probe arguments");
  CHKRET(post_device_check());
  constexpr int kBlocks = 1024, kThreads = 256;  // 16 waves per CU on 256 CUs
  double* d = nullptr;
  if (hipMalloc((void**)&d, (size_t)kBlocks * kThreads * sizeof(double)) != hipSuccess)
    return fail(nullptr, TDHF_ENOMEM, "hipMalloc failed");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  // a full-length warm-up launch (clocks ramp up under load), then the
  // fastest of three timed launches
  float t_ms = 0.0f;
  for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {
    e = hipEventRecord(e0, 0);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(tdhf::fma_probe_kernel, dim3(kBlocks), dim3(kThreads), 0, 0, d,
                         0.999999, 1e-7, iters);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (rep == 1 || (rep > 1 && t < t_ms)) t_ms = t;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(nullptr, TDHF_EDEVICE, "FMA probe: %s", hipGetErrorString(e));
  const double flops = 2.0 * tdhf::kProbeChains * (double)iters * kBlocks * kThreads;
  *tflops = flops / (t_ms * 1e-3) / 1e12;
  if (ms) *ms = t_ms;
  return 0;
}
