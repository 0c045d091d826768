// selftest.hip -- device self-test of the wave-level primitives that the
// mixed method's kernels are built on: the DPP reductions (wmin, wmax, wsum,
// wminf, wmaxf) and the wave prefix sum (wave_incl_scan), over one workgroup
// of 256 threads (four waves), right after lanes rejoin from a branch, for
// the host to check against plain sums.
#include "mixed_dev.h"

namespace shp {

__global__ __launch_bounds__(256) void k_wave_selftest(const int *in, int *out) {
  const int t = threadIdx.x;
  int v = in[t];
  // odd lanes take a branch of their own: the wave's EXEC is rewritten just
  // before the reductions, as where diverged lanes rejoin in the kernels
  if (v & 1) asm volatile("" : "+v"(v));
  int *o = out + 6 * t;
  o[0] = wave_incl_scan(v);
  o[1] = wmin(v);
  o[2] = wmax(v);
  o[3] = wsum(v);
  o[4] = __float_as_int(wminf((float)v));
  o[5] = __float_as_int(wmaxf((float)v));
}

}  // namespace shp

// in: 256 ints; out: 6 ints per thread (scan, min, max, sum, FP32 min bits,
// FP32 max bits), each over the thread's wave.  0 on success, -4 on a HIP
// error.
extern "C" int shape_hip_wave_selftest(const int *in, int *out) {
  int *d_in = nullptr, *d_out = nullptr;
  int rc = 0;
  if (hipMalloc((void **)&d_in, 256 * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&d_out, 6 * 256 * sizeof(int)) != hipSuccess)
    rc = -4;
  if (!rc && hipMemcpy(d_in, in, 256 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) rc = -4;
  if (!rc) {
    hipLaunchKernelGGL(shp::k_wave_selftest, dim3(1), dim3(256), 0, 0, d_in, d_out);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(out, d_out, 6 * 256 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      rc = -4;
  }
  if (d_in) (void)hipFree(d_in);
  if (d_out) (void)hipFree(d_out);
  return rc;
}
