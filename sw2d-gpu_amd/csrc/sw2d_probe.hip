// sw2d_probe.hip -- streaming reference kernels for the roofline of the step
// (measurement only; built as libsw2d_probe.so, not part of libsw2d.so).
//
// The step reads five arrays per cell (h, U, V, z, n) and writes three
// (h', U', V').  These kernels move the same bytes in the most favourable
// pattern, a grid-stride sweep over eight dense arrays, with a trivial
// combination so that no load is dead: what they reach on the device is the
// practical ceiling of a kernel with the step's bytes per cell.
#include <hip/hip_runtime.h>

namespace {

__global__ void k_stream5r3w(const float *__restrict__ a, const float *__restrict__ b,
                             const float *__restrict__ c, const float *__restrict__ d,
                             const float *__restrict__ e, float *__restrict__ x, float *__restrict__ y,
                             float *__restrict__ z, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) {
    const float bi = b[i], ci = c[i];
    x[i] = a[i] + 1e-7f * (d[i] + e[i]) + 1e-9f * (bi + ci);
    y[i] = bi * 0.999f;
    z[i] = ci * 0.999f;
  }
}

// the same with one float4 per thread and step (n4 = n / 4 vectors)
__global__ void k_stream5r3w_v4(const float4 *__restrict__ a, const float4 *__restrict__ b,
                                const float4 *__restrict__ c, const float4 *__restrict__ d,
                                const float4 *__restrict__ e, float4 *__restrict__ x,
                                float4 *__restrict__ y, float4 *__restrict__ z, long long n4) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += stride) {
    const float4 A = a[i], B = b[i], C = c[i], D = d[i], E = e[i];
    x[i] = make_float4(A.x + 1e-7f * (D.x + E.x) + 1e-9f * (B.x + C.x),
                       A.y + 1e-7f * (D.y + E.y) + 1e-9f * (B.y + C.y),
                       A.z + 1e-7f * (D.z + E.z) + 1e-9f * (B.z + C.z),
                       A.w + 1e-7f * (D.w + E.w) + 1e-9f * (B.w + C.w));
    y[i] = make_float4(B.x * 0.999f, B.y * 0.999f, B.z * 0.999f, B.w * 0.999f);
    z[i] = make_float4(C.x * 0.999f, C.y * 0.999f, C.z * 0.999f, C.w * 0.999f);
  }
}

}  // namespace

extern "C" {

// Times `reps` passes of the 5-read / 3-write sweep over arrays of n floats
// (eight device arrays the caller allocated, n a multiple of 4) on the null
// stream after 3 untimed passes; `vec4` selects the float4 form.  *ms_out is
// the mean milliseconds per pass.  Returns 0, or -1 on a bad argument or a
// HIP error.
int sw2d_probe_stream5r3w(float *const arrays[8], long long n, int blocks, int threads, int vec4, int reps,
                          double *ms_out) {
  if (!arrays || !ms_out || n <= 0 || n % 4 != 0 || blocks <= 0 || threads <= 0 || threads > 1024 || reps <= 0)
    return -1;
  for (int k = 0; k < 8; k++)
    if (!arrays[k]) return -1;
  auto launch = [&]() {
    if (vec4)
      hipLaunchKernelGGL(k_stream5r3w_v4, dim3(blocks), dim3(threads), 0, 0, (const float4 *)arrays[0],
                         (const float4 *)arrays[1], (const float4 *)arrays[2], (const float4 *)arrays[3],
                         (const float4 *)arrays[4], (float4 *)arrays[5], (float4 *)arrays[6], (float4 *)arrays[7],
                         n / 4);
    else
      hipLaunchKernelGGL(k_stream5r3w, dim3(blocks), dim3(threads), 0, 0, arrays[0], arrays[1], arrays[2],
                         arrays[3], arrays[4], arrays[5], arrays[6], arrays[7], n);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
  for (int k = 0; k < 3; k++) launch();
  float ms = 0.0f;
  bool ok = hipEventRecord(e0, 0) == hipSuccess;
  for (int k = 0; k < reps && ok; k++) launch();
  ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, 0) == hipSuccess &&
       hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!ok) return -1;
  *ms_out = ms / reps;
  return 0;
}

}  // extern "C"
