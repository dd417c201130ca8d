// Achievable HBM bandwidth on this GPU for the three access mixes the solver's
// kernels sit between: read-only, four reads per write, and a 1:1 copy.
// Each kernel walks float4 arrays of `mib` MiB with a grid-stride loop; the
// grid size is swept and the best rate per mix is reported (bytes moved /
// time over 20 back-to-back launches between HIP events).
//
//   fvk_bw_probe [MiB per array, default 512]
//
// Prints one JSON object: {"mib": M, "read": GB/s, "r4w1": GB/s, "copy": GB/s,
// "per_grid": {...}}.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

// the sum is stored only when it equals a value it never takes, so the
// loads stay live without a per-thread store
__global__ void k_read(const f4* __restrict__ a, long n, float* out) {
  f4 s = {0, 0, 0, 0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    s += a[i];
  if (s.x + s.y + s.z + s.w == -1.0f) out[0] = 1.0f;
}

__global__ void k_r4w1(const f4* __restrict__ a, const f4* __restrict__ b, const f4* __restrict__ c,
                       const f4* __restrict__ d, f4* __restrict__ o, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    o[i] = a[i] + b[i] + c[i] + d[i];
}

__global__ void k_copy(const f4* __restrict__ a, f4* __restrict__ o, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    o[i] = a[i];
}

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  fprintf(stderr, "fvk_bw_probe: %s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  const long mib = argc > 1 ? std::atol(argv[1]) : 512;
  if (mib < 1 || mib > 16384) {
    fprintf(stderr, "fvk_bw_probe: MiB per array must be in [1, 16384]\n");
    return 2;
  }
  const long bytes = mib << 20;
  const long n = bytes / (long)sizeof(f4);
  std::vector<f4*> p(6, nullptr);
  for (auto& x : p) {
    if (!ok(hipMalloc(&x, bytes), "hipMalloc")) return 1;
    if (!ok(hipMemset(x, 0, bytes), "hipMemset")) return 1;
  }
  float* out = nullptr;
  if (!ok(hipMalloc(&out, sizeof(float)), "hipMalloc")) return 1;
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate")) return 1;
  const char* names[3] = {"read", "r4w1", "copy"};
  const double moved[3] = {1.0, 5.0, 2.0};   // arrays of `bytes` touched per launch
  double best[3] = {0, 0, 0};
  std::string per_grid;
  const int reps = 20;
  for (int grid : {1024, 2048, 4096, 8192, 16384}) {
    for (int kind = 0; kind < 3; ++kind) {
      // launch r of a timed series: reads rotate over p[0..3], writes go to p[4] / p[5]
      auto launch = [&](int r) {
        if (kind == 0) hipLaunchKernelGGL(k_read, dim3(grid), dim3(256), 0, 0, p[r % 4], n, out);
        else if (kind == 1)
          hipLaunchKernelGGL(k_r4w1, dim3(grid), dim3(256), 0, 0, p[0], p[1], p[2], p[3], p[4], n);
        else hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, p[r % 4], p[5], n);
      };
      for (int r = 0; r < 3; ++r) launch(r);
      if (!ok(hipDeviceSynchronize(), "warm-up")) return 1;
      if (!ok(hipEventRecord(e0, 0), "hipEventRecord")) return 1;
      for (int r = 0; r < reps; ++r) launch(r);
      if (!ok(hipEventRecord(e1, 0), "hipEventRecord") || !ok(hipEventSynchronize(e1), "sync")) return 1;
      float ms = 0;
      if (!ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) return 1;
      const double gbs = moved[kind] * bytes * reps / (ms * 1e-3) / 1e9;
      if (gbs > best[kind]) best[kind] = gbs;
      char buf[96];
      snprintf(buf, sizeof buf, "%s\"%s_%d\": %.1f", per_grid.empty() ? "" : ", ", names[kind], grid, gbs);
      per_grid += buf;
    }
  }
  printf("{\"mib\": %ld, \"read\": %.1f, \"r4w1\": %.1f, \"copy\": %.1f, \"per_grid\": {%s}}\n", mib,
         best[0], best[1], best[2], per_grid.c_str());
  for (auto x : p) (void)hipFree(x);
  (void)hipFree(out);
  return 0;
}
