// Byte-counter calibration workload (a measurement tool, not part of
// libbve): copies N MiB between two device buffers with a plain streaming
// kernel, so that rocprofv3's FETCH_SIZE and WRITE_SIZE of this known
// traffic give the factors that turn the counters of other kernels into
// bytes.
//     copy_probe MIB
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_probe_kernel(const double2 *__restrict__ in,
                                                         double2 *__restrict__ out, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        out[i] = in[i];
}

int main(int argc, char **argv)
{
    const long mib = argc > 1 ? std::atol(argv[1]) : 256;
    if (mib < 1 || mib > 16384) {
        std::fprintf(stderr, "copy_probe: MIB must be in [1, 16384]\n");
        return 2;
    }
    const size_t bytes = (size_t)mib << 20;
    void *in = nullptr, *out = nullptr;
    if (hipMalloc(&in, bytes) != hipSuccess || hipMalloc(&out, bytes) != hipSuccess) {
        std::fprintf(stderr, "copy_probe: allocation failed\n");
        return 1;
    }
    hipError_t e = hipMemset(in, 1, bytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(copy_probe_kernel, dim3(4096), dim3(256), 0, 0, (const double2 *)in,
                           (double2 *)out, bytes / sizeof(double2));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(in);
    (void)hipFree(out);
    if (e != hipSuccess) {
        std::fprintf(stderr, "copy_probe: %s\n", hipGetErrorString(e));
        return 1;
    }
    std::printf("{\"copy_probe_bytes\": %zu}\n", bytes);
    return 0;
}
