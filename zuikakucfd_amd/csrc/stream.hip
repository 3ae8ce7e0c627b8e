// stream.hip -- zk_stream_bandwidth: the streaming bandwidth of the device for
// the read/write mixes of a stage kernel, the practical ceiling its roofline
// fraction is read against.  A copy (1 read + 1 write) and an add (2 reads +
// 1 write) over arrays of a given size, 16-byte accesses, each timed with HIP
// events as the best of five runs after a first one.
#include <hip/hip_runtime.h>

#include <string>

#include "zk_internal.h"

namespace zk {
std::string &null_handle_error();
}

namespace {

// two 16-byte elements per thread, the second a grid's width after the first
__global__ __launch_bounds__(256) void stream_copy_kernel(const double2 *__restrict__ a, double2 *__restrict__ c,
                                                          size_t n2)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x, s = (size_t)gridDim.x * 256;
    if (i < n2) c[i] = a[i];
    if (i + s < n2) c[i + s] = a[i + s];
}

__global__ __launch_bounds__(256) void stream_add_kernel(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                                         double2 *__restrict__ c, size_t n2)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x, s = (size_t)gridDim.x * 256;
    if (i < n2) c[i] = double2{a[i].x + b[i].x, a[i].y + b[i].y};
    if (i + s < n2) c[i + s] = double2{a[i + s].x + b[i + s].x, a[i + s].y + b[i + s].y};
}

__global__ void fill_kernel(double2 *p, size_t n2, double v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x)
        p[i] = double2{v, v};
}

}  // namespace

extern "C" zk_status zk_stream_bandwidth(int device, size_t array_bytes, double *copy_gbs, double *add_gbs)
{
    std::string &err = zk::null_handle_error();
    if (!copy_gbs || !add_gbs || array_bytes < 4096) {
        err = "zk_stream_bandwidth: null output or arrays below 4 KiB";
        return ZK_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        err = "zk_stream_bandwidth: device ordinal out of range";
        return ZK_ERR_ARG;
    }
    const size_t n2 = array_bytes / 16;
    double2 *a = nullptr, *b = nullptr, *c = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMalloc(&a, n2 * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n2 * 16);
    if (e == hipSuccess) e = hipMalloc(&c, n2 * 16);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, st, a, n2, 1.0);
        hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, st, b, n2, 2.0);
        const unsigned grid = (unsigned)((n2 + 511) / 512);
        for (int kind = 0; kind < 2 && e == hipSuccess; ++kind) {
            float best = -1.0f;
            for (int rep = 0; rep < 6 && e == hipSuccess; ++rep) {
                (void)hipEventRecord(e0, st);
                if (kind == 0) hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, st, a, c, n2);
                else hipLaunchKernelGGL(stream_add_kernel, dim3(grid), dim3(256), 0, st, a, b, c, n2);
                (void)hipEventRecord(e1, st);
                e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (rep > 0 && (best < 0.0f || ms < best)) best = ms;
            }
            const double bytes = (kind == 0 ? 2.0 : 3.0) * (double)n2 * 16.0;
            (kind == 0 ? *copy_gbs : *add_gbs) = best > 0.0f ? bytes / (best * 1e-3) / 1e9 : 0.0;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (c) (void)hipFree(c);
    if (e != hipSuccess) {
        err = std::string("zk_stream_bandwidth: ") + hipGetErrorString(e);
        return ZK_ERR_HIP;
    }
    return ZK_OK;
}
