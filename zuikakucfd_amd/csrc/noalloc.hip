// noalloc.hip -- zk_weno_step_noalloc: WENO5 steps on caller-owned device
// buffers with no device allocation, for callers that cannot allocate (code
// under HIP graph capture, serving loops).
//
// The stage buffers, the ghost-fill map, the updated-cell flags, the ghost-ring
// list and the step scalars live in `scratch`, one file-scope device array per
// GPU, sized for grids of up to ZK_NOALLOC_MAX_CELLS cells.  Its device address
// is looked up once per device (hipGetSymbolAddress) and kept; kernels receive
// that address, never the array's host-side name.  The launches are the ones
// zk_step enqueues for a WENO handle, so results are bit-identical to it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <string>

#include "fv_map.h"
#include "zk_internal.h"

namespace zk {
// the message zk_last_error(NULL) returns (defined in zk_api.hip)
std::string &null_handle_error();
}

namespace {

constexpr size_t kMaxCells = ZK_NOALLOC_MAX_CELLS;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
// layout: step scalars | stage state A | stage state B | ghost map | updated flags | ring list
constexpr size_t kScalBytes = align_up(sizeof(zk::StepScalars));
constexpr size_t kStateBytes = align_up(4 * kMaxCells * sizeof(double));
constexpr size_t kMapBytes = align_up(kMaxCells * sizeof(int32_t));
constexpr size_t kUpdBytes = align_up(kMaxCells);
constexpr size_t kRingBytes = align_up(kMaxCells * sizeof(int32_t));
constexpr size_t kScratchBytes = kScalBytes + 2 * kStateBytes + kMapBytes + kUpdBytes + kRingBytes;
static_assert(kScratchBytes == ZK_NOALLOC_SCRATCH_BYTES, "ZK_NOALLOC_SCRATCH_BYTES in zk.h is out of date");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};

// Ghost map, updated flags and ring list of one grid, one thread per cell
// (fv_map.h restates zk_load's host-side construction per cell).
__global__ __launch_bounds__(256) void noalloc_maps_kernel(const int32_t *__restrict__ bct, int ny, int nx,
                                                           int32_t *__restrict__ src, uint8_t *__restrict__ upd,
                                                           int32_t *__restrict__ ring)
{
    const size_t np = (size_t)ny * nx;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(k / nx), j = (int)(k % nx);
        const int32_t e = zk::fv_src_entry(bct, ny, nx, i, j);
        src[k] = e;
        upd[k] = (e & zk::kSrcUpdated) ? 1 : 0;
        const int64_t t = zk::ring_slot(ny, nx, i, j);
        if (t >= 0) ring[t] = (int32_t)k;
    }
}

// The step scalars zk_step sets: dt, and done = 0 (the stage kernels run).
__global__ void noalloc_scalars_kernel(zk::StepScalars *S, double dt)
{
    if (threadIdx.x == 0) {
        S->dt = dt;
        S->t = 0.0;
        S->lambda = 0.0;
        S->t_end = 0.0;
        S->cfl = 0.0;
        S->done = 0;
        S->steps = 0;
        S->bad = 0;
        S->pad = 0;
    }
}

// The device address of `scratch` on `device` (current), looked up at the
// first call on that device.
unsigned char *scratch_on(int device, hipError_t *err)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess) return nullptr;
        g_ptr[device] = static_cast<unsigned char *>(p);
    }
    *err = hipSuccess;
    return g_ptr[device];
}

}  // namespace

extern "C" {

zk_status zk_weno_step_noalloc(const zk_config *cfg, int ny, int nx, double da, double gamma, double *Q,
                               const int32_t *bctype, double dt, int nsteps, void *stream)
{
    std::string &g_err = zk::null_handle_error();
    g_err.clear();
    if (!cfg || !Q || !bctype) {
        g_err = "zk_weno_step_noalloc: null argument";
        return ZK_ERR_ARG;
    }
    if (cfg->method != ZK_METHOD_WENO || (cfg->precision != ZK_PREC_FP64 && cfg->precision != ZK_PREC_FP32_BULK) ||
        (cfg->flux != ZK_FLUX_HLLC && cfg->flux != ZK_FLUX_HLL) || cfg->positivity != 0) {
        g_err = "zk_weno_step_noalloc: a WENO configuration (FP64 or FP32 bulk, HLLC or HLL) without the "
                "positivity option";
        return ZK_ERR_ARG;
    }
    if (ny < 7 || nx < 7 || !(da > 0) || !(gamma > 1.0) || !(dt >= 0.0) || nsteps < 0) {
        g_err = "zk_weno_step_noalloc: need ny, nx >= 7, da > 0, gamma > 1, dt >= 0 and nsteps >= 0";
        return ZK_ERR_ARG;
    }
    if ((size_t)ny * nx > kMaxCells) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "zk_weno_step_noalloc: %d x %d cells exceed the maximum of %zu (ZK_NOALLOC_MAX_CELLS)",
                      ny, nx, kMaxCells);
        g_err = msg;
        return ZK_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev || cfg->device >= kMaxDevices) {
        g_err = "zk_weno_step_noalloc: device ordinal out of range";
        return ZK_ERR_ARG;
    }
    int prev = 0;
    (void)hipGetDevice(&prev);
    hipError_t e = hipSetDevice(cfg->device);
    unsigned char *base = e == hipSuccess ? scratch_on(cfg->device, &e) : nullptr;
    if (!base) {
        g_err = std::string("zk_weno_step_noalloc: device scratch: ") + hipGetErrorString(e);
        (void)hipSetDevice(prev);
        return ZK_ERR_HIP;
    }
    // A handle on the stack that names the caller's buffers and the scratch
    // partitions: the stage launchers read only these fields.
    zk::Handle H;
    H.cfg = *cfg;
    H.nm = 1;
    H.g.ny = ny; H.g.nx = nx; H.g.np = (size_t)ny * nx; H.g.da = da; H.g.gamma = gamma;
    H.lay.ps = H.g.np;  // plane after plane
    H.lay.nx = (uint32_t)nx;
    H.lay.rs = 0;
    H.stream = static_cast<hipStream_t>(stream);
    H.loaded = true;
    H.scal = reinterpret_cast<zk::StepScalars *>(base);
    H.qa = reinterpret_cast<double *>(base + kScalBytes);
    H.qb = reinterpret_cast<double *>(base + kScalBytes + kStateBytes);
    H.src = reinterpret_cast<int32_t *>(base + kScalBytes + 2 * kStateBytes);
    H.upd = base + kScalBytes + 2 * kStateBytes + kMapBytes;
    H.ring = reinterpret_cast<int32_t *>(base + kScalBytes + 2 * kStateBytes + kMapBytes + kUpdBytes);
    H.nring = (int)zk::ring_count(ny, nx);
    H.q = Q;
    int blocks = (int)((H.g.np + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(noalloc_maps_kernel, dim3(blocks), dim3(256), 0, H.stream, bctype, ny, nx, H.src, H.upd,
                       H.ring);
    hipLaunchKernelGGL(noalloc_scalars_kernel, dim3(1), dim3(64), 0, H.stream, H.scal, dt);
    for (int s = 0; s < nsteps; ++s) {
        zk::fv_stage(H, H.q, H.q, H.qa, 0);
        zk::fv_stage(H, H.q, H.qa, H.qb, 1);
        zk::fv_stage(H, H.q, H.qb, H.q, 2);
    }
    e = hipGetLastError();
    (void)hipSetDevice(prev);
    if (e != hipSuccess) {
        g_err = std::string("zk_weno_step_noalloc: kernel launch: ") + hipGetErrorString(e);
        return ZK_ERR_HIP;
    }
    return ZK_OK;
}

}  // extern "C"
