// zk_api.hip -- the C ABI of libzk.so (declared in include/zk.h).
//
// Host side of the drop-in boundary: owns device memory and one HIP stream
// per handle, builds the ghost-fill source map once at load, and enqueues the
// stage kernels.  The time loop stays on the device; the host only reads back
// what the caller asks for.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "zk_internal.h"

namespace zk {
hipError_t dg_upload_tables(int device);
zk_status mlp_features_of(Handle &H, zk_mlp *M, const int32_t *cells, int n, uint16_t *X, float *scores);
}

using zk::Handle;

namespace {

thread_local std::string g_create_err;

}  // namespace

namespace zk {
// The message zk_last_error(NULL) returns: that of the calling thread's last
// failed call made without a handle (zk_create, zk_weno_step_noalloc).
std::string &null_handle_error() { return g_create_err; }
}  // namespace zk

namespace {

#define ZK_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t _e = (call);                                                           \
        if (_e != hipSuccess) {                                                           \
            H->err = std::string(#call) + ": " + hipGetErrorString(_e);                   \
            return ZK_ERR_HIP;                                                            \
        }                                                                                 \
    } while (0)

__global__ void set_dt_kernel(zk::StepScalars *S, double dt)
{
    if (threadIdx.x == 0) {
        S->dt = dt;
        S->done = 0;
    }
}

__global__ void advance_time_kernel(zk::StepScalars *S)
{
    if (threadIdx.x == 0 && !S->done) {
        S->t += S->dt;
        S->steps += 1;
    }
}

// dt = cfl * da / lambda, clipped so that t + dt <= t_end; marks done once t
// has reached t_end so that the stage kernels of later steps become no-ops.
// A wave-speed maximum that is not finite (a broken state) stops the stepping
// the same way and sets bad.
__global__ void cfl_dt_kernel(zk::StepScalars *S, const unsigned long long *lam_bits, double da)
{
    if (threadIdx.x != 0) return;
    const double lam = __longlong_as_double((long long)*lam_bits);
    S->lambda = lam;
    if (!(lam <= 1.79769313486231570815e308)) {
        S->bad = 1;
        S->done = 1;
        S->dt = 0.0;
        return;
    }
    if (S->t >= S->t_end) {
        S->done = 1;
        S->dt = 0.0;
        return;
    }
    double dt = S->cfl * da / lam;
    if (S->t + dt > S->t_end) dt = S->t_end - S->t;
    S->dt = dt;
}

// Non-finite check of a state's means (mode 0 of the four variables, every
// cell): records the smallest cell*4 + variable found and the step count
// `step` in health[0], health[1] (atomic minima; the record starts at all
// ones).  Reads 32 bytes per cell, a fraction of one stage's traffic.
__global__ __launch_bounds__(256) void health_kernel(const double *__restrict__ Q, size_t np, zk::StLay L, int nm,
                                                     unsigned long long *health, unsigned long long step)
{
    unsigned long long key = ~0ull;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (size_t)gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double x = Q[L.at(c * nm, k)];
            if (!(fabs(x) <= 1.79769313486231570815e308)) {
                const unsigned long long kk = (unsigned long long)k * 4 + c;
                if (kk < key) key = kk;
            }
        }
    if (key != ~0ull) {
        atomicMin(&health[0], key);
        atomicMin(&health[1], step);
    }
}

// Copies between a state array in layout L and plane-major planes of np cells:
// plane i of the plane-major side is state plane i * pstep (blockIdx.y = i).
// to_state: pm -> state, else state -> pm.
__global__ __launch_bounds__(256) void relayout_kernel(double *__restrict__ state, double *__restrict__ pm, size_t np,
                                                       zk::StLay L, int pstep, bool to_state)
{
    const int i = blockIdx.y;
    double *plane = pm + (size_t)i * np;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (size_t)gridDim.x * blockDim.x) {
        const size_t s = L.at(i * pstep, k);
        if (to_state) state[s] = plane[k];
        else plane[k] = state[s];
    }
}

// Plane-major planes 0, pstep, 2 pstep, ... (n of them) of the state q to or
// from pm, on the handle's stream.
void relayout(Handle *H, double *pm, int n, int pstep, bool to_state)
{
    unsigned bx = (unsigned)((H->g.np + 255) / 256);
    if (bx > 1024) bx = 1024;
    hipLaunchKernelGGL(relayout_kernel, dim3(bx, (unsigned)n), dim3(256), 0, H->stream, H->q, pm, H->g.np, H->lay,
                       pstep, to_state);
}

void enqueue_health(Handle *H)
{
    int blocks = (int)((H->g.np + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(health_kernel, dim3(blocks), dim3(256), 0, H->stream, H->q, H->g.np, H->lay, H->nm, H->health,
                       (unsigned long long)H->steps_enqueued);
}

// Reads the non-finite record after the handle's stream has drained.  A
// non-finite state makes this and every later stepping call fail with
// ZK_ERR_NUMERIC (until zk_load or zk_set_state replaces the state).
zk_status check_health(Handle *H)
{
    unsigned long long v[2] = {~0ull, ~0ull};
    ZK_HIP(hipMemcpyAsync(v, H->health, sizeof(v), hipMemcpyDeviceToHost, H->stream));
    ZK_HIP(hipStreamSynchronize(H->stream));
    if (v[0] == ~0ull) return ZK_OK;
    const unsigned long long cell = v[0] / 4;
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "non-finite state: U%d (conserved variable %d) of cell (i=%llu, j=%llu) is not finite, found by "
                  "the check after step %llu",
                  (int)(v[0] % 4), (int)(v[0] % 4), cell / (unsigned long long)H->g.nx,
                  cell % (unsigned long long)H->g.nx, v[1]);
    H->err = msg;
    H->numeric_failed = true;
    return ZK_ERR_NUMERIC;
}

zk_status check_args(Handle *H)
{
    if (!H) return ZK_ERR_ARG;
    if (!H->loaded) {
        H->err = "state not loaded (call zk_load first)";
        return ZK_ERR_ARG;
    }
    return ZK_OK;
}

// Drops the captured step (its kernel arguments name the current buffers and
// tuning, so it is rebuilt after any change to them).
void drop_graph(Handle *H)
{
    if (H->step_exec) (void)hipGraphExecDestroy(H->step_exec);
    if (H->step_graph) (void)hipGraphDestroy(H->step_graph);
    H->step_exec = nullptr;
    H->step_graph = nullptr;
    if (H->cfl_exec) (void)hipGraphExecDestroy(H->cfl_exec);
    if (H->cfl_graph) (void)hipGraphDestroy(H->cfl_graph);
    H->cfl_exec = nullptr;
    H->cfl_graph = nullptr;
}

void free_all(Handle *H)
{
    drop_graph(H);
    void *ptrs[] = {H->q, H->qa, H->qb, H->src, H->srcy, H->upd, H->ring, H->gx, H->gy, H->means, H->flags,
                    H->scal, H->lam_bits, H->health, H->bct, H->stats, H->score, H->cand, H->bad, H->any, H->blist,
                    H->zsok, H->tlist, H->dcount, H->dcells};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    H->q = H->qa = H->qb = nullptr;
    H->src = nullptr; H->srcy = nullptr; H->upd = nullptr; H->ring = nullptr;
    H->gx = H->gy = H->means = nullptr; H->flags = nullptr;
    H->scal = nullptr; H->lam_bits = nullptr; H->health = nullptr; H->bct = nullptr;
    H->stats = nullptr; H->score = nullptr; H->cand = nullptr; H->bad = nullptr; H->any = nullptr; H->blist = nullptr; H->zsok = nullptr; H->tlist = nullptr;
    H->dcount = nullptr; H->dcells = nullptr;
    H->loaded = false;
}

// Ghost-fill source map: replay the row pass then the column pass of the
// reference fill (src_WENO/fluid.cpp:338-370, src_DG/fluid.cpp:581-635) on
// indices, so the last writer of every ghost is recorded; then periodic (-2)
// ghosts take their periodic image inside [3, n-3).
// Mirror slip walls of the finite-volume option (no reference counterpart),
// one pass: pass 0 along columns (wall normal y, negate rho*v), pass 1 along
// rows (negate rho*u).  A bctype-1 cell whose nearest updated cell along the
// pass is k <= 3 away (ties to the larger index) takes the cell mirrored
// across the wall between them, 2k-1 away, or that nearest updated cell where
// the mirrored one is not updated.
void slip_mirror_pass(const std::vector<int32_t> &bct, int ny, int nx, int pass, std::vector<int32_t> &src)
{
    const int n = pass == 0 ? ny : nx;
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            if (bct[(size_t)i * nx + j] != 1) continue;
            auto cell = [&](int a) { return pass == 0 ? (size_t)a * nx + j : (size_t)i * nx + a; };
            auto updated = [&](int a) {
                if (a < 3 || a >= n - 3) return false;
                if (pass == 0 ? (j < 3 || j >= nx - 3) : (i < 3 || i >= ny - 3)) return false;
                return bct[cell(a)] == 0;
            };
            const int g = pass == 0 ? i : j;
            int k = 0, s = 0;
            for (int d = 1; d <= 3 && !k; ++d) {
                if (g + d < n && updated(g + d)) { k = d; s = 1; }
                else if (g - d >= 0 && updated(g - d)) { k = d; s = -1; }
            }
            if (!k) continue;
            int m = g + s * (2 * k - 1);
            if (!(m >= 0 && m < n && updated(m))) m = g + s * k;
            src[(size_t)i * nx + j] = (int32_t)((cell(m) << zk::kSrcShift) | (pass == 0 ? 2 : 1));
        }
}

// slip: 0 leaves bctype-1 cells alone (the reference's finite-volume fill),
// 1 copies with the normal momentum negated (its DG fill), 2 mirror slip walls
void build_src_map(const std::vector<int32_t> &bct, int ny, int nx, int slip, std::vector<int32_t> &src)
{
    src.resize((size_t)ny * nx);
    for (size_t k = 0; k < src.size(); ++k) src[k] = (int32_t)(k << zk::kSrcShift);
    const int OFF[6] = {-3, -2, -1, 1, 2, 3};
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 3; i < ny - 3; ++i)
            for (int j = 3; j < nx - 3; ++j) {
                const size_t k = (size_t)i * nx + j;
                if (bct[k] != 0) continue;
                for (int d = 0; d < 6; ++d) {
                    int ni = i, nj = j;
                    if (pass == 0) {
                        ni = i + OFF[d];
                        if ((unsigned)ni >= (unsigned)ny) continue;
                    } else {
                        nj = j + OFF[d];
                        if ((unsigned)nj >= (unsigned)nx) continue;
                    }
                    const size_t dst = (size_t)ni * nx + nj;
                    const int bt = bct[dst];
                    if (bt == -1) src[dst] = (int32_t)(k << zk::kSrcShift);
                    else if (bt == 1 && slip == 1) src[dst] = (int32_t)((k << zk::kSrcShift) | (pass == 0 ? 2 : 1));
                }
            }
    if (slip == 2) {
        slip_mirror_pass(bct, ny, nx, 0, src);
        slip_mirror_pass(bct, ny, nx, 1, src);
    }
    auto per = [](int i, int n) { return i < 3 ? i + (n - 6) : (i >= n - 3 ? i - (n - 6) : i); };
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            const size_t k = (size_t)i * nx + j;
            if (bct[k] != -2) continue;
            src[k] = (int32_t)(((size_t)per(i, ny) * nx + per(j, nx)) << zk::kSrcShift);
        }
    for (int i = 3; i < ny - 3; ++i)
        for (int j = 3; j < nx - 3; ++j) {
            const size_t k = (size_t)i * nx + j;
            if (bct[k] == 0) src[k] |= zk::kSrcUpdated;
        }
}

// The y-face map of the slip-wall option: the ghost map with the column
// pass of the slip walls applied again, so that a wall cell near updated
// cells along both its row and its column gives the y faces its column mirror.
void build_srcy_map(const std::vector<int32_t> &bct, int ny, int nx, const std::vector<int32_t> &src,
                    std::vector<int32_t> &srcy)
{
    srcy = src;
    slip_mirror_pass(bct, ny, nx, 0, srcy);
}

}  // namespace

namespace zk {
void build_src_map_public(const std::vector<int32_t> &bct, int ny, int nx, bool slip, std::vector<int32_t> &src)
{
    build_src_map(bct, ny, nx, slip, src);
}
}  // namespace zk

namespace {

zk_status enqueue_step(Handle *H)
{
    if (H->cfg.method == ZK_METHOD_WENO) {
        zk::fv_stage(*H, H->q, H->q, H->qa, 0);
        zk::fv_stage(*H, H->q, H->qa, H->qb, 1);
        zk::fv_stage(*H, H->q, H->qb, H->q, 2);
    } else {
        zk::dg_stage(*H, H->q, H->q, H->qa, 0);
        zk::dg_stage(*H, H->q, H->qa, H->qb, 1);
        zk::dg_stage(*H, H->q, H->qb, H->q, 2);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        H->err = std::string("kernel launch: ") + hipGetErrorString(e);
        return ZK_ERR_HIP;
    }
    return ZK_OK;
}

// Enqueues nsteps steps with the dt in device memory: one captured step
// replayed, or the launches themselves.
// One step of CFL stepping: the wave-speed maximum, dt = cfl * da / lambda
// (clipped to t_end; done once there), the three stages, and the time
// advance, all from device memory.
zk_status enqueue_cfl_step(Handle *H)
{
    if (H->cfg.method == ZK_METHOD_WENO) zk::fv_max_speed(*H, H->q);
    else zk::dg_max_speed(*H, H->q);
    hipLaunchKernelGGL(cfl_dt_kernel, dim3(1), dim3(64), 0, H->stream, H->scal, H->lam_bits, H->g.da);
    const zk_status st = enqueue_step(H);
    if (st) return st;
    hipLaunchKernelGGL(advance_time_kernel, dim3(1), dim3(64), 0, H->stream, H->scal);
    return ZK_OK;
}

zk_status enqueue_steps(Handle *H, int nsteps)
{
    zk_status st = ZK_OK;
    // One step (dt is read from device memory) is captured once and replayed,
    // which removes the per-launch gaps (a DG step is about 15 launches).  Not
    // for the learned limiter or harvesting, which salt their hashes per stage.
    const bool salted = H->cfg.method == ZK_METHOD_DG && (H->cfg.limiter == ZK_LIMITER_LEARNED || H->harvest);
    if (!salted && H->use_graph && !H->prof && nsteps >= 2) {
        if (!H->step_exec) {
            // the first step runs as plain launches: every kernel of the step
            // is then loaded and launched once before any capture records it
            st = enqueue_step(H);
            if (st) return st;
            --nsteps;
            ZK_HIP(hipStreamBeginCapture(H->stream, hipStreamCaptureModeThreadLocal));
            st = enqueue_step(H);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(H->stream, &g);
            if (st) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            ZK_HIP(e);
            H->step_graph = g;
            ZK_HIP(hipGraphInstantiate(&H->step_exec, g, nullptr, nullptr, 0));
        }
        for (int s = 0; s < nsteps; ++s) ZK_HIP(hipGraphLaunch(H->step_exec, H->stream));
        return ZK_OK;
    }
    for (int s = 0; s < nsteps; ++s) {
        st = enqueue_step(H);
        if (st) return st;
    }
    return ZK_OK;
}

}  // namespace

extern "C" {

const char *zk_last_error(const zk_handle *h)
{
    if (!h) return g_create_err.c_str();
    return reinterpret_cast<const Handle *>(h)->err.c_str();
}

zk_status zk_create(const zk_config *cfg, zk_handle **out)
{
    g_create_err.clear();
    if (!cfg || !out) {
        g_create_err = "null argument";
        return ZK_ERR_ARG;
    }
    if (cfg->method != ZK_METHOD_WENO && cfg->method != ZK_METHOD_DG) {
        g_create_err = "unknown method";
        return ZK_ERR_ARG;
    }
    if (cfg->method == ZK_METHOD_DG && (cfg->P < 1 || cfg->P > 3)) {
        g_create_err = "DG degree P must be 1, 2 or 3";
        return ZK_ERR_ARG;
    }
    if (cfg->flux != ZK_FLUX_HLLC && cfg->flux != ZK_FLUX_HLL) {
        g_create_err = "unknown flux";
        return ZK_ERR_ARG;
    }
    if (cfg->method == ZK_METHOD_DG && (cfg->limiter < 0 || cfg->limiter > ZK_LIMITER_TVB_LINEAR)) {
        g_create_err = "unknown limiter";
        return ZK_ERR_ARG;
    }
    if (cfg->method == ZK_METHOD_DG && cfg->limiter == ZK_LIMITER_LEARNED && cfg->P != 2) {
        g_create_err = "the learned troubled-cell indicator is defined for DG P=2";
        return ZK_ERR_ARG;
    }
    if (cfg->precision != ZK_PREC_FP64 && cfg->precision != ZK_PREC_FP32_BULK) {
        g_create_err = "unknown precision";
        return ZK_ERR_ARG;
    }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        g_create_err = std::string("no HIP device: ") + hipGetErrorString(e);
        return ZK_ERR_HIP;
    }
    if (cfg->device < 0 || cfg->device >= ndev) {
        g_create_err = "device ordinal out of range";
        return ZK_ERR_ARG;
    }
    Handle *H = new Handle();
    H->cfg = *cfg;
    if (cfg->method == ZK_METHOD_DG && cfg->limiter == ZK_LIMITER_TVB_LINEAR) {
        // the TVB path with the linear-mode test alone (the face-average test off)
        H->cfg.limiter = ZK_LIMITER_TVB;
        H->tvb_linear = true;
    }
    H->nm = cfg->method == ZK_METHOD_DG ? (cfg->P + 1) * (cfg->P + 1) : 1;
    // measurement overrides: ZK_NO_MFMA=1 (VALU cell kernel), ZK_DMA=0..3 (memory pipeline of
    // the fused DG kernel's workgroup form), ZK_MF4=0 (its 16x16x4 products instead of the
    // 4x4x4 blocks), ZK_TVB_LANE=0 (the 16-cell TVB scan instead of one lane per cell),
    // ZK_TUNE=<dg_variant> (a value zk_set_tuning accepts; anything else is ignored),
    // ZK_FUSED_GENERIC=1 (the fused kernel form used for state arrays of 4 GiB and more),
    // ZK_FUSED_AL=0 (the fused kernel's 62-column runs at every stage), ZK_DG_PLANES=1 (DG
    // state planes one after the other instead of the planes of each row side by side)
    if (const char *v = std::getenv("ZK_NO_MFMA")) H->no_mfma = v[0] == '1';
    if (const char *v = std::getenv("ZK_FUSED_GENERIC")) H->fused_generic = v[0] == '1';
    if (const char *v = std::getenv("ZK_MF4")) H->mf4 = v[0] != '0';
    if (const char *v = std::getenv("ZK_TVB_LANE")) H->tvb_lane = v[0] != '0';
    if (const char *v = std::getenv("ZK_FUSED_AL")) H->fused_al = v[0] != '0';
    if (const char *v = std::getenv("ZK_DG_PLANES")) H->dg_planes = v[0] == '1';
    if (const char *v = std::getenv("ZK_DMA"))
        if (v[0] >= '0' && v[0] <= '3' && v[1] == 0) H->dma = v[0] - '0';
    if (const char *v = std::getenv("ZK_TUNE")) {
        const int t = std::atoi(v);
        if ((t >= 2 && t <= 4) || (t >= 14 && t <= 74) || t == 100 || (t >= 104 && t <= 164)) H->tune = t;
    }
    e = hipSetDevice(cfg->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&H->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_create_err = std::string("stream creation: ") + hipGetErrorString(e);
        delete H;
        return ZK_ERR_HIP;
    }
    if (cfg->method == ZK_METHOD_DG) {
        e = zk::dg_upload_tables(cfg->device);
        if (e != hipSuccess) {
            g_create_err = std::string("DG table upload: ") + hipGetErrorString(e);
            (void)hipStreamDestroy(H->stream);
            delete H;
            return ZK_ERR_HIP;
        }
    }
    *out = reinterpret_cast<zk_handle *>(H);
    return ZK_OK;
}

zk_status zk_destroy(zk_handle *h)
{
    if (!h) return ZK_OK;
    Handle *H = reinterpret_cast<Handle *>(h);
    (void)hipSetDevice(H->cfg.device);
    if (H->stream) (void)hipStreamSynchronize(H->stream);
    free_all(H);
    for (auto &r : H->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : H->pool) (void)hipEventDestroy(e);
    if (H->stream) (void)hipStreamDestroy(H->stream);
    delete H;
    return ZK_OK;
}

zk_status zk_load(zk_handle *h, int ny, int nx, double da, double gamma, const double *rho, const double *u,
                  const double *v, const double *p, const int32_t *bctype)
{
    if (!h) return ZK_ERR_ARG;
    Handle *H = reinterpret_cast<Handle *>(h);
    if (ny < 7 || nx < 7 || !(da > 0) || !(gamma > 1.0) || !rho || !u || !v || !p || !bctype) {
        H->err = "zk_load: need ny, nx >= 7, da > 0, gamma > 1 and non-null arrays";
        return ZK_ERR_ARG;
    }
    if ((size_t)ny * nx >= ((size_t)1 << (31 - zk::kSrcShift))) {
        H->err = "zk_load: grid too large for 32-bit source indices";
        return ZK_ERR_ARG;
    }
    ZK_HIP(hipSetDevice(H->cfg.device));
    if (H->stream) ZK_HIP(hipStreamSynchronize(H->stream));
    free_all(H);
    H->g.ny = ny; H->g.nx = nx; H->g.np = (size_t)ny * nx; H->g.da = da; H->g.gamma = gamma;
    const size_t np = H->g.np, nst = 4 * (size_t)H->nm * np;
    // DG: the 4 nm planes of each grid row side by side ([row][plane][column]);
    // WENO (and DG under ZK_DG_PLANES=1): plane after plane
    zk::StLay &L = H->lay;
    const bool rows = H->cfg.method == ZK_METHOD_DG && !H->dg_planes;
    L.ps = rows ? (size_t)nx : np;
    L.nx = (uint32_t)nx;
    L.rs = rows ? (uint32_t)((4 * H->nm - 1) * nx) : 0u;

    // conserved variables, src_WENO/fluid.cpp:86-93 (DG modes >= 1 are zero)
    std::vector<double> Q(nst, 0.0);
    for (size_t k = 0; k < np; ++k) {
        const double r = rho[k], ux = u[k], vy = v[k], pr = p[k];
        Q[L.at(0, k)] = r;
        Q[L.at(H->nm, k)] = r * ux;
        Q[L.at(2 * H->nm, k)] = r * vy;
        Q[L.at(3 * H->nm, k)] = pr / (gamma - 1.0) + 0.5 * r * (ux * ux + vy * vy);
    }
    H->h_bct.assign(bctype, bctype + np);
    std::vector<int32_t> src;
    build_src_map(H->h_bct, ny, nx, H->cfg.method == ZK_METHOD_DG ? 1 : (H->fv_slip ? 2 : 0), src);
    std::vector<uint8_t> upd(np, 0);
    std::vector<int32_t> ring;
    int64_t nupd = 0;
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            const size_t k = (size_t)i * nx + j;
            const bool in = i >= 3 && i < ny - 3 && j >= 3 && j < nx - 3;
            if (in && bctype[k] == 0) {
                upd[k] = 1;
                ++nupd;
            }
            if (!in) ring.push_back((int32_t)k);
        }
    H->n_updated = nupd;
    H->nring = (int)ring.size();
    // whether a neighbour of an updated cell takes its values from a ring cell
    // (a fixed cell, or a ghost filled from one): then the DG limiter reads
    // means the ring combination writes, and the two keep their order
    H->ring_read = false;
    auto in_ring = [&](size_t c) {
        const size_t i = c / nx, j = c % nx;
        return !(i >= 3 && i < (size_t)ny - 3 && j >= 3 && j < (size_t)nx - 3);
    };
    for (size_t k = 0; k < np && !H->ring_read; ++k) {
        if (!upd[k]) continue;
        const size_t nb[4] = {k - 1, k + 1, k - nx, k + nx};
        for (size_t n : nb)
            if (in_ring((size_t)(src[n] >> zk::kSrcShift))) H->ring_read = true;
    }

    ZK_HIP(hipMalloc(&H->q, nst * sizeof(double)));
    ZK_HIP(hipMalloc(&H->qa, nst * sizeof(double)));
    ZK_HIP(hipMalloc(&H->qb, nst * sizeof(double)));
    ZK_HIP(hipMalloc(&H->src, np * sizeof(int32_t)));
    if (H->cfg.method == ZK_METHOD_WENO) ZK_HIP(hipMalloc(&H->srcy, np * sizeof(int32_t)));
    ZK_HIP(hipMalloc(&H->upd, np));
    ZK_HIP(hipMalloc(&H->bct, np * sizeof(int32_t)));
    ZK_HIP(hipMalloc(&H->ring, ring.size() * sizeof(int32_t)));
    ZK_HIP(hipMalloc(&H->flags, np));
    ZK_HIP(hipMalloc(&H->scal, sizeof(zk::StepScalars)));
    ZK_HIP(hipMalloc(&H->lam_bits, sizeof(unsigned long long)));
    ZK_HIP(hipMalloc(&H->health, 2 * sizeof(unsigned long long)));
    ZK_HIP(hipMemset(H->health, 0xff, 2 * sizeof(unsigned long long)));
    H->steps_enqueued = 0;
    H->numeric_failed = false;
    if (H->cfg.method == ZK_METHOD_WENO && H->cfg.positivity) {
        ZK_HIP(hipMalloc(&H->bad, np));
        ZK_HIP(hipMemset(H->bad, 0, np));
        ZK_HIP(hipMalloc(&H->any, 3 * sizeof(unsigned int)));
        ZK_HIP(hipMemset(H->any, 0, 3 * sizeof(unsigned int)));
        ZK_HIP(hipMalloc(&H->blist, np * sizeof(int2)));
    }
    if (H->cfg.method == ZK_METHOD_DG) {
        const int NQ = H->cfg.P + 1;
        const size_t nxu = nx - 6, nyu = ny - 6;
        ZK_HIP(hipMalloc(&H->gx, 4 * (size_t)NQ * (nxu + 1) * nyu * sizeof(double)));
        ZK_HIP(hipMalloc(&H->gy, 4 * (size_t)NQ * nxu * (nyu + 1) * sizeof(double)));
        ZK_HIP(hipMalloc(&H->means, 4 * np * sizeof(double)));
        ZK_HIP(hipMalloc(&H->zsok, np));
        ZK_HIP(hipMemset(H->zsok, 0, np));
        if (H->cfg.limiter == ZK_LIMITER_TVB) {
            ZK_HIP(hipMalloc(&H->dcount, 3 * sizeof(unsigned)));
            ZK_HIP(hipMemset(H->dcount, 0, 3 * sizeof(unsigned)));
            ZK_HIP(hipMalloc(&H->dcells, (nxu * nyu + 1) * sizeof(int32_t)));
        }
        if (H->cfg.limiter == ZK_LIMITER_LEARNED) {
            ZK_HIP(hipMalloc(&H->tlist, (1 + nxu * nyu) * sizeof(int32_t)));
            ZK_HIP(hipMemset(H->tlist, 0, sizeof(int32_t)));
        }
        ZK_HIP(hipMalloc(&H->stats, 8 * sizeof(unsigned long long)));
        ZK_HIP(hipMemset(H->stats, 0, 8 * sizeof(unsigned long long)));
        if (H->cfg.limiter == ZK_LIMITER_LEARNED) {
            ZK_HIP(hipMalloc(&H->score, np * sizeof(float)));
            ZK_HIP(hipMemset(H->score, 0, np * sizeof(float)));
            ZK_HIP(hipMalloc(&H->cand, (1 + nxu * nyu) * sizeof(int32_t)));
            ZK_HIP(hipMemset(H->cand, 0, sizeof(int32_t)));
        }
        ZK_HIP(hipDeviceGetAttribute(&H->ncu, hipDeviceAttributeMultiprocessorCount, H->cfg.device));
    }
    ZK_HIP(hipMemcpy(H->q, Q.data(), nst * sizeof(double), hipMemcpyHostToDevice));
    ZK_HIP(hipMemcpy(H->qa, Q.data(), nst * sizeof(double), hipMemcpyHostToDevice));
    ZK_HIP(hipMemcpy(H->qb, Q.data(), nst * sizeof(double), hipMemcpyHostToDevice));
    ZK_HIP(hipMemcpy(H->src, src.data(), np * sizeof(int32_t), hipMemcpyHostToDevice));
    if (H->srcy) {
        std::vector<int32_t> srcy;
        build_srcy_map(H->h_bct, ny, nx, src, srcy);
        ZK_HIP(hipMemcpy(H->srcy, srcy.data(), np * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    ZK_HIP(hipMemcpy(H->upd, upd.data(), np, hipMemcpyHostToDevice));
    ZK_HIP(hipMemcpy(H->bct, bctype, np * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!ring.empty()) ZK_HIP(hipMemcpy(H->ring, ring.data(), ring.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    ZK_HIP(hipMemset(H->flags, 0, np));
    zk::StepScalars s0{};
    ZK_HIP(hipMemcpy(H->scal, &s0, sizeof(s0), hipMemcpyHostToDevice));
    ZK_HIP(hipMemset(H->lam_bits, 0, sizeof(unsigned long long)));
    ZK_HIP(hipDeviceSynchronize());
    H->loaded = true;
    enqueue_health(H);  // an input that is not finite is reported by the next synchronising call
    return ZK_OK;
}

zk_status zk_set_state(zk_handle *h, const double *Q)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    if (!Q) {
        H->err = "zk_set_state: null array";
        return ZK_ERR_ARG;
    }
    ZK_HIP(hipSetDevice(H->cfg.device));
    if (H->lay.rs == 0) {
        ZK_HIP(hipMemcpyAsync(H->q, Q, 4 * (size_t)H->nm * H->g.np * sizeof(double), hipMemcpyHostToDevice,
                              H->stream));
    } else {
        // through stage buffer A (rewritten by every stage before it is read)
        ZK_HIP(hipMemcpyAsync(H->qa, Q, 4 * (size_t)H->nm * H->g.np * sizeof(double), hipMemcpyHostToDevice,
                              H->stream));
        relayout(H, H->qa, 4 * H->nm, 1, true);
        ZK_HIP(hipGetLastError());
    }
    // a new state starts a new record
    ZK_HIP(hipMemsetAsync(H->health, 0xff, 2 * sizeof(unsigned long long), H->stream));
    H->numeric_failed = false;
    enqueue_health(H);
    return check_health(H);  // the state is set either way
}

zk_status zk_get_state(zk_handle *h, double *Q)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    if (!Q) {
        H->err = "zk_get_state: null array";
        return ZK_ERR_ARG;
    }
    ZK_HIP(hipSetDevice(H->cfg.device));
    const double *from = H->q;
    if (H->lay.rs != 0) {
        // plane-major copy in stage buffer A (rewritten by every stage before it is read)
        relayout(H, H->qa, 4 * H->nm, 1, false);
        ZK_HIP(hipGetLastError());
        from = H->qa;
    }
    ZK_HIP(hipMemcpyAsync(Q, from, 4 * (size_t)H->nm * H->g.np * sizeof(double), hipMemcpyDeviceToHost,
                          H->stream));
    ZK_HIP(hipStreamSynchronize(H->stream));
    return check_health(H);
}

zk_status zk_get_conserved(zk_handle *h, double *U0, double *U1, double *U2, double *U3)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    double *dst[4] = {U0, U1, U2, U3};
    for (int c = 0; c < 4; ++c)
        if (!dst[c]) {
            H->err = "zk_get_conserved: null array";
            return ZK_ERR_ARG;
        }
    ZK_HIP(hipSetDevice(H->cfg.device));
    ZK_HIP(hipStreamSynchronize(H->stream));
    const double *from = H->q;
    size_t pstride = (size_t)H->nm * H->g.np;
    if (H->lay.rs != 0) {
        // the four mean planes, plane-major in stage buffer A
        relayout(H, H->qa, 4, H->nm, false);
        ZK_HIP(hipGetLastError());
        from = H->qa;
        pstride = H->g.np;
    }
    for (int c = 0; c < 4; ++c)
        ZK_HIP(hipMemcpyAsync(dst[c], from + (size_t)c * pstride, H->g.np * sizeof(double), hipMemcpyDeviceToHost,
                              H->stream));
    ZK_HIP(hipStreamSynchronize(H->stream));
    return check_health(H);
}

zk_status zk_max_speed(zk_handle *h, double *out)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    if (!out) {
        H->err = "zk_max_speed: null output";
        return ZK_ERR_ARG;
    }
    ZK_HIP(hipSetDevice(H->cfg.device));
    if (H->cfg.method == ZK_METHOD_WENO) zk::fv_max_speed(*H, H->q);
    else zk::dg_max_speed(*H, H->q);
    ZK_HIP(hipGetLastError());
    unsigned long long bits = 0;
    ZK_HIP(hipMemcpyAsync(&bits, H->lam_bits, sizeof(bits), hipMemcpyDeviceToHost, H->stream));
    ZK_HIP(hipStreamSynchronize(H->stream));
    double lam;
    std::memcpy(&lam, &bits, sizeof(lam));
    *out = lam;
    return ZK_OK;
}

zk_status zk_step(zk_handle *h, double dt, int nsteps)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    if (H->cfg.method == ZK_METHOD_DG && H->cfg.limiter == ZK_LIMITER_LEARNED && !H->mlp) {
        H->err = "ZK_LIMITER_LEARNED needs a network: call zk_attach_mlp first";
        return ZK_ERR_ARG;
    }
    if (!(dt >= 0.0) || nsteps < 0) {
        H->err = "zk_step: need dt >= 0 and nsteps >= 0";
        return ZK_ERR_ARG;
    }
    if (H->numeric_failed) return ZK_ERR_NUMERIC;  // H->err still names the cell
    ZK_HIP(hipSetDevice(H->cfg.device));
    hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(64), 0, H->stream, H->scal, dt);
    st = enqueue_steps(H, nsteps);
    if (st) return st;
    H->steps_enqueued += nsteps;
    if (nsteps > 0) enqueue_health(H);
    return ZK_OK;
}

zk_status zk_step_cfl(zk_handle *h, double cfl, double t_end, int nsteps, double *t_out, int *steps_out)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    if (!(cfl > 0.0) || nsteps < 0) {
        H->err = "zk_step_cfl: need cfl > 0 and nsteps >= 0";
        return ZK_ERR_ARG;
    }
    if (H->cfg.method == ZK_METHOD_DG) {
        // the stable limit of SSP-RK3 DG P, 0.9 / (2 (2P + 1)); the reference
        // refuses to run above it (src_DG/solver.cpp:12-17, 245-260)
        const double lim = 0.9 / (2.0 * (2 * H->cfg.P + 1));
        if (cfl > lim) {
            char msg[160];
            std::snprintf(msg, sizeof(msg), "zk_step_cfl: CFL %.6g exceeds the DG P=%d stable limit %.6g", cfl,
                          H->cfg.P, lim);
            H->err = msg;
            return ZK_ERR_CFL;
        }
    }
    if (H->numeric_failed) return ZK_ERR_NUMERIC;  // H->err still names the cell
    ZK_HIP(hipSetDevice(H->cfg.device));
    zk::StepScalars s{};
    ZK_HIP(hipMemcpyAsync(&s, H->scal, sizeof(s), hipMemcpyDeviceToHost, H->stream));
    ZK_HIP(hipStreamSynchronize(H->stream));
    s.cfl = cfl;
    s.t_end = t_end;
    s.done = 0;
    const int steps0 = s.steps;
    ZK_HIP(hipMemcpyAsync(H->scal, &s, sizeof(s), hipMemcpyHostToDevice, H->stream));
    ZK_HIP(hipStreamSynchronize(H->stream));
    // One CFL step (wave speed, dt, the three stages, time advance) reads and
    // writes only device memory; it is captured once and replayed like
    // zk_step's, in chunks of 16 between host reads of the step scalars.  A
    // step after t_end (or after a wave speed that is not finite) has done = 1
    // and dt = 0: its stage kernels return at once, and the host stops.
    const bool salted = H->cfg.method == ZK_METHOD_DG && (H->cfg.limiter == ZK_LIMITER_LEARNED || H->harvest);
    const bool graph = !salted && H->use_graph && !H->prof;
    for (int k = 0; k < nsteps; k += 16) {
        const int chunk = nsteps - k < 16 ? nsteps - k : 16;
        for (int c = 0; c < chunk; ++c) {
            if (graph && H->cfl_exec) {
                ZK_HIP(hipGraphLaunch(H->cfl_exec, H->stream));
                continue;
            }
            if ((st = enqueue_cfl_step(H)) != ZK_OK) return st;
            if (graph) {
                // captured after the first step ran as plain launches, so
                // every kernel of the step is loaded before a capture records it
                ZK_HIP(hipStreamBeginCapture(H->stream, hipStreamCaptureModeThreadLocal));
                st = enqueue_cfl_step(H);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(H->stream, &g);
                if (st) {
                    if (g) (void)hipGraphDestroy(g);
                    return st;
                }
                ZK_HIP(e);
                H->cfl_graph = g;
                ZK_HIP(hipGraphInstantiate(&H->cfl_exec, g, nullptr, nullptr, 0));
            }
        }
        H->steps_enqueued += chunk;
        ZK_HIP(hipMemcpyAsync(&s, H->scal, sizeof(s), hipMemcpyDeviceToHost, H->stream));
        ZK_HIP(hipStreamSynchronize(H->stream));
        if (s.done) break;
    }
    ZK_HIP(hipMemcpyAsync(&s, H->scal, sizeof(s), hipMemcpyDeviceToHost, H->stream));
    ZK_HIP(hipStreamSynchronize(H->stream));
    if (t_out) *t_out = s.t;
    if (steps_out) *steps_out = s.steps - steps0;
    enqueue_health(H);
    st = check_health(H);
    if (st) return st;
    if (s.bad) {
        // finite conserved values whose wave speed is not (zero density): the
        // stepping stopped at the step that would have divided by it
        const int at = s.steps - steps0;
        s.bad = 0;
        ZK_HIP(hipMemcpyAsync(H->scal, &s, sizeof(s), hipMemcpyHostToDevice, H->stream));
        ZK_HIP(hipStreamSynchronize(H->stream));
        char msg[160];
        std::snprintf(msg, sizeof(msg), "zk_step_cfl: the wave-speed maximum is not finite after %d steps "
                      "(a cell with zero density)", at);
        H->err = msg;
        return ZK_ERR_NUMERIC;
    }
    return ZK_OK;
}

zk_status zk_sync(zk_handle *h)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    if (!H) return ZK_ERR_ARG;
    ZK_HIP(hipSetDevice(H->cfg.device));
    ZK_HIP(hipStreamSynchronize(H->stream));
    return H->loaded ? check_health(H) : ZK_OK;
}

zk_status zk_step_timed(zk_handle *h, double dt, int nsteps, float *ms)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    ZK_HIP(hipSetDevice(H->cfg.device));
    hipEvent_t e0, e1;
    ZK_HIP(hipEventCreate(&e0));
    ZK_HIP(hipEventCreate(&e1));
    ZK_HIP(hipEventRecord(e0, H->stream));
    st = zk_step(h, dt, nsteps);
    if (st) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return st;
    }
    ZK_HIP(hipEventRecord(e1, H->stream));
    ZK_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    ZK_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return check_health(H);
}

zk_status zk_get_flags(zk_handle *h, uint8_t *flags)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    if (!flags) {
        H->err = "zk_get_flags: null array";
        return ZK_ERR_ARG;
    }
    ZK_HIP(hipSetDevice(H->cfg.device));
    ZK_HIP(hipMemcpyAsync(flags, H->flags, H->g.np, hipMemcpyDeviceToHost, H->stream));
    ZK_HIP(hipStreamSynchronize(H->stream));
    return ZK_OK;
}

zk_status zk_profile(zk_handle *h, int enable)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    if (!H) return ZK_ERR_ARG;
    ZK_HIP(hipSetDevice(H->cfg.device));
    ZK_HIP(hipStreamSynchronize(H->stream));
    for (auto &r : H->recs) {
        H->pool.push_back(r.a);
        H->pool.push_back(r.b);
    }
    H->recs.clear();
    H->prof = enable != 0;
    return ZK_OK;
}

zk_status zk_profile_report(zk_handle *h, char *buf, size_t len)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    if (!H || !buf || len == 0) return ZK_ERR_ARG;
    ZK_HIP(hipSetDevice(H->cfg.device));
    ZK_HIP(hipStreamSynchronize(H->stream));
    std::vector<std::string> names;
    std::vector<double> tot;
    std::vector<long> cnt;
    for (auto &r : H->recs) {
        float ms = 0.f;
        ZK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        size_t i = 0;
        while (i < names.size() && names[i] != r.name) ++i;
        if (i == names.size()) {
            names.emplace_back(r.name);
            tot.push_back(0.0);
            cnt.push_back(0);
        }
        tot[i] += ms;
        cnt[i] += 1;
    }
    std::string js = "{";
    for (size_t i = 0; i < names.size(); ++i) {
        char tmp[256];
        std::snprintf(tmp, sizeof(tmp), "%s\"%s\": [%ld, %.6f]", i ? ", " : "", names[i].c_str(), cnt[i], tot[i]);
        js += tmp;
    }
    js += "}";
    if (js.size() + 1 > len) {
        H->err = "zk_profile_report: buffer too small";
        return ZK_ERR_ARG;
    }
    std::memcpy(buf, js.c_str(), js.size() + 1);
    return ZK_OK;
}

zk_status zk_set_slip_walls(zk_handle *h, int enable)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    if (!H) return ZK_ERR_ARG;
    if (H->cfg.method != ZK_METHOD_WENO || (enable != 0 && enable != 1)) {
        H->err = "zk_set_slip_walls: WENO handle and enable 0 or 1";
        return ZK_ERR_ARG;
    }
    H->fv_slip = enable != 0;
    if (!H->loaded) return ZK_OK;  // zk_load builds the ghost map with it
    // rebuild the ghost map of the loaded grid in place
    ZK_HIP(hipSetDevice(H->cfg.device));
    ZK_HIP(hipStreamSynchronize(H->stream));
    drop_graph(H);
    std::vector<int32_t> src, srcy;
    build_src_map(H->h_bct, H->g.ny, H->g.nx, H->fv_slip ? 2 : 0, src);
    build_srcy_map(H->h_bct, H->g.ny, H->g.nx, src, srcy);
    ZK_HIP(hipMemcpy(H->src, src.data(), src.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    ZK_HIP(hipMemcpy(H->srcy, srcy.data(), srcy.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    return ZK_OK;
}

zk_status zk_set_reconstruction(zk_handle *h, int kind)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    if (!H) return ZK_ERR_ARG;
    if (H->cfg.method != ZK_METHOD_WENO || kind < ZK_RECON_WENO5 || kind > ZK_RECON_MUSCL_SUPERBEE) {
        H->err = "zk_set_reconstruction: WENO handle and a ZK_RECON_* value";
        return ZK_ERR_ARG;
    }
    drop_graph(H);  // the captured step holds the old choice in its arguments
    H->recon = kind;
    return ZK_OK;
}

zk_status zk_set_tuning(zk_handle *h, int fv_tile, int dg_variant, int use_graph)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    if (!H) return ZK_ERR_ARG;
    const bool dg_ok = dg_variant == -1 || dg_variant == 0 || (dg_variant >= 2 && dg_variant <= 4) ||
                       (dg_variant >= 14 && dg_variant <= 74) || dg_variant == 100 ||
                       (dg_variant >= 104 && dg_variant <= 164);
    if (fv_tile < 0 || fv_tile > 2 || !dg_ok) {
        H->err = "zk_set_tuning: fv_tile in 0..2, dg_variant -1, 0, 2..4, 10+R or 100+R with R in 4..64 (100: automatic R)";
        return ZK_ERR_ARG;
    }
    drop_graph(H);
    H->fv_tile = fv_tile;
    H->no_mfma = dg_variant == -1;
    H->tune = dg_variant > 0 ? dg_variant : 0;
    H->use_graph = use_graph != 0;
    return ZK_OK;
}

zk_status zk_attach_mlp(zk_handle *h, zk_mlp *m, float lo, float hi, float audit)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    if (!H) return ZK_ERR_ARG;
    if (H->cfg.method != ZK_METHOD_DG || H->cfg.limiter != ZK_LIMITER_LEARNED) {
        H->err = "zk_attach_mlp: handle must be DG P=2 with ZK_LIMITER_LEARNED";
        return ZK_ERR_ARG;
    }
    if (!(lo >= 0.0f && lo <= hi && hi <= 1.0f && audit >= 0.0f && audit <= 1.0f)) {
        H->err = "zk_attach_mlp: need 0 <= lo <= hi <= 1 and 0 <= audit <= 1";
        return ZK_ERR_ARG;
    }
    H->mlp = m;
    H->mlp_lo = lo;
    H->mlp_hi = hi;
    H->mlp_audit = audit;
    return ZK_OK;
}

zk_status zk_harvest(zk_handle *h, zk_mlp *m, float keep_pos, float keep_neg)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    if (!H) return ZK_ERR_ARG;
    if (m && (H->cfg.method != ZK_METHOD_DG || H->cfg.P != 2 || H->cfg.limiter != ZK_LIMITER_TVB)) {
        H->err = "zk_harvest: handle must be DG P=2 with ZK_LIMITER_TVB";
        return ZK_ERR_ARG;
    }
    H->harvest = m;
    H->keep_pos = keep_pos;
    H->keep_neg = keep_neg;
    return ZK_OK;
}

zk_status zk_learned_stats(zk_handle *h, int64_t out[8], int reset)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    if (!H->stats) {
        H->err = "zk_learned_stats: DG handles only";
        return ZK_ERR_ARG;
    }
    ZK_HIP(hipSetDevice(H->cfg.device));
    unsigned long long v[8] = {0};
    ZK_HIP(hipMemcpyAsync(v, H->stats, sizeof(v), hipMemcpyDeviceToHost, H->stream));
    ZK_HIP(hipStreamSynchronize(H->stream));
    if (out)
        for (int k = 0; k < 8; ++k) out[k] = (int64_t)v[k];
    if (reset) {
        ZK_HIP(hipMemsetAsync(H->stats, 0, sizeof(v), H->stream));
        ZK_HIP(hipStreamSynchronize(H->stream));
    }
    return ZK_OK;
}

zk_status zk_mlp_features(zk_handle *h, zk_mlp *m, const int32_t *cells, int n, uint16_t *X, float *scores)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    if (H->cfg.method != ZK_METHOD_DG || H->cfg.P != 2 || !cells || n <= 0) {
        H->err = "zk_mlp_features: DG P=2 handle and a non-empty cell list";
        return ZK_ERR_ARG;
    }
    for (int t = 0; t < n; ++t) {
        const int i = cells[t] / H->g.nx, j = cells[t] % H->g.nx;
        if (cells[t] < 0 || i < 3 || i >= H->g.ny - 3 || j < 3 || j >= H->g.nx - 3) {
            H->err = "zk_mlp_features: cells must lie in [3,ny-3)x[3,nx-3)";
            return ZK_ERR_ARG;
        }
    }
    ZK_HIP(hipSetDevice(H->cfg.device));
    ZK_HIP(hipStreamSynchronize(H->stream));
    return zk::mlp_features_of(*H, m, cells, n, X, scores);
}

zk_status zk_updated_cells(zk_handle *h, int64_t *out)
{
    Handle *H = reinterpret_cast<Handle *>(h);
    zk_status st = check_args(H);
    if (st) return st;
    if (out) *out = H->n_updated;
    return ZK_OK;
}

}  // extern "C"
