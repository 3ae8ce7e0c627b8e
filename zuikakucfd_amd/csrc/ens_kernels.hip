// ens_kernels.hip -- batched ensembles of independent WENO5 problems.
//
// Instances of different grid sizes are packed one after another in shared
// state buffers (instance k: 4 planes of ny_k*nx_k doubles at offset off_k)
// and advanced together: one persistent launch per stage walks a work list of
// (instance, tile) items, so hundreds of small grids fill the GPU as one
// large one.  Every instance has its own dt = cfl_k * h_k / lambda_k,
// computed on the device each step from its own wave-speed maximum
// (computeMaxSpeed's rule, src_WENO/fluid.cpp:539-553).  A whole step (wave
// speeds, dt, three stages with their ghost-ring passes, time advance) is
// captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "fv_tile.h"
#include "tile_map.h"

namespace zk {
void build_src_map_public(const std::vector<int32_t> &bct, int ny, int nx, bool slip, std::vector<int32_t> &src);
}

namespace {

using zk::FvArgs;
using zk::StepScalars;

struct EnsInst {
    int ny, nx;
    int tiles_x, ntiles;
    size_t off;      // first state element (plane 0) of this instance in the packed buffers
    size_t cell_off; // first cell in the packed per-cell arrays (src, upd)
    size_t np;
    double da, gamma, cfl;
};

struct EnsArgs {
    const EnsInst *inst;
    const int2 *work;        // (instance, tile)
    int nwork;
    const int2 *ring;        // (instance, local cell) of cells outside [3,n-3)^2
    int nring;
    const int2 *chunks;      // (instance, first cell) chunks of 1024 cells for the wave-speed max
    int nchunks;
    const double *Un;
    const double *Uin;
    double *Uout;
    const int32_t *src;      // per-instance local source maps, packed
    const uint8_t *upd;
    StepScalars *S;          // one per instance
    unsigned long long *lam; // one per instance
    int ninst;
    int stage;
    double ca, cb;
    int flux, positivity;
    uint8_t *bad;            // cell positivity fallback: per-cell flags (packed), or null
    unsigned int *any;       // bad-cell counts, one per stage
    int2 *blist;             // the bad cells of the stage: (instance, local cell)
};

__device__ __forceinline__ FvArgs make_args(const EnsArgs &E, const EnsInst &I, int idx)
{
    FvArgs A;
    A.Un = E.Un + I.off;
    A.Uin = E.Uin + I.off;
    A.Uout = E.Uout + I.off;
    A.src = E.src + I.cell_off;
    A.srcy = nullptr;
    A.upd = E.upd + I.cell_off;
    A.ny = I.ny; A.nx = I.nx; A.np = I.np;
    A.da = I.da; A.gamma = I.gamma;
    A.ca = E.ca; A.cb = E.cb;
    A.stage = E.stage; A.flux = E.flux; A.positivity = E.positivity; A.recon = 0;
    A.tiles_x = I.tiles_x;
    A.ntiles = 0; A.ring = nullptr; A.nring = 0;  // the ensemble has its own ring launch
    A.bad = E.bad ? E.bad + I.cell_off : nullptr;
    A.any = E.any ? E.any + E.stage : nullptr;
    A.blist = E.blist;
    A.inst = idx;
    return A;
}

template <int TX, int TY, bool F32>
__global__ __launch_bounds__(zk::FV_THREADS) void ens_stage_kernel(EnsArgs E)
{
    __shared__ double s_u[4][TY + 6][TX + 6];
    __shared__ double s_f[4][(TX + 1) * (TY + 1)];
    for (int w = blockIdx.x; w < E.nwork; w += gridDim.x) {
        const int2 it = E.work[w];
        const EnsInst I = E.inst[it.x];
        FvArgs A = make_args(E, I, it.x);
        A.S = E.S + it.x;
        zk::fv_tile_body<TX, TY, F32>(A, it.y, s_u, s_f);
        __syncthreads();  // the next item restages the LDS tile
    }
}

// Cell positivity fallback, pass 2: a grid-strided launch over five work
// items per bad cell pass 1 listed, (instance, cell); ends at once when there
// are none.
template <bool F32>
__global__ __launch_bounds__(256) void ens_fallback_kernel(EnsArgs E)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) E.any[(E.stage + 1) % 3] = 0u;
    const size_t n = (size_t)E.any[E.stage] * 5;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int2 b = E.blist[t / 5];
        const EnsInst I = E.inst[b.x];
        FvArgs A = make_args(E, I, b.x);
        A.S = E.S + b.x;
        zk::fv_fallback_item<F32>(A, (size_t)b.y, (int)(t % 5));
    }
}

__global__ void ens_ring_kernel(EnsArgs E)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= E.nring) return;
    const int2 rc = E.ring[t];
    const EnsInst I = E.inst[rc.x];
    const size_t k = (size_t)rc.y;
    const int32_t e = E.src[I.cell_off + k];
    const size_t s = (size_t)(e >> zk::kSrcShift);
    for (int v = 0; v < 4; ++v) {
        const double un = E.Un[I.off + v * I.np + k];
        const double uf = E.Uin[I.off + v * I.np + s];
        E.Uout[I.off + v * I.np + k] = (E.stage == 0) ? un : E.ca * un + E.cb * (uf + 0.0);
    }
}

__global__ void ens_zero_kernel(unsigned long long *lam, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) lam[t] = 0ull;
}

// Per-instance max(|u| + c) over every cell of the instance, ghosts included.
__global__ __launch_bounds__(256) void ens_max_speed_kernel(EnsArgs E)
{
    const int2 ch = E.chunks[blockIdx.x];
    const EnsInst I = E.inst[ch.x];
    const double *U = E.Un + I.off;
    const size_t np = I.np;
    const double gamma = I.gamma;
    double m = 0.0;
    const size_t end = (size_t)ch.y + 1024 < np ? (size_t)ch.y + 1024 : np;
    for (size_t k = (size_t)ch.y + threadIdx.x; k < end; k += blockDim.x) {
        const double rho = U[k];
        const double u = U[np + k] / rho, v = U[2 * np + k] / rho;
        const double p = (gamma - 1.0) * (U[3 * np + k] - 0.5 * rho * (u * u + v * v));
        const double a = sqrt(gamma * p / rho);
        m = zk::speed_max(m, sqrt(u * u + v * v) + a, zk::finite4(rho, U[np + k], U[2 * np + k], U[3 * np + k]));
    }
    zk::block_max_to(m, E.lam + ch.x);
}

__global__ void ens_dt_kernel(EnsArgs E)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E.ninst) return;
    const double lam = __longlong_as_double((long long)E.lam[i]);
    StepScalars &S = E.S[i];
    S.lambda = lam;
    S.done = 0;
    if (S.bad || !(lam <= 1.79769313486231570815e308)) {
        // a state that is no longer finite (or a zero density): the instance
        // stops with dt = 0 and zk_ensemble_sync reports it
        S.bad = 1;
        S.dt = 0.0;
        return;
    }
    S.dt = E.inst[i].cfl * E.inst[i].da / lam;
}

__global__ void ens_advance_kernel(EnsArgs E)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E.ninst || E.S[i].bad) return;  // a stopped instance keeps its time and step count
    E.S[i].t += E.S[i].dt;
    E.S[i].steps += 1;
}

}  // namespace

struct zk_ensemble {
    zk_config cfg{};
    std::string err;
    hipStream_t stream = nullptr;
    // host-side instance data until finalize
    std::vector<EnsInst> inst;
    std::vector<double> h_state;
    std::vector<int32_t> h_src;
    std::vector<uint8_t> h_upd;
    std::vector<int2> work, ring, chunks;
    bool final = false;
    int tile = 1;        // 1 = 16x15, 2 = 32x15
    int use_graph = 1;
    size_t nstate = 0, ncells = 0;
    double *q = nullptr, *qa = nullptr, *qb = nullptr;
    int32_t *src = nullptr;
    uint8_t *upd = nullptr;
    EnsInst *d_inst = nullptr;
    int2 *d_work = nullptr, *d_ring = nullptr, *d_chunks = nullptr;
    StepScalars *S = nullptr;
    unsigned long long *lam = nullptr;
    uint8_t *bad = nullptr;       // cell positivity fallback (positivity on)
    unsigned int *any = nullptr;
    int2 *blist = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int persistent_blocks = 0;
};

namespace {

#define ENS_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t _e = (call);                                                           \
        if (_e != hipSuccess) {                                                           \
            E->err = std::string(#call) + ": " + hipGetErrorString(_e);                   \
            return ZK_ERR_HIP;                                                            \
        }                                                                                 \
    } while (0)

EnsArgs args_for(zk_ensemble *E, const double *Un, const double *Uin, double *Uout, int stage)
{
    EnsArgs A{};
    A.inst = E->d_inst;
    A.work = E->d_work; A.nwork = (int)E->work.size();
    A.ring = E->d_ring; A.nring = (int)E->ring.size();
    A.chunks = E->d_chunks; A.nchunks = (int)E->chunks.size();
    A.Un = Un; A.Uin = Uin; A.Uout = Uout;
    A.src = E->src; A.upd = E->upd; A.S = E->S; A.lam = E->lam;
    A.ninst = (int)E->inst.size();
    A.stage = stage;
    A.ca = stage == 1 ? 0.75 : (1.0 / 3.0);
    A.cb = stage == 1 ? 0.25 : (2.0 / 3.0);
    A.flux = E->cfg.flux;
    A.positivity = E->cfg.positivity;
    A.bad = E->bad;
    A.any = E->any;
    A.blist = E->blist;
    return A;
}

void enqueue_ens_step(zk_ensemble *E)
{
    const int ninst = (int)E->inst.size();
    EnsArgs A0 = args_for(E, E->q, E->q, E->qa, 0);
    hipLaunchKernelGGL(ens_zero_kernel, dim3((ninst + 255) / 256), dim3(256), 0, E->stream, E->lam, ninst);
    hipLaunchKernelGGL(ens_max_speed_kernel, dim3(A0.nchunks), dim3(256), 0, E->stream, A0);
    hipLaunchKernelGGL(ens_dt_kernel, dim3((ninst + 255) / 256), dim3(256), 0, E->stream, A0);
    const double *ins[3] = {E->q, E->qa, E->qb};
    double *outs[3] = {E->qa, E->qb, E->q};
    const bool f32 = E->cfg.precision == ZK_PREC_FP32_BULK;
    for (int s = 0; s < 3; ++s) {
        EnsArgs A = args_for(E, E->q, ins[s], outs[s], s);
        const dim3 grid(E->persistent_blocks), blk(zk::FV_THREADS);
        if (E->tile == 2) {
            if (f32) hipLaunchKernelGGL((ens_stage_kernel<32, 15, true>), grid, blk, 0, E->stream, A);
            else hipLaunchKernelGGL((ens_stage_kernel<32, 15, false>), grid, blk, 0, E->stream, A);
        } else {
            if (f32) hipLaunchKernelGGL((ens_stage_kernel<16, 15, true>), grid, blk, 0, E->stream, A);
            else hipLaunchKernelGGL((ens_stage_kernel<16, 15, false>), grid, blk, 0, E->stream, A);
        }
        if (A.bad) {
            const dim3 fg(256);  // grid-strided over the listed bad cells
            if (f32) hipLaunchKernelGGL(ens_fallback_kernel<true>, fg, dim3(256), 0, E->stream, A);
            else hipLaunchKernelGGL(ens_fallback_kernel<false>, fg, dim3(256), 0, E->stream, A);
        }
        if (A.nring > 0) hipLaunchKernelGGL(ens_ring_kernel, dim3((A.nring + 255) / 256), dim3(256), 0, E->stream, A);
    }
    hipLaunchKernelGGL(ens_advance_kernel, dim3((ninst + 255) / 256), dim3(256), 0, E->stream, A0);
}

void free_ens(zk_ensemble *E)
{
    if (E->exec) (void)hipGraphExecDestroy(E->exec);
    if (E->graph) (void)hipGraphDestroy(E->graph);
    E->exec = nullptr;
    E->graph = nullptr;
    void *ptrs[] = {E->q, E->qa, E->qb, E->src, E->upd, E->d_inst, E->d_work, E->d_ring, E->d_chunks, E->S, E->lam,
                    E->bad, E->any, E->blist};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
}

}  // namespace

extern "C" {

zk_status zk_ensemble_create(const zk_config *cfg, zk_ensemble **out)
{
    if (!cfg || !out) return ZK_ERR_ARG;
    if (cfg->method != ZK_METHOD_WENO) return ZK_ERR_ARG;
    zk_ensemble *E = new zk_ensemble();
    E->cfg = *cfg;
    if (hipSetDevice(cfg->device) != hipSuccess ||
        hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
        delete E;
        return ZK_ERR_HIP;
    }
    *out = E;
    return ZK_OK;
}

zk_status zk_ensemble_add(zk_ensemble *E, int ny, int nx, double da, double gamma, const double *rho,
                          const double *u, const double *v, const double *p, const int32_t *bctype, double cfl)
{
    if (!E) return ZK_ERR_ARG;
    if (E->final) {
        E->err = "zk_ensemble_add after zk_ensemble_finalize";
        return ZK_ERR_ARG;
    }
    if (ny < 7 || nx < 7 || !(da > 0) || !(gamma > 1.0) || !(cfl > 0) || !rho || !u || !v || !p || !bctype) {
        E->err = "zk_ensemble_add: bad arguments";
        return ZK_ERR_ARG;
    }
    EnsInst I{};
    I.ny = ny; I.nx = nx; I.np = (size_t)ny * nx; I.da = da; I.gamma = gamma; I.cfl = cfl;
    I.off = E->h_state.size();
    I.cell_off = E->h_src.size();
    E->h_state.resize(I.off + 4 * I.np);
    double *Q = E->h_state.data() + I.off;
    for (size_t k = 0; k < I.np; ++k) {
        const double r = rho[k], ux = u[k], vy = v[k], pr = p[k];
        Q[k] = r;
        Q[I.np + k] = r * ux;
        Q[2 * I.np + k] = r * vy;
        Q[3 * I.np + k] = pr / (gamma - 1.0) + 0.5 * r * (ux * ux + vy * vy);
    }
    std::vector<int32_t> bct(bctype, bctype + I.np), src;
    zk::build_src_map_public(bct, ny, nx, false, src);
    E->h_src.insert(E->h_src.end(), src.begin(), src.end());
    const int idx = (int)E->inst.size();
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            const size_t k = (size_t)i * nx + j;
            const bool in = i >= 3 && i < ny - 3 && j >= 3 && j < nx - 3;
            E->h_upd.push_back(in && bctype[k] == 0 ? 1 : 0);
            if (!in) E->ring.push_back(int2{idx, (int)k});
        }
    for (size_t c = 0; c < I.np; c += 1024) E->chunks.push_back(int2{idx, (int)c});
    E->inst.push_back(I);
    return ZK_OK;
}

zk_status zk_ensemble_set_tuning(zk_ensemble *E, int tile, int use_graph)
{
    if (!E) return ZK_ERR_ARG;
    if (tile != 1 && tile != 2) {
        E->err = "tile must be 1 (16x15) or 2 (32x15)";
        return ZK_ERR_ARG;
    }
    E->tile = tile;
    E->use_graph = use_graph;
    if (E->exec) (void)hipGraphExecDestroy(E->exec);
    if (E->graph) (void)hipGraphDestroy(E->graph);
    E->exec = nullptr;
    E->graph = nullptr;
    if (E->final) {
        E->work.clear();
        const int TX = tile == 2 ? 32 : 16, TY = 15;
        for (size_t k = 0; k < E->inst.size(); ++k) {
            EnsInst &I = E->inst[k];
            I.tiles_x = (I.nx - 6 + TX - 1) / TX;
            I.ntiles = I.tiles_x * ((I.ny - 6 + TY - 1) / TY);
            for (int t = 0; t < I.ntiles; ++t) E->work.push_back(int2{(int)k, t});
        }
        ENS_HIP(hipSetDevice(E->cfg.device));
        ENS_HIP(hipStreamSynchronize(E->stream));
        if (E->d_work) (void)hipFree(E->d_work);
        ENS_HIP(hipMalloc(&E->d_work, E->work.size() * sizeof(int2)));
        ENS_HIP(hipMemcpy(E->d_work, E->work.data(), E->work.size() * sizeof(int2), hipMemcpyHostToDevice));
        ENS_HIP(hipMemcpy(E->d_inst, E->inst.data(), E->inst.size() * sizeof(EnsInst), hipMemcpyHostToDevice));
    }
    return ZK_OK;
}

zk_status zk_ensemble_finalize(zk_ensemble *E)
{
    if (!E) return ZK_ERR_ARG;
    if (E->final) {
        E->err = "zk_ensemble_finalize: the ensemble is already finalized";
        return ZK_ERR_ARG;
    }
    if (E->inst.empty()) {
        E->err = "zk_ensemble_finalize: no instance added";
        return ZK_ERR_ARG;
    }
    ENS_HIP(hipSetDevice(E->cfg.device));
    E->final = true;
    const size_t ns = E->h_state.size(), nc = E->h_src.size();
    E->nstate = ns;
    E->ncells = nc;
    // work list in instance order: neighbouring tiles of one instance are
    // consecutive, so a persistent block sweeps them with warm L2
    const int TX = E->tile == 2 ? 32 : 16, TY = 15;
    E->work.clear();
    for (size_t k = 0; k < E->inst.size(); ++k) {
        EnsInst &I = E->inst[k];
        I.tiles_x = (I.nx - 6 + TX - 1) / TX;
        I.ntiles = I.tiles_x * ((I.ny - 6 + TY - 1) / TY);
        for (int t = 0; t < I.ntiles; ++t) E->work.push_back(int2{(int)k, t});
    }
    ENS_HIP(hipMalloc(&E->q, ns * sizeof(double)));
    ENS_HIP(hipMalloc(&E->qa, ns * sizeof(double)));
    ENS_HIP(hipMalloc(&E->qb, ns * sizeof(double)));
    ENS_HIP(hipMalloc(&E->src, nc * sizeof(int32_t)));
    ENS_HIP(hipMalloc(&E->upd, nc));
    ENS_HIP(hipMalloc(&E->d_inst, E->inst.size() * sizeof(EnsInst)));
    ENS_HIP(hipMalloc(&E->d_work, E->work.size() * sizeof(int2)));
    ENS_HIP(hipMalloc(&E->d_ring, (E->ring.size() + 1) * sizeof(int2)));
    ENS_HIP(hipMalloc(&E->d_chunks, E->chunks.size() * sizeof(int2)));
    ENS_HIP(hipMalloc(&E->S, E->inst.size() * sizeof(StepScalars)));
    ENS_HIP(hipMalloc(&E->lam, E->inst.size() * sizeof(unsigned long long)));
    if (E->cfg.positivity) {
        ENS_HIP(hipMalloc(&E->bad, nc));
        ENS_HIP(hipMemset(E->bad, 0, nc));
        ENS_HIP(hipMalloc(&E->any, 3 * sizeof(unsigned int)));
        ENS_HIP(hipMemset(E->any, 0, 3 * sizeof(unsigned int)));
        ENS_HIP(hipMalloc(&E->blist, nc * sizeof(int2)));
    }
    ENS_HIP(hipMemcpy(E->q, E->h_state.data(), ns * sizeof(double), hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(E->qa, E->h_state.data(), ns * sizeof(double), hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(E->qb, E->h_state.data(), ns * sizeof(double), hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(E->src, E->h_src.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(E->upd, E->h_upd.data(), nc, hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(E->d_inst, E->inst.data(), E->inst.size() * sizeof(EnsInst), hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(E->d_work, E->work.data(), E->work.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!E->ring.empty())
        ENS_HIP(hipMemcpy(E->d_ring, E->ring.data(), E->ring.size() * sizeof(int2), hipMemcpyHostToDevice));
    ENS_HIP(hipMemcpy(E->d_chunks, E->chunks.data(), E->chunks.size() * sizeof(int2), hipMemcpyHostToDevice));
    std::vector<StepScalars> s0(E->inst.size());
    std::memset(s0.data(), 0, s0.size() * sizeof(StepScalars));
    ENS_HIP(hipMemcpy(E->S, s0.data(), s0.size() * sizeof(StepScalars), hipMemcpyHostToDevice));
    int dev = 0, ncu = 0;
    ENS_HIP(hipGetDevice(&dev));
    ENS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // persistent grid: enough resident blocks to fill every CU several times
    E->persistent_blocks = ncu * 4;
    ENS_HIP(hipDeviceSynchronize());
    E->h_state.clear();
    E->h_state.shrink_to_fit();
    return ZK_OK;
}

zk_status zk_ensemble_step(zk_ensemble *E, int nsteps)
{
    if (!E || !E->final || nsteps < 0) return ZK_ERR_ARG;
    ENS_HIP(hipSetDevice(E->cfg.device));
    if (E->use_graph && nsteps > 0) {
        if (!E->exec) {
            // the first step runs as plain launches, so that its kernels are
            // loaded and launched once before the capture records them
            enqueue_ens_step(E);
            ENS_HIP(hipGetLastError());
            --nsteps;
            ENS_HIP(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
            enqueue_ens_step(E);
            ENS_HIP(hipStreamEndCapture(E->stream, &E->graph));
            ENS_HIP(hipGraphInstantiate(&E->exec, E->graph, nullptr, nullptr, 0));
        }
        for (int s = 0; s < nsteps; ++s) ENS_HIP(hipGraphLaunch(E->exec, E->stream));
    } else if (!E->use_graph) {
        for (int s = 0; s < nsteps; ++s) enqueue_ens_step(E);
        ENS_HIP(hipGetLastError());
    }
    return ZK_OK;
}

zk_status zk_ensemble_sync(zk_ensemble *E)
{
    if (!E) return ZK_ERR_ARG;
    ENS_HIP(hipSetDevice(E->cfg.device));
    ENS_HIP(hipStreamSynchronize(E->stream));
    if (!E->final) return ZK_OK;
    // instances whose wave-speed maximum was not finite have stopped
    std::vector<StepScalars> s(E->inst.size());
    ENS_HIP(hipMemcpy(s.data(), E->S, s.size() * sizeof(StepScalars), hipMemcpyDeviceToHost));
    int nbad = 0;
    std::string which;
    for (size_t k = 0; k < s.size(); ++k)
        if (s[k].bad) {
            if (nbad < 8) which += (nbad ? ", " : "") + std::to_string(k) + " (step " + std::to_string(s[k].steps) + ")";
            ++nbad;
        }
    if (!nbad) return ZK_OK;
    E->err = "non-finite state: " + std::to_string(nbad) + " instance(s) stopped, first: " + which;
    return ZK_ERR_NUMERIC;
}

zk_status zk_ensemble_size(zk_ensemble *E, int *ninst, int64_t *updated_cells)
{
    if (!E) return ZK_ERR_ARG;
    if (ninst) *ninst = (int)E->inst.size();
    if (updated_cells) {
        int64_t n = 0;
        for (uint8_t x : E->h_upd) n += x;
        *updated_cells = n;
    }
    return ZK_OK;
}

zk_status zk_ensemble_get(zk_ensemble *E, int idx, double *U, double *t_out)
{
    if (!E || !E->final || idx < 0 || idx >= (int)E->inst.size() || !U) return ZK_ERR_ARG;
    const EnsInst &I = E->inst[idx];
    ENS_HIP(hipSetDevice(E->cfg.device));
    ENS_HIP(hipMemcpyAsync(U, E->q + I.off, 4 * I.np * sizeof(double), hipMemcpyDeviceToHost, E->stream));
    StepScalars s{};
    ENS_HIP(hipMemcpyAsync(&s, E->S + idx, sizeof(s), hipMemcpyDeviceToHost, E->stream));
    ENS_HIP(hipStreamSynchronize(E->stream));
    if (t_out) *t_out = s.t;
    return ZK_OK;
}

const char *zk_ensemble_last_error(const zk_ensemble *E) { return E ? E->err.c_str() : "null ensemble"; }

zk_status zk_ensemble_destroy(zk_ensemble *E)
{
    if (!E) return ZK_OK;
    (void)hipSetDevice(E->cfg.device);
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    free_ens(E);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
    return ZK_OK;
}

}  // extern "C"
