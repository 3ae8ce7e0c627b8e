// zk_internal.h -- shared between the C-ABI layer and the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/zk.h"

struct zk_mlp;

namespace zk {

// Ghost-fill map entry: (source cell index << 3) | flags.  Bit 0 negates the
// x momentum, bit 1 the y momentum (slip ghosts); bit 2 marks an updated cell
// (bctype 0 inside the update range), so that kernels which stage the map
// rows in LDS have the flag without another global load.  A cell that is not
// filled maps to itself.
constexpr int kSrcShift = 3;
constexpr int32_t kSrcUpdated = 4;

// Layout of a state array: element (plane p, cell k) sits at
// p * ps + k + (k / nx) * rs.  WENO keeps its planes one after the other
// (ps = ny nx, rs = 0); DG puts the planes of a grid row side by side,
// [row][plane][column] (ps = nx, rs = (planes - 1) nx), so that the row of
// every plane a stage kernel reads or writes lies in one span of memory
// (fewer DRAM pages and translations per row than planes ny nx apart).
struct StLay {
    size_t ps = 0;
    uint32_t nx = 1, rs = 0;
    // offset of plane 0 of cell k (cell indices stay below 2^31)
    __host__ __device__ __forceinline__ size_t cell(size_t k) const
    {
        return k + (size_t)((uint32_t)k / nx) * rs;
    }
    __host__ __device__ __forceinline__ size_t at(int p, size_t k) const { return (size_t)p * ps + cell(k); }
    // offset of plane 0 of the cell in row i, column j (no division)
    __host__ __device__ __forceinline__ size_t rc(size_t i, size_t j) const { return i * (nx + rs) + j; }
};

// Per-step scalars that stage kernels read from device memory, so that a
// captured graph replays with a dt the device itself computed.
struct StepScalars {
    double dt;        // time step of the current step
    double t;         // physical time reached
    double lambda;    // last max wave speed
    double t_end;     // clip target for CFL stepping
    double cfl;       // CFL number for CFL stepping
    int32_t done;     // 1 once t >= t_end (CFL stepping stops updating)
    int32_t steps;    // steps taken by CFL stepping
    int32_t bad;      // 1 once a wave-speed maximum was not finite (CFL stepping stops)
    int32_t pad;
};

struct Grid {
    int ny = 0, nx = 0;
    size_t np = 0;        // ny*nx
    double da = 0, gamma = 0;
};

struct Handle {
    zk_config cfg{};
    std::string err;
    int nm = 1;           // modes per variable
    Grid g;
    hipStream_t stream = nullptr;
    bool loaded = false;
    bool no_mfma = false;     // force the VALU cell kernel (ZK_NO_MFMA=1), for comparison
    int tune = 0;             // kernel variant selector (ZK_TUNE), used by the autotuner
    int recon = 0;            // WENO face reconstruction: 0 WENO5, 1..3 MUSCL (zk_set_reconstruction)
    bool fv_slip = false;     // WENO: bctype-1 cells are mirror slip walls (zk_set_slip_walls)
    int dma = 3;              // workgroup fused DG kernel: 0 tables from constant memory, 1 tables in LDS,
                              // 2 tables in LDS and rows by global-to-LDS loads, 3 the same with two
                              // row buffers (ZK_DMA overrides)
    bool mf4 = true;          // fused DG kernel (two row buffers) on 4x4x4 matrix-core blocks (ZK_MF4=0: 16x16x4)
    bool fused_generic = false;  // ZK_FUSED_GENERIC=1: the fused kernel's form for state arrays of 4 GiB and more
                                 // (stage read at run time, 64-bit plane offsets) at every size
    StLay lay;                // layout of q, qa, qb
    bool dg_planes = false;   // DG state planes one after the other (ZK_DG_PLANES=1) instead of rows interleaved
    bool fused_al = true;     // fused DG kernel on aligned 64-column blocks (ZK_FUSED_AL=0: the 62-column runs)
    bool tvb_lane = true;     // TVB after the fused kernel: one lane per cell (ZK_TVB_LANE=0: the 16-cell scan)
    bool tvb_linear = false;  // ZK_LIMITER_TVB_LINEAR: the TVB test on the linear modes alone at every P
    int fv_tile = 0;          // WENO tile shape: 0 auto, 1 = 16x15, 2 = 32x15
    bool use_graph = true;    // WENO zk_step: replay one captured step as a HIP graph
    hipGraph_t step_graph = nullptr;
    hipGraphExec_t step_exec = nullptr;
    hipGraph_t cfl_graph = nullptr;      // one CFL step: wave speed, dt, three stages, time advance
    hipGraphExec_t cfl_exec = nullptr;
    int64_t n_updated = 0;

    // device buffers
    double *q = nullptr;      // state U^n (4*nm planes, laid out as lay says)
    double *qa = nullptr;     // stage buffer A
    double *qb = nullptr;     // stage buffer B
    int32_t *src = nullptr;   // ghost-fill map (ny*nx)
    int32_t *srcy = nullptr;  // WENO: ghost map of the y faces (column mirrors of slip walls)
    uint8_t *upd = nullptr;   // 1 = updated cell
    int32_t *ring = nullptr;  // indices of cells outside [3,ny-3)x[3,nx-3)
    int nring = 0;
    bool ring_read = false;   // a neighbour of an updated cell takes its values from a ring cell
    double *gx = nullptr;     // DG x-face projected fluxes
    double *gy = nullptr;     // DG y-face projected fluxes
    double *means = nullptr;  // DG stage means (4 planes) for the TVB limiter
    uint8_t *flags = nullptr; // limiter flags
    StepScalars *scal = nullptr;          // device
    unsigned long long *lam_bits = nullptr; // device max-reduce target
    // non-finite state check, run after every zk_step / zk_step_cfl / zk_set_state:
    // [0] = smallest cell*4 + variable of a non-finite mean, [1] = smallest step
    // count at which a check saw one; both all ones while the state is finite
    unsigned long long *health = nullptr;
    long long steps_enqueued = 0;   // steps enqueued since zk_load (the count the checks record)
    bool numeric_failed = false;    // a synchronising call found a non-finite state
    int32_t *bct = nullptr;   // bctype copy (device)
    std::vector<int32_t> h_bct;

    // learned troubled-cell indicator (DG P=2)
    zk_mlp *mlp = nullptr;           // network used by ZK_LIMITER_LEARNED
    float mlp_lo = 0.2f, mlp_hi = 0.8f, mlp_audit = 0.02f;
    zk_mlp *harvest = nullptr;       // dataset receiving (features, TVB flag) pairs
    float keep_pos = 1.0f, keep_neg = 0.05f;
    uint32_t stage_counter = 0;
    unsigned long long *stats = nullptr;  // audited, agree, miss, false flag, mlp flagged, band, candidates, audited TVB positives
    float *score = nullptr;
    int32_t *cand = nullptr;  // [0] = count, then the candidate cells of the stage
    uint8_t *bad = nullptr;       // WENO cell positivity fallback: bad-cell flags of the stage
    unsigned int *any = nullptr;  // bad-cell counts, one per stage
    int2 *blist = nullptr;        // the bad cells of the current stage
    uint8_t *zsok = nullptr;      // DG TVB: cells whose positivity scaling is certified a no-op
    int32_t *tlist = nullptr;     // DG learned: [0] = count, then the cells the limiter pass treats
    unsigned *dcount = nullptr;   // DG TVB: uncertified cells the lane pass deferred, one count per stage
    int32_t *dcells = nullptr;    // DG TVB: those cells
    int ncu = 0;              // compute units of the device

    // kernel timing: when prof is set every launcher brackets its kernel with
    // a pair of events on the handle's stream
    bool prof = false;
    struct Rec { const char *name; hipEvent_t a, b; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
};

// Brackets one kernel launch with events when profiling is enabled.
struct ProfScope {
    Handle &H;
    int idx = -1;
    ProfScope(Handle &h, const char *name) : H(h)
    {
        if (!H.prof) return;
        hipEvent_t a, b;
        if (H.pool.size() >= 2) {
            a = H.pool.back(); H.pool.pop_back();
            b = H.pool.back(); H.pool.pop_back();
        } else {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
        }
        H.recs.push_back({name, a, b});
        idx = (int)H.recs.size() - 1;
        (void)hipEventRecord(a, H.stream);
    }
    ~ProfScope()
    {
        if (idx >= 0) (void)hipEventRecord(H.recs[idx].b, H.stream);
    }
};

// ---- launchers (defined in the kernel files) ----
// WENO finite volume: one SSP-RK3 stage.  stage = 0, 1, 2.
void fv_stage(Handle &H, const double *Un, const double *Uin, double *Uout, int stage);
void fv_max_speed(Handle &H, const double *U);   // result in H.lam_bits
void dg_stage(Handle &H, const double *Qn, const double *Qin, double *Qout, int stage);
void dg_max_speed(Handle &H, const double *Q);   // result in H.lam_bits
// CFL step control: dt = cfl*da/lambda (clip to t_end) computed on device.
void cfl_dt(Handle &H, double cfl_limit_scale);
void advance_time(Handle &H);

}  // namespace zk
