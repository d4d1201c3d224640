// Solver context shared by the C-ABI translation units (api.hip, nn_api.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "tohip_internal.h"

struct tohip_nn;

// One level of the multigrid hierarchy.  Level 0 borrows the context's
// operator buffers (E32, fixmask, dinv32); coarser levels own theirs.
struct tohip_mg_level {
    tohip::Grid g{};
    tohip::ElemOp<double> op64{};
    tohip::ElemOp<float> op32{};
    double* E64 = nullptr;
    float* E32 = nullptr;
    uint8_t* fixmask = nullptr;
    float* dinv = nullptr;
    float* x = nullptr;   // pre-smoothed iterate
    float* rz = nullptr;  // residual after pre-smoothing, then the level's result z
    float* b = nullptr;   // right-hand side (levels >= 1; level 0 uses the PCG residual)
    float* pw = nullptr;  // kept iterate of the lambda_max(D^-1 K) power steps
    float* wd = nullptr;  // {weight, rescale of pw, estimate} in device memory
    float omega = 0.f;    // damped Jacobi weight from the design-independent bound
    double lbound = 0.0;  // upper bound of lambda(D^-1 K) on this level, any design
    tohip::MgLevelView view() const { return {g, op32, E32, fixmask, dinv, wd}; }
};

struct tohip_ctx {
    tohip::Grid g{};
    double E = 1, nu = 0.3, xmin = 1e-3;
    tohip::Penalty pen{0, 0, 1.0, 0.0};  // SIMP penalty (power 1 by default)
    double h[3] = {1, 1, 1};
    int pbi = 1;
    double meandiag = 0;
    tohip::ElemOp<double> op64{};
    tohip::ElemOp<float> op32{};
    std::vector<int64_t> perm;  // internal DOF -> caller DOF (0-based); empty = identity
    std::vector<uint8_t> fixed_internal;
    hipStream_t stream = nullptr;
    std::string err;

    // device buffers
    uint8_t* fixmask = nullptr;
    double *xt = nullptr, *E64 = nullptr;
    float* E32 = nullptr;
    double *f = nullptr, *u = nullptr, *b64 = nullptr, *r64 = nullptr, *t64 = nullptr,
           *dinv64 = nullptr, *p64[2] = {nullptr, nullptr}, *q64 = nullptr;
    float *b32 = nullptr, *x32 = nullptr, *r32 = nullptr, *dinv32 = nullptr,
          *p32[2] = {nullptr, nullptr}, *q32 = nullptr;
    float* r32b = nullptr;  // second residual buffer of the multigrid PCG graphs
    tohip::PcgScalars *sc64 = nullptr, *sc32 = nullptr;
    double* partials = nullptr;  // general per-block partials (energy, OC, sums)
    double* partA = nullptr;     // PCG direction kernel -> update kernel (p.q)
    double* partB = nullptr;     // PCG update/start/refresh -> direction kernel (r.z, r.r)
    double* dscal = nullptr;     // [0] obj, [1] ||r64||^2 (refinement), [2] volume,
                                 // [3..8] warm-start dots, [kWsCoef..kWsRatio] its
                                 // coefficients and ratio
    // filter / OC
    int* f_offs = nullptr;
    double* f_wts = nullptr;
    int f_count = 0;
    int f_halo = 0;  // stencil halo (largest node offset on any axis)
    double f_rmin = -1;
    double *f_invden = nullptr, *f_colsum = nullptr;
    double *xd = nullptr, *xnew = nullptr, *dc = nullptr, *ce = nullptr, *grad = nullptr,
           *cellbuf = nullptr;
    tohip::OcState* oc = nullptr;
    // pinned host mirrors
    tohip::PcgScalars* h_sc = nullptr;
    double* h_scal = nullptr;
    tohip::OcState* h_oc = nullptr;
    // captured PCG iteration graphs
    hipGraphExec_t g64 = nullptr, g32 = nullptr, g64s = nullptr, g32s = nullptr;
    bool have_u = false;
    bool have_load = false;  // tohip_set_load was called
    int device = 0;          // the GPU the context was created on
    tohip_nn* nn = nullptr;  // learned warm start (tohip_nn_attach)
    float* u0f = nullptr;    // network prediction (FP32, internal order)
    hipEvent_t ws_ev[2] = {nullptr, nullptr};  // solver -> network -> solver ordering
    // CG stopping rule (TOHIP_CONV_*); EnergyCriteria state on the device and
    // its pinned host mirror (current between calls), FP64 graphs with the check
    int conv_kind = 0;
    tohip::EnergyCrit* ecrit = nullptr;
    tohip::EnergyCrit* h_ecrit = nullptr;
    hipGraphExec_t g64e = nullptr, g64es = nullptr;
    // assembled K (CGAssemblySolver comparator): CSR in node-major DOF order
    int64_t csr_nnz = 0;
    int64_t* csr_rowptr = nullptr;
    int* csr_cols = nullptr;
    double *csr_vals = nullptr, *csr_diag = nullptr, *csr_dinv = nullptr, *csr_ke = nullptr,
           *csr_md = nullptr;
    int csr_safe = -1;  // safe flag of the values on the device (-1: not assembled)
    // MMA87 state: asymptotes, the two previous designs, the dual state
    double *mma_L = nullptr, *mma_U = nullptr, *mma_x1 = nullptr, *mma_x2 = nullptr;
    tohip::MmaState *mma_st = nullptr, *h_mma = nullptr;
    // dense coarsest-level operator of the multigrid V-cycle (ndof^2 floats,
    // symmetric); null when the iteration kernel is used instead
    float* mg_M = nullptr;
    hipGraphExec_t gcsr = nullptr;
    int last_replacements = 0;
    size_t dev_bytes = 0;  // device memory held by this context
    // multigrid preconditioner (built on first use by a TOHIP_MIXED_MG solve)
    std::vector<tohip_mg_level> mg;
    std::vector<uint8_t> nodemask;  // host copy of fixmask
    float* ones32 = nullptr;        // all-ones "diagonal" for the preconditioned PCG kernels
    hipGraphExec_t g32mg = nullptr, g32mgs = nullptr;
    double* ones64 = nullptr;  // all-ones diagonal of the FP64 multigrid-preconditioned PCG
    hipGraphExec_t g64mg = nullptr, g64mgs = nullptr;
    tohip::MgCheb mg_cheb{};  // coefficients of the coarsest-level Chebyshev solve
    // smoother weights in device memory: 4 floats per level ({w, rescale,
    // estimate, -}) and a trailing 1; adapted from the power steps when
    // mg_wadapt, and the power steps taken so far
    float* mg_wd = nullptr;
    bool mg_wadapt = false;
    int64_t mg_wsteps = 0;
    // the coarsest dense operator is rebuilt on a second stream while the
    // weight estimate and the solve's first FP64 residual run on `stream`:
    // mg_ev[0] orders the rebuild after the new coarsest moduli, mg_ev[1]
    // marks it done, and mg_pending says the next V-cycle must wait on it
    hipStream_t mg_stream = nullptr;
    hipEvent_t mg_ev[2] = {nullptr, nullptr};
    bool mg_pending = false;
    // BESO (tohip_beso_run), allocated on first use
    double *beso_sens[2] = {nullptr, nullptr};
    uint8_t* beso_mask = nullptr;
    tohip::BesoState* beso_st = nullptr;
    tohip::BesoState* h_beso = nullptr;
    int64_t mma_iter = 0;     // MMA iterations of the current run (asymptote schedule)
    // learned warm start in tohip_simp_run: predictions used this run, the
    // consecutive iterations whose prediction did not pay, and whether the
    // run has stopped consulting the network
    int64_t host_syncs = 0;  // host round trips of the solvers (measurement)
    bool time_graphs = false;  // measurement: GPU time of the PCG graph replays
    hipEvent_t graph_ev[2] = {nullptr, nullptr};
    double graph_ms = 0;
    int64_t graph_launches = 0;
    int64_t nn_calls = 0;
    int nn_misses = 0;
    bool nn_off = false;
    int mg_sweeps = 0;        // Jacobi sweeps of the coarsest level when it is too large
                              // for the one-workgroup kernel (= Chebyshev iterations)
};

namespace tohip {
// ctx->dscal slots of the learned warm start: the combination coefficients,
// then the residual ratio of the combination to the previous displacement.
constexpr int kWsCoef = 10;
constexpr int kWsRatio = kWsCoef + 2;
// tohip_simp_run stops consulting the network for the rest of a run once
// kWsMaxMisses iterations in a row had a prediction that left more than
// kWsPayRatio of the starting residual of the previous displacement alone,
// or at once when a solve started from a prediction took more than
// kWsBlowup times the previous solve's PCG iterations (a smaller starting
// residual does not guarantee fewer iterations).
constexpr double kWsPayRatio = 0.25;
constexpr int kWsMaxMisses = 2;
constexpr int64_t kWsBlowup = 2;
// Learned warm start for the current densities and load (nn_api.hip): writes
// the initial guess into ctx->u.  have_prev: combine it with the previous u.
int nn_warm_start(tohip_ctx* ctx, bool have_prev);
size_t nn_device_bytes(const tohip_nn* nn);
// Removes ctx from the network's attached list (the context is being destroyed
// or attached to another network).
void nn_detach_ctx(tohip_nn* nn, tohip_ctx* ctx);
}  // namespace tohip
