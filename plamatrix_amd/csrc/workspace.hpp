// Workspace layout and the kernel-call records used to replay solver chunks
// either as direct launches or as HIP graphs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <tuple>
#include <utility>
#include <vector>

#include "pm_common.hpp"

namespace pm
{

// Device-resident solver state.  Written only by workgroup 0; every value a
// kernel reads is written by an earlier kernel (parity slots t % 2 keep the
// previous iteration's rho / completed count intact while the current one is
// recorded), so no workgroup ever reads a value being written in the same
// launch.
struct SolverState
{
    double rho[2];      // r . z of the direction of iteration t, slot t % 2
    double alpha;       // last step length
    double rr;          // ||r||^2 at the last decision
    double rr0;         // initial ||r||^2
    double tol;         // absolute tolerance on ||r||
    double rel_tol;
    double abs_tol;
    int32_t iter[2];    // completed updates seen by iteration t, slot t % 2 (iter[1] = -1 at start)
    int32_t iteration;  // completed updates (host visible)
    int32_t max_iterations;
    int32_t done;       // see Done below
    int32_t tol_rule;   // 0: tol = max(abs, rel ||r0||); 1: tol = abs + rel ||r0||
    unsigned long long bad_row;  // first row with an unusable Jacobi diagonal (~0 if none)
    unsigned counters[8];
    double den[2];      // single-reduction iteration: p.Ap of iteration t's direction, slot t % 2
};

enum Done : int32_t
{
    kRunning = 0,
    kConverged = 1,
    kIterationLimit = 2,
    kBreakdown = 3,
    kNonFinite = 4,
    kInitNonFinite = 5,
    kBadDiagonal = 6,
};

// COO conversion status (deferred, lowest triplet position first), as the
// reference's CooStatus (coo_to_csr.cu:116-126).
struct CooStatus
{
    unsigned long long invalid_source;  // first triplet with an out-of-range coordinate
    unsigned long long invalid_value;   // first triplet (or run) with a non-finite value / sum
    long long nnz;                      // combined nnz
    long long expected_nnz;             // caller's output nnz (-1 when unchecked)
};

// One recorded kernel launch: the function, its shape and a private copy of
// its arguments, so that it can be launched now or added to a graph later.
struct KernelCall
{
    const void* func = nullptr;
    dim3 grid;
    dim3 block;
    std::shared_ptr<void> storage;
    std::vector<void*> args;
};

template <typename Tuple, size_t... I>
std::vector<void*> tuple_addresses(Tuple& t, std::index_sequence<I...>)
{
    return {static_cast<void*>(&std::get<I>(t))...};
}

template <typename... P, typename... A>
KernelCall make_call(void (*f)(P...), dim3 grid, dim3 block, A... a)
{
    static_assert(sizeof...(P) == sizeof...(A), "argument count mismatch");
    auto tup = std::make_shared<std::tuple<P...>>(static_cast<P>(a)...);
    KernelCall k;
    k.func = reinterpret_cast<const void*>(f);
    k.grid = grid;
    k.block = block;
    k.args = tuple_addresses(*tup, std::index_sequence_for<P...>{});
    k.storage = tup;
    return k;
}

int launch(const KernelCall& k, hipStream_t stream);

// A chunk of recorded calls, optionally instantiated as one HIP graph.
struct Plan
{
    std::vector<KernelCall> calls;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int instantiate();
    int run(hipStream_t stream, bool use_graph);
    void reset();
    ~Plan() { reset(); }
};

} // namespace pm

struct pm_workspace
{
    hipStream_t stream = nullptr;
    bool bound = false;
    int device = -1;

    pm::DeviceBuffer vectors;   // solver vectors
    pm::DeviceBuffer partials;  // kMaxPartials * 8 doubles
    pm::DeviceBuffer state;     // SolverState + small scalars
    pm::DeviceBuffer scratch;   // miscellaneous (relres, validation)
    pm::DeviceBuffer sort;      // COO conversion buffers
    pm::DeviceBuffer refine;    // FP64 residual / FP32 correction of the refinement loop
    pm::DeviceBuffer batch;     // batched solver: layout, per-system state, per-chunk partials
    pm::DeviceBuffer blocks;    // batched solver: per-system block sizes and block-Jacobi inverse rows
    pm::DeviceBuffer dense;     // dense Cholesky: n x n lower triangle, diagonal-block inverses
    pm::DeviceBuffer dia;       // diagonal layout of the solver's matrix
    pm::DeviceBuffer dia_g;     // ... of the factored preconditioner's G
    pm::DeviceBuffer dia_gt;    // ... and of G^T
    int32_t last_diagonals = 0; // diagonals of the last solve's layout (0: CSR kernels)
    int32_t last_value_streams = 0; // value arrays its row sums read (fewer than the diagonals when mirrored)
    int32_t last_form = 0;      // kernels per iteration of the last solve's form (1, 2 or 3)
    int32_t last_preconditioner = 0; // preconditioner the last solve ran (pm::Pre)
    int64_t capacity_rows = 0;

    pm::SolverState* host_state = nullptr;  // pinned, 2 slots
    hipEvent_t events[2] = {nullptr, nullptr};
    hipEvent_t done_event = nullptr;        // completion of the last enqueued work
    // second stream (default priority) for work that overlaps the bound
    // stream inside one call; every call joins it back before returning
    hipStream_t aux_stream = nullptr;
    hipEvent_t aux_events[2] = {nullptr, nullptr};

    // COO conversion in flight
    int64_t coo_rows = 0, coo_cols = 0, coo_count = 0;
    int32_t coo_dtype = 0, coo_flags = 0;
    int coo_key_bits = 0;
    const void* coo_values = nullptr;
    bool coo_prepared = false;
    bool coo_pending_status = false;
    pm::CooStatus* host_coo = nullptr;      // pinned
    unsigned long long* host_batch = nullptr;  // pinned, 2 slots: stopped-system counts

    // cached solver chunk plans
    struct PlanCache;
    std::unique_ptr<PlanCache> plans;

    int bind(hipStream_t s);
    int ensure_events();
    int ensure_aux();
    ~pm_workspace();
};
