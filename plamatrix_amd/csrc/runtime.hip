// Host runtime of the library: error state, device buffers, workspace
// lifecycle, kernel-call replay and HIP graph instantiation.
#include <cstdarg>
#include <cstring>
#include <map>
#include <string>

#include "plan_cache.hpp"
#include "workspace.hpp"

namespace pm
{

static thread_local std::string g_error;

void set_error(int code, const std::string& message)
{
    (void)code;
    g_error = message;
}

int fail(int code, const char* fmt, ...)
{
    char buffer[1024];
    va_list args;
    va_start(args, fmt);
    vsnprintf(buffer, sizeof(buffer), fmt, args);
    va_end(args);
    g_error = buffer;
    return code;
}

int check_hip(hipError_t err, const char* what, const char* file, int line)
{
    return fail(PM_DEVICE_ERROR, "HIP error %d (%s) in %s at %s:%d", static_cast<int>(err),
                hipGetErrorString(err), what, file, line);
}

int DeviceBuffer::reserve(size_t need)
{
    if (need <= bytes) return PM_OK;
    release();
    size_t grow = need + need / 8 + 256;
    grow = (grow + 255) & ~size_t(255);
    void* p = nullptr;
    hipError_t err = hipMalloc(&p, grow);
    if (err != hipSuccess)
    {
        (void)hipGetLastError();
        return fail(PM_RUNTIME_ERROR, "device allocation of %zu bytes failed: %s", grow, hipGetErrorString(err));
    }
    ptr = p;
    bytes = grow;
    return PM_OK;
}

void DeviceBuffer::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int launch(const KernelCall& k, hipStream_t stream)
{
    PM_HIP(hipLaunchKernel(k.func, k.grid, k.block, const_cast<void**>(k.args.data()), 0, stream));
    return PM_OK;
}

void Plan::reset()
{
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
}

int Plan::instantiate()
{
    reset();
    PM_HIP(hipGraphCreate(&graph, 0));
    hipGraphNode_t prev = nullptr;
    for (const KernelCall& k : calls)
    {
        hipKernelNodeParams p{};
        p.func = const_cast<void*>(k.func);
        p.gridDim = k.grid;
        p.blockDim = k.block;
        p.sharedMemBytes = 0;
        p.kernelParams = const_cast<void**>(k.args.data());
        p.extra = nullptr;
        hipGraphNode_t node = nullptr;
        PM_HIP(hipGraphAddKernelNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &p));
        prev = node;
    }
    PM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return PM_OK;
}

int Plan::run(hipStream_t stream, bool use_graph)
{
    if (use_graph)
    {
        if (!exec) PM_TRY(instantiate());
        PM_HIP(hipGraphLaunch(exec, stream));
        return PM_OK;
    }
    for (const KernelCall& k : calls) PM_TRY(launch(k, stream));
    return PM_OK;
}

} // namespace pm

int pm_workspace::bind(hipStream_t s)
{
    if (!bound)
    {
        stream = s;
        bound = true;
        PM_HIP(hipGetDevice(&device));
        return PM_OK;
    }
    if (s != stream)
        return pm::fail(PM_LOGIC_ERROR,
                        "workspace is bound to another stream; synchronize that stream and close the "
                        "workspace before reusing it on a different stream");
    int current = -1;
    PM_HIP(hipGetDevice(&current));
    if (current != device)
        return pm::fail(PM_LOGIC_ERROR, "workspace was created on device %d but device %d is current", device,
                        current);
    return PM_OK;
}

int pm_workspace::ensure_events()
{
    if (!events[0])
    {
        PM_HIP(hipEventCreateWithFlags(&events[0], hipEventDisableTiming));
        PM_HIP(hipEventCreateWithFlags(&events[1], hipEventDisableTiming));
        PM_HIP(hipEventCreateWithFlags(&done_event, hipEventDisableTiming));
    }
    if (!host_state)
    {
        // fine-grained: the solvers' publish kernels store the per-chunk
        // status records straight into it
        PM_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_state), 2 * sizeof(pm::SolverState),
                             hipHostMallocCoherent));
        PM_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_coo), sizeof(pm::CooStatus), hipHostMallocDefault));
    }
    return PM_OK;
}

int pm_workspace::ensure_aux()
{
    if (!aux_stream) PM_HIP(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
    for (hipEvent_t& e : aux_events)
        if (!e) PM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return PM_OK;
}

pm_workspace::~pm_workspace()
{
    plans.reset();
    vectors.release();
    partials.release();
    state.release();
    scratch.release();
    sort.release();
    refine.release();
    batch.release();
    dense.release();
    dia.release();
    dia_g.release();
    dia_gt.release();
    for (hipEvent_t& e : events)
        if (e) (void)hipEventDestroy(e);
    if (done_event) (void)hipEventDestroy(done_event);
    if (aux_stream)
    {
        (void)hipStreamSynchronize(aux_stream);
        (void)hipStreamDestroy(aux_stream);
    }
    for (hipEvent_t& e : aux_events)
        if (e) (void)hipEventDestroy(e);
    if (host_state) (void)hipHostFree(host_state);
    if (host_coo) (void)hipHostFree(host_coo);
    if (host_batch) (void)hipHostFree(host_batch);
}

extern "C" {

const char* pm_last_error(void) { return pm::g_error.c_str(); }

int pm_version(void) { return 100; }

int pm_workspace_create(pm_workspace** ws)
{
    if (!ws) return pm::fail(PM_INVALID_ARGUMENT, "pm_workspace_create: null output pointer");
    *ws = new pm_workspace();
    (*ws)->plans = std::make_unique<pm_workspace::PlanCache>();
    return PM_OK;
}

int pm_workspace_destroy(pm_workspace* ws)
{
    if (!ws) return PM_OK;
    if (ws->bound) (void)hipStreamSynchronize(ws->stream);
    delete ws;
    return PM_OK;
}

int pm_workspace_close(pm_workspace* ws)
{
    if (!ws) return pm::fail(PM_INVALID_ARGUMENT, "pm_workspace_close: null workspace");
    if (ws->coo_pending_status)
        return pm::fail(PM_LOGIC_ERROR, "SparseOpsWorkspace::closeAsyncAllocation requires checkStatus first");
    if (ws->bound)
    {
        // any work still queued on the bound stream may use this storage
        // (sparse_ops.cu:640-660 of the reference queries the stream too)
        const hipError_t q = hipStreamQuery(ws->stream);
        if (q == hipErrorNotReady)
            return pm::fail(PM_LOGIC_ERROR,
                            "closeAsyncAllocation requires the workspace stream to be synchronized first");
        if (q != hipSuccess) return pm::check_hip(q, "hipStreamQuery", __FILE__, __LINE__);
    }
    ws->plans = std::make_unique<pm_workspace::PlanCache>();
    ws->vectors.release();
    ws->partials.release();
    ws->state.release();
    ws->scratch.release();
    ws->sort.release();
    ws->refine.release();
    ws->batch.release();
    ws->dense.release();
    ws->dia.release();
    ws->dia_g.release();
    ws->dia_gt.release();
    ws->capacity_rows = 0;
    ws->bound = false;
    ws->stream = nullptr;
    ws->coo_prepared = false;
    return PM_OK;
}

size_t pm_workspace_capacity_bytes(const pm_workspace* ws)
{
    if (!ws) return 0;
    return ws->vectors.bytes + ws->partials.bytes + ws->state.bytes + ws->scratch.bytes + ws->sort.bytes +
           ws->refine.bytes + ws->batch.bytes + ws->dense.bytes +
           ws->dia.bytes + ws->dia_g.bytes + ws->dia_gt.bytes;
}

int64_t pm_workspace_capacity_rows(const pm_workspace* ws) { return ws ? ws->capacity_rows : 0; }

int32_t pm_workspace_solver_diagonals(const pm_workspace* ws) { return ws ? ws->last_diagonals : 0; }
int32_t pm_workspace_solver_value_streams(const pm_workspace* ws) { return ws ? ws->last_value_streams : 0; }
int32_t pm_workspace_solver_form(const pm_workspace* ws) { return ws ? ws->last_form : 0; }
int32_t pm_workspace_solver_preconditioner(const pm_workspace* ws) { return ws ? ws->last_preconditioner : 0; }

} // extern "C"
