// C++ drop-in API checks on the GPU, written like the reference's own tests
// (sparse_ops_test.cpp, iterative_solver_test.cpp, block_schur_test.cpp):
// known answers, closed forms and the exception classes.  Prints one line per
// case and "ALL PASSED" at the end; exits non-zero on the first failure.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <thread>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "plamatrix/optimization/block_schur.h"
#include "plamatrix/optimization/levenberg_marquardt.h"
#include "plamatrix/sparse/coo_matrix.h"
#include "plamatrix/sparse/iterative_solver.h"
#include "plamatrix/sparse/sparse_ops.h"

using namespace plamatrix;

static int g_failures = 0;

#define EXPECT(cond)                                                                  \
    do                                                                                \
    {                                                                                 \
        if (!(cond))                                                                  \
        {                                                                             \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);           \
            ++g_failures;                                                             \
        }                                                                             \
    } while (0)

template <typename E, typename F>
static bool throws(F&& f)
{
    try
    {
        f();
    }
    catch (const E&)
    {
        return true;
    }
    catch (...)
    {
        return false;
    }
    return false;
}

template <typename Scalar>
static DenseMatrix<Scalar, Device::GPU> upload(const std::vector<Scalar>& v)
{
    DenseMatrix<Scalar, Device::CPU> h(static_cast<Index>(v.size()), 1);
    for (std::size_t i = 0; i < v.size(); ++i) h.data()[i] = v[i];
    return h.toGpu();
}

template <typename Scalar>
static std::vector<Scalar> download(const DenseMatrix<Scalar, Device::GPU>& d)
{
    auto h = d.toCpu();
    return std::vector<Scalar>(h.data(), h.data() + h.size());
}

// sparse_ops_test.cpp:320-369: left fold of duplicates in input order.
template <typename Scalar>
static void cooCancellation(Scalar large)
{
    COOMatrix<Scalar, Device::GPU> coo(4, 5);
    const Index rows[] = {2, 0, 2, 1, 2, 0, 2};
    const Index cols[] = {4, 3, 1, 2, 1, 3, 1};
    const Scalar vals[] = {5, 2, large, 4, -large, -2, 1};
    for (int i = 0; i < 7; ++i) coo.add(rows[i], cols[i], vals[i]);
    auto csr = coo.toCsr().toCpu();
    const std::vector<Index> ro(csr.rowOffsets(), csr.rowOffsets() + 5);
    const std::vector<Index> ci(csr.colIndices(), csr.colIndices() + csr.nnz());
    const std::vector<Scalar> v(csr.values(), csr.values() + csr.nnz());
    EXPECT((ro == std::vector<Index>{0, 1, 2, 4, 4}));
    EXPECT((ci == std::vector<Index>{3, 2, 1, 4}));
    EXPECT((v == std::vector<Scalar>{0, 4, 1, 5}));
    EXPECT(throws<std::out_of_range>([&] { coo.add(4, 0, Scalar(1)); }));
}

// sparse_ops_test.cpp:636-670
template <typename Scalar>
static void spmvKat()
{
    COOMatrix<Scalar, Device::GPU> coo(3, 4);
    coo.add(2, 3, 5);
    coo.add(0, 2, 2);
    coo.add(1, 1, -1);
    coo.add(0, 0, 3);
    auto a = coo.toCsr();
    auto x = upload<Scalar>({1, 2, 3, 4});
    const auto y = download(spmv(a, x));
    EXPECT((y == std::vector<Scalar>{9, -2, 20}));
    DenseMatrix<Scalar, Device::GPU> wrong(2, 1);
    EXPECT(throws<std::runtime_error>([&] { spmv(a, x, wrong); }));
}

// sparse_ops_test.cpp:196-233 and 704-737: column-major SpMM, output reuse.
template <typename Scalar>
static void spmmKat()
{
    COOMatrix<Scalar, Device::GPU> coo(3, 3);
    coo.add(1, 1, 4);
    coo.add(0, 2, 2);
    coo.add(2, 0, -1);
    coo.add(0, 0, 3);
    auto a = coo.toCsr();
    auto b = upload<Scalar>({1, 2, 3, 10, 20, 30});  // 3 x 2, column-major
    DenseMatrix<Scalar, Device::GPU> bm(3, 2);
    detail::checkHip(hipMemcpy(bm.data(), b.data(), 6 * sizeof(Scalar), hipMemcpyDeviceToDevice), "hipMemcpy");
    const auto y = download(spmm(a, bm));
    EXPECT((y == std::vector<Scalar>{9, 8, -1, 90, 80, -10}));
    DenseMatrix<Scalar, Device::GPU> reuse(3, 2);
    Scalar* before = reuse.data();
    spmm(a, bm, reuse);
    EXPECT(reuse.data() == before && download(reuse) == y);
    DenseMatrix<Scalar, Device::GPU> wrong(2, 2), square(3, 3);
    EXPECT(throws<std::runtime_error>([&] { spmm(a, bm, wrong); }));
    SparseOpsWorkspace ws;
    EXPECT(throws<std::invalid_argument>([&] { spmmAsync(a, square, square, ws); }));
    auto x3 = upload<Scalar>({1, 2, 3});
    EXPECT(throws<std::invalid_argument>([&] { spmvAsync(a, x3, x3, ws); }));
}

// diag(2, 4) with the int64 arrays read back through const access.
template <typename Scalar>
static bool isDiag24(const CSRMatrix<Scalar, Device::GPU>& a)
{
    Index ro[3] = {}, ci[2] = {};
    Scalar v[2] = {};
    detail::checkHip(hipMemcpy(ro, a.rowOffsets(), sizeof(ro), hipMemcpyDeviceToHost), "hipMemcpy");
    detail::checkHip(hipMemcpy(ci, a.colIndices(), sizeof(ci), hipMemcpyDeviceToHost), "hipMemcpy");
    detail::checkHip(hipMemcpy(v, a.values(), sizeof(v), hipMemcpyDeviceToHost), "hipMemcpy");
    return ro[0] == 0 && ro[1] == 1 && ro[2] == 2 && ci[0] == 0 && ci[1] == 1 && v[0] == Scalar(2) &&
           v[1] == Scalar(4);
}

template <typename Scalar>
static CSRMatrix<Scalar, Device::GPU> poisson1d(Index n)
{
    COOMatrix<Scalar, Device::GPU> coo(n, n);
    for (Index r = 0; r < n; ++r)
    {
        if (r > 0) coo.add(r, r - 1, -1);
        coo.add(r, r, 2);
        if (r + 1 < n) coo.add(r, r + 1, -1);
    }
    return coo.toCsr();
}

// iterative_solver_test.cpp:782-820: poisson1d(8), b = 1 -> 4 iterations,
// x_i = (i + 1)(8 - i) / 2.
template <typename Scalar>
static void pcgClosedForm()
{
    auto a = poisson1d<Scalar>(8);
    auto b = upload<Scalar>(std::vector<Scalar>(8, 1));
    DenseMatrix<Scalar, Device::GPU> x(8, 1);
    IterativeSolverWorkspace<Scalar> ws;
    IterativeSolverOptions o;
    o.relativeTolerance = std::is_same_v<Scalar, double> ? 1e-12 : 1e-6;
    const auto rep = pcg(a, b, x, ws, o);
    EXPECT(rep.converged);
    EXPECT(rep.iterations == 4);
    const auto got = download(x);
    const double tol = std::is_same_v<Scalar, double> ? 1e-10 : 1e-4;
    for (int i = 0; i < 8; ++i) EXPECT(std::fabs(got[i] - (i + 1) * (8 - i) / 2.0) <= tol);

    // fixed-iteration asynchronous submission of the same 4 steps
    DenseMatrix<Scalar, Device::GPU> x2(8, 1);
    auto state = pcgFixedIterationsAsync(a, b, x2, 4, ws);
    (void)hipDeviceSynchronize();
    const auto frep = finalizeIterativeSolverReport(state, o);
    EXPECT(frep.iterations == 4);
    const auto got2 = download(x2);
    for (int i = 0; i < 8; ++i) EXPECT(std::fabs(got2[i] - (i + 1) * (8 - i) / 2.0) <= tol);

    // requireConvergence turns the iteration limit into runtime_error
    IterativeSolverOptions strict;
    strict.maxIterations = 1;
    strict.relativeTolerance = 0.0;
    strict.requireConvergence = true;
    DenseMatrix<Scalar, Device::GPU> x3(8, 1);
    EXPECT(throws<std::runtime_error>([&] { cg(a, b, x3, ws, strict); }));
    IterativeSolverOptions bad;
    bad.relativeTolerance = 2.0;
    EXPECT(throws<std::invalid_argument>([&] { pcg(a, b, x3, ws, bad); }));
}

// iterative_solver_test.cpp:394-424: exact block inverses converge in 1 step.
static void blockPcgOneIteration()
{
    COOMatrix<double, Device::GPU> coo(4, 4);
    const Index r[] = {0, 0, 1, 1, 2, 2, 3, 3}, c[] = {0, 1, 0, 1, 2, 3, 2, 3};
    const double v[] = {4, 1, 1, 3, 2, 0.5, 0.5, 1.5};
    for (int i = 0; i < 8; ++i) coo.add(r[i], c[i], v[i]);
    auto a = coo.toCsr();
    auto inv = upload<double>({3 / 11.0, -1 / 11.0, -1 / 11.0, 4 / 11.0, 1.5 / 2.75, -0.5 / 2.75, -0.5 / 2.75,
                               2 / 2.75});
    auto b = upload<double>({4 * 1 + 1 * 2, 1 * 1 + 3 * 2, 2 * -1 + 0.5 * 3, 0.5 * -1 + 1.5 * 3});
    DenseMatrix<double, Device::GPU> x(4, 1);
    IterativeSolverWorkspace<double> ws;
    IterativeSolverOptions o;
    o.relativeTolerance = 1e-12;
    const auto rep = blockPcg(a, b, x, inv, 2, ws, o);
    EXPECT(rep.converged && rep.iterations == 1);
    const auto got = download(x);
    const double expected[] = {1, 2, -1, 3};
    for (int i = 0; i < 4; ++i) EXPECT(std::fabs(got[i] - expected[i]) <= 1e-10);
}

// FP32 matrix, FP64 refinement to 1e-8 on a 2-D Laplacian.
static void refinement()
{
    const Index nx = 64, n = nx * nx;
    COOMatrix<float, Device::GPU> coo(n, n);
    for (Index i = 0; i < n; ++i)
    {
        const Index x = i % nx, y = i / nx;
        if (y > 0) coo.add(i, i - nx, -1);
        if (x > 0) coo.add(i, i - 1, -1);
        coo.add(i, i, 4);
        if (x + 1 < nx) coo.add(i, i + 1, -1);
        if (y + 1 < nx) coo.add(i, i + nx, -1);
    }
    auto a = coo.toCsr();
    std::vector<double> bh(static_cast<std::size_t>(n));
    for (Index i = 0; i < n; ++i) bh[static_cast<std::size_t>(i)] = std::sin(0.37 * static_cast<double>(i));
    auto b = upload<double>(bh);
    DenseMatrix<double, Device::GPU> x(n, 1);
    IterativeSolverWorkspace<float> ws;
    const auto rep = pcgRefine(a, b, x, ws, 1e-8, 10);
    EXPECT(rep.converged);
    EXPECT(rep.finalRelres <= 1e-8);
}

// FP32 batch with FP64 refinement per system: the same two Poisson systems,
// closed-form solutions to 1e-10 although the matrix is FP32.
static void batchedRefinement()
{
    COOMatrix<float, Device::GPU> coo(69, 69);
    for (Index r = 0; r < 8; ++r)
    {
        if (r > 0) coo.add(r, r - 1, -1);
        coo.add(r, r, 2);
        if (r + 1 < 8) coo.add(r, r + 1, -1);
    }
    for (Index r = 0; r < 5; ++r)
    {
        if (r > 0) coo.add(64 + r, 64 + r - 1, -2);
        coo.add(64 + r, 64 + r, 4);
        if (r + 1 < 5) coo.add(64 + r, 64 + r + 1, -2);
    }
    auto a = coo.toCsr();
    std::vector<double> bh(69, 0.0);
    for (int i = 0; i < 8; ++i) bh[static_cast<std::size_t>(i)] = 1.0;
    for (int i = 64; i < 69; ++i) bh[static_cast<std::size_t>(i)] = 2.0;
    auto b = upload(bh);
    DenseMatrix<double, Device::GPU> x(69, 1);
    IterativeSolverWorkspace<float> ws;
    const auto reps = batchedPcgRefine(a, {0, 64}, {8, 69}, b, x, ws, 1e-12, 10);
    EXPECT(reps.size() == 2 && reps[0].converged && reps[1].converged);
    EXPECT(reps[0].final_relres <= 1e-12 && reps[1].final_relres <= 1e-12);
    const auto got = download(x);
    for (int i = 0; i < 8; ++i) EXPECT(std::fabs(got[static_cast<std::size_t>(i)] - (i + 1) * (8 - i) / 2.0) <= 1e-10);
    for (int i = 0; i < 5; ++i) EXPECT(std::fabs(got[static_cast<std::size_t>(64 + i)] - (i + 1) * (5 - i) / 2.0) <= 1e-10);
}

// block_schur_test.cpp:550-569, then a short LM loop over a linear
// least-squares problem: damped device Schur solves with the damping the
// strategy picks, the pattern built once for the whole loop.
static void levenbergMarquardt()
{
    LevenbergMarquardtOptions<double> lo;
    lo.initialDamping = 1e-3;
    lo.minimumDamping = 1e-8;
    lo.maximumDamping = 1e4;
    lo.decreaseFactor = 0.25;
    lo.increaseFactor = 8.0;
    LevenbergMarquardtStrategy<double> kat(lo);
    kat.acceptStep();
    EXPECT(kat.damping() == 2.5e-4 && kat.acceptedSteps() == 1 && kat.rejectedSteps() == 0);
    kat.rejectStep();
    EXPECT(kat.damping() == 2e-3 && kat.acceptedSteps() == 1 && kat.rejectedSteps() == 1);
    lo.decreaseFactor = 1.5;
    EXPECT(throws<std::invalid_argument>([&] { LevenbergMarquardtStrategy<double> bad(lo); }));

    // r_i(x) = pj_i . x_p + ej_i x_e + c_i for 2 primaries (size 2), 2 eliminated (size 1)
    struct Res
    {
        Index p, e;
        double pj[2], ej, c;
    };
    const Res res[] = {{0, 0, {1.0, 2.0}, 0.5, -1.0},  {1, 0, {-0.5, 1.0}, 1.5, 0.25},
                       {0, 1, {2.0, -1.0}, -0.75, 2.0}, {1, 1, {1.25, 0.5}, 2.0, -0.5},
                       {0, 0, {0.3, -0.4}, 1.1, 0.7},   {1, 1, {-0.9, 0.2}, -0.6, 1.3}};
    std::vector<double> xp(4, 0.0), xe(2, 0.0);
    auto residual = [&](const Res& q, const std::vector<double>& p, const std::vector<double>& e) {
        return q.pj[0] * p[static_cast<std::size_t>(2 * q.p)] + q.pj[1] * p[static_cast<std::size_t>(2 * q.p + 1)] +
               q.ej * e[static_cast<std::size_t>(q.e)] + q.c;
    };
    auto cost = [&](const std::vector<double>& p, const std::vector<double>& e) {
        double c = 0.0;
        for (const auto& q : res)
        {
            const double r = residual(q, p, e);
            c += r * r;
        }
        return c;
    };
    SchurComplementSolverOptions<double> o;
    o.maxIterations = 50;
    o.relativeTolerance = 1e-13;
    o.absoluteTolerance = 1e-14;
    SchurComplementSolverWorkspace<double> ws;
    LevenbergMarquardtStrategy<double> lm;
    double current = cost(xp, xe);
    const double start = current;
    for (int it = 0; it < 12; ++it)
    {
        BlockNormalEquations<double> eq(2, 2, 2, 1);
        for (const auto& q : res)
        {
            const double r = residual(q, xp, xe);
            eq.addResidualBlock(q.p, q.e, q.pj, &q.ej, &r, 1, 1.0);
        }
        std::vector<double> dp, de;
        const auto rep = solveDampedSchurComplement(eq, lm.damping(), o, ws, &dp, &de);
        EXPECT(rep.converged);
        std::vector<double> tp(xp), te(xe);
        for (std::size_t i = 0; i < tp.size(); ++i) tp[i] += dp[i];
        for (std::size_t i = 0; i < te.size(); ++i) te[i] += de[i];
        const double trial = cost(tp, te);
        if (std::isfinite(trial) && trial < current)
        {
            xp = tp;
            xe = te;
            current = trial;
            lm.acceptStep();
        }
        else
        {
            lm.rejectStep();
        }
    }
    EXPECT(ws.patternBuildCount() == 1);
    EXPECT(lm.acceptedSteps() >= 1 && current < 0.5 * start);
    // at the least-squares minimum the gradient J^T r vanishes
    double g[6] = {0, 0, 0, 0, 0, 0};
    for (const auto& q : res)
    {
        const double r = residual(q, xp, xe);
        g[2 * q.p] += q.pj[0] * r;
        g[2 * q.p + 1] += q.pj[1] * r;
        g[4 + q.e] += q.ej * r;
    }
    for (double v : g) EXPECT(std::fabs(v) <= 1e-6);
}

// ---- async transfer and stream contract (csr_matrix_test.cpp:184-450,
// sparse_ops_test.cpp:838-950, iterative_solver_test.cpp:562-680) ----------

struct Stream
{
    hipStream_t s = nullptr;
    Stream() { detail::checkHip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags"); }
    ~Stream()
    {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    void sync() const { detail::checkHip(hipStreamSynchronize(s), "hipStreamSynchronize"); }
};

// 3 x 4 with nnz 4: rows {0: (0, 1.5) (3, -2)}, {1: (1, 3.25)}, {2: (2, 4)}.
template <typename Scalar>
static void fillTransferInput(CSRMatrix<Scalar, Device::CPU>& m)
{
    const Index ro[] = {0, 2, 3, 4};
    const Index ci[] = {0, 3, 1, 2};
    const Scalar v[] = {Scalar(1.5), Scalar(-2), Scalar(3.25), Scalar(4)};
    for (int i = 0; i < 4; ++i) m.rowOffsets()[i] = ro[i];
    for (int i = 0; i < 4; ++i)
    {
        m.colIndices()[i] = ci[i];
        m.values()[i] = v[i];
    }
    m.validateStructure();
}

template <typename Scalar>
static bool isTransferInput(const CSRMatrix<Scalar, Device::CPU>& m)
{
    const Index ro[] = {0, 2, 3, 4};
    const Index ci[] = {0, 3, 1, 2};
    const Scalar v[] = {Scalar(1.5), Scalar(-2), Scalar(3.25), Scalar(4)};
    if (m.rows() != 3 || m.cols() != 4 || m.nnz() != 4) return false;
    for (int i = 0; i < 4; ++i)
        if (m.rowOffsets()[i] != ro[i] || m.colIndices()[i] != ci[i] || m.values()[i] != v[i]) return false;
    return true;
}

template <typename Scalar>
static void csrTransfers()
{
    using Cpu = CSRMatrix<Scalar, Device::CPU>;
    using Gpu = CSRMatrix<Scalar, Device::GPU>;
    Cpu pageable(3, 4, 4);
    fillTransferInput(pageable);
    auto pinned = Cpu::pinned(3, 4, 4);
    fillTransferInput(pinned);
    EXPECT(pinned.isPinnedHost() && !pageable.isPinnedHost());

    // synchronous round trip; the upload is trusted
    {
        const auto gpu = pageable.toGpu();
        EXPECT(gpu.hasValidatedStructure());
        EXPECT(isTransferInput(gpu.toCpu()));
    }
    // asynchronous round trip on a non-default stream
    {
        Stream st;
        auto gpu = pinned.toGpuAsync(st.s);
        auto back = gpu.toCpuAsync(st.s);
        EXPECT(gpu.isAsyncAllocation() && gpu.asyncAllocationStream() == st.s);
        EXPECT(back.isPinnedHost());
        st.sync();
        gpu.closeAsyncAllocation();
        st.sync();
        EXPECT(isTransferInput(back));
    }
    // async copies keep the output's existing storage
    {
        Gpu gpu(3, 4, 4);
        auto back = Cpu::pinned(3, 4, 4);
        const Scalar* gv = static_cast<const Gpu&>(gpu).values();
        const Scalar* hv = static_cast<const Cpu&>(back).values();
        Stream st;
        pinned.copyToGpuAsync(gpu, st.s);
        gpu.copyToCpuAsync(back, st.s);
        st.sync();
        EXPECT(static_cast<const Gpu&>(gpu).values() == gv && static_cast<const Cpu&>(back).values() == hv);
        EXPECT(isTransferInput(back));
    }
    // a reused pinned download is revalidated after an async overwrite
    {
        auto valid = pageable.toGpu();
        auto invalid = pageable.toGpu();
        const Index bad[] = {0, 3, 2, 4};
        detail::checkHip(hipMemcpy(invalid.rowOffsets(), bad, sizeof(bad), hipMemcpyHostToDevice), "hipMemcpy");
        Stream st;
        auto reused = valid.toCpuAsync(st.s);
        st.sync();
        reused.validateStructure(st.s);
        EXPECT(reused.hasValidatedStructure());
        invalid.copyToCpuAsync(reused, st.s);
        st.sync();
        EXPECT(throws<std::invalid_argument>([&] { reused.validateStructure(st.s); }));
    }
    // pageable host storage cannot take part in async copies
    {
        Gpu gpu(3, 4, 4);
        EXPECT(throws<std::invalid_argument>([&] { pageable.copyToGpuAsync(gpu); }));
        EXPECT(throws<std::invalid_argument>([&] { (void)pageable.toGpuAsync(); }));
        const auto src = pageable.toGpu();
        Cpu out(3, 4, 4);
        EXPECT(throws<std::invalid_argument>([&] { src.copyToCpuAsync(out); }));
        Cpu wrong(3, 4, 3);
        EXPECT(throws<std::runtime_error>([&] { src.copyToCpuAsync(wrong); }));
    }
    // closeAsyncAllocation is for stream-ordered storage only
    {
        Gpu ordinary(3, 4, 4);
        EXPECT(!ordinary.isAsyncAllocation());
        EXPECT(throws<std::logic_error>([&] { ordinary.closeAsyncAllocation(); }));
    }
    // stream-ordered storage rejects copies on another stream
    {
        Stream owner, other;
        auto gpu = pinned.toGpuAsync(owner.s);
        EXPECT(throws<std::logic_error>([&] { (void)gpu.toCpuAsync(other.s); }));
        EXPECT(throws<std::logic_error>([&] { pinned.copyToGpuAsync(gpu, other.s); }));
        EXPECT(throws<std::logic_error>([&] { (void)gpu.toCpu(); }));
        owner.sync();
        gpu.closeAsyncAllocation();
        owner.sync();
    }
    // ordinary storage rejects overlapping writes from two streams
    {
        Gpu gpu(3, 4, 4);
        Stream first, second;
        pinned.copyToGpuAsync(gpu, first.s);
        EXPECT(!gpu.hasValidatedStructure() && gpu.isStructureUsableOnStream(first.s));
        EXPECT(throws<std::logic_error>([&] { pinned.copyToGpuAsync(gpu, second.s); }));
        EXPECT(throws<std::logic_error>([&] { gpu.validateStructure(second.s); }));
        first.sync();
        gpu.validateStructure(first.s);
        EXPECT(gpu.hasValidatedStructure());
        pinned.copyToGpuAsync(gpu, second.s);
        second.sync();
    }
    // a download may not overlap a pending upload from another stream
    {
        Gpu gpu(3, 4, 4);
        auto out = Cpu::pinned(3, 4, 4);
        Stream first, second;
        pinned.copyToGpuAsync(gpu, first.s);
        EXPECT(throws<std::logic_error>([&] { gpu.copyToCpuAsync(out, second.s); }));
        first.sync();
        gpu.validateStructure(first.s);
        gpu.copyToCpuAsync(out, second.s);
        second.sync();
        out.validateStructure(second.s);
        EXPECT(out.hasValidatedStructure() && isTransferInput(out));
    }
    // two downloads into one pinned matrix from two streams
    {
        const auto gpu = pageable.toGpu();
        auto out = Cpu::pinned(3, 4, 4);
        Stream first, second;
        gpu.copyToCpuAsync(out, first.s);
        EXPECT(throws<std::logic_error>([&] { gpu.copyToCpuAsync(out, second.s); }));
        first.sync();
        out.validateStructure(first.s);
        gpu.copyToCpuAsync(out, second.s);
        second.sync();
    }
    // ownership of stream-ordered storage moves; repeated close is harmless
    {
        Stream st;
        auto source = pinned.toGpuAsync(st.s);
        const Scalar* v = static_cast<const Gpu&>(source).values();
        Gpu moved(std::move(source));
        EXPECT(!source.isAsyncAllocation() && source.rows() == 0);
        EXPECT(moved.isAsyncAllocation() && static_cast<const Gpu&>(moved).values() == v);
        st.sync();
        moved.closeAsyncAllocation();
        moved.closeAsyncAllocation();
        EXPECT(moved.rows() == 0 && moved.nnz() == 0);
        st.sync();
    }
    // move assignment releases the target's storage and takes the source's
    {
        Stream st;
        auto source = pinned.toGpuAsync(st.s);
        auto target = pinned.toGpuAsync(st.s);
        const Scalar* v = static_cast<const Gpu&>(source).values();
        target = std::move(source);
        EXPECT(!source.isAsyncAllocation() && target.isAsyncAllocation());
        EXPECT(static_cast<const Gpu&>(target).values() == v);
        st.sync();
        target.closeAsyncAllocation();
        st.sync();
    }
    // dense stream-ordered storage
    {
        Stream st;
        auto d = DenseMatrix<Scalar, Device::GPU>::uninitializedAsync(4, 1, st.s);
        EXPECT(d.isAsyncAllocation() && d.asyncAllocationStream() == st.s);
        auto h = DenseMatrix<Scalar, Device::CPU>::pinned(4, 1);
        EXPECT(h.isPinnedHost());
        for (Index i = 0; i < 4; ++i) h(i, 0) = Scalar(i + 1);
        h.copyToGpuAsync(d, st.s);
        auto back = d.toCpuAsync(st.s);
        st.sync();
        EXPECT(back(3, 0) == Scalar(4) && back(0, 0) == Scalar(1));
        d.closeAsyncAllocation();
        d.closeAsyncAllocation();
        EXPECT(d.rows() == 0 && !d.isAsyncAllocation());
        DenseMatrix<Scalar, Device::GPU> ordinary(2, 1);
        EXPECT(throws<std::logic_error>([&] { ordinary.closeAsyncAllocation(); }));
        DenseMatrix<Scalar, Device::CPU> wrong(3, 1);
        EXPECT(throws<std::runtime_error>([&] { ordinary.copyToCpuAsync(wrong); }));
        st.sync();
    }
}

template <typename Scalar>
static void sparseOpsStreamRules()
{
    // products reject stream-ordered operands on another stream
    {
        auto csr = CSRMatrix<Scalar, Device::CPU>::pinned(2, 2, 2);
        csr.rowOffsets()[1] = 1;
        csr.rowOffsets()[2] = 2;
        csr.colIndices()[1] = 1;
        csr.values()[0] = Scalar(2);
        csr.values()[1] = Scalar(3);
        auto in = DenseMatrix<Scalar, Device::CPU>::pinned(2, 1);
        in(0, 0) = Scalar(4);
        in(1, 0) = Scalar(5);
        Stream owner, other;
        auto a = csr.toGpuAsync(owner.s);
        auto x = DenseMatrix<Scalar, Device::GPU>::uninitializedAsync(2, 1, owner.s);
        auto y = DenseMatrix<Scalar, Device::GPU>::uninitializedAsync(2, 1, owner.s);
        in.copyToGpuAsync(x, owner.s);
        SparseOpsWorkspace ws;
        EXPECT(throws<std::logic_error>([&] { spmvAsync(a, x, y, ws, other.s); }));
        spmvAsync(a, x, y, ws, owner.s);
        auto out = y.toCpuAsync(owner.s);
        owner.sync();
        EXPECT(out(0, 0) == Scalar(8) && out(1, 0) == Scalar(15));
        ws.closeAsyncAllocation();
        a.closeAsyncAllocation();
        x.closeAsyncAllocation();
        y.closeAsyncAllocation();
        owner.sync();
    }
    // a workspace is bound to its first stream until closed; moves carry it
    {
        COOMatrix<Scalar, Device::GPU> coo(2, 2);
        coo.add(0, 0, 2);
        coo.add(1, 1, 3);
        const auto a = coo.toCsr();
        const auto x = upload<Scalar>({4, 5});
        DenseMatrix<Scalar, Device::GPU> y(2, 1);
        Stream first, second;
        SparseOpsWorkspace ws;
        spmvAsync(a, x, y, ws, first.s);
        first.sync();
        SparseOpsWorkspace moved(std::move(ws));
        EXPECT(ws.capacityBytes() == 0);
        EXPECT(throws<std::logic_error>([&] { spmvAsync(a, x, y, moved, second.s); }));
        moved.closeAsyncAllocation();
        moved.closeAsyncAllocation();
        spmvAsync(a, x, y, moved, second.s);
        second.sync();
        EXPECT((download(y) == std::vector<Scalar>{8, 15}));
        moved.closeAsyncAllocation();
    }
    // closing while the bound stream still has work is a logic_error
    {
        COOMatrix<Scalar, Device::GPU> coo(1, 1);
        const auto a = coo.toCsr();
        const auto x = upload<Scalar>({3});
        DenseMatrix<Scalar, Device::GPU> y(1, 1);
        // destruction order: the guard releases the host function, the
        // workspace and the stream drain it, then the flag goes
        std::atomic<bool> release{false};
        Stream st;
        SparseOpsWorkspace ws;
        struct Releaser
        {
            std::atomic<bool>& flag;
            ~Releaser() { flag.store(true, std::memory_order_release); }
        } guard{release};
        spmvAsync(a, x, y, ws, st.s);
        detail::checkHip(hipLaunchHostFunc(
                             st.s,
                             [](void* p) {
                                 auto* flag = static_cast<std::atomic<bool>*>(p);
                                 while (!flag->load(std::memory_order_acquire)) std::this_thread::yield();
                             },
                             &release),
                         "hipLaunchHostFunc");
        EXPECT(throws<std::logic_error>([&] { ws.closeAsyncAllocation(); }));
        release.store(true, std::memory_order_release);
        st.sync();
        ws.closeAsyncAllocation();
    }
    // escaped mutable storage is no longer trusted by SpMV
    {
        auto a = poisson1d<Scalar>(3);
        (void)a.values();
        a.validateStructure();
        const auto x = upload<Scalar>({1, 1, 1});
        DenseMatrix<Scalar, Device::GPU> y(3, 1);
        EXPECT(!a.hasValidatedStructure());
        EXPECT(throws<std::logic_error>([&] { spmv(a, x, y); }));
    }
    // an async conversion's output is trusted once its status is checked
    {
        const auto r = upload<Index>({1, 0, 1});
        const auto c = upload<Index>({1, 0, 1});
        const auto v = upload<Scalar>({1, 2, 3});
        Stream st;
        CSRMatrix<Scalar, Device::GPU> out(2, 2, 2);
        SparseOpsWorkspace ws;
        cooToCsrAsync(2, 2, r, c, v, out, ws, st.s);
        EXPECT(!out.hasValidatedStructure() && out.isStructureUsableOnStream(nullptr) == false);
        st.sync();
        ws.checkStatus();
        EXPECT(out.hasValidatedStructure());
        EXPECT(isDiag24(out));
    }
}

template <typename Scalar>
static void solverTrustRules()
{
    // a device COO result feeds the fixed-iteration solver directly
    {
        const auto r = upload<Index>({0, 1});
        const auto c = upload<Index>({0, 1});
        const auto v = upload<Scalar>({2, 4});
        SparseOpsWorkspace cws;
        const auto a = cooToCsr<Scalar>(2, 2, r, c, v, cws);
        EXPECT(a.hasValidatedStructure());
        const auto b = upload<Scalar>({2, 4});
        DenseMatrix<Scalar, Device::GPU> x(2, 1);
        IterativeSolverWorkspace<Scalar> ws;
        auto state = pcgFixedIterationsAsync(a, b, x, 1, ws);
        detail::checkHip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        EXPECT(finalizeIterativeSolverReport(state).converged);
        state.closeAsyncAllocation();
        state.closeAsyncAllocation();
        ws.closeAsyncAllocation();
        cws.closeAsyncAllocation();
    }
    // malformed structure written through a raw pointer is caught before any launch
    {
        auto a = poisson1d<Scalar>(2);
        const Index bad[] = {0, 2, 1};
        detail::checkHip(hipMemcpy(a.rowOffsets(), bad, sizeof(bad), hipMemcpyHostToDevice), "hipMemcpy");
        const auto b = upload<Scalar>({1, 1});
        DenseMatrix<Scalar, Device::GPU> x(2, 1);
        IterativeSolverWorkspace<Scalar> ws;
        EXPECT(throws<std::invalid_argument>([&] { cg(a, b, x, ws); }));
        EXPECT((download(x) == std::vector<Scalar>{0, 0}));
    }
    // escaped storage is rejected by the fixed-iteration solvers even after validation
    {
        auto a = poisson1d<Scalar>(2);
        (void)a.rowOffsets();
        a.validateStructure();
        const auto b = upload<Scalar>({1, 1});
        DenseMatrix<Scalar, Device::GPU> x(2, 1);
        IterativeSolverWorkspace<Scalar> ws;
        EXPECT(throws<std::logic_error>([&] { (void)cgFixedIterationsAsync(a, b, x, 1, ws); }));
        // the adaptive solver revalidates and solves: [2 -1; -1 2] x = 1 -> x = 1
        const auto rep = cg(a, b, x, ws);
        EXPECT(rep.converged);
        const auto xv = download(x);
        EXPECT(std::fabs(xv[0] - 1) < 1e-5 && std::fabs(xv[1] - 1) < 1e-5);
        EXPECT(throws<std::invalid_argument>([&] { cg(a, b, const_cast<DenseMatrix<Scalar, Device::GPU>&>(b), ws); }));
    }
    // an async upload is consumed on its copy stream after validation there
    {
        auto src = CSRMatrix<Scalar, Device::CPU>::pinned(2, 2, 2);
        src.rowOffsets()[1] = 1;
        src.rowOffsets()[2] = 2;
        src.colIndices()[1] = 1;
        src.values()[0] = Scalar(2);
        src.values()[1] = Scalar(4);
        CSRMatrix<Scalar, Device::GPU> a(2, 2, 2);
        Stream copy, other;
        src.copyToGpuAsync(a, copy.s);
        EXPECT(throws<std::logic_error>([&] { a.validateStructure(other.s); }));
        copy.sync();
        EXPECT(isDiag24(a));
        a.validateStructure(copy.s);
        const auto b = upload<Scalar>({2, 4});
        DenseMatrix<Scalar, Device::GPU> x(2, 1);
        IterativeSolverWorkspace<Scalar> ws;
        auto state = pcgFixedIterationsAsync(a, b, x, 1, ws, copy.s);
        copy.sync();
        EXPECT(finalizeIterativeSolverReport(state).converged);
        EXPECT((download(x) == std::vector<Scalar>{1, 1}));
        state.closeAsyncAllocation();
        ws.closeAsyncAllocation();
    }
}

// Two systems in one batched launch.
static void batched()
{
    // block-diagonal: poisson1d(8) at rows [0, 8), 2 * poisson1d(5) at [64, 69)
    COOMatrix<double, Device::GPU> coo(69, 69);
    for (Index r = 0; r < 8; ++r)
    {
        if (r > 0) coo.add(r, r - 1, -1);
        coo.add(r, r, 2);
        if (r + 1 < 8) coo.add(r, r + 1, -1);
    }
    for (Index r = 0; r < 5; ++r)
    {
        if (r > 0) coo.add(64 + r, 64 + r - 1, -2);
        coo.add(64 + r, 64 + r, 4);
        if (r + 1 < 5) coo.add(64 + r, 64 + r + 1, -2);
    }
    auto a = coo.toCsr();
    std::vector<double> bh(69, 0.0);
    for (int i = 0; i < 8; ++i) bh[static_cast<std::size_t>(i)] = 1.0;
    for (int i = 64; i < 69; ++i) bh[static_cast<std::size_t>(i)] = 2.0;
    auto b = upload(bh);
    DenseMatrix<double, Device::GPU> x(69, 1);
    IterativeSolverWorkspace<double> ws;
    IterativeSolverOptions o;
    o.relativeTolerance = 1e-12;
    const auto reps = batchedPcg(a, {0, 64}, {8, 69}, b, x, ws, o);
    EXPECT(reps.size() == 2 && reps[0].converged && reps[1].converged);
    EXPECT(reps[0].iterations == 4);
    const auto got = download(x);
    for (int i = 0; i < 8; ++i) EXPECT(std::fabs(got[static_cast<std::size_t>(i)] - (i + 1) * (8 - i) / 2.0) <= 1e-10);
    for (int i = 0; i < 5; ++i) EXPECT(std::fabs(got[static_cast<std::size_t>(64 + i)] - (i + 1) * (5 - i) / 2.0) <= 1e-10);
}

// block_schur_test.cpp:41-75, 198-233 and 77-115.
static void schur()
{
    BlockNormalEquations<double> eq(2, 2, 2, 1);
    const struct
    {
        Index p, e;
        double pj[2], ej, r;
    } res[] = {{0, 0, {1.0, 2.0}, 0.5, -1.0},
               {1, 0, {-0.5, 1.0}, 1.5, 0.25},
               {0, 1, {2.0, -1.0}, -0.75, 2.0},
               {1, 1, {1.25, 0.5}, 2.0, -0.5}};
    for (const auto& q : res) eq.addResidualBlock(q.p, q.e, q.pj, &q.ej, &q.r, 1, 1.0);
    const double ppj[2] = {0.4, -0.2}, pr = 0.75;
    eq.addPrimaryResidualBlock(0, ppj, &pr, 1, 2.0);
    const double eej = 0.8, er = -0.3;
    eq.addEliminatedResidualBlock(1, &eej, &er, 1, 1.5);
    SchurComplementSolverOptions<double> o;
    o.maxIterations = 50;
    o.relativeTolerance = 1e-13;
    o.absoluteTolerance = 1e-14;
    SchurComplementSolverWorkspace<double> ws;
    std::vector<double> dp, de;
    const auto rep = solveDampedSchurComplement(eq, 0.1, o, ws, &dp, &de);
    EXPECT(rep.converged && rep.schurAssemblyOnDevice);
    const double expected[] = {-0.532196868252671, 0.717080672008074, 0.078633780070238, -0.210823022930113,
                               0.017900956933130, 0.253200998620430};
    for (int i = 0; i < 4; ++i) EXPECT(std::fabs(dp[static_cast<std::size_t>(i)] - expected[i]) <= 1e-11);
    for (int i = 0; i < 2; ++i) EXPECT(std::fabs(de[static_cast<std::size_t>(i)] - expected[4 + i]) <= 1e-11);
    // same topology: the pattern is reused (block_schur_test.cpp:160-196)
    const auto rep2 = solveDampedSchurComplement(eq, 0.1, o, ws, &dp, &de);
    EXPECT(rep2.schurPatternReused && ws.patternBuildCount() == 1);

    BlockNormalEquations<double> multi(2, 1, 1, 1);
    const double j0 = 1.0, j1 = 2.0, je = 3.0, r = 4.0;
    multi.addResidualBlocks({0, 1}, {&j0, &j1}, 0, &je, &r, 1);
    const double k0 = 0.5, k1 = -1.0, r2 = 0.2;
    multi.addPrimaryResidualBlocks({0, 1}, {&k0, &k1}, &r2, 1);
    const auto rep3 = solveDampedSchurComplement(multi, 0.1, o, &dp, &de);
    EXPECT(rep3.converged);
    EXPECT(std::fabs(dp[0] + 1.345185185185185) <= 1e-11 && std::fabs(dp[1] + 0.512592592592593) <= 1e-11);
    EXPECT(std::fabs(de[0] + 0.493827160493827) <= 1e-11);

    // the Cpu (default) and Hip backends on the synthetic and multi-primary
    // systems (block_schur_test.cpp:77-158 via expectSyntheticReference /
    // expectMultiPrimaryReference)
    for (const auto backend : {SchurComplementLinearBackend::Cpu, SchurComplementLinearBackend::Hip})
    {
        auto ob = o;
        ob.linearBackend = backend;
        const auto rb = solveDampedSchurComplement(eq, 0.1, ob, &dp, &de);
        if (!rb.converged) std::printf("  backend %d: %s\n", static_cast<int>(backend), rb.message.c_str());
        EXPECT(rb.converged);
        EXPECT(rb.linearBackend == backend);
        EXPECT(!rb.deviceName.empty());
        EXPECT(rb.linearSolveSeconds > 0.0);
        for (int i = 0; i < 4; ++i) EXPECT(std::fabs(dp[static_cast<std::size_t>(i)] - expected[i]) <= 1e-10);
        for (int i = 0; i < 2; ++i) EXPECT(std::fabs(de[static_cast<std::size_t>(i)] - expected[4 + i]) <= 1e-10);
        const auto rm = solveDampedSchurComplement(multi, 0.1, ob, &dp, &de);
        EXPECT(rm.converged);
        EXPECT(std::fabs(dp[0] + 1.345185185185185) <= 1e-11 && std::fabs(dp[1] + 0.512592592592593) <= 1e-11);
        EXPECT(std::fabs(de[0] + 0.493827160493827) <= 1e-11);
    }
    EXPECT(SchurComplementSolverOptions<double>{}.linearBackend == SchurComplementLinearBackend::Cpu);
    {
        // block_schur_test.cpp:444-466: useMixedPrecision converges to the FP64
        // answer (the FP32 seed converges on this well-conditioned system and
        // seeds the FP64 solve)
        auto om = o;
        om.linearBackend = SchurComplementLinearBackend::Hip;
        om.relativeTolerance = 1e-12;
        om.useMixedPrecision = true;
        const auto rm = solveDampedSchurComplement(eq, 0.1, om, &dp, &de);
        EXPECT(rm.converged);
        EXPECT(std::fabs(dp[0] - expected[0]) <= 1e-10 && std::fabs(de[0] - expected[4]) <= 1e-10);
        EXPECT(rm.mixedPrecisionUsed);
    }
    {
        // block_schur_test.cpp:531-548: float point-only system
        BlockNormalEquations<float> fe(0, 1, 6, 1);
        const float jf = 2.0f, rf = 4.0f;
        fe.addEliminatedResidualBlock(0, &jf, &rf, 1);
        SchurComplementSolverOptions<float> of;
        std::vector<float> fp, fde;
        const auto rf_rep = solveDampedSchurComplement(fe, 0.5f, of, &fp, &fde);
        EXPECT(rf_rep.converged && fp.empty() && fde.size() == 1);
        EXPECT(std::fabs(fde[0] + 4.0f / 3.0f) <= 1e-6f);
    }
    {
        // the synthetic system accumulated and solved with float blocks
        BlockNormalEquations<float> fe(2, 2, 2, 1);
        const struct
        {
            Index p, e;
            float pj[2], ej, r;
        } fres[] = {{0, 0, {1.0f, 2.0f}, 0.5f, -1.0f},
                    {1, 0, {-0.5f, 1.0f}, 1.5f, 0.25f},
                    {0, 1, {2.0f, -1.0f}, -0.75f, 2.0f},
                    {1, 1, {1.25f, 0.5f}, 2.0f, -0.5f}};
        for (const auto& q : fres) fe.addResidualBlock(q.p, q.e, q.pj, &q.ej, &q.r, 1, 1.0f);
        const float fpj[2] = {0.4f, -0.2f}, fpr = 0.75f;
        fe.addPrimaryResidualBlock(0, fpj, &fpr, 1, 2.0f);
        const float fej = 0.8f, fer = -0.3f;
        fe.addEliminatedResidualBlock(1, &fej, &fer, 1, 1.5f);
        SchurComplementSolverOptions<float> of;
        of.relativeTolerance = 1e-6f;
        std::vector<float> fp, fde;
        const auto rf_rep = solveDampedSchurComplement(fe, 0.1f, of, &fp, &fde);
        EXPECT(rf_rep.converged && fp.size() == 4 && fde.size() == 2);
        for (int i = 0; i < 4; ++i) EXPECT(std::fabs(fp[static_cast<std::size_t>(i)] - expected[i]) <= 1e-5);
        for (int i = 0; i < 2; ++i) EXPECT(std::fabs(fde[static_cast<std::size_t>(i)] - expected[4 + i]) <= 1e-5);
    }
    {
        auto bad = o;
        bad.linearBackend = SchurComplementLinearBackend::Cuda;
        EXPECT(throws<std::invalid_argument>([&] { solveDampedSchurComplement(eq, 0.1, bad, &dp, &de); }));
        bad.linearBackend = SchurComplementLinearBackend::OpenCl;
        EXPECT(throws<std::invalid_argument>([&] { solveDampedSchurComplement(eq, 0.1, bad, &dp, &de); }));
    }

    EXPECT(throws<std::out_of_range>([&] { eq.addResidualBlock(2, 0, ppj, &eej, &er, 1); }));
    const double inf = INFINITY;
    EXPECT(throws<std::invalid_argument>([&] { eq.addResidualBlock(0, 0, ppj, &eej, &er, 1, inf); }));
    EXPECT(throws<std::invalid_argument>([&] { BlockNormalEquations<double>(1, 1, 0, 1); }));
    EXPECT(std::fabs(evaluateHuberLoss(25.0, 3.0).weight - 0.6) == 0.0);
    EXPECT(evaluateHuberLoss(25.0, 3.0).cost == 10.5);
}

// block_schur_test.cpp:235-405: the DenseCpu backend (served on the device)
// on the synthetic and multi-primary systems, a coupled chain of 24 primary
// blocks solved twice bitwise equal, an eliminated-heavy system against the
// PCG backend, and a singular reduced system.
static void denseSchur()
{
    SchurComplementSolverOptions<double> o;
    o.linearBackend = SchurComplementLinearBackend::DenseCpu;
    o.relativeTolerance = 1e-12;
    o.absoluteTolerance = 1e-14;
    {
        BlockNormalEquations<double> eq(2, 2, 2, 1);
        const struct
        {
            Index p, e;
            double pj[2], ej, r;
        } res[] = {{0, 0, {1.0, 2.0}, 0.5, -1.0},
                   {1, 0, {-0.5, 1.0}, 1.5, 0.25},
                   {0, 1, {2.0, -1.0}, -0.75, 2.0},
                   {1, 1, {1.25, 0.5}, 2.0, -0.5}};
        for (const auto& q : res) eq.addResidualBlock(q.p, q.e, q.pj, &q.ej, &q.r, 1, 1.0);
        const double ppj[2] = {0.4, -0.2}, pr = 0.75;
        eq.addPrimaryResidualBlock(0, ppj, &pr, 1, 2.0);
        const double eej = 0.8, er = -0.3;
        eq.addEliminatedResidualBlock(1, &eej, &er, 1, 1.5);
        SchurComplementSolverWorkspace<double> ws;
        std::vector<double> dp, de;
        const auto rep = solveDampedSchurComplement(eq, 0.1, o, ws, &dp, &de);
        EXPECT(rep.converged && rep.iterations == 1);
        EXPECT(rep.linearBackend == SchurComplementLinearBackend::DenseCpu);
        EXPECT(rep.linearSolveSeconds >= rep.choleskyFactorizationSeconds + rep.triangularSolveSeconds +
                                             rep.residualCheckSeconds);
        const double expected[] = {-0.532196868252671, 0.717080672008074, 0.078633780070238, -0.210823022930113,
                                   0.017900956933130, 0.253200998620430};
        for (int i = 0; i < 4; ++i) EXPECT(std::fabs(dp[static_cast<std::size_t>(i)] - expected[i]) <= 1e-11);
        for (int i = 0; i < 2; ++i) EXPECT(std::fabs(de[static_cast<std::size_t>(i)] - expected[4 + i]) <= 1e-11);
        const auto again = solveDampedSchurComplement(eq, 0.1, o, ws, &dp, &de);
        EXPECT(again.converged && again.schurPatternReused && ws.patternBuildCount() == 1);
    }
    {
        BlockNormalEquations<double> multi(2, 1, 1, 1);
        const double j0 = 1.0, j1 = 2.0, je = 3.0, r = 4.0;
        multi.addResidualBlocks({0, 1}, {&j0, &j1}, 0, &je, &r, 1);
        const double k0 = 0.5, k1 = -1.0, r2 = 0.2;
        multi.addPrimaryResidualBlocks({0, 1}, {&k0, &k1}, &r2, 1);
        std::vector<double> dp, de;
        auto dense = o;
        dense.linearBackend = SchurComplementLinearBackend::DenseHip;
        const auto rep = solveDampedSchurComplement(multi, 0.1, dense, &dp, &de);
        EXPECT(rep.converged);
        EXPECT(std::fabs(dp[0] + 1.345185185185185) <= 1e-11 && std::fabs(dp[1] + 0.512592592592593) <= 1e-11);
        EXPECT(std::fabs(de[0] + 0.493827160493827) <= 1e-11);
    }
    {
        constexpr Index blocks = 24, size = 6;
        BlockNormalEquations<double> eq(blocks, 0, size, 1);
        double identity[size * size] = {};
        for (Index d = 0; d < size; ++d) identity[d * size + d] = 1.0;
        for (Index b = 0; b < blocks; ++b)
        {
            double r[size];
            for (Index c = 0; c < size; ++c) r[c] = 0.01 * static_cast<double>(1 + b + c);
            eq.addPrimaryResidualBlock(b, identity, r, size);
        }
        const double left[size] = {1.0, 0.2, -0.1, 0.3, 0.05, -0.2};
        const double right[size] = {-0.8, 0.1, 0.25, -0.2, 0.15, 0.3};
        for (Index b = 0; b + 1 < blocks; ++b)
        {
            const double r = 0.02 * static_cast<double>(b + 1);
            eq.addPrimaryResidualBlocks({b, b + 1}, {left, right}, &r, 1);
        }
        std::vector<double> p1, e1, p2, e2;
        const auto first = solveDampedSchurComplement(eq, 0.01, o, &p1, &e1);
        const auto second = solveDampedSchurComplement(eq, 0.01, o, &p2, &e2);
        EXPECT(first.converged && second.converged);
        EXPECT(p1.size() == static_cast<std::size_t>(blocks * size) && e1.empty());
        EXPECT(p1 == p2 && e1 == e2);
        for (double v : p1) EXPECT(std::isfinite(v));
    }
    {
        constexpr Index pc = 36, ec = 240, ps = 9, es = 3;
        BlockNormalEquations<double> eq(pc, ec, ps, es);
        double identity[ps * ps] = {};
        for (Index d = 0; d < ps; ++d) identity[d * ps + d] = 1.0;
        const double zero[ps] = {};
        for (Index p = 0; p < pc; ++p) eq.addPrimaryResidualBlock(p, identity, zero, ps, 0.25);
        for (Index e = 0; e < ec; ++e)
            for (Index obs = 0; obs < 4; ++obs)
            {
                const Index p = (e * 7 + obs * 5) % pc;
                double pj[2 * ps], ej[2 * es];
                for (Index i = 0; i < 2 * ps; ++i) pj[i] = 0.01 * static_cast<double>(1 + (i + e + obs) % 13);
                for (Index i = 0; i < 2 * es; ++i) ej[i] = 0.02 * static_cast<double>(1 + (i + e + 2 * obs) % 7);
                const double r[2] = {0.001 * static_cast<double>(e + 1), -0.002 * static_cast<double>(obs + 1)};
                eq.addResidualBlock(p, e, pj, ej, r, 2);
            }
        auto dense = o;
        dense.relativeTolerance = 1e-11;
        dense.absoluteTolerance = 1e-13;
        std::vector<double> dp, de, pp, pe;
        const auto rep = solveDampedSchurComplement(eq, 0.1, dense, &dp, &de);
        auto pcg = dense;
        pcg.linearBackend = SchurComplementLinearBackend::Hip;
        pcg.maxIterations = 500;
        const auto prep = solveDampedSchurComplement(eq, 0.1, pcg, &pp, &pe);
        EXPECT(rep.converged && prep.converged);
        double diff = 0, norm = 0;
        for (std::size_t i = 0; i < dp.size(); ++i)
        {
            diff = std::max(diff, std::fabs(dp[i] - pp[i]));
            norm = std::max(norm, std::fabs(pp[i]));
        }
        for (std::size_t i = 0; i < de.size(); ++i) diff = std::max(diff, std::fabs(de[i] - pe[i]));
        EXPECT(diff <= 1e-9 * std::max(1.0, norm));
    }
    {
        // primary 1 has no residual and there is no damping: S = diag(s0, 0)
        BlockNormalEquations<double> eq(2, 0, 1, 1);
        const double j = 2.0, r = 1.0;
        eq.addPrimaryResidualBlock(0, &j, &r, 1);
        std::vector<double> dp, de;
        const auto rep = solveDampedSchurComplement(eq, 0.0, o, &dp, &de);
        EXPECT(!rep.converged && rep.message == "reduced Schur complement is not positive definite");
    }
}

int main()
{
    const struct
    {
        const char* name;
        void (*fn)();
    } cases[] = {
        {"coo cancellation float", [] { cooCancellation<float>(1e8f); }},
        {"coo cancellation double", [] { cooCancellation<double>(1e16); }},
        {"spmv kat float", [] { spmvKat<float>(); }},
        {"spmv kat double", [] { spmvKat<double>(); }},
        {"spmm kat float", [] { spmmKat<float>(); }},
        {"spmm kat double", [] { spmmKat<double>(); }},
        {"pcg closed form double", [] { pcgClosedForm<double>(); }},
        {"pcg closed form float", [] { pcgClosedForm<float>(); }},
        {"block pcg one iteration", blockPcgOneIteration},
        {"fp64 refinement", refinement},
        {"batched pcg", batched},
        {"batched fp64 refinement", batchedRefinement},
        {"levenberg-marquardt loop", levenbergMarquardt},
        {"schur", schur},
        {"dense schur", denseSchur},
        {"csr transfers float", [] { csrTransfers<float>(); }},
        {"csr transfers double", [] { csrTransfers<double>(); }},
        {"sparse ops stream rules float", [] { sparseOpsStreamRules<float>(); }},
        {"sparse ops stream rules double", [] { sparseOpsStreamRules<double>(); }},
        {"solver trust rules float", [] { solverTrustRules<float>(); }},
        {"solver trust rules double", [] { solverTrustRules<double>(); }},
    };
    for (const auto& c : cases)
    {
        const int before = g_failures;
        try
        {
            c.fn();
        }
        catch (const std::exception& e)
        {
            std::printf("  unexpected exception: %s\n", e.what());
            ++g_failures;
        }
        std::printf("%s %s\n", g_failures == before ? "ok  " : "FAIL", c.name);
        std::fflush(stdout);
    }
    if (g_failures)
    {
        std::printf("%d FAILURES\n", g_failures);
        return 1;
    }
    std::printf("ALL PASSED\n");
    return 0;
}
