// C++ drop-in API checks on the GPU, written like the reference's own tests
// (sparse_ops_test.cpp, iterative_solver_test.cpp, block_schur_test.cpp):
// known answers, closed forms and the exception classes.  Prints one line per
// case and "ALL PASSED" at the end; exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
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

    EXPECT(throws<std::out_of_range>([&] { eq.addResidualBlock(2, 0, ppj, &eej, &er, 1); }));
    const double inf = INFINITY;
    EXPECT(throws<std::invalid_argument>([&] { eq.addResidualBlock(0, 0, ppj, &eej, &er, 1, inf); }));
    EXPECT(throws<std::invalid_argument>([&] { BlockNormalEquations<double>(1, 1, 0, 1); }));
    EXPECT(std::fabs(evaluateHuberLoss(25.0, 3.0).weight - 0.6) == 0.0);
    EXPECT(evaluateHuberLoss(25.0, 3.0).cost == 10.5);
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
