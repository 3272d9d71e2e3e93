// Host overloads of the C++ drop-in API (cooToCsr on std::vector triplets,
// spmv / spmm on CPU matrices, cg / pcg on CPU matrices), checked like the
// reference's own CPU tests: sparse_ops_test.cpp:21-292 and
// iterative_solver_test.cpp:103-335 (known answers, closed forms, exception
// classes).  No GPU call is made, so it runs on any host.
//
//   test_host_api               run the cases, print "ALL PASSED"
//   test_host_api dump IN OUT   convert / multiply / solve the system in IN
//                               and write the results to OUT (compared with
//                               the oracle by tests/test_host_api.py)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

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

static bool near(double a, double b, double tol) { return std::fabs(a - b) <= tol; }

// sparse_ops_test.cpp:21-45: stable order, left fold from the first value
static void cooFold()
{
    const auto csr = cooToCsr(4, 5, std::vector<Index>{2, 0, 2, 1, 2, 0, 2},
                              std::vector<Index>{4, 3, 1, 2, 1, 3, 1},
                              std::vector<double>{5.0, 2.0, 1.0e16, 4.0, -1.0e16, -2.0, 1.0});
    EXPECT(csr.nnz() == 4);
    const Index ro[] = {0, 1, 2, 4, 4};
    const Index ci[] = {3, 2, 1, 4};
    const double v[] = {0.0, 4.0, 1.0, 5.0};
    for (int i = 0; i < 5; ++i) EXPECT(csr.rowOffsets()[i] == ro[i]);
    for (int i = 0; i < 4; ++i) EXPECT(csr.colIndices()[i] == ci[i] && csr.values()[i] == v[i]);
    EXPECT(csr.hasValidatedStructure());
}

// :47-71: conversion into existing storage keeps the arrays
static void cooReuse()
{
    CSRMatrix<float, Device::CPU> out(2, 2, 2);
    float* v0 = out.values();
    Index* c0 = out.colIndices();
    Index* r0 = out.rowOffsets();
    cooToCsr(2, 2, std::vector<Index>{1, 0, 1}, std::vector<Index>{0, 1, 0}, std::vector<float>{2.f, 3.f, 4.f}, out);
    EXPECT(out.values() == v0 && out.colIndices() == c0 && out.rowOffsets() == r0);
    EXPECT(out.rowOffsets()[0] == 0 && out.rowOffsets()[1] == 1 && out.rowOffsets()[2] == 2);
    EXPECT(out.colIndices()[0] == 1 && out.values()[0] == 3.f);
    EXPECT(out.colIndices()[1] == 0 && out.values()[1] == 6.f);
}

// :73-84 and :86-118
static void cooEmptyAndErrors()
{
    const std::vector<Index> none;
    const std::vector<float> no_values;
    auto empty = cooToCsr(0, 5, none, none, no_values);
    EXPECT(empty.rows() == 0 && empty.cols() == 5 && empty.nnz() == 0 && empty.rowOffsets() != nullptr &&
           empty.rowOffsets()[0] == 0);
    const std::vector<Index> one{0};
    const std::vector<float> one_value{1.f};
    EXPECT(throws<std::invalid_argument>([&] { cooToCsr(-1, 1, none, none, no_values); }));
    EXPECT(throws<std::invalid_argument>([&] { cooToCsr(1, -1, none, none, no_values); }));
    EXPECT(throws<std::invalid_argument>([&] { cooToCsr(1, 1, one, none, one_value); }));
    EXPECT(throws<std::invalid_argument>([&] { cooToCsr(1, 1, one, one, no_values); }));
    EXPECT(throws<std::out_of_range>([&] { cooToCsr(1, 1, std::vector<Index>{-1}, one, one_value); }));
    EXPECT(throws<std::out_of_range>([&] { cooToCsr(1, 1, one, std::vector<Index>{-1}, one_value); }));
    EXPECT(throws<std::out_of_range>([&] { cooToCsr(1, 1, std::vector<Index>{1}, one, one_value); }));
    EXPECT(throws<std::out_of_range>([&] { cooToCsr(1, 1, one, std::vector<Index>{1}, one_value); }));
    const Index beyond = static_cast<Index>(std::numeric_limits<int32_t>::max()) + 1;
    EXPECT(throws<std::out_of_range>([&] { cooToCsr(2, 2, std::vector<Index>{beyond}, one, one_value); }));
    CSRMatrix<float, Device::CPU> wrong_dims(1, 2, 1);
    EXPECT(throws<std::runtime_error>([&] { cooToCsr(2, 2, one, one, one_value, wrong_dims); }));
    CSRMatrix<float, Device::CPU> wrong_nnz(1, 1, 2);
    EXPECT(throws<std::runtime_error>([&] { cooToCsr(1, 1, one, one, one_value, wrong_nnz); }));
}

// :120-193
static void spmvKats()
{
    auto csr = cooToCsr(3, 4, std::vector<Index>{2, 0, 1, 0}, std::vector<Index>{3, 2, 1, 0},
                        std::vector<float>{5.f, 2.f, -1.f, 3.f});
    DenseMatrix<float, Device::CPU> x(4, 1);
    for (int i = 0; i < 4; ++i) x(i, 0) = static_cast<float>(i + 1);
    auto y = spmv(csr, x);
    EXPECT(y.rows() == 3 && y.cols() == 1 && y(0, 0) == 9.f && y(1, 0) == -2.f && y(2, 0) == 20.f);
    DenseMatrix<float, Device::CPU> out(3, 1);
    out.fill(-100.f);
    float* before = out.data();
    spmv(csr, x, out);
    EXPECT(out.data() == before && out(0, 0) == 9.f && out(1, 0) == -2.f && out(2, 0) == 20.f);

    auto sparse = cooToCsr(3, 2, std::vector<Index>{1}, std::vector<Index>{0}, std::vector<double>{2.0});
    DenseMatrix<double, Device::CPU> xd(2, 1);
    xd(0, 0) = 3.0;
    xd(1, 0) = 4.0;
    auto yd = spmv(sparse, xd);
    EXPECT(yd(0, 0) == 0.0 && yd(1, 0) == 6.0 && yd(2, 0) == 0.0);
    DenseMatrix<double, Device::CPU> wrong_rows(3, 1), wrong_cols(2, 2), wrong_out(2, 1);
    EXPECT(throws<std::runtime_error>([&] { spmv(sparse, wrong_rows); }));
    EXPECT(throws<std::runtime_error>([&] { spmv(sparse, wrong_cols); }));
    EXPECT(throws<std::runtime_error>([&] { spmv(sparse, xd, wrong_out); }));

    auto diag = cooToCsr(2, 2, std::vector<Index>{0, 1}, std::vector<Index>{0, 1}, std::vector<float>{2.f, 3.f});
    DenseMatrix<float, Device::CPU> same(2, 1);
    EXPECT(throws<std::invalid_argument>([&] { spmv(diag, same, same); }));
}

// :196-268
static void spmmKats()
{
    auto csr = cooToCsr(3, 3, std::vector<Index>{1, 0, 2, 0}, std::vector<Index>{1, 2, 0, 0},
                        std::vector<double>{4.0, 2.0, -1.0, 3.0});
    DenseMatrix<double, Device::CPU> B(3, 2);
    const double b[] = {1, 2, 3, 10, 20, 30};
    for (int i = 0; i < 6; ++i) B.data()[i] = b[i];
    auto y = spmm(csr, B);
    const double want[] = {9, 8, -1, 90, 80, -10};
    for (int i = 0; i < 6; ++i) EXPECT(y.data()[i] == want[i]);
    DenseMatrix<double, Device::CPU> out(3, 2);
    out.fill(-100.0);
    double* before = out.data();
    spmm(csr, B, out);
    EXPECT(out.data() == before);
    for (int i = 0; i < 6; ++i) EXPECT(out.data()[i] == want[i]);
    EXPECT(throws<std::invalid_argument>([&] { spmm(csr, B, B); }));

    const std::vector<Index> none;
    auto empty = cooToCsr(0, 3, none, none, std::vector<float>{});
    DenseMatrix<float, Device::CPU> Bf(3, 2), wrong_B(2, 2), wrong_out(1, 2);
    auto ye = spmm(empty, Bf);
    EXPECT(ye.rows() == 0 && ye.cols() == 2);
    EXPECT(throws<std::runtime_error>([&] { spmm(empty, wrong_B); }));
    EXPECT(throws<std::runtime_error>([&] { spmm(empty, Bf, wrong_out); }));
}

// :270-292
static void cooMatrixCpu()
{
    COOMatrix<float, Device::CPU> m(2, 3);
    m.add(1, 2, 1.5f);
    m.add(0, 1, 2.0f);
    m.add(1, 2, 2.5f);
    auto csr = m.toCsr();
    EXPECT(csr.nnz() == 2 && csr.rowOffsets()[1] == 1 && csr.rowOffsets()[2] == 2);
    EXPECT(csr.colIndices()[0] == 1 && csr.values()[0] == 2.f && csr.colIndices()[1] == 2 && csr.values()[1] == 4.f);
}

template <typename S>
static CSRMatrix<S, Device::CPU> poisson1d(Index n, S scale = S(1))
{
    std::vector<Index> r, c;
    std::vector<S> v;
    for (Index i = 0; i < n; ++i)
    {
        if (i > 0)
        {
            r.push_back(i);
            c.push_back(i - 1);
            v.push_back(-scale);
        }
        r.push_back(i);
        c.push_back(i);
        v.push_back(S(2) * scale);
        if (i + 1 < n)
        {
            r.push_back(i);
            c.push_back(i + 1);
            v.push_back(-scale);
        }
    }
    return cooToCsr(n, n, r, c, v);
}

template <typename S>
static double tolerance()
{
    return std::is_same_v<S, float> ? 2.0e-4 : 1.0e-10;
}

// iterative_solver_test.cpp:103-123
template <typename S>
static void pcgDiagonal()
{
    auto m = cooToCsr(3, 3, std::vector<Index>{0, 1, 2}, std::vector<Index>{0, 1, 2}, std::vector<S>{2, 4, 8});
    DenseMatrix<S, Device::CPU> b(3, 1), x(3, 1);
    b(0, 0) = 2;
    b(1, 0) = 8;
    b(2, 0) = 24;
    x.fill(0);
    auto rep = pcg(m, b, x);
    EXPECT(rep.converged && rep.iterations == 1 && rep.finalResidual <= rep.initialResidual);
    for (int i = 0; i < 3; ++i) EXPECT(near(x(i, 0), i + 1, tolerance<S>()));
}

// :125-155
template <typename S>
static void poissonScales()
{
    for (const S scale : {S(1.0e-3), S(1), S(1.0e3)})
    {
        auto m = poisson1d<S>(8, scale);
        DenseMatrix<S, Device::CPU> expected(8, 1);
        for (Index i = 0; i < 8; ++i) expected(i, 0) = S(i + 1);
        auto b = spmv(m, expected);
        DenseMatrix<S, Device::CPU> xc(8, 1), xp(8, 1);
        xc.fill(0);
        xp.fill(0);
        IterativeSolverOptions o;
        o.relativeTolerance = std::is_same_v<S, float> ? 1.0e-5 : 1.0e-12;
        auto rc = cg(m, b, xc, o);
        auto rp = pcg(m, b, xp, o);
        EXPECT(rc.converged && rp.converged);
        for (Index i = 0; i < 8; ++i)
        {
            EXPECT(near(xc(i, 0), expected(i, 0), tolerance<S>() * (i + 1)));
            EXPECT(near(xp(i, 0), expected(i, 0), tolerance<S>() * (i + 1)));
        }
    }
}

// :157-176
template <typename S>
static void zeroIterations()
{
    auto m = poisson1d<S>(4);
    DenseMatrix<S, Device::CPU> b(4, 1), x(4, 1);
    b.fill(0);
    x.fill(0);
    auto r0 = cg(m, b, x);
    EXPECT(r0.converged && r0.iterations == 0 && r0.initialResidual == 0.0 && r0.finalResidual == 0.0);
    x.fill(2);
    auto exact = spmv(m, x);
    auto r1 = pcg(m, exact, x);
    EXPECT(r1.converged && r1.iterations == 0);
}

// :178-286
template <typename S>
static void solverErrors()
{
    auto non_square = cooToCsr(2, 3, std::vector<Index>{0}, std::vector<Index>{0}, std::vector<S>{1});
    DenseMatrix<S, Device::CPU> b(2, 1), x(2, 1);
    EXPECT(throws<std::invalid_argument>([&] { cg(non_square, b, x); }));
    auto sq = poisson1d<S>(2);
    DenseMatrix<S, Device::CPU> wrong_b(3, 1), wrong_x(2, 2);
    EXPECT(throws<std::invalid_argument>([&] { cg(sq, wrong_b, x); }));
    EXPECT(throws<std::invalid_argument>([&] { cg(sq, b, wrong_x); }));
    IterativeSolverOptions o;
    o.maxIterations = -1;
    EXPECT(throws<std::invalid_argument>([&] { cg(sq, b, x, o); }));
    o.maxIterations = 10;
    o.relativeTolerance = -1.0;
    EXPECT(throws<std::invalid_argument>([&] { cg(sq, b, x, o); }));
    o.relativeTolerance = 1.0e100;
    EXPECT(throws<std::invalid_argument>([&] { cg(sq, b, x, o); }));
    o.relativeTolerance = std::numeric_limits<double>::quiet_NaN();
    EXPECT(throws<std::invalid_argument>([&] { cg(sq, b, x, o); }));
    o.relativeTolerance = 1.0e-6;
    o.absoluteTolerance = std::numeric_limits<double>::infinity();
    EXPECT(throws<std::invalid_argument>([&] { cg(sq, b, x, o); }));

    b.fill(1);
    x.fill(0);
    auto missing = cooToCsr(2, 2, std::vector<Index>{0, 1}, std::vector<Index>{1, 0}, std::vector<S>{1, 1});
    EXPECT(throws<std::runtime_error>([&] { pcg(missing, b, x); }));
    auto zero = cooToCsr(2, 2, std::vector<Index>{0, 1}, std::vector<Index>{0, 1}, std::vector<S>{1, 0});
    EXPECT(throws<std::runtime_error>([&] { pcg(zero, b, x); }));

    const S small = std::is_same_v<S, float> ? S(1.0e-7F) : S(1.0e-20);
    auto tiny = cooToCsr(2, 2, std::vector<Index>{0, 1}, std::vector<Index>{0, 1}, std::vector<S>{small, S(2) * small});
    b(0, 0) = small;
    b(1, 0) = S(2) * small;
    x.fill(0);
    auto rs = pcg(tiny, b, x);
    EXPECT(rs.converged && rs.iterations == 1 && near(x(0, 0), 1, tolerance<S>()) && near(x(1, 0), 1, tolerance<S>()));

    CSRMatrix<S, Device::CPU> dup(2, 2, 4);
    const Index dro[] = {0, 2, 4}, dci[] = {0, 0, 1, 1};
    const S dv[] = {1, 1, 1, 3};
    for (int i = 0; i < 3; ++i) dup.rowOffsets()[i] = dro[i];
    for (int i = 0; i < 4; ++i)
    {
        dup.colIndices()[i] = dci[i];
        dup.values()[i] = dv[i];
    }
    b(0, 0) = 2;
    b(1, 0) = 4;
    x.fill(0);
    auto rd = pcg(dup, b, x);
    EXPECT(rd.converged && rd.iterations == 1);

    b.fill(1);
    x.fill(0);
    auto negative = cooToCsr(2, 2, std::vector<Index>{0, 1}, std::vector<Index>{0, 1}, std::vector<S>{-1, 1});
    EXPECT(throws<std::runtime_error>([&] { pcg(negative, b, x); }));
    CSRMatrix<S, Device::CPU> bad_offsets(2, 2, 2);
    bad_offsets.rowOffsets()[0] = 0;
    bad_offsets.rowOffsets()[1] = 2;
    bad_offsets.rowOffsets()[2] = 1;
    EXPECT(throws<std::invalid_argument>([&] { cg(bad_offsets, b, x); }));
    CSRMatrix<S, Device::CPU> bad_column(2, 2, 2);
    const Index bro[] = {0, 1, 2}, bci[] = {0, 2};
    for (int i = 0; i < 3; ++i) bad_column.rowOffsets()[i] = bro[i];
    for (int i = 0; i < 2; ++i)
    {
        bad_column.colIndices()[i] = bci[i];
        bad_column.values()[i] = 1;
    }
    EXPECT(throws<std::invalid_argument>([&] { cg(bad_column, b, x); }));
}

// :288-335
template <typename S>
static void limitsAndTolerances()
{
    auto m = poisson1d<S>(4);
    DenseMatrix<S, Device::CPU> b(4, 1), x(4, 1);
    b.fill(1);
    x.fill(0);
    IterativeSolverOptions o;
    o.maxIterations = 0;
    o.relativeTolerance = 0.0;
    auto r0 = cg(m, b, x, o);
    EXPECT(!r0.converged && r0.iterations == 0);
    o.maxIterations = 10;
    o.absoluteTolerance = 2.0;
    auto ra = cg(m, b, x, o);
    EXPECT(ra.converged && ra.iterations == 0);
    o.absoluteTolerance = 0.0;
    o.relativeTolerance = std::is_same_v<S, float> ? 1.0e-5 : 1.0e-12;
    o.useJacobiPreconditioner = false;
    EXPECT(pcg(m, b, x, o).converged);

    auto big = poisson1d<S>(16);
    DenseMatrix<S, Device::CPU> b16(16, 1), x16(16, 1);
    b16.fill(1);
    x16.fill(0);
    IterativeSolverOptions one;
    one.maxIterations = 1;
    one.relativeTolerance = 0.0;
    auto r1 = cg(big, b16, x16, one);
    EXPECT(!r1.converged && r1.iterations == 1 && std::isfinite(r1.initialResidual) && std::isfinite(r1.finalResidual));
    x16.fill(0);
    one.requireConvergence = true;
    EXPECT(throws<std::runtime_error>([&] { cg(big, b16, x16, one); }));
}

// :782-820: poisson1d(8) with rhs 1 takes 4 Jacobi-PCG iterations to
// x_i = (i + 1)(8 - i) / 2
template <typename S>
static void poissonClosedForm()
{
    auto m = poisson1d<S>(8);
    DenseMatrix<S, Device::CPU> b(8, 1), x(8, 1);
    b.fill(1);
    x.fill(0);
    IterativeSolverOptions o;
    o.relativeTolerance = std::is_same_v<S, float> ? 1.0e-5 : 1.0e-12;
    auto rep = pcg(m, b, x, o);
    EXPECT(rep.converged && rep.iterations == 4);
    for (Index i = 0; i < 8; ++i) EXPECT(near(x(i, 0), (i + 1) * (8 - i) / 2.0, tolerance<S>() * (i + 1)));
}

// ---- dump mode ---------------------------------------------------------------------

struct Reader
{
    FILE* f;
    template <typename T>
    T get()
    {
        T v{};
        if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("short input");
        return v;
    }
    template <typename T>
    std::vector<T> many(std::size_t n)
    {
        std::vector<T> v(n);
        if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short input");
        return v;
    }
};

template <typename T>
static void put(FILE* f, const T* p, std::size_t n)
{
    if (n) std::fwrite(p, sizeof(T), n, f);
}

template <typename S>
static void dumpSystem(Reader& in, Index rows, Index cols, Index count, FILE* out)
{
    const auto r = in.many<Index>(static_cast<std::size_t>(count));
    const auto c = in.many<Index>(static_cast<std::size_t>(count));
    const auto v = in.many<S>(static_cast<std::size_t>(count));
    const auto xin = in.many<S>(static_cast<std::size_t>(cols));
    const auto bin = in.many<S>(static_cast<std::size_t>(rows));
    const int64_t max_iterations = in.get<int64_t>();
    const double rel = in.get<double>();
    auto csr = cooToCsr(rows, cols, r, c, v);
    const Index nnz = csr.nnz();
    put(out, &nnz, 1);
    put(out, csr.rowOffsets(), static_cast<std::size_t>(rows) + 1);
    put(out, csr.colIndices(), static_cast<std::size_t>(nnz));
    put(out, csr.values(), static_cast<std::size_t>(nnz));
    DenseMatrix<S, Device::CPU> x(cols, 1);
    for (Index i = 0; i < cols; ++i) x.data()[i] = xin[static_cast<std::size_t>(i)];
    auto y = spmv(csr, x);
    put(out, y.data(), static_cast<std::size_t>(rows));
    if (rows != cols) return;
    DenseMatrix<S, Device::CPU> b(rows, 1);
    for (Index i = 0; i < rows; ++i) b.data()[i] = bin[static_cast<std::size_t>(i)];
    for (int pre = 0; pre < 2; ++pre)
    {
        DenseMatrix<S, Device::CPU> sol(rows, 1);
        sol.fill(0);
        IterativeSolverOptions o;
        o.maxIterations = static_cast<int>(max_iterations);
        o.relativeTolerance = rel;
        const auto rep = pre ? pcg(csr, b, sol, o) : cg(csr, b, sol, o);
        const int64_t head[2] = {rep.converged ? 1 : 0, rep.iterations};
        const double res[2] = {rep.initialResidual, rep.finalResidual};
        put(out, head, 2);
        put(out, res, 2);
        put(out, sol.data(), static_cast<std::size_t>(rows));
    }
}

static int dump(const char* in_path, const char* out_path)
{
    FILE* fin = std::fopen(in_path, "rb");
    FILE* fout = std::fopen(out_path, "wb");
    if (!fin || !fout) return 2;
    Reader in{fin};
    const Index rows = in.get<Index>(), cols = in.get<Index>(), count = in.get<Index>();
    const int64_t dtype = in.get<int64_t>();
    if (dtype == 1) dumpSystem<double>(in, rows, cols, count, fout);
    else dumpSystem<float>(in, rows, cols, count, fout);
    std::fclose(fin);
    std::fclose(fout);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 4 && std::strcmp(argv[1], "dump") == 0) return dump(argv[2], argv[3]);
    const struct
    {
        const char* name;
        void (*fn)();
    } cases[] = {
        {"coo fold", cooFold},
        {"coo reuse", cooReuse},
        {"coo empty and errors", cooEmptyAndErrors},
        {"spmv kats", spmvKats},
        {"spmm kats", spmmKats},
        {"coo matrix cpu", cooMatrixCpu},
        {"pcg diagonal float", pcgDiagonal<float>},
        {"pcg diagonal double", pcgDiagonal<double>},
        {"poisson scales float", poissonScales<float>},
        {"poisson scales double", poissonScales<double>},
        {"zero iterations float", zeroIterations<float>},
        {"zero iterations double", zeroIterations<double>},
        {"solver errors float", solverErrors<float>},
        {"solver errors double", solverErrors<double>},
        {"limits and tolerances float", limitsAndTolerances<float>},
        {"limits and tolerances double", limitsAndTolerances<double>},
        {"poisson closed form float", poissonClosedForm<float>},
        {"poisson closed form double", poissonClosedForm<double>},
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
    }
    if (g_failures)
    {
        std::printf("%d FAILURES\n", g_failures);
        return 1;
    }
    std::printf("ALL PASSED\n");
    return 0;
}
