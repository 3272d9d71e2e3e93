/*
 * Oracle (test infrastructure only): CPU restatement of PlaMatrix's block
 * normal equations and damped Schur-complement solve.  See oracle.h.
 *
 * Follows:
 *   src/optimization/block_normal_equations.cpp:13-411  accumulation
 *   src/optimization/block_schur.cpp:53-459            damped Schur solve
 *   src/optimization/block_schur_cpu_solver.h:14-82    matrix-free PCG
 *   src/optimization/block_schur_linear_algebra.h       Cholesky inverse,
 *                                                       small mat-vecs
 *   src/optimization/block_schur_sparse_assembly.h:41-272  pattern rebuild
 *   src/optimization/block_schur_sparse_assembly.h:817-903 CPU CSR values
 *
 * Pinning: the hard-coded damped dense normal-equation solutions of
 * block_schur_test.cpp:77-158, :482-548 (tests/test_oracle_schur.py), dense
 * numpy elimination of [[B, W], [W^T, C]] on random instances, bitwise
 * agreement with oracle/_ref (the reference's own CPU build) where built, and
 * pattern invariants (symmetry, sortedness, slot counts of ring topologies).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

int or_fail(int code, const char* fmt, ...);

struct or_ne {
    int64_t pc, ec, ps, es;
    double* B;  /* pc * ps * ps */
    double* C;  /* ec * es * es */
    double* gp; /* pc * ps */
    double* ge; /* ec * es */
    /* cross blocks W (primary x eliminated, row-major ps x es), creation order */
    int64_t ncross, cap_cross;
    int64_t* cross_primary;
    int64_t* cross_eliminated;
    double* cross_values;
    /* primary cross blocks (row < col), creation order */
    int64_t npcross, cap_pcross;
    int64_t* pcross_row;
    int64_t* pcross_col;
    double* pcross_values;
    /* per eliminated block: list of cross indices (creation order) */
    int64_t** adj;
    int64_t* adj_n;
    int64_t* adj_cap;
};

or_ne* or_ne_create(int64_t pc, int64_t ec, int64_t ps, int64_t es)
{
    if (pc < 0 || ec < 0 || ps <= 0 || es <= 0)
    {
        or_fail(OR_INVALID_ARGUMENT,
                "BlockNormalEquations: block counts must be non-negative and block sizes positive");
        return NULL;
    }
    or_ne* eq = (or_ne*)calloc(1, sizeof(or_ne));
    eq->pc = pc;
    eq->ec = ec;
    eq->ps = ps;
    eq->es = es;
    eq->B = (double*)calloc((size_t)(pc * ps * ps + 1), sizeof(double));
    eq->C = (double*)calloc((size_t)(ec * es * es + 1), sizeof(double));
    eq->gp = (double*)calloc((size_t)(pc * ps + 1), sizeof(double));
    eq->ge = (double*)calloc((size_t)(ec * es + 1), sizeof(double));
    eq->adj = (int64_t**)calloc((size_t)(ec + 1), sizeof(int64_t*));
    eq->adj_n = (int64_t*)calloc((size_t)(ec + 1), sizeof(int64_t));
    eq->adj_cap = (int64_t*)calloc((size_t)(ec + 1), sizeof(int64_t));
    return eq;
}

void or_ne_destroy(or_ne* eq)
{
    if (!eq) return;
    for (int64_t e = 0; e < eq->ec; ++e) free(eq->adj[e]);
    free(eq->adj);
    free(eq->adj_n);
    free(eq->adj_cap);
    free(eq->B);
    free(eq->C);
    free(eq->gp);
    free(eq->ge);
    free(eq->cross_primary);
    free(eq->cross_eliminated);
    free(eq->cross_values);
    free(eq->pcross_row);
    free(eq->pcross_col);
    free(eq->pcross_values);
    free(eq);
}

/* validateResidual: block_normal_equations.cpp:47-82 */
static int validate_residual(const double* jacobian, int64_t jacobian_size, const double* residual,
                             int64_t residual_size, double weight, const char* operation)
{
    if (!jacobian || !residual || residual_size <= 0)
        return or_fail(OR_INVALID_ARGUMENT, "%s: pointers must be non-null and residual size positive",
                       operation);
    if (!isfinite(weight) || weight < 0.0)
        return or_fail(OR_INVALID_ARGUMENT, "%s: weight must be finite and non-negative", operation);
    for (int64_t i = 0; i < jacobian_size; ++i)
        if (!isfinite(jacobian[i]))
            return or_fail(OR_INVALID_ARGUMENT, "%s: Jacobian values must be finite", operation);
    for (int64_t i = 0; i < residual_size; ++i)
        if (!isfinite(residual[i]))
            return or_fail(OR_INVALID_ARGUMENT, "%s: residual values must be finite", operation);
    return OR_OK;
}

/* findOrCreateCrossBlock: block_normal_equations.cpp:84-106 */
static int64_t find_or_create_cross(or_ne* eq, int64_t primary, int64_t eliminated)
{
    int64_t* list = eq->adj[eliminated];
    for (int64_t i = 0; i < eq->adj_n[eliminated]; ++i)
        if (eq->cross_primary[list[i]] == primary) return list[i];
    const int64_t block = eq->ps * eq->es;
    if (eq->ncross == eq->cap_cross)
    {
        eq->cap_cross = eq->cap_cross ? eq->cap_cross * 2 : 64;
        eq->cross_primary = (int64_t*)realloc(eq->cross_primary, (size_t)eq->cap_cross * sizeof(int64_t));
        eq->cross_eliminated = (int64_t*)realloc(eq->cross_eliminated, (size_t)eq->cap_cross * sizeof(int64_t));
        eq->cross_values = (double*)realloc(eq->cross_values, (size_t)(eq->cap_cross * block) * sizeof(double));
    }
    const int64_t index = eq->ncross++;
    eq->cross_primary[index] = primary;
    eq->cross_eliminated[index] = eliminated;
    memset(eq->cross_values + index * block, 0, (size_t)block * sizeof(double));
    if (eq->adj_n[eliminated] == eq->adj_cap[eliminated])
    {
        eq->adj_cap[eliminated] = eq->adj_cap[eliminated] ? eq->adj_cap[eliminated] * 2 : 4;
        eq->adj[eliminated] = (int64_t*)realloc(eq->adj[eliminated],
                                                (size_t)eq->adj_cap[eliminated] * sizeof(int64_t));
    }
    eq->adj[eliminated][eq->adj_n[eliminated]++] = index;
    return index;
}

/* findOrCreatePrimaryCrossBlock: block_normal_equations.cpp:108-133 */
static int64_t find_or_create_primary_cross(or_ne* eq, int64_t row_block, int64_t column_block)
{
    if (row_block >= column_block)
    {
        or_fail(OR_INVALID_ARGUMENT, "BlockNormalEquations: primary cross blocks must be upper triangular");
        return -1;
    }
    for (int64_t i = 0; i < eq->npcross; ++i)
        if (eq->pcross_row[i] == row_block && eq->pcross_col[i] == column_block) return i;
    const int64_t block = eq->ps * eq->ps;
    if (eq->npcross == eq->cap_pcross)
    {
        eq->cap_pcross = eq->cap_pcross ? eq->cap_pcross * 2 : 16;
        eq->pcross_row = (int64_t*)realloc(eq->pcross_row, (size_t)eq->cap_pcross * sizeof(int64_t));
        eq->pcross_col = (int64_t*)realloc(eq->pcross_col, (size_t)eq->cap_pcross * sizeof(int64_t));
        eq->pcross_values = (double*)realloc(eq->pcross_values, (size_t)(eq->cap_pcross * block) * sizeof(double));
    }
    const int64_t index = eq->npcross++;
    eq->pcross_row[index] = row_block;
    eq->pcross_col[index] = column_block;
    memset(eq->pcross_values + index * block, 0, (size_t)block * sizeof(double));
    return index;
}

/* addPrimaryTerms: block_normal_equations.cpp:135-222 */
static int add_primary_terms(or_ne* eq, int64_t count, const int64_t* blocks, const double* const* jacs,
                             const double* residual, int64_t rs, double weight)
{
    const int64_t ps = eq->ps;
    if (count <= 0 || !blocks || !jacs)
        return or_fail(OR_INVALID_ARGUMENT,
                       "BlockNormalEquations: primary block and Jacobian lists must have equal non-zero size");
    for (int64_t index = 0; index < count; ++index)
    {
        const int64_t block = blocks[index];
        if (block < 0 || block >= eq->pc)
            return or_fail(OR_OUT_OF_RANGE, "BlockNormalEquations: primary block index out of range");
        for (int64_t prior = 0; prior < index; ++prior)
            if (blocks[prior] == block)
                return or_fail(OR_INVALID_ARGUMENT, "BlockNormalEquations: primary block indices must be unique");
        const int status = validate_residual(jacs[index], rs * ps, residual, rs, weight,
                                             "BlockNormalEquations::addPrimaryTerms");
        if (status != OR_OK) return status;
    }
    for (int64_t bi = 0; bi < count; ++bi)
    {
        const int64_t block = blocks[bi];
        const double* jacobian = jacs[bi];
        double* diag = eq->B + block * ps * ps;
        double* grad = eq->gp + block * ps;
        for (int64_t row = 0; row < rs; ++row)
            for (int64_t column = 0; column < ps; ++column)
            {
                const double value = jacobian[row * ps + column];
                grad[column] += weight * value * residual[row];
                for (int64_t other = 0; other < ps; ++other)
                    diag[column * ps + other] += weight * value * jacobian[row * ps + other];
            }
    }
    for (int64_t left = 0; left < count; ++left)
        for (int64_t right = left + 1; right < count; ++right)
        {
            const int ordered = blocks[left] < blocks[right];
            const int64_t row_block = ordered ? blocks[left] : blocks[right];
            const int64_t column_block = ordered ? blocks[right] : blocks[left];
            const double* row_jac = ordered ? jacs[left] : jacs[right];
            const double* col_jac = ordered ? jacs[right] : jacs[left];
            const int64_t index = find_or_create_primary_cross(eq, row_block, column_block);
            if (index < 0) return OR_INVALID_ARGUMENT;
            double* cross = eq->pcross_values + index * ps * ps;
            for (int64_t rr = 0; rr < rs; ++rr)
                for (int64_t row = 0; row < ps; ++row)
                {
                    const double row_value = row_jac[rr * ps + row];
                    for (int64_t column = 0; column < ps; ++column)
                        cross[row * ps + column] += weight * row_value * col_jac[rr * ps + column];
                }
        }
    return OR_OK;
}

/* addResidualBlocks: block_normal_equations.cpp:243-312 */
int or_ne_add_residual_blocks(or_ne* eq, int64_t count, const int64_t* blocks, const double* const* jacs,
                              int64_t eliminated, const double* ej, const double* residual, int64_t rs,
                              double weight)
{
    const int64_t ps = eq->ps, es = eq->es;
    if (eliminated < 0 || eliminated >= eq->ec)
        return or_fail(OR_OUT_OF_RANGE, "BlockNormalEquations::addResidualBlocks: eliminated block out of range");
    int status = validate_residual(ej, rs * es, residual, rs, weight,
                                   "BlockNormalEquations::addResidualBlocks(eliminated)");
    if (status != OR_OK) return status;
    status = add_primary_terms(eq, count, blocks, jacs, residual, rs, weight);
    if (status != OR_OK) return status;
    for (int64_t pi = 0; pi < count; ++pi)
    {
        const int64_t index = find_or_create_cross(eq, blocks[pi], eliminated);
        double* cross = eq->cross_values + index * ps * es;
        const double* pj = jacs[pi];
        for (int64_t row = 0; row < rs; ++row)
            for (int64_t pcol = 0; pcol < ps; ++pcol)
            {
                const double pv = pj[row * ps + pcol];
                for (int64_t ecol = 0; ecol < es; ++ecol)
                    cross[pcol * es + ecol] += weight * pv * ej[row * es + ecol];
            }
    }
    double* diag = eq->C + eliminated * es * es;
    double* grad = eq->ge + eliminated * es;
    for (int64_t row = 0; row < rs; ++row)
        for (int64_t ecol = 0; ecol < es; ++ecol)
        {
            const double ev = ej[row * es + ecol];
            grad[ecol] += weight * ev * residual[row];
            for (int64_t other = 0; other < es; ++other)
                diag[ecol * es + other] += weight * ev * ej[row * es + other];
        }
    return OR_OK;
}

int or_ne_add_residual_batch(or_ne* eq, int64_t count, const int64_t* primary, const int64_t* eliminated,
                             const double* jp, const double* je, const double* r, int64_t rs,
                             const double* weights)
{
    for (int64_t i = 0; i < count; ++i)
    {
        const double* jac = jp + i * rs * eq->ps;
        const int status = or_ne_add_residual_blocks(eq, 1, primary + i, &jac, eliminated[i],
                                                     je + i * rs * eq->es, r + i * rs, rs,
                                                     weights ? weights[i] : 1.0);
        if (status != OR_OK) return status;
    }
    return OR_OK;
}

int or_ne_add_primary_residual_blocks(or_ne* eq, int64_t count, const int64_t* blocks,
                                      const double* const* jacs, const double* residual, int64_t rs,
                                      double weight)
{
    return add_primary_terms(eq, count, blocks, jacs, residual, rs, weight);
}

/* addEliminatedResidualBlock: block_normal_equations.cpp:337-373 */
int or_ne_add_eliminated_residual_block(or_ne* eq, int64_t eliminated, const double* ej,
                                        const double* residual, int64_t rs, double weight)
{
    const int64_t es = eq->es;
    if (eliminated < 0 || eliminated >= eq->ec)
        return or_fail(OR_OUT_OF_RANGE,
                       "BlockNormalEquations::addEliminatedResidualBlock: block index out of range");
    const int status = validate_residual(ej, rs * es, residual, rs, weight,
                                         "BlockNormalEquations::addEliminatedResidualBlock");
    if (status != OR_OK) return status;
    double* diag = eq->C + eliminated * es * es;
    double* grad = eq->ge + eliminated * es;
    for (int64_t row = 0; row < rs; ++row)
        for (int64_t col = 0; col < es; ++col)
        {
            const double value = ej[row * es + col];
            grad[col] += weight * value * residual[row];
            for (int64_t other = 0; other < es; ++other)
                diag[col * es + other] += weight * value * ej[row * es + other];
        }
    return OR_OK;
}

/* mergeFrom: block_normal_equations.cpp:375-411 */
int or_ne_merge(or_ne* eq, const or_ne* other)
{
    if (eq->pc != other->pc || eq->ec != other->ec || eq->ps != other->ps || eq->es != other->es)
        return or_fail(OR_INVALID_ARGUMENT, "BlockNormalEquations::mergeFrom: block layouts must match");
    for (int64_t i = 0; i < eq->pc * eq->ps * eq->ps; ++i) eq->B[i] += other->B[i];
    for (int64_t i = 0; i < eq->ec * eq->es * eq->es; ++i) eq->C[i] += other->C[i];
    for (int64_t i = 0; i < eq->pc * eq->ps; ++i) eq->gp[i] += other->gp[i];
    for (int64_t i = 0; i < eq->ec * eq->es; ++i) eq->ge[i] += other->ge[i];
    for (int64_t s = 0; s < other->npcross; ++s)
    {
        const int64_t t = find_or_create_primary_cross(eq, other->pcross_row[s], other->pcross_col[s]);
        for (int64_t i = 0; i < eq->ps * eq->ps; ++i)
            eq->pcross_values[t * eq->ps * eq->ps + i] += other->pcross_values[s * eq->ps * eq->ps + i];
    }
    for (int64_t s = 0; s < other->ncross; ++s)
    {
        const int64_t t = find_or_create_cross(eq, other->cross_primary[s], other->cross_eliminated[s]);
        for (int64_t i = 0; i < eq->ps * eq->es; ++i)
            eq->cross_values[t * eq->ps * eq->es + i] += other->cross_values[s * eq->ps * eq->es + i];
    }
    return OR_OK;
}

int64_t or_ne_cross_count(const or_ne* eq) { return eq->ncross; }
int64_t or_ne_primary_cross_count(const or_ne* eq) { return eq->npcross; }

void or_ne_export(const or_ne* eq, double* B, double* C, double* gp, double* ge, int64_t* cp, int64_t* ce,
                  double* cv, int64_t* pr, int64_t* pcc, double* pv)
{
    if (B) memcpy(B, eq->B, (size_t)(eq->pc * eq->ps * eq->ps) * sizeof(double));
    if (C) memcpy(C, eq->C, (size_t)(eq->ec * eq->es * eq->es) * sizeof(double));
    if (gp) memcpy(gp, eq->gp, (size_t)(eq->pc * eq->ps) * sizeof(double));
    if (ge) memcpy(ge, eq->ge, (size_t)(eq->ec * eq->es) * sizeof(double));
    if (cp) memcpy(cp, eq->cross_primary, (size_t)eq->ncross * sizeof(int64_t));
    if (ce) memcpy(ce, eq->cross_eliminated, (size_t)eq->ncross * sizeof(int64_t));
    if (cv) memcpy(cv, eq->cross_values, (size_t)(eq->ncross * eq->ps * eq->es) * sizeof(double));
    if (pr) memcpy(pr, eq->pcross_row, (size_t)eq->npcross * sizeof(int64_t));
    if (pcc) memcpy(pcc, eq->pcross_col, (size_t)eq->npcross * sizeof(int64_t));
    if (pv) memcpy(pv, eq->pcross_values, (size_t)(eq->npcross * eq->ps * eq->ps) * sizeof(double));
}

/* ---- small dense linear algebra: block_schur_linear_algebra.h ----------- */

static double dot_product(int64_t n, const double* a, const double* b)
{
    double result = 0.0;
    for (int64_t i = 0; i < n; ++i) result += a[i] * b[i];
    return result;
}

static double vector_norm(int64_t n, const double* v)
{
    const double d = dot_product(n, v, v);
    return sqrt(d > 0.0 ? d : 0.0);
}

/* invertPositiveDefiniteInto: block_schur_linear_algebra.h:92-147 (identical
 * arithmetic to invertPositiveDefinite :31-89) */
static int invert_pd(const double* matrix, int64_t size, double* inverse, double* lower, double* intermediate)
{
    memset(lower, 0, (size_t)(size * size) * sizeof(double));
    for (int64_t row = 0; row < size; ++row)
        for (int64_t column = 0; column <= row; ++column)
        {
            double value = matrix[row * size + column];
            for (int64_t inner = 0; inner < column; ++inner)
                value -= lower[row * size + inner] * lower[column * size + inner];
            if (row == column)
            {
                if (!(value > 0.0) || !isfinite(value)) return 0;
                lower[row * size + column] = sqrt(value);
            }
            else
            {
                lower[row * size + column] = value / lower[column * size + column];
            }
        }
    memset(inverse, 0, (size_t)(size * size) * sizeof(double));
    for (int64_t column = 0; column < size; ++column)
    {
        for (int64_t row = 0; row < size; ++row)
        {
            double value = row == column ? 1.0 : 0.0;
            for (int64_t inner = 0; inner < row; ++inner) value -= lower[row * size + inner] * intermediate[inner];
            intermediate[row] = value / lower[row * size + row];
        }
        for (int64_t offset = 0; offset < size; ++offset)
        {
            const int64_t row = size - 1 - offset;
            double value = intermediate[row];
            for (int64_t inner = row + 1; inner < size; ++inner)
                value -= lower[inner * size + row] * inverse[inner * size + column];
            inverse[row * size + column] = value / lower[row * size + row];
        }
    }
    return 1;
}

/* multiplyMatrixVector :148-163 */
static void mat_vec(const double* m, int64_t rows, int64_t cols, const double* v, double* result)
{
    for (int64_t row = 0; row < rows; ++row)
    {
        double value = 0.0;
        for (int64_t col = 0; col < cols; ++col) value += m[row * cols + col] * v[col];
        result[row] = value;
    }
}

/* addMatrixVector :165-181 */
static void add_mat_vec(const double* m, int64_t rows, int64_t cols, const double* v, double scale, double* result)
{
    for (int64_t row = 0; row < rows; ++row)
    {
        double value = 0.0;
        for (int64_t col = 0; col < cols; ++col) value += m[row * cols + col] * v[col];
        result[row] += scale * value;
    }
}

/* addTransposeMatrixVector :183-199 */
static void add_transpose_mat_vec(const double* m, int64_t rows, int64_t cols, const double* v, double* result)
{
    for (int64_t col = 0; col < cols; ++col)
    {
        double value = 0.0;
        for (int64_t row = 0; row < rows; ++row) value += m[row * cols + col] * v[row];
        result[col] += value;
    }
}

/* ---- damped Schur prologue: block_schur.cpp:92-250 --------------------- */

typedef struct prologue {
    double* B;       /* damped primary diagonal */
    double* C;       /* damped eliminated diagonal */
    double* Cinv;    /* eliminated inverses */
    double* rhs;     /* reduced rhs */
    double* Pinv;    /* preconditioner inverses */
    int status;      /* 0 ok, 1 eliminated not PD, 2 preconditioner not PD */
} prologue;

static void prologue_free(prologue* p)
{
    free(p->B);
    free(p->C);
    free(p->Cinv);
    free(p->rhs);
    free(p->Pinv);
}

static void apply_damping(double* blocks, int64_t count, int64_t size, double damping, double minimum)
{
    for (int64_t block = 0; block < count; ++block)
        for (int64_t d = 0; d < size; ++d)
        {
            double* value = blocks + block * size * size + d * size + d;
            const double magnitude = fabs(*value);
            /* std::max(|v|, minimumDiagonal) */
            *value += damping * (magnitude < minimum ? minimum : magnitude);
        }
}

static void run_prologue(const or_ne* eq, double damping, double minimum, int want_preconditioner, prologue* out)
{
    const int64_t pc = eq->pc, ec = eq->ec, ps = eq->ps, es = eq->es;
    memset(out, 0, sizeof(*out));
    out->B = (double*)malloc((size_t)(pc * ps * ps + 1) * sizeof(double));
    out->C = (double*)malloc((size_t)(ec * es * es + 1) * sizeof(double));
    out->Cinv = (double*)calloc((size_t)(ec * es * es + 1), sizeof(double));
    out->rhs = (double*)calloc((size_t)(pc * ps + 1), sizeof(double));
    out->Pinv = (double*)calloc((size_t)(pc * ps * ps + 1), sizeof(double));
    memcpy(out->B, eq->B, (size_t)(pc * ps * ps) * sizeof(double));
    memcpy(out->C, eq->C, (size_t)(ec * es * es) * sizeof(double));
    apply_damping(out->B, pc, ps, damping, minimum);
    apply_damping(out->C, ec, es, damping, minimum);

    double* lower = (double*)malloc((size_t)(es * es + ps * ps + 1) * sizeof(double));
    double* inter = (double*)malloc((size_t)(es + ps + 1) * sizeof(double));
    for (int64_t block = 0; block < ec; ++block)
        if (!invert_pd(out->C + block * es * es, es, out->Cinv + block * es * es, lower, inter))
        {
            out->status = 1;
            free(lower);
            free(inter);
            return;
        }

    for (int64_t i = 0; i < pc * ps; ++i) out->rhs[i] = -eq->gp[i];
    double* product = (double*)malloc((size_t)es * sizeof(double));
    for (int64_t block = 0; block < ec; ++block)
    {
        mat_vec(out->Cinv + block * es * es, es, es, eq->ge + block * es, product);
        for (int64_t k = 0; k < eq->adj_n[block]; ++k)
        {
            const int64_t c = eq->adj[block][k];
            add_mat_vec(eq->cross_values + c * ps * es, ps, es, product, 1.0,
                        out->rhs + eq->cross_primary[c] * ps);
        }
    }
    free(product);

    if (want_preconditioner)
    {
        /* primary adjacency: cross indices grouped by primary, ascending */
        int64_t* count = (int64_t*)calloc((size_t)(pc + 1), sizeof(int64_t));
        for (int64_t c = 0; c < eq->ncross; ++c) ++count[eq->cross_primary[c] + 1];
        for (int64_t b = 0; b < pc; ++b) count[b + 1] += count[b];
        int64_t* order = (int64_t*)malloc((size_t)(eq->ncross + 1) * sizeof(int64_t));
        int64_t* fill = (int64_t*)malloc((size_t)(pc + 1) * sizeof(int64_t));
        memcpy(fill, count, (size_t)pc * sizeof(int64_t));
        for (int64_t c = 0; c < eq->ncross; ++c) order[fill[eq->cross_primary[c]]++] = c;
        double* schur = (double*)malloc((size_t)(ps * ps) * sizeof(double));
        double* temp = (double*)malloc((size_t)(ps * es) * sizeof(double));
        for (int64_t block = 0; block < pc; ++block)
        {
            memcpy(schur, out->B + block * ps * ps, (size_t)(ps * ps) * sizeof(double));
            for (int64_t k = count[block]; k < count[block + 1]; ++k)
            {
                const int64_t c = order[k];
                const double* inverse = out->Cinv + eq->cross_eliminated[c] * es * es;
                const double* values = eq->cross_values + c * ps * es;
                if (ps == 9 && es == 3)
                {
                    /* transform9x3 :102-115 */
                    for (int64_t row = 0; row < 9; ++row)
                    {
                        const double x0 = values[row * 3], x1 = values[row * 3 + 1], x2 = values[row * 3 + 2];
                        temp[row * 3] = inverse[0] * x0 + inverse[1] * x1 + inverse[2] * x2;
                        temp[row * 3 + 1] = inverse[3] * x0 + inverse[4] * x1 + inverse[5] * x2;
                        temp[row * 3 + 2] = inverse[6] * x0 + inverse[7] * x1 + inverse[8] * x2;
                    }
                }
                else
                {
                    for (int64_t row = 0; row < ps; ++row)
                        mat_vec(inverse, es, es, values + row * es, temp + row * es);
                }
                for (int64_t row = 0; row < ps; ++row)
                    for (int64_t col = 0; col < ps; ++col)
                    {
                        const double* t = temp + row * es;
                        const double* s = values + col * es;
                        double value;
                        if (es == 3)
                        {
                            value = t[0] * s[0] + t[1] * s[1] + t[2] * s[2]; /* dot3 :117-121 */
                        }
                        else
                        {
                            value = 0.0;
                            for (int64_t inner = 0; inner < es; ++inner) value += t[inner] * s[inner];
                        }
                        schur[row * ps + col] -= value;
                    }
            }
            if (!invert_pd(schur, ps, out->Pinv + block * ps * ps, lower, inter))
            {
                out->status = 2;
                break;
            }
        }
        free(schur);
        free(temp);
        free(count);
        free(order);
        free(fill);
    }
    free(lower);
    free(inter);
}

int or_schur_prologue(const or_ne* eq, double damping, double minimum_diagonal, double* B, double* Cinv,
                      double* rhs, double* Pinv)
{
    prologue p;
    run_prologue(eq, damping, minimum_diagonal, 1, &p);
    const int status = p.status;
    if (B) memcpy(B, p.B, (size_t)(eq->pc * eq->ps * eq->ps) * sizeof(double));
    if (Cinv) memcpy(Cinv, p.Cinv, (size_t)(eq->ec * eq->es * eq->es) * sizeof(double));
    if (rhs) memcpy(rhs, p.rhs, (size_t)(eq->pc * eq->ps) * sizeof(double));
    if (Pinv) memcpy(Pinv, p.Pinv, (size_t)(eq->pc * eq->ps * eq->ps) * sizeof(double));
    prologue_free(&p);
    if (status == 1) return or_fail(OR_RUNTIME, "eliminated block is not positive definite");
    if (status == 2) return or_fail(OR_RUNTIME, "Schur preconditioner block is not positive definite");
    return OR_OK;
}

/* apply_schur lambda: block_schur.cpp:254-309 */
static void apply_schur(const or_ne* eq, const prologue* p, const double* input, double* output, double* point,
                        double* inverse_value)
{
    const int64_t pc = eq->pc, ps = eq->ps, es = eq->es;
    memset(output, 0, (size_t)(pc * ps) * sizeof(double));
    for (int64_t block = 0; block < pc; ++block)
        mat_vec(p->B + block * ps * ps, ps, ps, input + block * ps, output + block * ps);
    for (int64_t k = 0; k < eq->npcross; ++k)
    {
        const double* values = eq->pcross_values + k * ps * ps;
        add_mat_vec(values, ps, ps, input + eq->pcross_col[k] * ps, 1.0, output + eq->pcross_row[k] * ps);
        add_transpose_mat_vec(values, ps, ps, input + eq->pcross_row[k] * ps, output + eq->pcross_col[k] * ps);
    }
    for (int64_t e = 0; e < eq->ec; ++e)
    {
        memset(point, 0, (size_t)es * sizeof(double));
        for (int64_t k = 0; k < eq->adj_n[e]; ++k)
        {
            const int64_t c = eq->adj[e][k];
            add_transpose_mat_vec(eq->cross_values + c * ps * es, ps, es, input + eq->cross_primary[c] * ps, point);
        }
        mat_vec(p->Cinv + e * es * es, es, es, point, inverse_value);
        for (int64_t k = 0; k < eq->adj_n[e]; ++k)
        {
            const int64_t c = eq->adj[e][k];
            add_mat_vec(eq->cross_values + c * ps * es, ps, es, inverse_value, -1.0,
                        output + eq->cross_primary[c] * ps);
        }
    }
}

int or_schur_solve(const or_ne* eq, double damping, const or_schur_options* o, double* primary_step,
                   double* eliminated_step, or_schur_report* report)
{
    memset(report, 0, sizeof(*report));
    if (!primary_step || !eliminated_step || (void*)primary_step == (void*)eliminated_step)
    {
        if (eq->pc * eq->ps > 0 || eq->ec * eq->es > 0 || primary_step == eliminated_step)
            return or_fail(OR_INVALID_ARGUMENT,
                           "solveDampedSchurComplement: output vectors must be non-null and distinct");
    }
    if (!isfinite(damping) || damping < 0.0 || o->max_iterations <= 0 || !isfinite(o->relative_tolerance)
        || o->relative_tolerance < 0.0 || !isfinite(o->absolute_tolerance) || o->absolute_tolerance < 0.0
        || !isfinite(o->minimum_diagonal) || o->minimum_diagonal <= 0.0)
        return or_fail(OR_INVALID_ARGUMENT, "solveDampedSchurComplement: invalid solver options");
    const int64_t pc = eq->pc, ec = eq->ec, ps = eq->ps, es = eq->es;
    const int64_t pd = pc * ps, ed = ec * es;
    if (!o->use_initial_guess)
        for (int64_t i = 0; i < pd; ++i) primary_step[i] = 0.0;
    for (int64_t i = 0; i < ed; ++i) eliminated_step[i] = 0.0;

    prologue p;
    run_prologue(eq, damping, o->minimum_diagonal, 1, &p);
    if (p.status != 0)
    {
        report->status = p.status;
        prologue_free(&p);
        return OR_OK;
    }

    double* point = (double*)malloc((size_t)(es + 1) * sizeof(double));
    double* inverse_value = (double*)malloc((size_t)(es + 1) * sizeof(double));
    if (pd == 0)
    {
        report->converged = 1;
    }
    else
    {
        /* solveReducedSchurOnCpu: block_schur_cpu_solver.h:14-82 */
        const size_t bytes = (size_t)pd * sizeof(double);
        double* product = (double*)calloc((size_t)pd, sizeof(double));
        double* residual = (double*)malloc(bytes);
        double* preconditioned = (double*)calloc((size_t)pd, sizeof(double));
        double* direction = (double*)malloc(bytes);
        apply_schur(eq, &p, primary_step, product, point, inverse_value);
        memcpy(residual, p.rhs, bytes);
        for (int64_t i = 0; i < pd; ++i) residual[i] -= product[i];
        report->initial_residual_norm = vector_norm(pd, residual);
        report->final_residual_norm = report->initial_residual_norm;
        const double tolerance = o->absolute_tolerance + o->relative_tolerance * report->initial_residual_norm;
        if (report->initial_residual_norm <= tolerance)
        {
            report->converged = 1;
        }
        else
        {
            for (int64_t b = 0; b < pc; ++b)
                mat_vec(p.Pinv + b * ps * ps, ps, ps, residual + b * ps, preconditioned + b * ps);
            memcpy(direction, preconditioned, bytes);
            memset(product, 0, bytes);
            double residual_dot = dot_product(pd, residual, preconditioned);
            for (int iteration = 0; iteration < o->max_iterations; ++iteration)
            {
                apply_schur(eq, &p, direction, product, point, inverse_value);
                const double denominator = dot_product(pd, direction, product);
                if (!(denominator > 0.0) || !isfinite(denominator))
                {
                    report->status = 3;
                    break;
                }
                const double alpha = residual_dot / denominator;
                for (int64_t i = 0; i < pd; ++i)
                {
                    primary_step[i] += alpha * direction[i];
                    residual[i] -= alpha * product[i];
                }
                report->iterations = iteration + 1;
                report->final_residual_norm = vector_norm(pd, residual);
                if (report->final_residual_norm <= tolerance)
                {
                    report->converged = 1;
                    break;
                }
                for (int64_t b = 0; b < pc; ++b)
                    mat_vec(p.Pinv + b * ps * ps, ps, ps, residual + b * ps, preconditioned + b * ps);
                const double next = dot_product(pd, residual, preconditioned);
                const double beta = next / residual_dot;
                for (int64_t i = 0; i < pd; ++i) direction[i] = preconditioned[i] + beta * direction[i];
                residual_dot = next;
            }
        }
        free(product);
        free(residual);
        free(preconditioned);
        free(direction);
    }

    if (report->converged)
    {
        /* back-substitution: block_schur.cpp:422-454 */
        double* point_rhs = point;
        for (int64_t e = 0; e < ec; ++e)
        {
            for (int64_t i = 0; i < es; ++i) point_rhs[i] = -eq->ge[e * es + i];
            for (int64_t k = 0; k < eq->adj_n[e]; ++k)
            {
                const int64_t c = eq->adj[e][k];
                const double* values = eq->cross_values + c * ps * es;
                const double* primary = primary_step + eq->cross_primary[c] * ps;
                for (int64_t column = 0; column < es; ++column)
                {
                    double value = 0.0;
                    for (int64_t row = 0; row < ps; ++row) value += values[row * es + column] * primary[row];
                    point_rhs[column] -= value;
                }
            }
            mat_vec(p.Cinv + e * es * es, es, es, point_rhs, eliminated_step + e * es);
        }
    }
    else if (report->status == 0)
    {
        report->status = 4;
    }
    free(point);
    free(inverse_value);
    prologue_free(&p);
    return OR_OK;
}

/* ---- Schur CSR pattern: block_schur_sparse_assembly.h:41-272 ----------- */

typedef struct pattern {
    int64_t slots, values, terms;
    int64_t* slot_row;
    int64_t* slot_col;
    int64_t* slot_ordinal;
    int64_t* row_offsets;
    int64_t* col_indices;
    int64_t* term_offsets;
    int64_t* term_e;
    int64_t* term_l;
    int64_t* term_r;
    int64_t* pair_slot_offsets; /* per eliminated: start into pair_slots */
    int64_t* pair_slots;        /* k_e * k_e slot ids per eliminated */
    int64_t* primary_pair_slots;
    int64_t* diagonal_slots;
} pattern;

static int cmp_key(const void* a, const void* b)
{
    const int64_t* x = (const int64_t*)a;
    const int64_t* y = (const int64_t*)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

static int64_t find_slot(const pattern* p, int64_t row, int64_t col)
{
    int64_t lo = 0, hi = p->slots;
    while (lo < hi)
    {
        const int64_t mid = (lo + hi) / 2;
        if (p->slot_row[mid] < row || (p->slot_row[mid] == row && p->slot_col[mid] < col)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

static void pattern_free(pattern* p)
{
    free(p->slot_row);
    free(p->slot_col);
    free(p->slot_ordinal);
    free(p->row_offsets);
    free(p->col_indices);
    free(p->term_offsets);
    free(p->term_e);
    free(p->term_l);
    free(p->term_r);
    free(p->pair_slot_offsets);
    free(p->pair_slots);
    free(p->primary_pair_slots);
    free(p->diagonal_slots);
}

static void build_pattern(const or_ne* eq, pattern* p)
{
    const int64_t pc = eq->pc, ec = eq->ec, ps = eq->ps;
    memset(p, 0, sizeof(*p));
    /* collect keys: the std::map insertion set :52-76 */
    int64_t key_count = pc + 2 * eq->npcross;
    for (int64_t e = 0; e < ec; ++e) key_count += eq->adj_n[e] * eq->adj_n[e];
    int64_t* keys = (int64_t*)malloc((size_t)(2 * key_count + 2) * sizeof(int64_t));
    int64_t k = 0;
    for (int64_t b = 0; b < pc; ++b)
    {
        keys[2 * k] = b;
        keys[2 * k + 1] = b;
        ++k;
    }
    for (int64_t c = 0; c < eq->npcross; ++c)
    {
        keys[2 * k] = eq->pcross_row[c];
        keys[2 * k + 1] = eq->pcross_col[c];
        ++k;
        keys[2 * k] = eq->pcross_col[c];
        keys[2 * k + 1] = eq->pcross_row[c];
        ++k;
    }
    for (int64_t e = 0; e < ec; ++e)
        for (int64_t l = 0; l < eq->adj_n[e]; ++l)
            for (int64_t r = 0; r < eq->adj_n[e]; ++r)
            {
                keys[2 * k] = eq->cross_primary[eq->adj[e][l]];
                keys[2 * k + 1] = eq->cross_primary[eq->adj[e][r]];
                ++k;
            }
    qsort(keys, (size_t)key_count, 2 * sizeof(int64_t), cmp_key);
    int64_t unique = 0;
    for (int64_t i = 0; i < key_count; ++i)
        if (i == 0 || cmp_key(keys + 2 * i, keys + 2 * (i - 1)) != 0)
        {
            keys[2 * unique] = keys[2 * i];
            keys[2 * unique + 1] = keys[2 * i + 1];
            ++unique;
        }
    p->slots = unique;
    p->slot_row = (int64_t*)malloc((size_t)(unique + 1) * sizeof(int64_t));
    p->slot_col = (int64_t*)malloc((size_t)(unique + 1) * sizeof(int64_t));
    p->slot_ordinal = (int64_t*)malloc((size_t)(unique + 1) * sizeof(int64_t));
    int64_t current_row = -1, ordinal = 0;
    for (int64_t s = 0; s < unique; ++s)
    {
        p->slot_row[s] = keys[2 * s];
        p->slot_col[s] = keys[2 * s + 1];
        if (p->slot_row[s] != current_row)
        {
            current_row = p->slot_row[s];
            ordinal = 0;
        }
        p->slot_ordinal[s] = ordinal++;
    }
    free(keys);

    p->diagonal_slots = (int64_t*)malloc((size_t)(pc + 1) * sizeof(int64_t));
    for (int64_t b = 0; b < pc; ++b) p->diagonal_slots[b] = find_slot(p, b, b);
    p->primary_pair_slots = (int64_t*)malloc((size_t)(2 * eq->npcross + 1) * sizeof(int64_t));
    for (int64_t c = 0; c < eq->npcross; ++c)
    {
        p->primary_pair_slots[2 * c] = find_slot(p, eq->pcross_row[c], eq->pcross_col[c]);
        p->primary_pair_slots[2 * c + 1] = find_slot(p, eq->pcross_col[c], eq->pcross_row[c]);
    }
    p->pair_slot_offsets = (int64_t*)malloc((size_t)(ec + 1) * sizeof(int64_t));
    int64_t total_pairs = 0;
    for (int64_t e = 0; e < ec; ++e)
    {
        p->pair_slot_offsets[e] = total_pairs;
        total_pairs += eq->adj_n[e] * eq->adj_n[e];
    }
    p->pair_slot_offsets[ec] = total_pairs;
    p->pair_slots = (int64_t*)malloc((size_t)(total_pairs + 1) * sizeof(int64_t));
    for (int64_t e = 0; e < ec; ++e)
    {
        const int64_t n = eq->adj_n[e];
        for (int64_t l = 0; l < n; ++l)
            for (int64_t r = 0; r < n; ++r)
                p->pair_slots[p->pair_slot_offsets[e] + l * n + r]
                    = find_slot(p, eq->cross_primary[eq->adj[e][l]], eq->cross_primary[eq->adj[e][r]]);
    }

    /* scalar CSR pattern :129-154 */
    const int64_t dimension = pc * ps;
    int64_t* slots_in_row = (int64_t*)calloc((size_t)(pc + 1), sizeof(int64_t));
    for (int64_t s = 0; s < unique; ++s) ++slots_in_row[p->slot_row[s] + 1];
    for (int64_t b = 0; b < pc; ++b) slots_in_row[b + 1] += slots_in_row[b];
    p->values = unique * ps * ps;
    p->row_offsets = (int64_t*)malloc((size_t)(dimension + 1) * sizeof(int64_t));
    p->col_indices = (int64_t*)malloc((size_t)(p->values + 1) * sizeof(int64_t));
    int64_t cursor = 0;
    for (int64_t br = 0; br < pc; ++br)
        for (int64_t lr = 0; lr < ps; ++lr)
        {
            p->row_offsets[br * ps + lr] = cursor;
            for (int64_t s = slots_in_row[br]; s < slots_in_row[br + 1]; ++s)
                for (int64_t c = 0; c < ps; ++c) p->col_indices[cursor++] = p->slot_col[s] * ps + c;
        }
    p->row_offsets[dimension] = cursor;
    free(slots_in_row);

    /* per-slot term lists in eliminated order :223-258 */
    p->terms = total_pairs;
    p->term_offsets = (int64_t*)calloc((size_t)(unique + 1), sizeof(int64_t));
    for (int64_t i = 0; i < total_pairs; ++i) ++p->term_offsets[p->pair_slots[i] + 1];
    for (int64_t s = 0; s < unique; ++s) p->term_offsets[s + 1] += p->term_offsets[s];
    int64_t* fill = (int64_t*)malloc((size_t)(unique + 1) * sizeof(int64_t));
    memcpy(fill, p->term_offsets, (size_t)(unique + 1) * sizeof(int64_t));
    p->term_e = (int64_t*)malloc((size_t)(total_pairs + 1) * sizeof(int64_t));
    p->term_l = (int64_t*)malloc((size_t)(total_pairs + 1) * sizeof(int64_t));
    p->term_r = (int64_t*)malloc((size_t)(total_pairs + 1) * sizeof(int64_t));
    for (int64_t e = 0; e < ec; ++e)
    {
        const int64_t n = eq->adj_n[e];
        for (int64_t l = 0; l < n; ++l)
            for (int64_t r = 0; r < n; ++r)
            {
                const int64_t slot = p->pair_slots[p->pair_slot_offsets[e] + l * n + r];
                const int64_t t = fill[slot]++;
                p->term_e[t] = e;
                p->term_l[t] = eq->adj[e][l];
                p->term_r[t] = eq->adj[e][r];
            }
    }
    free(fill);
}

static int64_t value_position(const pattern* p, int64_t ps, int64_t slot, int64_t lr, int64_t lc)
{
    return p->row_offsets[p->slot_row[slot] * ps + lr] + p->slot_ordinal[slot] * ps + lc;
}

int or_schur_pattern_size(const or_ne* eq, int64_t* slot_count, int64_t* value_count, int64_t* term_count)
{
    pattern p;
    build_pattern(eq, &p);
    *slot_count = p.slots;
    *value_count = p.values;
    *term_count = p.terms;
    pattern_free(&p);
    return OR_OK;
}

int or_schur_pattern(const or_ne* eq, int64_t* slot_rows, int64_t* slot_cols, int64_t* row_offsets,
                     int64_t* col_indices, int64_t* term_offsets, int64_t* term_e, int64_t* term_l,
                     int64_t* term_r)
{
    pattern p;
    build_pattern(eq, &p);
    const int64_t dimension = eq->pc * eq->ps;
    if (slot_rows) memcpy(slot_rows, p.slot_row, (size_t)p.slots * sizeof(int64_t));
    if (slot_cols) memcpy(slot_cols, p.slot_col, (size_t)p.slots * sizeof(int64_t));
    if (row_offsets) memcpy(row_offsets, p.row_offsets, (size_t)(dimension + 1) * sizeof(int64_t));
    if (col_indices) memcpy(col_indices, p.col_indices, (size_t)p.values * sizeof(int64_t));
    if (term_offsets) memcpy(term_offsets, p.term_offsets, (size_t)(p.slots + 1) * sizeof(int64_t));
    if (term_e) memcpy(term_e, p.term_e, (size_t)p.terms * sizeof(int64_t));
    if (term_l) memcpy(term_l, p.term_l, (size_t)p.terms * sizeof(int64_t));
    if (term_r) memcpy(term_r, p.term_r, (size_t)p.terms * sizeof(int64_t));
    pattern_free(&p);
    return OR_OK;
}

/* CPU accumulation of Schur values: block_schur_sparse_assembly.h:817-903 */
int or_schur_csr_values(const or_ne* eq, double damping, double minimum_diagonal, double* values)
{
    const int64_t pc = eq->pc, ec = eq->ec, ps = eq->ps, es = eq->es;
    prologue pro;
    run_prologue(eq, damping, minimum_diagonal, 0, &pro);
    if (pro.status != 0)
    {
        prologue_free(&pro);
        return or_fail(OR_RUNTIME, "eliminated block is not positive definite");
    }
    pattern p;
    build_pattern(eq, &p);
    memset(values, 0, (size_t)p.values * sizeof(double));
    for (int64_t b = 0; b < pc; ++b)
    {
        const int64_t slot = p.diagonal_slots[b];
        for (int64_t row = 0; row < ps; ++row)
            for (int64_t col = 0; col < ps; ++col)
                values[value_position(&p, ps, slot, row, col)] += pro.B[b * ps * ps + row * ps + col];
    }
    for (int64_t c = 0; c < eq->npcross; ++c)
    {
        const int64_t fwd = p.primary_pair_slots[2 * c], tr = p.primary_pair_slots[2 * c + 1];
        for (int64_t row = 0; row < ps; ++row)
            for (int64_t col = 0; col < ps; ++col)
            {
                const double v = eq->pcross_values[c * ps * ps + row * ps + col];
                values[value_position(&p, ps, fwd, row, col)] += v;
                values[value_position(&p, ps, tr, col, row)] += v;
            }
    }
    double* transformed = NULL;
    int64_t transformed_cap = 0;
    for (int64_t e = 0; e < ec; ++e)
    {
        const int64_t n = eq->adj_n[e];
        if (n * ps * es > transformed_cap)
        {
            transformed_cap = n * ps * es;
            transformed = (double*)realloc(transformed, (size_t)transformed_cap * sizeof(double));
        }
        for (int64_t l = 0; l < n; ++l)
        {
            const double* cross = eq->cross_values + eq->adj[e][l] * ps * es;
            for (int64_t row = 0; row < ps; ++row)
                mat_vec(pro.Cinv + e * es * es, es, es, cross + row * es, transformed + l * ps * es + row * es);
        }
        for (int64_t l = 0; l < n; ++l)
            for (int64_t r = 0; r < n; ++r)
            {
                const double* right = eq->cross_values + eq->adj[e][r] * ps * es;
                const int64_t slot = p.pair_slots[p.pair_slot_offsets[e] + l * n + r];
                for (int64_t row = 0; row < ps; ++row)
                    for (int64_t col = 0; col < ps; ++col)
                    {
                        double value = 0.0;
                        for (int64_t inner = 0; inner < es; ++inner)
                            value += transformed[l * ps * es + row * es + inner] * right[col * es + inner];
                        values[value_position(&p, ps, slot, row, col)] += -value;
                    }
            }
    }
    free(transformed);
    pattern_free(&p);
    prologue_free(&pro);
    return OR_OK;
}
