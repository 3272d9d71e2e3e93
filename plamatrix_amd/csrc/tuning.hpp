// Lanes-per-row autotuning table shared by SpMV and the solver kernels.
#pragma once

#include "pm_common.hpp"

namespace pm
{

// SpMV variant for this matrix shape (rows, nnz, dtype, index width): 0 is the
// wave-stream kernel, 1..64 the lanes-per-row kernel; untuned shapes use the
// wave-stream kernel unless rows are long on average.
int tuned_lanes(const pm_csr* a);
void record_lanes(const pm_csr* a, int lanes);

int check_csr(const pm_csr* a, const char* op);
bool use_compact(const pm_csr* a);
int spmv_launch(const pm_csr* a, const void* x, void* y, int lanes, hipStream_t stream);
int relres_launch(const pm_csr* a, const void* b, const void* x, int32_t vec_dtype, double* partials,
                  unsigned* counter, double* out, hipStream_t stream);

} // namespace pm
