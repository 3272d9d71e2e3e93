// SpMV autotuner: times the wave-stream kernel and every lanes-per-row variant
// on the caller's matrix
// with HIP events and keeps the fastest in memory.  Only when the environment
// variable PLAMATRIX_AMD_TUNE_CACHE names a file is the table also read from
// and written to that file, so later processes start tuned; without it the
// library touches no file.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include <sys/stat.h>

#include "spmv_kernels.hpp"
#include "tuning.hpp"

namespace pm
{

namespace
{
using Key = std::tuple<long long, long long, int, int>;
std::mutex g_mutex;
std::map<Key, int> g_table;
bool g_loaded = false;

Key key_of(const pm_csr* a)
{
    return Key(static_cast<long long>(a->rows), static_cast<long long>(a->nnz), a->dtype, use_compact(a) ? 32 : 64);
}

std::string cache_path()
{
    const char* p = std::getenv("PLAMATRIX_AMD_TUNE_CACHE");
    return p != nullptr ? std::string(p) : std::string();
}

bool valid_lanes(int lanes) { return lanes >= 0 && lanes <= 64 && (lanes & (lanes - 1)) == 0; }

void load_locked()
{
    if (g_loaded) return;
    g_loaded = true;
    const std::string path = cache_path();
    if (path.empty()) return;
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return;
    long long rows, nnz;
    int dtype, width, lanes;
    // entries that no launch could use (a stale or damaged file) are dropped
    while (std::fscanf(f, "%lld %lld %d %d %d", &rows, &nnz, &dtype, &width, &lanes) == 5)
        if (valid_lanes(lanes) && rows >= 0 && nnz >= 0 && (dtype == PM_F32 || dtype == PM_F64) &&
            (width == 32 || width == 64))
            g_table[Key(rows, nnz, dtype, width)] = lanes;
    std::fclose(f);
}

void save_locked()
{
    const std::string path = cache_path();
    if (path.empty()) return;
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos)
    {
        std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); ++i)
            if (i == dir.size() || dir[i] == '/') ::mkdir(dir.substr(0, i).c_str(), 0755);
    }
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return;
    for (const auto& [k, lanes] : g_table)
        std::fprintf(f, "%lld %lld %d %d %d\n", std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k), lanes);
    std::fclose(f);
}
} // namespace

int tuned_lanes(const pm_csr* a)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    load_locked();
    auto it = g_table.find(key_of(a));
    if (it != g_table.end()) return it->second;
    const double mean = a->rows > 0 ? static_cast<double>(a->nnz) / static_cast<double>(a->rows) : 0.0;
    return mean <= 32.0 ? 0 : heuristic_lanes(a->rows, a->nnz);
}

void record_lanes(const pm_csr* a, int lanes)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    load_locked();
    g_table[key_of(a)] = lanes;
    save_locked();
}

} // namespace pm

using namespace pm;

extern "C" {

// Time every lanes-per-row variant (median of `repeats` launches each) on
// y = A x and keep the fastest for this matrix shape.  Synchronises `stream`.
int pm_autotune_spmv(const pm_csr* a, const void* x, void* y, int32_t repeats, int32_t* best_lanes,
                     float* best_ms, hipStream_t stream)
{
    PM_TRY(check_csr(a, "autotune"));
    if (a->rows == 0) return fail(PM_INVALID_ARGUMENT, "autotune: empty matrix");
    if (repeats <= 0) repeats = 10;
    hipEvent_t start, stop;
    PM_HIP(hipEventCreate(&start));
    PM_HIP(hipEventCreate(&stop));
    int best = 0;
    float best_time = 1e30f;
    for (int lanes = 0; lanes <= 64; lanes = lanes == 0 ? 1 : lanes * 2)
    {
        PM_TRY(spmv_launch(a, x, y, lanes, stream));  // warm-up
        float times[64];
        const int n = repeats < 64 ? repeats : 64;
        for (int rep = 0; rep < n; ++rep)
        {
            PM_HIP(hipEventRecord(start, stream));
            PM_TRY(spmv_launch(a, x, y, lanes, stream));
            PM_HIP(hipEventRecord(stop, stream));
            PM_HIP(hipEventSynchronize(stop));
            PM_HIP(hipEventElapsedTime(&times[rep], start, stop));
        }
        for (int i = 1; i < n; ++i)
            for (int j = i; j > 0 && times[j] < times[j - 1]; --j) std::swap(times[j], times[j - 1]);
        const float median = times[n / 2];
        if (median < best_time)
        {
            best_time = median;
            best = lanes;
        }
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    record_lanes(a, best);
    if (best_lanes) *best_lanes = best;
    if (best_ms) *best_ms = best_time;
    return PM_OK;
}

int pm_tune_record(const pm_csr* a, int32_t lanes)
{
    PM_TRY(check_csr(a, "tune record"));
    if (!valid_lanes(lanes))
        return fail(PM_INVALID_ARGUMENT, "lanes per row must be 0 (wave-stream) or a power of two in [1, 64], got %d",
                    lanes);
    record_lanes(a, lanes);
    return PM_OK;
}

int pm_tuned_lanes(const pm_csr* a)
{
    if (check_csr(a, "tuned lanes") != PM_OK) return -1;
    return tuned_lanes(a);
}

} // extern "C"
