// Calibration probe: the GPU's rate of relaxed device-scope 64-bit atomic
// adds, the resource the fast-mode trace kernel is bound by.  Each lane walks
// a cell index space as a droplet does -- one of the 8 neighbours per step on
// an NE-wide grid -- and at every step adds to the cell in `planes` separate
// int64 planes (the applying pass adds to four: height change, discharge,
// two momentum components).  Pattern 1 replaces the walk by uniformly random
// cells (no locality at all); pattern h >= 2 by a set of h hot cells (many
// lanes on one address, as on a river).  Not part of the erosion path.
#include "pg_erode.h"
#include "pg_internal.h"

namespace pg {
namespace {

typedef unsigned long long ull;

__device__ __forceinline__ ull mix64(ull x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

__global__ void atomic_probe_kernel(ull* t, int64_t n, int64_t NE, int planes, int steps, int pattern, ull seed) {
    const ull g = (ull)blockIdx.x * blockDim.x + threadIdx.x;
    ull st = mix64(seed ^ (g * 0x9e3779b97f4a7c15ull));
    int64_t pos = (int64_t)(st % (ull)(n - 2 * NE - 2)) + NE + 1;
    for (int s = 0; s < steps; s++) {
        st = mix64(st + 0x9e3779b97f4a7c15ull);
        if (pattern == 1) {
            pos = (int64_t)(st % (ull)n);
        } else if (pattern >= 2) {
            pos = (int64_t)(st % (ull)pattern);  // `pattern` hot cells: same-address traffic
        } else {
            const int k = (int)(st & 7);
            const int64_t di = (k <= 1 || k == 7) ? 1 : ((k >= 3 && k <= 5) ? -1 : 0);
            const int64_t dj = (k >= 1 && k <= 3) ? 1 : (k >= 5 ? -1 : 0);
            pos += di * NE + dj;
            if (pos < NE + 1 || pos >= n - NE - 1) pos = (int64_t)((st >> 8) % (ull)(n - 2 * NE - 2)) + NE + 1;
        }
        for (int p = 0; p < planes; p++) atomicAdd(&t[(int64_t)p * n + pos], 1ull);
    }
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" int pg_atomic_probe(int64_t* targets, int64_t n, int64_t NE, int32_t planes, int64_t lanes, int32_t steps,
                               int32_t pattern, uint64_t seed, pg_stream stream) {
    clear_error();
    PG_REQUIRE(targets && NE >= 3 && n > 4 * NE && planes >= 1 && planes <= 8 && lanes > 0 && steps > 0,
               "pg_atomic_probe: bad arguments");
    PG_REQUIRE(pattern >= 0 && pattern < n, "pg_atomic_probe: pattern must be 0 (walk), 1 (random) or a hot-set size < n");
    const int64_t blocks = (lanes + 255) / 256;
    PG_REQUIRE(blocks <= 0x7fffffff, "pg_atomic_probe: too many lanes");
    hipLaunchKernelGGL(atomic_probe_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (ull*)targets, n, NE, (int)planes, (int)steps, (int)pattern, (ull)seed);
    PG_LAUNCHED("atomic_probe_kernel");
    return 0;
}
