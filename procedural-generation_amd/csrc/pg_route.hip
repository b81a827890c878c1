// The routing surface: priority flood + epsilon (bake/globe/erosion/route.py:
// 30-117) on the device.
//
// The flood assigns every cell it reaches W(n) = max(s(n), W(p) + eps), p the
// first popped cell with an edge p -> n.  Pops come in non-decreasing order,
// so W(p) is the least W over all reached in-neighbours of n, and W is the
// fixed point of
//     W(seed) = s(seed),   W(n) = max(s(n), fl(min_{q -> n} W(q) + eps))
// (fl(x + eps) is monotone in x, so the min commutes with the rounding): the
// same bits, whatever order the cells are settled in.  The kernel relaxes
// that system in place from W = +inf -- every thread owns one cell, reads its
// in-neighbours' current values (any value read is >= the fixed point, so a
// race only delays convergence) and lowers its own -- until a sweep changes
// nothing.  Unreached cells keep the surface, as in the flood.
//
// In-edges: a cell's D8 neighbours on its own face (the owner table is the
// identity there), plus, for the few cells a cross-face halo cell maps onto
// (the owner table of the cube-sphere), an explicit list built on the host.
#include "pg_erode.h"
#include "pg_internal.h"

namespace pg {
namespace {

constexpr int SX = 64, SY = 4;
enum : uint8_t { R_NONE = 0, R_SEED = 1, R_FREE = 2, R_SPECIAL = 3 };

__device__ __forceinline__ int d8i(int k) { return (k <= 1 || k == 7) ? 1 : ((k >= 3 && k <= 5) ? -1 : 0); }
__device__ __forceinline__ int d8j(int k) { return (k >= 1 && k <= 3) ? 1 : (k >= 5 ? -1 : 0); }

inline dim3 grid3(int64_t F, int64_t NE) {
    return dim3((unsigned)((NE + SX - 1) / SX), (unsigned)((NE + SY - 1) / SY), (unsigned)F);
}

// W = s on seeds (static ones, and interior cells of the open sea / a basin
// floor: s < 0 and s <= base), +inf elsewhere
__global__ void flood_init_kernel(int64_t NE, int64_t H, int64_t N, const double* height, const double* sediment,
                                  const double* base, const uint8_t* cls, double* s_out, double* W, int* flag) {
    const int j = blockIdx.x * SX + threadIdx.x, i = blockIdx.y * SY + threadIdx.y;
    if (i >= NE || j >= NE) return;
    const int64_t c = ((int64_t)blockIdx.z * NE + i) * NE + j;
    const double s = height[c] + sediment[c];
    s_out[c] = s;
    const uint8_t k = cls[c];
    const bool inter = i >= H && i < H + N && j >= H && j < H + N;
    bool seed = k == R_SEED;
    if (!seed && (k == R_FREE || k == R_SPECIAL) && inter) seed = s < 0.0 && s <= base[c];
    W[c] = seed ? s : INFINITY;
    if (c == 0) *flag = 0;
}

__device__ __forceinline__ int64_t find_special(const int64_t* cells, int64_t n, int64_t c) {
    int64_t lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int64_t m = (lo + hi) >> 1;
        const int64_t v = cells[m];
        if (v == c) return m;
        if (v < c) lo = m + 1;
        else hi = m - 1;
    }
    return -1;
}

// one in-place relaxation sweep; `flag` set when a value changed and
// `record` (the last sweep of a batch) is set
__global__ void flood_sweep_kernel(int64_t NE, const double* s, const uint8_t* cls, double* W, double eps,
                                   const int64_t* spec_cell, const int32_t* spec_ptr, const int64_t* spec_src,
                                   int64_t n_spec, int* flag, int record) {
    const int j = blockIdx.x * SX + threadIdx.x, i = blockIdx.y * SY + threadIdx.y;
    if (i >= NE || j >= NE) return;
    const int64_t c = ((int64_t)blockIdx.z * NE + i) * NE + j;
    const uint8_t k = cls[c];
    if (k != R_FREE && k != R_SPECIAL) return;
    const double cur = W[c];
    const double sc = s[c];
    if (cur == sc) return;  // a seed of this refresh or already at its surface: cannot go lower
    double m = INFINITY;
    // free cells are never on the outer ring, so the 8 neighbours are in the array
    for (int d = 0; d < 8; d++) {
        const double w = W[c + (int64_t)d8i(d) * NE + d8j(d)];
        m = w < m ? w : m;
    }
    if (k == R_SPECIAL) {
        const int64_t at = find_special(spec_cell, n_spec, c);
        if (at >= 0)
            for (int32_t e = spec_ptr[at]; e < spec_ptr[at + 1]; e++) {
                const double w = W[spec_src[e]];
                m = w < m ? w : m;
            }
    }
    if (m == INFINITY) return;
    const double up = m + eps;
    const double v = sc < up ? up : sc;
    if (v < cur) {
        W[c] = v;
        if (record) *flag = 1;
    }
}

__global__ void flag_reset_kernel(int* flag) { *flag = 0; }

// unreached cells (and those the flood never assigns) keep the surface
__global__ void flood_finish_kernel(int64_t n, const double* s, const double* W, double* route) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const double w = W[c];
        route[c] = w == INFINITY ? s[c] : w;
    }
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" int pg_flood_init(int64_t F, int64_t NE, int64_t H, int64_t N, const double* height, const double* sediment,
                             const double* base, const uint8_t* cls, double* s_out, double* W, int32_t* flag,
                             pg_stream stream) {
    clear_error();
    PG_REQUIRE(F > 0 && NE >= 3 && H >= 0 && N > 0 && N + 2 * H == NE, "pg_flood_init: bad sizes");
    PG_REQUIRE(height && sediment && base && cls && s_out && W && flag, "pg_flood_init: NULL array");
    hipLaunchKernelGGL(flood_init_kernel, grid3(F, NE), dim3(SX, SY), 0, (hipStream_t)stream, NE, H, N, height,
                       sediment, base, cls, s_out, W, (int*)flag);
    PG_LAUNCHED("flood_init_kernel");
    return 0;
}

extern "C" int pg_flood_sweeps(int64_t F, int64_t NE, const double* s, const uint8_t* cls, double* W, double eps,
                               const int64_t* spec_cell, const int32_t* spec_ptr, const int64_t* spec_src,
                               int64_t n_spec, int32_t sweeps, int32_t* flag, pg_stream stream) {
    clear_error();
    PG_REQUIRE(F > 0 && NE >= 3 && s && cls && W && flag && sweeps > 0, "pg_flood_sweeps: bad arguments");
    PG_REQUIRE(n_spec == 0 || (spec_cell && spec_ptr && spec_src), "pg_flood_sweeps: NULL special-cell table");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(flag_reset_kernel, dim3(1), dim3(1), 0, st, (int*)flag);
    for (int32_t r = 0; r < sweeps; r++)
        hipLaunchKernelGGL(flood_sweep_kernel, grid3(F, NE), dim3(SX, SY), 0, st, NE, s, cls, W, eps, spec_cell, spec_ptr,
                           spec_src, n_spec, (int*)flag, (int)(r == sweeps - 1));
    PG_LAUNCHED("flood_sweep_kernel");
    return 0;
}

extern "C" int pg_flood_finish(int64_t n, const double* s, const double* W, double* route, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n > 0 && s && W && route, "pg_flood_finish: bad arguments");
    int blocks = (int)((n + 255) / 256);
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(flood_finish_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n, s, W, route);
    PG_LAUNCHED("flood_finish_kernel");
    return 0;
}
