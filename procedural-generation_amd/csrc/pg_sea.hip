// The sea mask on the device (bake/globe/erosion/maps.py:360-421,
// refresh_base, on hydro/run.py:64-83 below_sea_components): the cells whose
// surface is below the waterline, joined 8-connected across faces through
// the owner table into components; a component is open ocean when it covers
// at least min_fraction of the globe's area (the largest always is); every
// other one is a closed basin whose cells take the basin's deepest surface as
// their base level.  Halo cells copy their owners.
//
// The reference labels components by a depth-first search in cell order
// (hydro/lakes.py:19-46).  The neighbour relation through the owner table is
// symmetric, so the components are the undirected ones and a union-find
// finds the same partition: every below-sea cell hooks the larger of its and
// a neighbour's roots under the smaller (atomicMin), rounds of hooking and
// path compression repeat until a round hooks nothing, and each component's
// root is then its lowest cell id.  Areas are summed exactly in fixed point
// (the caller's per-cell integers), floors by atomicMin of an order-preserving
// key of the float64 surface: the result does not depend on scheduling.
#include <cstring>

#include "pg_erode.h"
#include "pg_internal.h"

namespace pg {
namespace {

typedef unsigned long long ull;
typedef long long ll;

constexpr int TPB = 256;
constexpr int MAX_ROUNDS = 64;

struct SeaScalars {
    int hooked;
    int argmax;
    ull max_area;
    ll sea_cells, closed_cells, closed_basins;
    ull deepest_key;
};

struct SeaWork {
    int* parent;
    ll* area;
    ull* floor_key;
    SeaScalars* sc;
};

__device__ __forceinline__ ull dkey(double x) {
    const ull b = (ull)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_value(ull k) {
    return __longlong_as_double((ll)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
}

__device__ __forceinline__ int64_t ext_of(int64_t c, int N, int H, int NE) {
    const int64_t f = c / ((int64_t)N * N), r = c - f * N * N;
    const int64_t i = r / N, j = r - i * N;
    return (f * NE + i + H) * NE + j + H;
}

// interior id of the k-th D8 neighbour (hydro/d8.py:71-77)
__device__ __forceinline__ int64_t neighbour(const int64_t* owner, int64_t e, int k, int N, int H, int NE) {
    const int di = (k <= 1 || k == 7) ? 1 : ((k >= 3 && k <= 5) ? -1 : 0);
    const int dj = (k >= 1 && k <= 3) ? 1 : (k >= 5 ? -1 : 0);
    const int64_t o = owner[e + (int64_t)di * NE + dj];
    const int64_t f2 = o / ((int64_t)NE * NE), r2 = o - f2 * NE * NE;
    return (f2 * N + r2 / NE - H) * N + (r2 % NE) - H;
}

__device__ __forceinline__ int find_root(const int* parent, int x) {
    int p = parent[x];
    while (p != x) {  // parent ids only decrease: the walk ends
        x = p;
        p = parent[x];
    }
    return x;
}

__global__ void sea_init_kernel(int64_t M, int N, int H, int NE, const double* height, const double* sediment, SeaWork w) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = ext_of(c, N, H, NE);
        // the waterline test on the float32 surface (maps.py:397)
        const bool below = (float)(height[e] + sediment[e]) < 0.0f;
        w.parent[c] = below ? (int)c : -1;
        w.area[c] = 0;
        w.floor_key[c] = dkey(0.0);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        w.sc->hooked = 0;
        w.sc->argmax = 0x7fffffff;
        w.sc->max_area = 0;
        w.sc->sea_cells = w.sc->closed_cells = w.sc->closed_basins = 0;
        w.sc->deepest_key = dkey(0.0);
    }
}

__global__ void sea_flag_kernel(SeaScalars* sc) { sc->hooked = 0; }

__global__ void sea_hook_kernel(int64_t M, int N, int H, int NE, const int64_t* owner, SeaWork w) {
    int hooked = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (int64_t)gridDim.x * blockDim.x) {
        if (w.parent[c] < 0) continue;
        const int64_t e = ext_of(c, N, H, NE);
        for (int k = 0; k < 8; k++) {
            const int64_t q = neighbour(owner, e, k, N, H, NE);
            if (w.parent[q] < 0) continue;
            const int ru = find_root(w.parent, (int)c), rv = find_root(w.parent, (int)q);
            if (ru == rv) continue;
            const int lo = ru < rv ? ru : rv, hi = ru < rv ? rv : ru;
            if (atomicMin(&w.parent[hi], lo) > lo) hooked = 1;
        }
    }
    if (__any(hooked) && (threadIdx.x & 63) == 0) atomicOr(&w.sc->hooked, 1);
}

__global__ void sea_compress_kernel(int64_t M, SeaWork w) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (int64_t)gridDim.x * blockDim.x)
        if (w.parent[c] >= 0) w.parent[c] = find_root(w.parent, (int)c);
}

// per component: the exact area and the lowest float64 surface
__global__ void sea_reduce_kernel(int64_t M, int N, int H, int NE, const double* height, const double* sediment,
                                  const int64_t* area_fx, SeaWork w) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (int64_t)gridDim.x * blockDim.x) {
        const int r = w.parent[c];
        if (r < 0) continue;
        const int64_t e = ext_of(c, N, H, NE);
        atomicAdd((ull*)&w.area[r], (ull)area_fx[c]);
        atomicMin(&w.floor_key[r], dkey(height[e] + sediment[e]));
    }
}

__global__ void sea_max_kernel(int64_t M, SeaWork w) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (int64_t)gridDim.x * blockDim.x)
        if (w.parent[c] == (int)c) atomicMax(&w.sc->max_area, (ull)w.area[c]);
}

// the largest component (the lowest root among equals, numpy's argmax)
__global__ void sea_argmax_kernel(int64_t M, SeaWork w) {
    const ull mx = w.sc->max_area;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (int64_t)gridDim.x * blockDim.x)
        if (w.parent[c] == (int)c && (ull)w.area[c] == mx) atomicMin(&w.sc->argmax, (int)c);
}

__global__ void sea_base_kernel(int64_t M, int N, int H, int NE, double threshold, double* base, SeaWork w) {
    const int amax = w.sc->argmax;
    ll sea = 0, closed = 0, basins = 0;
    ull deepest = dkey(0.0);
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < M; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = ext_of(c, N, H, NE);
        const int r = w.parent[c];
        double b = 0.0;
        if (r >= 0) {
            const bool keep = (double)w.area[r] >= threshold || r == amax;
            if (keep) {
                sea += 1;
            } else {
                closed += 1;
                b = dkey_value(w.floor_key[r]);
                if (r == (int)c) {
                    basins += 1;
                    deepest = w.floor_key[r] < deepest ? w.floor_key[r] : deepest;
                }
            }
        }
        base[e] = b;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sea += __shfl_xor(sea, o);
        closed += __shfl_xor(closed, o);
        basins += __shfl_xor(basins, o);
        const ull d = __shfl_xor(deepest, o);
        deepest = d < deepest ? d : deepest;
    }
    if ((threadIdx.x & 63) == 0) {
        if (sea) atomicAdd((ull*)&w.sc->sea_cells, (ull)sea);
        if (closed) atomicAdd((ull*)&w.sc->closed_cells, (ull)closed);
        if (basins) atomicAdd((ull*)&w.sc->closed_basins, (ull)basins);
        atomicMin(&w.sc->deepest_key, deepest);
    }
}

// halo cells take their owner's value (maps.py:414-416)
__global__ void sea_halo_kernel(int64_t n_ext, const int64_t* owner, double* base) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_ext; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = owner[e];
        if (o != e) base[e] = base[o];
    }
}

inline size_t up(size_t x) { return (x + 255) & ~(size_t)255; }

int blocks_of(int64_t n) {
    const int64_t b = (n + TPB - 1) / TPB;
    return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" int64_t pg_sea_work_bytes(int64_t N) {
    const size_t M = (size_t)(6 * N * N);
    return (int64_t)(up(4 * M) + up(8 * M) + up(8 * M) + up(sizeof(SeaScalars)));
}

extern "C" int pg_refresh_base(const double* height, const double* sediment, const int64_t* owner, const int64_t* area_fx,
                               double threshold_fx, int64_t N, int64_t H, double* base, void* work, int64_t work_bytes,
                               double* info5, pg_stream stream) {
    clear_error();
    PG_REQUIRE(height && sediment && owner && area_fx && base && work && info5, "pg_refresh_base: NULL argument");
    PG_REQUIRE(N >= 2 && H >= 1, "pg_refresh_base: bad grid N=%lld H=%lld", (long long)N, (long long)H);
    PG_REQUIRE(6 * N * N < (1ll << 31), "pg_refresh_base: grid too large for int32 cell ids");
    PG_REQUIRE(work_bytes >= pg_sea_work_bytes(N), "pg_refresh_base: work of %lld bytes, need %lld",
               (long long)work_bytes, (long long)pg_sea_work_bytes(N));
    const int64_t M = 6 * N * N, NE = N + 2 * H;
    char* p = (char*)work;
    SeaWork w;
    w.parent = (int*)p;
    p += up(4 * (size_t)M);
    w.area = (ll*)p;
    p += up(8 * (size_t)M);
    w.floor_key = (ull*)p;
    p += up(8 * (size_t)M);
    w.sc = (SeaScalars*)p;
    hipStream_t s = (hipStream_t)stream;
    const int nb = blocks_of(M);
    hipLaunchKernelGGL(sea_init_kernel, dim3(nb), dim3(TPB), 0, s, M, (int)N, (int)H, (int)NE, height, sediment, w);
    PG_LAUNCHED("sea_init_kernel");
    int hooked = 1, rounds = 0;
    while (hooked) {
        PG_REQUIRE(rounds < MAX_ROUNDS, "pg_refresh_base: components did not settle in %d rounds", MAX_ROUNDS);
        hipLaunchKernelGGL(sea_flag_kernel, dim3(1), dim3(1), 0, s, w.sc);
        hipLaunchKernelGGL(sea_hook_kernel, dim3(nb), dim3(TPB), 0, s, M, (int)N, (int)H, (int)NE, owner, w);
        hipLaunchKernelGGL(sea_compress_kernel, dim3(nb), dim3(TPB), 0, s, M, w);
        PG_LAUNCHED("sea_hook_kernel");
        PG_HIP(hipMemcpyAsync(&hooked, &w.sc->hooked, sizeof(int), hipMemcpyDeviceToHost, s));
        PG_HIP(hipStreamSynchronize(s));
        rounds += 1;
    }
    hipLaunchKernelGGL(sea_reduce_kernel, dim3(nb), dim3(TPB), 0, s, M, (int)N, (int)H, (int)NE, height, sediment,
                       area_fx, w);
    hipLaunchKernelGGL(sea_max_kernel, dim3(nb), dim3(TPB), 0, s, M, w);
    hipLaunchKernelGGL(sea_argmax_kernel, dim3(nb), dim3(TPB), 0, s, M, w);
    hipLaunchKernelGGL(sea_base_kernel, dim3(nb), dim3(TPB), 0, s, M, (int)N, (int)H, (int)NE, threshold_fx, base, w);
    hipLaunchKernelGGL(sea_halo_kernel, dim3(blocks_of(6 * NE * NE)), dim3(TPB), 0, s, 6 * NE * NE, owner, base);
    PG_LAUNCHED("sea_base_kernel");
    SeaScalars h;
    PG_HIP(hipMemcpyAsync(&h, w.sc, sizeof(SeaScalars), hipMemcpyDeviceToHost, s));
    PG_HIP(hipStreamSynchronize(s));
    info5[0] = (double)h.closed_cells;
    info5[1] = (double)h.closed_basins;
    info5[2] = (double)h.sea_cells;
    const ull k = h.deepest_key;
    const ull b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double d;
    memcpy(&d, &b, sizeof(d));
    info5[3] = d;
    info5[4] = (double)rounds;
    return 0;
}
