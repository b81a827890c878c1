// Per-cell passes of an erosion iteration: packing the per-cell record,
// folding the net change, the discharge/momentum EMA, uplift, the halo
// exchange, spawn-cell search and the exact order statistic of the datum
// hold.  All are bandwidth-bound elementwise or gather kernels: grid-stride
// loops over cells, 256 threads per block, float64 state.
//
// Bit parity: each output is one IEEE operation chain in the reference's
// order (built with -ffp-contract=off), so fold / EMA / uplift / pack agree
// with the reference bit for bit.  The halo gather sums its 16 sources in
// source order; the reference's numpy einsum does not state its order, so
// halos agree to rounding, not bit for bit.
#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"

namespace pg {
namespace {

constexpr int TPB = 256;

inline int grid_for(int64_t n) {
    int b = blocks_for(n, TPB);
    return b < 1 ? 1 : (b > 65535 ? 65535 : b);
}

#define GRID_STRIDE(i, n) for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// particle.py:164-191
__global__ void pack_kernel(float* samp, int64_t n, const double* height, const double* sediment, const double* route,
                            int use_route, const double* discharge, const double* momentum, const float* evap,
                            const float* hardness, const float* metric, const float* metric_inv, const double* base,
                            const uint8_t* lake_flag) {
    GRID_STRIDE(c, n) {
        float* s = samp + c * NS;
        const double sd = sediment[c];
        const double surf = height[c] + sd;
        const float fs = (float)surf;
        const float fr = (float)(use_route ? route[c] : surf);
        float4* s4 = reinterpret_cast<float4*>(s);
        s4[0] = make_float4(fs, fr, (float)discharge[c], (float)momentum[2 * c]);
        s4[1] = make_float4((float)momentum[2 * c + 1], evap[c], hardness[c], (float)sd);
        const uint8_t lf = lake_flag ? lake_flag[c] : 0;
        float flag;
        if (use_route && lf > 0) flag = lf > 1 ? 3.0f : 2.0f;
        else flag = (use_route && fr != fs) ? 1.0f : 0.0f;
        s4[2] = make_float4(metric[3 * c], metric[3 * c + 1], metric[3 * c + 2], metric_inv[3 * c]);
        s4[3] = make_float4(metric_inv[3 * c + 1], metric_inv[3 * c + 2], flag, (float)base[c]);
    }
}

// particle.py:1079-1101
__global__ void fold_kernel(int64_t n, double* height, double* sediment, double* acc, const uint8_t* mask) {
    GRID_STRIDE(c, n) {
        const double a = acc[c];
        acc[c] = 0.0;
        if (a == 0.0 || mask[c] != MASK_ACTIVE) continue;
        if (a > 0.0) {
            sediment[c] += a;
        } else {
            const double take = -a, s = sediment[c];
            if (s >= take) {
                sediment[c] = s - take;
            } else {
                height[c] -= take - s;
                sediment[c] = 0.0;
            }
        }
    }
}

// particle.py:1230-1242
__global__ void ema_kernel(int64_t n, double* discharge, double* momentum, double* disch_track, double* mom_track,
                           const uint8_t* mask, double ema) {
    GRID_STRIDE(c, n) {
        if (mask[c] == MASK_ACTIVE) {
            discharge[c] = (1.0 - ema) * discharge[c] + ema * disch_track[c];
            momentum[2 * c] = (1.0 - ema) * momentum[2 * c] + ema * mom_track[2 * c];
            momentum[2 * c + 1] = (1.0 - ema) * momentum[2 * c + 1] + ema * mom_track[2 * c + 1];
        }
        disch_track[c] = 0.0;
        mom_track[2 * c] = 0.0;
        mom_track[2 * c + 1] = 0.0;
    }
}

// maps.py:1132-1134 (the cap taken before the caller's mean, maps.py:1142)
__global__ void uplift_kernel(int64_t n, double* height, const double* uplift, const uint8_t* mask, double cap,
                              int use_cap, double mean) {
    GRID_STRIDE(c, n) {
        if (mask[c] != MASK_ACTIVE) continue;
        double u = uplift[c];
        if (use_cap) u = u < cap ? u : cap;
        height[c] += u - mean;
    }
}

// particle.py:1245-1260: the reference walks k forward over sorted targets;
// the result is the first index with cdf[k] > target, capped at M - 1,
// which a binary search finds independently per droplet.
__global__ void spawn_cells_kernel(const double* cdf, int64_t M, const double* u, int64_t P, int64_t* out) {
    const double total = cdf[M - 1];
    GRID_STRIDE(p, P) {
        const double target = u[p] * total;
        int64_t lo = 0, hi = M - 1;  // answer in [lo, hi]
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (cdf[mid] <= target) lo = mid + 1;
            else hi = mid;
        }
        out[p] = lo;
    }
}

// field.py:104-124: halo cells from their 16 interior sources
__global__ void halo_scalar_kernel(double* data, int64_t K, int64_t S, const int64_t* dst, const int64_t* src,
                                   const double* w) {
    GRID_STRIDE(k, K) {
        double acc = 0.0;
        for (int64_t s = 0; s < S; s++) acc += w[k * S + s] * data[src[k * S + s]];
        data[dst[k]] = acc;
    }
}

// one thread per source cell: the contributions of every halo entry that
// reads it, summed in entry order (deterministic), then added to the cell
__global__ void halo_scalar_adjoint_gather_kernel(double* adj, int64_t S, const int64_t* dst, const double* w,
                                                  int64_t n_src, const int64_t* src_cell, const int64_t* src_start,
                                                  const int64_t* src_entry) {
    GRID_STRIDE(i, n_src) {
        double v = 0.0;
        for (int64_t j = src_start[i]; j < src_start[i + 1]; j++) {
            const int64_t e = src_entry[j];
            v += w[e] * adj[dst[e / S]];
        }
        if (v != 0.0) adj[src_cell[i]] += v;
    }
}

__global__ void halo_clear_kernel(double* adj, int64_t K, const int64_t* dst) {
    GRID_STRIDE(k, K) { adj[dst[k]] = 0.0; }
}

__global__ void halo_vector_kernel(double* data, int64_t K, int64_t S, const int64_t* dst, const int64_t* src,
                                   const double* w, const double* rot) {
    GRID_STRIDE(k, K) {
        double a0 = 0.0, a1 = 0.0;
        for (int64_t s = 0; s < S; s++) {
            const int64_t c = src[k * S + s];
            const double x0 = data[2 * c], x1 = data[2 * c + 1], ws = w[k * S + s];
            const double* R = rot + (k * S + s) * 4;  // (2, 2) row-major: R[a][b]
            a0 += ws * (R[0] * x0 + R[1] * x1);
            a1 += ws * (R[2] * x0 + R[3] * x1);
        }
        data[2 * dst[k]] = a0;
        data[2 * dst[k] + 1] = a1;
    }
}

// ---------------------------------------------------------------------------
// exact k-th order statistic: 8-bit radix select over order-preserving keys
// ---------------------------------------------------------------------------
struct SelectWork {
    unsigned long long hist[256];
    unsigned long long prefix, known;  // key bits fixed so far, and which
    long long k_rem;
};

__device__ __forceinline__ unsigned long long order_key(double x) {
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double key_value(unsigned long long k) {
    unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

__global__ void select_init_kernel(SelectWork* w, long long k) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) w->hist[i] = 0;
    if (threadIdx.x == 0) {
        w->prefix = 0;
        w->known = 0;
        w->k_rem = k;
    }
}

__global__ void select_hist_kernel(const double* height, const double* sediment, int64_t F, int64_t N, int64_t H,
                                   int shift, SelectWork* w) {
    __shared__ unsigned int h[256];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const unsigned long long prefix = w->prefix, known = w->known;
    const int64_t NE = N + 2 * H, M = F * N * N;
    GRID_STRIDE(c, M) {
        const int64_t f = c / (N * N), r = c - f * N * N, i = r / N, j = r - i * N;
        const int64_t e = (f * NE + i + H) * NE + j + H;
        const unsigned long long key = order_key(height[e] + sediment[e]);
        if ((key & known) == prefix) atomicAdd(&h[(key >> shift) & 255], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 256; i += blockDim.x)
        if (h[i]) atomicAdd(&w->hist[i], (unsigned long long)h[i]);
}

__global__ void select_pick_kernel(SelectWork* w, int shift, double* q_out) {
    if (threadIdx.x != 0) return;
    long long k = w->k_rem;
    int b = 0;
    for (; b < 255; b++) {
        const long long c = (long long)w->hist[b];
        if (k < c) break;
        k -= c;
    }
    w->k_rem = k;
    w->prefix |= (unsigned long long)b << shift;
    w->known |= 255ull << shift;
    for (int i = 0; i < 256; i++) w->hist[i] = 0;
    if (shift == 0) *q_out = key_value(w->prefix);
}

// maps.py:1496-1507
__global__ void datum_shift_kernel(int64_t n, double* height, double* route, double* base, const double* qp) {
    const double q = *qp;
    if (q == 0.0) return;
    GRID_STRIDE(c, n) {
        height[c] -= q;
        if (route) route[c] -= q;
        if (base && base[c] < 0.0) base[c] -= q;
    }
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" int pg_pack_samples(float* samp, int64_t F, int64_t NE, const double* height, const double* sediment,
                               const double* route, int use_route, const double* discharge, const double* momentum,
                               const float* evap, const float* hardness, const float* metric, const float* metric_inv,
                               const double* base, const uint8_t* lake_flag, pg_stream stream) {
    clear_error();
    PG_REQUIRE(F > 0 && NE > 1, "pg_pack_samples: bad sizes");
    PG_REQUIRE(samp && height && sediment && discharge && momentum && evap && hardness && metric && metric_inv && base,
               "pg_pack_samples: NULL array");
    PG_REQUIRE(!use_route || route, "pg_pack_samples: use_route without route");
    PG_REQUIRE(((uintptr_t)samp & 15) == 0, "pg_pack_samples: samp must be 16-byte aligned");
    const int64_t n = F * NE * NE;
    hipLaunchKernelGGL(pack_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, samp, n, height, sediment,
                       route, use_route, discharge, momentum, evap, hardness, metric, metric_inv, base, lake_flag);
    PG_LAUNCHED("pack_kernel");
    return 0;
}

extern "C" int pg_fold_changes(int64_t n, double* height, double* sediment, double* acc, const uint8_t* mask, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && height && sediment && acc && mask, "pg_fold_changes: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(fold_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, height, sediment, acc, mask);
    PG_LAUNCHED("fold_kernel");
    return 0;
}

extern "C" int pg_ema_update(int64_t n, double* discharge, double* momentum, double* disch_track, double* mom_track,
                             const uint8_t* mask, double ema, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && discharge && momentum && disch_track && mom_track && mask, "pg_ema_update: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(ema_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, discharge, momentum,
                       disch_track, mom_track, mask, ema);
    PG_LAUNCHED("ema_kernel");
    return 0;
}

extern "C" int pg_apply_uplift(int64_t n, double* height, const double* uplift, const uint8_t* mask, double cap,
                               int use_cap, double mean, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && height && uplift && mask, "pg_apply_uplift: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(uplift_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, height, uplift, mask,
                       cap, use_cap, mean);
    PG_LAUNCHED("uplift_kernel");
    return 0;
}

extern "C" int pg_spawn_cells(const double* cdf, int64_t M, const double* u, int64_t P, int64_t* out, pg_stream stream) {
    clear_error();
    PG_REQUIRE(M > 0 && P >= 0, "pg_spawn_cells: bad sizes M=%lld P=%lld", (long long)M, (long long)P);
    if (P == 0) return 0;
    PG_REQUIRE(cdf && u && out, "pg_spawn_cells: NULL array");
    hipLaunchKernelGGL(spawn_cells_kernel, dim3(grid_for(P)), dim3(TPB), 0, (hipStream_t)stream, cdf, M, u, P, out);
    PG_LAUNCHED("spawn_cells_kernel");
    return 0;
}

extern "C" int pg_exchange_halos(double* data, int64_t ncomp, int64_t K, int64_t S, const int64_t* dst,
                                 const int64_t* src, const double* w, const double* rot, pg_stream stream) {
    clear_error();
    PG_REQUIRE(K >= 0 && S > 0 && data && dst && src && w, "pg_exchange_halos: bad arguments");
    PG_REQUIRE((ncomp == 1 && !rot) || (ncomp == 2 && rot), "pg_exchange_halos: ncomp 1 needs rot NULL, ncomp 2 needs rot");
    if (K == 0) return 0;
    if (ncomp == 1)
        hipLaunchKernelGGL(halo_scalar_kernel, dim3(grid_for(K)), dim3(TPB), 0, (hipStream_t)stream, data, K, S, dst, src, w);
    else
        hipLaunchKernelGGL(halo_vector_kernel, dim3(grid_for(K)), dim3(TPB), 0, (hipStream_t)stream, data, K, S, dst, src, w, rot);
    PG_LAUNCHED("halo_kernel");
    return 0;
}

extern "C" int pg_halo_scalar_adjoint(double* adj, int64_t K, int64_t S, const int64_t* dst, const double* w,
                                      int64_t n_src, const int64_t* src_cell, const int64_t* src_start,
                                      const int64_t* src_entry, pg_stream stream) {
    clear_error();
    PG_REQUIRE(K >= 0 && S > 0 && n_src >= 0 && adj && dst && w && src_cell && src_start && src_entry,
               "pg_halo_scalar_adjoint: bad arguments");
    if (K == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (n_src > 0) {
        hipLaunchKernelGGL(halo_scalar_adjoint_gather_kernel, dim3(grid_for(n_src)), dim3(TPB), 0, s, adj, S, dst, w,
                           n_src, src_cell, src_start, src_entry);
        PG_LAUNCHED("halo_scalar_adjoint_gather_kernel");
    }
    hipLaunchKernelGGL(halo_clear_kernel, dim3(grid_for(K)), dim3(TPB), 0, s, adj, K, dst);
    PG_LAUNCHED("halo_clear_kernel");
    return 0;
}

extern "C" int64_t pg_select_work_bytes(void) { return (int64_t)sizeof(SelectWork); }

extern "C" int pg_select_kth_surface(const double* height, const double* sediment, int64_t F, int64_t N, int64_t H,
                                     int64_t k, double* q_out, void* work, pg_stream stream) {
    clear_error();
    PG_REQUIRE(height && sediment && q_out && work, "pg_select_kth_surface: NULL argument");
    PG_REQUIRE(F > 0 && N > 0 && H >= 0, "pg_select_kth_surface: bad sizes");
    const int64_t M = F * N * N;
    PG_REQUIRE(k >= 0 && k < M, "pg_select_kth_surface: k=%lld outside [0, %lld)", (long long)k, (long long)M);
    hipStream_t s = (hipStream_t)stream;
    SelectWork* w = (SelectWork*)work;
    hipLaunchKernelGGL(select_init_kernel, dim3(1), dim3(256), 0, s, w, (long long)k);
    int blocks = grid_for(M);
    if (blocks > 2048) blocks = 2048;
    for (int shift = 56; shift >= 0; shift -= 8) {
        hipLaunchKernelGGL(select_hist_kernel, dim3(blocks), dim3(TPB), 0, s, height, sediment, F, N, H, shift, w);
        hipLaunchKernelGGL(select_pick_kernel, dim3(1), dim3(64), 0, s, w, shift, q_out);
    }
    PG_LAUNCHED("select_kernels");
    return 0;
}

extern "C" int pg_datum_shift(int64_t n, double* height, double* route, double* base, const double* q, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && height && q, "pg_datum_shift: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(datum_shift_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, height, route, base, q);
    PG_LAUNCHED("datum_shift_kernel");
    return 0;
}
