// Side passes of an erosion iteration around the droplet loop: mass wasting
// (the thermal talus pass and hillslope creep), soillib's slope-limited
// erosion cap, and the isostatic response (the flux-form Laplace-Beltrami
// smoother of the load, its mean-free application).  All are stencil or
// elementwise passes over the float64 state, bandwidth-bound: stencils on a
// (row segment, row, face) grid, elementwise passes as grid-stride loops.
//
// Bit parity: every output is the reference's IEEE operation chain in its
// order (the library is built with -ffp-contract=off): float32 metric and
// hardness widen exactly, sqrt and division round correctly, and the only
// float32 rounding is where the reference stores into float32 (the talus
// slope, the smoother's Laplacian).  Reductions (the smoother's norms, the
// rebound's mean) sum in a fixed tree order of their own: deterministic, and
// equal to numpy's pairwise sums to rounding.
#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"

namespace pg {
namespace {

constexpr int TPB = 256;
constexpr int RED_BLOCKS = 512;  // partial sums of a reduction (fixed: the order is part of the result)

inline int grid_for(int64_t n) {
    int b = blocks_for(n, TPB);
    return b < 1 ? 1 : (b > 65535 ? 65535 : b);
}

#define GRID_STRIDE(i, n) for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// D8 neighbour offsets (particle.py:101-102): (1,0) (1,1) (0,1) (-1,1)
// (-1,0) (-1,-1) (0,-1) (1,-1); the opposite of k is (k + 4) % 8
__device__ __forceinline__ int d8i(int k) { return (k <= 1 || k == 7) ? 1 : ((k >= 3 && k <= 5) ? -1 : 0); }
__device__ __forceinline__ int d8j(int k) { return (k >= 1 && k <= 3) ? 1 : (k >= 5 ? -1 : 0); }

// particle.py:1107-1125: raw outflow from cell c to its D8 neighbour n = c +
// (di, dj); metric float32 widened, offsets as float64 (numba's typing)
__device__ __forceinline__ double thermal_out(const double* h, const double* s, const float* metric, const double* talus,
                                              int64_t c, int64_t n, int di, int dj, double rate) {
    const double d = (h[c] + s[c]) - (h[n] + s[n]);
    if (d <= 0.0) return 0.0;
    const double g0 = (double)metric[3 * c], g1 = (double)metric[3 * c + 1], g2 = (double)metric[3 * c + 2];
    const double ddi = (double)di, ddj = (double)dj;
    const double dist = sqrt(g0 * ddi * ddi + 2.0 * g1 * ddi * ddj + g2 * ddj * ddj);
    const double ex = d - talus[c] * dist;
    if (ex <= 0.0) return 0.0;
    return 0.5 * rate * ex;
}

// stencil kernels run on a (j, i, face) grid of 64 x 4 blocks: one thread
// per extended cell, consecutive threads along a row (coalesced), no index
// division
constexpr int SX = 64, SY = 4;

inline dim3 stencil_grid(int64_t F, int64_t NE) {
    return dim3((unsigned)((NE + SX - 1) / SX), (unsigned)((NE + SY - 1) / SY), (unsigned)F);
}

struct Cell {
    int64_t c;
    int i, j;
    bool ok;
};

__device__ __forceinline__ Cell stencil_cell(int64_t NE) {
    Cell r;
    r.j = blockIdx.x * SX + threadIdx.x;
    r.i = blockIdx.y * SY + threadIdx.y;
    r.ok = r.i < NE && r.j < NE;
    r.c = ((int64_t)blockIdx.z * NE + r.i) * NE + r.j;
    return r;
}

__device__ __forceinline__ bool inner(const Cell& x, int64_t NE) {
    return x.i >= 1 && x.i < NE - 1 && x.j >= 1 && x.j < NE - 1;
}

// maps.py:1062: soft + (hard - soft) * hardness in float32 (the hardness
// array is float32 and numpy keeps the product there), widened
__global__ void talus_kernel(int64_t n, const float* hardness, float soft, float span, double* talus) {
    GRID_STRIDE(c, n) { talus[c] = (double)(soft + span * hardness[c]); }
}

// particle.py:1128-1160 (out_total optional: nothing downstream reads it)
__global__ void thermal_a_kernel(int64_t F, int64_t NE, const double* h, const double* s, const float* metric,
                                 const double* talus, const uint8_t* mask, double rate, double tmax, double* out_total,
                                 double* scale) {
    const Cell x = stencil_cell(NE);
    if (x.ok) {
        const int64_t c = x.c;
        double tot = 0.0, sc = 0.0;
        if (inner(x, NE) && mask[c] == MASK_ACTIVE) {
            double S = 0.0, mx = 0.0;
            for (int k = 0; k < 8; k++) {
                const int di = d8i(k), dj = d8j(k);
                const int64_t nb = c + (int64_t)di * NE + dj;
                if (mask[nb] != MASK_ACTIVE) continue;
                const double o = thermal_out(h, s, metric, talus, c, nb, di, dj, rate);
                S += o;
                if (o > mx) mx = o;
            }
            if (S > 0.0) {
                sc = 1.0;
                if (S > 2.0 * mx) sc = 2.0 * mx / S;
                if (tmax > 0.0 && S * sc > tmax) sc = tmax / S;
                tot = S * sc;
            } else {
                sc = 0.0;
            }
        }
        if (out_total) out_total[c] = tot;
        scale[c] = sc;
    }
}

// particle.py:1163-1187: a submerged receiver takes inflow up to its room
// below -dep_floor
__global__ void thermal_r_kernel(int64_t F, int64_t NE, const double* h, const double* s, const float* metric,
                                 const double* talus, const uint8_t* mask, double rate, const double* scale,
                                 double* rscale, double dep_floor) {
    const Cell x = stencil_cell(NE);
    if (x.ok) {
        const int64_t c = x.c;
        double r = 1.0;
        const double surf = h[c] + s[c];
        if (inner(x, NE) && !(surf >= 0.0 || mask[c] != MASK_ACTIVE)) {
            double inflow = 0.0;
            for (int k = 0; k < 8; k++) {
                const int di = d8i(k), dj = d8j(k);
                const int64_t nb = c + (int64_t)di * NE + dj;
                const double sc = scale[nb];
                if (sc <= 0.0 || mask[nb] == MASK_OUTSIDE) continue;
                const int ko = (k + 4) & 7;
                inflow += sc * thermal_out(h, s, metric, talus, nb, c, d8i(ko), d8j(ko), rate);
            }
            double room = -dep_floor - surf;
            if (room < 0.0) room = 0.0;
            if (inflow > room) r = room / inflow;
        }
        rscale[c] = r;
    }
}

// particle.py:1190-1227: losses (scaled by each receiver's rscale; sediment
// first) and gains (as sediment)
__global__ void thermal_b_kernel(int64_t F, int64_t NE, const double* h, const double* s, const float* metric,
                                 const double* talus, const uint8_t* mask, double rate, const double* scale,
                                 const double* rscale, double* new_h, double* new_s) {
    const Cell x = stencil_cell(NE);
    if (x.ok) {
        const int64_t c = x.c;
        double hh = h[c], ss = s[c];
        if (inner(x, NE) && mask[c] == MASK_ACTIVE) {
            double inflow = 0.0, out = 0.0;
            const double sc_self = scale[c];
            for (int k = 0; k < 8; k++) {
                const int di = d8i(k), dj = d8j(k);
                const int64_t nb = c + (int64_t)di * NE + dj;
                const uint8_t mn = mask[nb];
                if (mn == MASK_OUTSIDE) continue;
                if (sc_self > 0.0 && mn == MASK_ACTIVE)
                    out += sc_self * rscale[nb] * thermal_out(h, s, metric, talus, c, nb, di, dj, rate);
                const double sc = scale[nb];
                if (sc <= 0.0) continue;
                const int ko = (k + 4) & 7;
                inflow += sc * thermal_out(h, s, metric, talus, nb, c, d8i(ko), d8j(ko), rate);
            }
            inflow *= rscale[c];
            if (out > 0.0) {
                if (ss >= out) {
                    ss -= out;
                } else {
                    hh -= out - ss;
                    ss = 0.0;
                }
            }
            ss = ss + inflow;
        }
        new_h[c] = hh;
        new_s[c] = ss;
    }
}

// maps.py:1067: creep freezes the submerged cells
__global__ void creep_mask_kernel(int64_t n, const double* h, const double* s, const double* base, const uint8_t* mask,
                                  uint8_t* out) {
    GRID_STRIDE(c, n) { out[c] = (h[c] + s[c] < base[c]) ? (uint8_t)MASK_FROZEN : mask[c]; }
}

__global__ void copy2_kernel(int64_t n, const double* a, const double* b, double* a_out, double* b_out) {
    GRID_STRIDE(c, n) {
        a_out[c] = a[c];
        b_out[c] = b[c];
    }
}

// particle.py:1045-1076
__global__ void slope_cap_kernel(int64_t F, int64_t NE, const double* h, const double* s, const uint8_t* mask,
                                 double k, double exit_slope, double* out) {
    const Cell x = stencil_cell(NE);
    if (x.ok) {
        const int64_t c = x.c;
        if (mask[c] != MASK_ACTIVE) {
            out[c] = 0.0;
            return;
        }
        const double hc = h[c] + s[c];
        const double gxn = x.i == 0 ? exit_slope : hc - (h[c - NE] + s[c - NE]);
        const double gxp = x.i == NE - 1 ? exit_slope : hc - (h[c + NE] + s[c + NE]);
        const double gyn = x.j == 0 ? exit_slope : hc - (h[c - 1] + s[c - 1]);
        const double gyp = x.j == NE - 1 ? exit_slope : hc - (h[c + 1] + s[c + 1]);
        double gx = gxn > gxp ? gxn : gxp;
        double gy = gyn > gyp ? gyn : gyp;
        if (gx < 0.0) gx = 0.0;
        if (gy < 0.0) gy = 0.0;
        out[c] = k * sqrt(gx * gx + gy * gy);
    }
}

// field.py:372-404 at an inner cell: the flux-form operator before the
// division by the cell size squared
__device__ __forceinline__ double laplacian_at(const double* d, const float* ginv, const float* sq, int64_t c,
                                               int64_t NE) {
    const int64_t E = NE;
    auto D = [&](int64_t o) { return d[c + o]; };
    auto G = [&](int64_t o, int k) { return (double)ginv[3 * (c + o) + k]; };
    auto Q = [&](int64_t o) { return (double)sq[c + o]; };
    const double fj_c = 0.5 * (D(1) - D(-1));
    const double fj_p = 0.5 * (D(E + 1) - D(E - 1));
    const double fj_m = 0.5 * (D(-E + 1) - D(-E - 1));
    const double fi_c = 0.5 * (D(E) - D(-E));
    const double fi_p = 0.5 * (D(E + 1) - D(-E + 1));
    const double fi_m = 0.5 * (D(E - 1) - D(-E - 1));
    double gii = 0.5 * (G(0, 0) + G(E, 0));
    double gij = 0.5 * (G(0, 1) + G(E, 1));
    double q = 0.5 * (Q(0) + Q(E));
    const double Fi_p = q * (gii * (D(E) - D(0)) + gij * 0.5 * (fj_c + fj_p));
    gii = 0.5 * (G(0, 0) + G(-E, 0));
    gij = 0.5 * (G(0, 1) + G(-E, 1));
    q = 0.5 * (Q(0) + Q(-E));
    const double Fi_m = q * (gii * (D(0) - D(-E)) + gij * 0.5 * (fj_c + fj_m));
    double gjj = 0.5 * (G(0, 2) + G(1, 2));
    gij = 0.5 * (G(0, 1) + G(1, 1));
    q = 0.5 * (Q(0) + Q(1));
    const double Fj_p = q * (gjj * (D(1) - D(0)) + gij * 0.5 * (fi_c + fi_p));
    gjj = 0.5 * (G(0, 2) + G(-1, 2));
    gij = 0.5 * (G(0, 1) + G(-1, 1));
    q = 0.5 * (Q(0) + Q(-1));
    const double Fj_m = q * (gjj * (D(0) - D(-1)) + gij * 0.5 * (fi_c + fi_m));
    return (Fi_p - Fi_m + Fj_p - Fj_m) / Q(0);
}

// FaceField.laplacian (field.py:179-188) of d, in float32, then
// out = d_add + kappa * lap * c2 (the smoother's step, maps.py:1401) or,
// with d_add NULL, out = lap * c2 (the power iteration's field, maps.py:1356)
__global__ void laplacian_kernel(int64_t F, int64_t NE, const double* d, const float* ginv, const float* sq, double cs2,
                                 int mode, double kappa, double c2, double* out) {
    const Cell x = stencil_cell(NE);
    if (x.ok) {
        const int64_t c = x.c;
        double lap = 0.0;
        if (inner(x, NE)) lap = laplacian_at(d, ginv, sq, c, NE);
        const double L = (double)(float)(lap / cs2);
        if (mode == 0) out[c] = lap;                  // raw operator
        else if (mode == 1) out[c] = L * c2;          // Lv
        else out[c] = d[c] + kappa * L * c2;          // one smoothing step
    }
}

// sum of v (what 0), v^2 (1) or |v| (2), the count, max and min of v over
// interior cells (mask NULL: all of them; else mask == ACTIVE only)
__device__ __forceinline__ double red_term(double v, int what) {
    return what == 0 ? v : (what == 1 ? v * v : fabs(v));
}

__global__ void interior_partial_kernel(const double* d, const uint8_t* mask, int64_t F, int64_t N, int64_t H, int what,
                                        double* partial) {
    const int64_t NE = N + 2 * H, M = F * N * N;
    double acc = 0.0, cnt = 0.0, mx = -INFINITY, mn = INFINITY;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < M; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = t / (N * N), r = t - f * N * N, i = r / N, j = r - i * N;
        const int64_t c = (f * NE + i + H) * NE + j + H;
        if (mask && mask[c] != MASK_ACTIVE) continue;
        const double v = d[c];
        acc += red_term(v, what);
        cnt += 1.0;
        mx = fmax(mx, v);
        mn = fmin(mn, v);
    }
    __shared__ double sa[TPB], sc[TPB], sx[TPB], sn[TPB];
    const int tid = threadIdx.x;
    sa[tid] = acc;
    sc[tid] = cnt;
    sx[tid] = mx;
    sn[tid] = mn;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (tid < o) {
            sa[tid] += sa[tid + o];
            sc[tid] += sc[tid + o];
            sx[tid] = fmax(sx[tid], sx[tid + o]);
            sn[tid] = fmin(sn[tid], sn[tid + o]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        double* p = partial + 4 * blockIdx.x;
        p[0] = sa[0];
        p[1] = sc[0];
        p[2] = sx[0];
        p[3] = sn[0];
    }
}

// one block of TPB threads: out = {sum, count, max, min}
__global__ void interior_final_kernel(const double* partial, int nb, double* out) {
    __shared__ double sa[TPB], sc[TPB], sx[TPB], sn[TPB];
    const int tid = threadIdx.x;
    double a = 0.0, c = 0.0, mx = -INFINITY, mn = INFINITY;
    for (int b = tid; b < nb; b += TPB) {
        a += partial[4 * b];
        c += partial[4 * b + 1];
        mx = fmax(mx, partial[4 * b + 2]);
        mn = fmin(mn, partial[4 * b + 3]);
    }
    sa[tid] = a;
    sc[tid] = c;
    sx[tid] = mx;
    sn[tid] = mn;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (tid < o) {
            sa[tid] += sa[tid + o];
            sc[tid] += sc[tid + o];
            sx[tid] = fmax(sx[tid], sx[tid + o]);
            sn[tid] = fmin(sn[tid], sn[tid + o]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = sa[0];
        out[1] = sc[0];
        out[2] = sx[0];
        out[3] = sn[0];
    }
}

// maps.py:1405-1406: height += data - mean on active cells, the mean of
// data over the active interior read from sums = {sum, count, ...}
__global__ void rebound_kernel(int64_t n, double* height, const double* data, const uint8_t* mask, const double* sums) {
    const double mean = sums[1] > 0.0 ? sums[0] / sums[1] : 0.0;
    GRID_STRIDE(c, n) {
        if (mask[c] == MASK_ACTIVE) height[c] += data[c] - mean;
    }
}

// maps.py:1572-1586: the surface before the surface processes, and the
// load they moved: iso_acc += (height + sediment) - s0
__global__ void surface_kernel(int64_t n, const double* h, const double* s, double* out) {
    GRID_STRIDE(c, n) { out[c] = h[c] + s[c]; }
}

__global__ void iso_accumulate_kernel(int64_t n, double* iso, const double* h, const double* s, const double* s0) {
    GRID_STRIDE(c, n) { iso[c] += (h[c] + s[c]) - s0[c]; }
}

// out = a * x (maps.py:1390, the load -k * d); a == 0 fills zeros
__global__ void scale_kernel(int64_t n, const double* x, double a, double* out) {
    GRID_STRIDE(c, n) { out[c] = a == 0.0 ? 0.0 : a * x[c]; }
}

// out = x / v (maps.py:1363)
__global__ void divide_kernel(int64_t n, const double* x, double v, double* out) {
    GRID_STRIDE(c, n) { out[c] = x[c] / v; }
}

}  // namespace
}  // namespace pg

using namespace pg;

#define PG_STENCIL_ARGS(name)                                                                                         \
    PG_REQUIRE(F > 0 && NE >= 3, name ": bad sizes F=%lld NE=%lld", (long long)F, (long long)NE);                      \
    PG_REQUIRE(height && sediment && metric && talus && mask, name ": NULL array")

extern "C" int pg_talus(int64_t n, const float* hardness, double soft, double hard, double* talus, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && hardness && talus, "pg_talus: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(talus_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, hardness, (float)soft,
                       (float)(hard - soft), talus);
    PG_LAUNCHED("talus_kernel");
    return 0;
}

extern "C" int pg_thermal_pass_a(int64_t F, int64_t NE, const double* height, const double* sediment, const float* metric,
                                 const double* talus, const uint8_t* mask, double rate, double thermal_max,
                                 double* out_total, double* scale, pg_stream stream) {
    clear_error();
    PG_STENCIL_ARGS("pg_thermal_pass_a");
    PG_REQUIRE(scale, "pg_thermal_pass_a: NULL scale");
    hipLaunchKernelGGL(thermal_a_kernel, stencil_grid(F, NE), dim3(SX, SY), 0, (hipStream_t)stream, F, NE, height,
                       sediment, metric, talus, mask, rate, thermal_max, out_total, scale);
    PG_LAUNCHED("thermal_a_kernel");
    return 0;
}

extern "C" int pg_thermal_pass_r(int64_t F, int64_t NE, const double* height, const double* sediment, const float* metric,
                                 const double* talus, const uint8_t* mask, double rate, const double* scale,
                                 double* rscale, double dep_floor, pg_stream stream) {
    clear_error();
    PG_STENCIL_ARGS("pg_thermal_pass_r");
    PG_REQUIRE(scale && rscale, "pg_thermal_pass_r: NULL scale");
    hipLaunchKernelGGL(thermal_r_kernel, stencil_grid(F, NE), dim3(SX, SY), 0, (hipStream_t)stream, F, NE, height,
                       sediment, metric, talus, mask, rate, scale, rscale, dep_floor);
    PG_LAUNCHED("thermal_r_kernel");
    return 0;
}

extern "C" int pg_thermal_pass_b(int64_t F, int64_t NE, const double* height, const double* sediment, const float* metric,
                                 const double* talus, const uint8_t* mask, double rate, const double* scale,
                                 const double* rscale, double* new_height, double* new_sediment, pg_stream stream) {
    clear_error();
    PG_STENCIL_ARGS("pg_thermal_pass_b");
    PG_REQUIRE(scale && rscale && new_height && new_sediment, "pg_thermal_pass_b: NULL array");
    PG_REQUIRE(new_height != height && new_sediment != sediment, "pg_thermal_pass_b: outputs must not alias the inputs");
    hipLaunchKernelGGL(thermal_b_kernel, stencil_grid(F, NE), dim3(SX, SY), 0, (hipStream_t)stream, F, NE, height,
                       sediment, metric, talus, mask, rate, scale, rscale, new_height, new_sediment);
    PG_LAUNCHED("thermal_b_kernel");
    return 0;
}

extern "C" int pg_mass_wasting_pass(int64_t F, int64_t NE, double* height, double* sediment, const float* metric,
                                    const double* talus, const uint8_t* mask, double rate, double thermal_max,
                                    double dep_floor, double* work, pg_stream stream) {
    clear_error();
    PG_STENCIL_ARGS("pg_mass_wasting_pass");
    PG_REQUIRE(work, "pg_mass_wasting_pass: NULL work");
    const int64_t n = F * NE * NE;
    double *scale = work, *rscale = work + n, *nh = work + 2 * n, *ns = work + 3 * n;
    hipStream_t s = (hipStream_t)stream;
    const int g = grid_for(n);
    const dim3 sg = stencil_grid(F, NE), sb(SX, SY);
    hipLaunchKernelGGL(thermal_a_kernel, sg, sb, 0, s, F, NE, height, sediment, metric, talus, mask, rate,
                       thermal_max, (double*)nullptr, scale);
    hipLaunchKernelGGL(thermal_r_kernel, sg, sb, 0, s, F, NE, height, sediment, metric, talus, mask, rate,
                       scale, rscale, dep_floor);
    hipLaunchKernelGGL(thermal_b_kernel, sg, sb, 0, s, F, NE, height, sediment, metric, talus, mask, rate,
                       scale, rscale, nh, ns);
    hipLaunchKernelGGL(copy2_kernel, dim3(g), dim3(TPB), 0, s, n, nh, ns, height, sediment);
    PG_LAUNCHED("thermal kernels");
    return 0;
}

extern "C" int pg_creep_mask(int64_t n, const double* height, const double* sediment, const double* base,
                             const uint8_t* mask, uint8_t* out, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && height && sediment && base && mask && out, "pg_creep_mask: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(creep_mask_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, height, sediment,
                       base, mask, out);
    PG_LAUNCHED("creep_mask_kernel");
    return 0;
}

extern "C" int pg_slope_erode_cap(int64_t F, int64_t NE, const double* height, const double* sediment,
                                  const uint8_t* mask, double coef, double exit_slope, double* out, pg_stream stream) {
    clear_error();
    PG_REQUIRE(F > 0 && NE >= 1 && height && sediment && mask && out, "pg_slope_erode_cap: bad arguments");
    const double k = coef * sqrt(2.0);
    hipLaunchKernelGGL(slope_cap_kernel, stencil_grid(F, NE), dim3(SX, SY), 0, (hipStream_t)stream, F, NE, height,
                       sediment, mask, k, exit_slope, out);
    PG_LAUNCHED("slope_cap_kernel");
    return 0;
}

extern "C" int pg_laplacian(int64_t F, int64_t NE, const double* data, const float* metric_inv, const float* sqrt_det,
                            double cell_size_sq, int32_t mode, double kappa, double c2, double* out, pg_stream stream) {
    clear_error();
    PG_REQUIRE(F > 0 && NE >= 3 && data && metric_inv && sqrt_det && out, "pg_laplacian: bad arguments");
    PG_REQUIRE(mode >= 0 && mode <= 2, "pg_laplacian: mode must be 0, 1 or 2");
    PG_REQUIRE(out != data, "pg_laplacian: out must not alias data");
    hipLaunchKernelGGL(laplacian_kernel, stencil_grid(F, NE), dim3(SX, SY), 0, (hipStream_t)stream, F, NE, data,
                       metric_inv, sqrt_det, cell_size_sq, (int)mode, kappa, c2, out);
    PG_LAUNCHED("laplacian_kernel");
    return 0;
}

extern "C" int64_t pg_interior_sums_work_bytes(void) { return (int64_t)(4 * RED_BLOCKS * sizeof(double)); }

extern "C" int pg_interior_sums(const double* data, const uint8_t* mask, int64_t F, int64_t N, int64_t H, int32_t what,
                                double* work, double* out4, pg_stream stream) {
    clear_error();
    PG_REQUIRE(data && work && out4 && F > 0 && N > 0 && H >= 0, "pg_interior_sums: bad arguments");
    PG_REQUIRE(what >= 0 && what <= 2, "pg_interior_sums: what must be 0, 1 or 2");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(interior_partial_kernel, dim3(RED_BLOCKS), dim3(TPB), 0, s, data, mask, F, N, H, (int)what, work);
    hipLaunchKernelGGL(interior_final_kernel, dim3(1), dim3(TPB), 0, s, work, RED_BLOCKS, out4);
    PG_LAUNCHED("interior_sum_kernels");
    return 0;
}

extern "C" int pg_rebound(int64_t n, double* height, const double* data, const uint8_t* mask, const double* sums,
                          pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && height && data && mask && sums, "pg_rebound: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(rebound_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, height, data, mask, sums);
    PG_LAUNCHED("rebound_kernel");
    return 0;
}

extern "C" int pg_surface(int64_t n, const double* height, const double* sediment, double* out, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && height && sediment && out, "pg_surface: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(surface_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, height, sediment, out);
    PG_LAUNCHED("surface_kernel");
    return 0;
}

extern "C" int pg_iso_accumulate(int64_t n, double* iso_acc, const double* height, const double* sediment,
                                 const double* s0, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && iso_acc && height && sediment && s0, "pg_iso_accumulate: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(iso_accumulate_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, iso_acc, height,
                       sediment, s0);
    PG_LAUNCHED("iso_accumulate_kernel");
    return 0;
}

extern "C" int pg_scale(int64_t n, const double* x, double a, double* out, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && out && (x || a == 0.0), "pg_scale: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, x ? x : out, a, out);
    PG_LAUNCHED("scale_kernel");
    return 0;
}

extern "C" int pg_divide(int64_t n, const double* x, double v, double* out, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && x && out, "pg_divide: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(divide_kernel, dim3(grid_for(n)), dim3(TPB), 0, (hipStream_t)stream, n, x, v, out);
    PG_LAUNCHED("divide_kernel");
    return 0;
}
