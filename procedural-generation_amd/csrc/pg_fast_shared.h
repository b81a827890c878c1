// What the fast mode's translation unit shares with the differentiable mode:
// the per-iteration device scalars, the workspace layout, the spawn
// pipeline, and the droplet steering step (one cell move on the frozen
// stencil record).  Internal to the library.
#pragma once

#include <hip/hip_runtime.h>

#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"

namespace pg {
namespace fastmode {

typedef unsigned long long ull;
typedef long long ll;

constexpr int TPB = 256;
constexpr int TRACE_TPB = 256;

inline int grid_for(int64_t n, int cap = 65535) {
    int b = blocks_for(n, TPB);
    return b < 1 ? 1 : (b > cap ? cap : b);
}

#define GRID_STRIDE(i, n) for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// device scalars of an iteration
struct Scalars {
    double volume0, q_weight, total;
    ll eligible, pending, n_total, cdf_exact;
    int emin;
    int pad;
    float inv_sat, mom_scale;  // the iteration's discharge scales (in cells of rain when asked, maps.py:911-921)
    ll queue[1];  // followed by `chunks` queue counters
};

// ---------------------------------------------------------------------------
// workspace layout
// ---------------------------------------------------------------------------
struct Layout {
    size_t cell4, geo, dyn, rte, chunk_fx, acc_fx, dis_fx, mom_fx, pend_fx, ero_fx, fin_fx, floor_key, rscale, r1, wq, cdf_fx, flag,
        rank, comp_idx, comp_cdf, bbuf, bval,
        draws, u_sorted, d_cell, d_fx, d_fy, d_load, d_sa, d_sb, scalars, temp, total;
    size_t temp_bytes;
};

int layout_of(int64_t F, int64_t NE, int64_t P, Layout* L);

template <class T>
T* at(void* work, size_t off) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(work) + off);
}

struct Ctx {
    Layout L;
    hipStream_t s;
    double hscale, tscale;
    Scalars* sc;
    float4* cell4;
    float2* dyn;
    int *flag, *rank;
    ll *pend_fx, *st;
    void* temp;
    int chunks;
    int64_t N, H, NE, n;
    ll P;
};

// argument checks and workspace views of an iteration
int prepare(const pg_fast_fields* fl, const pg_trace_params* prm, const pg_fast_params* fp, void* work,
            int64_t work_bytes, int64_t* stats, pg_stream stream, Ctx* c);
// pack + spawn of `fl` (the droplets land in the workspace's d_* arrays)
int fast_spawn(const pg_fast_fields* fl, const pg_trace_params* prm, const pg_fast_params* fp, void* work, Ctx& c);
// fold of acc_fx into `fl`'s height / sediment, EMA of the tracks, stockpile mirror
int finish_iteration(const pg_fast_fields* fl, const pg_fast_params* fp, void* work, Ctx& c);
// persistent grid of a trace kernel: one block set per CU at full occupancy
int persistent_grid(const void* kernel, int threads, int64_t requested, int* grid);

// ---------------------------------------------------------------------------
// the 2x2 stencil of the packed record {surface, discharge, momentum a, b}
// ---------------------------------------------------------------------------
struct Stencil {
    float4 c00, c10, c01, c11;  // records at (i0, j0), (i0+1, j0), (i0, j0+1), (i0+1, j0+1)
    float wi, wj;
    int i0, j0;
};

__device__ __forceinline__ float lerp2(float v00, float v10, float v01, float v11, float wi, float wj) {
    return v00 * (1.0f - wi) * (1.0f - wj) + v10 * wi * (1.0f - wj) + v01 * (1.0f - wi) * wj + v11 * wi * wj;
}

// stencil corner indices and weights of a position given as (cell, fraction)
// on face f: the extended-array bilinear of particle.py:194-252, clamped
__device__ __forceinline__ void stencil_at(int NE, int H, int ix, int iy, float fx, float fy, int& i0, int& j0, float& wi,
                                           float& wj) {
    i0 = ix + H - (fx < 0.5f ? 1 : 0);
    j0 = iy + H - (fy < 0.5f ? 1 : 0);
    wi = fx < 0.5f ? fx + 0.5f : fx - 0.5f;
    wj = fy < 0.5f ? fy + 0.5f : fy - 0.5f;
    if (i0 < 0) { i0 = 0; wi = 0.0f; }
    if (i0 > NE - 2) { i0 = NE - 2; wi = 1.0f; }
    if (j0 < 0) { j0 = 0; wj = 0.0f; }
    if (j0 > NE - 2) { j0 = NE - 2; wj = 1.0f; }
}

__device__ __forceinline__ Stencil load_stencil(const float4* cell4, int NE, int H, int f, int ix, int iy, float fx,
                                                float fy) {
    Stencil s;
    stencil_at(NE, H, ix, iy, fx, fy, s.i0, s.j0, s.wi, s.wj);
    const int64_t b = ((int64_t)f * NE + s.i0) * NE + s.j0;
    s.c00 = cell4[b];
    s.c01 = cell4[b + 1];
    s.c10 = cell4[b + NE];
    s.c11 = cell4[b + NE + 1];
    return s;
}

// the record of extended cell (ei, ej): always one of the stencil's corners
// when the position lies in that cell
__device__ __forceinline__ float4 corner_of(const Stencil& s, int ei, int ej) {
    const bool ri = ei != s.i0, rj = ej != s.j0;
    return ri ? (rj ? s.c11 : s.c10) : (rj ? s.c01 : s.c00);
}

// what a droplet needs of a stencil: the bilinear surface and its gradient,
// the clamped discharge and the momentum -- six values, evaluated once when
// the stencil is read, so a droplet carries 6 registers instead of 16
struct Samp {
    float h, gi, gj, q, ma, mb;
    float hr;  // the routing height (= h without a routing surface)
};

__device__ __forceinline__ Samp sample_of(const Stencil& S) {
    Samp p;
    const float h00 = S.c00.x, h10 = S.c10.x, h01 = S.c01.x, h11 = S.c11.x, wi = S.wi, wj = S.wj;
    p.h = lerp2(h00, h10, h01, h11, wi, wj);
    p.gi = (h10 - h00) * (1.0f - wj) + (h11 - h01) * wj;
    p.gj = (h01 - h00) * (1.0f - wi) + (h11 - h10) * wi;
    p.q = fmaxf(lerp2(S.c00.y, S.c10.y, S.c01.y, S.c11.y, wi, wj), 0.0f);
    p.ma = lerp2(S.c00.z, S.c10.z, S.c01.z, S.c11.z, wi, wj);
    p.mb = lerp2(S.c00.w, S.c10.w, S.c01.w, S.c11.w, wi, wj);
    p.hr = p.h;
    return p;
}

// the sample with the routing surface (particle.py:450-460, 552-560): where
// any of the stencil's four cells has a route flag, the bilinear route and
// its gradient; above the terrain they steer (gi, gj) and set the routing
// height hr the drop dhr is taken on, while h stays the terrain's.  rte:
// per extended cell {route, flag} (flag 0 none, 1 route above terrain, 2 a
// lake cell)
__device__ __forceinline__ Samp sample_route(const Stencil& S, const float2* rte, int64_t f, int NE) {
    Samp p = sample_of(S);
    const int64_t b = ((int64_t)f * NE + S.i0) * NE + S.j0;
    const float2 r00 = rte[b], r01 = rte[b + 1], r10 = rte[b + NE], r11 = rte[b + NE + 1];
    p.hr = p.h;
    if (r00.y + r10.y + r01.y + r11.y != 0.0f) {
        const float wi = S.wi, wj = S.wj;
        const float r0 = lerp2(r00.x, r10.x, r01.x, r11.x, wi, wj);
        if (r0 > p.h) {
            p.hr = r0;
            p.gi = (r10.x - r00.x) * (1.0f - wj) + (r11.x - r01.x) * wj;
            p.gj = (r01.x - r00.x) * (1.0f - wi) + (r11.x - r10.x) * wi;
        }
    }
    return p;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ ll wave_sum(ll v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---------------------------------------------------------------------------
// one move of a droplet (particle.py:437-531): gradient of the bilinear
// surface, inertia, gravity (and the stream push when k_mom != 0), the step
// normalised to one cell in the metric, face transfer.  `K` supplies fr,
// k_mom, mom_scale, dt, density, slope_gain, slope_sat, H, spherical; N is
// the droplet's grid's interior size.
// Returns -1 (moved), DEATH_STOP (no finite direction) or DEATH_EXIT (left a
// window).  P = the stencil's sample at the current position; mom_scale the
// iteration's momentum scale (k.mom_scale is not read).  F32_TRANSFER:
// a face crossing in single precision (transfer_velocity_f: the fast mode's
// trace, ~50 fewer registers) or in double (the differentiable mode, whose
// float32 / float64 gradients are compared along the same trajectories).
// ---------------------------------------------------------------------------
template <bool F32_TRANSFER, class K>
__device__ __forceinline__ int steer_move(const K& k, int N, const Samp& P, const float4 g0, const float4 g1,
                                          float vol, float vol_rel, float mom_scale, float& sa, float& sb, int f,
                                          int ix, int iy, float fx, float fy, int& nf, int& nix, int& niy, float& nfx,
                                          float& nfy) {
    const float gi = P.gi, gj = P.gj, q = P.q;
    const float di = g0.w * gi + g1.x * gj;
    const float dj = g1.x * gi + g1.y * gj;
    const float s2 = fmaxf(gi * di + gj * dj, 0.0f);
    const float nrm = k.slope_sat > 0.0f ? k.slope_gain * rsqrtf(s2 + k.slope_sat * k.slope_sat)
                                        : k.slope_gain * rsqrtf(1.0f + s2);
    const float na = -di * nrm, nb = -dj * nrm;
    sa *= k.fr;
    sb *= k.fr;
    if (k.k_mom != 0.0f) {
        const float ma = P.ma, mb = P.mb;
        const float mlen = sqrtf(fmaxf(g0.x * ma * ma + 2.0f * g0.y * ma * mb + g0.z * mb * mb, 0.0f));
        const float slen = sqrtf(fmaxf(g0.x * sa * sa + 2.0f * g0.y * sa * sb + g0.z * sb * sb, 0.0f));
        if (mlen > 1e-12f && slen > 1e-12f) {
            const float cosang = (g0.x * ma * sa + g0.y * (ma * sb + mb * sa) + g0.z * mb * sb) / (mlen * slen);
            const float kk = k.k_mom * cosang * mom_scale / (vol + mom_scale * q);
            sa += kk * ma;
            sb += kk * mb;
        }
    }
    const float acc = k.dt / (vol_rel * k.density);
    sa += acc * na;
    sb += acc * nb;
    const float n2 = fmaxf(g0.x * sa * sa + 2.0f * g0.y * sa * sb + g0.z * sb * sb, 0.0f);
    // (a non-finite direction -- only from non-finite inputs -- stops the droplet
    // rather than turning into a cell index)
    if (!(n2 >= 1e-24f) || !(n2 <= 3.0e38f)) return DEATH_STOP;
    const float inv = rsqrtf(n2);
    sa *= inv;
    sb *= inv;
    // one cell step, across a face edge if need be
    nfx = fx + sa;
    nfy = fy + sb;
    const float flx = floorf(nfx), fly = floorf(nfy);
    nix = ix + (int)flx;
    niy = iy + (int)fly;
    nfx = fminf(nfx - flx, 0.99999994f);
    nfy = fminf(nfy - fly, 0.99999994f);
    nf = f;
    if (k.spherical && !F32_TRANSFER) {
        if (nix < 0 || nix >= N || niy < 0 || niy >= N) {
            // across a face edge in double precision (device libm)
            const double Nd = (double)N;
            const Transfer t = transfer_velocity<LibMath>(f, ((double)nix + nfx) / Nd, ((double)niy + nfy) / Nd,
                                                          (double)sa, (double)sb);
            nf = t.face;
            double nx = t.u * Nd, ny = t.v * Nd;
            if (nx >= Nd) nx = Nd * 0.999999999;
            if (ny >= Nd) ny = Nd * 0.999999999;
            if (!(nx >= 0.0)) nx = 0.0;
            if (!(ny >= 0.0)) ny = 0.0;
            const double fnx = floor(nx), fny = floor(ny);
            nix = (int)fnx;
            niy = (int)fny;
            nfx = fminf((float)(nx - fnx), 0.99999994f);
            nfy = fminf((float)(ny - fny), 0.99999994f);
            sa = (float)t.a;
            sb = (float)t.b;
        }
    } else if (k.spherical) {
        if (nix < 0 || nix >= N || niy < 0 || niy >= N) {
            // across a face edge, in single precision (transfer_velocity_f)
            const float Nf = (float)N;
            const TransferF t = transfer_velocity_f(f, ((float)nix + nfx) / Nf, ((float)niy + nfy) / Nf, sa, sb);
            nf = t.face;
            const float top = __int_as_float(__float_as_int(Nf) - 1);  // the float below N
            float nx = fminf(t.u * Nf, top), ny = fminf(t.v * Nf, top);
            if (!(nx >= 0.0f)) nx = 0.0f;
            if (!(ny >= 0.0f)) ny = 0.0f;
            const float fnx = floorf(nx), fny = floorf(ny);
            nix = (int)fnx;
            niy = (int)fny;
            nfx = fminf(nx - fnx, 0.99999994f);
            nfy = fminf(ny - fny, 0.99999994f);
            sa = t.a;
            sb = t.b;
        }
    } else if (nix < -k.H || nix >= N + k.H || niy < -k.H || niy >= N + k.H) {
        return DEATH_EXIT;
    }
    return -1;
}

// the same move from a stencil (h0 = the bilinear surface at the position)
template <bool F32_TRANSFER, class K>
__device__ __forceinline__ int steer_move(const K& k, int N, const Stencil& S, const float4 g0, const float4 g1, float q,
                                          float vol, float vol_rel, float& sa, float& sb, int f, int ix, int iy,
                                          float fx, float fy, int& nf, int& nix, int& niy, float& nfx, float& nfy,
                                          float& h0) {
    Samp P = sample_of(S);
    P.q = q;
    h0 = P.h;
    return steer_move<F32_TRANSFER>(k, N, P, g0, g1, vol, vol_rel, k.mom_scale, sa, sb, f, ix, iy, fx, fy, nf, nix, niy,
                                    nfx, nfy);
}

}  // namespace fastmode
}  // namespace pg

// ---------------------------------------------------------------------------
// the erosion brush (pg_brush.hip): a chunk's per-cell net change spread over
// a footprint, halo shares folded onto their owners
// ---------------------------------------------------------------------------
namespace pg {
namespace brush {

using fastmode::ll;
using fastmode::ull;

constexpr int MAX_APRON = 3;  // radius - 1: radius at most 4
constexpr int MAX_TAPS = (2 * MAX_APRON + 1) * (2 * MAX_APRON + 1);
constexpr int N_TILES = 3;    // LDS tiles 64 x 4, 32 x 8, 16 x 16

// integer shares in units of 2^-16 of the footprint's offsets, row-major over
// (2 apron + 1)^2, the centre's 0 (it keeps the remainder)
struct Weights {
    int w[MAX_TAPS];
    int apron;
};

// the weights of a radius (host)
int weights_of(int64_t radius, Weights* W);
// the per-cell target flags from the mask and the owner table (NULL: identity)
int enqueue_valid(int64_t n, const uint8_t* mask, const int64_t* owner, uint8_t* bval, hipStream_t s);
// dst = the spread of src (every cell of dst written), halo shares folded onto
// their owners when an owner table is given
int enqueue_spread(int64_t F, int64_t NE, const Weights& W, int64_t tile, const ll* src, const uint8_t* bval,
                   const int64_t* owner, ll* dst, hipStream_t s);

}  // namespace brush
}  // namespace pg
