// The erosion brush: each height change of a chunk spread over a footprint
// of cells around the cell where it happened.
//
// No counterpart in the reference, whose droplets change one cell per step
// (bake/globe/erosion/particle.py:660-667): the brush is an extension of the
// fast mode, off (radius 0) unless asked for.  A footprint of radius r holds
// the offsets o with |o| < r, weighted max(0, r - |o|) (|o| the Euclidean
// length in cells), so radius 2 is the 3 x 3 block and radius 3 the 5 x 5
// one; radius 1 is the cell alone.
//
// The spread is linear in the amount, so spreading each move's change and
// summing equals summing per cell and spreading once per chunk -- up to the
// fixed-point rounding of each share.  The fast mode therefore keeps one
// atomic per move and runs the brush as a gather stencil over the chunk's
// per-cell net change, before the per-cell resolve (which then caps each
// footprint cell's deposit by its own room):
//
//   share(d, o) = sign(d) floor(|d| W_o / 2^16)       W_o = round(2^16 w_o / sum w)
//   out[t]      = sum_{o != 0, t valid} share(d[t - o], o)
//               + d[t] - sum_{o != 0, t + o valid} share(d[t], o)
//
// A target is valid when it lies in the face's extended array and is an
// active cell, or a halo cell whose owner (the interior cell of the
// neighbouring face it stands for, bake/globe/cubesphere.py:665-686) is
// active; every share a source cannot give stays with it.  A halo cell's
// shares are then added to its owner (grid.owner remapping on the device)
// and the halo cell cleared.  All int64: the sum over the grid is exactly the
// sum before, and the result does not depend on the tile or the schedule.
//
// The stencil is HBM-bound: a tile of cells and its apron (r - 1) of
// amounts and target flags staged in LDS, 8 B read + 1 B flag + 8 B written
// per extended cell and chunk; three tile shapes (64 x 4, 32 x 8, 16 x 16)
// for the autotuner.
#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"
#include "pg_fast_shared.h"

namespace pg {
namespace brush {

int weights_of(int64_t radius, Weights* W) {
    PG_REQUIRE(radius >= 0 && radius <= MAX_APRON + 1, "brush: radius %lld outside [0, %d]", (long long)radius,
               MAX_APRON + 1);
    const int A = radius > 1 ? (int)radius - 1 : 0, K = 2 * A + 1;
    double wt[MAX_TAPS], sum = 0.0;
    for (int dy = -A; dy <= A; dy++)
        for (int dx = -A; dx <= A; dx++) {
            const double d = sqrt((double)(dx * dx + dy * dy));
            const double w = (double)radius - d;
            wt[(dy + A) * K + dx + A] = w > 0.0 ? w : 0.0;
            sum += w > 0.0 ? w : 0.0;
        }
    for (int k = 0; k < MAX_TAPS; k++) W->w[k] = 0;
    for (int k = 0; k < K * K; k++)
        W->w[k] = (k == A * K + A || sum <= 0.0) ? 0 : (int)floor(65536.0 * wt[k] / sum + 0.5);
    W->apron = A;
    return 0;
}

namespace {

__device__ __forceinline__ ll share(ll d, int w) {
    const ll m = d < 0 ? -d : d;
    const ll s = (m >> 16) * (ll)w + (((m & 0xffff) * (ll)w) >> 16);
    return d < 0 ? -s : s;
}

// 1: an active cell that keeps its shares; 2: a halo cell whose active owner
// takes them; 0: no target
__global__ void brush_valid_kernel(int64_t n, const uint8_t* mask, const int64_t* owner, uint8_t* bval) {
    GRID_STRIDE(c, n) {
        const int64_t o = owner ? owner[c] : c;
        bval[c] = o == c ? (mask[c] == MASK_ACTIVE ? 1 : 0) : (mask[o] == MASK_ACTIVE ? 2 : 0);
    }
}

template <int TX, int TY, int A>
__global__ void __launch_bounds__(TX * TY) brush_spread_kernel(int NE, const ll* __restrict__ src,
                                                               const uint8_t* __restrict__ bval, ll* __restrict__ dst,
                                                               Weights W) {
    constexpr int LX = TX + 2 * A, LY = TY + 2 * A, K = 2 * A + 1;
    __shared__ ll sd[LY][LX];
    __shared__ uint8_t sv[LY][LX];
    const int64_t fbase = (int64_t)blockIdx.z * NE * NE;
    const int i0 = (int)blockIdx.y * TY - A, j0 = (int)blockIdx.x * TX - A;
    for (int t = threadIdx.x; t < LX * LY; t += TX * TY) {
        const int li = t / LX, lj = t - li * LX;
        const int i = i0 + li, j = j0 + lj;
        const bool in = (unsigned)i < (unsigned)NE && (unsigned)j < (unsigned)NE;
        const int64_t c = fbase + (int64_t)i * NE + j;
        sd[li][lj] = in ? src[c] : 0;
        sv[li][lj] = in ? bval[c] : 0;
    }
    __syncthreads();
    const int ti = threadIdx.x / TX, tj = threadIdx.x - ti * TX;
    const int i = (int)blockIdx.y * TY + ti, j = (int)blockIdx.x * TX + tj;
    if (i >= NE || j >= NE) return;
    const int li = ti + A, lj = tj + A;
    ll out = 0;
    if (sv[li][lj] != 0) {
#pragma unroll
        for (int dy = -A; dy <= A; dy++)
#pragma unroll
            for (int dx = -A; dx <= A; dx++) {
                if (dy == 0 && dx == 0) continue;
                const ll d = sd[li - dy][lj - dx];
                if (d != 0) out += share(d, W.w[(dy + A) * K + dx + A]);
            }
    }
    const ll d = sd[li][lj];
    if (d != 0) {
        ll given = 0;
#pragma unroll
        for (int dy = -A; dy <= A; dy++)
#pragma unroll
            for (int dx = -A; dx <= A; dx++) {
                if (dy == 0 && dx == 0) continue;
                if (sv[li + dy][lj + dx] != 0) given += share(d, W.w[(dy + A) * K + dx + A]);
            }
        out += d - given;
    }
    dst[fbase + (int64_t)i * NE + j] = out;
}

// a halo cell's shares to its owner (integer adds: order free)
__global__ void brush_fold_kernel(int64_t n, const uint8_t* bval, const int64_t* owner, ll* dst) {
    GRID_STRIDE(c, n) {
        if (bval[c] == 2) {
            const ll v = dst[c];
            if (v != 0) atomicAdd((ull*)&dst[owner[c]], (ull)v);
            dst[c] = 0;
        }
    }
}

template <int TX, int TY>
int launch_tile(int A, int64_t F, int64_t NE, const ll* src, const uint8_t* bval, ll* dst, const Weights& W,
                hipStream_t s) {
    const dim3 grid((unsigned)((NE + TX - 1) / TX), (unsigned)((NE + TY - 1) / TY), (unsigned)F);
    switch (A) {
        case 0: hipLaunchKernelGGL((brush_spread_kernel<TX, TY, 0>), grid, dim3(TX * TY), 0, s, (int)NE, src, bval, dst, W); break;
        case 1: hipLaunchKernelGGL((brush_spread_kernel<TX, TY, 1>), grid, dim3(TX * TY), 0, s, (int)NE, src, bval, dst, W); break;
        case 2: hipLaunchKernelGGL((brush_spread_kernel<TX, TY, 2>), grid, dim3(TX * TY), 0, s, (int)NE, src, bval, dst, W); break;
        default: hipLaunchKernelGGL((brush_spread_kernel<TX, TY, 3>), grid, dim3(TX * TY), 0, s, (int)NE, src, bval, dst, W); break;
    }
    PG_LAUNCHED("brush_spread_kernel");
    return 0;
}

}  // namespace

int enqueue_valid(int64_t n, const uint8_t* mask, const int64_t* owner, uint8_t* bval, hipStream_t s) {
    hipLaunchKernelGGL(brush_valid_kernel, dim3(blocks_for(n, 256) < 65536 ? blocks_for(n, 256) : 65536), dim3(256), 0,
                       s, n, mask, owner, bval);
    PG_LAUNCHED("brush_valid_kernel");
    return 0;
}

int enqueue_spread(int64_t F, int64_t NE, const Weights& W, int64_t tile, const ll* src, const uint8_t* bval,
                   const int64_t* owner, ll* dst, hipStream_t s) {
    PG_REQUIRE(tile >= 0 && tile < N_TILES, "brush: tile %lld outside [0, %d)", (long long)tile, N_TILES);
    PG_REQUIRE(W.apron >= 0 && W.apron <= MAX_APRON, "brush: apron %d", W.apron);
    const int A = W.apron;
    int rc = tile == 0 ? launch_tile<64, 4>(A, F, NE, src, bval, dst, W, s)
           : tile == 1 ? launch_tile<32, 8>(A, F, NE, src, bval, dst, W, s)
                       : launch_tile<16, 16>(A, F, NE, src, bval, dst, W, s);
    if (rc) return rc;
    if (owner) {
        const int64_t n = F * NE * NE;
        hipLaunchKernelGGL(brush_fold_kernel, dim3(blocks_for(n, 256) < 65536 ? blocks_for(n, 256) : 65536), dim3(256),
                           0, s, n, bval, owner, dst);
        PG_LAUNCHED("brush_fold_kernel");
    }
    return 0;
}

}  // namespace brush
}  // namespace pg

extern "C" int pg_brush_weights(int64_t radius, int32_t* w, int64_t n) {
    pg::clear_error();
    PG_REQUIRE(w != nullptr, "pg_brush_weights: NULL output");
    pg::brush::Weights W;
    if (pg::brush::weights_of(radius, &W)) return -1;
    const int K = 2 * W.apron + 1;
    PG_REQUIRE(n >= K * K, "pg_brush_weights: room for %lld weights, need %d", (long long)n, K * K);
    for (int k = 0; k < K * K; k++) w[k] = W.w[k];
    return W.apron;
}

extern "C" int pg_brush_spread(const int64_t* src, int64_t* dst, const uint8_t* mask, const int64_t* owner, int64_t F,
                               int64_t NE, int64_t radius, int64_t tile, uint8_t* valid, pg_stream stream) {
    pg::clear_error();
    PG_REQUIRE(src && dst && mask && valid, "pg_brush_spread: NULL argument");
    PG_REQUIRE(src != dst, "pg_brush_spread: src and dst must differ");
    PG_REQUIRE(F >= 1 && NE >= 1 && F * NE * NE < (1ll << 31), "pg_brush_spread: bad grid F=%lld NE=%lld",
               (long long)F, (long long)NE);
    pg::brush::Weights W;
    if (pg::brush::weights_of(radius, &W)) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (pg::brush::enqueue_valid(F * NE * NE, mask, owner, valid, s)) return -1;
    return pg::brush::enqueue_spread(F, NE, W, tile, (const pg::brush::ll*)src, valid, owner, (pg::brush::ll*)dst, s);
}
