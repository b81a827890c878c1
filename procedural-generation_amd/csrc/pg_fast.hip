// Fast mode: one iteration's particle pass entirely on the device.
//
//   pack      -- per cell: the 16-byte stencil record {surface, discharge,
//                momentum a, b} and {sediment, base}; spawn eligibility.
//   spawn     -- the params.rng stream (numpy PCG64, jump-ahead per thread),
//                radix-sorted uniforms, inverse CDF over an exact
//                fixed-point prefix sum of the rain weights, offsets;
//                pending stockpiles re-injected after the rain in cell order.
//   K chunks  -- persistent trace kernel: one droplet per lane, float32
//                droplet state, lanes refilled from a device work queue as
//                droplets die; every step's effect accumulated with int64
//                fixed-point atomics (deterministic, order independent);
//                then a per-cell resolve applies the chunk against the live
//                terrain (deposit caps, surplus to the cell's stockpile with
//                the offshore write-off) and refreshes the live surface.
//   finish    -- fold, discharge/momentum EMA, stockpile mirror.
//
// Mass is exact in fixed point: every droplet carries its load as an int64
// and deposits exactly what it took, so per iteration
//   sum(applied) + new pending + lost offshore + exits + vanished = released.
//
// Semantic differences from the reference (bake/globe/erosion/particle.py),
// which applies each chunk's change list serially with caps that depend on
// apply order: here all droplets of a chunk act on the same frozen terrain
// and the caps are order independent -- a census pass sums each cell's
// erosion requests and the lowest surface stepped to from it, a per-cell
// scale limits the cell to its iteration budget and to that floor lowered by
// what the floor cell itself loses in the chunk (the reference's next-cell
// floor is live), the applying pass scales each entry (deficit carried to
// later deposits), clamps each deposit at the previous path cell + dep_floor
// -- with the droplet's own change there -- and the waterline (surplus
// carried on), and spreads a land death's load over its last SPREAD cells
// under the same ceilings; the per-cell resolve applies the deposit budget
// and stockpiles the excess.  Plain profile only.  Batched windows (pg_batch_*) run the same kernels over an
// arena of single-face windows, one persistent launch per chunk and pass.
#include <hipcub/hipcub.hpp>

#include <vector>

#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"

#pragma clang fp contract(fast)

#include "pg_fast_shared.h"

namespace pg {
namespace fastmode {
namespace {

// ---------------------------------------------------------------------------
// PCG64 (numpy's: 128-bit LCG, XSL-RR output, doubles from the top 53 bits)
// ---------------------------------------------------------------------------
typedef unsigned __int128 u128;
__device__ const u128 PCG_MULT = ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;

__device__ __forceinline__ u128 pcg_advance(u128 s, u128 inc, ull n) {
    u128 am = 1, ap = 0, cm = PCG_MULT, cp = inc;
    while (n) {
        if (n & 1) {
            am = am * cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm = cm * cm;
        n >>= 1;
    }
    return am * s + ap;
}

__device__ __forceinline__ double pcg_double(u128 s) {
    const ull hi = (ull)(s >> 64), lo = (ull)s;
    const ull x = hi ^ lo;
    const unsigned r = (unsigned)(s >> 122);
    const ull v = (x >> r) | (x << ((64 - r) & 63));
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

// out[k] = draw (start + k): each thread advances once, then steps L times
__global__ void pcg_fill_kernel(ull slo, ull shi, ull ilo, ull ihi, const ull* rng_dev, ll start, ll n, double* out) {
    constexpr int L = 16;
    if (rng_dev) {
        slo = rng_dev[0];
        shi = rng_dev[1];
        ilo = rng_dev[2];
        ihi = rng_dev[3];
    }
    const u128 s0 = ((u128)shi << 64) | slo, inc = ((u128)ihi << 64) | ilo;
    for (ll t = (ll)blockIdx.x * blockDim.x + threadIdx.x; t * L < n; t += (ll)gridDim.x * blockDim.x) {
        const ll k0 = t * L;
        u128 s = pcg_advance(s0, inc, (ull)(start + k0));
        for (int j = 0; j < L && k0 + j < n; j++) {
            s = s * PCG_MULT + inc;
            out[k0 + j] = pcg_double(s);
        }
    }
}

// ---------------------------------------------------------------------------
// setup: static geometry record, fixed-point stockpile
// ---------------------------------------------------------------------------
__global__ void setup_kernel(int64_t n, const float* metric, const float* metric_inv, const float* evap,
                             const float* hardness, float4* geo, const double* pending, ll* pend_fx, double hscale) {
    GRID_STRIDE(c, n) {
        geo[2 * c] = make_float4(metric[3 * c], metric[3 * c + 1], metric[3 * c + 2], metric_inv[3 * c]);
        geo[2 * c + 1] = make_float4(metric_inv[3 * c + 1], metric_inv[3 * c + 2], evap[c], hardness[c]);
        pend_fx[c] = llrint(pending[c] * hscale);
    }
}

// ---------------------------------------------------------------------------
// pack + spawn eligibility (maps.py:793-806, particle.py:164-191)
// ---------------------------------------------------------------------------
__global__ void pack_kernel(int64_t n, int64_t N, int64_t H, const double* height, const double* sediment,
                            const double* discharge, const double* momentum, const double* base, const float* precip,
                            const uint8_t* mask, float4* cell4, float2* dyn, int* eligible, Scalars* sc,
                            const double* route, const uint8_t* lake_flag, float2* rte) {
    const int64_t NE = N + 2 * H, NE2 = NE * NE;
    int my_emin = 1 << 30;
    GRID_STRIDE(c, n) {
        const double s = sediment[c], surf = height[c] + s;
        cell4[c] = make_float4((float)surf, (float)discharge[c], (float)momentum[2 * c], (float)momentum[2 * c + 1]);
        dyn[c] = make_float2((float)s, (float)base[c]);
        if (route) {
            // the routing channels of the packed record (particle.py:174-190): the route as float32 and its
            // flag -- 2 on an overflowing lake cell (3 with room for its rivers' load), else 1 where the
            // route differs from the float32 surface
            const float r = (float)route[c];
            const int lk = lake_flag ? (int)lake_flag[c] : 0;
            rte[c] = make_float2(r, lk > 0 ? (lk > 1 ? 3.0f : 2.0f) : (r != (float)surf ? 1.0f : 0.0f));
        }
        const int64_t r = c % NE2, i = r / NE, j = r - i * NE;
        const bool inter = i >= H && i < H + N && j >= H && j < H + N;
        const float w = precip[c];
        const bool ok = inter && mask[c] == MASK_ACTIVE && surf >= base[c] && w > 0.0f;
        eligible[c] = ok ? 1 : 0;
        if (ok) {
            int e;
            frexpf(w, &e);
            my_emin = e < my_emin ? e : my_emin;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int v = __shfl_xor(my_emin, o);
        my_emin = v < my_emin ? v : my_emin;
    }
    if ((threadIdx.x & 63) == 0 && my_emin < (1 << 30)) atomicMin(&sc->emin, my_emin);
}

__global__ void reset_scalars_kernel(Scalars* sc, ll* stats, int nstats, int chunks) {
    if (threadIdx.x == 0) {
        sc->volume0 = 0.0;
        sc->q_weight = 0.0;
        sc->total = 0.0;
        sc->eligible = sc->pending = sc->n_total = 0;
        sc->cdf_exact = 1;
        sc->emin = 1 << 30;
    }
    for (int k = threadIdx.x; k < nstats; k += blockDim.x) stats[k] = 0;
    for (int k = threadIdx.x; k < 2 * chunks; k += blockDim.x) sc->queue[k] = 0;
}

// rain weight in units of the smallest weight's float32 quantum: exact integers
__global__ void quantize_kernel(int64_t n, const float* precip, const int* eligible, const Scalars* sc, ll* wq) {
    const int emin = sc->emin;
    GRID_STRIDE(c, n) {
        ll v = 0;
        if (eligible[c]) v = (ll)ldexp((double)precip[c], 24 - emin);
        wq[c] = v;
    }
}

// compact the eligible cells in flat order with their cumulative weight
__global__ void compact_kernel(int64_t n, const int* eligible, const int* rank, const ll* cdf_fx, Scalars* sc,
                               int* comp_idx, double* comp_cdf) {
    const double q = ldexp(1.0, sc->emin - 24);
    GRID_STRIDE(c, n) {
        if (eligible[c]) {
            const int r = rank[c] - 1;
            comp_idx[r] = (int)c;
            comp_cdf[r] = (double)cdf_fx[c] * q;
        }
    }
}

// numpy's pairwise float64 sum (np.add.reduce on a contiguous float64
// array): blocks of < 8 summed in order, blocks up to 128 with 8 interleaved
// accumulators, longer ones split at a multiple of 8 below the half.
__device__ double pairwise_leaf(const float* w, const int* idx, int64_t n) {
    if (n < 8) {
        double r = 0.0;  // -0.0 only differs for an all -0.0 input, impossible for weights > 0
        for (int64_t i = 0; i < n; i++) r += (double)w[idx[i]];
        return r;
    }
    double r[8];
    for (int k = 0; k < 8; k++) r[k] = (double)w[idx[k]];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
        for (int k = 0; k < 8; k++) r[k] += (double)w[idx[i + k]];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += (double)w[idx[i]];
    return res;
}

__device__ double pairwise_sum(const float* w, const int* idx, int64_t n) {
    // iterative post-order walk of the split tree
    struct Frame {
        int64_t off, n;
        int state;
        double left;
    };
    Frame st[64];
    int sp = 0;
    st[0] = {0, n, 0, 0.0};
    double ret = 0.0;
    while (sp >= 0) {
        Frame& fr = st[sp];
        if (fr.n <= 128) {
            ret = pairwise_leaf(w, idx + fr.off, fr.n);
            sp--;
            continue;
        }
        int64_t n2 = fr.n / 2;
        n2 -= n2 % 8;
        if (fr.state == 0) {
            fr.state = 1;
            st[sp + 1] = {fr.off, n2, 0, 0.0};
            sp++;
        } else if (fr.state == 1) {
            fr.left = ret;
            fr.state = 2;
            st[sp + 1] = {fr.off + n2, fr.n - n2, 0, 0.0};
            sp++;
        } else {
            ret = fr.left + ret;
            sp--;
        }
    }
    return ret;
}

// spawn totals; if the fixed-point sum is not exact in float64, the serial
// path reproduces numpy's cumsum (sequential) and sum (pairwise) instead
__global__ void spawn_totals_kernel(int64_t n, const int* rank, const ll* cdf_fx, const int* comp_idx, double* comp_cdf,
                                    const float* precip, ll n_particles, Scalars* sc, ll* stats, float inv_sat,
                                    float mom_scale, double disc_cells, double mom_cells, double inflow) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const ll M = rank[n - 1];
    sc->eligible = M;
    stats[PG_FS_ELIGIBLE] = M;
    sc->inv_sat = inv_sat;
    sc->mom_scale = mom_scale;
    if (M == 0) {
        sc->volume0 = 0.0;
        return;
    }
    const ll tot = cdf_fx[n - 1];
    const bool exact = tot < (1ll << 53);
    sc->cdf_exact = exact;
    double total;
    if (exact) {
        total = (double)tot * ldexp(1.0, sc->emin - 24);
    } else {
        double acc = 0.0;
        for (ll r = 0; r < M; r++) {
            acc += (double)precip[comp_idx[r]];
            comp_cdf[r] = acc;
        }
        total = pairwise_sum(precip, comp_idx, M);
        stats[PG_FS_CDF_SERIAL] = 1;
    }
    sc->total = total;
    sc->volume0 = total / (double)n_particles;
    stats[PG_FS_VOLUME0_BITS] = __double_as_longlong(sc->volume0);
    stats[PG_FS_RAIN] = n_particles;
    if (disc_cells > 0.0 || mom_cells > 0.0) {
        // discharge scales in cells of upstream rain (maps.py:911-921): one cell's rain per iteration is
        // the spawn's volume less the inflow, over the cells it falls on
        const double vol = sc->volume0 * (double)n_particles;
        double rpc = fmax(vol - inflow, 0.0) / (double)M;
        if (!(rpc > 0.0)) rpc = vol / (double)M;
        if (disc_cells > 0.0) {
            const double ds = disc_cells * rpc;
            sc->inv_sat = ds > 0.0 ? (float)(1.0 / ds) : 0.0f;
        }
        if (mom_cells > 0.0) sc->mom_scale = (float)(sc->volume0 / (mom_cells * rpc));
    }
}

// rain droplets: the p-th smallest uniform picks the cell, the p-th offset
// pair places the droplet inside it (maps.py:821-833, particle.py:1245-1260)
__global__ void rain_drops_kernel(ll P, const double* u_sorted, const double* draws, const int* comp_idx,
                                  const double* comp_cdf, const Scalars* sc, int64_t N, int64_t H, int* d_cell,
                                  float* d_fx, float* d_fy, ll* d_load, float* d_sa, float* d_sb) {
    const ll M = sc->eligible;
    if (M == 0) return;
    const double total = comp_cdf[M - 1];
    GRID_STRIDE(p, P) {
        const double target = u_sorted[p] * total;
        ll lo = 0, hi = M - 1;
        while (lo < hi) {
            const ll mid = lo + (hi - lo) / 2;
            if (comp_cdf[mid] <= target) lo = mid + 1;
            else hi = mid;
        }
        d_cell[p] = comp_idx[lo];
        d_fx[p] = (float)draws[P + 2 * p];
        d_fy[p] = (float)draws[P + 2 * p + 1];
        d_load[p] = 0;
        d_sa[p] = 0.0f;
        d_sb[p] = 0.0f;
    }
}

__global__ void pending_flags_kernel(int64_t n, const ll* pend_fx, int* flag) {
    GRID_STRIDE(c, n) flag[c] = pend_fx[c] > 0 ? 1 : 0;
}

// stockpiles -> loaded droplets appended after the rain, in cell order, each
// starting at its cell centre in a direction drawn after the offsets
// (maps.py:886-903)
__global__ void pending_drops_kernel(int64_t n, const int* flag, const int* rank, ll* pend_fx, ll P, ull slo, ull shi,
                                     ull ilo, ull ihi, const ull* rng_dev, const Scalars* sc, int* d_cell, float* d_fx, float* d_fy,
                                     ll* d_load, float* d_sa, float* d_sb, ll* stats) {
    if (rng_dev) {
        slo = rng_dev[0];
        shi = rng_dev[1];
        ilo = rng_dev[2];
        ihi = rng_dev[3];
    }
    const u128 s0 = ((u128)shi << 64) | slo, inc = ((u128)ihi << 64) | ilo;
    ll released = 0;
    const bool any = sc->eligible > 0;
    GRID_STRIDE(c, n) {
        if (!flag[c]) continue;
        const ll r = rank[c] - 1;
        const ll load = pend_fx[c];
        if (any) {
            const ll d = P + r;
            const double u = pcg_double(pcg_advance(s0, inc, (ull)(3 * P + r + 1)));
            double sn, cs;
            sincos(u * (2.0 * PI_D), &sn, &cs);
            d_cell[d] = (int)c;
            d_fx[d] = 0.5f;
            d_fy[d] = 0.5f;
            d_load[d] = load;
            d_sa[d] = (float)cs;
            d_sb[d] = (float)sn;
            pend_fx[c] = 0;
            released += load;
        }
    }
    for (int o = 32; o > 0; o >>= 1) released += __shfl_xor(released, o);
    if ((threadIdx.x & 63) == 0 && released) atomicAdd((ull*)&stats[PG_FS_RELEASED_FX], (ull)released);
}

__global__ void totals_kernel(int64_t n, const int* rank, ll P, Scalars* sc, ll* stats, int chunks) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const ll k = sc->eligible > 0 ? (ll)rank[n - 1] : 0;
    sc->pending = k;
    sc->n_total = sc->eligible > 0 ? P + k : 0;
    stats[PG_FS_PARTICLES] = sc->n_total;
    const ll C = (sc->n_total + chunks - 1) / chunks;
    for (int c = 0; c < chunks; c++) sc->queue[2 * c] = sc->queue[2 * c + 1] = c * C;
}

// ---------------------------------------------------------------------------
// the persistent trace kernel
// ---------------------------------------------------------------------------
// one window of a batch: its cells start at arena cell `coff`, its droplets at
// `dbase`; the per-droplet scalars derived from its spawn as the single-grid
// kernel derives them
struct WinDesc {
    ll coff, dbase, n_total, C;
    int N, NE, max_steps, pad;
    float inv_v0, min_volume, conc_scale, v0f;
    float inv_sat, mom_scale, pad2, pad3;
};

struct TraceCfg {
    const float4* cell4;
    const float4* geo;
    const float2* dyn;
    const uint8_t* mask;
    const float* rscale;  // per-cell erosion scale of the chunk (pass 1 only)
    int N, H, NE, spherical, deposit_on_exit, chunk, chunks, pass;
    const int* d_cell;
    const float *d_fx, *d_fy, *d_sa, *d_sb;
    const ll* d_load;
    Scalars* sc;
    // parameters (float32)
    float dt, density, fr, rate, dt_evap, k_mom, mom_scale, k_disc, inv_sat, disc_exponent, slope_gain, slope_sat,
        erodibility, cover_depth, max_erode, ocean_rate, fan_room, dep_floor, fan_slope, min_volume_frac;
    int max_steps, pit_steps, ocean_steps;
    double hscale, tscale, conc_chunk;
    ll* chunk_fx;   // pass 1: net applied change per cell
    ll* ero_fx;     // pass 0: erosion requested per cell
    ll* fin_fx;     // pass 0: final loads of the land deaths that stop on a cell
    ull* floor_key;       // pass 0: lowest (surface key << 32 | cell) stepped to from a cell
    ll* dis_fx;
    ll* mom_fx;    // momentum track a; track b in its own plane (a move's atomics on separate lines)
    ll* mom_fx_b;
    ll* stats;
    // the applying pass's final deposits: the iteration's net change so far,
    // the stockpiles, the iteration's deposit budget (fixed point)
    const ll* acc_fx;
    ll* pend_fx;
    ll dep_budget;
    // the census's move tape (single grid only; NULL = the applying pass
    // traces again): per wave of the persistent grid, `tape_rows` rows of 64
    // lane records (two float4 each) -- row t holds what every lane of the
    // wave did in its t-th loop iteration -- and the rows used (tape_len)
    // the routing surface (ROUTE kernels): per extended cell {route, flag}
    float2* rte;
    float lake_trap;
    float lateral_rate;  // bank cuts on the outside of a bend (0: off)
    float4* tape;
    ll* tape_len;
    ll tape_rows;
    int* tape_last;  // per droplet of the chunk: the stop cell of its final deposit (a row can hold a move and an end)
    ll tape_waves;   // waves of the census grid (the replay's grid strides over them)
    // batched windows: per-window geometry and spawn totals, per-chunk union
    // offsets (n_win + 1 per chunk), the batch's queue counters
    const WinDesc* wins;
    const ll* win_prefix;
    int n_win;
    ll* queue_base;
};

// what a droplet reads at the cell it stands in
struct Here {
    int cell;  // extended cell id (grids and batch arenas hold < 2^31 cells)
    int m;
    float h, base, sthick, rs;
    float4 g0, g1;
    float2 rte;  // {route, flag} of the cell (ROUTE kernels)
};

// order-preserving key of a float (for atomicMin of surfaces)
__device__ __forceinline__ unsigned fkey(float x) {
    const unsigned b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fkey_value(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// ceiling of a step deposit on the frozen terrain as a float (3e38: none)
// and in fixed point: the previous path cell (+ dep_floor on land,
// - fan_slope under water) and the waterline
__device__ __forceinline__ ll ceiling_fx(float lim, double hscale) {
    if (lim >= 3.0e38f) return (ll)9.0e18;
    return lim > 0.0f ? llrint((double)lim * hscale) : 0;
}

// a land step deposit's ceiling is the previous path cell's live surface
// (particle.py:986-991): the frozen ceiling plus what this droplet itself
// changed there one move earlier (its own erosion lowers it, its own deposit
// raises it); other droplets' changes are not seen (order independent)
__device__ __forceinline__ ll own_ceiling(ll lim, int64_t cell, ll own_prev, int own_cell) {
    if (lim >= (ll)9.0e18 || own_cell < 0 || (int64_t)own_cell == cell) return lim;
    return lim + own_prev;
}

// the last SPREAD active cells a droplet stood in, newest first
// (particle.py:443-447 and the stop cell, :786-793)
struct Ring {
    int c[SPREAD];
    int n;
    __device__ __forceinline__ void push(int cell) {
#pragma unroll
        for (int r = SPREAD - 1; r > 0; r--) c[r] = c[r - 1];
        c[0] = cell;
        n = n < SPREAD ? n + 1 : SPREAD;
    }
};

// the final deposits of a droplet that stopped on land (particle.py:794-815,
// and apply_changes' final entries, :964-1002): the newest ring cell is
// offered the whole load, each older one what the newer ones refused; a slot
// takes at most the room left of the iteration's deposit budget and never
// rises above the next-older ring cell + dep_floor (the oldest slot has no
// such ceiling), nor above the waterline less dep_floor when submerged.
// Surfaces and budgets are the chunk's frozen ones (a cell listed twice is
// offered once); the per-cell resolve then applies the live budget.
// Returns the load no slot took (parked on the stop cell's stockpile).
__device__ __forceinline__ ll final_cascade(const TraceCfg& k, const Ring& ring, ll load) {
    ll left = load;
#pragma unroll
    for (int r = 0; r < SPREAD; r++) {
        if (r < ring.n && left > 0) {
            const int c = ring.c[r];
            bool seen = false;
#pragma unroll
            for (int q = 0; q < r; q++) seen |= ring.c[q] == c;
            if (!seen) {
                const float s_c = k.cell4[c].x;
                const float b_c = k.dyn[c].y;
                float lim = r + 1 < ring.n ? k.cell4[ring.c[r + 1]].x + k.dep_floor - s_c : 3.0e38f;
                if (s_c < b_c) lim = fminf(lim, b_c - k.dep_floor - s_c);
                ll dep = left;
                ll room = k.dep_budget - k.acc_fx[c];
                // the loads that stop on this cell share its room in proportion (the stop cell only:
                // fin_fx counts the loads that stop there, not the overflow reaching older ring cells)
                const ll demand = k.fin_fx[c];
                if (r == 0 && room > 0 && demand > room)
                    room = (ll)((double)room * fmin((double)load / (double)demand, 1.0));
                if (dep > room) dep = room;
                const ll cap = ceiling_fx(lim, k.hscale);
                if (dep > cap) dep = cap;
                if (dep > 0) {
                    atomicAdd((ull*)&k.chunk_fx[c], (ull)dep);
                    left -= dep;
                }
            }
        }
    }
    return left;
}

// tape record: flags word = transfer kind (bits 0-1: none / sea deposit /
// land step), MOVE (bit 2), START of a droplet (bit 3), END (bit 4), the end's
// cause (bits 8-15), exited (16), final deposit (17), the end's step stood
// on an active cell without recording a move (18), a bank cut (19: its cell
// and amount in the second float4, beside a start's droplet ids), final
// entries (20-23)
enum : int { REC_NONE = 0, REC_SEA = 1, REC_LAND = 2 };
enum : int { RF_MOVE = 1 << 2, RF_START = 1 << 3, RF_END = 1 << 4, RF_EXITED = 1 << 16, RF_DEPOSIT = 1 << 17,
             RF_STOOD = 1 << 18, RF_LATERAL = 1 << 19 };

// WPE: the register budget as waves per SIMD (trace_waves: the census's
// natural allocation fits 4, the traced applying pass's 3; one or two more
// waves trade spills for latency hiding; the autotuner picks).
// APPLY: false = the census pass (erosion requests and floors; with a tape,
// every move recorded for replay_kernel), true = the applying pass traced
// again (batched windows, or no tape).  ROUTE: steer on the routing surface
// and keep the lakes (particle.py:450-460, 552-560, 626-640, 751-769, 951-955).
// a bank cut's erosion request at `bcell` (fixed point), at most the room
// above the bank's own floor -- its base level, or its surface when
// submerged, or a lake's water level -- on the frozen terrain; `cut` gets
// the part refused (particle.py:941-966, the floor of the cell stepped to
// being the per-cell scale's)
__device__ __forceinline__ ll bank_request(const TraceCfg& k, int bcell, float amount, ll& cut) {
    const float hb = k.cell4[bcell].x, bb = k.dyn[bcell].y;
    float fl = hb >= bb ? bb : hb;
    if (k.rte && k.rte[bcell].y >= 1.5f) fl = fmaxf(fl, k.rte[bcell].x);
    const float room = hb - fl;
    const ll lim = room > 0.0f ? llrint((double)room * k.hscale) : 0;
    ll dq = llrint((double)amount * k.hscale);
    cut = 0;
    if (dq > lim) {
        cut = dq - lim;
        dq = lim;
    }
    return dq;
}

template <bool BATCH, int WPE, bool APPLY, bool ROUTE>
__global__ void __launch_bounds__(TRACE_TPB) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) trace_kernel(TraceCfg k) {
    const int lane = threadIdx.x & 63;
    constexpr bool apply = APPLY;
    const bool taping = !APPLY && k.tape != nullptr;
    // per-lane geometry and droplet scalars: the grid's (one launch = one
    // grid), or, batched, those of the window of the lane's droplet
    int N = k.N, NE = k.NE, max_steps = k.max_steps;
    int64_t coff = 0;
    float inv_v0 = 0.0f, min_volume = 0.0f, conc_scale = 0.0f, v0f = 0.0f;
    float inv_sat = 0.0f, mom_scale = 0.0f;  // the iteration's discharge scales (per window, batched)
    ll d_end, d_begin = 0;
    ll* queue;
    if constexpr (BATCH) {
        d_end = k.win_prefix[(ll)k.chunk * (k.n_win + 1) + k.n_win];
        queue = &k.queue_base[2 * k.chunk + k.pass];
    } else {
        const ll n_total = k.sc->n_total;
        const ll C = (n_total + k.chunks - 1) / k.chunks;
        d_begin = k.chunk * C;
        d_end = (k.chunk + 1) * C < n_total ? (k.chunk + 1) * C : n_total;
        const double volume0 = k.sc->volume0;
        if (volume0 <= 0.0) return;
        inv_v0 = (float)(1.0 / volume0);
        min_volume = (float)(k.min_volume_frac * volume0);
        // expected concurrent visits of a cell: the reference's chunk (conc_chunk
        // droplets) unless none is given, then this kernel's own chunk
        conc_scale = (float)((k.conc_chunk > 0.0 ? k.conc_chunk : (double)C) / (volume0 * (double)n_total));
        v0f = (float)volume0;
        inv_sat = k.sc->inv_sat;
        mom_scale = k.sc->mom_scale;
        queue = &k.sc->queue[2 * k.chunk + k.pass];
    }

    // lane state
    bool alive = false, exhausted = false;
    int f = 0, ix = 0, iy = 0;
    float fx = 0, fy = 0, vol = 0, sed = 0, sa = 0, sb = 0, h_prev = 0;
    bool has_prev = false, in_sea = false, in_lake = false;
    int step = 0, uphill = 0, sea_steps = 0, pushes = 0, last_cell = -1;
    ll real = 0, deficit = 0;  // load actually held (fixed point) and load cancelled by erosion caps
    ll surplus = 0;            // deposits a ceiling refused, offered to the droplet's next deposits
    ll own_prev = 0;           // what the droplet changed at its previous path cell (applying pass)
    int own_cell = -1;         // that cell
    Ring ring;                 // applying pass: the cells of the final deposits
    ring.n = 0;
    int dl = 0;                // the droplet's slot in the chunk (a chunk holds < 2^31 droplets)
    // the wave's tape rows (census with a tape): wave-uniform row counter
    const ll gw = (ll)blockIdx.x * (TRACE_TPB / 64) + (threadIdx.x >> 6);
    float4* wtape = taping ? k.tape + 2 * gw * k.tape_rows * 64 : nullptr;
    ll t_row = 0;
    Samp S;  // the sample of the stencil at the droplet's position
    Here cur;
    // lane totals
    ll n_entries = 0, n_moves = 0, exit_fx = 0, vanished_fx = 0, lost_fx = 0, to_pend_fx = 0, lat_fx = 0;
    ll deaths[N_DEATHS] = {0, 0, 0, 0, 0, 0};

    auto load_here = [&](int ff, int cx, int cy, Here& h, const float4* rec) {
        const int ci = clampi(cx, 0, N - 1), cj = clampi(cy, 0, N - 1);
        h.cell = (int)(coff + ((int64_t)ff * NE + ci + k.H) * NE + cj + k.H);
        h.m = k.mask[h.cell];
        const float2 dy = k.dyn[h.cell];
        h.sthick = dy.x;
        h.base = dy.y;
        h.h = rec ? rec->x : k.cell4[h.cell].x;
        h.g0 = k.geo[2 * h.cell];
        h.g1 = k.geo[2 * h.cell + 1];
        h.rs = apply ? k.rscale[h.cell] : 1.0f;
        if constexpr (ROUTE) h.rte = k.rte[h.cell];
        else h.rte = make_float2(0.0f, 0.0f);
    };
    // deposit request `d` (fixed point) at `cell`: with the surplus of earlier
    // refused deposits, less the deficit, at most `lim` (particle.py:969-1002)
    auto deposit = [&](int64_t cell, ll d, ll lim) -> ll {
        ll want = d + surplus;
        surplus = 0;
        ll a = want < deficit ? want : deficit;
        if (a < 0) a = 0;
        deficit -= a;
        want -= a;
        ll dep = want < lim ? want : lim;
        if (dep < 0) dep = 0;
        surplus = want - dep;
        if (dep) atomicAdd((ull*)&k.chunk_fx[cell], (ull)dep);
        real -= dep;
        return dep;
    };
    // ceiling of a step deposit on the frozen terrain: the previous path cell
    // (+ dep_floor on land, - fan_slope under water) and the waterline
    auto ceiling_lim = [&](float h_here, float b_here) -> float {
        float lim = 3.0e38f;
        if (has_prev) {
            float ceil_ = h_prev + k.dep_floor;
            if (h_here < b_here) ceil_ -= k.dep_floor + k.fan_slope;
            lim = ceil_ - h_here;
        }
        if (h_here < b_here) lim = fminf(lim, b_here - k.dep_floor - h_here);
        return lim;
    };
    auto step_ceiling = [&](float h_here, float b_here) -> ll { return ceiling_fx(ceiling_lim(h_here, b_here), k.hscale); };

    while (true) {
        // this iteration's tape record
        int r_flags = 0, r_cell = 0, r_dl = 0, r_d = 0, r_bank = 0;
        float r_x = 0.0f, r_aux = 0.0f, r_lat = 0.0f;
        // a wave whose tape rows could not hold one more droplet takes none
        if (taping && !exhausted && t_row + k.max_steps + 2 > k.tape_rows) exhausted = true;
        // refill the lanes whose droplet died from the chunk's queue
        if (!exhausted) {
            const ull dead = __ballot(!alive);
            if (dead) {
                const int leader = __ffsll((ll)dead) - 1;
                ll base = 0;
                if (lane == leader) base = (ll)atomicAdd((ull*)queue, (ull)__popcll(dead));
                base = __shfl(base, leader);
                if (base >= d_end) exhausted = true;
                if (!alive) {
                    const ll u = base + __popcll(dead & ((1ull << lane) - 1));
                    ll d = u;
                    if constexpr (BATCH) {
                        if (u < d_end) {
                            // the window whose share of this chunk holds u
                            const ll* pre = k.win_prefix + (ll)k.chunk * (k.n_win + 1);
                            int lo = 0, hi = k.n_win - 1;
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (pre[mid] <= u) lo = mid;
                                else hi = mid - 1;
                            }
                            const WinDesc wd = k.wins[lo];
                            coff = wd.coff;
                            N = wd.N;
                            NE = wd.NE;
                            max_steps = wd.max_steps;
                            inv_v0 = wd.inv_v0;
                            min_volume = wd.min_volume;
                            conc_scale = wd.conc_scale;
                            v0f = wd.v0f;
                            inv_sat = wd.inv_sat;
                            mom_scale = wd.mom_scale;
                            d = wd.dbase + (ll)k.chunk * wd.C + (u - pre[lo]);
                        }
                    }
                    if (u < d_end) {
                        dl = (int)(u - d_begin);
                        if (taping) {
                            r_flags |= RF_START;
                            r_dl = (int)dl;
                            r_d = (int)d;
                        }
                        const int NE2 = NE * NE;
                        const int cell = k.d_cell[d];
                        f = cell / NE2;
                        const int r = cell - f * NE2;
                        ix = r / NE - k.H;
                        iy = r % NE - k.H;
                        fx = k.d_fx[d];
                        fy = k.d_fy[d];
                        real = k.d_load[d];
                        deficit = 0;
                        surplus = 0;
                        vol = v0f;
                        sed = (float)((double)real / k.hscale);
                        sa = k.d_sa[d];
                        sb = k.d_sb[d];
                        has_prev = false;
                        own_prev = 0;
                        own_cell = -1;
                        ring.n = 0;
                        in_sea = false;
                        in_lake = false;
                        step = uphill = sea_steps = pushes = 0;
                        last_cell = -1;
                        if constexpr (ROUTE)
                            S = sample_route(load_stencil(k.cell4 + coff, NE, k.H, f, ix, iy, fx, fy), k.rte + coff, f, NE);
                        else
                            S = sample_of(load_stencil(k.cell4 + coff, NE, k.H, f, ix, iy, fx, fy));
                        load_here(f, ix, iy, cur, nullptr);
                        alive = true;
                    }
                }
            }
        }
        if (!__any(alive)) break;
        do {
        if (!alive) break;

        // ---- one step of this lane's droplet (particle.py:421-769)
        int cause = -1;
        bool exited = false, stood = false;
        if (step >= max_steps) {
            cause = DEATH_AGE;
        } else if (cur.m == MASK_OUTSIDE) {
            cause = DEATH_EXIT;
            exited = true;
        } else {
            const float h_here = cur.h, b_here = cur.base;
            if (step == 0 && h_here < b_here) in_sea = true;
            if (in_sea) {
                if (h_here >= b_here || sed * vol * inv_v0 < 1e-4f || sea_steps >= k.ocean_steps) cause = DEATH_OCEAN;
                else sea_steps += 1;
            }
            if (cause < 0) {
                if (cur.m == MASK_ACTIVE) {
                    last_cell = (int)cur.cell;
                    pushes += 1;
                    stood = true;
                    if (apply) ring.push((int)cur.cell);
                }
                const float4 g0 = cur.g0, g1 = cur.g1;
                const float q = S.q;
                const float vol_rel = vol * inv_v0;
                int nf, nix, niy;
                float nfx, nfy;
                const float h0 = S.h;
                const float psa = sa, psb = sb;  // the direction before this step (a bank cut's turn)
                cause = steer_move<true>(k, N, S, g0, g1, vol, vol_rel, mom_scale, sa, sb, f, ix, iy, fx, fy, nf, nix, niy,
                                         nfx, nfy);
                exited = cause == DEATH_EXIT;
                {
                    if (cause < 0) {
                        // everything the next step reads, issued together
                        const int nei = nix + k.H, nej = niy + k.H;
                        const int64_t ncell = coff + ((int64_t)nf * NE + nei) * NE + nej;
                        const Stencil S1 = load_stencil(k.cell4 + coff, NE, k.H, nf, nix, niy, nfx, nfy);
                        const int nm = k.mask[ncell];
                        const float2 ndy = k.dyn[ncell];
                        const float4 nrec = corner_of(S1, nei, nej);
                        Samp P1;
                        float nflag = 0.0f;  // the route flag of the cell stepped to
                        if constexpr (ROUTE) {
                            P1 = sample_route(S1, k.rte + coff, nf, NE);
                            nflag = k.rte[ncell].y;
                        } else {
                            P1 = sample_of(S1);
                        }
                        Here nxt;
                        const bool same = nix >= 0 && nix < N && niy >= 0 && niy < N;
                        load_here(nf, nix, niy, nxt, same ? &nrec : nullptr);
                        if (nm == MASK_OUTSIDE) {
                            cause = DEATH_EXIT;
                            exited = true;
                        } else {
                            const float h1 = P1.h;
                            const float b1 = ndy.y;
                            const float dh = h0 - fmaxf(h1, b1);
                            int rk = REC_NONE;
                            float rx = 0.0f, rroom = 0.0f;
                            ll chg = 0;  // applying pass: the change made at this cell
                            // mass transfer at the cell being left
                            if (cur.m == MASK_ACTIVE && in_sea) {
                                float cdiff = -k.ocean_rate * sed;
                                float room = has_prev ? h_prev - h_here : k.fan_room;
                                room = fmaxf(room, k.fan_room);
                                room = fminf(room, b_here - k.dep_floor - h_here);
                                room = fmaxf(room, 0.0f);
                                if (-cdiff * vol_rel > room) cdiff = -room / vol_rel;
                                sed += cdiff;
                                if (apply) {
                                    chg = deposit(cur.cell, llrint((double)(-cdiff * vol_rel) * k.hscale),
                                                  step_ceiling(h_here, b_here));
                                    n_entries += 1;
                                } else if (taping) {
                                    rk = REC_SEA;
                                    rx = -cdiff * vol_rel;
                                }
                            } else if (cur.m == MASK_ACTIVE) {
                                float qe;
                                if (inv_sat <= 0.0f) qe = 0.0f;
                                else if (k.disc_exponent == 0.5f) qe = sqrtf(q * inv_sat);
                                else if (k.disc_exponent > 0.0f) qe = powf(q * inv_sat, k.disc_exponent);
                                else qe = erff(q * inv_sat);
                                const float c_eq = fmaxf(k.erodibility * dh * (1.0f + k.k_disc * qe), 0.0f);
                                float k_e = k.rate;
                                if (c_eq > sed) {
                                    const float w_cover = k.cover_depth > 0.0f ? fminf(cur.sthick / k.cover_depth, 1.0f)
                                                                               : (cur.sthick > 0.0f ? 1.0f : 0.0f);
                                    k_e = k.rate * (w_cover + (1.0f - w_cover) * (1.0f - g1.w));
                                }
                                float cdiff = k_e * (c_eq - sed);
                                if constexpr (ROUTE) {
                                    // crossing a lake: no exchange with the bed; entering one: the load
                                    // drops at the shore (particle.py:626-640)
                                    if (in_lake) cdiff = 0.0f;
                                    else if (sed > 0.0f && nflag >= 1.5f) cdiff = fminf(cdiff, -k.lake_trap * sed);
                                }
                                const float kexp = 1.0f + q * conc_scale;
                                if (cdiff > 0.0f) {
                                    float ecap = fminf(dh / kexp, k.max_erode);
                                    if (ecap > h_here) ecap = fmaxf(h_here, 0.0f);
                                    if (cdiff * vol_rel > ecap) cdiff = ecap / vol_rel;
                                } else if (cdiff < 0.0f) {
                                    float room = has_prev ? h_prev - h_here : k.dep_floor;
                                    room = fmaxf(room, k.dep_floor);
                                    const float dcap = room / kexp;
                                    if (-cdiff * vol_rel > dcap) cdiff = -dcap / vol_rel;
                                }
                                sed += cdiff;
                                const float amount = cdiff * vol_rel;
                                ll dq = llrint((double)amount * k.hscale);  // > 0: erosion
                                ll cut = 0;
                                if (taping) {
                                    rk = REC_LAND;
                                    rx = amount;
                                }
                                if (dq > 0) {
                                    // an entry never takes the cell below its base level (particle.py:951);
                                    // the floor of the cell it steps to, which the reference takes live
                                    // (:956-960), is the per-cell scale's (census floor key, two sweeps)
                                    float fl = h_here >= b_here ? b_here : h_here;
                                    // a lake bed is never eroded below its water level (particle.py:952-955)
                                    if (ROUTE && cur.rte.y >= 1.5f) fl = fmaxf(fl, cur.rte.x);
                                    const float room = h_here - fl;
                                    rroom = room;
                                    const ll lim = room > 0.0f ? llrint((double)room * k.hscale) : 0;
                                    if (dq > lim) {
                                        cut = dq - lim;
                                        dq = lim;
                                    }
                                }
                                if (!apply) {
                                    if (dq > 0) {
                                        atomicAdd((ull*)&k.ero_fx[cur.cell], (ull)dq);
                                        atomicMin(&k.floor_key[cur.cell], ((ull)fkey(nrec.x) << 32) | (ull)(unsigned)ncell);
                                    }
                                    if (taping) {
                                        // the tracks do not depend on the erosion scale: added here, where the
                                        // trace leaves the atomic units idle, not in the replay
                                        atomicAdd((ull*)&k.dis_fx[cur.cell], (ull)llrint((double)vol * k.tscale));
                                        atomicAdd((ull*)&k.mom_fx[cur.cell], (ull)llrint((double)(vol * sa) * k.tscale));
                                        atomicAdd((ull*)&k.mom_fx_b[cur.cell], (ull)llrint((double)(vol * sb) * k.tscale));
                                    }
                                } else {
                                    deficit += cut;
                                    if (dq > 0) {
                                        const ll taken = (ll)floor((double)dq * (double)cur.rs);
                                        deficit += dq - taken;
                                        real += taken;
                                        chg = -taken;
                                        if (taken) atomicAdd((ull*)&k.chunk_fx[cur.cell], (ull)(-taken));
                                    } else if (dq < 0) {
                                        chg = deposit(cur.cell, -dq,
                                                      own_ceiling(step_ceiling(h_here, b_here), cur.cell, own_prev, own_cell));
                                    }
                                    atomicAdd((ull*)&k.dis_fx[cur.cell], (ull)llrint((double)vol * k.tscale));
                                    atomicAdd((ull*)&k.mom_fx[cur.cell], (ull)llrint((double)(vol * sa) * k.tscale));
                                    atomicAdd((ull*)&k.mom_fx_b[cur.cell], (ull)llrint((double)(vol * sb) * k.tscale));
                                    n_entries += 1;
                                }
                                if (k.lateral_rate > 0.0f && cdiff > 0.0f) {
                                    // the bank on the outside of the bend is cut as well, by the turn's
                                    // sine times the bed's erosion, at most the bank's height over the
                                    // bed (particle.py:707-740); it erodes like a step's request, its
                                    // floor the cell stepped to
                                    const float cross = psa * sb - psb * sa;
                                    const float turn = fabsf(cross);
                                    if (turn > 1e-6f) {
                                        const float sgn = cross > 0.0f ? 1.0f : -1.0f;
                                        const int bei = ix + k.H + (int)floorf(fx + sgn * sb);
                                        const int bej = iy + k.H + (int)floorf(fy - sgn * sa);
                                        if (bei >= 0 && bei < NE && bej >= 0 && bej < NE) {
                                            const int bcell = (int)(coff + ((int64_t)f * NE + bei) * NE + bej);
                                            const float bank = k.cell4[bcell].x - h_here;
                                            if (k.mask[bcell] == MASK_ACTIVE && bank > 0.0f) {
                                                float lat = k.lateral_rate * turn * cdiff;
                                                const float lcap = fminf(bank / kexp, k.max_erode);
                                                if (lat * vol_rel > lcap) lat = lcap / vol_rel;
                                                if (lat > 0.0f) {
                                                    sed += lat;
                                                    const float lamount = lat * vol_rel;
                                                    if (taping) {
                                                        r_flags |= RF_LATERAL;
                                                        r_bank = bcell;
                                                        r_lat = lamount;
                                                    }
                                                    ll cutb = 0;
                                                    const ll dqb = bank_request(k, bcell, lamount, cutb);
                                                    if (!apply) {
                                                        if (dqb > 0) {
                                                            atomicAdd((ull*)&k.ero_fx[bcell], (ull)dqb);
                                                            atomicMin(&k.floor_key[bcell],
                                                                      ((ull)fkey(nrec.x) << 32) | (ull)(unsigned)ncell);
                                                        }
                                                    } else {
                                                        deficit += cutb;
                                                        if (dqb > 0) {
                                                            const ll taken = (ll)floor((double)dqb * (double)k.rscale[bcell]);
                                                            deficit += dqb - taken;
                                                            real += taken;
                                                            lat_fx += taken;
                                                            if (taken) atomicAdd((ull*)&k.chunk_fx[bcell], (ull)(-taken));
                                                        }
                                                        n_entries += 1;
                                                    }
                                                }
                                            }
                                        }
                                    }
                                }
                            }
                            if (apply) {
                                own_prev = chg;
                                own_cell = (int)cur.cell;
                            }
                            if (taping) {
                                // the move as the applying pass needs it: cell, kind, amount, and the
                                // floor room of an erosion or the ceiling of a deposit; the track inputs
                                r_flags |= RF_MOVE | rk;
                                r_cell = (int)cur.cell;
                                r_x = rx;
                                r_aux = (rk == REC_LAND && rx > 0.0f) ? rroom
                                        : (rk != REC_NONE ? ceiling_lim(h_here, b_here) : 0.0f);
                            }
                            // evaporation, advance, stop tests
                            const float ev = fmaxf(1.0f - k.dt_evap * g1.z, 0.01f);
                            vol *= ev;
                            sed /= ev;
                            f = nf;
                            ix = nix;
                            iy = niy;
                            fx = nfx;
                            fy = nfy;
                            const float dhr = S.hr - P1.hr;  // the drop on the routing surface
                            S = P1;
                            cur = nxt;
                            h_prev = h_here;
                            if constexpr (ROUTE) in_lake = nflag >= 1.5f;
                            has_prev = true;
                            n_moves += 1;
                            step += 1;
                            if (nrec.x < ndy.y && !in_sea) {
                                in_sea = true;
                                if (sed <= 0.0f) cause = DEATH_OCEAN;
                            }
                            if (cause < 0 && vol < min_volume) cause = in_sea ? DEATH_OCEAN : DEATH_EVAP;
                            if (cause < 0) {
                                if (dhr < 0.0f && !in_lake) {
                                    uphill += 1;
                                    if (uphill > k.pit_steps) cause = in_sea ? DEATH_OCEAN : DEATH_PIT;
                                } else {
                                    uphill = 0;
                                }
                            }
                            if (cause < 0 && step >= max_steps) cause = DEATH_AGE;
                        }
                    }
                }
            }
        }
        if (cause < 0) break;

        // ---- the droplet stopped: final deposit (particle.py:770-815)
        if (taping) {
            int cz = cause;
            if (in_sea && (cz == DEATH_AGE || cz == DEATH_STOP)) cz = DEATH_OCEAN;
            int lc = last_cell, pu = pushes;
            if (!exited && cur.m == MASK_ACTIVE && !(in_sea && cur.h >= cur.base) && (pu == 0 || lc != (int)cur.cell)) {
                lc = (int)cur.cell;
                pu += 1;
            }
            const bool dodep = pu > 0 && (!exited || k.deposit_on_exit);
            const int nfin = dodep && sed > 0.0f ? (cz == DEATH_OCEAN ? 1 : (pu < SPREAD ? pu : SPREAD)) : 0;
            r_flags |= RF_END | (cz << 8) | (exited ? RF_EXITED : 0) | (dodep ? RF_DEPOSIT : 0) | (nfin << 20) |
                       (stood && !(r_flags & RF_MOVE) ? RF_STOOD : 0);
            k.tape_last[dl] = lc;
        }
        if (!apply) {
            // a land death's final load, on the cell it stops in: the applying pass shares each
            // cell's deposit room among the loads that stop there
            int cz = cause;
            if (in_sea && (cz == DEATH_AGE || cz == DEATH_STOP)) cz = DEATH_OCEAN;
            int lc = last_cell, pu = pushes;
            if (!exited && cur.m == MASK_ACTIVE && !(in_sea && cur.h >= cur.base) && (pu == 0 || lc != (int)cur.cell)) {
                lc = (int)cur.cell;
                pu += 1;
            }
            if (pu > 0 && (!exited || k.deposit_on_exit) && cz != DEATH_OCEAN && sed > 0.0f) {
                const ll load = llrint((double)(sed * vol * inv_v0) * k.hscale);
                if (load > 0) atomicAdd((ull*)&k.fin_fx[lc], (ull)load);
            }
        }
        if (apply) {
            if (in_sea && (cause == DEATH_AGE || cause == DEATH_STOP)) cause = DEATH_OCEAN;
            deaths[cause] += 1;
            if (!exited) {
                // the stop cell is `cur` (the cell of the current position)
                if (cur.m == MASK_ACTIVE && !(in_sea && cur.h >= cur.base)) {
                    if (pushes == 0 || last_cell != (int)cur.cell) {
                        last_cell = (int)cur.cell;
                        pushes += 1;
                        ring.push((int)cur.cell);
                    }
                }
            }
            if (pushes > 0 && (!exited || k.deposit_on_exit)) {
                if (sed > 0.0f) n_entries += cause == DEATH_OCEAN ? 1 : (pushes < SPREAD ? pushes : SPREAD);
                if (cause != DEATH_OCEAN && real > 0) {
                    const ll left = final_cascade(k, ring, real);
                    if (left) {
                        atomicAdd((ull*)&k.pend_fx[last_cell], (ull)left);
                        to_pend_fx += left;
                    }
                } else if (real) {
                    atomicAdd((ull*)&k.chunk_fx[last_cell], (ull)real);
                }
            } else if (exited) {
                exit_fx += real;
            } else {
                vanished_fx += real;
            }
            lost_fx += deficit;
        }
        real = 0;
        deficit = 0;
        surplus = 0;
        alive = false;
        } while (false);
        if (taping) {
            // one row per loop iteration: the wave's 64 records side by side
            float4* row = wtape + 2 * (t_row * 64 + lane);
            row[0] = make_float4(__int_as_float(r_flags), __int_as_float(r_cell), r_x, r_aux);
            if (r_flags & (RF_START | RF_LATERAL))
                row[1] = make_float4(__int_as_float(r_bank), r_lat, __int_as_float(r_d), __int_as_float(r_dl));
        }
        t_row += 1;
    }
    if (taping && lane == 0) k.tape_len[gw] = t_row;
    if (!apply) return;
    // lane totals -> device statistics (one atomic per wave and counter)
    ll v[7 + N_DEATHS] = {n_entries, n_moves, exit_fx, vanished_fx, lost_fx, to_pend_fx, lat_fx};
    for (int c = 0; c < N_DEATHS; c++) v[7 + c] = deaths[c];
    for (int c = 0; c < 7 + N_DEATHS; c++) v[c] = wave_sum(v[c]);
    if (lane == 0) {
        const int slot[7] = {PG_FS_ENTRIES, PG_FS_MOVES, PG_FS_EXIT_FX, PG_FS_VANISHED_FX, PG_FS_DEFICIT_FX,
                             PG_FS_TO_PENDING_FX, PG_FS_LATERAL_FX};
        for (int c = 0; c < 7; c++)
            if (v[c]) atomicAdd((ull*)&k.stats[slot[c]], (ull)v[c]);
        for (int c = 0; c < N_DEATHS; c++)
            if (v[7 + c]) atomicAdd((ull*)&k.stats[PG_FS_DEATHS + c], (ull)v[7 + c]);
    }
}

// waves per SIMD of a register-budget variant (0: the kernel's natural
// budget -- 121 VGPRs for the census, 141 for the traced applying pass --
// 1 and 2: one and two waves more)
template <bool APPLY>
constexpr int trace_waves(int variant) {
    return (APPLY ? 3 : 4) + variant;
}

// the trace kernel of a register-budget variant and pass (the persistent
// grid is sized on the kernel without the routing surface, for both)
template <bool BATCH, bool APPLY>
const void* trace_fn(int64_t variant) {
    if (variant == 1) return (const void*)trace_kernel<BATCH, trace_waves<APPLY>(1), APPLY, false>;
    if (variant == 2) return (const void*)trace_kernel<BATCH, trace_waves<APPLY>(2), APPLY, false>;
    return (const void*)trace_kernel<BATCH, trace_waves<APPLY>(0), APPLY, false>;
}

template <bool BATCH, bool APPLY, bool ROUTE>
void launch_trace_v(int64_t variant, int grid, hipStream_t s, const TraceCfg& k) {
    if (variant == 1)
        hipLaunchKernelGGL((trace_kernel<BATCH, trace_waves<APPLY>(1), APPLY, ROUTE>), dim3(grid), dim3(TRACE_TPB), 0, s, k);
    else if (variant == 2)
        hipLaunchKernelGGL((trace_kernel<BATCH, trace_waves<APPLY>(2), APPLY, ROUTE>), dim3(grid), dim3(TRACE_TPB), 0, s, k);
    else
        hipLaunchKernelGGL((trace_kernel<BATCH, trace_waves<APPLY>(0), APPLY, ROUTE>), dim3(grid), dim3(TRACE_TPB), 0, s, k);
}

// the routing-surface kernels when the launch carries a route (k.rte)
template <bool BATCH, bool APPLY>
void launch_trace(int64_t variant, int grid, hipStream_t s, const TraceCfg& k) {
    if (k.rte) launch_trace_v<BATCH, APPLY, true>(variant, grid, s, k);
    else launch_trace_v<BATCH, APPLY, false>(variant, grid, s, k);
}

// ---------------------------------------------------------------------------
// the applying pass from the census's tape: every wave of the census's grid
// replays its own rows -- the same lanes, droplets and moves in the same
// order, each row one coalesced read -- with the chunk's per-cell erosion
// scales: the fixed-point arithmetic of the traced applying pass
// (bit-identical results) without tracing again.  The discharge and
// momentum tracks, which the scales do not touch, were added by the census.
// ---------------------------------------------------------------------------
template <bool BATCH>
__global__ void __launch_bounds__(TRACE_TPB) __attribute__((amdgpu_waves_per_eu(7))) replay_kernel(TraceCfg k) {
    const int lane = threadIdx.x & 63;
    const ll rw = (ll)blockIdx.x * (TRACE_TPB / 64) + (threadIdx.x >> 6);
    const ll rwaves = (ll)gridDim.x * (TRACE_TPB / 64);
    // the chunk's droplets (the census's queue counted them out) and the end of its range
    ll d_end;
    const ll* queue;
    if constexpr (BATCH) {
        d_end = k.win_prefix[(ll)k.chunk * (k.n_win + 1) + k.n_win];
        queue = &k.queue_base[2 * k.chunk];
    } else {
        const ll n_total = k.sc->n_total;
        const ll C = (n_total + k.chunks - 1) / k.chunks;
        d_end = (k.chunk + 1) * C < n_total ? (k.chunk + 1) * C : n_total;
        queue = &k.sc->queue[2 * k.chunk];
        if (k.sc->volume0 <= 0.0) return;  // the census traced nothing (and wrote no rows)
    }
    ll real = 0, deficit = 0, surplus = 0;
    ll own_prev = 0;  // what the droplet changed at its previous path cell
    int own_cell = -1, dl = 0;
    Ring ring;        // the cells of its final deposits
    ring.n = 0;
    int n_entries = 0, n_moves = 0;  // per lane: below 2^31
    ll exit_fx = 0, vanished_fx = 0, lost_fx = 0, to_pend_fx = 0, lat_fx = 0;
    int deaths[N_DEATHS] = {0, 0, 0, 0, 0, 0};
    // a deposit request: the amount laid down (the caller adds it to the cell)
    auto deposit = [&](ll d, ll lim) -> ll {
        ll want = d + surplus;
        surplus = 0;
        ll a = want < deficit ? want : deficit;
        if (a < 0) a = 0;
        deficit -= a;
        want -= a;
        ll dep = want < lim ? want : lim;
        if (dep < 0) dep = 0;
        surplus = want - dep;
        real -= dep;
        return dep;
    };
    // each replay wave takes census waves rw, rw + rwaves, ...: every census
    // wave's rows are replayed by one wave, lane for lane, in order
    for (ll gw = rw; gw < k.tape_waves; gw += rwaves) {
        const float4* wtape = k.tape + 2 * gw * k.tape_rows * 64;
        const ll rows = k.tape_len[gw];
        for (ll t = 0; t < rows; t++) {
            const float4* row = wtape + 2 * (t * 64 + lane);
            const float4 A = row[0];
            const int fl = __float_as_int(A.x);
            if (!fl) continue;
            if (fl & RF_START) {
                const float4 B = row[1];
                dl = __float_as_int(B.w);
                real = k.d_load[__float_as_int(B.z)];
                deficit = surplus = 0;
                own_prev = 0;
                own_cell = -1;
                ring.n = 0;
            }
            if (fl & RF_MOVE) {
                // one height-change atomic per row for erosion and deposit alike (full lanes per instruction)
                const int64_t cell = (int64_t)__float_as_int(A.y);
                const int kind = fl & 3;
                ll change = 0;
                if (kind == REC_SEA) {
                    change = deposit(llrint((double)A.z * k.hscale), ceiling_fx(A.w, k.hscale));
                    n_entries += 1;
                } else if (kind == REC_LAND) {
                    ll dq = llrint((double)A.z * k.hscale);
                    ll cut = 0;
                    if (dq > 0) {
                        const ll lim = A.w > 0.0f ? llrint((double)A.w * k.hscale) : 0;
                        if (dq > lim) {
                            cut = dq - lim;
                            dq = lim;
                        }
                    }
                    deficit += cut;
                    if (dq > 0) {
                        const ll taken = (ll)floor((double)dq * (double)k.rscale[cell]);
                        deficit += dq - taken;
                        real += taken;
                        change = -taken;
                    } else if (dq < 0) {
                        change = deposit(-dq, own_ceiling(ceiling_fx(A.w, k.hscale), cell, own_prev, own_cell));
                    }
                    n_entries += 1;
                }
                own_prev = change;
                own_cell = (int)cell;
                if (kind != REC_NONE) ring.push((int)cell);  // a move from an active cell
                if (change) atomicAdd((ull*)&k.chunk_fx[cell], (ull)change);
                n_moves += 1;
            }
            if (fl & RF_LATERAL) {
                // the move's bank cut, after the move's own change (as the traced pass)
                const float4 B = row[1];
                const int bcell = __float_as_int(B.x);
                ll cutb = 0;
                const ll dqb = bank_request(k, bcell, B.y, cutb);
                deficit += cutb;
                if (dqb > 0) {
                    const ll taken = (ll)floor((double)dqb * (double)k.rscale[bcell]);
                    deficit += dqb - taken;
                    real += taken;
                    lat_fx += taken;
                    if (taken) atomicAdd((ull*)&k.chunk_fx[bcell], (ull)(-taken));
                }
                n_entries += 1;
            }
            if (fl & RF_END) {
                // the droplet's end (particle.py:770-815): cause, stop cell and final entries as the census saw them
                deaths[(fl >> 8) & 0xff] += 1;
                n_entries += (fl >> 20) & 0xf;
                if (fl & RF_DEPOSIT) {
                    // the ring as the droplet left it: the end's own step, then the stop cell
                    const int last = k.tape_last[dl];
                    if (fl & RF_STOOD) ring.push(last);
                    if (!(fl & RF_EXITED) && (ring.n == 0 || ring.c[0] != last)) ring.push(last);
                    if (((fl >> 8) & 0xff) != DEATH_OCEAN && real > 0) {
                        const ll left = final_cascade(k, ring, real);
                        if (left) {
                            atomicAdd((ull*)&k.pend_fx[last], (ull)left);
                            to_pend_fx += left;
                        }
                    } else if (real) {
                        atomicAdd((ull*)&k.chunk_fx[last], (ull)real);
                    }
                } else if (fl & RF_EXITED) {
                    exit_fx += real;
                } else {
                    vanished_fx += real;
                }
                lost_fx += deficit;
                real = deficit = surplus = 0;
            }
        }
    }
    ll v[7 + N_DEATHS] = {n_entries, n_moves, exit_fx, vanished_fx, lost_fx, to_pend_fx, lat_fx};
    for (int c = 0; c < N_DEATHS; c++) v[7 + c] = deaths[c];
    for (int c = 0; c < 7 + N_DEATHS; c++) v[c] = wave_sum(v[c]);
    if (lane == 0) {
        const int slot[7] = {PG_FS_ENTRIES, PG_FS_MOVES, PG_FS_EXIT_FX, PG_FS_VANISHED_FX, PG_FS_DEFICIT_FX,
                             PG_FS_TO_PENDING_FX, PG_FS_LATERAL_FX};
        for (int c = 0; c < 7; c++)
            if (v[c]) atomicAdd((ull*)&k.stats[slot[c]], (ull)v[c]);
        for (int c = 0; c < N_DEATHS; c++)
            if (v[7 + c]) atomicAdd((ull*)&k.stats[PG_FS_DEATHS + c], (ull)v[7 + c]);
    }
    // droplets the census left untraced because every wave's rows were full: must be none
    if (rw == 0 && lane == 0) {
        const ll taken = *queue;
        if (taken < d_end) atomicAdd((ull*)&k.stats[PG_FS_TAPE_SHORT], (ull)(d_end - taken));
    }
}

// ---------------------------------------------------------------------------
// per-cell erosion scale of a chunk (between the two trace passes): a cell
// may lose at most its iteration budget (iter_erode + this iteration's net
// change so far) and never sink below its base level (land) nor below the
// lowest surface a droplet stepped to from it -- the reference's per-entry
// floor, taken over the whole chunk (particle.py:937-968)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float erosion_scale(int64_t c, const double* height, const double* sediment,
                                               const double* base, const ll* acc_fx, ll e, double fl_next,
                                               double iter_erode, const double* ecap, double inv, const float2* rte) {
    const double acc = (double)acc_fx[c] * inv;
    const double s_c = height[c] + sediment[c] + acc;
    const double b = base[c];
    double fl = s_c >= b ? b : s_c;
    if (rte && rte[c].y >= 1.5f) fl = fmax(fl, (double)rte[c].x);  // a lake bed stays under its water
    fl = fmax(fl, fl_next);
    double room = iter_erode + acc;
    if (ecap) {  // the slope limit: at most erode_cap + this iteration's net change (particle.py:944-947)
        const double r2 = ecap[c] + acc;
        if (r2 < room) room = r2;
    }
    double allowed = fmin(room, s_c - fl);
    if (allowed < 0.0) allowed = 0.0;
    const double want = (double)e * inv;
    return want > allowed ? (float)(allowed / want) : 1.0f;
}

// sweep 1: floor = the lowest surface stepped to (frozen)
__global__ void erosion_scale_kernel(int64_t n, const double* height, const double* sediment, const double* base,
                                     const ll* ero_fx, const ull* floor_key, const ll* acc_fx, float* r1,
                                     double iter_erode, const double* ecap, double hscale, const float2* rte) {
    const double inv = 1.0 / hscale;
    GRID_STRIDE(c, n) {
        const ll e = ero_fx[c];
        const ull fk = floor_key[c];
        r1[c] = e > 0 ? erosion_scale(c, height, sediment, base, acc_fx, e,
                                      fk != ~0ull ? (double)fkey_value((unsigned)(fk >> 32)) : -1.0e300, iter_erode, ecap,
                                      inv, rte)
                      : 1.0f;
    }
}

// sweep 2: that floor lowered by what sweep 1 lets the floor cell lose
__global__ void erosion_scale2_kernel(int64_t n, const double* height, const double* sediment, const double* base,
                                      const ll* ero_fx, const ull* floor_key, const ll* acc_fx, const float* r1,
                                      float* rscale, double iter_erode, const double* ecap, double hscale,
                                      const float2* rte) {
    const double inv = 1.0 / hscale;
    GRID_STRIDE(c, n) {
        const ll e = ero_fx[c];
        const ull fk = floor_key[c];
        float r = 1.0f;
        if (e > 0) {
            double fl_next = -1.0e300;
            if (fk != ~0ull) {
                const int64_t nc = (int64_t)(unsigned)(fk & 0xffffffffull);
                fl_next = (double)fkey_value((unsigned)(fk >> 32)) - (double)ero_fx[nc] * inv * (double)r1[nc];
            }
            r = erosion_scale(c, height, sediment, base, acc_fx, e, fl_next, iter_erode, ecap, inv, rte);
        }
        rscale[c] = r;
    }
}

// ---------------------------------------------------------------------------
// per-cell resolve of a chunk against the live terrain
// ---------------------------------------------------------------------------
__global__ void resolve_kernel(int64_t n, const double* height, const double* sediment, const double* base, ll* chunk_fx,
                               ll* acc_fx, ll* pend_fx, float4* cell4, float2* dyn, double iter_deposit,
                               double dep_floor, double writeoff, double hscale, ll* stats, ll* ero_fx, ull* floor_key,
                               float2* rte, ll* fin_fx, ll* zero_also) {
    const double inv = 1.0 / hscale;
    ll clamped = 0, to_pend = 0, lost_off = 0;
    GRID_STRIDE(c, n) {
        ero_fx[c] = 0;
        fin_fx[c] = 0;
        floor_key[c] = ~0ull;
        if (zero_also) zero_also[c] = 0;  // the brushed chunk's unbrushed change
        const ll net = chunk_fx[c];
        if (net == 0) continue;
        chunk_fx[c] = 0;
        const ll acc = acc_fx[c];
        const double s_c = height[c] + sediment[c] + (double)acc * inv;
        ll applied = net;
        if (net > 0) {
            const double b = base[c];
            double room = iter_deposit - (double)acc * inv;
            const bool sub = s_c < b;
            if (sub) room = fmin(room, b - dep_floor - s_c);
            const ll room_fx = room > 0.0 ? (ll)floor(room * hscale) : 0;
            if (net > room_fx) {
                const ll surplus = net - room_fx;
                applied = room_fx;
                clamped += 1;
                ll keep = surplus;
                if (sub) {
                    keep = (ll)floor((double)surplus * (1.0 - writeoff));
                    if ((double)keep * inv < 1e-4) keep = 0;
                    lost_off += surplus - keep;
                }
                pend_fx[c] += keep;
                to_pend += keep;
            }
        }
        const ll a2 = acc + applied;
        acc_fx[c] = a2;
        const double live = height[c] + sediment[c] + (double)a2 * inv;
        const double sd = sediment[c] + (double)a2 * inv;
        cell4[c].x = (float)live;
        dyn[c].x = (float)(sd > 0.0 ? sd : 0.0);
        if (rte && rte[c].y < 1.5f) {
            // the route flag follows the live surface; a lake cell stays one (particle.py:1010-1012)
            rte[c].y = rte[c].x != (float)live ? 1.0f : 0.0f;
        }
    }
    clamped = wave_sum(clamped);
    to_pend = wave_sum(to_pend);
    lost_off = wave_sum(lost_off);
    if ((threadIdx.x & 63) == 0) {
        if (clamped) atomicAdd((ull*)&stats[PG_FS_CLAMPED], (ull)clamped);
        if (to_pend) atomicAdd((ull*)&stats[PG_FS_TO_PENDING_FX], (ull)to_pend);
        if (lost_off) atomicAdd((ull*)&stats[PG_FS_LOST_OFFSHORE_FX], (ull)lost_off);
    }
}

// fold (particle.py:1079-1101) + EMA (particle.py:1230-1242) + pending mirror
__global__ void finish_kernel(int64_t n, double* height, double* sediment, double* discharge, double* momentum,
                              double* pending, const uint8_t* mask, ll* acc_fx, ll* dis_fx, ll* mom_fx, const ll* pend_fx,
                              double ema, double hscale, double tscale, ll* stats) {
    const double hinv = 1.0 / hscale, tinv = 1.0 / tscale;
    ll eroded = 0, deposited = 0;
    GRID_STRIDE(c, n) {
        const ll af = acc_fx[c];
        acc_fx[c] = 0;
        const bool active = mask[c] == MASK_ACTIVE;
        if (af != 0 && active) {
            const double a = (double)af * hinv;
            if (af > 0) {
                sediment[c] += a;
                deposited += af;
            } else {
                eroded -= af;
                const double take = -a, s = sediment[c];
                if (s >= take) {
                    sediment[c] = s - take;
                } else {
                    height[c] -= take - s;
                    sediment[c] = 0.0;
                }
            }
        }
        const ll d = dis_fx[c], m0 = mom_fx[c], m1 = mom_fx[n + c];
        if (active) {
            discharge[c] = (1.0 - ema) * discharge[c] + ema * ((double)d * tinv);
            momentum[2 * c] = (1.0 - ema) * momentum[2 * c] + ema * ((double)m0 * tinv);
            momentum[2 * c + 1] = (1.0 - ema) * momentum[2 * c + 1] + ema * ((double)m1 * tinv);
        }
        dis_fx[c] = 0;
        mom_fx[c] = 0;
        mom_fx[n + c] = 0;
        pending[c] = (double)pend_fx[c] * hinv;
    }
    eroded = wave_sum(eroded);
    deposited = wave_sum(deposited);
    if ((threadIdx.x & 63) == 0) {
        if (eroded) atomicAdd((ull*)&stats[PG_FS_ERODED_FX], (ull)eroded);
        if (deposited) atomicAdd((ull*)&stats[PG_FS_DEPOSITED_FX], (ull)deposited);
    }
}

// ---------------------------------------------------------------------------
// batched windows: the per-window table and the per-chunk union offsets,
// from each window's spawn totals
// ---------------------------------------------------------------------------
struct WinStatic {
    ll coff, dbase;
    int N, NE, max_steps, pad;
    double conc_chunk;
};

__global__ void batch_plan_kernel(int W, int chunks, const WinStatic* ws, const char* sc_base, size_t sc_slot,
                                  float min_volume_frac, WinDesc* wd, ll* prefix, ll* queue, ll* stats, int nstats) {
    for (int w = threadIdx.x; w < W; w += blockDim.x) {
        const Scalars* sc = reinterpret_cast<const Scalars*>(sc_base + (size_t)w * sc_slot);
        const double volume0 = sc->volume0;
        const ll n_total = volume0 > 0.0 ? sc->n_total : 0;
        WinDesc d;
        d.coff = ws[w].coff;
        d.dbase = ws[w].dbase;
        d.n_total = n_total;
        d.C = (n_total + chunks - 1) / chunks;
        d.N = ws[w].N;
        d.NE = ws[w].NE;
        d.max_steps = ws[w].max_steps;
        d.pad = 0;
        d.inv_v0 = n_total > 0 ? (float)(1.0 / volume0) : 0.0f;
        d.min_volume = (float)(min_volume_frac * volume0);
        d.conc_scale = n_total > 0 ? (float)((ws[w].conc_chunk > 0.0 ? ws[w].conc_chunk : (double)d.C) /
                                             (volume0 * (double)n_total))
                                   : 0.0f;
        d.v0f = (float)volume0;
        d.inv_sat = sc->inv_sat;
        d.mom_scale = sc->mom_scale;
        d.pad2 = d.pad3 = 0.0f;
        wd[w] = d;
    }
    for (int k = threadIdx.x; k < 2 * chunks; k += blockDim.x) queue[k] = 0;
    for (int k = threadIdx.x; k < nstats; k += blockDim.x) stats[k] = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < chunks; c += blockDim.x) {
        ll* pre = prefix + (ll)c * (W + 1);
        ll acc = 0;
        pre[0] = 0;
        for (int w = 0; w < W; w++) {
            const ll lo = (ll)c * wd[w].C, hi0 = (ll)(c + 1) * wd[w].C;
            const ll hi = hi0 < wd[w].n_total ? hi0 : wd[w].n_total;
            acc += hi > lo ? hi - lo : 0;
            pre[w + 1] = acc;
        }
    }
}

// trace-launch timing (pg_fast_trace_time): event pairs reused across calls
struct TraceTimer {
    hipEvent_t start[4096], stop[4096];
    int pass[4096];
    int created = 0, recorded = 0;
};
TraceTimer g_timer;

int timer_record(hipStream_t s, bool is_start, int pass) {
    TraceTimer& t = g_timer;
    const int i = t.recorded;
    PG_REQUIRE(i < 4096, "trace timing: more than 4096 launches recorded without pg_fast_trace_time");
    if (i >= t.created) {
        PG_HIP(hipEventCreate(&t.start[i]));
        PG_HIP(hipEventCreate(&t.stop[i]));
        t.created = i + 1;
    }
    if (is_start) {
        PG_HIP(hipEventRecord(t.start[i], s));
        t.pass[i] = pass;
    } else {
        PG_HIP(hipEventRecord(t.stop[i], s));
        t.recorded = i + 1;
    }
    return 0;
}

int check_fields(const pg_fast_fields* fl) {
    PG_REQUIRE(fl != nullptr, "fast mode: fields is NULL");
    PG_REQUIRE(fl->F > 0 && fl->N >= 2 && fl->H >= 1, "fast mode: bad grid F=%lld N=%lld H=%lld", (long long)fl->F,
               (long long)fl->N, (long long)fl->H);
    PG_REQUIRE((fl->F * (fl->N + 2 * fl->H) * (fl->N + 2 * fl->H)) < (1ll << 31), "fast mode: grid too large for int32 cell ids");
    PG_REQUIRE(fl->height && fl->sediment && fl->discharge && fl->momentum && fl->pending && fl->base && fl->precip &&
                   fl->evap && fl->hardness && fl->metric && fl->metric_inv && fl->mask,
               "fast mode: NULL field pointer");
    return 0;
}

}  // namespace

inline size_t up(size_t x) { return (x + 255) & ~(size_t)255; }

int layout_of(int64_t F, int64_t NE, int64_t P, Layout* L) {
    const size_t n = (size_t)(F * NE * NE), p = (size_t)P, nd = p + n;
    size_t o = 0;
    auto take = [&](size_t bytes) { size_t at = o; o += up(bytes); return at; };
    L->cell4 = take(16 * n);
    L->geo = take(32 * n);
    L->dyn = take(8 * n);
    L->rte = take(8 * n);
    L->chunk_fx = take(8 * n);
    L->acc_fx = take(8 * n);
    L->dis_fx = take(8 * n);
    L->mom_fx = take(16 * n);
    L->pend_fx = take(8 * n);
    L->ero_fx = take(8 * n);
    L->fin_fx = take(8 * n);
    L->floor_key = take(8 * n);
    L->rscale = take(4 * n);
    L->r1 = take(4 * n);
    L->wq = take(8 * n);
    L->cdf_fx = take(8 * n);
    L->flag = take(4 * n);
    L->rank = take(4 * n);
    L->comp_idx = take(4 * n);
    L->comp_cdf = take(8 * n);
    L->bbuf = take(8 * n);
    L->bval = take(n);
    L->draws = take(8 * 3 * p);
    L->u_sorted = take(8 * p);
    L->d_cell = take(4 * nd);
    L->d_fx = take(4 * nd);
    L->d_fy = take(4 * nd);
    L->d_load = take(8 * nd);
    L->d_sa = take(4 * nd);
    L->d_sb = take(4 * nd);
    L->scalars = take(sizeof(Scalars) + 8 * 2 * 4096);
    size_t t1 = 0, t2 = 0, t3 = 0;
    if (hipcub::DeviceRadixSort::SortKeys(nullptr, t1, (const double*)nullptr, (double*)nullptr, (int)(p > 0 ? p : 1)) != hipSuccess)
        return -1;
    if (hipcub::DeviceScan::InclusiveSum(nullptr, t2, (const ll*)nullptr, (ll*)nullptr, (int)n) != hipSuccess) return -1;
    if (hipcub::DeviceScan::InclusiveSum(nullptr, t3, (const int*)nullptr, (int*)nullptr, (int)n) != hipSuccess) return -1;
    size_t t = t1 > t2 ? t1 : t2;
    t = t > t3 ? t : t3;
    L->temp_bytes = t;
    L->temp = take(t + 256);
    L->total = o;
    return 0;
}

int persistent_grid(const void* kernel, int threads, int64_t requested, int* grid) {
    if (requested > 0) {
        *grid = (int)requested;
        return 0;
    }
    int dev = 0, cus = 0, per = 0;
    PG_HIP(hipGetDevice(&dev));
    PG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0));
    *grid = cus * (per > 0 ? per : 1);
    return 0;
}

}  // namespace fastmode
}  // namespace pg

using namespace pg;
using namespace pg::fastmode;

namespace pg {
namespace fastmode {
// rows per wave of the census tape for a grid of `grid` blocks
// (cmax: the most droplets a chunk can hold -- the spawn plus a stockpile
// re-injected from every extended cell -- over the chunks)
int64_t tape_rows_c(int64_t cmax, int64_t max_steps, int grid) {
    const int64_t waves = (int64_t)grid * (TRACE_TPB / 64);
    const int64_t share = (cmax * (max_steps + 1) + 64 * waves - 1) / (64 * waves);
    return 2 * share + max_steps + 3;
}
int64_t tape_bytes_c(int64_t cmax, int64_t max_steps, int grid) {
    const int64_t waves = (int64_t)grid * (TRACE_TPB / 64);
    return waves * tape_rows_c(cmax, max_steps, grid) * 64 * 32 + waves * 8 + cmax * 4;
}
int64_t tape_rows(int64_t n_particles, int64_t F, int64_t NE, int64_t chunks, int64_t max_steps, int grid) {
    return tape_rows_c((n_particles + F * NE * NE + chunks - 1) / chunks, max_steps, grid);
}
// the union of every window's share of a chunk
int64_t batch_chunk_max(const pg_fast_fields* wins, const pg_fast_params* fps, int64_t n_win, int64_t chunks) {
    int64_t cmax = 0;
    for (int64_t w = 0; w < n_win; w++) {
        const int64_t NE = wins[w].N + 2 * wins[w].H;
        cmax += (fps[w].n_particles + NE * NE + chunks - 1) / chunks;
    }
    return cmax;
}
}  // namespace fastmode
}  // namespace pg

extern "C" int64_t pg_fast_tape_bytes(int64_t n_particles, int64_t F, int64_t NE, int64_t chunks, int64_t max_steps,
                                      int64_t grid_blocks, int64_t variant) {
    clear_error();
    if (n_particles < 0 || F <= 0 || NE <= 0 || chunks <= 0 || max_steps <= 0 || grid_blocks < 0 || variant < 0 ||
        variant > 2) {
        set_error("pg_fast_tape_bytes: bad arguments");
        return -1;
    }
    int grid = 0;
    if (persistent_grid(trace_fn<false, false>(variant), TRACE_TPB, grid_blocks, &grid)) return -1;
    return tape_bytes_c((n_particles + F * NE * NE + chunks - 1) / chunks, max_steps, grid);
}

extern "C" int64_t pg_batch_tape_bytes(const pg_fast_fields* wins, const pg_fast_params* fps, int64_t n_win,
                                       int64_t max_steps, int64_t grid_blocks, int64_t variant) {
    clear_error();
    if (!wins || !fps || n_win <= 0 || max_steps <= 0 || grid_blocks < 0 || variant < 0 || variant > 2 ||
        fps[0].chunks <= 0) {
        set_error("pg_batch_tape_bytes: bad arguments");
        return -1;
    }
    int grid = 0;
    if (persistent_grid(trace_fn<true, false>(variant), TRACE_TPB, grid_blocks, &grid)) return -1;
    return tape_bytes_c(batch_chunk_max(wins, fps, n_win, fps[0].chunks), max_steps, grid);
}

extern "C" int64_t pg_fast_workspace_bytes(int64_t F, int64_t NE, int64_t n_particles) {
    clear_error();
    Layout L;
    if (F <= 0 || NE <= 1 || n_particles < 0) {
        set_error("pg_fast_workspace_bytes: bad sizes");
        return -1;
    }
    if (layout_of(F, NE, n_particles, &L) != 0) {
        set_error("pg_fast_workspace_bytes: hipcub temp-size query failed");
        return -1;
    }
    return (int64_t)L.total;
}

extern "C" int pg_pcg64_random(uint64_t slo, uint64_t shi, uint64_t ilo, uint64_t ihi, int64_t start, int64_t n,
                               double* out, pg_stream stream) {
    clear_error();
    PG_REQUIRE(n >= 0 && start >= 0 && (out || n == 0), "pg_pcg64_random: bad arguments");
    if (n == 0) return 0;
    const int64_t threads = (n + 15) / 16;
    hipLaunchKernelGGL(pcg_fill_kernel, dim3(grid_for(threads, 8192)), dim3(TPB), 0, (hipStream_t)stream, slo, shi, ilo,
                       ihi, (const ull*)nullptr, (ll)start, (ll)n, out);
    PG_LAUNCHED("pcg_fill_kernel");
    return 0;
}

extern "C" int pg_fast_setup(const pg_fast_fields* fl, const pg_fast_params* fp, void* work, int64_t work_bytes,
                             pg_stream stream) {
    clear_error();
    if (check_fields(fl)) return -1;
    PG_REQUIRE(fp && work, "pg_fast_setup: NULL argument");
    const int64_t NE = fl->N + 2 * fl->H, n = fl->F * NE * NE;
    Layout L;
    PG_REQUIRE(layout_of(fl->F, NE, fp->n_particles, &L) == 0, "pg_fast_setup: hipcub temp-size query failed");
    PG_REQUIRE((int64_t)L.total <= work_bytes, "pg_fast_setup: workspace of %lld bytes, need %lld", (long long)work_bytes,
               (long long)L.total);
    PG_REQUIRE(fp->height_bits > 0 && fp->height_bits < 62, "pg_fast_setup: height_bits out of range");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(setup_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, fl->metric, fl->metric_inv, fl->evap,
                       fl->hardness, at<float4>(work, L.geo), fl->pending, at<ll>(work, L.pend_fx),
                       ldexp(1.0, (int)fp->height_bits));
    PG_LAUNCHED("setup_kernel");
    PG_HIP(hipMemsetAsync(at<char>(work, L.chunk_fx), 0, L.pend_fx - L.chunk_fx, s));
    PG_HIP(hipMemsetAsync(at<char>(work, L.ero_fx), 0, L.floor_key - L.ero_fx, s));
    PG_HIP(hipMemsetAsync(at<char>(work, L.floor_key), 0xff, L.rscale - L.floor_key, s));
    return 0;
}

namespace pg {
namespace fastmode {

int prepare(const pg_fast_fields* fl, const pg_trace_params* prm, const pg_fast_params* fp, void* work,
            int64_t work_bytes, int64_t* stats, pg_stream stream, Ctx* c) {
    if (check_fields(fl)) return -1;
    PG_REQUIRE(prm && fp && work && stats, "fast mode: NULL argument");
    PG_REQUIRE(fp->n_particles >= 1 && fp->n_particles < (1ll << 30), "fast mode: n_particles out of range");
    PG_REQUIRE(fp->chunks >= 1 && fp->chunks <= 4096, "fast mode: chunks must be in [1, 4096]");
    PG_REQUIRE(fp->conc_chunk >= 0.0 && fp->iter_erode >= 0.0, "fast mode: conc_chunk and iter_erode must be >= 0");
    PG_REQUIRE(prm->max_steps >= 0 && prm->max_steps < (1ll << 30), "fast mode: bad max_steps");
    PG_REQUIRE(prm->lake_park <= 0.0, "fast mode: lake parks are outside the fast mode");
    PG_REQUIRE(fp->height_bits > 0 && fp->height_bits < 62 && fp->track_bits > 0 && fp->track_bits < 62,
               "fast mode: fixed-point bits out of range");
    c->N = fl->N;
    c->H = fl->H;
    c->NE = fl->N + 2 * fl->H;
    c->n = fl->F * c->NE * c->NE;
    c->P = fp->n_particles;
    PG_REQUIRE(layout_of(fl->F, c->NE, c->P, &c->L) == 0, "fast mode: hipcub temp-size query failed");
    PG_REQUIRE((int64_t)c->L.total <= work_bytes, "fast mode: workspace of %lld bytes, need %lld",
               (long long)work_bytes, (long long)c->L.total);
    c->s = (hipStream_t)stream;
    c->hscale = ldexp(1.0, (int)fp->height_bits);
    c->tscale = ldexp(1.0, (int)fp->track_bits);
    c->sc = at<Scalars>(work, c->L.scalars);
    c->cell4 = at<float4>(work, c->L.cell4);
    c->dyn = at<float2>(work, c->L.dyn);
    c->flag = at<int>(work, c->L.flag);
    c->rank = at<int>(work, c->L.rank);
    c->pend_fx = at<ll>(work, c->L.pend_fx);
    c->st = (ll*)stats;
    c->temp = at<void>(work, c->L.temp);
    c->chunks = (int)fp->chunks;
    return 0;
}

int fast_spawn(const pg_fast_fields* fl, const pg_trace_params* prm, const pg_fast_params* fp, void* work, Ctx& c) {
    const int64_t N = c.N, H = c.H, n = c.n;
    const ll P = c.P;
    hipStream_t s = c.s;
    Scalars* sc = c.sc;
    float4* cell4 = c.cell4;
    float2* dyn = c.dyn;
    int *flag = c.flag, *rank = c.rank;
    ll *pend_fx = c.pend_fx, *st = c.st;
    void* temp = c.temp;
    const Layout& L = c.L;
    size_t tb = L.temp_bytes;
    const int chunks = c.chunks;
    hipLaunchKernelGGL(reset_scalars_kernel, dim3(1), dim3(256), 0, s, sc, st, (int)PG_FAST_NSTATS, chunks);
    hipLaunchKernelGGL(pack_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, N, H, fl->height, fl->sediment, fl->discharge,
                       fl->momentum, fl->base, fl->precip, fl->mask, cell4, dyn, flag, sc, fp->route, fp->lake_flag,
                       at<float2>(work, L.rte));
    hipLaunchKernelGGL(quantize_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, fl->precip, flag, sc, at<ll>(work, L.wq));
    PG_LAUNCHED("pack/quantize");
    PG_HIP(hipcub::DeviceScan::InclusiveSum(temp, tb, at<ll>(work, L.wq), at<ll>(work, L.cdf_fx), (int)n, s));
    tb = L.temp_bytes;
    PG_HIP(hipcub::DeviceScan::InclusiveSum(temp, tb, flag, rank, (int)n, s));
    hipLaunchKernelGGL(compact_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, flag, rank, at<ll>(work, L.cdf_fx), sc,
                       at<int>(work, L.comp_idx), at<double>(work, L.comp_cdf));
    hipLaunchKernelGGL(spawn_totals_kernel, dim3(1), dim3(64), 0, s, n, rank, at<ll>(work, L.cdf_fx),
                       at<int>(work, L.comp_idx), at<double>(work, L.comp_cdf), fl->precip, P, sc, st,
                       (float)(prm->disc_saturation > 0.0 ? 1.0 / prm->disc_saturation : 0.0), (float)prm->mom_scale,
                       fp->disc_saturation_cells, fp->momentum_saturation_cells, fp->inflow_volume);
    // the rain draws: P uniforms (sorted) then P offset pairs
    double* draws = at<double>(work, L.draws);
    hipLaunchKernelGGL(pcg_fill_kernel, dim3(grid_for((3 * P + 15) / 16, 8192)), dim3(TPB), 0, s, fp->rng_state_lo,
                       fp->rng_state_hi, fp->rng_inc_lo, fp->rng_inc_hi, (const ull*)fp->rng_dev, (ll)0, 3 * P, draws);
    PG_LAUNCHED("spawn kernels");
    tb = L.temp_bytes;
    PG_HIP(hipcub::DeviceRadixSort::SortKeys(temp, tb, (const double*)draws, at<double>(work, L.u_sorted), (int)P, 0, 64, s));
    int* d_cell = at<int>(work, L.d_cell);
    float *d_fx = at<float>(work, L.d_fx), *d_fy = at<float>(work, L.d_fy), *d_sa = at<float>(work, L.d_sa),
          *d_sb = at<float>(work, L.d_sb);
    ll* d_load = at<ll>(work, L.d_load);
    hipLaunchKernelGGL(rain_drops_kernel, dim3(grid_for(P)), dim3(TPB), 0, s, P, at<double>(work, L.u_sorted), draws,
                       at<int>(work, L.comp_idx), at<double>(work, L.comp_cdf), sc, N, H, d_cell, d_fx, d_fy, d_load,
                       d_sa, d_sb);
    // pending stockpiles after the rain
    hipLaunchKernelGGL(pending_flags_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, pend_fx, flag);
    tb = L.temp_bytes;
    PG_HIP(hipcub::DeviceScan::InclusiveSum(temp, tb, flag, rank, (int)n, s));
    hipLaunchKernelGGL(pending_drops_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, flag, rank, pend_fx, P,
                       fp->rng_state_lo, fp->rng_state_hi, fp->rng_inc_lo, fp->rng_inc_hi, (const ull*)fp->rng_dev, sc,
                       d_cell, d_fx, d_fy,
                       d_load, d_sa, d_sb, st);
    hipLaunchKernelGGL(totals_kernel, dim3(1), dim3(64), 0, s, n, rank, P, sc, st, chunks);
    PG_LAUNCHED("pending kernels");

    return 0;
}

int fast_particles(const pg_fast_fields* fl, const pg_trace_params* prm, const pg_fast_params* fp, void* work, Ctx& c) {
    const int64_t N = c.N, H = c.H, NE = c.NE, n = c.n;
    hipStream_t s = c.s;
    Scalars* sc = c.sc;
    float4* cell4 = c.cell4;
    float2* dyn = c.dyn;
    ll *pend_fx = c.pend_fx, *st = c.st;
    const Layout& L = c.L;
    const int chunks = c.chunks;
    const double hscale = c.hscale, tscale = c.tscale;
    int* d_cell = at<int>(work, L.d_cell);
    float *d_fx = at<float>(work, L.d_fx), *d_fy = at<float>(work, L.d_fy), *d_sa = at<float>(work, L.d_sa),
          *d_sb = at<float>(work, L.d_sb);
    ll* d_load = at<ll>(work, L.d_load);
    // chunks: persistent trace, then resolve against the live terrain
    TraceCfg k;
    k.cell4 = cell4;
    k.geo = at<float4>(work, L.geo);
    k.dyn = dyn;
    k.mask = fl->mask;
    k.N = (int)N;
    k.H = (int)H;
    k.NE = (int)NE;
    k.spherical = (int)fl->spherical;
    k.deposit_on_exit = (int)fl->deposit_on_exit;
    k.chunks = chunks;
    k.d_cell = d_cell;
    k.d_fx = d_fx;
    k.d_fy = d_fy;
    k.d_sa = d_sa;
    k.d_sb = d_sb;
    k.d_load = d_load;
    k.sc = sc;
    k.dt = (float)prm->dt;
    k.density = (float)prm->density;
    double fr = 1.0 - prm->dt * prm->friction;
    k.fr = (float)(fr < 0.0 ? 0.0 : fr);
    k.rate = (float)prm->deposition_rate;
    k.dt_evap = (float)(prm->dt * prm->evap_rate);
    k.k_mom = (float)prm->k_mom;
    k.mom_scale = (float)prm->mom_scale;
    k.k_disc = (float)prm->k_disc;
    k.inv_sat = (float)(prm->disc_saturation > 0.0 ? 1.0 / prm->disc_saturation : 0.0);
    k.disc_exponent = (float)prm->disc_exponent;
    k.slope_gain = (float)prm->slope_gain;
    k.slope_sat = (float)prm->slope_sat;
    k.erodibility = (float)prm->erodibility;
    k.cover_depth = (float)prm->cover_depth;
    k.max_erode = (float)prm->max_erode;
    k.ocean_rate = (float)prm->ocean_rate;
    k.fan_room = (float)prm->fan_room;
    k.dep_floor = (float)prm->dep_floor;
    k.min_volume_frac = (float)fp->min_volume_frac;
    k.max_steps = (int)prm->max_steps;
    k.pit_steps = (int)prm->pit_steps;
    k.ocean_steps = (int)prm->ocean_steps;
    k.hscale = hscale;
    k.tscale = tscale;
    k.chunk_fx = at<ll>(work, L.chunk_fx);
    k.ero_fx = at<ll>(work, L.ero_fx);
    k.fin_fx = at<ll>(work, L.fin_fx);
    k.floor_key = at<ull>(work, L.floor_key);
    k.fan_slope = (float)fp->fan_slope;
    k.rscale = at<float>(work, L.rscale);
    k.conc_chunk = fp->conc_chunk;
    k.dis_fx = at<ll>(work, L.dis_fx);
    k.mom_fx = at<ll>(work, L.mom_fx);
    k.mom_fx_b = k.mom_fx + n;
    k.stats = st;
    k.acc_fx = at<ll>(work, L.acc_fx);
    k.pend_fx = at<ll>(work, L.pend_fx);
    k.dep_budget = (ll)fmin(floor(fp->iter_deposit * hscale), 4.0e18);
    int grid = 0, rgrid = 0;
    PG_REQUIRE(fp->variant >= 0 && fp->variant <= 2, "fast mode: variant must be 0, 1 or 2");
    if (persistent_grid(trace_fn<false, false>(fp->variant), TRACE_TPB, fp->grid_blocks, &grid)) return -1;
    if (fp->tape && persistent_grid((const void*)replay_kernel<false>, TRACE_TPB, 0, &rgrid)) return -1;
    k.rte = fp->route ? at<float2>(work, L.rte) : nullptr;
    k.lake_trap = (float)prm->lake_trap;
    k.lateral_rate = (float)(prm->lateral_rate > 0.0 ? prm->lateral_rate : 0.0);
    k.tape = nullptr;
    k.tape_len = nullptr;
    k.tape_rows = 0;
    k.tape_last = nullptr;
    k.tape_waves = 0;
    if (fp->tape) {
        const int64_t waves = (int64_t)grid * (TRACE_TPB / 64);
        const int64_t cmax = (fp->n_particles + fl->F * NE * NE + chunks - 1) / chunks;
        const int64_t rows = tape_rows_c(cmax, prm->max_steps, grid);
        const int64_t need = tape_bytes_c(cmax, prm->max_steps, grid);
        PG_REQUIRE(fp->tape_bytes >= need, "fast mode: tape of %lld bytes, need %lld", (long long)fp->tape_bytes,
                   (long long)need);
        k.tape = reinterpret_cast<float4*>(fp->tape);
        k.tape_rows = rows;
        k.tape_len = reinterpret_cast<ll*>(k.tape + 2 * waves * rows * 64);
        k.tape_last = reinterpret_cast<int*>(k.tape_len + waves);
        k.tape_waves = waves;
    }
    // the brush (radius >= 1; radius 1 is the identity): its weights and the per-cell target flags
    const bool brushed = fp->brush_radius >= 1;
    brush::Weights BW;
    ll* bbuf = at<ll>(work, L.bbuf);
    uint8_t* bval = at<uint8_t>(work, L.bval);
    if (brushed) {
        if (brush::weights_of(fp->brush_radius, &BW)) return -1;
        PG_REQUIRE((fp->brush_owner != nullptr) == (fl->spherical != 0),
                   "fast mode: a brush needs the owner table on the globe, and none on a window");
        if (brush::enqueue_valid(n, fl->mask, fp->brush_owner, bval, s)) return -1;
    }
    for (int c = 0; c < chunks; c++) {
        k.chunk = c;
        // pass 0: erosion census on the frozen terrain (and, with a tape, every move recorded)
        k.pass = 0;
        if (fp->time_trace && timer_record(s, true, k.pass)) return -1;
        launch_trace<false, false>(fp->variant, grid, s, k);
        if (fp->time_trace && timer_record(s, false, k.pass)) return -1;
        hipLaunchKernelGGL(erosion_scale_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, fl->height, fl->sediment, fl->base,
                           k.ero_fx, k.floor_key, at<ll>(work, L.acc_fx), at<float>(work, L.r1), fp->iter_erode,
                           fp->erode_cap, hscale, k.rte);
        hipLaunchKernelGGL(erosion_scale2_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, fl->height, fl->sediment,
                           fl->base, k.ero_fx, k.floor_key, at<ll>(work, L.acc_fx), at<float>(work, L.r1),
                           at<float>(work, L.rscale), fp->iter_erode, fp->erode_cap, hscale, k.rte);
        // pass 1: the same trajectories, erosion scaled per cell, deficits cancel deposits
        k.pass = 1;
        if (fp->time_trace && timer_record(s, true, k.pass)) return -1;
        if (k.tape) hipLaunchKernelGGL(replay_kernel<false>, dim3(rgrid), dim3(TRACE_TPB), 0, s, k);
        else launch_trace<false, true>(fp->variant, grid, s, k);
        if (fp->time_trace && timer_record(s, false, k.pass)) return -1;
        // the brush: the chunk's net change spread over each cell's footprint before the resolve
        ll* net = k.chunk_fx;
        if (brushed) {
            if (fp->time_trace && timer_record(s, true, 2)) return -1;
            if (brush::enqueue_spread(fl->F, NE, BW, fp->brush_tile, k.chunk_fx, bval, fp->brush_owner, bbuf, s))
                return -1;
            if (fp->time_trace && timer_record(s, false, 2)) return -1;
            net = bbuf;
        }
        hipLaunchKernelGGL(resolve_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, fl->height, fl->sediment, fl->base,
                           net, at<ll>(work, L.acc_fx), pend_fx, cell4, dyn, fp->iter_deposit, prm->dep_floor,
                           fp->offshore_writeoff, hscale, st, k.ero_fx, k.floor_key, k.rte, k.fin_fx,
                           brushed ? k.chunk_fx : nullptr);
    }
    PG_LAUNCHED("trace/resolve");
    return finish_iteration(fl, fp, work, c);
}

int finish_iteration(const pg_fast_fields* fl, const pg_fast_params* fp, void* work, Ctx& c) {
    hipLaunchKernelGGL(finish_kernel, dim3(grid_for(c.n)), dim3(TPB), 0, c.s, c.n, fl->height, fl->sediment,
                       fl->discharge, fl->momentum, fl->pending, fl->mask, at<ll>(work, c.L.acc_fx),
                       at<ll>(work, c.L.dis_fx), at<ll>(work, c.L.mom_fx), c.pend_fx, fp->ema, c.hscale, c.tscale, c.st);
    PG_LAUNCHED("finish_kernel");
    return 0;
}
}  // namespace fastmode
}  // namespace pg

extern "C" int pg_fast_spawn(const pg_fast_fields* fl, const pg_trace_params* prm, const pg_fast_params* fp, void* work,
                             int64_t work_bytes, int64_t* stats, pg_stream stream) {
    clear_error();
    Ctx c;
    if (prepare(fl, prm, fp, work, work_bytes, stats, stream, &c)) return -1;
    return fast_spawn(fl, prm, fp, work, c);
}

extern "C" int pg_fast_particles(const pg_fast_fields* fl, const pg_trace_params* prm, const pg_fast_params* fp,
                                 void* work, int64_t work_bytes, int64_t* stats, pg_stream stream) {
    clear_error();
    Ctx c;
    if (prepare(fl, prm, fp, work, work_bytes, stats, stream, &c)) return -1;
    return fast_particles(fl, prm, fp, work, c);
}

extern "C" int pg_fast_iterate(const pg_fast_fields* fl, const pg_trace_params* prm, const pg_fast_params* fp, void* work,
                               int64_t work_bytes, int64_t* stats, pg_stream stream) {
    clear_error();
    Ctx c;
    if (prepare(fl, prm, fp, work, work_bytes, stats, stream, &c)) return -1;
    if (fast_spawn(fl, prm, fp, work, c)) return -1;
    return fast_particles(fl, prm, fp, work, c);
}

namespace {
// timed launches read back so far, per kind: census, applying pass, brush
double g_bucket_ms[3];
int64_t g_bucket_n[3];

int drain_timer() {
    for (int i = 0; i < g_timer.recorded; i++) {
        PG_HIP(hipEventSynchronize(g_timer.stop[i]));
        float t = 0.0f;
        PG_HIP(hipEventElapsedTime(&t, g_timer.start[i], g_timer.stop[i]));
        const int p = g_timer.pass[i] >= 2 ? 2 : g_timer.pass[i] ? 1 : 0;
        g_bucket_ms[p] += t;
        g_bucket_n[p] += 1;
    }
    g_timer.recorded = 0;
    return 0;
}
}  // namespace

extern "C" int pg_fast_trace_time_passes(double* ms, int64_t* launches) {
    clear_error();
    PG_REQUIRE(ms && launches, "pg_fast_trace_time_passes: NULL argument");
    if (drain_timer()) return -1;
    for (int p = 0; p < 2; p++) {
        ms[p] = g_bucket_ms[p];
        launches[p] = g_bucket_n[p];
        g_bucket_ms[p] = 0.0;
        g_bucket_n[p] = 0;
    }
    return 0;
}

extern "C" int pg_fast_brush_time(double* ms, int64_t* launches) {
    clear_error();
    PG_REQUIRE(ms && launches, "pg_fast_brush_time: NULL argument");
    if (drain_timer()) return -1;
    *ms = g_bucket_ms[2];
    *launches = g_bucket_n[2];
    g_bucket_ms[2] = 0.0;
    g_bucket_n[2] = 0;
    return 0;
}

extern "C" int pg_fast_trace_time(double* ms_total, int64_t* launches) {
    PG_REQUIRE(ms_total && launches, "pg_fast_trace_time: NULL argument");
    double ms[2];
    int64_t n[2];
    if (pg_fast_trace_time_passes(ms, n)) return -1;
    *ms_total = ms[0] + ms[1];
    *launches = n[0] + n[1];
    return 0;
}

extern "C" int pg_fast_layout(int64_t F, int64_t NE, int64_t n_particles, int64_t* offsets, int n) {
    clear_error();
    PG_REQUIRE(offsets && n >= 16, "pg_fast_layout: need room for 16 offsets");
    Layout L;
    PG_REQUIRE(layout_of(F, NE, n_particles, &L) == 0, "pg_fast_layout: hipcub temp-size query failed");
    const size_t v[16] = {L.cell4, L.geo, L.dyn, L.chunk_fx, L.acc_fx, L.dis_fx, L.mom_fx, L.pend_fx,
                          L.d_cell, L.d_fx, L.d_fy, L.d_load, L.d_sa, L.d_sb, L.scalars, L.total};
    for (int k = 0; k < 16; k++) offsets[k] = (int64_t)v[k];
    return 0;
}

// ---------------------------------------------------------------------------
// batched windows
// ---------------------------------------------------------------------------
namespace pg {
namespace fastmode {
namespace {

constexpr int MAX_CHUNKS = 4096;
constexpr size_t SC_SLOT = (sizeof(Scalars) + 8 * 2 * MAX_CHUNKS + 255) & ~(size_t)255;

// the batch arena: per-cell arrays over all windows' cells (window order),
// per-droplet arrays over all windows' droplet slots, one scalar slot per
// window, shared sort/scan scratch, the window tables
struct BLayout {
    Layout base;  // arena offsets (per-cell, per-droplet, scalars = first slot, temp)
    size_t wstatic, wdesc, prefix, queue, total;
    int64_t ncells;
};

int batch_layout(int64_t W, const int64_t* NE, const int64_t* P, BLayout* B) {
    int64_t n = 0, p = 0, nd = 0;
    size_t tmax = 0;
    for (int64_t w = 0; w < W; w++) {
        Layout Lw;
        if (layout_of(1, NE[w], P[w], &Lw) != 0) return -1;
        tmax = Lw.temp_bytes > tmax ? Lw.temp_bytes : tmax;
        n += NE[w] * NE[w];
        p += P[w];
        nd += P[w] + NE[w] * NE[w];
    }
    Layout& L = B->base;
    size_t o = 0;
    auto up256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    auto take = [&](size_t bytes) { size_t a = o; o += up256(bytes); return a; };
    L.cell4 = take(16 * n);
    L.geo = take(32 * n);
    L.dyn = take(8 * n);
    L.rte = take(8 * n);
    L.chunk_fx = take(8 * n);
    L.acc_fx = take(8 * n);
    L.dis_fx = take(8 * n);
    L.mom_fx = take(16 * n);
    L.pend_fx = take(8 * n);
    L.ero_fx = take(8 * n);
    L.fin_fx = take(8 * n);
    L.floor_key = take(8 * n);
    L.rscale = take(4 * n);
    L.r1 = take(4 * n);
    L.wq = take(8 * n);
    L.cdf_fx = take(8 * n);
    L.flag = take(4 * n);
    L.rank = take(4 * n);
    L.comp_idx = take(4 * n);
    L.comp_cdf = take(8 * n);
    L.bbuf = take(8 * n);
    L.bval = take(n);
    L.draws = take(8 * 3 * p);
    L.u_sorted = take(8 * p);
    L.d_cell = take(4 * nd);
    L.d_fx = take(4 * nd);
    L.d_fy = take(4 * nd);
    L.d_load = take(8 * nd);
    L.d_sa = take(4 * nd);
    L.d_sb = take(4 * nd);
    L.scalars = take(SC_SLOT * (size_t)W);
    L.temp_bytes = tmax;
    L.temp = take(tmax + 256);
    B->wstatic = take(sizeof(WinStatic) * (size_t)W);
    B->wdesc = take(sizeof(WinDesc) * (size_t)W);
    B->prefix = take(8 * (size_t)MAX_CHUNKS * (size_t)(W + 1));
    B->queue = take(8 * 2 * (size_t)MAX_CHUNKS);
    L.total = B->total = o;
    B->ncells = n;
    return 0;
}

// window w's view of the arena, as the single-grid code addresses it
void window_layout(const BLayout& B, int64_t coff, int64_t poff, int64_t doff, int64_t w, Layout* Lw) {
    const Layout& L = B.base;
    *Lw = L;
    const size_t c = (size_t)coff;
    Lw->cell4 = L.cell4 + 16 * c;
    Lw->geo = L.geo + 32 * c;
    Lw->dyn = L.dyn + 8 * c;
    Lw->rte = L.rte + 8 * c;
    Lw->chunk_fx = L.chunk_fx + 8 * c;
    Lw->acc_fx = L.acc_fx + 8 * c;
    Lw->dis_fx = L.dis_fx + 8 * c;
    Lw->mom_fx = L.mom_fx + 16 * c;
    Lw->pend_fx = L.pend_fx + 8 * c;
    Lw->ero_fx = L.ero_fx + 8 * c;
    Lw->fin_fx = L.fin_fx + 8 * c;
    Lw->floor_key = L.floor_key + 8 * c;
    Lw->rscale = L.rscale + 4 * c;
    Lw->r1 = L.r1 + 4 * c;
    Lw->wq = L.wq + 8 * c;
    Lw->cdf_fx = L.cdf_fx + 8 * c;
    Lw->flag = L.flag + 4 * c;
    Lw->rank = L.rank + 4 * c;
    Lw->comp_idx = L.comp_idx + 4 * c;
    Lw->comp_cdf = L.comp_cdf + 8 * c;
    Lw->bbuf = L.bbuf + 8 * c;
    Lw->bval = L.bval + c;
    Lw->draws = L.draws + 8 * 3 * (size_t)poff;
    Lw->u_sorted = L.u_sorted + 8 * (size_t)poff;
    Lw->d_cell = L.d_cell + 4 * (size_t)doff;
    Lw->d_fx = L.d_fx + 4 * (size_t)doff;
    Lw->d_fy = L.d_fy + 4 * (size_t)doff;
    Lw->d_load = L.d_load + 8 * (size_t)doff;
    Lw->d_sa = L.d_sa + 4 * (size_t)doff;
    Lw->d_sb = L.d_sb + 4 * (size_t)doff;
    Lw->scalars = L.scalars + SC_SLOT * (size_t)w;
}

// the windows' fields must lie back to back in window order (one arena per
// field); returns the per-window cell offsets
int check_windows(const pg_fast_fields* wins, int64_t W, std::vector<int64_t>& coff) {
    PG_REQUIRE(wins && W >= 1 && W <= 65536, "batch: need 1..65536 windows");
    coff.assign((size_t)W + 1, 0);
    for (int64_t w = 0; w < W; w++) {
        const pg_fast_fields& f = wins[w];
        if (check_fields(&f)) return -1;
        PG_REQUIRE(f.F == 1 && !f.spherical, "batch: window %lld is not a single-face window", (long long)w);
        PG_REQUIRE(f.H == wins[0].H && f.deposit_on_exit == wins[0].deposit_on_exit,
                   "batch: windows differ in halo width or deposit_on_exit");
        const int64_t NE = f.N + 2 * f.H;
        coff[w + 1] = coff[w] + NE * NE;
        const int64_t o = coff[w];
        PG_REQUIRE(f.height == wins[0].height + o && f.sediment == wins[0].sediment + o &&
                       f.discharge == wins[0].discharge + o && f.momentum == wins[0].momentum + 2 * o &&
                       f.pending == wins[0].pending + o && f.base == wins[0].base + o &&
                       f.precip == wins[0].precip + o && f.evap == wins[0].evap + o &&
                       f.hardness == wins[0].hardness + o && f.metric == wins[0].metric + 3 * o &&
                       f.metric_inv == wins[0].metric_inv + 3 * o && f.mask == wins[0].mask + o,
                   "batch: window %lld's fields do not follow window %lld's in one arena", (long long)w,
                   (long long)(w - 1));
    }
    PG_REQUIRE(coff[W] < (1ll << 31), "batch: arena too large for int32 cell ids");
    return 0;
}

}  // namespace
}  // namespace fastmode
}  // namespace pg

extern "C" int64_t pg_batch_workspace_bytes(int64_t n_win, const int64_t* NE, const int64_t* n_particles) {
    clear_error();
    if (n_win < 1 || !NE || !n_particles) {
        set_error("pg_batch_workspace_bytes: bad arguments");
        return -1;
    }
    BLayout B;
    if (batch_layout(n_win, NE, n_particles, &B) != 0) {
        set_error("pg_batch_workspace_bytes: hipcub temp-size query failed");
        return -1;
    }
    return (int64_t)B.total;
}

extern "C" int pg_batch_setup(const pg_fast_fields* wins, int64_t n_win, const pg_fast_params* fps,
                              const int64_t* max_steps, void* work, int64_t work_bytes, pg_stream stream) {
    clear_error();
    std::vector<int64_t> coff;
    if (check_windows(wins, n_win, coff)) return -1;
    PG_REQUIRE(fps && max_steps && work, "pg_batch_setup: NULL argument");
    std::vector<int64_t> NE(n_win), P(n_win);
    for (int64_t w = 0; w < n_win; w++) {
        NE[w] = wins[w].N + 2 * wins[w].H;
        P[w] = fps[w].n_particles;
        PG_REQUIRE(P[w] >= 1 && max_steps[w] >= 0, "pg_batch_setup: window %lld: bad droplet count or lifetime",
                   (long long)w);
    }
    BLayout B;
    PG_REQUIRE(batch_layout(n_win, NE.data(), P.data(), &B) == 0, "pg_batch_setup: hipcub temp-size query failed");
    PG_REQUIRE((int64_t)B.total <= work_bytes, "pg_batch_setup: workspace of %lld bytes, need %lld",
               (long long)work_bytes, (long long)B.total);
    PG_REQUIRE(fps[0].height_bits > 0 && fps[0].height_bits < 62, "pg_batch_setup: height_bits out of range");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = coff[n_win];
    const Layout& L = B.base;
    hipLaunchKernelGGL(setup_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, wins[0].metric, wins[0].metric_inv,
                       wins[0].evap, wins[0].hardness, at<float4>(work, L.geo), wins[0].pending, at<ll>(work, L.pend_fx),
                       ldexp(1.0, (int)fps[0].height_bits));
    PG_LAUNCHED("setup_kernel");
    PG_HIP(hipMemsetAsync(at<char>(work, L.chunk_fx), 0, L.pend_fx - L.chunk_fx, s));
    PG_HIP(hipMemsetAsync(at<char>(work, L.ero_fx), 0, L.floor_key - L.ero_fx, s));
    PG_HIP(hipMemsetAsync(at<char>(work, L.floor_key), 0xff, L.rscale - L.floor_key, s));
    std::vector<WinStatic> ws(n_win);
    int64_t doff = 0;
    for (int64_t w = 0; w < n_win; w++) {
        ws[w].coff = coff[w];
        ws[w].dbase = doff;
        ws[w].N = (int)wins[w].N;
        ws[w].NE = (int)NE[w];
        ws[w].max_steps = (int)max_steps[w];
        ws[w].pad = 0;
        ws[w].conc_chunk = fps[w].conc_chunk;
        doff += P[w] + NE[w] * NE[w];
    }
    PG_HIP(hipMemcpyAsync(at<char>(work, B.wstatic), ws.data(), sizeof(WinStatic) * (size_t)n_win,
                          hipMemcpyHostToDevice, s));
    PG_HIP(hipStreamSynchronize(s));
    return 0;
}

extern "C" int pg_batch_iterate(const pg_fast_fields* wins, int64_t n_win, const pg_trace_params* prm,
                                const pg_fast_params* fps, void* work, int64_t work_bytes, int64_t* stats,
                                int64_t* win_stats, pg_stream stream) {
    clear_error();
    std::vector<int64_t> coff;
    if (check_windows(wins, n_win, coff)) return -1;
    PG_REQUIRE(prm && fps && work && stats && win_stats, "pg_batch_iterate: NULL argument");
    const pg_fast_params& f0 = fps[0];
    PG_REQUIRE(f0.chunks >= 1 && f0.chunks <= MAX_CHUNKS, "pg_batch_iterate: chunks must be in [1, %d]", MAX_CHUNKS);
    PG_REQUIRE(prm->lake_park <= 0.0, "pg_batch_iterate: lake parks are outside the fast mode");
    std::vector<int64_t> NE(n_win), P(n_win);
    for (int64_t w = 0; w < n_win; w++) {
        NE[w] = wins[w].N + 2 * wins[w].H;
        P[w] = fps[w].n_particles;
        PG_REQUIRE(fps[w].chunks == f0.chunks, "pg_batch_iterate: windows differ in chunks");
        PG_REQUIRE(P[w] >= 1 && P[w] < (1ll << 30), "pg_batch_iterate: window %lld: n_particles out of range",
                   (long long)w);
    }
    BLayout B;
    PG_REQUIRE(batch_layout(n_win, NE.data(), P.data(), &B) == 0, "pg_batch_iterate: hipcub temp-size query failed");
    PG_REQUIRE((int64_t)B.total <= work_bytes, "pg_batch_iterate: workspace of %lld bytes, need %lld",
               (long long)work_bytes, (long long)B.total);
    hipStream_t s = (hipStream_t)stream;
    const double hscale = ldexp(1.0, (int)f0.height_bits), tscale = ldexp(1.0, (int)f0.track_bits);
    const int chunks = (int)f0.chunks;
    const int64_t n = coff[n_win];
    const Layout& L = B.base;
    // per-window spawn into the window's view of the arena
    int64_t poff = 0, doff = 0;
    for (int64_t w = 0; w < n_win; w++) {
        Ctx c;
        window_layout(B, coff[w], poff, doff, w, &c.L);
        c.s = s;
        c.hscale = hscale;
        c.tscale = tscale;
        c.sc = at<Scalars>(work, c.L.scalars);
        c.cell4 = at<float4>(work, c.L.cell4);
        c.dyn = at<float2>(work, c.L.dyn);
        c.flag = at<int>(work, c.L.flag);
        c.rank = at<int>(work, c.L.rank);
        c.pend_fx = at<ll>(work, c.L.pend_fx);
        c.st = (ll*)win_stats + w * PG_FAST_NSTATS;
        c.temp = at<void>(work, c.L.temp);
        c.chunks = chunks;
        c.N = wins[w].N;
        c.H = wins[w].H;
        c.NE = NE[w];
        c.n = NE[w] * NE[w];
        c.P = P[w];
        if (fast_spawn(&wins[w], prm, &fps[w], work, c)) return -1;
        poff += P[w];
        doff += P[w] + NE[w] * NE[w];
    }
    ll* qbase = at<ll>(work, B.queue);
    ll* prefix = at<ll>(work, B.prefix);
    WinDesc* wd = at<WinDesc>(work, B.wdesc);
    hipLaunchKernelGGL(batch_plan_kernel, dim3(1), dim3(256), 0, s, (int)n_win, chunks, at<WinStatic>(work, B.wstatic),
                       at<char>(work, L.scalars), SC_SLOT, (float)f0.min_volume_frac, wd, prefix, qbase, (ll*)stats,
                       (int)PG_FAST_NSTATS);
    PG_LAUNCHED("batch_plan_kernel");
    // chunks: batched persistent trace over every window's share, per-cell passes over the arena
    TraceCfg k = {};
    k.cell4 = at<float4>(work, L.cell4);
    k.geo = at<float4>(work, L.geo);
    k.dyn = at<float2>(work, L.dyn);
    k.mask = wins[0].mask;
    k.rscale = at<float>(work, L.rscale);
    k.H = (int)wins[0].H;
    k.spherical = 0;
    k.deposit_on_exit = (int)wins[0].deposit_on_exit;
    k.chunks = chunks;
    k.d_cell = at<int>(work, L.d_cell);
    k.d_fx = at<float>(work, L.d_fx);
    k.d_fy = at<float>(work, L.d_fy);
    k.d_sa = at<float>(work, L.d_sa);
    k.d_sb = at<float>(work, L.d_sb);
    k.d_load = at<ll>(work, L.d_load);
    k.dt = (float)prm->dt;
    k.density = (float)prm->density;
    const double fr = 1.0 - prm->dt * prm->friction;
    k.fr = (float)(fr < 0.0 ? 0.0 : fr);
    k.rate = (float)prm->deposition_rate;
    k.dt_evap = (float)(prm->dt * prm->evap_rate);
    k.k_mom = (float)prm->k_mom;
    k.mom_scale = (float)prm->mom_scale;
    k.k_disc = (float)prm->k_disc;
    k.inv_sat = (float)(prm->disc_saturation > 0.0 ? 1.0 / prm->disc_saturation : 0.0);
    k.disc_exponent = (float)prm->disc_exponent;
    k.slope_gain = (float)prm->slope_gain;
    k.slope_sat = (float)prm->slope_sat;
    k.erodibility = (float)prm->erodibility;
    k.cover_depth = (float)prm->cover_depth;
    k.max_erode = (float)prm->max_erode;
    k.ocean_rate = (float)prm->ocean_rate;
    k.fan_room = (float)prm->fan_room;
    k.dep_floor = (float)prm->dep_floor;
    k.min_volume_frac = (float)f0.min_volume_frac;
    k.pit_steps = (int)prm->pit_steps;
    k.ocean_steps = (int)prm->ocean_steps;
    k.hscale = hscale;
    k.tscale = tscale;
    k.chunk_fx = at<ll>(work, L.chunk_fx);
    k.ero_fx = at<ll>(work, L.ero_fx);
    k.fin_fx = at<ll>(work, L.fin_fx);
    k.floor_key = at<ull>(work, L.floor_key);
    k.fan_slope = (float)f0.fan_slope;
    k.dis_fx = at<ll>(work, L.dis_fx);
    k.mom_fx = at<ll>(work, L.mom_fx);
    k.mom_fx_b = k.mom_fx + n;
    k.stats = (ll*)stats;
    k.acc_fx = at<ll>(work, L.acc_fx);
    k.pend_fx = at<ll>(work, L.pend_fx);
    k.dep_budget = (ll)fmin(floor(f0.iter_deposit * hscale), 4.0e18);
    k.wins = wd;
    k.win_prefix = prefix;
    k.n_win = (int)n_win;
    k.queue_base = qbase;
    int grid = 0, rgrid = 0;
    PG_REQUIRE(f0.variant >= 0 && f0.variant <= 2, "pg_batch_iterate: variant must be 0, 1 or 2");
    if (persistent_grid(trace_fn<true, false>(f0.variant), TRACE_TPB, f0.grid_blocks, &grid)) return -1;
    // the routing surface: every window's or none
    for (int64_t w = 0; w < n_win; w++)
        PG_REQUIRE((fps[w].route != nullptr) == (f0.route != nullptr),
                   "pg_batch_iterate: either every window has a routing surface or none");
    k.rte = f0.route ? at<float2>(work, L.rte) : nullptr;
    k.lake_trap = (float)prm->lake_trap;
    k.lateral_rate = (float)(prm->lateral_rate > 0.0 ? prm->lateral_rate : 0.0);
    // the move tape: prm->max_steps bounds every window's lifetime
    k.max_steps = (int)prm->max_steps;
    if (f0.tape) {
        PG_REQUIRE(prm->max_steps > 0, "pg_batch_iterate: a tape needs the windows' largest max_steps in prm");
        if (persistent_grid((const void*)replay_kernel<true>, TRACE_TPB, 0, &rgrid)) return -1;
        const int64_t cmax = batch_chunk_max(wins, fps, n_win, chunks);
        const int64_t waves = (int64_t)grid * (TRACE_TPB / 64);
        const int64_t rows = tape_rows_c(cmax, prm->max_steps, grid);
        const int64_t need = tape_bytes_c(cmax, prm->max_steps, grid);
        PG_REQUIRE(f0.tape_bytes >= need, "pg_batch_iterate: tape of %lld bytes, need %lld", (long long)f0.tape_bytes,
                   (long long)need);
        k.tape = reinterpret_cast<float4*>(f0.tape);
        k.tape_rows = rows;
        k.tape_len = reinterpret_cast<ll*>(k.tape + 2 * waves * rows * 64);
        k.tape_last = reinterpret_cast<int*>(k.tape_len + waves);
        k.tape_waves = waves;
    }
    const pg_fast_fields& a = wins[0];  // arena bases
    // the brush: every window's radius and tile are the first's
    const bool brushed = f0.brush_radius >= 1;
    brush::Weights BW;
    ll* bbuf = at<ll>(work, L.bbuf);
    uint8_t* bval = at<uint8_t>(work, L.bval);
    for (int64_t w = 0; w < n_win; w++)
        PG_REQUIRE(fps[w].brush_radius == f0.brush_radius && fps[w].brush_owner == nullptr,
                   "pg_batch_iterate: windows differ in brush radius, or have an owner table");
    if (brushed) {
        if (brush::weights_of(f0.brush_radius, &BW)) return -1;
        if (brush::enqueue_valid(n, a.mask, nullptr, bval, s)) return -1;
    }
    for (int c = 0; c < chunks; c++) {
        k.chunk = c;
        k.pass = 0;
        if (f0.time_trace && timer_record(s, true, k.pass)) return -1;
        launch_trace<true, false>(f0.variant, grid, s, k);
        if (f0.time_trace && timer_record(s, false, k.pass)) return -1;
        hipLaunchKernelGGL(erosion_scale_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, a.height, a.sediment, a.base,
                           k.ero_fx, k.floor_key, at<ll>(work, L.acc_fx), at<float>(work, L.r1), f0.iter_erode,
                           f0.erode_cap, hscale, k.rte);
        hipLaunchKernelGGL(erosion_scale2_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, a.height, a.sediment, a.base,
                           k.ero_fx, k.floor_key, at<ll>(work, L.acc_fx), at<float>(work, L.r1),
                           at<float>(work, L.rscale), f0.iter_erode, f0.erode_cap, hscale, k.rte);
        k.pass = 1;
        if (f0.time_trace && timer_record(s, true, k.pass)) return -1;
        if (k.tape) hipLaunchKernelGGL(replay_kernel<true>, dim3(rgrid), dim3(TRACE_TPB), 0, s, k);
        else launch_trace<true, true>(f0.variant, grid, s, k);
        if (f0.time_trace && timer_record(s, false, k.pass)) return -1;
        ll* net = k.chunk_fx;
        if (brushed) {
            // each window's own footprints: a window is a single face with no owner table
            if (f0.time_trace && timer_record(s, true, 2)) return -1;
            for (int64_t w = 0; w < n_win; w++)
                if (brush::enqueue_spread(1, NE[w], BW, f0.brush_tile, k.chunk_fx + coff[w], bval + coff[w], nullptr,
                                          bbuf + coff[w], s))
                    return -1;
            if (f0.time_trace && timer_record(s, false, 2)) return -1;
            net = bbuf;
        }
        hipLaunchKernelGGL(resolve_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, a.height, a.sediment, a.base,
                           net, at<ll>(work, L.acc_fx), at<ll>(work, L.pend_fx), at<float4>(work, L.cell4),
                           at<float2>(work, L.dyn), f0.iter_deposit, prm->dep_floor, f0.offshore_writeoff, hscale,
                           (ll*)stats, k.ero_fx, k.floor_key, k.rte, k.fin_fx, brushed ? k.chunk_fx : nullptr);
    }
    PG_LAUNCHED("batch trace/resolve");
    hipLaunchKernelGGL(finish_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, n, a.height, a.sediment, a.discharge,
                       a.momentum, a.pending, a.mask, at<ll>(work, L.acc_fx), k.dis_fx, k.mom_fx, at<ll>(work, L.pend_fx),
                       f0.ema, hscale, tscale, (ll*)stats);
    PG_LAUNCHED("finish_kernel");
    return 0;
}
