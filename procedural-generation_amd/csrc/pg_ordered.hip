// The serial change-list apply (apply_changes, bake/globe/erosion/particle.py:
// 821-1042) computed in parallel, bit for bit.
//
// The reference applies a chunk's lists one droplet after another, entry
// after entry, against the live terrain.  Every entry's effect is a function
// of three things: the live `acc` of its own cell just before it, the live
// `acc` of one neighbour just before it (the next path cell for an erosion
// floor, the previous / next path cell for a deposit ceiling), and the
// droplet's carries (surplus, deficit) from its earlier entries.  Call the
// values an entry reads R and the amount it adds to its cell W.  Given all W,
// every R is a prefix: acc0 + the earlier W of that cell, added in serial
// order.  Given all R, every W (and every carry) follows by walking each
// droplet's list.  The serial result is the one solution of that system: it
// is triangular in serial order, so it has exactly one, and alternating the
// two halves reaches it in finitely many sweeps (at most the depth of the
// dependency chain; 10-70 on the earth globe) -- once a sweep changes no W,
// the values are that solution, bit for bit.
//
// One chunk (all launches on the caller's stream, no host synchronisation):
//  1. build (one wave per droplet): the per-entry constants in the change
//     list's own droplet-major layout, the first R (the chunk's starting
//     acc), and each entry's accesses -- (cell, kind) pairs in serial order:
//     self read, neighbour read, write; then the read of the death cell;
//  2. a stable radix sort of the accesses by cell: each touched cell's
//     accesses form one segment in serial order;
//  3. sweeps, each two launches (a launch boundary is the cheapest grid-wide
//     barrier on this part: ~2 us against ~17 us for a cooperative grid
//     sync of 128 blocks):
//     - the droplet half, one wave per droplet: every lane evaluates its
//       entry's limits from R in parallel, then the wave walks the entries
//       that touch the carries (deposits, clamped erosions) in order with
//       the carries held uniform -- the only serial arithmetic of a droplet;
//       changed amounts are written both per entry and at their sorted slot;
//     - the cell half, one wave per range of 64 sorted slots: the running
//       value of each segment over its writes in serial order (the same
//       float64 additions as the reference, one lane after the other), each
//       read's value then taken in parallel from the last write before it;
//     a droplet half that changes nothing marks the chunk done and every
//     later launch of it returns at once.  The first `budget` sweeps are
//     enqueued as plain launches; a chunk that needs more finishes in one
//     cooperative launch with grid barriers;
//  4. finish (one wave per range): per segment the final acc, the discharge
//     and momentum tracks and the stockpiles in serial order (float64 sums
//     in the reference's order), the packed record of every written cell;
//  5. the chunk's 4-tuple, summed in droplet order.
//
// Lake parks (volume -3 entries, emitted only with lake_park > 0) are not
// handled here, and a chunk may need more than max_sweeps sweeps (a cell that
// one chunk's droplets enter tens of thousands of times, as in a pit of a
// clamp-dominated window, chains that many amounts).  Standalone, the status
// word then reports it and nothing is applied; in the chunk loop the chunk
// goes to the serial apply (one lane, the reference's walk) instead.
//
// Built with -ffp-contract=off: every expression rounds where the
// reference's does (its operand order is kept below).
#include <hip/hip_cooperative_groups.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <mutex>
#include <vector>

#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"

namespace cg = cooperative_groups;

namespace pg {
namespace {

using ull = unsigned long long;

// The accesses' sort: an LSD radix sort (stable, as the segments' serial order
// needs) by onesweep at every size.  rocprim's default takes a block sort plus
// merge passes below 2^20 items -- 21 launches for a chunk's ~1 M slots --
// where onesweep needs a histogram, its scan and one pass per 8 key bits.
using AccessSortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                    rocprim::default_config, 0>;

constexpr int WAVE = 64;
constexpr int BLOCK = 256;                  // 4 waves
constexpr int WAVES_PER_BLOCK = BLOCK / WAVE;
// sorted slots per cell-half wave: one batch.  A wave walks the segments that
// start in its range to their ends, so a narrower range puts fewer segments
// on one wave's serial path (measured at N_c=1024: 64 slots 9 % faster per
// iteration than 256, 32 and 128 in between)
constexpr int RANGE = 64;
constexpr uint8_t EF_ACTIVE = 1, EF_ERODE = 2, EF_STEP = 4, EF_NB = 8, EF_RFLOOR = 16, EF_TRACK = 32;
constexpr uint32_t K_SELF = 0, K_NB = 1, K_WRITE = 2, K_END = 3;
constexpr uint32_t KIND_SHIFT = 30, IDX_MASK = (1u << KIND_SHIFT) - 1u;
constexpr uint32_t NONE = 0xffffffffu;
// control words: done, sweeps taken, the serial fallback's gate, then one change count per sweep
constexpr int CTL_DONE = 0, CTL_SWEEPS = 1, CTL_FALLBACK = 2, CTL_CHANGED = 3;

struct Layout {
    size_t flags, d, hs, hsn, ec, b, r, rs, rn, w, posw, cd, cs, cn;   // per entry (P * cap)
    size_t death, hsd, bd, re, drop, pend, pendf, dend, fc;          // per droplet
    size_t dflag;                                                    // per entry, padded rows
    size_t keys_in, vals_in, keys_out, vals_out, sv;                 // per access slot
    size_t ownerw;                                                   // per entry
    size_t rmin;                                                     // per range
    size_t ctl, temp;
    size_t temp_bytes, total;
};

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int layout_of(int64_t P, int64_t cap, int max_sweeps, Layout* L) {
    const size_t ne = (size_t)P * (size_t)cap, na = (size_t)P * (size_t)(3 * cap + 1);
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t r = o;
        o = align_up(o + (bytes ? bytes : 1));
        return r;
    };
    L->flags = take(ne);
    L->d = take(ne * 8);
    L->hs = take(ne * 8);
    L->hsn = take(ne * 8);
    L->ec = take(ne * 8);
    L->b = take(ne * 4);
    L->r = take(ne * 4);
    L->rs = take(ne * 8);
    L->rn = take(ne * 8);
    L->w = take(ne * 8);
    L->posw = take(ne * 4);
    L->cd = take(ne * 8);
    L->cs = take(ne * 8);
    L->cn = take(ne * 4);
    L->death = take((size_t)P * 4);
    L->hsd = take((size_t)P * 8);
    L->bd = take((size_t)P * 4);
    L->re = take((size_t)P * 8);
    L->drop = take((size_t)P * 4 * 8);
    L->pend = take((size_t)P * 8);
    L->pendf = take((size_t)P);
    L->dend = take((size_t)P);
    L->dflag = take((size_t)P * (size_t)((cap + 3) & ~(int64_t)3));
    L->fc = take((size_t)P * 24);
    L->keys_in = take(na * 4);
    L->vals_in = take(na * 4);
    L->keys_out = take(na * 4);
    L->vals_out = take(na * 4);
    L->sv = take(na * 8);
    L->ownerw = take(ne * 4);
    L->rmin = take(((na + RANGE - 1) / RANGE + 1) * 4);
    L->ctl = take((size_t)(CTL_CHANGED + max_sweeps + 1) * 4);
    size_t tb = 0;
    if (rocprim::radix_sort_pairs<AccessSortConfig>(nullptr, tb, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                    (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                    (unsigned int)(na > 0 ? na : 1), 0u, 32u) != hipSuccess)
        return -1;
    L->temp_bytes = tb;
    L->temp = take(tb);
    L->total = o;
    return 0;
}

template <class T>
__host__ __device__ inline T* at(void* base, size_t off) {
    return (T*)((char*)base + off);
}

struct OrdArgs {
    // the change list (reference layout: droplet p's entries at p * cap + k)
    const int32_t* cl_cell;
    const double* cl_delta;
    const float* cl_vol;
    const float* cl_mom;
    const int64_t* cl_count;
    int64_t P, cap;
    // state
    const double *h, *s;
    double *acc, *pending;
    float* samp;
    double *disch, *mom;
    const uint8_t* mask;
    const double* ecap;
    int use_ecap, use_route;
    double iter_erode, iter_deposit, fan_slope, dep_floor, writeoff;
    // workspace
    uint8_t* flags;
    double *d, *hs, *hsn, *ec;
    float *b, *r;
    double *rs, *rn, *w;
    uint32_t* posw;
    double *cd, *cs;   // carries (deficit, surplus) entering each entry
    int32_t* cn;       // clamped entries before each entry
    int32_t* death;
    double* hsd;
    float* bd;
    double *re, *drop, *pend;
    uint8_t* pendf;
    // per droplet p, a row of capw = cap rounded up to 4 bytes: byte k is set
    // when a read of entry k changed value (the droplet walks again from its
    // first set byte); dend[p] when its end read did.  Plain byte stores, no
    // atomics: the cell half that sets them and the droplet half that reads
    // and clears them are separate launches.
    uint8_t* dflag;
    uint8_t* dend;
    int64_t capw;
    double* fc;        // per droplet: final carries (deficit, surplus, clamped count)
    uint32_t *keys_in, *vals_in, *keys_out, *vals_out;
    double* sv;  // per sorted slot: a write's amount W, a read's value R (coalesced copies)
    uint32_t* ownerw;  // per entry: the range whose wave walks the segment of its write
    uint32_t* rmin;    // per range: the first slot of a write that changed (NONE: clean)
    uint32_t* ctl;
    int32_t* status;  // [0] status bits (1 lake park, 2 no convergence), [1] most sweeps, [2] sweeps summed,
                      // [3] chunks applied by the serial fallback
    int fallback;     // lake parks and non-convergence hand the chunk to the serial apply instead of failing
    uint32_t hole;
    int64_t n_slots, n_ranges;
    int max_sweeps;
    double* out4;
    unsigned long long* moves;  // steps traced (entries of volume >= 0) added here, if not NULL
};

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }

// lane j's value of x (j uniform across the wave)
__device__ __forceinline__ double rl(double x, int j) {
    const long long bits = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)bits, j);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), j);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ int rli(int x, int j) { return __builtin_amdgcn_readlane(x, j); }

__device__ __forceinline__ bool same_bits(double x, double y) {
    return __double_as_longlong(x) == __double_as_longlong(y);
}

// highest set bit of m below bit `lane` (-1 if none)
__device__ __forceinline__ int last_below(ull m, int lane) {
    const ull x = lane == 0 ? 0ull : (m & ((1ull << lane) - 1ull));
    return x ? 63 - __clzll(x) : -1;
}
// highest set bit of m at or below bit `lane` (-1 if none)
__device__ __forceinline__ int last_upto(ull m, int lane) {
    const ull x = lane == 63 ? m : (m & ((2ull << lane) - 1ull));
    return x ? 63 - __clzll(x) : -1;
}

// ---------------------------------------------------------------------------
// 1. build: one wave per droplet
// ---------------------------------------------------------------------------
// 0. the chunk's control words start at zero (their own launch: build may set them)
__global__ void __launch_bounds__(BLOCK) ord_reset_kernel(OrdArgs a) {
    for (int i = threadIdx.x; i < CTL_CHANGED + a.max_sweeps + 1; i += BLOCK) a.ctl[i] = 0;
}

__global__ void __launch_bounds__(BLOCK) ord_build_kernel(OrdArgs a) {
    for (int64_t r = (int64_t)blockIdx.x * BLOCK + threadIdx.x; r < a.n_ranges; r += (int64_t)gridDim.x * BLOCK) a.rmin[r] = NONE;
    const int lane = lane_id();
    const int64_t p = (int64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x / WAVE);
    if (p >= a.P) return;
    const int64_t cap = a.cap, base = p * cap, cnt = a.cl_count[p], sbase = p * (3 * cap + 1);
    int prev = -1, dth = -1;  // uniform
    bool bad = false;
    int n_moves = 0;  // uniform
    for (int64_t k0 = 0; k0 < cap; k0 += WAVE) {
        const int64_t k = k0 + lane;
        const bool in = k < cap, live = k < cnt;
        const int c = live ? a.cl_cell[base + k] : 0;
        const float v = live ? a.cl_vol[base + k] : 0.0f;
        const double d = live ? a.cl_delta[base + k] : 0.0;
        const int cnext = (live && k + 1 < cnt) ? a.cl_cell[base + k + 1] : -1;
        const bool lake = live && v == -3.0f;
        n_moves += __popcll(__ballot(live && v >= 0.0f));
        bad |= lake;
        const bool act = live && !lake && a.mask[c] == MASK_ACTIVE;
        const bool ero = act && d < 0.0;
        const bool step = act && v >= 0.0f;
        // the previous step cell: the last active step entry before this one
        const ull sm = __ballot(step);
        const int src = last_below(sm, lane);
        const int csrc = __shfl(c, src < 0 ? lane : src);
        const int prev_here = src < 0 ? prev : csrc;
        int nb = -1;
        if (act) nb = ero ? cnext : (step ? prev_here : cnext);
        // the death cell: the first active final deposit
        const ull fm = __ballot(act && !ero && !step);
        if (dth < 0 && fm) dth = __shfl(c, __ffsll((long long)fm) - 1);
        if (sm) prev = __shfl(c, 63 - __clzll(sm));
        if (in) {
            const int64_t e = base + k;
            uint32_t* ks = a.keys_in + sbase + 3 * k;
            uint32_t* vs = a.vals_in + sbase + 3 * k;
            uint8_t fl = 0;
            if (act) {
                const float* rec = a.samp + (int64_t)c * NS;
                fl = EF_ACTIVE | (ero ? EF_ERODE : 0) | (step ? EF_STEP : 0) | ((step && v > 0.0f) ? EF_TRACK : 0) |
                     (nb >= 0 ? EF_NB : 0);
                if (a.use_route && (double)rec[S_RFLAG] >= 1.5) fl |= EF_RFLOOR;
                a.d[e] = d;
                a.hs[e] = a.h[c] + a.s[c];
                a.b[e] = rec[S_BASE];
                a.r[e] = rec[S_ROUTE];
                a.ec[e] = a.use_ecap ? a.ecap[c] : 0.0;
                a.rs[e] = a.acc[c];
                if (nb >= 0) {
                    a.hsn[e] = a.h[nb] + a.s[nb];
                    a.rn[e] = a.acc[nb];
                }
                ks[0] = (uint32_t)c;
                vs[0] = (K_SELF << KIND_SHIFT) | (uint32_t)e;
                ks[1] = nb >= 0 ? (uint32_t)nb : a.hole;
                vs[1] = (K_NB << KIND_SHIFT) | (uint32_t)e;
                ks[2] = (uint32_t)c;
                vs[2] = (K_WRITE << KIND_SHIFT) | (uint32_t)e;
            } else {
                ks[0] = ks[1] = ks[2] = a.hole;
                vs[0] = vs[1] = vs[2] = 0;
            }
            a.flags[e] = fl;
            a.w[e] = 0.0;
        }
    }
    // no dirty marks: sweep 0 walks every droplet from its first entry
    for (int64_t wi = lane; wi < a.capw / 4; wi += WAVE) ((uint32_t*)(a.dflag + p * a.capw))[wi] = 0u;
    if (lane == 0) {
        a.dend[p] = 0;
        a.death[p] = dth;
        uint32_t* ks = a.keys_in + sbase + 3 * cap;
        uint32_t* vs = a.vals_in + sbase + 3 * cap;
        if (dth >= 0) {
            a.hsd[p] = a.h[dth] + a.s[dth];
            a.bd[p] = a.samp[(int64_t)dth * NS + S_BASE];
            a.re[p] = a.acc[dth];
            *ks = (uint32_t)dth;
            *vs = (K_END << KIND_SHIFT) | (uint32_t)p;
        } else {
            *ks = a.hole;
            *vs = 0;
        }
    }
    if (__ballot(bad) && lane == 0) {
        if (a.fallback) {
            // lake parks: the serial apply takes the chunk, the sweeps return at once
            a.ctl[CTL_FALLBACK] = 1;
            a.ctl[CTL_DONE] = 1;
        } else {
            atomicOr(a.status, 1);
        }
    }
    if (a.moves && lane == 0 && n_moves) atomicAdd(a.moves, (unsigned long long)n_moves);
}

// 2b. sorted slot of every entry's write (the droplet half writes its amounts
// there), and the slot values: a write's amount starts at 0, every read's
// value at the chunk's starting acc of its cell (the value build gave it)
__global__ void __launch_bounds__(BLOCK) ord_posw_kernel(OrdArgs a) {
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.n_slots) return;
    const uint32_t key = a.keys_out[i];
    if (key == a.hole) return;
    const uint32_t val = a.vals_out[i];
    if ((val >> KIND_SHIFT) == K_WRITE) {
        a.posw[val & IDX_MASK] = (uint32_t)i;
        a.sv[i] = 0.0;
    } else {
        a.sv[i] = a.acc[key];
    }
}

// ---------------------------------------------------------------------------
// 3a. the droplet half: one wave per droplet
// ---------------------------------------------------------------------------
// One batch of a droplet's entries, every field loaded without waiting on the
// flags (the workspace holds all P * cap entries), so the next batch's loads
// can be in flight while this batch's carries are walked.
struct Ent {
    uint8_t fl;
    double d, hs, ac, an, hsn, ec, wold;
    float b, r;
    uint32_t slot, owner;
};

__device__ __forceinline__ Ent load_ent(const OrdArgs& a, int64_t base, int64_t k0, int64_t k_first, int64_t cnt,
                                        int sweep) {
    // unconditional loads at an index clamped into the droplet's slice (see load_slots)
    const int64_t k = k0 + lane_id();
    const int64_t e = base + (k < a.cap ? k : a.cap - 1);
    Ent t;
    t.fl = a.flags[e];
    t.d = a.d[e];
    t.hs = a.hs[e];
    t.ac = a.rs[e];
    t.an = a.rn[e];
    t.hsn = a.hsn[e];
    t.ec = a.ec[e];
    t.wold = a.w[e];
    t.b = a.b[e];
    t.r = a.r[e];
    t.slot = a.posw[e];
    t.owner = a.ownerw[e];
    if (!(k >= k_first && k < cnt)) t.fl = 0;
    if (!a.use_ecap) t.ec = 0.0;
    if (sweep == 0) t.owner = 0u;
    return t;
}

__device__ void droplet_wave(const OrdArgs& a, int64_t p, int sweep) {
    const int lane = lane_id();
    const int64_t cap = a.cap, base = p * cap, cnt = a.cl_count[p];
    // the first entry whose read value changed since the last walk (uniform):
    // the marks of all entries (4 to a word) and of the end read, in one round
    // trip; the walk consumes every mark, so they are cleared here
    int64_t fd = 0;
    if (sweep > 0) {
        uint32_t* fw = (uint32_t*)(a.dflag + p * a.capw);
        const int64_t nw = a.capw / 4;
        const uint8_t de = a.dend[p];
        fd = -1;
        for (int64_t w0 = 0; w0 < nw; w0 += WAVE) {
            const int64_t wi = w0 + lane;
            const uint32_t wd = wi < nw ? fw[wi] : 0u;
            const ull m = __ballot(wd != 0u);
            if (m && fd < 0) {
                const int j = __ffsll((long long)m) - 1;
                fd = 4 * (w0 + j) + (__builtin_ctz((uint32_t)rli((int)wd, j)) >> 3);
            }
            if (wd) fw[wi] = 0u;
        }
        if (fd < 0) {
            if (!de) return;  // no read value changed since the last walk
            fd = cnt;
        }
        if (de && lane == 0) a.dend[p] = 0;
    }
    // the stockpile's inputs, in flight with the entries' (used by lane 0 at the end)
    const int32_t dc = a.death[p];
    double hsd = 0.0, red = 0.0;
    float bdd = 0.0f;
    if (lane == 0 && dc >= 0) {
        hsd = a.hsd[p];
        red = a.re[p];
        bdd = a.bd[p];
    }
    // the carries entering entry fd: stored by the last walk (fd == 0: none yet)
    double D = 0.0, S = 0.0;  // deficit, surplus: uniform
    int n_clamp = 0;          // uniform
    const int64_t k_first = fd < cnt ? fd : cnt;
    if (k_first > 0) {
        if (k_first < cnt) {
            D = a.cd[base + k_first];
            S = a.cs[base + k_first];
            n_clamp = a.cn[base + k_first];
        } else {
            D = a.fc[3 * p];
            S = a.fc[3 * p + 1];
            n_clamp = (int)a.fc[3 * p + 2];
        }
    }
    bool changed = false;
    const int64_t kb0 = (k_first / WAVE) * WAVE;
    // two batches ahead in flight: a walk is a chain of batches whose loads come from HBM or the MALL
    Ent en = load_ent(a, base, kb0, k_first, cnt, sweep);
    Ent nx = load_ent(a, base, kb0 + WAVE, k_first, cnt, sweep);
    for (int64_t k0 = kb0; k0 < cnt; k0 += WAVE) {
        const int64_t k = k0 + lane, e = base + k;
        const Ent nx2 = load_ent(a, base, k0 + 2 * WAVE, k_first, cnt, sweep);  // in flight during this batch's walk
        const uint8_t fl = en.fl;
        const bool act = fl & EF_ACTIVE, ero = fl & EF_ERODE;
        double d = 0.0, w = 0.0, x = 0.0, U = 0.0;
        const double wold = en.wold;
        const uint32_t slot = en.slot, owner = en.owner;
        bool clamped = false;
        if (act) {
            d = en.d;
            const double hs = en.hs, ac = en.ac, an = en.an;
            const float b_c = en.b;
            const double s_c = hs + ac;
            const bool sub = s_c < (double)b_c;
            if (ero) {
                const double ev = -d;
                double e_act = ev;
                double rm = a.iter_erode + ac;
                if (a.use_ecap) {
                    const double r2 = en.ec + ac;
                    if (r2 < rm) rm = r2;
                }
                if (e_act > rm) e_act = rm;
                double floor_ = s_c >= (double)b_c ? (double)b_c : s_c;
                if (fl & EF_RFLOOR) {
                    const float r_c = en.r;
                    if ((double)r_c > floor_) floor_ = (double)r_c;
                }
                if (fl & EF_NB) {
                    const double ns = en.hsn + an;
                    if (ns > floor_) floor_ = ns;
                }
                if (e_act > s_c - floor_) e_act = s_c - floor_;
                if (e_act < 0.0) e_act = 0.0;
                clamped = e_act < ev;
                x = ev - e_act;
                w = -e_act;
            } else {
                // the deposit's limits, one bound: the reference clamps by the
                // budget, then the ceiling, then the waterline; with no NaN the
                // successive "if (x > lim) x = lim" give the least of them
                U = a.iter_deposit - ac;
                if (fl & EF_NB) {
                    double ceil_ = en.hsn + an + a.dep_floor;
                    if (sub && (fl & EF_STEP)) ceil_ -= a.dep_floor + a.fan_slope;
                    const double lim = ceil_ - s_c;
                    if (lim < U) U = lim;
                }
                if (sub) {
                    const double lim = (double)b_c - a.dep_floor - s_c;
                    if (lim < U) U = lim;
                }
            }
        }
        // the carries, entry after entry (deposits and clamped erosions only)
        const ull rel = __ballot(act && (!ero || clamped));
        ull m = rel;
        double Da = 0.0, Sa = 0.0;  // carries after this lane's entry
        int Na = 0;
        const double D_in = D, S_in = S;
        const int N_in = n_clamp;
        while (m) {
            const int j = __ffsll((long long)m) - 1;
            m &= m - 1;
            if (__builtin_amdgcn_readlane((int)ero, j)) {
                D = D + rl(x, j);
                n_clamp += 1;
            } else {
                double want = rl(d, j) + S;
                S = 0.0;
                if (D > 0.0) {
                    const double t = want < D ? want : D;
                    want -= t;
                    D -= t;
                }
                const double u = rl(U, j);
                double d_act = want > u ? u : want;
                if (d_act < 0.0) d_act = 0.0;
                if (d_act < want) n_clamp += 1;
                S = want - d_act;
                if (lane == j) w = d_act;
            }
            if (lane == j) {
                Da = D;
                Sa = S;
                Na = n_clamp;
            }
        }
        // the carries entering each entry: after the last carry-changing entry before it
        const int lr = last_below(rel, lane);
        const double Db = __shfl(Da, lr < 0 ? lane : lr), Sb = __shfl(Sa, lr < 0 ? lane : lr);
        const int Nb = __shfl(Na, lr < 0 ? lane : lr);
        if (act) {
            a.cd[e] = lr < 0 ? D_in : Db;
            a.cs[e] = lr < 0 ? S_in : Sb;
            a.cn[e] = lr < 0 ? N_in : Nb;
            if (!same_bits(w, wold)) {
                a.w[e] = w;
                a.sv[slot] = w;
                if (sweep > 0) atomicMin(a.rmin + owner, slot);
                changed = true;
            }
        }
        en = nx;
        nx = nx2;
    }
    // the stockpile of what the droplet could not place (particle.py:1019-1041)
    if (lane == 0) {
        a.fc[3 * p] = D;
        a.fc[3 * p + 1] = S;
        a.fc[3 * p + 2] = (double)n_clamp;
        double tp = 0.0, lo = 0.0, pend = 0.0;
        uint8_t pf = 0;
        if (S > 0.0 && dc >= 0) {
            pf = 1;
            if (hsd + red < (double)bdd) {
                double keep = S * (1.0 - a.writeoff);
                if (keep < 1e-4) keep = 0.0;
                pend = keep;
                tp = keep;
                lo = S - keep;
            } else {
                pend = S;
                tp = S;
            }
        }
        double* o = a.drop + 4 * p;
        o[0] = (double)n_clamp;
        o[1] = tp;
        o[2] = D;
        o[3] = lo;
        a.pend[p] = pend;
        a.pendf[p] = pf;
    }
    if (__ballot(changed) && lane == 0) atomicAdd(&a.ctl[CTL_CHANGED + sweep], 1u);
}

// ---------------------------------------------------------------------------
// 3b. the cell half: one wave per range of sorted slots
// ---------------------------------------------------------------------------
// The wave takes every segment that starts in [r * RANGE, (r + 1) * RANGE)
// and walks it to its end (possibly past the range).  Per batch of 64 slots:
// the slots' keys, kinds and values (coalesced; the next batch's loads are
// issued before this batch's serial part), then the segment's running value
// over its writes in order (one lane after the other), then each read's value
// from the last write before it (or the segment's start value, or the
// running value carried in).  A segment's first access is always a read (an
// entry's self read precedes its write), and no write precedes it: its value
// is the segment's start value, the chunk's starting acc of the cell.
struct Slots {
    uint32_t key, val;
    double v;  // the slot's value: a write's amount, a read's value
};

// Unconditional loads (the index clamped into the slots, the values past the
// end masked afterwards): with no branch around them the compiler can keep
// the next batch's loads in flight while it waits for this batch's.
__device__ __forceinline__ Slots load_slots(const OrdArgs& a, int64_t base) {
    const int64_t i = base + lane_id();
    const int64_t ic = i < a.n_slots ? i : a.n_slots - 1;
    Slots b;
    b.key = a.keys_out[ic];
    b.val = a.vals_out[ic];
    b.v = a.sv[ic];
    if (i >= a.n_slots) {
        b.key = a.hole;
        b.val = 0;
        b.v = 0.0;
    }
    return b;
}

__device__ __forceinline__ bool first_start(const OrdArgs& a, int64_t r, int64_t* out) {
    const int lane = lane_id();
    const int64_t lo = r * RANGE, hi = (r + 1) * RANGE < a.n_slots ? (r + 1) * RANGE : a.n_slots;
    for (int64_t i0 = lo; i0 < hi; i0 += WAVE) {
        const int64_t i = i0 + lane;
        bool st = false;
        if (i < hi) {
            const uint32_t k = a.keys_out[i];
            st = k != a.hole && (i == 0 || a.keys_out[i - 1] != k);
        }
        const ull m = __ballot(st);
        if (m) {
            *out = i0 + __ffsll((long long)m) - 1;
            return true;
        }
    }
    return false;
}

template <bool FINISH>
__device__ void range_wave(const OrdArgs& a, int64_t r, int sweep) {
    const int lane = lane_id();
    int64_t i0;
    double v = 0.0;  // running value of the current segment (uniform)
    uint32_t cur = 0xffffffffu;
    if (!FINISH && sweep > 0) {
        // only from the first write of this range's segments that changed: every
        // access before it keeps its value, and the value entering that write is
        // the read just before it (its own entry's: same cell, no write between)
        const uint32_t first = a.rmin[r];
        if (first == NONE) return;
        if (lane == 0) a.rmin[r] = NONE;
        i0 = first;
        cur = a.keys_out[first];
        v = a.sv[first - 1];
    } else if (!first_start(a, r, &i0)) {
        return;
    }
    const int64_t hi = (r + 1) * RANGE;  // segments starting at or past hi belong to the next range
    uint32_t carry_key = i0 > 0 ? a.keys_out[i0 - 1] : 0xffffffffu;  // the key of the slot before the batch
    double dv = 0.0, m0 = 0.0, m1 = 0.0, pv = 0.0;  // finish: tracks and stockpile running values
    bool wrote = false, tracked = false, parked = false;
    auto write_back = [&]() {
        if (!FINISH || cur == 0xffffffffu || lane != 0) return;
        if (wrote) {
            a.acc[cur] = v;
            float* rec = a.samp + (int64_t)cur * NS;
            rec[S_SURF] = (float)(a.h[cur] + a.s[cur] + v);
            const double sd = a.s[cur] + v;
            rec[S_SED] = (float)(sd > 0.0 ? sd : 0.0);
            if (a.use_route && (double)rec[S_RFLAG] < 1.5) rec[S_RFLAG] = rec[S_ROUTE] != rec[S_SURF] ? 1.0f : 0.0f;
        }
        if (tracked) {
            a.disch[cur] = dv;
            a.mom[2 * (int64_t)cur] = m0;
            a.mom[2 * (int64_t)cur + 1] = m1;
        }
        if (parked) a.pending[cur] = pv;
    };
    // two batches ahead in flight (a long segment is a chain of batches)
    Slots b = load_slots(a, i0);
    Slots nb = load_slots(a, i0 + WAVE);
    for (int64_t base = i0;; base += WAVE) {
        const int64_t i = base + lane;
        const bool more = base + WAVE < a.n_slots;
        const Slots nb2 = load_slots(a, base + 2 * WAVE);  // in flight during this batch's serial part
        const uint32_t key = b.key, val = b.val;
        uint32_t prevkey = __shfl_up(key, 1);
        if (lane == 0) prevkey = carry_key;
        carry_key = (uint32_t)__builtin_amdgcn_readlane((int)key, WAVE - 1);
        // (resuming: the first slot continues its segment)
        const bool start = key != a.hole && (i == 0 || prevkey != key) && !(base == i0 && lane == 0 && cur != 0xffffffffu);
        // stop at the first start at or past hi, or at a hole (the tail of the sort)
        const ull stop_m = __ballot(key == a.hole || (start && i >= hi));
        const int stop = stop_m ? __ffsll((long long)stop_m) - 1 : WAVE;
        const bool valid = lane < stop;
        const uint32_t kind = val >> KIND_SHIFT, idx = val & IDX_MASK;
        const bool write = valid && kind == K_WRITE;
        const double wv = write ? b.v : 0.0, a0 = (valid && start) ? b.v : 0.0;
        double tv = 0.0, t0 = 0.0, t1 = 0.0, pk = 0.0, d0 = 0.0, mm0 = 0.0, mm1 = 0.0, p0 = 0.0;
        bool trk = false, pnd = false;
        if (FINISH) {
            if (write && (a.flags[idx] & EF_TRACK)) {
                trk = true;
                tv = (double)a.cl_vol[idx];
                t0 = (double)a.cl_mom[2 * (int64_t)idx];
                t1 = (double)a.cl_mom[2 * (int64_t)idx + 1];
            }
            if (valid && kind == K_END && a.pendf[idx]) {
                pnd = true;
                pk = a.pend[idx];
            }
            if (valid && start) {
                d0 = a.disch[key];
                mm0 = a.mom[2 * (int64_t)key];
                mm1 = a.mom[2 * (int64_t)key + 1];
                p0 = a.pending[key];
            }
        }
        const ull sm = __ballot(valid && start), wm = __ballot(write);
        const ull tm = FINISH ? __ballot(trk) : 0ull, pm = FINISH ? __ballot(pnd) : 0ull;
        const double v_in = v;
        // the serial part: segment starts and writes (and the finish's sums), in slot order
        double vafter = 0.0;
        ull m = sm | wm | tm | pm;
        while (m) {
            const int j = __ffsll((long long)m) - 1;
            m &= m - 1;
            const ull bit = 1ull << j;
            if (sm & bit) {
                write_back();
                cur = (uint32_t)rli((int)key, j);
                v = rl(a0, j);
                wrote = tracked = parked = false;
                if (FINISH) {
                    dv = rl(d0, j);
                    m0 = rl(mm0, j);
                    m1 = rl(mm1, j);
                    pv = rl(p0, j);
                }
            }
            if (wm & bit) {
                v = v + rl(wv, j);
                wrote = true;
                if (lane == j) vafter = v;
            }
            if (FINISH && (tm & bit)) {
                dv = dv + rl(tv, j);
                m0 = m0 + rl(t0, j);
                m1 = m1 + rl(t1, j);
                tracked = true;
            }
            if (FINISH && (pm & bit)) {
                pv = pv + rl(pk, j);
                parked = true;
            }
        }
        if (!FINISH) {
            // each read's value: the running value just before it -- after the last
            // write before it in its segment, else its segment's start value, else
            // the value carried into the batch
            const int lw = last_below(wm, lane), ls = last_upto(sm, lane);
            const double from_w = __shfl(vafter, lw < 0 ? lane : lw);
            const double from_s = __shfl(a0, ls < 0 ? lane : ls);
            if (valid) {
                if (sweep == 0 && kind == K_WRITE) a.ownerw[idx] = (uint32_t)r;
                if (kind != K_WRITE) {
                    const double rv = (lw >= 0 && lw > ls) ? from_w : (ls >= 0 ? from_s : v_in);
                    if (!same_bits(rv, b.v)) {
                        a.sv[i] = rv;
                        double* dst = kind == K_SELF ? a.rs + idx : (kind == K_NB ? a.rn + idx : a.re + idx);
                        *dst = rv;
                        // the droplet reading it walks again from this entry
                        if (kind == K_END) {
                            a.dend[idx] = 1;
                        } else {
                            const int64_t dp = (int64_t)idx / a.cap;
                            a.dflag[dp * a.capw + ((int64_t)idx - dp * a.cap)] = 1;
                        }
                    }
                }
            }
        }
        if (stop < WAVE || !more) break;
        b = nb;
        nb = nb2;
    }
    write_back();
}

// 3. the sweep launches: a droplet half that changed nothing ends the chunk
__global__ void __launch_bounds__(BLOCK) ord_droplets_kernel(OrdArgs a, int sweep) {
    if (a.ctl[CTL_DONE]) return;
    const int64_t p = (int64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x / WAVE);
    if (p < a.P) droplet_wave(a, p, sweep);
}

__global__ void __launch_bounds__(BLOCK) ord_cells_kernel(OrdArgs a, int sweep) {
    if (a.ctl[CTL_DONE]) return;
    if (a.ctl[CTL_CHANGED + sweep] == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.ctl[CTL_DONE] = 1;
            a.ctl[CTL_SWEEPS] = (uint32_t)sweep + 1;
        }
        return;
    }
    const int64_t r = (int64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x / WAVE);
    if (r < a.n_ranges) range_wave<false>(a, r, sweep);
}

// The sweeps past the launched ones, in one cooperative launch.  Every block
// passes the same grid barriers: the stop test reads a word all blocks see
// after the same barrier, so the grid leaves the loop together.
__global__ void __launch_bounds__(BLOCK) ord_tail_kernel(OrdArgs a, int first) {
    cg::grid_group grid = cg::this_grid();
    if (a.ctl[CTL_DONE]) return;  // read by every block before any block can set it below
    const int64_t wv = (int64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x / WAVE);
    const int64_t nw = (int64_t)gridDim.x * WAVES_PER_BLOCK;
    int sweep = first;
    for (; sweep < a.max_sweeps; sweep++) {
        for (int64_t p = wv; p < a.P; p += nw) droplet_wave(a, p, sweep);
        grid.sync();
        if (__hip_atomic_load(&a.ctl[CTL_CHANGED + sweep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        for (int64_t r = wv; r < a.n_ranges; r += nw) range_wave<false>(a, r, sweep);
        grid.sync();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (sweep < a.max_sweeps) {
            a.ctl[CTL_DONE] = 1;
            a.ctl[CTL_SWEEPS] = (uint32_t)sweep + 1;
        } else if (a.fallback) {
            a.ctl[CTL_FALLBACK] = 1;
        } else {
            atomicOr(a.status, 2);
        }
    }
}

// 4. finish: final acc, tracks, stockpiles and packed record per segment
__global__ void __launch_bounds__(BLOCK) ord_finish_kernel(OrdArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int sw = a.ctl[CTL_DONE] ? (int)a.ctl[CTL_SWEEPS] : a.max_sweeps;
        atomicMax(&a.status[1], sw);
        atomicAdd(&a.status[2], sw);
    }
    if (a.status[0] != 0 || a.ctl[CTL_FALLBACK]) return;
    const int64_t r = (int64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x / WAVE);
    if (r < a.n_ranges) range_wave<true>(a, r, 0);
}

// 5. the chunk's (clamped, to pending, deficit lost, lost offshore), summed in
// droplet order: the block stages up to STAT_TILE droplets' values in LDS
// (coalesced loads), then one lane adds them in order.
constexpr int STAT_TILE = 1024;
__global__ void __launch_bounds__(BLOCK) ord_stats_kernel(OrdArgs a) {
    __shared__ double tile[3 * STAT_TILE];
    __shared__ unsigned long long n_clamp;
    if (threadIdx.x == 0) n_clamp = 0;
    double tp = 0.0, lost = 0.0, lo = 0.0;
    for (int64_t p0 = 0; p0 < a.P; p0 += STAT_TILE) {
        const int64_t m = a.P - p0 < STAT_TILE ? a.P - p0 : STAT_TILE;
        __syncthreads();
        unsigned long long nc = 0;
        for (int64_t j = threadIdx.x; j < m; j += blockDim.x) {
            const double* o = a.drop + 4 * (p0 + j);
            nc += (unsigned long long)o[0];
            tile[3 * j] = o[1];
            tile[3 * j + 1] = o[2];
            tile[3 * j + 2] = o[3];
        }
        if (nc) atomicAdd(&n_clamp, nc);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int64_t j = 0; j < m; j++) {
                tp += tile[3 * j];
                lost += tile[3 * j + 1];
                lo += tile[3 * j + 2];
            }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.out4[0] = (double)n_clamp;
        a.out4[1] = tp;
        a.out4[2] = lost;
        a.out4[3] = lo;
    }
}

__global__ void ord_tail_kernel_flag(OrdArgs a) {
    if (a.ctl[CTL_DONE]) return;
    if (a.fallback) a.ctl[CTL_FALLBACK] = 1;
    else atomicOr(a.status, 2);
}

int tail_grid(int* blocks) {
    static int cached = 0;
    if (cached == 0) {
        int dev = 0, per_cu = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)ord_tail_kernel, BLOCK, 0) != hipSuccess)
            return -1;
        const int n = per_cu * prop.multiProcessorCount;
        cached = n < 64 ? n : 64;
        if (cached < 1) cached = 1;
    }
    *blocks = cached;
    return 0;
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" int64_t pg_ordered_workspace_bytes(int64_t P, int64_t cap, int max_sweeps) {
    clear_error();
    if (P < 0 || cap <= 0 || max_sweeps < 1) {
        set_error("pg_ordered_workspace_bytes: bad sizes P=%lld cap=%lld max_sweeps=%d", (long long)P, (long long)cap, max_sweeps);
        return -1;
    }
    Layout L;
    if (layout_of(P, cap, max_sweeps, &L) != 0) {
        set_error("pg_ordered_workspace_bytes: sort temp-size query failed");
        return -1;
    }
    return (int64_t)L.total;
}

namespace pg {
namespace {
// drop_out: NULL -> the 4-tuple is summed into out4 here; else each droplet's
// contributions land in drop_out[4 * p ...] (the caller sums them in order)
int ordered_enqueue(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol, const float* cl_mom,
                    const int64_t* cl_count, int64_t P, int64_t cap, const double* height, const double* sediment,
                    double* acc, double* pending, float* samp, double* disch_track, double* mom_track,
                    const uint8_t* mask, int64_t n_cells, double iter_erode, double iter_deposit, double fan_slope,
                    int use_route, double dep_floor, double offshore_writeoff, const double* erode_cap,
                    int use_erode_cap, double* out4, int32_t* status, int max_sweeps, int launched_sweeps, void* work,
                    int64_t work_bytes, hipStream_t stream, double* drop_out, int64_t* moves, int fallback = 0,
                    double* lake_load = nullptr, const int32_t* lake_id = nullptr, double* lake_room = nullptr) {
    PG_REQUIRE(P >= 0 && cap > 0 && n_cells > 0, "pg_apply_changes_ordered: bad sizes P=%lld cap=%lld n_cells=%lld",
               (long long)P, (long long)cap, (long long)n_cells);
    PG_REQUIRE(max_sweeps >= 1 && launched_sweeps >= 0, "pg_apply_changes_ordered: max_sweeps must be >= 1, launched_sweeps >= 0");
    PG_REQUIRE((out4 || drop_out) && status, "pg_apply_changes_ordered: out4 / status is NULL");
    PG_REQUIRE(!use_erode_cap || erode_cap, "pg_apply_changes_ordered: use_erode_cap without erode_cap");
    PG_REQUIRE((uint64_t)n_cells < 0x7fffffffull, "pg_apply_changes_ordered: %lld cells exceed 31-bit cell ids",
               (long long)n_cells);
    PG_REQUIRE((uint64_t)P * (uint64_t)cap <= IDX_MASK, "pg_apply_changes_ordered: P * cap = %lld exceeds %u entries",
               (long long)(P * cap), IDX_MASK);
    PG_REQUIRE((uint64_t)P * (uint64_t)(3 * cap + 1) < 0x7fffffffull, "pg_apply_changes_ordered: too many accesses");
    Layout L;
    PG_REQUIRE(layout_of(P, cap, max_sweeps, &L) == 0, "pg_apply_changes_ordered: sort temp-size query failed");
    PG_REQUIRE(work && work_bytes >= (int64_t)L.total, "pg_apply_changes_ordered: workspace of %lld bytes, %lld needed",
               (long long)work_bytes, (long long)L.total);
    hipStream_t s = stream;
    if (P == 0) {
        if (out4) PG_HIP(hipMemsetAsync(out4, 0, 4 * sizeof(double), s));
        return 0;
    }
    PG_REQUIRE(cl_cell && cl_delta && cl_vol && cl_mom && cl_count && height && sediment && acc && pending && samp &&
                   disch_track && mom_track && mask,
               "pg_apply_changes_ordered: NULL array");
    int nbits = 1;
    while (nbits < 31 && ((uint64_t)1 << nbits) <= (uint64_t)n_cells) nbits++;
    OrdArgs a{};
    a.cl_cell = cl_cell;
    a.cl_delta = cl_delta;
    a.cl_vol = cl_vol;
    a.cl_mom = cl_mom;
    a.cl_count = cl_count;
    a.P = P;
    a.cap = cap;
    a.h = height;
    a.s = sediment;
    a.acc = acc;
    a.pending = pending;
    a.samp = samp;
    a.disch = disch_track;
    a.mom = mom_track;
    a.mask = mask;
    a.ecap = erode_cap;
    a.use_ecap = use_erode_cap;
    a.use_route = use_route;
    a.iter_erode = iter_erode;
    a.iter_deposit = iter_deposit;
    a.fan_slope = fan_slope;
    a.dep_floor = dep_floor;
    a.writeoff = offshore_writeoff;
    a.flags = at<uint8_t>(work, L.flags);
    a.d = at<double>(work, L.d);
    a.hs = at<double>(work, L.hs);
    a.hsn = at<double>(work, L.hsn);
    a.ec = at<double>(work, L.ec);
    a.b = at<float>(work, L.b);
    a.r = at<float>(work, L.r);
    a.rs = at<double>(work, L.rs);
    a.rn = at<double>(work, L.rn);
    a.w = at<double>(work, L.w);
    a.posw = at<uint32_t>(work, L.posw);
    a.cd = at<double>(work, L.cd);
    a.cs = at<double>(work, L.cs);
    a.cn = at<int32_t>(work, L.cn);
    a.dflag = at<uint8_t>(work, L.dflag);
    a.dend = at<uint8_t>(work, L.dend);
    a.capw = (cap + 3) & ~(int64_t)3;
    a.fc = at<double>(work, L.fc);
    a.ownerw = at<uint32_t>(work, L.ownerw);
    a.rmin = at<uint32_t>(work, L.rmin);
    a.death = at<int32_t>(work, L.death);
    a.hsd = at<double>(work, L.hsd);
    a.bd = at<float>(work, L.bd);
    a.re = at<double>(work, L.re);
    a.drop = drop_out ? drop_out : at<double>(work, L.drop);
    a.pend = at<double>(work, L.pend);
    a.pendf = at<uint8_t>(work, L.pendf);
    a.keys_in = at<uint32_t>(work, L.keys_in);
    a.vals_in = at<uint32_t>(work, L.vals_in);
    a.keys_out = at<uint32_t>(work, L.keys_out);
    a.vals_out = at<uint32_t>(work, L.vals_out);
    a.sv = at<double>(work, L.sv);
    a.ctl = at<uint32_t>(work, L.ctl);
    a.status = status;
    a.hole = (uint32_t)(((uint64_t)1 << nbits) - 1);
    a.n_slots = P * (3 * cap + 1);
    a.n_ranges = (a.n_slots + RANGE - 1) / RANGE;
    a.max_sweeps = max_sweeps;
    a.fallback = fallback;
    a.out4 = out4;
    a.moves = (unsigned long long*)moves;
    const int dblocks = (int)((P + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK);
    const int rblocks = (int)((a.n_ranges + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK);
    PG_REQUIRE(!fallback || drop_out, "pg_apply_changes_ordered: the serial fallback needs per-droplet outputs");
    hipLaunchKernelGGL(ord_reset_kernel, dim3(1), dim3(BLOCK), 0, s, a);
    PG_LAUNCHED("ord_reset_kernel");
    hipLaunchKernelGGL(ord_build_kernel, dim3(dblocks), dim3(BLOCK), 0, s, a);
    PG_LAUNCHED("ord_build_kernel");
    size_t tb = L.temp_bytes;
    PG_HIP(rocprim::radix_sort_pairs<AccessSortConfig>(at<void>(work, L.temp), tb, a.keys_in, a.keys_out, a.vals_in,
                                                       a.vals_out, (unsigned int)a.n_slots, 0u, (unsigned int)nbits,
                                                       s));
    hipLaunchKernelGGL(ord_posw_kernel, dim3(blocks_for(a.n_slots, BLOCK)), dim3(BLOCK), 0, s, a);
    PG_LAUNCHED("ord_posw_kernel");
    const int launched = launched_sweeps < max_sweeps ? launched_sweeps : max_sweeps;
    for (int sw = 0; sw < launched; sw++) {
        hipLaunchKernelGGL(ord_droplets_kernel, dim3(dblocks), dim3(BLOCK), 0, s, a, sw);
        hipLaunchKernelGGL(ord_cells_kernel, dim3(rblocks), dim3(BLOCK), 0, s, a, sw);
    }
    PG_LAUNCHED("ord_droplets_kernel / ord_cells_kernel");
    if (launched < max_sweeps) {
        int blocks = 0;
        PG_REQUIRE(tail_grid(&blocks) == 0, "pg_apply_changes_ordered: occupancy query failed");
        int first = launched;
        void* kargs[] = {&a, &first};
        PG_HIP(hipLaunchCooperativeKernel((const void*)ord_tail_kernel, dim3(blocks), dim3(BLOCK), kargs, 0, s));
    } else {
        // every sweep was launched: a chunk still running after them did not converge
        hipLaunchKernelGGL(ord_tail_kernel_flag, dim3(1), dim3(1), 0, s, a);
        PG_LAUNCHED("ord_tail_kernel_flag");
    }
    hipLaunchKernelGGL(ord_finish_kernel, dim3(rblocks), dim3(BLOCK), 0, s, a);
    PG_LAUNCHED("ord_finish_kernel");
    if (fallback) {
        // a chunk with lake parks, or one that did not converge within max_sweeps (nothing of it
        // applied above), goes through the serial apply: the reference's own walk, one lane
        if (enqueue_serial_apply(cl_cell, cl_delta, cl_vol, cl_mom, cl_count, P, cap, height, sediment, acc, pending,
                                 samp, disch_track, mom_track, mask, iter_erode, iter_deposit, fan_slope, use_route,
                                 dep_floor, offshore_writeoff, erode_cap, use_erode_cap, lake_load, lake_id, lake_room,
                                 a.ctl + CTL_FALLBACK, drop_out, status + 3, s) != 0)
            return -1;
    }
    if (!drop_out) {
        hipLaunchKernelGGL(ord_stats_kernel, dim3(1), dim3(BLOCK), 0, s, a);
        PG_LAUNCHED("ord_stats_kernel");
    }
    return 0;
}

// Each chunk's 4-tuple from its droplets' contributions, in droplet order:
// one block per chunk (stage in LDS, one lane adds in order), chunks in parallel.
__global__ void __launch_bounds__(BLOCK) ord_chunk_sums_kernel(const double* drop, int64_t P, int64_t chunk, double* outs) {
    __shared__ double tile[3 * STAT_TILE];
    __shared__ unsigned long long n_clamp;
    const int64_t c0 = (int64_t)blockIdx.x * chunk;
    const int64_t m_all = P - c0 < chunk ? P - c0 : chunk;
    if (threadIdx.x == 0) n_clamp = 0;
    double tp = 0.0, lost = 0.0, lo = 0.0;
    for (int64_t p0 = 0; p0 < m_all; p0 += STAT_TILE) {
        const int64_t m = m_all - p0 < STAT_TILE ? m_all - p0 : STAT_TILE;
        __syncthreads();
        unsigned long long nc = 0;
        for (int64_t j = threadIdx.x; j < m; j += blockDim.x) {
            const double* o = drop + 4 * (c0 + p0 + j);
            nc += (unsigned long long)o[0];
            tile[3 * j] = o[1];
            tile[3 * j + 1] = o[2];
            tile[3 * j + 2] = o[3];
        }
        if (nc) atomicAdd(&n_clamp, nc);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int64_t j = 0; j < m; j++) {
                tp += tile[3 * j];
                lost += tile[3 * j + 1];
                lo += tile[3 * j + 2];
            }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* o = outs + 4 * blockIdx.x;
        o[0] = (double)n_clamp;
        o[1] = tp;
        o[2] = lost;
        o[3] = lo;
    }
}
// Trace-launch timing of pg_check_chunks: event pairs reused across calls,
// summed and released by pg_check_trace_time.
struct CheckTimer {
    std::vector<hipEvent_t> start, stop;
    size_t recorded = 0;
};
CheckTimer g_check_timer;

int timer_mark(hipStream_t s, bool is_start) {
    CheckTimer& t = g_check_timer;
    const size_t i = t.recorded;
    if (is_start && i >= t.start.size()) {
        hipEvent_t a, b;
        PG_HIP(hipEventCreate(&a));
        PG_HIP(hipEventCreate(&b));
        t.start.push_back(a);
        t.stop.push_back(b);
    }
    if (is_start) {
        PG_HIP(hipEventRecord(t.start[i], s));
    } else {
        PG_HIP(hipEventRecord(t.stop[i], s));
        t.recorded = i + 1;
    }
    return 0;
}
}  // namespace
}  // namespace pg

extern "C" int pg_apply_changes_ordered(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol,
                                        const float* cl_mom, const int64_t* cl_count, int64_t P, int64_t cap,
                                        const double* height, const double* sediment, double* acc, double* pending,
                                        float* samp, double* disch_track, double* mom_track, const uint8_t* mask,
                                        int64_t n_cells, double iter_erode, double iter_deposit, double fan_slope,
                                        int use_route, double dep_floor, double offshore_writeoff,
                                        const double* erode_cap, int use_erode_cap, double* out4, int32_t* status,
                                        int max_sweeps, int launched_sweeps, void* work, int64_t work_bytes,
                                        pg_stream stream) {
    clear_error();
    PG_REQUIRE(out4 != nullptr, "pg_apply_changes_ordered: out4 is NULL");
    return ordered_enqueue(cl_cell, cl_delta, cl_vol, cl_mom, cl_count, P, cap, height, sediment, acc, pending, samp,
                           disch_track, mom_track, mask, n_cells, iter_erode, iter_deposit, fan_slope, use_route,
                           dep_floor, offshore_writeoff, erode_cap, use_erode_cap, out4, status, max_sweeps,
                           launched_sweeps, work, work_bytes, (hipStream_t)stream, nullptr, nullptr);
}

// ---- the same apply with no allocation and no workspace argument ----------
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops): the workspace is a file-scope device array, held by
// every GPU the library runs on until the process ends.  Calls that overlap
// on one GPU share it and race.  Nothing in it needs clearing between calls:
// the build kernel initialises everything the apply reads.  Every sweep is a
// plain launch (no cooperative launch), max_sweeps of them at most.
namespace pg {
namespace {
constexpr size_t ORDERED_SCRATCH_BYTES = size_t(96) << 20;
std::mutex g_scratch_mu;
void* g_scratch_ptr[64];
}  // namespace
}  // namespace pg

static __device__ __align__(256) unsigned char scratch[pg::ORDERED_SCRATCH_BYTES];

namespace pg {
namespace {
// The array's device address on the current device, looked up once per device.
int scratch_address(void** out) {
    int dev = 0;
    PG_HIP(hipGetDevice(&dev));
    PG_REQUIRE(dev >= 0 && dev < 64, "scratch: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        PG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        PG_REQUIRE(ptr != nullptr, "scratch: hipGetSymbolAddress returned NULL");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace
}  // namespace pg

extern "C" int64_t pg_ordered_scratch_bytes(void) { return (int64_t)ORDERED_SCRATCH_BYTES; }

extern "C" int pg_apply_changes_ordered_noalloc(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol,
                                                const float* cl_mom, const int64_t* cl_count, int64_t P, int64_t cap,
                                                const double* height, const double* sediment, double* acc,
                                                double* pending, float* samp, double* disch_track, double* mom_track,
                                                const uint8_t* mask, int64_t n_cells, double iter_erode,
                                                double iter_deposit, double fan_slope, int use_route, double dep_floor,
                                                double offshore_writeoff, const double* erode_cap, int use_erode_cap,
                                                double* out4, int32_t* status, int max_sweeps, pg_stream stream) {
    clear_error();
    PG_REQUIRE(out4 != nullptr, "pg_apply_changes_ordered_noalloc: out4 is NULL");
    PG_REQUIRE(P >= 0 && cap > 0 && max_sweeps >= 1 && max_sweeps <= 4096,
               "pg_apply_changes_ordered_noalloc: bad sizes P=%lld cap=%lld max_sweeps=%d", (long long)P, (long long)cap,
               max_sweeps);
    Layout L;
    PG_REQUIRE(layout_of(P, cap, max_sweeps, &L) == 0, "pg_apply_changes_ordered_noalloc: sort temp-size query failed");
    PG_REQUIRE(L.total <= ORDERED_SCRATCH_BYTES,
               "pg_apply_changes_ordered_noalloc: P=%lld cap=%lld max_sweeps=%d need %llu bytes of scratch, the maximum "
               "is %llu (chunks of up to 2048 droplets with lifetimes up to 128 fit)",
               (long long)P, (long long)cap, max_sweeps, (unsigned long long)L.total,
               (unsigned long long)ORDERED_SCRATCH_BYTES);
    void* ptr = nullptr;
    if (scratch_address(&ptr) != 0) return -1;
    return ordered_enqueue(cl_cell, cl_delta, cl_vol, cl_mom, cl_count, P, cap, height, sediment, acc, pending, samp,
                           disch_track, mom_track, mask, n_cells, iter_erode, iter_deposit, fan_slope, use_route,
                           dep_floor, offshore_writeoff, erode_cap, use_erode_cap, out4, status, max_sweeps, max_sweeps,
                           ptr, (int64_t)ORDERED_SCRATCH_BYTES, (hipStream_t)stream, nullptr, nullptr);
}

extern "C" int pg_check_chunks(int64_t P, int64_t chunk, const int64_t* sp_face, const double* sp_x, const double* sp_y,
                               const double* sp_sed, const double* sp_dir, double volume0, float* samp, int use_route,
                               const uint8_t* mask, int64_t N, int64_t H, int spherical, const pg_trace_params* prm,
                               const float* roots, int use_roots, int32_t* cl_cell, double* cl_delta, float* cl_vol,
                               float* cl_mom, int64_t* cl_count, int64_t* cl_lo, int64_t* cl_hi, int8_t* sp_death,
                               const double* height, const double* sediment, double* acc, double* pending,
                               double* disch_track, double* mom_track, double iter_erode, double iter_deposit,
                               double fan_slope, double dep_floor, double offshore_writeoff, const double* erode_cap,
                               int use_erode_cap, double* lake_load, const int32_t* lake_id, double* lake_room,
                               double* outs, int64_t* counts7, int32_t* status, int max_sweeps,
                               int launched_sweeps, void* work, int64_t work_bytes, double* drop_out,
                               int64_t* moves, int timing, pg_stream stream) {
    clear_error();
    PG_REQUIRE(P >= 0 && chunk >= 1 && N >= 2 && H >= 1, "pg_check_chunks: bad sizes P=%lld chunk=%lld", (long long)P,
               (long long)chunk);
    PG_REQUIRE(prm && outs && counts7 && status && drop_out, "pg_check_chunks: NULL argument");
    const int64_t NE = N + 2 * H, F = spherical ? 6 : 1;
    const int64_t cap = pg_change_list_cap(prm->max_steps, prm->lateral_rate > 0.0);
    for (int64_t c0 = 0, ci = 0; c0 < P; c0 += chunk, ci++) {
        const int64_t m = P - c0 < chunk ? P - c0 : chunk;
        if (timing && timer_mark((hipStream_t)stream, true) != 0) return -1;
        if (pg_trace_particles(m, sp_face + c0, sp_x + c0, sp_y + c0, sp_sed + c0, sp_dir + c0, volume0, samp, use_route,
                               mask, N, H, spherical, prm, roots, use_roots, cl_cell, cl_delta, cl_vol, cl_mom, cl_count,
                               cl_lo, cl_hi, sp_death, stream) != 0)
            return -1;
        if (timing && timer_mark((hipStream_t)stream, false) != 0) return -1;
        if (ordered_enqueue(cl_cell, cl_delta, cl_vol, cl_mom, cl_count, m, cap, height, sediment, acc, pending, samp,
                            disch_track, mom_track, mask, F * NE * NE, iter_erode, iter_deposit, fan_slope, use_route,
                            dep_floor, offshore_writeoff, erode_cap, use_erode_cap, nullptr, status, max_sweeps,
                            launched_sweeps, work, work_bytes, (hipStream_t)stream, drop_out + 4 * c0, moves, 1,
                            lake_load, lake_id, lake_room) != 0)
            return -1;
        if (pg_chunk_stats(cl_count, sp_death, m, counts7, stream) != 0) return -1;
        (void)ci;
    }
    if (P > 0) {
        const int64_t n_chunks = (P + chunk - 1) / chunk;
        hipLaunchKernelGGL(ord_chunk_sums_kernel, dim3((unsigned)n_chunks), dim3(BLOCK), 0, (hipStream_t)stream, drop_out, P,
                           chunk, outs);
        PG_LAUNCHED("ord_chunk_sums_kernel");
    }
    return 0;
}

extern "C" int pg_check_trace_time(double* ms_total, int64_t* launches) {
    clear_error();
    PG_REQUIRE(ms_total && launches, "pg_check_trace_time: NULL argument");
    CheckTimer& t = g_check_timer;
    double ms = 0.0;
    if (t.recorded) PG_HIP(hipEventSynchronize(t.stop[t.recorded - 1]));
    for (size_t i = 0; i < t.recorded; i++) {
        float x = 0.0f;
        PG_HIP(hipEventElapsedTime(&x, t.start[i], t.stop[i]));
        ms += x;
    }
    *ms_total = ms;
    *launches = (int64_t)t.recorded;
    t.recorded = 0;
    return 0;
}
