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
// droplet's list.  The serial result is the one solution of that system: the
// system is triangular in serial order, so it has exactly one, and an
// iteration that alternates the two halves reaches it in finitely many
// sweeps (at most the depth of the dependency chain) -- and once a sweep
// changes nothing, the values are that solution, bit for bit.
//
// One chunk:
//  1. build: one lane per droplet lays out its entries (per-entry constants
//     in an entry-major layout so lanes of a wave read adjacent words), the
//     first R (the chunk's starting acc) and its accesses -- (cell, kind)
//     pairs in serial order: self read, neighbour read, write per entry, the
//     final read of the death cell;
//  2. a stable radix sort of the accesses by cell: each cell's accesses in
//     serial order, one segment per touched cell;
//  3. sweeps (one cooperative launch, grid barriers between the halves):
//     the droplet pass walks each droplet's list from its R and yields W,
//     the wants of its deposits and its carries; the cell scan walks each
//     segment in serial order with the live value, recomputing each write
//     from it (same-cell chains resolve within one scan) and refreshing
//     the R of every read.  A droplet pass that changes no W and no want
//     ends the loop;
//  4. finish: per segment the final acc, the discharge / momentum tracks
//     and the stockpile adds in serial order (float64 sums in the
//     reference's order), the packed record (S_SURF, S_SED, S_RFLAG) of
//     every written cell;
//  5. the chunk's 4-tuple, summed in droplet order.
//
// Lake parks (volume -3 entries, emitted only with lake_park > 0) are not
// handled here: the status word reports them and nothing is applied.
//
// Built with -ffp-contract=off: every expression rounds where the
// reference's does (its operand order is kept below).
#include <hip/hip_cooperative_groups.h>
#include <hipcub/hipcub.hpp>

#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"

namespace cg = cooperative_groups;

namespace pg {
namespace {

constexpr uint8_t EF_ACTIVE = 1, EF_ERODE = 2, EF_STEP = 4, EF_NB = 8, EF_RFLOOR = 16, EF_TRACK = 32;
constexpr uint32_t K_SELF = 0, K_NB = 1, K_WRITE = 2, K_END = 3;
constexpr uint32_t KIND_SHIFT = 30, IDX_MASK = (1u << KIND_SHIFT) - 1u;
constexpr int SWEEP_THREADS = 256;

struct Layout {
    size_t flags, d, hs, hsn, ec, b, r, rs, rn, rn2, w, want;   // per entry (P * cap)
    size_t death, hsd, bd, re, drop, pend, pendf;          // per droplet
    size_t keys_in, vals_in, keys_out, vals_out;            // per access slot
    size_t seg, nseg, changes, status, temp;
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
    L->rn2 = take(ne * 8);
    L->w = take(ne * 8);
    L->want = take(ne * 8);
    L->death = take((size_t)P * 4);
    L->hsd = take((size_t)P * 8);
    L->bd = take((size_t)P * 4);
    L->re = take((size_t)P * 8);
    L->drop = take((size_t)P * 4 * 8);
    L->pend = take((size_t)P * 8);
    L->pendf = take((size_t)P);
    L->keys_in = take(na * 4);
    L->vals_in = take(na * 4);
    L->keys_out = take(na * 4);
    L->vals_out = take(na * 4);
    L->seg = take(na * 4);
    L->nseg = take(8);
    L->changes = take((size_t)(max_sweeps + 1) * 4);
    L->status = take(8);
    size_t tb = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)(na > 0 ? na : 1)) != hipSuccess)
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
    double *rs, *w, *want;
    double* rn[2];  // neighbour reads: the scan reads one buffer and fills the other
    int32_t* death;
    double* hsd;
    float* bd;
    double *re, *drop, *pend;
    uint8_t* pendf;
    uint32_t *keys_in, *vals_in, *keys_out, *vals_out, *seg, *nseg, *changes;
    int32_t* status;  // [0] status bits (1 lake park, 2 no convergence), [1] most sweeps, [2] sweeps summed
    uint32_t hole;
    int64_t n_slots;
    int max_sweeps;
    double* out4;
};

// One entry's amount given its cell's live acc `ac`, the neighbour's live acc
// `an` and (deposits) the droplet's want; the reference's operand order
// (particle.py:934-1000).  *clamped: the entry was limited.
__device__ __forceinline__ double entry_amount(const OrdArgs& a, uint8_t fl, double d, double hs, double hsn, float b_c,
                                               float r_c, double ec, double ac, double an, double want, bool* clamped) {
    const double s_c = hs + ac;
    if (fl & EF_ERODE) {
        const double e = -d;
        double e_act = e;
        double room = a.iter_erode + ac;
        if (a.use_ecap) {
            const double r2 = ec + ac;
            if (r2 < room) room = r2;
        }
        if (e_act > room) e_act = room;
        double floor_ = s_c >= (double)b_c ? (double)b_c : s_c;
        if ((fl & EF_RFLOOR) && (double)r_c > floor_) floor_ = (double)r_c;
        if (fl & EF_NB) {
            const double ns = hsn + an;
            if (ns > floor_) floor_ = ns;
        }
        if (e_act > s_c - floor_) e_act = s_c - floor_;
        if (e_act < 0.0) e_act = 0.0;
        *clamped = e_act < e;
        return -e_act;
    }
    double d_act = want;
    const double room = a.iter_deposit - ac;
    if (d_act > room) d_act = room;
    if (fl & EF_NB) {
        double ceil_ = hsn + an + a.dep_floor;
        if (s_c < (double)b_c && (fl & EF_STEP)) ceil_ -= a.dep_floor + a.fan_slope;
        const double lim = ceil_ - s_c;
        if (d_act > lim) d_act = lim;
    }
    if (s_c < (double)b_c) {
        const double lim = (double)b_c - a.dep_floor - s_c;
        if (d_act > lim) d_act = lim;
    }
    if (d_act < 0.0) d_act = 0.0;
    *clamped = d_act < want;
    return d_act;
}

__device__ __forceinline__ bool same_bits(double x, double y) {
    return __double_as_longlong(x) == __double_as_longlong(y);
}

// 1. per-droplet layout of the entries and their accesses
__global__ void __launch_bounds__(256) ord_build_kernel(OrdArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        *a.nseg = 0;
        for (int i = 0; i <= a.max_sweeps; i++) a.changes[i] = 0;
    }
    if (p >= a.P) return;
    const int64_t P = a.P, cap = a.cap, base = p * cap, cnt = a.cl_count[p];
    const int64_t sbase = p * (3 * cap + 1);
    int64_t prev = -1, dth = -1;
    int32_t bad = 0;
    for (int64_t k = 0; k < cap; k++) {
        const int64_t t = k * P + p;
        uint32_t* ks = a.keys_in + sbase + 3 * k;
        uint32_t* vs = a.vals_in + sbase + 3 * k;
        ks[0] = ks[1] = ks[2] = a.hole;
        vs[0] = vs[1] = vs[2] = 0;
        a.flags[t] = 0;
        a.w[t] = 0.0;
        a.want[t] = 0.0;
        if (k >= cnt) continue;
        const int64_t c = a.cl_cell[base + k];
        const float v = a.cl_vol[base + k];
        if (v == -3.0f) {
            bad = 1;
            continue;
        }
        if (a.mask[c] != MASK_ACTIVE) continue;
        const double d = a.cl_delta[base + k];
        const bool is_step = v >= 0.0f;
        int64_t nb = -1;
        uint8_t fl = EF_ACTIVE;
        if (d < 0.0) {
            fl |= EF_ERODE;
            if (k + 1 < cnt) nb = a.cl_cell[base + k + 1];
        } else {
            if (is_step) nb = prev;
            else if (k + 1 < cnt) nb = a.cl_cell[base + k + 1];
            if (!is_step && dth < 0) dth = c;
        }
        if (is_step) fl |= EF_STEP;
        if (is_step && v > 0.0f) fl |= EF_TRACK;
        const float* rec = a.samp + c * NS;
        if (a.use_route && (double)rec[S_RFLAG] >= 1.5) fl |= EF_RFLOOR;
        a.d[t] = d;
        a.hs[t] = a.h[c] + a.s[c];
        a.b[t] = rec[S_BASE];
        a.r[t] = rec[S_ROUTE];
        a.ec[t] = a.use_ecap ? a.ecap[c] : 0.0;
        a.rs[t] = a.acc[c];
        ks[0] = (uint32_t)c;
        vs[0] = (K_SELF << KIND_SHIFT) | (uint32_t)t;
        if (nb >= 0) {
            fl |= EF_NB;
            a.hsn[t] = a.h[nb] + a.s[nb];
            a.rn[0][t] = a.acc[nb];
            ks[1] = (uint32_t)nb;
            vs[1] = (K_NB << KIND_SHIFT) | (uint32_t)t;
        } else {
            a.hsn[t] = 0.0;
            a.rn[0][t] = 0.0;
            a.rn[1][t] = 0.0;
        }
        ks[2] = (uint32_t)c;
        vs[2] = (K_WRITE << KIND_SHIFT) | (uint32_t)t;
        a.flags[t] = fl;
        if (is_step) prev = c;
    }
    a.death[p] = (int32_t)dth;
    a.keys_in[sbase + 3 * cap] = a.hole;
    a.vals_in[sbase + 3 * cap] = 0;
    if (dth >= 0) {
        a.hsd[p] = a.h[dth] + a.s[dth];
        a.bd[p] = a.samp[dth * NS + S_BASE];
        a.re[p] = a.acc[dth];
        a.keys_in[sbase + 3 * cap] = (uint32_t)dth;
        a.vals_in[sbase + 3 * cap] = (K_END << KIND_SHIFT) | (uint32_t)p;
    }
    if (bad) atomicOr(a.status, 1);
}

// 2b. segment starts of the sorted accesses (one per touched cell; list order free)
__global__ void __launch_bounds__(256) ord_segments_kernel(OrdArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_slots) return;
    const uint32_t k = a.keys_out[i];
    if (k == a.hole) return;
    if (i == 0 || a.keys_out[i - 1] != k) a.seg[atomicAdd(a.nseg, 1u)] = (uint32_t)i;
}

// The droplet half: walk droplet p's list from the current R.  Returns true
// when a W or a want changed.  On the converged pass the per-droplet results
// (clamped entries, stockpile, deficit, offshore loss) are final.
__device__ bool droplet_pass(const OrdArgs& a, const double* rn, int64_t p) {
    const int64_t P = a.P, cnt = a.cl_count[p];
    double deficit = 0.0, surplus = 0.0;
    int64_t n_clamp = 0;
    bool changed = false;
    for (int64_t k = 0; k < cnt; k++) {
        const int64_t t = k * P + p;
        const uint8_t fl = a.flags[t];
        if (!(fl & EF_ACTIVE)) continue;
        const double d = a.d[t];
        double wn = 0.0;
        if (!(fl & EF_ERODE)) {
            wn = d + surplus;
            surplus = 0.0;
            if (deficit > 0.0) {
                const double x = wn < deficit ? wn : deficit;
                wn -= x;
                deficit -= x;
            }
            if (!same_bits(wn, a.want[t])) {
                a.want[t] = wn;
                changed = true;
            }
        }
        bool clamped;
        const double w = entry_amount(a, fl, d, a.hs[t], a.hsn[t], a.b[t], a.r[t], a.ec[t], a.rs[t], rn[t], wn, &clamped);
        if (clamped) n_clamp += 1;
        if (fl & EF_ERODE) {
            if (clamped) deficit += -d - (-w);
        } else {
            surplus = wn - w;
        }
        if (!same_bits(w, a.w[t])) {
            a.w[t] = w;
            changed = true;
        }
    }
    // the stockpile of what the droplet could not place (particle.py:1019-1041)
    double tp = 0.0, lo = 0.0, pend = 0.0;
    uint8_t pf = 0;
    const int32_t dc = a.death[p];
    if (surplus > 0.0 && dc >= 0) {
        pf = 1;
        if (a.hsd[p] + a.re[p] < (double)a.bd[p]) {
            double keep = surplus * (1.0 - a.writeoff);
            if (keep < 1e-4) keep = 0.0;
            pend = keep;
            tp = keep;
            lo = surplus - keep;
        } else {
            pend = surplus;
            tp = surplus;
        }
    }
    double* o = a.drop + 4 * p;
    o[0] = (double)n_clamp;
    o[1] = tp;
    o[2] = deficit;
    o[3] = lo;
    a.pend[p] = pend;
    a.pendf[p] = pf;
    return changed;
}

// The cell half: walk one segment in serial order with the live value,
// recomputing each write from it and handing every read its value.
__device__ void cell_scan(const OrdArgs& a, const double* rn_in, double* rn_out, uint32_t g) {
    uint32_t i = a.seg[g];
    const uint32_t cell = a.keys_out[i];
    double v = a.acc[cell];
    for (; i < (uint32_t)a.n_slots && a.keys_out[i] == cell; i++) {
        const uint32_t val = a.vals_out[i];
        const uint32_t kind = val >> KIND_SHIFT, idx = val & IDX_MASK;
        if (kind == K_SELF) {
            a.rs[idx] = v;
        } else if (kind == K_NB) {
            rn_out[idx] = v;
        } else if (kind == K_END) {
            a.re[idx] = v;
        } else {
            bool clamped;
            const double w = entry_amount(a, a.flags[idx], a.d[idx], a.hs[idx], a.hsn[idx], a.b[idx], a.r[idx], a.ec[idx], v,
                                          rn_in[idx], a.want[idx], &clamped);
            a.w[idx] = w;
            v = v + w;
        }
    }
}

// 3. the sweeps, in one cooperative launch.  Every block runs the same number
// of grid barriers (the stop test reads a counter all blocks see after the
// same barrier), so the grid leaves the loop together.
__global__ void __launch_bounds__(SWEEP_THREADS) ord_sweeps_kernel(OrdArgs a) {
    cg::grid_group grid = cg::this_grid();
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    __shared__ int any;
    int sweep = 0;
    for (; sweep < a.max_sweeps; sweep++) {
        if (threadIdx.x == 0) any = 0;
        __syncthreads();
        bool ch = false;
        const double* rn = a.rn[sweep & 1];
        for (int64_t p = tid; p < a.P; p += nth) ch |= droplet_pass(a, rn, p);
        if (ch) any = 1;
        __syncthreads();
        if (threadIdx.x == 0 && any) atomicAdd(&a.changes[sweep], 1u);
        grid.sync();
        if (__hip_atomic_load(&a.changes[sweep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        const uint32_t nseg = *a.nseg;
        for (int64_t g = tid; g < (int64_t)nseg; g += nth) cell_scan(a, rn, a.rn[(sweep & 1) ^ 1], (uint32_t)g);
        grid.sync();
    }
    if (tid == 0) {
        atomicMax(&a.status[1], sweep + 1);
        atomicAdd(&a.status[2], sweep + 1);
        if (sweep >= a.max_sweeps) atomicOr(a.status, 2);
    }
}

// 4. per segment: final acc, tracks and stockpiles in serial order, packed record
__device__ void finish_segment(const OrdArgs& a, int64_t g) {
    uint32_t i = a.seg[g];
    const int64_t c = a.keys_out[i];
    double v = a.acc[c];
    bool wrote = false;
    for (; i < (uint32_t)a.n_slots && a.keys_out[i] == (uint32_t)c; i++) {
        const uint32_t val = a.vals_out[i];
        const uint32_t kind = val >> KIND_SHIFT, idx = val & IDX_MASK;
        if (kind == K_WRITE) {
            v = v + a.w[idx];
            wrote = true;
            if (a.flags[idx] & EF_TRACK) {
                const int64_t p = idx % a.P, k = idx / a.P, e = p * a.cap + k;
                a.disch[c] += (double)a.cl_vol[e];
                a.mom[2 * c] += (double)a.cl_mom[2 * e];
                a.mom[2 * c + 1] += (double)a.cl_mom[2 * e + 1];
            }
        } else if (kind == K_END) {
            if (a.pendf[idx]) a.pending[c] += a.pend[idx];
        }
    }
    if (!wrote) return;
    a.acc[c] = v;
    float* rec = a.samp + c * NS;
    rec[S_SURF] = (float)(a.h[c] + a.s[c] + v);
    const double sd = a.s[c] + v;
    rec[S_SED] = (float)(sd > 0.0 ? sd : 0.0);
    if (a.use_route && (double)rec[S_RFLAG] < 1.5) rec[S_RFLAG] = rec[S_ROUTE] != rec[S_SURF] ? 1.0f : 0.0f;
}

__global__ void __launch_bounds__(256) ord_finish_kernel(OrdArgs a) {
    if (a.status[0] != 0) return;
    const int64_t n = (int64_t)*a.nseg;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x)
        finish_segment(a, g);
}

// 5. the chunk's (clamped, to pending, deficit lost, lost offshore), summed in
// droplet order: the block stages up to STAT_TILE droplets' values in LDS
// (coalesced loads), then one lane adds them in order.
constexpr int STAT_TILE = 1024;
__global__ void __launch_bounds__(256) ord_stats_kernel(OrdArgs a) {
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

int sweep_grid(int* blocks) {
    static int cached = 0;
    if (cached == 0) {
        int dev = 0, per_cu = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)ord_sweeps_kernel, SWEEP_THREADS, 0) != hipSuccess)
            return -1;
        int n = per_cu * prop.multiProcessorCount;
        cached = n < 128 ? n : 128;
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
        set_error("pg_ordered_workspace_bytes: hipcub temp-size query failed");
        return -1;
    }
    return (int64_t)L.total;
}

extern "C" int pg_apply_changes_ordered(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol,
                                        const float* cl_mom, const int64_t* cl_count, int64_t P, int64_t cap,
                                        const double* height, const double* sediment, double* acc, double* pending,
                                        float* samp, double* disch_track, double* mom_track, const uint8_t* mask,
                                        int64_t n_cells, double iter_erode, double iter_deposit, double fan_slope,
                                        int use_route, double dep_floor, double offshore_writeoff,
                                        const double* erode_cap, int use_erode_cap, double* out4, int32_t* status,
                                        int max_sweeps, void* work, int64_t work_bytes, pg_stream stream) {
    clear_error();
    PG_REQUIRE(P >= 0 && cap > 0 && n_cells > 0, "pg_apply_changes_ordered: bad sizes P=%lld cap=%lld n_cells=%lld",
               (long long)P, (long long)cap, (long long)n_cells);
    PG_REQUIRE(max_sweeps >= 1, "pg_apply_changes_ordered: max_sweeps must be >= 1");
    PG_REQUIRE(out4 && status, "pg_apply_changes_ordered: out4 / status is NULL");
    PG_REQUIRE(!use_erode_cap || erode_cap, "pg_apply_changes_ordered: use_erode_cap without erode_cap");
    PG_REQUIRE((uint64_t)n_cells < 0xffffffffull, "pg_apply_changes_ordered: %lld cells exceed 32-bit cell ids",
               (long long)n_cells);
    PG_REQUIRE((uint64_t)P * (uint64_t)cap <= IDX_MASK, "pg_apply_changes_ordered: P * cap = %lld exceeds %u entries",
               (long long)(P * cap), IDX_MASK);
    PG_REQUIRE((uint64_t)P * (uint64_t)(3 * cap + 1) < 0x7fffffffull, "pg_apply_changes_ordered: too many accesses");
    Layout L;
    PG_REQUIRE(layout_of(P, cap, max_sweeps, &L) == 0, "pg_apply_changes_ordered: hipcub temp-size query failed");
    PG_REQUIRE(work && work_bytes >= (int64_t)L.total, "pg_apply_changes_ordered: workspace of %lld bytes, %lld needed",
               (long long)work_bytes, (long long)L.total);
    hipStream_t s = (hipStream_t)stream;
    if (P == 0) {
        PG_HIP(hipMemsetAsync(out4, 0, 4 * sizeof(double), s));
        return 0;
    }
    PG_REQUIRE(cl_cell && cl_delta && cl_vol && cl_mom && cl_count && height && sediment && acc && pending && samp &&
                   disch_track && mom_track && mask,
               "pg_apply_changes_ordered: NULL array");
    int nbits = 1;
    while (nbits < 32 && ((uint64_t)1 << nbits) <= (uint64_t)n_cells) nbits++;
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
    a.rn[0] = at<double>(work, L.rn);
    a.rn[1] = at<double>(work, L.rn2);
    a.w = at<double>(work, L.w);
    a.want = at<double>(work, L.want);
    a.death = at<int32_t>(work, L.death);
    a.hsd = at<double>(work, L.hsd);
    a.bd = at<float>(work, L.bd);
    a.re = at<double>(work, L.re);
    a.drop = at<double>(work, L.drop);
    a.pend = at<double>(work, L.pend);
    a.pendf = at<uint8_t>(work, L.pendf);
    a.keys_in = at<uint32_t>(work, L.keys_in);
    a.vals_in = at<uint32_t>(work, L.vals_in);
    a.keys_out = at<uint32_t>(work, L.keys_out);
    a.vals_out = at<uint32_t>(work, L.vals_out);
    a.seg = at<uint32_t>(work, L.seg);
    a.nseg = at<uint32_t>(work, L.nseg);
    a.changes = at<uint32_t>(work, L.changes);
    a.status = status;
    a.hole = (uint32_t)(((uint64_t)1 << nbits) - 1);
    a.n_slots = P * (3 * cap + 1);
    a.max_sweeps = max_sweeps;
    a.out4 = out4;
    hipLaunchKernelGGL(ord_build_kernel, dim3(blocks_for(P, 256)), dim3(256), 0, s, a);
    PG_LAUNCHED("ord_build_kernel");
    size_t tb = L.temp_bytes;
    PG_HIP(hipcub::DeviceRadixSort::SortPairs(at<void>(work, L.temp), tb, a.keys_in, a.keys_out, a.vals_in, a.vals_out,
                                              (int)a.n_slots, 0, nbits, s));
    hipLaunchKernelGGL(ord_segments_kernel, dim3(blocks_for(a.n_slots, 256)), dim3(256), 0, s, a);
    PG_LAUNCHED("ord_segments_kernel");
    int blocks = 0;
    PG_REQUIRE(sweep_grid(&blocks) == 0, "pg_apply_changes_ordered: occupancy query failed");
    void* kargs[] = {&a};
    PG_HIP(hipLaunchCooperativeKernel((const void*)ord_sweeps_kernel, dim3(blocks), dim3(SWEEP_THREADS), kargs, 0, s));
    // the finish pass may touch every segment; its grid covers the largest possible count
    int fblocks = blocks_for(a.n_slots, 256);
    if (fblocks > 512) fblocks = 512;
    hipLaunchKernelGGL(ord_finish_kernel, dim3(fblocks), dim3(256), 0, s, a);
    PG_LAUNCHED("ord_finish_kernel");
    hipLaunchKernelGGL(ord_stats_kernel, dim3(1), dim3(256), 0, s, a);
    PG_LAUNCHED("ord_stats_kernel");
    return 0;
}
