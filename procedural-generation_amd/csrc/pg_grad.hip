// Differentiable mode: the erosion iteration as a function of the surface,
// a per-cell erodibility field and six scalar parameters, with the droplet
// trajectories frozen, and its hand-written adjoint.
//
// Forward (grad_forward_kernel): the fast mode's persistent droplet loop with
// one frozen-terrain chunk per iteration and only the trace-time caps
// (particle.py:596-660: c_eq, k_e, relief / expected-visit caps, the sea
// walk's settling); every droplet's exchange goes straight into the int64
// fixed-point change (no apply-time budget, floors or ceilings).  Two
// surfaces are read:
//   - the steering record (the packed float32 {surface, discharge, momentum}
//     of a nominal state): directions, cells visited, death -- the
//     trajectory;
//   - the exchange surface (float or double): h0, h1, the cell's own height
//     and the previous cell's -- everything the mass exchange depends on.
// Run on one state with both from that state, it is the nominal iteration.
// Run with the steering record of a nominal state, the nominal run's move
// counts (fixed_moves) and a perturbed exchange surface, it is the
// frozen-trajectory map whose derivative the adjoint computes (and that
// finite differences check).
//
// Tape: per droplet, per move, the exchange's inputs (position, cached
// cell scalars, vol, sed, dh, own height) -- enough for the reverse sweep
// to rebuild every quantity of the forward without reading the surface.
//
// Reverse (grad_reverse_kernel): one droplet per lane, moves in reverse.
// With g = dL/d(surface after the particle pass), a move's load change m_k
// (taken from cell_k, laid down where the droplet stops) has
// dL/dm_k = g[stop cell] - g[cell_k]; the adjoints of sed and vol are
// carried backwards through evaporation, the caps and c_eq; surface
// adjoints are scattered (double atomics) onto the bilinear stencils and
// cells the forward read; the erodibility field's gradient onto cell_k; the
// six scalars' gradients are reduced per wave.
#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"

#pragma clang fp contract(fast)

#include "pg_fast_shared.h"

namespace pg {
namespace gradmode {
namespace {

using namespace pg::fastmode;

constexpr int F_LAND = 1, F_SEA = 2, F_H1 = 4, F_PREV = 8;
constexpr int N_THETA = 6;  // erodibility, deposition_rate, evap_rate, k_disc, ocean_rate, dep_floor

template <typename Real>
struct TapeRec {
    int cell, flags;                 // cell of the scalars (clamped to the interior), flags
    float fx, fy, q, base, erod, evap, blend;
    int pos;                         // extended cell of the position (its stencil)
    Real vol, sed, dh, hh;
};

struct TapeDroplet {
    int moves, last, deposited, cause;
    double sed0, pad;
};

struct TapeHead {
    double volume0, conc_scale;
    ll n_total, capacity, stride;
    int real, pad;
    double pad2[2];
};

__host__ __device__ inline size_t up(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename Real>
size_t tape_bytes(ll capacity, ll stride) {
    return up(sizeof(TapeHead)) + up((size_t)capacity * sizeof(TapeDroplet)) +
           (size_t)capacity * (size_t)stride * sizeof(TapeRec<Real>);
}

template <typename Real>
struct Tape {
    TapeHead* head;
    TapeDroplet* drops;
    TapeRec<Real>* recs;
};

template <typename Real>
__host__ __device__ inline Tape<Real> tape_view(char* base, ll capacity) {
    Tape<Real> t;
    t.head = reinterpret_cast<TapeHead*>(base);
    t.drops = reinterpret_cast<TapeDroplet*>(base + up(sizeof(TapeHead)));
    t.recs = reinterpret_cast<TapeRec<Real>*>(base + up(sizeof(TapeHead)) + up((size_t)capacity * sizeof(TapeDroplet)));
    return t;
}

template <typename Real>
struct GradCfg {
    const float4* cell4;  // steering record of the nominal state
    const float4* geo;
    const float2* dyn;    // sediment thickness, base (nominal)
    const uint8_t* mask;
    const Real* surf;     // exchange surface
    const float* erod;    // per-cell erodibility factor (NULL = 1)
    int N, H, NE, spherical, deposit_on_exit;
    const int* d_cell;
    const float *d_fx, *d_fy, *d_sa, *d_sb;
    const ll* d_load;
    Scalars* sc;
    // steering / fixed parameters (float32, as the fast mode)
    float dt, density, fr, k_mom, mom_scale, inv_sat, disc_exponent, slope_gain, slope_sat, cover_depth, max_erode,
        fan_room, min_volume_frac, dt_evap_steer;
    int max_steps, pit_steps, ocean_steps;
    // the differentiable scalars
    Real erodibility, rate, evap_rate, k_disc, ocean_rate, dep_floor;
    double hscale, tscale, conc_chunk;
    ll *acc_fx, *dis_fx, *mom_fx, *stats;  // momentum track b one plane (n_cells) after a
    int64_t n_cells;
    const int* fixed_moves;
    int* moves_out;
    char* tape;
    ll tape_capacity, tape_stride;
    ll d_lo, d_hi;  // the droplets traced: [d_lo, d_hi) of the spawn (d_hi <= d_lo: all); droplet d_lo + i is
                    // tape slot i
    int apply;      // 0: record the tape only (a range), no cell accumulators
};

template <typename Real>
__device__ __forceinline__ Real disc_factor(float q, float inv_sat, float disc_exponent) {
    const Real x = (Real)q * (Real)inv_sat;
    if (inv_sat <= 0.0f) return (Real)0;
    if (disc_exponent == 0.5f) return sqrt(x);
    if (disc_exponent > 0.0f) return pow(x, (Real)disc_exponent);
    return erf(x);
}

// the land exchange of one move (particle.py:587-660 without lakes and
// lateral cuts): returns cdiff
template <typename Real>
__device__ __forceinline__ Real land_exchange(Real sed, Real dh, Real hh, Real h_prev, bool has_prev, Real vr, float q,
                                              float erod, float blend, Real conc_scale, Real erodibility, Real rate,
                                              Real k_disc, Real dep_floor, float inv_sat, float disc_exponent,
                                              float max_erode) {
    const Real qe = disc_factor<Real>(q, inv_sat, disc_exponent);
    const Real c_raw = erodibility * (Real)erod * dh * ((Real)1 + k_disc * qe);
    const Real c_eq = c_raw > (Real)0 ? c_raw : (Real)0;
    const Real k_e = c_eq > sed ? rate * (Real)blend : rate;
    const Real u = k_e * (c_eq - sed);
    const Real kexp = (Real)1 + (Real)q * conc_scale;
    if (u > (Real)0) {
        Real ecap = dh / kexp;
        if (ecap > (Real)max_erode) ecap = (Real)max_erode;
        if (ecap > hh) ecap = hh > (Real)0 ? hh : (Real)0;
        if (u * vr > ecap) return ecap / vr;
    } else if (u < (Real)0) {
        Real room = has_prev ? h_prev - hh : dep_floor;
        if (room < dep_floor) room = dep_floor;
        const Real dcap = room / kexp;
        if (-u * vr > dcap) return -dcap / vr;
    }
    return u;
}

// the seafloor walk's settling (particle.py:562-585)
template <typename Real>
__device__ __forceinline__ Real sea_exchange(Real sed, Real hh, Real h_prev, bool has_prev, Real vr, float b_here,
                                             Real ocean_rate, Real dep_floor, float fan_room) {
    const Real u = -ocean_rate * sed;
    Real room = has_prev ? h_prev - hh : (Real)fan_room;
    if (room < (Real)fan_room) room = (Real)fan_room;
    const Real wl = (Real)b_here - dep_floor - hh;
    if (wl < room) room = wl;
    if (room < (Real)0) room = (Real)0;
    if (-u * vr > room) return -room / vr;
    return u;
}

template <typename Real>
__device__ __forceinline__ Real bilerp(const Real* surf, int64_t b, int NE, float wi, float wj) {
    const Real a = (Real)wi, c = (Real)wj;
    return surf[b] * ((Real)1 - a) * ((Real)1 - c) + surf[b + NE] * a * ((Real)1 - c) + surf[b + 1] * ((Real)1 - a) * c +
           surf[b + NE + 1] * a * c;
}

template <typename Real>
__global__ void __launch_bounds__(TRACE_TPB) grad_forward_kernel(GradCfg<Real> k) {
    const int lane = threadIdx.x & 63;
    // the spawn's droplet count sets the trace caps' concentration; a range traces only its droplets
    const ll n_spawn = k.sc->n_total;
    const ll d_lo = k.d_hi > k.d_lo ? (k.d_lo < n_spawn ? k.d_lo : n_spawn) : 0;
    const ll n_total = k.d_hi > k.d_lo ? (k.d_hi < n_spawn ? k.d_hi : n_spawn) - d_lo : n_spawn;
    const double volume0 = k.sc->volume0;
    if (volume0 <= 0.0) return;
    const float inv_v0 = (float)(1.0 / volume0);
    const float min_volume = (float)(k.min_volume_frac * volume0);
    const Real conc_scale = (Real)((k.conc_chunk > 0.0 ? k.conc_chunk : (double)n_spawn) / (volume0 * (double)n_spawn));
    const Real v0 = (Real)volume0;
    const float v0f = (float)volume0;
    ll* queue = &k.sc->queue[0];
    const int NE2 = k.NE * k.NE;
    Tape<Real> tp;
    const bool taping = k.tape != nullptr;
    if (taping) {
        tp = tape_view<Real>(k.tape, k.tape_capacity);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            tp.head->volume0 = volume0;
            tp.head->conc_scale = (double)conc_scale;
            tp.head->n_total = n_total;
            tp.head->capacity = k.tape_capacity;
            tp.head->stride = k.tape_stride;
            tp.head->real = sizeof(Real) == 8 ? 1 : 0;
        }
    }
    const bool replay = k.fixed_moves != nullptr;

    bool alive = false, exhausted = false;
    ll d = 0;
    int f = 0, ix = 0, iy = 0, fixed = -1, fixed_cause = 0;
    float fx = 0, fy = 0, vol_s = 0, sa = 0, sb = 0;
    Real vol = 0, sed = 0, h_prev = 0;
    bool has_prev = false, in_sea = false;
    int step = 0, uphill = 0, sea_steps = 0, pushes = 0, last_cell = -1;
    ll real = 0;
    Stencil S;
    int64_t cell = 0;
    int m = 0;
    float b_here = 0, sthick = 0, hs_here = 0;
    float4 g0, g1;
    ll n_entries = 0, n_moves = 0, exit_fx = 0, vanished_fx = 0;
    ll deaths[N_DEATHS] = {0, 0, 0, 0, 0, 0};

    auto load_cell = [&](int ff, int cx, int cy) {
        const int ci = clampi(cx, 0, k.N - 1), cj = clampi(cy, 0, k.N - 1);
        cell = ((int64_t)ff * k.NE + ci + k.H) * k.NE + cj + k.H;
        m = k.mask[cell];
        hs_here = k.cell4[cell].x;
        const float2 dy = k.dyn[cell];
        sthick = dy.x;
        b_here = dy.y;
        g0 = k.geo[2 * cell];
        g1 = k.geo[2 * cell + 1];
    };

    while (true) {
        if (!exhausted) {
            const ull dead = __ballot(!alive);
            if (dead) {
                const int leader = __ffsll((ll)dead) - 1;
                ll base = 0;
                if (lane == leader) base = (ll)atomicAdd((ull*)queue, (ull)__popcll(dead));
                base = __shfl(base, leader);
                if (base >= n_total) exhausted = true;
                if (!alive) {
                    d = base + __popcll(dead & ((1ull << lane) - 1));
                    if (d < n_total) {
                        const ll dd = d_lo + d;  // the droplet's index in the spawn (d: its tape slot)
                        const int c0 = k.d_cell[dd];
                        f = c0 / NE2;
                        const int r = c0 - f * NE2;
                        ix = r / k.NE - k.H;
                        iy = r % k.NE - k.H;
                        fx = k.d_fx[dd];
                        fy = k.d_fy[dd];
                        real = k.d_load[dd];
                        vol_s = v0f;
                        vol = v0;
                        sed = (Real)((double)real / k.hscale);
                        sa = k.d_sa[dd];
                        sb = k.d_sb[dd];
                        has_prev = false;
                        in_sea = false;
                        step = uphill = sea_steps = pushes = 0;
                        last_cell = -1;
                        if (replay) {
                            fixed = k.fixed_moves[dd] & 0xffffff;
                            fixed_cause = (k.fixed_moves[dd] >> 24) & 0xff;
                        }
                        if (taping && d < k.tape_capacity) tp.drops[d].sed0 = (double)sed;
                        S = load_stencil(k.cell4, k.NE, k.H, f, ix, iy, fx, fy);
                        load_cell(f, ix, iy);
                        alive = true;
                    }
                }
            }
        }
        if (!__any(alive)) break;
        if (!alive) continue;

        int cause = -1;
        bool exited = false;
        if (replay && step == fixed) {
            cause = fixed_cause;
            exited = cause == DEATH_EXIT;
        } else if (!replay && step >= k.max_steps) {
            cause = DEATH_AGE;
        } else if (m == MASK_OUTSIDE) {
            cause = DEATH_EXIT;
            exited = true;
        } else {
            if (step == 0 && hs_here < b_here) in_sea = true;
            if (in_sea) {
                if (!replay &&
                    (hs_here >= b_here || (double)sed * (double)vol / volume0 < 1e-4 || sea_steps >= k.ocean_steps))
                    cause = DEATH_OCEAN;
                else
                    sea_steps += 1;
            }
            if (cause < 0) {
                if (m == MASK_ACTIVE) {
                    last_cell = (int)cell;
                    pushes += 1;
                }
                const float q = fmaxf(lerp2(S.c00.y, S.c10.y, S.c01.y, S.c11.y, S.wi, S.wj), 0.0f);
                int nf, nix, niy;
                float nfx, nfy, h0s;
                cause = steer_move<false>(k, k.N, S, g0, g1, q, vol_s, vol_s * inv_v0, sa, sb, f, ix, iy, fx, fy, nf, nix, niy, nfx,
                                   nfy, h0s);
                exited = cause == DEATH_EXIT;
                if (cause < 0) {
                    const int nei = nix + k.H, nej = niy + k.H;
                    const int64_t ncell = ((int64_t)nf * k.NE + nei) * k.NE + nej;
                    const Stencil S1 = load_stencil(k.cell4, k.NE, k.H, nf, nix, niy, nfx, nfy);
                    const int nm = k.mask[ncell];
                    const float2 ndy = k.dyn[ncell];
                    const float4 nrec = corner_of(S1, nei, nej);
                    if (nm == MASK_OUTSIDE) {
                        cause = DEATH_EXIT;
                        exited = true;
                    } else {
                        const float h1s = lerp2(S1.c00.x, S1.c10.x, S1.c01.x, S1.c11.x, S1.wi, S1.wj);
                        // the exchange, on the exchange surface
                        const Real h0 = bilerp(k.surf, ((int64_t)f * k.NE + S.i0) * k.NE + S.j0, k.NE, S.wi, S.wj);
                        const Real h1 = bilerp(k.surf, ((int64_t)nf * k.NE + S1.i0) * k.NE + S1.j0, k.NE, S1.wi, S1.wj);
                        const Real b1 = (Real)ndy.y;
                        const Real dh = h0 - (h1 > b1 ? h1 : b1);
                        const Real hh = k.surf[cell];
                        const Real vr = vol / v0;
                        const float erod = k.erod ? k.erod[cell] : 1.0f;
                        float blend = 1.0f;
                        int flags = (h1 > b1 ? F_H1 : 0) | (has_prev ? F_PREV : 0);
                        Real cdiff = 0;
                        if (m == MASK_ACTIVE && in_sea) {
                            cdiff = sea_exchange<Real>(sed, hh, h_prev, has_prev, vr, b_here, k.ocean_rate, k.dep_floor,
                                                       k.fan_room);
                            flags |= F_SEA;
                        } else if (m == MASK_ACTIVE) {
                            const float w_cover = k.cover_depth > 0.0f ? fminf(sthick / k.cover_depth, 1.0f)
                                                                       : (sthick > 0.0f ? 1.0f : 0.0f);
                            blend = w_cover + (1.0f - w_cover) * (1.0f - g1.w);
                            cdiff = land_exchange<Real>(sed, dh, hh, h_prev, has_prev, vr, q, erod, blend, conc_scale,
                                                        k.erodibility, k.rate, k.k_disc, k.dep_floor, k.inv_sat,
                                                        k.disc_exponent, k.max_erode);
                            flags |= F_LAND;
                            if (!replay && k.apply) {
                                atomicAdd((ull*)&k.dis_fx[cell], (ull)llrint((double)vol_s * k.tscale));
                                atomicAdd((ull*)&k.mom_fx[cell], (ull)llrint((double)(vol_s * sa) * k.tscale));
                                atomicAdd((ull*)&k.mom_fx[k.n_cells + cell], (ull)llrint((double)(vol_s * sb) * k.tscale));
                            }
                        }
                        if (taping && d < k.tape_capacity) {
                            TapeRec<Real> rc;
                            rc.cell = (int)cell;
                            rc.flags = flags;
                            rc.fx = fx;
                            rc.fy = fy;
                            rc.q = q;
                            rc.base = b_here;
                            rc.erod = erod;
                            rc.evap = g1.z;
                            rc.blend = blend;
                            rc.pos = (int)(((int64_t)f * k.NE + ix + k.H) * k.NE + iy + k.H);
                            rc.vol = vol;
                            rc.sed = sed;
                            rc.dh = dh;
                            rc.hh = hh;
                            tp.recs[d * k.tape_stride + step] = rc;
                        }
                        if (flags & (F_LAND | F_SEA)) {
                            const ll dq = llrint((double)(cdiff * vr) * k.hscale);  // > 0: taken from the cell
                            if (dq && k.apply) atomicAdd((ull*)&k.acc_fx[cell], (ull)(-dq));
                            real += dq;
                            sed += cdiff;
                            n_entries += 1;
                        }
                        // evaporation: the steering volume with the nominal rate, the
                        // exchange's with the differentiable one
                        const float ev_s = fmaxf(1.0f - k.dt_evap_steer * g1.z, 0.01f);
                        vol_s *= ev_s;
                        Real ev = (Real)1 - (Real)k.dt * k.evap_rate * (Real)g1.z;
                        if (ev < (Real)0.01) ev = (Real)0.01;
                        vol *= ev;
                        sed /= ev;
                        f = nf;
                        ix = nix;
                        iy = niy;
                        fx = nfx;
                        fy = nfy;
                        S = S1;
                        load_cell(nf, nix, niy);
                        h_prev = hh;
                        has_prev = true;
                        n_moves += 1;
                        step += 1;
                        if (nrec.x < ndy.y && !in_sea) {
                            in_sea = true;
                            if (!replay && sed <= (Real)0) cause = DEATH_OCEAN;
                        }
                        if (!replay) {
                            if (cause < 0 && vol_s < min_volume) cause = in_sea ? DEATH_OCEAN : DEATH_EVAP;
                            if (cause < 0) {
                                if (h0s - h1s < 0.0f) {
                                    uphill += 1;
                                    if (uphill > k.pit_steps) cause = in_sea ? DEATH_OCEAN : DEATH_PIT;
                                } else {
                                    uphill = 0;
                                }
                            }
                            if (cause < 0 && step >= k.max_steps) cause = DEATH_AGE;
                        }
                    }
                }
            }
        }
        if (cause < 0) continue;

        // ---- the droplet stopped: its load is laid down where it stopped
        if (!replay && in_sea && (cause == DEATH_AGE || cause == DEATH_STOP)) cause = DEATH_OCEAN;
        deaths[cause] += 1;
        if (!exited) {
            if (m == MASK_ACTIVE && !(in_sea && hs_here >= b_here)) {
                if (pushes == 0 || last_cell != (int)cell) {
                    last_cell = (int)cell;
                    pushes += 1;
                }
            }
        }
        int deposited = 0;
        if (pushes > 0 && (!exited || k.deposit_on_exit)) {
            if (real && k.apply) atomicAdd((ull*)&k.acc_fx[last_cell], (ull)real);
            deposited = 1;
        } else if (exited) {
            exit_fx += real;
        } else {
            vanished_fx += real;
        }
        if (k.moves_out) k.moves_out[d_lo + d] = step | (cause << 24);
        if (taping && d < k.tape_capacity) {
            TapeDroplet td;
            td.moves = step;
            td.last = last_cell;
            td.deposited = deposited;
            td.cause = cause;
            td.sed0 = tp.drops[d].sed0;
            td.pad = 0.0;
            tp.drops[d] = td;
            // the stop position closes the last move's stencil
            TapeRec<Real> rc = {};
            rc.cell = (int)cell;
            rc.pos = (int)(((int64_t)f * k.NE + ix + k.H) * k.NE + iy + k.H);
            rc.fx = fx;
            rc.fy = fy;
            rc.vol = vol;
            rc.sed = sed;
            tp.recs[d * k.tape_stride + step] = rc;
        }
        real = 0;
        alive = false;
    }
    ll v[4 + N_DEATHS] = {n_entries, n_moves, exit_fx, vanished_fx};
    for (int c = 0; c < N_DEATHS; c++) v[4 + c] = deaths[c];
    for (int c = 0; c < 4 + N_DEATHS; c++) v[c] = wave_sum(v[c]);
    if (lane == 0) {
        const int slot[4] = {PG_FS_ENTRIES, PG_FS_MOVES, PG_FS_EXIT_FX, PG_FS_VANISHED_FX};
        for (int c = 0; c < 4; c++)
            if (v[c]) atomicAdd((ull*)&k.stats[slot[c]], (ull)v[c]);
        for (int c = 0; c < N_DEATHS; c++)
            if (v[4 + c]) atomicAdd((ull*)&k.stats[PG_FS_DEATHS + c], (ull)v[4 + c]);
    }
}

// ---------------------------------------------------------------------------
// reverse sweep
// ---------------------------------------------------------------------------
// Deterministic accumulation: every contribution is added as a signed 128-bit
// fixed-point number of unit 2^-64 (two 64-bit integer atomics with the
// carry), so the sum does not depend on the order the lanes arrive in; each
// contribution is truncated toward zero at 2^-64, |contribution| < 2^62.
struct Fx128 {
    ull lo;
    ll hi;
};

// v as a two's-complement 128-bit integer (H, L) of unit 2^-64
__device__ __forceinline__ void fx_split(double v, ull& H, ull& L) {
    const double m = fabs(v);
    const double fh = floor(m);
    H = (ull)fh;
    L = (ull)((m - fh) * 18446744073709551616.0);  // m - fh is exact; the scaling by 2^64 too
    if (v < 0.0) {  // two's complement of (H, L)
        L = ~L + 1ull;
        H = ~H + (L == 0ull ? 1ull : 0ull);
    }
}

// (H, L) += (h, l) modulo 2^128: integer sums, so any grouping gives the same bits
__device__ __forceinline__ void fx_acc(ull& H, ull& L, ull h, ull l) {
    const ull s = L + l;
    H += h + (s < L ? 1ull : 0ull);
    L = s;
}

__device__ __forceinline__ void fx_add_raw(Fx128* a, ull H, ull L) {
    const ull old = L ? atomicAdd(&a->lo, L) : 0ull;
    const ull carry = old + L < old ? 1ull : 0ull;
    if (H + carry) atomicAdd((ull*)&a->hi, H + carry);
}

__device__ __forceinline__ void fx_add(Fx128* a, double v) {
    if (v == 0.0) return;
    ull H, L;
    fx_split(v, H, L);
    fx_add_raw(a, H, L);
}

// the accumulated sum, in units of 1: hi + lo 2^-64
__device__ __forceinline__ double fx_value(const Fx128& a) {
    return (double)a.hi + (double)a.lo * 5.421010862427522e-20;
}

// the reverse sweep's accumulators: surface and erodibility adjoints per
// extended cell, the six scalars (and two spare)
int64_t reverse_work_bytes(int64_t n) { return (2 * n + 8) * (int64_t)sizeof(Fx128); }

template <typename Real>
struct RevCfg {
    const char* tape;
    ll capacity;
    const double* g;   // dL/d(surface after the particle pass), extended cells
    Fx128* lam;        // the particle pass's share of dL/d(surface before it)
    Fx128* grad_erod;  // dL/d(erodibility factor), extended cells
    Fx128* grad_theta;
    int N, H, NE, feedback;
    float dt, inv_sat, disc_exponent, max_erode, fan_room;
    Real erodibility, rate, evap_rate, k_disc, ocean_rate, dep_floor;
};

__device__ __forceinline__ void add_d(Fx128* p, double v) { fx_add(p, v); }

__global__ void queue_reset_kernel(Scalars* sc) {
    if (threadIdx.x < 2) sc->queue[threadIdx.x] = 0;
}

__global__ void fx_zero_kernel(int64_t n, Fx128* acc) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        acc[c].lo = 0ull;
        acc[c].hi = 0;
    }
}

// out[c] += the accumulated sum
__global__ void fx_flush_kernel(int64_t n, const Fx128* acc, double* out) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const Fx128 a = acc[c];
        if (a.lo != 0ull || a.hi != 0) out[c] += fx_value(a);
    }
}

// scatter `st` over the bilinear stencil of a taped position and `ca` onto its cell
template <typename Real>
__device__ __forceinline__ void scatter(const RevCfg<Real>& r, const TapeRec<Real>& rec, double st, double ca) {
    const int NE2 = r.NE * r.NE;
    const int cell = rec.cell;
    const int f = rec.pos / NE2, rr = rec.pos - f * NE2;
    const int ix = rr / r.NE - r.H, iy = rr % r.NE - r.H;
    int i0, j0;
    float wi, wj;
    stencil_at(r.NE, r.H, ix, iy, rec.fx, rec.fy, i0, j0, wi, wj);
    const int64_t b = ((int64_t)f * r.NE + i0) * r.NE + j0;
    const double a = wi, c = wj;
    double v00 = st * (1.0 - a) * (1.0 - c), v10 = st * a * (1.0 - c), v01 = st * (1.0 - a) * c, v11 = st * a * c;
    if (cell == b) v00 += ca;
    else if (cell == b + r.NE) v10 += ca;
    else if (cell == b + 1) v01 += ca;
    else if (cell == b + r.NE + 1) v11 += ca;
    else add_d(&r.lam[cell], ca);
    add_d(&r.lam[b], v00);
    add_d(&r.lam[b + r.NE], v10);
    add_d(&r.lam[b + 1], v01);
    add_d(&r.lam[b + r.NE + 1], v11);
}

template <typename Real>
__global__ void __launch_bounds__(TPB) grad_reverse_kernel(RevCfg<Real> r) {
    const Tape<Real> tp = tape_view<Real>(const_cast<char*>(r.tape), r.capacity);
    const ll n_total = tp.head->n_total < r.capacity ? tp.head->n_total : r.capacity;
    const ll stride = tp.head->stride;
    const Real v0 = (Real)tp.head->volume0, conc_scale = (Real)tp.head->conc_scale;
    // the scalars' adjoints: summed per droplet in Real, then as 128-bit integers (fixed point), so the
    // totals do not depend on which droplets share a lane or a wave (a tape in pieces gives the same bits)
    ull tH[N_THETA] = {0, 0, 0, 0, 0, 0}, tL[N_THETA] = {0, 0, 0, 0, 0, 0};
    for (ll d = (ll)blockIdx.x * blockDim.x + threadIdx.x; d < n_total; d += (ll)gridDim.x * blockDim.x) {
        const TapeDroplet td = tp.drops[d];
        const int K = td.moves;
        if (K <= 0) continue;
        Real gt[N_THETA] = {0, 0, 0, 0, 0, 0};
        const TapeRec<Real>* rec = tp.recs + d * stride;
        const Real mu_final = td.deposited ? (Real)r.g[td.last] : (Real)0;
        Real a_sed = 0, a_vol = 0;
        double st_adj = 0.0, cell_adj = 0.0, prev_adj = 0.0;
        TapeRec<Real> rn = rec[K];
        TapeRec<Real> rc = rec[K - 1];
        for (int k = K - 1; k >= 0; k--) {
            const TapeRec<Real> rp = k > 0 ? rec[k - 1] : rc;
            const bool has_prev = (rc.flags & F_PREV) != 0;
            const Real h_prev = has_prev ? rp.hh : (Real)0;
            const Real vol_k = rc.vol, sed_k = rc.sed, sed_next = rn.sed;
            const Real ev_raw = (Real)1 - (Real)r.dt * r.evap_rate * (Real)rc.evap;
            const Real ev = ev_raw < (Real)0.01 ? (Real)0.01 : ev_raw;
            Real a_dh = 0, a_hh = 0, a_hprev = 0, a_sed_k = a_sed / ev, a_vr = 0;
            if (rc.flags & (F_LAND | F_SEA)) {
                const Real vr = vol_k / v0;
                const Real mu = mu_final - (Real)r.g[rc.cell];
                const Real hh = rc.hh, dh = rc.dh;
                Real cdiff;
                if (rc.flags & F_SEA)
                    cdiff = sea_exchange<Real>(sed_k, hh, h_prev, has_prev, vr, rc.base, r.ocean_rate, r.dep_floor,
                                               r.fan_room);
                else
                    cdiff = land_exchange<Real>(sed_k, dh, hh, h_prev, has_prev, vr, rc.q, rc.erod, rc.blend, conc_scale,
                                                r.erodibility, r.rate, r.k_disc, r.dep_floor, r.inv_sat,
                                                r.disc_exponent, r.max_erode);
                const Real A_c = mu * vr + a_sed / ev;  // adjoint of cdiff
                a_vr += mu * cdiff;
                if (rc.flags & F_SEA) {
                    const Real u = -r.ocean_rate * sed_k;
                    Real room0 = has_prev ? h_prev - hh : (Real)r.fan_room;
                    const bool from_prev = has_prev && room0 > (Real)r.fan_room;
                    const Real r1 = from_prev ? room0 : (Real)r.fan_room;
                    const Real wl = (Real)rc.base - r.dep_floor - hh;
                    const bool from_wl = wl < r1;
                    const Real r2 = from_wl ? wl : r1;
                    const Real r3 = r2 > (Real)0 ? r2 : (Real)0;
                    if (-u * vr > r3) {
                        a_vr += A_c * r3 / (vr * vr);
                        const Real A_r = r2 > (Real)0 ? -A_c / vr : (Real)0;
                        if (from_wl) {
                            gt[5] -= A_r;
                            a_hh -= A_r;
                        } else if (from_prev) {
                            a_hprev += A_r;
                            a_hh -= A_r;
                        }
                    } else {
                        a_sed_k += -r.ocean_rate * A_c;
                        gt[4] += -sed_k * A_c;
                    }
                } else {
                    const Real qe = disc_factor<Real>(rc.q, r.inv_sat, r.disc_exponent);
                    const Real amp = (Real)1 + r.k_disc * qe;
                    const Real c_raw = r.erodibility * (Real)rc.erod * dh * amp;
                    const Real c_eq = c_raw > (Real)0 ? c_raw : (Real)0;
                    const bool ero_br = c_eq > sed_k;
                    const Real kb = ero_br ? (Real)rc.blend : (Real)1;
                    const Real k_e = r.rate * kb;
                    const Real u = k_e * (c_eq - sed_k);
                    const Real kexp = (Real)1 + (Real)rc.q * conc_scale;
                    bool capped = false;
                    if (u > (Real)0) {
                        const Real e1 = dh / kexp;
                        Real ecap = e1;
                        int src = 0;  // 0: relief, 1: max_erode, 2: own height, 3: zero
                        if (ecap > (Real)r.max_erode) {
                            ecap = (Real)r.max_erode;
                            src = 1;
                        }
                        if (ecap > hh) {
                            ecap = hh > (Real)0 ? hh : (Real)0;
                            src = hh > (Real)0 ? 2 : 3;
                        }
                        if (u * vr > ecap) {
                            capped = true;
                            a_vr += -A_c * ecap / (vr * vr);
                            const Real A_e = A_c / vr;
                            if (src == 0) a_dh += A_e / kexp;
                            else if (src == 2) a_hh += A_e;
                        }
                    } else if (u < (Real)0) {
                        const Real room0 = has_prev ? h_prev - hh : r.dep_floor;
                        const bool from_prev = has_prev && room0 > r.dep_floor;
                        const Real room = from_prev ? room0 : r.dep_floor;
                        const Real dcap = room / kexp;
                        if (-u * vr > dcap) {
                            capped = true;
                            a_vr += A_c * dcap / (vr * vr);
                            const Real A_room = -A_c / (vr * kexp);
                            if (from_prev) {
                                a_hprev += A_room;
                                a_hh -= A_room;
                            } else {
                                gt[5] += A_room;
                            }
                        }
                    }
                    if (!capped) {
                        a_sed_k += -k_e * A_c;
                        gt[1] += A_c * (c_eq - sed_k) * kb;
                        if (c_raw > (Real)0) {
                            const Real A_ceq = k_e * A_c;
                            gt[0] += A_ceq * (Real)rc.erod * dh * amp;
                            gt[3] += A_ceq * r.erodibility * (Real)rc.erod * dh * qe;
                            a_dh += A_ceq * r.erodibility * (Real)rc.erod * amp;
                            add_d(&r.grad_erod[rc.cell], (double)(A_ceq * r.erodibility * dh * amp));
                        }
                    }
                }
            }
            // evaporation: sed_{k+1} = (sed_k + cdiff) / ev, vol_{k+1} = vol_k ev
            const Real a_ev = -a_sed * sed_next / ev + a_vol * vol_k;
            if (ev_raw > (Real)0.01) gt[2] += a_ev * (-(Real)r.dt * (Real)rc.evap);
            const Real a_vol_k = a_vol * ev + a_vr / v0;
            // surface adjoints: h1 of this move sits on the next position's stencil
            const double a_h1 = (rc.flags & F_H1) ? -(double)a_dh : 0.0;
            if (r.feedback) scatter(r, rn, st_adj + a_h1, cell_adj);
            st_adj = (double)a_dh;
            cell_adj = (double)a_hh + prev_adj;
            prev_adj = (double)a_hprev;
            a_sed = a_sed_k;
            a_vol = a_vol_k;
            rn = rc;
            rc = rp;
        }
        if (r.feedback) scatter(r, rn, st_adj, cell_adj);
        for (int c = 0; c < N_THETA; c++) {
            ull h, l;
            fx_split((double)gt[c], h, l);
            fx_acc(tH[c], tL[c], h, l);
        }
    }
    for (int c = 0; c < N_THETA; c++) {
        ull h = tH[c], l = tL[c];
        for (int o = 32; o > 0; o >>= 1) fx_acc(h, l, __shfl_xor(h, o), __shfl_xor(l, o));
        if ((threadIdx.x & 63) == 0 && (h | l)) fx_add_raw(&r.grad_theta[c], h, l);
    }
}

template <typename Real>
int launch_forward(const pg_fast_fields* steer, const pg_trace_params* prm, const pg_fast_params* fp,
                   const pg_grad_args* ga, void* work, Ctx& c) {
    GradCfg<Real> k;
    k.cell4 = c.cell4;
    k.geo = at<float4>(work, c.L.geo);
    k.dyn = c.dyn;
    k.mask = steer->mask;
    k.surf = reinterpret_cast<const Real*>(ga->surf);
    k.erod = ga->erod;
    k.N = (int)c.N;
    k.H = (int)c.H;
    k.NE = (int)c.NE;
    k.spherical = (int)steer->spherical;
    k.deposit_on_exit = (int)steer->deposit_on_exit;
    k.d_cell = at<int>(work, c.L.d_cell);
    k.d_fx = at<float>(work, c.L.d_fx);
    k.d_fy = at<float>(work, c.L.d_fy);
    k.d_sa = at<float>(work, c.L.d_sa);
    k.d_sb = at<float>(work, c.L.d_sb);
    k.d_load = at<ll>(work, c.L.d_load);
    k.sc = c.sc;
    k.dt = (float)prm->dt;
    k.density = (float)prm->density;
    const double fr = 1.0 - prm->dt * prm->friction;
    k.fr = (float)(fr < 0.0 ? 0.0 : fr);
    k.k_mom = (float)prm->k_mom;
    k.mom_scale = (float)prm->mom_scale;
    k.inv_sat = (float)(prm->disc_saturation > 0.0 ? 1.0 / prm->disc_saturation : 0.0);
    k.disc_exponent = (float)prm->disc_exponent;
    k.slope_gain = (float)prm->slope_gain;
    k.slope_sat = (float)prm->slope_sat;
    k.cover_depth = (float)prm->cover_depth;
    k.max_erode = (float)prm->max_erode;
    k.fan_room = (float)prm->fan_room;
    k.min_volume_frac = (float)fp->min_volume_frac;
    k.dt_evap_steer = (float)(prm->dt * prm->evap_rate);
    k.max_steps = (int)prm->max_steps;
    k.pit_steps = (int)prm->pit_steps;
    k.ocean_steps = (int)prm->ocean_steps;
    k.erodibility = (Real)ga->theta[0];
    k.rate = (Real)ga->theta[1];
    k.evap_rate = (Real)ga->theta[2];
    k.k_disc = (Real)ga->theta[3];
    k.ocean_rate = (Real)ga->theta[4];
    k.dep_floor = (Real)ga->theta[5];
    k.hscale = c.hscale;
    k.tscale = c.tscale;
    k.conc_chunk = fp->conc_chunk;
    k.acc_fx = at<ll>(work, c.L.acc_fx);
    k.dis_fx = at<ll>(work, c.L.dis_fx);
    k.mom_fx = at<ll>(work, c.L.mom_fx);
    k.n_cells = c.n;
    k.stats = c.st;
    k.fixed_moves = ga->fixed_moves;
    k.moves_out = ga->moves_out;
    k.tape = reinterpret_cast<char*>(ga->tape);
    k.tape_stride = (ll)prm->max_steps + 1;
    k.tape_capacity = 0;
    if (ga->tape) {
        // capacity from the size handed over
        const size_t per = sizeof(TapeDroplet) + (size_t)k.tape_stride * sizeof(TapeRec<Real>);
        ll cap = (ll)(((size_t)ga->tape_bytes - up(sizeof(TapeHead))) / per);
        while (cap > 0 && tape_bytes<Real>(cap, k.tape_stride) > (size_t)ga->tape_bytes) cap--;
        const ll want = ga->d_hi > ga->d_lo ? (ll)(ga->d_hi - ga->d_lo) : (ll)c.P;
        PG_REQUIRE(cap >= want, "grad mode: a tape of %lld bytes holds %lld droplets, %lld are traced",
                   (long long)ga->tape_bytes, (long long)cap, (long long)want);
        k.tape_capacity = cap;
    }
    PG_REQUIRE(ga->d_lo >= 0, "grad mode: d_lo < 0");
    k.d_lo = ga->d_lo;
    k.d_hi = ga->d_hi;
    k.apply = ga->d_hi > ga->d_lo ? 0 : 1;
    int grid = 0;
    if (persistent_grid((const void*)grad_forward_kernel<Real>, TRACE_TPB, fp->grid_blocks, &grid)) return -1;
    hipLaunchKernelGGL(grad_forward_kernel<Real>, dim3(grid), dim3(TRACE_TPB), 0, c.s, k);
    PG_LAUNCHED("grad_forward_kernel");
    return 0;
}

template <typename Real>
int launch_reverse(int64_t F, int64_t N, int64_t H, const pg_trace_params* prm, const pg_grad_args* ga, const double* g,
                   double* lam, double* grad_erod, double* grad_theta, hipStream_t s) {
    RevCfg<Real> r;
    const ll stride = (ll)prm->max_steps + 1;
    const size_t per = sizeof(TapeDroplet) + (size_t)stride * sizeof(TapeRec<Real>);
    ll cap = (ll)(((size_t)ga->tape_bytes - up(sizeof(TapeHead))) / per);
    while (cap > 0 && tape_bytes<Real>(cap, stride) > (size_t)ga->tape_bytes) cap--;
    PG_REQUIRE(cap > 0, "pg_grad_reverse: tape too small");
    const int64_t n = F * (N + 2 * H) * (N + 2 * H);
    PG_REQUIRE(ga->rev_work && ga->rev_work_bytes >= reverse_work_bytes(n),
               "pg_grad_reverse: reverse workspace of %lld bytes, need %lld", (long long)ga->rev_work_bytes,
               (long long)reverse_work_bytes(n));
    Fx128* acc = reinterpret_cast<Fx128*>(ga->rev_work);
    r.tape = reinterpret_cast<const char*>(ga->tape);
    r.capacity = cap;
    r.g = g;
    r.lam = acc;
    r.grad_erod = acc + n;
    r.grad_theta = acc + 2 * n;
    r.N = (int)N;
    r.H = (int)H;
    r.NE = (int)(N + 2 * H);
    r.feedback = ga->feedback != 0;
    r.dt = (float)prm->dt;
    r.inv_sat = (float)(prm->disc_saturation > 0.0 ? 1.0 / prm->disc_saturation : 0.0);
    r.disc_exponent = (float)prm->disc_exponent;
    r.max_erode = (float)prm->max_erode;
    r.fan_room = (float)prm->fan_room;
    r.erodibility = (Real)ga->theta[0];
    r.rate = (Real)ga->theta[1];
    r.evap_rate = (Real)ga->theta[2];
    r.k_disc = (Real)ga->theta[3];
    r.ocean_rate = (Real)ga->theta[4];
    r.dep_floor = (Real)ga->theta[5];
    // rev_flags bit 0: keep the accumulators of the previous call (a tape in pieces); bit 1: leave them
    // unflushed for the next call -- the pieces' sum is then exactly the whole tape's
    if (!(ga->rev_flags & 1)) {
        const int zgrid = grid_for(2 * n + N_THETA, 65535);
        hipLaunchKernelGGL(fx_zero_kernel, dim3(zgrid), dim3(TPB), 0, s, 2 * n + N_THETA, acc);
    }
    const int grid = grid_for(cap, 65535);
    hipLaunchKernelGGL(grad_reverse_kernel<Real>, dim3(grid), dim3(TPB), 0, s, r);
    PG_LAUNCHED("grad_reverse_kernel");
    if (!(ga->rev_flags & 2)) {
        const int fgrid = grid_for(n, 65535);
        hipLaunchKernelGGL(fx_flush_kernel, dim3(fgrid), dim3(TPB), 0, s, n, acc, lam);
        hipLaunchKernelGGL(fx_flush_kernel, dim3(fgrid), dim3(TPB), 0, s, n, acc + n, grad_erod);
        hipLaunchKernelGGL(fx_flush_kernel, dim3(1), dim3(64), 0, s, (int64_t)N_THETA, acc + 2 * n, grad_theta);
        PG_LAUNCHED("fx_flush_kernel");
    }
    return 0;
}

}  // namespace
}  // namespace gradmode
}  // namespace pg

using namespace pg;
using namespace pg::fastmode;
using namespace pg::gradmode;

extern "C" int64_t pg_grad_reverse_work_bytes(int64_t F, int64_t N, int64_t H) {
    clear_error();
    PG_REQUIRE(F > 0 && N >= 2 && H >= 1, "pg_grad_reverse_work_bytes: bad grid");
    return reverse_work_bytes(F * (N + 2 * H) * (N + 2 * H));
}

extern "C" int pg_grad_tape_bytes(int64_t capacity, int64_t max_steps, int32_t real, int64_t* bytes) {
    clear_error();
    PG_REQUIRE(bytes && capacity >= 0 && max_steps >= 0 && (real == 0 || real == 1), "pg_grad_tape_bytes: bad arguments");
    const ll stride = (ll)max_steps + 1;
    *bytes = (int64_t)(real ? tape_bytes<double>(capacity, stride) : tape_bytes<float>(capacity, stride));
    return 0;
}

extern "C" int pg_grad_iterate(const pg_fast_fields* steer, const pg_fast_fields* target, const pg_trace_params* prm,
                               const pg_fast_params* fp, const pg_grad_args* ga, void* work, int64_t work_bytes,
                               int64_t* stats, pg_stream stream) {
    clear_error();
    PG_REQUIRE(target && ga && ga->surf, "pg_grad_iterate: NULL argument");
    PG_REQUIRE(ga->real == 0 || ga->real == 1, "pg_grad_iterate: real must be 0 (float) or 1 (double)");
    PG_REQUIRE(fp && fp->chunks == 1, "pg_grad_iterate: the differentiable mode runs one chunk per iteration");
    PG_REQUIRE(prm && prm->lateral_rate <= 0.0, "pg_grad_iterate: lateral cuts are outside the differentiable mode");
    PG_REQUIRE(!fp->erode_cap, "pg_grad_iterate: the slope limit (erode_cap) is not differentiated");
    PG_REQUIRE(steer && target->F == steer->F && target->N == steer->N && target->H == steer->H,
               "pg_grad_iterate: steering and target states differ in shape");
    Ctx c;
    if (prepare(steer, prm, fp, work, work_bytes, stats, stream, &c)) return -1;
    PG_REQUIRE(target->height && target->sediment && target->discharge && target->momentum && target->pending &&
                   target->mask,
               "pg_grad_iterate: NULL target field");
    PG_REQUIRE(ga->d_hi <= ga->d_lo || ga->tape, "pg_grad_iterate: a droplet range records a tape and needs one");
    if (ga->reuse_spawn) {
        // the workspace holds this iteration's record and spawn from the previous call: only the queue restarts
        hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, c.s, c.sc);
        PG_LAUNCHED("queue_reset_kernel");
    } else if (fast_spawn(steer, prm, fp, work, c)) {
        return -1;
    }
    if (ga->real ? launch_forward<double>(steer, prm, fp, ga, work, c) : launch_forward<float>(steer, prm, fp, ga, work, c))
        return -1;
    if (ga->d_hi > ga->d_lo) return 0;  // a range records its tape only: the target is not written
    return finish_iteration(target, fp, work, c);
}

extern "C" int pg_grad_reverse(int64_t F, int64_t N, int64_t H, const pg_trace_params* prm, const pg_grad_args* ga,
                               const double* g, double* lam, double* grad_erod, double* grad_theta, pg_stream stream) {
    clear_error();
    PG_REQUIRE(prm && ga && ga->tape && g && lam && grad_erod && grad_theta, "pg_grad_reverse: NULL argument");
    PG_REQUIRE(F > 0 && N >= 2 && H >= 1, "pg_grad_reverse: bad grid");
    PG_REQUIRE(ga->real == 0 || ga->real == 1, "pg_grad_reverse: real must be 0 or 1");
    hipStream_t s = (hipStream_t)stream;
    return ga->real ? launch_reverse<double>(F, N, H, prm, ga, g, lam, grad_erod, grad_theta, s)
                    : launch_reverse<float>(F, N, H, prm, ga, g, lam, grad_erod, grad_theta, s);
}
