// Check mode: the droplet trace and the serial change-list apply with the
// reference's exact semantics and float64 droplet state.
//
// trace_particles (bake/globe/erosion/particle.py:275-818): one lane per
// droplet of a chunk, traced on the frozen packed record; each lane fills its
// own slice of the change list (cap entries) with one entry per step at the
// cell it leaves, then its final deposits.  apply_changes (particle.py:
// 821-1042) is inherently serial -- later entries see the live terrain the
// earlier ones left -- so it runs as a single lane walking the chunk's
// lists in droplet order.
//
// Arithmetic follows numba's typing of the reference: float32 samples
// combined with float32 stay float32 (h10 - h00, the route-flag sum, the
// bank height), anything touching a float64 is float64, and float64 values
// stored into the float32 record round to nearest.  Built with
// -ffp-contract=off: no fused multiply-adds.
#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"

namespace pg {
namespace {

static_assert((SPREAD & (SPREAD - 1)) == 0, "the ring index wraps with a mask");

// Cell ids are 32-bit (F * NE * NE < 2^31, checked on the host); only the
// record's address is formed in 64 bits.
struct Record {
    const float* s;
    int NE;
    __device__ __forceinline__ float at(int f, int i, int j, int c) const {
        return s[(size_t)((f * NE + i) * NE + j) * NS + c];
    }
};

// Bilinear corner and weights at a fractional extended index (particle.py:108-130).
struct Corner {
    int i0, j0;
    double wi, wj;
};

__device__ __forceinline__ Corner corner_of(double fi, double fj, int NE) {
    int i0 = (int)floor(fi), j0 = (int)floor(fj);
    i0 = i0 < 0 ? 0 : (i0 > NE - 2 ? NE - 2 : i0);
    j0 = j0 < 0 ? 0 : (j0 > NE - 2 ? NE - 2 : j0);
    double wi = fi - (double)i0, wj = fj - (double)j0;
    wi = wi < 0.0 ? 0.0 : (wi > 1.0 ? 1.0 : wi);
    wj = wj < 0.0 ? 0.0 : (wj > 1.0 ? 1.0 : wj);
    return {i0, j0, wi, wj};
}

// particle.py:133-140
__device__ __forceinline__ int cell_of(double x, int N) {
    int i = (int)floor(x);
    return i < 0 ? 0 : (i > N - 1 ? N - 1 : i);
}

// The four float32 corner values of one channel.
struct Quad {
    float v00, v10, v01, v11;
};

__device__ __forceinline__ Quad quad(const Record& R, int f, const Corner& k, int c) {
    return {R.at(f, k.i0, k.j0, c), R.at(f, k.i0 + 1, k.j0, c), R.at(f, k.i0, k.j0 + 1, c), R.at(f, k.i0 + 1, k.j0 + 1, c)};
}

// Bilinear value, summed in the reference's order (particle.py:199-204).
__device__ __forceinline__ double bilinear(const Quad& q, const Corner& k) {
    return (double)q.v00 * (1.0 - k.wi) * (1.0 - k.wj) + (double)q.v10 * k.wi * (1.0 - k.wj)
         + (double)q.v01 * (1.0 - k.wi) * k.wj + (double)q.v11 * k.wi * k.wj;
}

// Exact gradient of the bilinear interpolant (particle.py:221-223): the
// corner differences are float32 subtractions.
__device__ __forceinline__ void gradient(const Quad& q, const Corner& k, double* gi, double* gj) {
    float d_i0 = q.v10 - q.v00, d_i1 = q.v11 - q.v01, d_j0 = q.v01 - q.v00, d_j1 = q.v11 - q.v10;
    *gi = (double)d_i0 * (1.0 - k.wj) + (double)d_i1 * k.wj;
    *gj = (double)d_j0 * (1.0 - k.wi) + (double)d_j1 * k.wi;
}

// Any routing flag in the stencil (particle.py:227-233), float32 sum.
__device__ __forceinline__ bool route_flagged(const Record& R, int f, const Corner& k) {
    Quad q = quad(R, f, k, S_RFLAG);
    float s = q.v00 + q.v10;
    s = s + q.v01;
    s = s + q.v11;
    return s != 0.0f;
}

// A cell's whole 16-channel record: 64 bytes, four float4 loads.
struct Rec16 {
    float4 q[4];
};
__device__ __forceinline__ Rec16 load_rec(const float* s, int cell) {
    const float4* r = (const float4*)(s + (size_t)cell * NS);
    return {{r[0], r[1], r[2], r[3]}};
}
__device__ __forceinline__ float chan(const Rec16& r, int c) {
    const float4& v = r.q[c >> 2];
    return (c & 3) == 0 ? v.x : (c & 3) == 1 ? v.y : (c & 3) == 2 ? v.z : v.w;
}

// The corner channels a step interpolates -- surface, route, discharge,
// momentum (both components) and the routing flag -- of the four corners of
// a bilinear stencil, in quad()'s order.  A step reads its stencil from these
// registers, and the stencil it steps to is loaded once, with the next
// cell's record, all in flight together, and becomes the next step's.  The
// record is frozen during the trace, so holding it changes no value.
struct Corner6 {
    float4 a;  // S_SURF, S_ROUTE, S_DISCH, S_MOMA
    float momb, rflag;
};
struct Sten {
    Corner6 c[4];
};
__device__ __forceinline__ Corner6 load_corner(const float* s, int cell, bool route) {
    const float* c = s + (size_t)cell * NS;
    return {*(const float4*)c, c[S_MOMB], route ? c[S_RFLAG] : 0.0f};
}
__device__ __forceinline__ Sten load_sten(const Record& R, int f, const Corner& k, bool route) {
    const int NE = R.NE, b = (f * NE + k.i0) * NE + k.j0;
    return {{load_corner(R.s, b, route), load_corner(R.s, b + NE, route), load_corner(R.s, b + 1, route),
             load_corner(R.s, b + NE + 1, route)}};
}
__device__ __forceinline__ float corner_chan(const Corner6& c, int ch) {
    return ch == S_SURF ? c.a.x : ch == S_ROUTE ? c.a.y : ch == S_DISCH ? c.a.z : ch == S_MOMA ? c.a.w : ch == S_MOMB ? c.momb : c.rflag;
}
__device__ __forceinline__ Quad squad(const Sten& S, int ch) {
    return {corner_chan(S.c[0], ch), corner_chan(S.c[1], ch), corner_chan(S.c[2], ch), corner_chan(S.c[3], ch)};
}
// Any routing flag in the stencil (particle.py:227-233), float32 sum.
__device__ __forceinline__ bool sflagged(const Sten& S) {
    const Quad q = squad(S, S_RFLAG);
    float sum = q.v00 + q.v10;
    sum = sum + q.v01;
    sum = sum + q.v11;
    return sum != 0.0f;
}
__device__ __forceinline__ Rec16 cell_rec(const Record& R, int f, int i, int j) {
    return load_rec(R.s, (f * R.NE + i) * R.NE + j);
}

// Metric norm and dot product at a cell (particle.py:236-252).
struct Metric {
    double g0, g1, g2;
    __device__ __forceinline__ double norm(double a, double b) const {
        double n2 = g0 * a * a + 2.0 * g1 * a * b + g2 * b * b;
        return sqrt(n2 < 0.0 ? 0.0 : n2);
    }
    __device__ __forceinline__ double dot(double a, double b, double c, double d) const {
        return g0 * a * c + g1 * (a * d + b * c) + g2 * b * d;
    }
};

// The face transfer out of line: a wave runs its double-double tan / atan
// only for the lanes that cross a face edge, and the step loop keeps its
// registers.
__device__ __noinline__ Transfer transfer_out_of_line(int face, double u, double v, double a, double b) {
    return transfer_velocity(face, u, v, a, b);
}

struct TraceArgs {
    int64_t P;
    const int64_t* sp_face;
    const double *sp_x, *sp_y, *sp_sed, *sp_dir;
    double volume0;
    Record R;
    int use_route;
    const uint8_t* mask;
    int64_t N, H;
    int spherical;
    pg_trace_params prm;
    const float* roots;
    int use_roots;
    int32_t* cl_cell;
    double* cl_delta;
    float* cl_vol;
    float* cl_mom;
    int64_t *cl_count, *cl_lo, *cl_hi;
    int8_t* sp_death;
    int lanes;  // droplets per wave (the rest of the wave idles)
};

// Droplets per wave of the trace (pg_check_set_lanes): a wave steps through
// the union of the branches its droplets take, so fewer droplets per wave
// can mean fewer instructions per step; a chunk's 2048 droplets leave most
// SIMDs idle either way.  Library-wide, like the launch timers.
int g_trace_lanes = 64;

// A droplet's slice of the change list.
struct Slice {
    int32_t* cell;
    double* delta;
    float* vol;
    float* mom;
    int64_t base;
    int n, lo, hi;
    __device__ __forceinline__ void put(int c, double d, float v, float m0, float m1) {
        const int64_t e = base + n;
        cell[e] = (int32_t)c;
        delta[e] = d;
        vol[e] = v;
        mom[2 * e] = m0;
        mom[2 * e + 1] = m1;
        n += 1;
        lo = c < lo ? c : lo;
        hi = c > hi ? c : hi;
    }
};

// A loop-invariant double held in vector registers.  The trace loop reads
// ~25 parameters; as scalars they exceed the SGPR file and the compiler
// spills them to VGPR lanes, paying a v_readlane (and its wait states) at
// every use.  A lane has room for them (one wave per SIMD here), and a VGPR
// operand costs nothing extra.
__device__ __forceinline__ double vreg(double x) {
    asm volatile("" : "+v"(x));
    return x;
}
template <class T>
__device__ __forceinline__ T* vptr(T* x) {
    asm volatile("" : "+v"(x));
    return x;
}

__global__ void __launch_bounds__(64) trace_check_kernel(TraceArgs a) {
    if ((int)threadIdx.x >= a.lanes) return;
    const int64_t p = (int64_t)blockIdx.x * a.lanes + threadIdx.x;
    if (p >= a.P) return;
    pg_trace_params P = a.prm;
    P.dt = vreg(P.dt);
    P.density = vreg(P.density);
    P.friction = vreg(P.friction);
    P.deposition_rate = vreg(P.deposition_rate);
    P.evap_rate = vreg(P.evap_rate);
    P.k_mom = vreg(P.k_mom);
    P.mom_scale = vreg(P.mom_scale);
    P.k_disc = vreg(P.k_disc);
    P.disc_exponent = vreg(P.disc_exponent);
    P.slope_gain = vreg(P.slope_gain);
    P.slope_sat = vreg(P.slope_sat);
    P.erodibility = vreg(P.erodibility);
    P.cover_depth = vreg(P.cover_depth);
    P.min_volume = vreg(P.min_volume);
    P.max_erode = vreg(P.max_erode);
    P.conc_scale = vreg(P.conc_scale);
    P.ocean_rate = vreg(P.ocean_rate);
    P.fan_room = vreg(P.fan_room);
    P.dep_floor = vreg(P.dep_floor);
    P.lake_trap = vreg(P.lake_trap);
    P.lateral_rate = vreg(P.lateral_rate);
    a.cl_cell = vptr(a.cl_cell);
    a.cl_delta = vptr(a.cl_delta);
    a.cl_vol = vptr(a.cl_vol);
    a.cl_mom = vptr(a.cl_mom);
    a.mask = vptr(a.mask);
    a.roots = vptr(a.roots);
    a.R.s = vptr(a.R.s);
    const Record& R = a.R;
    const int N = (int)a.N, H = (int)a.H, NE = N + 2 * H;
    const int64_t cap = change_list_cap(P.max_steps, P.lateral_rate > 0.0);
    const double Nf = (double)N;
    const double inv_sat = P.disc_saturation > 0.0 ? 1.0 / P.disc_saturation : 0.0;
    auto maskat = [&](int f, int i, int j) -> int { return a.mask[(f * NE + i) * NE + j]; };

    int f = (int)a.sp_face[p];
    double x = a.sp_x[p], y = a.sp_y[p];
    double vol = a.volume0;
    double sed = a.sp_sed[p];
    double sa = 0.0, sb = 0.0;
    if (sed > 0.0) pg_sincos(a.sp_dir[p], &sb, &sa);

    Slice L{a.cl_cell, a.cl_delta, a.cl_vol, a.cl_mom, p * cap, 0, 2147483647, -1};
    // ring of the last SPREAD active cells, in the slice's tail (particle.py:407-412)
    const int64_t rbase = L.base + cap - SPREAD;
    int nring = 0, rpos = 0;  // rpos in [0, SPREAD)
    double h_prev = -1e300;
    bool exited = false, in_sea = false, in_lake = false;
    int64_t uphill = 0, sea_steps = 0;
    int parks = 0;
    int cause = DEATH_AGE;

    // the first step's stencil, cell record and mask; later steps get theirs
    // from the step before
    // (and the step's cell and bilinear corner, which are the previous step's
    // next cell and next corner: the same expressions of the same position)
    int ei = cell_of(x, N) + H, ej = cell_of(y, N) + H;
    Corner k0 = corner_of(x - 0.5 + (double)H, y - 0.5 + (double)H, NE);
    Sten S = load_sten(R, f, k0, a.use_route);
    Rec16 cr = cell_rec(R, f, ei, ej);
    int m = maskat(f, ei, ej);
    for (int64_t step = 0; step < P.max_steps; step++) {
        if (m == MASK_OUTSIDE) { exited = true; cause = DEATH_EXIT; break; }
        const int cell = (f * NE + ei) * NE + ej;
        const float h_here = chan(cr, S_SURF);
        const float b_here = chan(cr, S_BASE);
        if (step == 0 && h_here < b_here) in_sea = true;
        if (in_sea) {
            if (h_here >= b_here || sed * vol / a.volume0 < 1e-4 || sea_steps >= P.ocean_steps) { cause = DEATH_OCEAN; break; }
            sea_steps += 1;
        }
        if (m == MASK_ACTIVE) {
            a.cl_cell[rbase + rpos] = (int32_t)cell;
            rpos = (rpos + 1) & (SPREAD - 1);
            if (nring < SPREAD) nring += 1;
        }
        // -- forces (particle.py:450-511)
        const Quad hs = squad(S, S_SURF);
        const double h0 = bilinear(hs, k0);
        double gi, gj;
        gradient(hs, k0, &gi, &gj);
        double h0r = h0;
        if (a.use_route && sflagged(S)) {
            const Quad rs = squad(S, S_ROUTE);
            const double r0 = bilinear(rs, k0);
            if (r0 > h0) {
                h0r = r0;
                gradient(rs, k0, &gi, &gj);
            }
        }
        const double q0 = chan(cr, S_GINV), q1 = chan(cr, S_GINV + 1), q2 = chan(cr, S_GINV + 2);
        const double di = q0 * gi + q1 * gj;
        const double dj = q1 * gi + q2 * gj;
        double s2 = gi * di + gj * dj;
        if (s2 < 0.0) s2 = 0.0;
        const double nrm = P.slope_sat > 0.0 ? P.slope_gain / sqrt(s2 + P.slope_sat * P.slope_sat)
                                             : P.slope_gain / sqrt(1.0 + s2);
        const double na = -di * nrm, nb = -dj * nrm;
        const double ma = bilinear(squad(S, S_MOMA), k0);
        const double mb = bilinear(squad(S, S_MOMB), k0);
        double q = bilinear(squad(S, S_DISCH), k0);
        if (q < 0.0) q = 0.0;
        const double vol_rel = vol / a.volume0;
        const double psa = sa, psb = sb;
        double fr = 1.0 - P.dt * P.friction;
        if (fr < 0.0) fr = 0.0;
        sa *= fr;
        sb *= fr;
        const Metric g{(double)chan(cr, S_G), (double)chan(cr, S_G + 1), (double)chan(cr, S_G + 2)};
        const double mlen = g.norm(ma, mb);
        double slen = g.norm(sa, sb);
        if (mlen > 1e-12 && slen > 1e-12) {
            const double cosang = g.dot(ma, mb, sa, sb) / (mlen * slen);
            const double kk = P.k_mom * cosang * P.mom_scale / (vol + P.mom_scale * q);
            sa += kk * ma;
            sb += kk * mb;
        }
        const double acc = P.dt / (vol_rel * P.density);
        sa += acc * na;
        sb += acc * nb;
        slen = g.norm(sa, sb);
        if (slen < 1e-12) { cause = DEATH_STOP; break; }
        sa /= slen;
        sb /= slen;
        // -- one cell step, across a face edge if need be (particle.py:513-549)
        double nx = x + sa, ny = y + sb;
        int nf = f;
        if (a.spherical) {
            if (nx < 0.0 || nx >= Nf || ny < 0.0 || ny >= Nf) {
                const Transfer t = transfer_out_of_line((int)f, nx / Nf, ny / Nf, sa, sb);
                nf = t.face;
                nx = t.u * Nf;
                ny = t.v * Nf;
                sa = t.a;
                sb = t.b;
                if (nx >= Nf) nx = Nf * 0.999999999;
                if (ny >= Nf) ny = Nf * 0.999999999;
                if (nx < 0.0) nx = 0.0;
                if (ny < 0.0) ny = 0.0;
            }
        } else if (nx < (double)(-H) || nx >= Nf + (double)H || ny < (double)(-H) || ny >= Nf + (double)H) {
            exited = true; cause = DEATH_EXIT; break;
        }
        const int nei = (int)floor(nx) + H, nej = (int)floor(ny) + H;
        if (nei < 0 || nei >= NE || nej < 0 || nej >= NE) { exited = true; cause = DEATH_EXIT; break; }
        // everything the rest of the step and the next step read, in flight at once
        const double nfi = nx - 0.5 + (double)H, nfj = ny - 0.5 + (double)H;
        const Corner k1 = corner_of(nfi, nfj, NE);
        const Sten Sn = load_sten(R, nf, k1, a.use_route);
        const int ei2 = cell_of(nx, N) + H, ej2 = cell_of(ny, N) + H;
        const Rec16 crn = cell_rec(R, nf, ei2, ej2);
        const int m2 = maskat(nf, ei2, ej2);
        // the cell stepped to: on the globe the position is inside its face, so
        // it is the next step's cell; a window's can lie in the halo
        const int nmask = a.spherical ? m2 : maskat(nf, nei, nej);
        const Rec16 nrec = a.spherical ? crn : cell_rec(R, nf, nei, nej);
        if (nmask == MASK_OUTSIDE) { exited = true; cause = DEATH_EXIT; break; }
        const double h1 = bilinear(squad(Sn, S_SURF), k1);
        const float b1 = chan(nrec, S_BASE);
        const double dh = h0 - (h1 > (double)b1 ? h1 : (double)b1);
        double h1r = h1;
        if (a.use_route && sflagged(Sn)) {
            const double r1 = bilinear(squad(Sn, S_ROUTE), k1);
            if (r1 > h1) h1r = r1;
        }
        const double dhr = h0r - h1r;
        const float nflag = chan(nrec, S_RFLAG);
        // -- mass transfer at the cell being left (particle.py:561-740)
        if (m == MASK_ACTIVE && in_sea) {
            double cdiff = -P.ocean_rate * sed;
            double room = h_prev > -1e300 ? h_prev - (double)h_here : P.fan_room;
            if (room < P.fan_room) room = P.fan_room;
            const double lim = (double)b_here - P.dep_floor - (double)h_here;
            if (lim < room) room = lim;
            if (room < 0.0) room = 0.0;
            if (-cdiff * vol_rel > room) cdiff = -room / vol_rel;
            sed += cdiff;
            L.put(cell, -cdiff * vol_rel, 0.0f, 0.0f, 0.0f);
        } else if (m == MASK_ACTIVE) {
            double qe;
            if (inv_sat <= 0.0) qe = 0.0;
            else if (P.disc_exponent == 0.5) qe = sqrt(q * inv_sat);
            else if (P.disc_exponent > 0.0) qe = pow(q * inv_sat, P.disc_exponent);
            else qe = erf(q * inv_sat);
            double c_eq = P.erodibility * dh * (1.0 + P.k_disc * qe);
            if (c_eq < 0.0) c_eq = 0.0;
            double k_e;
            if (c_eq > sed) {
                const float sthick = chan(cr, S_SED);
                double w_cover;
                if (P.cover_depth > 0.0) {
                    w_cover = (double)sthick / P.cover_depth;
                    if (w_cover > 1.0) w_cover = 1.0;
                } else {
                    w_cover = sthick > 0.0f ? 1.0 : 0.0;
                }
                k_e = P.deposition_rate * (w_cover + (1.0 - w_cover) * (1.0 - (double)chan(cr, S_HARD)));
            } else {
                k_e = P.deposition_rate;
            }
            if (a.use_roots) k_e *= 1.0 - (double)a.roots[(f * NE + ei) * NE + ej];
            double cdiff = k_e * (c_eq - sed);
            double trapped = 0.0;
            if (in_lake) {
                cdiff = 0.0;
            } else if (sed > 0.0 && (double)nflag >= 1.5) {
                const double trap = -P.lake_trap * sed;
                if (trap < cdiff) cdiff = trap;
                if ((double)nflag >= 2.5) trapped = P.lake_trap * sed;
            }
            const double kexp = 1.0 + q * P.conc_scale;
            if (cdiff > 0.0) {
                double ecap = dh / kexp;
                if (ecap > P.max_erode) ecap = P.max_erode;
                if (ecap > (double)h_here) ecap = h_here > 0.0f ? (double)h_here : 0.0;
                if (cdiff * vol_rel > ecap) cdiff = ecap / vol_rel;
            } else if (cdiff < 0.0) {
                double room = h_prev > -1e300 ? h_prev - (double)h_here : P.dep_floor;
                if (room < P.dep_floor) room = P.dep_floor;
                const double dcap = room / kexp;
                if (-cdiff * vol_rel > dcap) cdiff = -dcap / vol_rel;
            }
            sed += cdiff;
            L.put(cell, -cdiff * vol_rel, (float)vol, (float)(vol * sa), (float)(vol * sb));
            if (P.lake_park > 0.0 && trapped > 0.0 && parks < LAKE_PARKS) {
                const double rest = trapped + cdiff;
                if (rest > 0.0) {
                    parks += 1;
                    sed -= rest;
                    L.put((nf * NE + nei) * NE + nej, rest * vol_rel, -3.0f, 0.0f, 0.0f);
                }
            }
            if (P.lateral_rate > 0.0 && cdiff > 0.0 && L.n + 2 * SPREAD + 2 < cap) {
                const double cross = psa * sb - psb * sa;
                const double turn = cross > 0.0 ? cross : -cross;
                if (turn > 1e-6) {
                    const double sgn = cross > 0.0 ? 1.0 : -1.0;
                    const double bx = x + sgn * sb, by = y - sgn * sa;
                    const int bei = (int)floor(bx) + H, bej = (int)floor(by) + H;
                    if (0 <= bei && bei < NE && 0 <= bej && bej < NE && maskat(f, bei, bej) == MASK_ACTIVE) {
                        const float bank = R.at(f, bei, bej, S_SURF) - h_here;
                        if (bank > 0.0f) {
                            double lat = P.lateral_rate * turn * cdiff;
                            double lcap = (double)bank / kexp;
                            if (lcap > P.max_erode) lcap = P.max_erode;
                            if (lat * vol_rel > lcap) lat = lcap / vol_rel;
                            if (lat > 0.0) {
                                sed += lat;
                                L.put((f * NE + bei) * NE + bej, -lat * vol_rel, -2.0f, 0.0f, 0.0f);
                            }
                        }
                    }
                }
            }
        }
        // -- evaporation and stop tests (particle.py:741-769)
        double ev = 1.0 - P.dt * P.evap_rate * (double)chan(cr, S_EVAP);
        if (ev < 0.01) ev = 0.01;
        vol *= ev;
        sed /= ev;
        x = nx;
        y = ny;
        f = nf;
        ei = ei2;
        ej = ej2;
        k0 = k1;
        S = Sn;
        cr = crn;
        m = m2;
        h_prev = (double)h_here;
        in_lake = (double)nflag >= 1.5;
        if (chan(nrec, S_SURF) < chan(nrec, S_BASE) && !in_sea) {
            in_sea = true;
            if (sed <= 0.0) { cause = DEATH_OCEAN; break; }
        }
        if (vol < P.min_volume) { cause = in_sea ? DEATH_OCEAN : DEATH_EVAP; break; }
        if (dhr < 0.0 && !in_lake) {
            uphill += 1;
            if (uphill > P.pit_steps) { cause = in_sea ? DEATH_OCEAN : DEATH_PIT; break; }
        } else {
            uphill = 0;
        }
    }
    // -- final deposits (particle.py:770-818)
    if (in_sea && (cause == DEATH_AGE || cause == DEATH_STOP)) cause = DEATH_OCEAN;
    a.sp_death[p] = (int8_t)cause;
    if (!exited) {
        const int fei = cell_of(x, N) + H, fej = cell_of(y, N) + H;
        if (maskat(f, fei, fej) == MASK_ACTIVE && !(in_sea && R.at(f, fei, fej, S_SURF) >= R.at(f, fei, fej, S_BASE))) {
            const int cell = (f * NE + fei) * NE + fej;
            const int newest = (rpos + SPREAD - 1) & (SPREAD - 1);
            if (nring == 0 || a.cl_cell[rbase + newest] != cell) {
                a.cl_cell[rbase + rpos] = (int32_t)cell;
                rpos = (rpos + 1) & (SPREAD - 1);
                if (nring < SPREAD) nring += 1;
            }
        }
    }
    if (sed > 0.0 && nring > 0 && (!exited || P.deposit_on_exit)) {
        const double load = sed * vol / a.volume0;
        const int nout = cause == DEATH_OCEAN ? 1 : nring;
        for (int r = 0; r < nout; r++) {
            const int c = a.cl_cell[rbase + ((rpos + SPREAD - 1 - r) & (SPREAD - 1))];
            L.put(c, r == 0 ? load : 0.0, -1.0f, 0.0f, 0.0f);
        }
    }
    a.cl_count[p] = L.n;
    a.cl_lo[p] = L.lo;
    a.cl_hi[p] = L.hi;
}

struct ApplyArgs {
    const int32_t* cl_cell;
    const double* cl_delta;
    const float* cl_vol;
    const float* cl_mom;
    const int64_t* cl_count;
    int64_t P, cap;
    const double *h, *s;
    double *acc, *pending;
    float* samp;
    double *disch, *mom;
    const uint8_t* mask;
    double iter_erode, iter_deposit, fan_slope;
    int use_route;
    double dep_floor, writeoff;
    const double* ecap;
    int use_ecap;
    double* lake_load;
    const int32_t* lake_id;
    double* lake_room;
    double* out4;          // the chunk's 4-tuple (NULL: not written)
    const uint32_t* gate;  // run only if *gate != 0 (NULL: always)
    double* drop;          // per droplet: its share of the 4-tuple (NULL: not written)
    int32_t* count;        // + 1 when the kernel runs (NULL: not counted)
};

// One lane walks the chunk in droplet order (particle.py:902-1041).
__global__ void __launch_bounds__(64) apply_serial_kernel(ApplyArgs a) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (a.gate && *a.gate == 0) return;
    if (a.count) atomicAdd(a.count, 1);
    int64_t n_clamp = 0;
    double to_pending = 0.0, lost = 0.0, lost_offshore = 0.0;
    auto live = [&](int64_t c) { return a.h[c] + a.s[c] + a.acc[c]; };
    auto rec = [&](int64_t c, int k) -> float& { return a.samp[c * NS + k]; };
    for (int64_t p = 0; p < a.P; p++) {
        const int64_t base = p * a.cap, cnt = a.cl_count[p];
        double deficit = 0.0, surplus = 0.0, refused = 0.0;
        int64_t prev = -1, death = -1;
        const int64_t clamp0 = n_clamp;
        double tp_p = 0.0, lo_p = 0.0;
        for (int64_t k = 0; k < cnt; k++) {
            const int64_t c = a.cl_cell[base + k];
            const float v = a.cl_vol[base + k];
            if (v == -3.0f) {
                // a load for a lake to keep, up to its room (particle.py:912-933)
                double take = a.cl_delta[base + k], rest = 0.0;
                const int64_t lk = a.lake_id[c];
                if (lk < 0) {
                    rest = take;
                    take = 0.0;
                } else if (take > a.lake_room[lk]) {
                    rest = take - a.lake_room[lk];
                    take = a.lake_room[lk];
                }
                if (take > 0.0) {
                    a.lake_load[c] += take;
                    a.lake_room[lk] -= take;
                }
                refused += rest;
                continue;
            }
            if (a.mask[c] != MASK_ACTIVE) continue;
            const double d = a.cl_delta[base + k];
            const bool is_step = v >= 0.0f;
            const double s_c = live(c);
            const float b_c = rec(c, S_BASE);
            if (d < 0.0) {
                // erosion under the iteration budget and the floors
                const double e = -d;
                double e_act = e;
                double room = a.iter_erode + a.acc[c];
                if (a.use_ecap) {
                    const double r2 = a.ecap[c] + a.acc[c];
                    if (r2 < room) room = r2;
                }
                if (e_act > room) e_act = room;
                double fl = s_c >= (double)b_c ? (double)b_c : s_c;
                if (a.use_route && (double)rec(c, S_RFLAG) >= 1.5) {
                    const float r_c = rec(c, S_ROUTE);
                    if ((double)r_c > fl) fl = (double)r_c;
                }
                if (k + 1 < cnt) {
                    const double ns = live(a.cl_cell[base + k + 1]);
                    if (ns > fl) fl = ns;
                }
                if (e_act > s_c - fl) e_act = s_c - fl;
                if (e_act < 0.0) e_act = 0.0;
                if (e_act < e) {
                    n_clamp += 1;
                    deficit += e - e_act;
                }
                a.acc[c] -= e_act;
            } else {
                // deposition: request + surplus - deficit, under the budget and ceilings
                double want = d + surplus;
                surplus = 0.0;
                if (deficit > 0.0) {
                    const double t = want < deficit ? want : deficit;
                    want -= t;
                    deficit -= t;
                }
                double d_act = want;
                const double room = a.iter_deposit - a.acc[c];
                if (d_act > room) d_act = room;
                int64_t ref = -1;
                if (is_step) ref = prev;
                else if (k + 1 < cnt) ref = a.cl_cell[base + k + 1];
                if (ref >= 0) {
                    double ceil_ = live(ref) + a.dep_floor;
                    if (s_c < (double)b_c && is_step) ceil_ -= a.dep_floor + a.fan_slope;
                    const double lim = ceil_ - s_c;
                    if (d_act > lim) d_act = lim;
                }
                if (s_c < (double)b_c) {
                    const double lim = (double)b_c - a.dep_floor - s_c;
                    if (d_act > lim) d_act = lim;
                }
                if (d_act < 0.0) d_act = 0.0;
                if (d_act < want) n_clamp += 1;
                surplus = want - d_act;
                a.acc[c] += d_act;
                if (!is_step && death < 0) death = c;
            }
            const double a2 = a.acc[c];
            rec(c, S_SURF) = (float)(a.h[c] + a.s[c] + a2);
            const double sd = a.s[c] + a2;
            rec(c, S_SED) = (float)(sd > 0.0 ? sd : 0.0);
            if (a.use_route && (double)rec(c, S_RFLAG) < 1.5) rec(c, S_RFLAG) = rec(c, S_ROUTE) != rec(c, S_SURF) ? 1.0f : 0.0f;
            if (is_step) {
                if (v > 0.0f) {
                    a.disch[c] += (double)v;
                    a.mom[2 * c] += (double)a.cl_mom[2 * (base + k)];
                    a.mom[2 * c + 1] += (double)a.cl_mom[2 * (base + k) + 1];
                }
                prev = c;
            }
        }
        surplus += refused;
        if (surplus > 0.0 && death >= 0) {
            if (live(death) < (double)rec(death, S_BASE)) {
                double keep = surplus * (1.0 - a.writeoff);
                if (keep < 1e-4) keep = 0.0;
                a.pending[death] += keep;
                tp_p = keep;
                lo_p = surplus - keep;
            } else {
                a.pending[death] += surplus;
                tp_p = surplus;
            }
        }
        to_pending += tp_p;
        lost_offshore += lo_p;
        lost += deficit;
        if (a.drop) {
            double* o = a.drop + 4 * p;
            o[0] = (double)(n_clamp - clamp0);
            o[1] = tp_p;
            o[2] = deficit;
            o[3] = lo_p;
        }
    }
    if (a.out4) {
        a.out4[0] = (double)n_clamp;
        a.out4[1] = to_pending;
        a.out4[2] = lost;
        a.out4[3] = lost_offshore;
    }
}

// Entries and death causes of a chunk, added to counts7 (integers: exact).
__global__ void chunk_stats_kernel(const int64_t* cl_count, const int8_t* sp_death, int64_t P, unsigned long long* counts7) {
    __shared__ unsigned long long s[7];
    if (threadIdx.x < 7) s[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long n = 0;
    unsigned long long d[N_DEATHS] = {0, 0, 0, 0, 0, 0};
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (int64_t)gridDim.x * blockDim.x) {
        n += (unsigned long long)cl_count[p];
        const int c = sp_death[p];
        if (c >= 0 && c < N_DEATHS) d[c] += 1;
    }
    atomicAdd(&s[0], n);
    for (int c = 0; c < N_DEATHS; c++)
        if (d[c]) atomicAdd(&s[1 + c], d[c]);
    __syncthreads();
    if (threadIdx.x < 7 && s[threadIdx.x]) atomicAdd(&counts7[threadIdx.x], s[threadIdx.x]);
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" int pg_trace_particles(int64_t P, const int64_t* sp_face, const double* sp_x, const double* sp_y,
                                  const double* sp_sed, const double* sp_dir, double volume0, const float* samp,
                                  int use_route, const uint8_t* mask, int64_t N, int64_t H, int spherical,
                                  const pg_trace_params* prm, const float* roots, int use_roots, int32_t* cl_cell,
                                  double* cl_delta, float* cl_vol, float* cl_mom, int64_t* cl_count, int64_t* cl_lo,
                                  int64_t* cl_hi, int8_t* sp_death, pg_stream stream) {
    clear_error();
    PG_REQUIRE(prm != nullptr, "pg_trace_particles: prm is NULL");
    PG_REQUIRE(P >= 0 && N >= 2 && H >= 1, "pg_trace_particles: bad sizes P=%lld N=%lld H=%lld", (long long)P, (long long)N, (long long)H);
    PG_REQUIRE(prm->max_steps >= 0, "pg_trace_particles: max_steps < 0");
    PG_REQUIRE(!use_roots || roots, "pg_trace_particles: use_roots without roots");
    PG_REQUIRE(volume0 > 0.0, "pg_trace_particles: volume0 must be > 0");
    if (P == 0) return 0;
    PG_REQUIRE(sp_face && sp_x && sp_y && sp_sed && sp_dir && samp && mask && cl_cell && cl_delta && cl_vol && cl_mom &&
                   cl_count && cl_lo && cl_hi && sp_death,
               "pg_trace_particles: NULL array");
    // 32-bit cell ids: six faces of NE x NE cells (a window has one) below 2^31
    PG_REQUIRE(6 * (N + 2 * H) * (N + 2 * H) < ((int64_t)1 << 31), "pg_trace_particles: N=%lld too large for 32-bit cell ids",
               (long long)N);
    TraceArgs a{P, sp_face, sp_x, sp_y, sp_sed, sp_dir, volume0, Record{samp, (int)(N + 2 * H)}, use_route, mask, N, H,
                spherical, *prm, roots, use_roots, cl_cell, cl_delta, cl_vol, cl_mom, cl_count, cl_lo, cl_hi, sp_death,
                g_trace_lanes};
    hipLaunchKernelGGL(trace_check_kernel, dim3(blocks_for(P, g_trace_lanes)), dim3(64), 0, (hipStream_t)stream, a);
    PG_LAUNCHED("trace_check_kernel");
    return 0;
}

extern "C" int pg_check_set_lanes(int lanes) {
    clear_error();
    PG_REQUIRE(lanes >= 1 && lanes <= 64 && (lanes & (lanes - 1)) == 0, "pg_check_set_lanes: %d is not a power of two "
               "in [1, 64]", lanes);
    g_trace_lanes = lanes;
    return 0;
}

extern "C" int64_t pg_change_list_cap(int64_t max_steps, int lateral) { return change_list_cap(max_steps, lateral != 0); }

extern "C" int pg_apply_changes(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol, const float* cl_mom,
                                const int64_t* cl_count, int64_t P, int64_t cap, const double* height,
                                const double* sediment, double* acc, double* pending, float* samp, double* disch_track,
                                double* mom_track, const uint8_t* mask, double iter_erode, double iter_deposit,
                                double fan_slope, int use_route, double dep_floor, double offshore_writeoff,
                                const double* erode_cap, int use_erode_cap, double* lake_load, const int32_t* lake_id,
                                double* lake_room, double* out4, pg_stream stream) {
    clear_error();
    PG_REQUIRE(P >= 0 && cap > 0, "pg_apply_changes: bad sizes P=%lld cap=%lld", (long long)P, (long long)cap);
    PG_REQUIRE(out4 != nullptr, "pg_apply_changes: out4 is NULL");
    PG_REQUIRE(!use_erode_cap || erode_cap, "pg_apply_changes: use_erode_cap without erode_cap");
    PG_REQUIRE(cl_cell && cl_delta && cl_vol && cl_mom && cl_count && height && sediment && acc && pending && samp &&
                   disch_track && mom_track && mask,
               "pg_apply_changes: NULL array");
    ApplyArgs a{cl_cell, cl_delta, cl_vol, cl_mom, cl_count, P, cap, height, sediment, acc, pending, samp,
                disch_track, mom_track, mask, iter_erode, iter_deposit, fan_slope, use_route, dep_floor,
                offshore_writeoff, erode_cap, use_erode_cap, lake_load, lake_id, lake_room, out4, nullptr, nullptr,
                nullptr};
    hipLaunchKernelGGL(apply_serial_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, a);
    PG_LAUNCHED("apply_serial_kernel");
    return 0;
}

namespace pg {
int enqueue_serial_apply(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol, const float* cl_mom,
                         const int64_t* cl_count, int64_t P, int64_t cap, const double* height, const double* sediment,
                         double* acc, double* pending, float* samp, double* disch_track, double* mom_track,
                         const uint8_t* mask, double iter_erode, double iter_deposit, double fan_slope, int use_route,
                         double dep_floor, double offshore_writeoff, const double* erode_cap, int use_erode_cap,
                         double* lake_load, const int32_t* lake_id, double* lake_room, const uint32_t* gate,
                         double* drop, int32_t* count, hipStream_t stream) {
    ApplyArgs a{cl_cell, cl_delta, cl_vol, cl_mom, cl_count, P, cap, height, sediment, acc, pending, samp,
                disch_track, mom_track, mask, iter_erode, iter_deposit, fan_slope, use_route, dep_floor,
                offshore_writeoff, erode_cap, use_erode_cap, lake_load, lake_id, lake_room, nullptr, gate, drop, count};
    hipLaunchKernelGGL(apply_serial_kernel, dim3(1), dim3(64), 0, stream, a);
    PG_LAUNCHED("apply_serial_kernel (gated)");
    return 0;
}
}  // namespace pg

extern "C" int pg_chunk_stats(const int64_t* cl_count, const int8_t* sp_death, int64_t P, int64_t* counts7, pg_stream stream) {
    clear_error();
    PG_REQUIRE(P >= 0 && counts7, "pg_chunk_stats: bad arguments");
    if (P == 0) return 0;
    PG_REQUIRE(cl_count && sp_death, "pg_chunk_stats: NULL array");
    int blocks = blocks_for(P, 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(chunk_stats_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, cl_count, sp_death, P,
                       (unsigned long long*)counts7);
    PG_LAUNCHED("chunk_stats_kernel");
    return 0;
}
