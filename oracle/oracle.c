/*
 * ORACLE -- test infrastructure only.
 *
 * A CPU restatement, in plain C, of the reference's particle-erosion kernels
 * (bake/globe/erosion/particle.py) and of the cube-sphere scalar geometry they
 * call (bake/globe/cubesphere.py).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline; the product (the HIP library) never links or calls it.
 *
 * Type rules follow numba's, which the reference is compiled with:
 *   - float32 (op) float32 stays float32 (e.g. `h10 - h00`, particle.py:222;
 *     the S_RFLAG sum, particle.py:233; `bank = hb - h_here`, particle.py:720);
 *   - float32 (op) float64, including Python float literals, is float64;
 *   - float64 stores into float32 arrays round to nearest even.
 * Build with -ffp-contract=off -fno-fast-math (no FMA contraction), so every
 * expression rounds exactly where numba's (fastmath=False) does.
 *
 * Parity status of each function (how it is pinned, see tests/):
 *   to_sphere / from_sphere  -- bit-exact against the reference's golden
 *                               vector file (5000 records, glibc tan/atan).
 *   jacobian / transfer_*    -- closed forms: 3-D velocity preserved, Jacobian
 *                               against finite differences, edge behaviour.
 *   sampling helpers         -- closed forms (bilinear reproduces linear
 *                               functions; the gradient is the interpolant's).
 *   trace / apply            -- hand-checked single-particle traces (flat plane,
 *                               tilt, pit) and the reference's own invariants
 *                               (mass conservation, offshore split, change-list
 *                               layout, stockpile bound, exit loses mass).
 *                               Lateral cuts: the reference's own test
 *                               (test_erosion.py:1757-1791), each cut's amount
 *                               (rate x |sin turn| x the step's erosion) and
 *                               side read back from the change lists, and
 *                               closed-window conservation with cuts on.
 *                               Equality with numba's own codegen is NOT pinned:
 *                               numba is not installable here.
 *   fold / ema / pack / spawn_cells -- closed forms (elementwise / brute force).
 *   thermal passes a / r / b  -- closed forms (a spike on a plane: D8 offsets,
 *                               diagonal distance, the 2 x max and thermal_max
 *                               caps, sediment before bedrock), exact
 *                               conservation between active cells, the
 *                               submerged receiver's room, mirror symmetry.
 *   slope_erode_cap          -- the reference's own known-answer test
 *                               (test_erosion.py:1599-1619) and a tilted plane.
 *   laplacian                -- quadratics under a constant metric (each
 *                               inverse-metric component's role) and the
 *                               reference's spherical-harmonic eigenfunction
 *                               test (test_cubesphere.py:314-336).
 *   label_components         -- the reference's closed-basin tests
 *                               (test_hydro.py:292-309, test_erosion.py:
 *                               1109-1136) on the sea classification built on
 *                               it, and a union-find over the same neighbour
 *                               graph (another algorithm) on random masks.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------- */
/* cube-sphere geometry (cubesphere.py:37-220)                             */
/* ---------------------------------------------------------------------- */
/* BASES[f] = (right, up, normal), cubesphere.py:37-48 */
static const double BASES[6][3][3] = {
    {{0.0, 0.0, -1.0}, {0.0, -1.0, 0.0}, {1.0, 0.0, 0.0}},
    {{0.0, 0.0, 1.0}, {0.0, -1.0, 0.0}, {-1.0, 0.0, 0.0}},
    {{1.0, 0.0, 0.0}, {0.0, 0.0, 1.0}, {0.0, 1.0, 0.0}},
    {{1.0, 0.0, 0.0}, {0.0, 0.0, -1.0}, {0.0, -1.0, 0.0}},
    {{1.0, 0.0, 0.0}, {0.0, -1.0, 0.0}, {0.0, 0.0, 1.0}},
    {{-1.0, 0.0, 0.0}, {0.0, -1.0, 0.0}, {0.0, 0.0, -1.0}},
};
#define OR_PI 3.141592653589793

/* tan/atan of the geometry: glibc's (the reference's, through numba) by
 * default.  or_set_libm swaps in other implementations -- a test uses it to
 * run the oracle with correctly rounded tan/atan, which isolates the one
 * place where glibc's last-bit rounding differs from the device's. */
typedef double (*or_fn1)(double);
typedef void (*or_sincos_fn)(double, double *, double *);
static or_fn1 or_tan = tan;
static or_fn1 or_atan = atan;
/* the initial direction of a droplet that carries a load (particle.py:400-401:
 * math.cos / math.sin of its angle): glibc's by default, or a swapped-in
 * sincos(x, &sin, &cos) */
static or_sincos_fn or_sincos = 0;
void or_set_libm(or_fn1 t, or_fn1 a) {
    or_tan = t ? t : tan;
    or_atan = a ? a : atan;
}
void or_set_sincos(or_sincos_fn f) { or_sincos = f; }
static const double HALF_PI = OR_PI / 2.0;     /* cubesphere.py:50 */
static const double TWO_OVER_PI = 2.0 / OR_PI; /* cubesphere.py:51 */

/* cubesphere.py:70-80 */
void or_to_sphere(int face, double u, double v, double *ox, double *oy, double *oz) {
    double s = or_tan((u - 0.5) * HALF_PI);
    double t = or_tan((v - 0.5) * HALF_PI);
    const double *r = BASES[face][0], *up = BASES[face][1], *n = BASES[face][2];
    double x = n[0] + s * r[0] + t * up[0];
    double y = n[1] + s * r[1] + t * up[1];
    double z = n[2] + s * r[2] + t * up[2];
    double inv = 1.0 / sqrt(x * x + y * y + z * z);
    *ox = x * inv;
    *oy = y * inv;
    *oz = z * inv;
}

/* cubesphere.py:83-93: argmax |component| with sign, ties X > Y > Z */
int or_face_of(double x, double y, double z) {
    double ax = fabs(x), ay = fabs(y), az = fabs(z);
    if (ax >= ay && ax >= az) return x >= 0.0 ? 0 : 1;
    if (ay >= az) return y >= 0.0 ? 2 : 3;
    return z >= 0.0 ? 4 : 5;
}

/* cubesphere.py:96-112 */
void or_from_sphere(double x, double y, double z, int *oface, double *ou, double *ov) {
    int face = or_face_of(x, y, z);
    const double *r = BASES[face][0], *up = BASES[face][1], *n = BASES[face][2];
    double d = x * n[0] + y * n[1] + z * n[2];
    double s = (x * r[0] + y * r[1] + z * r[2]) / d;
    double t = (x * up[0] + y * up[1] + z * up[2]) / d;
    *oface = face;
    *ou = or_atan(s) * TWO_OVER_PI + 0.5;
    *ov = or_atan(t) * TWO_OVER_PI + 0.5;
}

/* cubesphere.py:130-167; j = (jux, juy, juz, jvx, jvy, jvz) */
void or_jacobian(int face, double u, double v, double *j) {
    double a = (u - 0.5) * HALF_PI;
    double b = (v - 0.5) * HALF_PI;
    double s = or_tan(a);
    double t = or_tan(b);
    double dsdu = HALF_PI * (1.0 + s * s);
    double dtdv = HALF_PI * (1.0 + t * t);
    const double *r = BASES[face][0], *up = BASES[face][1], *n = BASES[face][2];
    double qx = n[0] + s * r[0] + t * up[0];
    double qy = n[1] + s * r[1] + t * up[1];
    double qz = n[2] + s * r[2] + t * up[2];
    double ql = sqrt(qx * qx + qy * qy + qz * qz);
    double px = qx / ql, py = qy / ql, pz = qz / ql;
    double dux = r[0] * dsdu, duy = r[1] * dsdu, duz = r[2] * dsdu;
    double pd = px * dux + py * duy + pz * duz;
    j[0] = (dux - px * pd) / ql;
    j[1] = (duy - py * pd) / ql;
    j[2] = (duz - pz * pd) / ql;
    double dvx = up[0] * dtdv, dvy = up[1] * dtdv, dvz = up[2] * dtdv;
    pd = px * dvx + py * dvy + pz * dvz;
    j[3] = (dvx - px * pd) / ql;
    j[4] = (dvy - py * pd) / ql;
    j[5] = (dvz - pz * pd) / ql;
}

/* cubesphere.py:181-201 */
void or_transfer_vector(int fs, double us, double vs, double a, double b, int fd, double ud, double vd,
                        double *oa, double *ob) {
    double J[6], K[6];
    or_jacobian(fs, us, vs, J);
    double wx = a * J[0] + b * J[3];
    double wy = a * J[1] + b * J[4];
    double wz = a * J[2] + b * J[5];
    or_jacobian(fd, ud, vd, K);
    double guu = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
    double guv = K[0] * K[3] + K[1] * K[4] + K[2] * K[5];
    double gvv = K[3] * K[3] + K[4] * K[4] + K[5] * K[5];
    double ru = K[0] * wx + K[1] * wy + K[2] * wz;
    double rv = K[3] * wx + K[4] * wy + K[5] * wz;
    double det = guu * gvv - guv * guv;
    *oa = (gvv * ru - guv * rv) / det;
    *ob = (guu * rv - guv * ru) / det;
}

/* cubesphere.py:204-220 */
void or_transfer_velocity(int face, double u, double v, double a, double b, int *of, double *ou, double *ov,
                          double *oa, double *ob) {
    if (u >= 0.0 && u < 1.0 && v >= 0.0 && v < 1.0) {
        *of = face; *ou = u; *ov = v; *oa = a; *ob = b;
        return;
    }
    double x, y, z, u2, v2;
    int f2;
    or_to_sphere(face, u, v, &x, &y, &z);
    or_from_sphere(x, y, z, &f2, &u2, &v2);
    if (f2 == face) {
        /* min(max(u2, 0.0), 0.999999999) with Python's max/min */
        u2 = (0.0 > u2) ? 0.0 : u2;
        u2 = (0.999999999 < u2) ? 0.999999999 : u2;
        v2 = (0.0 > v2) ? 0.0 : v2;
        v2 = (0.999999999 < v2) ? 0.999999999 : v2;
        *of = face; *ou = u2; *ov = v2; *oa = a; *ob = b;
        return;
    }
    double a2, b2;
    or_transfer_vector(face, u, v, a, b, f2, u2, v2, &a2, &b2);
    *of = f2; *ou = u2; *ov = v2; *oa = a2; *ob = b2;
}

/* ---------------------------------------------------------------------- */
/* particle kernels (particle.py)                                          */
/* ---------------------------------------------------------------------- */
#define MASK_OUTSIDE 0
#define MASK_ACTIVE 1
#define SPREAD 8
#define LAKE_PARKS 8
#define DEATH_AGE 0
#define DEATH_OCEAN 1
#define DEATH_EXIT 2
#define DEATH_PIT 3
#define DEATH_EVAP 4
#define DEATH_STOP 5
/* channels, particle.py:149-161 */
#define S_SURF 0
#define S_ROUTE 1
#define S_DISCH 2
#define S_MOMA 3
#define S_MOMB 4
#define S_EVAP 5
#define S_HARD 6
#define S_SED 7
#define S_G 8
#define S_GINV 11
#define S_RFLAG 14
#define S_BASE 15
#define NS 16

/* particle.py:258-272 */
int64_t or_change_list_cap(int64_t max_steps, int lateral) {
    return (lateral ? 2 * max_steps : max_steps) + 2 * SPREAD + 1 + SPREAD + LAKE_PARKS;
}

#define SAMP(f, i, j, c) samp[(((size_t)(f) * NE + (size_t)(i)) * NE + (size_t)(j)) * NS + (c)]

/* particle.py:108-130 */
static inline void clamp_ij(double fi, double fj, int64_t NE, int64_t *i0, int64_t *j0, double *wi, double *wj) {
    int64_t a = (int64_t)floor(fi);
    int64_t b = (int64_t)floor(fj);
    if (a < 0) a = 0;
    if (a > NE - 2) a = NE - 2;
    if (b < 0) b = 0;
    if (b > NE - 2) b = NE - 2;
    double x = fi - (double)a;
    double y = fj - (double)b;
    if (x < 0.0) x = 0.0;
    if (x > 1.0) x = 1.0;
    if (y < 0.0) y = 0.0;
    if (y > 1.0) y = 1.0;
    *i0 = a; *j0 = b; *wi = x; *wj = y;
}

/* particle.py:133-140 */
static inline int64_t cell_of(double x, int64_t N) {
    int64_t i = (int64_t)floor(x);
    if (i < 0) i = 0;
    if (i > N - 1) i = N - 1;
    return i;
}

/* particle.py:194-204 */
static inline double sbilin(const float *samp, int64_t NE, int64_t f, double fi, double fj, int c) {
    int64_t i0, j0;
    double wi, wj;
    clamp_ij(fi, fj, NE, &i0, &j0, &wi, &wj);
    return (double)SAMP(f, i0, j0, c) * (1.0 - wi) * (1.0 - wj)
         + (double)SAMP(f, i0 + 1, j0, c) * wi * (1.0 - wj)
         + (double)SAMP(f, i0, j0 + 1, c) * (1.0 - wi) * wj
         + (double)SAMP(f, i0 + 1, j0 + 1, c) * wi * wj;
}

/* particle.py:207-224 */
static inline void sbilin_grad(const float *samp, int64_t NE, int64_t f, double fi, double fj, int c,
                               double *val, double *gi, double *gj) {
    int64_t i0, j0;
    double wi, wj;
    clamp_ij(fi, fj, NE, &i0, &j0, &wi, &wj);
    float h00 = SAMP(f, i0, j0, c);
    float h10 = SAMP(f, i0 + 1, j0, c);
    float h01 = SAMP(f, i0, j0 + 1, c);
    float h11 = SAMP(f, i0 + 1, j0 + 1, c);
    *val = (double)h00 * (1.0 - wi) * (1.0 - wj) + (double)h10 * wi * (1.0 - wj)
         + (double)h01 * (1.0 - wi) * wj + (double)h11 * wi * wj;
    float d1 = h10 - h00, d2 = h11 - h01, d3 = h01 - h00, d4 = h11 - h10; /* float32 ops */
    *gi = (double)d1 * (1.0 - wj) + (double)d2 * wj;
    *gj = (double)d3 * (1.0 - wi) + (double)d4 * wi;
}

/* particle.py:227-233 */
static inline int rflag(const float *samp, int64_t NE, int64_t f, double fi, double fj) {
    int64_t i0, j0;
    double wi, wj;
    clamp_ij(fi, fj, NE, &i0, &j0, &wi, &wj);
    float s = SAMP(f, i0, j0, S_RFLAG) + SAMP(f, i0 + 1, j0, S_RFLAG);
    s = s + SAMP(f, i0, j0 + 1, S_RFLAG);
    s = s + SAMP(f, i0 + 1, j0 + 1, S_RFLAG);
    return s != 0.0f;
}

/* particle.py:236-244 */
static inline double snorm(const float *samp, int64_t NE, int64_t f, int64_t ei, int64_t ej, double a, double b) {
    double g0 = SAMP(f, ei, ej, S_G), g1 = SAMP(f, ei, ej, S_G + 1), g2 = SAMP(f, ei, ej, S_G + 2);
    double n2 = g0 * a * a + 2.0 * g1 * a * b + g2 * b * b;
    if (n2 < 0.0) n2 = 0.0;
    return sqrt(n2);
}

/* particle.py:247-252 */
static inline double sdot(const float *samp, int64_t NE, int64_t f, int64_t ei, int64_t ej, double a, double b,
                          double c, double d) {
    double g0 = SAMP(f, ei, ej, S_G), g1 = SAMP(f, ei, ej, S_G + 1), g2 = SAMP(f, ei, ej, S_G + 2);
    return g0 * a * c + g1 * (a * d + b * c) + g2 * b * d;
}

/* Scalar parameters of trace_particles, in the reference's order
 * (particle.py:292-318). */
typedef struct {
    double dt, density, friction, deposition_rate, evap_rate, k_mom, mom_scale, k_disc, disc_saturation,
        disc_exponent, slope_gain, slope_sat, erodibility, cover_depth, min_volume;
    int64_t max_steps;
    double max_erode, conc_scale;
    int64_t pit_steps;
    int64_t deposit_on_exit;
    double ocean_rate;
    int64_t ocean_steps;
    double fan_room, dep_floor, lake_trap, lake_park, lateral_rate;
} or_trace_params;

/* particle.py:275-818, one particle */
static void trace_one(int64_t p, const int64_t *sp_face, const double *sp_x, const double *sp_y,
                      const double *sp_sed, const double *sp_dir, double volume0, const float *samp, int use_route,
                      const uint8_t *mask, int64_t N, int64_t H, int spherical, const or_trace_params *P,
                      const float *roots, int use_roots, int32_t *cl_cell, double *cl_delta, float *cl_vol,
                      float *cl_mom, int64_t *cl_count, int64_t *cl_lo, int64_t *cl_hi, int8_t *sp_death) {
    const int64_t NE = N + 2 * H;
    const int64_t cap = or_change_list_cap(P->max_steps, P->lateral_rate > 0.0);
    const double Nf = (double)N;
    const double inv_sat = P->disc_saturation > 0.0 ? 1.0 / P->disc_saturation : 0.0;
#define MASK(f, i, j) mask[((size_t)(f) * NE + (size_t)(i)) * NE + (size_t)(j)]
#define ROOTS(f, i, j) roots[((size_t)(f) * NE + (size_t)(i)) * NE + (size_t)(j)]
    int64_t f = sp_face[p];
    double x = sp_x[p], y = sp_y[p];
    double vol = volume0;
    double sed = sp_sed[p];
    double sa = 0.0, sb = 0.0;
    if (sed > 0.0) {
        if (or_sincos) {
            or_sincos(sp_dir[p], &sb, &sa);
        } else {
            sa = cos(sp_dir[p]);
            sb = sin(sp_dir[p]);
        }
    }
    int64_t base = p * cap;
    int64_t n = 0;
    int64_t uphill = 0;
    int64_t lo = 2147483647, hi = -1;
    int64_t rbase = base + cap - SPREAD;
    int64_t nring = 0, rpos = 0;
    double h_prev = -1e300;
    int exited = 0, in_sea = 0, in_lake = 0;
    int64_t parks = 0, sea_steps = 0;
    int cause = DEATH_AGE;
    for (int64_t step = 0; step < P->max_steps; step++) {
        int64_t ci = cell_of(x, N), cj = cell_of(y, N);
        int64_t ei = ci + H, ej = cj + H;
        uint8_t m = MASK(f, ei, ej);
        if (m == MASK_OUTSIDE) { exited = 1; cause = DEATH_EXIT; break; }
        int64_t cell = (f * NE + ei) * NE + ej;
        float h_here = SAMP(f, ei, ej, S_SURF);
        float b_here = SAMP(f, ei, ej, S_BASE);
        if (step == 0 && h_here < b_here) in_sea = 1;
        if (in_sea) {
            if (h_here >= b_here || sed * vol / volume0 < 1e-4 || sea_steps >= P->ocean_steps) {
                cause = DEATH_OCEAN;
                break;
            }
            sea_steps += 1;
        }
        if (m == MASK_ACTIVE) {
            cl_cell[rbase + rpos] = (int32_t)cell;
            rpos = (rpos + 1) % SPREAD;
            if (nring < SPREAD) nring += 1;
        }
        double fi = x - 0.5 + (double)H;
        double fj = y - 0.5 + (double)H;
        double h0, gi, gj;
        sbilin_grad(samp, NE, f, fi, fj, S_SURF, &h0, &gi, &gj);
        double h0r = h0;
        if (use_route && rflag(samp, NE, f, fi, fj)) {
            double r0, rgi, rgj;
            sbilin_grad(samp, NE, f, fi, fj, S_ROUTE, &r0, &rgi, &rgj);
            if (r0 > h0) { h0r = r0; gi = rgi; gj = rgj; }
        }
        double q0 = SAMP(f, ei, ej, S_GINV), q1 = SAMP(f, ei, ej, S_GINV + 1), q2 = SAMP(f, ei, ej, S_GINV + 2);
        double di = q0 * gi + q1 * gj;
        double dj = q1 * gi + q2 * gj;
        double s2 = gi * di + gj * dj;
        if (s2 < 0.0) s2 = 0.0;
        double nrm;
        if (P->slope_sat > 0.0) nrm = P->slope_gain / sqrt(s2 + P->slope_sat * P->slope_sat);
        else nrm = P->slope_gain / sqrt(1.0 + s2);
        double na = -di * nrm;
        double nb = -dj * nrm;
        double ma = sbilin(samp, NE, f, fi, fj, S_MOMA);
        double mb = sbilin(samp, NE, f, fi, fj, S_MOMB);
        double q = sbilin(samp, NE, f, fi, fj, S_DISCH);
        if (q < 0.0) q = 0.0;
        double vol_rel = vol / volume0;
        double psa = sa, psb = sb;
        double fr = 1.0 - P->dt * P->friction;
        if (fr < 0.0) fr = 0.0;
        sa *= fr;
        sb *= fr;
        double mlen = snorm(samp, NE, f, ei, ej, ma, mb);
        double slen = snorm(samp, NE, f, ei, ej, sa, sb);
        if (mlen > 1e-12 && slen > 1e-12) {
            double cosang = sdot(samp, NE, f, ei, ej, ma, mb, sa, sb) / (mlen * slen);
            double k = P->k_mom * cosang * P->mom_scale / (vol + P->mom_scale * q);
            sa += k * ma;
            sb += k * mb;
        }
        double acc = P->dt / (vol_rel * P->density);
        sa += acc * na;
        sb += acc * nb;
        slen = snorm(samp, NE, f, ei, ej, sa, sb);
        if (slen < 1e-12) { cause = DEATH_STOP; break; }
        sa /= slen;
        sb /= slen;
        double nx = x + sa, ny = y + sb;
        int64_t nf = f;
        if (spherical) {
            if (nx < 0.0 || nx >= Nf || ny < 0.0 || ny >= Nf) {
                int f2;
                double u2, v2, a2, b2;
                or_transfer_velocity((int)f, nx / Nf, ny / Nf, sa, sb, &f2, &u2, &v2, &a2, &b2);
                nf = f2;
                nx = u2 * Nf;
                ny = v2 * Nf;
                sa = a2;
                sb = b2;
                if (nx >= Nf) nx = Nf * 0.999999999;
                if (ny >= Nf) ny = Nf * 0.999999999;
                if (nx < 0.0) nx = 0.0;
                if (ny < 0.0) ny = 0.0;
            }
        } else {
            if (nx < (double)(-H) || nx >= Nf + (double)H || ny < (double)(-H) || ny >= Nf + (double)H) {
                exited = 1; cause = DEATH_EXIT; break;
            }
        }
        int64_t nci = (int64_t)floor(nx), ncj = (int64_t)floor(ny);
        int64_t nei = nci + H, nej = ncj + H;
        if (nei < 0 || nei >= NE || nej < 0 || nej >= NE) { exited = 1; cause = DEATH_EXIT; break; }
        uint8_t nm = MASK(nf, nei, nej);
        if (nm == MASK_OUTSIDE) { exited = 1; cause = DEATH_EXIT; break; }
        double nfi = nx - 0.5 + (double)H, nfj = ny - 0.5 + (double)H;
        double h1 = sbilin(samp, NE, nf, nfi, nfj, S_SURF);
        float b1 = SAMP(nf, nei, nej, S_BASE);
        double dh = h0 - (h1 > (double)b1 ? h1 : (double)b1);
        double h1r = h1;
        if (use_route && rflag(samp, NE, nf, nfi, nfj)) {
            double r1 = sbilin(samp, NE, nf, nfi, nfj, S_ROUTE);
            if (r1 > h1) h1r = r1;
        }
        double dhr = h0r - h1r;
        if (m == MASK_ACTIVE && in_sea) {
            double cdiff = -P->ocean_rate * sed;
            double room = h_prev > -1e300 ? h_prev - (double)h_here : P->fan_room;
            if (room < P->fan_room) room = P->fan_room;
            double lim = (double)b_here - P->dep_floor - (double)h_here;
            if (lim < room) room = lim;
            if (room < 0.0) room = 0.0;
            if (-cdiff * vol_rel > room) cdiff = -room / vol_rel;
            sed += cdiff;
            cl_cell[base + n] = (int32_t)cell;
            cl_delta[base + n] = -cdiff * vol_rel;
            cl_vol[base + n] = 0.0f;
            cl_mom[2 * (base + n)] = 0.0f;
            cl_mom[2 * (base + n) + 1] = 0.0f;
            n += 1;
            if (cell < lo) lo = cell;
            if (cell > hi) hi = cell;
        } else if (m == MASK_ACTIVE) {
            double qe;
            if (inv_sat <= 0.0) qe = 0.0;
            else if (P->disc_exponent == 0.5) qe = sqrt(q * inv_sat);
            else if (P->disc_exponent > 0.0) qe = pow(q * inv_sat, P->disc_exponent);
            else qe = erf(q * inv_sat);
            double c_eq = P->erodibility * dh * (1.0 + P->k_disc * qe);
            if (c_eq < 0.0) c_eq = 0.0;
            double k_e;
            if (c_eq > sed) {
                float sthick = SAMP(f, ei, ej, S_SED);
                double w_cover;
                if (P->cover_depth > 0.0) {
                    w_cover = (double)sthick / P->cover_depth;
                    if (w_cover > 1.0) w_cover = 1.0;
                } else {
                    w_cover = sthick > 0.0f ? 1.0 : 0.0;
                }
                k_e = P->deposition_rate * (w_cover + (1.0 - w_cover) * (1.0 - (double)SAMP(f, ei, ej, S_HARD)));
            } else {
                k_e = P->deposition_rate;
            }
            if (use_roots) k_e *= 1.0 - (double)ROOTS(f, ei, ej);
            double cdiff = k_e * (c_eq - sed);
            double trapped = 0.0;
            if (in_lake) {
                cdiff = 0.0;
            } else if (sed > 0.0 && (double)SAMP(nf, nei, nej, S_RFLAG) >= 1.5) {
                double trap = -P->lake_trap * sed;
                if (trap < cdiff) cdiff = trap;
                if ((double)SAMP(nf, nei, nej, S_RFLAG) >= 2.5) trapped = P->lake_trap * sed;
            }
            double kexp = 1.0 + q * P->conc_scale;
            if (cdiff > 0.0) {
                double ecap = dh / kexp;
                if (ecap > P->max_erode) ecap = P->max_erode;
                if (ecap > (double)h_here) ecap = h_here > 0.0f ? (double)h_here : 0.0;
                if (cdiff * vol_rel > ecap) cdiff = ecap / vol_rel;
            } else if (cdiff < 0.0) {
                double room = h_prev > -1e300 ? h_prev - (double)h_here : P->dep_floor;
                if (room < P->dep_floor) room = P->dep_floor;
                double dcap = room / kexp;
                if (-cdiff * vol_rel > dcap) cdiff = -dcap / vol_rel;
            }
            sed += cdiff;
            cl_cell[base + n] = (int32_t)cell;
            cl_delta[base + n] = -cdiff * vol_rel;
            cl_vol[base + n] = (float)vol;
            cl_mom[2 * (base + n)] = (float)(vol * sa);
            cl_mom[2 * (base + n) + 1] = (float)(vol * sb);
            n += 1;
            if (cell < lo) lo = cell;
            if (cell > hi) hi = cell;
            if (P->lake_park > 0.0 && trapped > 0.0 && parks < LAKE_PARKS) {
                double rest = trapped + cdiff;
                if (rest > 0.0) {
                    parks += 1;
                    sed -= rest;
                    int64_t lcell = (nf * NE + nei) * NE + nej;
                    cl_cell[base + n] = (int32_t)lcell;
                    cl_delta[base + n] = rest * vol_rel;
                    cl_vol[base + n] = -3.0f;
                    cl_mom[2 * (base + n)] = 0.0f;
                    cl_mom[2 * (base + n) + 1] = 0.0f;
                    n += 1;
                    if (lcell < lo) lo = lcell;
                    if (lcell > hi) hi = lcell;
                }
            }
            if (P->lateral_rate > 0.0 && cdiff > 0.0 && n + 2 * SPREAD + 2 < cap) {
                double cross = psa * sb - psb * sa;
                double turn = cross > 0.0 ? cross : -cross;
                if (turn > 1e-6) {
                    double sgn = cross > 0.0 ? 1.0 : -1.0;
                    double bx = x + sgn * sb;
                    double by = y - sgn * sa;
                    int64_t bei = (int64_t)floor(bx) + H, bej = (int64_t)floor(by) + H;
                    if (0 <= bei && bei < NE && 0 <= bej && bej < NE && MASK(f, bei, bej) == MASK_ACTIVE) {
                        float hb = SAMP(f, bei, bej, S_SURF);
                        float bank = hb - h_here; /* float32 op */
                        if (bank > 0.0f) {
                            double lat = P->lateral_rate * turn * cdiff;
                            double lcap = (double)bank / kexp;
                            if (lcap > P->max_erode) lcap = P->max_erode;
                            if (lat * vol_rel > lcap) lat = lcap / vol_rel;
                            if (lat > 0.0) {
                                sed += lat;
                                int64_t bcell = (f * NE + bei) * NE + bej;
                                cl_cell[base + n] = (int32_t)bcell;
                                cl_delta[base + n] = -lat * vol_rel;
                                cl_vol[base + n] = -2.0f;
                                cl_mom[2 * (base + n)] = 0.0f;
                                cl_mom[2 * (base + n) + 1] = 0.0f;
                                n += 1;
                                if (bcell < lo) lo = bcell;
                                if (bcell > hi) hi = bcell;
                            }
                        }
                    }
                }
            }
        }
        double ev = 1.0 - P->dt * P->evap_rate * (double)SAMP(f, ei, ej, S_EVAP);
        if (ev < 0.01) ev = 0.01;
        vol *= ev;
        sed /= ev;
        x = nx;
        y = ny;
        f = nf;
        h_prev = (double)h_here;
        in_lake = (double)SAMP(nf, nei, nej, S_RFLAG) >= 1.5;
        if (SAMP(nf, nei, nej, S_SURF) < SAMP(nf, nei, nej, S_BASE) && !in_sea) {
            in_sea = 1;
            if (sed <= 0.0) { cause = DEATH_OCEAN; break; }
        }
        if (vol < P->min_volume) { cause = in_sea ? DEATH_OCEAN : DEATH_EVAP; break; }
        if (dhr < 0.0 && !in_lake) {
            uphill += 1;
            if (uphill > P->pit_steps) { cause = in_sea ? DEATH_OCEAN : DEATH_PIT; break; }
        } else {
            uphill = 0;
        }
    }
    if (in_sea && cause == DEATH_AGE) cause = DEATH_OCEAN;
    if (in_sea && cause == DEATH_STOP) cause = DEATH_OCEAN;
    sp_death[p] = (int8_t)cause;
    if (!exited) {
        int64_t ci = cell_of(x, N), cj = cell_of(y, N);
        int64_t ei = ci + H, ej = cj + H;
        if (MASK(f, ei, ej) == MASK_ACTIVE && !(in_sea && SAMP(f, ei, ej, S_SURF) >= SAMP(f, ei, ej, S_BASE))) {
            int64_t cell = (f * NE + ei) * NE + ej;
            int64_t newest = (rpos + SPREAD - 1) % SPREAD;
            if (nring == 0 || cl_cell[rbase + newest] != cell) {
                cl_cell[rbase + rpos] = (int32_t)cell;
                rpos = (rpos + 1) % SPREAD;
                if (nring < SPREAD) nring += 1;
            }
        }
    }
    if (sed > 0.0 && nring > 0) {
        if (!exited || P->deposit_on_exit) {
            double load = sed * vol / volume0;
            int64_t nout = cause == DEATH_OCEAN ? 1 : nring;
            for (int64_t r = 0; r < nout; r++) {
                int64_t cell = cl_cell[rbase + (rpos + SPREAD - 1 - r) % SPREAD];
                cl_cell[base + n] = (int32_t)cell;
                cl_delta[base + n] = r == 0 ? load : 0.0;
                cl_vol[base + n] = -1.0f;
                cl_mom[2 * (base + n)] = 0.0f;
                cl_mom[2 * (base + n) + 1] = 0.0f;
                n += 1;
                if (cell < lo) lo = cell;
                if (cell > hi) hi = cell;
            }
        }
    }
    cl_count[p] = n;
    cl_lo[p] = lo;
    cl_hi[p] = hi;
#undef MASK
#undef ROOTS
}

/* particle.py:275-818: trace one chunk (P particles).  `nthreads` <= 0 keeps
 * OpenMP's default; dynamic scheduling in blocks of 4 like the reference's
 * parallel_chunksize(4) (maps.py:957).  Results do not depend on threads. */
void or_trace_particles(int64_t P, const int64_t *sp_face, const double *sp_x, const double *sp_y,
                        const double *sp_sed, const double *sp_dir, double volume0, const float *samp, int use_route,
                        const uint8_t *mask, int64_t N, int64_t H, int spherical, const or_trace_params *prm,
                        const float *roots, int use_roots, int32_t *cl_cell, double *cl_delta, float *cl_vol,
                        float *cl_mom, int64_t *cl_count, int64_t *cl_lo, int64_t *cl_hi, int8_t *sp_death,
                        int nthreads) {
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
#endif
    for (int64_t p = 0; p < P; p++)
        trace_one(p, sp_face, sp_x, sp_y, sp_sed, sp_dir, volume0, samp, use_route, mask, N, H, spherical, prm,
                  roots, use_roots, cl_cell, cl_delta, cl_vol, cl_mom, cl_count, cl_lo, cl_hi, sp_death);
    (void)nthreads;
}

/* particle.py:821-1042.  Arrays are flat views of (F, NE, NE[, 2]) state.
 * out[0..3] = (n_clamp, to_pending, lost, lost_offshore). */
void or_apply_changes(const int32_t *cl_cell, const double *cl_delta, const float *cl_vol, const float *cl_mom,
                      const int64_t *cl_count, int64_t P, int64_t cap, const double *hflat, const double *sflat,
                      double *aflat, double *pdflat, float *samp, double *dflat, double *mflat,
                      const uint8_t *kflat, double iter_erode, double iter_deposit, double fan_slope, int use_route,
                      double dep_floor, double offshore_writeoff, const double *ecflat, int use_erode_cap,
                      double *lkflat, const int32_t *idflat, double *lake_room, double *out) {
    int64_t n_clamp = 0;
    double to_pending = 0.0, lost = 0.0, lost_offshore = 0.0;
#define PF(c, k) samp[(size_t)(c) * NS + (k)]
    for (int64_t p = 0; p < P; p++) {
        int64_t base = p * cap;
        int64_t cnt = cl_count[p];
        double deficit = 0.0, surplus = 0.0, refused = 0.0;
        int64_t prev = -1, death = -1;
        for (int64_t k = 0; k < cnt; k++) {
            int64_t c = cl_cell[base + k];
            if (cl_vol[base + k] == -3.0f) {
                double take = cl_delta[base + k];
                double rest = 0.0;
                int64_t lk = idflat[c];
                if (lk < 0) {
                    rest = take;
                    take = 0.0;
                } else if (take > lake_room[lk]) {
                    rest = take - lake_room[lk];
                    take = lake_room[lk];
                }
                if (take > 0.0) {
                    lkflat[c] += take;
                    lake_room[lk] -= take;
                }
                refused += rest;
                continue;
            }
            if (kflat[c] != MASK_ACTIVE) continue;
            double d = cl_delta[base + k];
            float v = cl_vol[base + k];
            int is_step = v >= 0.0f;
            double s_c = hflat[c] + sflat[c] + aflat[c];
            float b_c = PF(c, S_BASE);
            if (d < 0.0) {
                double e = -d;
                double e_act = e;
                double room = iter_erode + aflat[c];
                if (use_erode_cap) {
                    double r2 = ecflat[c] + aflat[c];
                    if (r2 < room) room = r2;
                }
                if (e_act > room) e_act = room;
                double floor_ = s_c >= (double)b_c ? (double)b_c : s_c;
                if (use_route && (double)PF(c, S_RFLAG) >= 1.5) {
                    float r_c = PF(c, S_ROUTE);
                    if ((double)r_c > floor_) floor_ = (double)r_c;
                }
                if (k + 1 < cnt) {
                    int64_t nc = cl_cell[base + k + 1];
                    double ns = hflat[nc] + sflat[nc] + aflat[nc];
                    if (ns > floor_) floor_ = ns;
                }
                if (e_act > s_c - floor_) e_act = s_c - floor_;
                if (e_act < 0.0) e_act = 0.0;
                if (e_act < e) {
                    n_clamp += 1;
                    deficit += e - e_act;
                }
                aflat[c] -= e_act;
            } else {
                double want = d + surplus;
                surplus = 0.0;
                if (deficit > 0.0) {
                    double a = want < deficit ? want : deficit;
                    want -= a;
                    deficit -= a;
                }
                double d_act = want;
                double room = iter_deposit - aflat[c];
                if (d_act > room) d_act = room;
                int64_t ref = -1;
                if (is_step) ref = prev;
                else if (k + 1 < cnt) ref = cl_cell[base + k + 1];
                if (ref >= 0) {
                    double ceil_ = hflat[ref] + sflat[ref] + aflat[ref] + dep_floor;
                    if (s_c < (double)b_c && is_step) ceil_ -= dep_floor + fan_slope;
                    double lim = ceil_ - s_c;
                    if (d_act > lim) d_act = lim;
                }
                if (s_c < (double)b_c) {
                    double lim = (double)b_c - dep_floor - s_c;
                    if (d_act > lim) d_act = lim;
                }
                if (d_act < 0.0) d_act = 0.0;
                if (d_act < want) n_clamp += 1;
                surplus = want - d_act;
                aflat[c] += d_act;
                if (!is_step && death < 0) death = c;
            }
            double a2 = aflat[c];
            double s_new = hflat[c] + sflat[c] + a2;
            PF(c, S_SURF) = (float)s_new;
            double sd = sflat[c] + a2;
            PF(c, S_SED) = (float)(sd > 0.0 ? sd : 0.0);
            if (use_route && (double)PF(c, S_RFLAG) < 1.5)
                PF(c, S_RFLAG) = PF(c, S_ROUTE) != PF(c, S_SURF) ? 1.0f : 0.0f;
            if (is_step) {
                if (v > 0.0f) {
                    dflat[c] += (double)v;
                    mflat[2 * c] += (double)cl_mom[2 * (base + k)];
                    mflat[2 * c + 1] += (double)cl_mom[2 * (base + k) + 1];
                }
                prev = c;
            }
        }
        surplus += refused;
        if (surplus > 0.0) {
            if (death >= 0) {
                if (hflat[death] + sflat[death] + aflat[death] < (double)PF(death, S_BASE)) {
                    double keep = surplus * (1.0 - offshore_writeoff);
                    if (keep < 1e-4) keep = 0.0;
                    pdflat[death] += keep;
                    to_pending += keep;
                    lost_offshore += surplus - keep;
                } else {
                    pdflat[death] += surplus;
                    to_pending += surplus;
                }
            }
        }
        lost += deficit;
    }
#undef PF
    out[0] = (double)n_clamp;
    out[1] = to_pending;
    out[2] = lost;
    out[3] = lost_offshore;
}

/* particle.py:164-191 */
void or_pack_samples(float *samp, int64_t F, int64_t NE, const double *height, const double *sediment,
                     const double *route, int use_route, const double *discharge, const double *momentum,
                     const float *evap, const float *hardness, const float *metric, const float *metric_inv,
                     const double *base, const uint8_t *lake_flag) {
    int64_t n = F * NE * NE;
    for (int64_t c = 0; c < n; c++) {
        float *s = samp + (size_t)c * NS;
        double h = height[c], sd = sediment[c];
        double surf = h + sd;
        double r = use_route ? route[c] : surf;
        s[S_SURF] = (float)surf;
        s[S_ROUTE] = (float)r;
        s[S_DISCH] = (float)discharge[c];
        s[S_MOMA] = (float)momentum[2 * c];
        s[S_MOMB] = (float)momentum[2 * c + 1];
        s[S_EVAP] = evap[c];
        s[S_HARD] = hardness[c];
        s[S_SED] = (float)sd;
        for (int k = 0; k < 3; k++) {
            s[S_G + k] = metric[3 * c + k];
            s[S_GINV + k] = metric_inv[3 * c + k];
        }
        if (use_route && lake_flag[c] > 0) s[S_RFLAG] = lake_flag[c] > 1 ? 3.0f : 2.0f;
        else s[S_RFLAG] = (use_route && s[S_ROUTE] != s[S_SURF]) ? 1.0f : 0.0f;
        s[S_BASE] = (float)base[c];
    }
}

/* particle.py:1079-1101 */
void or_fold_changes(int64_t n, double *height, double *sediment, double *acc, const uint8_t *mask) {
    for (int64_t c = 0; c < n; c++) {
        double a = acc[c];
        acc[c] = 0.0;
        if (a == 0.0 || mask[c] != MASK_ACTIVE) continue;
        if (a > 0.0) {
            sediment[c] += a;
        } else {
            double take = -a;
            double s = sediment[c];
            if (s >= take) {
                sediment[c] = s - take;
            } else {
                height[c] -= take - s;
                sediment[c] = 0.0;
            }
        }
    }
}

/* particle.py:1230-1242 */
void or_ema_update(int64_t n, double *discharge, double *momentum, double *disch_track, double *mom_track,
                   const uint8_t *mask, double ema) {
    for (int64_t c = 0; c < n; c++) {
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

/* particle.py:1245-1260 */
void or_spawn_cells(const double *cdf, int64_t M, const double *u, int64_t P, int64_t *out) {
    double total = cdf[M - 1];
    int64_t k = 0;
    for (int64_t p = 0; p < P; p++) {
        double target = u[p] * total;
        while (k < M - 1 && cdf[k] <= target) k += 1;
        out[p] = k;
    }
}

/* ---------------------------------------------------------------------- */
/* mass wasting: thermal erosion and hillslope creep (particle.py:1104-1227) */
/* ---------------------------------------------------------------------- */
/* D8 neighbour order (particle.py:101-102); the opposite of k is (k + 4) % 8 */
static const int64_t OR_D8[8][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, -1}};

/* particle.py:1107-1125: raw outflow from cell c = (f, ei, ej) to its k-th
 * D8 neighbour.  numba types: metric float32 x int64 offsets -> float64. */
static double or_thermal_out(const double *h, const double *s, const float *metric, const double *talus, int64_t NE,
                             int64_t f, int64_t ei, int64_t ej, int k, double rate) {
    const int64_t di = OR_D8[k][0], dj = OR_D8[k][1];
    const int64_t c = (f * NE + ei) * NE + ej, n = (f * NE + ei + di) * NE + ej + dj;
    const double d = (h[c] + s[c]) - (h[n] + s[n]);
    if (d <= 0.0) return 0.0;
    const double g0 = metric[3 * c], g1 = metric[3 * c + 1], g2 = metric[3 * c + 2];
    const double ddi = (double)di, ddj = (double)dj;
    const double dist = sqrt(g0 * ddi * ddi + 2.0 * g1 * ddi * ddj + g2 * ddj * ddj);
    const double ex = d - talus[c] * dist;
    if (ex <= 0.0) return 0.0;
    return 0.5 * rate * ex;
}

/* particle.py:1128-1160: per-cell total outflow and cap scale */
void or_thermal_pass_a(int64_t F, int64_t NE, const double *h, const double *s, const float *metric, const double *talus,
                       const uint8_t *mask, double rate, double tmax, double *out_total, double *scale) {
    for (int64_t f = 0; f < F; f++)
        for (int64_t ei = 1; ei < NE - 1; ei++)
            for (int64_t ej = 1; ej < NE - 1; ej++) {
                const int64_t c = (f * NE + ei) * NE + ej;
                out_total[c] = 0.0;
                scale[c] = 0.0;
                if (mask[c] != MASK_ACTIVE) continue;
                double S = 0.0, mx = 0.0;
                for (int k = 0; k < 8; k++) {
                    if (mask[(f * NE + ei + OR_D8[k][0]) * NE + ej + OR_D8[k][1]] != MASK_ACTIVE) continue;
                    const double o = or_thermal_out(h, s, metric, talus, NE, f, ei, ej, k, rate);
                    S += o;
                    if (o > mx) mx = o;
                }
                if (S <= 0.0) continue;
                double sc = 1.0;
                if (S > 2.0 * mx) sc = 2.0 * mx / S;
                if (tmax > 0.0 && S * sc > tmax) sc = tmax / S;
                out_total[c] = S * sc;
                scale[c] = sc;
            }
}

/* particle.py:1163-1187: receiver scale of submerged cells */
void or_thermal_pass_r(int64_t F, int64_t NE, const double *h, const double *s, const float *metric, const double *talus,
                       const uint8_t *mask, double rate, const double *scale, double *rscale, double dep_floor) {
    for (int64_t f = 0; f < F; f++)
        for (int64_t ei = 1; ei < NE - 1; ei++)
            for (int64_t ej = 1; ej < NE - 1; ej++) {
                const int64_t c = (f * NE + ei) * NE + ej;
                rscale[c] = 1.0;
                const double surf = h[c] + s[c];
                if (surf >= 0.0 || mask[c] != MASK_ACTIVE) continue;
                double inflow = 0.0;
                for (int k = 0; k < 8; k++) {
                    const int64_t ni = ei + OR_D8[k][0], nj = ej + OR_D8[k][1];
                    const int64_t n = (f * NE + ni) * NE + nj;
                    const double sc = scale[n];
                    if (sc <= 0.0 || mask[n] == MASK_OUTSIDE) continue;
                    inflow += sc * or_thermal_out(h, s, metric, talus, NE, f, ni, nj, (k + 4) % 8, rate);
                }
                double room = -dep_floor - surf;
                if (room < 0.0) room = 0.0;
                if (inflow > room) rscale[c] = room / inflow;
            }
}

/* particle.py:1190-1227: gather of outflows (sediment first) and inflows */
void or_thermal_pass_b(int64_t F, int64_t NE, const double *h, const double *s, const float *metric, const double *talus,
                       const uint8_t *mask, double rate, const double *scale, const double *rscale, double *new_h,
                       double *new_s) {
    for (int64_t f = 0; f < F; f++)
        for (int64_t ei = 1; ei < NE - 1; ei++)
            for (int64_t ej = 1; ej < NE - 1; ej++) {
                const int64_t c = (f * NE + ei) * NE + ej;
                double hh = h[c], ss = s[c];
                if (mask[c] != MASK_ACTIVE) {
                    new_h[c] = hh;
                    new_s[c] = ss;
                    continue;
                }
                double inflow = 0.0, out = 0.0;
                const double sc_self = scale[c];
                for (int k = 0; k < 8; k++) {
                    const int64_t ni = ei + OR_D8[k][0], nj = ej + OR_D8[k][1];
                    const int64_t n = (f * NE + ni) * NE + nj;
                    if (mask[n] == MASK_OUTSIDE) continue;
                    if (sc_self > 0.0 && mask[n] == MASK_ACTIVE)
                        out += sc_self * rscale[n] * or_thermal_out(h, s, metric, talus, NE, f, ei, ej, k, rate);
                    const double sc = scale[n];
                    if (sc <= 0.0) continue;
                    inflow += sc * or_thermal_out(h, s, metric, talus, NE, f, ni, nj, (k + 4) % 8, rate);
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
                new_h[c] = hh;
                new_s[c] = ss + inflow;
            }
}

/* particle.py:1045-1076: soillib's pit-free erosion limit per cell */
void or_slope_erode_cap(int64_t F, int64_t NE, const double *h, const double *s, const uint8_t *mask, double coef,
                        double exit_slope, double *out) {
    const double k = coef * sqrt(2.0);
#define SURF(i, j) (h[(f * NE + (i)) * NE + (j)] + s[(f * NE + (i)) * NE + (j)])
    for (int64_t f = 0; f < F; f++)
        for (int64_t i = 0; i < NE; i++)
            for (int64_t j = 0; j < NE; j++) {
                const int64_t c = (f * NE + i) * NE + j;
                if (mask[c] != MASK_ACTIVE) {
                    out[c] = 0.0;
                    continue;
                }
                const double hc = h[c] + s[c];
                const double gxn = i == 0 ? exit_slope : hc - SURF(i - 1, j);
                const double gxp = i == NE - 1 ? exit_slope : hc - SURF(i + 1, j);
                const double gyn = j == 0 ? exit_slope : hc - SURF(i, j - 1);
                const double gyp = j == NE - 1 ? exit_slope : hc - SURF(i, j + 1);
                double gx = gxn > gxp ? gxn : gxp;
                double gy = gyn > gyp ? gyn : gyp;
                if (gx < 0.0) gx = 0.0;
                if (gy < 0.0) gy = 0.0;
                out[c] = k * sqrt(gx * gx + gy * gy);
            }
#undef SURF
}

/* ---------------------------------------------------------------------- */
/* Laplace-Beltrami operator, flux form (field.py:372-404): the raw kernel  */
/* (before field.py:187's division by the cell size squared); out is zero  */
/* on the outermost ring                                                    */
/* ---------------------------------------------------------------------- */
void or_laplacian(int64_t F, int64_t NE, const double *d, const double *ginv, const double *sq, double *out) {
#define D(i, j) d[(f * NE + (i)) * NE + (j)]
#define G(i, j, k) ginv[((f * NE + (i)) * NE + (j)) * 3 + (k)]
#define Q(i, j) sq[(f * NE + (i)) * NE + (j)]
    for (int64_t c = 0; c < F * NE * NE; c++) out[c] = 0.0;
    for (int64_t f = 0; f < F; f++)
        for (int64_t i = 1; i < NE - 1; i++)
            for (int64_t j = 1; j < NE - 1; j++) {
                const double fj_c = 0.5 * (D(i, j + 1) - D(i, j - 1));
                const double fj_p = 0.5 * (D(i + 1, j + 1) - D(i + 1, j - 1));
                const double fj_m = 0.5 * (D(i - 1, j + 1) - D(i - 1, j - 1));
                const double fi_c = 0.5 * (D(i + 1, j) - D(i - 1, j));
                const double fi_p = 0.5 * (D(i + 1, j + 1) - D(i - 1, j + 1));
                const double fi_m = 0.5 * (D(i + 1, j - 1) - D(i - 1, j - 1));
                double gii = 0.5 * (G(i, j, 0) + G(i + 1, j, 0));
                double gij = 0.5 * (G(i, j, 1) + G(i + 1, j, 1));
                double s = 0.5 * (Q(i, j) + Q(i + 1, j));
                const double Fi_p = s * (gii * (D(i + 1, j) - D(i, j)) + gij * 0.5 * (fj_c + fj_p));
                gii = 0.5 * (G(i, j, 0) + G(i - 1, j, 0));
                gij = 0.5 * (G(i, j, 1) + G(i - 1, j, 1));
                s = 0.5 * (Q(i, j) + Q(i - 1, j));
                const double Fi_m = s * (gii * (D(i, j) - D(i - 1, j)) + gij * 0.5 * (fj_c + fj_m));
                double gjj = 0.5 * (G(i, j, 2) + G(i, j + 1, 2));
                gij = 0.5 * (G(i, j, 1) + G(i, j + 1, 1));
                s = 0.5 * (Q(i, j) + Q(i, j + 1));
                const double Fj_p = s * (gjj * (D(i, j + 1) - D(i, j)) + gij * 0.5 * (fi_c + fi_p));
                gjj = 0.5 * (G(i, j, 2) + G(i, j - 1, 2));
                gij = 0.5 * (G(i, j, 1) + G(i, j - 1, 1));
                s = 0.5 * (Q(i, j) + Q(i, j - 1));
                const double Fj_m = s * (gjj * (D(i, j) - D(i, j - 1)) + gij * 0.5 * (fi_c + fi_m));
                out[(f * NE + i) * NE + j] = (Fi_p - Fi_m + Fj_p - Fj_m) / Q(i, j);
            }
#undef D
#undef G
#undef Q
}

/* ---------------------------------------------------------------------- */
/* routing surface: priority flood + epsilon (route.py:30-117)            */
/* ---------------------------------------------------------------------- */
typedef struct {
    double v;
    int64_t c;
} or_hitem;

static int or_hless(or_hitem a, or_hitem b) { return a.v < b.v || (a.v == b.v && a.c < b.c); }

static void or_hpush(or_hitem *h, int64_t *n, or_hitem x) {
    int64_t i = (*n)++;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (!or_hless(x, h[p])) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = x;
}

static or_hitem or_hpop(or_hitem *h, int64_t *n) {
    or_hitem top = h[0], x = h[--(*n)];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, m = i;
        or_hitem best = x;
        if (l < *n && or_hless(h[l], best)) { best = h[l]; m = l; }
        if (r < *n && or_hless(h[r], best)) { best = h[r]; m = r; }
        if (m == i) break;
        h[i] = h[m];
        i = m;
    }
    if (*n > 0) h[i] = x;
    return top;
}

/* route.py:30-117.  filled: out (F, NE, NE); visited: scratch uint8[F NE NE];
 * heap: scratch of F NE NE items (each cell is pushed at most once). */
void or_priority_flood_eps(int64_t F, int64_t NE, int64_t H, int64_t N, const double *surface, const uint8_t *mask,
                           const int64_t *owner, double eps, const double *base, double *filled, uint8_t *visited,
                           void *heap_mem) {
    const int64_t total = F * NE * NE;
    or_hitem *heap = (or_hitem *)heap_mem;
    int64_t hn = 0;
    for (int64_t c = 0; c < total; c++) filled[c] = surface[c];
    memset(visited, 0, (size_t)total);
    for (int64_t f = 0; f < F; f++)
        for (int64_t ei = 0; ei < NE; ei++)
            for (int64_t ej = 0; ej < NE; ej++) {
                const int64_t c = (f * NE + ei) * NE + ej;
                if (mask[c] == MASK_OUTSIDE || ei == 0 || ej == 0 || ei == NE - 1 || ej == NE - 1) visited[c] = 1;
            }
    for (int64_t f = 0; f < F; f++)
        for (int64_t ei = H; ei < H + N; ei++)
            for (int64_t ej = H; ej < H + N; ej++) {
                const int64_t c = (f * NE + ei) * NE + ej;
                if (visited[c]) continue;
                int seed = surface[c] < 0.0 && surface[c] <= base[c];
                if (!seed) {
                    for (int k = 0; k < 8; k++) {
                        const int64_t n = owner[(f * NE + ei + OR_D8[k][0]) * NE + ej + OR_D8[k][1]];
                        if (mask[n] == MASK_OUTSIDE) {
                            seed = 1;
                            break;
                        }
                        const int64_t nf = n / (NE * NE), rem = n - nf * NE * NE, nei = rem / NE, nej = rem - nei * NE;
                        if (nei < H || nei >= H + N || nej < H || nej >= H + N) {
                            seed = 1;
                            break;
                        }
                    }
                }
                if (seed) {
                    visited[c] = 1;
                    or_hitem it = {surface[c], c};
                    or_hpush(heap, &hn, it);
                }
            }
    while (hn > 0) {
        const or_hitem top = or_hpop(heap, &hn);
        const double h = top.v;
        const int64_t c = top.c, f = c / (NE * NE), rem = c - f * NE * NE, ei = rem / NE, ej = rem - ei * NE;
        for (int k = 0; k < 8; k++) {
            const int64_t n = owner[(f * NE + ei + OR_D8[k][0]) * NE + ej + OR_D8[k][1]];
            if (visited[n]) continue;
            visited[n] = 1;
            double v = surface[n];
            if (v < h + eps) v = h + eps;
            filled[n] = v;
            or_hitem it = {v, n};
            or_hpush(heap, &hn, it);
        }
    }
}

/* ---------------------------------------------------------------------- */
/* connected components of a mask of interior cells (bake/globe/hydro/     */
/* lakes.py:19-46 with the level all equal, as below_sea_components calls  */
/* it, hydro/run.py:64-83): 8-connected across faces through the owner     */
/* table (hydro/d8.py:71-77), numbered in order of their lowest cell id.   */
/* ---------------------------------------------------------------------- */
int64_t or_label_components(const uint8_t *mask, const int64_t *owner, int64_t N, int64_t H, int32_t *labels,
                            int64_t *stack) {
    const int64_t NE = N + 2 * H, M = 6 * N * N;
    for (int64_t c = 0; c < M; c++) labels[c] = -1;
    int64_t n = 0;
    for (int64_t s = 0; s < M; s++) {
        if (!mask[s] || labels[s] >= 0) continue;
        labels[s] = (int32_t)n;
        int64_t sp = 0;
        stack[sp++] = s;
        while (sp > 0) {
            const int64_t c = stack[--sp];
            const int64_t f = c / (N * N), r = c - f * N * N, i = r / N, j = r - i * N;
            for (int k = 0; k < 8; k++) {
                const int64_t o = owner[(f * NE + i + H + OR_D8[k][0]) * NE + j + H + OR_D8[k][1]];
                const int64_t f2 = o / (NE * NE), r2 = o - f2 * NE * NE;
                const int64_t q = (f2 * N + r2 / NE - H) * N + (r2 % NE) - H;
                if (mask[q] && labels[q] < 0) {
                    labels[q] = (int32_t)n;
                    stack[sp++] = q;
                }
            }
        }
        n += 1;
    }
    return n;
}
