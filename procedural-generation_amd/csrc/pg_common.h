// Shared device/host code of the erosion library: constants of the particle
// model, double-double arithmetic, correctly rounded tan/atan, and the
// equi-angular cube-sphere geometry the droplets use to cross faces.
//
// Bit parity.  The check mode must reproduce the reference's float64
// arithmetic (bake/globe/cubesphere.py:70-220 under glibc's libm).  Every
// expression here keeps the reference's operand order and the build uses
// -ffp-contract=off, so products and sums round exactly where the
// reference's do.  The only libm functions on the path are tan and atan;
// they are computed in double-double (about 104 significant bits) and rounded
// once, i.e. correctly rounded except when the true value lies within ~2^-100
// relative of a rounding boundary.  glibc's tan/atan are correctly rounded on
// the same inputs except in rare cases; both are checked against the
// reference's golden vector file and against glibc on random inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PG_HD __host__ __device__ __forceinline__

namespace pg {

// mask codes (particle.py:53-56)
constexpr int MASK_OUTSIDE = 0;
constexpr int MASK_ACTIVE = 1;
constexpr int MASK_FROZEN = 2;
// a dying droplet's overflow slots and lake parks (particle.py:62-67)
constexpr int SPREAD = 8;
constexpr int LAKE_PARKS = 8;
// why a droplet stopped (particle.py:69-76)
constexpr int DEATH_AGE = 0, DEATH_OCEAN = 1, DEATH_EXIT = 2, DEATH_PIT = 3, DEATH_EVAP = 4, DEATH_STOP = 5;
constexpr int N_DEATHS = 6;
// channels of the packed per-cell record (particle.py:149-161)
constexpr int S_SURF = 0, S_ROUTE = 1, S_DISCH = 2, S_MOMA = 3, S_MOMB = 4, S_EVAP = 5, S_HARD = 6, S_SED = 7,
              S_G = 8, S_GINV = 11, S_RFLAG = 14, S_BASE = 15, NS = 16;

// entries of a droplet's change-list slice (particle.py:258-272)
PG_HD int64_t change_list_cap(int64_t max_steps, bool lateral) {
    return (lateral ? 2 * max_steps : max_steps) + 2 * SPREAD + 1 + SPREAD + LAKE_PARKS;
}

// ---------------------------------------------------------------------------
// double-double arithmetic: a value is hi + lo with |lo| <= ulp(hi) / 2
// ---------------------------------------------------------------------------
struct dd {
    double hi, lo;
};

PG_HD dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
PG_HD dd quick_two_sum(double a, double b) {
    double s = a + b;
    return {s, b - (s - a)};
}
PG_HD dd two_prod(double a, double b) {
    double p = a * b;
    return {p, fma(a, b, -p)};
}
PG_HD dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}
PG_HD dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
PG_HD dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
PG_HD dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p.hi, p.lo);
}
PG_HD dd dd_div(dd a, dd b) {
    // long division: q1 + q2 + q3 with a remainder update per quotient digit
    double q1 = a.hi / b.hi;
    dd r = dd_sub(a, dd_mul(b, dd{q1, 0.0}));
    double q2 = r.hi / b.hi;
    r = dd_sub(r, dd_mul(b, dd{q2, 0.0}));
    double q3 = r.hi / b.hi;
    dd q = quick_two_sum(q1, q2);
    return dd_add(q, dd{q3, 0.0});
}

// Constant tables (generated with mpmath at 300 bits, rounded to double-double).
struct Tables {
    // atan(k / 64), k = 0..64
    static constexpr double atan_tab[65][2] = {
        {0.0, 0.0},
        {0.015623728620476831, -4.913600136566304e-19},
        {0.031239833430268277, -1.188442711587748e-18},
        {0.046840712915969654, -1.655677442254952e-19},
        {0.06241880999595735, -1.5490756308295046e-18},
        {0.0779666338315423, 5.804551873143357e-18},
        {0.09347678115858947, -6.2844725995420954e-18},
        {0.10894195698986579, 6.8267122072409585e-18},
        {0.12435499454676144, -3.1253241424539383e-18},
        {0.13970887428916365, -2.9579864247315813e-18},
        {0.15499674192394097, 9.585415594114324e-18},
        {0.1702119252854744, -3.541164079802125e-18},
        {0.18534794999569476, 4.180692268843079e-18},
        {0.2003985538258785, 3.1399542871844493e-18},
        {0.21535769969773805, 4.738160130078733e-19},
        {0.23021958727684372, 1.2313404529142703e-17},
        {0.24497866312686414, 1.0698755618734451e-17},
        {0.2596296294082575, 1.9238754924615304e-17},
        {0.2741674511196588, 8.261353575163773e-18},
        {0.2885873618940774, -1.428369957377257e-17},
        {0.3028848683749714, -1.1010827903001369e-17},
        {0.31705575320914703, -1.893928924292642e-17},
        {0.3310960767041321, -7.952610375793799e-18},
        {0.34500217720710513, -2.2938804755578304e-17},
        {0.35877067027057225, -2.4623815582638635e-17},
        {0.3723984466767542, 1.9612311504845653e-17},
        {0.38588266939807375, 2.378822732491941e-17},
        {0.39922076957525254, 2.246598105617042e-17},
        {0.4124104415973873, -1.587652227770689e-17},
        {0.42544963737004227, 2.3315530741892885e-17},
        {0.43833655985795783, -2.494277030626541e-17},
        {0.4510696559885235, -2.2703795229420475e-17},
        {0.4636476090008061, 2.2698777452961687e-17},
        {0.4760693303227612, 1.4654487332256713e-17},
        {0.48833395105640554, -1.1373236189329585e-17},
        {0.5004408131472942, -4.7181675085518756e-17},
        {0.5123894603107377, -2.5462781472855804e-17},
        {0.5241796287829132, 5.520094119641666e-18},
        {0.5358112379604637, -4.0637956834825575e-18},
        {0.5472843809874369, 4.923709671396255e-17},
        {0.5585993153435624, -5.4556305485916264e-18},
        {0.5697564534829784, 1.2255062085054184e-17},
        {0.5807563535676704, -1.441464378193067e-17},
        {0.5915997103351114, 4.920495453686772e-17},
        {0.6022873461349642, 2.950430737228402e-17},
        {0.6128202021652414, -3.1552061848586226e-17},
        {0.6231993299340659, 2.672403885140095e-17},
        {0.6334258829691446, -2.7290767436015276e-17},
        {0.6435011087932844, 1.5834785051444286e-17},
        {0.6534263411807619, 3.5800634857340095e-17},
        {0.6632029927060933, -3.076054864429649e-17},
        {0.6728325475937632, -1.899315009714705e-17},
        {0.6823165548747481, 6.943223671560008e-18},
        {0.6916566218531999, -8.117151192285796e-18},
        {0.7008544078844502, -1.987626234335816e-17},
        {0.7099116184635249, -4.597166450584887e-17},
        {0.7188299996216245, -2.1478388444456983e-17},
        {0.7276113326265107, 2.569325697391839e-18},
        {0.7362574289814281, 3.473937648299457e-17},
        {0.7447701257160751, 3.708315849135547e-17},
        {0.7531512809621944, -2.4256934659182068e-17},
        {0.7614027698055784, 9.850030332752822e-18},
        {0.7695264804056583, -3.704991905602721e-17},
        {0.7775243103733478, -2.6676490951944502e-17},
        {0.7853981633974483, 3.061616997868383e-17},
    };
    // tan(k / 32), k = 0..40
    static constexpr double tan_tab[41][2] = {
        {0.0, 0.0},
        {0.031260176501255954, 2.3803927834128917e-18},
        {0.06258150756627502, -6.767531942779112e-18},
        {0.09402562724573195, 6.5762523927164225e-18},
        {0.12565513657513097, -3.439960680352485e-18},
        {0.15753410732527162, -1.1061539275293055e-17},
        {0.18972861071805913, 9.78955701743985e-19},
        {0.22230728055343132, 6.328821377607695e-18},
        {0.25534192122103627, -5.589358343564783e-18},
        {0.2889081724405147, 2.70817328270223e-17},
        {0.32308624435174554, -2.389398809095344e-17},
        {0.357961738848017, -2.086600346579411e-17},
        {0.39362657592563277, -1.2972931096830582e-17},
        {0.43018004746423005, -4.113275164307763e-18},
        {0.4677300254523918, 1.5519502793263498e-17},
        {0.5063943574962299, -4.0096068550680074e-17},
        {0.5463024898437905, 2.9096576216837176e-17},
        {0.5875973675914432, 1.211599079372634e-17},
        {0.6304376738358848, -1.5183875765700744e-17},
        {0.6750004851442429, -2.666904806431612e-17},
        {0.7214844409909045, -5.4243984813454325e-17},
        {0.7701135513442087, 3.6147512759166313e-17},
        {0.8211418015898941, -1.6135619305114956e-17},
        {0.8748587605544823, 4.305785584054271e-17},
        {0.9315964599440725, -1.3547381396593036e-17},
        {0.9917378983632686, 3.625934281680031e-17},
        {1.05572763941192, -3.117795528517418e-17},
        {1.1240851347045608, 1.041957172133486e-16},
        {1.197421629234348, -5.600854085021751e-17},
        {1.2764618289823835, -2.7516464280290155e-17},
        {1.3620719763762281, 1.0932946789732315e-16},
        {1.4552966624690729, 5.552137562595018e-17},
        {1.5574077246549023, -6.186464176037592e-17},
        {1.6699701303536016, 2.3349811023765545e-17},
        {1.794932157265411, 1.3094532666512137e-18},
        {1.934751011916104, -6.676528971124806e-17},
        {2.092571276372179, 1.2547979245442145e-16},
        {2.272484060247449, 1.7137874176760215e-16},
        {2.4799129175567587, 1.0330706833705156e-16},
        {2.722205296368711, -1.8685788855683481e-16},
        {3.0095696738628313, 3.6503648506807936e-17},
    };
    // 1 / (2n + 1), n = 0..11
    static constexpr double inv_odd[12][2] = {
        {1.0, 0.0},
        {0.3333333333333333, 1.850371707708594e-17},
        {0.2, -1.1102230246251566e-17},
        {0.14285714285714285, 7.93016446160826e-18},
        {0.1111111111111111, 6.1679056923619804e-18},
        {0.09090909090909091, -2.523234146875356e-18},
        {0.07692307692307693, -4.270088556250602e-18},
        {0.06666666666666667, 9.251858538542971e-19},
        {0.058823529411764705, 8.163404592832033e-19},
        {0.05263157894736842, 2.921639538487254e-18},
        {0.047619047619047616, 2.64338815386942e-18},
        {0.043478260869565216, 1.206764157201257e-18},
    };
    // 1 / n!, n = 0..29
    static constexpr double inv_fact[30][2] = {
        {1.0, 0.0},
        {1.0, 0.0},
        {0.5, 0.0},
        {0.16666666666666666, 9.25185853854297e-18},
        {0.041666666666666664, 2.3129646346357427e-18},
        {0.008333333333333333, 1.1564823173178714e-19},
        {0.001388888888888889, -5.300543954373577e-20},
        {0.0001984126984126984, 1.7209558293420705e-22},
        {2.48015873015873e-05, 2.1511947866775882e-23},
        {2.7557319223985893e-06, -1.858393274046472e-22},
        {2.755731922398589e-07, 2.3767714622250297e-23},
        {2.505210838544172e-08, -1.448814070935912e-24},
        {2.08767569878681e-09, -1.20734505911326e-25},
        {1.6059043836821613e-10, 1.2585294588752098e-26},
        {1.1470745597729725e-11, 2.0655512752830745e-28},
        {7.647163731819816e-13, 7.03872877733453e-30},
        {4.779477332387385e-14, 4.399205485834081e-31},
        {2.8114572543455206e-15, 1.6508842730861433e-31},
        {1.5619206968586225e-16, 1.1910679660273754e-32},
        {8.22063524662433e-18, 2.2141894119604265e-34},
        {4.110317623312165e-19, 1.4412973378659527e-36},
        {1.9572941063391263e-20, -1.3643503830087908e-36},
        {8.896791392450574e-22, -7.911402614872376e-38},
        {3.868170170630684e-23, -8.843177655482344e-40},
        {1.6117375710961184e-24, -3.6846573564509766e-41},
        {6.446950284384474e-26, -1.9330404233703465e-42},
        {2.4795962632247976e-27, -1.2953730964765229e-43},
        {9.183689863795546e-29, 1.4303150396787322e-45},
        {3.279889237069838e-30, 1.5117542744029879e-46},
        {1.1309962886447716e-31, 1.0498015412959506e-47},
    };
    static constexpr double half_pi[2] = {1.5707963267948966, 6.123233995736766e-17};
};

PG_HD dd tab(const double (&t)[2]) { return {t[0], t[1]}; }

// atan(x), correctly rounded with overwhelming probability.  Reduction:
// |x| > 1 -> pi/2 - atan(1/|x|); then y = c + (y - c) with c = k/64 nearest
// and atan(y) = atan(c) + atan((y - c) / (1 + y c)), |t| <= 2^-7, series of
// t - t^3/3 + ... in double-double (12 terms: t^24 < 2^-168).
PG_HD double pg_atan(double x) {
    if (x != x) return x;
    bool neg = x < 0.0;
    double ax = neg ? -x : x;
    if (ax == 0.0) return x;
    if (ax > 1.0e300) return neg ? -Tables::half_pi[0] : Tables::half_pi[0];
    bool inv = ax > 1.0;
    dd y = inv ? dd_div(dd{1.0, 0.0}, dd{ax, 0.0}) : dd{ax, 0.0};
    int k = (int)(y.hi * 64.0 + 0.5);
    if (k > 64) k = 64;
    dd c = {k / 64.0, 0.0};
    dd t = dd_div(dd_sub(y, c), dd_add(dd{1.0, 0.0}, dd_mul(y, c)));
    dd t2 = dd_mul(t, t);
    // atan(t) / t = 1 - t^2/3 + t^4/5 - ..., Horner in t^2
    dd s = tab(Tables::inv_odd[11]);
    for (int n = 10; n >= 0; n--) s = dd_sub(tab(Tables::inv_odd[n]), dd_mul(t2, s));
    dd r = dd_add(tab(Tables::atan_tab[k]), dd_mul(t, s));
    if (inv) r = dd_sub(tab(Tables::half_pi), r);
    double out = r.hi + r.lo;
    return neg ? -out : out;
}

// tan(x) for |x| <= 1.25 (the face-crossing path needs |x| < pi/4 + 2 pi/N),
// correctly rounded with overwhelming probability.  x = c + d with
// c = k/32 nearest (d exact, |d| <= 2^-6); tan(d) = sin(d)/cos(d) by their
// series in double-double; tan(x) = (T + t) / (1 - T t), T = tan(c).
PG_HD double pg_tan(double x) {
    if (x != x) return x;
    bool neg = x < 0.0;
    double ax = neg ? -x : x;
    if (ax == 0.0) return x;
    int k = (int)(ax * 32.0 + 0.5);
    if (k > 40) k = 40;
    double d = ax - k / 32.0;  // exact: |d| < 2^-5 and both operands are multiples of ulp(ax)
    dd dv = {d, 0.0};
    dd d2 = dd_mul(dv, dv);
    // sin(d)/d = sum (-1)^n d^2n / (2n+1)!, cos(d) = sum (-1)^n d^2n / (2n)!
    dd ss = tab(Tables::inv_fact[17]);
    for (int n = 15; n >= 1; n -= 2) ss = dd_sub(tab(Tables::inv_fact[n]), dd_mul(d2, ss));
    dd cc = tab(Tables::inv_fact[16]);
    for (int n = 14; n >= 0; n -= 2) cc = dd_sub(tab(Tables::inv_fact[n]), dd_mul(d2, cc));
    dd t = dd_div(dd_mul(dv, ss), cc);
    dd T = tab(Tables::tan_tab[k]);
    dd r = dd_div(dd_add(T, t), dd_sub(dd{1.0, 0.0}, dd_mul(T, t)));
    double out = r.hi + r.lo;
    return neg ? -out : out;
}

// sin(x) and cos(x) for |x| < 2^20, correctly rounded with overwhelming
// probability away from the zeros: x = k pi/2 + r with pi/2 in triple-double,
// |r| <= pi/4, then the Taylor series of sin r and cos r in double-double.
PG_HD void pg_sincos(double x, double* os, double* oc) {
    const double p1 = 1.5707963267948966, p2 = 6.123233995736766e-17, p3 = -1.4973849048591698e-33;
    double kf = rint(x * 0.6366197723675814);
    dd r = dd_sub(dd{x, 0.0}, two_prod(kf, p1));
    r = dd_sub(r, two_prod(kf, p2));
    r = dd_sub(r, dd{kf * p3, 0.0});
    dd r2 = dd_mul(r, r);
    // |r| <= pi/4: r^29 / 29! < 2^-110
    dd ss = tab(Tables::inv_fact[29]);
    for (int n = 27; n >= 1; n -= 2) ss = dd_sub(tab(Tables::inv_fact[n]), dd_mul(r2, ss));
    ss = dd_mul(r, ss);
    dd cc = tab(Tables::inv_fact[28]);
    for (int n = 26; n >= 0; n -= 2) cc = dd_sub(tab(Tables::inv_fact[n]), dd_mul(r2, cc));
    double s = ss.hi + ss.lo, c = cc.hi + cc.lo;
    int q = ((int)(int64_t)kf) & 3;
    if (q == 0) { *os = s; *oc = c; }
    else if (q == 1) { *os = c; *oc = -s; }
    else if (q == 2) { *os = -s; *oc = -c; }
    else { *os = -c; *oc = s; }
}

// ---------------------------------------------------------------------------
// cube-sphere geometry (bake/globe/cubesphere.py:37-220), float64
// ---------------------------------------------------------------------------
struct Bases {
    // (right, up, normal) per face; faces +X, -X, +Y, -Y, +Z, -Z
    static constexpr double b[6][3][3] = {
        {{0.0, 0.0, -1.0}, {0.0, -1.0, 0.0}, {1.0, 0.0, 0.0}},
        {{0.0, 0.0, 1.0}, {0.0, -1.0, 0.0}, {-1.0, 0.0, 0.0}},
        {{1.0, 0.0, 0.0}, {0.0, 0.0, 1.0}, {0.0, 1.0, 0.0}},
        {{1.0, 0.0, 0.0}, {0.0, 0.0, -1.0}, {0.0, -1.0, 0.0}},
        {{1.0, 0.0, 0.0}, {0.0, -1.0, 0.0}, {0.0, 0.0, 1.0}},
        {{-1.0, 0.0, 0.0}, {0.0, -1.0, 0.0}, {0.0, 0.0, -1.0}},
    };
};
constexpr double PI_D = 3.141592653589793;
constexpr double HALF_PI = PI_D / 2.0;
constexpr double TWO_OVER_PI = 2.0 / PI_D;

struct Vec3 {
    double x, y, z;
};

// tan/atan used by the geometry: correctly rounded (check mode, bit parity)
// or the device libm (fast mode: a face crossing is rare and within a
// tolerance; the double-double path costs ~500 flops per call).
struct CRMath {
    static PG_HD double tan(double x) { return pg_tan(x); }
    static PG_HD double atan(double x) { return pg_atan(x); }
};
struct LibMath {
    static PG_HD double tan(double x) { return ::tan(x); }
    static PG_HD double atan(double x) { return ::atan(x); }
};

template <class M = CRMath>
PG_HD Vec3 to_sphere(int face, double u, double v) {
    double s = M::tan((u - 0.5) * HALF_PI);
    double t = M::tan((v - 0.5) * HALF_PI);
    const double(&B)[3][3] = Bases::b[face];
    double x = B[2][0] + s * B[0][0] + t * B[1][0];
    double y = B[2][1] + s * B[0][1] + t * B[1][1];
    double z = B[2][2] + s * B[0][2] + t * B[1][2];
    double inv = 1.0 / sqrt(x * x + y * y + z * z);
    return {x * inv, y * inv, z * inv};
}

PG_HD int face_of(double x, double y, double z) {
    double ax = fabs(x), ay = fabs(y), az = fabs(z);
    if (ax >= ay && ax >= az) return x >= 0.0 ? 0 : 1;
    if (ay >= az) return y >= 0.0 ? 2 : 3;
    return z >= 0.0 ? 4 : 5;
}

struct FaceUV {
    int face;
    double u, v;
};

template <class M = CRMath>
PG_HD FaceUV from_sphere(double x, double y, double z) {
    int face = face_of(x, y, z);
    const double(&B)[3][3] = Bases::b[face];
    double d = x * B[2][0] + y * B[2][1] + z * B[2][2];
    double s = (x * B[0][0] + y * B[0][1] + z * B[0][2]) / d;
    double t = (x * B[1][0] + y * B[1][1] + z * B[1][2]) / d;
    return {face, M::atan(s) * TWO_OVER_PI + 0.5, M::atan(t) * TWO_OVER_PI + 0.5};
}

// dp/du and dp/dv on the unit sphere: j[0..2] = J_u, j[3..5] = J_v
template <class M = CRMath>
PG_HD void jacobian(int face, double u, double v, double* j) {
    double s = M::tan((u - 0.5) * HALF_PI);
    double t = M::tan((v - 0.5) * HALF_PI);
    double dsdu = HALF_PI * (1.0 + s * s);
    double dtdv = HALF_PI * (1.0 + t * t);
    const double(&B)[3][3] = Bases::b[face];
    double qx = B[2][0] + s * B[0][0] + t * B[1][0];
    double qy = B[2][1] + s * B[0][1] + t * B[1][1];
    double qz = B[2][2] + s * B[0][2] + t * B[1][2];
    double ql = sqrt(qx * qx + qy * qy + qz * qz);
    double px = qx / ql, py = qy / ql, pz = qz / ql;
    double dux = B[0][0] * dsdu, duy = B[0][1] * dsdu, duz = B[0][2] * dsdu;
    double pd = px * dux + py * duy + pz * duz;
    j[0] = (dux - px * pd) / ql;
    j[1] = (duy - py * pd) / ql;
    j[2] = (duz - pz * pd) / ql;
    double dvx = B[1][0] * dtdv, dvy = B[1][1] * dtdv, dvz = B[1][2] * dtdv;
    pd = px * dvx + py * dvy + pz * dvz;
    j[3] = (dvx - px * pd) / ql;
    j[4] = (dvy - py * pd) / ql;
    j[5] = (dvz - pz * pd) / ql;
}

// Re-express a tangent vector (a, b) given in parameter components on
// (fs, us, vs) as components on (fd, ud, vd): the 2x2 normal equations.
template <class M = CRMath>
PG_HD void transfer_vector(int fs, double us, double vs, double a, double b, int fd, double ud, double vd,
                           double* oa, double* ob) {
    double J[6], K[6];
    jacobian<M>(fs, us, vs, J);
    double wx = a * J[0] + b * J[3];
    double wy = a * J[1] + b * J[4];
    double wz = a * J[2] + b * J[5];
    jacobian<M>(fd, ud, vd, K);
    double guu = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
    double guv = K[0] * K[3] + K[1] * K[4] + K[2] * K[5];
    double gvv = K[3] * K[3] + K[4] * K[4] + K[5] * K[5];
    double ru = K[0] * wx + K[1] * wy + K[2] * wz;
    double rv = K[3] * wx + K[4] * wy + K[5] * wz;
    double det = guu * gvv - guv * guv;
    *oa = (gvv * ru - guv * rv) / det;
    *ob = (guu * rv - guv * ru) / det;
}

struct Transfer {
    int face;
    double u, v, a, b;
};

// A droplet position (u, v) that may have left [0, 1)^2 on `face`, with its
// velocity: the face, position and velocity components it has on the face it
// is actually on (cubesphere.py:204-220).
template <class M = CRMath>
PG_HD Transfer transfer_velocity(int face, double u, double v, double a, double b) {
    if (u >= 0.0 && u < 1.0 && v >= 0.0 && v < 1.0) return {face, u, v, a, b};
    Vec3 p = to_sphere<M>(face, u, v);
    FaceUV q = from_sphere<M>(p.x, p.y, p.z);
    if (q.face == face) {
        double u2 = 0.0 > q.u ? 0.0 : q.u;
        u2 = 0.999999999 < u2 ? 0.999999999 : u2;
        double v2 = 0.0 > q.v ? 0.0 : q.v;
        v2 = 0.999999999 < v2 ? 0.999999999 : v2;
        return {face, u2, v2, a, b};
    }
    double a2, b2;
    transfer_vector<M>(face, u, v, a, b, q.face, q.u, q.v, &a2, &b2);
    return {q.face, q.u, q.v, a2, b2};
}

// ---------------------------------------------------------------------------
// The face transfer in single precision, for the fast mode (whose droplets
// carry float32 positions and directions): transfer_velocity's geometry with
// float tan / atan.  Two things differ from the double version, neither in
// value: the projected point is not normalised (face_of and the projection
// ratios do not depend on its length), and the destination face's Jacobian
// takes the projection ratios (s2, t2) directly instead of
// tan(atan(s2)) -- 2 tan and 2 atan per crossing instead of 6 and 2, and
// ~50 fewer registers in the trace loop.
// ---------------------------------------------------------------------------
struct TransferF {
    int face;
    float u, v, a, b;
};

PG_HD int face_of_f(float x, float y, float z) {
    const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
    if (ax >= ay && ax >= az) return x >= 0.0f ? 0 : 1;
    if (ay >= az) return y >= 0.0f ? 2 : 3;
    return z >= 0.0f ? 4 : 5;
}

// dp/du, dp/dv on the unit sphere at the point with projection ratios (s, t)
// on `face`: j[0..2] = J_u, j[3..5] = J_v
PG_HD void jacobian_st_f(int face, float s, float t, float* j) {
    constexpr float HP = 1.57079632679489662f;
    const float dsdu = HP * (1.0f + s * s), dtdv = HP * (1.0f + t * t);
    const double(&B)[3][3] = Bases::b[face];
    const float qx = (float)B[2][0] + s * (float)B[0][0] + t * (float)B[1][0];
    const float qy = (float)B[2][1] + s * (float)B[0][1] + t * (float)B[1][1];
    const float qz = (float)B[2][2] + s * (float)B[0][2] + t * (float)B[1][2];
    const float il = 1.0f / sqrtf(qx * qx + qy * qy + qz * qz);
    const float px = qx * il, py = qy * il, pz = qz * il;
    const float dux = (float)B[0][0] * dsdu, duy = (float)B[0][1] * dsdu, duz = (float)B[0][2] * dsdu;
    float pd = px * dux + py * duy + pz * duz;
    j[0] = (dux - px * pd) * il;
    j[1] = (duy - py * pd) * il;
    j[2] = (duz - pz * pd) * il;
    const float dvx = (float)B[1][0] * dtdv, dvy = (float)B[1][1] * dtdv, dvz = (float)B[1][2] * dtdv;
    pd = px * dvx + py * dvy + pz * dvz;
    j[3] = (dvx - px * pd) * il;
    j[4] = (dvy - py * pd) * il;
    j[5] = (dvz - pz * pd) * il;
}

PG_HD TransferF transfer_velocity_f(int face, float u, float v, float a, float b) {
    if (u >= 0.0f && u < 1.0f && v >= 0.0f && v < 1.0f) return {face, u, v, a, b};
    constexpr float HP = 1.57079632679489662f, TWO_OVER_PI_F = 0.636619772367581343f;
    const float s = tanf((u - 0.5f) * HP), t = tanf((v - 0.5f) * HP);
    const double(&B)[3][3] = Bases::b[face];
    const float x = (float)B[2][0] + s * (float)B[0][0] + t * (float)B[1][0];
    const float y = (float)B[2][1] + s * (float)B[0][1] + t * (float)B[1][1];
    const float z = (float)B[2][2] + s * (float)B[0][2] + t * (float)B[1][2];
    const int fd = face_of_f(x, y, z);
    const double(&D)[3][3] = Bases::b[fd];
    const float d = x * (float)D[2][0] + y * (float)D[2][1] + z * (float)D[2][2];
    const float s2 = (x * (float)D[0][0] + y * (float)D[0][1] + z * (float)D[0][2]) / d;
    const float t2 = (x * (float)D[1][0] + y * (float)D[1][1] + z * (float)D[1][2]) / d;
    float u2 = atanf(s2) * TWO_OVER_PI_F + 0.5f, v2 = atanf(t2) * TWO_OVER_PI_F + 0.5f;
    if (fd == face) {
        u2 = fminf(fmaxf(u2, 0.0f), 0.99999994f);
        v2 = fminf(fmaxf(v2, 0.0f), 0.99999994f);
        return {face, u2, v2, a, b};
    }
    float J[6], K[6];
    jacobian_st_f(face, s, t, J);
    jacobian_st_f(fd, s2, t2, K);
    const float wx = a * J[0] + b * J[3], wy = a * J[1] + b * J[4], wz = a * J[2] + b * J[5];
    const float guu = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
    const float guv = K[0] * K[3] + K[1] * K[4] + K[2] * K[5];
    const float gvv = K[3] * K[3] + K[4] * K[4] + K[5] * K[5];
    const float ru = K[0] * wx + K[1] * wy + K[2] * wz;
    const float rv = K[3] * wx + K[4] * wy + K[5] * wz;
    const float det = guu * gvv - guv * guv;
    return {fd, u2, v2, (gvv * ru - guv * rv) / det, (guu * rv - guv * ru) / det};
}

}  // namespace pg
