// rh_math.hpp -- field integrals of uniformly magnetized elements.
//
// Shared by the gfx950 kernels and by the host build (RH_HD = host+device).
// Everything here is per (observation point, element) and branch-uniform
// across a wavefront, because every lane of a wave sweeps the same element
// for its own point.
//
// Formulation (own design; the reference evaluates each polygon face in its
// own 2D frame edge by edge, cpp/src/core/radplnr2.cpp:29-386).  For a convex
// polyhedron with outward unit face normals n_f, surface charge
// sigma_f = n_f . M, the field per unit magnetization is the symmetric block
//
//     Q(p) = 1/(4 pi) * [ sum_edges L_e W_e  -  sum_faces Omega_f n_f n_f^T ]
//
// with  L_e     = integral over the edge of dl/|x - p|
//                = ln((R_b + s_b)/(R_a + s_a)), s = (v - p).t_e,
//       W_e     = sum over the two faces meeting at e of m_{e,f} n_f^T,
//                 m_{e,f} = in-plane outward normal of face f at edge e,
//       Omega_f = signed solid angle integral of (x-p).n_f/|x-p|^3 over f,
//                 from the Van Oosterom-Strackee triangle formula with the
//                 fan of each face multiplied as complex numbers so that a
//                 face costs one atan2.
// W_e is stored symmetrized: the antisymmetric parts cancel in the sum since
// Q is a Hessian.  Each edge's log is shared by both of its faces, so a
// tetrahedron costs 4 sqrt, 6 log, 6 div and 4 atan2 per point.
//
// The rectangular block uses the same decomposition specialised to an
// axis-aligned box: 3 logs (one per off-diagonal pair, as products of edge
// ratios), 6 atan2 (one per face), 8 sqrt -- the quantities of the closed form
// in cpp/src/core/radrec.cpp:49-315.
//
// H = Q M, and B = H + M when the point is strictly inside the element
// (cpp/src/core/radvlpgn.cpp:497-503, radrec.cpp:307-312).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define RH_HD __host__ __device__ __forceinline__

namespace rh {

constexpr double kInv4Pi = 0.079577471545947667884441881686257181;  // 1/(4 pi)

// Element record kinds in the packed element data.
enum ElemKind : int32_t { EK_REC = 0, EK_TET = 1, EK_GEN = 3 };

// Record sizes (doubles).
constexpr int kRecSize = 6;             // centre(3), half-dims(3)
constexpr int kTetEdge = 10;            // t(3), len, W sym (xx yy zz xy xz yz)
constexpr int kTetFaceRec = 4;         // n(3), 2*area
constexpr int kTetSize = 12 + 6 * kTetEdge + 4 * kTetFaceRec;  // verts, edges, faces = 88
constexpr int kGenHdr = 4;              // nv, ne, nf, nfv
constexpr int kGenEdge = 12;            // a, b, t(3), len, W sym(6)
constexpr int kGenFace = 5;             // n(3), fv begin, fv count

// Canonical tetrahedron topology: vertices 0..3 positively oriented
// (det(v1-v0, v2-v0, v3-v0) > 0); faces counter-clockwise about the outward
// normal; edge k joins kTetEdgeA[k] -> kTetEdgeB[k].
constexpr int kTetEdgeA[6] = {0, 0, 0, 1, 1, 2};
constexpr int kTetEdgeB[6] = {1, 2, 3, 2, 3, 3};
constexpr int kTetFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

template <class R>
struct V3 {
    R x, y, z;
};
template <class R>
RH_HD V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R>
RH_HD R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// a^2 + b^2 + c^2 with its fused multiply-adds written out (the same bits in
// every kernel)
RH_HD double sq3(double a, double b, double c) { return fma(a, a, fma(b, b, c * c)); }
RH_HD float sq3(float a, float b, float c) { return fmaf(a, a, fmaf(b, b, c * c)); }
template <class R>
RH_HD V3<R> cross(const V3<R>& a, const V3<R>& b)
{
    return V3<R>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

RH_HD double rh_log(double x) { return log(x); }
RH_HD float rh_log(float x) { return logf(x); }
// FP32 square root and reciprocal: the hardware instructions (about 1 ulp) on
// the device instead of the correctly rounded library sequences.
RH_HD float rh_sqrt(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
RH_HD float rh_rcp(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
// FP32 atan2: the octant reduction of the FP64 form below with a degree-4
// polynomial in t^2 (Chebyshev interpolant of (atan(t)/t - 1)/t^2 on
// |t| <= tan(pi/8), within 1.8e-8 of atan t in float arithmetic).
RH_HD float rh_atan2(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const bool big = mn > 0.41421356f * mx;
    const float t0 = (big ? mn - mx : mn) * rh_rcp(big ? mn + mx : mx);
    const float t = (mx > 0.0f) ? t0 : 0.0f;
    const float u = t * t;
    const float p =
        -0.3333333134651184f +
        u * (0.19999539852142334f + u * (-0.14263956248760223f + u * (0.10743731260299683f + u * -0.06451927870512009f)));
    float a = fmaf(t * u, p, t);
    a = big ? a + 0.78539816f : a;
    a = (ay > ax) ? 1.5707963f - a : a;
    a = (x < 0.0f) ? 3.1415927f - a : a;
    return copysignf(a, y);
}

// ln(1 + x) for x >= 0.  FP64: the library routine.  FP32: with u = fl(1 + x),
// ln(u) * x / (u - 1) is accurate to a few ulp because u - 1 is exact and
// ln(1 + y)/y varies slowly (Goldberg, "What every computer scientist should
// know about floating-point arithmetic", 1991, Theorem 4).
RH_HD double rh_log1p(double x) { return log1p(x); }
RH_HD float rh_log1p(float x)
{
    const float u = 1.0f + x;
    const float d = u - 1.0f;
#ifdef __HIP_DEVICE_COMPILE__
    // u >= 1 is never subnormal: the hardware log2 (about 1 ulp) times ln 2
    const float lu = __builtin_amdgcn_logf(u) * 0.69314718f;
#else
    const float lu = logf(u);
#endif
    const float r = lu * (x * rh_rcp(d));
    return (d == 0.0f) ? x : ((u < 3.0e38f) ? r : u);
}

// ---------------------------------------------------------------- FP64 elementary functions
// The FP64 sweeps need these to about 1e-16 absolute on positive, finite,
// normal arguments.  The library routines round correctly over the whole
// domain and cost about 90 (log), 100 (atan2) and 15 (sqrt) vector
// instructions each on gfx950; the forms below cost about a third:
//  * reciprocal: the hardware seed refined by two Newton-Raphson steps;
//  * square root: the hardware reciprocal-square-root seed, one Goldschmidt
//    step and one Newton correction;
//  * ln(num/den): exponents split off, ln(a/b) = 2 atanh(s) for the
//    mantissas a, b in [1/2, 1) with s = (a - b)/(a + b) exact in the
//    numerator, |s| < 1/3, and a degree-9 polynomial in s^2;
//  * atan2: the octant reduced to |t| <= tan(pi/8) (atan r = pi/4 +
//    atan((r - 1)/(r + 1)) above it, one division either way) and a
//    degree-10 polynomial in t^2.
// The polynomials are Chebyshev interpolants of (2 atanh(s)/s - 2)/s^2 and
// (atan(t)/t - 1)/t^2 on those ranges, fitted at 60 digits; evaluated in
// double they are within 6e-17 absolute of ln(a/b) and 3e-17 of atan t.  The host build
// runs the same code with exactly rounded seeds.
constexpr double kLogP[10] = {0.6666666666666666,  0.40000000000014585, 0.2857142856710876,
                              0.22222222717427864, 0.18181789436132859, 0.15385570064958623,
                              0.13314144209654935, 0.12001072809590713, 0.08800738486144674,
                              0.1616404561194353};
constexpr double kAtanP[11] = {-0.3333333333333333,  0.1999999999999552,   -0.14285714284666542,
                               0.11111111015256361,  -0.09090904578123903, 0.07692183190826087,
                               -0.06664511447381948, 0.0585814891280221,   -0.0508544973794026,
                               0.03923165829558719,  -0.01917688711906226};
#ifdef __HIP_DEVICE_COMPILE__
// The same coefficients in constant memory: the kernels then hold them in
// scalar registers and the Horner steps take them as the scalar operand of
// v_fma_f64 (as immediates they are copied into vector registers per step).
static __constant__ double d_log_p[10] = {0.6666666666666666,  0.40000000000014585, 0.2857142856710876,
                                          0.22222222717427864, 0.18181789436132859, 0.15385570064958623,
                                          0.13314144209654935, 0.12001072809590713, 0.08800738486144674,
                                          0.1616404561194353};
static __constant__ double d_atan_p[11] = {
    -0.3333333333333333,  0.1999999999999552,   -0.14285714284666542, 0.11111111015256361,
    -0.09090904578123903, 0.07692183190826087,  -0.06664511447381948, 0.0585814891280221,
    -0.0508544973794026,  0.03923165829558719,  -0.01917688711906226};
// (atan(t) - t)/t^3 on |t| <= tan(pi/16) in u = t^2: Chebyshev interpolant of
// degree 6 fitted at 60 digits (within 7.3e-18 of atan t)
static __constant__ double d_atan16_p[7] = {-0.33333333333333237, 0.19999999999760989, -0.14285714188882722,
                                            0.11111096372965646,  -0.090898379007095286, 0.076521309422432607,
                                            -0.059098572131161474};
#endif
constexpr double kTanPi16 = 0.19891236737965800;
constexpr double kLn2Hi = 0.6931471803691238;  // ln 2 to 32 bits: k * kLn2Hi is exact
constexpr double kLn2Lo = 1.9082149292705877e-10;
constexpr double kTanPi8 = 0.41421356237309503;
constexpr double kPi4 = 0.7853981633974483, kPi2 = 1.5707963267948966, kPi = 3.141592653589793;

RH_HD double rh_rcp(double d)
{
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(d);
#else
    double r = 1.0 / d;
#endif
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

RH_HD double rh_sqrt(double x)
{
#ifdef __HIP_DEVICE_COMPILE__
    const double y = __builtin_amdgcn_rsq(x);
#else
    const double y = 1.0 / sqrt(x);
#endif
    // the seed is within 2^-24 (measured on gfx950): the Goldschmidt step
    // leaves 2^-48, the Newton correction the rounding error
    double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, x);
    return fma(d, h, g);
}

RH_HD void rh_frexp(double x, double& m, int& e)
{
#ifdef __HIP_DEVICE_COMPILE__
    m = __builtin_amdgcn_frexp_mant(x);
    e = __builtin_amdgcn_frexp_exp(x);
#else
    m = frexp(x, &e);
#endif
}

// ln(num/den), num and den positive, finite and normal: with the mantissas
// a, b in [1/2, 1), s = (a - b)/(a + b) (numerator exact), |s| < 1/3.
RH_HD double rh_log_ratio(double num, double den)
{
    double a, b;
    int ea, eb;
    rh_frexp(num, a, ea);
    rh_frexp(den, b, eb);
    const double k = double(ea - eb);
    const double s = (a - b) * rh_rcp(a + b);
    const double v = s * s;
#ifdef __HIP_DEVICE_COMPILE__
    const double* c = d_log_p;
#else
    const double* c = kLogP;
#endif
    double p = c[9];
#pragma unroll
    for (int i = 8; i >= 0; i--) p = fma(p, v, c[i]);
    const double r = fma(s * v, p, 2.0 * s);
    return fma(k, kLn2Hi, fma(k, kLn2Lo, r));
}
RH_HD float rh_log_ratio(float num, float den) { return logf(num / den); }

RH_HD double rh_atan2(double y, double x)
{
    const double ax = fabs(x), ay = fabs(y);
    const double mx = fmax(ax, ay), mn = fmin(ax, ay);
    const bool big = mn > kTanPi8 * mx;
    const double t0 = (big ? mn - mx : mn) * rh_rcp(big ? mn + mx : mx);
    const double t = (mx > 0.0) ? t0 : 0.0;
    const double u = t * t;
#ifdef __HIP_DEVICE_COMPILE__
    const double* c = d_atan_p;
#else
    const double* c = kAtanP;
#endif
    double p = c[10];
#pragma unroll
    for (int i = 9; i >= 0; i--) p = fma(p, u, c[i]);
    double a = fma(t * u, p, t);
    a = big ? a + kPi4 : a;
    a = (ay > ax) ? kPi2 - a : a;
    a = (x < 0.0) ? kPi - a : a;
    return copysign(a, y);
}

// atan2 for sweeps in which most angles are small.  Three tiers, chosen per
// wave by a vote over its active lanes:
// - every lane has x > 0 and |y| < tan(pi/16) x: a degree-6 polynomial in t^2
//   on t = y/x fitted to that range (within 7.4e-18 of atan t in exact
//   arithmetic);
// - every lane has x > 0 and |y| < tan(pi/8) x: the degree-10 polynomial of
//   rh_atan2 on t = y/x, the operations rh_atan2 performs for such a lane
//   (its octant reduction is the identity there), so this tier and rh_atan2
//   give identical results;
// - any other wave: rh_atan2.
// A lane's result therefore depends on which points share its wave: a lane
// below pi/16 gets the short polynomial or the long one as its wave-mates
// decide, and the two differ by rounding (about 1e-17 absolute).  A sweep's
// result for one point can change at that level with the order or batching
// of the points.
RH_HD double rh_atan2_wave(double y, double x)
{
#ifdef __HIP_DEVICE_COMPILE__
    // waves whose angles are all below pi/16 (distant faces) take a shorter
    // polynomial (same accuracy on that range)
    if (__all(fabs(y) < kTanPi16 * x)) {
        const double t = y * rh_rcp(x);
        const double u = t * t;
        const double* c = d_atan16_p;
        double p = c[6];
#pragma unroll
        for (int i = 5; i >= 0; i--) p = fma(p, u, c[i]);
        return fma(t * u, p, t);
    }
    if (__all(fabs(y) < kTanPi8 * x)) {
        const double t = y * rh_rcp(x);
        const double u = t * t;
        const double* c = d_atan_p;
        double p = c[10];
#pragma unroll
        for (int i = 9; i >= 0; i--) p = fma(p, u, c[i]);
        return fma(t * u, p, t);
    }
#endif
    return rh_atan2(y, x);
}
// FP32 counterpart: degree 2 below pi/16 (within 9e-9 of atan t in float
// arithmetic, 1.6e-9 in exact arithmetic) and degree 4 below pi/8 (1.8e-8);
// the tiers differ by float rounding, so a point's result can change by a few
// 1e-8 of an angle with the order or batching of the points.
RH_HD float rh_atan2_wave(float y, float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    // below pi/16: degree 2 in t^2 (within 9e-9 of atan t in float arithmetic)
    if (__all(fabsf(y) < 0.19891237f * x)) {
        const float t = y * rh_rcp(x);
        const float u = t * t;
        const float p = -0.3333331346511841f + u * (0.19990657269954681f + u * -0.13649502396583557f);
        return fmaf(t * u, p, t);
    }
    if (__all(fabsf(y) < 0.41421356f * x)) {
        const float t = y * rh_rcp(x);
        const float u = t * t;
        const float p = -0.3333333134651184f +
                        u * (0.19999539852142334f +
                             u * (-0.14263956248760223f + u * (0.10743731260299683f + u * -0.06451927870512009f)));
        return fmaf(t * u, p, t);
    }
#endif
    return rh_atan2(y, x);
}

// Vertex relative to the observation point, formed in FP64 then rounded to R.
template <class R>
RH_HD V3<R> rel(const double* v, const double p[3])
{
    return V3<R>{R(v[0] - p[0]), R(v[1] - p[1]), R(v[2] - p[2])};
}

// ln of the edge integral of 1/|x-p| as (numerator, denominator) without
// cancellation: whichever end has the larger projection along the edge gives
// R + s directly; the other end uses R + s = d^2/(R - s) when s < 0.
template <class R>
RH_HD void edge_ratio(const V3<R>& ra, R Ra, const V3<R>& rb, R Rb, const V3<R>& t, R& num,
                      R& den)
{
    R sa = dot(ra, t), sb = dot(rb, t);
    bool pos = (sa + sb) >= R(0);
    R Rn = pos ? Rb : Ra, sn = pos ? sb : -sa;
    R Rm = pos ? Ra : Rb, sm = pos ? sa : -sb;
    V3<R> c = cross(ra, t);
    R d2 = dot(c, c);
    bool good = sm >= R(0);
    num = (Rn + sn) * (good ? R(1) : (Rm - sm));
    den = good ? (Rm + sm) : d2;
}

// The same edge integral as ln(1 + delta) with delta >= 0 formed without
// cancellation: num - den = len * (1 + |s_a + s_b| / (R_a + R_b)) in either
// orientation, and 1/den = (R_m - s_m)/d^2 when s_m < 0.  Used by the FP32
// evaluation, where a ratio near 1 would otherwise carry eps32 absolute error.
template <class R>
RH_HD R edge_log1p(const V3<R>& ra, R Ra, const V3<R>& rb, R Rb, const V3<R>& t, R len)
{
    const R sa = dot(ra, t), sb = dot(rb, t);
    const R ssum = sa + sb;
    const bool pos = ssum >= R(0);
    const R Rm = pos ? Ra : Rb, sm = pos ? sa : -sb;
    const R g = len * (R(1) + fabs(ssum) * rh_rcp(Ra + Rb));
    const V3<R> c = cross(ra, t);
    const bool good = sm >= R(0);
    const R num = good ? g : g * (Rm - sm);
    const R den = good ? (Rm + sm) : dot(c, c);
    return rh_log1p(num * rh_rcp(den));
}

// An empty volatile asm statement: no instruction, but the compiler may not
// execute the block it sits in speculatively, so a wave-uniform branch around
// rarely needed work stays a branch instead of becoming selects.
#ifdef __HIP_DEVICE_COMPILE__
#define RH_KEEP_BRANCH() asm volatile("")
#else
#define RH_KEEP_BRANCH() ((void)0)
#endif

// The edge integral of 1/|x-p| from the end distances and the length alone:
// with s = Ra + Rb and l = |rb - ra|, (s + l)(s - l) = 2 (Ra Rb + ra . rb),
// so
//     L = ln((s + l)/(s - l)) = ln(1 + l (l + s) / (Ra Rb + ra . rb)),
// whose terms are all positive when ra . rb >= 0 (the point sees the edge
// under an angle below pi/2).  For ra . rb < 0 the denominator is
// |ra x rb|^2 / (Ra Rb - ra . rb), again without cancellation.  A wave whose
// lanes all have ra . rb >= 0 skips the cross product; each lane's result
// does not depend on the branch its wave takes.
RH_HD double edge_L(const V3<double>& ra, double Ra, const V3<double>& rb, double Rb, double dab, double len)
{
    const double q = len * (len + (Ra + Rb));
    double den = fma(Ra, Rb, dab);
    double num = den + q;
#ifdef __HIP_DEVICE_COMPILE__
    if (!__all(dab >= 0.0))
#endif
    {
        RH_KEEP_BRANCH();
        if (dab < 0.0) {
            const V3<double> c = cross(ra, rb);
            den = dot(c, c);
            num = fma(q, fma(Ra, Rb, -dab), den);
        }
    }
    return rh_log_ratio(num, den);
}

// FP32: the same ratio as ln(1 + delta), delta = l (l + s) / (Ra Rb + ra . rb).
RH_HD float edge_L(const V3<float>& ra, float Ra, const V3<float>& rb, float Rb, float dab, float len)
{
    const float q = len * (len + (Ra + Rb));
    float delta = q * rh_rcp(fmaf(Ra, Rb, dab));
#ifdef __HIP_DEVICE_COMPILE__
    if (!__all(dab >= 0.0f))
#endif
    {
        RH_KEEP_BRANCH();
        if (dab < 0.0f) {
            const V3<float> c = cross(ra, rb);
            delta = (q * fmaf(Ra, Rb, -dab)) * rh_rcp(dot(c, c));
        }
    }
    return rh_log1p(delta);
}

template <class R>
RH_HD R edge_L(const V3<R>& ra, R Ra, const V3<R>& rb, R Rb, R len)
{
    return edge_L(ra, Ra, rb, Rb, dot(ra, rb), len);
}

// Solid-angle factor of triangle (r1, r2, r3) as a complex number whose
// argument is half the signed solid angle (Van Oosterom & Strackee 1983).
template <class R>
RH_HD void tri_factor(const V3<R>& r1, R R1, const V3<R>& r2, R R2, const V3<R>& r3, R R3, R& re,
                      R& im)
{
    im = dot(r1, cross(r2, r3));
    re = R1 * R2 * R3 + dot(r1, r2) * R3 + dot(r1, r3) * R2 + dot(r2, r3) * R1;
}

// q6 layout: xx, yy, zz, xy, xz, yz.
RH_HD void q6_to_q9(const double q6[6], double q9[9])
{
    q9[0] = q6[0]; q9[1] = q6[3]; q9[2] = q6[4];
    q9[3] = q6[3]; q9[4] = q6[1]; q9[5] = q6[5];
    q9[6] = q6[4]; q9[7] = q6[5]; q9[8] = q6[2];
}

// ---------------------------------------------------------------- block
// rec: centre c, half-dims h.  Corner (i,j,k) in {0,1}^3 sits at
// c + ((2i-1)hx, (2j-1)hy, (2k-1)hz).
template <class R>
RH_HD bool eval_rec(const double* __restrict__ rec, const double p[3], double q6[6])
{
    const double cx = rec[0] - p[0], cy = rec[1] - p[1], cz = rec[2] - p[2];
    const R X[2] = {R(cx - rec[3]), R(cx + rec[3])};
    const R Y[2] = {R(cy - rec[4]), R(cy + rec[4])};
    const R Z[2] = {R(cz - rec[5]), R(cz + rec[5])};
    R Rc[2][2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int k = 0; k < 2; k++)
                Rc[i][j][k] = rh_sqrt(X[i] * X[i] + Y[j] * Y[j] + Z[k] * Z[k]);

    // Off-diagonals: Q_yz from the four x-edges, Q_xz from y-edges, Q_xy from z-edges.
    // edge (j,k) along x contributes s_j s_k ln(ratio), s_0 = -1, s_1 = +1.
    double lx, ly, lz;
    {
        R nP = R(1), dP = R(1);
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                R num, den;
                edge_ratio(v3(X[0], Y[j], Z[k]), Rc[0][j][k], v3(X[1], Y[j], Z[k]), Rc[1][j][k],
                           v3(R(1), R(0), R(0)), num, den);
                if (j == k) { nP *= num; dP *= den; } else { nP *= den; dP *= num; }
            }
        lx = double(rh_log_ratio(nP, dP));
    }
    {
        R nP = R(1), dP = R(1);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                R num, den;
                edge_ratio(v3(X[i], Y[0], Z[k]), Rc[i][0][k], v3(X[i], Y[1], Z[k]), Rc[i][1][k],
                           v3(R(0), R(1), R(0)), num, den);
                if (i == k) { nP *= num; dP *= den; } else { nP *= den; dP *= num; }
            }
        ly = double(rh_log_ratio(nP, dP));
    }
    {
        R nP = R(1), dP = R(1);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                R num, den;
                edge_ratio(v3(X[i], Y[j], Z[0]), Rc[i][j][0], v3(X[i], Y[j], Z[1]), Rc[i][j][1],
                           v3(R(0), R(0), R(1)), num, den);
                if (i == j) { nP *= num; dP *= den; } else { nP *= den; dP *= num; }
            }
        lz = double(rh_log_ratio(nP, dP));
    }
    // Diagonals: Q_aa = -(Omega(-a face) + Omega(+a face)); faces as two triangles.
    double om[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double s = 0.;
#pragma unroll
        for (int side = 0; side < 2; side++) {
            // corners of the face counter-clockwise about its outward normal
            int c[4][3];
            if (a == 0) {
                int i = side;
                int cc[4][3] = {{i, 0, 0}, {i, 0, 1}, {i, 1, 1}, {i, 1, 0}};
                int cp[4][3] = {{i, 0, 0}, {i, 1, 0}, {i, 1, 1}, {i, 0, 1}};
                for (int u = 0; u < 4; u++)
                    for (int w = 0; w < 3; w++) c[u][w] = side ? cp[u][w] : cc[u][w];
            } else if (a == 1) {
                int j = side;
                int cc[4][3] = {{0, j, 0}, {1, j, 0}, {1, j, 1}, {0, j, 1}};
                int cp[4][3] = {{0, j, 0}, {0, j, 1}, {1, j, 1}, {1, j, 0}};
                for (int u = 0; u < 4; u++)
                    for (int w = 0; w < 3; w++) c[u][w] = side ? cp[u][w] : cc[u][w];
            } else {
                int k = side;
                int cc[4][3] = {{0, 0, k}, {0, 1, k}, {1, 1, k}, {1, 0, k}};
                int cp[4][3] = {{0, 0, k}, {1, 0, k}, {1, 1, k}, {0, 1, k}};
                for (int u = 0; u < 4; u++)
                    for (int w = 0; w < 3; w++) c[u][w] = side ? cp[u][w] : cc[u][w];
            }
            V3<R> r[4];
            R rr[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                r[u] = v3(X[c[u][0]], Y[c[u][1]], Z[c[u][2]]);
                rr[u] = Rc[c[u][0]][c[u][1]][c[u][2]];
            }
            R re1, im1, re2, im2;
            tri_factor(r[0], rr[0], r[1], rr[1], r[2], rr[2], re1, im1);
            tri_factor(r[0], rr[0], r[2], rr[2], r[3], rr[3], re2, im2);
            R re = re1 * re2 - im1 * im2, im = re1 * im2 + im1 * re2;
            s += 2.0 * double(rh_atan2(im, re));
        }
        om[a] = s;
    }
    q6[0] = -kInv4Pi * om[0];
    q6[1] = -kInv4Pi * om[1];
    q6[2] = -kInv4Pi * om[2];
    q6[3] = kInv4Pi * lz;  // xy from z-edges
    q6[4] = kInv4Pi * ly;  // xz from y-edges
    q6[5] = kInv4Pi * lx;  // yz from x-edges
    return (X[0] * X[1] < R(0)) && (Y[0] * Y[1] < R(0)) && (Z[0] * Z[1] < R(0));
}

// ---------------------------------------------------------------- tetrahedron
template <class R>
RH_HD bool eval_tet(const double* __restrict__ rec, const double p[3], double q6[6])
{
    V3<R> r[4];
    R Rv[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        r[v] = rel<R>(rec + 3 * v, p);
        Rv[v] = rh_sqrt(dot(r[v], r[v]));
    }
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double* ed = rec + 12;
#pragma unroll
    for (int e = 0; e < 6; e++) {
        const double* E = ed + kTetEdge * e;
        const int a = kTetEdgeA[e], b = kTetEdgeB[e];
        V3<R> t = v3(R(E[0]), R(E[1]), R(E[2]));
        double L;
        if (sizeof(R) == 4) {
            L = double(edge_log1p(r[a], Rv[a], r[b], Rv[b], t, R(E[3])));
        } else {
            R num, den;
            edge_ratio(r[a], Rv[a], r[b], Rv[b], t, num, den);
            L = double(rh_log_ratio(num, den));
        }
#pragma unroll
        for (int k = 0; k < 6; k++) acc[k] += L * E[4 + k];
    }
    const double* fn = rec + 12 + 6 * kTetEdge;
    // signed plane distances (r_a . n_f) decide the inside test and the sign of
    // every solid angle, so the two stay consistent
    R hd[4];
    bool inside = true;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double* F = fn + kTetFaceRec * f;
        hd[f] = dot(r[kTetFace[f][0]], v3(R(F[0]), R(F[1]), R(F[2])));
        inside = inside && (hd[f] > R(0));
    }
    // the solid angles of a closed surface sum to 4 pi inside and 0 outside:
    // three atan2 give the fourth
    double om_sum = 0.0;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double* F = fn + kTetFaceRec * f;
        double om;
        if (f < 3) {
            const int a = kTetFace[f][0], b = kTetFace[f][1], c = kTetFace[f][2];
            // r_a.(r_b x r_c) = 2*area*(r_a . n_f) exactly
            const R im = R(F[3]) * hd[f];
            const R re = Rv[a] * Rv[b] * Rv[c] + dot(r[a], r[b]) * Rv[c] + dot(r[a], r[c]) * Rv[b] +
                         dot(r[b], r[c]) * Rv[a];
            om = 2.0 * double(rh_atan2(im, re));
            om_sum += om;
        } else {
            om = (inside ? 4.0 * 3.14159265358979323846 : 0.0) - om_sum;
        }
        const double nx = F[0], ny = F[1], nz = F[2];
        acc[0] -= om * nx * nx;
        acc[1] -= om * ny * ny;
        acc[2] -= om * nz * nz;
        acc[3] -= om * nx * ny;
        acc[4] -= om * nx * nz;
        acc[5] -= om * ny * nz;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) q6[k] = kInv4Pi * acc[k];
    return inside;
}

// ---------------------------------------------------------------- charged tetrahedron
// Field-only record for a given magnetization M (element frame), built per
// field call: verts(12) | edges x6: t(3), len, w = W_e M (3) | faces x4: n(3),
// 2*area, g = (n.M) n (3) | M(3), longest edge.  H = 1/(4pi) [sum L_e w_e - sum Omega_f g_f].
// 86 doubles = 43 x 16 bytes.
constexpr int kChargedM = 12 + 6 * 7 + 4 * 7;  // offset of M
constexpr int kChargedTet = kChargedM + 4;     // 86

RH_HD void charge_tet(const double* __restrict__ rec, const double M[3], double* __restrict__ out)
{
    for (int k = 0; k < 12; k++) out[k] = rec[k];
    for (int e = 0; e < 6; e++) {
        const double* E = rec + 12 + kTetEdge * e;
        double* o = out + 12 + 7 * e;
        const double* W = E + 4;  // xx yy zz xy xz yz
        o[0] = E[0]; o[1] = E[1]; o[2] = E[2]; o[3] = E[3];
        o[4] = W[0] * M[0] + W[3] * M[1] + W[4] * M[2];
        o[5] = W[3] * M[0] + W[1] * M[1] + W[5] * M[2];
        o[6] = W[4] * M[0] + W[5] * M[1] + W[2] * M[2];
    }
    for (int f = 0; f < 4; f++) {
        const double* F = rec + 12 + 6 * kTetEdge + kTetFaceRec * f;
        double* o = out + 12 + 6 * 7 + 7 * f;
        const double sg = F[0] * M[0] + F[1] * M[1] + F[2] * M[2];
        o[0] = F[0]; o[1] = F[1]; o[2] = F[2]; o[3] = F[3];
        o[4] = sg * F[0]; o[5] = sg * F[1]; o[6] = sg * F[2];
    }
    out[kChargedM] = M[0];
    out[kChargedM + 1] = M[1];
    out[kChargedM + 2] = M[2];
    // the longest edge: the FP32 sweep's near test
    double lmax = 0.0;
    for (int e = 0; e < 6; e++) lmax = fmax(lmax, rec[12 + kTetEdge * e + 3]);
    out[kChargedM + 3] = lmax;
}

// Near test of the FP32 sweeps: FP32 offsets carry eps32 of the distance, so
// a term's relative error grows like eps32 * size / distance.  A wave with a
// point within 1.5 longest edges of a vertex evaluates the element in FP64.
constexpr float kNear32 = 1.5f;

// The near test of a charged tetrahedron for the FP32 sweep (the distances as
// field_tet<float> forms them); on the device true for the whole wave when
// any lane is near.
RH_HD bool tet_near32(const double* __restrict__ rec, const double p[3])
{
    float d = 3.0e38f;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const V3<float> r = rel<float>(rec + 3 * v, p);
        d = fminf(d, rh_sqrt(sq3(r.x, r.y, r.z)));
    }
    bool near = d < kNear32 * float(rec[kChargedM + 3]);
#ifdef __HIP_DEVICE_COMPILE__
    near = __any(near);
#endif
    return near;
}

template <class R>
RH_HD bool field_tet(const double* __restrict__ rec, const double p[3], double H[3])
{
    V3<R> r[4];
    R Rv[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        r[v] = rel<R>(rec + 3 * v, p);
        Rv[v] = rh_sqrt(dot(r[v], r[v]));
    }
    double h0 = 0.0, h1 = 0.0, h2 = 0.0;
#pragma unroll
    for (int e = 0; e < 6; e++) {
        const double* E = rec + 12 + 7 * e;
        const int a = kTetEdgeA[e], b = kTetEdgeB[e];
        const double L = double(edge_L(r[a], Rv[a], r[b], Rv[b], R(E[3])));
        h0 += L * E[4];
        h1 += L * E[5];
        h2 += L * E[6];
    }
    const double* fn = rec + 12 + 6 * 7;
    R hd[4];
    bool inside = true;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double* F = fn + 7 * f;
        hd[f] = dot(r[kTetFace[f][0]], v3(R(F[0]), R(F[1]), R(F[2])));
        inside = inside && (hd[f] > R(0));
    }
    double om_sum = 0.0;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double* F = fn + 7 * f;
        double om;
        if (f < 3) {
            const int a = kTetFace[f][0], b = kTetFace[f][1], c = kTetFace[f][2];
            const R im = R(F[3]) * hd[f];
            const R re = Rv[a] * Rv[b] * Rv[c] + dot(r[a], r[b]) * Rv[c] + dot(r[a], r[c]) * Rv[b] +
                         dot(r[b], r[c]) * Rv[a];
            om = 2.0 * double(rh_atan2_wave(im, re));
            om_sum += om;
        } else {
            om = (inside ? 4.0 * 3.14159265358979323846 : 0.0) - om_sum;
        }
        h0 -= om * F[4];
        h1 -= om * F[5];
        h2 -= om * F[6];
    }
    H[0] = kInv4Pi * h0;
    H[1] = kInv4Pi * h1;
    H[2] = kInv4Pi * h2;
    return inside;
}

// ---------------------------------------------------------------- general convex polyhedron
// rec: nv, ne, nf, nfv | verts[3nv] | edges[12 ne] | faces[5 nf] | fv[nfv]
template <class R>
RH_HD bool eval_gen(const double* __restrict__ rec, const double p[3], double q6[6])
{
    const int nv = int(rec[0]), ne = int(rec[1]), nf = int(rec[2]);
    const double* vt = rec + kGenHdr;
    const double* ed = vt + 3 * nv;
    const double* fc = ed + kGenEdge * ne;
    const double* fv = fc + kGenFace * nf;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int e = 0; e < ne; e++) {
        const double* E = ed + kGenEdge * e;
        const int a = int(E[0]), b = int(E[1]);
        V3<R> ra = rel<R>(vt + 3 * a, p), rb = rel<R>(vt + 3 * b, p);
        R Ra = rh_sqrt(dot(ra, ra)), Rb = rh_sqrt(dot(rb, rb));
        const V3<R> t = v3(R(E[2]), R(E[3]), R(E[4]));
        double L;
        if (sizeof(R) == 4) {
            L = double(edge_log1p(ra, Ra, rb, Rb, t, R(E[5])));
        } else {
            R num, den;
            edge_ratio(ra, Ra, rb, Rb, t, num, den);
            L = double(rh_log_ratio(num, den));
        }
        for (int k = 0; k < 6; k++) acc[k] += L * E[6 + k];
    }
    bool inside = true;
    for (int f = 0; f < nf; f++) {
        const double* F = fc + kGenFace * f;
        const int b0 = int(F[3]), cnt = int(F[4]);
        const int i0 = int(fv[b0]);
        V3<R> r0 = rel<R>(vt + 3 * i0, p);
        R R0 = rh_sqrt(dot(r0, r0));
        V3<R> r1 = rel<R>(vt + 3 * int(fv[b0 + 1]), p);
        R R1 = rh_sqrt(dot(r1, r1));
        R zr = R(1), zi = R(0);
        for (int k = 2; k < cnt; k++) {
            V3<R> r2 = rel<R>(vt + 3 * int(fv[b0 + k]), p);
            R R2 = rh_sqrt(dot(r2, r2));
            R re, im;
            tri_factor(r0, R0, r1, R1, r2, R2, re, im);
            R s = R(1) / (R0 * R1 * R2);  // keep the running product O(1)
            re *= s;
            im *= s;
            R nr = zr * re - zi * im, ni = zr * im + zi * re;
            zr = nr;
            zi = ni;
            r1 = r2;
            R1 = R2;
        }
        double om = 2.0 * double(rh_atan2(zi, zr));
        const double nx = F[0], ny = F[1], nz = F[2];
        acc[0] -= om * nx * nx;
        acc[1] -= om * ny * ny;
        acc[2] -= om * nz * nz;
        acc[3] -= om * nx * ny;
        acc[4] -= om * nx * nz;
        acc[5] -= om * ny * nz;
        const double* v0 = vt + 3 * i0;
        double h = (v0[0] - p[0]) * nx + (v0[1] - p[1]) * ny + (v0[2] - p[2]) * nz;
        inside = inside && (h > 0.0);
    }
    for (int k = 0; k < 6; k++) q6[k] = kInv4Pi * acc[k];
    return inside;
}

template <class R>
RH_HD bool eval_elem(int kind, const double* __restrict__ rec, const double p[3], double q6[6])
{
    if (kind == EK_TET) return eval_tet<R>(rec, p, q6);
    if (kind == EK_REC) return eval_rec<R>(rec, p, q6);
    return eval_gen<R>(rec, p, q6);
}

}  // namespace rh
