// pair_math.h -- per (station, facet) arithmetic of the polyhedral gz / TMI forward.
//
// One facet contributes, at station o (reference: faceNagyContribution,
// modules/litho-forward/src/forward/gravity_forward.cpp:13-42, and magneticFacet,
// modules/litho-forward/src/forward/magnetic_forward.cpp:312-347):
//   gz_f  = kz * ( sum_e dperp_e * L_e  -  h * Omega )          kz = G*drho*1e8 * n_z
//   tmi_f = wOm * Omega + sum_e w_e * L_e
// with Omega the solid angle (geometry.cpp:10-30), L_e = ln((s2+d2)/(s1+d1)) the edge
// line integral (geometry.cpp:32-52), h = (a-o).n, dperp_e = (o-p1).(n x e_hat).
// The face constants (kz, wOm, w_e, unit normal, edge frames) do not depend on the
// station and are built once per facet (facet_frame below).
//
// Two formulations:
//  * FP64 ("plain"): the reference's own operation sequence and guards.
//  * Frame form (FP32C forward, both adjoints): the station offset P = o - a is formed
//    in FP64 and rounded once, then expressed in the facet's orthonormal frame
//    (u along edge a->b, v = n x u, n): p = (px, py, pz).  The three vertex distances,
//    the solid-angle numerator (2A*h, h = -pz) and denominator, and every edge's
//    (s1, s2, dperp, rho^2) follow from p and nine in-plane shape numbers, and the
//    edge log is evaluated without cancellation (edge_ratio, atanh2_series,
//    ln_ratio_near).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lfk {


// ---------------------------------------------------------------------------- records

// Frame-form facet record.  FaceFrame<float> is 128 bytes; the FP32C forward streams it
// through LDS.  The station offset in the frame is p = R (o - a) = R (o - a') + d: the
// record's origin a' is vertex a moved by a small per-facet offset (FP32C records; a' = a
// for FP64 ones) and d = R (a' - a) puts it back, so p is the offset from a itself.
template <class T>
struct __attribute__((aligned(16))) FaceFrame {
    double ax, ay, az;                         // origin a', FP64
    T ux, uy, uz, vx, vy, vz, nx, ny, nz;      // facet frame
    T l0, cu, cv;                              // |b-a|; (c-a) in the frame (2A = l0 cv)
    T e1u, e1v, l1, e2u, e2v, l2;              // unit edges b->c, c->a in the frame; lengths
    T K1;                                      // gz  = K1*(sum_e dperp_e*lnQ_e + 2 pz atan2(N,D))
    T W0, W1, W2, W3;                          // tmi = W0*atan2(N,D) + sum_e W_e*lnQ_e
    T du, dv, dn;                              // R (a' - a)
};
using FaceF32 = FaceFrame<float>;
static_assert(sizeof(FaceF32) == 128, "FaceF32 must be 128 bytes");

// FP64 facet record in the reference's terms.
struct FaceF64 {
    double a[3], b[3], c[3];
    double n[3];
    double eh[3][3];
    double perp[3][3];
    double kz, wom, we[3];
    double pad;
};
static_assert(sizeof(FaceF64) == 288, "FaceF64 layout");

// ---------------------------------------------------------------------- primitives

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ double fast_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double fast_log2(double x) { return log2(x); }
__device__ __forceinline__ double fast_rsqrt(double x) { return 1.0 / sqrt(x); }

// The hardware FP32 reciprocal, square root and log2 are faithful (within 1 ulp) but not
// unbiased: measured over every input of a binade on gfx950 their mean errors are about
// -0.07 ulp (rcp), -0.15 ulp (sqrt), -0.09 ulp (rsq) and -0.2 to -0.4 ulp (log2).  A sum of
// millions of same-signed terms (the reference TMI formula's edge terms all share one sign)
// turns such a bias into an error of the total at that relative size, so the FP32C forward
// uses the forms below: one Newton step after the hardware estimate leaves only the
// rounding of the final fma, which is unbiased.
__device__ __forceinline__ float sqrt_nr(float x)
{
    // x >= 0; x = 0 gives 0 (the clamp keeps rsq finite).  The clamp is not folded into
    // the squared distance as an fma floor (pz * pz + 1e-30): a positive addend breaks
    // every exact rounding tie of pz * pz upwards, and ties are common where station and
    // vertex coordinates sit on grids (offsets in quarter metres: pz with 13 significant
    // bits has a square exactly halfway between two floats), which biased the sums.
    const float rs = fast_rsqrt(fmaxf(x, 1e-30f));
    const float d = x * rs;
    return fmaf(fmaf(-d, d, x), 0.5f * rs, d);
}
__device__ __forceinline__ float rcp_nr(float x)
{
    const float r = fast_rcp(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}

// atan2 in FP32: argument reduction to a = min(|x|,|y|)/max(|x|,|y|) in [0, 1], odd
// minimax polynomial atan(a) ~= a*P(a^2) of degree 7 in a^2 (max abs error 3.7e-8 rad;
// coefficients from csrc/gen_atan_poly.py, a discrete Remez/Lawson fit), then the
// octant reflections pi/2 - r, pi - r and the sign of y.
__device__ __forceinline__ float atan2_t(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mn = fminf(ax, ay), mx = fmaxf(ax, ay);
    const float a = mn * fast_rcp(mx);
    const float s = a * a;
    float p = -4.054563001e-03f;
    p = fmaf(p, s, 2.186295371e-02f);
    p = fmaf(p, s, -5.591234242e-02f);
    p = fmaf(p, s, 9.642200692e-02f);
    p = fmaf(p, s, -1.390863213e-01f);
    p = fmaf(p, s, 1.994656656e-01f);
    p = fmaf(p, s, -3.332986092e-01f);
    p = fmaf(p, s, 9.999993356e-01f);
    float r = p * a;
    r = (ay > ax) ? (1.57079632679489662f - r) : r;
    r = (x < 0.0f) ? (3.14159265358979324f - r) : r;
    return copysignf(r, y);
}
__device__ __forceinline__ double atan2_t(double y, double x) { return atan2(y, x); }

// atan(a) for |a| <= kSmallAngle: the Taylor series a - a^3/3 + ... + a^9/9 (the first
// omitted term, a^11/11, is below 1e-10 of a at |a| = 1/8), written as a + (a s) q so that
// the result carries one final rounding.  The solid angle of a facet seen from more than
// a few edge lengths away has tan(Omega/2) = N/D far inside this range, where atan2_t's
// octant reduction and its minimax polynomial (whose constant term 1 - 6.6e-7 biases
// small angles) are not needed.
constexpr float kSmallAngle = 0.125f;
__device__ __forceinline__ float atan_small(float a)
{
    const float s = a * a;
    float q = fmaf(s, 1.0f / 9.0f, -1.0f / 7.0f);
    q = fmaf(s, q, 1.0f / 5.0f);
    q = fmaf(s, q, -1.0f / 3.0f);
    return fmaf(a * s, q, a);
}
// d atan(a) / da = 1 / (1 + s), s = a^2 <= 1/64, by its series to s^4 (truncation below
// 1e-9 relative).
__device__ __forceinline__ float datan_small(float s)
{
    float g = s - 1.0f;
    g = fmaf(s, g, 1.0f);
    g = fmaf(s, g, -1.0f);
    g = fmaf(s, g, 1.0f);
    return g;
}

// Kahan-compensated accumulation: s holds the sum, c the running compensation;
// the compensated total is s - c.
__device__ __forceinline__ void kahan_add(float& s, float& c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// --------------------------------------------------------------- stable edge log

// ln(1 + y) for y >= 0.  u = fl(1 + y); u - 1 is exact for 1 <= u < 2^24 and so is
// c = y - (u - 1), the rounding error of u; ln(1 + y) = ln u + ln(1 + c/u) ~ ln u + c/u.
// c/u is replaced by c * sat(2 - u) (sat clamps to [0, 1], the instruction's clamp bit):
// for 1 <= u < 2, |1/u - (2 - u)| = (u - 1)^2 / u <= 1/2, so the correction is off by at
// most |c|/2 (relative 3e-8); for u >= 2 the term is below 1e-7 of ln u and dropped.
// ln u from the hardware base-2 log (1 ulp).
__device__ __forceinline__ float log1p_t(float y)
{
    const float u = 1.0f + y;
    const float c = y - (u - 1.0f);
    return fmaf(0.693147180559945f, fast_log2(u), c * __builtin_amdgcn_fmed3f(2.0f - u, 0.0f, 1.0f));
}
__device__ __forceinline__ double log1p_t(double y) { return log1p(y); }

// The edge log Q = (s2+d2)/(s1+d1) for an edge of length ell (s2 = s1 + ell), X = |s| + d,
// is 1 + O(ell/d) for a distant edge, so ln of a rounded ratio would lose eps*d/ell.  With
// dd = d2 - d1 = ell (2 s1 + ell) / (d1 + d2) (exact, since d2^2 - d1^2 = 2 s1 ell + ell^2):
//   s1 >= 0      : Q - 1 = (ell + dd)/X1 = ell (X1 + X2) / (X1 (d1 + d2))
//   s2 <  0      : Q = X1/X2 (identity (s+d)(d-s) = rho^2 at both ends),
//                  Q - 1 = (ell - dd)/X2 = ell (X1 + X2) / (X2 (d1 + d2))
//   s1 < 0 <= s2 : Q = X1 X2 / rho^2 (station projects inside the edge)
// so every case is ln(1 + y) with y >= 0 formed from positive products.
//
// Numerator and denominator of Q - 1 = num / den, with the middle case in a form without
// cancellation: from (|s| + d)(d - |s|) = rho^2 and d2 - d1 = ell (s1 + s2) / (d1 + d2),
//   X1 X2 - rho^2 = ell (X1 X2 + rho^2) / (d1 + d2),
// so every case is num = ell * A, den = (d1 + d2) * B with A, B sums or products of
// positive numbers:
//   outside the edge : A = X1 + X2,           B = min(X1, X2)
//   inside           : A = X1 X2 + rho^2,     B = rho^2
__device__ __forceinline__ void edge_ratio(float s1, float ell, float d1, float d2, float rho2,
                                           float& num, float& den)
{
    const float s2 = s1 + ell;
    const float X1 = fabsf(s1) + d1;
    const float X2 = fabsf(s2) + d2;
    const bool mid = (s1 < 0.f) && (s2 >= 0.f);
    const float A = mid ? fmaf(X1, X2, rho2) : X1 + X2;
    const float B = mid ? rho2 : fminf(X1, X2);
    num = ell * A;
    den = (d1 + d2) * B;
}

// Edge log from z = ell / (d1 + d2): in all three cases above Q = (d1 + d2 + ell) /
// (d1 + d2 - ell), i.e. ln Q = 2 atanh(z) (from X1 X2 - rho^2 = ell (X1 X2 + rho^2) /
// (d1 + d2) and the end-case identities).  For z <= kZSeries (the station more than about
// five edge lengths from the edge) the series 2 z (1 + z^2/3 + ... + z^8/9) (truncation
// below 1e-11 relative at z = 0.1) uses only fmas, so no biased hardware transcendental
// reaches the result; nearer edges take ln_ratio_near, whose middle case has no 1 - z
// cancellation for a station close to the edge's line.
constexpr float kZSeries = 0.1f;
__device__ __forceinline__ float atanh2_series(float z)
{
    // z (2 + 2z^2/3 + ... + 2z^8/9): the factor 2 sits in the coefficients
    const float s = z * z;
    float p = fmaf(s, 2.0f / 9.0f, 2.0f / 7.0f);
    p = fmaf(s, p, 2.0f / 5.0f);
    p = fmaf(s, p, 2.0f / 3.0f);
    p = fmaf(s, p, 2.0f);
    return z * p;
}
__device__ __forceinline__ float ln_ratio_near(float num, float den)
{
    return log1p_t(num * rcp_nr(den));
}

// The same logarithm in FP64 from original coordinates (used by the FP32C forward for
// edges whose line passes close to the station).
__device__ __forceinline__ float edge_ln_f64(double ox, double oy, double oz, double p1x,
                                             double p1y, double p1z, double p2x, double p2y,
                                             double p2z)
{
    const double ex = p2x - p1x, ey = p2y - p1y, ez = p2z - p1z;
    const double il = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
    const double hx = ex * il, hy = ey * il, hz = ez * il;
    const double r1x = p1x - ox, r1y = p1y - oy, r1z = p1z - oz;
    const double r2x = p2x - ox, r2y = p2y - oy, r2z = p2z - oz;
    const double s1 = r1x * hx + r1y * hy + r1z * hz;
    const double s2 = r2x * hx + r2y * hy + r2z * hz;
    const double d1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z);
    const double d2 = sqrt(r2x * r2x + r2y * r2y + r2z * r2z);
    const double cx = r1y * hz - r1z * hy, cy = r1z * hx - r1x * hz, cz = r1x * hy - r1y * hx;
    const double rho2 = cx * cx + cy * cy + cz * cz;
    const double X1 = fabs(s1) + d1, X2 = fabs(s2) + d2;
    double num, den;
    if (s1 >= 0.0) { num = X2; den = X1; }
    else if (s2 < 0.0) { num = X1; den = X2; }
    else { num = X1 * X2; den = rho2; }
    // a station on the segment itself (rho = 0): the integral diverges; the reference's
    // edge term is 0 there (geometry.cpp:48-50)
    if (!(den > 0.0)) return 0.0f;
    return (float)log(num / den);
}

// ---------------------------------------------------------- forward-mode dual number

// value + one directional derivative; used to push adjoint seeds of the facet constants
// through their definition (once per facet and direction).
struct Dual {
    double v, d;
    __device__ Dual() : v(0), d(0) {}
    __device__ Dual(double x) : v(x), d(0) {}
    __device__ Dual(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual dsqrt(Dual a)
{
    const double s = sqrt(a.v);
    return Dual(s, s > 0.0 ? 0.5 * a.d / s : 0.0);
}
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ double val(double a) { return a; }
__device__ __forceinline__ double val(Dual a) { return a.v; }

// ------------------------------------------------------- facet constants (one source)

template <class S>
struct FrameVals {
    S u[3], v[3], n[3];
    S l0, cu, cv, e1u, e1v, l1, e2u, e2v, l2;
    S K1, W0, W1, W2, W3, twoA;
    S kz, wom, we[3];
    S eh[3][3], perp[3][3], len[3];
    bool degen, edge_ok[3], bad;
};

// Facet constants from the vertex positions P[3][3] (a, b, c), the density contrast
// drho = rho_front - rho_back, magnetization difference Md = M_front - M_back and sum
// Ms = M_front + M_back, the field direction F and the TMI edge-coefficient mode.
// Guards: zero-area facet (gravity_forward.cpp:19) and zero-length edge (:33) -> zero
// contribution.  Scale constants: gravity_forward.h:37-41; magnetic -100/(4 pi):
// magnetic_forward.cpp:384-385.
template <class S>
__device__ __forceinline__ void facet_frame(const S P[3][3], double drho, const double Md[3],
                                            const double Ms[3], const double F[3],
                                            bool textbook, FrameVals<S>& o)
{
    const double kEps = 1e-15;
    const double kGrav = 6.67430e-11 * 1000.0 * 1e5;
    const double C = -100.0 / (4.0 * 3.14159265358979323846);
    const S e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
    const S e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
    S n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
              e1[0] * e2[1] - e1[1] * e2[0]};
    const S area2 = dsqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    o.degen = val(area2) < kEps;
    for (int j = 0; j < 3; ++j) o.n[j] = o.degen ? S(0.0) : n[j] / area2;
    o.bad = o.degen;
    for (int k = 0; k < 3; ++k) {
        const S* p1 = P[k];
        const S* p2 = P[(k + 1) % 3];
        const S e[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
        o.len[k] = dsqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        o.edge_ok[k] = val(o.len[k]) >= kEps;
        o.bad = o.bad || !o.edge_ok[k];
        for (int j = 0; j < 3; ++j) o.eh[k][j] = o.edge_ok[k] ? e[j] / o.len[k] : S(0.0);
        const S* nn = o.n;
        o.perp[k][0] = nn[1] * o.eh[k][2] - nn[2] * o.eh[k][1];
        o.perp[k][1] = nn[2] * o.eh[k][0] - nn[0] * o.eh[k][2];
        o.perp[k][2] = nn[0] * o.eh[k][1] - nn[1] * o.eh[k][0];
    }
    const S fn = o.n[0] * F[0] + o.n[1] * F[1] + o.n[2] * F[2];
    const S mdn = o.n[0] * Md[0] + o.n[1] * Md[1] + o.n[2] * Md[2];
    o.kz = o.degen ? S(0.0) : S(kGrav * drho) * o.n[2];
    o.wom = o.degen ? S(0.0) : S(C) * mdn * fn;
    for (int k = 0; k < 3; ++k) {
        const S* t = o.perp[k];
        const S ft = t[0] * F[0] + t[1] * F[1] + t[2] * F[2];
        const S mt = textbook ? mdn : t[0] * Ms[0] + t[1] * Ms[1] + t[2] * Ms[2];
        o.we[k] = (o.degen || !o.edge_ok[k]) ? S(0.0) : S(C) * mt * ft;
    }
    for (int j = 0; j < 3; ++j) {
        o.u[j] = o.eh[0][j];
        o.v[j] = o.perp[0][j];
    }
    o.l0 = o.len[0];
    o.cu = e2[0] * o.u[0] + e2[1] * o.u[1] + e2[2] * o.u[2];
    o.cv = e2[0] * o.v[0] + e2[1] * o.v[1] + e2[2] * o.v[2];
    o.e1u = o.eh[1][0] * o.u[0] + o.eh[1][1] * o.u[1] + o.eh[1][2] * o.u[2];
    o.e1v = o.eh[1][0] * o.v[0] + o.eh[1][1] * o.v[1] + o.eh[1][2] * o.v[2];
    o.l1 = o.len[1];
    o.e2u = o.eh[2][0] * o.u[0] + o.eh[2][1] * o.u[1] + o.eh[2][2] * o.u[2];
    o.e2v = o.eh[2][0] * o.v[0] + o.eh[2][1] * o.v[1] + o.eh[2][2] * o.v[2];
    o.l2 = o.len[2];
    o.K1 = o.kz;
    o.W0 = S(2.0) * o.wom;
    o.W1 = o.we[0];
    o.W2 = o.we[1];
    o.W3 = o.we[2];
    o.twoA = o.l0 * o.cv;
}

// ------------------------------------------------- corner order of the frame records

// The facet integral does not depend on which corner is called a (a cyclic relabelling
// keeps the winding), but the FP32 frame does: with u along an axis-parallel edge, the
// frame of a voxel-face triangle (two axis-parallel legs and a diagonal) is exactly
// (0, +-1) in FP32, while u along the diagonal rounds every projection by the same
// relative amount -- an error common to all facets of that shape, which a sum over
// millions of them does not average out.  frame_rotation returns the cyclic shift k
// (corner c of the record = corner (c + k) % 3 of the facet) whose frame (u, v; n is
// shared) has the smallest FP32 representation error; ties keep k = 0.  A function of
// the FP64 vertices only, so the facet setup and the adjoint's derivative matrix pick
// the same k.
__device__ __forceinline__ double f32_rep_err(double x) { return fabs((double)(float)x - x); }

__device__ __forceinline__ int frame_rotation(const double P[3][3])
{
    const double e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
    const double e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    const double an = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (!(an >= 1e-15)) return 0;
    for (int j = 0; j < 3; ++j) n[j] /= an;
    int best = 0;
    double best_err = 1e300;
    for (int k = 0; k < 3; ++k) {
        const double* p1 = P[k];
        const double* p2 = P[(k + 1) % 3];
        const double e[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
        const double l = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        if (!(l >= 1e-15)) continue;
        const double u[3] = {e[0] / l, e[1] / l, e[2] / l};
        const double v[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2],
                             n[0] * u[1] - n[1] * u[0]};
        double err = 0.0;
        for (int j = 0; j < 3; ++j) err += f32_rep_err(u[j]) + f32_rep_err(v[j]);
        if (err < best_err) {
            best_err = err;
            best = k;
        }
    }
    return best;
}

__device__ __forceinline__ void rotate_corners(const double P[3][3], int k, double Q[3][3])
{
    for (int c = 0; c < 3; ++c)
        for (int j = 0; j < 3; ++j) Q[c][j] = P[(c + k) % 3][j];
}

// ------------------------------------------------------- rounding of record constants

// FP32 rounding of a facet constant that is unbiased across facets: round to one of the
// two neighbouring floats with the probability that makes the expected value exact, from
// a fixed hash of (facet, field).  Round-to-nearest gives every facet of one shape the
// same representation error (voxel models have a handful of distinct edge lengths,
// directions and coefficients), which a sum over millions of facets accumulates; with
// this rounding the errors of different facets are independent with zero mean, and the
// records are still a deterministic function of the vertices and properties.
__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float round_unbiased(double x, uint32_t key)
{
    const float r = (float)x;
    const double err = x - (double)r;
    if (err == 0.0 || !(fabs(x) < 3.0e38)) return r;
    const float nb = nextafterf(r, err > 0.0 ? 3.4e38f : -3.4e38f);
    const double gap = fabs((double)nb - (double)r);
    const double uni = (double)(mix32(key) >> 8) * (1.0 / 16777216.0);  // [0, 1)
    return uni * gap < fabs(err) ? nb : r;
}

template <class T>
struct RecRound {
    uint32_t base;
    __device__ T operator()(double x, int field) const { return T(x); }
};
template <>
struct RecRound<float> {
    uint32_t base;
    __device__ float operator()(double x, int field) const
    {
        return round_unbiased(x, base * 32u + (uint32_t)field);
    }
};

// Per-facet origin offset of the FP32C records: each component uniform in
// [-2^-9, 2^-9) m from a fixed hash of the facet.  The FP32 rounding of o - a has the
// same error for every facet whose vertex a shares a coordinate with another (a voxel
// layer shares its z, a column its x), and a sum over millions of facets adds those
// errors as if the station sat slightly off in each layer; from an offset origin the
// roundings of different facets are independent.  2^-9 m is two units in the last
// place of FP32 at 8-16 km, the largest station-facet distances of a survey; d = R (a'-a)
// restores the exact offset, so the offset changes no result beyond rounding.
constexpr double kOriginJitter = 1.0 / 512.0;

// Writes the frame-form record; a facet with zero area or a zero-length edge gets a
// harmless unit right triangle at a with all coefficients zero.  fid seeds the unbiased
// rounding and the origin offset of FP32 records (jitter = true).
template <class T>
__device__ __forceinline__ void store_frame(const FrameVals<double>& fv, const double a[3],
                                            FaceFrame<T>& r, uint32_t fid = 0,
                                            bool jitter = false)
{
    double dl[3] = {0.0, 0.0, 0.0};
    if (jitter && !fv.bad)
        for (int j = 0; j < 3; ++j)
            dl[j] = ((double)(mix32(fid * 32u + 29u + (uint32_t)j) >> 8) * (2.0 / 16777216.0) -
                     1.0) * kOriginJitter;
    r.ax = a[0] + dl[0];
    r.ay = a[1] + dl[1];
    r.az = a[2] + dl[2];
    if (fv.bad) {
        r.ux = T(1); r.uy = T(0); r.uz = T(0);
        r.vx = T(0); r.vy = T(1); r.vz = T(0);
        r.nx = T(0); r.ny = T(0); r.nz = T(1);
        r.l0 = T(1); r.cu = T(0); r.cv = T(1);
        r.e1u = T(-0.70710678118654752); r.e1v = T(0.70710678118654752); r.l1 = T(1.41421356237309505);
        r.e2u = T(0); r.e2v = T(-1); r.l2 = T(1);
        r.K1 = r.W0 = r.W1 = r.W2 = r.W3 = T(0);
        r.du = r.dv = r.dn = T(0);
    } else {
        const RecRound<T> q{fid};
        r.ux = q(fv.u[0], 0); r.uy = q(fv.u[1], 1); r.uz = q(fv.u[2], 2);
        r.vx = q(fv.v[0], 3); r.vy = q(fv.v[1], 4); r.vz = q(fv.v[2], 5);
        r.nx = q(fv.n[0], 6); r.ny = q(fv.n[1], 7); r.nz = q(fv.n[2], 8);
        r.l0 = q(fv.l0, 9); r.cu = q(fv.cu, 10); r.cv = q(fv.cv, 11);
        r.e1u = q(fv.e1u, 12); r.e1v = q(fv.e1v, 13); r.l1 = q(fv.l1, 14);
        r.e2u = q(fv.e2u, 15); r.e2v = q(fv.e2v, 16); r.l2 = q(fv.l2, 17);
        r.K1 = q(fv.K1, 18);
        r.W0 = q(fv.W0, 20); r.W1 = q(fv.W1, 21); r.W2 = q(fv.W2, 22); r.W3 = q(fv.W3, 23);
        r.du = T(fv.u[0] * dl[0] + fv.u[1] * dl[1] + fv.u[2] * dl[2]);
        r.dv = T(fv.v[0] * dl[0] + fv.v[1] * dl[1] + fv.v[2] * dl[2]);
        r.dn = T(fv.n[0] * dl[0] + fv.n[1] * dl[1] + fv.n[2] * dl[2]);
    }
}

}  // namespace lfk
