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
//    edge log is evaluated without cancellation (edge_ln).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lfk {


// ---------------------------------------------------------------------------- records

// Frame-form facet record.  FaceFrame<float> is 128 bytes and is read with wave-uniform
// (scalar) loads by the FP32C forward.
template <class T>
struct __attribute__((aligned(16))) FaceFrame {
    double ax, ay, az;                         // vertex a, FP64
    T ux, uy, uz, vx, vy, vz, nx, ny, nz;      // facet frame
    T l0, cu, cv;                              // |b-a|; (c-a) in the frame
    T e1u, e1v, l1, e2u, e2v, l2;              // unit edges b->c, c->a in the frame; lengths
    T K1, K2;                                  // gz  = K1*sum_e dperp_e*lnQ_e + K2*pz*atan2(N,D)
    T W0, W1, W2, W3;                          // tmi = W0*atan2(N,D) + sum_e W_e*lnQ_e
    T twoA;                                    // 2*area = l0*cv
    T zfb;  // max edge length / 60: no edge can need the FP64 branch when |pz| >= zfb
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

// ln((s2+d2)/(s1+d1)) for an edge of length ell (s2 = s1 + ell), X = |s| + d, without
// cancellation even when the edge is far away (the ratio is then 1 + O(ell/d)).  With
// dd = d2 - d1 = ell (2 s1 + ell) / (d1 + d2) (exact, since d2^2 - d1^2 = 2 s1 ell + ell^2):
//   s1 >= 0      : Q - 1 = (ell + dd)/X1 = ell (X1 + X2) / (X1 (d1 + d2))
//   s2 <  0      : Q = X1/X2 (identity (s+d)(d-s) = rho^2 at both ends),
//                  Q - 1 = (ell - dd)/X2 = ell (X1 + X2) / (X2 (d1 + d2))
//   s1 < 0 <= s2 : Q = X1 X2 / rho^2 (station projects inside the edge)
// so every case is ln(1 + y) with y >= 0 formed from positive products and one reciprocal.
// Also returns X1, X2 and the case flags for the derivative.
template <class T>
__device__ __forceinline__ T edge_ln(T s1, T ell, T d1, T d2, T rho2, T& X1, T& X2, bool& pos1,
                                     bool& neg2)
{
    const T s2 = s1 + ell;
    X1 = fabs(s1) + d1;
    X2 = fabs(s2) + d2;
    pos1 = s1 >= T(0);
    neg2 = s2 < T(0);
    const bool mid = !(pos1 || neg2);
    const T num = mid ? X1 * X2 - rho2 : ell * (X1 + X2);
    // outside the edge the nearer end has the smaller X (X1 when s1 >= 0, X2 when s2 < 0)
    const T den = mid ? rho2 : fmin(X1, X2) * (d1 + d2);
    return log1p_t(num * fast_rcp(den));
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
    S K1, K2, W0, W1, W2, W3, twoA;
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
    o.K2 = S(2.0) * o.kz;
    o.W0 = S(2.0) * o.wom;
    o.W1 = o.we[0];
    o.W2 = o.we[1];
    o.W3 = o.we[2];
    o.twoA = o.l0 * o.cv;
}

// Writes the frame-form record; a facet with zero area or a zero-length edge gets a
// harmless unit right triangle at a with all coefficients zero.
template <class T>
__device__ __forceinline__ void store_frame(const FrameVals<double>& fv, const double a[3],
                                            FaceFrame<T>& r)
{
    r.ax = a[0];
    r.ay = a[1];
    r.az = a[2];
    if (fv.bad) {
        r.ux = T(1); r.uy = T(0); r.uz = T(0);
        r.vx = T(0); r.vy = T(1); r.vz = T(0);
        r.nx = T(0); r.ny = T(0); r.nz = T(1);
        r.l0 = T(1); r.cu = T(0); r.cv = T(1);
        r.e1u = T(-0.70710678118654752); r.e1v = T(0.70710678118654752); r.l1 = T(1.41421356237309505);
        r.e2u = T(0); r.e2v = T(-1); r.l2 = T(1);
        r.K1 = r.K2 = r.W0 = r.W1 = r.W2 = r.W3 = T(0);
        r.twoA = T(1);
        r.zfb = T(0);
    } else {
        r.ux = T(fv.u[0]); r.uy = T(fv.u[1]); r.uz = T(fv.u[2]);
        r.vx = T(fv.v[0]); r.vy = T(fv.v[1]); r.vz = T(fv.v[2]);
        r.nx = T(fv.n[0]); r.ny = T(fv.n[1]); r.nz = T(fv.n[2]);
        r.l0 = T(fv.l0); r.cu = T(fv.cu); r.cv = T(fv.cv);
        r.e1u = T(fv.e1u); r.e1v = T(fv.e1v); r.l1 = T(fv.l1);
        r.e2u = T(fv.e2u); r.e2v = T(fv.e2v); r.l2 = T(fv.l2);
        r.K1 = T(fv.K1); r.K2 = T(fv.K2);
        r.W0 = T(fv.W0); r.W1 = T(fv.W1); r.W2 = T(fv.W2); r.W3 = T(fv.W3);
        r.twoA = T(fv.twoA);
        // A straddling edge needs the FP64 branch only if rho^2 * 1024 < d1 d2; with
        // d1 d2 <= ((len + 2 rho)/2)^2 that requires rho < len/62, and rho >= |pz|.
        const double lmax = fmax(fv.len[0], fmax(fv.len[1], fv.len[2]));
        r.zfb = T(lmax / 60.0);
    }
}

}  // namespace lfk
