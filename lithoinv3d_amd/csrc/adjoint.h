// adjoint.h -- matrix-free adjoint of the polyhedral gz / TMI forward.
//
// Replaces the reference's dense per-face one-sided-FD Jacobian followed by -J^T W^2 r
// (gravity_forward.cpp:184-292, magnetic_forward.cpp:196-306, objective.cpp:42-55) with
// the exact derivative of the frame-form facet integral, in one pass over all
// (station, facet) pairs:
//
//   grad_v = sum_i [ w_gz_i * d gz_i / d x_v  +  w_tmi_i * d TMI_i / d x_v ]
//
// Structure.  A facet's contribution at station o is F(p; theta): p = R (o - a) is the
// station offset in the facet frame R = (u, v, n) and theta are the facet constants
// (seven in-plane shape numbers: l0, cu, cv and the length and angle of edges 1 and 2;
// 2A; five coefficients K1, W0..W3).  Per pair the kernel runs the frame-form forward and
// its reverse, and accumulates over the stations of a chunk:
//   Gp    = sum w dF/dp                                  (3)
//   A_jk  = sum w (dF/dp_j p_k - dF/dp_k p_j), j<k     (3)  -- the frame's rotation
//   tbar  = sum w dF/dtheta                              (13)
// 19 numbers per facet, lane = facet, stations wave-uniform (scalar loads).  A finalize
// kernel then pushes them through R(a,b,c) and theta(a,b,c) once per facet with
// forward-mode duals (9 directions):
//   dPhi/dx_k = sum_{j<k'} (dR_j/dx_k . R_k') A_jk' + tbar . dtheta/dx_k - [x_k in a] (R^T Gp)_k
// and a gather kernel sums facet corners into vertices through a vertex -> corner CSR
// list built once from the topology (deterministic, no atomics).
//
// Edge logs: 2 atanh(ell / (d1 + d2)) by its series for edges more than about five edge
// lengths from the station, the cancellation-free three-case form of the forward for the
// nearer ones; the FP32 adjoint has no FP64 edge branch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "pair_math.h"

namespace lfk {

// S_PHI1 / S_PHI2: the in-plane directions (e1u, e1v), (e2u, e2v) of edges 1 and 2 are
// unit vectors, so only their angle has a seed: (se_u, se_v) . (-e_v, e_u).  S_K1 is the
// seed of the gravity coefficient K1 (gz = K1 (sum_e dperp_e L_e + 2 pz th)).
constexpr int kSeeds = 19;
enum SeedIdx {
    S_GPX = 0, S_GPY, S_GPZ, S_AXY, S_AXZ, S_AYZ,
    S_L0, S_CU, S_CV, S_PHI1, S_L1, S_PHI2, S_L2, S_TWOA,
    S_K1, S_W0, S_W1, S_W2, S_W3
};

struct StationW {
    double x, y, z, wg, wt;
    int64_t skip;  // 1 when both weights are zero (tested with scalar integer ops)
    double pad0, pad1;
};

struct AdjointWork {
    int64_t n_v = 0, n_f = 0, n_s = 0;
    int precision = 0;
    int n_chunks = 1;
    int64_t st_per_chunk = 1;
    void* d_seeds = nullptr;
    StationW* d_st = nullptr;
    double* d_face_grad = nullptr;
    int64_t* d_csr_off = nullptr;
    int32_t* d_csr_ent = nullptr;
    double* d_wg = nullptr;
    double* d_wt = nullptr;
    double* d_grad = nullptr;
    double* d_zero = nullptr;  // n_s zeros for a missing weight vector
};

struct AdjointLaunch {
    hipStream_t stream;
    int64_t n_v, n_f, n_s;
    int precision;
    const FaceFrame<float>* f32;
    const FaceFrame<double>* f64;
    const double *sx, *sy, *sz;
    const double *wg, *wt;
    double* grad;
};

// ln(Q) of an edge and its derivatives with respect to (s1, s2, d1, d2, rho^2) in the
// form of each case: Q = X2/X1, X1/X2 or X1 X2/rho^2 with X = |s| + d.  The forms agree
// where d^2 = s^2 + rho^2, which the frame parametrization keeps, so the chain rule
// through it is exact whichever form is differentiated.  Used for edges near the station
// (z > kZSeries), where 2 atanh(z) would lose accuracy to the 1 - z cancellation next to
// the edge's line; 1/X1, 1/X2 and the ratio's reciprocal come from one reciprocal of
// X1 X2 (d1 + d2), and 1/rho^2 from a second.
template <class T>
__device__ __forceinline__ T edge_ln_grad(T s1, T ell, T d1, T d2, T rho2, T& gs1, T& gs2,
                                          T& gd1, T& gd2, T& grho)
{
    const T s2 = s1 + ell;
    const T X1 = fabs(s1) + d1;
    const T X2 = fabs(s2) + d2;
    const bool pos1 = s1 >= T(0);
    const bool neg2 = s2 < T(0);
    // a station on the edge's line inside the segment (rho = 0) has no finite
    // ln(X1 X2 / rho^2); it takes the end form, finite there (the integral itself diverges;
    // the reference's edge term is 0 at such a point, geometry.cpp:48-50)
    const bool mid = !(pos1 || neg2) && rho2 > T(0);
    const T dsum = d1 + d2;
    const T X12 = X1 * X2;
    const T r = fast_rcp(X12 * dsum);
    const T irho = fast_rcp(rho2);  // used by the middle case only
    const T iX1 = X2 * dsum * r;
    const T iX2 = X1 * dsum * r;
    const T y = mid ? (X12 - rho2) * irho : ell * (X1 + X2) * (pos1 ? X2 : X1) * r;
    gs1 = -iX1;
    gs2 = iX2;
    gd1 = pos1 ? -iX1 : iX1;
    gd2 = neg2 ? -iX2 : iX2;
    grho = mid ? -irho : T(0);
    return log1p_t(y);
}

// L = 2 atanh(z) and dL/dz = 2 / (1 - z^2) for an edge with z = ell / (d1 + d2) <= kZSeries:
// in FP32 by their series (truncation below 2e-9 and 2e-8 relative at z = 0.1: the
// gradient needs no more), in FP64 from the library functions.
__device__ __forceinline__ float atanh2_dz(float z, float& g)
{
    const float s = z * z;
    float p = fmaf(s, 2.0f / 7.0f, 2.0f / 5.0f);
    p = fmaf(s, p, 2.0f / 3.0f);
    p = fmaf(s, p, 2.0f);
    float q = fmaf(s, 2.0f, 2.0f);  // 2 + 2s
    q = fmaf(s, q, 2.0f);           // 2 + 2s + 2s^2
    g = fmaf(s, q, 2.0f);           // 2 + 2s + 2s^2 + 2s^3
    return z * p;
}
__device__ __forceinline__ double atanh2_dz(double z, double& g)
{
    // the derivative of a near edge (z > kZSeries) is taken out again by the near branch,
    // so it only has to be finite: z = 1 for a station on the edge itself
    const double zc = fmin(z, 0.5);
    g = 2.0 / (1.0 - zc * zc);
    return 2.0 * atanh(z);
}

// th = atan2(N, D) with its partial derivatives dN = D / (N^2 + D^2), dD = -N / (N^2 + D^2).
// FP32: a solid angle seen from more than a few edge lengths away (|N| <= D / 8) takes the
// series of atan(a), a = N / D, and of its derivative 1 / (1 + a^2), with one reciprocal;
// larger angles the octant-reduced polynomial.
__device__ __forceinline__ void atan2_grad(float N, float D, float& th, float& dN, float& dD)
{
    if (fabsf(N) <= kSmallAngle * D) {
        const float rD = fast_rcp(D);
        const float a = N * rD;
        th = atan_small(a);
        dN = datan_small(a * a) * rD;
        dD = -dN * a;
    } else {
        th = atan2_t(N, D);
        const float q = fast_rcp(N * N + D * D);
        dN = D * q;
        dD = -N * q;
    }
}
__device__ __forceinline__ void atan2_grad(double N, double D, double& th, double& dN,
                                           double& dD)
{
    th = atan2(N, D);
    const double q = 1.0 / (N * N + D * D);
    dN = D * q;
    dD = -N * q;
}

// Forward quantities of a pair that the reverse sweep reads.
template <class T>
struct PairFwd {
    T px, py, pz, t, dcx, dcy, rda, rdb, rdc, s1b, s1c, dp1, dp2;
};

// The part of the reverse sweep that every seed source shares: from the seeds of the
// edge-frame numbers (s1, dperp of edges 1 and 2), of the solid angle's products
// rab, rac, rbc and of the three distances down to p and the in-plane shape numbers,
// adding the pair's p, torque and shape seeds into acc.  The sweep is linear in its
// seeds, so it runs once on the seeds of the common path (FAR: no near-edge seeds, whose
// zero terms are left out rather than computed) and, in the rare branch with an edge near
// the station, once more on that branch's own seeds (no solid-angle seeds).
template <class T, bool FAR>
__device__ __forceinline__ void reverse_tail(const FaceFrame<T>& F, const PairFwd<T>& p,
                                             T spx, T spy, T spz, T st_, T sda, T sdb,
                                             T sdc, T srho0, T spz2, T ss1b, T ss1c,
                                             T sdp1, T sdp2, T srab, T srac, T srbc, T sL0,
                                             T sL1, T sL2, T* acc)
{
    T sdcx, sdcy, Sda2, scu, scv, sl0c;
    if constexpr (FAR) {
        // s1b = t e1u - py e1v, dp1 = t e1v + py e1u; s1c = dcx e2u + dcy e2v,
        // dp2 = dcx e2v - dcy e2u (only dp1, dp2 carry seeds here)
        st_ = sdp1 * F.e1v;
        spy += sdp1 * F.e1u;
        sdcx = sdp2 * F.e2v;
        sdcy = -sdp2 * F.e2u;
        acc[S_PHI1] = fma(sdp1, p.s1b, acc[S_PHI1]);
        acc[S_PHI2] = fma(sdp2, p.s1c, acc[S_PHI2]);
        // rbc = t dcx - py dcy + pz2
        st_ += srbc * p.dcx;
        sdcx += srbc * p.t;
        spy -= srbc * p.dcy;
        sdcy -= srbc * p.py;
        spz2 = srbc;
        // rac = da2 - cu px - cv py ; rab = da2 - l0 px
        scu = -srac * p.px;
        scv = -srac * p.py;
        spx = -(srac * F.cu + srab * F.l0);
        spy -= srac * F.cv;
        sl0c = -srab * p.px;
        Sda2 = T(2) * (srac + srab) + sda * p.rda;
    } else {
        st_ += ss1b * F.e1u + sdp1 * F.e1v;
        spy += sdp1 * F.e1u - ss1b * F.e1v;
        sdcx = ss1c * F.e2u + sdp2 * F.e2v;
        sdcy = ss1c * F.e2v - sdp2 * F.e2u;
        // edge angles: (s_e1u, s_e1v) = ss1b (t, -py) + sdp1 (py, t) projected on
        // (-e1v, e1u) is sdp1 s1b - ss1b dp1; likewise sdp2 s1c - ss1c dp2 for edge 2
        acc[S_PHI1] = fma(sdp1, p.s1b, acc[S_PHI1]);
        acc[S_PHI1] = fma(-ss1b, p.dp1, acc[S_PHI1]);
        acc[S_PHI2] = fma(sdp2, p.s1c, acc[S_PHI2]);
        acc[S_PHI2] = fma(-ss1c, p.dp2, acc[S_PHI2]);
        Sda2 = sda * p.rda;
        scu = T(0);
        scv = T(0);
        sl0c = T(0);
    }
    // square roots and squares, with the seeds of the squared quantities doubled
    // (2 dF/d(d^2) = dF/dd / d, and d(x^2)/dx = 2x takes the factor back): hb = 2 s_db2,
    // hc = 2 s_dc2, Sda2 = 2 s_da2, Srho0 = 2 s_rho0, Spz2 = 2 s_pz2
    const T hb = sdb * p.rdb;
    const T hc = sdc * p.rdc;
    // dc2 = dcx^2 + dcy^2 + pz2
    sdcx += p.dcx * hc;
    sdcy += p.dcy * hc;
    // dcx = cu - px, dcy = cv - py
    scu += sdcx;
    spx -= sdcx;
    scv += sdcy;
    spy -= sdcy;
    // db2 = t^2 + rho0
    st_ += p.t * hb;
    // t = l0 - px
    sl0c += st_;
    spx -= st_;
    // da2 = px^2 + rho0 ; rho0 = py^2 + pz2 ; pz2 = pz^2
    spx += p.px * Sda2;
    T Srho0 = hb + Sda2;
    if constexpr (!FAR) Srho0 += T(2) * srho0;
    spy += p.py * Srho0;
    const T Spz2 = T(2) * spz2 + hc + Srho0;
    spz += p.pz * Spz2;
    // ---- accumulate
    acc[S_GPX] += spx;
    acc[S_GPY] += spy;
    acc[S_GPZ] += spz;
    acc[S_AXY] = fma(-spy, p.px, fma(spx, p.py, acc[S_AXY]));
    acc[S_AXZ] = fma(-spz, p.px, fma(spx, p.pz, acc[S_AXZ]));
    acc[S_AYZ] = fma(-spz, p.py, fma(spy, p.pz, acc[S_AYZ]));
    acc[S_L0] += sl0c + sL0;
    acc[S_CU] += scu;
    acc[S_CV] += scv;
    acc[S_L1] += sL1;
    acc[S_L2] += sL2;
}

// One (station, facet) pair: frame-form forward and its reverse sweep for the weights
// (lam on gz, mu on TMI), adding the pair's 19 facet seeds into acc (see the layout in
// SeedIdx).  Linear in (lam, mu).  Edge logs are L = 2 atanh(z), z = ell / (d1 + d2),
// which depends on the edge length and the two end distances only; an edge with
// z > kZSeries (station within about five edge lengths) takes the case form of
// edge_ln_grad, differentiated through (s1, s2, d1, d2, rho^2), in a branch that is rare
// for a wave.
template <class T>
__device__ __forceinline__ void pair_seeds(const FaceFrame<T>& F, double ox, double oy,
                                           double oz, T lam, T mu, T* acc)
{
    const T Px = T(ox - F.ax), Py = T(oy - F.ay), Pz = T(oz - F.az);
    PairFwd<T> p;
    // sums written as nested fmas: the compiler keeps the association as written
    p.px = fma(Px, F.ux, fma(Py, F.uy, fma(Pz, F.uz, F.du)));
    p.py = fma(Px, F.vx, fma(Py, F.vy, fma(Pz, F.vz, F.dv)));
    p.pz = fma(Px, F.nx, fma(Py, F.ny, fma(Pz, F.nz, F.dn)));
    const T px = p.px, py = p.py, pz = p.pz;
    // ---- forward
    const T pz2 = pz * pz;
    const T rho0 = fma(py, py, pz2);
    const T da2 = fma(px, px, rho0);
    const T t = F.l0 - px;
    const T db2 = fma(t, t, rho0);
    const T dcx = F.cu - px, dcy = F.cv - py;
    const T dc2 = fma(dcx, dcx, fma(dcy, dcy, pz2));
    p.t = t;
    p.dcx = dcx;
    p.dcy = dcy;
    // one reciprocal square root per distance gives both d and the 1/d of the reverse
    const T rda = fast_rsqrt(da2), rdb = fast_rsqrt(db2), rdc = fast_rsqrt(dc2);
    p.rda = rda;
    p.rdb = rdb;
    p.rdc = rdc;
    const T da = da2 * rda, db = db2 * rdb, dc = dc2 * rdc;
    const T rab = da2 - F.l0 * px;
    const T rac = da2 - (F.cu * px + F.cv * py);
    const T rbc = fma(t, dcx, fma(-py, dcy, pz2));
    const T D = da * db * dc + rab * dc + rac * db + rbc * da;
    const T twoA = F.l0 * F.cv;
    const T N = -twoA * pz;
    T th, dthN, dthD;
    atan2_grad(N, D, th, dthN, dthD);
    const T dp1 = t * F.e1v + py * F.e1u;
    const T dp2 = dcx * F.e2v - dcy * F.e2u;
    const T s1b = t * F.e1u - py * F.e1v;
    const T s1c = dcx * F.e2u + dcy * F.e2v;
    p.dp1 = dp1;
    p.dp2 = dp2;
    p.s1b = s1b;
    p.s1c = s1c;
    // edges: z = ell q, q = 1 / (d1 + d2)
    const T q0 = fast_rcp(da + db), q1 = fast_rcp(db + dc), q2 = fast_rcp(dc + da);
    const T z0 = F.l0 * q0, z1 = F.l1 * q1, z2 = F.l2 * q2;
    T g0, g1, g2;
    T lam0 = atanh2_dz(z0, g0);
    T lam1 = atanh2_dz(z1, g1);
    T lam2 = atanh2_dz(z2, g2);
    const bool n0 = z0 > T(kZSeries), n1 = z1 > T(kZSeries), n2 = z2 > T(kZSeries);
    const bool any_near = n0 || n1 || n2;
    if (any_near) {
        T a1, a2, a3, a4, a5;
        if (n0) lam0 = edge_ln_grad(-px, F.l0, da, db, rho0, a1, a2, a3, a4, a5);
        if (n1) lam1 = edge_ln_grad(s1b, F.l1, db, dc, dp1 * dp1 + pz2, a1, a2, a3, a4, a5);
        if (n2) lam2 = edge_ln_grad(s1c, F.l2, dc, da, dp2 * dp2 + pz2, a1, a2, a3, a4, a5);
    }
    const T es = py * lam0 + dp1 * lam1 + dp2 * lam2;
    const T pzth = pz * th;
    // ---- coefficient seeds
    acc[S_K1] += lam * (es + pzth + pzth);
    acc[S_W0] += mu * th;
    acc[S_W1] += mu * lam0;
    acc[S_W2] += mu * lam1;
    acc[S_W3] += mu * lam2;
    // ---- reverse
    const T ses = lam * F.K1;
    const T sth = T(2) * ses * pz + mu * F.W0;
    const T sl0 = ses * py + mu * F.W1;
    const T sl1 = ses * dp1 + mu * F.W2;
    const T sl2 = ses * dp2 + mu * F.W3;
    // far edges: L(ell, d1, d2); w = sL dL/dz q, dz/dell = q, dz/dd = -z q.  Taken for
    // every edge (finite for near ones too); the near branch removes the near edges' share
    // again, which keeps selects off the common path
    const T w0 = sl0 * g0 * q0;
    const T w1 = sl1 * g1 * q1;
    const T w2 = sl2 * g2 * q2;
    const T wz0 = w0 * z0, wz1 = w1 * z1, wz2 = w2 * z2;
    // th = atan2(N, D)
    const T sN = sth * dthN;
    const T sD = sth * dthD;
    acc[S_TWOA] -= sN * pz;
    // D = da db dc + rab dc + rac db + rbc da
    const T sda = sD * (db * dc + rbc) - (wz0 + wz2);
    const T sdb = sD * (da * dc + rac) - (wz0 + wz1);
    const T sdc = sD * (da * db + rab) - (wz1 + wz2);
    const T spz = T(2) * ses * th - sN * twoA;
    reverse_tail<T, true>(F, p, T(0), ses * lam0, spz, T(0), sda, sdb, sdc, T(0), T(0), T(0),
                          T(0), ses * lam1, ses * lam2, sD * dc, sD * db, sD * da, w0, w1, w2,
                          acc);
    if (any_near) {
        // the near edges' seeds, through the same tail, less the far form's seeds that the
        // common path took for those edges
        T spx = T(0), st_ = T(0), sdan = T(0), sdbn = T(0), sdcn = T(0), srho0 = T(0);
        T spz2 = T(0), ss1b = T(0), ss1c = T(0), sdp1 = T(0), sdp2 = T(0);
        T sL0 = T(0), sL1 = T(0), sL2 = T(0);
        T gs1, gs2, gd1, gd2, gr;
        if (n0) {
            // s1 = -px, s2 = t, d1 = da, d2 = db
            edge_ln_grad(-px, F.l0, da, db, rho0, gs1, gs2, gd1, gd2, gr);
            spx -= sl0 * gs1;
            st_ += sl0 * gs2;
            sdan += sl0 * gd1 + wz0;
            sdbn += sl0 * gd2 + wz0;
            srho0 += sl0 * gr;
            sL0 -= w0;
        }
        if (n1) {
            // s1 = s1b, s2 = s1b + l1, d1 = db, d2 = dc, rho^2 = dp1^2 + pz^2
            edge_ln_grad(s1b, F.l1, db, dc, dp1 * dp1 + pz2, gs1, gs2, gd1, gd2, gr);
            ss1b += sl1 * (gs1 + gs2);
            sL1 += sl1 * gs2 - w1;
            sdbn += sl1 * gd1 + wz1;
            sdcn += sl1 * gd2 + wz1;
            sdp1 += T(2) * dp1 * (sl1 * gr);
            spz2 += sl1 * gr;
        }
        if (n2) {
            edge_ln_grad(s1c, F.l2, dc, da, dp2 * dp2 + pz2, gs1, gs2, gd1, gd2, gr);
            ss1c += sl2 * (gs1 + gs2);
            sL2 += sl2 * gs2 - w2;
            sdcn += sl2 * gd1 + wz2;
            sdan += sl2 * gd2 + wz2;
            sdp2 += T(2) * dp2 * (sl2 * gr);
            spz2 += sl2 * gr;
        }
        reverse_tail<T, false>(F, p, spx, T(0), T(0), st_, sdan, sdbn, sdcn, srho0, spz2,
                               ss1b, ss1c, sdp1, sdp2, T(0), T(0), T(0), sL0, sL1, sL2, acc);
    }
}

// Seed sums.  The per-pair seeds of one facet over the stations of a chunk have random
// signs (station weights are residuals), so an FP32 running sum of n of them carries a
// relative error growing like sqrt(n) eps, and the finalize cancels the seeds against
// each other by about R/L (a far facet's edge and area terms of size L/R combine into a
// vertex sensitivity of size (L/R)^2).  The FP32 sums therefore run over at most
// kFlushStations stations and are then added into FP64 sums, one LDS column per thread
// (19 x 256 doubles, 38 KiB per block), written to the chunk's seeds at the end.  On the
// 192x192x64 contact model with 65,536 stations in 8 station chunks the FP32 adjoint's
// largest deviation from the FP64 one is 8.9e-6 of the largest component with 32 stations
// per flush, 6.5e-6 with 16 (1 % more time) and 6.0e-6 with 8 (3.5 % more): below 16 the
// per-pair roundings, not the sums, set the error (with 9 chunks and 16 stations: 7.4e-6;
// the chunk boundaries decide which stations share an FP32 run).
#ifndef LF_ADJOINT_FLUSH
#define LF_ADJOINT_FLUSH 16
#endif
constexpr int kFlushStations = LF_ADJOINT_FLUSH;
constexpr int kAdjointBlock = 256;

// FP32 pair kernel occupancy (waves per SIMD): the compiler's own choice for this loop
// (about 180 VGPRs) leaves 2; asking for more trades a few registers for issue slots.
#ifndef LF_ADJOINT_WAVES_PER_EU
#define LF_ADJOINT_WAVES_PER_EU 4
#endif

template <class T>
__device__ __forceinline__ void adjoint_pairs_body(const FaceFrame<T>* __restrict__ faces,
                                                   int64_t n_f, const StationW* __restrict__ st,
                                                   int64_t n_s, int64_t st_per_chunk,
                                                   double* __restrict__ seeds,
                                                   double (*sum64)[kAdjointBlock])
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    const bool active = f < n_f;
    const FaceFrame<T> F = faces[active ? f : n_f - 1];
    const int64_t i0 = (int64_t)chunk * st_per_chunk;
    const int64_t i1 = min(n_s, i0 + st_per_chunk);
    // 32-bit counts keep the loop tests on the scalar unit (it has no 64-bit less-than)
    const StationW* __restrict__ sp = st + (i0 < i1 ? i0 : 0);
    const int cnt = i0 < i1 ? (int)(i1 - i0) : 0;
    // each thread reads and writes only its own column: no barrier needed.  A compiler
    // memory barrier after each flush keeps the column in LDS: promoted to registers
    // (legal, since no other thread reads it) the 19 doubles would take 38 VGPRs from the
    // pair loop.
    double(*col)[kAdjointBlock] = sum64;
#pragma unroll
    for (int k = 0; k < kSeeds; ++k) col[k][threadIdx.x] = 0.0;
    __asm__ volatile("" ::: "memory");
    for (int b0 = 0; b0 < cnt; b0 += kFlushStations) {
        const int b1 = min(cnt, b0 + kFlushStations);
        T acc[kSeeds];
#pragma unroll
        for (int k = 0; k < kSeeds; ++k) acc[k] = T(0);
        // the next station record is requested (scalar loads) while this one is used
        StationW Sn = sp[b0];
        for (int i = b0; i < b1; ++i) {
            const StationW S = Sn;
            Sn = sp[i + 1 < b1 ? i + 1 : i];
            if (S.skip) continue;
            // T(w) from the scalar FP64 weight lands in a vector register, where the
            // products that use it issue faster than with a scalar-register operand
            pair_seeds(F, S.x, S.y, S.z, T(S.wg), T(S.wt), acc);
        }
#pragma unroll
        for (int k = 0; k < kSeeds; ++k) col[k][threadIdx.x] += (double)acc[k];
        __asm__ volatile("" ::: "memory");
    }
    if (active) {
        double* __restrict__ out = seeds + (int64_t)chunk * kSeeds * n_f + f;
#pragma unroll
        for (int k = 0; k < kSeeds; ++k) out[(int64_t)k * n_f] = col[k][threadIdx.x];
    }
}

__global__ __launch_bounds__(kAdjointBlock) __attribute__((amdgpu_waves_per_eu(LF_ADJOINT_WAVES_PER_EU)))
void k_adjoint_pairs_f32(const FaceFrame<float>* __restrict__ faces, int64_t n_f,
                         const StationW* __restrict__ st, int64_t n_s, int64_t st_per_chunk,
                         double* __restrict__ seeds)
{
    __shared__ double sum64[kSeeds][kAdjointBlock];
    adjoint_pairs_body<float>(faces, n_f, st, n_s, st_per_chunk, seeds, sum64);
}

__global__ __launch_bounds__(kAdjointBlock) void k_adjoint_pairs_f64(
    const FaceFrame<double>* __restrict__ faces, int64_t n_f, const StationW* __restrict__ st,
    int64_t n_s, int64_t st_per_chunk, double* __restrict__ seeds)
{
    __shared__ double sum64[kSeeds][kAdjointBlock];
    adjoint_pairs_body<double>(faces, n_f, st, n_s, st_per_chunk, seeds, sum64);
}

// Per-facet derivative matrix T (19 x 9): the derivative of the facet's contribution
// with respect to vertex coordinate k (a.x, a.y, a.z, b.x, ..., c.z of the facet's own
// corner order) is sum_m seed_m * T[m][k] for the per-pair (or accumulated) seeds of
// SeedIdx.  Built from the facet constants with forward-mode duals along the nine
// directions, in the corner order of the facet's frame record (frame_rotation), with the
// columns returned in the facet's own order.  facet_T_columns hands each column to
// column(k, col) as soon as it is formed (col[m] = T[m][k]), so that a caller that only
// contracts T with seeds keeps no 19 x 9 matrix; it returns false, calling nothing, for a
// degenerate facet (T = 0).
template <class Column>
__device__ __forceinline__ bool facet_T_columns(const double P0[3][3], double drho,
                                                const double Md[3], const double Ms[3],
                                                const double Fd[3], bool textbook,
                                                Column column)
{
    const int rot = frame_rotation(P0);
    double P[3][3];
    rotate_corners(P0, rot, P);
    FrameVals<double> fv;
    facet_frame(P, drho, Md, Ms, Fd, textbook, fv);
    if (fv.bad) return false;
    for (int kr = 0; kr < 9; ++kr) {
        // direction kr of the record's corners is column k of the facet's corners
        const int k = 3 * ((kr / 3 + rot) % 3) + kr % 3;
        Dual Q[3][3];
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < 3; ++j) Q[c][j] = Dual(P[c][j], (c * 3 + j == kr) ? 1.0 : 0.0);
        FrameVals<Dual> fd;
        facet_frame(Q, drho, Md, Ms, Fd, textbook, fd);
        double wxy = 0, wxz = 0, wyz = 0;
        for (int j = 0; j < 3; ++j) {
            wxy += fd.u[j].d * fv.v[j];
            wxz += fd.u[j].d * fv.n[j];
            wyz += fd.v[j].d * fv.n[j];
        }
        double col[kSeeds];
        if (kr < 3) {
            // p = R (o - a): moving a moves p by -R da
            col[S_GPX] = -fv.u[kr];
            col[S_GPY] = -fv.v[kr];
            col[S_GPZ] = -fv.n[kr];
        } else {
            col[S_GPX] = col[S_GPY] = col[S_GPZ] = 0.0;
        }
        col[S_AXY] = wxy;
        col[S_AXZ] = wxz;
        col[S_AYZ] = wyz;
        col[S_L0] = fd.l0.d;
        col[S_CU] = fd.cu.d;
        col[S_CV] = fd.cv.d;
        // d(angle) of the unit edge directions: (e_u, e_v) x d(e_u, e_v)
        col[S_PHI1] = fv.e1u * fd.e1v.d - fv.e1v * fd.e1u.d;
        col[S_L1] = fd.l1.d;
        col[S_PHI2] = fv.e2u * fd.e2v.d - fv.e2v * fd.e2u.d;
        col[S_L2] = fd.l2.d;
        col[S_TWOA] = fd.twoA.d;
        col[S_K1] = fd.K1.d;
        col[S_W0] = fd.W0.d;
        col[S_W1] = fd.W1.d;
        col[S_W2] = fd.W2.d;
        col[S_W3] = fd.W3.d;
        column(k, col);
    }
    return true;
}

__device__ __forceinline__ bool facet_T(const double P0[3][3], double drho, const double Md[3],
                                        const double Ms[3], const double Fd[3], bool textbook,
                                        double Tm[kSeeds][9])
{
    for (int m = 0; m < kSeeds; ++m)
        for (int k = 0; k < 9; ++k) Tm[m][k] = 0.0;
    return facet_T_columns(P0, drho, Md, Ms, Fd, textbook, [&](int k, const double* col) {
        for (int m = 0; m < kSeeds; ++m) Tm[m][k] = col[m];
    });
}

__device__ __forceinline__ void load_facet(int64_t f, const double* __restrict__ verts,
                                           const uint32_t* __restrict__ tri,
                                           const int32_t* __restrict__ front,
                                           const int32_t* __restrict__ back,
                                           const double* __restrict__ density,
                                           const double* __restrict__ mag, double P[3][3],
                                           double& drho, double Md[3], double Ms[3])
{
    for (int k = 0; k < 3; ++k) {
        const uint32_t vi = tri[3 * f + k];
        for (int j = 0; j < 3; ++j) P[k][j] = verts[3 * (int64_t)vi + j];
    }
    const int uf = front[f], ub = back[f];
    drho = density[uf] - (ub >= 0 ? density[ub] : 0.0);
    for (int j = 0; j < 3; ++j) {
        const double mf = mag[3 * uf + j], mb = ub >= 0 ? mag[3 * ub + j] : 0.0;
        Md[j] = mf - mb;
        Ms[j] = mf + mb;
    }
}

// Per facet: sum the chunks' FP64 seeds, then apply the facet's derivative matrix.
__global__ __launch_bounds__(128) void k_adjoint_finalize(
    const double* __restrict__ seeds, int n_chunks, int64_t n_f, const double* __restrict__ verts,
    const uint32_t* __restrict__ tri, const int32_t* __restrict__ front,
    const int32_t* __restrict__ back, const double* __restrict__ density,
    const double* __restrict__ mag, double fx, double fy, double fz, int tmi_mode,
    double* __restrict__ face_grad)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_f) return;
    double s[kSeeds];
    for (int k = 0; k < kSeeds; ++k) {
        double a = 0.0;
        for (int c = 0; c < n_chunks; ++c) a += seeds[((int64_t)c * kSeeds + k) * n_f + f];
        s[k] = a;
    }
    double P[3][3], drho, Md[3], Ms[3];
    load_facet(f, verts, tri, front, back, density, mag, P, drho, Md, Ms);
    const double Fd[3] = {fx, fy, fz};
    double* out = face_grad + 9 * f;
    const bool ok = facet_T_columns(P, drho, Md, Ms, Fd, tmi_mode == 1,
                                    [&](int k, const double* col) {
                                        double g = 0.0;
                                        for (int m = 0; m < kSeeds; ++m) g += s[m] * col[m];
                                        out[k] = g;
                                    });
    if (!ok)
        for (int k = 0; k < 9; ++k) out[k] = 0.0;
}

// ------------------------------------------------- Jacobian column norms (per DOF)

// T restricted to one axis: Tz[f][m][corner] = T[m][3*corner + axis].
template <class T>
__global__ __launch_bounds__(128) void k_facet_T_axis(
    int64_t n_f, const double* __restrict__ verts, const uint32_t* __restrict__ tri,
    const int32_t* __restrict__ front, const int32_t* __restrict__ back,
    const double* __restrict__ density, const double* __restrict__ mag, double fx, double fy,
    double fz, int tmi_mode, int axis, T* __restrict__ Tax)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_f) return;
    double P[3][3], drho, Md[3], Ms[3];
    load_facet(f, verts, tri, front, back, density, mag, P, drho, Md, Ms);
    const double Fd[3] = {fx, fy, fz};
    double Tm[kSeeds][9];
    facet_T(P, drho, Md, Ms, Fd, tmi_mode == 1, Tm);
    T* out = Tax + (int64_t)f * kSeeds * 3;
    for (int m = 0; m < kSeeds; ++m)
        for (int c = 0; c < 3; ++c) out[m * 3 + c] = (T)Tm[m][3 * c + axis];
}

// lane = facet; for each station of the block (wave-uniform) the per-pair seeds for a
// unit weight on gz and on TMI, contracted with T: the derivative of gz_s and TMI_s with
// respect to the chosen coordinate of each facet corner, G[s][3 f + corner].
// blockIdx.y splits the block's stations into groups of kStationGroup.
constexpr int kStationGroup = 16;
template <class T>
__global__ __launch_bounds__(256) void k_station_facet_grad(
    const FaceFrame<T>* __restrict__ faces, const T* __restrict__ Tax, int64_t n_f,
    const StationW* __restrict__ st, int n_blk, float* __restrict__ Gg, float* __restrict__ Gt)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_f) return;
    const FaceFrame<T> F = faces[f];
    const T* tf = Tax + (int64_t)f * kSeeds * 3;
    const int s0 = blockIdx.y * kStationGroup;
    const int s1 = min(n_blk, s0 + kStationGroup);
    for (int s = s0; s < s1; ++s) {
        const StationW S = st[s];
        T ag[kSeeds], at[kSeeds];
#pragma unroll
        for (int k = 0; k < kSeeds; ++k) ag[k] = at[k] = T(0);
        pair_seeds(F, S.x, S.y, S.z, T(1), T(0), ag);
        pair_seeds(F, S.x, S.y, S.z, T(0), T(1), at);
        for (int c = 0; c < 3; ++c) {
            T dg = 0, dt = 0;
#pragma unroll
            for (int m = 0; m < kSeeds; ++m) {
                dg += ag[m] * tf[m * 3 + c];
                dt += at[m] * tf[m * 3 + c];
            }
            Gg[(int64_t)s * 3 * n_f + 3 * f + c] = (float)dg;
            Gt[(int64_t)s * 3 * n_f + 3 * f + c] = (float)dt;
        }
    }
}

// ------------------------------------------------- Jacobian-vector products

// J v along a vertex tangent dX (3 per vertex): the facet constants of the records move
// by dq[m] = sum_k T[m][k] dX_k (T as in the finalize, columns in the facet's own corner
// order), and a pair's contribution to gz_s moves by sum_m seeds_g[m] dq[m], where
// seeds_g are the pair's seeds for a unit weight on gz (likewise TMI).
constexpr int kSeedsPad = 20;

template <class T>
__global__ __launch_bounds__(128) void k_facet_tangent(
    int64_t n_f, const double* __restrict__ verts, const uint32_t* __restrict__ tri,
    const int32_t* __restrict__ front, const int32_t* __restrict__ back,
    const double* __restrict__ density, const double* __restrict__ mag, double fx, double fy,
    double fz, int tmi_mode, const double* __restrict__ vtan, T* __restrict__ dq)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_f) return;
    double P[3][3], drho, Md[3], Ms[3];
    load_facet(f, verts, tri, front, back, density, mag, P, drho, Md, Ms);
    double dX[9];
    for (int k = 0; k < 3; ++k) {
        const uint32_t vi = tri[3 * f + k];
        for (int j = 0; j < 3; ++j) dX[3 * k + j] = vtan[3 * (int64_t)vi + j];
    }
    const double Fd[3] = {fx, fy, fz};
    double acc[kSeeds];
    for (int m = 0; m < kSeeds; ++m) acc[m] = 0.0;
    const bool ok = facet_T_columns(P, drho, Md, Ms, Fd, tmi_mode == 1,
                                    [&](int k, const double* col) {
                                        for (int m = 0; m < kSeeds; ++m) acc[m] += col[m] * dX[k];
                                    });
    T* out = dq + f * kSeedsPad;
    for (int m = 0; m < kSeeds; ++m) out[m] = ok ? (T)acc[m] : T(0);
    out[kSeeds] = T(0);
}

// One pair's change of gz and TMI along the facet constants' tangent dq (forward mode):
// the forward of pair_seeds, operation for operation, with each quantity's tangent
// beside it, so that dgz = sum_m seeds_g[m] dq[m] for pair_seeds' own seeds.  The
// station offset moves with the translation and rotation tangents,
//   dp = dq[GP] + (py w_xy + pz w_xz, -px w_xy + pz w_yz, -px w_xz - py w_yz)
// (the transposes of the torque seeds A_jk = s_j p_k - s_k p_j), an in-plane edge
// direction turns by its angle tangent (d e = dphi (-e_v, e_u)), and N takes 2A's own
// tangent.  Near edges use the case forms' partial derivatives from edge_ln_grad.
template <class T>
__device__ __forceinline__ void pair_tangent(const FaceFrame<T>& F, double ox, double oy,
                                             double oz, const T* __restrict__ q, T& dgz,
                                             T& dtmi)
{
    const T Px = T(ox - F.ax), Py = T(oy - F.ay), Pz = T(oz - F.az);
    const T px = fma(Px, F.ux, fma(Py, F.uy, fma(Pz, F.uz, F.du)));
    const T py = fma(Px, F.vx, fma(Py, F.vy, fma(Pz, F.vz, F.dv)));
    const T pz = fma(Px, F.nx, fma(Py, F.ny, fma(Pz, F.nz, F.dn)));
    const T dpx = fma(py, q[S_AXY], fma(pz, q[S_AXZ], q[S_GPX]));
    const T dpy = fma(pz, q[S_AYZ], fma(-px, q[S_AXY], q[S_GPY]));
    const T dpz = fma(-py, q[S_AYZ], fma(-px, q[S_AXZ], q[S_GPZ]));
    const T pz2 = pz * pz, dpz2 = T(2) * pz * dpz;
    const T rho0 = fma(py, py, pz2), drho0 = fma(T(2) * py, dpy, dpz2);
    const T da2 = fma(px, px, rho0), dda2 = fma(T(2) * px, dpx, drho0);
    const T t = F.l0 - px, dt = q[S_L0] - dpx;
    const T db2 = fma(t, t, rho0), ddb2 = fma(T(2) * t, dt, drho0);
    const T dcx = F.cu - px, dcy = F.cv - py;
    const T ddcx = q[S_CU] - dpx, ddcy = q[S_CV] - dpy;
    const T dc2 = fma(dcx, dcx, fma(dcy, dcy, pz2));
    const T ddc2 = fma(T(2) * dcx, ddcx, fma(T(2) * dcy, ddcy, dpz2));
    const T rda = fast_rsqrt(da2), rdb = fast_rsqrt(db2), rdc = fast_rsqrt(dc2);
    const T da = da2 * rda, db = db2 * rdb, dc = dc2 * rdc;
    const T dda = T(0.5) * dda2 * rda, ddb = T(0.5) * ddb2 * rdb, ddc = T(0.5) * ddc2 * rdc;
    const T rab = da2 - F.l0 * px;
    const T drab = dda2 - (q[S_L0] * px + F.l0 * dpx);
    const T rac = da2 - (F.cu * px + F.cv * py);
    const T drac = dda2 - (q[S_CU] * px + F.cu * dpx + q[S_CV] * py + F.cv * dpy);
    const T rbc = fma(t, dcx, fma(-py, dcy, pz2));
    const T drbc = dt * dcx + t * ddcx - dpy * dcy - py * ddcy + dpz2;
    const T D = da * db * dc + rab * dc + rac * db + rbc * da;
    const T dD = dda * (db * dc + rbc) + ddb * (da * dc + rac) + ddc * (da * db + rab) +
                 drab * dc + drac * db + drbc * da;
    const T twoA = F.l0 * F.cv;
    const T N = -twoA * pz;
    const T dN = -(q[S_TWOA] * pz + twoA * dpz);
    T th, dthN, dthD;
    atan2_grad(N, D, th, dthN, dthD);
    const T dth = dthN * dN + dthD * dD;
    const T dp1 = t * F.e1v + py * F.e1u;
    const T dp2 = dcx * F.e2v - dcy * F.e2u;
    const T s1b = t * F.e1u - py * F.e1v;
    const T s1c = dcx * F.e2u + dcy * F.e2v;
    const T ddp1 = dt * F.e1v + dpy * F.e1u + s1b * q[S_PHI1];
    const T ddp2 = ddcx * F.e2v - ddcy * F.e2u + s1c * q[S_PHI2];
    const T q0 = fast_rcp(da + db), q1 = fast_rcp(db + dc), q2 = fast_rcp(dc + da);
    const T z0 = F.l0 * q0, z1 = F.l1 * q1, z2 = F.l2 * q2;
    T g0, g1, g2;
    T lam0 = atanh2_dz(z0, g0);
    T lam1 = atanh2_dz(z1, g1);
    T lam2 = atanh2_dz(z2, g2);
    // dz = q (dell - z (dd1 + dd2))
    T dlam0 = g0 * q0 * (q[S_L0] - z0 * (dda + ddb));
    T dlam1 = g1 * q1 * (q[S_L1] - z1 * (ddb + ddc));
    T dlam2 = g2 * q2 * (q[S_L2] - z2 * (ddc + dda));
    if (z0 > T(kZSeries) || z1 > T(kZSeries) || z2 > T(kZSeries)) {
        T gs1, gs2, gd1, gd2, gr;
        if (z0 > T(kZSeries)) {
            // s1 = -px, s2 = t, d1 = da, d2 = db, rho^2 = rho0
            lam0 = edge_ln_grad(-px, F.l0, da, db, rho0, gs1, gs2, gd1, gd2, gr);
            dlam0 = -gs1 * dpx + gs2 * dt + gd1 * dda + gd2 * ddb + gr * drho0;
        }
        if (z1 > T(kZSeries)) {
            // s1 = s1b, s2 = s1b + l1, d1 = db, d2 = dc, rho^2 = dp1^2 + pz^2
            const T ds1 = dt * F.e1u - dpy * F.e1v - dp1 * q[S_PHI1];
            lam1 = edge_ln_grad(s1b, F.l1, db, dc, dp1 * dp1 + pz2, gs1, gs2, gd1, gd2, gr);
            dlam1 = gs1 * ds1 + gs2 * (ds1 + q[S_L1]) + gd1 * ddb + gd2 * ddc +
                    gr * (T(2) * dp1 * ddp1 + dpz2);
        }
        if (z2 > T(kZSeries)) {
            // s1 = s1c, s2 = s1c + l2, d1 = dc, d2 = da, rho^2 = dp2^2 + pz^2
            const T ds1 = ddcx * F.e2u + ddcy * F.e2v - dp2 * q[S_PHI2];
            lam2 = edge_ln_grad(s1c, F.l2, dc, da, dp2 * dp2 + pz2, gs1, gs2, gd1, gd2, gr);
            dlam2 = gs1 * ds1 + gs2 * (ds1 + q[S_L2]) + gd1 * ddc + gd2 * dda +
                    gr * (T(2) * dp2 * ddp2 + dpz2);
        }
    }
    const T es = py * lam0 + dp1 * lam1 + dp2 * lam2;
    const T des = dpy * lam0 + py * dlam0 + ddp1 * lam1 + dp1 * dlam1 + ddp2 * lam2 + dp2 * dlam2;
    const T pzth = pz * th;
    dgz = q[S_K1] * (es + pzth + pzth) + F.K1 * (des + T(2) * (dpz * th + pz * dth));
    dtmi = q[S_W0] * th + F.W0 * dth + q[S_W1] * lam0 + F.W1 * dlam0 + q[S_W2] * lam1 +
           F.W2 * dlam1 + q[S_W3] * lam2 + F.W3 * dlam2;
}

// lane = station, facets of the block's chunk wave-uniform: per pair the forward-mode
// change of gz and TMI along the facet's dq (pair_tangent), added into the lane's FP64
// sums (a closed surface's facet terms cancel against each other); one partial per
// (facet chunk, station), summed by k_reduce_chunks.
template <class T>
__global__ __launch_bounds__(256) void k_jvp_pairs(
    const FaceFrame<T>* __restrict__ faces, const T* __restrict__ dq, int64_t n_f,
    int64_t faces_per_chunk, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, int64_t n_s, double* __restrict__ part_gz,
    double* __restrict__ part_tmi)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    const bool active = s < n_s;
    const int64_t sc = active ? s : n_s - 1;
    const double ox = sx[sc], oy = sy[sc], oz = sz[sc];
    const int64_t f0 = (int64_t)chunk * faces_per_chunk;
    const int64_t f1 = min(n_f, f0 + faces_per_chunk);
    double g = 0.0, t = 0.0;
    for (int64_t f = f0; f < f1; ++f) {
        const FaceFrame<T> F = faces[f];
        const T* __restrict__ q = dq + f * kSeedsPad;
        T pg, pt;
        pair_tangent(F, ox, oy, oz, q, pg, pt);
        g += (double)pg;
        t += (double)pt;
    }
    if (active) {
        part_gz[(int64_t)chunk * n_s + s] = g;
        part_tmi[(int64_t)chunk * n_s + s] = t;
    }
}

// D[s][v] = sum of G[s][corner entries of v]
__global__ __launch_bounds__(256) void k_gather_block(const float* __restrict__ G,
                                                      const int64_t* __restrict__ off,
                                                      const int32_t* __restrict__ ent,
                                                      int64_t n_v, int64_t n_f,
                                                      float* __restrict__ D)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (v >= n_v) return;
    const float* g = G + (int64_t)s * 3 * n_f;
    float acc = 0.f;
    for (int64_t e = off[v]; e < off[v + 1]; ++e) acc += g[ent[e]];
    D[(int64_t)s * n_v + v] = acc;
}

// out[j] += sum_s (sum_e w_e D[s][vertex_e])^2 over the entries of DOF j.  One wave per
// DOF, one lane per station of the block (n_blk <= 64); the squares are summed across the
// wave in a fixed butterfly order and one lane adds them to out[j].
__global__ __launch_bounds__(256) void k_dof_colnorm(const float* __restrict__ D, int n_blk,
                                                     int64_t n_v, int64_t n_dof,
                                                     const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ vert,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ out)
{
    const int64_t j = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int s = threadIdx.x % 64;
    if (j >= n_dof) return;  // uniform across the wave
    double val = 0.0;
    if (s < n_blk) {
        const float* d = D + (int64_t)s * n_v;
        for (int64_t e = off[j]; e < off[j + 1]; ++e) val += w[e] * (double)d[vert[e]];
    }
    double sq = val * val;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) sq += __shfl_xor(sq, m, 64);
    if (s == 0) out[j] += sq;
}

__global__ __launch_bounds__(256) void k_vertex_gather(const double* __restrict__ face_grad,
                                                       const int64_t* __restrict__ off,
                                                       const int32_t* __restrict__ ent,
                                                       int64_t n_v, double* __restrict__ grad)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_v) return;
    double gx = 0, gy = 0, gz = 0;
    for (int64_t e = off[v]; e < off[v + 1]; ++e) {
        const int64_t fc = ent[e];
        const double* g = face_grad + 9 * (fc / 3) + 3 * (fc % 3);
        gx += g[0];
        gy += g[1];
        gz += g[2];
    }
    grad[3 * v] = gx;
    grad[3 * v + 1] = gy;
    grad[3 * v + 2] = gz;
}

__global__ __launch_bounds__(256) void k_pack_stations(const double* __restrict__ sx,
                                                       const double* __restrict__ sy,
                                                       const double* __restrict__ sz,
                                                       const double* __restrict__ wg,
                                                       const double* __restrict__ wt,
                                                       int64_t n_s, StationW* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_s) return;
    StationW s;
    s.x = sx[i];
    s.y = sy[i];
    s.z = sz[i];
    s.wg = wg[i];
    s.wt = wt[i];
    s.skip = (s.wg == 0.0 && s.wt == 0.0) ? 1 : 0;
    s.pad0 = s.pad1 = 0.0;
    out[i] = s;
}

__global__ __launch_bounds__(256) void k_fill_zero(double* __restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ----------------------------------------------------------------------- host side

inline int adjoint_init(AdjointWork& w, int64_t n_v, int64_t n_f, int64_t n_s,
                        const uint32_t* faces, int precision)
{
    w.n_v = n_v;
    w.n_f = n_f;
    w.n_s = n_s;
    w.precision = precision;
    // station chunks (grid = facet blocks x chunks): at least 16 stations per chunk and at
    // most 8 GiB of FP64 seeds (a seed buffer of one chunk always fits), and among those
    // the count that minimizes a cost model in units of one station's work for one round
    // of resident blocks:
    //   rounds(c) * (stations per chunk + kBlockFixed) + c * n_f * kFinalizePerFacet,
    // rounds(c) = ceil(facet blocks * c / resident blocks).  The first term is the pair
    // kernel (whose last round may be partly idle, and which pays a fixed set-up per
    // block), the second the finalize's reads of every chunk's seeds.  Measured at
    // 140,384 facets x 16,384 stations: 120 chunks of 137 stations (the earlier rule of
    // at least 65,536 blocks) took 9.8 ms in the pair kernel and 3.0 ms in the finalize.
    const int64_t face_blocks = std::max<int64_t>(1, (n_f + kAdjointBlock - 1) / kAdjointBlock);
    const size_t seed_bytes = (size_t)kSeeds * std::max<int64_t>(n_f, 1) * sizeof(double);
    int64_t c_max = std::min<int64_t>(std::max<int64_t>(1, (n_s + 15) / 16), 65535);
    c_max = std::max<int64_t>(1, std::min<int64_t>(c_max, (int64_t)((8ull << 30) / seed_bytes)));
    int dev = 0, cus = 256, per_cu = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
    }
    if (precision == 1)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adjoint_pairs_f32, kAdjointBlock, 0);
    else
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adjoint_pairs_f64, kAdjointBlock, 0);
    const int64_t slots = (int64_t)cus * std::max(1, per_cu);
    constexpr double kBlockFixed = 4.0;          // record load, column set-up, seed write
    constexpr double kFinalizePerFacet = 1.8e-4; // 152 bytes of seeds per facet and chunk
    int64_t chunks = 1;
    double best = 1e300;
    for (int64_t c = 1; c <= c_max; ++c) {
        const int64_t spc = (n_s + c - 1) / c;
        const int64_t rounds = (face_blocks * c + slots - 1) / slots;
        const double cost = (double)rounds * ((double)spc + kBlockFixed) +
                            (double)c * (double)n_f * kFinalizePerFacet;
        if (cost < best) {
            best = cost;
            chunks = c;
        }
    }
    w.st_per_chunk = std::max<int64_t>(1, (n_s + chunks - 1) / chunks);
    w.n_chunks = (int)std::max<int64_t>(1, (n_s + w.st_per_chunk - 1) / w.st_per_chunk);
    const size_t nseed = (size_t)w.n_chunks * kSeeds * std::max<int64_t>(n_f, 1);
    if (hipMalloc(&w.d_seeds, nseed * sizeof(double)) != hipSuccess) return -1;
    if (hipMalloc((void**)&w.d_st, sizeof(StationW) * std::max<int64_t>(n_s, 1)) != hipSuccess) return -1;
    if (hipMalloc((void**)&w.d_face_grad, sizeof(double) * 9 * std::max<int64_t>(n_f, 1)) != hipSuccess) return -1;
    if (hipMalloc((void**)&w.d_csr_off, sizeof(int64_t) * (n_v + 1)) != hipSuccess) return -1;
    if (hipMalloc((void**)&w.d_csr_ent, sizeof(int32_t) * std::max<int64_t>(3 * n_f, 1)) != hipSuccess) return -1;
    if (hipMalloc((void**)&w.d_wg, sizeof(double) * std::max<int64_t>(n_s, 1)) != hipSuccess) return -1;
    if (hipMalloc((void**)&w.d_wt, sizeof(double) * std::max<int64_t>(n_s, 1)) != hipSuccess) return -1;
    if (hipMalloc((void**)&w.d_zero, sizeof(double) * std::max<int64_t>(n_s, 1)) != hipSuccess) return -1;
    if (hipMalloc((void**)&w.d_grad, sizeof(double) * std::max<int64_t>(3 * n_v, 1)) != hipSuccess) return -1;
    // vertex -> (facet, corner) CSR, in facet order
    std::vector<int64_t> off(n_v + 1, 0);
    for (int64_t i = 0; i < 3 * n_f; ++i) off[faces[i] + 1]++;
    for (int64_t v = 0; v < n_v; ++v) off[v + 1] += off[v];
    std::vector<int32_t> ent(std::max<int64_t>(3 * n_f, 1));
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    for (int64_t i = 0; i < 3 * n_f; ++i) ent[pos[faces[i]]++] = (int32_t)i;
    if (hipMemcpy(w.d_csr_off, off.data(), sizeof(int64_t) * (n_v + 1), hipMemcpyHostToDevice) != hipSuccess) return -1;
    if (n_f && hipMemcpy(w.d_csr_ent, ent.data(), sizeof(int32_t) * 3 * n_f, hipMemcpyHostToDevice) != hipSuccess) return -1;
    std::vector<double> z(std::max<int64_t>(n_s, 1), 0.0);
    if (hipMemcpy(w.d_zero, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice) != hipSuccess) return -1;
    return 0;
}

inline void adjoint_free(AdjointWork& w)
{
    void* bufs[] = {w.d_seeds, w.d_st, w.d_face_grad, w.d_csr_off, w.d_csr_ent,
                    w.d_wg,    w.d_wt, w.d_zero,      w.d_grad};
    for (void* p : bufs)
        if (p) hipFree(p);
    w = AdjointWork();
}

inline hipError_t adjoint_pairs(AdjointWork& w, const AdjointLaunch& L)
{
    if (L.n_s == 0 || L.n_f == 0) return hipSuccess;
    const double* wg = L.wg ? L.wg : w.d_zero;
    const double* wt = L.wt ? L.wt : w.d_zero;
    hipLaunchKernelGGL(k_pack_stations, dim3((unsigned)((L.n_s + 255) / 256)), dim3(256), 0,
                       L.stream, L.sx, L.sy, L.sz, wg, wt, L.n_s, w.d_st);
    const dim3 grid((unsigned)((L.n_f + kAdjointBlock - 1) / kAdjointBlock), (unsigned)w.n_chunks);
    if (L.precision == 1)
        hipLaunchKernelGGL(k_adjoint_pairs_f32, grid, dim3(kAdjointBlock), 0, L.stream, L.f32, L.n_f,
                           w.d_st, L.n_s, w.st_per_chunk, (double*)w.d_seeds);
    else
        hipLaunchKernelGGL(k_adjoint_pairs_f64, grid, dim3(kAdjointBlock), 0, L.stream, L.f64, L.n_f,
                           w.d_st, L.n_s, w.st_per_chunk, (double*)w.d_seeds);
    return hipGetLastError();
}

inline hipError_t adjoint_finalize(AdjointWork& w, const AdjointLaunch& L, const double* verts,
                                   const uint32_t* tri, const int32_t* front,
                                   const int32_t* back, const double* density,
                                   const double* mag, const double* fdir, int tmi_mode)
{
    if (L.n_f == 0) return hipSuccess;
    const dim3 grid((unsigned)((L.n_f + 127) / 128));
    const int nch = L.n_s == 0 ? 0 : w.n_chunks;
    hipLaunchKernelGGL(k_adjoint_finalize, grid, dim3(128), 0, L.stream,
                       (const double*)w.d_seeds, nch, L.n_f, verts, tri, front, back, density,
                       mag, fdir[0], fdir[1], fdir[2], tmi_mode, w.d_face_grad);
    return hipGetLastError();
}

inline hipError_t adjoint_gather(AdjointWork& w, const AdjointLaunch& L)
{
    if (L.n_v == 0) return hipSuccess;
    if (L.n_f == 0) {
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((3 * L.n_v + 255) / 256)), dim3(256), 0,
                           L.stream, L.grad, 3 * L.n_v);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_vertex_gather, dim3((unsigned)((L.n_v + 255) / 256)), dim3(256), 0,
                       L.stream, w.d_face_grad, w.d_csr_off, w.d_csr_ent, L.n_v, L.grad);
    return hipGetLastError();
}

}  // namespace lfk
