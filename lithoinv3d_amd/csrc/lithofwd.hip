// lithofwd.hip -- gfx950 kernels and the C ABI (include/lithofwd.h) of the polyhedral
// gz / TMI forward model and its adjoint.
//
// Kernels
//   k_face_setup      one thread per facet: builds the station-independent facet record
//                     (normal, edge frame, coefficients) from FP64 vertices
//                     (the per-pair recomputation in gravity_forward.cpp:17-35 and
//                     magnetic_forward.cpp:317-341 hoisted out of the station loop).
//   k_forward_f32c    lane = station, wave-uniform facet: facet records stream through
//                     LDS in stages of 64 and are read at a uniform address; each lane
//                     keeps its own accumulators (FP32 sums over groups of 16 facets,
//                     Kahan-compensated across groups), so the facet reduction needs no
//                     cross-lane traffic; two stations per lane share each record
//                     read.  Grid = station tiles x facet chunks, about 128 blocks per CU.
//   k_forward_f64     the same map in FP64 with the reference's operation sequence;
//                     records arrive through scalar loads.
//   k_reduce_chunks   sums the per-chunk partials of each station in a fixed order.
//   k_adjoint_*       see adjoint.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lithofwd.h"
#include "pair_math.h"
#include "adjoint.h"

using namespace lfk;

namespace {

constexpr double kEps = 1e-15;                         // geometry.cpp:5
constexpr int kBlock = 256;

// ------------------------------------------------------------------- face setup

__global__ __launch_bounds__(128) void k_face_setup(
    const double* __restrict__ verts, const uint32_t* __restrict__ tri,
    const int32_t* __restrict__ front, const int32_t* __restrict__ back,
    const double* __restrict__ density, const double* __restrict__ mag, double fx,
    double fy, double fz, int tmi_mode, int64_t n_f, int64_t n_v, int n_u,
    FaceF32* __restrict__ out32, FaceF64* __restrict__ out64,
    FaceFrame<double>* __restrict__ outfr64, double* __restrict__ geo64)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_f) return;
    // a facet with a vertex or unit id out of range (possible only for models passed in
    // device memory, which the host cannot check) becomes a zero-area facet with no
    // properties: it contributes nothing
    const uint32_t v0 = tri[3 * f], v1 = tri[3 * f + 1], v2 = tri[3 * f + 2];
    int uf = front[f], ub = back[f];
    const bool valid = (int64_t)v0 < n_v && (int64_t)v1 < n_v && (int64_t)v2 < n_v &&
                       uf >= 0 && uf < n_u && ub >= -1 && ub < n_u;
    double P[3][3];
    const uint32_t vid[3] = {v0, v1, v2};
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) P[k][j] = valid ? verts[3 * (int64_t)vid[k] + j] : 0.0;
    if (!valid) {
        uf = 0;
        ub = -1;
    }
    const double drho = valid ? density[uf] - (ub >= 0 ? density[ub] : 0.0) : 0.0;
    double Md[3], Ms[3];
    for (int j = 0; j < 3; ++j) {
        const double mf = valid ? mag[3 * uf + j] : 0.0;
        const double mb = ub >= 0 ? mag[3 * ub + j] : 0.0;
        Md[j] = mf - mb;  // the Omega term flips with the winding
        Ms[j] = mf + mb;  // the reference edge term does not: t = n x e is reversal-invariant
    }
    const double Fd[3] = {fx, fy, fz};
    const bool textbook = tmi_mode == LF_TMI_TEXTBOOK;
    FrameVals<double> fv;
    facet_frame(P, drho, Md, Ms, Fd, textbook, fv);
    // frame-form records use the corner order that makes the frame exact in FP32 where
    // the facet allows it (the FP64 reference-form record keeps the facet's own order)
    const int rot = frame_rotation(P);
    double Q[3][3];
    rotate_corners(P, rot, Q);
    FrameVals<double> fq;
    if (rot) facet_frame(Q, drho, Md, Ms, Fd, textbook, fq);
    const FrameVals<double>& fr = rot ? fq : fv;
    if (out64) {
        FaceF64 r;
        for (int j = 0; j < 3; ++j) {
            r.a[j] = P[0][j];
            r.b[j] = P[1][j];
            r.c[j] = P[2][j];
            r.n[j] = fv.n[j];
            for (int k = 0; k < 3; ++k) {
                r.eh[k][j] = fv.eh[k][j];
                r.perp[k][j] = fv.edge_ok[k] ? fv.perp[k][j] : 0.0;
            }
        }
        r.kz = fv.kz;
        r.wom = fv.wom;
        for (int k = 0; k < 3; ++k) r.we[k] = fv.we[k];
        r.pad = 0.0;
        out64[f] = r;
    }
    if (outfr64) {
        FaceFrame<double> r;
        store_frame(fr, Q[0], r);
        outfr64[f] = r;
    }
    if (out32) {
        FaceF32 r;
        store_frame(fr, Q[0], r, (uint32_t)f, true);
        out32[f] = r;
    }
    if (geo64) {
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < 3; ++j) geo64[9 * f + 3 * c + j] = Q[c][j];
    }
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

// Block -> (station tile, facet chunk).  Consecutive block ids are dealt round-robin to
// the 8 XCDs; remap so that each XCD works through a contiguous range of chunks and the
// facet records of those chunks stay in that XCD's L2.
__device__ __forceinline__ void block_coords(int n_tiles, int& tile, int& chunk)
{
    int64_t nb = gridDim.x;
    int64_t b = blockIdx.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    tile = (int)(b % n_tiles);
    chunk = (int)(b / n_tiles);
}

// ----------------------------------------------------------------- FP32C forward

// Optional occupancy target for the pair kernel (waves per SIMD), a build-time tuning knob.
#ifdef LF_FORWARD_WAVES_PER_EU
#define LF_FORWARD_OCC __attribute__((amdgpu_waves_per_eu(LF_FORWARD_WAVES_PER_EU)))
#else
#define LF_FORWARD_OCC
#endif

// One (station, facet) pair of the FP32C forward, added into the station's FP32 sums of
// the current facet group (gs, ts).  F is the facet record (wave-uniform), f its index
// (for the FP64 branch).
template <bool GZ, bool TMI>
__device__ __forceinline__ void forward_pair(const FaceF32& F, const double* __restrict__ geo64,
                                             int64_t f, double ox, double oy, double oz,
                                             float& gs, float& ts, unsigned& nfb)
{
    // station offset from the record's origin: FP64 difference, one rounding; in the
    // frame, plus the origin's offset d, it is the offset from vertex a
    const float Px = (float)(ox - F.ax);
    const float Py = (float)(oy - F.ay);
    const float Pz = (float)(oz - F.az);
    const float px = fmaf(Px, F.ux, fmaf(Py, F.uy, fmaf(Pz, F.uz, F.du)));
    const float py = fmaf(Px, F.vx, fmaf(Py, F.vy, fmaf(Pz, F.vz, F.dv)));
    const float pz = fmaf(Px, F.nx, fmaf(Py, F.ny, fmaf(Pz, F.nz, F.dn)));  // h = -pz
    const float pz2 = pz * pz;
    const float rho0 = fmaf(py, py, pz2);
    const float da2 = fmaf(px, px, rho0);
    const float t = F.l0 - px;  // (b - o).u
    const float db2 = fmaf(t, t, rho0);
    const float dcx = F.cu - px, dcy = F.cv - py;
    const float dc2 = fmaf(dcx, dcx, fmaf(dcy, dcy, pz2));
    const float da = sqrt_nr(da2), db = sqrt_nr(db2), dc = sqrt_nr(dc2);
    // solid angle: N = r_a.(r_b x r_c) = 2A*h, D from the pairwise dots
    const float rab = fmaf(-F.l0, px, da2);
    const float rac = da2 - fmaf(F.cu, px, F.cv * py);
    const float rbc = fmaf(t, dcx, fmaf(-py, dcy, pz2));
    const float D = fmaf(da * db, dc, fmaf(rab, dc, fmaf(rac, db, rbc * da)));
    const float N = -(F.l0 * F.cv) * pz;
    // edge logs 2 atanh(ell / (d1 + d2)) by the series for every edge more than about five
    // edge lengths away (d1 + d2 >= 10 ell); the few nearer ones from (s1, rho^2) of the
    // edge
    // the three 1 / (d1 + d2) from one Newton-refined reciprocal of their product
    // (S0 S1 S2 <= (2R)^3 stays far inside FP32 range for any survey geometry)
    const float S0 = da + db, S1 = db + dc, S2 = dc + da;
    const float S01 = S0 * S1;
    const float r012 = rcp_nr(S01 * S2);
    const float r01 = S2 * r012;  // 1 / (S0 S1)
    const float z0 = F.l0 * (S1 * r01);
    const float z1 = F.l1 * (S0 * r01);
    const float z2 = F.l2 * (S01 * r012);
    // in-plane distances of the station from the lines of edges 1 and 2 (gz weights)
    const float dp1 = fmaf(t, F.e1v, py * F.e1u);
    const float dp2 = fmaf(dcx, F.e2v, -dcy * F.e2u);
    float th, lam0, lam1, lam2;  // Omega = 2 th
    if (fmaxf(z0, fmaxf(z1, z2)) <= kZSeries && fabsf(N) <= kSmallAngle * D) {
        // the common case: a station far from the facet compared with its size
        th = atan_small(N * fast_rcp(D));
        lam0 = atanh2_series(z0);
        lam1 = atanh2_series(z1);
        lam2 = atanh2_series(z2);
    } else {
        th = atan2_t(N, D);
        lam0 = atanh2_series(z0);
        lam1 = atanh2_series(z1);
        lam2 = atanh2_series(z2);
        const float s1b = fmaf(t, F.e1u, -py * F.e1v);
        const float s1c = fmaf(dcx, F.e2u, dcy * F.e2v);
        const float rho1 = fmaf(dp1, dp1, pz2);
        const float rho2 = fmaf(dp2, dp2, pz2);
        float num, den;
        if (z0 > kZSeries) {
            edge_ratio(-px, F.l0, da, db, rho0, num, den);
            lam0 = ln_ratio_near(num, den);
        }
        if (z1 > kZSeries) {
            edge_ratio(s1b, F.l1, db, dc, rho1, num, den);
            lam1 = ln_ratio_near(num, den);
        }
        if (z2 > kZSeries) {
            edge_ratio(s1c, F.l2, dc, da, rho2, num, den);
            lam2 = ln_ratio_near(num, den);
        }
        // FP64 for edges whose line passes close to the station while the station
        // projects inside the edge (rho < sqrt(d1 d2)/32): there the FP32 offsets'
        // absolute error would dominate the log.  Only possible near the facet plane
        // (a straddling edge needs rho^2 * 1024 < d1 d2 <= ((len + 2 rho)/2)^2, i.e.
        // rho < len/62, and rho >= |pz|), where the edge's z is above kZSeries.
        if (fabsf(pz) * 60.f < fmaxf(F.l0, fmaxf(F.l1, F.l2))) {
            const bool fb0 = (-px < 0.f) && (t >= 0.f) && (rho0 * 1024.f < da * db);
            const bool fb1 = (s1b < 0.f) && (s1b + F.l1 >= 0.f) && (rho1 * 1024.f < db * dc);
            const bool fb2 = (s1c < 0.f) && (s1c + F.l2 >= 0.f) && (rho2 * 1024.f < dc * da);
            if (fb0 | fb1 | fb2) {
                const double* g = geo64 + 9 * f;  // the record's corners a, b, c in FP64
                if (fb0) lam0 = edge_ln_f64(ox, oy, oz, g[0], g[1], g[2], g[3], g[4], g[5]);
                if (fb1) lam1 = edge_ln_f64(ox, oy, oz, g[3], g[4], g[5], g[6], g[7], g[8]);
                if (fb2) lam2 = edge_ln_f64(ox, oy, oz, g[6], g[7], g[8], g[0], g[1], g[2]);
                nfb += (unsigned)fb0 + (unsigned)fb1 + (unsigned)fb2;
            }
        }
    }
    if (GZ) {
        const float es = fmaf(py, lam0, fmaf(dp1, lam1, dp2 * lam2));
        gs = fmaf(F.K1, fmaf(pz + pz, th, es), gs);
    }
    if (TMI) ts += fmaf(F.W0, th, fmaf(F.W1, lam0, fmaf(F.W2, lam1, F.W3 * lam2)));
}

// Lane = station; the facet records of the block's chunk stream through LDS: the block
// copies kStage records at a time with coalesced vector loads (issued one stage ahead and
// held in registers while the current stage is computed), and every wave reads the
// current record from LDS at a uniform address, so its fields reach the VALU as VGPR
// operands (VALU instructions with a scalar-register operand issue at half the rate on
// gfx950).  Each lane keeps its own Kahan sums: the facet reduction needs no cross-lane
// traffic.
constexpr int kStage = 64;
static_assert(kStage * sizeof(FaceF32) == kBlock * 32, "one 32-byte piece per thread");
// Facet sums: each lane adds the pairs of kGroup consecutive facets in plain FP32 and
// adds each group's sum into its Kahan-compensated FP32 total.  The group sum is small
// against the total, so its own roundings are small against one rounding of the total,
// and the compensation costs four instructions per group instead of per pair.
#ifndef LF_FORWARD_GROUP
#define LF_FORWARD_GROUP 16
#endif
constexpr int kGroup = LF_FORWARD_GROUP;
static_assert(kStage % kGroup == 0, "groups tile a stage");
// stations per lane: each record read from LDS serves this many pairs.  Two halve the
// LDS-to-VGPR traffic of the records (8 uniform 16-byte reads per record per wave) at 142
// VGPRs (3 waves per SIMD instead of 4): measured on the 192x192x64 contact model with
// 65,536 stations, 378 ms with one station per lane, 347 ms with two, 366 ms with three.
#ifndef LF_FORWARD_SPL
#define LF_FORWARD_SPL 2
#endif
constexpr int kSPL = LF_FORWARD_SPL;

template <bool GZ, bool TMI>
__global__ __launch_bounds__(256) LF_FORWARD_OCC void k_forward_f32c(
    const FaceF32* __restrict__ faces, const double* __restrict__ geo64, int64_t n_f,
    int64_t faces_per_chunk, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, int64_t n_s, int n_tiles, double* __restrict__ part_gz,
    double* __restrict__ part_tmi, unsigned long long* __restrict__ fb_count)
{
    __shared__ FaceF32 stage[2][kStage];
    int tile, chunk;
    block_coords(n_tiles, tile, chunk);
    double ox[kSPL], oy[kSPL], oz[kSPL];
#pragma unroll
    for (int k = 0; k < kSPL; ++k) {
        const int64_t s = ((int64_t)tile * kSPL + k) * kBlock + threadIdx.x;
        const int64_t si = s < n_s ? s : n_s - 1;
        ox[k] = sx[si];
        oy[k] = sy[si];
        oz[k] = sz[si];
    }
    const int64_t f0 = (int64_t)chunk * faces_per_chunk;
    const int64_t f1 = min(n_f, f0 + faces_per_chunk);

    // thread t copies bytes [32 (t % 4), +32) of record t / 4 of a stage
    const int rec = threadIdx.x >> 2;
    const int piece = (threadIdx.x & 3) * 2;  // in float4 units
    float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
    auto fetch = [&](int64_t fs) {
        const int64_t fr = fs + rec;
        if (fr < f1) {
            const float4* src = reinterpret_cast<const float4*>(faces + fr) + piece;
            q0 = src[0];
            q1 = src[1];
        }
    };
    auto park = [&](int b) {
        float4* dst = reinterpret_cast<float4*>(&stage[b][rec]) + piece;
        dst[0] = q0;
        dst[1] = q1;
    };

    // per station: Kahan-compensated FP32 sums (s, c) of the facet-group sums
    float gs[kSPL], gc[kSPL], ts[kSPL], tc[kSPL];
#pragma unroll
    for (int k = 0; k < kSPL; ++k) gs[k] = gc[k] = ts[k] = tc[k] = 0.f;
    unsigned nfb = 0;
    if (f0 < f1) {
        fetch(f0);
        park(0);
    }
    __syncthreads();
    int b = 0;
    for (int64_t fs = f0; fs < f1; fs += kStage) {
        const bool more = fs + kStage < f1;
        if (more) fetch(fs + kStage);
        const int nst = (int)min((int64_t)kStage, f1 - fs);
        for (int j0 = 0; j0 < nst; j0 += kGroup) {
            const int j1 = min(nst, j0 + kGroup);
            float g[kSPL], m[kSPL];
#pragma unroll
            for (int k = 0; k < kSPL; ++k) g[k] = m[k] = 0.f;
            for (int j = j0; j < j1; ++j) {
                const FaceF32 F = stage[b][j];
#pragma unroll
                for (int k = 0; k < kSPL; ++k)
                    forward_pair<GZ, TMI>(F, geo64, fs + j, ox[k], oy[k], oz[k], g[k], m[k], nfb);
            }
#pragma unroll
            for (int k = 0; k < kSPL; ++k) {
                if (GZ) kahan_add(gs[k], gc[k], g[k]);
                if (TMI) kahan_add(ts[k], tc[k], m[k]);
            }
        }
        // the other buffer was last read before the previous barrier
        if (more) park(b ^ 1);
        __syncthreads();
        b ^= 1;
    }
#pragma unroll
    for (int k = 0; k < kSPL; ++k) {
        const int64_t s = ((int64_t)tile * kSPL + k) * kBlock + threadIdx.x;
        if (s < n_s) {
            if (GZ) part_gz[(int64_t)chunk * n_s + s] = (double)gs[k] - (double)gc[k];
            if (TMI) part_tmi[(int64_t)chunk * n_s + s] = (double)ts[k] - (double)tc[k];
        }
    }
    if (nfb) atomicAdd(fb_count, (unsigned long long)nfb);
}

// ------------------------------------------------------------------ FP64 forward

template <bool GZ, bool TMI>
__global__ __launch_bounds__(256) void k_forward_f64(
    const FaceF64* __restrict__ faces, int64_t n_f, int64_t faces_per_chunk,
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, int64_t n_s, int n_tiles, double* __restrict__ part_gz,
    double* __restrict__ part_tmi)
{
    int tile, chunk;
    block_coords(n_tiles, tile, chunk);
    const int64_t s = (int64_t)tile * kBlock + threadIdx.x;
    const bool active = s < n_s;
    const int64_t si = active ? s : n_s - 1;
    const double o[3] = {sx[si], sy[si], sz[si]};
    const int64_t f0 = (int64_t)chunk * faces_per_chunk;
    const int64_t f1 = min(n_f, f0 + faces_per_chunk);
    double gsum = 0.0, tsum = 0.0;
    for (int64_t f = f0; f < f1; ++f) {
        const FaceF64& F = faces[f];
        double r[3][3], d[3];
        for (int j = 0; j < 3; ++j) {
            r[0][j] = F.a[j] - o[j];
            r[1][j] = F.b[j] - o[j];
            r[2][j] = F.c[j] - o[j];
        }
        for (int k = 0; k < 3; ++k)
            d[k] = sqrt(r[k][0] * r[k][0] + r[k][1] * r[k][1] + r[k][2] * r[k][2]);
        // solidAngle, geometry.cpp:10-30
        double omega = 0.0;
        if (!(d[0] < kEps || d[1] < kEps || d[2] < kEps)) {
            const double cx = r[1][1] * r[2][2] - r[1][2] * r[2][1];
            const double cy = r[1][2] * r[2][0] - r[1][0] * r[2][2];
            const double cz = r[1][0] * r[2][1] - r[1][1] * r[2][0];
            const double num = r[0][0] * cx + r[0][1] * cy + r[0][2] * cz;
            const double dab = r[0][0] * r[1][0] + r[0][1] * r[1][1] + r[0][2] * r[1][2];
            const double dac = r[0][0] * r[2][0] + r[0][1] * r[2][1] + r[0][2] * r[2][2];
            const double dbc = r[1][0] * r[2][0] + r[1][1] * r[2][1] + r[1][2] * r[2][2];
            const double den = d[0] * d[1] * d[2] + dab * d[2] + dac * d[1] + dbc * d[0];
            omega = 2.0 * atan2(num, den);
        }
        const double h = r[0][0] * F.n[0] + r[0][1] * F.n[1] + r[0][2] * F.n[2];
        double dsum = 0.0, tm = F.wom * omega;
        for (int k = 0; k < 3; ++k) {
            const int k2 = (k + 1) % 3;
            // lineIntegralTerm, geometry.cpp:32-52
            const double s1 = r[k][0] * F.eh[k][0] + r[k][1] * F.eh[k][1] + r[k][2] * F.eh[k][2];
            const double s2 = r[k2][0] * F.eh[k][0] + r[k2][1] * F.eh[k][1] + r[k2][2] * F.eh[k][2];
            const double num = s2 + d[k2];
            const double den = s1 + d[k];
            double L = 0.0;
            if (!(fabs(den) < kEps) && num > 0.0 && den > 0.0) L = log(num / den);
            // (o - p1).(n x e_hat)
            const double dperp =
                -(r[k][0] * F.perp[k][0] + r[k][1] * F.perp[k][1] + r[k][2] * F.perp[k][2]);
            dsum += dperp * L;
            tm += F.we[k] * L;
        }
        if (GZ) gsum += F.kz * (dsum - h * omega);
        if (TMI) tsum += tm;
    }
    if (active) {
        if (GZ) part_gz[(int64_t)chunk * n_s + s] = gsum;
        if (TMI) part_tmi[(int64_t)chunk * n_s + s] = tsum;
    }
}

// ------------------------------------------------------------------------ reduce

__global__ __launch_bounds__(256) void k_reduce_chunks(const double* __restrict__ part,
                                                       int n_chunks, int64_t n_s,
                                                       double* __restrict__ out)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_s) return;
    double acc = 0.0;
    for (int c = 0; c < n_chunks; ++c) acc += part[(int64_t)c * n_s + s];
    out[s] = acc;
}

__global__ __launch_bounds__(256) void k_split_xyz(const double* __restrict__ xyz, int64_t n,
                                                   double* __restrict__ x,
                                                   double* __restrict__ y,
                                                   double* __restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = xyz[3 * i];
    y[i] = xyz[3 * i + 1];
    z[i] = xyz[3 * i + 2];
}

}  // namespace

// =========================================================================== context

static thread_local std::string g_create_error;

struct lf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n_v = 0, n_f = 0, n_s = 0;
    int32_t n_u = 0;
    int precision = LF_FP64, tmi_mode = LF_TMI_REFERENCE, want = 3;
    double fdir[3] = {0, 0, 1};
    std::string err;
    bool faces_dirty = true;

    // model
    double* d_verts = nullptr;  // n_v*3
    uint32_t* d_tri = nullptr;
    int32_t* d_front = nullptr;
    int32_t* d_back = nullptr;
    double* d_density = nullptr;
    double* d_mag = nullptr;
    double* d_sx = nullptr;
    double* d_sy = nullptr;
    double* d_sz = nullptr;
    // facet records
    FaceF32* d_f32 = nullptr;
    FaceF64* d_f64 = nullptr;
    FaceFrame<double>* d_fr64 = nullptr;
    double* d_geo64 = nullptr;
    // forward work
    int n_chunks = 1, n_tiles = 1;
    int64_t faces_per_chunk = 0;
    double* d_part_gz = nullptr;
    double* d_part_tmi = nullptr;
    double* d_out_gz = nullptr;
    double* d_out_tmi = nullptr;
    unsigned long long* d_fb = nullptr;
    // adjoint work
    AdjointWork adj;
    // column-norm work (lf_colnorm2): the model-sized buffers are allocated on the first
    // call, the per-parameter ones grow to the largest parameter map seen
    struct ColnormWork {
        void* dT = nullptr;  // per-facet derivative matrix restricted to one axis
        StationW* dst = nullptr;
        float *Gg = nullptr, *Gt = nullptr, *Dg = nullptr, *Dt = nullptr;
        int64_t* doff = nullptr;
        int32_t* dvert = nullptr;
        double *dw = nullptr, *og = nullptr, *ot = nullptr;
        int64_t cap_dof = -1, cap_ent = -1;
    } cn;
    // Jacobian-vector products (lf_jvp): the facets' constant tangents and a vertex
    // tangent copied from the host, allocated on the first call
    void* d_dq = nullptr;
    double* d_vtan = nullptr;
    // timing
    bool timing = false;
    struct Ev { int cls; hipEvent_t a, b; };
    std::vector<Ev> events;
    double ms[8] = {0};
    int64_t launches[8] = {0};
};

#define LF_TRY(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);            \
            return LF_EHIP;                                                            \
        }                                                                              \
    } while (0)

namespace {

int set_err(lf_ctx* c, const std::string& m, int code)
{
    c->err = m;
    return code;
}

template <class T>
hipError_t dalloc(T** p, size_t count)
{
    return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

struct TimedLaunch {
    lf_ctx* c;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(lf_ctx* c_, int cls_) : c(c_), cls(cls_)
    {
        if (c->timing) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, c->stream);
        }
    }
    ~TimedLaunch()
    {
        if (c->timing) {
            hipEventRecord(b, c->stream);
            c->events.push_back({cls, a, b});
        }
    }
};

// Station tiles x facet chunks of the forward grid: about 128 blocks per CU, which lets
// the dispatcher balance blocks that near pairs slow down and keeps the last, partly
// filled round of resident blocks short; at least kFwdMinFacets facets per chunk, so that
// a block's set-up (its stations, the first record stage, the partial-sum writes) stays
// small against its pairs; at most max_chunks chunks (the per-chunk partial sums take
// 16 bytes per station and chunk).  Measured on the 192x192x64 contact model with 65,536
// stations (128 station tiles), FP32C forward per launch: 24 facet chunks 353 ms, 32:
// 349 ms, 96: 339 ms, 256: 336 ms, 512: 334 ms (the chunk reduction 0.03 -> 0.4 ms).
constexpr int64_t kFwdBlocksPerCU = 128;
constexpr int64_t kFwdMinFacets = 512;
void chunk_plan(int64_t n_s, int64_t n_f, int precision, int dev_cus, int64_t max_chunks,
                int& n_tiles, int64_t& faces_per_chunk, int& n_chunks)
{
    const int64_t tile_st = (int64_t)kBlock * (precision == LF_FP32C ? kSPL : 1);
    n_tiles = (int)std::max<int64_t>(1, (n_s + tile_st - 1) / tile_st);
    const int64_t target = (int64_t)std::max(1, dev_cus) * kFwdBlocksPerCU;
    int64_t chunks = std::max<int64_t>(1, (target + n_tiles - 1) / n_tiles);
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, n_f / kFwdMinFacets));
    chunks = std::max<int64_t>(1, std::min<int64_t>(chunks, max_chunks));
    faces_per_chunk = std::max<int64_t>(1, (n_f + chunks - 1) / chunks);
    n_chunks = (int)std::max<int64_t>(1, (n_f + faces_per_chunk - 1) / faces_per_chunk);
}

int choose_chunks(lf_ctx* c)
{
    int dev_cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        dev_cus = prop.multiProcessorCount;
    chunk_plan(c->n_s, c->n_f, c->precision, dev_cus, INT64_MAX, c->n_tiles,
               c->faces_per_chunk, c->n_chunks);
    return LF_OK;
}

int ensure_faces(lf_ctx* c)
{
    if (!c->faces_dirty) return LF_OK;
    if (c->n_f > 0) {
        TimedLaunch tl(c, 0);
        const int64_t blocks = (c->n_f + 127) / 128;
        hipLaunchKernelGGL(k_face_setup, dim3((unsigned)blocks), dim3(128), 0, c->stream,
                           c->d_verts, c->d_tri, c->d_front, c->d_back, c->d_density, c->d_mag,
                           c->fdir[0], c->fdir[1], c->fdir[2], c->tmi_mode, c->n_f, c->n_v,
                           (int)c->n_u, c->d_f32, c->d_f64, c->d_fr64, c->d_geo64);
        LF_TRY(c, hipGetLastError());
    }
    c->faces_dirty = false;
    return LF_OK;
}

template <bool GZ, bool TMI>
void launch_forward(lf_ctx* c)
{
    const dim3 grid((unsigned)((int64_t)c->n_tiles * c->n_chunks));
    if (c->precision == LF_FP32C)
        hipLaunchKernelGGL((k_forward_f32c<GZ, TMI>), grid, dim3(kBlock), 0, c->stream, c->d_f32,
                           c->d_geo64, c->n_f, c->faces_per_chunk, c->d_sx, c->d_sy, c->d_sz,
                           c->n_s, c->n_tiles, c->d_part_gz, c->d_part_tmi, c->d_fb);
    else
        hipLaunchKernelGGL((k_forward_f64<GZ, TMI>), grid, dim3(kBlock), 0, c->stream, c->d_f64,
                           c->n_f, c->faces_per_chunk, c->d_sx, c->d_sy, c->d_sz, c->n_s,
                           c->n_tiles, c->d_part_gz, c->d_part_tmi);
}

int forward_device(lf_ctx* c, double* d_gz, double* d_tmi)
{
    int rc = ensure_faces(c);
    if (rc) return rc;
    const bool gz = d_gz && (c->want & LF_WANT_GZ);
    const bool tm = d_tmi && (c->want & LF_WANT_TMI);
    if (c->n_s == 0) return LF_OK;
    const unsigned sblocks = (unsigned)((c->n_s + 255) / 256);
    if (c->n_f == 0) {
        if (gz) hipLaunchKernelGGL(k_fill_zero, dim3(sblocks), dim3(256), 0, c->stream, d_gz, c->n_s);
        if (tm) hipLaunchKernelGGL(k_fill_zero, dim3(sblocks), dim3(256), 0, c->stream, d_tmi, c->n_s);
        LF_TRY(c, hipGetLastError());
        return LF_OK;
    }
    if (!gz && !tm) return LF_OK;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, c->stream, c->d_fb);
    {
        TimedLaunch tl(c, 1);
        if (gz && tm) launch_forward<true, true>(c);
        else if (gz) launch_forward<true, false>(c);
        else launch_forward<false, true>(c);
        LF_TRY(c, hipGetLastError());
    }
    {
        TimedLaunch tl(c, 2);
        const unsigned blocks = (unsigned)((c->n_s + 255) / 256);
        if (gz)
            hipLaunchKernelGGL(k_reduce_chunks, dim3(blocks), dim3(256), 0, c->stream,
                               c->d_part_gz, c->n_chunks, c->n_s, d_gz);
        if (tm)
            hipLaunchKernelGGL(k_reduce_chunks, dim3(blocks), dim3(256), 0, c->stream,
                               c->d_part_tmi, c->n_chunks, c->n_s, d_tmi);
        LF_TRY(c, hipGetLastError());
    }
    return LF_OK;
}

int check_ctx(const lf_ctx* c) { return c ? LF_OK : LF_EINVAL; }

}  // namespace

// ============================================================================== ABI

extern "C" {

const char* lf_last_error(const lf_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

void lf_destroy(lf_ctx* c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& e : c->events) {
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    void* bufs[] = {c->d_verts, c->d_tri,  c->d_front,    c->d_back,       c->d_density,
                    c->d_mag,   c->d_sx,   c->d_sy,       c->d_sz,         c->d_f32,
                    c->d_f64,   c->d_fr64, c->d_geo64, c->d_part_gz, c->d_part_tmi, c->d_out_gz,
                    c->d_out_tmi, c->d_fb};
    for (void* p : bufs)
        if (p) hipFree(p);
    void* cnb[] = {c->cn.dT, c->cn.dst, c->cn.Gg,  c->cn.Gt, c->cn.Dg, c->cn.Dt,
                   c->cn.doff, c->cn.dvert, c->cn.dw, c->cn.og, c->cn.ot};
    for (void* p : cnb)
        if (p) hipFree(p);
    if (c->d_dq) hipFree(c->d_dq);
    if (c->d_vtan) hipFree(c->d_vtan);
    adjoint_free(c->adj);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int lf_create(lf_ctx** out, const lf_desc* d)
{
    g_create_error.clear();
    if (!out || !d) {
        g_create_error = "lf_create: null argument";
        return LF_EINVAL;
    }
    *out = nullptr;
    if (d->n_vertices < 0 || d->n_faces < 0 || d->n_stations < 0 || d->n_units <= 0 ||
        (d->n_faces > 0 && (!d->vertices || !d->faces || !d->front || !d->back)) ||
        (d->n_stations > 0 && !d->stations) || !d->density || !d->magnetization ||
        (d->precision != LF_FP64 && d->precision != LF_FP32C) ||
        (d->tmi_mode != LF_TMI_REFERENCE && d->tmi_mode != LF_TMI_TEXTBOOK) ||
        (d->want & ~3) != 0 || d->n_vertices > 0xffffffffLL) {
        g_create_error = "lf_create: invalid descriptor";
        return LF_EINVAL;
    }
    for (int64_t f = 0; f < d->n_faces; ++f) {
        for (int k = 0; k < 3; ++k)
            if ((int64_t)d->faces[3 * f + k] >= d->n_vertices) {
                g_create_error = "lf_create: face " + std::to_string(f) + " references vertex " +
                                 std::to_string(d->faces[3 * f + k]) + " >= n_vertices";
                return LF_EINVAL;
            }
        if (d->front[f] < 0 || d->front[f] >= d->n_units || d->back[f] < -1 ||
            d->back[f] >= d->n_units) {
            g_create_error = "lf_create: face " + std::to_string(f) + " has an invalid unit id";
            return LF_EINVAL;
        }
    }
    lf_ctx* c = new lf_ctx();
    c->device = d->device;
    c->n_v = d->n_vertices;
    c->n_f = d->n_faces;
    c->n_s = d->n_stations;
    c->n_u = d->n_units;
    c->precision = d->precision;
    c->tmi_mode = d->tmi_mode;
    c->want = d->want ? d->want : 3;
    for (int j = 0; j < 3; ++j) c->fdir[j] = d->field_dir[j];

    auto fail = [&](int rc) {
        g_create_error = c->err;
        lf_destroy(c);
        return rc;
    };
#define CR(expr)                                                                          \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            c->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
            return fail(LF_EHIP);                                                         \
        }                                                                                 \
    } while (0)
    CR(hipSetDevice(c->device));
    CR(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CR(dalloc(&c->d_verts, 3 * c->n_v));
    CR(dalloc(&c->d_tri, 3 * c->n_f));
    CR(dalloc(&c->d_front, c->n_f));
    CR(dalloc(&c->d_back, c->n_f));
    CR(dalloc(&c->d_density, c->n_u));
    CR(dalloc(&c->d_mag, 3 * c->n_u));
    CR(dalloc(&c->d_sx, c->n_s));
    CR(dalloc(&c->d_sy, c->n_s));
    CR(dalloc(&c->d_sz, c->n_s));
    if (c->precision == LF_FP32C) {
        CR(dalloc(&c->d_f32, c->n_f));
        CR(dalloc(&c->d_geo64, 9 * c->n_f));
    } else {
        CR(dalloc(&c->d_f64, c->n_f));
        CR(dalloc(&c->d_fr64, c->n_f));
    }
    choose_chunks(c);
    CR(dalloc(&c->d_part_gz, (size_t)c->n_chunks * c->n_s));
    CR(dalloc(&c->d_part_tmi, (size_t)c->n_chunks * c->n_s));
    CR(dalloc(&c->d_out_gz, c->n_s));
    CR(dalloc(&c->d_out_tmi, c->n_s));
    CR(dalloc(&c->d_fb, 1));
    if (c->n_v) CR(hipMemcpy(c->d_verts, d->vertices, sizeof(double) * 3 * c->n_v, hipMemcpyHostToDevice));
    if (c->n_f) {
        CR(hipMemcpy(c->d_tri, d->faces, sizeof(uint32_t) * 3 * c->n_f, hipMemcpyHostToDevice));
        CR(hipMemcpy(c->d_front, d->front, sizeof(int32_t) * c->n_f, hipMemcpyHostToDevice));
        CR(hipMemcpy(c->d_back, d->back, sizeof(int32_t) * c->n_f, hipMemcpyHostToDevice));
    }
    CR(hipMemcpy(c->d_density, d->density, sizeof(double) * c->n_u, hipMemcpyHostToDevice));
    CR(hipMemcpy(c->d_mag, d->magnetization, sizeof(double) * 3 * c->n_u, hipMemcpyHostToDevice));
    if (c->n_s) {
        std::vector<double> tmp(3 * c->n_s);
        for (int64_t i = 0; i < c->n_s; ++i) {
            tmp[i] = d->stations[3 * i];
            tmp[c->n_s + i] = d->stations[3 * i + 1];
            tmp[2 * c->n_s + i] = d->stations[3 * i + 2];
        }
        CR(hipMemcpy(c->d_sx, tmp.data(), sizeof(double) * c->n_s, hipMemcpyHostToDevice));
        CR(hipMemcpy(c->d_sy, tmp.data() + c->n_s, sizeof(double) * c->n_s, hipMemcpyHostToDevice));
        CR(hipMemcpy(c->d_sz, tmp.data() + 2 * c->n_s, sizeof(double) * c->n_s, hipMemcpyHostToDevice));
    }
    {
        int rc = adjoint_init(c->adj, c->n_v, c->n_f, c->n_s, d->faces, c->precision);
        if (rc) {
            c->err = "adjoint workspace: " + std::string(hipGetErrorString(hipGetLastError()));
            return fail(LF_EHIP);
        }
    }
#undef CR
    c->faces_dirty = true;
    *out = c;
    return LF_OK;
}

int lf_set_vertices(lf_ctx* c, const double* xyz)
{
    if (check_ctx(c)) return LF_EINVAL;
    if (!xyz && c->n_v) return set_err(c, "lf_set_vertices: null vertices", LF_EINVAL);
    LF_TRY(c, hipSetDevice(c->device));
    if (c->n_v)
        LF_TRY(c, hipMemcpyAsync(c->d_verts, xyz, sizeof(double) * 3 * c->n_v,
                                 hipMemcpyHostToDevice, c->stream));
    LF_TRY(c, hipStreamSynchronize(c->stream));
    c->faces_dirty = true;
    return LF_OK;
}

int lf_set_vertices_device(lf_ctx* c, const double* d_xyz)
{
    if (check_ctx(c)) return LF_EINVAL;
    if (!d_xyz && c->n_v) return set_err(c, "lf_set_vertices_device: null vertices", LF_EINVAL);
    LF_TRY(c, hipSetDevice(c->device));
    if (c->n_v)
        LF_TRY(c, hipMemcpyAsync(c->d_verts, d_xyz, sizeof(double) * 3 * c->n_v,
                                 hipMemcpyDeviceToDevice, c->stream));
    c->faces_dirty = true;
    return LF_OK;
}

int lf_set_properties(lf_ctx* c, const double* density, const double* mag)
{
    if (check_ctx(c)) return LF_EINVAL;
    LF_TRY(c, hipSetDevice(c->device));
    if (density)
        LF_TRY(c, hipMemcpyAsync(c->d_density, density, sizeof(double) * c->n_u,
                                 hipMemcpyHostToDevice, c->stream));
    if (mag)
        LF_TRY(c, hipMemcpyAsync(c->d_mag, mag, sizeof(double) * 3 * c->n_u,
                                 hipMemcpyHostToDevice, c->stream));
    LF_TRY(c, hipStreamSynchronize(c->stream));
    c->faces_dirty = true;
    return LF_OK;
}

int lf_forward_device(lf_ctx* c, double* d_gz, double* d_tmi)
{
    if (check_ctx(c)) return LF_EINVAL;
    LF_TRY(c, hipSetDevice(c->device));
    return forward_device(c, d_gz, d_tmi);
}

int lf_forward(lf_ctx* c, double* gz, double* tmi)
{
    if (check_ctx(c)) return LF_EINVAL;
    LF_TRY(c, hipSetDevice(c->device));
    int rc = forward_device(c, gz ? c->d_out_gz : nullptr, tmi ? c->d_out_tmi : nullptr);
    if (rc) return rc;
    if (gz && c->n_s)
        LF_TRY(c, hipMemcpyAsync(gz, c->d_out_gz, sizeof(double) * c->n_s,
                                 hipMemcpyDeviceToHost, c->stream));
    if (tmi && c->n_s)
        LF_TRY(c, hipMemcpyAsync(tmi, c->d_out_tmi, sizeof(double) * c->n_s,
                                 hipMemcpyDeviceToHost, c->stream));
    LF_TRY(c, hipStreamSynchronize(c->stream));
    return LF_OK;
}

int lf_adjoint_device(lf_ctx* c, const double* d_wg, const double* d_wt, double* d_grad)
{
    if (check_ctx(c)) return LF_EINVAL;
    if (!d_grad) return set_err(c, "lf_adjoint_device: null gradient", LF_EINVAL);
    LF_TRY(c, hipSetDevice(c->device));
    int rc = ensure_faces(c);
    if (rc) return rc;
    const double* wg = (c->want & LF_WANT_GZ) ? d_wg : nullptr;
    const double* wt = (c->want & LF_WANT_TMI) ? d_wt : nullptr;
    AdjointLaunch L{c->stream, c->n_v, c->n_f, c->n_s, c->precision, c->d_f32, c->d_fr64,
                    c->d_sx, c->d_sy, c->d_sz, wg, wt, d_grad};
    hipError_t e;
    {
        TimedLaunch t3(c, 3);
        e = adjoint_pairs(c->adj, L);
    }
    if (e != hipSuccess) return set_err(c, std::string("adjoint pairs: ") + hipGetErrorString(e), LF_EHIP);
    {
        TimedLaunch t4(c, 4);
        e = adjoint_finalize(c->adj, L, c->d_verts, c->d_tri, c->d_front, c->d_back,
                             c->d_density, c->d_mag, c->fdir, c->tmi_mode);
    }
    if (e != hipSuccess) return set_err(c, std::string("adjoint finalize: ") + hipGetErrorString(e), LF_EHIP);
    {
        TimedLaunch t5(c, 5);
        e = adjoint_gather(c->adj, L);
    }
    if (e != hipSuccess) return set_err(c, std::string("adjoint gather: ") + hipGetErrorString(e), LF_EHIP);
    return LF_OK;
}

int lf_adjoint(lf_ctx* c, const double* w_gz, const double* w_tmi, double* grad)
{
    if (check_ctx(c)) return LF_EINVAL;
    if (!grad) return set_err(c, "lf_adjoint: null gradient", LF_EINVAL);
    LF_TRY(c, hipSetDevice(c->device));
    double* dwg = c->adj.d_wg;
    double* dwt = c->adj.d_wt;
    if (w_gz && c->n_s)
        LF_TRY(c, hipMemcpyAsync(dwg, w_gz, sizeof(double) * c->n_s, hipMemcpyHostToDevice, c->stream));
    if (w_tmi && c->n_s)
        LF_TRY(c, hipMemcpyAsync(dwt, w_tmi, sizeof(double) * c->n_s, hipMemcpyHostToDevice, c->stream));
    int rc = lf_adjoint_device(c, w_gz ? dwg : nullptr, w_tmi ? dwt : nullptr, c->adj.d_grad);
    if (rc) return rc;
    if (c->n_v)
        LF_TRY(c, hipMemcpyAsync(grad, c->adj.d_grad, sizeof(double) * 3 * c->n_v,
                                 hipMemcpyDeviceToHost, c->stream));
    LF_TRY(c, hipStreamSynchronize(c->stream));
    return LF_OK;
}

int lf_colnorm2(lf_ctx* c, int axis, int64_t n_dof, const int64_t* dof_off,
                const int32_t* dof_vertex, const double* dof_weight, double* out_gz,
                double* out_tmi)
{
    if (check_ctx(c)) return LF_EINVAL;
    if (axis < 0 || axis > 2 || n_dof < 0 || (n_dof > 0 && (!dof_off || !dof_vertex || !dof_weight)))
        return set_err(c, "lf_colnorm2: invalid arguments", LF_EINVAL);
    if (n_dof > 0 && dof_off[0] != 0) return set_err(c, "lf_colnorm2: dof_off[0] != 0", LF_EINVAL);
    for (int64_t j = 0; j < n_dof; ++j)
        if (dof_off[j + 1] < dof_off[j]) return set_err(c, "lf_colnorm2: dof_off not monotone", LF_EINVAL);
    const int64_t n_ent = n_dof > 0 ? dof_off[n_dof] : 0;
    for (int64_t e = 0; e < n_ent; ++e)
        if (dof_vertex[e] < 0 || dof_vertex[e] >= c->n_v)
            return set_err(c, "lf_colnorm2: DOF entry references a vertex out of range", LF_EINVAL);
    LF_TRY(c, hipSetDevice(c->device));
    int rc = ensure_faces(c);
    if (rc) return rc;
    const int kBlk = 64;
    const size_t tsz = c->precision == LF_FP32C ? sizeof(float) : sizeof(double);
    auto& W = c->cn;
    hipError_t e;
    auto fail = [&](hipError_t err, const char* what) {
        return set_err(c, std::string("lf_colnorm2 ") + what + ": " + hipGetErrorString(err), LF_EHIP);
    };
#define CN(expr, what) do { e = (expr); if (e != hipSuccess) return fail(e, what); } while (0)
    if (!(W.dT && W.dst && W.Gg && W.Gt && W.Dg && W.Dt)) {
        // first call, or a first call that failed part-way
        for (void* q : {W.dT, (void*)W.dst, (void*)W.Gg, (void*)W.Gt, (void*)W.Dg, (void*)W.Dt})
            if (q) hipFree(q);
        W.dT = nullptr;
        W.dst = nullptr;
        W.Gg = W.Gt = W.Dg = W.Dt = nullptr;
        const int64_t nf = std::max<int64_t>(c->n_f, 1), nv = std::max<int64_t>(c->n_v, 1);
        CN(hipMalloc(&W.dT, tsz * kSeeds * 3 * nf), "alloc");
        CN(hipMalloc((void**)&W.dst, sizeof(StationW) * kBlk), "alloc");
        CN(hipMalloc((void**)&W.Gg, sizeof(float) * kBlk * 3 * nf), "alloc");
        CN(hipMalloc((void**)&W.Gt, sizeof(float) * kBlk * 3 * nf), "alloc");
        CN(hipMalloc((void**)&W.Dg, sizeof(float) * kBlk * nv), "alloc");
        CN(hipMalloc((void**)&W.Dt, sizeof(float) * kBlk * nv), "alloc");
    }
    if (n_dof > W.cap_dof) {
        // grow; the previous call ended with a stream synchronization, so nothing reads
        // the old buffers
        CN(hipStreamSynchronize(c->stream), "sync");
        for (void* q : {(void*)W.doff, (void*)W.og, (void*)W.ot})
            if (q) hipFree(q);
        W.doff = nullptr;
        W.og = W.ot = nullptr;
        W.cap_dof = -1;
        CN(hipMalloc((void**)&W.doff, sizeof(int64_t) * (n_dof + 1)), "alloc");
        CN(hipMalloc((void**)&W.og, sizeof(double) * std::max<int64_t>(n_dof, 1)), "alloc");
        CN(hipMalloc((void**)&W.ot, sizeof(double) * std::max<int64_t>(n_dof, 1)), "alloc");
        W.cap_dof = n_dof;
    }
    if (n_ent > W.cap_ent) {
        CN(hipStreamSynchronize(c->stream), "sync");
        if (W.dvert) hipFree(W.dvert);
        if (W.dw) hipFree(W.dw);
        W.dvert = nullptr;
        W.dw = nullptr;
        W.cap_ent = -1;
        CN(hipMalloc((void**)&W.dvert, sizeof(int32_t) * std::max<int64_t>(n_ent, 1)), "alloc");
        CN(hipMalloc((void**)&W.dw, sizeof(double) * std::max<int64_t>(n_ent, 1)), "alloc");
        W.cap_ent = n_ent;
    }
    void* dT = W.dT;
    StationW* dst = W.dst;
    float *Gg = W.Gg, *Gt = W.Gt, *Dg = W.Dg, *Dt = W.Dt;
    int64_t* doff = W.doff;
    int32_t* dvert = W.dvert;
    double *dw = W.dw, *og = W.og, *ot = W.ot;
    if (n_dof > 0) {
        CN(hipMemcpyAsync(doff, dof_off, sizeof(int64_t) * (n_dof + 1), hipMemcpyHostToDevice, c->stream), "copy");
        if (n_ent) {
            CN(hipMemcpyAsync(dvert, dof_vertex, sizeof(int32_t) * n_ent, hipMemcpyHostToDevice, c->stream), "copy");
            CN(hipMemcpyAsync(dw, dof_weight, sizeof(double) * n_ent, hipMemcpyHostToDevice, c->stream), "copy");
        }
        const unsigned jb = (unsigned)((n_dof + 255) / 256);
        hipLaunchKernelGGL(k_fill_zero, dim3(jb), dim3(256), 0, c->stream, og, n_dof);
        hipLaunchKernelGGL(k_fill_zero, dim3(jb), dim3(256), 0, c->stream, ot, n_dof);
        if (c->n_f > 0 && c->n_s > 0) {
            const unsigned fb = (unsigned)((c->n_f + 127) / 128);
            if (c->precision == LF_FP32C)
                hipLaunchKernelGGL(k_facet_T_axis<float>, dim3(fb), dim3(128), 0, c->stream, c->n_f,
                                   c->d_verts, c->d_tri, c->d_front, c->d_back, c->d_density,
                                   c->d_mag, c->fdir[0], c->fdir[1], c->fdir[2], c->tmi_mode, axis,
                                   (float*)dT);
            else
                hipLaunchKernelGGL(k_facet_T_axis<double>, dim3(fb), dim3(128), 0, c->stream, c->n_f,
                                   c->d_verts, c->d_tri, c->d_front, c->d_back, c->d_density,
                                   c->d_mag, c->fdir[0], c->fdir[1], c->fdir[2], c->tmi_mode, axis,
                                   (double*)dT);
            CN(hipGetLastError(), "facet T");
            for (int64_t b0 = 0; b0 < c->n_s; b0 += kBlk) {
                const int nb = (int)std::min<int64_t>(kBlk, c->n_s - b0);
                hipLaunchKernelGGL(k_pack_stations, dim3(1), dim3(256), 0, c->stream, c->d_sx + b0,
                                   c->d_sy + b0, c->d_sz + b0, c->adj.d_zero, c->adj.d_zero,
                                   (int64_t)nb, dst);
                const dim3 gb((unsigned)((c->n_f + 255) / 256),
                              (unsigned)((nb + kStationGroup - 1) / kStationGroup));
                if (c->precision == LF_FP32C)
                    hipLaunchKernelGGL(k_station_facet_grad<float>, gb, dim3(256), 0, c->stream,
                                       c->d_f32, (const float*)dT, c->n_f, dst, nb, Gg, Gt);
                else
                    hipLaunchKernelGGL(k_station_facet_grad<double>, gb, dim3(256), 0, c->stream,
                                       c->d_fr64, (const double*)dT, c->n_f, dst, nb, Gg, Gt);
                const dim3 vg((unsigned)((c->n_v + 255) / 256), (unsigned)nb);
                hipLaunchKernelGGL(k_gather_block, vg, dim3(256), 0, c->stream, Gg, c->adj.d_csr_off,
                                   c->adj.d_csr_ent, c->n_v, c->n_f, Dg);
                hipLaunchKernelGGL(k_gather_block, vg, dim3(256), 0, c->stream, Gt, c->adj.d_csr_off,
                                   c->adj.d_csr_ent, c->n_v, c->n_f, Dt);
                const unsigned cb = (unsigned)((n_dof + 3) / 4);  // one wave per DOF
                hipLaunchKernelGGL(k_dof_colnorm, dim3(cb), dim3(256), 0, c->stream, Dg, nb, c->n_v,
                                   n_dof, doff, dvert, dw, og);
                hipLaunchKernelGGL(k_dof_colnorm, dim3(cb), dim3(256), 0, c->stream, Dt, nb, c->n_v,
                                   n_dof, doff, dvert, dw, ot);
                CN(hipGetLastError(), "column-norm kernels");
            }
        }
        if (out_gz) CN(hipMemcpyAsync(out_gz, og, sizeof(double) * n_dof, hipMemcpyDeviceToHost, c->stream), "copy");
        if (out_tmi) CN(hipMemcpyAsync(out_tmi, ot, sizeof(double) * n_dof, hipMemcpyDeviceToHost, c->stream), "copy");
    }
    CN(hipStreamSynchronize(c->stream), "sync");
#undef CN
    return LF_OK;
}

int lf_jvp_device(lf_ctx* c, const double* d_vtan, double* d_gz, double* d_tmi)
{
    if (check_ctx(c)) return LF_EINVAL;
    if (!d_vtan && c->n_v) return set_err(c, "lf_jvp_device: null tangent", LF_EINVAL);
    LF_TRY(c, hipSetDevice(c->device));
    int rc = ensure_faces(c);
    if (rc) return rc;
    if (c->n_s == 0 || (!d_gz && !d_tmi)) return LF_OK;
    const unsigned sblocks = (unsigned)((c->n_s + 255) / 256);
    if (c->n_f == 0) {
        if (d_gz) hipLaunchKernelGGL(k_fill_zero, dim3(sblocks), dim3(256), 0, c->stream, d_gz, c->n_s);
        if (d_tmi) hipLaunchKernelGGL(k_fill_zero, dim3(sblocks), dim3(256), 0, c->stream, d_tmi, c->n_s);
        LF_TRY(c, hipGetLastError());
        return LF_OK;
    }
    const bool f32 = c->precision == LF_FP32C;
    if (!c->d_dq)
        LF_TRY(c, hipMalloc(&c->d_dq, (f32 ? sizeof(float) : sizeof(double)) * kSeedsPad * c->n_f));
    {
        TimedLaunch tl(c, 6);
        const unsigned fb = (unsigned)((c->n_f + 127) / 128);
        const dim3 grid(sblocks, (unsigned)c->n_chunks);
        if (f32) {
            hipLaunchKernelGGL(k_facet_tangent<float>, dim3(fb), dim3(128), 0, c->stream, c->n_f,
                               c->d_verts, c->d_tri, c->d_front, c->d_back, c->d_density, c->d_mag,
                               c->fdir[0], c->fdir[1], c->fdir[2], c->tmi_mode, d_vtan,
                               (float*)c->d_dq);
            hipLaunchKernelGGL(k_jvp_pairs<float>, grid, dim3(256), 0, c->stream, c->d_f32,
                               (const float*)c->d_dq, c->n_f, c->faces_per_chunk, c->d_sx, c->d_sy,
                               c->d_sz, c->n_s, c->d_part_gz, c->d_part_tmi);
        } else {
            hipLaunchKernelGGL(k_facet_tangent<double>, dim3(fb), dim3(128), 0, c->stream, c->n_f,
                               c->d_verts, c->d_tri, c->d_front, c->d_back, c->d_density, c->d_mag,
                               c->fdir[0], c->fdir[1], c->fdir[2], c->tmi_mode, d_vtan,
                               (double*)c->d_dq);
            hipLaunchKernelGGL(k_jvp_pairs<double>, grid, dim3(256), 0, c->stream, c->d_fr64,
                               (const double*)c->d_dq, c->n_f, c->faces_per_chunk, c->d_sx,
                               c->d_sy, c->d_sz, c->n_s, c->d_part_gz, c->d_part_tmi);
        }
        LF_TRY(c, hipGetLastError());
    }
    {
        TimedLaunch tl(c, 2);
        if (d_gz)
            hipLaunchKernelGGL(k_reduce_chunks, dim3(sblocks), dim3(256), 0, c->stream,
                               c->d_part_gz, c->n_chunks, c->n_s, d_gz);
        if (d_tmi)
            hipLaunchKernelGGL(k_reduce_chunks, dim3(sblocks), dim3(256), 0, c->stream,
                               c->d_part_tmi, c->n_chunks, c->n_s, d_tmi);
        LF_TRY(c, hipGetLastError());
    }
    return LF_OK;
}

int lf_jvp(lf_ctx* c, const double* vtan, double* gz, double* tmi)
{
    if (check_ctx(c)) return LF_EINVAL;
    if (!vtan && c->n_v) return set_err(c, "lf_jvp: null tangent", LF_EINVAL);
    LF_TRY(c, hipSetDevice(c->device));
    if (!c->d_vtan && c->n_v) LF_TRY(c, hipMalloc((void**)&c->d_vtan, sizeof(double) * 3 * c->n_v));
    if (c->n_v)
        LF_TRY(c, hipMemcpyAsync(c->d_vtan, vtan, sizeof(double) * 3 * c->n_v,
                                 hipMemcpyHostToDevice, c->stream));
    int rc = lf_jvp_device(c, c->d_vtan, gz ? c->d_out_gz : nullptr, tmi ? c->d_out_tmi : nullptr);
    if (rc) return rc;
    if (gz && c->n_s)
        LF_TRY(c, hipMemcpyAsync(gz, c->d_out_gz, sizeof(double) * c->n_s, hipMemcpyDeviceToHost,
                                 c->stream));
    if (tmi && c->n_s)
        LF_TRY(c, hipMemcpyAsync(tmi, c->d_out_tmi, sizeof(double) * c->n_s,
                                 hipMemcpyDeviceToHost, c->stream));
    LF_TRY(c, hipStreamSynchronize(c->stream));
    return LF_OK;
}

int lf_sync(lf_ctx* c)
{
    if (check_ctx(c)) return LF_EINVAL;
    LF_TRY(c, hipSetDevice(c->device));
    LF_TRY(c, hipStreamSynchronize(c->stream));
    return LF_OK;
}

int lf_set_timing(lf_ctx* c, int enable)
{
    if (check_ctx(c)) return LF_EINVAL;
    c->timing = enable != 0;
    return LF_OK;
}

int lf_timing(lf_ctx* c, double* ms, int64_t* launches, int n_classes)
{
    if (check_ctx(c)) return LF_EINVAL;
    LF_TRY(c, hipSetDevice(c->device));
    LF_TRY(c, hipStreamSynchronize(c->stream));
    for (auto& e : c->events) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, e.a, e.b) == hipSuccess && e.cls >= 0 && e.cls < 8) {
            c->ms[e.cls] += t;
            c->launches[e.cls] += 1;
        }
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    c->events.clear();
    for (int k = 0; k < n_classes && k < 8; ++k) {
        if (ms) ms[k] = c->ms[k];
        if (launches) launches[k] = c->launches[k];
        c->ms[k] = 0.0;
        c->launches[k] = 0;
    }
    return LF_OK;
}

int lf_stats(lf_ctx* c, int64_t* fp64_edges)
{
    if (check_ctx(c)) return LF_EINVAL;
    LF_TRY(c, hipSetDevice(c->device));
    unsigned long long v = 0;
    LF_TRY(c, hipMemcpyAsync(&v, c->d_fb, sizeof v, hipMemcpyDeviceToHost, c->stream));
    LF_TRY(c, hipStreamSynchronize(c->stream));
    if (fp64_edges) *fp64_edges = (int64_t)v;
    return LF_OK;
}

}  // extern "C"

// ============================================================== forward, no allocation

namespace {

// Scratch of lf_forward_noalloc: one per GPU, part of the code object, so it exists from
// the library's load on that GPU until the process ends.  Host code never uses the
// symbol's name as an address (on the host it is a stand-in): it uses the per-device
// address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[LF_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
int g_scratch_cus[kMaxDevices] = {};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct NoallocLayout {
    size_t recs, geo, st, part, fb, total;
};

// Byte offsets in the scratch: facet records (FaceF32 + FP64 corners for FP32C, FaceF64
// for FP64), station x/y/z arrays, per-chunk gz and TMI partials, the FP64-edge counter.
NoallocLayout noalloc_layout(int64_t n_f, int64_t n_s, int precision, int n_chunks)
{
    NoallocLayout L;
    L.recs = 0;
    const size_t rec = precision == LF_FP32C ? sizeof(FaceF32) : sizeof(FaceF64);
    L.geo = align256((size_t)n_f * rec);
    const size_t geo = precision == LF_FP32C ? (size_t)n_f * 9 * sizeof(double) : 0;
    L.st = align256(L.geo + geo);
    L.part = align256(L.st + 3 * (size_t)n_s * sizeof(double));
    L.fb = align256(L.part + 2 * (size_t)n_chunks * n_s * sizeof(double));
    L.total = L.fb + sizeof(unsigned long long);
    return L;
}

}  // namespace

extern "C" int lf_forward_noalloc(const lf_desc* d, double* d_gz, double* d_tmi, void* stream)
{
    g_create_error.clear();
    if (!d || d->n_vertices < 0 || d->n_faces < 0 || d->n_stations < 0 || d->n_units <= 0 ||
        (d->n_faces > 0 && (!d->vertices || !d->faces || !d->front || !d->back)) ||
        (d->n_stations > 0 && !d->stations) || !d->density || !d->magnetization ||
        (d->precision != LF_FP64 && d->precision != LF_FP32C) ||
        (d->tmi_mode != LF_TMI_REFERENCE && d->tmi_mode != LF_TMI_TEXTBOOK) ||
        (d->want & ~3) != 0 || d->n_vertices > 0xffffffffLL) {
        g_create_error = "lf_forward_noalloc: invalid descriptor";
        return LF_EINVAL;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != d->device || dev < 0 || dev >= kMaxDevices) {
        g_create_error = "lf_forward_noalloc: desc->device must be the current HIP device";
        return LF_EINVAL;
    }
    unsigned char* base = nullptr;
    int cus = 0;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                g_create_error = std::string("lf_forward_noalloc: hipGetSymbolAddress: ") +
                                 hipGetErrorString(e);
                return LF_EHIP;
            }
            int n = 0;
            if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                n <= 0)
                n = 256;
            g_scratch_cus[dev] = n;
            g_scratch_ptr[dev] = (unsigned char*)p;
        }
        base = g_scratch_ptr[dev];
        cus = g_scratch_cus[dev];
    }
    const int64_t n_f = d->n_faces, n_s = d->n_stations;
    const int want = d->want ? d->want : 3;
    const bool gz = d_gz && (want & LF_WANT_GZ);
    const bool tm = d_tmi && (want & LF_WANT_TMI);
    hipStream_t st = (hipStream_t)stream;
    if (n_s == 0 || (!gz && !tm)) return LF_OK;
    const unsigned sblocks = (unsigned)((n_s + 255) / 256);
    if (n_f == 0) {
        if (gz) hipLaunchKernelGGL(k_fill_zero, dim3(sblocks), dim3(256), 0, st, d_gz, n_s);
        if (tm) hipLaunchKernelGGL(k_fill_zero, dim3(sblocks), dim3(256), 0, st, d_tmi, n_s);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            g_create_error = std::string("lf_forward_noalloc: ") + hipGetErrorString(e);
            return LF_EHIP;
        }
        return LF_OK;
    }
    int n_tiles = 1, n_chunks = 1;
    int64_t fpc = 1;
    // partial sums of at most 8 MiB, so that the scratch holds the largest models
    const int64_t max_chunks = std::max<int64_t>(1, (8ll << 20) / (16 * std::max<int64_t>(n_s, 1)));
    chunk_plan(n_s, n_f, d->precision, cus, max_chunks, n_tiles, fpc, n_chunks);
    const NoallocLayout L = noalloc_layout(n_f, n_s, d->precision, n_chunks);
    if (L.total > (size_t)LF_NOALLOC_SCRATCH_BYTES) {
        g_create_error = "lf_forward_noalloc: the model needs " + std::to_string(L.total) +
                         " bytes of scratch, more than LF_NOALLOC_SCRATCH_BYTES (" +
                         std::to_string((unsigned long long)LF_NOALLOC_SCRATCH_BYTES) + ")";
        return LF_ENOMEM;
    }
    FaceF32* f32 = d->precision == LF_FP32C ? (FaceF32*)(base + L.recs) : nullptr;
    FaceF64* f64 = d->precision == LF_FP64 ? (FaceF64*)(base + L.recs) : nullptr;
    double* geo = d->precision == LF_FP32C ? (double*)(base + L.geo) : nullptr;
    double* sx = (double*)(base + L.st);
    double* sy = sx + n_s;
    double* sz = sy + n_s;
    double* pg = (double*)(base + L.part);
    double* pt = pg + (size_t)n_chunks * n_s;
    unsigned long long* fb = (unsigned long long*)(base + L.fb);

    hipLaunchKernelGGL(k_face_setup, dim3((unsigned)((n_f + 127) / 128)), dim3(128), 0, st,
                       d->vertices, d->faces, d->front, d->back, d->density, d->magnetization,
                       d->field_dir[0], d->field_dir[1], d->field_dir[2], d->tmi_mode, n_f,
                       d->n_vertices, (int)d->n_units, f32, f64,
                       (FaceFrame<double>*)nullptr, geo);
    hipLaunchKernelGGL(k_split_xyz, dim3(sblocks), dim3(256), 0, st, d->stations, n_s, sx, sy, sz);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, fb);
    const dim3 grid((unsigned)((int64_t)n_tiles * n_chunks));
#define LF_FWD_NOALLOC(G, T)                                                                  \
    do {                                                                                     \
        if (d->precision == LF_FP32C)                                                        \
            hipLaunchKernelGGL((k_forward_f32c<G, T>), grid, dim3(kBlock), 0, st, f32, geo,  \
                               n_f, fpc, sx, sy, sz, n_s, n_tiles, pg, pt, fb);             \
        else                                                                                 \
            hipLaunchKernelGGL((k_forward_f64<G, T>), grid, dim3(kBlock), 0, st, f64, n_f,   \
                               fpc, sx, sy, sz, n_s, n_tiles, pg, pt);                       \
    } while (0)
    if (gz && tm) LF_FWD_NOALLOC(true, true);
    else if (gz) LF_FWD_NOALLOC(true, false);
    else LF_FWD_NOALLOC(false, true);
#undef LF_FWD_NOALLOC
    if (gz)
        hipLaunchKernelGGL(k_reduce_chunks, dim3(sblocks), dim3(256), 0, st, pg, n_chunks, n_s, d_gz);
    if (tm)
        hipLaunchKernelGGL(k_reduce_chunks, dim3(sblocks), dim3(256), 0, st, pt, n_chunks, n_s, d_tmi);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_create_error = std::string("lf_forward_noalloc: ") + hipGetErrorString(e);
        return LF_EHIP;
    }
    return LF_OK;
}
