// litho_host.cpp -- implementation of the C++ host layer (see litho_host.hpp).
#include "litho_host.hpp"

#include <algorithm>
#include <array>
#include <thread>
#include <tuple>
#include <chrono>
#include <cmath>
#include <limits>
#include <map>
#include <random>
#include <set>
#include <stdexcept>
#include <unordered_map>

#include "../../include/lithofwd.h"

namespace lithofwd {

namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr double kEps = 1e-15;

double dot(const Vec& a, const Vec& b)
{
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

void fieldUnitVector(double inc_deg, double dec_deg, double out[3])
{
    // magnetic_forward.cpp:61-66
    const double I = inc_deg * kPi / 180.0, D = dec_deg * kPi / 180.0;
    out[0] = std::cos(I) * std::cos(D);
    out[1] = std::cos(I) * std::sin(D);
    out[2] = -std::sin(I);
}

void check(int rc, lf_ctx* c, const char* what)
{
    if (rc != 0)
        throw std::runtime_error(std::string(what) + ": " + lf_last_error(c));
}
}  // namespace

// =========================================================================== model

LithologyModel::LithologyModel(std::vector<double> vertices, std::vector<uint32_t> faces,
                               std::vector<int32_t> front, std::vector<int32_t> back,
                               std::vector<LithoGroup> groups)
    : m_vertices(std::move(vertices)), m_faces(std::move(faces)), m_front(std::move(front)),
      m_back(std::move(back)), m_groups(std::move(groups))
{
    if (m_vertices.size() % 3 || m_faces.size() != 3 * m_front.size() ||
        m_front.size() != m_back.size())
        throw std::invalid_argument("LithologyModel: inconsistent array sizes");
    m_rest = m_vertices;
    m_freedom.assign(vertexCount(), VertexFreedom::FIXED);
    m_moveVector.assign(3 * vertexCount(), 0.0);
    m_isCP.assign(vertexCount(), 1);
    rebuildDofs();
}

void LithologyModel::setVertexFreedom(size_t v, VertexFreedom f)
{
    m_freedom.at(v) = f;
    rebuildDofs();
}

void LithologyModel::setMoveVector(size_t v, const double mv[3])
{
    if (v >= vertexCount()) throw std::out_of_range("setMoveVector: vertex out of range");
    // Vector3d::normalized() (surface_mesh.cpp:195, 218); zero stays zero
    const double n = std::sqrt(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2]);
    for (int a = 0; a < 3; ++a) m_moveVector[3 * v + a] = n > 0.0 ? mv[a] / n : 0.0;
}

void LithologyModel::fixExteriorVertices(VertexFreedom interior)
{
    std::vector<uint8_t> ext(vertexCount(), 0);
    for (size_t f = 0; f < m_front.size(); ++f)
        if (m_back[f] < 0)
            for (int k = 0; k < 3; ++k) ext[m_faces[3 * f + k]] = 1;
    for (size_t v = 0; v < vertexCount(); ++v)
        m_freedom[v] = ext[v] ? VertexFreedom::FIXED : interior;
    rebuildDofs();
}

void LithologyModel::fixExteriorFaces(VertexFreedom freedom)
{
    const size_t nv = vertexCount();
    m_freedom.assign(nv, freedom);
    // per vertex: bounding facets by dominant normal class, and contact facets
    std::vector<uint8_t> tb(nv, 0), ns(nv, 0), ew(nv, 0), contact(nv, 0);
    for (size_t f = 0; f < m_front.size(); ++f) {
        const uint32_t* t = &m_faces[3 * f];
        if (m_back[f] >= 0) {
            for (int k = 0; k < 3; ++k) contact[t[k]] = 1;
            continue;
        }
        const double* a = &m_vertices[3 * (size_t)t[0]];
        const double* b = &m_vertices[3 * (size_t)t[1]];
        const double* c = &m_vertices[3 * (size_t)t[2]];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                             e1[0] * e2[1] - e1[1] * e2[0]};
        if (std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) < 1e-30) continue;
        const double ax = std::abs(n[0]), ay = std::abs(n[1]), az = std::abs(n[2]);
        std::vector<uint8_t>& cls = (az >= ax && az >= ay) ? tb : (ay >= ax ? ns : ew);
        for (int k = 0; k < 3; ++k) cls[t[k]] = 1;
    }
    for (size_t v = 0; v < nv; ++v) {
        const int nClass = tb[v] + ns[v] + ew[v];
        if (nClass == 0) continue;
        VertexFreedom f;
        if (nClass >= 3)
            f = VertexFreedom::FIXED;
        else if (nClass == 2)
            f = (tb[v] && ns[v]) ? VertexFreedom::X_ONLY
                                 : (tb[v] && ew[v]) ? VertexFreedom::Y_ONLY : VertexFreedom::Z_ONLY;
        else
            f = tb[v] ? VertexFreedom::XY_FREE : (ns[v] ? VertexFreedom::X_ONLY : VertexFreedom::Y_ONLY);
        if (contact[v] && !tb[v]) {
            switch (f) {
                case VertexFreedom::FIXED: f = VertexFreedom::XYZ_FREE; break;
                case VertexFreedom::X_ONLY: f = VertexFreedom::XZ_FREE; break;
                case VertexFreedom::Y_ONLY: f = VertexFreedom::YZ_FREE; break;
                case VertexFreedom::XY_FREE: f = VertexFreedom::XYZ_FREE; break;
                default: break;
            }
        }
        m_freedom[v] = f;
    }
    rebuildDofs();
}

void LithologyModel::setAxisScale(int axis, double s)
{
    if (axis >= 0 && axis < 3 && s > 0) m_axisScale[axis] = s;
}

bool LithologyModel::isControlPoint(size_t v) const { return m_cpStride <= 0 || m_isCP[v]; }

namespace {
// axes a vertex of this freedom may move along
bool axisFree(VertexFreedom f, int a)
{
    switch (f) {
        case VertexFreedom::XYZ_FREE: return true;
        case VertexFreedom::X_ONLY: return a == 0;
        case VertexFreedom::Y_ONLY: return a == 1;
        case VertexFreedom::Z_ONLY: return a == 2;
        case VertexFreedom::XY_FREE: return a != 2;
        case VertexFreedom::XZ_FREE: return a != 1;
        case VertexFreedom::YZ_FREE: return a != 0;
        default: return false;
    }
}
}  // namespace

void LithologyModel::projectFree(size_t v, double d[3]) const
{
    if (m_freedom[v] == VertexFreedom::ALONG_VECTOR) {
        const double* u = &m_moveVector[3 * v];
        const double t = d[0] * u[0] + d[1] * u[1] + d[2] * u[2];
        for (int a = 0; a < 3; ++a) d[a] = t * u[a];
        return;
    }
    for (int a = 0; a < 3; ++a)
        if (!axisFree(m_freedom[v], a)) d[a] = 0.0;
}

namespace {
// models with at least this many vertices map parameters to vertices and pull gradients
// back on OpenMP threads (a few microseconds per parallel region; single-threaded, the
// two loops took about 4.5 ms per L-BFGS-B iteration on a 67,054-vertex model)
constexpr int64_t kParallelVertices = 16384;

// body(i) for i in [0, n) on up to 16 threads
template <class F>
void parallel_range(int64_t n, F body)
{
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(
        {(int64_t)std::max(1u, std::thread::hardware_concurrency()), 16, n / 2048 + 1}));
    if (nt == 1) {
        for (int64_t i = 0; i < n; ++i) body(i);
        return;
    }
    std::vector<std::thread> th;
    for (int64_t t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int64_t i = t; i < n; i += nt) body(i);
        });
    for (auto& x : th) x.join();
}

// Control points of one unit in a dense grid of cubic buckets, for K-nearest queries.
struct CpBuckets {
    double lo[3] = {0, 0, 0}, h = 1.0;
    int64_t n[3] = {0, 0, 0};
    std::vector<int64_t> off;
    std::vector<int32_t> cps;

    void build(const std::vector<int32_t>& cpList, const std::vector<double>& V, double bucket)
    {
        h = bucket;
        double hi[3] = {-1e300, -1e300, -1e300};
        for (int a = 0; a < 3; ++a) lo[a] = 1e300;
        for (int32_t c : cpList)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], V[3 * (size_t)c + a]);
                hi[a] = std::max(hi[a], V[3 * (size_t)c + a]);
            }
        // coarsen the buckets if the grid would be large against the CP count
        for (;;) {
            int64_t total = 1;
            for (int a = 0; a < 3; ++a) {
                n[a] = (int64_t)std::floor((hi[a] - lo[a]) / h) + 1;
                total *= n[a];
            }
            if (total <= 8 * (int64_t)cpList.size() + 64) break;
            h *= 2.0;
        }
        off.assign((size_t)(n[0] * n[1] * n[2]) + 1, 0);
        auto idx = [&](int32_t c) {
            int64_t b[3];
            for (int a = 0; a < 3; ++a)
                b[a] = std::clamp((int64_t)std::floor((V[3 * (size_t)c + a] - lo[a]) / h),
                                  (int64_t)0, n[a] - 1);
            return (b[0] * n[1] + b[1]) * n[2] + b[2];
        };
        for (int32_t c : cpList) ++off[(size_t)idx(c) + 1];
        for (size_t i = 1; i < off.size(); ++i) off[i] += off[i - 1];
        cps.resize(cpList.size());
        std::vector<int64_t> fill(off.begin(), off.end() - 1);
        for (int32_t c : cpList) cps[(size_t)fill[(size_t)idx(c)]++] = c;
    }

    // the K nearest control points to p (fewer if the unit has fewer), nearest first
    void nearest(const double* p, const std::vector<double>& V, int K,
                 std::vector<std::pair<double, int32_t>>& out) const
    {
        out.clear();
        if (cps.empty()) return;
        int64_t b[3];
        for (int a = 0; a < 3; ++a) b[a] = (int64_t)std::floor((p[a] - lo[a]) / h);
        // distance from p to the grid box; shells beyond it are searched until the K-th
        // distance is inside the searched cube
        const int64_t rmax = std::max({n[0], n[1], n[2]}) + 2 +
                             std::max({std::abs(b[0]), std::abs(b[1]), std::abs(b[2]),
                                       std::abs(b[0] - n[0]), std::abs(b[1] - n[1]),
                                       std::abs(b[2] - n[2])});
        const size_t want = std::min<size_t>((size_t)K, cps.size());
        for (int64_t r = 0; r <= rmax; ++r) {
            // add the shell at Chebyshev distance r
            for (int64_t dx = -r; dx <= r; ++dx)
                for (int64_t dy = -r; dy <= r; ++dy)
                    for (int64_t dz = -r; dz <= r; ++dz) {
                        if (std::max({std::abs(dx), std::abs(dy), std::abs(dz)}) != r) continue;
                        const int64_t x = b[0] + dx, y = b[1] + dy, z = b[2] + dz;
                        if (x < 0 || y < 0 || z < 0 || x >= n[0] || y >= n[1] || z >= n[2]) continue;
                        const size_t cell = (size_t)((x * n[1] + y) * n[2] + z);
                        for (int64_t k = off[cell]; k < off[cell + 1]; ++k) {
                            const int32_t c = cps[(size_t)k];
                            const double* q = &V[3 * (size_t)c];
                            const double d2 = (q[0] - p[0]) * (q[0] - p[0]) +
                                              (q[1] - p[1]) * (q[1] - p[1]) +
                                              (q[2] - p[2]) * (q[2] - p[2]);
                            out.push_back({d2, c});
                        }
                    }
            if (out.size() >= want) {
                std::nth_element(out.begin(), out.begin() + (want - 1), out.end());
                // every point within r*h of p lies in the searched cube
                const double rh = (double)r * h;
                if (out.size() == cps.size() || out[want - 1].first <= rh * rh) break;
            }
        }
        const size_t keep = std::min(want, out.size());
        std::partial_sort(out.begin(), out.begin() + keep, out.end());
        out.resize(keep);
        for (auto& e : out) e.first = std::sqrt(e.first);
    }
};
}  // namespace

uint32_t LithologyModel::setControlPointStride(int stride, double cellSize)
{
    const size_t nv = vertexCount();
    m_cpStride = stride;
    m_wOff.clear();
    m_wCp.clear();
    m_w.clear();
    m_depOff.clear();
    m_depV.clear();
    m_depW.clear();
    m_isCP.assign(nv, 1);
    if (stride <= 0 || nv < 4) {
        m_cpStride = 0;
        rebuildDofs();
        return totalDofCount();
    }
    // quantized positions (1 um), lexicographic order like the reference's std::map
    struct Q { int64_t x, y, z; };
    std::vector<Q> q(nv);
    for (size_t v = 0; v < nv; ++v)
        q[v] = {(int64_t)std::llround(m_vertices[3 * v] * 1e6),
                (int64_t)std::llround(m_vertices[3 * v + 1] * 1e6),
                (int64_t)std::llround(m_vertices[3 * v + 2] * 1e6)};
    auto lexLess = [&](int32_t a, int32_t b) {
        if (q[a].x != q[b].x) return q[a].x < q[b].x;
        if (q[a].y != q[b].y) return q[a].y < q[b].y;
        return q[a].z < q[b].z;
    };
    if (cellSize <= 0.0) {
        // lithology_model.cpp:916-934: the 300 lexicographically first positions
        std::vector<int32_t> first(nv);
        for (size_t v = 0; v < nv; ++v) first[v] = (int32_t)v;
        const size_t ns = std::min<size_t>(nv, 300);
        std::partial_sort(first.begin(), first.begin() + ns, first.end(), lexLess);
        double mn[3] = {1e12, 1e12, 1e12};
        for (size_t i = 0; i < ns; ++i)
            for (size_t j = i + 1; j < ns; ++j) {
                const Q& a = q[first[i]];
                const Q& b = q[first[j]];
                const double d[3] = {std::abs((double)(a.x - b.x)) * 1e-6,
                                     std::abs((double)(a.y - b.y)) * 1e-6,
                                     std::abs((double)(a.z - b.z)) * 1e-6};
                for (int k = 0; k < 3; ++k)
                    if (d[k] > 0 && d[k] < mn[k]) mn[k] = d[k];
            }
        cellSize = std::max({mn[0], mn[1], mn[2], 1.0});
    }
    // lithology_model.cpp:952-989: one control point per populated cell of a stride*cell
    // grid, the position closest to the cell centre (ties: the lexicographically first)
    const int64_t strideQ = std::max<int64_t>(1, (int64_t)stride * (int64_t)(cellSize * 1e6));
    auto cellOf = [strideQ](int64_t v) { return v >= 0 ? v / strideQ : (v - strideQ + 1) / strideQ; };
    struct Cand { int64_t cx, cy, cz, d2; int32_t v; };
    std::vector<Cand> cand(nv);
    parallel_range((int64_t)nv, [&](int64_t i) {
        const size_t v = (size_t)i;
        const int64_t cx = cellOf(q[v].x), cy = cellOf(q[v].y), cz = cellOf(q[v].z);
        const int64_t ox = cx * strideQ + strideQ / 2, oy = cy * strideQ + strideQ / 2,
                      oz = cz * strideQ + strideQ / 2;
        const int64_t d2 = (q[v].x - ox) * (q[v].x - ox) + (q[v].y - oy) * (q[v].y - oy) +
                           (q[v].z - oz) * (q[v].z - oz);
        cand[v] = {cx, cy, cz, d2, (int32_t)v};
    });
    std::sort(cand.begin(), cand.end(), [&](const Cand& a, const Cand& b) {
        if (a.cx != b.cx) return a.cx < b.cx;
        if (a.cy != b.cy) return a.cy < b.cy;
        if (a.cz != b.cz) return a.cz < b.cz;
        if (a.d2 != b.d2) return a.d2 < b.d2;
        return lexLess(a.v, b.v);
    });
    m_isCP.assign(nv, 0);
    for (size_t i = 0; i < nv; ++i)
        if (i == 0 || cand[i].cx != cand[i - 1].cx || cand[i].cy != cand[i - 1].cy ||
            cand[i].cz != cand[i - 1].cz)
            m_isCP[(size_t)cand[i].v] = 1;
    std::vector<Cand>().swap(cand);

    // SurfaceMesh::setControlPointsExplicit (surface_mesh.cpp:544-601) per unit surface: a
    // vertex follows the K = 4 nearest control points on that surface with inverse-distance
    // weights; a vertex on several surfaces takes the mean of their interpolations, as the
    // reference's shared-vertex averaging does (lithology_model.cpp:230-247).  A surface
    // with fewer than two control points has none (the reference then gives each of its
    // vertices a parameter): its vertices become control points here.
    const int nu = groupCount();
    std::vector<std::vector<int32_t>> unitVerts(nu);
    {
        std::vector<uint8_t> on((size_t)nu * nv, 0);  // on[u * nv + v]: v on u's surface
        for (size_t f = 0; f < faceCount(); ++f)
            for (int side = 0; side < 2; ++side) {
                const int u = side == 0 ? m_front[f] : m_back[f];
                if (u < 0) continue;
                for (int k = 0; k < 3; ++k) on[(size_t)u * nv + m_faces[3 * f + k]] = 1;
            }
        for (int u = 0; u < nu; ++u)
            for (size_t v = 0; v < nv; ++v)
                if (on[(size_t)u * nv + v]) unitVerts[u].push_back((int32_t)v);
    }
    std::vector<std::vector<int32_t>> unitCps(nu);
    for (int u = 0; u < nu; ++u) {
        for (int32_t v : unitVerts[u])
            if (m_isCP[(size_t)v]) unitCps[u].push_back(v);
        if (unitCps[u].size() < 2)
            for (int32_t v : unitVerts[u]) m_isCP[(size_t)v] = 1;
    }
    for (int u = 0; u < nu; ++u) {
        unitCps[u].clear();
        for (int32_t v : unitVerts[u])
            if (m_isCP[(size_t)v]) unitCps[u].push_back(v);
    }
    const double bucket = std::max(1.0, stride * cellSize);
    std::vector<CpBuckets> grids(nu);
    for (int u = 0; u < nu; ++u) grids[u].build(unitCps[u], m_vertices, bucket);
    // units of each vertex (CSR)
    std::vector<int64_t> vuOff(nv + 1, 0);
    for (int u = 0; u < nu; ++u)
        for (int32_t v : unitVerts[u]) ++vuOff[(size_t)v + 1];
    for (size_t v = 0; v < nv; ++v) vuOff[v + 1] += vuOff[v];
    std::vector<int32_t> vuList(vuOff[nv]);
    {
        std::vector<int64_t> fill(vuOff.begin(), vuOff.end() - 1);
        for (int u = 0; u < nu; ++u)
            for (int32_t v : unitVerts[u]) vuList[(size_t)fill[(size_t)v]++] = u;
    }
    const int K = 4;
    std::vector<int64_t> slot(nv + 1, 0);
    for (size_t v = 0; v < nv; ++v) slot[v + 1] = slot[v] + (m_isCP[v] ? 0 : K * (vuOff[v + 1] - vuOff[v]));
    std::vector<int32_t> wc(slot[nv], -1);
    std::vector<double> ww(slot[nv], 0.0);
    std::vector<int32_t> cnt(nv, 0);
    parallel_range((int64_t)nv, [&](int64_t i) {
        const size_t v = (size_t)i;
        if (m_isCP[v]) return;
        const int64_t nUnits = vuOff[v + 1] - vuOff[v];
        if (nUnits == 0) return;
        thread_local std::vector<std::pair<double, int32_t>> nb;
        thread_local std::vector<std::pair<int32_t, double>> acc;
        acc.clear();
        for (int64_t k = vuOff[v]; k < vuOff[v + 1]; ++k) {
            grids[vuList[(size_t)k]].nearest(&m_vertices[3 * v], m_vertices, K, nb);
            double tot = 0.0;
            for (auto& e : nb) tot += 1.0 / std::max(e.first, 1e-12);
            for (auto& e : nb)
                acc.push_back({e.second, (1.0 / std::max(e.first, 1e-12)) / tot / (double)nUnits});
        }
        std::sort(acc.begin(), acc.end());
        int32_t m = 0;
        for (size_t j = 0; j < acc.size(); ++j) {
            if (m > 0 && wc[(size_t)slot[v] + m - 1] == acc[j].first) {
                ww[(size_t)slot[v] + m - 1] += acc[j].second;
                continue;
            }
            wc[(size_t)slot[v] + m] = acc[j].first;
            ww[(size_t)slot[v] + m] = acc[j].second;
            ++m;
        }
        cnt[v] = m;
    });
    m_wOff.assign(nv + 1, 0);
    for (size_t v = 0; v < nv; ++v) m_wOff[v + 1] = m_wOff[v] + cnt[v];
    m_wCp.resize((size_t)m_wOff[nv]);
    m_w.resize((size_t)m_wOff[nv]);
    for (size_t v = 0; v < nv; ++v)
        for (int32_t j = 0; j < cnt[v]; ++j) {
            m_wCp[(size_t)m_wOff[v] + j] = wc[(size_t)slot[v] + j];
            m_w[(size_t)m_wOff[v] + j] = ww[(size_t)slot[v] + j];
        }
    buildDependants();
    rebuildDofs();
    return totalDofCount();
}

void LithologyModel::buildDependants()
{
    const size_t nv = vertexCount();
    m_depOff.assign(nv + 1, 0);
    for (size_t e = 0; e < m_wCp.size(); ++e) m_depOff[(size_t)m_wCp[e] + 1]++;
    for (size_t v = 0; v < nv; ++v) m_depOff[v + 1] += m_depOff[v];
    m_depV.resize(m_wCp.size());
    m_depW.resize(m_wCp.size());
    std::vector<int64_t> pos(m_depOff.begin(), m_depOff.end() - 1);
    for (size_t v = 0; v < nv; ++v)  // dependants in vertex order
        for (int64_t e = m_wOff[v]; e < m_wOff[v + 1]; ++e) {
            const int64_t k = pos[(size_t)m_wCp[e]]++;
            m_depV[(size_t)k] = (int32_t)v;
            m_depW[(size_t)k] = m_w[e];
        }
}

void LithologyModel::rebuildDofs()
{
    m_dofVertex.clear();
    m_dofAxis.clear();
    for (size_t v = 0; v < vertexCount(); ++v) {
        if (m_cpStride > 0 && !m_isCP[v]) continue;
        // one parameter per free axis, x before y before z; one for ALONG_VECTOR, recorded
        // on axis 0 as the reference does (surface_mesh.cpp:91-172)
        if (m_freedom[v] == VertexFreedom::ALONG_VECTOR) {
            m_dofVertex.push_back((int)v);
            m_dofAxis.push_back(0);
            continue;
        }
        for (int a = 0; a < 3; ++a)
            if (axisFree(m_freedom[v], a)) {
                m_dofVertex.push_back((int)v);
                m_dofAxis.push_back(a);
            }
    }
}

bool LithologyModel::dofAlongVector(size_t d) const
{
    return m_freedom[(size_t)m_dofVertex[d]] == VertexFreedom::ALONG_VECTOR;
}

Vec LithologyModel::assembleParameterVector() const
{
    // extractParams (surface_mesh.cpp:211-225)
    Vec p(m_dofVertex.size());
    for (size_t d = 0; d < p.size(); ++d) {
        const double* x = &m_vertices[3 * (size_t)m_dofVertex[d]];
        if (dofAlongVector(d)) {
            const double* u = &m_moveVector[3 * (size_t)m_dofVertex[d]];
            p[d] = x[0] * u[0] + x[1] * u[1] + x[2] * u[2];
        } else {
            p[d] = x[m_dofAxis[d]] / m_axisScale[m_dofAxis[d]];
        }
    }
    return p;
}

void LithologyModel::applyParameterVector(const Vec& params)
{
    if (params.size() != m_dofVertex.size())
        throw std::invalid_argument("applyParameterVector: wrong parameter count");
    // applyParams (surface_mesh.cpp:188-209)
    for (size_t d = 0; d < params.size(); ++d) {
        double* x = &m_vertices[3 * (size_t)m_dofVertex[d]];
        if (dofAlongVector(d)) {
            const double* u = &m_moveVector[3 * (size_t)m_dofVertex[d]];
            const double t = params[d] - (x[0] * u[0] + x[1] * u[1] + x[2] * u[2]);
            for (int a = 0; a < 3; ++a) x[a] += t * u[a];
        } else {
            x[m_dofAxis[d]] = params[d] * m_axisScale[m_dofAxis[d]];
        }
    }
    if (m_cpStride <= 0) return;
    // every dependant reads only control-point vertices, which this loop does not write
    const int64_t nv = (int64_t)vertexCount();
#pragma omp parallel for schedule(static) if (nv >= kParallelVertices)
    for (int64_t vi = 0; vi < nv; ++vi) {
        const size_t v = (size_t)vi;
        if (m_isCP[v]) continue;
        double* x = &m_vertices[3 * v];
        if (m_cpInterp == ControlPointInterpolation::Position) {
            // interpolateFromControlPoints (surface_mesh.cpp:610-623)
            double pos[3] = {0.0, 0.0, 0.0};
            for (int64_t e = m_wOff[v]; e < m_wOff[v + 1]; ++e) {
                const size_t c = (size_t)m_wCp[e];
                for (int a = 0; a < 3; ++a) pos[a] += m_w[e] * m_vertices[3 * c + a];
            }
            for (int a = 0; a < 3; ++a) x[a] = pos[a];
        } else {
            // the control points' weighted displacement, along the vertex's own free
            // directions (a FIXED vertex never moves)
            double disp[3] = {0.0, 0.0, 0.0};
            for (int64_t e = m_wOff[v]; e < m_wOff[v + 1]; ++e) {
                const size_t c = (size_t)m_wCp[e];
                for (int a = 0; a < 3; ++a)
                    disp[a] += m_w[e] * (m_vertices[3 * c + a] - m_rest[3 * c + a]);
            }
            projectFree(v, disp);
            for (int a = 0; a < 3; ++a) x[a] = m_rest[3 * v + a] + disp[a];
        }
    }
}

Vec LithologyModel::vertexTangent(const Vec& dp) const
{
    if (dp.size() != m_dofVertex.size())
        throw std::invalid_argument("vertexTangent: wrong parameter count");
    Vec t(3 * vertexCount(), 0.0);
    for (size_t d = 0; d < dp.size(); ++d) {
        double* x = &t[3 * (size_t)m_dofVertex[d]];
        if (dofAlongVector(d)) {
            const double* u = &m_moveVector[3 * (size_t)m_dofVertex[d]];
            for (int a = 0; a < 3; ++a) x[a] += dp[d] * u[a];
        } else {
            x[m_dofAxis[d]] = dp[d] * m_axisScale[m_dofAxis[d]];
        }
    }
    if (m_cpStride <= 0) return t;
    const int64_t nv = (int64_t)vertexCount();
#pragma omp parallel for schedule(static) if (nv >= kParallelVertices)
    for (int64_t vi = 0; vi < nv; ++vi) {
        const size_t v = (size_t)vi;
        if (m_isCP[v]) continue;
        double d3[3] = {0.0, 0.0, 0.0};
        for (int64_t e = m_wOff[v]; e < m_wOff[v + 1]; ++e) {
            const size_t c = (size_t)m_wCp[e];
            for (int a = 0; a < 3; ++a) d3[a] += m_w[e] * t[3 * c + a];
        }
        if (m_cpInterp == ControlPointInterpolation::Displacement) projectFree(v, d3);
        for (int a = 0; a < 3; ++a) t[3 * v + a] = d3[a];
    }
    return t;
}

Vec LithologyModel::pullBackVertexGradient(const Vec& vg) const
{
    // dPhi/dx_cp = dPhi/dx_cp (direct) + sum over its dependants w * dPhi/dx_v (projected on
    // the dependant's free directions in displacement form), gathered per parameter through
    // the dependants' lists, so that parameters are independent of each other
    const int64_t nd = (int64_t)m_dofVertex.size();
    const bool cp = m_cpStride > 0;
    const bool displacement = m_cpInterp == ControlPointInterpolation::Displacement;
    Vec g((size_t)nd);
#pragma omp parallel for schedule(dynamic, 64) if (cp && (int64_t)vertexCount() >= kParallelVertices)
    for (int64_t d = 0; d < nd; ++d) {
        const size_t c = (size_t)m_dofVertex[(size_t)d];
        double ge[3] = {vg[3 * c], vg[3 * c + 1], vg[3 * c + 2]};
        if (cp)
            for (int64_t k = m_depOff[c]; k < m_depOff[c + 1]; ++k) {
                const size_t v = (size_t)m_depV[(size_t)k];
                double gv[3] = {vg[3 * v], vg[3 * v + 1], vg[3 * v + 2]};
                if (displacement) projectFree(v, gv);
                for (int a = 0; a < 3; ++a) ge[a] += m_depW[(size_t)k] * gv[a];
            }
        if (dofAlongVector((size_t)d)) {
            const double* u = &m_moveVector[3 * c];
            g[(size_t)d] = ge[0] * u[0] + ge[1] * u[1] + ge[2] * u[2];
        } else {
            g[(size_t)d] = ge[m_dofAxis[(size_t)d]] * m_axisScale[m_dofAxis[(size_t)d]];
        }
    }
    return g;
}

void LithologyModel::dofVertexMap(std::vector<int64_t>& off, std::vector<int32_t>& vert,
                                  std::vector<double>& w) const
{
    const size_t nd = m_dofVertex.size();
    for (size_t d = 0; d < nd; ++d)
        if (dofAlongVector(d))
            throw std::invalid_argument(
                "dofVertexMap: ALONG_VECTOR parameters move a vertex along a vector, not one axis");
    // dependants of each control-point vertex (transpose of the IDW lists)
    std::vector<std::vector<std::pair<int32_t, double>>> dep;
    if (m_cpStride > 0) {
        dep.resize(vertexCount());
        for (size_t v = 0; v < vertexCount(); ++v)
            for (int64_t e = m_wOff[v]; e < m_wOff[v + 1]; ++e)
                dep[(size_t)m_wCp[e]].push_back({(int32_t)v, m_w[e]});
    }
    const bool position = m_cpInterp == ControlPointInterpolation::Position;
    off.assign(nd + 1, 0);
    vert.clear();
    w.clear();
    for (size_t d = 0; d < nd; ++d) {
        const double s = m_axisScale[m_dofAxis[d]];
        vert.push_back(m_dofVertex[d]);
        w.push_back(s);
        if (m_cpStride > 0)
            for (auto& pr : dep[(size_t)m_dofVertex[d]]) {
                if (!position && !axisFree(m_freedom[(size_t)pr.first], m_dofAxis[d])) continue;
                vert.push_back(pr.first);
                w.push_back(pr.second * s);
            }
        off[d + 1] = (int64_t)vert.size();
    }
}

void LithologyModel::getBounds(const Vec& params, double margin_m, Vec& lower, Vec& upper) const
{
    lower.resize(params.size());
    upper.resize(params.size());
    for (size_t d = 0; d < params.size(); ++d) {
        if (dofAlongVector(d)) {
            // "basically unbounded" along a vector (surface_mesh.cpp:234-237)
            lower[d] = -1e6;
            upper[d] = 1e6;
            continue;
        }
        const double m = margin_m / m_axisScale[m_dofAxis[d]];
        lower[d] = params[d] - m;
        upper[d] = params[d] + m;
    }
}

void LithologyModel::getModelBounds(double marginMultiplier, Vec& lower, Vec& upper) const
{
    double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
    for (size_t v = 0; v < vertexCount(); ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], m_vertices[3 * v + a]);
            hi[a] = std::max(hi[a], m_vertices[3 * v + a]);
        }
    const double margin = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1.0}) * 0.01 *
                          marginMultiplier;
    lower.resize(m_dofVertex.size());
    upper.resize(m_dofVertex.size());
    for (size_t d = 0; d < m_dofVertex.size(); ++d) {
        const int a = m_dofAxis[d];
        if (dofAlongVector(d)) {
            lower[d] = -1e6;
            upper[d] = 1e6;
            continue;
        }
        lower[d] = (lo[a] - margin) / m_axisScale[a];
        upper[d] = (hi[a] + margin) / m_axisScale[a];
    }
}

// ================================================================== ForwardModel

Mat ForwardModel::computeJacobian(const Vec& params)
{
    const size_t nd = dataCount(), np = parameterCount();
    Mat J(nd, np);
    for (size_t j = 0; j < np; ++j) {
        Vec pp = params, pm = params;
        pp[j] += m_fdStep;
        pm[j] -= m_fdStep;
        const Vec fp = compute(pp), fm = compute(pm);
        for (size_t i = 0; i < nd; ++i) J(i, j) = (fp[i] - fm[i]) / (2.0 * m_fdStep);
    }
    compute(params);
    return J;
}

void ForwardModel::computeBoth(const Vec& params, Vec& predicted, Mat& jacobian)
{
    predicted = compute(params);
    jacobian = computeJacobian(params);
}

Vec ForwardModel::computeJTr(const Vec& params, const Vec& w)
{
    const Mat J = computeJacobian(params);
    Vec g(J.cols, 0.0);
    for (size_t j = 0; j < J.cols; ++j)
        for (size_t i = 0; i < J.rows; ++i) g[j] += J(i, j) * w[i];
    return g;
}

// ================================================================ GPU forwards

GpuForwardBase::GpuForwardBase(std::shared_ptr<LithologyModel> model,
                               std::vector<double> stations, int want, Precision prec,
                               const double fdir[3], int device)
    : m_model(std::move(model)), m_nData(stations.size() / 3), m_want(want),
      m_stations(std::move(stations)), m_prec(prec), m_device(device)
{
    for (int k = 0; k < 3; ++k) m_fdir[k] = fdir[k];
}

// The device context (model, facet records, stations resident on the GPU) is created on
// first use, so objects that are constructed but never evaluated hold no device memory.
lf_ctx* GpuForwardBase::ctx()
{
    if (m_ctx) return m_ctx;
    lf_desc d{};
    d.n_vertices = (int64_t)m_model->vertexCount();
    d.n_faces = (int64_t)m_model->faceCount();
    d.n_stations = (int64_t)m_nData;
    d.n_units = m_model->groupCount();
    d.precision = m_prec == Precision::FP32C ? LF_FP32C : LF_FP64;
    d.tmi_mode = m_textbook ? LF_TMI_TEXTBOOK : LF_TMI_REFERENCE;
    d.want = m_want;
    d.device = m_device;
    d.vertices = m_model->vertices().data();
    d.faces = m_model->faces().data();
    d.front = m_model->front().data();
    d.back = m_model->back().data();
    d.stations = m_stations.data();
    std::vector<double> dens, mag;
    buildProperties(dens, mag);
    d.density = dens.data();
    d.magnetization = mag.data();
    for (int k = 0; k < 3; ++k) d.field_dir[k] = m_fdir[k];
    const int rc = lf_create(&m_ctx, &d);
    if (rc != 0) {
        m_ctx = nullptr;
        throw std::runtime_error(std::string("lf_create: ") + lf_last_error(nullptr));
    }
    return m_ctx;
}

GpuForwardBase::~GpuForwardBase()
{
    if (m_ctx) lf_destroy(m_ctx);
}

void GpuForwardBase::pushProperties()
{
    if (!m_ctx) return;  // the context takes the current properties when it is created
    std::vector<double> dens, mag;
    buildProperties(dens, mag);
    check(lf_set_properties(m_ctx, dens.data(), mag.data()), m_ctx, "lf_set_properties");
}

void GpuForwardBase::upload(const Vec& params)
{
    m_model->applyParameterVector(params);
    lf_ctx* c = ctx();
    check(lf_set_vertices(c, m_model->vertices().data()), c, "lf_set_vertices");
}

void GpuForwardBase::syncGeometry()
{
    lf_ctx* c = ctx();
    check(lf_set_vertices(c, m_model->vertices().data()), c, "lf_set_vertices");
}

Vec GpuForwardBase::forwardRaw(int which)
{
    Vec out(m_nData);
    lf_ctx* c = ctx();
    check(lf_forward(c, which == 0 ? out.data() : nullptr, which == 1 ? out.data() : nullptr),
          c, "lf_forward");
    return out;
}

Vec GpuForwardBase::computeJTr(const Vec& params, const Vec& w)
{
    upload(params);
    Vec vg(3 * m_model->vertexCount());
    const bool gz = (m_want & LF_WANT_GZ) != 0;
    check(lf_adjoint(m_ctx, gz ? w.data() : nullptr, gz ? nullptr : w.data(), vg.data()),
          m_ctx, "lf_adjoint");  // upload() created the context
    return m_model->pullBackVertexGradient(vg);
}

Mat GpuForwardBase::computeJacobian(const Vec& params)
{
    const size_t nd = dataCount(), np = parameterCount();
    Mat J(nd, np);
    Vec e(nd, 0.0);
    for (size_t i = 0; i < nd; ++i) {
        e[i] = 1.0;
        const Vec row = computeJTr(params, e);
        e[i] = 0.0;
        for (size_t j = 0; j < np; ++j) J(i, j) = row[j];
    }
    return J;
}

GravityForward::GravityForward(std::shared_ptr<LithologyModel> model, const GravityData& data,
                               Precision prec, int device)
    : GpuForwardBase(model,
                     [&] {
                         std::vector<double> s;
                         for (auto& p : data) s.insert(s.end(), p.position, p.position + 3);
                         return s;
                     }(),
                     LF_WANT_GZ, prec, (const double[3]){0.0, 0.0, 1.0}, device)
{
    pushProperties();
}

void GravityForward::buildProperties(std::vector<double>& dens, std::vector<double>& mag) const
{
    const int ng = m_model->groupCount();
    dens.assign(ng, 0.0);
    mag.assign(3 * ng, 0.0);
    for (int g = 0; g < ng; ++g)
        dens[g] = (m_unitOnly < 0) ? m_model->group(g).density : (g == m_unitOnly ? 1.0 : 0.0);
}

Vec GravityForward::compute(const Vec& params)
{
    upload(params);
    return forwardRaw(0);
}

Vec GravityForward::computeGroupUnitResponse(int g)
{
    m_unitOnly = g;
    pushProperties();
    Vec r = forwardRaw(0);
    m_unitOnly = -1;
    pushProperties();
    return r;
}

MagneticForward::MagneticForward(std::shared_ptr<LithologyModel> model, const MagneticData& data,
                                 double inc_deg, double dec_deg, double field_nT, Precision prec,
                                 bool textbook, int device)
    : GpuForwardBase(model,
                     [&] {
                         std::vector<double> s;
                         for (auto& p : data) s.insert(s.end(), p.position, p.position + 3);
                         return s;
                     }(),
                     LF_WANT_TMI, prec,
                     [&] {
                         static thread_local double f[3];
                         fieldUnitVector(inc_deg, dec_deg, f);
                         return (const double*)f;
                     }(),
                     device),
      m_field(field_nT)
{
    fieldUnitVector(inc_deg, dec_deg, m_fhat);
    m_textbook = textbook;
    pushProperties();
}

void groupMagnetization(const LithoGroup& G, RemanentMagnetizationMode mode, double field_nT,
                        const double fhat[3], double& amp, double dir[3])
{
    // magnetic_forward.cpp:100-134
    const double induced = G.susceptibility * field_nT / 100.0;
    double M[3] = {induced * fhat[0], induced * fhat[1], induced * fhat[2]};
    if (mode == RemanentMagnetizationMode::FixedVectorPerGroup) {
        double R[3];
        fieldUnitVector(G.remanence_inclination, G.remanence_declination, R);
        for (int k = 0; k < 3; ++k) M[k] += G.remanence_magnitude * R[k];
    } else if (mode == RemanentMagnetizationMode::VectorPerGroup) {
        for (int k = 0; k < 3; ++k) M[k] += G.magnetization[k];
    }
    amp = std::sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
    for (int k = 0; k < 3; ++k) dir[k] = amp < kEps ? fhat[k] : M[k] / amp;
}

void MagneticForward::buildGroupMagnetization(int g, double& amp, double dir[3]) const
{
    groupMagnetization(m_model->group(g), m_mode, m_field, m_fhat, amp, dir);
}

void MagneticForward::buildProperties(std::vector<double>& dens, std::vector<double>& mag) const
{
    const int ng = m_model->groupCount();
    dens.assign(ng, 0.0);
    mag.assign(3 * ng, 0.0);
    for (int g = 0; g < ng; ++g) {
        double amp, dir[3];
        if (m_unitOnly >= 0) {
            // a unit response: only group m_unitOnly, with the requested magnetization
            amp = g == m_unitOnly ? m_unitAmp : 0.0;
            for (int k = 0; k < 3; ++k) dir[k] = m_unitDir[k];
        } else {
            buildGroupMagnetization(g, amp, dir);
        }
        for (int k = 0; k < 3; ++k) mag[3 * g + k] = amp * dir[k];
    }
}

Vec MagneticForward::compute(const Vec& params)
{
    upload(params);
    return forwardRaw(1);
}

Vec MagneticForward::unitResponse(int g, double amp, const double dir[3])
{
    if (g < 0 || g >= m_model->groupCount())
        throw std::out_of_range("MagneticForward: group index out of range");
    m_unitOnly = g;
    m_unitAmp = amp;
    for (int k = 0; k < 3; ++k) m_unitDir[k] = dir[k];
    try {
        pushProperties();
        Vec r = forwardRaw(1);
        m_unitOnly = -1;
        pushProperties();
        return r;
    } catch (...) {
        m_unitOnly = -1;
        throw;
    }
}

Vec MagneticForward::computeGroupUnitResponse(int g)
{
    // magneticClosedMesh(obs, mesh, 1.0, F-hat, F): unit susceptibility, induced
    // (magnetic_forward.cpp:148-162)
    return unitResponse(g, m_field / 100.0, m_fhat);
}

Vec MagneticForward::computeGroupUnitResponseDirection(int g, const double magDirection[3])
{
    // magneticClosedMeshVector(obs, mesh, 1.0, magDirection, F-hat)
    // (magnetic_forward.cpp:164-178)
    return unitResponse(g, 1.0, magDirection);
}

Vec MagneticForward::computeGroupUnitRemanenceResponse(int g)
{
    // magnetic_forward.cpp:180-190
    if (g < 0 || g >= m_model->groupCount())
        throw std::out_of_range("MagneticForward: group index out of range");
    const LithoGroup& G = m_model->group(g);
    double R[3];
    fieldUnitVector(G.remanence_inclination, G.remanence_declination, R);
    return computeGroupUnitResponseDirection(g, R);
}

// ============================================================= regularization

double ReferenceModelRegularization::evaluate(const Vec& params)
{
    if (!m_captured) return 0.0;
    double s = 0.0;
    for (size_t i = 0; i < params.size(); ++i) s += (params[i] - m_ref[i]) * (params[i] - m_ref[i]);
    return 0.5 * m_weight * s;
}

Vec ReferenceModelRegularization::gradient(const Vec& params)
{
    Vec g(params.size(), 0.0);
    if (!m_captured) return g;
    for (size_t i = 0; i < params.size(); ++i) g[i] = m_weight * (params[i] - m_ref[i]);
    return g;
}

// ================================================================ smoothness

SurfaceSmoothness::SurfaceSmoothness(std::shared_ptr<LithologyModel> model)
    : m_model(std::move(model))
{
    const int nu = m_model->groupCount();
    const auto& F = m_model->faces();
    const auto& fr = m_model->front();
    const auto& bk = m_model->back();
    std::vector<std::vector<std::pair<int32_t, int32_t>>> edges(nu);
    for (size_t f = 0; f < fr.size(); ++f) {
        const int32_t a = (int32_t)F[3 * f], b = (int32_t)F[3 * f + 1], c = (int32_t)F[3 * f + 2];
        for (int32_t u : {fr[f], bk[f]}) {
            if (u < 0 || u >= nu) continue;
            auto& E = edges[u];
            E.insert(E.end(), {{a, b}, {a, c}, {b, a}, {b, c}, {c, a}, {c, b}});
        }
    }
    m_unitVerts.resize(nu);
    m_nbrOff.resize(nu);
    m_nbr.resize(nu);
    for (int u = 0; u < nu; ++u) {
        auto& E = edges[u];
        std::sort(E.begin(), E.end());
        E.erase(std::unique(E.begin(), E.end()), E.end());
        auto& V = m_unitVerts[u];
        auto& off = m_nbrOff[u];
        auto& nb = m_nbr[u];
        for (size_t i = 0; i < E.size(); ++i) {
            if (V.empty() || V.back() != E[i].first) {
                V.push_back(E[i].first);
                off.push_back((int64_t)nb.size());
            }
            nb.push_back(E[i].second);
        }
        off.push_back((int64_t)nb.size());
    }
}

double SurfaceSmoothness::penalty(std::vector<double>* vgrad) const
{
    const auto& X = m_model->vertices();
    if (vgrad) vgrad->assign(X.size(), 0.0);
    double pen = 0.0;
    for (size_t u = 0; u < m_unitVerts.size(); ++u) {
        const auto& V = m_unitVerts[u];
        const auto& off = m_nbrOff[u];
        const auto& nb = m_nbr[u];
        for (size_t i = 0; i < V.size(); ++i) {
            const int32_t v = V[i];
            const VertexFreedom fv = m_model->vertexFreedom(v);
            if (fv == VertexFreedom::FIXED) continue;
            double mean[3] = {0.0, 0.0, 0.0};
            int n = 0;
            for (int64_t e = off[i]; e < off[i + 1]; ++e)
                if (m_model->vertexFreedom(nb[e]) == fv) {
                    for (int a = 0; a < 3; ++a) mean[a] += X[3 * (size_t)nb[e] + a];
                    ++n;
                }
            if (n == 0) continue;
            double d[3];
            for (int a = 0; a < 3; ++a) d[a] = X[3 * (size_t)v + a] - mean[a] / n;
            // the free axes' components, or the part along the move vector
            // (smoothness.cpp:54-98)
            m_model->projectFree((size_t)v, d);
            for (int a = 0; a < 3; ++a) pen += 0.5 * d[a] * d[a];
            if (!vgrad) continue;
            // own term d, and -d/n to each same-freedom neighbour
            for (int a = 0; a < 3; ++a) (*vgrad)[3 * (size_t)v + a] += d[a];
            for (int64_t e = off[i]; e < off[i + 1]; ++e)
                if (m_model->vertexFreedom(nb[e]) == fv)
                    for (int a = 0; a < 3; ++a) (*vgrad)[3 * (size_t)nb[e] + a] -= d[a] / n;
        }
    }
    return pen;
}

double SurfaceSmoothness::evaluate(const Vec& params)
{
    m_model->applyParameterVector(params);
    return m_weight * penalty(nullptr);
}

Vec SurfaceSmoothness::gradient(const Vec& params)
{
    m_model->applyParameterVector(params);
    std::vector<double> vg;
    penalty(&vg);
    Vec g = m_model->pullBackVertexGradient(vg);
    for (auto& x : g) x *= m_weight;
    return g;
}

// ================================================================ boreholes

BoreholeContactTerm::BoreholeContactTerm(std::shared_ptr<LithologyModel> model,
                                         std::vector<ContactPick> picks, double searchRadius)
    : m_model(std::move(model)), m_picks(std::move(picks))
{
    const auto& X = m_model->vertices();
    const auto& F = m_model->faces();
    const auto& fr = m_model->front();
    const auto& bk = m_model->back();
    m_cand.resize(m_picks.size());
    for (const ContactPick& p : m_picks)
        if (!(p.sigma > 0.0)) throw std::invalid_argument("BoreholeContactTerm: sigma must be > 0");
    const size_t nf = fr.size();
    if (nf == 0 || m_picks.empty()) return;
    // the facets of any picked unit pair, binned in (x, y) by their bounding boxes; a
    // pick's candidates are the facets of its pair whose box grown by the search radius
    // holds the pick (the bins only narrow the search)
    std::set<std::pair<int, int>> pairs;
    for (const ContactPick& p : m_picks) pairs.insert({std::min(p.above, p.below), std::max(p.above, p.below)});
    std::vector<int64_t> sel;
    std::vector<double> box;  // lo x, lo y, hi x, hi y per selected facet
    double ext = 0.0, x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
    for (size_t f = 0; f < nf; ++f) {
        if (!pairs.count({std::min(fr[f], bk[f]), std::max(fr[f], bk[f])})) continue;
        double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 2; ++a) {
                const double v = X[3 * (size_t)F[3 * f + c] + a];
                lo[a] = std::min(lo[a], v);
                hi[a] = std::max(hi[a], v);
            }
        sel.push_back((int64_t)f);
        box.insert(box.end(), {lo[0], lo[1], hi[0], hi[1]});
        ext += (hi[0] - lo[0]) + (hi[1] - lo[1]);
        x0 = std::min(x0, lo[0]);
        y0 = std::min(y0, lo[1]);
        x1 = std::max(x1, hi[0]);
        y1 = std::max(y1, hi[1]);
    }
    if (sel.empty()) return;
    const double h = std::max({ext / (double)sel.size(), (x1 - x0) / 2048.0, (y1 - y0) / 2048.0, 1e-6});
    const int64_t nx = (int64_t)std::floor((x1 - x0) / h) + 1, ny = (int64_t)std::floor((y1 - y0) / h) + 1;
    auto bx = [&](double x) { return std::clamp((int64_t)std::floor((x - x0) / h), (int64_t)0, nx - 1); };
    auto by = [&](double y) { return std::clamp((int64_t)std::floor((y - y0) / h), (int64_t)0, ny - 1); };
    std::vector<int64_t> off((size_t)(nx * ny) + 1, 0);
    for (size_t i = 0; i < sel.size(); ++i)
        for (int64_t ix = bx(box[4 * i]); ix <= bx(box[4 * i + 2]); ++ix)
            for (int64_t iy = by(box[4 * i + 1]); iy <= by(box[4 * i + 3]); ++iy) ++off[(size_t)(ix * ny + iy) + 1];
    for (size_t i = 1; i < off.size(); ++i) off[i] += off[i - 1];
    std::vector<int32_t> ent((size_t)off.back());
    {
        std::vector<int64_t> fill(off.begin(), off.end() - 1);
        for (size_t i = 0; i < sel.size(); ++i)
            for (int64_t ix = bx(box[4 * i]); ix <= bx(box[4 * i + 2]); ++ix)
                for (int64_t iy = by(box[4 * i + 1]); iy <= by(box[4 * i + 3]); ++iy)
                    ent[(size_t)fill[(size_t)(ix * ny + iy)]++] = (int32_t)i;
    }
    for (size_t k = 0; k < m_picks.size(); ++k) {
        const ContactPick& p = m_picks[k];
        const int ua = std::min(p.above, p.below), ub = std::max(p.above, p.below);
        auto& c = m_cand[k];
        for (int64_t ix = bx(p.x - searchRadius); ix <= bx(p.x + searchRadius); ++ix)
            for (int64_t iy = by(p.y - searchRadius); iy <= by(p.y + searchRadius); ++iy)
                for (int64_t e = off[(size_t)(ix * ny + iy)]; e < off[(size_t)(ix * ny + iy) + 1]; ++e) {
                    const size_t i = (size_t)ent[(size_t)e];
                    const int64_t f = sel[i];
                    if (std::min(fr[f], bk[f]) != ua || std::max(fr[f], bk[f]) != ub) continue;
                    if (p.x >= box[4 * i] - searchRadius && p.x <= box[4 * i + 2] + searchRadius &&
                        p.y >= box[4 * i + 1] - searchRadius && p.y <= box[4 * i + 3] + searchRadius)
                        c.push_back(f);
                }
        // in facet order, once each (a facet spans several bins)
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
    }
}

double BoreholeContactTerm::misfit(std::vector<double>* vgrad, std::vector<double>* zmodel) const
{
    const auto& X = m_model->vertices();
    const auto& F = m_model->faces();
    if (vgrad) vgrad->assign(X.size(), 0.0);
    if (zmodel) zmodel->assign(m_picks.size(), std::numeric_limits<double>::quiet_NaN());
    double total = 0.0;
    for (size_t k = 0; k < m_picks.size(); ++k) {
        const ContactPick& p = m_picks[k];
        double best = std::numeric_limits<double>::infinity(), zb = 0.0;
        int64_t fb = -1;
        double lb[3] = {0, 0, 0}, sxb = 0, syb = 0;
        for (int64_t f : m_cand[k]) {
            const double* P0 = &X[3 * (size_t)F[3 * f]];
            const double* P1 = &X[3 * (size_t)F[3 * f + 1]];
            const double* P2 = &X[3 * (size_t)F[3 * f + 2]];
            const double x10 = P1[0] - P0[0], y10 = P1[1] - P0[1];
            const double x20 = P2[0] - P0[0], y20 = P2[1] - P0[1];
            const double d = x10 * y20 - x20 * y10;  // twice the signed horizontal area
            const double scale = std::abs(x10 * y20) + std::abs(x20 * y10);
            if (!(std::abs(d) > 1e-12 * scale)) continue;  // vertical facet
            const double qx = p.x - P0[0], qy = p.y - P0[1];
            const double l1 = (qx * y20 - x20 * qy) / d;
            const double l2 = (x10 * qy - qx * y10) / d;
            const double l0 = 1.0 - l1 - l2;
            const double tol = -1e-12;
            if (l0 < tol || l1 < tol || l2 < tol) continue;
            const double z = l0 * P0[2] + l1 * P1[2] + l2 * P2[2];
            if (std::abs(z - p.z) < best) {
                best = std::abs(z - p.z);
                zb = z;
                fb = f;
                lb[0] = l0;
                lb[1] = l1;
                lb[2] = l2;
                // slope of the facet plane, dz/dx and dz/dy at fixed vertices
                const double z10 = P1[2] - P0[2], z20 = P2[2] - P0[2];
                sxb = (z10 * y20 - z20 * y10) / d;
                syb = (x10 * z20 - x20 * z10) / d;
            }
        }
        if (fb < 0) continue;
        if (zmodel) (*zmodel)[k] = zb;
        const double r = (zb - p.z) / p.sigma;
        total += 0.5 * r * r;
        if (!vgrad) continue;
        // z = sum_j l_j z_j; moving vertex j horizontally by delta at fixed (x, y) shifts
        // the barycentrics like moving the query point by -l_j delta
        const double s = r / p.sigma;
        for (int c = 0; c < 3; ++c) {
            const size_t v = (size_t)F[3 * fb + c];
            (*vgrad)[3 * v] += s * (-lb[c] * sxb);
            (*vgrad)[3 * v + 1] += s * (-lb[c] * syb);
            (*vgrad)[3 * v + 2] += s * lb[c];
        }
    }
    return total;
}

double BoreholeContactTerm::evaluate(const Vec& params)
{
    m_model->applyParameterVector(params);
    return m_weight * misfit(nullptr, nullptr);
}

Vec BoreholeContactTerm::gradient(const Vec& params)
{
    m_model->applyParameterVector(params);
    std::vector<double> vg;
    misfit(&vg, nullptr);
    Vec g = m_model->pullBackVertexGradient(vg);
    for (auto& x : g) x *= m_weight;
    return g;
}

// ================================================================== objectives

ObjectiveFunction::ObjectiveFunction(std::shared_ptr<ForwardModel> forward, const GravityData& data)
    : m_forward(std::move(forward)), m_data(data)
{
}

double ObjectiveFunction::dataMisfit(const Vec& params)
{
    const Vec pred = m_forward->compute(params);
    double mis = 0.0;
    for (size_t i = 0; i < m_data.size(); ++i) {
        const double w = m_data[i].g_std > 0 ? 1.0 / m_data[i].g_std : 1.0;
        const double r = w * (m_data[i].g_obs - pred[i]);
        mis += 0.5 * r * r;
    }
    return mis;
}

Vec ObjectiveFunction::residuals(const Vec& params)
{
    const Vec pred = m_forward->compute(params);
    Vec r(m_data.size());
    for (size_t i = 0; i < m_data.size(); ++i) r[i] = m_data[i].g_obs - pred[i];
    return r;
}

Vec ObjectiveFunction::dataMisfitGradient(const Vec& params)
{
    // -J^T W^2 r  (objective.cpp:42-55), matrix-free
    const Vec pred = m_forward->compute(params);
    Vec w(m_data.size());
    for (size_t i = 0; i < m_data.size(); ++i) {
        const double wi = m_data[i].g_std > 0 ? 1.0 / m_data[i].g_std : 1.0;
        w[i] = -wi * wi * (m_data[i].g_obs - pred[i]);
    }
    return m_forward->computeJTr(params, w);
}

Vec ObjectiveFunction::gradient(const Vec& params)
{
    Vec g = dataMisfitGradient(params);
    for (auto& r : m_regs) {
        const Vec gr = r->gradient(params);
        for (size_t i = 0; i < g.size(); ++i) g[i] += gr[i];
    }
    return g;
}

ObjectiveFunction::Components ObjectiveFunction::evaluateComponents(const Vec& params)
{
    Components c;
    c.dataMisfit = dataMisfit(params);
    for (auto& r : m_regs) c.regularization += r->evaluate(params);
    c.total = c.dataMisfit + c.regularization + c.constraintPenalty;
    return c;
}

void JointObjective::setMagnetic(std::shared_ptr<ForwardModel> fwd, const MagneticData& data,
                                 double alpha)
{
    m_mag = std::move(fwd);
    m_magData = &data;
    m_alpha = alpha;
}

Vec JointObjective::gradient(const Vec& params)
{
    Vec g = m_grav->gradient(params);
    if (m_mag && m_magData && !m_magData->empty() && m_alpha > 0.0) {
        // -alpha J_mag^T W^2 r  (joint_objective.cpp:47-58)
        const Vec pred = m_mag->compute(params);
        Vec w(m_magData->size());
        for (size_t i = 0; i < w.size(); ++i) {
            const auto& d = (*m_magData)[i];
            const double wi = d.t_std > 0 ? 1.0 / d.t_std : 1.0;
            w[i] = -m_alpha * wi * wi * (d.t_obs - pred[i]);
        }
        const Vec gm = m_mag->computeJTr(params, w);
        for (size_t i = 0; i < g.size(); ++i) g[i] += gm[i];
    }
    return g;
}

JointObjective::JointComponents JointObjective::evaluateComponents(const Vec& params)
{
    JointComponents c;
    const auto gc = m_grav->evaluateComponents(params);
    c.gravityMisfit = gc.dataMisfit;
    c.regularization = gc.regularization;
    c.constraintPenalty = gc.constraintPenalty;
    if (m_mag && m_magData && !m_magData->empty() && m_alpha > 0.0) {
        const Vec pred = m_mag->compute(params);
        double mis = 0.0;
        for (size_t i = 0; i < pred.size(); ++i) {
            const auto& d = (*m_magData)[i];
            const double wi = d.t_std > 0 ? 1.0 / d.t_std : 1.0;
            const double r = wi * (d.t_obs - pred[i]);
            mis += 0.5 * r * r;
        }
        c.magneticMisfit = mis;
    }
    c.total = c.gravityMisfit + m_alpha * c.magneticMisfit + c.regularization + c.constraintPenalty;
    return c;
}

// ======================================================== fused joint objective

FusedJointObjective::FusedJointObjective(std::shared_ptr<LithologyModel> model,
                                         const GravityData& gdata, const MagneticData& mdata,
                                         double inc_deg, double dec_deg, double field_nT,
                                         double alpha_mag, Precision prec, int device)
    : m_model(std::move(model)), m_g(gdata), m_m(mdata), m_alpha(alpha_mag)
{
    if (gdata.size() != mdata.size())
        throw std::invalid_argument("FusedJointObjective: gravity and magnetic stations differ");
    for (size_t i = 0; i < gdata.size(); ++i)
        for (int k = 0; k < 3; ++k)
            if (gdata[i].position[k] != mdata[i].position[k])
                throw std::invalid_argument("FusedJointObjective: gravity and magnetic stations differ");
    double fh[3];
    fieldUnitVector(inc_deg, dec_deg, fh);
    m_field = field_nT;
    for (int k = 0; k < 3; ++k) m_fhat[k] = fh[k];
    std::vector<double> st;
    for (auto& p : gdata) st.insert(st.end(), p.position, p.position + 3);
    const int ng = m_model->groupCount();
    std::vector<double> dens(ng), mag(3 * ng);
    for (int g = 0; g < ng; ++g) {
        dens[g] = m_model->group(g).density;
        double amp, dir[3];
        groupMagnetization(m_model->group(g), m_mode, m_field, m_fhat, amp, dir);
        for (int k = 0; k < 3; ++k) mag[3 * g + k] = amp * dir[k];
    }
    lf_desc d{};
    d.n_vertices = (int64_t)m_model->vertexCount();
    d.n_faces = (int64_t)m_model->faceCount();
    d.n_stations = (int64_t)gdata.size();
    d.n_units = ng;
    d.precision = prec == Precision::FP32C ? LF_FP32C : LF_FP64;
    d.tmi_mode = LF_TMI_REFERENCE;
    d.want = LF_WANT_GZ | LF_WANT_TMI;
    d.device = device;
    d.vertices = m_model->vertices().data();
    d.faces = m_model->faces().data();
    d.front = m_model->front().data();
    d.back = m_model->back().data();
    d.stations = st.data();
    d.density = dens.data();
    d.magnetization = mag.data();
    for (int k = 0; k < 3; ++k) d.field_dir[k] = fh[k];
    if (lf_create(&m_ctx, &d) != 0)
        throw std::runtime_error(std::string("lf_create: ") + lf_last_error(nullptr));
}

FusedJointObjective::~FusedJointObjective()
{
    if (m_ctx) lf_destroy(m_ctx);
}

void FusedJointObjective::refreshProperties()
{
    const int ng = m_model->groupCount();
    std::vector<double> dens(ng), mag(3 * ng);
    for (int g = 0; g < ng; ++g) {
        dens[g] = m_model->group(g).density;
        double amp, dir[3];
        groupMagnetization(m_model->group(g), m_mode, m_field, m_fhat, amp, dir);
        for (int k = 0; k < 3; ++k) mag[3 * g + k] = amp * dir[k];
    }
    check(lf_set_properties(m_ctx, dens.data(), mag.data()), m_ctx, "lf_set_properties");
    m_cacheValid = false;
}

Vec FusedJointObjective::jtw(const Vec& params, const Vec& wg, const Vec& wt)
{
    m_model->applyParameterVector(params);
    check(lf_set_vertices(m_ctx, m_model->vertices().data()), m_ctx, "lf_set_vertices");
    m_cacheValid = false;
    std::vector<double> vg(3 * m_model->vertexCount());
    check(lf_adjoint(m_ctx, wg.empty() ? nullptr : wg.data(), wt.empty() ? nullptr : wt.data(),
                     vg.data()),
          m_ctx, "lf_adjoint");
    ++m_nAdjoint;
    return m_model->pullBackVertexGradient(vg);
}

void FusedJointObjective::jvp(const Vec& params, const Vec& dp, Vec& gz, Vec& tmi)
{
    m_model->applyParameterVector(params);
    check(lf_set_vertices(m_ctx, m_model->vertices().data()), m_ctx, "lf_set_vertices");
    m_cacheValid = false;
    const Vec vt = m_model->vertexTangent(dp);
    gz.assign(m_g.size(), 0.0);
    tmi.assign(m_m.size(), 0.0);
    check(lf_jvp(m_ctx, vt.data(), gz.data(), tmi.data()), m_ctx, "lf_jvp");
}

CalibrationResult calibrateMagneticWeight(std::shared_ptr<LithologyModel> model,
                                          const GravityData& gdata, const MagneticData& mdata,
                                          double inc_deg, double dec_deg, double field_nT,
                                          const Vec& params, double alpha_mag, uint64_t seed,
                                          double eps, int device, RemanentMagnetizationMode mode,
                                          Precision prec)
{
    CalibrationResult out;
    out.alpha = alpha_mag;
    if (mdata.empty() || alpha_mag <= 0.0) return out;
    const bool exact = !(eps > 0.0);
    FusedJointObjective f(model, gdata, mdata, inc_deg, dec_deg, field_nT, alpha_mag,
                          exact ? prec : Precision::FP64, device);
    if (mode != RemanentMagnetizationMode::EffectiveSusceptibility) f.setRemanenceMode(mode);
    const size_t np = params.size(), ns = gdata.size();
    Vec wgt_g(ns), wgt_m(ns);
    for (size_t i = 0; i < ns; ++i) {
        wgt_g[i] = gdata[i].g_std > 0 ? 1.0 / (gdata[i].g_std * gdata[i].g_std) : 1.0;
        wgt_m[i] = mdata[i].t_std > 0 ? 1.0 / (mdata[i].t_std * mdata[i].t_std) : 1.0;
    }
    // W^2 (J v) for one data type, J v exact or by central differences along v
    auto wjv = [&](const Vec& v, bool mag) {
        Vec w(ns);
        if (exact) {
            Vec jg, jt;
            f.jvp(params, v, jg, jt);
            for (size_t i = 0; i < ns; ++i) w[i] = mag ? jt[i] * wgt_m[i] : jg[i] * wgt_g[i];
            return w;
        }
        Vec pp(np), pm(np), gp, tp, gm, tm;
        for (size_t j = 0; j < np; ++j) {
            pp[j] = params[j] + eps * v[j];
            pm[j] = params[j] - eps * v[j];
        }
        f.forward(pp, gp, tp);
        f.forward(pm, gm, tm);
        for (size_t i = 0; i < ns; ++i)
            w[i] = (mag ? (tp[i] - tm[i]) * wgt_m[i] : (gp[i] - gm[i]) * wgt_g[i]) / (2.0 * eps);
        return w;
    };
    auto jt = [&](const Vec& w, bool mag) {
        return mag ? f.jtw(params, Vec(), w) : f.jtw(params, w, Vec());
    };
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> uni(-1.0, 1.0);
    auto power = [&](bool mag, int& iters) {
        Vec v(np);
        for (auto& x : v) x = uni(rng);
        double nv = std::sqrt(dot(v, v));
        for (auto& x : v) x /= nv;
        double lamOld = 0.0;
        iters = 0;
        for (int it = 0; it < 50; ++it) {
            iters = it + 1;
            v = jt(wjv(v, mag), mag);
            const double lam = std::sqrt(dot(v, v));
            if (lam < 1e-60) return 0.0;
            for (auto& x : v) x /= lam;
            if (std::abs(lam - lamOld) / lam < 1e-8) break;
            lamOld = lam;
        }
        const Vec hv = jt(wjv(v, mag), mag);
        return std::max(dot(v, hv), 0.0);
    };
    out.eigGravity = power(false, out.iterationsGravity);
    out.eigMagnetic = power(true, out.iterationsMagnetic);
    out.ratio = out.eigMagnetic > 1e-60 ? out.eigGravity / out.eigMagnetic : 0.0;
    if (std::isfinite(out.ratio) && out.ratio > 0.0) {
        out.alpha = alpha_mag * std::max(1e-6, std::min(out.ratio, 1e6));
    } else {
        // gradient-norm ratio of the weighted data misfits
        Vec gz, tmi;
        f.forward(params, gz, tmi);
        Vec rg(ns), rm(ns);
        for (size_t i = 0; i < ns; ++i) {
            rg[i] = wgt_g[i] * (gdata[i].g_obs - gz[i]);
            rm[i] = wgt_m[i] * (mdata[i].t_obs - tmi[i]);
        }
        const Vec gg = jt(rg, false), gm = jt(rm, true);
        const double ng = std::sqrt(dot(gg, gg)), nm = std::sqrt(dot(gm, gm));
        out.fallback = true;
        if (ng >= 1e-60 && nm > 1e-60) out.alpha = alpha_mag * ng / nm;
    }
    model->applyParameterVector(params);
    return out;
}

namespace {
// min over p in [lo, hi] of 0.5 |w (obs - U p)|^2 / s^2 + 0.5 lambda |p - p0|^2 with the
// reference's optimizer settings; U column-major (n_obs x n_p)
Vec bounded_unit_fit(const std::vector<double>& U, size_t nObs, const Vec& obs, const Vec& sd,
                     const Vec& p0, const Vec& lo, const Vec& hi, const PropertyInversionOptions& o)
{
    const size_t np = p0.size();
    double var = 0.0;
    for (double v : obs) var += v * v;
    var /= std::max<size_t>(nObs, 1);
    const double s2 = std::max(var, 1.0);  // rms floored at 1
    auto pred = [&](const Vec& p) {
        Vec r(nObs, 0.0);
        for (size_t j = 0; j < np; ++j)
            for (size_t i = 0; i < nObs; ++i) r[i] += U[j * nObs + i] * p[j];
        return r;
    };
    auto wt = [&](size_t i) { return sd[i] > 0.0 ? 1.0 / sd[i] : 1.0; };
    auto f = [&](const Vec& p) {
        const Vec q = pred(p);
        double m = 0.0;
        for (size_t i = 0; i < nObs; ++i) {
            const double r = wt(i) * (obs[i] - q[i]);
            m += 0.5 * r * r;
        }
        m /= s2;
        if (o.damping > 0.0)
            for (size_t j = 0; j < np; ++j) m += 0.5 * o.damping * (p[j] - p0[j]) * (p[j] - p0[j]);
        return m;
    };
    auto g = [&](const Vec& p) {
        const Vec q = pred(p);
        Vec gr(np, 0.0);
        for (size_t i = 0; i < nObs; ++i) {
            const double w = wt(i);
            const double wr = w * w * (obs[i] - q[i]) / s2;
            for (size_t j = 0; j < np; ++j) gr[j] -= U[j * nObs + i] * wr;
        }
        if (o.damping > 0.0)
            for (size_t j = 0; j < np; ++j) gr[j] += o.damping * (p[j] - p0[j]);
        return gr;
    };
    LBFGSBOptimizer opt;
    opt.setMaxIterations(o.maxIter);
    opt.setHistorySize(std::min<int>(5, (int)np));
    opt.setTolerance(o.tolerance);
    return opt.minimize(f, g, p0, lo, hi).params;
}
}  // namespace

Vec invertUnitDensities(GravityForward& fwd, const GravityData& data, LithologyModel& model,
                        const PropertyInversionOptions& o)
{
    const int ng = model.groupCount();
    const size_t n = data.size();
    std::vector<double> U((size_t)ng * n);
    fwd.syncGeometry();  // the responses of the current geometry
    for (int g = 0; g < ng; ++g) {
        const Vec col = fwd.computeGroupUnitResponse(g);
        std::copy(col.begin(), col.end(), U.begin() + (size_t)g * n);
    }
    Vec obs(n), sd(n), p0(ng);
    for (size_t i = 0; i < n; ++i) {
        obs[i] = data[i].g_obs;
        sd[i] = data[i].g_std;
    }
    for (int g = 0; g < ng; ++g) p0[g] = model.group(g).density;
    const Vec p = bounded_unit_fit(U, n, obs, sd, p0, Vec(ng, o.lower), Vec(ng, o.upper), o);
    for (int g = 0; g < ng; ++g) model.group(g).density = p[g];
    fwd.refreshProperties();
    return p;
}

Vec invertUnitSusceptibilities(MagneticForward& fwd, const MagneticData& data,
                               LithologyModel& model, const PropertyInversionOptions& o)
{
    // runner.cpp:801-935: chi per unit, plus the remanent parameters of the mode
    const RemanentMagnetizationMode mode = fwd.remanenceMode();
    const int per = mode == RemanentMagnetizationMode::FixedVectorPerGroup ? 2
                    : mode == RemanentMagnetizationMode::VectorPerGroup    ? 4
                                                                           : 1;
    const int ng = model.groupCount();
    const size_t n = data.size(), np = (size_t)per * ng;
    std::vector<double> U(np * n);
    fwd.syncGeometry();  // the responses of the current geometry
    auto put = [&](size_t col, const Vec& r) { std::copy(r.begin(), r.end(), U.begin() + col * n); };
    const double axes[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    Vec p0(np), lo(np), hi(np);
    for (int g = 0; g < ng; ++g) {
        const size_t c = (size_t)per * g;
        const LithoGroup& G = model.group(g);
        put(c, fwd.computeGroupUnitResponse(g));
        p0[c] = G.susceptibility;
        lo[c] = o.lower;
        hi[c] = o.upper;
        if (per == 2) {
            put(c + 1, fwd.computeGroupUnitRemanenceResponse(g));
            p0[c + 1] = G.remanence_magnitude;
            lo[c + 1] = o.remanenceLower;
            hi[c + 1] = o.remanenceUpper;
        } else if (per == 4) {
            for (int k = 0; k < 3; ++k) {
                put(c + 1 + k, fwd.computeGroupUnitResponseDirection(g, axes[k]));
                p0[c + 1 + k] = G.magnetization[k];
                lo[c + 1 + k] = o.componentLower;
                hi[c + 1 + k] = o.componentUpper;
            }
        }
    }
    Vec obs(n), sd(n);
    for (size_t i = 0; i < n; ++i) {
        obs[i] = data[i].t_obs;
        sd[i] = data[i].t_std;
    }
    const Vec p = bounded_unit_fit(U, n, obs, sd, p0, lo, hi, o);
    for (int g = 0; g < ng; ++g) {
        const size_t c = (size_t)per * g;
        LithoGroup& G = model.group(g);
        G.susceptibility = p[c];
        if (per == 2) G.remanence_magnitude = p[c + 1];
        if (per == 4)
            for (int k = 0; k < 3; ++k) G.magnetization[k] = p[c + 1 + k];
    }
    fwd.refreshProperties();
    return p;
}

void FusedJointObjective::ensureForward(const Vec& params)
{
    if (m_cacheValid && params == m_cacheParams) return;
    m_model->applyParameterVector(params);
    check(lf_set_vertices(m_ctx, m_model->vertices().data()), m_ctx, "lf_set_vertices");
    m_gz.assign(m_g.size(), 0.0);
    m_tmi.assign(m_m.size(), 0.0);
    check(lf_forward(m_ctx, m_gz.data(), m_tmi.data()), m_ctx, "lf_forward");
    ++m_nForward;
    m_cacheParams = params;
    m_cacheValid = true;
}

void FusedJointObjective::forward(const Vec& params, Vec& gz, Vec& tmi)
{
    ensureForward(params);
    gz = m_gz;
    tmi = m_tmi;
}

JointObjective::JointComponents FusedJointObjective::evaluateComponents(const Vec& params)
{
    ensureForward(params);
    JointObjective::JointComponents c;
    for (size_t i = 0; i < m_g.size(); ++i) {
        const double w = m_g[i].g_std > 0 ? 1.0 / m_g[i].g_std : 1.0;
        const double r = w * (m_g[i].g_obs - m_gz[i]);
        c.gravityMisfit += 0.5 * r * r;
    }
    if (m_alpha > 0.0)
        for (size_t i = 0; i < m_m.size(); ++i) {
            const double w = m_m[i].t_std > 0 ? 1.0 / m_m[i].t_std : 1.0;
            const double r = w * (m_m[i].t_obs - m_tmi[i]);
            c.magneticMisfit += 0.5 * r * r;
        }
    for (auto& r : m_regs) c.regularization += r->evaluate(params);
    c.total = c.gravityMisfit + m_alpha * c.magneticMisfit + c.regularization;
    return c;
}

Vec FusedJointObjective::gradient(const Vec& params)
{
    ensureForward(params);
    const size_t n = m_g.size();
    Vec wg(n), wt(n, 0.0);
    for (size_t i = 0; i < n; ++i) {
        const double a = m_g[i].g_std > 0 ? 1.0 / m_g[i].g_std : 1.0;
        wg[i] = -a * a * (m_g[i].g_obs - m_gz[i]);
        if (m_alpha > 0.0) {
            const double b = m_m[i].t_std > 0 ? 1.0 / m_m[i].t_std : 1.0;
            wt[i] = -m_alpha * b * b * (m_m[i].t_obs - m_tmi[i]);
        }
    }
    Vec vg(3 * m_model->vertexCount());
    check(lf_adjoint(m_ctx, wg.data(), wt.data(), vg.data()), m_ctx, "lf_adjoint");
    ++m_nAdjoint;
    Vec g = m_model->pullBackVertexGradient(vg);
    for (auto& r : m_regs) {
        const Vec gr = r->gradient(params);
        for (size_t i = 0; i < g.size(); ++i) g[i] += gr[i];
    }
    return g;
}

// =================================================================== L-BFGS-B

namespace {
Vec project(const Vec& x, const Vec& lb, const Vec& ub)
{
    Vec p(x.size());
    for (size_t i = 0; i < x.size(); ++i) p[i] = std::max(lb[i], std::min(ub[i], x[i]));
    return p;
}
}  // namespace

Vec LBFGSBOptimizer::direction(const Vec& gin, const std::vector<uint8_t>& freeMask) const
{
    // lbfgsb_optimizer.cpp:33-72 with H0 = gamma * D
    const size_t n = gin.size();
    const bool diag = m_h0.size() == n;
    Vec g = gin;
    if (!freeMask.empty())
        for (size_t i = 0; i < n; ++i)
            if (!freeMask[i]) g[i] = 0.0;
    if (m_s.empty()) {
        Vec d(n);
        for (size_t i = 0; i < n; ++i) d[i] = -(diag ? m_h0[i] : 1.0) * g[i];
        return d;
    }
    const Vec& sl = m_s.back();
    const Vec& yl = m_y.back();
    double gamma;
    if (diag) {
        double yDy = 0.0;
        for (size_t i = 0; i < n; ++i) yDy += yl[i] * m_h0[i] * yl[i];
        gamma = yDy > std::numeric_limits<double>::epsilon() ? dot(sl, yl) / yDy : 1.0;
    } else {
        const double yy = dot(yl, yl);
        gamma = yy > std::numeric_limits<double>::epsilon() ? dot(sl, yl) / yy : 1.0;
    }
    const int m = (int)m_s.size();
    std::vector<double> alpha(m);
    Vec q = g;
    for (int i = m - 1; i >= 0; --i) {
        alpha[i] = m_rho[i] * dot(m_s[i], q);
        for (size_t k = 0; k < n; ++k) q[k] -= alpha[i] * m_y[i][k];
    }
    Vec r(n);
    for (size_t k = 0; k < n; ++k) r[k] = gamma * (diag ? m_h0[k] : 1.0) * q[k];
    for (int i = 0; i < m; ++i) {
        const double beta = m_rho[i] * dot(m_y[i], r);
        for (size_t k = 0; k < n; ++k) r[k] += m_s[i][k] * (alpha[i] - beta);
    }
    for (size_t k = 0; k < n; ++k) r[k] = (!freeMask.empty() && !freeMask[k]) ? 0.0 : -r[k];
    return r;
}

OptimizerResult LBFGSBOptimizer::minimize(Objective objective, Gradient gradient, const Vec& x0,
                                          const Vec& lb, const Vec& ub)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    OptimizerResult res;
    m_iter = 0;
    if (m_clearOnMinimize) {
        m_s.clear();
        m_y.clear();
        m_rho.clear();
    }
    for (const Vec& v : m_s)
        if (v.size() != x0.size())
            throw std::invalid_argument("LBFGSBOptimizer: history of another parameter count");
    m_hist.clear();
    int nObj = 0, nGrad = 0;
    auto fobj = [&](const Vec& x) { ++nObj; return objective(x); };
    auto fgrad = [&](const Vec& x) { ++nGrad; return gradient(x); };
    Vec x = project(x0, lb, ub);
    Vec g = fgrad(x);
    // with the line search off the reference never evaluates f inside the loop
    // (lbfgsb_optimizer.cpp:90-96): the records carry NaN until the final value
    const double nan = std::numeric_limits<double>::quiet_NaN();
    double f = m_noLineSearch ? nan : fobj(x);
    double pg0 = 0.0;
    int stall = 0;
    auto pgnorm = [&](const Vec& xx, const Vec& gg) {
        Vec t(xx.size());
        for (size_t i = 0; i < xx.size(); ++i) t[i] = xx[i] - gg[i];
        const Vec p = project(t, lb, ub);
        double s = 0.0;
        for (size_t i = 0; i < xx.size(); ++i) s += (xx[i] - p[i]) * (xx[i] - p[i]);
        return std::sqrt(s);
    };
    m_hist.push_back({0, f, pgnorm(x, g), 0.0,
                      std::chrono::duration<double>(clk::now() - t0).count(), nObj, nGrad});
    if (m_onIteration) m_onIteration(m_hist.back());
    for (int k = 0; k < m_maxIter; ++k) {
        m_iter = k + 1;
        if (m_hasTarget && !m_noLineSearch && f <= m_target) {
            res.converged = true;
            res.message = "Reached target objective";
            break;
        }
        const double pg = pgnorm(x, g);
        if (k == 0) pg0 = pg;
        if (k > 0 && pg < m_ftol * std::max(1.0, pg0)) {
            res.converged = true;
            res.message = "Converged: projected gradient below tolerance";
            break;
        }
        std::vector<uint8_t> freeMask;
        if (m_activeSet) {
            freeMask.assign(x.size(), 1);
            for (size_t i = 0; i < x.size(); ++i)
                if ((x[i] <= lb[i] && g[i] > 0.0) || (x[i] >= ub[i] && g[i] < 0.0)) freeMask[i] = 0;
        }
        const Vec d = direction(g, freeMask);
        // projected Armijo backtracking (lbfgsb_optimizer.cpp:85-144); f(x) is carried
        // from the previous accepted step instead of being re-evaluated (:100), which
        // saves one objective call per iteration and leaves the iterates unchanged
        double alpha = 1.0;
        const double gd = dot(g, d);
        Vec xn, gn;
        double fn = f;
        bool accepted = false;
        if (m_noLineSearch) {
            // the full projected step, always (:90-96)
            xn.resize(x.size());
            for (size_t i = 0; i < x.size(); ++i) xn[i] = x[i] + d[i];
            xn = project(xn, lb, ub);
            fn = nan;
            accepted = true;
        } else if (gd >= 0.0) {
            res.converged = true;
            res.message = "Converged: no step taken";
            break;
        }
        for (int it = 0; !accepted && it < m_lsMax && alpha >= 1e-16; ++it) {
            Vec xt(x.size());
            for (size_t i = 0; i < x.size(); ++i) xt[i] = x[i] + alpha * d[i];
            xt = project(xt, lb, ub);
            // the trial is evaluated before its projected decrease is checked, as
            // there (:117-125), so the objective sees the same points
            const double ft = fobj(xt);
            double dec = 0.0;
            for (size_t i = 0; i < x.size(); ++i) dec += g[i] * (xt[i] - x[i]);
            if (dec >= 0.0) {
                alpha *= 0.5;
                continue;
            }
            if (ft <= f + m_c1 * dec) {
                xn = xt;
                fn = ft;
                accepted = true;
                break;
            }
            alpha *= 0.5;
        }
        if (!accepted) {
            alpha = 1e-16;
            xn.resize(x.size());
            for (size_t i = 0; i < x.size(); ++i) xn[i] = x[i] + alpha * d[i];
            xn = project(xn, lb, ub);
            fn = fobj(xn);
        }
        gn = fgrad(xn);
        if (alpha < 1e-8) {
            if (++stall >= 3) {
                res.converged = true;
                res.message = "Converged: line search stalled (alpha < 1e-8 for " +
                              std::to_string(stall) + " consecutive iterations)";
                break;
            }
        } else {
            stall = 0;
        }
        Vec s(x.size()), y(x.size());
        for (size_t i = 0; i < x.size(); ++i) {
            s[i] = xn[i] - x[i];
            y[i] = gn[i] - g[i];
        }
        const double sy = dot(s, y);
        if (sy > 1e-16) {
            if ((int)m_s.size() >= m_historySize) {
                m_s.erase(m_s.begin());
                m_y.erase(m_y.begin());
                m_rho.erase(m_rho.begin());
            }
            m_s.push_back(s);
            m_y.push_back(y);
            m_rho.push_back(1.0 / sy);
        }
        x = xn;
        g = gn;
        f = fn;
        m_hist.push_back({m_iter, f, pgnorm(x, g), alpha,
                          std::chrono::duration<double>(clk::now() - t0).count(), nObj, nGrad});
        if (m_onIteration) m_onIteration(m_hist.back());
    }
    if (res.message.empty()) res.message = "Maximum iterations reached";
    if (m_noLineSearch) {
        f = fobj(x);  // :255
        m_hist.back().f = f;
    }
    res.params = x;
    res.finalValue = f;
    res.iterations = m_iter;
    return res;
}

void LBFGSBOptimizer::setHistory(const std::vector<Vec>& s, const std::vector<Vec>& y,
                                 const std::vector<double>& rho)
{
    if (s.size() != y.size() || s.size() != rho.size())
        throw std::invalid_argument("LBFGSBOptimizer::setHistory: s, y and rho differ in length");
    m_s = s;
    m_y = y;
    m_rho = rho;
}

void LBFGSBOptimizer::restart(const Vec&)
{
    // lbfgsb_optimizer.cpp:265-272; the next minimize() starts from its own x0
    m_s.clear();
    m_y.clear();
    m_rho.clear();
    m_iter = 0;
}

// =================================================================== GN-CG

Vec GNCGOptimizer::cgSolve(const std::function<Vec(const Vec&)>& Hv, const Vec& b,
                           int& iters) const
{
    // gncg_optimizer.cpp:39-77: CG from x = 0.  A direction with p.Hp <= 1e-16 takes the
    // step |r|^2 / |p|^2 and still updates the residual with Hp.  Hv runs before the
    // |p|^2 check, as there.
    const size_t n = b.size();
    Vec x(n, 0.0), r = b, p = b;
    double rs = dot(r, r);
    iters = 0;
    if (rs < 1e-30) return x;
    for (int j = 0; j < m_cgMaxIter; ++j) {
        const Vec Hp = Hv(p);
        ++iters;
        const double p2 = dot(p, p);
        if (p2 < 1e-30) break;
        const double pHp = dot(p, Hp);
        const double a = pHp > 1e-16 ? rs / pHp : rs / p2;
        for (size_t i = 0; i < n; ++i) {
            x[i] += a * p[i];
            r[i] -= a * Hp[i];
        }
        const double rsNew = dot(r, r);
        if (rsNew < m_cgTol * m_cgTol) break;
        const double beta = rsNew / rs;
        for (size_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        rs = rsNew;
    }
    return x;
}

OptimizerResult GNCGOptimizer::minimize(Objective objective, Gradient gradient, const Vec& x0,
                                        const Vec& lb, const Vec& ub)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    OptimizerResult res;
    m_iter = 0;
    m_cgIters.clear();
    m_hist.clear();
    int nObj = 0, nGrad = 0;
    auto fobj = [&](const Vec& x) { ++nObj; return objective(x); };
    auto fgrad = [&](const Vec& x) { ++nGrad; return gradient(x); };
    const size_t n = x0.size();
    // gncg_optimizer.cpp:24-37: a component at a bound (within 1e-12) whose gradient
    // points out of the box does not count
    auto pgnorm = [&](const Vec& x, const Vec& g) {
        double s = 0.0;
        for (size_t i = 0; i < n; ++i) {
            double pg = g[i];
            if (x[i] <= lb[i] + 1e-12 && g[i] > 0) pg = 0.0;
            if (x[i] >= ub[i] - 1e-12 && g[i] < 0) pg = 0.0;
            s += pg * pg;
        }
        return std::sqrt(s);
    };
    auto record = [&](int k, double f, double pg, double alpha) {
        m_hist.push_back({k, f, pg, alpha,
                          std::chrono::duration<double>(clk::now() - t0).count(), nObj, nGrad});
        if (m_onIteration) m_onIteration(m_hist.back());
    };
    Vec x = project(x0, lb, ub);
    double pg0 = 0.0, fx = 0.0, lastAlpha = 0.0;
    bool recorded = false;  // x's row is in the history
    int stall = 0;
    for (int k = 0; k < m_maxIter; ++k) {
        m_iter = k + 1;
        const Vec g = fgrad(x);
        const double pg = pgnorm(x, g);
        if (k == 0) pg0 = pg;
        if (k > 0) {
            record(k, fx, pg, lastAlpha);
            recorded = true;
            if (pg < m_ftol * std::max(1.0, pg0)) {
                res.converged = true;
                res.message = "Converged: projected gradient below tolerance";
                break;
            }
        }
        // H v by central differences of the gradient (:134-147)
        auto Hv = [&](const Vec& v) {
            const double vn = std::sqrt(dot(v, v));
            if (vn < 1e-15) return Vec(n, 0.0);
            const double h = m_fdStep / vn;
            Vec xp(n), xm(n);
            for (size_t i = 0; i < n; ++i) {
                xp[i] = x[i] + h * v[i];
                xm[i] = x[i] - h * v[i];
            }
            const Vec gp = fgrad(project(xp, lb, ub));
            const Vec gm = fgrad(project(xm, lb, ub));
            Vec out(n);
            for (size_t i = 0; i < n; ++i) out[i] = (gp[i] - gm[i]) / (2.0 * h);
            return out;
        };
        Vec b(n);
        for (size_t i = 0; i < n; ++i) b[i] = -g[i];
        int cgIt = 0;
        Vec d = cgSolve(Hv, b, cgIt);
        m_cgIters.push_back(cgIt);
        if (dot(d, g) >= 0.0) d = b;  // not a descent direction: steepest descent (:151-157)
        // backtracking from 1 by halves until the projected point does not raise the
        // objective at x, evaluated afresh (:79-103)
        const double f0 = fobj(x);
        if (k == 0) {
            record(0, f0, pg, 0.0);
            recorded = true;
        }
        double alpha = 1.0, fn = f0;
        Vec xn;
        bool found = false;
        for (int i = 0; i < m_lsMax; ++i) {
            Vec xt(n);
            for (size_t j = 0; j < n; ++j) xt[j] = x[j] + alpha * d[j];
            xt = project(xt, lb, ub);
            const double ft = fobj(xt);
            if (ft <= f0) {
                xn = std::move(xt);
                fn = ft;
                found = true;
                break;
            }
            alpha *= 0.5;
        }
        if (!found) {
            res.converged = true;
            res.message = "Converged: no step taken";
            break;
        }
        double step = 0.0;
        for (size_t i = 0; i < n; ++i) step += (xn[i] - x[i]) * (xn[i] - x[i]);
        if (std::sqrt(step) < 1e-8) {
            if (++stall >= 3) {
                res.converged = true;
                res.message = "Converged: step stalled";
                break;
            }
        } else {
            stall = 0;
        }
        x = std::move(xn);
        fx = fn;
        lastAlpha = alpha;
        recorded = false;
    }
    res.params = x;
    res.finalValue = fobj(x);
    res.iterations = m_iter;
    if (!recorded)
        record(m_iter, res.finalValue, std::numeric_limits<double>::quiet_NaN(), lastAlpha);
    if (!res.converged && m_iter >= m_maxIter) res.message = "Maximum iterations reached";
    return res;
}

}  // namespace lithofwd
