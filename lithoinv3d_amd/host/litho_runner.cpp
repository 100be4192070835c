// litho_runner.cpp -- the inversion loop, its history statistics and the validation gate.
#include "litho_runner.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace lithofwd {

namespace {

double dw1d(const Vec& r)
{
    const size_t n = r.size();
    if (n < 2) return 2.0;
    double den = 0.0, num = 0.0;
    for (double v : r) den += v * v;
    if (den < 1e-30) return 2.0;
    for (size_t i = 1; i < n; ++i) {
        const double d = r[i] - r[i - 1];
        num += d * d;
    }
    return num / den;
}

// body(i) for i in [0, n) over the machine's threads (at most 16)
template <class F>
void parallel_for(int64_t n, F body)
{
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(
        {(int64_t)std::max(1u, std::thread::hardware_concurrency()), 16, n / 4096 + 1}));
    if (nt == 1) {
        for (int64_t i = 0; i < n; ++i) body(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int64_t i = t; i < n; i += nt) body(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace

DurbinWatson durbinWatson(const std::vector<double>& xyz, const Vec& residuals)
{
    const size_t n = residuals.size();
    if (n < 2 || xyz.size() != 3 * n) return {};
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    auto byAxis = [&](int a, int b) {
        return [&, a, b](size_t i, size_t j) {
            if (std::abs(xyz[3 * i + a] - xyz[3 * j + a]) > 1e-9) return xyz[3 * i + a] < xyz[3 * j + a];
            return xyz[3 * i + b] < xyz[3 * j + b];
        };
    };
    Vec sorted(n);
    DurbinWatson out;
    std::sort(order.begin(), order.end(), byAxis(0, 1));
    for (size_t i = 0; i < n; ++i) sorted[i] = residuals[order[i]];
    out.x = dw1d(sorted);
    std::sort(order.begin(), order.end(), byAxis(1, 0));
    for (size_t i = 0; i < n; ++i) sorted[i] = residuals[order[i]];
    out.y = dw1d(sorted);
    return out;
}

// ------------------------------------------------------------------ classification

PointClassifier::PointClassifier(const LithologyModel& model, const LabelGridData* grid,
                                 double bottomDepth)
    : m_model(model), m_grid(grid), m_bottom(bottomDepth)
{
    const auto& V = model.vertices();
    const auto& F = model.faces();
    const size_t nf = model.faceCount();
    if (nf == 0) return;
    double y0 = 1e300, y1 = -1e300, z0 = 1e300, z1 = -1e300;
    for (size_t v = 0; v < model.vertexCount(); ++v) {
        y0 = std::min(y0, V[3 * v + 1]);
        y1 = std::max(y1, V[3 * v + 1]);
        z0 = std::min(z0, V[3 * v + 2]);
        z1 = std::max(z1, V[3 * v + 2]);
    }
    // bins about twice the mean facet extent, so a facet covers a few bins (finer bins
    // multiply the entries of large facets; coarser ones lengthen every ray's list)
    double ey = 0.0, ez = 0.0;
    for (size_t f = 0; f < nf; ++f) {
        double a = 1e300, b = -1e300, c = 1e300, d = -1e300;
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = F[3 * f + k];
            a = std::min(a, V[3 * v + 1]);
            b = std::max(b, V[3 * v + 1]);
            c = std::min(c, V[3 * v + 2]);
            d = std::max(d, V[3 * v + 2]);
        }
        ey += b - a;
        ez += d - c;
    }
    m_y0 = y0;
    m_z0 = z0;
    m_hy = std::max({2.0 * ey / (double)nf, (y1 - y0) / 4096.0, 1e-9});
    m_hz = std::max({2.0 * ez / (double)nf, (z1 - z0) / 4096.0, 1e-9});
    m_ny = (int)std::min<double>(4096.0, std::floor((y1 - y0) / m_hy) + 1.0);
    m_nz = (int)std::min<double>(4096.0, std::floor((z1 - z0) / m_hz) + 1.0);
    auto cellRange = [&](size_t f, int& ya, int& yb, int& za, int& zb) {
        double lo_y = 1e300, hi_y = -1e300, lo_z = 1e300, hi_z = -1e300;
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = F[3 * f + k];
            lo_y = std::min(lo_y, V[3 * v + 1]);
            hi_y = std::max(hi_y, V[3 * v + 1]);
            lo_z = std::min(lo_z, V[3 * v + 2]);
            hi_z = std::max(hi_z, V[3 * v + 2]);
        }
        ya = std::clamp((int)std::floor((lo_y - m_y0) / m_hy), 0, m_ny - 1);
        yb = std::clamp((int)std::floor((hi_y - m_y0) / m_hy), 0, m_ny - 1);
        za = std::clamp((int)std::floor((lo_z - m_z0) / m_hz), 0, m_nz - 1);
        zb = std::clamp((int)std::floor((hi_z - m_z0) / m_hz), 0, m_nz - 1);
    };
    m_binOff.assign((size_t)m_ny * m_nz + 1, 0);
    for (size_t f = 0; f < nf; ++f) {
        int ya, yb, za, zb;
        cellRange(f, ya, yb, za, zb);
        for (int iy = ya; iy <= yb; ++iy)
            for (int iz = za; iz <= zb; ++iz) ++m_binOff[(size_t)iy * m_nz + iz + 1];
    }
    for (size_t i = 1; i < m_binOff.size(); ++i) m_binOff[i] += m_binOff[i - 1];
    m_binFace.resize(m_binOff.back());
    std::vector<int64_t> fill(m_binOff.begin(), m_binOff.end() - 1);
    for (size_t f = 0; f < nf; ++f) {
        int ya, yb, za, zb;
        cellRange(f, ya, yb, za, zb);
        for (int iy = ya; iy <= yb; ++iy)
            for (int iz = za; iz <= zb; ++iz) m_binFace[fill[(size_t)iy * m_nz + iz]++] = (int32_t)f;
    }
}

int PointClassifier::classify(double x, double y, double z) const
{
    const int nu = m_model.groupCount();
    if (nu == 0 || m_model.faceCount() == 0) return -1;
    if (z < m_bottom) return -1;
    if (m_grid && !m_grid->labels.empty()) {
        const int ix = (int)((x - m_grid->x0) / m_grid->dx);
        const int iy = (int)((y - m_grid->y0) / m_grid->dy);
        const int iz = (int)((z - m_grid->z0) / m_grid->dz);
        if (ix >= 0 && ix < m_grid->nx && iy >= 0 && iy < m_grid->ny && iz >= 0 && iz < m_grid->nz) {
            const int32_t gid = m_grid->labels[(size_t)iz * m_grid->ny * m_grid->nx +
                                               (size_t)iy * m_grid->nx + ix];
            if (gid >= 0) return gid;
        }
    }
    const double oy = y + 1.234567e-6, oz = z + 2.345678e-6;
    const int iy = (int)std::floor((oy - m_y0) / m_hy), iz = (int)std::floor((oz - m_z0) / m_hz);
    // beyond every facet's (y, z) extent (the far edges, index m_ny or m_nz, can still
    // touch a facet and fall into the last bin)
    if (iy < 0 || iz < 0 || iy > m_ny || iz > m_nz) return -1;
    const int cy = std::clamp(iy, 0, m_ny - 1), cz = std::clamp(iz, 0, m_nz - 1);
    const auto& V = m_model.vertices();
    const auto& F = m_model.faces();
    const auto& front = m_model.front();
    const auto& back = m_model.back();
    std::vector<int> hits(nu, 0);
    const size_t bin = (size_t)cy * m_nz + cz;
    for (int64_t k = m_binOff[bin]; k < m_binOff[bin + 1]; ++k) {
        const int32_t f = m_binFace[k];
        const double* v0 = &V[3 * F[3 * f]];
        const double* v1 = &V[3 * F[3 * f + 1]];
        const double* v2 = &V[3 * F[3 * f + 2]];
        const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        // h = (1, 0, 0) x e2
        const double h[3] = {0.0, -e2[2], e2[1]};
        const double a = e1[0] * h[0] + e1[1] * h[1] + e1[2] * h[2];
        if (std::abs(a) < 1e-30) continue;
        const double fa = 1.0 / a;
        const double s[3] = {x - v0[0], oy - v0[1], oz - v0[2]};
        const double u = fa * (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]);
        if (u < 0.0 || u > 1.0) continue;
        const double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2],
                             s[0] * e1[1] - s[1] * e1[0]};
        const double vv = fa * q[0];
        if (vv < 0.0 || u + vv > 1.0) continue;
        const double t = fa * (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]);
        if (t > 1e-12) {
            ++hits[front[f]];
            if (back[f] >= 0) ++hits[back[f]];
        }
    }
    for (int u = 0; u < nu; ++u)
        if (hits[u] & 1) return u;
    return -1;
}

// ------------------------------------------------------------------ validation

UnitNeighbours unitNeighbours(const LithologyModel& model)
{
    const int nu = model.groupCount();
    const auto& F = model.faces();
    const auto& front = model.front();
    const auto& back = model.back();
    UnitNeighbours nb;
    nb.verts.resize(nu);
    nb.off.resize(nu);
    nb.nbr.resize(nu);
    const size_t nv = model.vertexCount();
    std::vector<std::vector<std::pair<int32_t, int32_t>>> edges(nu);
    for (size_t f = 0; f < model.faceCount(); ++f) {
        for (int side = 0; side < 2; ++side) {
            const int u = side == 0 ? front[f] : back[f];
            if (u < 0) continue;
            for (int k = 0; k < 3; ++k) {
                const int32_t a = (int32_t)F[3 * f + k], b = (int32_t)F[3 * f + (k + 1) % 3];
                edges[u].push_back({a, b});
                edges[u].push_back({b, a});
            }
        }
    }
    std::vector<int32_t> local(nv, -1);
    for (int u = 0; u < nu; ++u) {
        auto& e = edges[u];
        std::sort(e.begin(), e.end());
        e.erase(std::unique(e.begin(), e.end()), e.end());
        auto& verts = nb.verts[u];
        for (auto& p : e)
            if (verts.empty() || verts.back() != p.first) verts.push_back(p.first);
        nb.off[u].assign(verts.size() + 1, 0);
        nb.nbr[u].reserve(e.size());
        size_t k = 0;
        for (size_t i = 0; i < verts.size(); ++i) {
            while (k < e.size() && e[k].first == verts[i]) nb.nbr[u].push_back(e[k++].second);
            nb.off[u][i + 1] = (int64_t)nb.nbr[u].size();
        }
    }
    return nb;
}

OverlapCounts countOverlaps(const LithologyModel& model, const PointClassifier& cls)
{
    const UnitNeighbours nb = unitNeighbours(model);
    const auto& V = model.vertices();
    OverlapCounts out;
    for (int g = 0; g < model.groupCount(); ++g) {
        const auto& verts = nb.verts[g];
        std::vector<int32_t> c(verts.size());
        parallel_for((int64_t)verts.size(), [&](int64_t i) {
            const int32_t v = verts[i];
            c[i] = cls.classify(V[3 * v], V[3 * v + 1], V[3 * v + 2]);
        });
        for (int32_t k : c) {
            if (k < 0 || k == g) continue;
            if (std::abs(k - g) == 1)
                ++out.adjacent;
            else
                ++out.nonAdjacent;
        }
    }
    return out;
}

ValidationResult validateModel(const LithologyModel& model, const PointClassifier& cls,
                               int64_t overlapBaseline, int64_t adjacentOverlapBaseline)
{
    ValidationResult r;
    if (model.groupCount() < 2) return r;
    const auto& V = model.vertices();
    double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
    for (size_t v = 0; v < model.vertexCount(); ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], V[3 * v + a]);
            hi[a] = std::max(hi[a], V[3 * v + a]);
        }
    const double modelDim = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1.0});

    const OverlapCounts oc = countOverlaps(model, cls);
    r.adjacentOverlaps = oc.adjacent;
    r.nonAdjacentOverlaps = oc.nonAdjacent;
    const int64_t adjLimit = adjacentOverlapBaseline + 20;
    if (oc.adjacent > adjLimit) {
        r.passed = false;
        r.worstViolation = (double)(oc.adjacent - adjacentOverlapBaseline);
        r.failureReason = "Adjacent overlap: " + std::to_string(oc.adjacent) + " (baseline=" +
                          std::to_string(adjacentOverlapBaseline) + ", limit=" +
                          std::to_string(adjLimit) + ")";
        return r;
    }
    const int64_t limit = overlapBaseline + 120;
    if (oc.nonAdjacent > limit) {
        r.passed = false;
        r.worstViolation = (double)(oc.nonAdjacent - overlapBaseline);
        r.failureReason = "Non-adjacent overlap: " + std::to_string(oc.nonAdjacent) +
                          " (baseline=" + std::to_string(overlapBaseline) + ", limit=" +
                          std::to_string(limit) + ")";
        return r;
    }
    const UnitNeighbours nb = unitNeighbours(model);
    const double spike = modelDim / 5.0;
    for (int g = 0; g < model.groupCount(); ++g) {
        const auto& verts = nb.verts[g];
        for (size_t i = 0; i < verts.size(); ++i) {
            const int64_t a = nb.off[g][i], b = nb.off[g][i + 1];
            if (a == b) continue;
            double m[3] = {0, 0, 0};
            for (int64_t k = a; k < b; ++k)
                for (int ax = 0; ax < 3; ++ax) m[ax] += V[3 * nb.nbr[g][k] + ax];
            double d2 = 0.0;
            for (int ax = 0; ax < 3; ++ax) {
                const double d = V[3 * verts[i] + ax] - m[ax] / (double)(b - a);
                d2 += d * d;
            }
            const double mag = std::sqrt(d2);
            if (mag > spike) {
                r.passed = false;
                r.worstViolation = mag;
                std::ostringstream os;
                os << "Spike detected: vertex " << verts[i] << " of unit " << g
                   << " Laplacian=" << mag << "m (threshold=" << spike << "m)";
                r.failureReason = os.str();
                return r;
            }
        }
    }
    return r;
}

Vec smoothGradient(const LithologyModel& model, const UnitNeighbours& nb, const Vec& g, double w)
{
    if (w <= 0.0) return g;
    const size_t nv = model.vertexCount();
    std::vector<int32_t> dofAt(3 * nv, -1);
    for (uint32_t d = 0; d < model.totalDofCount(); ++d)
        dofAt[3 * (size_t)model.dofVertex(d) + model.dofAxis(d)] = (int32_t)d;
    Vec out = g;
    for (int u = 0; u < model.groupCount(); ++u) {
        const auto& verts = nb.verts[u];
        for (size_t i = 0; i < verts.size(); ++i) {
            for (int a = 0; a < 3; ++a) {
                const int32_t d = dofAt[3 * (size_t)verts[i] + a];
                if (d < 0) continue;
                double s = 0.0;
                int n = 0;
                for (int64_t k = nb.off[u][i]; k < nb.off[u][i + 1]; ++k) {
                    const int32_t e = dofAt[3 * (size_t)nb.nbr[u][k] + a];
                    if (e >= 0) {
                        s += g[e];
                        ++n;
                    }
                }
                out[d] = n > 0 ? (1.0 - w) * g[d] + w * s / n : g[d];
            }
        }
    }
    return out;
}

// ------------------------------------------------------------------ the loop

RunnerResult runInversion(std::shared_ptr<LithologyModel> model, RunnerProblem& problem,
                          const RunnerOptions& opt, const LabelGridData* labelGrid)
{
    if (!problem.components || !problem.gradient || !problem.predict || !problem.gdata)
        throw std::invalid_argument("runInversion: components, gradient, predict and "
                                    "gravity data are required");
    RunnerResult res;
    std::ostringstream log;
    const GravityData& gd = *problem.gdata;
    const bool hasMag = problem.mdata && !problem.mdata->empty();
    std::vector<double> gxyz(3 * gd.size()), mxyz;
    for (size_t i = 0; i < gd.size(); ++i)
        for (int k = 0; k < 3; ++k) gxyz[3 * i + k] = gd[i].position[k];
    if (hasMag) {
        mxyz.resize(3 * problem.mdata->size());
        for (size_t i = 0; i < problem.mdata->size(); ++i)
            for (int k = 0; k < 3; ++k) mxyz[3 * i + k] = (*problem.mdata)[i].position[k];
    }

    Vec params = model->assembleParameterVector();
    Vec lower, upper;
    if (opt.modelBounds)
        model->getModelBounds(opt.modelBoundsMultiplier, lower, upper);
    else
        model->getBounds(params, opt.boundsMargin, lower, upper);
    if (opt.depthBounds)
        for (size_t i = 0; i < params.size(); ++i) {
            lower[i] = std::max(lower[i], params[i] - opt.depthMargin);
            upper[i] = std::min(upper[i], params[i] + opt.depthMargin);
        }

    auto entry = [&](const Vec& p, const JointObjective::JointComponents& c) {
        RunnerIteration it;
        it.iteration = (int)res.history.size();
        it.gravityMisfit = c.gravityMisfit;
        it.magneticMisfit = c.magneticMisfit;
        it.regularization = c.regularization;
        it.total = c.total;
        Vec gz, tmi;
        problem.predict(p, gz, tmi);
        Vec rg(gd.size());
        double ss = 0.0;
        for (size_t i = 0; i < gd.size(); ++i) {
            rg[i] = gd[i].g_obs - gz[i];
            ss += rg[i] * rg[i];
        }
        it.rms = gd.empty() ? 0.0 : std::sqrt(ss / (double)gd.size());
        const DurbinWatson dg = durbinWatson(gxyz, rg);
        it.dwGravityX = dg.x;
        it.dwGravityY = dg.y;
        if (hasMag && tmi.size() == problem.mdata->size()) {
            Vec rm(tmi.size());
            for (size_t i = 0; i < tmi.size(); ++i) rm[i] = (*problem.mdata)[i].t_obs - tmi[i];
            const DurbinWatson dm = durbinWatson(mxyz, rm);
            it.dwMagneticX = dm.x;
            it.dwMagneticY = dm.y;
        }
        return it;
    };

    res.history.push_back(entry(params, problem.components(params)));
    if (problem.onIteration) problem.onIteration(res.history.back());
    if (hasMag && opt.eigenvalueScaling && problem.calibrate) {
        res.alphaMag = problem.calibrate(params);
        log << "calibrated alpha_mag = " << res.alphaMag << "\n";
    }

    Vec lastRecorded = params;
    auto geoObj = [&](const Vec& p) {
        const auto c = problem.components(p);
        double d2 = 0.0;
        for (size_t i = 0; i < p.size(); ++i) d2 += (p[i] - lastRecorded[i]) * (p[i] - lastRecorded[i]);
        if (res.history.empty() || d2 > 1e-14) {
            lastRecorded = p;
            res.history.push_back(entry(p, c));
            if (problem.onIteration) problem.onIteration(res.history.back());
        }
        return c.total;
    };
    UnitNeighbours nb;
    if (opt.gradientSmoothing > 0.0) nb = unitNeighbours(*model);
    auto geoGrad = [&](const Vec& p) {
        const Vec g = problem.gradient(p);
        return opt.gradientSmoothing > 0.0 ? smoothGradient(*model, nb, g, opt.gradientSmoothing) : g;
    };

    OverlapCounts base;
    if (opt.validate && opt.geometry) {
        PointClassifier cls(*model, labelGrid);
        base = countOverlaps(*model, cls);
        log << "overlap baselines: adjacent " << base.adjacent << ", non-adjacent "
            << base.nonAdjacent << "\n";
    }

    // the geometry solver of the configuration (runner.cpp:191-212), one object for every
    // phase as there; each minimize starts afresh (L-BFGS-B clears its pairs)
    std::unique_ptr<Optimizer> geoOptimizer;
    if (opt.solver == RunnerOptions::Solver::GNCG) {
        auto gn = std::make_unique<GNCGOptimizer>();
        gn->setCGMaxIter(opt.gncgCGMaxIter);
        gn->setCGTolerance(opt.gncgCGTolerance);
        gn->setFDStep(opt.fdStep);
        gn->setArmijoC1(opt.armijoC1);
        gn->setLineSearchMaxIter(opt.lineSearchMaxIter);
        gn->setTolerance(opt.tolerance);
        geoOptimizer = std::move(gn);
    } else {
        auto lb = std::make_unique<LBFGSBOptimizer>();
        lb->setHistorySize(opt.lbfgsHistory);
        lb->setTolerance(opt.tolerance);
        lb->setArmijoC1(opt.armijoC1);
        lb->setLineSearchMaxIter(opt.lineSearchMaxIter);
        lb->setDisableLineSearch(opt.disableLineSearch);
        geoOptimizer = std::move(lb);
    }
    int total = 0;
    int exportCounter = 0;
    bool converged = false;
    while (total < opt.maxIterations && !converged &&
           (int)res.history.size() < 3 * opt.maxIterations) {
        if (opt.geometry) {
            const int phase = std::min(opt.propertyInterval, opt.maxIterations - total);
            if (phase <= 0) break;
            geoOptimizer->setMaxIterations(phase);
            const OptimizerResult geo = geoOptimizer->minimize(geoObj, geoGrad, params, lower, upper);
            log << "geometry phase"
                << (opt.solver == RunnerOptions::Solver::GNCG ? " (gncg)" : "") << ": "
                << geo.iterations << " iterations, f = " << geo.finalValue
                << " (" << geo.message << ")\n";
            const Vec prev = params;
            bool accepted = false;
            Vec best = geo.params;
            for (int retry = 0; retry <= 5; ++retry) {
                Vec trial = geo.params;
                if (retry > 0) {
                    const double s = 1.0 / (double)(1 << retry);
                    for (size_t i = 0; i < trial.size(); ++i)
                        trial[i] = prev[i] + s * (geo.params[i] - prev[i]);
                }
                model->applyParameterVector(trial);
                if (!opt.validate) {
                    best = trial;
                    accepted = true;
                    break;
                }
                PointClassifier cls(*model, labelGrid);
                const ValidationResult v = validateModel(*model, cls, base.nonAdjacent, base.adjacent);
                if (v.passed) {
                    best = trial;
                    accepted = true;
                    break;
                }
                ++res.validationRetries;
                log << "validation failed (retry " << retry << "): " << v.failureReason
                    << " (violation=" << v.worstViolation << ")\n";
            }
            if (!accepted) {
                model->applyParameterVector(prev);
                ++res.validationRejections;
                log << "all validation retries failed, reverting to the previous model\n";
            } else {
                params = best;
                // the model holds the accepted geometry (runner.cpp:571-575)
                if (!opt.iterationExportDir.empty()) {
                    char sub[24];
                    std::snprintf(sub, sizeof(sub), "iter_%03d", exportCounter++);
                    exportUnitVolumes(*model, opt.iterationExportDir + "/" + sub, "invert",
                                      opt.groupExportNames);
                }
            }
            if (geo.iterations == 0) {
                converged = true;
                break;
            }
            if (accepted) {
                converged = geo.converged;
                total += geo.iterations;
            }
        } else {
            converged = true;
        }
        if (opt.properties && total < opt.maxIterations && problem.invertProperties) {
            model->applyParameterVector(params);
            problem.invertProperties(0, opt.densityOptions);
            if (hasMag) problem.invertProperties(1, opt.susceptibilityOptions);
            log << "property phase\n";
            converged = !opt.geometry;
        }
    }

    Vec gz, tmi;
    problem.predict(params, gz, tmi);
    res.predictedGravity = gz;
    double ss = 0.0;
    for (size_t i = 0; i < gd.size(); ++i) ss += (gd[i].g_obs - gz[i]) * (gd[i].g_obs - gz[i]);
    res.finalRMS = gd.empty() ? 0.0 : std::sqrt(ss / (double)gd.size());
    res.finalMisfit = problem.components(params).total;
    res.converged = converged;
    res.totalIterations = total;
    res.params = params;
    for (int g = 0; g < model->groupCount(); ++g) {
        res.densities.push_back(model->group(g).density);
        res.susceptibilities.push_back(model->group(g).susceptibility);
    }
    res.log = log.str();
    return res;
}

ConstraintHandler::ConstraintHandler(std::shared_ptr<LithologyModel> model,
                                     std::vector<ConstraintRecord> c, const LabelGridData* grid,
                                     double bottomDepth)
    : m_model(std::move(model)), m_c(std::move(c)), m_hasGrid(grid != nullptr),
      m_bottom(bottomDepth)
{
    if (grid) m_grid = *grid;
}

Vec ConstraintHandler::violations() const
{
    PointClassifier cls(*m_model, m_hasGrid ? &m_grid : nullptr, m_bottom);
    Vec v(m_c.size(), 0.0);
    for (size_t k = 0; k < m_c.size(); ++k) {
        const ConstraintRecord& c = m_c[k];
        const int unit = cls.classify(c.x, c.y, -0.5 * (c.z_top + c.z_bottom));
        if (unit != c.unit) v[k] = std::abs(c.z_bottom - c.z_top);
    }
    return v;
}

double ConstraintHandler::evaluate(const Vec& params)
{
    m_model->applyParameterVector(params);
    double pen = 0.0;
    for (double v : violations()) pen += v * v;
    return m_weight * pen;
}

Vec ConstraintHandler::gradient(const Vec& params)
{
    const double h = 1.0;
    Vec p = params, g(params.size(), 0.0);
    for (size_t j = 0; j < p.size(); ++j) {
        p[j] = params[j] + h;
        const double fp = evaluate(p);
        p[j] = params[j] - h;
        const double fm = evaluate(p);
        p[j] = params[j];
        g[j] = m_weight * (fp - fm) / (2.0 * h);
    }
    m_model->applyParameterVector(params);
    return g;
}

SurfaceMeshData unitClosedMesh(const LithologyModel& model, int unit)
{
    const auto& F = model.faces();
    const auto& front = model.front();
    const auto& back = model.back();
    const auto& V = model.vertices();
    std::vector<uint32_t> tris;
    for (size_t f = 0; f < model.faceCount(); ++f) {
        if (front[f] == unit) tris.insert(tris.end(), {F[3 * f], F[3 * f + 1], F[3 * f + 2]});
    }
    for (size_t f = 0; f < model.faceCount(); ++f) {
        if (back[f] == unit) tris.insert(tris.end(), {F[3 * f], F[3 * f + 2], F[3 * f + 1]});
    }
    std::vector<uint32_t> used(tris);
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    SurfaceMeshData m;
    m.vertices.reserve(3 * used.size());
    for (uint32_t v : used) m.vertices.insert(m.vertices.end(), {V[3 * v], V[3 * v + 1], V[3 * v + 2]});
    m.triangles.resize(tris.size());
    for (size_t i = 0; i < tris.size(); ++i)
        m.triangles[i] = (uint32_t)(std::lower_bound(used.begin(), used.end(), tris[i]) - used.begin());
    return m;
}

std::vector<std::string> exportUnitVolumes(const LithologyModel& model, const std::string& dir,
                                           const std::string& base,
                                           const std::vector<std::string>& names)
{
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) throw std::runtime_error("exportUnitVolumes: cannot create " + dir + ": " + ec.message());
    std::vector<std::string> written;
    for (int g = 0; g < model.groupCount(); ++g) {
        const SurfaceMeshData m = unitClosedMesh(model, g);
        if (m.vertexCount() == 0) continue;
        const std::string name = g < (int)names.size() ? names[g] : "group_" + std::to_string(g);
        const std::string path = dir + "/" + base + "_" + name + ".ts";
        writeTSurf(m, path, base + "_" + name);
        written.push_back(path);
    }
    return written;
}

}  // namespace lithofwd
