// litho_capi.cpp -- the C entry points of include/lithohost.h over the C++ host layer,
// for the Python drivers and tests (ctypes).  Handles are opaque; every call returns 0
// on success and a negative code on failure with li_last_error() describing it.
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <sstream>
#include <string>

#include "../../include/lithohost.h"
#include "litho_host.hpp"
#include "litho_io.hpp"
#include "litho_runner.hpp"

using namespace lithofwd;

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f)
{
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

struct Problem {
    std::shared_ptr<LithologyModel> model;
    GravityData gdata;
    MagneticData mdata;
    std::shared_ptr<GravityForward> grav;
    std::shared_ptr<MagneticForward> mag;
    std::shared_ptr<ObjectiveFunction> gobj;
    std::shared_ptr<ReferenceModelRegularization> ref;
    std::unique_ptr<JointObjective> joint;
    std::unique_ptr<FusedJointObjective> fused;  // used when present and enabled
    bool useFused = true;
    double inc = 0.0, dec = 0.0, field = 0.0;
    int device = 0;
    Precision prec = Precision::FP64;
    // magnetic property phase: remanence mode and the bounds of its parameters
    // (runner.h:98-107)
    RemanentMagnetizationMode remMode = RemanentMagnetizationMode::EffectiveSusceptibility;
    double remLower = 0.0, remUpper = 10.0, compLower = -10.0, compUpper = 10.0;
    li_iter_cb iterCb = nullptr;  // li_problem_set_iteration_callback
    void* iterUser = nullptr;

    void applyRemanenceBounds(PropertyInversionOptions& o) const
    {
        o.remanenceLower = remLower;
        o.remanenceUpper = remUpper;
        o.componentLower = compLower;
        o.componentUpper = compUpper;
    }
    MagneticForward& magnetic()
    {
        if (!mag) throw std::runtime_error("problem has no magnetic data");
        return *mag;
    }

    JointObjective::JointComponents components(const Vec& p)
    {
        return (fused && useFused) ? fused->evaluateComponents(p) : joint->evaluateComponents(p);
    }
    Vec grad(const Vec& p) { return (fused && useFused) ? fused->gradient(p) : joint->gradient(p); }
};

struct ModelHandle {
    std::shared_ptr<LithologyModel> m;
};

struct MeshHandle {
    SurfaceMeshData m;
    std::string warnings;
};
struct IniHandle {
    IniConfig c;
};
struct StationsHandle {
    std::vector<StationRecord> s;
    std::string warnings;
};
struct GridHandle {
    LabelGridData g;
};
struct AssembledHandle {
    AssembledModel a;
};
}  // namespace

extern "C" {

const char* li_last_error(void) { return g_err.c_str(); }

int li_model_create(void** out, int64_t n_v, const double* verts, int64_t n_f,
                    const uint32_t* faces, const int32_t* front, const int32_t* back,
                    int32_t n_u, const double* density, const double* susceptibility)
{
    return guard([&] {
        std::vector<LithoGroup> groups(n_u);
        for (int g = 0; g < n_u; ++g) {
            groups[g].id = g;
            groups[g].name = "unit" + std::to_string(g);
            groups[g].density = density[g];
            groups[g].susceptibility = susceptibility[g];
        }
        auto* h = new ModelHandle;
        h->m = std::make_shared<LithologyModel>(
            std::vector<double>(verts, verts + 3 * n_v),
            std::vector<uint32_t>(faces, faces + 3 * n_f), std::vector<int32_t>(front, front + n_f),
            std::vector<int32_t>(back, back + n_f), std::move(groups));
        *out = h;
    });
}

void li_model_destroy(void* h) { delete (ModelHandle*)h; }

int li_model_properties(void* h, double* density, double* susceptibility)
{
    return guard([&] {
        const auto& m = *((ModelHandle*)h)->m;
        for (int g = 0; g < m.groupCount(); ++g) {
            density[g] = m.group(g).density;
            susceptibility[g] = m.group(g).susceptibility;
        }
    });
}

// LithoGroup's remanence fields (litho_group.h:18-24): per unit magnitude (A/m),
// inclination and declination (degrees) and the vector (3 per unit); NULL keeps a field
int li_model_set_remanence(void* h, const double* magnitude, const double* inc_deg,
                           const double* dec_deg, const double* vector)
{
    return guard([&] {
        auto& m = *((ModelHandle*)h)->m;
        for (int g = 0; g < m.groupCount(); ++g) {
            LithoGroup& G = m.group(g);
            if (magnitude) G.remanence_magnitude = magnitude[g];
            if (inc_deg) G.remanence_inclination = inc_deg[g];
            if (dec_deg) G.remanence_declination = dec_deg[g];
            if (vector)
                for (int k = 0; k < 3; ++k) G.magnetization[k] = vector[3 * g + k];
        }
    });
}

int li_model_remanence(void* h, double* magnitude, double* inc_deg, double* dec_deg,
                       double* vector)
{
    return guard([&] {
        const auto& m = *((ModelHandle*)h)->m;
        for (int g = 0; g < m.groupCount(); ++g) {
            const LithoGroup& G = m.group(g);
            if (magnitude) magnitude[g] = G.remanence_magnitude;
            if (inc_deg) inc_deg[g] = G.remanence_inclination;
            if (dec_deg) dec_deg[g] = G.remanence_declination;
            if (vector)
                for (int k = 0; k < 3; ++k) vector[3 * g + k] = G.magnetization[k];
        }
    });
}

int li_model_set_move_vectors(void* h, const double* move_vectors)
{
    return guard([&] {
        auto& m = *((ModelHandle*)h)->m;
        for (size_t v = 0; v < m.vertexCount(); ++v) m.setMoveVector(v, move_vectors + 3 * v);
    });
}

int li_model_set_cp_interpolation(void* h, int mode)
{
    return guard([&] {
        if (mode != 0 && mode != 1)
            throw std::invalid_argument("li_model_set_cp_interpolation: mode must be 0 or 1");
        ((ModelHandle*)h)->m->setControlPointInterpolation((ControlPointInterpolation)mode);
    });
}

int li_model_fix_exterior(void* h, int interior_freedom)
{
    return guard([&] { ((ModelHandle*)h)->m->fixExteriorVertices((VertexFreedom)interior_freedom); });
}

int li_model_fix_exterior_faces(void* h, int freedom)
{
    return guard([&] { ((ModelHandle*)h)->m->fixExteriorFaces((VertexFreedom)freedom); });
}

int li_model_freedom(void* h, uint8_t* freedom)
{
    return guard([&] {
        const auto& m = *((ModelHandle*)h)->m;
        for (size_t v = 0; v < m.vertexCount(); ++v) freedom[v] = (uint8_t)m.vertexFreedom(v);
    });
}

int li_model_export_units(void* h, const char* dir, const char* base, const char* const* names,
                          int32_t n_names, int32_t* n_written)
{
    return guard([&] {
        if (!dir || !*dir) throw std::invalid_argument("li_model_export_units: empty directory");
        std::vector<std::string> nm;
        for (int32_t i = 0; i < n_names; ++i) nm.emplace_back(names[i] ? names[i] : "");
        const auto w = exportUnitVolumes(*((ModelHandle*)h)->m, dir, base ? base : "invert", nm);
        if (n_written) *n_written = (int32_t)w.size();
    });
}

int li_model_set_freedom(void* h, const uint8_t* freedom)
{
    return guard([&] {
        auto& m = *((ModelHandle*)h)->m;
        for (size_t v = 0; v < m.vertexCount(); ++v) m.setVertexFreedom(v, (VertexFreedom)freedom[v]);
    });
}

int li_model_set_axis_scale(void* h, int axis, double s)
{
    return guard([&] { ((ModelHandle*)h)->m->setAxisScale(axis, s); });
}

int li_model_set_cp_stride(void* h, int stride, double cell, int64_t* n_dof)
{
    return guard([&] { *n_dof = ((ModelHandle*)h)->m->setControlPointStride(stride, cell); });
}

int64_t li_model_dof_count(void* h) { return ((ModelHandle*)h)->m->totalDofCount(); }

int li_model_dofs(void* h, int32_t* vertex, int32_t* axis, uint8_t* is_cp)
{
    return guard([&] {
        auto& m = *((ModelHandle*)h)->m;
        for (uint32_t d = 0; d < m.totalDofCount(); ++d) {
            vertex[d] = m.dofVertex(d);
            axis[d] = m.dofAxis(d);
        }
        if (is_cp)
            for (size_t v = 0; v < m.vertexCount(); ++v) is_cp[v] = m.isControlPoint(v) ? 1 : 0;
    });
}

int li_model_params(void* h, double* out)
{
    return guard([&] {
        const Vec p = ((ModelHandle*)h)->m->assembleParameterVector();
        std::memcpy(out, p.data(), sizeof(double) * p.size());
    });
}

int li_model_apply(void* h, const double* params, double* vertices_out)
{
    return guard([&] {
        auto& m = *((ModelHandle*)h)->m;
        m.applyParameterVector(Vec(params, params + m.totalDofCount()));
        if (vertices_out)
            std::memcpy(vertices_out, m.vertices().data(), sizeof(double) * m.vertices().size());
    });
}

int li_model_pullback(void* h, const double* vgrad, double* out)
{
    return guard([&] {
        auto& m = *((ModelHandle*)h)->m;
        const Vec g = m.pullBackVertexGradient(Vec(vgrad, vgrad + 3 * m.vertexCount()));
        std::memcpy(out, g.data(), sizeof(double) * g.size());
    });
}

int li_model_tangent(void* h, const double* dparams, double* out)
{
    return guard([&] {
        auto& m = *((ModelHandle*)h)->m;
        const Vec t = m.vertexTangent(Vec(dparams, dparams + m.totalDofCount()));
        std::memcpy(out, t.data(), sizeof(double) * t.size());
    });
}

// parameter -> vertex map as CSR; call with vert/w NULL to get the entry count first
int li_model_dof_map(void* h, int64_t* off, int32_t* vert, double* w, int64_t* n_entries)
{
    return guard([&] {
        std::vector<int64_t> o;
        std::vector<int32_t> v;
        std::vector<double> ww;
        ((ModelHandle*)h)->m->dofVertexMap(o, v, ww);
        *n_entries = (int64_t)v.size();
        if (off) std::memcpy(off, o.data(), sizeof(int64_t) * o.size());
        if (vert) std::memcpy(vert, v.data(), sizeof(int32_t) * v.size());
        if (w) std::memcpy(w, ww.data(), sizeof(double) * ww.size());
    });
}

int li_model_bounds(void* h, const double* params, double margin_m, double* lo, double* hi)
{
    return guard([&] {
        auto& m = *((ModelHandle*)h)->m;
        Vec l, u;
        m.getBounds(Vec(params, params + m.totalDofCount()), margin_m, l, u);
        std::memcpy(lo, l.data(), sizeof(double) * l.size());
        std::memcpy(hi, u.data(), sizeof(double) * u.size());
    });
}

// Joint gravity + magnetic problem on one station set.  *_std <= 0 means unweighted.
// SurfaceSmoothness of the model at params: weight * penalty and its parameter gradient.
int li_model_model_bounds(void* h, double margin_multiplier, double* lo, double* hi)
{
    return guard([&] {
        Vec l, u;
        ((ModelHandle*)h)->m->getModelBounds(margin_multiplier, l, u);
        std::memcpy(lo, l.data(), sizeof(double) * l.size());
        std::memcpy(hi, u.data(), sizeof(double) * u.size());
    });
}

int li_model_smoothness(void* h, const double* params, double weight, double* value,
                        double* grad)
{
    return guard([&] {
        auto& m = ((ModelHandle*)h)->m;
        SurfaceSmoothness s(m);
        s.setWeight(weight);
        const Vec p(params, params + m->totalDofCount());
        *value = s.evaluate(p);
        if (grad) {
            const Vec g = s.gradient(p);
            std::memcpy(grad, g.data(), sizeof(double) * g.size());
        }
    });
}

int li_problem_create(void** out, void* model, int64_t n_s, const double* stations,
                      const double* gz_obs, const double* gz_std, const double* tmi_obs,
                      const double* tmi_std, double inc, double dec, double field,
                      double alpha_mag, int precision, double ref_weight, int device)
{
    return guard([&] {
        auto P = std::make_unique<Problem>();
        P->model = ((ModelHandle*)model)->m;
        P->gdata.resize(n_s);
        P->mdata.resize(n_s);
        for (int64_t i = 0; i < n_s; ++i) {
            for (int k = 0; k < 3; ++k) {
                P->gdata[i].position[k] = stations[3 * i + k];
                P->mdata[i].position[k] = stations[3 * i + k];
            }
            P->gdata[i].g_obs = gz_obs ? gz_obs[i] : 0.0;
            P->gdata[i].g_std = gz_std ? gz_std[i] : 0.0;
            P->mdata[i].t_obs = tmi_obs ? tmi_obs[i] : 0.0;
            P->mdata[i].t_std = tmi_std ? tmi_std[i] : 0.0;
        }
        const Precision prec = precision == 1 ? Precision::FP32C : Precision::FP64;
        P->inc = inc;
        P->dec = dec;
        P->field = field;
        P->device = device;
        P->prec = prec;
        P->grav = std::make_shared<GravityForward>(P->model, P->gdata, prec, device);
        P->gobj = std::make_shared<ObjectiveFunction>(P->grav, P->gdata);
        if (ref_weight > 0.0) {
            P->ref = std::make_shared<ReferenceModelRegularization>();
            P->ref->setWeight(ref_weight);
            P->ref->captureReference(P->model->assembleParameterVector());
            P->gobj->addRegularization(P->ref);
        }
        P->joint = std::make_unique<JointObjective>(P->gobj);
        if (tmi_obs && alpha_mag > 0.0) {
            P->mag = std::make_shared<MagneticForward>(P->model, P->mdata, inc, dec, field, prec);
            P->joint->setMagnetic(P->mag, P->mdata, alpha_mag);
            P->fused = std::make_unique<FusedJointObjective>(P->model, P->gdata, P->mdata, inc,
                                                             dec, field, alpha_mag, prec, device);
            if (P->ref) P->fused->addRegularization(P->ref);
        }
        *out = P.release();
    });
}

void li_problem_destroy(void* h) { delete (Problem*)h; }

namespace {
std::vector<ContactPick> make_picks(int64_t n, const double* xyz, const int32_t* above,
                                    const int32_t* below, const double* sigma)
{
    std::vector<ContactPick> picks((size_t)n);
    for (int64_t k = 0; k < n; ++k)
        picks[(size_t)k] = {xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], above[k], below[k], sigma[k]};
    return picks;
}
}  // namespace

// BoreholeContactTerm of the model at params: weight * misfit, its parameter gradient (if
// grad) and the model elevation of each pick (if zmodel; NaN where the line misses).
int li_model_boreholes(void* h, int64_t n, const double* xyz, const int32_t* above,
                       const int32_t* below, const double* sigma, double radius,
                       const double* params, double weight, double* value, double* grad,
                       double* zmodel)
{
    return guard([&] {
        auto& m = ((ModelHandle*)h)->m;
        BoreholeContactTerm b(m, make_picks(n, xyz, above, below, sigma), radius);
        b.setWeight(weight);
        const Vec p(params, params + m->totalDofCount());
        *value = b.evaluate(p);
        if (grad) {
            const Vec g = b.gradient(p);
            std::memcpy(grad, g.data(), sizeof(double) * g.size());
        }
        if (zmodel) {
            std::vector<double> z;
            b.misfit(nullptr, &z);
            std::memcpy(zmodel, z.data(), sizeof(double) * z.size());
        }
    });
}

// Adds a BoreholeContactTerm with this weight to the problem's objectives.
int li_problem_add_boreholes(void* h, int64_t n, const double* xyz, const int32_t* above,
                             const int32_t* below, const double* sigma, double radius,
                             double weight)
{
    return guard([&] {
        auto* P = (Problem*)h;
        auto b = std::make_shared<BoreholeContactTerm>(P->model,
                                                       make_picks(n, xyz, above, below, sigma),
                                                       radius);
        b->setWeight(weight);
        P->gobj->addRegularization(b);
        if (P->fused) P->fused->addRegularization(b);
    });
}

namespace {
std::vector<ConstraintRecord> make_constraints(int64_t n, const double* xy, const double* z_top,
                                               const double* z_bottom, const int32_t* unit)
{
    std::vector<ConstraintRecord> c((size_t)n);
    for (int64_t k = 0; k < n; ++k)
        c[(size_t)k] = {xy[2 * k], xy[2 * k + 1], z_top[k], z_bottom[k], unit[k]};
    return c;
}
}  // namespace

int li_model_constraint_penalty(void* h, void* label_grid, double bottom_depth, int64_t n,
                                const double* xy, const double* z_top, const double* z_bottom,
                                const int32_t* unit, double weight, const double* params,
                                double* value, double* violations, double* grad)
{
    return guard([&] {
        auto& m = ((ModelHandle*)h)->m;
        ConstraintHandler c(m, make_constraints(n, xy, z_top, z_bottom, unit),
                            label_grid ? &((GridHandle*)label_grid)->g : nullptr, bottom_depth);
        c.setWeight(weight);
        const Vec p(params, params + m->totalDofCount());
        *value = c.evaluate(p);
        if (violations) {
            const Vec v = c.violations();
            std::memcpy(violations, v.data(), sizeof(double) * v.size());
        }
        if (grad) {
            const Vec g = c.gradient(p);
            std::memcpy(grad, g.data(), sizeof(double) * g.size());
        }
    });
}

int li_problem_add_constraints(void* h, void* label_grid, double bottom_depth, int64_t n,
                               const double* xy, const double* z_top, const double* z_bottom,
                               const int32_t* unit, double weight)
{
    return guard([&] {
        auto* P = (Problem*)h;
        auto c = std::make_shared<ConstraintHandler>(
            P->model, make_constraints(n, xy, z_top, z_bottom, unit),
            label_grid ? &((GridHandle*)label_grid)->g : nullptr, bottom_depth);
        c->setWeight(weight);
        P->gobj->addRegularization(c);
        if (P->fused) P->fused->addRegularization(c);
    });
}

// The property phase (runner.cpp:708-935): which = 0 densities from gravity, 1
// susceptibilities from TMI; the new unit properties go to out (n_units) and into the
// model, and every forward of the problem takes them.
int li_problem_invert_properties(void* h, int which, double damping, double lower, double upper,
                                 int max_iter, double tol, double* out)
{
    return guard([&] {
        auto* P = (Problem*)h;
        PropertyInversionOptions o;
        o.damping = damping;
        o.lower = lower;
        o.upper = upper;
        o.maxIter = max_iter;
        o.tolerance = tol;
        P->applyRemanenceBounds(o);
        Vec p;
        if (which == 0) {
            p = invertUnitDensities(*P->grav, P->gdata, *P->model, o);
        } else if (which == 1) {
            p = invertUnitSusceptibilities(P->magnetic(), P->mdata, *P->model, o);
        } else {
            throw std::invalid_argument("li_problem_invert_properties: which must be 0 or 1");
        }
        if (P->mag) P->mag->refreshProperties();
        P->grav->refreshProperties();
        if (P->fused) P->fused->refreshProperties();
        std::memcpy(out, p.data(), sizeof(double) * p.size());
    });
}

// JointObjective::calibrateDataWeights (joint_objective.cpp:281-393): sets alpha_mag of the
// problem from the ratio of the largest eigenvalues of J^T W^2 J of gravity and magnetics
// at params.  out[0..6] = gravity eigenvalue, magnetic eigenvalue, their ratio, the new
// alpha_mag, gravity and magnetic power iterations, 1 if the gradient-norm fallback was used.
int li_problem_calibrate(void* h, const double* params, uint64_t seed, double eps, double* out)
{
    return guard([&] {
        auto* P = (Problem*)h;
        if (!P->mag) throw std::runtime_error("problem has no magnetic data");
        const Vec p(params, params + P->model->totalDofCount());
        const double a0 = P->joint->alphaMag();
        const CalibrationResult r = calibrateMagneticWeight(P->model, P->gdata, P->mdata, P->inc,
                                                            P->dec, P->field, p, a0, seed, eps,
                                                            P->device, P->remMode, P->prec);
        P->joint->setAlphaMag(r.alpha);
        if (P->fused) P->fused->setAlphaMag(r.alpha);
        const double o[7] = {r.eigGravity, r.eigMagnetic, r.ratio, r.alpha,
                             (double)r.iterationsGravity, (double)r.iterationsMagnetic,
                             r.fallback ? 1.0 : 0.0};
        std::memcpy(out, o, sizeof(o));
    });
}

int li_problem_jvp(void* h, const double* params, const double* dparams, double* gz_out,
                   double* tmi_out)
{
    return guard([&] {
        auto* P = (Problem*)h;
        if (!P->fused) throw std::runtime_error("li_problem_jvp: problem has no magnetic data");
        const size_t n = P->model->totalDofCount();
        Vec gz, tmi;
        P->fused->jvp(Vec(params, params + n), Vec(dparams, dparams + n), gz, tmi);
        if (gz_out) std::memcpy(gz_out, gz.data(), sizeof(double) * gz.size());
        if (tmi_out) std::memcpy(tmi_out, tmi.data(), sizeof(double) * tmi.size());
    });
}

namespace {
void copy_text(const std::string& t, char* out, int64_t cap)
{
    if (!out || cap <= 0) return;
    const size_t n = std::min<size_t>(t.size(), (size_t)cap - 1);
    std::memcpy(out, t.data(), n);
    out[n] = 0;
}
}  // namespace

// InversionRunner::run (runner.cpp:217-641) over this problem: calibration, geometry
// phases with the validation gate, property phases.  hist rows of 10: iteration,
// gravity misfit, magnetic misfit, regularization, total, gravity RMS, DW gravity x/y,
// DW magnetic x/y.  summary[8]: converged, geometry iterations, final objective, final
// gravity RMS, alpha_mag after calibration (0 if none), validation retries, rejections,
// history entries in total.
// a runner history entry as the 10 values of li_problem_run's hist rows
static void runner_row(const RunnerIteration& it, double* row)
{
    const double v[10] = {(double)it.iteration, it.gravityMisfit, it.magneticMisfit,
                          it.regularization, it.total, it.rms, it.dwGravityX,
                          it.dwGravityY, it.dwMagneticX, it.dwMagneticY};
    std::memcpy(row, v, sizeof(v));
}

int li_problem_run(void* h, const LiRunnerOptions* o, void* label_grid, double* params_out,
                   double* hist, int64_t hist_cap, int64_t* n_hist, double* summary, char* log,
                   int64_t log_cap)
{
    return guard([&] {
        auto* P = (Problem*)h;
        RunnerOptions ro;
        ro.maxIterations = o->max_iterations;
        ro.propertyInterval = o->property_interval;
        ro.geometry = o->geometry != 0;
        ro.properties = o->properties != 0;
        ro.eigenvalueScaling = o->eigenvalue_scaling != 0;
        ro.validate = o->validate != 0;
        ro.depthBounds = o->depth_bounds != 0;
        ro.depthMargin = o->depth_margin;
        ro.boundsMargin = o->bounds_margin;
        ro.modelBounds = o->model_bounds != 0;
        ro.modelBoundsMultiplier = o->model_bounds_multiplier;
        ro.gradientSmoothing = o->gradient_smoothing;
        ro.lbfgsHistory = o->lbfgs_history;
        ro.tolerance = o->tolerance;
        ro.armijoC1 = o->armijo_c1;
        ro.lineSearchMaxIter = o->line_search_max_iter;
        if (o->solver != 0 && o->solver != 1)
            throw std::invalid_argument("li_problem_run: solver must be 0 (lbfgsb) or 1 (gncg)");
        ro.solver = (RunnerOptions::Solver)o->solver;
        ro.gncgCGMaxIter = o->gncg_cg_max_iter;
        ro.gncgCGTolerance = o->gncg_cg_tolerance;
        ro.fdStep = o->fd_step;
        ro.disableLineSearch = o->disable_line_search != 0;
        if (o->iteration_export_dir) ro.iterationExportDir = o->iteration_export_dir;
        for (int32_t i = 0; i < o->n_group_export_names; ++i)
            ro.groupExportNames.emplace_back(o->group_export_names[i] ? o->group_export_names[i] : "");
        ro.densityOptions = {o->density_damping, o->density_lower, o->density_upper,
                             o->property_max_iter, o->property_tolerance};
        ro.susceptibilityOptions = {o->susceptibility_damping, o->susceptibility_lower,
                                    o->susceptibility_upper, o->property_max_iter,
                                    o->property_tolerance};
        RunnerProblem rp;
        rp.gdata = &P->gdata;
        rp.mdata = P->mag ? &P->mdata : nullptr;
        rp.components = [P](const Vec& p) { return P->components(p); };
        rp.gradient = [P](const Vec& p) { return P->grad(p); };
        rp.predict = [P](const Vec& p, Vec& gz, Vec& tmi) {
            if (P->fused && P->useFused) {
                P->fused->forward(p, gz, tmi);
                return;
            }
            gz = P->grav->compute(p);
            tmi.clear();
            if (P->mag) tmi = P->mag->compute(p);
        };
        rp.invertProperties = [P](int which, const PropertyInversionOptions& po0) {
            PropertyInversionOptions po = po0;
            P->applyRemanenceBounds(po);
            if (which == 0)
                invertUnitDensities(*P->grav, P->gdata, *P->model, po);
            else if (P->mag)
                invertUnitSusceptibilities(*P->mag, P->mdata, *P->model, po);
            if (P->mag) P->mag->refreshProperties();
            P->grav->refreshProperties();
            if (P->fused) P->fused->refreshProperties();
        };
        rp.calibrate = [P](const Vec& p) {
            const CalibrationResult r = calibrateMagneticWeight(
                P->model, P->gdata, P->mdata, P->inc, P->dec, P->field, p, P->joint->alphaMag(),
                1, 0.0, P->device, P->remMode, P->prec);
            P->joint->setAlphaMag(r.alpha);
            if (P->fused) P->fused->setAlphaMag(r.alpha);
            return r.alpha;
        };
        if (P->iterCb) {
            rp.onIteration = [P](const RunnerIteration& it) {
                double row[10];
                runner_row(it, row);
                P->iterCb(1, row, 10, P->iterUser);
            };
        }
        const LabelGridData* grid = label_grid ? &((GridHandle*)label_grid)->g : nullptr;
        const RunnerResult r = runInversion(P->model, rp, ro, grid);
        std::memcpy(params_out, r.params.data(), sizeof(double) * r.params.size());
        const int64_t nh = std::min<int64_t>((int64_t)r.history.size(), hist_cap);
        for (int64_t k = 0; k < nh; ++k) runner_row(r.history[(size_t)k], hist + 10 * k);
        *n_hist = nh;
        const double sm[8] = {r.converged ? 1.0 : 0.0, (double)r.totalIterations, r.finalMisfit,
                              r.finalRMS, r.alphaMag, (double)r.validationRetries,
                              (double)r.validationRejections, (double)r.history.size()};
        std::memcpy(summary, sm, sizeof(sm));
        copy_text(r.log, log, log_cap);
    });
}

// classifyPoint (lithology_model.cpp:288-362) for n points; label_grid may be NULL
int li_model_classify(void* h, void* label_grid, double bottom_depth, int64_t n,
                      const double* xyz, int32_t* out)
{
    return guard([&] {
        const auto& m = *((ModelHandle*)h)->m;
        const LabelGridData* grid = label_grid ? &((GridHandle*)label_grid)->g : nullptr;
        PointClassifier cls(m, grid, bottom_depth);
        for (int64_t i = 0; i < n; ++i)
            out[i] = cls.classify(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    });
}

// LithologyModel::validate (lithology_model.cpp:624-798) at the model's current geometry.
// counts[2] = adjacent and non-adjacent overlaps; with baselines < 0 only the counts
// are computed (the baselines of the starting model, runner.cpp:493-516).
int li_model_validate(void* h, void* label_grid, double bottom_depth, int64_t overlap_baseline,
                      int64_t adjacent_baseline, int32_t* passed, double* worst,
                      int64_t* counts, char* reason, int64_t reason_cap)
{
    return guard([&] {
        const auto& m = *((ModelHandle*)h)->m;
        const LabelGridData* grid = label_grid ? &((GridHandle*)label_grid)->g : nullptr;
        PointClassifier cls(m, grid, bottom_depth);
        if (overlap_baseline < 0 || adjacent_baseline < 0) {
            const OverlapCounts c = countOverlaps(m, cls);
            counts[0] = c.adjacent;
            counts[1] = c.nonAdjacent;
            *passed = 1;
            *worst = 0.0;
            copy_text("", reason, reason_cap);
            return;
        }
        const ValidationResult v = validateModel(m, cls, overlap_baseline, adjacent_baseline);
        counts[0] = v.adjacentOverlaps;
        counts[1] = v.nonAdjacentOverlaps;
        *passed = v.passed ? 1 : 0;
        *worst = v.worstViolation;
        copy_text(v.failureReason, reason, reason_cap);
    });
}

// computeDurbinWatson (stats.cpp:31-85): out[2] = DW along x and along y
int li_durbin_watson(int64_t n, const double* xyz, const double* residuals, double* out)
{
    return guard([&] {
        const DurbinWatson d = durbinWatson(std::vector<double>(xyz, xyz + 3 * n),
                                            Vec(residuals, residuals + n));
        out[0] = d.x;
        out[1] = d.y;
    });
}

// Adds SurfaceSmoothness with this weight to the problem's objectives.
int li_problem_add_smoothness(void* h, double weight)
{
    return guard([&] {
        auto* P = (Problem*)h;
        auto s = std::make_shared<SurfaceSmoothness>(P->model);
        s->setWeight(weight);
        P->gobj->addRegularization(s);
        if (P->fused) P->fused->addRegularization(s);
    });
}

// 1: the fused one-context objective (default when gravity and magnetic data share the
// stations); 0: the reference-shaped JointObjective over separate forwards
int li_problem_set_fused(void* h, int on)
{
    return guard([&] { ((Problem*)h)->useFused = on != 0; });
}

// forward / adjoint passes made by the fused objective so far
int li_problem_pass_counts(void* h, int64_t* fwd, int64_t* adj)
{
    return guard([&] {
        auto* P = (Problem*)h;
        *fwd = P->fused ? P->fused->forwardPasses() : 0;
        *adj = P->fused ? P->fused->adjointPasses() : 0;
    });
}

// components: gravity misfit, magnetic misfit, regularization, total
int li_problem_evaluate(void* h, const double* params, double* comp)
{
    return guard([&] {
        auto* P = (Problem*)h;
        const auto c = P->components(Vec(params, params + P->model->totalDofCount()));
        comp[0] = c.gravityMisfit;
        comp[1] = c.magneticMisfit;
        comp[2] = c.regularization;
        comp[3] = c.total;
    });
}

int li_problem_gradient(void* h, const double* params, double* out)
{
    return guard([&] {
        auto* P = (Problem*)h;
        const Vec g = P->grad(Vec(params, params + P->model->totalDofCount()));
        std::memcpy(out, g.data(), sizeof(double) * g.size());
    });
}

int li_problem_forward(void* h, const double* params, double* gz, double* tmi)
{
    return guard([&] {
        auto* P = (Problem*)h;
        const Vec p(params, params + P->model->totalDofCount());
        if (gz && tmi && P->mag && P->fused && P->useFused) {
            // both fields from one pass over the station x facet pairs
            Vec g, t;
            P->fused->forward(p, g, t);
            std::memcpy(gz, g.data(), sizeof(double) * g.size());
            std::memcpy(tmi, t.data(), sizeof(double) * t.size());
            return;
        }
        if (gz) {
            const Vec g = P->grav->compute(p);
            std::memcpy(gz, g.data(), sizeof(double) * g.size());
        }
        if (tmi) {
            if (!P->mag) throw std::runtime_error("problem has no magnetic data");
            const Vec t = P->mag->compute(p);
            std::memcpy(tmi, t.data(), sizeof(double) * t.size());
        }
    });
}

// Dense Jacobians through the C++ ForwardModel surface (small problems).
int li_problem_jacobian(void* h, const double* params, int which, double* J_colmajor)
{
    return guard([&] {
        auto* P = (Problem*)h;
        const Vec p(params, params + P->model->totalDofCount());
        Mat J = which == 0 ? P->grav->computeJacobian(p) : P->mag->computeJacobian(p);
        std::memcpy(J_colmajor, J.data.data(), sizeof(double) * J.data.size());
    });
}

int li_problem_unit_response(void* h, int which, int group, double* out)
{
    return guard([&] {
        auto* P = (Problem*)h;
        const Vec r = which == 0 ? P->grav->computeGroupUnitResponse(group)
                                 : P->magnetic().computeGroupUnitResponse(group);
        std::memcpy(out, r.data(), sizeof(double) * r.size());
    });
}

int li_problem_unit_response_direction(void* h, int group, const double* direction, double* out)
{
    return guard([&] {
        const Vec r = ((Problem*)h)->magnetic().computeGroupUnitResponseDirection(group, direction);
        std::memcpy(out, r.data(), sizeof(double) * r.size());
    });
}

int li_problem_unit_remanence_response(void* h, int group, double* out)
{
    return guard([&] {
        const Vec r = ((Problem*)h)->magnetic().computeGroupUnitRemanenceResponse(group);
        std::memcpy(out, r.data(), sizeof(double) * r.size());
    });
}

int li_problem_set_remanence_mode(void* h, int mode)
{
    return guard([&] {
        auto* P = (Problem*)h;
        if (mode < 0 || mode > 2)
            throw std::invalid_argument("li_problem_set_remanence_mode: mode must be 0, 1 or 2");
        P->remMode = (RemanentMagnetizationMode)mode;
        if (P->mag) P->mag->setRemanenceMode(P->remMode);
        if (P->fused) P->fused->setRemanenceMode(P->remMode);
    });
}

int li_problem_set_remanence_bounds(void* h, double magnitude_lower, double magnitude_upper,
                                    double component_lower, double component_upper)
{
    return guard([&] {
        auto* P = (Problem*)h;
        P->remLower = magnitude_lower;
        P->remUpper = magnitude_upper;
        P->compLower = component_lower;
        P->compUpper = component_upper;
    });
}

// hist rows: iteration, f, |proj grad|, alpha, seconds, objective calls, gradient calls
// an optimizer record as the 7 values of the minimize hist rows
static void minimize_row(const LBFGSBOptimizer::IterationRecord& r, double* row)
{
    row[0] = r.iteration;
    row[1] = r.f;
    row[2] = r.pgNorm;
    row[3] = r.alpha;
    row[4] = r.seconds;
    row[5] = r.objectiveCalls;
    row[6] = r.gradientCalls;
}

static int run_minimize(LBFGSBOptimizer::Objective obj, LBFGSBOptimizer::Gradient grad,
                        int64_t n, const double* x0, const double* lb, const double* ub,
                        const LiOptions* o, const double* h0, double* x_out, double* hist,
                        int64_t hist_cap, int64_t* n_hist, int32_t* iterations, char* msg,
                        int64_t msg_cap, li_iter_cb cb = nullptr, void* cb_user = nullptr,
                        LBFGSBOptimizer* shared = nullptr)
{
    return guard([&] {
        LBFGSBOptimizer local;
        LBFGSBOptimizer& opt = shared ? *shared : local;
        if (cb) {
            opt.setIterationCallback([cb, cb_user](const LBFGSBOptimizer::IterationRecord& r) {
                double row[7];
                minimize_row(r, row);
                cb(0, row, 7, cb_user);
            });
        } else {
            opt.setIterationCallback(LBFGSBOptimizer::IterationCallback());
        }
        opt.setMaxIterations(o->max_iter);
        opt.setHistorySize(o->history);
        opt.setTolerance(o->ftol);
        opt.setArmijoC1(o->armijo_c1);
        if (!std::isnan(o->target)) opt.setTargetObjective(o->target);
        if (o->has_h0 && h0) opt.setInitialHessianDiagonal(Vec(h0, h0 + n));
        opt.setActiveSetAware(o->active_set != 0);
        opt.setDisableLineSearch(o->disable_line_search != 0);
        const auto r = opt.minimize(obj, grad, Vec(x0, x0 + n), Vec(lb, lb + n), Vec(ub, ub + n));
        std::memcpy(x_out, r.params.data(), sizeof(double) * n);
        const auto& H = opt.history();
        const int64_t nh = std::min<int64_t>((int64_t)H.size(), hist_cap);
        for (int64_t k = 0; k < nh; ++k) minimize_row(H[k], hist + 7 * k);
        *n_hist = nh;
        *iterations = r.iterations;
        if (msg && msg_cap > 0) {
            std::strncpy(msg, r.message.c_str(), (size_t)msg_cap - 1);
            msg[msg_cap - 1] = 0;
        }
    });
}

int li_minimize_callbacks(int64_t n, const double* x0, const double* lb, const double* ub,
                          li_obj_cb f, li_grad_cb g, void* user, const LiOptions* o,
                          const double* h0, double* x_out, double* hist, int64_t hist_cap,
                          int64_t* n_hist, int32_t* iterations, char* msg, int64_t msg_cap)
{
    auto obj = [&](const Vec& x) { return f(x.data(), (int64_t)x.size(), user); };
    auto grad = [&](const Vec& x) {
        Vec out(x.size());
        g(x.data(), (int64_t)x.size(), out.data(), user);
        return out;
    };
    return run_minimize(obj, grad, n, x0, lb, ub, o, h0, x_out, hist, hist_cap, n_hist, iterations,
                        msg, msg_cap);
}

struct LbfgsbHandle {
    LBFGSBOptimizer opt;
};

int li_lbfgsb_create(void** out)
{
    return guard([&] { *out = new LbfgsbHandle(); });
}

void li_lbfgsb_destroy(void* h) { delete (LbfgsbHandle*)h; }

int li_lbfgsb_set_clear_on_minimize(void* h, int32_t clear)
{
    return guard([&] { ((LbfgsbHandle*)h)->opt.setClearOnMinimize(clear != 0); });
}

int li_lbfgsb_minimize_callbacks(void* h, int64_t n, const double* x0, const double* lb,
                                 const double* ub, li_obj_cb f, li_grad_cb g, void* user,
                                 const LiOptions* o, const double* h0, double* x_out,
                                 double* hist, int64_t hist_cap, int64_t* n_hist,
                                 int32_t* iterations, char* msg, int64_t msg_cap)
{
    auto obj = [&](const Vec& x) { return f(x.data(), (int64_t)x.size(), user); };
    auto grad = [&](const Vec& x) {
        Vec out(x.size());
        g(x.data(), (int64_t)x.size(), out.data(), user);
        return out;
    };
    return run_minimize(obj, grad, n, x0, lb, ub, o, h0, x_out, hist, hist_cap, n_hist, iterations,
                        msg, msg_cap, nullptr, nullptr, &((LbfgsbHandle*)h)->opt);
}

int li_lbfgsb_get_history(void* h, int64_t n, int32_t* count, double* s, double* y, double* rho)
{
    return guard([&] {
        std::vector<Vec> S, Y;
        std::vector<double> R;
        ((LbfgsbHandle*)h)->opt.getHistory(S, Y, R);
        *count = (int32_t)S.size();
        for (size_t k = 0; k < S.size(); ++k) {
            if ((int64_t)S[k].size() != n)
                throw std::invalid_argument("li_lbfgsb_get_history: n differs from the history's");
            if (s) std::memcpy(s + k * n, S[k].data(), sizeof(double) * n);
            if (y) std::memcpy(y + k * n, Y[k].data(), sizeof(double) * n);
            if (rho) rho[k] = R[k];
        }
    });
}

int li_lbfgsb_set_history(void* h, int64_t n, int32_t count, const double* s, const double* y,
                          const double* rho)
{
    return guard([&] {
        if (count < 0 || n < 0 || (count > 0 && (!s || !y || !rho)))
            throw std::invalid_argument("li_lbfgsb_set_history: invalid arguments");
        std::vector<Vec> S(count), Y(count);
        std::vector<double> R(count);
        for (int32_t k = 0; k < count; ++k) {
            S[k].assign(s + k * n, s + (k + 1) * n);
            Y[k].assign(y + k * n, y + (k + 1) * n);
            R[k] = rho[k];
        }
        ((LbfgsbHandle*)h)->opt.setHistory(S, Y, R);
    });
}

int li_lbfgsb_restart(void* h)
{
    return guard([&] { ((LbfgsbHandle*)h)->opt.restart(); });
}

int li_problem_minimize(void* h, const double* x0, const double* lb, const double* ub,
                        const LiOptions* o, const double* h0, double* x_out, double* hist,
                        int64_t hist_cap, int64_t* n_hist, int32_t* iterations, char* msg,
                        int64_t msg_cap)
{
    auto* P = (Problem*)h;
    const int64_t n = P->model->totalDofCount();
    auto obj = [&](const Vec& x) { return P->components(x).total; };
    auto grad = [&](const Vec& x) { return P->grad(x); };
    return run_minimize(obj, grad, n, x0, lb, ub, o, h0, x_out, hist, hist_cap, n_hist, iterations,
                        msg, msg_cap, P->iterCb, P->iterUser);
}

static int run_gncg(GNCGOptimizer::Objective obj, GNCGOptimizer::Gradient grad, int64_t n,
                    const double* x0, const double* lb, const double* ub,
                    const LiGncgOptions* o, double* x_out, double* hist, int64_t hist_cap,
                    int64_t* n_hist, int32_t* cg_iters, int64_t cg_cap, int32_t* iterations,
                    char* msg, int64_t msg_cap, li_iter_cb cb = nullptr, void* cb_user = nullptr)
{
    return guard([&] {
        if (o->max_iter < 0 || o->cg_max_iter < 0 || o->line_search_max_iter < 0)
            throw std::invalid_argument("li_gncg: iteration limits must be >= 0");
        if (!(o->fd_step > 0.0)) throw std::invalid_argument("li_gncg: fd_step must be > 0");
        GNCGOptimizer opt;
        if (cb) {
            opt.setIterationCallback([cb, cb_user](const GNCGOptimizer::IterationRecord& r) {
                double row[7];
                minimize_row(r, row);
                cb(0, row, 7, cb_user);
            });
        }
        opt.setMaxIterations(o->max_iter);
        opt.setCGMaxIter(o->cg_max_iter);
        opt.setCGTolerance(o->cg_tolerance);
        opt.setFDStep(o->fd_step);
        opt.setTolerance(o->ftol);
        opt.setArmijoC1(o->armijo_c1);
        opt.setLineSearchMaxIter(o->line_search_max_iter);
        const auto r = opt.minimize(obj, grad, Vec(x0, x0 + n), Vec(lb, lb + n), Vec(ub, ub + n));
        std::memcpy(x_out, r.params.data(), sizeof(double) * n);
        const auto& H = opt.history();
        const int64_t nh = std::min<int64_t>((int64_t)H.size(), hist_cap);
        for (int64_t k = 0; k < nh; ++k) minimize_row(H[k], hist + 7 * k);
        *n_hist = nh;
        const auto& C = opt.cgIterations();
        for (int64_t k = 0; k < std::min<int64_t>((int64_t)C.size(), cg_cap); ++k)
            cg_iters[k] = C[k];
        *iterations = r.iterations;
        if (msg && msg_cap > 0) {
            std::strncpy(msg, r.message.c_str(), (size_t)msg_cap - 1);
            msg[msg_cap - 1] = 0;
        }
    });
}

int li_gncg_minimize_callbacks(int64_t n, const double* x0, const double* lb, const double* ub,
                               li_obj_cb f, li_grad_cb g, void* user, const LiGncgOptions* o,
                               double* x_out, double* hist, int64_t hist_cap, int64_t* n_hist,
                               int32_t* cg_iters, int64_t cg_cap, int32_t* iterations, char* msg,
                               int64_t msg_cap)
{
    auto obj = [&](const Vec& x) { return f(x.data(), (int64_t)x.size(), user); };
    auto grad = [&](const Vec& x) {
        Vec out(x.size());
        g(x.data(), (int64_t)x.size(), out.data(), user);
        return out;
    };
    return run_gncg(obj, grad, n, x0, lb, ub, o, x_out, hist, hist_cap, n_hist, cg_iters, cg_cap,
                    iterations, msg, msg_cap);
}

int li_problem_gncg_minimize(void* h, const double* x0, const double* lb, const double* ub,
                             const LiGncgOptions* o, double* x_out, double* hist,
                             int64_t hist_cap, int64_t* n_hist, int32_t* cg_iters,
                             int64_t cg_cap, int32_t* iterations, char* msg, int64_t msg_cap)
{
    auto* P = (Problem*)h;
    const int64_t n = P->model->totalDofCount();
    auto obj = [&](const Vec& x) { return P->components(x).total; };
    auto grad = [&](const Vec& x) { return P->grad(x); };
    return run_gncg(obj, grad, n, x0, lb, ub, o, x_out, hist, hist_cap, n_hist, cg_iters, cg_cap,
                    iterations, msg, msg_cap, P->iterCb, P->iterUser);
}

int li_problem_set_iteration_callback(void* h, li_iter_cb cb, void* user)
{
    auto* P = (Problem*)h;
    P->iterCb = cb;
    P->iterUser = cb ? user : nullptr;
    return 0;
}

// ------------------------------------------------------------------ file formats

int li_ini_load(const char* path, void** out)
{
    return guard([&] {
        auto h = std::make_unique<IniHandle>();
        if (!path || !h->c.load(path))
            throw std::runtime_error(std::string("li_ini_load: cannot open ") + (path ? path : "(null)"));
        *out = h.release();
    });
}

int li_ini_parse(const char* text, void** out)
{
    return guard([&] {
        auto h = std::make_unique<IniHandle>();
        if (text) {
            std::istringstream in(text);
            h->c.parse(in);
        }
        *out = h.release();
    });
}

void li_ini_destroy(void* h) { delete (IniHandle*)h; }

int li_ini_get(void* h, const char* section, const char* key, char* value, int64_t cap,
               int32_t* found)
{
    return guard([&] {
        const auto& c = ((IniHandle*)h)->c;
        const bool f = c.has(section, key);
        if (found) *found = f ? 1 : 0;
        copy_text(f ? c.getString(section, key) : std::string(), value, cap);
    });
}

int li_ini_get_double(void* h, const char* section, const char* key, double def, double* out)
{
    return guard([&] { *out = ((IniHandle*)h)->c.getDouble(section, key, def); });
}

int li_ini_get_int(void* h, const char* section, const char* key, int32_t def, int32_t* out)
{
    return guard([&] { *out = ((IniHandle*)h)->c.getInt(section, key, def); });
}

int li_ini_get_bool(void* h, const char* section, const char* key, int32_t def, int32_t* out)
{
    return guard([&] { *out = ((IniHandle*)h)->c.getBool(section, key, def != 0) ? 1 : 0; });
}

int li_ini_set(void* h, const char* section, const char* key, const char* value)
{
    return guard([&] { ((IniHandle*)h)->c.setString(section, key, value ? value : ""); });
}

int li_ini_save(void* h, const char* path)
{
    return guard([&] { ((IniHandle*)h)->c.save(path); });
}

int li_ini_dump(void* h, char* buf, int64_t cap, int64_t* need)
{
    return guard([&] {
        std::string t;
        for (const auto& sec : ((IniHandle*)h)->c.data())
            for (const auto& kv : sec.second) t += sec.first + "\t" + kv.first + "\t" + kv.second + "\n";
        if (need) *need = (int64_t)t.size() + 1;
        copy_text(t, buf, cap);
    });
}

int li_mesh_load(const char* path, void** out)
{
    return guard([&] {
        auto h = std::make_unique<MeshHandle>();
        h->m = loadSurfaceMesh(path, &h->warnings);
        *out = h.release();
    });
}

int li_mesh_counts(void* h, int64_t* n_vertices, int64_t* n_triangles)
{
    auto* H = (MeshHandle*)h;
    *n_vertices = (int64_t)H->m.vertexCount();
    *n_triangles = (int64_t)H->m.triangleCount();
    return 0;
}

int li_mesh_get(void* h, double* vertices, uint32_t* triangles)
{
    auto* H = (MeshHandle*)h;
    std::memcpy(vertices, H->m.vertices.data(), sizeof(double) * H->m.vertices.size());
    std::memcpy(triangles, H->m.triangles.data(), sizeof(uint32_t) * H->m.triangles.size());
    return 0;
}

const char* li_mesh_warnings(void* h) { return ((MeshHandle*)h)->warnings.c_str(); }

void li_mesh_destroy(void* h) { delete (MeshHandle*)h; }

int li_mesh_write_ts(const char* path, const char* name, int64_t n_vertices,
                     const double* vertices, int64_t n_triangles, const uint32_t* triangles)
{
    return guard([&] {
        SurfaceMeshData m;
        m.vertices.assign(vertices, vertices + 3 * n_vertices);
        m.triangles.assign(triangles, triangles + 3 * n_triangles);
        for (uint32_t t : m.triangles)
            if ((int64_t)t >= n_vertices)
                throw std::invalid_argument("li_mesh_write_ts: triangle index out of range");
        writeTSurf(m, path, name);
    });
}

int li_stations_load(const char* path, void** out)
{
    return guard([&] {
        auto h = std::make_unique<StationsHandle>();
        h->s = loadStationCSV(path, &h->warnings);
        *out = h.release();
    });
}

int64_t li_stations_count(void* h) { return (int64_t)((StationsHandle*)h)->s.size(); }

int li_stations_get(void* h, double* xyz, double* obs, double* stdev)
{
    const auto& S = ((StationsHandle*)h)->s;
    for (size_t i = 0; i < S.size(); ++i) {
        xyz[3 * i] = S[i].x;
        xyz[3 * i + 1] = S[i].y;
        xyz[3 * i + 2] = S[i].z;
        obs[i] = S[i].obs;
        stdev[i] = S[i].stdev;
    }
    return 0;
}

const char* li_stations_warnings(void* h) { return ((StationsHandle*)h)->warnings.c_str(); }

void li_stations_destroy(void* h) { delete (StationsHandle*)h; }

struct ConstraintsHandle {
    std::vector<ConstraintRecord> c;
    std::string warnings;
};

int li_constraints_load(const char* path, void** out)
{
    return guard([&] {
        auto h = std::make_unique<ConstraintsHandle>();
        h->c = loadConstraintCSV(path, &h->warnings);
        *out = h.release();
    });
}

int64_t li_constraints_count(void* h) { return (int64_t)((ConstraintsHandle*)h)->c.size(); }

int li_constraints_get(void* h, double* xy, double* z_top, double* z_bottom, int32_t* unit)
{
    const auto& C = ((ConstraintsHandle*)h)->c;
    for (size_t i = 0; i < C.size(); ++i) {
        xy[2 * i] = C[i].x;
        xy[2 * i + 1] = C[i].y;
        z_top[i] = C[i].z_top;
        z_bottom[i] = C[i].z_bottom;
        unit[i] = C[i].unit;
    }
    return 0;
}

const char* li_constraints_warnings(void* h) { return ((ConstraintsHandle*)h)->warnings.c_str(); }

void li_constraints_destroy(void* h) { delete (ConstraintsHandle*)h; }

int li_label_grid_load(const char* path, void** out)
{
    return guard([&] {
        auto h = std::make_unique<GridHandle>();
        h->g = loadLabelGrid(path);
        *out = h.release();
    });
}

// a label grid from memory: dims nx, ny, nz; labels with i fastest
int li_label_grid_create(const int32_t* dims, const double* origin, const double* cell,
                         const int32_t* labels, void** out)
{
    return guard([&] {
        if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0 || cell[0] <= 0 || cell[1] <= 0 ||
            cell[2] <= 0)
            throw std::invalid_argument("li_label_grid_create: dims and cell sizes must be > 0");
        auto h = std::make_unique<GridHandle>();
        h->g.nx = dims[0];
        h->g.ny = dims[1];
        h->g.nz = dims[2];
        h->g.x0 = origin[0];
        h->g.y0 = origin[1];
        h->g.z0 = origin[2];
        h->g.dx = cell[0];
        h->g.dy = cell[1];
        h->g.dz = cell[2];
        h->g.labels.assign(labels, labels + (size_t)dims[0] * dims[1] * dims[2]);
        *out = h.release();
    });
}

int li_label_grid_info(void* h, int32_t* dims, double* origin, double* cell)
{
    const auto& g = ((GridHandle*)h)->g;
    dims[0] = g.nx;
    dims[1] = g.ny;
    dims[2] = g.nz;
    origin[0] = g.x0;
    origin[1] = g.y0;
    origin[2] = g.z0;
    cell[0] = g.dx;
    cell[1] = g.dy;
    cell[2] = g.dz;
    return 0;
}

int li_label_grid_get(void* h, int32_t* labels)
{
    const auto& g = ((GridHandle*)h)->g;
    std::memcpy(labels, g.labels.data(), sizeof(int32_t) * g.labels.size());
    return 0;
}

void li_label_grid_destroy(void* h) { delete (GridHandle*)h; }

int li_label_grid_write(const char* path, const int32_t* dims, const double* origin,
                        const double* cell, const int32_t* labels)
{
    return guard([&] {
        LabelGridData g;
        g.nx = dims[0];
        g.ny = dims[1];
        g.nz = dims[2];
        if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0)
            throw std::invalid_argument("li_label_grid_write: non-positive grid size");
        g.x0 = origin[0];
        g.y0 = origin[1];
        g.z0 = origin[2];
        g.dx = cell[0];
        g.dy = cell[1];
        g.dz = cell[2];
        g.labels.assign(labels, labels + (size_t)g.nx * g.ny * g.nz);
        writeLabelGrid(g, path);
    });
}

// Units' meshes concatenated: unit u owns vertices [v_off[u], v_off[u+1]) of verts and
// triangles [t_off[u], t_off[u+1]) of tris, whose indices are local to the unit.
int li_assemble(int32_t n_units, const int64_t* v_off, const double* verts, const int64_t* t_off,
                const uint32_t* tris, void** out)
{
    return guard([&] {
        if (n_units < 0) throw std::invalid_argument("li_assemble: negative unit count");
        std::vector<SurfaceMeshData> meshes(n_units);
        for (int32_t u = 0; u < n_units; ++u) {
            const int64_t v0 = v_off[u], v1 = v_off[u + 1], t0 = t_off[u], t1 = t_off[u + 1];
            if (v1 < v0 || t1 < t0) throw std::invalid_argument("li_assemble: offsets not monotone");
            meshes[u].vertices.assign(verts + 3 * v0, verts + 3 * v1);
            meshes[u].triangles.assign(tris + 3 * t0, tris + 3 * t1);
            for (uint32_t t : meshes[u].triangles)
                if ((int64_t)t >= v1 - v0)
                    throw std::invalid_argument("li_assemble: unit " + std::to_string(u) +
                                                ": triangle index out of range");
        }
        auto h = std::make_unique<AssembledHandle>();
        h->a = assembleUnitMeshes(meshes);
        *out = h.release();
    });
}

int li_assembled_counts(void* h, int64_t* n_vertices, int64_t* n_faces, int64_t* n_contacts,
                        int64_t* n_merged)
{
    const auto& a = ((AssembledHandle*)h)->a;
    *n_vertices = (int64_t)(a.vertices.size() / 3);
    *n_faces = (int64_t)a.front.size();
    *n_contacts = a.contacts;
    *n_merged = a.mergedVertices;
    return 0;
}

int li_assembled_get(void* h, double* vertices, uint32_t* faces, int32_t* front, int32_t* back)
{
    const auto& a = ((AssembledHandle*)h)->a;
    std::memcpy(vertices, a.vertices.data(), sizeof(double) * a.vertices.size());
    std::memcpy(faces, a.faces.data(), sizeof(uint32_t) * a.faces.size());
    std::memcpy(front, a.front.data(), sizeof(int32_t) * a.front.size());
    std::memcpy(back, a.back.data(), sizeof(int32_t) * a.back.size());
    return 0;
}

void li_assembled_destroy(void* h) { delete (AssembledHandle*)h; }

}  // extern "C"
