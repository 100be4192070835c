// litho_runner.hpp -- the inversion loop around the GPU objective: the phased geometry /
// property loop of InversionRunner::run with its per-evaluation history (misfits, RMS,
// Durbin-Watson statistics of the residuals), the magnetic-weight calibration at the
// start, optional Laplacian smoothing of the gradient, and the model validation gate
// that damps or rejects a geometry step which breaks the model.
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "litho_host.hpp"
#include "litho_io.hpp"

namespace lithofwd {

// modules/litho-core/src/core/stats.cpp:8-85: DW = sum (r_i - r_{i-1})^2 / sum r_i^2 over
// the residuals ordered by x then y (x), and by y then x (y); coordinates closer than 1e-9
// count as equal; 2 for fewer than two residuals or sum r_i^2 < 1e-30.
struct DurbinWatson {
    double x = 2.0, y = 2.0;
};
DurbinWatson durbinWatson(const std::vector<double>& xyz, const Vec& residuals);

// Point classification (lithology_model.cpp:288-362): the label grid cell's unit when the
// grid holds one (>= 0) at the point, else the lowest unit whose closed surface a +x ray
// from the point (offset by 1.234567e-6 in y and 2.345678e-6 in z) crosses an odd number
// of times (Moller-Trumbore, t > 1e-12), else -1; -1 below bottomDepth.  The rays test
// only the facets whose (y, z) bounding box holds the ray (a uniform (y, z) bin grid),
// which finds the same crossings as testing every facet.
class PointClassifier {
public:
    PointClassifier(const LithologyModel& model, const LabelGridData* grid = nullptr,
                    double bottomDepth = -5000.0);
    int classify(double x, double y, double z) const;

private:
    const LithologyModel& m_model;
    const LabelGridData* m_grid;
    double m_bottom;
    double m_y0 = 0, m_z0 = 0, m_hy = 1, m_hz = 1;
    int m_ny = 0, m_nz = 0;
    std::vector<int64_t> m_binOff;
    std::vector<int32_t> m_binFace;
};

// ConstraintHandler (modules/litho-inversion/src/inversion/constraint_handler.cpp:15-66),
// the reference's discrete borehole penalty: weight * sum over intervals of v^2, with
// v = |z_bottom - z_top| when the interval's midpoint (x, y, -(z_top + z_bottom) / 2;
// depths positive down, z up) does not classify (PointClassifier at the applied
// parameters) into the interval's unit, else 0.  Its gradient: central differences of the
// weighted penalty with a step of 1 parameter unit over every parameter, multiplied by the
// weight once more, as there (the weight enters the gradient squared).  With a label grid
// the classification does not follow the geometry and the gradient is 0.  The model is
// left at params.  BoreholeContactTerm is the differentiable alternative.
class ConstraintHandler : public Regularization {
public:
    ConstraintHandler(std::shared_ptr<LithologyModel> model, std::vector<ConstraintRecord> c,
                      const LabelGridData* grid = nullptr, double bottomDepth = -5000.0);
    double evaluate(const Vec& params) override;
    Vec gradient(const Vec& params) override;
    // the unweighted violations v at the model's current geometry
    Vec violations() const;

private:
    std::shared_ptr<LithologyModel> m_model;
    std::vector<ConstraintRecord> m_c;
    bool m_hasGrid = false;
    LabelGridData m_grid;  // a copy: the term may outlive the caller's grid
    double m_bottom;
};

// LithologyModel::validate (lithology_model.cpp:624-798).  With one shared vertex array
// the shared-vertex-gap and edge-tearing checks (1, 4) hold by construction, and the
// bounding-box check (5) compares each vertex with a box grown from the same vertices,
// which no vertex can leave; the checks that can fail are the overlaps (2, 3): vertices
// of unit g classified into unit c != g, |c - g| == 1 (adjacent, limit baseline + 20)
// or > 1 (limit baseline + 120), and spikes (6): a vertex farther than modelDim / 5 from
// the mean of its neighbours on a unit surface (modelDim = largest box side, >= 1 m).
struct ValidationResult {
    bool passed = true;
    double worstViolation = 0.0;
    std::string failureReason;
    int64_t adjacentOverlaps = 0, nonAdjacentOverlaps = 0;
};
struct OverlapCounts {
    int64_t adjacent = 0, nonAdjacent = 0;
};
OverlapCounts countOverlaps(const LithologyModel& model, const PointClassifier& cls);
ValidationResult validateModel(const LithologyModel& model, const PointClassifier& cls,
                               int64_t overlapBaseline, int64_t adjacentOverlapBaseline);

// Unique neighbours of every vertex on each unit's closed surface (CSR per unit): the
// vertex lists the Laplacian checks and the gradient smoothing average over.
struct UnitNeighbours {
    std::vector<std::vector<int32_t>> verts;   // per unit: its vertices
    std::vector<std::vector<int64_t>> off;     // per unit: CSR offsets, one row per vertex
    std::vector<std::vector<int32_t>> nbr;
};
UnitNeighbours unitNeighbours(const LithologyModel& model);

// InversionRunner::smoothGradient (runner.cpp:643-706): per unit surface, a parameter on
// axis a of vertex v becomes (1 - w) g + w * mean of the parameters on axis a of v's
// neighbours that have one (unchanged if none); a vertex on several units takes the last
// unit's value, as the reference's loop over group meshes does.
Vec smoothGradient(const LithologyModel& model, const UnitNeighbours& nb, const Vec& g, double w);

// One unit's closed surface out of the shared facet set, as the reference holds it per
// group mesh: the facets with the unit in front as stored, those with the unit behind
// reversed (v0, v2, v1), and the vertices they use in ascending shared index.
SurfaceMeshData unitClosedMesh(const LithologyModel& model, int unit);
// InversionExporter::exportClosedVolume over the units (exporters.cpp:769-813, as the
// runner's iteration export calls it, runner.cpp:319-348): <dir>/<base>_<name>.ts with
// name = names[g] where the list has one, else "group_<g>"; units without facets are
// skipped; dir is created.  Returns the files written.
std::vector<std::string> exportUnitVolumes(const LithologyModel& model, const std::string& dir,
                                           const std::string& base,
                                           const std::vector<std::string>& names);

// InversionConfig (runner.h:27-170), the fields the loop here uses.
struct RunnerOptions {
    int maxIterations = 500;
    int propertyInterval = 50;       // geometry iterations between property phases
    bool geometry = true;            // false: property-only, one round
    bool properties = false;         // alternate with property phases
    PropertyInversionOptions densityOptions, susceptibilityOptions;
    bool eigenvalueScaling = true;   // calibrate alpha_mag at the start (magnetics only)
    double gradientSmoothing = 0.0;  // smoothGradient weight, 0 = off
    double boundsMargin = 200.0;     // getBounds margin (m)
    bool modelBounds = false;        // bounds from the model box (applyModelBounds) instead
    double modelBoundsMultiplier = 1.0;
    bool depthBounds = false;        // also clamp to the start -/+ depthMargin (param units)
    double depthMargin = 500.0;
    bool validate = true;            // the validation gate after each geometry phase
    int lbfgsHistory = 10;
    double tolerance = 1e-6;
    double armijoC1 = 1e-4;
    int lineSearchMaxIter = 50;
    // geometry solver (runner.h:49-55, runner.cpp:192-212): L-BFGS-B, or Gauss-Newton
    // with an inner CG solve whose Hessian-vector products are central differences of
    // the gradient with step fdStep / |v|
    enum class Solver { LBFGSB = 0, GNCG = 1 } solver = Solver::LBFGSB;
    int gncgCGMaxIter = 50;
    double gncgCGTolerance = 1e-6;
    double fdStep = 1.0;
    bool disableLineSearch = false;  // L-BFGS-B: the full projected step (runner.h:115-116)
    // iterationExportDir / groupExportNames (runner.h:139-155): after each accepted
    // geometry phase every unit's closed surface goes to <dir>/iter_NNN/invert_<name>.ts
    std::string iterationExportDir;
    std::vector<std::string> groupExportNames;
};

// InversionIteration (inversion_result.h:10-28), gravity and magnetics.
struct RunnerIteration {
    int iteration = 0;
    double gravityMisfit = 0, magneticMisfit = 0, regularization = 0, total = 0, rms = 0;
    double dwGravityX = 2, dwGravityY = 2, dwMagneticX = 2, dwMagneticY = 2;
};

struct RunnerResult {
    std::vector<RunnerIteration> history;
    Vec params;
    bool converged = false;
    int totalIterations = 0;
    double finalMisfit = 0.0, finalRMS = 0.0;
    Vec predictedGravity;
    std::vector<double> densities, susceptibilities;
    double alphaMag = 0.0;
    int validationRetries = 0, validationRejections = 0;
    std::string log;  // one line per phase, calibration and validation event
};

// The objective side the loop drives (a joint gravity + magnetic problem): the
// components and gradient of the objective, the predicted data at a parameter vector,
// the property phase, the calibration and alpha_mag.
struct RunnerProblem {
    std::function<JointObjective::JointComponents(const Vec&)> components;
    std::function<Vec(const Vec&)> gradient;
    std::function<void(const Vec&, Vec& gz, Vec& tmi)> predict;   // tmi empty: no magnetics
    std::function<void(int which, const PropertyInversionOptions&)> invertProperties;
    std::function<double(const Vec&)> calibrate;   // returns the new alpha_mag
    const GravityData* gdata = nullptr;
    const MagneticData* mdata = nullptr;           // null: gravity only
    // called with each history entry as it is appended (InversionRunner's iteration
    // callback, runner.h:200-201, runner.cpp:297, 447, 472)
    std::function<void(const RunnerIteration&)> onIteration;
};

// InversionRunner::run (runner.cpp:217-641): start parameters and bounds, history entry
// 0, calibration, then up to maxIterations geometry iterations in phases of
// propertyInterval, each followed by the validation gate (the phase's step scaled by
// 1, 1/2, ..., 1/32 until the model passes, else reverted) and, when enabled, a density
// and a susceptibility phase; every objective evaluation at new parameters (|dp|^2 >
// 1e-14, line-search trials included) appends a history entry.
RunnerResult runInversion(std::shared_ptr<LithologyModel> model, RunnerProblem& problem,
                          const RunnerOptions& options, const LabelGridData* labelGrid = nullptr);

}  // namespace lithofwd
