// litho_host.hpp -- C++ host layer over the lithofwd C ABI, keeping the call surface
// that LithoInv3D's litho-inversion module uses (ForwardModel::compute /
// computeJacobian / computeBoth / dataCount / parameterCount, GravityForward and
// MagneticForward constructors, ObjectiveFunction, JointObjective, LBFGSBOptimizer,
// ReferenceModelRegularization), so that the GPU path drops in under the same names.
//
// Vectors are std::vector<double> and matrices column-major (the layout of Eigen's
// MatrixXd), so an Eigen caller maps its VectorXd / MatrixXd storage onto them directly.
//
// Differences from the reference, by design:
//   * the model keeps ONE shared vertex array (the reference keeps one per group mesh
//     and re-synchronizes shared positions after every update,
//     lithology_model.cpp:230-247), and stores a contact triangle once with the unit on
//     each side;
//   * control-point interpolation moves non-control vertices by the IDW-weighted
//     DISPLACEMENT of their K = 4 nearest control points; the reference writes the
//     IDW-weighted control-point POSITIONS (surface_mesh.cpp:610-623), which moves
//     every non-control vertex even for an unchanged parameter vector;
//   * gradients are exact: ForwardModel::computeJTr runs the GPU adjoint and maps vertex
//     gradients to parameters through the transposed parameter -> vertex map, including
//     the interpolation weights the reference's Jacobian drops
//     (gravity_forward.cpp:269-288).
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

struct lf_ctx;

namespace lithofwd {

using Vec = std::vector<double>;

// Column-major dense matrix (Eigen::MatrixXd layout).
struct Mat {
    size_t rows = 0, cols = 0;
    std::vector<double> data;
    Mat() = default;
    Mat(size_t r, size_t c) : rows(r), cols(c), data(r * c, 0.0) {}
    double& operator()(size_t r, size_t c) { return data[c * rows + r]; }
    double operator()(size_t r, size_t c) const { return data[c * rows + r]; }
};

// modules/litho-core/include/litho_invert/core/common.h:16-38
struct GravityPoint {
    double position[3] = {0, 0, 0};
    double g_obs = 0.0;
    double g_std = 0.0;
};
struct MagneticPoint {
    double position[3] = {0, 0, 0};
    double t_obs = 0.0;
    double t_std = 0.0;
};
using GravityData = std::vector<GravityPoint>;
using MagneticData = std::vector<MagneticPoint>;

// modules/litho-surface/include/litho_invert/surface/surface_mesh.h (VertexFreedom);
// ALONG_VECTOR: one parameter, the vertex's projection on its (normalized) move vector
// (surface_mesh.cpp:125-128, 194-196, 217-218)
enum class VertexFreedom : uint8_t {
    FIXED = 0, Z_ONLY = 1, X_ONLY = 2, Y_ONLY = 3, ALONG_VECTOR = 4, XY_FREE = 5, XZ_FREE = 6,
    YZ_FREE = 7, XYZ_FREE = 8
};

// How vertices that are not control points follow the control points:
//   Displacement: the IDW-weighted displacement of their control points, along their own
//                 free directions (a FIXED vertex stays put); unchanged parameters leave
//                 the geometry as it is
//   Position:     the reference's interpolateFromControlPoints (surface_mesh.cpp:610-623):
//                 every non-control vertex, FIXED ones included, becomes the IDW-weighted
//                 average of its control points' positions
enum class ControlPointInterpolation : uint8_t { Displacement = 0, Position = 1 };

// modules/litho-model/include/litho_invert/litho/litho_group.h:7-34
enum class RemanentMagnetizationMode { EffectiveSusceptibility = 0, FixedVectorPerGroup = 1, VectorPerGroup = 2 };
struct LithoGroup {
    int id = -1;
    std::string name;
    double density = 0.0;         // g/cc (signed contrasts allowed here)
    double susceptibility = 0.0;  // SI
    double remanence_magnitude = 0.0;
    double remanence_inclination = 0.0;
    double remanence_declination = 0.0;
    double magnetization[3] = {0, 0, 0};
};

// magnetic_forward.cpp:100-134: the unit's magnetization M = chi F/100 along F-hat (induced),
// plus the remanent part of the mode (FixedVectorPerGroup: remanence_magnitude along the
// remanence inclination / declination, VectorPerGroup: the magnetization vector), returned
// as amplitude |M| and direction (F-hat when |M| < 1e-15).  Units as in the reference:
// the remanent part is added to the chi F/100 amplitude as it is given.
void groupMagnetization(const LithoGroup& G, RemanentMagnetizationMode mode, double field_nT,
                        const double fhat[3], double& amp, double dir[3]);

// Lithology model: shared vertices, per-facet (front, back) units, DOF map.
class LithologyModel {
public:
    LithologyModel(std::vector<double> vertices, std::vector<uint32_t> faces,
                   std::vector<int32_t> front, std::vector<int32_t> back,
                   std::vector<LithoGroup> groups);

    int groupCount() const { return (int)m_groups.size(); }
    LithoGroup& group(int g) { return m_groups[g]; }
    const LithoGroup& group(int g) const { return m_groups[g]; }
    size_t vertexCount() const { return m_vertices.size() / 3; }
    size_t faceCount() const { return m_front.size(); }
    const std::vector<double>& vertices() const { return m_vertices; }
    const std::vector<uint32_t>& faces() const { return m_faces; }
    const std::vector<int32_t>& front() const { return m_front; }
    const std::vector<int32_t>& back() const { return m_back; }

    void setVertexFreedom(size_t v, VertexFreedom f);
    VertexFreedom vertexFreedom(size_t v) const { return m_freedom[v]; }
    // the move vector of an ALONG_VECTOR vertex (surface_mesh.h Vertex::moveVector; used
    // normalized); a zero vector gives the vertex no motion
    void setMoveVector(size_t v, const double mv[3]);
    const double* moveVector(size_t v) const { return &m_moveVector[3 * v]; }
    // d -> the part of d the vertex may move along: the free axes' components, or the
    // projection on the unit move vector (ALONG_VECTOR), or 0 (FIXED)
    void projectFree(size_t v, double d[3]) const;
    void setControlPointInterpolation(ControlPointInterpolation m) { m_cpInterp = m; }
    ControlPointInterpolation controlPointInterpolation() const { return m_cpInterp; }
    // Every vertex of a facet that bounds the model (back == -1) becomes FIXED, the
    // rest takes `interior` (the role of LithologyModel::fixExteriorFaces,
    // lithology_model.cpp:369-579, for closed voxel models).
    void fixExteriorVertices(VertexFreedom interior);
    // LithologyModel::fixExteriorFaces after finalizeModelSetup's freedom reset
    // (lithology_model.cpp:369-579, apps/forrestania/model_setup.h:9-41): every vertex takes
    // `freedom`, then each vertex on model-bounding facets (back == -1) is classed by the
    // dominant normal axes of those facets (z; +-y; +-x): three classes FIXED; two: X_ONLY
    // (z and y), Y_ONLY (z and x), Z_ONLY (x and y); one: XY_FREE (z), X_ONLY (y), Y_ONLY (x);
    // a vertex that is also on a contact facet and on no z-class facet gets z back
    // (FIXED -> XYZ_FREE, X_ONLY -> XZ_FREE, Y_ONLY -> YZ_FREE, XY_FREE -> XYZ_FREE).
    // The reference classes each unit mesh's copy of a shared vertex by that mesh's own
    // facets; here one vertex is classed by all its bounding facets.
    void fixExteriorFaces(VertexFreedom freedom);
    void setAxisScale(int axis, double s);
    double axisScale(int axis) const { return m_axisScale[axis]; }

    // Control points on a 3-D grid of stride*cellSize cells, one per populated cell,
    // the vertex closest to the cell centre (lithology_model.cpp:864-1040); cellSize <= 0
    // estimates it like the reference (smallest non-zero coordinate step among the
    // first 300 positions in lexicographic order, at least 1 m).  Non-control vertices
    // follow their K = 4 nearest control points with inverse-distance weights
    // (surface_mesh.cpp:544-601).  stride <= 0 turns control points off.
    uint32_t setControlPointStride(int stride, double cellSize = 0.0);
    int controlPointStride() const { return m_cpStride; }
    bool isControlPoint(size_t v) const;

    uint32_t totalDofCount() const { return (uint32_t)m_dofVertex.size(); }
    int dofVertex(uint32_t d) const { return m_dofVertex[d]; }
    int dofAxis(uint32_t d) const { return m_dofAxis[d]; }
    Vec assembleParameterVector() const;
    // param -> vertices (lithology_model.cpp:210-248, surface_mesh.cpp:188-209): a vertex
    // moves along each of its parameters' unit directions u (an axis, or an ALONG_VECTOR
    // move vector) until x.u = p s (s: the axis scale, 1 along a vector), then the
    // control-point interpolation of the chosen form
    void applyParameterVector(const Vec& params);
    // vertex gradient (3*n_v) -> parameter gradient (transpose of the map above)
    Vec pullBackVertexGradient(const Vec& vgrad) const;
    // parameter tangent -> vertex tangent (3*n_v): the linear part of
    // applyParameterVector, whose transpose pullBackVertexGradient is
    Vec vertexTangent(const Vec& dparams) const;
    // The linear map itself, per parameter: parameter d moves coordinate dofAxis(d) of
    // vertex vert[e] by w[e] (metres per parameter unit) for e in [off[d], off[d+1]).
    // Defined for axis parameters only (an ALONG_VECTOR parameter moves a vertex along a
    // vector, not one coordinate): throws if the model has one.
    void dofVertexMap(std::vector<int64_t>& off, std::vector<int32_t>& vert,
                      std::vector<double>& w) const;
    // params -/+ margin/axisScale (runner.cpp:241-248)
    void getBounds(const Vec& params, double margin_m, Vec& lower, Vec& upper) const;
    // applyModelBounds + getBounds (lithology_model.cpp:249-282, 802-846): every
    // parameter inside the model's bounding box of the current geometry grown by
    // marginMultiplier percent of its largest side (at least 1 m), per axis
    void getModelBounds(double marginMultiplier, Vec& lower, Vec& upper) const;

private:
    void rebuildDofs();
    bool dofAlongVector(size_t d) const;
    std::vector<double> m_vertices, m_rest;  // current and initial positions
    std::vector<uint32_t> m_faces;
    std::vector<int32_t> m_front, m_back;
    std::vector<LithoGroup> m_groups;
    std::vector<VertexFreedom> m_freedom;
    std::vector<double> m_moveVector;  // 3 per vertex, unit length or zero
    ControlPointInterpolation m_cpInterp = ControlPointInterpolation::Displacement;
    double m_axisScale[3] = {1.0, 1.0, 1.0};
    int m_cpStride = 0;
    std::vector<uint8_t> m_isCP;
    // non-CP vertex -> (cp vertex, weight) x K, CSR
    std::vector<int64_t> m_wOff;
    std::vector<int32_t> m_wCp;
    std::vector<double> m_w;
    // its transpose: CP vertex -> (dependent vertex, weight), CSR (for the gradient's pull-back)
    std::vector<int64_t> m_depOff;
    std::vector<int32_t> m_depV;
    std::vector<double> m_depW;
    void buildDependants();
    std::vector<int32_t> m_dofVertex, m_dofAxis;
};

// modules/litho-forward/include/litho_invert/forward/forward_model.h:6-33
class ForwardModel {
public:
    virtual ~ForwardModel() = default;
    virtual Vec compute(const Vec& params) = 0;
    // Default: central finite differences over params, step m_fdStep
    // (gravity_forward.cpp:298-321).
    virtual Mat computeJacobian(const Vec& params);
    virtual void computeBoth(const Vec& params, Vec& predicted, Mat& jacobian);
    // J^T w without forming J (new; the default forms J).
    virtual Vec computeJTr(const Vec& params, const Vec& w);
    void setFiniteDifferenceStep(double h) { m_fdStep = h; }
    double finiteDifferenceStep() const { return m_fdStep; }
    virtual size_t dataCount() const = 0;
    virtual size_t parameterCount() const = 0;

protected:
    double m_fdStep = 0.01;
};

enum class Precision { FP64 = 0, FP32C = 1 };

// GPU-backed potential-field forward shared by gravity and magnetics.
class GpuForwardBase : public ForwardModel {
public:
    ~GpuForwardBase() override;
    size_t dataCount() const override { return m_nData; }
    size_t parameterCount() const override { return m_model->totalDofCount(); }
    // J via the adjoint, one station at a time (for small problems and tests).
    Mat computeJacobian(const Vec& params) override;
    Vec computeJTr(const Vec& params, const Vec& w) override;
    const std::string& lastError() const { return m_err; }
    // re-read the unit properties from the model (after they were changed)
    void refreshProperties() { pushProperties(); }
    // upload the model's current vertex positions
    void syncGeometry();

protected:
    GpuForwardBase(std::shared_ptr<LithologyModel> model, std::vector<double> stations,
                   int want, Precision prec, const double fdir[3], int device);
    lf_ctx* ctx();  // creates the device context on first use
    void upload(const Vec& params);
    void pushProperties();
    Vec forwardRaw(int which);  // 0 = gz, 1 = tmi
    virtual void buildProperties(std::vector<double>& dens, std::vector<double>& mag) const = 0;

    std::shared_ptr<LithologyModel> m_model;
    lf_ctx* m_ctx = nullptr;
    size_t m_nData = 0;
    int m_want = 0;
    std::vector<double> m_stations;
    Precision m_prec = Precision::FP64;
    double m_fdir[3] = {0.0, 0.0, 1.0};
    int m_device = 0;
    bool m_textbook = false;
    std::string m_err;
};

// modules/litho-forward/include/litho_invert/forward/gravity_forward.h:8-52
class GravityForward : public GpuForwardBase {
public:
    GravityForward(std::shared_ptr<LithologyModel> model, const GravityData& data,
                   Precision prec = Precision::FP64, int device = 0);
    Vec compute(const Vec& params) override;
    // unit-density response of one group at the current geometry (gravity_forward.cpp:132-145)
    Vec computeGroupUnitResponse(int groupIndex);
    static constexpr double G_SI = 6.67430e-11;
    static constexpr double DENSITY_SCALE = 1000.0;
    static constexpr double M_S2_TO_MGAL = 1e5;

protected:
    void buildProperties(std::vector<double>& dens, std::vector<double>& mag) const override;
    int m_unitOnly = -1;
};

// modules/litho-forward/include/litho_invert/forward/magnetic_forward.h:8-77
class MagneticForward : public GpuForwardBase {
public:
    MagneticForward(std::shared_ptr<LithologyModel> model, const MagneticData& data,
                    double inc_deg, double dec_deg, double field_nT,
                    Precision prec = Precision::FP64, bool textbook = false, int device = 0);
    Vec compute(const Vec& params) override;
    // unit-susceptibility (induced) response of one group at the current geometry
    // (magnetic_forward.cpp:148-162: amplitude F/100 along F-hat)
    Vec computeGroupUnitResponse(int groupIndex);
    // response of one group magnetized with unit amplitude along magDirection
    // (magnetic_forward.cpp:164-178)
    Vec computeGroupUnitResponseDirection(int groupIndex, const double magDirection[3]);
    // the same along the group's remanence direction (magnetic_forward.cpp:180-190)
    Vec computeGroupUnitRemanenceResponse(int groupIndex);
    void buildGroupMagnetization(int g, double& amp, double dir[3]) const;
    void setRemanenceMode(RemanentMagnetizationMode m) { m_mode = m; pushProperties(); }
    RemanentMagnetizationMode remanenceMode() const { return m_mode; }
    const double* earthFieldUnitVector() const { return m_fhat; }
    double fieldStrength() const { return m_field; }

protected:
    void buildProperties(std::vector<double>& dens, std::vector<double>& mag) const override;
    Vec unitResponse(int g, double amp, const double dir[3]);
    double m_fhat[3];
    double m_field;
    RemanentMagnetizationMode m_mode = RemanentMagnetizationMode::EffectiveSusceptibility;
    int m_unitOnly = -1;           // >= 0: only this group is magnetized, m_unitAmp along m_unitDir
    double m_unitAmp = 0.0;
    double m_unitDir[3] = {0.0, 0.0, 1.0};
};

// modules/litho-regularization/include/litho_invert/regularization/regularization.h:6-15
class Regularization {
public:
    virtual ~Regularization() = default;
    virtual double evaluate(const Vec& params) = 0;
    virtual Vec gradient(const Vec& params) = 0;
    void setWeight(double w) { m_weight = w; }
    double weight() const { return m_weight; }

protected:
    double m_weight = 1.0;
};

// modules/litho-regularization/src/regularization/reference_model.cpp:16-25
class ReferenceModelRegularization : public Regularization {
public:
    void captureReference(const Vec& params) { m_ref = params; m_captured = true; }
    double evaluate(const Vec& params) override;
    Vec gradient(const Vec& params) override;

private:
    Vec m_ref;
    bool m_captured = false;
};

// modules/litho-regularization/src/regularization/smoothness.cpp:8-110 (penalty) and
// :112-388 (gradient): for every unit's closed surface and every non-FIXED vertex on it,
// 0.5 |v - mean of its neighbours on that surface with the same freedom|^2 over the
// vertex's free axes, summed over units (a contact vertex counts once per unit it bounds).
// The gradient is exact through the control-point map (the reference reads the
// per-vertex gradient at control points only).
class SurfaceSmoothness : public Regularization {
public:
    explicit SurfaceSmoothness(std::shared_ptr<LithologyModel> model);
    double evaluate(const Vec& params) override;
    Vec gradient(const Vec& params) override;
    // unweighted penalty at the model's current geometry; vgrad (3 per vertex) if given
    double penalty(std::vector<double>* vgrad) const;

private:
    std::shared_ptr<LithologyModel> m_model;
    // per unit: its vertices and their unique neighbours on the unit's surface (CSR)
    std::vector<std::vector<int32_t>> m_unitVerts;
    std::vector<std::vector<int64_t>> m_nbrOff;
    std::vector<std::vector<int32_t>> m_nbr;
};

// A depth-to-contact pick from a borehole: at collar (x, y) the contact between unit
// `above` and unit `below` lies at elevation z (z up, metres) with standard deviation
// sigma.
struct ContactPick {
    double x, y, z;
    int above, below;
    double sigma;
};

// A differentiable stand-in for the reference's borehole penalty, which classifies the
// interval midpoint through the label grid and differentiates that step function with
// 1 m finite differences over every parameter (constraint_handler.cpp:15-66):
// 0.5 sum ((z_model - z)/sigma)^2, z_model the elevation at (x, y) of the model's
// above|below contact surface (a vertical line through the contact facets, the
// intersection nearest z), with the exact gradient through the intersected facet's
// vertices.  Picks whose line misses the contact add nothing.
class BoreholeContactTerm : public Regularization {
public:
    // searchRadius: how far (m, horizontally) the candidate facets of a pick may lie from
    // its collar at the starting geometry
    BoreholeContactTerm(std::shared_ptr<LithologyModel> model, std::vector<ContactPick> picks,
                        double searchRadius = 250.0);
    double evaluate(const Vec& params) override;
    Vec gradient(const Vec& params) override;
    // unweighted misfit at the current geometry; vgrad (3 per vertex) and the model
    // elevation of each pick (NaN where the line misses) if given
    double misfit(std::vector<double>* vgrad, std::vector<double>* zmodel) const;

private:
    std::shared_ptr<LithologyModel> m_model;
    std::vector<ContactPick> m_picks;
    std::vector<std::vector<int64_t>> m_cand;  // candidate contact facets per pick
};

// modules/litho-inversion/src/inversion/objective.cpp:6-103 (gravity data misfit +
// regularization); the gradient uses computeJTr instead of forming J.
class ObjectiveFunction {
public:
    struct Components {
        double dataMisfit = 0.0, regularization = 0.0, constraintPenalty = 0.0, total = 0.0;
    };
    ObjectiveFunction(std::shared_ptr<ForwardModel> forward, const GravityData& data);
    void addRegularization(std::shared_ptr<Regularization> r) { m_regs.push_back(std::move(r)); }
    double dataMisfit(const Vec& params);
    Vec residuals(const Vec& params);
    Vec dataMisfitGradient(const Vec& params);
    double evaluate(const Vec& params) { return evaluateComponents(params).total; }
    Vec gradient(const Vec& params);
    Components evaluateComponents(const Vec& params);

private:
    std::shared_ptr<ForwardModel> m_forward;
    const GravityData& m_data;
    std::vector<std::shared_ptr<Regularization>> m_regs;
};

// modules/litho-inversion/src/inversion/joint_objective.cpp:42-173 (gravity + alpha *
// magnetic misfit + regularization)
class JointObjective {
public:
    struct JointComponents {
        double gravityMisfit = 0.0, magneticMisfit = 0.0, regularization = 0.0,
               constraintPenalty = 0.0, total = 0.0;
    };
    JointObjective(std::shared_ptr<ObjectiveFunction> grav) : m_grav(std::move(grav)) {}
    void setMagnetic(std::shared_ptr<ForwardModel> fwd, const MagneticData& data, double alpha);
    double evaluate(const Vec& params) { return evaluateComponents(params).total; }
    Vec gradient(const Vec& params);
    JointComponents evaluateComponents(const Vec& params);
    double alphaMag() const { return m_alpha; }
    void setAlphaMag(double a) { m_alpha = a; }

private:
    std::shared_ptr<ObjectiveFunction> m_grav;
    std::shared_ptr<ForwardModel> m_mag;
    const MagneticData* m_magData = nullptr;
    double m_alpha = 0.0;
};

// JointObjective on one GPU context for gravity and magnetic data observed at the same
// stations: one forward pass yields gz and TMI (the pair geometry is shared) and one
// adjoint pass takes both weight vectors.  The last forward is cached, so gradient(x)
// right after evaluate(x) -- what the line search does on an accepted step -- costs
// only the adjoint.  Same objective as JointObjective (joint_objective.cpp:113-173).
class FusedJointObjective {
public:
    FusedJointObjective(std::shared_ptr<LithologyModel> model, const GravityData& gdata,
                        const MagneticData& mdata, double inc_deg, double dec_deg,
                        double field_nT, double alpha_mag, Precision prec, int device = 0);
    ~FusedJointObjective();
    void addRegularization(std::shared_ptr<Regularization> r) { m_regs.push_back(std::move(r)); }
    JointObjective::JointComponents evaluateComponents(const Vec& params);
    double evaluate(const Vec& params) { return evaluateComponents(params).total; }
    Vec gradient(const Vec& params);
    void forward(const Vec& params, Vec& gz, Vec& tmi);
    int forwardPasses() const { return m_nForward; }
    int adjointPasses() const { return m_nAdjoint; }
    // re-read the unit properties from the model (after they were changed)
    void refreshProperties();
    double alphaMag() const { return m_alpha; }
    void setAlphaMag(double a) { m_alpha = a; }
    // how unit magnetizations are built (groupMagnetization); re-reads the properties
    void setRemanenceMode(RemanentMagnetizationMode m) { m_mode = m; refreshProperties(); }
    // sum_i wg[i] dgz_i/dp + wt[i] dTMI_i/dp at params (either weight vector may be empty)
    Vec jtw(const Vec& params, const Vec& wg, const Vec& wt);
    // J dp at params: the change of gz and TMI along the parameter tangent dp (lf_jvp)
    void jvp(const Vec& params, const Vec& dp, Vec& gz, Vec& tmi);

private:
    void ensureForward(const Vec& params);
    double m_field = 50000.0;
    double m_fhat[3] = {0.0, 0.0, 1.0};
    RemanentMagnetizationMode m_mode = RemanentMagnetizationMode::EffectiveSusceptibility;
    std::shared_ptr<LithologyModel> m_model;
    const GravityData& m_g;
    const MagneticData& m_m;
    double m_alpha;
    lf_ctx* m_ctx = nullptr;
    std::vector<std::shared_ptr<Regularization>> m_regs;
    Vec m_cacheParams, m_gz, m_tmi;
    bool m_cacheValid = false;
    int m_nForward = 0, m_nAdjoint = 0;
};

// The property phase of the reference's inversion loop, modules/litho-inversion/src/
// inversion/runner.cpp:708-800 (densities from gravity) and :801-935 (magnetic properties
// from TMI): unit responses U (one forward per unit and column with a unit property), then
// a bounded L-BFGS on
//   0.5 |W (d_obs - U p)|^2 / rms(d_obs)^2 + 0.5 damping |p - p_start|^2
// (rms floored at 1, W = 1/std or 1), history min(5, n_params).  The result is written into
// the model's units.  Magnetic parameters per unit follow the forward's remanence mode
// (runner.cpp:808-857): chi (EffectiveSusceptibility); chi and the remanence magnitude
// (FixedVectorPerGroup, columns: unit-chi response, unit response along the unit's
// remanence direction); chi and Mx, My, Mz (VectorPerGroup, columns: unit-chi response,
// unit responses along x, y, z).  Bounds: [lower, upper] for densities and
// susceptibilities, [remanenceLower, remanenceUpper] for magnitudes, [componentLower,
// componentUpper] for components (runner.h:94-107 defaults 0..10 and -10..10).
struct PropertyInversionOptions {
    double damping = 0.0;
    double lower = -1e300, upper = 1e300;
    int maxIter = 50;
    double tolerance = 1e-6;
    double remanenceLower = 0.0, remanenceUpper = 10.0;
    double componentLower = -10.0, componentUpper = 10.0;
};
Vec invertUnitDensities(GravityForward& fwd, const GravityData& data, LithologyModel& model,
                        const PropertyInversionOptions& opt);
// returns the parameters in the layout above (params per unit x units)
Vec invertUnitSusceptibilities(MagneticForward& fwd, const MagneticData& data,
                               LithologyModel& model, const PropertyInversionOptions& opt);

// JointObjective::calibrateDataWeights (joint_objective.cpp:281-393) without forming J:
// power iteration (at most 50 steps, relative tolerance 1e-8) for the largest eigenvalue
// of J^T W^2 J of gravity and of magnetics (W = 1/std, or 1 where std is 0), J^T u from
// the adjoint and J v either exactly, from the Jacobian-vector product kernel in the
// given precision (eps <= 0), or from central differences of an FP64 forward (step eps in
// parameter units along the unit vector v; then the FP64 adjoint too, whatever the
// precision); alpha_mag is multiplied by the eigenvalue ratio
// clipped to [1e-6, 1e6], or, if the ratio is unusable, by the ratio of the two data-misfit
// gradient norms.  The power iterations start from a seeded random unit vector (the
// reference uses Eigen's unseeded Random()).
struct CalibrationResult {
    double eigGravity = 0.0, eigMagnetic = 0.0, ratio = 0.0, alpha = 0.0;
    int iterationsGravity = 0, iterationsMagnetic = 0;
    bool fallback = false;
};
CalibrationResult calibrateMagneticWeight(std::shared_ptr<LithologyModel> model,
                                          const GravityData& gdata, const MagneticData& mdata,
                                          double inc_deg, double dec_deg, double field_nT,
                                          const Vec& params, double alpha_mag,
                                          uint64_t seed = 1, double eps = 0.0, int device = 0,
                                          RemanentMagnetizationMode mode =
                                              RemanentMagnetizationMode::EffectiveSusceptibility,
                                          Precision prec = Precision::FP64);

// modules/litho-inversion/include/litho_invert/inversion/optimizer.h:7-36
struct OptimizerResult {
    Vec params;
    double finalValue = 0.0;
    int iterations = 0;
    bool converged = false;
    std::string message;
};

// The interface of the geometry solvers (optimizer.h:16-36): minimize over a box, restart
// (drops what the solver remembers; the next minimize starts from its own x0, so
// newParams changes nothing else, as in the reference), the iteration limit and count,
// the projected-gradient tolerance.
class Optimizer {
public:
    using Objective = std::function<double(const Vec&)>;
    using Gradient = std::function<Vec(const Vec&)>;
    virtual ~Optimizer() = default;
    virtual OptimizerResult minimize(Objective objective, Gradient gradient, const Vec& x0,
                                     const Vec& lb, const Vec& ub) = 0;
    virtual void restart(const Vec& newParams = Vec()) = 0;
    virtual void setMaxIterations(int n) = 0;
    virtual int iterationCount() const = 0;
    virtual void setTolerance(double t) = 0;
};

// Projected L-BFGS with Armijo backtracking: modules/litho-inversion/src/inversion/
// lbfgsb_optimizer.cpp:13-263, with an optional diagonal initial inverse Hessian D:
// H0 = gamma * D, gamma = s^T y / (y^T D y) (gamma * I with D = I, :40-46), and a first
// step along -D g.
class LBFGSBOptimizer : public Optimizer {
public:
    struct IterationRecord {
        int iteration;
        double f, pgNorm, alpha, seconds;
        int objectiveCalls, gradientCalls;
    };
    OptimizerResult minimize(Objective objective, Gradient gradient, const Vec& x0,
                             const Vec& lb, const Vec& ub) override;
    void setMaxIterations(int n) override { m_maxIter = n; }
    void setTolerance(double t) override { m_ftol = t; }
    void setHistorySize(int m) { m_historySize = m; }
    void setArmijoC1(double c) { m_c1 = c; }
    void setLineSearchMaxIter(int n) { m_lsMax = n; }
    void setInitialHessianDiagonal(const Vec& d) { m_h0 = d; }
    // stop once f <= target (used to time "seconds to a target misfit")
    void setTargetObjective(double f) { m_target = f; m_hasTarget = true; }
    // Variables sitting on a bound with the gradient pushing outward are left out of the
    // two-loop recursion and the step (off by default: the reference projects the full
    // L-BFGS step, which can stop descending once bounds are active).
    void setActiveSetAware(bool on) { m_activeSet = on; }
    // InversionConfig::disableLineSearch (lbfgsb_optimizer.cpp:90-96): every iteration
    // takes the full projected step without evaluating the objective; f is evaluated once
    // at the end (history records before that carry NaN; a target is not checked)
    void setDisableLineSearch(bool on) { m_noLineSearch = on; }
    // Warm start (lbfgsb_optimizer.h:24-37): with clearOnMinimize false the (s, y, rho)
    // pairs survive minimize() and the next run's first direction comes from them;
    // getHistory / setHistory move them between optimizers (oldest first); restart
    // clears them.  setReferenceObjective is kept for the interface: the reference stores
    // the value and its convergence test does not read it (lbfgsb_optimizer.cpp:186-189).
    void setClearOnMinimize(bool v) { m_clearOnMinimize = v; }
    bool clearOnMinimize() const { return m_clearOnMinimize; }
    void setHistory(const std::vector<Vec>& s, const std::vector<Vec>& y,
                    const std::vector<double>& rho);
    void getHistory(std::vector<Vec>& s, std::vector<Vec>& y, std::vector<double>& rho) const
    {
        s = m_s;
        y = m_y;
        rho = m_rho;
    }
    bool hasHistory() const { return !m_s.empty(); }
    void restart(const Vec& newParams = Vec()) override;
    int historySize() const { return m_historySize; }
    void setReferenceObjective(double f) { m_fRef = f; }
    double referenceObjective() const { return m_fRef; }
    // called with each iteration's record as it is appended (record 0: the start)
    using IterationCallback = std::function<void(const IterationRecord&)>;
    void setIterationCallback(IterationCallback cb) { m_onIteration = std::move(cb); }
    int iterationCount() const override { return m_iter; }
    const std::vector<IterationRecord>& history() const { return m_hist; }

private:
    Vec direction(const Vec& g, const std::vector<uint8_t>& freeMask) const;
    IterationCallback m_onIteration;
    std::vector<Vec> m_s, m_y;
    std::vector<double> m_rho;
    Vec m_h0;
    int m_maxIter = 500, m_historySize = 10, m_lsMax = 50, m_iter = 0;
    double m_ftol = 1e-6, m_c1 = 1e-4, m_target = 0.0;
    bool m_hasTarget = false;
    bool m_activeSet = false;
    bool m_noLineSearch = false;
    bool m_clearOnMinimize = true;
    double m_fRef = 1.0;
    std::vector<IterationRecord> m_hist;
};

// Projected Gauss-Newton with an inner conjugate-gradient solve:
// modules/litho-inversion/src/inversion/gncg_optimizer.cpp:13-194 (the runner's
// solver = "gncg", runner.cpp:192-202).  Each outer iteration solves H d = -g by CG, where
// H v is the central difference of the gradient at x -/+ (fd_step / |v|) v (both points
// projected onto the bounds), falls back to -g when d is not a descent direction, and
// halves the step from 1 until the projected point does not raise the objective.  The
// Armijo constant is kept for the interface; the reference's line search does not use it.
class GNCGOptimizer : public Optimizer {
public:
    // one row per iterate x_k once both f(x_k) and the gradient there are known; the
    // final row, written when the loop ends at a point whose gradient was not evaluated,
    // has pgNorm = NaN.  alpha: the step that reached x_k.
    using IterationRecord = LBFGSBOptimizer::IterationRecord;
    using IterationCallback = std::function<void(const IterationRecord&)>;
    OptimizerResult minimize(Objective objective, Gradient gradient, const Vec& x0,
                             const Vec& lb, const Vec& ub) override;
    void restart(const Vec& = Vec()) override { m_iter = 0; }
    void setMaxIterations(int n) override { m_maxIter = n; }
    void setTolerance(double t) override { m_ftol = t; }
    void setCGMaxIter(int n) { m_cgMaxIter = n; }
    void setCGTolerance(double t) { m_cgTol = t; }
    void setFDStep(double h) { m_fdStep = h; }
    void setArmijoC1(double c) { m_c1 = c; }
    void setLineSearchMaxIter(int n) { m_lsMax = n; }
    void setIterationCallback(IterationCallback cb) { m_onIteration = std::move(cb); }
    int iterationCount() const override { return m_iter; }
    // inner CG iterations of each outer iteration, in order
    const std::vector<int>& cgIterations() const { return m_cgIters; }
    const std::vector<IterationRecord>& history() const { return m_hist; }

private:
    Vec cgSolve(const std::function<Vec(const Vec&)>& Hv, const Vec& b, int& iters) const;
    IterationCallback m_onIteration;
    int m_maxIter = 500, m_cgMaxIter = 50, m_lsMax = 30, m_iter = 0;
    double m_ftol = 1e-6, m_cgTol = 1e-6, m_fdStep = 2.0, m_c1 = 1e-4;
    std::vector<int> m_cgIters;
    std::vector<IterationRecord> m_hist;
};

}  // namespace lithofwd
