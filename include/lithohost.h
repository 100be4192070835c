/*
 * lithohost.h -- C ABI of the host layer (liblithohost.so) that sits above lithofwd.h: the
 * lithology model with its parameter vector, the reference's ForwardModel /
 * ObjectiveFunction / JointObjective / L-BFGS-B surface over the GPU forward and adjoint,
 * the regularizers, the property phase, and the file formats around the forward path.
 *
 * Every handle is an opaque pointer made by a *_create / *_load / li_assemble call and
 * released by the matching *_destroy.  Entry points returning int give 0 on success and
 * -1 on failure, li_last_error() then describes it (the message of the C++ exception,
 * thread-local).  Arrays are host pointers, row-major, 0-based indices; xyz triples are
 * metres with z positive up.  A handle is not thread-safe.
 */
#ifndef LITHOHOST_H
#define LITHOHOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* li_last_error(void);

/* ---- lithology model (modules/litho-model/src/litho/lithology_model.cpp) ---- */

/* LithologyModel from a shared-vertex facet set: faces oriented outward for `front`, `back`
 * the unit on the other side (-1 outside), per-unit density contrast (g/cc) and
 * susceptibility (SI); litho_group.h:7-34 holds the same unit properties. */
int li_model_create(void** out, int64_t n_v, const double* verts, int64_t n_f,
                    const uint32_t* faces, const int32_t* front, const int32_t* back,
                    int32_t n_u, const double* density, const double* susceptibility);
void li_model_destroy(void* h);
/* current per-unit density and susceptibility (n_units each; LithoGroup, litho_group.h) */
int li_model_properties(void* h, double* density, double* susceptibility);
/* LithoGroup's remanence fields (litho_group.h:18-24), n_units each (vector: 3 per unit):
 * magnitude (A/m, reference units), inclination and declination (degrees, inclination
 * positive down), and the remanent vector; a NULL array keeps / skips that field */
int li_model_set_remanence(void* h, const double* magnitude, const double* inc_deg,
                           const double* dec_deg, const double* vector);
int li_model_remanence(void* h, double* magnitude, double* inc_deg, double* dec_deg,
                       double* vector);
/* fixExteriorFaces (lithology_model.cpp:369-579): vertices of model-bounding facets FIXED,
 * the rest `interior_freedom` (VertexFreedom of surface_mesh.h: 0 FIXED, 1 Z_ONLY,
 * 2 X_ONLY, 3 Y_ONLY, 8 XYZ_FREE). */
int li_model_fix_exterior(void* h, int interior_freedom);
/* finalizeModelSetup's freedom reset then fixExteriorFaces (apps/forrestania/model_setup.h:9-41,
 * lithology_model.cpp:369-579): every vertex takes `freedom`, then vertices on
 * model-bounding facets are classed by those facets' dominant normal axes (corners FIXED,
 * edges one axis, faces two; contact vertices off the top and bottom get z back);
 * codes 5 XY_FREE, 6 XZ_FREE, 7 YZ_FREE besides those above */
int li_model_fix_exterior_faces(void* h, int freedom);
/* the current freedom of each vertex (n_v values) */
int li_model_freedom(void* h, uint8_t* freedom);
/* InversionExporter::exportClosedVolume over the units at the current geometry
 * (exporters.cpp:769-813): each unit's closed surface -- its front facets as stored, its
 * back facets reversed, its vertices in ascending shared index -- as the GOCAD TSurf
 * <dir>/<base>_<name>.ts (header name <base>_<name>), name = names[g] for g < n_names, else
 * "group_<g>"; units without facets are skipped; dir is created.  n_written may be NULL. */
int li_model_export_units(void* h, const char* dir, const char* base, const char* const* names,
                          int32_t n_names, int32_t* n_written);
/* setVertexFreedom per vertex (n_v values, same codes, plus 4 ALONG_VECTOR: one
 * parameter, the vertex's projection on its move vector, surface_mesh.cpp:125-128,
 * 194-196) */
int li_model_set_freedom(void* h, const uint8_t* freedom);
/* the move vectors of ALONG_VECTOR vertices (3*n_v; used normalized, zero = no motion;
 * surface_mesh.h Vertex::moveVector) */
int li_model_set_move_vectors(void* h, const double* move_vectors);
/* how non-control vertices follow the control points: 0 displacement form (default: the
 * control points' IDW-weighted displacement along the vertex's own free directions), 1 the
 * reference's position form (interpolateFromControlPoints, surface_mesh.cpp:610-623: every
 * non-control vertex, FIXED ones included, becomes the IDW-weighted average of its control
 * points' positions, even for an unchanged parameter vector) */
int li_model_set_cp_interpolation(void* h, int mode);
/* axis scale of the parameter vector (runner.cpp:241-248: params in units of s metres) */
int li_model_set_axis_scale(void* h, int axis, double s);
/* control points on a stride*cell grid (lithology_model.cpp:864-1040, IDW K = 4 of
 * surface_mesh.cpp:544-601); stride <= 0 turns them off; *n_dof = new parameter count */
int li_model_set_cp_stride(void* h, int stride, double cell, int64_t* n_dof);
int64_t li_model_dof_count(void* h);
/* per DOF: its vertex and axis; per vertex: is it a control point */
int li_model_dofs(void* h, int32_t* vertex, int32_t* axis, uint8_t* is_cp);
/* assembleParameterVector (lithology_model.cpp:194-209) */
int li_model_params(void* h, double* out);
/* applyParameterVector (lithology_model.cpp:210-248); vertices_out (3*n_v) may be NULL */
int li_model_apply(void* h, const double* params, double* vertices_out);
/* vertex gradient (3*n_v) -> parameter gradient, the transpose of li_model_apply's map */
int li_model_pullback(void* h, const double* vgrad, double* out);
/* parameter tangent -> vertex tangent (3*n_v): the linear part of li_model_apply's map,
 * whose transpose li_model_pullback is (no counterpart in the reference, which forms
 * Jacobian columns by moving one parameter at a time) */
int li_model_tangent(void* h, const double* dparams, double* out);
/* that map as CSR per parameter; call with vert and w NULL for the entry count first */
int li_model_dof_map(void* h, int64_t* off, int32_t* vert, double* w, int64_t* n_entries);
/* geometry bounds params -/+ margin/axisScale (runner.cpp:241-248) */
int li_model_bounds(void* h, const double* params, double margin_m, double* lo, double* hi);
/* applyModelBounds + getBounds (lithology_model.cpp:249-282, 802-846): the model box of
 * the current geometry grown by margin_multiplier percent of its largest side */
int li_model_model_bounds(void* h, double margin_multiplier, double* lo, double* hi);
/* SmoothnessRegularization (smoothness.cpp:8-388): weight * penalty at params and its
 * parameter gradient (grad may be NULL) */
int li_model_smoothness(void* h, const double* params, double weight, double* value,
                        double* grad);
/* Depth-to-contact picks (constraint_handler.cpp:15-66, made differentiable): weight *
 * 0.5 sum ((z_model - z)/sigma)^2 at params, its gradient and each pick's model
 * elevation (NaN where the vertical line misses the contact); grad, zmodel may be NULL */
int li_model_boreholes(void* h, int64_t n, const double* xyz, const int32_t* above,
                       const int32_t* below, const double* sigma, double radius,
                       const double* params, double weight, double* value, double* grad,
                       double* zmodel);

/* ---- joint gravity + magnetic problem (modules/litho-inversion) ---- */

/* GravityForward + ObjectiveFunction (objective.cpp:6-103), MagneticForward and
 * JointObjective (joint_objective.cpp:42-173) when tmi_obs is given and alpha_mag > 0, on
 * one station set (n_s xyz triples); *_std <= 0 means unweighted; inc/dec in degrees,
 * field in nT; precision 0 FP64, 1 FP32C; ref_weight > 0 adds ReferenceModelRegularization
 * (reference_model.cpp:16-25) at the model's current parameters. */
int li_problem_create(void** out, void* model, int64_t n_s, const double* stations,
                      const double* gz_obs, const double* gz_std, const double* tmi_obs,
                      const double* tmi_std, double inc, double dec, double field,
                      double alpha_mag, int precision, double ref_weight, int device);
void li_problem_destroy(void* h);
/* adds SmoothnessRegularization / the depth-to-contact term to the objective */
int li_problem_add_smoothness(void* h, double weight);
int li_problem_add_boreholes(void* h, int64_t n, const double* xyz, const int32_t* above,
                             const int32_t* below, const double* sigma, double radius,
                             double weight);
/* the property phase (runner.cpp:708-800 densities, :801-935 magnetic properties): which 0
 * densities from gravity (out: n_units values), 1 magnetic properties from TMI in the
 * problem's remanence mode (out: n_units x {chi} | {chi, remanence magnitude} |
 * {chi, Mx, My, Mz}, runner.cpp:808-857); the model and every forward of the problem take
 * the new values.  lower/upper bound densities and susceptibilities; remanence parameters
 * take li_problem_set_remanence_bounds. */
int li_problem_invert_properties(void* h, int which, double damping, double lower, double upper,
                                 int max_iter, double tol, double* out);
/* MagneticForward::setRemanenceMode (magnetic_forward.h:61; litho_group.h:7-11): 0
 * EffectiveSusceptibility, 1 FixedVectorPerGroup, 2 VectorPerGroup, for every magnetic
 * forward, objective, calibration and property phase of the problem */
int li_problem_set_remanence_mode(void* h, int mode);
/* bounds of the remanence magnitude and of the remanent vector's components in the
 * magnetic property phase (runner.h:102-107; defaults 0..10 and -10..10) */
int li_problem_set_remanence_bounds(void* h, double magnitude_lower, double magnitude_upper,
                                    double component_lower, double component_upper);
/* JointObjective::calibrateDataWeights (joint_objective.cpp:281-393), matrix-free: sets
 * alpha_mag from the ratio of the largest eigenvalues of J^T W^2 J of gravity and of
 * magnetics at params.  out[7]: both eigenvalues, their ratio, the new alpha_mag, the two
 * power-iteration counts, 1 if the gradient-norm fallback ran.  eps <= 0: J v exactly,
 * one Jacobian-vector product pass (lf_jvp) in the problem's precision; eps > 0: J v by
 * central differences of two FP64 forwards with that step in parameter units (and J^T u
 * in FP64).  seed: the power iterations' start vectors.  li_problem_run calibrates with
 * eps = 0 and seed 1. */
int li_problem_calibrate(void* h, const double* params, uint64_t seed, double eps, double* out);
/* J dp at params for a problem with magnetic data: gz_out and tmi_out (n_stations each)
 * the change of the data along the parameter tangent dp, one lf_jvp pass (the product the
 * reference's calibrateDataWeights forms with its dense Jacobians, joint_objective.cpp:
 * 281-393) */
int li_problem_jvp(void* h, const double* params, const double* dparams, double* gz_out,
                   double* tmi_out);
/* 1 (default with magnetic data): one GPU context for gz and TMI, one adjoint pass and a
 * cached forward; 0: the reference-shaped JointObjective over separate forwards */
int li_problem_set_fused(void* h, int on);
/* forward and adjoint passes run so far */
int li_problem_pass_counts(void* h, int64_t* fwd, int64_t* adj);
/* JointObjective::evaluateComponents: comp[4] = gravity misfit, magnetic misfit (before
 * alpha_mag), regularization, total */
int li_problem_evaluate(void* h, const double* params, double* comp);
/* JointObjective::gradient (joint_objective.cpp:113-173) */
int li_problem_gradient(void* h, const double* params, double* out);
/* ForwardModel::compute (forward_model.h:6-33): gz (mGal) and TMI (nT), tmi may be NULL */
int li_problem_forward(void* h, const double* params, double* gz, double* tmi);
/* ForwardModel::computeJacobian (gravity_forward.cpp:269-288): which 0 gz, 1 TMI;
 * n_s x n_dof, column-major */
int li_problem_jacobian(void* h, const double* params, int which, double* J_colmajor);
/* unit-property response of one group at the current geometry: which 0 unit density
 * (gravity_forward.cpp:132-145), 1 unit susceptibility (magnetic_forward.cpp:148-162) */
int li_problem_unit_response(void* h, int which, int group, double* out);
/* MagneticForward::computeGroupUnitResponseDirection (magnetic_forward.cpp:164-178): TMI of
 * one group magnetized with unit amplitude along direction[3] */
int li_problem_unit_response_direction(void* h, int group, const double* direction, double* out);
/* MagneticForward::computeGroupUnitRemanenceResponse (magnetic_forward.cpp:180-190): the
 * same along the group's remanence inclination / declination */
int li_problem_unit_remanence_response(void* h, int group, double* out);

/* ---- the inversion loop (runner.cpp:217-641) ---- */

/* InversionConfig (runner.h:27-170), the fields the loop uses */
typedef struct {
    int32_t max_iterations;      /* geometry iterations in total */
    int32_t property_interval;   /* geometry iterations between property phases */
    int32_t geometry;            /* 0: property phases only, one round */
    int32_t properties;          /* alternate with density / susceptibility phases */
    int32_t eigenvalue_scaling;  /* calibrate alpha_mag first (li_problem_calibrate) */
    int32_t validate;            /* validation gate after each geometry phase */
    int32_t depth_bounds;        /* also bound each parameter to start -/+ depth_margin */
    int32_t lbfgs_history;
    int32_t line_search_max_iter;
    int32_t property_max_iter;
    int32_t model_bounds;        /* 1: bounds from the model box grown by
                                    model_bounds_multiplier % (applyModelBounds);
                                    0: start -/+ bounds_margin */
    double model_bounds_multiplier;
    double depth_margin;         /* parameter units */
    double bounds_margin;        /* metres (li_model_bounds) */
    double gradient_smoothing;   /* smoothGradient weight (runner.cpp:643-706), 0 = off */
    double tolerance;
    double armijo_c1;
    double density_lower, density_upper, density_damping;
    double susceptibility_lower, susceptibility_upper, susceptibility_damping;
    double property_tolerance;
    int32_t solver;              /* geometry solver (runner.h:49-51, runner.cpp:192-212):
                                    0 L-BFGS-B, 1 Gauss-Newton CG (li_gncg_*) */
    int32_t gncg_cg_max_iter;    /* inner CG iterations (runner.h:54, default 50) */
    double gncg_cg_tolerance;    /* CG residual tolerance (runner.h:55, default 1e-6) */
    double fd_step;              /* Hessian-vector difference step (runner.h:39, default 1) */
    int32_t disable_line_search; /* L-BFGS-B takes the full projected step (runner.h:115-116) */
    /* iterationExportDir / groupExportNames (runner.h:139-155, runner.cpp:319-348): when
     * non-NULL and non-empty, every accepted geometry phase writes each unit's closed
     * surface to <dir>/iter_NNN/invert_<name>.ts, name = group_export_names[g] for g below
     * n_group_export_names, else "group_<g>" (li_model_export_units) */
    const char* iteration_export_dir;
    const char* const* group_export_names;
    int32_t n_group_export_names;
} LiRunnerOptions;
/* hist rows of 10: iteration, gravity misfit, magnetic misfit, regularization, total,
 * gravity RMS, Durbin-Watson gravity x, y, magnetic x, y (InversionIteration,
 * inversion_result.h:10-28), one per objective evaluation at new parameters;
 * summary[8]: converged, geometry iterations, final objective, final gravity RMS,
 * calibrated alpha_mag (0: none), validation retries, rejections, history entries in
 * total; params_out: the final parameters; log: one line per phase and event.
 * label_grid (a li_label_grid_load handle) may be NULL. */
int li_problem_run(void* h, const LiRunnerOptions* o, void* label_grid, double* params_out,
                   double* hist, int64_t hist_cap, int64_t* n_hist, double* summary, char* log,
                   int64_t log_cap);
/* classifyPoint (lithology_model.cpp:288-362): the label grid's unit where it holds one,
 * else the lowest unit whose surface a +x ray crosses an odd number of times, else -1
 * (and -1 below bottom_depth, the reference's default -5000) */
int li_model_classify(void* h, void* label_grid, double bottom_depth, int64_t n,
                      const double* xyz, int32_t* out);
/* LithologyModel::validate (lithology_model.cpp:624-798) at the current geometry:
 * counts[2] = adjacent and non-adjacent overlaps; negative baselines: counts only */
int li_model_validate(void* h, void* label_grid, double bottom_depth, int64_t overlap_baseline,
                      int64_t adjacent_baseline, int32_t* passed, double* worst,
                      int64_t* counts, char* reason, int64_t reason_cap);
/* computeDurbinWatson (modules/litho-core/src/core/stats.cpp:31-85): out[2] = x, y */
int li_durbin_watson(int64_t n, const double* xyz, const double* residuals, double* out);

/* ---- L-BFGS-B (lbfgsb_optimizer.cpp:13-263, optimizer.h:7-36) ---- */

/* Iteration callback of a problem (InversionRunner::setIterationCallback, runner.h:200-201):
 * called on the calling thread with each history row as it is recorded -- from
 * li_problem_minimize the 7 values of its hist rows, from li_problem_run the 10 values of
 * its hist rows; kind is 0 for a minimize row and 1 for a runner row. */
typedef void (*li_iter_cb)(int32_t kind, const double* row, int64_t n_row, void* user);
typedef double (*li_obj_cb)(const double* x, int64_t n, void* user);
typedef void (*li_grad_cb)(const double* x, int64_t n, double* g, void* user);
typedef struct {
    int32_t max_iter;
    int32_t history;
    double ftol;
    double armijo_c1;
    double target;  /* stop when f <= target; NaN: none */
    int32_t has_h0; /* h0 holds a diagonal initial inverse Hessian */
    int32_t active_set;
    int32_t disable_line_search; /* the full projected step, no objective calls in the
                                    loop (lbfgsb_optimizer.cpp:90-96); target ignored */
} LiOptions;
/* hist rows of 7: iteration, f, |projected gradient|, step, seconds, objective calls,
 * gradient calls; msg: the stop reason */
int li_minimize_callbacks(int64_t n, const double* x0, const double* lb, const double* ub,
                          li_obj_cb f, li_grad_cb g, void* user, const LiOptions* o,
                          const double* h0, double* x_out, double* hist, int64_t hist_cap,
                          int64_t* n_hist, int32_t* iterations, char* msg, int64_t msg_cap);
/* An L-BFGS-B optimizer that keeps its (s, y, rho) history between calls: the warm start
 * of lbfgsb_optimizer.h:24-37.  set_clear_on_minimize(0) (setClearOnMinimize(false))
 * keeps the history through li_lbfgsb_minimize_callbacks (options and outputs as
 * li_minimize_callbacks); get/set_history move it between optimizers as count pairs,
 * s and y count x n doubles oldest first (get: pass NULL arrays to read the count);
 * restart clears it. */
int li_lbfgsb_create(void** out);
void li_lbfgsb_destroy(void* h);
int li_lbfgsb_set_clear_on_minimize(void* h, int32_t clear);
int li_lbfgsb_minimize_callbacks(void* h, int64_t n, const double* x0, const double* lb,
                                 const double* ub, li_obj_cb f, li_grad_cb g, void* user,
                                 const LiOptions* o, const double* h0, double* x_out,
                                 double* hist, int64_t hist_cap, int64_t* n_hist,
                                 int32_t* iterations, char* msg, int64_t msg_cap);
int li_lbfgsb_get_history(void* h, int64_t n, int32_t* count, double* s, double* y, double* rho);
int li_lbfgsb_set_history(void* h, int64_t n, int32_t count, const double* s, const double* y,
                          const double* rho);
int li_lbfgsb_restart(void* h);
/* sets (cb non-NULL) or clears the problem's iteration callback */
int li_problem_set_iteration_callback(void* h, li_iter_cb cb, void* user);
int li_problem_minimize(void* h, const double* x0, const double* lb, const double* ub,
                        const LiOptions* o, const double* h0, double* x_out, double* hist,
                        int64_t hist_cap, int64_t* n_hist, int32_t* iterations, char* msg,
                        int64_t msg_cap);

/* ---- Gauss-Newton CG (gncg_optimizer.cpp:13-194, gncg_optimizer.h:6-65) ---- */

/* GNCGOptimizer's settings: outer iterations, inner CG iterations and residual tolerance,
 * the step h = fd_step / |v| of the central-difference Hessian-vector product
 * (g(x + h v) - g(x - h v)) / 2h, the projected-gradient tolerance (relative to
 * max(1, its first value)), the backtracking limit; armijo_c1 is kept for the
 * interface, the reference's line search accepts any step that does not raise f */
typedef struct {
    int32_t max_iter;
    int32_t cg_max_iter;
    double cg_tolerance;
    double fd_step;
    double ftol;
    double armijo_c1;
    int32_t line_search_max_iter;
} LiGncgOptions;
/* GNCGOptimizer::minimize (gncg_optimizer.cpp:105-194) on callbacks.  hist rows of 7 as
 * li_minimize_callbacks, one per iterate once f and the gradient there are known (a final
 * iterate whose gradient was never evaluated has NaN as its projected-gradient norm);
 * cg_iters[k]: the inner CG iterations of outer iteration k (up to cg_cap) */
int li_gncg_minimize_callbacks(int64_t n, const double* x0, const double* lb, const double* ub,
                               li_obj_cb f, li_grad_cb g, void* user, const LiGncgOptions* o,
                               double* x_out, double* hist, int64_t hist_cap, int64_t* n_hist,
                               int32_t* cg_iters, int64_t cg_cap, int32_t* iterations, char* msg,
                               int64_t msg_cap);
/* the same on the problem's objective and GPU gradient; rows go to its iteration callback */
int li_problem_gncg_minimize(void* h, const double* x0, const double* lb, const double* ub,
                             const LiGncgOptions* o, double* x_out, double* hist,
                             int64_t hist_cap, int64_t* n_hist, int32_t* cg_iters,
                             int64_t cg_cap, int32_t* iterations, char* msg, int64_t msg_cap);

/* ConstraintHandler (constraint_handler.cpp:15-66), the reference's discrete borehole
 * penalty: weight * sum of v^2, v = |z_bottom - z_top| when the interval's midpoint
 * (x, y, -(z_top + z_bottom) / 2; depths positive down) does not classify (as
 * li_model_classify, label_grid may be NULL) into `unit`, else 0; gradient by central
 * differences with a step of 1 parameter unit, times the weight again as there.
 * xy: n (x, y) collars.  li_model_constraint_penalty evaluates it at params (violations
 * and grad may be NULL); li_problem_add_constraints adds it to the problem's objective
 * (the grid is copied) */
int li_model_constraint_penalty(void* h, void* label_grid, double bottom_depth, int64_t n,
                                const double* xy, const double* z_top, const double* z_bottom,
                                const int32_t* unit, double weight, const double* params,
                                double* value, double* violations, double* grad);
int li_problem_add_constraints(void* h, void* label_grid, double bottom_depth, int64_t n,
                               const double* xy, const double* z_top, const double* z_bottom,
                               const int32_t* unit, double weight);

/* ---- run configuration files (modules/litho-io/src/io/ini_config.cpp) ---- */

/* IniConfig::load ("[section]", "key = value" or "key: value", '#' / ';' comments);
 * a file that cannot be opened is an error */
int li_ini_load(const char* path, void** out);
/* the same parser on text in memory; text NULL gives an empty configuration */
int li_ini_parse(const char* text, void** out);
void li_ini_destroy(void* h);
/* IniConfig::getString: *found 0 and value "" when absent; value cut to cap - 1 bytes */
int li_ini_get(void* h, const char* section, const char* key, char* value, int64_t cap,
               int32_t* found);
/* getDouble / getInt / getBool: def for a missing or empty value or (numbers) no leading
 * number; booleans true for "true", "yes", "1", "on" in any case */
int li_ini_get_double(void* h, const char* section, const char* key, double def, double* out);
int li_ini_get_int(void* h, const char* section, const char* key, int32_t def, int32_t* out);
int li_ini_get_bool(void* h, const char* section, const char* key, int32_t def, int32_t* out);
/* IniConfig::setString / save */
int li_ini_set(void* h, const char* section, const char* key, const char* value);
int li_ini_save(void* h, const char* path);
/* every entry as "section<TAB>key<TAB>value<LF>" lines, sections and keys sorted; *need
 * gets the size with the terminating zero (call with cap 0 first) */
int li_ini_dump(void* h, char* buf, int64_t cap, int64_t* need);

/* ---- file formats (modules/litho-io) ---- */

/* .ts (surface_loader.cpp:93-148) or .obj (:9-91) by extension (:150-168) */
int li_mesh_load(const char* path, void** out);
int li_mesh_counts(void* h, int64_t* n_vertices, int64_t* n_triangles);
int li_mesh_get(void* h, double* vertices, uint32_t* triangles);
const char* li_mesh_warnings(void* h);
void li_mesh_destroy(void* h);
/* InversionExporter::exportTS (exporters.cpp:116-150) */
int li_mesh_write_ts(const char* path, const char* name, int64_t n_vertices,
                     const double* vertices, int64_t n_triangles, const uint32_t* triangles);
/* gravity_loader.cpp:8-69 / magnetic_loader.cpp:8-69: x, y, z, observation[, std] */
int li_stations_load(const char* path, void** out);
int64_t li_stations_count(void* h);
int li_stations_get(void* h, double* xyz, double* obs, double* stdev);
const char* li_stations_warnings(void* h);
void li_stations_destroy(void* h);
/* constraint_loader.cpp:8-66: x, y, z_top, z_bottom, unit */
int li_constraints_load(const char* path, void** out);
int64_t li_constraints_count(void* h);
int li_constraints_get(void* h, double* xy, double* z_top, double* z_bottom, int32_t* unit);
const char* li_constraints_warnings(void* h);
void li_constraints_destroy(void* h);
/* label grid of python/lithoseed/src/lithoseed/_surfaces.py:659-671 */
int li_label_grid_load(const char* path, void** out);
int li_label_grid_create(const int32_t* dims, const double* origin, const double* cell,
                         const int32_t* labels, void** out);
int li_label_grid_info(void* h, int32_t* dims, double* origin, double* cell);
int li_label_grid_get(void* h, int32_t* labels);
void li_label_grid_destroy(void* h);
int li_label_grid_write(const char* path, const int32_t* dims, const double* origin,
                        const double* cell, const int32_t* labels);
/* per-unit closed meshes (unit u: vertices v_off[u]..v_off[u+1], triangles t_off[u]..) ->
 * one shared-vertex facet set, vertices merged on the 1 um keys of rebuildGlobalDofs
 * (lithology_model.cpp:113-168), a triangle whose reversed copy is in another unit one
 * facet with (front, back) */
int li_assemble(int32_t n_units, const int64_t* v_off, const double* verts, const int64_t* t_off,
                const uint32_t* tris, void** out);
int li_assembled_counts(void* h, int64_t* n_vertices, int64_t* n_faces, int64_t* n_contacts,
                        int64_t* n_merged);
int li_assembled_get(void* h, double* vertices, uint32_t* faces, int32_t* front, int32_t* back);
void li_assembled_destroy(void* h);

#ifdef __cplusplus
}
#endif
#endif /* LITHOHOST_H */
