/*
 * lithofwd_oracle.h -- FP64 CPU restatement of LithoInv3D's polyhedral gravity (gz) and
 * total-field magnetic (TMI) forward model and its per-face finite-difference vertex
 * Jacobian.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker that the HIP path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library (liblithofwd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * reference repository root).  Conventions follow the reference: z is positive up,
 * gz is the downward attraction in mGal, TMI in nT, density in g/cc.
 *
 * Pinning (which independent check fixes each function) is listed per function in
 * lithofwd_oracle.c.
 */
#ifndef LITHOFWD_ORACLE_H
#define LITHOFWD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* modules/litho-core/src/core/geometry.cpp:10-30 */
double lfo_solid_angle(const double a[3], const double b[3], const double c[3],
                       const double o[3]);
/* modules/litho-core/src/core/geometry.cpp:32-52 */
double lfo_line_integral(const double a[3], const double b[3], const double o[3]);
/* modules/litho-forward/src/forward/gravity_forward.cpp:13-42 */
double lfo_face_nagy(const double o[3], const double a[3], const double b[3],
                     const double c[3]);
/* modules/litho-forward/src/forward/magnetic_forward.cpp:312-347 */
double lfo_magnetic_facet(const double o[3], const double a[3], const double b[3],
                          const double c[3], const double mdir[3], const double fdir[3]);
/* Textbook polyhedron TMI facet factor (edge coefficient (M.n)(F.t) instead of the
 * reference's (M.t)(F.t)); used only by tests to separate the two formulas. */
double lfo_magnetic_facet_textbook(const double o[3], const double a[3],
                                   const double b[3], const double c[3],
                                   const double mdir[3], const double fdir[3]);
/* modules/litho-forward/src/forward/magnetic_forward.cpp:57-67 */
void lfo_field_unit_vector(double inc_deg, double dec_deg, double out[3]);
/* modules/litho-forward/src/forward/magnetic_forward.cpp:100-134.
 * mode 0 = EffectiveSusceptibility, 1 = FixedVectorPerGroup, 2 = VectorPerGroup. */
void lfo_group_magnetization(int mode, double susceptibility, double field_nT,
                             const double fhat[3], double rem_magnitude,
                             double rem_inc_deg, double rem_dec_deg,
                             const double rem_vector[3], double* amp, double mdir[3]);

/* A lithology model the way the reference holds it: one closed triangle mesh per group
 * (modules/litho-model/src/litho/lithology_model.cpp, one SurfaceMesh per LithoGroup).
 * Vertex positions are shared: every group's triangles index one global vertex array,
 * which is what LithologyModel::applyParameterVector's shared-vertex sync guarantees
 * (lithology_model.cpp:230-247). */
typedef struct {
    int32_t n_groups;
    int64_t n_vertices;
    const double* verts;            /* n_vertices * 3, xyz interleaved */
    const uint32_t* tri;            /* all groups' triangles, 3 vertex ids each */
    const int64_t* group_tri_begin; /* n_groups + 1 offsets into tri (in triangles) */
    const double* density;          /* n_groups, g/cc (signed contrasts allowed) */
    const double* mag_amp;          /* n_groups, reference "A/m" amplitude (chi*F/100) */
    const double* mag_dir;          /* n_groups * 3, unit magnetization directions */
} lfo_model;

/* GravityForward::compute loop: gravity_forward.cpp:105-118, 124-130, 162-178 */
void lfo_gravity_forward(const lfo_model* m, int64_t n_st, const double* st, double* gz,
                         int threads);
/* MagneticForward::compute loop: magnetic_forward.cpp:81-94, 136-146, 367-386.
 * textbook != 0 swaps in lfo_magnetic_facet_textbook. */
void lfo_magnetic_forward(const lfo_model* m, const double fhat[3], int64_t n_st,
                          const double* st, double* tmi, int textbook, int threads);

/* Per-face one-sided FD (h = 1e-4) vertex gradients, accumulated per global vertex:
 * gravity_forward.cpp:48-85 + 233-267, magnetic_forward.cpp:14-50 + 244-281.
 * Output J[i][v][axis] (n_st * n_vertices * 3), mGal/m or nT/m. */
void lfo_gravity_vertex_jacobian(const lfo_model* m, int64_t n_st, const double* st,
                                 double* J, int threads);
void lfo_magnetic_vertex_jacobian(const lfo_model* m, const double fhat[3], int64_t n_st,
                                  const double* st, double* J, int threads);
/* The same Jacobians contracted with per-station weights without storing them:
 * grad[v][axis] = sum_i w[i] * J[i][v][axis]  (objective.cpp:42-55 forms -J^T W^2 r). */
void lfo_gravity_vertex_jtw(const lfo_model* m, int64_t n_st, const double* st,
                            const double* w, double* grad, int threads);
void lfo_magnetic_vertex_jtw(const lfo_model* m, const double fhat[3], int64_t n_st,
                             const double* st, const double* w, double* grad,
                             int threads);

#ifdef __cplusplus
}
#endif
#endif
