/*
 * lithofwd.h -- C ABI of the MI355X (gfx950) polyhedral gravity / magnetic forward model
 * and its adjoint.
 *
 * The library replaces the per-station x per-facet loops of LithoInv3D's litho-forward
 * module.  A context holds one lithology model resident in HBM:
 *   - vertices  (shared by every unit; the reference keeps them identical across group
 *     meshes, modules/litho-model/src/litho/lithology_model.cpp:230-247),
 *   - triangles stored once each, oriented outward for their `front` unit, with the unit
 *     on the other side in `back` (-1: outside the model).  A reference contact face that
 *     two group meshes carry with opposite winding is one triangle here,
 *   - per-unit density contrast (g/cc) and magnetization vector (reference units:
 *     amplitude chi*F/100 times the unit direction, magnetic_forward.cpp:100-134),
 *   - the stations (gravity and magnetic data share one station set per context).
 *
 * Conventions: z positive up; gz = downward attraction in mGal; TMI in nT; arrays are
 * host pointers unless the entry point name ends in _device.  Every entry point returns
 * 0 on success and a negative LF_E* code on failure; lf_last_error() then describes it.
 * A context is not thread-safe (neither is the reference's ForwardModel) and all its
 * work runs on one HIP stream.
 */
#ifndef LITHOFWD_H
#define LITHOFWD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LITHOFWD_ABI_VERSION 1

/* precision */
#define LF_FP64 0  /* plain FP64 path, the reference's arithmetic */
#define LF_FP32C 1 /* FP32 in station-centred face frames, compensated facet sums,
                      FP64 for edges whose line passes next to the station */
/* TMI edge coefficient */
#define LF_TMI_REFERENCE 0 /* (M.t)(F.t), magnetic_forward.cpp:338-343 */
#define LF_TMI_TEXTBOOK 1  /* (M.n)(F.t), the uniformly-magnetized polyhedron field */
/* outputs wanted */
#define LF_WANT_GZ 1
#define LF_WANT_TMI 2

/* errors */
#define LF_OK 0
#define LF_EINVAL -1
#define LF_EHIP -2
#define LF_ENOMEM -3
#define LF_ESTATE -4

typedef struct lf_ctx lf_ctx;

typedef struct {
    int64_t n_vertices;
    int64_t n_faces;
    int64_t n_stations;
    int32_t n_units;
    int32_t precision; /* LF_FP64 | LF_FP32C */
    int32_t tmi_mode;  /* LF_TMI_REFERENCE | LF_TMI_TEXTBOOK */
    int32_t want;      /* LF_WANT_GZ | LF_WANT_TMI */
    int32_t device;    /* HIP device ordinal */
    int32_t reserved;
    const double* vertices;  /* n_vertices*3, xyz interleaved (metres) */
    const uint32_t* faces;   /* n_faces*3 vertex ids, outward for the front unit */
    const int32_t* front;    /* n_faces unit ids (>= 0) */
    const int32_t* back;     /* n_faces unit ids, -1 = outside */
    const double* stations;  /* n_stations*3, xyz interleaved (metres) */
    const double* density;   /* n_units, g/cc */
    const double* magnetization; /* n_units*3: amp * direction (amp = chi*F/100) */
    double field_dir[3];     /* inducing-field unit vector (magnetic_forward.cpp:61-66) */
} lf_desc;

/* Creates a context and uploads the model.  Replaces the construction of GravityForward
 * (modules/litho-forward/include/litho_invert/forward/gravity_forward.h:10-11) and
 * MagneticForward (magnetic_forward.h:10-12): the forward keeps the model and stations. */
int lf_create(lf_ctx** ctx, const lf_desc* desc);
void lf_destroy(lf_ctx* ctx);
/* Last error message of ctx (or of the last failed lf_create when ctx is NULL). */
const char* lf_last_error(const lf_ctx* ctx);

/* New vertex positions (n_vertices*3, host).  Replaces the vertex write-back of
 * LithologyModel::applyParameterVector (lithology_model.cpp:210-248) that
 * GravityForward::compute performs first (gravity_forward.cpp:106). */
int lf_set_vertices(lf_ctx* ctx, const double* xyz);
/* Same, from device memory already on this context's device. */
int lf_set_vertices_device(lf_ctx* ctx, const double* d_xyz);
/* New unit properties (either pointer may be NULL to keep the current values). */
int lf_set_properties(lf_ctx* ctx, const double* density, const double* magnetization);

/* gz (mGal) and TMI (nT) at every station; either output may be NULL.
 * Replaces GravityForward::compute (gravity_forward.cpp:105-118) and
 * MagneticForward::compute (magnetic_forward.cpp:81-94) after the parameter write. */
int lf_forward(lf_ctx* ctx, double* gz, double* tmi);
/* Same into device buffers of n_stations doubles; enqueued on the context stream, returns
 * without waiting. */
int lf_forward_device(lf_ctx* ctx, double* d_gz, double* d_tmi);

/* Matrix-free adjoint: grad[v*3+axis] = sum_i (w_gz[i] dgz_i/dx_{v,axis}
 *                                          + w_tmi[i] dTMI_i/dx_{v,axis})
 * for every vertex coordinate.  With w = -W^2 r this is the data-misfit gradient that
 * ObjectiveFunction::dataMisfitGradient forms densely as -J^T W^2 r
 * (modules/litho-inversion/src/inversion/objective.cpp:42-55,
 * joint_objective.cpp:47-58) from computeGravityAnalyticJacobian /
 * computeMagneticAnalyticJacobian (gravity_forward.cpp:184-292,
 * magnetic_forward.cpp:196-306), before the vertex -> DOF column mapping.
 * Either weight vector may be NULL (treated as zero). */
int lf_adjoint(lf_ctx* ctx, const double* w_gz, const double* w_tmi, double* grad);
/* Device-pointer variant (n_stations weights each, 3*n_vertices gradient doubles). */
int lf_adjoint_device(lf_ctx* ctx, const double* d_w_gz, const double* d_w_tmi,
                      double* d_grad);

/* Squared Jacobian column norms for parameters that are linear combinations of one
 * vertex coordinate axis: parameter j moves coordinate `axis` of vertex
 * dof_vertex[e] by dof_weight[e] per unit, for e in [dof_off[j], dof_off[j+1]).
 *   out_gz[j]  = sum_i (d gz_i  / d p_j)^2,   out_tmi[j] = sum_i (d TMI_i / d p_j)^2
 * i.e. the columns of the reference's Jacobian (gravity_forward.cpp:184-292,
 * magnetic_forward.cpp:196-306) including control-point interpolation weights, without
 * forming it.  Used as the exact targets of the learned L-BFGS preconditioner (the
 * Gauss-Newton diagonal).  Either output may be NULL. */
int lf_colnorm2(lf_ctx* ctx, int axis, int64_t n_dof, const int64_t* dof_off,
                const int32_t* dof_vertex, const double* dof_weight, double* out_gz,
                double* out_tmi);

/* Jacobian-vector product along a vertex tangent: gz[i] = sum_v dgz_i/dx_v . vtan_v and
 * tmi[i] likewise (vtan: 3*n_vertices doubles, xyz per vertex), at the context's current
 * geometry and properties, in one pass over the station x facet pairs: the matrix-free
 * counterpart of J v with the reference's Jacobian (computeGravityAnalyticJacobian /
 * computeMagneticAnalyticJacobian, gravity_forward.cpp:184-292,
 * magnetic_forward.cpp:196-306, before the vertex -> DOF mapping), which its magnetic
 * weight calibration multiplies densely (joint_objective.cpp:281-393).  Per pair the
 * forward of lf_adjoint's pair arithmetic with tangents (forward mode), equal to the
 * contraction of that pair's adjoint seeds with the facet constants' tangents; FP64 sums
 * over facets per station.  Either output may be NULL. */
int lf_jvp(lf_ctx* ctx, const double* vtan, double* gz, double* tmi);
/* Device-pointer variant (3*n_vertices tangent doubles, n_stations outputs each). */
int lf_jvp_device(lf_ctx* ctx, const double* d_vtan, double* d_gz, double* d_tmi);

/* Waits for the context stream. */
int lf_sync(lf_ctx* ctx);

/* Per-kernel HIP-event timing on the context stream.  When enabled, each launch is
 * bracketed by events; lf_timing() returns, for kernel class k (0 = face setup,
 * 1 = forward, 2 = forward reduce, 3 = adjoint pairs, 4 = adjoint finalize,
 * 5 = vertex gather, 6 = Jacobian-vector product), the summed milliseconds and launch
 * count since the last reset. */
int lf_set_timing(lf_ctx* ctx, int enable);
int lf_timing(lf_ctx* ctx, double* ms, int64_t* launches, int n_classes);

/* Diagnostics of the last forward: edges whose station-edge geometry took the FP64
 * branch in FP32C (0 in FP64 mode). */
int lf_stats(lf_ctx* ctx, int64_t* fp64_edges);

/* ---- forward with no device allocation ---------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * forward of lf_create + lf_forward_device for a model given entirely in device memory,
 * with no context.  desc's array pointers are DEVICE pointers on desc->device, which must
 * be the current HIP device; d_gz / d_tmi are device buffers of n_stations doubles (either
 * may be NULL); stream is a hipStream_t (NULL: the null stream).  The work is enqueued and
 * the call returns without waiting.
 * The facet records, station arrays and per-chunk partial sums live in a static device
 * buffer of LF_NOALLOC_SCRATCH_BYTES per GPU, part of the library's code object (held from
 * the library's first use on a GPU until the process ends).  A model needs
 *   ~ n_faces * (200 B FP32C | 288 B FP64) + n_stations * 24 B
 *     + 16 B * n_stations * n_chunks   (n_chunks <= n_faces / 512 and <= 8 MiB / (16 B *
 *                                        n_stations): at most 8 MiB of partial sums)
 * and a larger one is refused with LF_ENOMEM (lf_last_error(NULL) gives the bytes).
 * Every call on one GPU uses the same buffer: calls that overlap on a GPU race, so
 * serialize them (one stream, or wait for the previous call).  Facet vertex and unit ids
 * cannot be checked on the host here; a facet with an id out of range contributes
 * nothing.  Errors: lf_last_error(NULL). */
#define LF_NOALLOC_SCRATCH_BYTES (64ull << 20)
int lf_forward_noalloc(const lf_desc* desc, double* d_gz, double* d_tmi, void* stream);

/* ---- learned diagonal initial inverse Hessian for L-BFGS-B --------------------------
 * No counterpart in the reference: it replaces H0 = gamma*I of
 * modules/litho-inversion/src/inversion/lbfgsb_optimizer.cpp:40-62 by gamma*D with D
 * predicted per parameter.  A 16 -> 64 -> 64 -> 1 GELU MLP; W1 and W2 are used in BF16
 * on MFMA with FP32 accumulation; the FP32 master parameters are laid out as
 * W1[16][64], b1[64], W2[64][64], b2[64], w3[64], b3 (lf_mlp_param_count() floats).
 * Inputs x are n*16 floats (row-major), targets and outputs n floats; host pointers. */
typedef struct lf_mlp lf_mlp;
int64_t lf_mlp_param_count(void);
int lf_mlp_create(lf_mlp** mlp, int device, uint64_t seed);
void lf_mlp_destroy(lf_mlp* mlp);
const char* lf_mlp_last_error(const lf_mlp* mlp);
int lf_mlp_set_params(lf_mlp* mlp, const float* params);
int lf_mlp_get_params(lf_mlp* mlp, float* params);
int lf_mlp_forward(lf_mlp* mlp, int64_t n, const float* x, float* y);
/* loss = sum_s (y_s - t_s)^2 / n and its gradient with respect to every parameter */
int lf_mlp_loss_grad(lf_mlp* mlp, int64_t n, const float* x, const float* t, float* loss,
                     float* grad);
/* `steps` full-batch Adam steps (beta 0.9/0.999) on the MSE; losses[k] = loss before step k */
int lf_mlp_train(lf_mlp* mlp, int64_t n, const float* x, const float* t, int steps, float lr,
                 float* losses);

#ifdef __cplusplus
}
#endif
#endif
