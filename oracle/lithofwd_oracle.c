/*
 * lithofwd_oracle.c -- FP64 CPU restatement of the reference's forward path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See lithofwd_oracle.h.
 *
 * The arithmetic follows the reference operation by operation (Eigen's norm() is
 * sqrt(squaredNorm()), dot() is the plain sum of products); summation order inside a
 * dot product may differ from Eigen's by an ulp.
 *
 * How each function is pinned (the tests named here live in tests/):
 *   lfo_solid_angle      sum over a closed outward mesh is 4*pi inside / 0 outside; exact
 *                        on-axis value for a right triangle corner; antisymmetric under
 *                        reversed winding                         (test_oracle_kat.py)
 *   lfo_line_integral    closed form asinh(s2/rho) - asinh(s1/rho) (test_oracle_kat.py)
 *   lfo_face_nagy and    rectangular-prism closed form (Plouff/Nagy antiderivative),
 *   lfo_gravity_forward  point-mass far field, volume-matched icosphere, linearity in
 *                        density, coplanar re-triangulation invariance (test_oracle_kat.py)
 *   lfo_magnetic_facet_textbook
 *                        Poisson's relation: TMI of a uniformly magnetized body equals
 *                        the directional derivative of the gravity field of the same
 *                        body (pinned gravity oracle, rotated), for arbitrary polyhedra
 *   lfo_magnetic_facet   (reference edge coefficient (M.t)(F.t)) = textbook result plus
 *                        the asinh closed forms of the box edges whose coefficient differs,
 *                        for a box in a vertical field where every coefficient is 0, 1 or
 *                        known by hand                             (test_oracle_kat.py)
 *   lfo_*_vertex_jacobian / _jtw
 *                        central finite differences (0.01 m) of the pinned forwards on
 *                        every vertex coordinate                   (test_oracle_kat.py)
 *   lfo_group_magnetization
 *                        closed-form direction cosines; in all three modes: amplitudes
 *                        that add, subtract or meet at right angles, exact cancellation
 *                        (the |M| < 1e-15 fallback to F-hat), each mode ignoring the
 *                        other's inputs; linearity of the TMI in M  (test_oracle_kat.py)
 */
#include "lithofwd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* geometry.cpp:4-6 and gravity_forward.cpp:9, magnetic_forward.cpp:9 */
static const double kEps = 1e-15;
static const double kPi = 3.14159265358979323846;
/* gravity_forward.h:37-41 */
static const double G_SI = 6.67430e-11;
static const double DENSITY_SCALE = 1000.0;
static const double M_S2_TO_MGAL = 1e5;

static inline void sub3(const double* a, const double* b, double* o)
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static inline double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double norm3(const double* a) { return sqrt(dot3(a, a)); }

/* geometry.cpp:10-30 (van Oosterom & Strackee) */
double lfo_solid_angle(const double a[3], const double b[3], const double c[3],
                       const double o[3])
{
    double ra[3], rb[3], rc[3], bxc[3];
    sub3(a, o, ra);
    sub3(b, o, rb);
    sub3(c, o, rc);
    double la = norm3(ra), lb = norm3(rb), lc = norm3(rc);
    if (la < kEps || lb < kEps || lc < kEps) return 0.0; /* geometry.cpp:21 */
    cross3(rb, rc, bxc);
    double num = dot3(ra, bxc); /* tripleProduct, geometry.h:8 */
    double den = la * lb * lc + dot3(ra, rb) * lc + dot3(ra, rc) * lb + dot3(rb, rc) * la;
    return 2.0 * atan2(num, den);
}

/* geometry.cpp:32-52 */
double lfo_line_integral(const double a[3], const double b[3], const double o[3])
{
    double ra[3], rb[3], e[3];
    sub3(a, o, ra);
    sub3(b, o, rb);
    sub3(b, a, e);
    double L = norm3(e);
    if (L < kEps) return 0.0;
    double s1 = dot3(ra, e) / L;
    double s2 = dot3(rb, e) / L;
    double d1 = norm3(ra);
    double d2 = norm3(rb);
    double num = s2 + d2;
    double den = s1 + d1;
    if (fabs(den) < kEps && fabs(num) < kEps) return 0.0;
    if (fabs(den) < kEps) return 0.0;
    if (num <= 0.0 || den <= 0.0) return 0.0;
    return log(num / den);
}

/* gravity_forward.cpp:13-42 (faceNagyContribution) */
double lfo_face_nagy(const double o[3], const double a[3], const double b[3],
                     const double c[3])
{
    double e1[3], e2[3], n[3];
    sub3(b, a, e1);
    sub3(c, a, e2);
    cross3(e1, e2, n);
    double area_x2 = norm3(n);
    if (area_x2 < kEps) return 0.0;
    double un[3] = {n[0] / area_x2, n[1] / area_x2, n[2] / area_x2};
    double nz = un[2];
    double omega = lfo_solid_angle(a, b, c, o);
    double amo[3];
    sub3(a, o, amo);
    double h = dot3(amo, un);
    const double* v[3] = {a, b, c};
    double d_sum = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double* p1 = v[i];
        const double* p2 = v[(i + 1) % 3];
        double e[3];
        sub3(p2, p1, e);
        double L = norm3(e);
        if (L < kEps) continue;
        double eh[3] = {e[0] / L, e[1] / L, e[2] / L};
        double perp[3], omp1[3];
        cross3(un, eh, perp);
        sub3(o, p1, omp1);
        double d_perp = dot3(omp1, perp);
        double Lambda = lfo_line_integral(p1, p2, o);
        d_sum += d_perp * Lambda;
    }
    return nz * (d_sum - h * omega);
}

static double magnetic_facet_impl(const double o[3], const double a[3], const double b[3],
                                  const double c[3], const double mdir[3],
                                  const double fdir[3], int textbook)
{
    double e1[3], e2[3], n[3];
    sub3(b, a, e1);
    sub3(c, a, e2);
    cross3(e1, e2, n);
    double area_x2 = norm3(n);
    if (area_x2 < kEps) return 0.0;
    double nh[3] = {n[0] / area_x2, n[1] / area_x2, n[2] / area_x2};
    double omega = lfo_solid_angle(a, b, c, o);
    double mn = dot3(mdir, nh);
    double fn = dot3(fdir, nh);
    double contribution = mn * fn * omega;
    const double* v[3] = {a, b, c};
    for (int i = 0; i < 3; ++i) {
        const double* p1 = v[i];
        const double* p2 = v[(i + 1) % 3];
        double e[3];
        sub3(p2, p1, e);
        double L = norm3(e);
        if (L < kEps) continue;
        double eh[3] = {e[0] / L, e[1] / L, e[2] / L};
        double t[3];
        cross3(nh, eh, t);
        double mt = textbook ? mn : dot3(mdir, t);
        double ft = dot3(fdir, t);
        double Lambda = lfo_line_integral(p1, p2, o);
        contribution += mt * ft * Lambda;
    }
    return contribution;
}

/* magnetic_forward.cpp:312-347 */
double lfo_magnetic_facet(const double o[3], const double a[3], const double b[3],
                          const double c[3], const double mdir[3], const double fdir[3])
{
    return magnetic_facet_impl(o, a, b, c, mdir, fdir, 0);
}

double lfo_magnetic_facet_textbook(const double o[3], const double a[3],
                                   const double b[3], const double c[3],
                                   const double mdir[3], const double fdir[3])
{
    return magnetic_facet_impl(o, a, b, c, mdir, fdir, 1);
}

/* magnetic_forward.cpp:61-66 */
void lfo_field_unit_vector(double inc_deg, double dec_deg, double out[3])
{
    double I = inc_deg * kPi / 180.0;
    double D = dec_deg * kPi / 180.0;
    double cosI = cos(I);
    out[0] = cosI * cos(D);
    out[1] = cosI * sin(D);
    out[2] = -sin(I);
}

/* magnetic_forward.cpp:100-134 */
void lfo_group_magnetization(int mode, double susceptibility, double field_nT,
                             const double fhat[3], double rem_magnitude,
                             double rem_inc_deg, double rem_dec_deg,
                             const double rem_vector[3], double* amp, double mdir[3])
{
    double induced = susceptibility * field_nT / 100.0;
    double M[3] = {induced * fhat[0], induced * fhat[1], induced * fhat[2]};
    if (mode == 1) {
        double R[3];
        lfo_field_unit_vector(rem_inc_deg, rem_dec_deg, R);
        for (int k = 0; k < 3; ++k) M[k] += rem_magnitude * R[k];
    } else if (mode == 2) {
        for (int k = 0; k < 3; ++k) M[k] += rem_vector[k];
    }
    double a = norm3(M);
    *amp = a;
    if (a < kEps) {
        mdir[0] = fhat[0];
        mdir[1] = fhat[1];
        mdir[2] = fhat[2];
    } else {
        mdir[0] = M[0] / a;
        mdir[1] = M[1] / a;
        mdir[2] = M[2] / a;
    }
}

static inline const double* vtx(const lfo_model* m, uint32_t i) { return m->verts + 3 * (int64_t)i; }

static void set_threads(int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

/* gravity_forward.cpp:105-118 (compute), 124-130 (computeUnitGravity),
 * 162-178 (gravityClosedMesh) */
void lfo_gravity_forward(const lfo_model* m, int64_t n_st, const double* st, double* gz,
                         int threads)
{
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n_st; ++i) {
        const double* o = st + 3 * i;
        double g = 0.0;
        for (int grp = 0; grp < m->n_groups; ++grp) {
            int64_t t0 = m->group_tri_begin[grp], t1 = m->group_tri_begin[grp + 1];
            if (t1 == t0) continue;
            double total = 0.0;
            for (int64_t t = t0; t < t1; ++t) {
                const uint32_t* tr = m->tri + 3 * t;
                total += lfo_face_nagy(o, vtx(m, tr[0]), vtx(m, tr[1]), vtx(m, tr[2]));
            }
            double scale = G_SI * m->density[grp] * DENSITY_SCALE * M_S2_TO_MGAL;
            g += total * scale;
        }
        gz[i] = g;
    }
}

/* magnetic_forward.cpp:81-94 (compute), 136-146 (computeUnitMagnetic),
 * 367-386 (magneticClosedMeshVector) */
void lfo_magnetic_forward(const lfo_model* m, const double fhat[3], int64_t n_st,
                          const double* st, double* tmi, int textbook, int threads)
{
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n_st; ++i) {
        const double* o = st + 3 * i;
        double dt = 0.0;
        for (int grp = 0; grp < m->n_groups; ++grp) {
            int64_t t0 = m->group_tri_begin[grp], t1 = m->group_tri_begin[grp + 1];
            if (t1 == t0) continue;
            const double* md = m->mag_dir + 3 * grp;
            double geom = 0.0;
            for (int64_t t = t0; t < t1; ++t) {
                const uint32_t* tr = m->tri + 3 * t;
                geom += magnetic_facet_impl(o, vtx(m, tr[0]), vtx(m, tr[1]), vtx(m, tr[2]),
                                            md, fhat, textbook);
            }
            double scale = m->mag_amp[grp] * 100.0 / (4.0 * kPi);
            dt += -scale * geom;
        }
        tmi[i] = dt;
    }
}

/* One-sided FD of one facet factor on all nine vertex coordinates, h = 1e-4:
 * gravity_forward.cpp:48-85 and magnetic_forward.cpp:14-50.  kind 0 = gravity,
 * 1 = magnetic. */
static void facet_fd_gradient(int kind, const double* o, const double* a, const double* b,
                              const double* c, const double* md, const double* fh,
                              double g[9])
{
    const double h = 1e-4;
    double v[3][3];
    memcpy(v[0], a, sizeof v[0]);
    memcpy(v[1], b, sizeof v[1]);
    memcpy(v[2], c, sizeof v[2]);
    double value = kind == 0 ? lfo_face_nagy(o, v[0], v[1], v[2])
                             : lfo_magnetic_facet(o, v[0], v[1], v[2], md, fh);
    for (int k = 0; k < 3; ++k) {
        for (int ax = 0; ax < 3; ++ax) {
            double save = v[k][ax];
            v[k][ax] = save + h;
            double f = kind == 0 ? lfo_face_nagy(o, v[0], v[1], v[2])
                                 : lfo_magnetic_facet(o, v[0], v[1], v[2], md, fh);
            v[k][ax] = save;
            g[3 * k + ax] = (f - value) / h;
        }
    }
}

/* Per-station vertex gradient row (gravity_forward.cpp:233-267,
 * magnetic_forward.cpp:244-281) added into row[n_vertices*3] with weight w. */
static void vertex_gradient_row(int kind, const lfo_model* m, const double* fh,
                                const double* o, double w, double* row)
{
    const double kInv4Pi = 1.0 / (4.0 * kPi);
    for (int grp = 0; grp < m->n_groups; ++grp) {
        int64_t t0 = m->group_tri_begin[grp], t1 = m->group_tri_begin[grp + 1];
        double faceScale;
        if (kind == 0) {
            faceScale = G_SI * DENSITY_SCALE * M_S2_TO_MGAL * m->density[grp];
        } else {
            if (m->mag_amp[grp] < kEps) continue; /* magnetic_forward.cpp:259 */
            faceScale = -m->mag_amp[grp] * 100.0 * kInv4Pi;
        }
        const double* md = kind == 1 ? m->mag_dir + 3 * grp : NULL;
        for (int64_t t = t0; t < t1; ++t) {
            const uint32_t* tr = m->tri + 3 * t;
            double g[9];
            facet_fd_gradient(kind, o, vtx(m, tr[0]), vtx(m, tr[1]), vtx(m, tr[2]), md, fh, g);
            for (int k = 0; k < 3; ++k)
                for (int ax = 0; ax < 3; ++ax)
                    row[3 * (int64_t)tr[k] + ax] += w * g[3 * k + ax] * faceScale;
        }
    }
}

static void vertex_jacobian(int kind, const lfo_model* m, const double* fh, int64_t n_st,
                            const double* st, double* J, int threads)
{
    set_threads(threads);
    const int64_t nv3 = 3 * m->n_vertices;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < n_st; ++i) {
        double* row = J + i * nv3;
        memset(row, 0, sizeof(double) * (size_t)nv3);
        vertex_gradient_row(kind, m, fh, st + 3 * i, 1.0, row);
    }
}

static void vertex_jtw(int kind, const lfo_model* m, const double* fh, int64_t n_st,
                       const double* st, const double* w, double* grad, int threads)
{
    set_threads(threads);
    const int64_t nv3 = 3 * m->n_vertices;
    memset(grad, 0, sizeof(double) * (size_t)nv3);
#pragma omp parallel
    {
        double* acc = (double*)calloc((size_t)nv3, sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < n_st; ++i) {
            if (w[i] != 0.0) vertex_gradient_row(kind, m, fh, st + 3 * i, w[i], acc);
        }
#pragma omp critical
        for (int64_t k = 0; k < nv3; ++k) grad[k] += acc[k];
        free(acc);
    }
}

void lfo_gravity_vertex_jacobian(const lfo_model* m, int64_t n_st, const double* st,
                                 double* J, int threads)
{
    vertex_jacobian(0, m, NULL, n_st, st, J, threads);
}

void lfo_magnetic_vertex_jacobian(const lfo_model* m, const double fhat[3], int64_t n_st,
                                  const double* st, double* J, int threads)
{
    vertex_jacobian(1, m, fhat, n_st, st, J, threads);
}

void lfo_gravity_vertex_jtw(const lfo_model* m, int64_t n_st, const double* st,
                            const double* w, double* grad, int threads)
{
    vertex_jtw(0, m, NULL, n_st, st, w, grad, threads);
}

void lfo_magnetic_vertex_jtw(const lfo_model* m, const double fhat[3], int64_t n_st,
                             const double* st, const double* w, double* grad,
                             int threads)
{
    vertex_jtw(1, m, fhat, n_st, st, w, grad, threads);
}
