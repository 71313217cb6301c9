/*
 * ORACLE — test infrastructure, not product code.
 *
 * A from-scratch CPU restatement of PyCPET's two hot paths, used only as the
 * reference that results are compared with, by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
 * path (pycpet_amd/) never loads this library.
 *
 * Two families:
 *   *_f64  FP64 restatement of the reference semantics on the FP32 inputs the
 *          reference sees (the "FP64 oracle" all GPU results are measured
 *          against).
 *   *_f32  FP32 port with the reference's arithmetic order (one pass per point,
 *          sequential FP32 accumulation), used as the restated CPU baseline when
 *          the reference build under oracle/_ref is not present.
 *
 * Reference semantics followed (paths relative to the reference repo):
 *   field   CPET/utils/math_module.c:405-451  compute_looped_field
 *           E(p) = k * sum_j q_j (p - x_j) * max(r^2, 1e-6f)^(-3/2)
 *   esp     CPET/utils/math_module.c:453-486  calc_esp_base
 *           phi(p) = k * sum_j q_j / |p - x_j|            (no softening)
 *   point   CPET/utils/math_module.c:255-293  calc_field  (no softening)
 *   stream  CPET/utils/math_module.c:296-333  calc_field_base (no softening)
 *           CPET/utils/math_module.c:489-503  propagate_topo: x += h E/|E|
 *           CPET/utils/math_module.c:523-591  thread_operation:
 *             up to n_iter Euler steps, stop after the first step whose point
 *             is strictly outside +-dims (that point is the final point),
 *             then x0+, x0++, xf+, xf++ and
 *             curvature  CPET/utils/math_module.c:108-121 |v' x v''| / |v'|^3
 *             distance   CPET/utils/math_module.c:99-105  |x0 - xf|
 *   k = 14.3996451 (CPET/utils/math_module.c:305,412,462)
 *
 * Parity pins for each function are listed in the test files that exercise
 * them (tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

#define ORACLE_K 14.3996451

/* ------------------------------------------------------------------ FP64 */

/* E and/or phi at M arbitrary points.  soft != 0 floors r^2 at 1e-6f for E
 * only (math_module.c:407,433); phi is never softened (math_module.c:475). */
void oracle_field_f64(int M, int N, const float *pts, const float *x,
                      const float *q, int soft, double *E, double *phi)
{
    const double floor2 = (double)1e-6f;
#pragma omp parallel for schedule(dynamic, 64)
    for (int m = 0; m < M; ++m) {
        const double px = pts[3 * m], py = pts[3 * m + 1], pz = pts[3 * m + 2];
        double ex = 0.0, ey = 0.0, ez = 0.0, ph = 0.0;
        for (int j = 0; j < N; ++j) {
            const double dx = px - (double)x[3 * j];
            const double dy = py - (double)x[3 * j + 1];
            const double dz = pz - (double)x[3 * j + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double qj = (double)q[j];
            if (E) {
                const double r2e = (soft && r2 < floor2) ? floor2 : r2;
                const double s = qj / (r2e * sqrt(r2e));
                ex += s * dx;
                ey += s * dy;
                ez += s * dz;
            }
            if (phi)
                ph += qj / sqrt(r2);
        }
        if (E) {
            E[3 * m] = ORACLE_K * ex;
            E[3 * m + 1] = ORACLE_K * ey;
            E[3 * m + 2] = ORACLE_K * ez;
        }
        if (phi)
            phi[m] = ORACLE_K * ph;
    }
}

/* Unsoftened field at one FP64 point (math_module.c:296-333). */
static void field_at_f64(const double p[3], int N, const float *x,
                         const float *q, double E[3])
{
    double ex = 0.0, ey = 0.0, ez = 0.0;
    for (int j = 0; j < N; ++j) {
        const double dx = p[0] - (double)x[3 * j];
        const double dy = p[1] - (double)x[3 * j + 1];
        const double dz = p[2] - (double)x[3 * j + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double s = (double)q[j] / (r2 * sqrt(r2));
        ex += s * dx;
        ey += s * dy;
        ez += s * dz;
    }
    E[0] = ORACLE_K * ex;
    E[1] = ORACLE_K * ey;
    E[2] = ORACLE_K * ez;
}

/* Field of the N sources at p: charges (src = q) or dipoles (src = mu). */
typedef void (*field_fn)(const double p[3], int N, const float *x, const float *src, int opt,
                         double E[3]);

static void charge_field(const double p[3], int N, const float *x, const float *q, int opt,
                         double E[3])
{
    (void)opt;
    field_at_f64(p, N, x, q, E);
}

/* Field of N point dipoles (calc_field_base_dipole, math_module.c:335-371):
 *   E = k sum_j r^-5 (c R - mu_j r^2),  R = p - x_j,
 * with c = 3 mu_x R_x + mu_y R_y + mu_z R_z as the reference writes it
 * (opt = 0), or the physical c = 3 (mu . R) (opt = 1). */
static void dipole_field(const double p[3], int N, const float *x, const float *mu, int opt,
                         double E[3])
{
    double e0 = 0.0, e1 = 0.0, e2 = 0.0;
    for (int j = 0; j < N; ++j) {
        const double R0 = p[0] - x[3 * j], R1 = p[1] - x[3 * j + 1], R2 = p[2] - x[3 * j + 2];
        const double m0 = mu[3 * j], m1 = mu[3 * j + 1], m2 = mu[3 * j + 2];
        const double r2 = R0 * R0 + R1 * R1 + R2 * R2;
        const double r = sqrt(r2);
        const double r5 = 1.0 / (r2 * r2 * r);
        const double c = opt ? 3.0 * (m0 * R0 + m1 * R1 + m2 * R2)
                             : 3.0 * m0 * R0 + m1 * R1 + m2 * R2;
        e0 += r5 * (c * R0 - m0 * r2);
        e1 += r5 * (c * R1 - m1 * r2);
        e2 += r5 * (c * R2 - m2 * r2);
    }
    E[0] = ORACLE_K * e0;
    E[1] = ORACLE_K * e1;
    E[2] = ORACLE_K * e2;
}

/* One Euler step along the unit field direction (math_module.c:489-503,
 * dipoles :506-520). */
static void step_f64(field_fn fld, int opt, const double p[3], double out[3], int N,
                     const float *x, const float *src, double h)
{
    double E[3];
    fld(p, N, x, src, opt, E);
    const double n = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
    out[0] = p[0] + h * E[0] / n;
    out[1] = p[1] + h * E[1] / n;
    out[2] = p[2] + h * E[2] / n;
}

/* One classical RK4 step (weights 1, 2, 2, 1 over 6) of dx/ds = E/|E|, the
 * reference's unit-direction equation with its Euler step replaced (an
 * extension: the reference integrates with Euler only).  The weighted sum
 * is accumulated u1, +2 u2, +2 u3, +u4, then scaled by h/6. */
static void unit_dir(field_fn fld, int opt, const double p[3], double u[3], int N,
                     const float *x, const float *src)
{
    double E[3];
    fld(p, N, x, src, opt, E);
    const double n = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
    u[0] = E[0] / n;
    u[1] = E[1] / n;
    u[2] = E[2] / n;
}

static void step_rk4_f64(field_fn fld, int opt, const double p[3], double out[3], int N,
                         const float *x, const float *src, double h)
{
    double u[3], q[3], acc[3];
    unit_dir(fld, opt, p, u, N, x, src);
    for (int i = 0; i < 3; ++i) {
        acc[i] = u[i];
        q[i] = p[i] + 0.5 * h * u[i];
    }
    unit_dir(fld, opt, q, u, N, x, src);
    for (int i = 0; i < 3; ++i) {
        acc[i] += 2.0 * u[i];
        q[i] = p[i] + 0.5 * h * u[i];
    }
    unit_dir(fld, opt, q, u, N, x, src);
    for (int i = 0; i < 3; ++i) {
        acc[i] += 2.0 * u[i];
        q[i] = p[i] + h * u[i];
    }
    unit_dir(fld, opt, q, u, N, x, src);
    for (int i = 0; i < 3; ++i) {
        acc[i] += u[i];
        out[i] = p[i] + (h / 6.0) * acc[i];
    }
}

typedef void (*step_fn)(field_fn fld, int opt, const double p[3], double out[3], int N,
                        const float *x, const float *src, double h);

/* |v' x v''| / |v'|^3 (math_module.c:71-121). */
static double curve_f64(const double a[3], const double b[3])
{
    const double c0 = a[1] * b[2] - a[2] * b[1];
    const double c1 = a[2] * b[0] - a[0] * b[2];
    const double c2 = a[0] * b[1] - a[1] * b[0];
    const double nc = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    const double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    return nc / (na * na * na);
}

/* Streamlines (math_module.c:523-591), one per start point.
 *   feats[2l]   = |x0 - xf|, feats[2l+1] = (kappa_init + kappa_final) / 2
 *   steps[l]    = number of Euler steps taken (<= n_iter[l])
 *   exited[l]   = 1 if the walk ended on a box exit, 0 on the step limit
 *   pts6 (opt.) = the six points x0, x0+, x0++, xf, xf+, xf++ (L x 6 x 3)  */
static void walk_f64(field_fn fld, int opt, step_fn step_f64, int N, int L,
                     const float *starts, const int32_t *n_iter, float h32, const float *dims,
                     const float *x, const float *q, double *feats, int32_t *steps,
                     int32_t *exited, double *pts6)
{
    const double h = (double)h32;
    const double d0 = dims[0], d1 = dims[1], d2 = dims[2];
#pragma omp parallel for schedule(dynamic, 1)
    for (int l = 0; l < L; ++l) {
        double x0[3] = {starts[3 * l], starts[3 * l + 1], starts[3 * l + 2]};
        double p[3] = {x0[0], x0[1], x0[2]};
        int s = 0, out = 0;
        for (int i = 0; i < n_iter[l]; ++i) {
            double nx[3];
            step_f64(fld, opt, p, nx, N, x, q, h);
            p[0] = nx[0];
            p[1] = nx[1];
            p[2] = nx[2];
            s = i + 1;
            if (p[0] < -d0 || p[0] > d0 || p[1] < -d1 || p[1] > d1 ||
                p[2] < -d2 || p[2] > d2) {
                out = 1;
                break;
            }
        }
        double a1[3], a2[3], b1[3], b2[3];
        step_f64(fld, opt, x0, a1, N, x, q, h);
        step_f64(fld, opt, a1, a2, N, x, q, h);
        step_f64(fld, opt, p, b1, N, x, q, h);
        step_f64(fld, opt, b1, b2, N, x, q, h);
        double v1[3], v2[3], w1[3], w2[3];
        for (int i = 0; i < 3; ++i) {
            v1[i] = a1[i] - x0[i];
            v2[i] = a2[i] - 2.0 * a1[i] + x0[i];
            w1[i] = b1[i] - p[i];
            w2[i] = b2[i] - 2.0 * b1[i] + p[i];
        }
        const double ki = curve_f64(v1, v2);
        const double kf = curve_f64(w1, w2);
        const double dd0 = x0[0] - p[0], dd1 = x0[1] - p[1], dd2 = x0[2] - p[2];
        feats[2 * l] = sqrt(dd0 * dd0 + dd1 * dd1 + dd2 * dd2);
        feats[2 * l + 1] = 0.5 * (ki + kf);
        if (steps)
            steps[l] = s;
        if (exited)
            exited[l] = out;
        if (pts6) {
            const double *src[6] = {x0, a1, a2, p, b1, b2};
            for (int k = 0; k < 6; ++k)
                for (int i = 0; i < 3; ++i)
                    pts6[(size_t)l * 18 + 3 * k + i] = src[k][i];
        }
    }
}

void oracle_topo_f64(int N, int L, const float *starts, const int32_t *n_iter,
                     float h32, const float *dims, const float *x,
                     const float *q, double *feats, int32_t *steps,
                     int32_t *exited, double *pts6)
{
    walk_f64(charge_field, 0, step_f64, N, L, starts, n_iter, h32, dims, x, q, feats, steps,
             exited, pts6);
}

/* The same walk and features with RK4 steps (step_rk4_f64), for the main
 * walk and the four curvature-defining steps alike. */
void oracle_topo_rk4_f64(int N, int L, const float *starts, const int32_t *n_iter,
                         float h32, const float *dims, const float *x,
                         const float *q, double *feats, int32_t *steps,
                         int32_t *exited, double *pts6)
{
    walk_f64(charge_field, 0, step_rk4_f64, N, L, starts, n_iter, h32, dims, x, q, feats, steps,
             exited, pts6);
}

/* Dipole streamlines (thread_operation_dipole, math_module.c:593-659): the
 * same walk and features with the dipole field; physical selects the field
 * form (see dipole_field). */
void oracle_topo_dipole_f64(int N, int L, const float *starts, const int32_t *n_iter,
                            float h32, const float *dims, const float *x, const float *mu,
                            int physical, double *feats, int32_t *steps, int32_t *exited)
{
    walk_f64(dipole_field, physical, step_f64, N, L, starts, n_iter, h32, dims, x, mu, feats,
             steps, exited, NULL);
}

/* ------------------------------------------------------------------ FP32 */

/* FP32 port of compute_looped_field / calc_esp_base arithmetic: one pass
 * over the charges per point, sequential FP32 accumulation, k applied per
 * term as the reference does (math_module.c:440). */
void oracle_field_f32(int M, int N, const float *pts, const float *x,
                      const float *q, int soft, float *E, float *phi)
{
    const float k = (float)ORACLE_K;
#pragma omp parallel for schedule(dynamic, 64)
    for (int m = 0; m < M; ++m) {
        const float px = pts[3 * m], py = pts[3 * m + 1], pz = pts[3 * m + 2];
        float ex = 0.f, ey = 0.f, ez = 0.f, ph = 0.f;
        for (int j = 0; j < N; ++j) {
            const float dx = px - x[3 * j];
            const float dy = py - x[3 * j + 1];
            const float dz = pz - x[3 * j + 2];
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (E) {
                const float r2e = (soft && r2 < 1e-6f) ? 1e-6f : r2;
                const float ri = 1.0f / sqrtf(r2e);
                const float sc = k * q[j] * (ri * ri * ri);
                ex += sc * dx;
                ey += sc * dy;
                ez += sc * dz;
            }
            if (phi)
                ph += k / sqrtf(r2) * q[j];
        }
        if (E) {
            E[3 * m] = ex;
            E[3 * m + 1] = ey;
            E[3 * m + 2] = ez;
        }
        if (phi)
            phi[m] = ph;
    }
}

static void step_f32(const float p[3], float out[3], int N, const float *x,
                     const float *q, float h)
{
    const float k = (float)ORACLE_K;
    float ex = 0.f, ey = 0.f, ez = 0.f;
    for (int j = 0; j < N; ++j) {
        const float dx = p[0] - x[3 * j];
        const float dy = p[1] - x[3 * j + 1];
        const float dz = p[2] - x[3 * j + 2];
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        const float s = k / (r * r * r) * q[j];
        ex += s * dx;
        ey += s * dy;
        ez += s * dz;
    }
    const float n = sqrtf(ex * ex + ey * ey + ez * ez);
    out[0] = p[0] + h * ex / n;
    out[1] = p[1] + h * ey / n;
    out[2] = p[2] + h * ez / n;
}

static float curve_f32(const float a[3], const float b[3])
{
    const float c0 = a[1] * b[2] - a[2] * b[1];
    const float c1 = a[2] * b[0] - a[0] * b[2];
    const float c2 = a[0] * b[1] - a[1] * b[0];
    const float nc = sqrtf(c0 * c0 + c1 * c1 + c2 * c2);
    const float na = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    return nc / (na * na * na);
}

/* FP32 port of thread_operation over L streamlines (OpenMP over streamlines,
 * the restated counterpart of the reference's Pool over streamlines). */
void oracle_topo_f32(int N, int L, const float *starts, const int32_t *n_iter,
                     float h, const float *dims, const float *x,
                     const float *q, float *feats, int32_t *steps)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int l = 0; l < L; ++l) {
        float x0[3] = {starts[3 * l], starts[3 * l + 1], starts[3 * l + 2]};
        float p[3] = {x0[0], x0[1], x0[2]};
        int s = 0;
        for (int i = 0; i < n_iter[l]; ++i) {
            float nx[3];
            step_f32(p, nx, N, x, q, h);
            p[0] = nx[0];
            p[1] = nx[1];
            p[2] = nx[2];
            s = i + 1;
            if (p[0] < -dims[0] || p[0] > dims[0] || p[1] < -dims[1] ||
                p[1] > dims[1] || p[2] < -dims[2] || p[2] > dims[2])
                break;
        }
        float a1[3], a2[3], b1[3], b2[3];
        step_f32(x0, a1, N, x, q, h);
        step_f32(a1, a2, N, x, q, h);
        step_f32(p, b1, N, x, q, h);
        step_f32(b1, b2, N, x, q, h);
        float v1[3], v2[3], w1[3], w2[3];
        for (int i = 0; i < 3; ++i) {
            v1[i] = a1[i] - x0[i];
            v2[i] = a2[i] - 2.f * a1[i] + x0[i];
            w1[i] = b1[i] - p[i];
            w2[i] = b2[i] - 2.f * b1[i] + p[i];
        }
        const float ki = curve_f32(v1, v2), kf = curve_f32(w1, w2);
        const float dd0 = x0[0] - p[0], dd1 = x0[1] - p[1], dd2 = x0[2] - p[2];
        feats[2 * l] = sqrtf(dd0 * dd0 + dd1 * dd1 + dd2 * dd2);
        feats[2 * l + 1] = (ki + kf) / 2.f;
        if (steps)
            steps[l] = s;
    }
}
