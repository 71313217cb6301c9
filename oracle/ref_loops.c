/* CPU-baseline driver — test infrastructure, not product code.
 *
 * compute_ESP_on_grid (CPET/utils/calculator.py:450-475) calls the
 * reference's calc_esp_base once per grid point from a Python loop
 * (:468-471).  ref_esp_points runs the same per-point calls from C, so the
 * baseline can be timed with and without the Python/FFI overhead.  The
 * Makefile links this file with the reference's own math_module.c (compiled
 * from its source) into _ref/libref_math.so; nothing here restates the
 * reference arithmetic. */

void calc_esp_base(float ESP[1], float x_init[3], int n_charges, float x[n_charges][3],
                   float Q[n_charges]);

void ref_esp_points(int M, const float *pts, int N, const float *x, const float *Q,
                    float *out)
{
    for (int m = 0; m < M; ++m) {
        float e[1] = {0.0f};
        float p[3] = {pts[3 * m], pts[3 * m + 1], pts[3 * m + 2]};
        calc_esp_base(e, p, N, (float (*)[3])x, (float *)Q);
        out[m] = e[0];
    }
}
