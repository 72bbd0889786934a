/*
 * ORACLE (test infrastructure only; never linked into the product library).
 *
 * Restatement of the reference's deterministic f32 distance:
 *   native/fast_cosine.c:4-13, compiled with `-ffast-math -O3` (build.rs:1-10).
 * This translation unit is compiled with the same flags so the compiler picks
 * the same kind of reassociated reduction as the reference build does with gcc.
 * Returns -dot when dot > 0, else FLT_MIN (1.17549435e-38f).
 */
float orc_fast_cosine_c(const float *x, const float *y, int dim)
{
    float acc = 0.0f;
    for (int i = 0; i < dim; i++)
        acc += x[i] * y[i];
    return acc > 0.0f ? -acc : 1.17549435e-38f;
}
