/*
 * Test infrastructure (oracle): the combined rows of an advice opening (RFC-0011 steps 2 to 5),
 * restated from crates/tessera-pcs/src/opening.rs. Not part of the product.
 */
#include <stdlib.h>
#include "oracle.h"

/* RowLayout::row_coefficients (opening.rs:116-158) with add_scaled_eq (opening.rs:161-200):
 * for group g, c_g(row_i) = Sum_{claims in g} gamma eq(row_point, row_i) over the N rows of the
 * layout in layout order; then xi_row = eq(xi_point, row_i). eq(r, x) = Prod_t (x_t ? r_t : 1 - r_t)
 * with x_t bit t of the row index. Claims are listed group after group (group_sizes[g] each),
 * each a row point of row_vars Fp2 and a gamma. out: (groups + 1) x N Fp2. */
void orc_row_coefficients(uint32_t row_vars, const uint64_t *rows, uint64_t N,
                          const orc_fp2 *row_points, const orc_fp2 *gammas,
                          const uint32_t *group_sizes, uint32_t groups, const orc_fp2 *xi_point,
                          orc_fp2 *out) {
    const orc_fp2 one = {1, 0};
    uint64_t claim = 0;
    for (uint32_t g = 0; g <= groups; g++) {
        const uint32_t nc = g < groups ? group_sizes[g] : 1;
        for (uint64_t i = 0; i < N; i++) out[(uint64_t)g * N + i] = (orc_fp2){0, 0};
        for (uint32_t c = 0; c < nc; c++) {
            const orc_fp2 *pt = g < groups ? row_points + (claim + c) * row_vars : xi_point;
            const orc_fp2 scale = g < groups ? gammas[claim + c] : one;
            for (uint64_t i = 0; i < N; i++) {
                orc_fp2 e = scale;
                for (uint32_t t = 0; t < row_vars; t++)
                    e = fp2_mul(e, (rows[i] >> t) & 1 ? pt[t] : fp2_sub(one, pt[t]));
                out[(uint64_t)g * N + i] = fp2_add(out[(uint64_t)g * N + i], e);
            }
        }
        if (g < groups) claim += nc;
    }
}

/* combine (opening.rs:241-330): out[g][i] = Sum_r coefficients[g][r] matrix[r][i] over the N rows
 * of a row-major N x width matrix of Fp; out is count x width Fp2. */
void orc_combine(const orc_fp2 *coefficients, uint32_t count, const uint64_t *matrix, uint64_t N,
                 uint64_t width, orc_fp2 *out) {
    for (uint32_t g = 0; g < count; g++)
        for (uint64_t i = 0; i < width; i++) {
            orc_fp2 acc = {0, 0};
            for (uint64_t r = 0; r < N; r++)
                acc = fp2_add(acc, fp2_mul_base(coefficients[(uint64_t)g * N + r], matrix[r * width + i]));
            out[(uint64_t)g * width + i] = acc;
        }
}
