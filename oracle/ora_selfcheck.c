/* ora_selfcheck.c -- TEST INFRASTRUCTURE: a driver that runs the oracle's code paths on small
 * cases, built with AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers_cpu.py.
 *
 * Cases: a 2D two-block ContactHuang collision (IMPLICIT and EXPLICIT) with CartesianBox,
 * a 3D block on a CartesianFrictionalBox floor with CPDI off, a 2D Sand_SachithLocal column
 * on a CartesianSmoothBox, and a block that leaves the domain (no boundary), each stepped
 * serially and with the threaded twin. Exits 0 when every step completes and the final
 * state is finite (the left-domain case excepted); the sanitizers abort on any memory or
 * undefined-behaviour error. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_oracle.h"

static void lattice(int dim, const double *lo, const double *hi, double h, int k, double rho, int64_t *n_out,
                    double **m, double **v, double **x) {
    int cnt[3] = {1, 1, 1};
    for (int d = 0; d < dim; d++) cnt[d] = (int)((hi[d] - lo[d]) / h * k + 0.5);
    const int64_t n = (int64_t)cnt[0] * cnt[1] * cnt[2];
    *m = calloc((size_t)n, sizeof(double));
    *v = calloc((size_t)n, sizeof(double));
    *x = calloc((size_t)n * 3, sizeof(double));
    const double vol = pow(h / k, dim);
    int64_t p = 0;
    for (int c = 0; c < cnt[2]; c++)
        for (int b = 0; b < cnt[1]; b++)
            for (int a = 0; a < cnt[0]; a++, p++) {
                const int idx[3] = {a, b, c};
                for (int d = 0; d < dim; d++) (*x)[3 * p + d] = lo[d] + (idx[d] + 0.5) * h / k;
                (*m)[p] = rho * vol;
                (*v)[p] = vol;
            }
    *n_out = n;
}

static int64_t add_block(ora_sim *s, int dim, const double *lo, const double *hi, double h, double vx,
                         int material, int boundary, double bc_mu) {
    int64_t n;
    double *m, *v, *x;
    lattice(dim, lo, hi, h, 2, 1000.0, &n, &m, &v, &x);
    double *xt = calloc((size_t)n * 3, sizeof(double));
    int32_t *act = malloc((size_t)n * sizeof(int32_t));
    for (int64_t p = 0; p < n; p++) {
        xt[3 * p] = vx;
        act[p] = 1;
    }
    const int b = ora_add_body(s, n, m, v, x, xt, act, material, 1e5, 0.3, boundary, bc_mu);
    free(m);
    free(v);
    free(x);
    free(xt);
    free(act);
    return b < 0 ? -1 : n;
}

static int finite_state(const ora_sim *s, int body, int64_t n) {
    double *x = malloc((size_t)n * 3 * sizeof(double)), *xt = malloc((size_t)n * 3 * sizeof(double));
    double *T = malloc((size_t)n * 9 * sizeof(double));
    ora_get_points(s, body, x, NULL, xt, NULL, T, NULL, NULL, NULL, NULL);
    int ok = 1;
    for (int64_t i = 0; i < 3 * n; i++) ok = ok && isfinite(x[i]) && isfinite(xt[i]);
    for (int64_t i = 0; i < 9 * n; i++) ok = ok && isfinite(T[i]);
    free(x);
    free(xt);
    free(T);
    return ok;
}

static int run_case(int which, int nthreads) {
    const double L[3] = {1.0, 1.0, 1.0};
    const int N2[3] = {20, 20, 1}, N3[3] = {12, 12, 12};
    ora_sim *s = NULL;
    int64_t counts[2] = {0, 0};
    int nb = 0, expect_finite = 1;
    const double g2[3] = {0.0, -9.81, 0.0}, g3[3] = {0.0, 0.0, -9.81}, g0[3] = {0.0, 0.0, 0.0};
    if (which == 0 || which == 1) {  /* 2D collision, IMPLICIT (0) / EXPLICIT (1) */
        s = ora_create(2, L, N2, 1, which == 0 ? 1 : 0);
        const double lo1[2] = {0.2, 0.4}, hi1[2] = {0.45, 0.6}, lo2[2] = {0.5, 0.4}, hi2[2] = {0.75, 0.6};
        counts[0] = add_block(s, 2, lo1, hi1, 0.05, 1.0, 1, 1, 0.0);
        counts[1] = add_block(s, 2, lo2, hi2, 0.05, -1.0, 1, 1, 0.0);
        ora_add_contact(s, 0, 1, 0.4);
        nb = 2;
        if (ora_init(s)) return 1;
        ora_set_gravity(s, g0);
    } else if (which == 2) {  /* 3D frictional floor, CPDI off */
        s = ora_create(3, L, N3, 0, 1);
        const double lo[3] = {0.3, 0.3, 0.0}, hi[3] = {0.6, 0.6, 0.25};
        counts[0] = add_block(s, 3, lo, hi, 1.0 / 12, 0.5, 1, 3, 0.3);
        nb = 1;
        if (ora_init(s)) return 1;
        ora_set_gravity(s, g3);
    } else if (which == 3) {  /* 2D sand column on a smooth box */
        s = ora_create(2, L, N2, 1, 1);
        const double lo[2] = {0.3, 0.0}, hi[2] = {0.5, 0.4};
        counts[0] = add_block(s, 2, lo, hi, 0.05, 0.0, 2, 2, 0.0);
        const double sand[6] = {0.01, 0.577, 1650.0, 1000.0, 0.8, 0.3};
        if (ora_set_sand(s, 0, sand, 1)) return 1;
        nb = 1;
        if (ora_init(s)) return 1;
        ora_set_gravity(s, g2);
    } else {  /* a fast block leaving the domain, no boundary */
        s = ora_create(2, L, N2, 1, 1);
        const double lo[2] = {0.7, 0.4}, hi[2] = {0.95, 0.6};
        counts[0] = add_block(s, 2, lo, hi, 0.05, 60.0, 1, 0, 0.0);
        nb = 1;
        expect_finite = 0;
        if (ora_init(s)) return 1;
        ora_set_gravity(s, g2);
    }
    for (int b = 0; b < nb; b++)
        if (counts[b] <= 0) return 1;
    ora_step(s, 1e-4, 40, nthreads);
    int ok = 1;
    if (expect_finite)
        for (int b = 0; b < nb; b++) ok = ok && finite_state(s, b, counts[b]);
    ora_destroy(s);
    return ok ? 0 : 2;
}

int main(void) {
    for (int which = 0; which < 5; which++)
        for (int t = 1; t <= 3; t += 2) {
            const int rc = run_case(which, t);
            if (rc) {
                fprintf(stderr, "case %d with %d thread(s) failed (%d)\n", which, t, rc);
                return rc;
            }
        }
    printf("ora_selfcheck: all cases ran\n");
    return 0;
}
