/*
 * mpm_oracle.c -- CPU restatement of mpm_3d's explicit USL step (TEST INFRASTRUCTURE).
 *
 * Oracle only: loaded by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py. The product path never links or calls this file.
 *
 * Each function cites the reference file:line it restates. Operation order follows the
 * reference's operators, which matters for bitwise agreement:
 *   - vector / scalar is computed as vector * (1.0/scalar)          (mpm_vector.hpp:600-615)
 *   - scalar * vector and vector * scalar are per-component products (mpm_vector.hpp:584-598)
 *   - MaterialTensor * KinematicVector sums j < DIM from 0           (mpm_vector.hpp:816-824)
 *   - tensor products sum k = 0,1,2 left to right                   (mpm_tensor.hpp:1060-1070)
 *   - COO operate() zeroes a dense result then accumulates tuples in push order
 *                                                                   (mpm_sparse.cpp:30-44)
 * Build with -ffp-contract=off so no multiply-add is fused (the reference is built
 * without FMA contraction).
 *
 * Pins (tests/test_oracle_pins.py): the reference ships no golden vectors for this path
 * (its examples hold inputs only), so every function here is pinned by closed forms and
 * invariants chosen so that a dropped term, wrong sign, wrong index or transposed operand
 * fails: brute-force element search, partition of unity and linear reproduction of S and
 * gradS, hand values of the shipped blockdrop2D fixture, the CPDI-OFF gradient against
 * finite differences, mass/momentum conservation of P2G, free fall and rigid translation
 * trajectories in closed form, a linear velocity field (L, density, Hooke stress),
 * hand-computed Jaumann updates (uniaxial, volumetric, shear, spin), the CartesianBox mask,
 * ContactHuang's momentum conservation, normals and friction limits, and the smooth and
 * frictional boxes (face-normal mask against a boundary-free twin, free-slip momentum
 * conservation, mu = 0 equal to the smooth box, the Coulomb impulse per face node, and a
 * sliding block's deceleration near mu g).
 */
#define _GNU_SOURCE
#include "mpm_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* -------------------------------------------------------------------------- */
/* growable COO tuple list (include/mpm_sparse.hpp:18-170)                      */
typedef struct {
    int64_t size, cap;
    int ncomp;      /* 1 for S, 3 for gradS */
    int32_t *i;     /* node */
    int32_t *j;     /* point */
    double *val;    /* ncomp per tuple */
} coo_t;

static void coo_clear(coo_t *c) { c->size = 0; }

static void coo_reserve(coo_t *c, int64_t need) {
    if (need <= c->cap) return;
    int64_t cap = c->cap ? c->cap : 1024;
    while (cap < need) cap *= 2;
    c->i = (int32_t *)realloc(c->i, sizeof(int32_t) * cap);
    c->j = (int32_t *)realloc(c->j, sizeof(int32_t) * cap);
    c->val = (double *)realloc(c->val, sizeof(double) * cap * c->ncomp);
    c->cap = cap;
}

static void coo_push(coo_t *c, int32_t node, int32_t point, const double *v) {
    coo_reserve(c, c->size + 1);
    c->i[c->size] = node;
    c->j[c->size] = point;
    for (int d = 0; d < c->ncomp; d++) c->val[c->size * c->ncomp + d] = v[d];
    c->size++;
}

static void coo_free(coo_t *c) {
    free(c->i);
    free(c->j);
    free(c->val);
    memset(c, 0, sizeof(*c));
}

/* -------------------------------------------------------------------------- */
typedef struct {
    int64_t n;
    double *x, *u, *xt, *mxt, *b; /* 3n */
    double *L, *T;                /* 9n */
    double *m, *v, *v0, *extent;  /* n */
    int32_t *active;

    int has_material;  /* 1 IsotropicLinearElasticity, 2 Sand_SachithLocal */
    double E, nu, Gmod, K, lambda;
    /* Sand_SachithLocal {grain diameter, mu_s, grains_rho, rho_critical, mu_2, I_0} and
       USE_DENSITY_FOR_PRESSURE (sand_sachithlocal.cpp:58-117); plastic strain and its rate */
    double sand[6];
    int sand_dp;
    double *gammap, *gammap_dot;
    int has_boundary;  /* 1 CartesianBox, 2 CartesianSmoothBox, 3 CartesianFrictionalBox */
    double bc_mu;      /* CartesianFrictionalBox mu_f */
    double *bc_mask;   /* 3 per node (cartesian_box.cpp:53-59, cartesian_smooth_box.cpp:40-50,
                          cartesian_frictional_box.cpp:55-77) */
    double *bc_force;  /* 3 per node: bcNodalForce of the frictional box */

    /* DefaultNodes, dense over the grid (default_nodes.cpp:29-63) */
    double *nm;                                  /* node_count */
    double *nx, *nu_, *nxt, *ndxt, *nmxt, *nf;  /* 3 * node_count */

    coo_t S, G;
} ora_body;

typedef struct {
    int b1, b2;
    double mu;
    double *normal; /* 3 * node_count */
    double *force;  /* 3 * node_count */
} ora_contact;

struct ora_sim {
    int dim;
    double L[3];
    int N[3];
    double h[3];
    int64_t node_count, element_count;
    int npe;      /* nodes per element, 2^dim */
    int ncorner;  /* GIMP corners, 2^dim */
    int Apt[8][3];  /* point corner signs (default_points.cpp:59-84) */
    int Ael[8][3];  /* element local node offsets (cartesian.cpp:91-112) */
    int cpdi_spec, contact_spec;
    double g[3];
    double dt;
    int nbody, ncontact;
    ora_body *bodies;
    ora_contact *contacts;
    int nthreads;
};

/* -------------------------------------------------------------------------- */
/* CartesianLinear                                                              */

/* cartesian.cpp:245-253 */
static int in_domain(const ora_sim *s, const double *x) {
    for (int i = 0; i < s->dim; i++) {
        if (!(x[i] <= s->L[i] && x[i] >= 0)) return 0;
    }
    return 1;
}

/* cartesian.cpp:219-240. The element id is accumulated in double and truncated to int
 * at each step, as the reference's `int elementID += double` does. */
static int which_element(const ora_sim *s, const double *x) {
    for (int i = 0; i < s->dim; i++) {
        if (!(x[i] < s->L[i] && x[i] >= 0)) return -1;
    }
    int e = (int)floor(x[0] / s->L[0] * s->N[0]);
    if (s->dim >= 2) e = (int)((double)e + floor(x[1] / s->L[1] * s->N[1]) * (s->N[0]));
    if (s->dim >= 3) e = (int)((double)e + floor(x[2] / s->L[2] * s->N[2]) * (s->N[1]) * (s->N[0]));
    return e;
}

/* element -> i,j,k (cartesian.cpp:119-126) */
static void element_ijk(const ora_sim *s, int e, int *ijk) {
    int tmp = e;
    for (int i = 0; i < s->dim; i++) {
        ijk[i] = tmp % s->N[i];
        tmp = tmp / s->N[i];
    }
}

/* nodeIDs(e, n) (cartesian.cpp:128-140) and the node's i,j,k */
static int element_node(const ora_sim *s, const int *eijk, int n, int *nijk) {
    int id = 0;
    for (int i = 0; i < s->dim; i++) {
        nijk[i] = eijk[i] + s->Ael[n][i];
        if (i == 0) id += nijk[i];
        else if (i == 1) id += nijk[i] * (s->N[0] + 1);
        else id += nijk[i] * (s->N[0] + 1) * (s->N[1] + 1);
    }
    return id;
}

/* nodeIDToPosition (cartesian.cpp:258-274): h * ijk */
static double node_coord(const ora_sim *s, int i, int ijk_i) { return s->h[i] * ijk_i; }

/* An element id can exceed the table when (x/L)*N rounds up to N for x just below L;
 * the reference then reads past its element table. The oracle reports no nodes for such
 * an element (and the GPU kernels do the same), which is the only departure. */
static int element_valid(const ora_sim *s, int e) { return e >= 0 && e < s->element_count; }

/* evaluateBasisFnValue (cartesian.cpp:280-302): appends npe (node, value) pairs */
static int basis_value(const ora_sim *s, const double *x, int32_t *nid, double *val) {
    int e = which_element(s, x);
    if (!element_valid(s, e)) return 0;
    int eijk[3] = {0, 0, 0}, nijk[3] = {0, 0, 0};
    element_ijk(s, e, eijk);
    for (int n = 0; n < s->npe; n++) {
        int id = element_node(s, eijk, n, nijk);
        double tmp = 1.0;
        for (int i = 0; i < s->dim; i++) {
            double r = (x[i] - node_coord(s, i, nijk[i])) / s->h[i];
            tmp *= (1 - fabs(r));
        }
        nid[n] = id;
        val[n] = tmp;
    }
    return s->npe;
}

/* evaluateBasisFnGradient (cartesian.cpp:307-357); 0/0 = NaN when r = 0 is kept */
static int basis_gradient(const ora_sim *s, const double *x, int32_t *nid, double *grad3) {
    int e = which_element(s, x);
    if (!element_valid(s, e)) return 0;
    int eijk[3] = {0, 0, 0}, nijk[3] = {0, 0, 0};
    element_ijk(s, e, eijk);
    for (int n = 0; n < s->npe; n++) {
        int id = element_node(s, eijk, n, nijk);
        double r[3] = {0, 0, 0};
        double tmp = 1.0;
        for (int i = 0; i < s->dim; i++) {
            r[i] = (x[i] - node_coord(s, i, nijk[i])) / s->h[i];
            tmp *= (1 - fabs(r[i]));
        }
        for (int i = 0; i < 3; i++) grad3[3 * n + i] = 0.0;
        for (int i = 0; i < s->dim; i++) {
            double gi = -tmp;
            gi = gi / (1 - fabs(r[i]));
            gi = gi * r[i];
            gi = gi / fabs(r[i]);
            gi = gi / s->h[i];
            if (fabs(r[i]) >= 1) gi = 0.0;
            grad3[3 * n + i] = gi;
        }
        nid[n] = id;
    }
    return s->npe;
}

/* -------------------------------------------------------------------------- */
/* DefaultPoints::generateMap (default_points.cpp:196-330) for points [p0, p1)  */

static void corner_position(const ora_sim *s, const ora_body *B, int64_t p, int c, double *xc) {
    xc[0] = xc[1] = xc[2] = 0.0;
    for (int pos = 0; pos < s->dim; pos++) {
        /* x_i[pos] + 0.5*extent(i)*A(c,pos): (0.5*extent)*A, then added */
        xc[pos] = B->x[3 * p + pos] + 0.5 * B->extent[p] * (double)s->Apt[c][pos];
    }
}

static void generate_map_range(const ora_sim *s, ora_body *B, int64_t p0, int64_t p1, coo_t *S,
                               coo_t *G) {
    const int nc = s->ncorner;
    int32_t nvec[64];
    double valvec[64];
    double gradvec[64 * 3];
    int32_t nbuf[8];
    double vbuf[8], gbuf[24];
    for (int64_t i = p0; i < p1; i++) {
        const double *xi = &B->x[3 * i];
        int cpdi = s->cpdi_spec;
        if (B->active[i] == 0) {
            continue;
        } else if (!in_domain(s, xi)) {
            B->active[i] = 0;
            continue;
        } else if (cpdi == 1) {
            for (int c = 0; c < nc; c++) {
                double xc[3];
                corner_position(s, B, i, c, xc);
                if (!in_domain(s, xc)) {
                    cpdi = 0;
                    break;
                }
            }
        }
        if (cpdi == 0) {
            int k = basis_value(s, xi, nbuf, vbuf);
            for (int j = 0; j < k; j++) coo_push(S, nbuf[j], (int32_t)i, &vbuf[j]);
            k = basis_gradient(s, xi, nbuf, gbuf);
            for (int j = 0; j < k; j++) coo_push(G, nbuf[j], (int32_t)i, &gbuf[3 * j]);
        } else if (cpdi == 1) {
            int cnt = 0;
            for (int c = 0; c < nc; c++) {
                double xc[3];
                corner_position(s, B, i, c, xc);
                cnt += basis_value(s, xc, &nvec[cnt], &valvec[cnt]);
            }
            for (int j = 0; j < cnt; j++) {
                double sv = valvec[j] / (double)nc; /* valvec[j] / A.rows() */
                coo_push(S, nvec[j], (int32_t)i, &sv);
            }
            cnt = 0;
            int gcnt = 0;
            for (int c = 0; c < nc; c++) {
                double xc[3];
                corner_position(s, B, i, c, xc);
                int k = basis_value(s, xc, &nvec[cnt], valvec);
                cnt += k;
                for (int v = 0; v < k; v++) {
                    double *tg = &gradvec[3 * gcnt];
                    tg[0] = tg[1] = tg[2] = 0.0;
                    for (int pos = 0; pos < s->dim; pos++) {
                        /* valvec[v] / (A.rows() * 0.5 * extent(i)) * A(c,pos) */
                        tg[pos] = valvec[v] / ((double)nc * 0.5 * B->extent[i]) * (double)s->Apt[c][pos];
                    }
                    gcnt++;
                }
            }
            for (int j = 0; j < cnt; j++) coo_push(G, nvec[j], (int32_t)i, &gradvec[3 * j]);
        }
    }
}

/* -------------------------------------------------------------------------- */
/* threading helper: a persistent worker pool, as include/threadpool.hpp:19-146 keeps one
 * (workers started once, woken per job, the caller waiting until every job of a batch has
 * finished, as ThreadPoolExplicitUSL waits on its jobs). A batch is one range per worker:
 * thread t takes [n t / T, n (t + 1) / T). The pool lives until the process ends; a call
 * with another thread count rebuilds it. */
typedef void (*range_fn)(void *ctx, int tid, int64_t lo, int64_t hi);

#define ORA_MAX_THREADS 256
static struct {
    int nthreads;            /* workers alive (0: none) */
    pthread_t th[ORA_MAX_THREADS];
    pthread_mutex_t mu;
    pthread_cond_t wake, done;
    unsigned long gen;       /* batch generation, bumped per parallel_for */
    int pending;             /* workers still running the current batch */
    int quit;
    range_fn fn;
    void *ctx;
    int64_t n;
} g_pool = {0, {0}, PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, PTHREAD_COND_INITIALIZER, 0, 0, 0, NULL,
            NULL, 0};

static void *pool_worker(void *arg) {
    const int tid = (int)(intptr_t)arg;
    unsigned long seen = 0;
    pthread_mutex_lock(&g_pool.mu);
    for (;;) {
        while (!g_pool.quit && g_pool.gen == seen) pthread_cond_wait(&g_pool.wake, &g_pool.mu);
        if (g_pool.quit) break;
        seen = g_pool.gen;
        const range_fn fn = g_pool.fn;
        void *ctx = g_pool.ctx;
        const int64_t n = g_pool.n;
        const int T = g_pool.nthreads;
        pthread_mutex_unlock(&g_pool.mu);
        fn(ctx, tid, n * tid / T, n * (tid + 1) / T);
        pthread_mutex_lock(&g_pool.mu);
        if (--g_pool.pending == 0) pthread_cond_signal(&g_pool.done);
    }
    pthread_mutex_unlock(&g_pool.mu);
    return NULL;
}

static void pool_resize(int nthreads) {
    if (g_pool.nthreads == nthreads) return;
    if (g_pool.nthreads > 0) {
        pthread_mutex_lock(&g_pool.mu);
        g_pool.quit = 1;
        pthread_cond_broadcast(&g_pool.wake);
        pthread_mutex_unlock(&g_pool.mu);
        for (int t = 0; t < g_pool.nthreads; t++) pthread_join(g_pool.th[t], NULL);
        g_pool.quit = 0;
        g_pool.nthreads = 0;
    }
    g_pool.gen = 0;
    g_pool.nthreads = nthreads;
    for (int t = 0; t < nthreads; t++) pthread_create(&g_pool.th[t], NULL, pool_worker, (void *)(intptr_t)t);
}

/* every thread id 0..nthreads-1 runs exactly once (ranges may be empty) */
static void parallel_for(int nthreads, int64_t n, range_fn fn, void *ctx) {
    if (nthreads <= 1) {
        fn(ctx, 0, 0, n);
        return;
    }
    if (nthreads > ORA_MAX_THREADS) nthreads = ORA_MAX_THREADS;
    pool_resize(nthreads);
    pthread_mutex_lock(&g_pool.mu);
    g_pool.fn = fn;
    g_pool.ctx = ctx;
    g_pool.n = n;
    g_pool.pending = nthreads;
    g_pool.gen++;
    pthread_cond_broadcast(&g_pool.wake);
    while (g_pool.pending > 0) pthread_cond_wait(&g_pool.done, &g_pool.mu);
    pthread_mutex_unlock(&g_pool.mu);
}

/* ThreadPoolPoints::generateMap (threadpool_points.cpp:202-349): per-thread COO over
 * contiguous point ranges, concatenated in thread order (same tuple order as serial). */
typedef struct {
    const ora_sim *s;
    ora_body *B;
    coo_t *Sp, *Gp;
} map_ctx;

static void map_worker(void *vctx, int tid, int64_t lo, int64_t hi) {
    map_ctx *c = (map_ctx *)vctx;
    coo_clear(&c->Sp[tid]);
    coo_clear(&c->Gp[tid]);
    generate_map_range(c->s, c->B, lo, hi, &c->Sp[tid], &c->Gp[tid]);
}

static void coo_append(coo_t *dst, const coo_t *src) {
    coo_reserve(dst, dst->size + src->size);
    memcpy(dst->i + dst->size, src->i, sizeof(int32_t) * src->size);
    memcpy(dst->j + dst->size, src->j, sizeof(int32_t) * src->size);
    memcpy(dst->val + dst->size * dst->ncomp, src->val, sizeof(double) * src->size * src->ncomp);
    dst->size += src->size;
}

static void generate_map(ora_sim *s, ora_body *B) {
    coo_clear(&B->S);
    coo_clear(&B->G);
    if (s->nthreads <= 1) {
        generate_map_range(s, B, 0, B->n, &B->S, &B->G);
        return;
    }
    int nt = s->nthreads;
    coo_t *Sp = (coo_t *)calloc(nt, sizeof(coo_t));
    coo_t *Gp = (coo_t *)calloc(nt, sizeof(coo_t));
    for (int t = 0; t < nt; t++) {
        Sp[t].ncomp = 1;
        Gp[t].ncomp = 3;
    }
    map_ctx ctx = {s, B, Sp, Gp};
    if (B->n < 2 * nt) {
        generate_map_range(s, B, 0, B->n, &B->S, &B->G);
    } else {
        parallel_for(nt, B->n, map_worker, &ctx);
        for (int t = 0; t < nt; t++) {
            coo_append(&B->S, &Sp[t]);
            coo_append(&B->G, &Gp[t]);
        }
    }
    for (int t = 0; t < nt; t++) {
        coo_free(&Sp[t]);
        coo_free(&Gp[t]);
    }
    free(Sp);
    free(Gp);
}

/* -------------------------------------------------------------------------- */
/* COO operators (mpm_sparse.cpp)                                               */

enum { OP_S_SCALAR, OP_S_VEC, OP_G_TENSOR, OP_G_SCALAR };

/* scatter of one tuple range into a dense node-sized output */
static void scatter_range(const ora_sim *s, int op, const coo_t *C, const double *rhs, double *out,
                          int64_t k0, int64_t k1) {
    const int D = s->dim;
    for (int64_t k = k0; k < k1; k++) {
        int64_t i = C->i[k], j = C->j[k];
        if (op == OP_S_SCALAR) {
            /* tmp[i] += buffer[k] * rhs[j]   (mpm_sparse.cpp:30-44) */
            out[i] += C->val[k] * rhs[j];
        } else if (op == OP_S_VEC) {
            /* KinematicVector: buffer[k] * rhs[j], then += (mpm_sparse.cpp:66-81) */
            for (int d = 0; d < D; d++) out[3 * i + d] = out[3 * i + d] + rhs[3 * j + d] * C->val[k];
        } else if (op == OP_G_TENSOR) {
            /* left_multiply_by_tensor: tmp[i] += T_j * g_k (mpm_sparse.cpp:691-706,
             * MaterialTensor*KinematicVector mpm_vector.hpp:816-824) */
            const double *T = &rhs[9 * j];
            const double *g = &C->val[3 * k];
            for (int r = 0; r < 3; r++) {
                double t = 0;
                for (int c = 0; c < D; c++) t += T[3 * r + c] * g[c];
                out[3 * i + r] = out[3 * i + r] + t;
            }
        } else { /* OP_G_SCALAR: gradS * m (KinematicVectorSparseMatrix::operate 171-186) */
            const double *g = &C->val[3 * k];
            for (int d = 0; d < D; d++) out[3 * i + d] = out[3 * i + d] + g[d] * rhs[j];
        }
    }
}

static int op_width(int op) { return op == OP_S_SCALAR ? 1 : 3; }

typedef struct {
    const ora_sim *s;
    int op;
    const coo_t *C;
    const double *rhs;
    double *partials; /* nthreads * width * rows */
    double *out;
    int64_t rows;
    int width;
    int nt;
} scat_ctx;

static void scat_worker(void *vctx, int tid, int64_t lo, int64_t hi) {
    scat_ctx *c = (scat_ctx *)vctx;
    double *part = c->partials + (int64_t)tid * c->rows * c->width;
    memset(part, 0, sizeof(double) * c->rows * c->width);
    scatter_range(c->s, c->op, c->C, c->rhs, part, lo, hi);
}

static void reduce_worker(void *vctx, int tid, int64_t lo, int64_t hi) {
    (void)tid;
    scat_ctx *c = (scat_ctx *)vctx;
    int64_t stride = c->rows * c->width;
    for (int64_t e = lo * c->width; e < hi * c->width; e++) {
        double acc = 0;
        for (int t = 0; t < c->nt; t++) acc += c->partials[t * stride + e];
        c->out[e] = acc;
    }
}

/* operate(NORMAL): zero a dense result, then accumulate (serial: COO order; threaded:
 * per-thread partials over tuple ranges reduced in thread order, as
 * ThreadPoolExplicitUSL::parallelMultiply does, threadpool_explicit_usl.cpp:276-412) */
static void scatter(const ora_sim *s, int op, const coo_t *C, const double *rhs, double *out,
                    int64_t rows) {
    int w = op_width(op);
    int nt = s->nthreads;
    if (nt <= 1 || C->size < 4096) {
        memset(out, 0, sizeof(double) * rows * w);
        scatter_range(s, op, C, rhs, out, 0, C->size);
        return;
    }
    scat_ctx c = {s, op, C, rhs, NULL, out, rows, w, nt};
    c.partials = (double *)malloc(sizeof(double) * rows * w * nt);
    parallel_for(nt, C->size, scat_worker, &c);
    parallel_for(nt, rows, reduce_worker, &c);
    free(c.partials);
}

/* operate(TRANSPOSED) gathers: per point, tuples in push order */
enum { GA_S_VEC, GA_G_TENSOR };

static void gather_range(const ora_sim *s, int op, const coo_t *C, const double *rhs, double *out,
                         int64_t k0, int64_t k1) {
    const int D = s->dim;
    for (int64_t k = k0; k < k1; k++) {
        int64_t node = C->i[k], p = C->j[k];
        if (op == GA_S_VEC) {
            for (int d = 0; d < D; d++) out[3 * p + d] = out[3 * p + d] + rhs[3 * node + d] * C->val[k];
        } else {
            /* tensor_product_transpose: tmp[p] += v_node (x) g  (mpm_sparse.cpp:673-688,
             * KinematicVector::tensor mpm_vector.hpp:684-701, KinematicTensor += 698-711) */
            const double *v = &rhs[3 * node];
            const double *g = &C->val[3 * k];
            for (int r = 0; r < D; r++)
                for (int c = 0; c < D; c++) out[9 * p + 3 * r + c] = out[9 * p + 3 * r + c] + v[r] * g[c];
        }
    }
}

typedef struct {
    const ora_sim *s;
    int op;
    const coo_t *C;
    const double *rhs;
    double *out;
    const int64_t *cut;
} gath_ctx;

static void gath_worker(void *vctx, int tid, int64_t lo, int64_t hi) {
    (void)lo;
    (void)hi;
    gath_ctx *c = (gath_ctx *)vctx;
    gather_range(c->s, c->op, c->C, c->rhs, c->out, c->cut[tid], c->cut[tid + 1]);
}

static void gather(const ora_sim *s, int op, const coo_t *C, const double *rhs, double *out,
                   int64_t npts) {
    int w = op == GA_S_VEC ? 3 : 9;
    memset(out, 0, sizeof(double) * npts * w);
    int nt = s->nthreads;
    if (nt <= 1 || C->size < 4096) {
        gather_range(s, op, C, rhs, out, 0, C->size);
        return;
    }
    /* tuples are grouped by point, so cut the tuple range at point boundaries: each
     * thread owns whole points and the per-point order is the serial one */
    int64_t cut[257];
    if (nt > 256) nt = 256; /* parallel_for caps at the same count */
    cut[0] = 0;
    for (int t = 1; t < nt; t++) {
        int64_t k = C->size * t / nt;
        if (k < cut[t - 1]) k = cut[t - 1];
        while (k > 0 && k < C->size && C->j[k] == C->j[k - 1]) k++;
        cut[t] = k;
    }
    cut[nt] = C->size;
    gath_ctx c = {s, op, C, rhs, out, cut};
    parallel_for(nt, nt, gath_worker, &c);
}

/* -------------------------------------------------------------------------- */
/* ExplicitUSL phases (explicit_usl.cpp)                                        */

/* mapPointsToNodes (explicit_usl.cpp:112-190) */
static void map_points_to_nodes(ora_sim *s, ora_body *B) {
    const int64_t nn = s->node_count, n = B->n;
    const int D = s->dim;
    scatter(s, OP_S_SCALAR, &B->S, B->m, B->nm, nn);
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) B->mxt[3 * i + d] = B->xt[3 * i + d] * B->m[i];
    scatter(s, OP_S_VEC, &B->S, B->mxt, B->nmxt, nn);
    for (int64_t i = 0; i < nn; i++) {
        if (B->nm[i] > 0) {
            double r = 1.0 / B->nm[i];
            for (int d = 0; d < D; d++) B->nxt[3 * i + d] = B->nmxt[3 * i + d] * r;
        } else {
            for (int d = 0; d < D; d++) B->nxt[3 * i + d] = 0;
        }
    }
    double *pvec = (double *)calloc(3 * n, sizeof(double));
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) pvec[3 * i + d] = B->b[3 * i + d] * B->m[i];
    scatter(s, OP_S_VEC, &B->S, pvec, B->nf, nn);
    free(pvec);
    double *tmpMat = (double *)malloc(sizeof(double) * 9 * n);
    for (int64_t i = 0; i < n; i++)
        for (int c = 0; c < 9; c++) tmpMat[9 * i + c] = B->T[9 * i + c] * B->v[i];
    double *nvec = (double *)malloc(sizeof(double) * 3 * nn);
    scatter(s, OP_G_TENSOR, &B->G, tmpMat, nvec, nn);
    for (int64_t i = 0; i < nn; i++)
        for (int d = 0; d < D; d++) B->nf[3 * i + d] = B->nf[3 * i + d] - nvec[3 * i + d];
    free(nvec);
    free(tmpMat);
    /* CartesianFrictionalBox::generateRules/applyRules recompute exactly this force with the
     * same operations in the same order before contacts act (cartesian_frictional_box.cpp:
     * 85-117); keep it */
    if (B->has_boundary == 3) memcpy(B->bc_force, B->nf, sizeof(double) * 3 * nn);
}

static double dot3(const double *a, const double *b, int D) {
    double t = 0;
    for (int i = 0; i < D; i++) t += a[i] * b[i];
    return t;
}

/* ContactHuang::generateRules (contact_huang.cpp:80-99) */
static void contact_generate(ora_sim *s, ora_contact *C) {
    ora_body *B1 = &s->bodies[C->b1];
    const int D = s->dim;
    scatter(s, OP_G_SCALAR, &B1->G, B1->m, C->normal, s->node_count);
    for (int64_t i = 0; i < s->node_count; i++) {
        double *nrm = &C->normal[3 * i];
        double len = sqrt(dot3(nrm, nrm, D));
        double r = 1.0 / len;
        for (int d = 0; d < D; d++) nrm[d] *= r;
    }
}

/* ContactHuang::applyRules (contact_huang.cpp:104-197) */
static void contact_apply(ora_sim *s, ora_contact *C, int spec) {
    ora_body *B1 = &s->bodies[C->b1], *B2 = &s->bodies[C->b2];
    const int D = s->dim;
    const double dt = s->dt;
    for (int64_t i = 0; i < s->node_count; i++) {
        double *fc = &C->force[3 * i];
        fc[0] = fc[1] = fc[2] = 0;
        if (!(B1->nm[i] > 0 && B2->nm[i] > 0)) continue;
        const double *nrm = &C->normal[3 * i];
        double m1 = B1->nm[i], m2 = B2->nm[i];
        double mv1[3] = {0, 0, 0}, mv2[3] = {0, 0, 0}, vcm[3] = {0, 0, 0}, tmp[3] = {0, 0, 0};
        double s1[3] = {0, 0, 0};
        for (int d = 0; d < D; d++) {
            if (spec == 1) {
                mv1[d] = B1->nmxt[3 * i + d] + B1->nf[3 * i + d] * dt;
                mv2[d] = B2->nmxt[3 * i + d] + B2->nf[3 * i + d] * dt;
            } else {
                mv1[d] = B1->nmxt[3 * i + d];
                mv2[d] = B2->nmxt[3 * i + d];
            }
        }
        double rm12 = 1.0 / (m1 + m2);
        for (int d = 0; d < D; d++) vcm[d] = (mv1[d] + mv2[d]) * rm12;
        double rm1 = 1.0 / m1;
        for (int d = 0; d < D; d++) tmp[d] = mv1[d] * rm1 - vcm[d];
        if (dot3(tmp, nrm, D) > 0) {
            double fn = m1 / dt * (dot3(vcm, nrm, D) - dot3(mv1, nrm, D) / m1);
            double a = m1 / dt;
            for (int d = 0; d < D; d++) s1[d] = (vcm[d] - mv1[d] * rm1) * a - nrm[d] * fn;
            double ft = sqrt(dot3(s1, s1, D));
            if (ft != 0) {
                double rft = 1.0 / ft;
                for (int d = 0; d < D; d++) s1[d] *= rft;
            }
            double fnc = (fn < 0.0) ? fn : 0.0;             /* std::min(0.0, fn) */
            double fa = C->mu * fabs(fn), fb = fabs(ft);
            double fsc = (fb < fa) ? fb : fa;                /* std::min(mu|fn|, |ft|) */
            for (int d = 0; d < D; d++) fc[d] = nrm[d] * fnc + s1[d] * fsc;
            for (int d = 0; d < D; d++) {
                B1->nf[3 * i + d] = B1->nf[3 * i + d] + fc[d];
                B2->nf[3 * i + d] = B2->nf[3 * i + d] - fc[d];
            }
        }
    }
}

/* CartesianBox::applyRules (cartesian_box.cpp:73-85) and CartesianSmoothBox::applyRules
 * (cartesian_smooth_box.cpp:70-84): zero the masked components */
static void boundary_apply(ora_sim *s, ora_body *B) {
    const int D = s->dim;
    for (int64_t i = 0; i < s->node_count; i++) {
        for (int d = 0; d < D; d++) {
            if (B->bc_mask[3 * i + d] == 1) {
                B->nxt[3 * i + d] = 0;
                B->nmxt[3 * i + d] = 0;
                B->nf[3 * i + d] = 0;
            }
        }
    }
}

static double norm3(const double *a, int D) {
    double t = 0;
    for (int i = 0; i < D; i++) t += a[i] * a[i];
    return sqrt(t);
}

/* CartesianFrictionalBox::applyRules (cartesian_frictional_box.cpp:98-151). bcNodalForce
 * = S (m b) - gradS (V T) (105-117) is the P2G force before contact, kept in bc_force. */
static void frictional_apply(ora_sim *s, ora_body *B) {
    const int D = s->dim;
    const double dt = s->dt;
    for (int64_t i = 0; i < s->node_count; i++) {
        double dm[3] = {0, 0, 0};
        double *mxt = &B->nmxt[3 * i], *f = &B->nf[3 * i];
        const double *F = &B->bc_force[3 * i];
        for (int d = 0; d < D; d++) {
            const double mk = B->bc_mask[3 * i + d];
            if (mk == -1) {
                const double t = mxt[d] + dt * F[d];
                dm[d] = -(t < 0.0 ? t : 0.0);  /* -min(0, .) */
                B->nxt[3 * i + d] = 0;
                mxt[d] = 0;
                f[d] = 0;
            } else if (mk == 1) {
                const double t = mxt[d] + dt * F[d];
                dm[d] = -(t > 0.0 ? t : 0.0);  /* -max(0, .) */
                B->nxt[3 * i + d] = 0;
                mxt[d] = 0;
                f[d] = 0;
            } else {
                dm[d] = 0;
            }
        }
        const double f_normal = norm3(dm, D) / dt;
        for (int d = 0; d < D; d++) dm[d] = mxt[d] + dt * f[d];
        const double len = norm3(dm, D);
        if (len > 0) {
            const double a = len / dt, b = f_normal * B->bc_mu;
            const double mag = a < b ? a : b;
            const double r = 1.0 / len;
            for (int d = 0; d < D; d++) f[d] -= (mag * dm[d]) * r;
        }
    }
}

/* moveGrid (explicit_usl.cpp:233-264) */
static void move_grid(ora_sim *s, ora_body *B) {
    const int64_t nn = s->node_count;
    const int D = s->dim;
    const double dt = s->dt;
    for (int64_t e = 0; e < 3 * nn; e++) B->nmxt[e] = B->nmxt[e] + B->nf[e] * dt;
    for (int64_t i = 0; i < nn; i++) {
        if (B->nm[i] > 0) {
            double r = 1.0 / B->nm[i];
            for (int d = 0; d < D; d++) B->nxt[3 * i + d] = B->nmxt[3 * i + d] * r;
        } else {
            for (int d = 0; d < D; d++) B->nxt[3 * i + d] = 0;
        }
    }
    for (int64_t e = 0; e < 3 * nn; e++) B->nu_[e] = B->nxt[e] * dt;
    for (int64_t i = 0; i < nn; i++) {
        if (B->nm[i] > 0) {
            double r = 1.0 / B->nm[i];
            for (int d = 0; d < D; d++) B->ndxt[3 * i + d] = (B->nf[3 * i + d] * dt) * r;
        } else {
            for (int d = 0; d < D; d++) B->ndxt[3 * i + d] = 0;
        }
    }
}

/* movePoints (explicit_usl.cpp:266-300) */
static void move_points(ora_sim *s, ora_body *B) {
    const int64_t n = B->n;
    double *tmp = (double *)malloc(sizeof(double) * 3 * n);
    gather(s, GA_S_VEC, &B->S, B->nu_, tmp, n);
    for (int64_t e = 0; e < 3 * n; e++) B->x[e] = B->x[e] + tmp[e];
    gather(s, GA_S_VEC, &B->S, B->nu_, tmp, n);
    for (int64_t e = 0; e < 3 * n; e++) B->u[e] = B->u[e] + tmp[e];
    gather(s, GA_S_VEC, &B->S, B->ndxt, tmp, n);
    for (int64_t e = 0; e < 3 * n; e++) B->xt[e] = B->xt[e] + tmp[e];
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < s->dim; d++) B->mxt[3 * i + d] = B->xt[3 * i + d] * B->m[i];
    free(tmp);
}

/* calculateStrainRate (explicit_usl.cpp:302-330) */
static void strain_rate(ora_sim *s, ora_body *B) { gather(s, GA_G_TENSOR, &B->G, B->nxt, B->L, B->n); }

/* updateDensity (explicit_usl.cpp:332-345) */
static void update_density(ora_sim *s, ora_body *B) {
    for (int64_t i = 0; i < B->n; i++) {
        const double *L = &B->L[9 * i];
        double tr = L[0] + L[4] + L[8];
        B->v[i] *= exp(s->dt * tr);
    }
}

/* IsotropicLinearElasticity::calculateStress on one point (isolin.cpp:50-78) */
static void isolin_point(double G, double lambda, double dt, const double *L, double *T) {
    double D[9], W[9], tmp[9], WT[9], TW[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            D[3 * r + c] = (L[3 * r + c] + L[3 * c + r]) * 0.5;
            W[3 * r + c] = (L[3 * r + c] - L[3 * c + r]) * 0.5;
        }
    double trD = D[0] + D[4] + D[8];
    double twoG = 2 * G;
    double lt = lambda * trD;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            WT[3 * r + c] = W[3 * r + 0] * T[0 + c] + W[3 * r + 1] * T[3 + c] + W[3 * r + 2] * T[6 + c];
            TW[3 * r + c] = T[3 * r + 0] * W[0 + c] + T[3 * r + 1] * W[3 + c] + T[3 * r + 2] * W[6 + c];
        }
    for (int k = 0; k < 9; k++) {
        double Ik = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
        tmp[k] = ((D[k] * twoG + Ik * lt) + WT[k]) - TW[k];
    }
    for (int k = 0; k < 9; k++) T[k] = T[k] + tmp[k] * dt;
}

/* Sand_SachithLocal::negative_root (sand_sachithlocal.cpp:44-53): the smaller root of
 * a x^2 + b x + c without cancellation */
static double negative_root(double a, double b, double c) {
    double x;
    if (b > 0) x = (-b - sqrt(b * b - 4 * a * c)) / (2 * a);
    else x = (2 * c) / (-b + sqrt(b * b - 4 * a * c));
    return x;
}

/* Sand_SachithLocal::calculateStress on one point, SPEC = UPDATE (sand_sachithlocal.cpp:
 * 144-232): the Jaumann trial stress of isolin_point, then the local mu(I) return map.
 * m / v is the point's density after this step's volume update. Returns nup_tau. */
static double sand_point(double G, double K, double lambda, const double *sp, int use_dp, double dt, double m,
                         double v, const double *L, double *T) {
    const double c = 0;
    const double d = sp[0], MU_S = sp[1], GRAINS_RHO = sp[2], RHO_CRITICAL = sp[3], MU_2 = sp[4], I_0 = sp[5];
    double s_tr[9], t0_tr[9];
    memcpy(s_tr, T, sizeof(s_tr));
    isolin_point(G, lambda, dt, L, s_tr);  /* s_tr = T + dt (2G D + lambda trD I + W T - T W) */
    double p_tr = -(s_tr[0] + s_tr[4] + s_tr[8]) / 3.0;
    double nn = 0;
    for (int k = 0; k < 9; k++) {
        const double Ik = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
        t0_tr[k] = s_tr[k] + p_tr * Ik;
    }
    for (int k = 0; k < 9; k++) nn += t0_tr[k] * t0_tr[k];
    const double tau_tr = sqrt(nn) / sqrt(2.0);
    if (use_dp) p_tr = K * ((m / v - RHO_CRITICAL) / RHO_CRITICAL);
    const int density_flag = (m / v) < RHO_CRITICAL;
    double nup_tau;
    if (density_flag || p_tr <= c) {
        nup_tau = tau_tr / (G * dt);
        for (int k = 0; k < 9; k++) T[k] = 0;
    } else {
        double tau_tau, scale_factor;
        const double S0 = MU_S * p_tr;
        if (tau_tr <= S0) {
            tau_tau = tau_tr;
            scale_factor = 1.0;
        } else {
            const double S2 = MU_2 * p_tr;
            const double alpha = G * I_0 * dt * sqrt(p_tr / GRAINS_RHO) / d;
            const double Bq = -(S2 + tau_tr + alpha);
            const double H = S2 * tau_tr + S0 * alpha;
            tau_tau = negative_root(1.0, Bq, H);
            scale_factor = (tau_tau / tau_tr);
        }
        nup_tau = ((tau_tr - tau_tau) / G) / dt;
        for (int k = 0; k < 9; k++) {
            const double Ik = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
            T[k] = scale_factor * t0_tr[k] - p_tr * Ik;
        }
    }
    return nup_tau;
}

static void update_stress(ora_sim *s, ora_body *B) {
    if (!B->has_material) return;
    for (int64_t i = 0; i < B->n; i++) {
        if (B->active[i] == 0) continue;
        if (B->has_material == 2) {
            const double nup = sand_point(B->Gmod, B->K, B->lambda, B->sand, B->sand_dp, s->dt, B->m[i], B->v[i],
                                          &B->L[9 * i], &B->T[9 * i]);
            B->gammap[i] += nup * s->dt;
            B->gammap_dot[i] = nup;
        } else {
            isolin_point(B->Gmod, B->lambda, s->dt, &B->L[9 * i], &B->T[9 * i]);
        }
    }
}

/* DefaultDriver::applyGravity (default_driver.cpp:109-119) */
static void apply_gravity(ora_sim *s) {
    for (int b = 0; b < s->nbody; b++) {
        ora_body *B = &s->bodies[b];
        for (int64_t i = 0; i < B->n; i++)
            for (int d = 0; d < s->dim; d++) B->b[3 * i + d] = s->g[d];
    }
}

/* ExplicitUSL::step (explicit_usl.cpp:48-87) */
static void usl_step(ora_sim *s) {
    for (int b = 0; b < s->nbody; b++) generate_map(s, &s->bodies[b]);
    for (int b = 0; b < s->nbody; b++) map_points_to_nodes(s, &s->bodies[b]);
    /* generateLoads/applyLoads: no-ops for DefaultBody/DefaultPoints/DefaultNodes */
    for (int c = 0; c < s->ncontact; c++) contact_generate(s, &s->contacts[c]);
    for (int c = 0; c < s->ncontact; c++) contact_apply(s, &s->contacts[c], s->contact_spec);
    for (int b = 0; b < s->nbody; b++) {
        ora_body *B = &s->bodies[b];
        if (B->has_boundary == 3) frictional_apply(s, B);
        else if (B->has_boundary) boundary_apply(s, B);
    }
    for (int b = 0; b < s->nbody; b++) move_grid(s, &s->bodies[b]);
    for (int b = 0; b < s->nbody; b++) move_points(s, &s->bodies[b]);
    for (int b = 0; b < s->nbody; b++) strain_rate(s, &s->bodies[b]);
    for (int b = 0; b < s->nbody; b++) update_density(s, &s->bodies[b]);
    apply_gravity(s);
    for (int b = 0; b < s->nbody; b++) update_stress(s, &s->bodies[b]);
}

/* -------------------------------------------------------------------------- */
/* public API                                                                   */

ora_sim *ora_create(int dim, const double L[3], const int N[3], int cpdi_spec, int contact_spec) {
    if (dim < 2 || dim > 3) return NULL;
    ora_sim *s = (ora_sim *)calloc(1, sizeof(ora_sim));
    s->dim = dim;
    s->node_count = 1;
    s->element_count = 1;
    s->npe = 1;
    for (int i = 0; i < 3; i++) {
        s->L[i] = i < dim ? L[i] : 0;
        s->N[i] = i < dim ? N[i] : 0;
        s->h[i] = i < dim ? L[i] / N[i] : 0;
    }
    for (int i = 0; i < dim; i++) {
        s->node_count *= (N[i] + 1);
        s->element_count *= N[i];
        s->npe *= 2;
    }
    s->ncorner = s->npe;
    /* corner signs: dim 0 fastest, -1 then +1 (default_points.cpp:70-84);
     * element node offsets: dim 0 fastest, 0 then 1 (cartesian.cpp:97-112) */
    for (int c = 0; c < s->npe; c++)
        for (int i = 0; i < 3; i++) {
            int bit = (c >> i) & 1;
            s->Apt[c][i] = i < dim ? (bit ? 1 : -1) : 0;
            s->Ael[c][i] = i < dim ? bit : 0;
        }
    s->cpdi_spec = cpdi_spec;
    s->contact_spec = contact_spec;
    s->nthreads = 1;
    return s;
}

int ora_add_body(ora_sim *s, int64_t n, const double *m, const double *v, const double *x,
                 const double *xt, const int32_t *active, int has_material, double E, double nu,
                 int has_boundary, double bc_mu) {
    s->bodies = (ora_body *)realloc(s->bodies, sizeof(ora_body) * (s->nbody + 1));
    ora_body *B = &s->bodies[s->nbody];
    memset(B, 0, sizeof(*B));
    B->n = n;
    B->x = (double *)calloc(3 * n, sizeof(double));
    B->u = (double *)calloc(3 * n, sizeof(double));
    B->xt = (double *)calloc(3 * n, sizeof(double));
    B->mxt = (double *)calloc(3 * n, sizeof(double));
    B->b = (double *)calloc(3 * n, sizeof(double));
    B->L = (double *)calloc(9 * n, sizeof(double));
    B->T = (double *)calloc(9 * n, sizeof(double));
    B->m = (double *)calloc(n, sizeof(double));
    B->v = (double *)calloc(n, sizeof(double));
    B->v0 = (double *)calloc(n, sizeof(double));
    B->extent = (double *)calloc(n, sizeof(double));
    B->active = (int32_t *)calloc(n, sizeof(int32_t));
    B->gammap = (double *)calloc(n, sizeof(double));
    B->gammap_dot = (double *)calloc(n, sizeof(double));
    memcpy(B->m, m, sizeof(double) * n);
    memcpy(B->v, v, sizeof(double) * n);
    memcpy(B->x, x, sizeof(double) * 3 * n);
    memcpy(B->xt, xt, sizeof(double) * 3 * n);
    memcpy(B->active, active, sizeof(int32_t) * n);
    B->has_material = has_material;
    B->E = E;
    B->nu = nu;
    B->has_boundary = has_boundary;
    B->bc_mu = bc_mu;
    if (has_boundary < 0 || has_boundary > 3) return -1;
    B->S.ncomp = 1;
    B->G.ncomp = 3;
    return s->nbody++;
}

int ora_add_contact(ora_sim *s, int body1, int body2, double mu) {
    s->contacts = (ora_contact *)realloc(s->contacts, sizeof(ora_contact) * (s->ncontact + 1));
    ora_contact *C = &s->contacts[s->ncontact];
    memset(C, 0, sizeof(*C));
    C->b1 = body1;
    C->b2 = body2;
    C->mu = mu;
    return s->ncontact++;
}

int ora_init(ora_sim *s) {
    const int64_t nn = s->node_count;
    for (int b = 0; b < s->nbody; b++) {
        ora_body *B = &s->bodies[b];
        /* DefaultPoints::init (default_points.cpp:35-57) */
        memcpy(B->v0, B->v, sizeof(double) * B->n);
        for (int64_t i = 0; i < B->n; i++)
            B->extent[i] = s->dim == 2 ? 0.5 * sqrt(B->v[i]) : 0.5 * cbrt(B->v[i]);
        /* DefaultNodes::init (default_nodes.cpp:29-63) */
        B->nm = (double *)calloc(nn, sizeof(double));
        B->nx = (double *)calloc(3 * nn, sizeof(double));
        B->nu_ = (double *)calloc(3 * nn, sizeof(double));
        B->nxt = (double *)calloc(3 * nn, sizeof(double));
        B->ndxt = (double *)calloc(3 * nn, sizeof(double));
        B->nmxt = (double *)calloc(3 * nn, sizeof(double));
        B->nf = (double *)calloc(3 * nn, sizeof(double));
        for (int64_t i = 0; i < nn; i++) {
            int64_t tmp = i;
            for (int d = 0; d < s->dim; d++) {
                int ijk = (int)(tmp % (s->N[d] + 1));
                tmp /= (s->N[d] + 1);
                B->nx[3 * i + d] = node_coord(s, d, ijk);
            }
        }
        /* IsotropicLinearElasticity::init (isolin.cpp:26-46) */
        if (B->has_material) {
            B->Gmod = B->E / (2.0 * (1.0 + B->nu));
            B->K = B->E / (3.0 * (1.0 - 2 * B->nu));
            B->lambda = B->K - 2.0 * B->Gmod / 3.0;
        }
        /* CartesianBox::init (cartesian_box.cpp:23-60), CartesianSmoothBox::init
         * (cartesian_smooth_box.cpp:22-54), CartesianFrictionalBox::init (20-80) */
        if (B->has_boundary) {
            double Lx[3] = {0, 0, 0};
            for (int64_t i = 0; i < nn; i++)
                for (int d = 0; d < s->dim; d++)
                    if (B->nx[3 * i + d] > Lx[d]) Lx[d] = B->nx[3 * i + d];
            B->bc_mask = (double *)calloc(3 * nn, sizeof(double));
            for (int64_t i = 0; i < nn; i++) {
                if (B->has_boundary == 1) {
                    for (int d = 0; d < s->dim; d++) {
                        if (B->nx[3 * i + d] == 0 || B->nx[3 * i + d] == Lx[d]) {
                            for (int e = 0; e < s->dim; e++) B->bc_mask[3 * i + e] = 1;
                            break;
                        }
                    }
                } else if (B->has_boundary == 2) {
                    for (int d = 0; d < s->dim; d++)
                        if (B->nx[3 * i + d] == 0 || B->nx[3 * i + d] == Lx[d]) B->bc_mask[3 * i + d] = 1;
                } else {
                    int edge = 0;
                    for (int d = 0; d < s->dim; d++) {
                        if (B->nx[3 * i + d] == 0) {
                            B->bc_mask[3 * i + d] = -1;
                            edge = 1;
                        } else if (B->nx[3 * i + d] == Lx[d]) {
                            B->bc_mask[3 * i + d] = 1;
                            edge = 1;
                        }
                    }
                    for (int d = 0; d < s->dim; d++)
                        if (edge && B->bc_mask[3 * i + d] != 1 && B->bc_mask[3 * i + d] != -1) B->bc_mask[3 * i + d] = 2;
                }
            }
            if (B->has_boundary == 3) B->bc_force = (double *)calloc(3 * nn, sizeof(double));
        }
    }
    for (int c = 0; c < s->ncontact; c++) {
        ora_contact *C = &s->contacts[c];
        if (C->b1 < 0 || C->b1 >= s->nbody || C->b2 < 0 || C->b2 >= s->nbody) return -1;
        C->normal = (double *)calloc(3 * nn, sizeof(double));
        C->force = (double *)calloc(3 * nn, sizeof(double));
    }
    return 0;
}

void ora_set_gravity(ora_sim *s, const double g[3]) {
    for (int d = 0; d < 3; d++) s->g[d] = d < s->dim ? g[d] : 0;
    apply_gravity(s);
}

void ora_step(ora_sim *s, double dt, int nsteps, int nthreads) {
    s->dt = dt;
    s->nthreads = nthreads < 1 ? 1 : (nthreads > 256 ? 256 : nthreads);
    for (int k = 0; k < nsteps; k++) usl_step(s);
    s->nthreads = 1;
}

void ora_get_points(const ora_sim *s, int body, double *x, double *u, double *xt, double *mxt,
                    double *T9, double *L9, double *v, double *extent, int32_t *active) {
    const ora_body *B = &s->bodies[body];
    int64_t n = B->n;
    if (x) memcpy(x, B->x, sizeof(double) * 3 * n);
    if (u) memcpy(u, B->u, sizeof(double) * 3 * n);
    if (xt) memcpy(xt, B->xt, sizeof(double) * 3 * n);
    if (mxt) memcpy(mxt, B->mxt, sizeof(double) * 3 * n);
    if (T9) memcpy(T9, B->T, sizeof(double) * 9 * n);
    if (L9) memcpy(L9, B->L, sizeof(double) * 9 * n);
    if (v) memcpy(v, B->v, sizeof(double) * n);
    if (extent) memcpy(extent, B->extent, sizeof(double) * n);
    if (active) memcpy(active, B->active, sizeof(int32_t) * n);
}

void ora_set_points(ora_sim *s, int body, const double *x, const double *xt, const double *T9,
                    const double *v, const int32_t *active) {
    ora_body *B = &s->bodies[body];
    int64_t n = B->n;
    if (x) memcpy(B->x, x, sizeof(double) * 3 * n);
    if (xt) memcpy(B->xt, xt, sizeof(double) * 3 * n);
    if (T9) memcpy(B->T, T9, sizeof(double) * 9 * n);
    if (v) memcpy(B->v, v, sizeof(double) * n);
    if (active) memcpy(B->active, active, sizeof(int32_t) * n);
}

void ora_get_nodes(const ora_sim *s, int body, double *m, double *mxt, double *f, double *xt,
                   double *u, double *dxt) {
    const ora_body *B = &s->bodies[body];
    int64_t nn = s->node_count;
    if (m && B->nm) memcpy(m, B->nm, sizeof(double) * nn);
    if (mxt && B->nmxt) memcpy(mxt, B->nmxt, sizeof(double) * 3 * nn);
    if (f && B->nf) memcpy(f, B->nf, sizeof(double) * 3 * nn);
    if (xt && B->nxt) memcpy(xt, B->nxt, sizeof(double) * 3 * nn);
    if (u && B->nu_) memcpy(u, B->nu_, sizeof(double) * 3 * nn);
    if (dxt && B->ndxt) memcpy(dxt, B->ndxt, sizeof(double) * 3 * nn);
}

void ora_get_contact(const ora_sim *s, int contact, double *normal, double *force) {
    const ora_contact *C = &s->contacts[contact];
    if (normal) memcpy(normal, C->normal, sizeof(double) * 3 * s->node_count);
    if (force) memcpy(force, C->force, sizeof(double) * 3 * s->node_count);
}

int64_t ora_node_count(const ora_sim *s) { return s->node_count; }
int64_t ora_element_count(const ora_sim *s) { return s->element_count; }

int64_t ora_get_map(ora_sim *s, int body, int which, int32_t *node, int32_t *point, double *val,
                    int64_t cap) {
    ora_body *B = &s->bodies[body];
    generate_map(s, B);
    const coo_t *C = which == 0 ? &B->S : &B->G;
    int64_t k = C->size < cap ? C->size : cap;
    if (node) memcpy(node, C->i, sizeof(int32_t) * k);
    if (point) memcpy(point, C->j, sizeof(int32_t) * k);
    if (val) memcpy(val, C->val, sizeof(double) * k * C->ncomp);
    return C->size;
}

void ora_connectivity(const ora_sim *s, int body, int32_t *corner_elem, uint8_t *code) {
    const ora_body *B = &s->bodies[body];
    const int nc = s->ncorner;
    for (int64_t i = 0; i < B->n; i++) {
        const double *xi = &B->x[3 * i];
        for (int c = 0; c < nc; c++) corner_elem[i * nc + c] = -1;
        if (B->active[i] == 0 || !in_domain(s, xi)) {
            code[i] = ORA_CONN_INACTIVE;
            continue;
        }
        int cpdi = s->cpdi_spec;
        if (cpdi == 1) {
            for (int c = 0; c < nc; c++) {
                double xc[3];
                corner_position(s, B, i, c, xc);
                if (!in_domain(s, xc)) {
                    cpdi = 0;
                    break;
                }
            }
        }
        if (cpdi == 1) {
            code[i] = ORA_CONN_CPDI_ON;
            for (int c = 0; c < nc; c++) {
                double xc[3];
                corner_position(s, B, i, c, xc);
                corner_elem[i * nc + c] = which_element(s, xc);
            }
        } else {
            code[i] = ORA_CONN_CPDI_OFF;
            corner_elem[i * nc] = which_element(s, xi);
        }
    }
}

int ora_which_element(const ora_sim *s, const double *x) { return which_element(s, x); }
int ora_in_domain(const ora_sim *s, const double *x) { return in_domain(s, x); }
int ora_basis_value(const ora_sim *s, const double *x, int32_t *nid, double *val) {
    return basis_value(s, x, nid, val);
}
int ora_basis_gradient(const ora_sim *s, const double *x, int32_t *nid, double *grad3) {
    return basis_gradient(s, x, nid, grad3);
}

int ora_set_sand(ora_sim *s, int body, const double *props6, int use_density_pressure) {
    if (body < 0 || body >= s->nbody) return -1;
    ora_body *B = &s->bodies[body];
    if (!B->has_material) return -1;
    B->has_material = 2;
    memcpy(B->sand, props6, sizeof(B->sand));
    B->sand_dp = use_density_pressure;
    return 0;
}

void ora_get_plastic(const ora_sim *s, int body, double *gammap, double *gammap_dot) {
    const ora_body *B = &s->bodies[body];
    for (int64_t i = 0; i < B->n; i++) {
        if (gammap) gammap[i] = B->gammap[i];
        if (gammap_dot) gammap_dot[i] = B->gammap_dot[i];
    }
}

double ora_sand_update(double E, double nu, const double *props6, int use_density_pressure, double dt, double m,
                       double v, const double *L9, double *T9) {
    double G = E / (2.0 * (1.0 + nu));
    double K = E / (3.0 * (1.0 - 2 * nu));
    double lambda = K - 2.0 * G / 3.0;
    return sand_point(G, K, lambda, props6, use_density_pressure, dt, m, v, L9, T9);
}

void ora_isolin_update(double E, double nu, double dt, const double *L9, double *T9) {
    double G = E / (2.0 * (1.0 + nu));
    double K = E / (3.0 * (1.0 - 2 * nu));
    double lambda = K - 2.0 * G / 3.0;
    isolin_point(G, lambda, dt, L9, T9);
}

void ora_destroy(ora_sim *s) {
    if (!s) return;
    for (int b = 0; b < s->nbody; b++) {
        ora_body *B = &s->bodies[b];
        free(B->x); free(B->u); free(B->xt); free(B->mxt); free(B->b);
        free(B->L); free(B->T); free(B->m); free(B->v); free(B->v0); free(B->extent);
        free(B->active); free(B->bc_mask); free(B->bc_force); free(B->gammap); free(B->gammap_dot);
        free(B->nm); free(B->nx); free(B->nu_); free(B->nxt); free(B->ndxt); free(B->nmxt); free(B->nf);
        coo_free(&B->S);
        coo_free(&B->G);
    }
    for (int c = 0; c < s->ncontact; c++) {
        free(s->contacts[c].normal);
        free(s->contacts[c].force);
    }
    free(s->bodies);
    free(s->contacts);
    free(s);
}
