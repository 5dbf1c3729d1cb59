/*
 * mpm_oracle.h -- CPU restatement of mpm_3d's explicit USL step (TEST INFRASTRUCTURE).
 *
 * This library is the correctness oracle for the HIP solver. Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the product
 * path (the HIP library and the host framework) never links or calls it.
 *
 * It restates, in plain C and FP64, the reference's serial ExplicitUSL step for the
 * plugin set CartesianLinear / DefaultPoints / DefaultNodes / IsotropicLinearElasticity /
 * CartesianBox / ContactHuang / DefaultDriver gravity, keeping the reference's COO
 * push order, its dense-temporary scatter order and its floating-point operation order.
 * A second entry (nthreads > 1) restates the threaded ThreadPoolExplicitUSL +
 * ThreadPoolPoints path for CPU timing; its scatter summation order differs from the
 * serial path exactly as the reference's does.
 *
 * Layout: vectors are AoS with 3 doubles per point/node (the third is 0 in 2D),
 * tensors are row-major 3x3 (9 doubles), as in the reference's arrays.
 *
 * Pins (tests/test_oracle_pins.py): the reference ships no golden vectors for this path
 * and cannot be compiled here (it needs Eigen 3), so each function is pinned by an
 * independent statement -- ora_which_element / ora_in_domain by brute force and edge
 * cases; ora_basis_value / ora_basis_gradient by hand products, partition of unity,
 * linear reproduction, central differences and the 0/0 NaN; ora_get_map by hand values
 * of the shipped blockdrop2D fixture and the separable identity; ora_isolin_update by
 * uniaxial, volumetric, shear and spin closed forms; ora_step by free fall, rigid
 * translation, a linear velocity field, conservation laws, the CartesianBox mask and
 * ContactHuang's momentum, normal and friction limits; the threaded twin against the
 * serial path. Parity unpinned: bitwise multi-step agreement with the reference binary's
 * own summation order (that needs the reference itself).
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_sim ora_sim;

/* connectivity codes returned by ora_connectivity */
#define ORA_CONN_CPDI_OFF 0 /* plain linear weights at the point (cartesian.cpp:280-357) */
#define ORA_CONN_CPDI_ON 1  /* GIMP/CPDI corners (default_points.cpp:264-324) */
#define ORA_CONN_INACTIVE 2 /* inactive, or deactivated for leaving the domain */

ora_sim *ora_create(int dim, const double L[3], const int N[3], int cpdi_spec, int contact_spec);
/* returns the body index; arrays are copied. x, xt are n*3 AoS. E/nu used if has_material. */
/* has_boundary: 0 none, 1 CartesianBox, 2 CartesianSmoothBox, 3 CartesianFrictionalBox
 * (with mu_f = bc_mu) */
int ora_add_body(ora_sim *s, int64_t n, const double *m, const double *v, const double *x,
                 const double *xt, const int32_t *active, int has_material, double E, double nu,
                 int has_boundary, double bc_mu);
int ora_add_contact(ora_sim *s, int body1, int body2, double mu);
/* Job::init order: points (extent, corner table), nodes, material, boundary, contacts. */
int ora_init(ora_sim *s);
/* DefaultDriver::generateGravity/applyGravity with an explicit gravity vector. */
void ora_set_gravity(ora_sim *s, const double g[3]);
/* nsteps calls of ExplicitUSL::step. nthreads <= 1: serial; > 1: threaded twin. */
void ora_step(ora_sim *s, double dt, int nsteps, int nthreads);

void ora_get_points(const ora_sim *s, int body, double *x, double *u, double *xt, double *mxt,
                    double *T9, double *L9, double *v, double *extent, int32_t *active);
/* overwrite point state (n*3 AoS / n*9); NULL pointers leave a field unchanged */
void ora_set_points(ora_sim *s, int body, const double *x, const double *xt, const double *T9,
                    const double *v, const int32_t *active);
/* nodal fields after the last step (dense, node_count entries) */
void ora_get_nodes(const ora_sim *s, int body, double *m, double *mxt, double *f, double *xt,
                   double *u, double *dxt);
void ora_get_contact(const ora_sim *s, int contact, double *normal, double *force);
int64_t ora_node_count(const ora_sim *s);
int64_t ora_element_count(const ora_sim *s);

/* Build the maps for the current state (DefaultPoints::generateMap), then copy the COO
 * tuples of S (which = 0: one value each) or gradS (which = 1: three values each).
 * Returns the tuple count; copies at most cap tuples. Deactivates points that left the
 * domain, exactly as generateMap does. */
int64_t ora_get_map(ora_sim *s, int body, int which, int32_t *node, int32_t *point, double *val,
                    int64_t cap);
/* Per-point connectivity of the current state: code (ORA_CONN_*) and the element of each
 * GIMP corner (2^dim per point, -1 where none); CPDI_OFF points report their own element
 * in slot 0. Does not modify the state. */
void ora_connectivity(const ora_sim *s, int body, int32_t *corner_elem, uint8_t *code);

/* CartesianLinear primitives (cartesian.cpp:219-357), exposed for direct checks */
int ora_which_element(const ora_sim *s, const double *x);
int ora_in_domain(const ora_sim *s, const double *x);
int ora_basis_value(const ora_sim *s, const double *x, int32_t *nid, double *val);
int ora_basis_gradient(const ora_sim *s, const double *x, int32_t *nid, double *grad3);
/* IsotropicLinearElasticity::calculateStress on a single tensor (isolin.cpp:50-78) */
void ora_isolin_update(double E, double nu, double dt, const double *L9, double *T9);

/* Sand_SachithLocal (sand_sachithlocal.cpp): make body `body` (added with a material, before
 * ora_init) a sand with props6 = {grain diameter, mu_s, grains_rho, rho_critical, mu_2, I_0}
 * and USE_DENSITY_FOR_PRESSURE; its plastic strain gammap and rate gammap_dot (per point);
 * one point's update (returns nup_tau; T9 updated in place). */
int ora_set_sand(ora_sim *s, int body, const double *props6, int use_density_pressure);
void ora_get_plastic(const ora_sim *s, int body, double *gammap, double *gammap_dot);
double ora_sand_update(double E, double nu, const double *props6, int use_density_pressure, double dt, double m,
                       double v, const double *L9, double *T9);

void ora_destroy(ora_sim *s);

#ifdef __cplusplus
}
#endif
#endif
