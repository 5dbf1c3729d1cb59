/*
 * mpm_gpu.h -- C-ABI of the MI355X explicit MPM step (libmpm_gpu.so).
 *
 * One handle holds one or more independent simulations ("sims"), each a CartesianLinear
 * grid with its bodies and contacts, resident in HBM. Every entry point returns 0 on
 * success or a negative MPM_ERR_* code; mpm_gpu_last_error() describes the last failure
 * of the calling thread. A handle owns one HIP stream and is not thread-safe; host
 * buffers belong to the caller and are not retained after a call returns.
 *
 * Array layout across the boundary is structure-of-arrays: a vector field of a body
 * with n points is `dim` planes of n doubles (x[d*n + p]); a stress is 6 planes in the
 * order xx, yy, zz, xy, xz, yz (the reference's symmetric MaterialTensor); a velocity
 * gradient is 9 planes, row-major (L[(3*r + c)*n + p]).
 *
 * The reference has no FFI for this path: its ExplicitUSL calls the Points/Nodes/Grid/
 * Material/Boundary/Contact virtuals in-process (src/objects/solvers/explicit_usl.cpp:48-87).
 * Each entry point below names the reference interface it stands in for.
 */
#ifndef MPM_GPU_H
#define MPM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_OK 0
#define MPM_ERR_ARG -1      /* invalid argument or unsupported configuration */
#define MPM_ERR_HIP -2      /* HIP runtime failure (allocation, launch, copy) */
#define MPM_ERR_STATE -3    /* call out of order (e.g. step before upload) */
#define MPM_ERR_NOMEM -4    /* device memory exhausted */

/* precision forms */
#define MPM_PREC_FP64 0     /* plain form: particle state and grid in FP64 */
#define MPM_PREC_FP32 1     /* particle state and stress in FP32, grid accumulation in FP64 */

/* connectivity codes (mpm_gpu_connectivity) */
#define MPM_CONN_CPDI_OFF 0 /* linear weights at the point (cartesian.cpp:280-357) */
#define MPM_CONN_CPDI_ON 1  /* GIMP/CPDI corners (default_points.cpp:264-324) */
#define MPM_CONN_INACTIVE 2 /* inactive or outside the domain */

#define MPM_MAX_BODIES_PER_SIM 4

typedef struct mpm_gpu_handle mpm_gpu_handle;

/* Grid and solver settings of one simulation.
 * Replaces: CartesianLinear::init properties {Lx..}, int-properties {Nx..}
 * (src/objects/grids/cartesian.cpp:24-67), Job TYPE (include/job.hpp:58-73),
 * ExplicitUSL int-properties {cpdi_spec, contact_spec} (explicit_usl.cpp:25-44) and the
 * driver's gravity (default_driver.cpp:90-105). */
typedef struct {
    int dim;              /* 2 or 3 (TYPE 2 / TYPE 3) */
    double L[3];          /* domain lengths */
    int N[3];             /* elements per dimension */
    int cpdi_spec;        /* 1: GIMP/CPDI corners; 0: linear basis at the point */
    int contact_spec;     /* 1: IMPLICIT; 0: EXPLICIT (mpm_objects.hpp:217-218) */
    double gravity[3];    /* body force per unit mass, constant over the run */
} mpm_grid_desc;

/* One body. Replaces: DefaultBody with DefaultPoints/DefaultNodes, its material
 *   has_material 0  none
 *                1  IsotropicLinearElasticity {E, nu} (isolin.cpp:26-46)
 *                2  Sand_SachithLocal {E, nu, grain diameter, mu_s, grains_rho, rho_critical
 *                   [, mu_2, I_0]} with sand = {grain diameter, mu_s, grains_rho, rho_critical,
 *                   mu_2, I_0} (mu_2 = mu_s and I_0 = 1 when the reference is given six
 *                   properties) and sand_density_pressure for its USE_DENSITY_FOR_PRESSURE
 *                   flag (sand_sachithlocal.cpp:58-117)
 * and the body's boundary:
 *   has_boundary 0  none
 *                1  CartesianBox, all components fixed on the box faces (cartesian_box.cpp:23-60)
 *                2  CartesianSmoothBox, the face-normal component fixed (cartesian_smooth_box.cpp:24-80)
 *                3  CartesianFrictionalBox, face-normal fixed plus Coulomb friction with
 *                   mu_f = bc_mu against the normal impulse (cartesian_frictional_box.cpp:23-151) */
typedef struct {
    int64_t n_points;
    int has_material;
    double E, nu;
    int has_boundary;
    double bc_mu;
    double sand[6];
    int sand_density_pressure;
} mpm_body_desc;

/* ContactHuang between two bodies of the same sim, {mu_f} and {body_1, body_2}
 * (contact_huang.cpp:24-75). body indices are sim-local. */
typedef struct {
    int body1, body2;
    double mu;
} mpm_contact_desc;

/* One simulation of a batch: grid settings plus its bodies and contacts. */
typedef struct {
    mpm_grid_desc grid;
    int nbody;
    const mpm_body_desc *bodies;
    int ncontact;
    const mpm_contact_desc *contacts;
} mpm_sim_desc;

/* Create a handle for one simulation. Replaces Job::init of the grid, bodies and
 * contacts (include/job.hpp:76-105) for the supported plugin set. */
int mpm_gpu_create(const mpm_grid_desc *grid, const mpm_body_desc *bodies, int nbody,
                   const mpm_contact_desc *contacts, int ncontact, int precision,
                   mpm_gpu_handle **out);

/* Create a handle holding nsim independent simulations advanced together. Bodies are
 * numbered globally in sim order (sim 0's bodies first). */
int mpm_gpu_create_batch(const mpm_sim_desc *sims, int nsim, int precision,
                         mpm_gpu_handle **out);

/* Upload the state of one body (global body index). x, v: dim planes; T: 6 planes or
 * NULL for zero stress; V volume, m mass, extent the GIMP half-width computed on the
 * host as 0.5*sqrt(V) (2D) or 0.5*cbrt(V) (3D) (default_points.cpp:39-57); active 0/1.
 * Replaces DefaultPoints::readFromFile + init (default_points.cpp:35-193). */
int mpm_gpu_upload(mpm_gpu_handle *h, int body, const double *x, const double *v,
                   const double *T, const double *V, const double *m, const double *extent,
                   const int32_t *active);

/* Advance every sim of the handle by nsteps steps of size dt. Replaces nsteps calls of
 * ExplicitUSL::step
 * (explicit_usl.cpp:48-87) with DefaultDriver gravity (default_driver.cpp:46-86). */
int mpm_gpu_step(mpm_gpu_handle *h, double dt, int nsteps);

/* Serving a stream of simulations from one batch handle: each simulation of the handle
 * is a slot whose point, body, contact and grid sizes at creation are its capacity.
 * mpm_gpu_load_sim reconfigures slot `sim` for a new simulation `d` (grid L and N <= the
 * slot's N, up to its bodies and contacts, n_points per body <= the body's capacity),
 * deactivating whatever the slot held; then mpm_gpu_upload fills each body (n_points
 * entries). mpm_gpu_clear_sim empties a slot. Other slots keep their state and can keep
 * stepping between these calls. Replaces building a new Job per request
 * (config.cpp:176-206 + Job::init, job.hpp:76-106); no counterpart in the reference. */
int mpm_gpu_load_sim(mpm_gpu_handle *h, int sim, const mpm_sim_desc *d);
int mpm_gpu_clear_sim(mpm_gpu_handle *h, int sim);

/* Batched form of mpm_gpu_step: every simulation of the handle advances nsteps steps,
 * simulation s with its own time step dt[s] (nsim entries, in create order). Replaces
 * ExplicitUSL::step of each simulation's Job (explicit_usl.cpp:48-87), in one launch
 * sequence for all. */
int mpm_gpu_step_batch(mpm_gpu_handle *h, const double *dt, int nsim, int nsteps);

/* Download one body in its original point order. Any pointer may be NULL.
 * Replaces reads of Points::x, x_t, T, v, active (mpm_objects.hpp:137-159). */
int mpm_gpu_download(mpm_gpu_handle *h, int body, double *x, double *v, double *T,
                     double *V, int32_t *active);

/* Displacement u (dim planes; accumulated x - x0 since upload) and the velocity gradient
 * L of the last step (9 planes; zero for points without connectivity).
 * Replaces reads of Points::u and Points::L. */
int mpm_gpu_download_aux(mpm_gpu_handle *h, int body, double *u, double *L);

/* Sand_SachithLocal's per-point gammap (accumulated plastic shear strain) and gammap_dot
 * (its rate in the last step) of one body, original order; zeros for other materials.
 * Either pointer may be NULL. Replaces reads of Sand_SachithLocal::gammap, gammap_dot, which
 * its writeFrame serializes (sand_sachithlocal.cpp:121-125). */
int mpm_gpu_download_plastic(mpm_gpu_handle *h, int body, double *gammap, double *gammap_dot);

/* Node fields of one body after the last step call, in the reference's node numbering
 * (CartesianLinear: n = i + (N0+1)(j + (N1+1)k), (N0+1)(N1+1)[(N2+1)] nodes for the body's
 * sim), vectors as dim planes of that many doubles. m (mass), mxt (momentum after
 * moveGrid) and f (force after contact and boundary) need MPM_OPT_NODE_RECORD on during
 * that call; xt and dxt (velocity and its increment, moveGrid) are always available. Nodes
 * the last grid update did not reach (no point near them) report 0, as the reference's
 * massless nodes. Any pointer may be NULL. MPM_ERR_STATE before the first step after an
 * upload or load. Replaces reads of Nodes::m, mx_t, f, x_t, diff_x_t (mpm_objects.hpp:164-167),
 * which DefaultNodes::writeFrame serializes (default_nodes.cpp:68-82). */
int mpm_gpu_download_nodes(mpm_gpu_handle *h, int body, double *m, double *mxt, double *f, double *xt,
                           double *dxt);

/* ContactHuang's contact_normal and contact_force (contacts.hpp:56) of contact `contact`
 * (global index, create order) over its sim's nodes, numbered as above, after the last step
 * call (recorded under MPM_OPT_NODE_RECORD): the normal is body 1's normalised mass gradient
 * (NaN where it is zero, as the reference's 0/0), the force the contact force on body 1 where
 * both bodies have mass (contact_huang.cpp:80-197), else 0. Either pointer may be NULL. */
int mpm_gpu_download_contact(mpm_gpu_handle *h, int contact, double *normal, double *force);

/* Per-point connectivity of the current state, original order: code (MPM_CONN_*) and the
 * element of each of the 2^dim GIMP corners (-1 where none; CPDI-off points report their
 * own element in slot 0). Replaces DefaultPoints::generateMap's element lookups
 * (default_points.cpp:196-330, cartesian.cpp:219-253). */
int mpm_gpu_connectivity(mpm_gpu_handle *h, int body, int32_t *corner_elem, uint8_t *code);

/* No counterpart in the reference. The same result as mpm_gpu_connectivity, for callers that
 * cannot allocate device memory (code between HIP graph captures, serving loops that must not
 * call hipMalloc): it calls no hipMalloc/hipFree or allocating library routine. Its kernel
 * writes into a fixed device array of the library (MPM_NOALLOC_MAX_POINTS points x 2^dim
 * corner elements + codes, 16.5 MiB held on each device the library runs on until the process
 * ends) instead of a per-call buffer. Bodies with more than MPM_NOALLOC_MAX_POINTS points are
 * refused with MPM_ERR_ARG. Every call on a device shares that device's array: calls that
 * overlap on one device (other threads, other handles) race and must be serialised by the
 * caller. The handle's device (current at create) must be current (else MPM_ERR_STATE). */
#define MPM_NOALLOC_MAX_POINTS 524288
int mpm_gpu_connectivity_noalloc(mpm_gpu_handle *h, int body, int32_t *corner_elem, uint8_t *code);

/* Points re-sorted by grid block every `interval` steps, counted across mpm_gpu_step
 * calls (default 16; 0 = at the start of every mpm_gpu_step call). An upload always
 * forces a sort before the next step. Affects speed only. */
int mpm_gpu_set_sort_interval(mpm_gpu_handle *h, int interval);

/* Execution options (speed only; results agree to rounding):
 *   MPM_OPT_SORT_INTERVAL  as mpm_gpu_set_sort_interval
 *   MPM_OPT_FUSED_PASS     1: one particle pass per step doing G2P and the next step's
 *                          P2G; 0 (default): separate G2P and P2G passes */
#define MPM_OPT_SORT_INTERVAL 1
#define MPM_OPT_FUSED_PASS 2
/*   MPM_OPT_GRAPH          1: replay each sort period (a sort + sort_interval steps) as
 *                          a captured HIP graph (split passes, timing off); 0 (default)
 *   MPM_OPT_AUTOTUNE       1: the next step call first times candidate launch shapes
 *                          (persistent particle-pass grids, grid-update grid) and then, when
 *                          it has 57 steps to spare, sort intervals 8, 16 and 32, on its own
 *                          first steps, and keeps the fastest; those steps count. Reads 2
 *                          once tuned
 *   MPM_OPT_PASS_GRID      workgroups of the persistent G2P and fused particle passes
 *                          (default 1024); setting it sets MPM_OPT_P2G_GRID too
 *   MPM_OPT_P2G_GRID       workgroups of the persistent P2G pass (default: as many as the
 *                          device holds at once for that pass, e.g. 1024 for the 3D
 *                          precision form and 768 for 3D FP64 on MI355X)
 *   MPM_OPT_GRID_WGS       workgroups of the grid update (default 2048)
 *   MPM_OPT_GRAPH_CAPTURES (get only) graphs captured so far
 *   MPM_OPT_NODE_RECORD    1: every grid update also records nodal mass, momentum and force
 *                          and the contacts' normals and forces, for mpm_gpu_download_nodes
 *                          and mpm_gpu_download_contact ((1 + 2 dim) doubles per node plus
 *                          2 dim per node of each contact's sim, allocated on first use and
 *                          not counted by mpm_gpu_footprint); graph replay is off meanwhile
 *   MPM_OPT_KEEP_L         1 (default): the last step of every step call writes the velocity
 *                          gradient for mpm_gpu_download_aux (9 doubles per point slot,
 *                          allocated at the first step and not counted by mpm_gpu_footprint);
 *                          0: not kept, the plane is freed and mpm_gpu_download_aux's L
 *                          returns MPM_ERR_STATE
 *   MPM_OPT_CELL_SORT      1 (default where it fits): the sort also orders each particle
 *                          block's points by dual cell within their octant, so that the
 *                          particle passes' LDS accesses fall on distinct banks; it keeps 256
 *                          bytes of counters per (block, octant) of every body's
 *                          particle-block grid plus a byte per point slot (counted by
 *                          mpm_gpu_footprint), and is off (setting 1 returns MPM_ERR_ARG) on
 *                          handles where the counters would exceed 4 GiB; 0: blocks and
 *                          octants only */
#define MPM_OPT_GRAPH 3
#define MPM_OPT_AUTOTUNE 4
#define MPM_OPT_PASS_GRID 5
#define MPM_OPT_GRID_WGS 6
#define MPM_OPT_GRAPH_CAPTURES 7
#define MPM_OPT_P2G_GRID 8
#define MPM_OPT_NODE_RECORD 9
#define MPM_OPT_KEEP_L 10
#define MPM_OPT_CELL_SORT 11
int mpm_gpu_set_option(mpm_gpu_handle *h, int key, int value);
int mpm_gpu_get_option(mpm_gpu_handle *h, int key, int *value);

/* Kernel timing with HIP events on the handle's stream: enable, then read the number of
 * timed launches and their summed milliseconds per kernel class
 * (0 = fused particle pass G2P+P2G, 1 = grid update, 2 = sort, 3 = P2G pass, 4 = G2P pass). */
int mpm_gpu_set_timing(mpm_gpu_handle *h, int enable);
int mpm_gpu_get_timing(mpm_gpu_handle *h, int kernel_class, int64_t *launches, double *ms);

/* Number of grid nodes that held mass in the last step (summed over bodies) and the
 * number of active points. */
int mpm_gpu_stats(mpm_gpu_handle *h, int64_t *active_nodes, int64_t *active_points);

/* Testing aid with no counterpart in the reference: evaluates, on the host, the same
 * shape-function code the kernels run for one point (position x, GIMP half-width
 * extent) and writes its node tuples -- node i,j,k (3 ints each), S, and gradS (3
 * doubles each) -- up to max_tuples. Returns the tuple count, -100 when the point is
 * outside the domain, or a negative MPM_ERR_* code. Needs no GPU. */
int mpm_gpu_host_weights(const mpm_grid_desc *grid, const double *x, double extent, int32_t *node_ijk,
                         double *S, double *G, int max_tuples);

/* Measurement aid with no counterpart in the reference: times `reps` launches of a
 * plain device copy of nbytes (read + write) and returns the achieved HBM GB/s. Its
 * dispatches also calibrate the FETCH_SIZE / WRITE_SIZE counters. */
int mpm_gpu_copy_probe(int64_t nbytes, int reps, double *gbps);

int mpm_gpu_synchronize(mpm_gpu_handle *h);
int mpm_gpu_destroy(mpm_gpu_handle *h);
const char *mpm_gpu_last_error(void);
/* bytes of device memory a handle with these sims allocates at create (the velocity gradient
 * of MPM_OPT_KEEP_L and the node record of MPM_OPT_NODE_RECORD are allocated on first use
 * and not counted) */
int64_t mpm_gpu_footprint(const mpm_sim_desc *sims, int nsim, int precision);

#ifdef __cplusplus
}
#endif
#endif
