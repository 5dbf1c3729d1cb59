// hip_explicit_usl.cpp -- HipExplicitUSL: the explicit USL step (explicit_usl.cpp:48-87) on
// the GPU through the C-ABI of include/mpm_gpu.h.
//
// init reads the ExplicitUSL properties (explicit_usl.cpp:25-44). Job::init calls the
// solver before the grid and bodies, so the device state is created on the first step:
// the plugin set is checked (anything the GPU path does not implement is refused with the
// reference's convention, a message on stderr and exit(0)), the AoS point arrays are
// converted to SoA planes and uploaded once. A step is queued; the queued steps run as one
// mpm_gpu_step call when the host arrays are needed (syncHost: a frame, the end of the
// run, a saveState), so the point state is copied back only then.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>

#include "mpm_gpu.h"
#include "mpm_host.hpp"

namespace mpmhost {

static void refuse(const char *func, const std::string &what) {
    fprintf(stderr, "%s:%s: %s\n", __FILE__, func, what.c_str());
    exit(0);
}

static void check(int rc, const char *what) {
    if (rc != MPM_OK) throw std::runtime_error(std::string(what) + ": " + mpm_gpu_last_error());
}

HipExplicitUSL::~HipExplicitUSL() {
    if (handle) mpm_gpu_destroy(handle);
}

void HipExplicitUSL::init(Job *) {
    cpdi_spec = 1;     // DefaultBody::CPDI_ON
    contact_spec = 1;  // Contact::IMPLICIT
    if (int_props.size() >= 1) cpdi_spec = int_props[0];
    if (int_props.size() >= 2) contact_spec = int_props[1];
    if (int_props.size() >= 3) precision = int_props[2];
    if (int_props.size() >= 4) sort_interval = int_props[3];
    printf("Solver properties (cpdi_spec = %i, contact_spec = %i, precision = %i).\n", cpdi_spec, contact_spec,
           precision);
    std::cout << "Solver Initialized." << std::endl;
}

void HipThreadedAlias::init(Job *) {
    cpdi_spec = 1;
    contact_spec = 1;
    num_threads = 1;
    if (int_props.size() >= 1) cpdi_spec = int_props[0];
    if (int_props.size() >= 2) contact_spec = int_props[1];
    if (int_props.size() >= 3) num_threads = int_props[2];
    printf("Solver properties (cpdi_spec = %i, contact_spec = %i, num_threads = %i).\n", cpdi_spec, contact_spec,
           num_threads);
    std::cout << "Solver Initialized." << std::endl;
}

void HipExplicitUSL::upload(Job *job) {
    if (job->JOB_TYPE != Job::JOB_2D && job->JOB_TYPE != Job::JOB_3D)
        refuse(__func__, "HipExplicitUSL supports TYPE = 2 and 3 only.");
    auto *grid = dynamic_cast<CartesianLinear *>(job->grid.get());
    if (!grid) refuse(__func__, "HipExplicitUSL needs the CartesianLinear grid.");
    if (precision != MPM_PREC_FP64 && precision != MPM_PREC_FP32) refuse(__func__, "precision must be 0 or 1.");
    const int dim = job->DIM;
    mpm_grid_desc g{};
    g.dim = dim;
    for (int d = 0; d < dim; d++) {
        g.L[d] = grid->L[d];
        g.N[d] = grid->N[d];
    }
    g.cpdi_spec = cpdi_spec;
    g.contact_spec = contact_spec;
    for (int d = 0; d < 3; d++) g.gravity[d] = job->driver->gravity[d];

    std::vector<mpm_body_desc> bd;
    for (auto &bp : job->bodies) {
        if (!bp || !bp->points) refuse(__func__, "every body needs points.");
        mpm_body_desc b{};
        b.n_points = (int64_t)bp->points->size();
        if (bp->activeMaterial) {
            if (auto *mat = dynamic_cast<IsotropicLinearElasticity *>(bp->material.get())) {
                b.has_material = 1;
                b.E = mat->E;
                b.nu = mat->nu;
            } else if (auto *sand = dynamic_cast<Sand_SachithLocal *>(bp->material.get())) {
                b.has_material = 2;
                b.E = sand->E;
                b.nu = sand->nu;
                const double sp[6] = {sand->grains_d, sand->MU_S, sand->GRAINS_RHO, sand->RHO_CRITICAL, sand->MU_2,
                                      sand->I_0};
                for (int k = 0; k < 6; k++) b.sand[k] = sp[k];
                b.sand_density_pressure = sand->use_density_to_determine_pressure ? 1 : 0;
            } else {
                refuse(__func__, "HipExplicitUSL supports the IsotropicLinearElasticity and Sand_SachithLocal "
                                 "materials only.");
            }
        }
        if (bp->activeBoundary) {
            Boundary *bc = bp->boundary.get();
            if (dynamic_cast<CartesianBox *>(bc)) b.has_boundary = 1;
            else if (dynamic_cast<CartesianSmoothBox *>(bc)) b.has_boundary = 2;
            else if (auto *fb = dynamic_cast<CartesianFrictionalBox *>(bc)) {
                b.has_boundary = 3;
                b.bc_mu = fb->mu_f;
            } else {
                refuse(__func__, "HipExplicitUSL supports the CartesianBox, CartesianSmoothBox and "
                                 "CartesianFrictionalBox boundaries only.");
            }
        }
        bd.push_back(b);
    }
    if (bd.size() > MPM_MAX_BODIES_PER_SIM) refuse(__func__, "at most 4 bodies per simulation.");
    std::vector<mpm_contact_desc> cd;
    for (auto &cp : job->contacts) {
        auto *c = dynamic_cast<ContactHuang *>(cp.get());
        if (!c) refuse(__func__, "HipExplicitUSL supports ContactHuang contacts only.");
        cd.push_back(mpm_contact_desc{c->bodyIDs[0], c->bodyIDs[1], c->mu_f});
    }
    check(mpm_gpu_create(&g, bd.data(), (int)bd.size(), cd.empty() ? nullptr : cd.data(), (int)cd.size(), precision,
                         &handle),
          "mpm_gpu_create");
    check(mpm_gpu_set_sort_interval(handle, sort_interval), "mpm_gpu_set_sort_interval");
    for (size_t b = 0; b < job->bodies.size(); b++) {
        const Points &P = *job->bodies[b]->points;
        const size_t n = P.size();
        std::vector<double> x(dim * n), v(dim * n), T(6 * n);
        std::vector<int32_t> act(n);
        static const int t6[6] = {0, 4, 8, 1, 2, 5};  // xx yy zz xy xz yz in a row-major 3x3
        for (size_t i = 0; i < n; i++) {
            for (int d = 0; d < dim; d++) {
                x[d * n + i] = P.x[3 * i + d];
                v[d * n + i] = P.x_t[3 * i + d];
            }
            for (int k = 0; k < 6; k++) T[k * n + i] = P.T[9 * i + t6[k]];
            act[i] = P.active[i];
        }
        check(mpm_gpu_upload(handle, (int)b, x.data(), v.data(), T.data(), P.v.data(), P.m.data(), P.extent.data(),
                             act.data()),
              "mpm_gpu_upload");
    }
}

// Issue the queued steps. record_last: the last of them records the node fields
// (MPM_OPT_NODE_RECORD), which a frame with node files reads.
void HipExplicitUSL::flush(Job *job, bool record_last) {
    if (pending > 0) {
        long long left = pending - (record_last ? 1 : 0);
        while (left > 0) {
            const int k = (int)std::min<long long>(left, 1 << 30);
            check(mpm_gpu_step(handle, job->dt, k), "mpm_gpu_step");
            left -= k;
        }
        if (record_last) {
            check(mpm_gpu_set_option(handle, MPM_OPT_NODE_RECORD, 1), "mpm_gpu_set_option");
            check(mpm_gpu_step(handle, job->dt, 1), "mpm_gpu_step");
            check(mpm_gpu_set_option(handle, MPM_OPT_NODE_RECORD, 0), "mpm_gpu_set_option");
        }
        pending = 0;
        host_stale = true;
        nodes_stale = true;
    }
}

void HipExplicitUSL::step(Job *job) {
    if (!handle) upload(job);
    pending++;
}

void HipExplicitUSL::syncHost(Job *job, bool nodes) {
    if (!handle) return;
    flush(job, nodes);
    if (nodes && nodes_stale) syncNodes(job);
    if (!host_stale) return;
    const int dim = job->DIM;
    for (size_t b = 0; b < job->bodies.size(); b++) {
        Points &P = *job->bodies[b]->points;
        const size_t n = P.size();
        if (n == 0) continue;
        std::vector<double> x(dim * n), v(dim * n), T(6 * n), V(n), u(dim * n), L(9 * n);
        std::vector<int32_t> act(n);
        check(mpm_gpu_download(handle, (int)b, x.data(), v.data(), T.data(), V.data(), act.data()), "mpm_gpu_download");
        check(mpm_gpu_download_aux(handle, (int)b, u.data(), L.data()), "mpm_gpu_download_aux");
        static const int r6[6] = {0, 1, 2, 0, 0, 1}, c6[6] = {0, 1, 2, 1, 2, 2};
        for (size_t i = 0; i < n; i++) {
            for (int d = 0; d < dim; d++) {
                P.x[3 * i + d] = x[d * n + i];
                P.x_t[3 * i + d] = v[d * n + i];
                P.u[3 * i + d] = u[d * n + i];
            }
            for (int d = 0; d < 3; d++) P.mx_t[3 * i + d] = P.m[i] * P.x_t[3 * i + d];
            for (int k = 0; k < 6; k++) {
                P.T[9 * i + 3 * r6[k] + c6[k]] = T[k * n + i];
                P.T[9 * i + 3 * c6[k] + r6[k]] = T[k * n + i];
            }
            for (int k = 0; k < 9; k++) P.L[9 * i + k] = L[k * n + i];
            P.v[i] = V[i];
            P.active[i] = act[i];
        }
        if (auto *sand = dynamic_cast<Sand_SachithLocal *>(job->bodies[b]->material.get())) {
            sand->gammap.assign(n, 0.0);
            sand->gammap_dot.assign(n, 0.0);
            check(mpm_gpu_download_plastic(handle, (int)b, sand->gammap.data(), sand->gammap_dot.data()),
                  "mpm_gpu_download_plastic");
        }
    }
    host_stale = false;
}

// Nodes::m, mx_t, f, x_t, diff_x_t and u = dt x_t (moveGrid, explicit_usl.cpp:233-264) of
// every body and ContactHuang's contact_normal / contact_force, from the recorded last step.
void HipExplicitUSL::syncNodes(Job *job) {
    const int dim = job->DIM;
    for (size_t b = 0; b < job->bodies.size(); b++) {
        Nodes &N = *job->bodies[b]->nodes;
        const size_t nn = N.size();
        std::vector<double> m(nn), mxt(dim * nn), f(dim * nn), xt(dim * nn), dxt(dim * nn);
        check(mpm_gpu_download_nodes(handle, (int)b, m.data(), mxt.data(), f.data(), xt.data(), dxt.data()),
              "mpm_gpu_download_nodes");
        for (size_t i = 0; i < nn; i++) {
            N.m[i] = m[i];
            for (int d = 0; d < dim; d++) {
                N.mx_t[3 * i + d] = mxt[d * nn + i];
                N.f[3 * i + d] = f[d * nn + i];
                N.x_t[3 * i + d] = xt[d * nn + i];
                N.diff_x_t[3 * i + d] = dxt[d * nn + i];
                N.u[3 * i + d] = job->dt * N.x_t[3 * i + d];
            }
        }
    }
    for (size_t c = 0; c < job->contacts.size(); c++) {
        auto *h = dynamic_cast<ContactHuang *>(job->contacts[c].get());
        if (!h) continue;
        const size_t nn = job->bodies[h->bodyIDs[0]]->nodes->size();
        std::vector<double> nrm(dim * nn), frc(dim * nn);
        check(mpm_gpu_download_contact(handle, (int)c, nrm.data(), frc.data()), "mpm_gpu_download_contact");
        h->contact_normal.assign(3 * nn, 0.0);
        h->contact_force.assign(3 * nn, 0.0);
        for (size_t i = 0; i < nn; i++)
            for (int d = 0; d < dim; d++) {
                h->contact_normal[3 * i + d] = nrm[d * nn + i];
                h->contact_force[3 * i + d] = frc[d * nn + i];
            }
    }
    nodes_stale = false;
}

// The reference's ExplicitUSL returns "err" / 0 here (explicit_usl.cpp:91-99); this solver
// brings the host arrays up to date so a serializer can save them.
std::string HipExplicitUSL::saveState(Job *job, Serializer *, std::string) {
    syncHost(job);
    return "err";
}

int HipExplicitUSL::loadState(Job *, Serializer *, std::string) { return 0; }

}  // namespace mpmhost
