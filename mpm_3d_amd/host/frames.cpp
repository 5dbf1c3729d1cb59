// frames.cpp -- the reference's legacy-VTK frame output for the host side of the HIP solver:
// DefaultVTK::writeFrame (default_vtk.cpp:121-200) with its header and array writers
// (default_vtk.cpp:53-332), MinimalVTK::writeFrame (minimal_vtk.cpp:23-112), the node
// header of CartesianLinear (cartesian.cpp:459-520) and the nodal volume of its hiddenInit
// (cartesian.cpp:70-185). File names, field names, field order, the 6-significant-digit
// ASCII numbers of a default std::ofstream, and the rule that values of inactive points
// and non-finite values are written as 0 follow those files.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "mpm_host.hpp"

namespace mpmhost {

// ------------------------------------------------------------------ grid and nodes
long long CartesianLinear::node_count() const {
    long long n = 1;
    for (int d = 0; d < GRID_DIM; d++) n *= N[d] + 1;
    return n;
}

long long CartesianLinear::element_count() const {
    long long n = 1;
    for (int d = 0; d < GRID_DIM; d++) n *= N[d];
    return n;
}

int CartesianLinear::nodeID(const int *ijk) const {
    long long id = 0, stride = 1;
    for (int d = 0; d < GRID_DIM; d++) {
        id += ijk[d] * stride;
        stride *= N[d] + 1;
    }
    return (int)id;
}

// node n of element e: offsets A(n, d) count in binary with dimension 0 fastest
// (cartesian.cpp:91-138)
void CartesianLinear::elementNodes(long long e, int *ids) const {
    int ijk[3] = {0, 0, 0};
    for (int d = 0; d < GRID_DIM; d++) {
        ijk[d] = (int)(e % N[d]);
        e /= N[d];
    }
    for (int n = 0; n < (1 << GRID_DIM); n++) {
        int c[3] = {0, 0, 0};
        for (int d = 0; d < GRID_DIM; d++) c[d] = ijk[d] + ((n >> d) & 1);
        ids[n] = nodeID(c);
    }
}

void Nodes::init(Job *job, Body *) {
    auto *g = dynamic_cast<CartesianLinear *>(job->grid.get());
    const size_t nn = g ? (size_t)g->node_count() : 0;
    x.assign(3 * nn, 0.0);
    u.assign(3 * nn, 0.0);
    x_t.assign(3 * nn, 0.0);
    diff_x_t.assign(3 * nn, 0.0);
    mx_t.assign(3 * nn, 0.0);
    f.assign(3 * nn, 0.0);
    m.assign(nn, 0.0);
    active.assign(nn, 1);
    for (size_t i = 0; i < nn; i++) {
        long long t = (long long)i;
        for (int d = 0; d < g->GRID_DIM; d++) {
            const int ijk = (int)(t % (g->N[d] + 1));
            t /= g->N[d] + 1;
            x[3 * i + d] = g->hx[d] * ijk;  // nodeIDToPosition (cartesian.cpp:258-275)
        }
    }
}

namespace {

// ------------------------------------------------------------------ writers
struct VtkFrame {
    std::ofstream pfile, nfile;
    long long plen = 0, nlen = 0;
    const Points *P = nullptr;  // the body whose files are open (currentBody)
    const Nodes *Nd = nullptr;

    // writeScalarArray: to each open file whose length matches
    void scalars(const std::vector<double> &a, const std::string &name) {
        const long long n = (long long)a.size();
        if (pfile.is_open() && n == plen) {
            pfile << "SCALARS " << name << " double 1\n";
            pfile << "LOOKUP_TABLE default\n";
            for (long long i = 0; i < plen; i++) {
                if (P->active[i] == 1 && std::isfinite(a[i])) pfile << a[i] << "\n";
                else pfile << "0" << "\n";
            }
        }
        if (nfile.is_open() && n == nlen) {
            nfile << "SCALARS " << name << " double 1\n";
            nfile << "LOOKUP_TABLE default\n";
            for (long long i = 0; i < nlen; i++) {
                if (Nd->active[i] == 1 && std::isfinite(a[i])) nfile << a[i] << "\n";
                else nfile << "0" << "\n";
            }
        }
    }

    // writeVectorArray: a holds 3 doubles per entry
    void vectors(const std::vector<double> &a, const std::string &name) {
        const long long n = (long long)a.size() / 3;
        if (pfile.is_open() && n == plen) {
            pfile << "VECTORS " << name << " double\n";
            for (long long i = 0; i < plen; i++) {
                for (int pos = 0; pos < 3; pos++) {
                    if (P->active[i] == 1 && std::isfinite(a[3 * i + pos])) pfile << a[3 * i + pos] << " ";
                    else pfile << "0 ";
                }
                pfile << "\n";
            }
        }
        if (nfile.is_open() && n == nlen) {
            nfile << "VECTORS " << name << " double\n";
            for (long long i = 0; i < nlen; i++) {
                for (int pos = 0; pos < 3; pos++) {
                    if (Nd->active[i] == 1 && std::isfinite(a[3 * i + pos])) nfile << a[3 * i + pos] << " ";
                    else nfile << "0 ";
                }
                nfile << "\n";
            }
        }
    }

    // writeTensorArray: a holds 9 doubles per entry, row-major
    void tensors(const std::vector<double> &a, const std::string &name) {
        const long long n = (long long)a.size() / 9;
        auto one = [&](std::ofstream &f, long long i, bool zero) {
            if (zero) {
                f << "0 0 0\n";
                f << "0 0 0\n";
                f << "0 0 0\n";
                f << "\n";
            } else {
                f << a[9 * i + 0] << " " << a[9 * i + 1] << " " << a[9 * i + 2] << "\n";
                f << a[9 * i + 3] << " " << a[9 * i + 4] << " " << a[9 * i + 5] << "\n";
                f << a[9 * i + 6] << " " << a[9 * i + 7] << " " << a[9 * i + 8] << "\n";
                f << "\n";
            }
        };
        auto finite = [&](long long i) {
            for (int k = 0; k < 9; k++)
                if (!std::isfinite(a[9 * i + k])) return false;
            return true;
        };
        if (pfile.is_open() && n == plen) {
            pfile << "TENSORS " << name << " double\n";
            for (long long i = 0; i < plen; i++) one(pfile, i, P->active[i] == 0 || !finite(i));
        }
        if (nfile.is_open() && n == nlen) {
            // the reference writes every node tensor three times (default_vtk.cpp:303-324)
            nfile << "TENSORS " << name << " double\n";
            for (long long i = 0; i < nlen; i++)
                for (int pos = 0; pos < 3; pos++) one(nfile, i, !finite(i));
        }
    }

    // writeDefaultPointHeader (default_vtk.cpp:53-85)
    void pointHeader(int dim) {
        pfile << "ASCII\n";
        pfile << "DATASET UNSTRUCTURED_GRID\n";
        pfile << "POINTS " << plen << " double\n";
        for (long long i = 0; i < plen; i++) {
            for (int pos = 0; pos < 3; pos++) {
                const double v = P->x[3 * i + pos];
                if (pos < dim && P->active[i] != 0 && std::isfinite(v)) pfile << v << " ";
                else pfile << "0 ";
            }
            pfile << "\n";
        }
        pfile << "CELLS " << plen << " " << 2 * plen << "\n";
        for (long long i = 0; i < plen; i++) pfile << "1 " << i << "\n";
        pfile << "CELL_TYPES " << plen << "\n";
        for (long long i = 0; i < plen; i++) pfile << "1\n";
        pfile << "POINT_DATA " << plen << "\n";
    }

    // CartesianLinear::writeHeader (cartesian.cpp:459-520)
    void nodeHeader(const CartesianLinear &g, int dim) {
        nfile << "ASCII\n";
        nfile << "DATASET UNSTRUCTURED_GRID\n";
        nfile << "POINTS " << nlen << " double\n";
        for (long long i = 0; i < nlen; i++) {
            for (int pos = 0; pos < 3; pos++) {
                const double v = Nd->x[3 * i + pos];
                if (pos < dim && Nd->active[i] != 0 && std::isfinite(v)) nfile << v << " ";
                else nfile << "0 ";
            }
            nfile << "\n";
        }
        const long long ne = g.element_count();
        const int npe = 1 << g.GRID_DIM;
        int ids[8];
        nfile << "CELLS " << ne << " " << (npe + 1) * ne << "\n";
        for (long long e = 0; e < ne; e++) {
            g.elementNodes(e, ids);
            nfile << npe;
            for (int n = 0; n < npe; n++) nfile << " " << ids[n];
            nfile << "\n";
        }
        nfile << "CELL_TYPES " << ne << "\n";
        const char *type = g.GRID_DIM == 3 ? "11\n" : (g.GRID_DIM == 2 ? "8\n" : "3\n");
        for (long long e = 0; e < ne; e++) nfile << type;
        nfile << "POINT_DATA " << nlen << "\n";
    }

    // open both files of body b and write their headers; frame: sampledFrames after the increment
    bool open(Job *job, Body &b, const std::string &dir, int frame) {
        auto name = [&](const char *pre) {
            std::stringstream ss;
            ss << pre << "." << b.id << "." << b.name << "." << std::setw(10) << std::setfill('0') << (frame - 1)
               << ".vtk";
            return ss.str();
        };
        const std::string pn = name("fpd"), nn = name("fnd");
        pfile.open(dir + pn, std::ios::trunc);
        nfile.open(dir + nn, std::ios::trunc);
        P = b.points.get();
        Nd = b.nodes.get();
        plen = (long long)P->size();
        nlen = (long long)Nd->size();
        auto *g = dynamic_cast<CartesianLinear *>(job->grid.get());
        if (pfile.is_open()) {
            pfile << "# vtk DataFile Version 3.0\n";
            pfile << "Frame: " << (frame - 1) << ", Time: " << job->t << "\n";
            pointHeader(job->DIM);
        } else {
            std::cerr << "Could not open point frame: " << dir + pn << " !" << std::endl;
        }
        if (nfile.is_open()) {
            nfile << "# vtk DataFile Version 3.0\n";
            nfile << "Frame: " << frame << ", Time: " << job->t << "\n";
            if (g) nodeHeader(*g, job->DIM);
        } else {
            std::cerr << "Could not open node frame: " << dir + nn << " !" << std::endl;
        }
        return true;
    }

    void close() {
        pfile.close();
        pfile.clear();
        nfile.close();
        nfile.clear();
    }
};

std::vector<double> as_double(const std::vector<int> &a) { return std::vector<double>(a.begin(), a.end()); }

// bcNodalMask of CartesianBox (cartesian_box.cpp:43-56) and CartesianSmoothBox
// (cartesian_smooth_box.cpp:38-50): faces are the nodes at 0 and at the largest coordinate
std::vector<double> box_mask(Job *job, const Nodes &Nd, bool all_components) {
    auto *g = dynamic_cast<CartesianLinear *>(job->grid.get());
    const size_t nn = Nd.size();
    std::vector<double> mask(3 * nn, 0.0);
    double Lx[3] = {0, 0, 0};
    for (size_t i = 0; i < nn; i++)
        for (int pos = 0; pos < g->GRID_DIM; pos++) Lx[pos] = std::max(Lx[pos], Nd.x[3 * i + pos]);
    for (size_t i = 0; i < nn; i++)
        for (int pos = 0; pos < g->GRID_DIM; pos++)
            if (Nd.x[3 * i + pos] == 0 || Nd.x[3 * i + pos] == Lx[pos]) {
                if (all_components) {
                    for (int d = 0; d < job->DIM; d++) mask[3 * i + d] = 1;  // setOnes over DIM components
                    break;
                }
                mask[3 * i + pos] = 1;
            }
    return mask;
}

// the contacts' frame arrays (ContactHuang::writeFrame, contact_huang.cpp:202-206)
void write_contacts(Job *job, VtkFrame &fr) {
    for (auto &cp : job->contacts) {
        auto *h = dynamic_cast<ContactHuang *>(cp.get());
        if (!h) continue;
        const std::string n1 = job->bodies[h->bodyIDs[0]]->name, n2 = job->bodies[h->bodyIDs[1]]->name;
        fr.vectors(h->contact_normal, "contact_normal_" + n1 + "_" + n2);
        fr.vectors(h->contact_force, "contact_force_by_" + n2 + "_on_" + n1);
    }
}

}  // namespace

// ------------------------------------------------------------------ DefaultVTK
void Serializer::init(Job *) {
    if (str_props.size() < 3 || fp64_props.size() < 1) {
        std::cout << str_props.size() << "\n";
        fprintf(stderr,
                "%s:%s: Need at least 4 properties defined ({sampleRate}, {frameDirectory, outputDirectory, outputName}).\n",
                __FILE__, __func__);
        exit(0);
    }
    auto dir = [](const std::string &s) { return s.empty() || s.back() == '/' ? s : s + "/"; };
    frameDirectory = dir(str_props[0]);
    outputDirectory = dir(str_props[1]);
    simulationName = str_props[2];
    sampleRate = fp64_props[0];
    sampledFrames = 0;
    t_last_frame = 0;
    printf("Serializer properties (frameDirectory = %s, outputDirectory = %s, outputName = %s, sampleRate = %g).\n",
           frameDirectory.c_str(), outputDirectory.c_str(), simulationName.c_str(), sampleRate);
    std::cout << "Serializer Initialized." << std::endl;
}

int Serializer::writeFrame(Job *job) {
    if (!((job->t - job->t0) >= (sampledFrames / sampleRate) || sampledFrames == 0)) return 0;
    t_last_frame = job->t0 + sampledFrames / sampleRate;
    sampledFrames += 1;
    job->solver->syncHost(job, true);
    for (size_t b = 0; b < job->bodies.size(); b++) {
        Body &B = *job->bodies[b];
        Points &P = *B.points;
        Nodes &Nd = *B.nodes;
        VtkFrame fr;
        fr.open(job, B, frameDirectory, sampledFrames);
        if (job->activeBodies[b] != 0) {
            // DefaultPoints::writeFrame (default_points.cpp:341-357)
            fr.vectors(P.u, "displacement");
            fr.vectors(P.x_t, "velocity");
            fr.scalars(P.m, "mass");
            fr.scalars(P.v, "volume");
            fr.vectors(P.mx_t, "momentum");
            fr.vectors(P.b, "body_force");
            fr.tensors(P.T, "cauchy_stress");
            fr.tensors(P.L, "velocity_gradient");
            fr.scalars(as_double(P.active), "active");
            fr.scalars(P.extent, "extent");
            // DefaultNodes::writeFrame (default_nodes.cpp:68-82)
            fr.vectors(Nd.u, "displacement");
            fr.vectors(Nd.x_t, "velocity");
            fr.vectors(Nd.diff_x_t, "diff_velocity");
            fr.scalars(Nd.m, "mass");
            fr.vectors(Nd.mx_t, "momentum");
            fr.vectors(Nd.f, "force");
            fr.scalars(as_double(Nd.active), "active");
            // Sand_SachithLocal::writeFrame (sand_sachithlocal.cpp:121-125); isolin writes nothing
            if (B.activeMaterial != 0)
                if (auto *sand = dynamic_cast<Sand_SachithLocal *>(B.material.get())) {
                    fr.scalars(sand->gammap, "gammap");
                    fr.scalars(sand->gammap_dot, "gammap_dot");
                }
            // CartesianFrictionalBox writes nothing; the two other boxes their mask
            // (cartesian_box.cpp:90-93, cartesian_smooth_box.cpp:86-89)
            if (B.activeBoundary != 0) {
                if (dynamic_cast<CartesianBox *>(B.boundary.get())) fr.vectors(box_mask(job, Nd, true), "bc_nodal_mask");
                else if (dynamic_cast<CartesianSmoothBox *>(B.boundary.get()))
                    fr.vectors(box_mask(job, Nd, false), "bc_nodal_mask");
            }
        }
        write_contacts(job, fr);
        fr.close();
    }
    return 1;
}

// ------------------------------------------------------------------ MinimalVTK
int MinimalVTK::writeFrame(Job *job) {
    if (!((job->t - job->t0) >= (sampledFrames / sampleRate) || sampledFrames == 0)) return 0;
    t_last_frame = job->t0 + sampledFrames / sampleRate;
    sampledFrames += 1;
    job->solver->syncHost(job, true);
    auto *g = dynamic_cast<CartesianLinear *>(job->grid.get());
    if (g && g->v_n.empty()) {
        // nodal volume: each element adds v_e / 2^GRID_DIM to its nodes, in element order
        // (cartesian.cpp:140-153)
        double v_e = 1;
        for (int pos = 0; pos < g->GRID_DIM; pos++) v_e *= g->hx[pos];
        const int npe = 1 << g->GRID_DIM;
        g->v_n.assign((size_t)g->node_count(), 0.0);
        int ids[8];
        for (long long e = 0; e < g->element_count(); e++) {
            g->elementNodes(e, ids);
            for (int c = 0; c < npe; c++) g->v_n[ids[c]] += v_e / npe;
        }
    }
    for (size_t b = 0; b < job->bodies.size(); b++) {
        Body &B = *job->bodies[b];
        Points &P = *B.points;
        Nodes &Nd = *B.nodes;
        VtkFrame fr;
        fr.open(job, B, frameDirectory, sampledFrames);
        const size_t n = P.size();
        fr.vectors(P.x_t, "velocity");
        fr.vectors(P.u, "displacement");
        std::vector<double> tmp(n);
        // pressure = -1/3 tr(T)
        for (size_t i = 0; i < n; i++) {
            const double *T = &P.T[9 * i];
            tmp[i] = -1.0 / 3.0 * (T[0] + T[4] + T[8]);
        }
        fr.scalars(tmp, "pressure");
        // tau_bar = |T - 1/3 tr(T) I| / sqrt(2)
        for (size_t i = 0; i < n; i++) {
            const double *T = &P.T[9 * i];
            const double p = 1.0 / 3.0 * (T[0] + T[4] + T[8]);
            double s2 = 0;
            for (int k = 0; k < 9; k++) {
                const double v = T[k] - (k % 4 == 0 ? p : 0.0);
                s2 += v * v;
            }
            tmp[i] = std::sqrt(s2) / std::sqrt(2.);
        }
        fr.scalars(tmp, "tau_bar");
        // gamma_dot = |sym(L) - 1/3 tr(L) I| sqrt(2)
        for (size_t i = 0; i < n; i++) {
            const double *L = &P.L[9 * i];
            const double tr = 1.0 / 3.0 * (L[0] + L[4] + L[8]);
            double s2 = 0;
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    const double v = (0.5 * L[3 * r + c] + 0.5 * L[3 * c + r]) - (r == c ? tr : 0.0);
                    s2 += v * v;
                }
            tmp[i] = std::sqrt(s2) * std::sqrt(2.);
        }
        fr.scalars(tmp, "gamma_dot");
        for (size_t i = 0; i < n; i++) tmp[i] = P.m[i] / P.v[i];
        fr.scalars(tmp, "density");
        // nodal velocity and density m / v_n
        fr.vectors(Nd.x_t, "velocity");
        std::vector<double> rho(Nd.m);
        for (size_t i = 0; i < rho.size(); i++) rho[i] /= g->v_n[i];
        fr.scalars(rho, "density");
        fr.close();
    }
    return 1;
}

}  // namespace mpmhost
