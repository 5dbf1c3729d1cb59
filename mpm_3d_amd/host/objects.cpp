// objects.cpp -- configuration grammar, registry and the plugin objects around the HIP
// solver: points I/O, material/grid/contact/driver initialisation, the driver loop and
// point frames. The step itself is in hip_explicit_usl.cpp.
#include <sys/stat.h>
#include <time.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "mpm_host.hpp"

namespace mpmhost {

// ------------------------------------------------------------------ parser (parser.hpp)
namespace parser {
std::string removeSpaces(std::string s) {
    std::string out;
    for (char c : s)
        if (!std::isspace((unsigned char)c)) out.push_back(c);
    return out;
}

// std::getline semantics: a trailing delimiter yields no empty last element
std::vector<std::string> splitString(const std::string &s, char delim) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, delim)) out.push_back(tok);
    return out;
}

std::string removeComments(const std::string &s) {
    auto v = splitString(s, '#');
    return v.empty() ? std::string() : v[0];
}

static std::string drop(const std::string &s, char c) {
    std::string out;
    for (char ch : s)
        if (ch != c) out.push_back(ch);
    return out;
}

std::string removeBraces(const std::string &s) { return drop(drop(s, '{'), '}'); }
std::string removeQuotes(const std::string &s) { return drop(drop(s, '\''), '"'); }
}  // namespace parser

// ------------------------------------------------------------------ registry
template <class T, class U> static std::pair<std::string, std::function<std::unique_ptr<T>()>> reg(const char *n) {
    return {n, [] { return std::unique_ptr<T>(new U()); }};
}

template <> Registry<Serializer>::Registry() {
    table.insert(reg<Serializer, Serializer>("DefaultVTK"));
    table.insert(reg<Serializer, MinimalVTK>("MinimalVTK"));
    table.insert(reg<Serializer, NullSerializer>("NullSerializer"));
}
template <> Registry<Driver>::Registry() {
    table.insert(reg<Driver, Driver>("DefaultDriver"));
    table.insert(reg<Driver, UserDefinedGravityDriver>("UserDefinedGravityDriver"));
}
// "ExplicitUSL" resolves to the HIP solver in this host so that unchanged configuration
// files run on the GPU; inside the reference the class registers as HipExplicitUSL only.
template <> Registry<Solver>::Registry() {
    table.insert(reg<Solver, HipExplicitUSL>("HipExplicitUSL"));
    table.insert(reg<Solver, HipExplicitUSL>("ExplicitUSL"));
    table.insert(reg<Solver, HipThreadedAlias>("ParallelExplicitUSL"));
    table.insert(reg<Solver, HipThreadedAlias>("ThreadPoolExplicitUSL"));
}
template <> Registry<Grid>::Registry() { table.insert(reg<Grid, CartesianLinear>("CartesianLinear")); }
template <> Registry<Body>::Registry() { table.insert(reg<Body, Body>("DefaultBody")); }
template <> Registry<Points>::Registry() {
    table.insert(reg<Points, Points>("DefaultPoints"));
    table.insert(reg<Points, Points>("ThreadPoolPoints"));
}
template <> Registry<Nodes>::Registry() {
    table.insert(reg<Nodes, Nodes>("DefaultNodes"));
    table.insert(reg<Nodes, Nodes>("ThreadPoolNodes"));
}
template <> Registry<Contact>::Registry() { table.insert(reg<Contact, ContactHuang>("ContactHuang")); }
template <> Registry<Material>::Registry() {
    table.insert(reg<Material, IsotropicLinearElasticity>("IsotropicLinearElasticity"));
    table.insert(reg<Material, Sand_SachithLocal>("Sand_SachithLocal"));
}
template <> Registry<Boundary>::Registry() {
    table.insert(reg<Boundary, CartesianBox>("CartesianBox"));
    table.insert(reg<Boundary, CartesianSmoothBox>("CartesianSmoothBox"));
    table.insert(reg<Boundary, CartesianFrictionalBox>("CartesianFrictionalBox"));
}

// ------------------------------------------------------------------ configurator (config.cpp)
Configurator::Configurator(std::string fileIN) : file(std::move(fileIN)) {
    const size_t k = file.find_last_of('/');
    directory = k == std::string::npos ? std::string() : file.substr(0, k + 1);
}

template <class T> static T stringTo(const std::string &a);
template <> std::string stringTo<std::string>(const std::string &a) { return a; }
template <> int stringTo<int>(const std::string &a) { return std::stoi(a); }
template <> double stringTo<double>(const std::string &a) { return std::stod(a); }

template <class T> static std::vector<T> parsePropertyList(const std::string &s) {
    std::vector<T> out;
    for (auto &tok : parser::splitString(s, ',')) out.push_back(stringTo<T>(tok));
    return out;
}

// key = value lines of one braced section; a repeated key keeps the last value
static std::map<std::string, std::string> sectionPairs(std::ifstream &fin) {
    std::map<std::string, std::string> kv;
    std::string line;
    std::getline(fin, line);
    line = parser::removeSpaces(parser::removeComments(line));
    if (line != "{") return kv;
    while (std::getline(fin, line)) {
        line = parser::removeSpaces(parser::removeComments(line));
        if (line == "}") break;
        line = parser::removeQuotes(parser::removeBraces(line));
        auto tok = parser::splitString(line, '=');
        if (tok.size() > 1) {
            if (tok.size() != 2)
                std::cerr << "Unexpected split of line \"" << line
                          << "\" into more than 2 tokens, but continuing anyways.\n";
            if (kv.count(tok[0]))
                std::cerr << "Already have seen key \"" << tok[0] << "\", but overriding old seen value \""
                          << kv[tok[0]] << "\" with new value \"" << tok[1] << "\".";
            kv[tok[0]] = tok[1];
        }
    }
    return kv;
}

using Setters = std::map<std::string, std::function<void(const std::string &)>>;

static void applyPairs(const std::map<std::string, std::string> &kv, const Setters &known) {
    for (auto &p : kv) {
        auto it = known.find(p.first);
        if (it != known.end()) it->second(p.second);
        else std::cerr << "Unknown key \"" << p.first << "\" (with value \"" << p.second << "\") provided. Ignoring.\n";
    }
}

static Setters propSetters(MPMObject &o, const std::string &prefix = "", bool long_names = true) {
    const std::string p = prefix;
    Setters s;
    s[p + (long_names ? "properties" : "props")] = [&o](const std::string &v) { o.fp64_props = parsePropertyList<double>(v); };
    s[p + (long_names ? "int-properties" : "int-props")] = [&o](const std::string &v) { o.int_props = parsePropertyList<int>(v); };
    s[p + (long_names ? "str-properties" : "str-props")] = [&o](const std::string &v) {
        o.str_props = parsePropertyList<std::string>(v);
    };
    return s;
}

template <class T>
static void copyProps(T &dst, const MPMObject &src) {
    dst.fp64_props = src.fp64_props;
    dst.int_props = src.int_props;
    dst.str_props = src.str_props;
}

template <class T>
static void simpleSection(std::ifstream &fin, const Registry<T> &reg, std::unique_ptr<T> &slot, const char *what) {
    auto kv = sectionPairs(fin);
    MPMObject tmp;
    Setters s = propSetters(tmp);
    s["class"] = [&](const std::string &v) { slot = reg.get_object(v); };
    applyPairs(kv, s);
    if (!slot) throw std::runtime_error(std::string(what) + " was nullptr -- was the class set in the config file?");
    copyProps(*slot, tmp);
}

static std::string resolvePath(const std::string &dir, const std::string &f) {
    struct stat st;
    if (f.empty() || f[0] == '/' || stat(f.c_str(), &st) == 0) return f;
    return dir + f;  // relative to the configuration file
}

int Configurator::configureJob(Job *job) {
    std::ifstream fin(file);
    if (!fin.is_open()) {
        std::cout << "ERROR: Unable to open file: " << file << std::endl;
        return 0;
    }
    std::string line;
    while (std::getline(fin, line)) {
        line = parser::removeSpaces(parser::removeComments(line));
        if (line.empty()) continue;
        if (line == "job") {
            auto kv = sectionPairs(fin);
            applyPairs(kv, Setters{
                          {"dt", [&](const std::string &v) { job->dt = std::stod(v); }},
                          {"t", [&](const std::string &v) { job->t = job->t0 = std::stod(v); }},
                          {"TYPE", [&](const std::string &v) { job->assignJobType(std::stoi(v)); }},
                          {"threads", [&](const std::string &v) { job->thread_count = std::stoi(v); }},
                      });
        } else if (line == "serializer") {
            simpleSection(fin, serializer_registry, job->serializer, "Serializer");
        } else if (line == "driver") {
            simpleSection(fin, driver_registry, job->driver, "Driver");
        } else if (line == "solver") {
            simpleSection(fin, solver_registry, job->solver, "Solver");
        } else if (line == "grid") {
            simpleSection(fin, grid_registry, job->grid, "Grid");
            job->grid->GRID_DIM = job->JOB_TYPE == Job::JOB_3D ? 3 : (job->JOB_TYPE == Job::JOB_1D ? 1 : 2);
        } else if (line == "body") {
            auto kv = sectionPairs(fin);
            job->bodies.emplace_back(nullptr);
            job->activeBodies.push_back(0);
            const size_t id = job->bodies.size() - 1;
            MPMObject tmp, pt, nd, mt, bd;
            std::string name, pfile;
            Setters s = propSetters(tmp);
            s["class"] = [&](const std::string &v) {
                job->bodies[id] = body_registry.get_object(v);
                job->bodies[id]->id = (int)id;
            };
            s["name"] = [&](const std::string &v) { name = v; };
            s["point-file"] = [&](const std::string &v) { pfile = v; };
            s["point-class"] = [&](const std::string &v) { pt.object_name = v; };
            s["node-class"] = [&](const std::string &v) { nd.object_name = v; };
            s["material-class"] = [&](const std::string &v) { mt.object_name = v; };
            s["boundary-class"] = [&](const std::string &v) { bd.object_name = v; };
            using Pre = std::pair<MPMObject *, const char *>;
            for (const Pre &o : {Pre(&pt, "point-"), Pre(&nd, "node-"), Pre(&mt, "material-"), Pre(&bd, "boundary-")}) {
                Setters ps = propSetters(*o.first, o.second, false);
                s.insert(ps.begin(), ps.end());
            }
            applyPairs(kv, s);
            if (!job->bodies[id]) throw std::runtime_error("Body was nullptr -- was the class set in the config file?");
            Body &B = *job->bodies[id];
            B.name = name;
            copyProps(B, tmp);
            if (pt.object_name.empty()) {
                std::cerr << "No Points!" << std::endl;
            } else {
                B.points = points_registry.get_object(pt.object_name);
                copyProps(*B.points, pt);
                B.points->readFromFile(job, &B, resolvePath(directory, pfile));
                job->activeBodies[id] = 1;
            }
            if (nd.object_name.empty()) std::cerr << "No Nodes!" << std::endl;
            else {
                B.nodes = nodes_registry.get_object(nd.object_name);
                copyProps(*B.nodes, nd);
            }
            if (mt.object_name.empty()) std::cerr << "No Material!" << std::endl;
            else {
                B.material = material_registry.get_object(mt.object_name);
                copyProps(*B.material, mt);
                B.activeMaterial = 1;
            }
            if (bd.object_name.empty()) std::cerr << "No Boundary!" << std::endl;
            else {
                B.boundary = boundary_registry.get_object(bd.object_name);
                copyProps(*B.boundary, bd);
                B.activeBoundary = 1;
            }
        } else if (line == "contact") {
            auto kv = sectionPairs(fin);
            job->contacts.emplace_back(nullptr);
            job->activeContacts.push_back(0);
            const size_t id = job->contacts.size() - 1;
            MPMObject tmp;
            std::string name;
            Setters s = propSetters(tmp);
            s["class"] = [&](const std::string &v) {
                job->contacts[id] = contact_registry.get_object(v);
                job->contacts[id]->id = (int)id;
            };
            s["name"] = [&](const std::string &v) { name = v; };
            applyPairs(kv, s);
            if (!job->contacts[id])
                throw std::runtime_error("Contacts[" + std::to_string(id) + "] was nullptr -- was the class set in the config file?");
            job->activeContacts[id] = 1;
            job->contacts[id]->name = name;
            copyProps(*job->contacts[id], tmp);
        } else {
            std::cerr << "Unknown header \"" << line << "\", ignoring it.\n";
        }
    }
    return 1;
}

// ------------------------------------------------------------------ job
void Job::assignJobType(int type) {
    JOB_TYPE = type;
    if (type == JOB_1D) DIM = 1;
    else if (type == JOB_2D) DIM = 2;
    else if (type == JOB_3D) DIM = 3;
    else if (type == JOB_2D_OOP || type == JOB_AXISYM) DIM = 3;
    else std::cerr << "Job doesn't have defined type for input " << type << "." << std::endl;
}

void Job::init() {
    if (!serializer || !driver || !solver || !grid)
        throw std::runtime_error("job needs serializer, driver, solver and grid sections");
    serializer->init(this);
    driver->init(this);
    solver->init(this);
    grid->init(this);
    for (auto &b : bodies) {
        if (!b) continue;
        if (b->points) b->points->init(this, b.get());
        if (b->nodes) b->nodes->init(this, b.get());
        if (b->activeMaterial) b->material->init(this);
        if (b->activeBoundary) b->boundary->init(this, b.get());
    }
    for (auto &c : contacts)
        if (c) c->init(this);
}

// ------------------------------------------------------------------ points (default_points.cpp)
void Points::readFromFile(Job *job, Body *, const std::string &fileIN) {
    file = fileIN;
    std::ifstream fin(file);
    if (!fin.is_open()) {
        std::cerr << "ERROR: Unable to open file: " << file << std::endl;
        return;
    }
    std::string line;
    std::getline(fin, line);
    const int len = std::stoi(line);
    x.assign(3 * len, 0.0);
    u.assign(3 * len, 0.0);
    x_t.assign(3 * len, 0.0);
    mx_t.assign(3 * len, 0.0);
    b.assign(3 * len, 0.0);
    T.assign(9 * len, 0.0);
    L.assign(9 * len, 0.0);
    m.assign(len, 0.0);
    v.assign(len, 0.0);
    v0.assign(len, 0.0);
    extent.assign(len, 0.0);
    active.assign(len, 0);
    const int gd = job->grid ? job->grid->GRID_DIM : job->DIM;
    for (int i = 0; i < len; i++) {
        std::getline(fin, line);
        auto s = parser::splitString(line, ' ');
        int d;
        if ((int)s.size() == 3 + 2 * job->DIM) d = job->DIM;
        else if ((int)s.size() == 3 + 2 * gd) d = gd;
        else {
            std::cerr << "ERROR: Unable to read line: " << file << std::endl;
            return;
        }
        m[i] = std::stod(s[0]);
        v[i] = std::stod(s[1]);
        for (int k = 0; k < d; k++) x[3 * i + k] = std::stod(s[2 + k]);
        for (int k = 0; k < d; k++) x_t[3 * i + k] = std::stod(s[2 + d + k]);
        active[i] = (int)std::stod(s[2 + 2 * d]);
    }
}

void Points::init(Job *job, Body *) {
    // DefaultPoints::init (default_points.cpp:35-57): v0 = v, extent fixed for the run
    const size_t n = size();
    for (size_t i = 0; i < n; i++) {
        v0[i] = v[i];
        if (job->DIM == 1) extent[i] = 0.5 * v[i];
        else if (job->DIM == 2) extent[i] = 0.5 * std::sqrt(v[i]);
        else extent[i] = 0.5 * std::cbrt(v[i]);
        for (int k = 0; k < 3; k++) mx_t[3 * i + k] = m[i] * x_t[3 * i + k];
    }
}

void Points::writeToFile(Job *job, const std::string &f) const {
    std::ofstream out(f, std::ios::trunc);
    out << size() << "\n" << std::setprecision(17);
    for (size_t i = 0; i < size(); i++) {
        out << m[i] << " " << v[i];
        for (int k = 0; k < job->DIM; k++) out << " " << x[3 * i + k];
        for (int k = 0; k < job->DIM; k++) out << " " << x_t[3 * i + k];
        out << " " << active[i] << "\n";
    }
}

// ------------------------------------------------------------------ plugins
void IsotropicLinearElasticity::init(Job *) {
    if (fp64_props.size() < 2) {
        fprintf(stderr, "%s:%s: Need at least 2 properties defined ({E, nu}).\n", __FILE__, __func__);
        exit(0);
    }
    E = fp64_props[0];
    nu = fp64_props[1];
    G = E / (2.0 * (1.0 + nu));
    K = E / (3.0 * (1.0 - 2 * nu));
    lambda = K - 2.0 * G / 3.0;
    printf("Material properties (E = %g, nu = %g, G = %g, K = %g).\n", E, nu, G, K);
}

// Sand_SachithLocal::init (sand_sachithlocal.cpp:58-117), including its messages and its
// exit(EXIT_FAILURE) on fewer than six properties
void Sand_SachithLocal::init(Job *) {
    if (fp64_props.size() < 6) {
        fprintf(stderr,
                "%s:%s: Need at least 6 properties defined (E, nu, grain diameter, mu_s, grains_rho, rho_critical).\n",
                __FILE__, __func__);
        exit(EXIT_FAILURE);
    }
    E = fp64_props[0];
    nu = fp64_props[1];
    grains_d = fp64_props[2];
    G = E / (2.0 * (1.0 + nu));
    K = E / (3.0 * (1.0 - 2 * nu));
    lambda = K - 2.0 * G / 3.0;
    MU_S = fp64_props[3];
    GRAINS_RHO = fp64_props[4];
    RHO_CRITICAL = fp64_props[5];
    if (fp64_props.size() < 8) {
        MU_2 = MU_S;
        DELTA_MU = MU_2 - MU_S;
        I_0 = 1;
        printf("Material properties (E = %g, nu = %g, G = %g, K = %g, grain diameter = %g, mu_s = %g, grains_rho = %g, "
               "rho_critical = %g).\n",
               E, nu, G, K, grains_d, MU_S, GRAINS_RHO, RHO_CRITICAL);
    } else {
        MU_2 = fp64_props[6];
        DELTA_MU = MU_2 - MU_S;
        I_0 = fp64_props[7];
        printf("Material properties (E = %g, nu = %g, G = %g, K = %g, grain diameter = %g, mu_s = %g, grains_rho = %g, "
               "rho_critical = %g, mu_2 = %g, i_0 = %g).\n",
               E, nu, G, K, grains_d, MU_S, GRAINS_RHO, RHO_CRITICAL, MU_2, I_0);
    }
    for (const std::string &sp : str_props)
        if (sp == "USE_DENSITY_FOR_PRESSURE") {
            use_density_to_determine_pressure = true;
            std::cout << "Using density to determine pressure." << std::endl;
        }
}

void CartesianLinear::init(Job *job) {
    if ((int)fp64_props.size() < GRID_DIM || (int)int_props.size() < GRID_DIM) {
        fprintf(stderr, "%s:%s: Need at least %i dimensions defined.\n", __FILE__, __func__, GRID_DIM);
        exit(0);
    }
    (void)job;
    for (int d = 0; d < GRID_DIM; d++) {
        L[d] = fp64_props[d];
        N[d] = int_props[d];
        hx[d] = L[d] / N[d];
    }
    v_n.clear();
}

// CartesianFrictionalBox::init (cartesian_frictional_box.cpp:26-37): mu_f is fp64-properties[0]
void CartesianFrictionalBox::init(Job *, Body *) {
    if (fp64_props.empty()) {
        std::cout << fp64_props.size() << "\n";
        fprintf(stderr, "%s:%s: Need at least 1 property defined (mu_f).\n", __FILE__, __func__);
        exit(0);
    }
    mu_f = fp64_props[0];
    std::cout << "Boundary properties ( mu_f = " << mu_f << ")." << std::endl;
}

void ContactHuang::init(Job *job) {
    auto need = [&]() {
        fprintf(stderr, "%s:%s: Need at least 3 properties defined ({mu_f},{body_1,body_2}).\n", __FILE__, __func__);
        exit(0);
    };
    if (fp64_props.empty() || (str_props.size() < 2 && int_props.size() < 2)) need();
    mu_f = fp64_props[0];
    if (str_props.size() == 2)
        for (int i = 0; i < 2; i++)
            for (size_t b = 0; b < job->bodies.size(); b++)
                if (job->bodies[b] && str_props[i] == job->bodies[b]->name) {
                    bodyIDs[i] = (int)b;
                    break;
                }
    for (int i = 0; i < 2; i++)
        if (bodyIDs[i] < 0) {
            if (int_props.size() == 2) {
                bodyIDs[0] = int_props[0];
                bodyIDs[1] = int_props[1];
            } else {
                need();
            }
            break;
        }
    printf("Contact properties (mu_f = %g, {%i, %i}).\n", mu_f, bodyIDs[0], bodyIDs[1]);
}

// ------------------------------------------------------------------ drivers (default_driver.cpp)
void Driver::init(Job *job) {
    if (fp64_props.empty()) {
        fprintf(stderr, "%s:%s: Need at least 1 property defined ({stop_time}).\n", __FILE__, __func__);
        exit(0);
    }
    stop_time = fp64_props[0];
    generateGravity(job);
}

void Driver::generateGravity(Job *job) {
    gravity[0] = gravity[1] = gravity[2] = 0.0;
    if (job->JOB_TYPE == Job::JOB_1D) gravity[0] = -9.81;
    else if (job->JOB_TYPE == Job::JOB_3D) gravity[2] = -9.81;
    else gravity[1] = -9.81;
}

void UserDefinedGravityDriver::init(Job *job) {
    if ((int)fp64_props.size() < 1 + job->DIM) {
        fprintf(stderr, "%s:%s: Need at least %i properties defined ({stop_time, <gravity vector>}).\n", __FILE__,
                __func__, job->DIM);
        exit(0);
    }
    stop_time = fp64_props[0];
    for (int d = 0; d < 3; d++) gravity[d] = d < job->DIM ? fp64_props[1 + d] : 0.0;
}

void Driver::applyGravity(Job *job) {
    for (auto &b : job->bodies) {
        if (!b || !b->points) continue;
        for (size_t i = 0; i < b->points->size(); i++)
            for (int k = 0; k < 3; k++) b->points->b[3 * i + k] = gravity[k];
    }
}

void Driver::run(Job *job) {
    timespec t0, tf, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    tf = t0;
    int frames = 0;
    applyGravity(job);
    while (job->t <= stop_time) {
        job->solver->step(job);
        if (job->serializer->writeFrame(job) == 1) {
            clock_gettime(CLOCK_MONOTONIC, &t1);
            const double dfr = (t1.tv_sec - tf.tv_sec) + (t1.tv_nsec - tf.tv_nsec) * 1e-9;
            const double dsim = (t1.tv_sec - t0.tv_sec) + (t1.tv_nsec - t0.tv_nsec) * 1e-9;
            tf = t1;
            std::cout << "Frame Written [" << ++frames << "]. Time/Frame [" << dfr << " s]. Elapsed Time [" << dsim
                      << " s]." << std::endl;
        }
        job->t += job->dt;
    }
    job->solver->syncHost(job);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    std::cout << "Simulation Complete. Elapsed Time [" << (t1.tv_sec - t0.tv_sec) + (t1.tv_nsec - t0.tv_nsec) * 1e-9
              << "s]." << std::endl;
}

}  // namespace mpmhost
