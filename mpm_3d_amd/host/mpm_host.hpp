// mpm_host.hpp -- the host side of the HIP explicit MPM solver: the reference's object
// model for this path (Job, Body, Points, Nodes, Material, Boundary, Contact, Grid, Driver,
// Serializer, Solver with their fp64/int/str property lists), its configuration grammar,
// its registry and the HipExplicitUSL solver that runs the step through the C-ABI in
// include/mpm_gpu.h.
//
// Interfaces kept like the reference's (same names, argument meaning, error behaviour):
//   MPMObject property lists        mpm_objects.hpp:33-41
//   Solver::init/step/saveState/loadState   mpm_objects.hpp:97-104
//   Registry<T>::get_object (unknown class -> std::invalid_argument)  registry.hpp:24-35
//   Configurator::configureJob and the section/key grammar           config.cpp:176-600
//   DefaultPoints::readFromFile (.points files)                      default_points.cpp:104-193
//   DefaultDriver::run (while t <= stop_time)                        default_driver.cpp:46-86
//   bad properties: message on stderr, exit(0)                       e.g. isolin.cpp:26-32
// Point and node arrays are AoS with 3 doubles per vector and 9 per tensor, as the
// reference's KinematicVectorArray / MaterialTensorArray store them.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct mpm_gpu_handle;

namespace mpmhost {

struct Job;
struct Body;
struct Serializer;

struct MPMObject {
    std::string object_name;
    std::vector<double> fp64_props;
    std::vector<int> int_props;
    std::vector<std::string> str_props;
    virtual ~MPMObject() = default;
};

// DefaultPoints / ThreadPoolPoints state (points.hpp:60-84)
struct Points : MPMObject {
    std::string file;
    std::vector<double> x, u, x_t, mx_t, b;  // n x 3
    std::vector<double> L, T;                // n x 9, row-major
    std::vector<double> m, v, v0, extent;
    std::vector<int> active;
    size_t size() const { return m.size(); }
    virtual void readFromFile(Job *job, Body *body, const std::string &file);
    virtual void init(Job *job, Body *body);
    // writes m v x x_t active with 17 significant digits
    void writeToFile(Job *job, const std::string &file) const;
};

// DefaultNodes state (mpm_objects.hpp:164-167): positions h * ijk and all nodes active
// (default_nodes.cpp:29-63); the other fields are those of the last step, filled by the
// solver when a frame needs them
struct Nodes : MPMObject {
    std::vector<double> x, u, x_t, diff_x_t, mx_t, f;  // n x 3
    std::vector<double> m;
    std::vector<int> active;
    size_t size() const { return m.size(); }
    virtual void init(Job *job, Body *body);
};

struct Material : MPMObject {
    virtual void init(Job *job) = 0;
};

struct IsotropicLinearElasticity : Material {
    double E = 0, nu = 0, G = 0, K = 0, lambda = 0;
    void init(Job *job) override;
};

// Sand_SachithLocal (materials.hpp:85-120): {E, nu, grain diameter, mu_s, grains_rho,
// rho_critical[, mu_2, I_0]}, str-properties USE_DENSITY_FOR_PRESSURE; gammap and
// gammap_dot are filled by the solver when a frame needs them
struct Sand_SachithLocal : Material {
    double E = 0, nu = 0, G = 0, K = 0, lambda = 0, grains_d = 0;
    double MU_S = 0, GRAINS_RHO = 0, RHO_CRITICAL = 0, MU_2 = 0, DELTA_MU = 0, I_0 = 1;
    bool use_density_to_determine_pressure = false;
    std::vector<double> gammap, gammap_dot;
    void init(Job *job) override;
};

struct Boundary : MPMObject {
    virtual void init(Job *, Body *) {}
};
struct CartesianBox : Boundary {};        // every component fixed on the box faces
struct CartesianSmoothBox : Boundary {};  // the face-normal component fixed (free slip)
struct CartesianFrictionalBox : Boundary { // face-normal fixed, Coulomb friction mu_f
    double mu_f = 0;
    void init(Job *job, Body *body) override;
};

struct Body : MPMObject {
    int id = 0;
    std::string name;
    std::unique_ptr<Points> points;
    std::unique_ptr<Nodes> nodes;
    std::unique_ptr<Material> material;
    std::unique_ptr<Boundary> boundary;
    int activeMaterial = 0, activeBoundary = 0;
};

struct Contact : MPMObject {
    int id = 0;
    std::string name;
    virtual void init(Job *job) = 0;
};

struct ContactHuang : Contact {
    double mu_f = 0;
    int bodyIDs[2] = {-1, -1};
    std::vector<double> contact_normal, contact_force;  // n x 3 per node, last step (contacts.hpp:56)
    void init(Job *job) override;
};

struct Grid : MPMObject {
    int GRID_DIM = 0;
    virtual void init(Job *job) = 0;
};

struct CartesianLinear : Grid {
    double L[3] = {0, 0, 0};
    int N[3] = {0, 0, 0};
    double hx[3] = {0, 0, 0};  // L / N (cartesian.cpp:24-67)
    std::vector<double> v_n;   // nodal volume, summed over elements as hiddenInit does
    void init(Job *job) override;
    long long node_count() const;
    long long element_count() const;
    int nodeID(const int *ijk) const;               // i + (Nx+1)(j + (Ny+1) k)
    void elementNodes(long long e, int *ids) const;  // the 2^GRID_DIM node ids of element e
};

struct Solver : MPMObject {
    virtual void init(Job *job) = 0;
    virtual void step(Job *job) = 0;
    virtual std::string saveState(Job *job, Serializer *serializer, std::string filepath) = 0;
    virtual int loadState(Job *job, Serializer *serializer, std::string fullpath) = 0;
    // bring the host Points arrays up to date with the solver's device state; nodes: also
    // the Nodes and ContactHuang arrays of the last step
    virtual void syncHost(Job *, bool nodes = false) { (void)nodes; }
};

struct Driver : MPMObject {
    double stop_time = 0;
    double gravity[3] = {0, 0, 0};
    virtual void init(Job *job);
    virtual void run(Job *job);
    virtual void generateGravity(Job *job);
    void applyGravity(Job *job);
};

struct UserDefinedGravityDriver : Driver {
    void init(Job *job) override;
    void generateGravity(Job *) override {}
};

// Frame writers. The base class is DefaultVTK (default_vtk.cpp): at the sample rate
// (fp64_props[0] frames/s) it writes, per body, a legacy-VTK point file fpd.<id>.<name>.<frame>.vtk
// and a node file fnd.<...>.vtk into str_props[0], with the fields of DefaultPoints,
// DefaultNodes, the boundary and the contacts (their writeFrame methods). MinimalVTK
// (minimal_vtk.cpp) writes point velocity, displacement, pressure, tau_bar, gamma_dot and
// density and node velocity and density into the same files. "NullSerializer" writes
// nothing. writeFrame returns 1 when a frame was written.
struct Serializer : MPMObject {
    double sampleRate = 60, t_last_frame = 0;
    int sampledFrames = 0;
    std::string frameDirectory, outputDirectory, simulationName;
    virtual void init(Job *job);
    virtual int writeFrame(Job *job);
    virtual bool writesNodes() const { return true; }  // frames need the solver's node arrays
};
struct MinimalVTK : Serializer {
    int writeFrame(Job *job) override;
};
struct NullSerializer : Serializer {
    void init(Job *) override {}
    int writeFrame(Job *) override { return 0; }
    bool writesNodes() const override { return false; }
};

// The solver over the C-ABI. int-properties {cpdi_spec, contact_spec[, precision[, sort_interval]]}
// with the reference's ExplicitUSL meaning for the first two (explicit_usl.cpp:25-44);
// precision 0 = FP64 (default), 1 = FP32 state with FP64 accumulation.
struct HipExplicitUSL : Solver {
    int cpdi_spec = 1, contact_spec = 1, precision = 0, sort_interval = 16;
    mpm_gpu_handle *handle = nullptr;
    long long pending = 0;  // steps issued since the host arrays were refreshed
    bool host_stale = false;
    bool nodes_stale = false;  // the Nodes / contact arrays predate the last step
    ~HipExplicitUSL() override;
    void init(Job *job) override;
    void step(Job *job) override;
    void syncHost(Job *job, bool nodes = false) override;
    std::string saveState(Job *job, Serializer *serializer, std::string filepath) override;
    int loadState(Job *job, Serializer *serializer, std::string fullpath) override;

   private:
    void flush(Job *job, bool record_last = false);
    void syncNodes(Job *job);
    void upload(Job *job);
};

// ParallelExplicitUSL / ThreadPoolExplicitUSL configurations: the same step on the GPU; their
// int-properties {cpdi_spec, contact_spec, num_threads[, debug]} (parallel_explicit_usl.cpp:
// 117-153) keep their meaning, the CPU thread count has nothing left to do
struct HipThreadedAlias : HipExplicitUSL {
    int num_threads = 1;
    void init(Job *job) override;
};

struct Job {
    enum { JOB_1D = 1, JOB_2D = 2, JOB_3D = 3, JOB_2D_OOP = 4, JOB_AXISYM = 5 };
    double dt = 1e-3, t = 0, t0 = 0;
    int JOB_TYPE = 0, DIM = 0, thread_count = 1;
    std::vector<std::unique_ptr<Body>> bodies;
    std::vector<int> activeBodies;
    std::vector<std::unique_ptr<Contact>> contacts;
    std::vector<int> activeContacts;
    std::unique_ptr<Grid> grid;
    std::unique_ptr<Solver> solver;
    std::unique_ptr<Driver> driver;
    std::unique_ptr<Serializer> serializer;
    void assignJobType(int type);
    // Job::init order: grid, bodies (points, nodes, material, boundary), contacts,
    // serializer, driver, solver (job.cpp)
    void init();
};

template <class T> class Registry {
   public:
    using Factory = std::function<std::unique_ptr<T>()>;
    Registry();
    std::unique_ptr<T> get_object(const std::string &name) const {
        auto it = table.find(name);
        if (it == table.end()) throw std::invalid_argument("No class named '" + name + "' in registry.");
        std::unique_ptr<T> p = it->second();
        p->object_name = name;
        return p;
    }
    std::map<std::string, Factory> table;
};

struct Configurator {
    std::string file;
    Registry<Serializer> serializer_registry;
    Registry<Driver> driver_registry;
    Registry<Solver> solver_registry;
    Registry<Grid> grid_registry;
    Registry<Body> body_registry;
    Registry<Points> points_registry;
    Registry<Nodes> nodes_registry;
    Registry<Contact> contact_registry;
    Registry<Material> material_registry;
    Registry<Boundary> boundary_registry;
    std::string directory;  // the config file's folder; point files are read relative to it
    explicit Configurator(std::string fileIN);
    // returns 1 on success, 0 if the file cannot be opened (config.cpp:176-206)
    int configureJob(Job *job);
};

// parser helpers with the reference's semantics (parser.hpp:18-108)
namespace parser {
std::string removeSpaces(std::string s);
std::vector<std::string> splitString(const std::string &s, char delim);
std::string removeComments(const std::string &s);
std::string removeBraces(const std::string &s);
std::string removeQuotes(const std::string &s);
}  // namespace parser

}  // namespace mpmhost
