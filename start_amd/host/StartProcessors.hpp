// C++ host side of the MI355X particle step: the two static entry points of
// Start's step loop, with the reference's names, argument order and error
// behaviour, forwarding to the C ABI of libstart_gpu (include/start_gpu.h).
//
//   ParticleDynamicsProcessor::process   ParticleDynamicsProcessor.hpp:154-170
//   NodeChargeDensityProcessor::gather   NodeChargeDensityProcessor.hpp:39-46
//
// The reference passes the mesh through three objects (CubicGrid, GSMAssembler,
// SurfaceMesh) whose CGAL/Trilinos types are not part of this library; here one
// ParticleKernelDevice takes their place.  It owns the device context, so the
// particles stay resident in HBM between calls; the ParticleVector is updated
// through setters after each step (never copied: copies of the reference's
// Particle get new ids, Particle.cpp:113-155).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <set>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "start_gpu.h"

namespace start_amd {

// constants::particle_types::ParticleType (Constants.hpp:78-93)
enum ParticleType { O2, Ar, Ne, He, Ti, Al, Sn, W, Au, Cu, Ni, Ag, Unknown };

// Mass, radius, charge and the gas parameters of a type (Constants.hpp:20-73,
// ParticlePropertiesManager.hpp:73-217).  omega and alpha are the reference's
// float constants.
struct TypeProperties {
    double mass;      // kg
    double radius;    // m
    double charge;    // C
    double omega;     // VHS/VSS viscosity index (gases)
    double alpha;     // VSS deflection parameter (gases)
};
TypeProperties typeProperties(ParticleType t);
ParticleType typeFromName(std::string_view name);  // "Ar", "Ti", ... ; Unknown otherwise

// The fields of Particle (Particle.hpp:25-30) the step reads and writes.
class Particle {
public:
    Particle(size_t id, ParticleType type, double x, double y, double z, double vx, double vy, double vz);
    size_t getId() const { return m_id; }
    ParticleType getType() const { return m_type; }
    std::array<double, 3> getCentre() const { return m_centre; }
    std::array<double, 3> getVelocity() const { return m_velocity; }
    double getEnergy_J() const { return m_energy; }
    double getMass() const;
    double getCharge() const;
    void setCentre(double x, double y, double z) { m_centre = {x, y, z}; }
    void setVelocity(double vx, double vy, double vz);  // energy = m |v|^2 / 2 (ParticleDynamicUtils.hpp:136-140)

private:
    size_t m_id;
    ParticleType m_type;
    std::array<double, 3> m_centre, m_velocity;
    double m_energy;
};

// PICTypes.hpp:18, :31, :41, :65
using ParticleVector = std::vector<Particle>;
using ParticlesIDSet = std::set<size_t>;
using NodeChargeDensitiesMap = std::map<size_t, double>;
using ParticleTrackerMap = std::map<double, std::map<size_t, std::vector<size_t>>>;
using ParticleMovementMap = std::map<size_t, std::vector<std::array<double, 3>>>;

// StopSubject / StopFlagObserver (StopModelingObserver.cpp:3-18)
class StopSubject {
public:
    void notifyStopRequested() { m_stop = true; }
    bool stopRequested() const { return m_stop; }

private:
    bool m_stop = false;
};

struct DeviceKernelOptions {
    int precision = START_PREC_FP64;   // START_PREC_FP64 | START_PREC_MIXED
    bool emPush = true;                // false: the reference's literal loop (no push, Q1)
    int densityMode = START_DENSITY_BARYCENTRIC;
    int sortInterval = 20;             // steps between sorts of the particles by tet
    uint64_t seed = 1;
    uint32_t run = 0;
    double bboxExpansion = 1.01;       // Constants.hpp:55
    // Update the ParticleVector after every process().  Off by default: the
    // reference's loop never reads particle state between steps (gather and
    // process read the device) nor after it (ModelingMainDriver.cpp:79-91
    // writes triangle counts and movements); syncParticles() reads it on demand.
    bool writeBackEachStep = false;
    // particleMovementMap of process(): the first movementMaxParticles moving
    // particles are recorded (ParticleMovementTracker.hpp:43); 0 records none
    size_t movementMaxParticles = 200000;
    // Collision diameters r_p + r_t by particle type; types not listed use the
    // reference's radii sum (HSModel.cpp:8-12, VHSModel.cpp:14).
    std::map<ParticleType, double> diameterOverride;
    // Charges by particle type, in C; types not listed use the reference's
    // charge of the type (ParticlePropertiesManager.hpp:191-217).
    std::map<ParticleType, double> chargeOverride;
};

// Options from the run's JSON configuration file (the file ConfigParser reads,
// ConfigParser.cpp:18-226), applied over `options`.  Optional keys the
// reference does not know and ignores:
//   "Device Precision": "FP64" | "Mixed"
//   "Device EM Push": true | false           (false: the reference's literal loop, Q1)
//   "Device Seed": integer                    (variate stream key)
//   "Device Density Mode": "Barycentric" | "TetAverage"
//   "Device Sort Interval": integer           (steps between sorts by tet; 0 never)
//   "Device Write Back Each Step": true | false
//   "Device Movement Max Particles": integer
//   "Device Charges": {"Ti": 2, ...}          (elementary charges by particle type)
//   "Device Diameters": {"Ti": 2.1e-10, ...}  (collision diameters r_p + r_t in m)
// The reference's own keys are skipped; an unknown "Device ..." key, a value of
// the wrong kind or a malformed file throws std::runtime_error.
DeviceKernelOptions deviceOptionsFromConfig(const std::string& path, DeviceKernelOptions options = {});

// The mesh as the step needs it: the tetrahedra of VolumetricMesh and the
// surface triangles of SurfaceMesh, with their Gmsh tags (GmshUtils.cpp:116-136).
struct DeviceMeshInput {
    std::vector<double> nodes;      // 3 per node
    std::vector<int32_t> tets;      // 4 node rows per tet
    std::vector<int32_t> triangles; // 3 node rows per surface triangle
    std::vector<size_t> nodeTags;   // tag of each node row (empty: row + 1)
    std::vector<size_t> triangleTags; // tag of each triangle row (empty: row)
    std::vector<size_t> tetTags;    // tag of each tet row (empty: rows ascend with the tags)
    // Tets and triangles are reordered by tag on construction: the device
    // breaks owner and hit ties by the smaller row, the reference by the
    // smaller element tag (NodeChargeDensityProcessor.cpp:51-56).
};

class ParticleKernelDevice {
public:
    ParticleKernelDevice(DeviceMeshInput mesh, DeviceKernelOptions options = {});
    ~ParticleKernelDevice();
    ParticleKernelDevice(const ParticleKernelDevice&) = delete;
    ParticleKernelDevice& operator=(const ParticleKernelDevice&) = delete;

    // Field side (ChargeDensityEquationSolver.cpp:17-46 hands over potentials
    // by node tag, MatrixEquationSolver.cpp:53-64).  Takes effect at once if
    // the device context exists, else when process() creates it.
    void setNodePotential(const std::map<size_t, double>& potentialByTag);
    void setTetField(const std::vector<double>& E);  // 3 per tet

    // TriangleCell::count of every surface triangle, by triangle tag.
    std::map<size_t, uint64_t> triangleCounts();
    // Copies the device particle state into the ParticleVector (setters).
    void syncParticles(ParticleVector& particles);
    start_ctx* context() { return m_ctx; }
    int64_t stepsTaken() const { return m_steps; }

private:
    friend class ParticleDynamicsProcessor;
    friend class NodeChargeDensityProcessor;
    // The context is created at the first process() / gather() call with one
    // species per ParticleType; the gas and time step follow each process().
    void ensureContext();
    // Reads the "Device ..." keys of the configuration file once, before the
    // context exists (process() and gather() pass their config_filename).
    void useConfig(std::string_view path);
    void configure(std::string_view scatteringModel, std::string_view gasName, double gasConcentration,
                   double timeStep);
    // Loads the ParticleVector at the first call; afterwards appends the
    // particles inserted at its end (continuous sputtering spawns every step)
    // and throws on any other change of size.
    void ensureParticles(const ParticleVector& particles);
    void check(int rc, const char* what);
    void step(ParticleVector& particles, ParticlesIDSet& settled, StopSubject& stop);
    void densities(NodeChargeDensitiesMap& out);
    // movement records: started before the first step of a population, then
    // each step's points appended to the caller's map
    void startMovements();
    void appendMovements(ParticleMovementMap& map, std::mutex& mutex);

    DeviceMeshInput m_mesh;
    DeviceKernelOptions m_opt;
    std::vector<size_t> m_rowsByTag;     // node rows in ascending tag order
    std::vector<uint64_t> m_settledIds;  // ids settled in the last step (reused)
    start_ctx* m_ctx = nullptr;
    bool m_configured = false;
    bool m_config_read = false;
    double m_dt = 0.0;
    std::string m_model, m_gas;
    double m_conc = 0.0;
    std::vector<double> m_phi_rows, m_E;  // field given before the context existed
    size_t m_n = 0;
    bool m_loaded = false;  // the device holds the population (appends extend it)
    uint64_t m_settled_seen = 0;
    int64_t m_steps = 0;
    bool m_charge_fresh = false;  // device node charge belongs to the current positions
    bool m_tracking = false;      // movement records on for the loaded population
};

class ParticleDynamicsProcessor {
public:
    // One step of every particle (ParticleDynamicsProcessor.cpp:213-277).  As
    // in the reference it never throws: failures are logged to stderr and the
    // call returns (:202-209, :268-276).  particleMovementMap receives the
    // step's movement records (the start of the step and any hit point) of the
    // tracked particles, under mutex_particlesMovementMap.  numThreads,
    // particleTracker and the other mutex are accepted for the signature; the
    // device keeps the owners itself and runs race-free.
    static void process(std::string_view config_filename, ParticleVector& particles, double timeMoment,
                        double timeStep, unsigned int numThreads, ParticleKernelDevice& device,
                        ParticleTrackerMap& particleTracker, ParticlesIDSet& settledParticlesIds,
                        std::shared_mutex& sh_mutex_settledParticlesCounterMap,
                        std::mutex& mutex_particlesMovementMap, ParticleMovementMap& particleMovementMap,
                        StopSubject& stopSubject, std::string_view scatteringModel, std::string_view gasName,
                        double gasConcentration);
};

class NodeChargeDensityException : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class NodeChargeDensityProcessor {
public:
    // Node charge densities of the current particle positions into
    // nodeChargeDensityMap, keyed by node tag (NodeChargeDensityProcessor.cpp:169-328).
    // Throws NodeChargeDensityException on failure, as the reference throws
    // (:103-112, :318-327).
    static void gather(double timeMoment, std::string_view configFilename, ParticleKernelDevice& device,
                       ParticleVector const& particles, ParticlesIDSet& settledParticleIDs,
                       ParticleTrackerMap& particleTrackerMap, NodeChargeDensitiesMap& nodeChargeDensityMap);
};

}  // namespace start_amd
