// Host side of the step loop's two entry points over libstart_gpu's C ABI.
#include "StartProcessors.hpp"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <iostream>
#include <iterator>
#include <utility>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace start_amd {

namespace {
// threads for the ordered map passes: the OpenMP default (the host's cores or
// OMP_NUM_THREADS), at most 16
int mapThreads()
{
#ifdef _OPENMP
    return std::max(1, std::min(16, omp_get_max_threads()));
#else
    return 1;
#endif
}
}  // namespace

namespace {

constexpr double kECharge = 1.602176565e-19;  // Constants.hpp:14

double f32(double v) { return (double)(float)v; }

struct TypeRow {
    const char* name;
    TypeProperties p;
};

// Constants.hpp:20-73 (masses, radii, VTI omega and VSS alpha as float
// constants); charges ParticlePropertiesManager.hpp:191-217.
const TypeRow kTypes[] = {
    {"O2", {53.1e-27, 2.035e-10, 0.0, f32(0.77), f32(1.40)}},
    {"Ar", {6.6335209e-26, 71e-12, 0.0, f32(0.81), f32(1.40)}},
    {"Ne", {3.3509177e-26, 160e-12, 0.0, f32(0.66), f32(1.31)}},
    {"He", {6.6464731e-27, 120e-12, 0.0, f32(0.66), f32(1.26)}},
    {"Ti", {7.9485017e-26, 140e-12, 2 * kECharge, 0.0, 0.0}},
    {"Al", {4.4803831e-26, 125e-12, 3 * kECharge, 0.0, 0.0}},
    {"Sn", {1.9712258e-25, 145e-12, 2 * kECharge, 0.0, 0.0}},
    {"W", {8.4590343e-26, 135e-12, 6 * kECharge, 0.0, 0.0}},
    {"Au", {3.2707137e-25, 135e-12, 3 * kECharge, 0.0, 0.0}},
    {"Cu", {1.0552061e-25, 135e-12, 1 * kECharge, 0.0, 0.0}},
    {"Ni", {9.7462675e-26, 135e-12, 2 * kECharge, 0.0, 0.0}},
    {"Ag", {1.7911901e-25, 160e-12, 1 * kECharge, 0.0, 0.0}},
};
constexpr int kNumTypes = (int)Unknown;
static_assert(sizeof(kTypes) / sizeof(kTypes[0]) == (size_t)Unknown, "one row per particle type");

std::string upper(std::string_view s)
{
    std::string u(s);
    for (auto& ch : u) ch = (char)std::toupper((unsigned char)ch);
    return u;
}

// CollisionModelFactory::create dispatches on the upper-cased name (CollisionModelFactory.cpp:23-36).
int modelCode(std::string_view name)
{
    std::string u = upper(name);
    if (u == "HS") return START_MODEL_HS;
    if (u == "VHS") return START_MODEL_VHS;
    if (u == "VSS") return START_MODEL_VSS;
    throw std::runtime_error("unknown scattering model '" + std::string(name) + "'");
}

}  // namespace

TypeProperties typeProperties(ParticleType t)
{
    if (t < 0 || t >= Unknown) throw std::runtime_error("unknown particle type");
    return kTypes[t].p;
}

ParticleType typeFromName(std::string_view name)
{
    for (int t = 0; t < kNumTypes; ++t)
        if (name == kTypes[t].name) return (ParticleType)t;
    return Unknown;
}

Particle::Particle(size_t id, ParticleType type, double x, double y, double z, double vx, double vy, double vz)
    : m_id(id), m_type(type), m_centre{x, y, z}, m_velocity{0.0, 0.0, 0.0}, m_energy(0.0)
{
    setVelocity(vx, vy, vz);
}

double Particle::getMass() const { return typeProperties(m_type).mass; }
double Particle::getCharge() const { return typeProperties(m_type).charge; }

void Particle::setVelocity(double vx, double vy, double vz)
{
    m_velocity = {vx, vy, vz};
    m_energy = 0.5 * getMass() * (vx * vx + vy * vy + vz * vz);
}

// ---------------------------------------------------------------------------

ParticleKernelDevice::ParticleKernelDevice(DeviceMeshInput mesh, DeviceKernelOptions options)
    : m_mesh(std::move(mesh)), m_opt(std::move(options))
{
    if (m_mesh.nodes.size() % 3 || m_mesh.tets.size() % 4 || m_mesh.triangles.size() % 3)
        throw std::invalid_argument("mesh arrays are not whole rows");
    size_t M = m_mesh.nodes.size() / 3, S = m_mesh.triangles.size() / 3;
    if (m_mesh.nodeTags.empty())
        for (size_t n = 0; n < M; ++n) m_mesh.nodeTags.push_back(n + 1);  // Gmsh tags 1..M (Q19)
    if (m_mesh.triangleTags.empty())
        for (size_t s = 0; s < S; ++s) m_mesh.triangleTags.push_back(s);
    if (m_mesh.nodeTags.size() != M || m_mesh.triangleTags.size() != S ||
        (!m_mesh.tetTags.empty() && m_mesh.tetTags.size() != m_mesh.tets.size() / 4))
        throw std::invalid_argument("tag arrays do not match the mesh");
    // rows in ascending tag order (ties of owner and hit go to the smaller tag)
    auto by_tag = [](std::vector<size_t>& tag, std::vector<int32_t>& conn, size_t per) {
        std::vector<size_t> ord(tag.size());
        for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return tag[a] < tag[b]; });
        std::vector<size_t> t2(tag.size());
        std::vector<int32_t> c2(conn.size());
        for (size_t k = 0; k < ord.size(); ++k) {
            t2[k] = tag[ord[k]];
            if (k && t2[k] == t2[k - 1]) throw std::invalid_argument("repeated element tag " + std::to_string(t2[k]));
            for (size_t q = 0; q < per; ++q) c2[per * k + q] = conn[per * ord[k] + q];
        }
        tag.swap(t2);
        conn.swap(c2);
    };
    if (!m_mesh.tetTags.empty()) by_tag(m_mesh.tetTags, m_mesh.tets, 4);
    by_tag(m_mesh.triangleTags, m_mesh.triangles, 3);
    m_rowsByTag.resize(M);
    for (size_t n = 0; n < M; ++n) m_rowsByTag[n] = n;
    std::sort(m_rowsByTag.begin(), m_rowsByTag.end(),
              [&](size_t a, size_t b) { return m_mesh.nodeTags[a] < m_mesh.nodeTags[b]; });
}

void ParticleKernelDevice::useConfig(std::string_view path)
{
    if (m_config_read || m_ctx || path.empty()) return;
    m_opt = deviceOptionsFromConfig(std::string(path), m_opt);
    m_config_read = true;
}

ParticleKernelDevice::~ParticleKernelDevice()
{
    if (m_ctx) start_destroy(m_ctx);
}

void ParticleKernelDevice::check(int rc, const char* what)
{
    if (rc != START_OK) throw std::runtime_error(std::string(what) + ": " + start_last_error(m_ctx));
}

void ParticleKernelDevice::ensureContext()
{
    if (m_ctx) return;
    std::vector<start_species_desc> sp((size_t)kNumTypes);
    for (int t = 0; t < kNumTypes; ++t) {
        TypeProperties p = kTypes[t].p;
        auto q = m_opt.chargeOverride.find((ParticleType)t);
        sp[(size_t)t] = start_species_desc{p.mass, q != m_opt.chargeOverride.end() ? q->second : p.charge, 0.0};
    }
    start_mesh_desc md{};
    md.n_nodes = (int64_t)(m_mesh.nodes.size() / 3);
    md.nodes = m_mesh.nodes.data();
    md.n_tets = (int64_t)(m_mesh.tets.size() / 4);
    md.tets = m_mesh.tets.data();
    md.n_tris = (int64_t)(m_mesh.triangles.size() / 3);
    md.tris = m_mesh.triangles.data();
    start_params prm{};
    prm.dt = m_dt;
    prm.bbox_expansion = m_opt.bboxExpansion;
    prm.seed = m_opt.seed;
    prm.run = m_opt.run;
    prm.em_push = m_opt.emPush ? 1 : 0;
    prm.precision = m_opt.precision;
    prm.charge_every_step = 1;
    prm.n_species = kNumTypes;
    prm.steps_per_launch = 1;
    prm.species = sp.data();
    prm.gas = start_gas_desc{0.0, 0.0, 0.0, 0.0, START_MODEL_NONE, 0};
    prm.sort_interval = m_opt.sortInterval;
    start_ctx* c = nullptr;
    int rc = start_create(&c, &md, &prm);
    if (rc != START_OK) throw std::runtime_error(std::string("device set-up: ") + start_last_error(nullptr));
    m_ctx = c;
    if (!m_E.empty()) check(start_set_tet_field(m_ctx, m_E.data()), "tet field");
    if (!m_phi_rows.empty()) check(start_set_node_potential(m_ctx, m_phi_rows.data()), "node potential");
}

void ParticleKernelDevice::configure(std::string_view scatteringModel, std::string_view gasName,
                                     double gasConcentration, double timeStep)
{
    if (m_configured && m_model == scatteringModel && m_gas == gasName && m_conc == gasConcentration &&
        m_dt == timeStep)
        return;
    int code = modelCode(scatteringModel);
    ParticleType g = typeFromName(gasName);
    if (g == Unknown) throw std::runtime_error("unknown gas '" + std::string(gasName) + "'");
    TypeProperties gp = typeProperties(g);
    start_gas_desc gd{gp.mass, gp.omega, gp.alpha, gasConcentration, code, 0};
    std::vector<double> diam((size_t)kNumTypes);
    for (int t = 0; t < kNumTypes; ++t) {
        auto d = m_opt.diameterOverride.find((ParticleType)t);
        diam[(size_t)t] = d != m_opt.diameterOverride.end() ? d->second : kTypes[t].p.radius + gp.radius;
    }
    check(start_set_gas(m_ctx, &gd, diam.data()), "gas");
    check(start_set_time_step(m_ctx, timeStep), "time step");
    m_model = std::string(scatteringModel);
    m_gas = std::string(gasName);
    m_conc = gasConcentration;
    m_dt = timeStep;
    m_configured = true;
}

void ParticleKernelDevice::ensureParticles(const ParticleVector& particles)
{
    const size_t n = particles.size();
    if (m_loaded && n == m_n) return;
    // The device keeps every particle it was given, settled ones included, and
    // the reference only ever inserts new particles at the end of the vector
    // (ModelingMainDriver.cpp:64-67): any other change of size is an error.
    if (m_loaded && n < m_n)
        throw std::runtime_error("the particle vector shrank from " + std::to_string(m_n) + " to " +
                                 std::to_string(n) + " particles; only appending at the end is supported");
    const size_t first = m_loaded ? m_n : 0, m = n - first;
    std::vector<double> x(m), y(m), z(m), vx(m), vy(m), vz(m);
    std::vector<int32_t> sp(m);
    std::vector<uint64_t> ids(m);
    for (size_t i = 0; i < m; ++i) {
        const Particle& p = particles[first + i];
        if (p.getType() == Unknown) throw std::runtime_error("particle of unknown type");
        auto c = p.getCentre();
        auto v = p.getVelocity();
        x[i] = c[0];
        y[i] = c[1];
        z[i] = c[2];
        vx[i] = v[0];
        vy[i] = v[1];
        vz[i] = v[2];
        sp[i] = (int32_t)p.getType();
        ids[i] = (uint64_t)p.getId();
    }
    if (!m_loaded) {
        check(start_load_particles(m_ctx, (int64_t)m, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(),
                                   sp.data(), ids.data()),
              "particle upload");
        m_settled_seen = 0;
        m_tracking = false;  // a new population ends the device's records
    } else {
        // spawned particles join the population; everything loaded is kept
        check(start_append_particles(m_ctx, (int64_t)m, x.data(), y.data(), z.data(), vx.data(), vy.data(),
                                     vz.data(), sp.data(), ids.data()),
              "particle append");
    }
    m_n = n;
    m_loaded = true;
    m_charge_fresh = false;  // the node charge does not hold the new particles yet
}

void ParticleKernelDevice::startMovements()
{
    if (m_tracking || m_opt.movementMaxParticles == 0 || m_n == 0) return;
    // at most a start and a hit point per step, read after every step
    check(start_track_movements(m_ctx, (int64_t)m_opt.movementMaxParticles, 2), "movement tracking");
    m_tracking = true;
}

void ParticleKernelDevice::appendMovements(ParticleMovementMap& map, std::mutex& mutex)
{
    if (!m_tracking) return;
    int64_t n = 0;
    check(start_read_movements(m_ctx, &n, nullptr, nullptr, nullptr), "movement count");
    if (n == 0) return;
    std::vector<uint64_t> ids((size_t)n);
    std::vector<int32_t> counts((size_t)n);
    std::vector<double> xyz((size_t)n * 2 * 3);
    check(start_read_movements(m_ctx, &n, ids.data(), counts.data(), xyz.data()), "movements");
    check(start_clear_movements(m_ctx), "movement reset");
    std::lock_guard<std::mutex> lock(mutex);  // ParticleMovementTracker.cpp:10
    // After a particle's first movement its key is in the map.  When every
    // particle with points this step already has its key, the points are
    // appended on several threads, each over a contiguous range of slots:
    // first a read-only pass finds the entries (slots come in load order, so
    // for ids loaded in ascending order a forward walk from the range's first
    // key finds each in O(1); otherwise a lookup), then each thread appends to
    // its own particles' vectors.  Any missing key sends the step to the
    // serial loop below, which inserts.
    if (n >= 16384 && !map.empty()) {
        // 8 threads: the appends allocate, and 16 measured slower (7.1 against
        // 4.7-5.0 ms per step for 200,000 tracked particles)
        const int T = std::min(8, mapThreads());
        std::vector<ParticleMovementMap::iterator> where((size_t)n, map.end());
        std::atomic<bool> missing{false};
#pragma omp parallel for num_threads(T) schedule(static)
        for (int c = 0; c < T; ++c) {
            const int64_t s0 = n * c / T, s1 = n * (c + 1) / T;
            auto it = map.lower_bound((size_t)ids[(size_t)s0]);
            for (int64_t q = s0; q < s1 && !missing.load(std::memory_order_relaxed); ++q) {
                if (counts[(size_t)q] == 0) continue;
                const size_t id = (size_t)ids[(size_t)q];
                while (it != map.end() && it->first < id) ++it;
                if (it == map.end() || it->first != id) it = map.find(id);
                if (it == map.end()) {
                    missing = true;
                    break;
                }
                where[(size_t)q] = it;
            }
        }
        if (!missing) {
#pragma omp parallel for num_threads(T) schedule(static)
            for (int c = 0; c < T; ++c) {
                const int64_t s0 = n * c / T, s1 = n * (c + 1) / T;
                for (int64_t q = s0; q < s1; ++q) {
                    const int32_t k = std::min<int32_t>(counts[(size_t)q], 2);
                    for (int32_t j = 0; j < k; ++j) {
                        const double* p = &xyz[((size_t)q * 2 + (size_t)j) * 3];
                        where[(size_t)q]->second.push_back({p[0], p[1], p[2]});
                    }
                }
            }
            return;
        }
    }
    // Slots come in load order, which is ascending id order for a population
    // loaded with ascending ids: the position after the previous particle's
    // entry is then the right hint, and each lookup costs O(1) on average
    // (any other order still finds or inserts the right entry).
    auto hint = map.begin();
    for (int64_t s = 0; s < n; ++s) {
        const int32_t k = std::min<int32_t>(counts[(size_t)s], 2);
        if (k == 0) continue;
        auto it = map.try_emplace(hint, (size_t)ids[(size_t)s]);
        auto& pts = it->second;
        for (int32_t j = 0; j < k; ++j) {
            const double* q = &xyz[((size_t)s * 2 + (size_t)j) * 3];
            pts.push_back({q[0], q[1], q[2]});
        }
        hint = std::next(it);
    }
}

void ParticleKernelDevice::step(ParticleVector& particles, ParticlesIDSet& settled, StopSubject& stop)
{
    check(start_step(m_ctx, 1), "step");
    ++m_steps;
    m_charge_fresh = true;  // the step scatters the end-of-step charge
    // the statistics and the ids settled since the last step, in one device sync
    start_stats s{};
    int64_t k = (int64_t)m_n - (int64_t)m_settled_seen;  // at most the particles not yet settled
    if (m_settledIds.size() < (size_t)std::max<int64_t>(k, 1)) m_settledIds.resize((size_t)std::max<int64_t>(k, 1));
    check(start_read_stats_settled(m_ctx, &s, (int64_t)m_settled_seen, m_settledIds.data(), &k),
          "step statistics and settled ids");
    if (k > 0) {
        settled.insert(m_settledIds.begin(), m_settledIds.begin() + k);
        m_settled_seen = s.settled;
    }
    if (s.stop) stop.notifyStopRequested();  // ParticleSettler.cpp:24-25
    if (m_opt.writeBackEachStep) syncParticles(particles);
}

void ParticleKernelDevice::syncParticles(ParticleVector& particles)
{
    if (!m_ctx || m_n == 0) return;
    if (particles.size() != m_n) throw std::runtime_error("particle vector size differs from the device's");
    std::vector<double> x(m_n), y(m_n), z(m_n), vx(m_n), vy(m_n), vz(m_n);
    check(start_read_particles(m_ctx, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(), nullptr,
                               nullptr, nullptr),
          "particle read-back");
    for (size_t i = 0; i < m_n; ++i) {
        particles[i].setCentre(x[i], y[i], z[i]);
        particles[i].setVelocity(vx[i], vy[i], vz[i]);
    }
}

void ParticleKernelDevice::densities(NodeChargeDensitiesMap& out)
{
    if (m_opt.densityMode == START_DENSITY_BARYCENTRIC && !m_charge_fresh) {
        check(start_gather_charge(m_ctx), "charge scatter");
        m_charge_fresh = true;
    }
    // barycentric: every node is written, so the written flags are not read.
    // The values are read in place from the library's page-locked buffers.
    const bool all = m_opt.densityMode == START_DENSITY_BARYCENTRIC;
    const double* rho = nullptr;
    const int32_t* written = nullptr;
    check(start_node_charge_density_view(m_ctx, m_opt.densityMode, &rho, all ? nullptr : &written),
          "node charge density");
    // The reference's main loop keeps one map for the whole run and never
    // clears it (ModelingMainDriver.cpp:122), so after the first step every
    // key is already there and only its value changes.  Then the ordered pass runs in
    // chunks on several threads: each finds its first key, walks its part of
    // the map and writes values only (no insertion, so the tree is only read);
    // a written node without a key sends the whole pass to the serial loop
    // below, which inserts.  Keys of other nodes are kept.
    const size_t M = m_rowsByTag.size();
    if (out.size() >= M / 2 && M >= 16384) {
        const int T = mapThreads();
        std::atomic<bool> missing{false};
#pragma omp parallel for num_threads(T) schedule(static)
        for (int c = 0; c < T; ++c) {
            const size_t k0 = M * (size_t)c / T, k1 = M * (size_t)(c + 1) / T;
            auto it = out.lower_bound(m_mesh.nodeTags[m_rowsByTag[k0]]);
            for (size_t k = k0; k < k1 && !missing.load(std::memory_order_relaxed); ++k) {
                const size_t r = m_rowsByTag[k];
                if (!all && !written[r]) continue;
                const size_t tag = m_mesh.nodeTags[r];
                while (it != out.end() && it->first < tag) ++it;
                if (it == out.end() || it->first != tag) {
                    missing = true;
                    break;
                }
                it->second = rho[r];
                ++it;
            }
        }
        if (!missing) return;
    }
    auto it = out.begin();
    for (size_t r : m_rowsByTag) {
        if (!all && !written[r]) continue;
        const size_t tag = m_mesh.nodeTags[r];
        while (it != out.end() && it->first < tag) ++it;
        if (it != out.end() && it->first == tag) {
            it->second = rho[r];
            ++it;
        } else {
            it = std::next(out.emplace_hint(it, tag, rho[r]));
        }
    }
}

void ParticleKernelDevice::setNodePotential(const std::map<size_t, double>& potentialByTag)
{
    size_t M = m_mesh.nodes.size() / 3;
    std::map<size_t, size_t> row;
    for (size_t n = 0; n < M; ++n) row[m_mesh.nodeTags[n]] = n;
    std::vector<double> phi(M, 0.0);
    for (const auto& [tag, v] : potentialByTag) {
        auto r = row.find(tag);
        if (r == row.end()) throw std::invalid_argument("potential for an unknown node tag");
        phi[r->second] = v;
    }
    m_E.clear();
    if (m_ctx) check(start_set_node_potential(m_ctx, phi.data()), "node potential");
    else m_phi_rows = std::move(phi);
}

void ParticleKernelDevice::setTetField(const std::vector<double>& E)
{
    if (E.size() != m_mesh.tets.size() / 4 * 3) throw std::invalid_argument("field must hold 3 values per tet");
    m_phi_rows.clear();
    if (m_ctx) check(start_set_tet_field(m_ctx, E.data()), "tet field");
    else m_E = E;
}

std::map<size_t, uint64_t> ParticleKernelDevice::triangleCounts()
{
    std::map<size_t, uint64_t> out;
    size_t S = m_mesh.triangles.size() / 3;
    if (!m_ctx) {
        for (size_t s = 0; s < S; ++s) out[m_mesh.triangleTags[s]] = 0;
        return out;
    }
    std::vector<uint64_t> c(std::max<size_t>(S, 1));
    check(start_read_tri_counts(m_ctx, c.data()), "triangle counts");
    for (size_t s = 0; s < S; ++s) out[m_mesh.triangleTags[s]] = c[s];
    return out;
}

// ---------------------------------------------------------------------------

void ParticleDynamicsProcessor::process(std::string_view config_filename, ParticleVector& particles,
                                        double /*timeMoment*/, double timeStep, unsigned int /*numThreads*/,
                                        ParticleKernelDevice& device, ParticleTrackerMap& /*particleTracker*/,
                                        ParticlesIDSet& settledParticlesIds,
                                        std::shared_mutex& /*sh_mutex_settledParticlesCounterMap*/,
                                        std::mutex& mutex_particlesMovementMap,
                                        ParticleMovementMap& particleMovementMap, StopSubject& stopSubject,
                                        std::string_view scatteringModel, std::string_view gasName,
                                        double gasConcentration)
{
    try {
        device.useConfig(config_filename);
        device.ensureContext();
        device.configure(scatteringModel, gasName, gasConcentration, timeStep);
        device.ensureParticles(particles);
        device.startMovements();
        device.step(particles, settledParticlesIds, stopSubject);
        device.appendMovements(particleMovementMap, mutex_particlesMovementMap);
    } catch (std::exception const& ex) {
        std::cerr << "ERROR: particle step on the GPU failed: " << ex.what() << std::endl;
    } catch (...) {
        std::cerr << "ERROR: particle step on the GPU failed" << std::endl;
    }
}

void NodeChargeDensityProcessor::gather(double /*timeMoment*/, std::string_view configFilename,
                                        ParticleKernelDevice& device, ParticleVector const& particles,
                                        ParticlesIDSet& /*settledParticleIDs*/,
                                        ParticleTrackerMap& /*particleTrackerMap*/,
                                        NodeChargeDensitiesMap& nodeChargeDensityMap)
{
    try {
        device.useConfig(configFilename);
        device.ensureContext();
        device.ensureParticles(particles);
        device.densities(nodeChargeDensityMap);
    } catch (std::exception const& ex) {
        throw NodeChargeDensityException(std::string("gathering node charge densities failed: ") + ex.what());
    }
}

}  // namespace start_amd
