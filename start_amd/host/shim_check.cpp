// Drives ParticleDynamicsProcessor::process and NodeChargeDensityProcessor::gather
// the way Start's step loop does (ModelingMainDriver.cpp:132-220: gather, then
// process, once per time moment, until the stop flag), on inputs read from a
// directory, and writes what the loop leaves behind.
//
//   shim_check <dir>
// Inputs in <dir> (raw little-endian arrays):
//   nodes.f64 (3M)  tets.i32 (4T)  tris.i32 (3S)  phi.f64 (M)
//   pos.f64 (3N)  vel.f64 (3N)  type.i32 (N)  id.u64 (N)
//   config.txt    one "key value" per line: dt, steps, model, gas, pressure,
//                 temperature, seed, precision (fp64|mixed), density
//                 (barycentric|tet_average), em_push (0|1), diameter_<Type>,
//                 charge_<Type> (in units of e); optional spawn_every m and
//                 spawn_count k: at the start of every m-th step after the
//                 first, k particles of spawn_pos.f64 (3K) spawn_vel.f64 (3K)
//                 spawn_type.i32 (K) spawn_id.u64 (K), in order, are inserted
//                 at the end of the ParticleVector (continuous sputtering,
//                 ModelingMainDriver.cpp:37-71, :141-142); optional shrink_at s:
//                 the last particle is removed before step s (an error);
//                 optional movements N (particles tracked in the movement map,
//                 200,000 by default as ParticleMovementTracker.hpp:43; 0: none),
//                 timing_warmup W (steps before the timed window)
//   config.json   optional: the run's JSON configuration, passed to gather and
//                 process as config_filename (its "Device ..." keys apply)
// Outputs in <dir>:
//   out_pos.f64 out_vel.f64  particle state after the loop (ParticleVector order)
//   out_settled.u64           settled ids, ascending
//   out_tri_counts.u64        TriangleCell::count by triangle tag (row order)
//   out_density.f64 out_density_written.i32   the last gather's map, by node row
//   out_summary.txt           steps run, stop flag
//   out_timing.txt            wall seconds of the loop after timing_warmup
//                             steps, of its gather and process calls, and
//                             the particle-steps processed in that window
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "StartProcessors.hpp"

using namespace start_amd;

template <typename T>
static std::vector<T> read_array(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    f.seekg(0, std::ios::end);
    size_t bytes = (size_t)f.tellg();
    f.seekg(0);
    std::vector<T> v(bytes / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    return v;
}

template <typename T>
static void write_array(const std::string& path, const std::vector<T>& v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    if (!f) throw std::runtime_error("cannot write " + path);
}

// --print-config <file.json>: the device options deviceOptionsFromConfig reads
// from a run configuration, one "name value" line each (no GPU needed)
int print_config(const char* path)
{
    try {
        const DeviceKernelOptions o = deviceOptionsFromConfig(path);
        std::cout.precision(17);
        std::cout << "precision " << (o.precision == START_PREC_MIXED ? "mixed" : "fp64") << "\n"
                  << "em_push " << (int)o.emPush << "\n"
                  << "seed " << o.seed << "\n"
                  << "density " << (o.densityMode == START_DENSITY_TET_AVERAGE ? "tet_average" : "barycentric") << "\n"
                  << "sort_interval " << o.sortInterval << "\n"
                  << "write_back " << (int)o.writeBackEachStep << "\n"
                  << "movement_max_particles " << o.movementMaxParticles << "\n";
        for (const auto& [t, q] : o.chargeOverride) std::cout << "charge " << (int)t << " " << q << "\n";
        for (const auto& [t, d] : o.diameterOverride) std::cout << "diameter " << (int)t << " " << d << "\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "ERROR: " << e.what() << std::endl;
        return 1;
    }
}

int main(int argc, char** argv)
{
    if (argc == 3 && std::string(argv[1]) == "--print-config") return print_config(argv[2]);
    if (argc != 2) {
        std::cerr << "usage: shim_check <dir> | shim_check --print-config <file.json>" << std::endl;
        return 2;
    }
    std::string dir = argv[1];
    // the run's JSON configuration handed to gather and process as the
    // reference's loop does (its "Device ..." keys apply over config.txt);
    // none when <dir>/config.json is absent
    const std::string run_config = std::ifstream(dir + "/config.json").good() ? dir + "/config.json" : "";
    try {
        std::map<std::string, std::string> cfg;
        {
            std::ifstream f(dir + "/config.txt");
            std::string k, v;
            while (f >> k >> v) cfg[k] = v;
        }
        auto num = [&](const std::string& k) { return std::stod(cfg.at(k)); };
        DeviceMeshInput mesh;
        mesh.nodes = read_array<double>(dir + "/nodes.f64");
        mesh.tets = read_array<int32_t>(dir + "/tets.i32");
        mesh.triangles = read_array<int32_t>(dir + "/tris.i32");
        DeviceKernelOptions opt;
        opt.seed = (uint64_t)std::stoull(cfg.at("seed"));
        opt.precision = cfg.at("precision") == "mixed" ? START_PREC_MIXED : START_PREC_FP64;
        opt.densityMode = cfg.at("density") == "tet_average" ? START_DENSITY_TET_AVERAGE : START_DENSITY_BARYCENTRIC;
        opt.emPush = cfg.count("em_push") ? std::stoi(cfg.at("em_push")) != 0 : true;
        if (cfg.count("movements")) opt.movementMaxParticles = (size_t)std::stoull(cfg.at("movements"));
        for (const auto& [k, v] : cfg) {
            if (k.rfind("diameter_", 0) == 0) opt.diameterOverride[typeFromName(k.substr(9))] = std::stod(v);
            if (k.rfind("charge_", 0) == 0) opt.chargeOverride[typeFromName(k.substr(7))] = std::stod(v) * 1.602176565e-19;
        }
        size_t M = mesh.nodes.size() / 3;
        ParticleKernelDevice device(mesh, opt);
        std::vector<double> phi = read_array<double>(dir + "/phi.f64");
        std::map<size_t, double> phiByTag;
        for (size_t n = 0; n < M; ++n) phiByTag[n + 1] = phi[n];
        device.setNodePotential(phiByTag);

        std::vector<double> pos = read_array<double>(dir + "/pos.f64"), vel = read_array<double>(dir + "/vel.f64");
        std::vector<int32_t> type = read_array<int32_t>(dir + "/type.i32");
        std::vector<uint64_t> id = read_array<uint64_t>(dir + "/id.u64");
        ParticleVector particles;
        particles.reserve(type.size());
        for (size_t i = 0; i < type.size(); ++i)
            particles.emplace_back((size_t)id[i], (ParticleType)type[i], pos[3 * i], pos[3 * i + 1], pos[3 * i + 2],
                                   vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);

        const int spawn_every = cfg.count("spawn_every") ? std::stoi(cfg.at("spawn_every")) : 0;
        const size_t spawn_count = cfg.count("spawn_count") ? (size_t)std::stoul(cfg.at("spawn_count")) : 0;
        const int shrink_at = cfg.count("shrink_at") ? std::stoi(cfg.at("shrink_at")) : -1;
        std::vector<double> spos, svel;
        std::vector<int32_t> stype;
        std::vector<uint64_t> sid;
        if (spawn_every > 0) {
            spos = read_array<double>(dir + "/spawn_pos.f64");
            svel = read_array<double>(dir + "/spawn_vel.f64");
            stype = read_array<int32_t>(dir + "/spawn_type.i32");
            sid = read_array<uint64_t>(dir + "/spawn_id.u64");
        }
        size_t spawned = 0;

        const double dt = num("dt");
        const int steps = std::stoi(cfg.at("steps"));
        const double conc = start_gas_concentration(num("pressure"), num("temperature"));
        ParticlesIDSet settled;
        ParticleTrackerMap tracker;
        ParticleMovementMap movements;
        NodeChargeDensitiesMap density;
        StopSubject stop;
        std::shared_mutex sh;
        std::mutex mx;
        const int warm = cfg.count("timing_warmup") ? std::stoi(cfg.at("timing_warmup")) : 0;
        using clk = std::chrono::steady_clock;
        double t_gather = 0.0, t_process = 0.0;
        clk::time_point t_begin = clk::now();
        uint64_t processed0 = 0;
        auto processed_now = [&]() {
            start_stats st{};
            if (!device.context() || start_read_stats(device.context(), &st) != START_OK) return (uint64_t)0;
            return st.processed;
        };
        int k = 0;
        double t = 0.0;
        for (; k < steps && !stop.stopRequested(); ++k, t += dt) {
            if (k == warm) {
                processed0 = processed_now();
                t_begin = clk::now();
                t_gather = t_process = 0.0;
            }
            if (spawn_every > 0 && k > 0 && k % spawn_every == 0)
                for (size_t j = 0; j < spawn_count && spawned < sid.size(); ++j, ++spawned)
                    particles.emplace_back((size_t)sid[spawned], (ParticleType)stype[spawned], spos[3 * spawned],
                                           spos[3 * spawned + 1], spos[3 * spawned + 2], svel[3 * spawned],
                                           svel[3 * spawned + 1], svel[3 * spawned + 2]);
            if (k == shrink_at) particles.pop_back();
            const clk::time_point a = clk::now();
            NodeChargeDensityProcessor::gather(t, run_config, device, particles, settled, tracker, density);
            const clk::time_point b = clk::now();
            ParticleDynamicsProcessor::process(run_config, particles, t, dt, 16, device, tracker, settled,
                                               sh, mx, movements, stop, cfg.at("model"), cfg.at("gas"), conc);
            const clk::time_point e = clk::now();
            t_gather += std::chrono::duration<double>(b - a).count();
            t_process += std::chrono::duration<double>(e - b).count();
        }
        const double t_loop = std::chrono::duration<double>(clk::now() - t_begin).count();
        const uint64_t processed1 = processed_now();
        std::ofstream(dir + "/out_timing.txt") << "timed_steps " << std::max(0, k - warm) << "\nloop_s " << t_loop
                                               << "\ngather_s " << t_gather << "\nprocess_s " << t_process
                                               << "\nparticle_steps " << (processed1 - processed0) << "\n";
        density.clear();
        NodeChargeDensityProcessor::gather(t, run_config, device, particles, settled, tracker, density);
        device.syncParticles(particles);

        std::vector<double> op, ov;
        for (const auto& p : particles) {
            auto c = p.getCentre();
            auto v = p.getVelocity();
            op.insert(op.end(), c.begin(), c.end());
            ov.insert(ov.end(), v.begin(), v.end());
        }
        write_array(dir + "/out_pos.f64", op);
        write_array(dir + "/out_vel.f64", ov);
        write_array(dir + "/out_settled.u64", std::vector<uint64_t>(settled.begin(), settled.end()));
        std::vector<uint64_t> mv_ids;
        std::vector<int32_t> mv_counts;
        std::vector<double> mv_xyz;
        for (const auto& [id, pts] : movements) {
            mv_ids.push_back(id);
            mv_counts.push_back((int32_t)pts.size());
            for (const auto& p : pts) mv_xyz.insert(mv_xyz.end(), p.begin(), p.end());
        }
        write_array(dir + "/out_mv_ids.u64", mv_ids);
        write_array(dir + "/out_mv_counts.i32", mv_counts);
        write_array(dir + "/out_mv_xyz.f64", mv_xyz);
        std::vector<uint64_t> counts;
        for (const auto& [tag, c] : device.triangleCounts()) counts.push_back(c);
        write_array(dir + "/out_tri_counts.u64", counts);
        std::vector<double> rho(M, 0.0);
        std::vector<int32_t> wr(M, 0);
        for (const auto& [tag, v] : density) {
            rho[tag - 1] = v;
            wr[tag - 1] = 1;
        }
        write_array(dir + "/out_density.f64", rho);
        write_array(dir + "/out_density_written.i32", wr);
        std::ofstream(dir + "/out_summary.txt") << "steps " << k << "\nstop " << (stop.stopRequested() ? 1 : 0)
                                                << "\nparticles " << particles.size() << "\n";
        std::cout << "shim_check: " << k << " steps, " << settled.size() << " settled" << std::endl;
    } catch (std::exception const& ex) {
        std::cerr << "shim_check: " << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
