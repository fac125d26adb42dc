// headless.cpp - command-line driver of the step: the reference's application loop
// (PCISPH/src/main.cpp:22-258) without a window.  OpenGL rendering and PNG recording
// are replaced by particle-frame dumps, which double as checkpoints.
//
//   pcisph_headless -i <scene> [-o] [-d ms] [-n steps] [--out-dir dir] [--final file]
//                   [--resume frame] [--precision fast|ieee|fp64] [--max-iterations k]
//                   [--emit-scene file]
//
// The reference's arguments keep their meaning (main.cpp:36-54):
//   -i <scene>  the scene: dam_break_16k | dam_break_1m | xraw:<dir>:<dambreak|cube_splash|simple_drop>
//   -o          record: a frame whenever at least 1/60 s of simulated time has passed since
//               the last one (the first after the first step), <out-dir>/frame_<k>.pcsf for
//               k = 0, 1, ... (main.cpp:193-222, where frames are out/frame_<k>.png)
//   -d <ms>     stop once the simulated time reaches ms / 1000 s (main.cpp:43-45,224)
// Additions: -n caps the step count (the scene's own step count when neither -n nor -d
// is given: there is no window to close); --out-dir (default "out"); --final writes the
// last state as a frame; --resume continues from a frame (its step count and simulated
// time); --emit-scene writes the generated scene without stepping it.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "../../include/pcisph.h"
#include "scene_gen.h"

namespace {

// Frame file: "PCSF", u32 version, u32 fluid, u32 boundary, u64 step, then
// float[3N] positions, float[3N] velocities, float[N] densities, float[3Nb] boundary.
constexpr uint32_t kVersion = 1;

void check(pcisph_solver* s, int rc, const char* what) {
    if (rc) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, pcisph_last_error(s));
        std::exit(1);
    }
}

std::vector<float> download(pcisph_solver* s, pcisph_field f) {
    size_t n = 0;
    check(s, pcisph_field_size(s, f, &n), "pcisph_field_size");
    std::vector<float> v(n);
    check(s, pcisph_download(s, f, v.data(), n), "pcisph_download");
    return v;
}

void write_frame(pcisph_solver* s, const std::string& path, uint64_t step) {
    Simulation_Params p;
    pcisph_get_params(s, &p);
    auto x = download(s, PCISPH_FLUID_POSITIONS), v = download(s, PCISPH_FLUID_VELOCITIES),
         rho = download(s, PCISPH_FLUID_DENSITIES), b = download(s, PCISPH_BOUNDARY_POSITIONS);
    std::ofstream f(path, std::ios::binary);
    f.write("PCSF", 4);
    f.write((const char*)&kVersion, 4);
    f.write((const char*)&p.fluid_count, 4);
    f.write((const char*)&p.boundary_count, 4);
    f.write((const char*)&step, 8);
    for (auto* a : {&x, &v, &rho, &b}) f.write((const char*)a->data(), a->size() * sizeof(float));
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(1);
    }
}

bool read_frame(const std::string& path, scene_gen::Scene& sc, std::vector<float>& vel, uint64_t& step) {
    std::ifstream f(path, std::ios::binary);
    char magic[4];
    uint32_t ver, n, nb;
    if (!f.read(magic, 4) || std::memcmp(magic, "PCSF", 4) || !f.read((char*)&ver, 4) || ver != kVersion)
        return false;
    f.read((char*)&n, 4);
    f.read((char*)&nb, 4);
    f.read((char*)&step, 8);
    sc.fluid.resize(3ull * n);
    vel.resize(3ull * n);
    std::vector<float> rho(n);
    sc.boundary.resize(3ull * nb);
    f.read((char*)sc.fluid.data(), sc.fluid.size() * 4);
    f.read((char*)vel.data(), vel.size() * 4);
    f.read((char*)rho.data(), rho.size() * 4);
    f.read((char*)sc.boundary.data(), sc.boundary.size() * 4);
    return (bool)f;
}

}  // namespace

int main(int argc, char** argv) {
    std::string scene_name, outdir = "out", resume, emit, precision = "fast", final_path;
    uint32_t steps = 0, max_it = 0;
    bool recording = false;
    float duration = std::numeric_limits<float>::infinity();
    const float interval = 1.f / 60.f;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        try {
            if (a == "-i") scene_name = next();
            else if (a == "-o") recording = true;
            else if (a == "-d") duration = 0.001f * (float)std::stoi(next());
            else if (a == "-n") steps = (uint32_t)std::stoul(next());
            else if (a == "--out-dir") outdir = next();
            else if (a == "--final") final_path = next();
            else if (a == "--resume") resume = next();
            else if (a == "--precision") precision = next();
            else if (a == "--max-iterations") max_it = (uint32_t)std::stoul(next());
            else if (a == "--emit-scene") emit = next();
            else {
                std::printf("Invalid argument: %s\n", a.c_str());
                return 2;
            }
        } catch (const std::exception&) {
            std::printf("Invalid value for %s\n", a.c_str());
            return 2;
        }
    }
    if (scene_name.empty()) {
        std::printf("-i <scene_name>\n");
        return 2;
    }
    scene_gen::Scene sc;
    try {
        if (scene_name == "dam_break_16k") sc = scene_gen::dam_break_16k();
        else if (scene_name == "dam_break_1m") sc = scene_gen::dam_break_1m();
        else if (scene_name.rfind("xraw:", 0) == 0) {
            const auto c = scene_name.find(':', 5);
            if (c == std::string::npos) throw std::runtime_error("use xraw:<dir>:<name>");
            sc = scene_gen::reference_scene(scene_name.substr(5, c - 5), scene_name.substr(c + 1));
        } else throw std::runtime_error("unknown scene " + scene_name);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    if (!emit.empty()) {  // scene only, no device: fluid then boundary, float32
        std::ofstream f(emit, std::ios::binary);
        const uint32_t n = (uint32_t)(sc.fluid.size() / 3), nb = (uint32_t)(sc.boundary.size() / 3);
        f.write((const char*)&n, 4);
        f.write((const char*)&nb, 4);
        f.write((const char*)sc.fluid.data(), sc.fluid.size() * 4);
        f.write((const char*)sc.boundary.data(), sc.boundary.size() * 4);
        std::printf("scene %s: %u fluid, %u boundary particles\n", sc.name.c_str(), n, nb);
        return f ? 0 : 1;
    }
    if (!steps && !std::isfinite(duration)) steps = sc.steps;
    std::vector<float> vel;
    uint64_t step0 = 0;
    if (!resume.empty() && !read_frame(resume, sc, vel, step0)) {
        std::fprintf(stderr, "cannot read frame %s\n", resume.c_str());
        return 1;
    }

    pcisph_solver* s = nullptr;
    if (int rc = pcisph_create(0, nullptr, &s)) {
        std::fprintf(stderr, "pcisph_create failed (%d): %s\n", rc, pcisph_last_error(nullptr));
        return 1;
    }
    const int prec = precision == "ieee" ? PCISPH_PRECISION_IEEE : precision == "fp64" ? PCISPH_PRECISION_FP64
                                                                                      : PCISPH_PRECISION_FAST;
    check(s, pcisph_set_precision(s, prec), "pcisph_set_precision");
    // scene::load order (PCISPH/src/scenes.cpp:217-226,166-168)
    check(s, pcisph_set_delta_t(s, sc.delta_t), "set_delta_t");
    check(s, pcisph_set_rest_density(s, sc.rest_density), "set_rest_density");
    check(s, pcisph_set_viscosity(s, sc.viscosity), "set_viscosity");
    check(s, pcisph_set_surface_tension(s, sc.surface_tension), "set_surface_tension");
    check(s, pcisph_set_gravity(s, sc.gravity), "set_gravity");
    check(s, pcisph_set_density_variation_threshold(s, sc.threshold), "set_threshold");
    check(s, pcisph_set_particle_radius(s, sc.particle_radius), "set_particle_radius");
    check(s, pcisph_set_fluid_count(s, (uint32_t)(sc.fluid.size() / 3)), "set_fluid_count");
    check(s, pcisph_set_boundary_count(s, (uint32_t)(sc.boundary.size() / 3)), "set_boundary_count");
    pcisph_solver_options opt{};
    check(s, pcisph_get_solver_options(s, &opt), "get_solver_options");  // library defaults
    opt.min_iterations = sc.min_iterations;
    opt.max_iterations = max_it ? max_it : sc.max_iterations;
    check(s, pcisph_set_solver_options(s, &opt), "set_solver_options");
    check(s, pcisph_upload(s, PCISPH_FLUID_POSITIONS, sc.fluid.data(), sc.fluid.size()), "upload positions");
    if (!sc.boundary.empty())
        check(s, pcisph_upload(s, PCISPH_BOUNDARY_POSITIONS, sc.boundary.data(), sc.boundary.size()),
              "upload boundary");
    if (!vel.empty()) {
        // a resumed run keeps its velocities: apply the setters first, then upload
        check(s, pcisph_prepare(s), "pcisph_prepare");
        check(s, pcisph_upload(s, PCISPH_FLUID_VELOCITIES, vel.data(), vel.size()), "upload velocities");
    }
    std::printf("scene %s: %zu fluid, %zu boundary particles\n", sc.name.c_str(), sc.fluid.size() / 3,
                sc.boundary.size() / 3);
    // main.cpp:157-225: step, advance the simulated time; while recording, step until a
    // frame interval has passed, then record; stop at the duration
    if (recording) {
        std::error_code ec;
        std::filesystem::create_directories(outdir, ec);
        if (ec) {
            std::fprintf(stderr, "cannot create %s: %s\n", outdir.c_str(), ec.message().c_str());
            return 1;
        }
    }
    float sim_time = (float)step0 * sc.delta_t;
    float last_frame = -std::numeric_limits<float>::infinity();
    uint32_t frame = 0, done = 0;
    double it_sum = 0, err_max = 0;
    auto limit = [&] { return (steps && done >= steps) || sim_time >= duration; };
    const auto t0 = std::chrono::steady_clock::now();
    while (!limit()) {
        do {
            pcisph_step_stats st;
            check(s, pcisph_step(s, &st), "pcisph_step");
            ++done;
            sim_time += sc.delta_t;
            it_sum += st.iterations;
            err_max = std::max(err_max, (double)st.max_density_error);
            if (st.nonfinite) {
                std::fprintf(stderr, "non-finite state at step %llu\n", (unsigned long long)(step0 + done));
                return 3;
            }
        } while (recording && sim_time - last_frame <= interval && !limit());
        if (recording && sim_time - last_frame >= interval) {
            write_frame(s, outdir + "/frame_" + std::to_string(frame) + ".pcsf", step0 + done);
            ++frame;
            last_frame = sim_time;
        }
    }
    if (!final_path.empty()) write_frame(s, final_path, step0 + done);
    check(s, pcisph_sync(s), "pcisph_sync");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("%u steps (simulated time %.4f s, %u frames) in %.3f s: %.1f steps/s, %.4g particle-steps/s, "
                "mean %.2f iterations, max error %.4f\n",
                done, sim_time, frame, secs, done / secs, (sc.fluid.size() / 3) * done / secs,
                done ? it_sum / done : 0.0, err_max);
    pcisph_destroy(s);
    return 0;
}
