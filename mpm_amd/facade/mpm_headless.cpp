// mpm_headless.cpp -- headless counterpart of the reference demo programs
// (src/demo/mpm-pic-flip/main.cu, src/demo/mpm-apic/main.cu: parse the YAML
// given as the only positional argument, build the solver, loop
// simulateOneStep).  No window: positions are written to a binary file.
//
// usage: mpm_headless <config.yaml> [--apic] [--steps N] [--out positions.f32]
//        [--every K] [--collide] [--density DIR]
// --out writes N x 4 float32 (x, y, z, 1) after the last step (and every K
// steps to <out>.<step> when --every is given).
// --density writes the density raster (writeToOpenVDB) after every K-th
// step (every step without --every) to DIR/<step, 5 digits>.mpmd, as the
// reference's offline mode writes DIR/<step>.vdb each frame
// (src/demo/mpm-apic/main.cu:198-202); a config with Render.Offline: true
// and no --density writes to output/<program>-<config>.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "mpm_solver3d_builder.h"

static std::string frame_path(const std::string& dir, int step) {
    std::string num = std::to_string(step);
    return dir + "/" + std::string(num.size() < 5 ? 5 - num.size() : 0, '0') + num + ".mpmd";
}

template <class Solver>
static int run(Solver& solver, int n, int steps, const std::string& out, int every, bool collide,
               const std::string& density_dir) {
    if (collide) solver.switchIfEnableParticlesCollision();
    std::vector<float> buf((size_t)n * 4);
    auto dump = [&](const std::string& path) {
        solver.copyPositionsToHost(buf.data());
        std::ofstream f(path, std::ios::binary);
        f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(buf.size() * sizeof(float)));
    };
    solver.synchronize();
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 1; s <= steps; ++s) {
        solver.simulateOneStep();
        if (every > 0 && !out.empty() && s % every == 0) dump(out + "." + std::to_string(s));
        if (!density_dir.empty() && (every <= 0 || s % every == 0)) solver.writeToOpenVDB(frame_path(density_dir, s));
    }
    solver.synchronize();
    auto t1 = std::chrono::steady_clock::now();
    double sec = std::chrono::duration<double>(t1 - t0).count();
    std::printf("[INFO] %d particles, %d steps in %.3f s (%.3e particle-updates/s)\n", n, steps, sec,
                steps > 0 && sec > 0 ? (double)n * steps / sec : 0.0);
    if (!out.empty()) dump(out);
    return 0;
}

int main(int argc, const char* argv[]) {
    if (argc < 2) {
        std::printf("Usage: %s path/to/configuration/file [--apic] [--steps N] [--out file] [--every K] [--collide]"
                    " [--density DIR]\n", argv[0]);
        return 1;
    }
    std::string config_path = argv[1];
    bool apic = false, collide = false;
    int steps = -1, every = 0;
    std::string out, density_dir;
    for (int i = 2; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--apic") apic = true;
        else if (a == "--collide") collide = true;
        else if (a == "--steps" && i + 1 < argc) steps = std::atoi(argv[++i]);
        else if (a == "--every" && i + 1 < argc) every = std::atoi(argv[++i]);
        else if (a == "--out" && i + 1 < argc) out = argv[++i];
        else if (a == "--density" && i + 1 < argc) density_dir = argv[++i];
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 1;
        }
    }
    try {
        chains::MPM3DConfiguration config = chains::parseYAML(config_path);
        if (steps < 0) steps = config.steps > 0 ? config.steps : 100;
        if (config.scheme == 1) apic = true;
        if (density_dir.empty() && config.offline)
            density_dir = "output/" + std::filesystem::path(argv[0]).stem().string() + "-" +
                          std::filesystem::path(config_path).stem().string();
        if (!density_dir.empty()) std::filesystem::create_directories(density_dir);
        int n = -1;
        if (apic) {
            chains::APICMPMSolver3D solver = chains::buildAPICMPMSolver3DFromYAML(config, n);
            return run(solver, n, steps, out, every, collide, density_dir);
        }
        chains::MPMSolver3D solver = chains::buildMPMSolver3DFromYAML(config, n);
        return run(solver, n, steps, out, every, collide, density_dir);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[ERROR] %s\n", e.what());
        return 2;
    }
}
