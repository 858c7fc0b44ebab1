// mpm_solver3d_builder.cpp -- reference src/mpm_solver3d_builder.cu:10-129.
// Particle groups are read in order, positions scaled by the grid stride,
// one MaterialPoint3D per point with the group's parameters, groups
// stacked.  As in the reference, a group whose file cannot be read is
// reported and skipped, and relative paths resolve against the working
// directory.
#include "mpm_solver3d_builder.h"

#include <fstream>
#include <iostream>
#include <sstream>

namespace chains {

bool readXYZFile(const std::string& filePath, std::vector<Vector3f>& points) {
    std::ifstream in(filePath);
    if (!in.is_open()) {
        std::cerr << "[ERROR] cannot open file " << filePath << std::endl;
        return false;
    }
    std::string line;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        Vector3f p;
        if (!(ss >> p.x() >> p.y() >> p.z())) {
            std::cerr << "[ERROR] cannot parse line: " << line << " when trying to parse file " << filePath << std::endl;
            return false;
        }
        points.push_back(p);
    }
    return true;
}

template <class P>
static std::vector<P> build_points(const MPM3DConfiguration& config, int& particlesNumber) {
    std::vector<P> particles;
    const size_t groups = config.group_paths.size();
    std::cout << "[INFO] building " << groups << " group(s) of particle on host" << std::endl;
    for (size_t i = 0; i < groups; ++i) {
        std::vector<Vector3f> coords;
        readXYZFile(config.group_paths[i], coords);
        for (const Vector3f& c : coords)
            particles.emplace_back(config.stride * c, config.velocities[i], config.masses[i], config.hardenings[i],
                                   config.youngs[i], config.poissons[i], config.compressions[i], config.stretches[i],
                                   config.constitutive_models[i]);
        std::cout << "[INFO] particle group " << i << " has " << coords.size() << " particles" << std::endl;
    }
    particlesNumber = (int)particles.size();
    std::cout << "[INFO] stacking all " << particlesNumber << " particles on host" << std::endl;
    return particles;
}

MPMSolver3D buildMPMSolver3DFromYAML(MPM3DConfiguration config, int& particlesNumber) {
    std::vector<MaterialPoint3D> particles = build_points<MaterialPoint3D>(config, particlesNumber);
    std::cout << "[INFO] building MPM Solver 3D" << std::endl;
    return MPMSolver3D(particles, config.origin, config.resolution, config.stride, config.friction_coeff,
                       config.blend_coeff, config.interp_type, config.delta_time, config.gravity);
}

APICMPMSolver3D buildAPICMPMSolver3DFromYAML(MPM3DConfiguration config, int& particlesNumber) {
    std::vector<APICMaterialPoint3D> particles = build_points<APICMaterialPoint3D>(config, particlesNumber);
    std::cout << "[INFO] building MPM Solver 3D" << std::endl;
    return APICMPMSolver3D(particles, config.origin, config.resolution, config.stride, config.friction_coeff,
                           config.interp_type, config.delta_time, config.gravity);
}

}  // namespace chains
