// mpm_solver3d.h -- drop-in for the reference's PIC-FLIP solver class
// (include/mpm_solver3d.h:21-96): same constructor arguments, public field
// and method names, over the C-ABI of libmpm_hip.so.
//
// Differences a caller can see (the reference is GPU-only CUDA/Thrust):
//  * errors throw std::runtime_error instead of exit(EXIT_FAILURE);
//  * the class is move-only (the reference's implicit copy double-frees);
//  * steps are asynchronous; every read synchronises;
//  * the OpenGL methods are replaced by host-buffer equivalents
//    (copyPositionsToHost, readParticles);
//  * writeToOpenVDB keeps its name and density raster but writes the
//    density block file of mpm_write_density (OpenVDB is not used).
#ifndef CHAINS_MPM_SOLVER_3D_H_
#define CHAINS_MPM_SOLVER_3D_H_

#include <string>
#include <vector>

#include "enums.h"
#include "material_point3d.h"
#include "small_linalg.h"

namespace chains {

// Common handle wrapper of both solver classes.
class MPMSolverHandle {
public:
    bool _enable_particles_collision = false;

    MPMSolverHandle(MPMSolverHandle&& o) noexcept;
    MPMSolverHandle& operator=(MPMSolverHandle&& o) noexcept;
    MPMSolverHandle(const MPMSolverHandle&) = delete;
    MPMSolverHandle& operator=(const MPMSolverHandle&) = delete;
    ~MPMSolverHandle();

    void simulateOneStep();
    void switchIfEnableParticlesCollision();
    void initialize();
    void resetGrid();
    void particlesToGrid();
    void computeExternalForces();
    void updateGridVelocities(float deltaTimeInSeconds);
    void solveLinearSystem(float deltaTimeInSeconds);
    void advectParticles(float deltaTimeInSeconds);
    void computeGravityForces();
    void gridCollision(float deltaTimeInSeconds);
    void particlesCollision(float deltaTimeInSeconds);

    // Replacement of updateGLBufferByCPU / updateGLBufferWithCUDA
    // (src/mpm_solver3d.cu:457-508): positions packed as (x, y, z, 1).
    void copyPositionsToHost(float* xyz1) const;
    // writeToOpenVDB (src/mpm_solver3d.cu:510-572): the density raster
    // m w / h^3 of the current positions, computed on the device and
    // written as the 4x4x4-block file of mpm_write_density.
    void writeToOpenVDB(std::string filePath);
    void synchronize() const;
    int particleCount() const { return _n; }
    void* handle() const { return _h; }

protected:
    MPMSolverHandle(const std::vector<MaterialPoint3D>& particles, const std::vector<Matrix3f>* B,
                    Vector3f gridOrigin, Vector3i gridResolution, float gridStride,
                    float gridBoundaryFrictionCoefficient, float blendCoefficient, InterpolationType interpolationType,
                    float deltaTime, int transfer, float gravity);
    void readState(std::vector<MaterialPoint3D>& out, std::vector<Matrix3f>* B) const;
    void syncCollisionFlag();

    void* _h = nullptr;
    int _n = 0;
    float _blend_coefficient = 0.0f;
    float _delta_time = 0.0f;
};

/* PIC-FLIP 3D Material Point Method Solver */
class MPMSolver3D : public MPMSolverHandle {
public:
    MPMSolver3D(const std::vector<MaterialPoint3D>& particles, Vector3f gridOrigin, Vector3i gridResolution,
                float gridStride, float gridBoundaryFrictionCoefficient, float blendCoefficient,
                InterpolationType interpolationType, float deltaTime, float gravity = 9.80665f);
    void gridToParticles(float deltaTimeInSeconds, float blendCoefficient);
    void readParticles(std::vector<MaterialPoint3D>& out) const { readState(out, nullptr); }
};

}  // namespace chains

#endif  // CHAINS_MPM_SOLVER_3D_H_
