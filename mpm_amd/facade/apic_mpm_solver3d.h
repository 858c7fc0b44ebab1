// apic_mpm_solver3d.h -- drop-in for the reference's APIC solver class
// (include/apic_mpm_solver3d.h:20-92).  See mpm_solver3d.h for the
// behaviour shared with the PIC-FLIP class.
#ifndef CHAINS_APIC_MPM_SOLVER_3D_H_
#define CHAINS_APIC_MPM_SOLVER_3D_H_

#include "mpm_solver3d.h"

namespace chains {

/* APIC 3D Material Point Method Solver */
class APICMPMSolver3D : public MPMSolverHandle {
public:
    APICMPMSolver3D(const std::vector<APICMaterialPoint3D>& particles, Vector3f gridOrigin, Vector3i gridResolution,
                    float gridStride, float gridBoundaryFrictionCoefficient, InterpolationType interpolationType,
                    float deltaTime, float gravity = 9.80665f);
    void gridToParticles(float deltaTimeInSeconds);
    void readParticles(std::vector<APICMaterialPoint3D>& out) const;
};

}  // namespace chains

#endif  // CHAINS_APIC_MPM_SOLVER_3D_H_
