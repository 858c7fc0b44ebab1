// material_point3d.h -- drop-in for the reference's particle types
// (include/material_point3d.h:14-72, include/apic_material_point3d.h:14-74):
// the same field names and constructor arguments.  The device-only methods
// of the reference (stress, F update) run inside the solver's kernels.
#ifndef CHAINS_MATERIAL_POINT3D_H_
#define CHAINS_MATERIAL_POINT3D_H_

#include "enums.h"
#include "small_linalg.h"

namespace chains {

struct MaterialPoint3D {
    float volume0 = 0.0f, mass = 0.0f;
    Vector3f position, velocity;
    Matrix3f deformation_gradient_elastic, deformation_gradient_plastic;
    float mu0 = 0.0f, lambda0 = 0.0f;
    float hardening_coefficient = 0.0f;
    float critical_compression = 0.0f, critical_stretch = 0.0f;
    ConstitutiveModel constitutive_model = ConstitutiveModel::FIXED_COROTATED;

    MaterialPoint3D() = default;
    // Reference src/material_point3d.cu:10-30 (Lame parameters from E, nu).
    MaterialPoint3D(Vector3f particlePosition, Vector3f particleVelocity, float particleMass,
                    float hardeningCoefficient, float YoungModulus, float PoissonRatio, float criticalCompression,
                    float criticleStretch, ConstitutiveModel constitutiveModel);
};

struct APICMaterialPoint3D : MaterialPoint3D {
    Matrix3f B, D, Dinv;  // reference include/apic_material_point3d.h:23

    APICMaterialPoint3D() = default;
    // Reference src/apic_material_point3d.cu:10-34 (B = D = Dinv = 0).
    APICMaterialPoint3D(Vector3f particlePosition, Vector3f particleVelocity, float particleMass,
                        float hardeningCoefficient, float YoungModulus, float PoissonRatio, float criticalCompression,
                        float criticleStretch, ConstitutiveModel constitutiveModel);
};

}  // namespace chains

#endif  // CHAINS_MATERIAL_POINT3D_H_
