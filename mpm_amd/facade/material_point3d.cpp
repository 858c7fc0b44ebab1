#include "material_point3d.h"

namespace chains {

MaterialPoint3D::MaterialPoint3D(Vector3f particlePosition, Vector3f particleVelocity, float particleMass,
                                 float hardeningCoefficient, float YoungModulus, float PoissonRatio,
                                 float criticalCompression, float criticleStretch,
                                 ConstitutiveModel constitutiveModel)
    : mass(particleMass),
      position(particlePosition),
      velocity(particleVelocity),
      hardening_coefficient(hardeningCoefficient),
      critical_compression(criticalCompression),
      critical_stretch(criticleStretch),
      constitutive_model(constitutiveModel) {
    mu0 = (float)(YoungModulus / (2.0 * (1.0 + PoissonRatio)));
    lambda0 = (float)((YoungModulus * PoissonRatio) / ((1.0 + PoissonRatio) * (1.0 - 2.0 * PoissonRatio)));
    deformation_gradient_elastic.setIdentity();
    deformation_gradient_plastic.setIdentity();
}

APICMaterialPoint3D::APICMaterialPoint3D(Vector3f particlePosition, Vector3f particleVelocity, float particleMass,
                                         float hardeningCoefficient, float YoungModulus, float PoissonRatio,
                                         float criticalCompression, float criticleStretch,
                                         ConstitutiveModel constitutiveModel)
    : MaterialPoint3D(particlePosition, particleVelocity, particleMass, hardeningCoefficient, YoungModulus,
                      PoissonRatio, criticalCompression, criticleStretch, constitutiveModel) {
    B.setZero();
    D.setZero();
    Dinv.setZero();
}

}  // namespace chains
