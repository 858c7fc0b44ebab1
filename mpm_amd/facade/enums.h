// enums.h -- drop-in for the reference's include/enums.h:6-21 (same names,
// same enumerator order, so values match mpm_interp / mpm_model).
#ifndef CHAINS_ENUMS_H_
#define CHAINS_ENUMS_H_

namespace chains {

enum class InterpolationType { LINEAR, QUADRATIC_BSPLINE, CUBIC_BSPLINE };

enum class ConstitutiveModel { FIXED_COROTATED, NEO_HOOKEAN };

enum class IntegrationType { EXPLICIT, SEMI_IMPLICIT, FULL_IMPLICIT };

}  // namespace chains

#endif  // CHAINS_ENUMS_H_
