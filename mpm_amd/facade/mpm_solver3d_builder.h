// mpm_solver3d_builder.h -- drop-in for the reference's builders
// (include/mpm_solver3d_builder.h:10-12): YAML configuration -> solver,
// returned by value (the facade classes are move-only).
#ifndef CHAINS_MPM_SOLVER3D_BUILDER_H_
#define CHAINS_MPM_SOLVER3D_BUILDER_H_

#include <string>
#include <vector>

#include "apic_mpm_solver3d.h"
#include "mpm3d_parser.h"
#include "mpm_solver3d.h"

namespace chains {

// Reference src/mpm_solver3d_builder.cu:10-30: whitespace-separated x y z
// per line; false (and a message on stderr) if the file cannot be opened or
// a line does not parse.
bool readXYZFile(const std::string& filePath, std::vector<Vector3f>& points);

MPMSolver3D buildMPMSolver3DFromYAML(MPM3DConfiguration config, int& particlesNumber);
APICMPMSolver3D buildAPICMPMSolver3DFromYAML(MPM3DConfiguration config, int& particlesNumber);

}  // namespace chains

#endif  // CHAINS_MPM_SOLVER3D_BUILDER_H_
