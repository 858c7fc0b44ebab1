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

// Text particle group: one point "x y z" per line (blanks between and
// before the numbers, anything after the third ignored).  Reading stops at
// the first line without three numbers -- a blank line included -- and the
// points read before it stay in `points` (badLine gets that line).
enum class TextGroupStatus { kComplete, kStopped, kUnreadable };
TextGroupStatus readTextGroup(const std::string& filePath, std::vector<Vector3f>& points,
                              std::string* badLine = nullptr);

// Reference src/mpm_solver3d_builder.cu:10-30: appends the file's points;
// false (and the reference's message on stderr) if the file cannot be
// opened or reading stopped at a bad line, whose prefix is kept.
bool readXYZFile(const std::string& filePath, std::vector<Vector3f>& points);

// Binary particle group (no reference counterpart: the text reader above is
// the bottleneck of multi-million-particle scenes).  Little-endian:
//    0  char[8]  "MPMXYZB1"
//    8  int64    count n
//   16  int32    fields: 1 velocity, 2 Young's modulus, 4 mass,
//                8 positions in world units (else cell units, scaled by
//                Grid.Stride like an .xyz file)
//   20  int32    0
//   24  float32  position[n][3], then in this order when flagged:
//                velocity[n][3], young[n], mass[n]
// Per-particle fields replace the group's YAML values.  A group File whose
// name ends in ".xyzb" is read this way; false (and a message on stderr)
// if the file cannot be opened or is malformed.
struct ParticleGroupData {
    std::vector<Vector3f> position;
    std::vector<Vector3f> velocity;  // empty unless flagged
    std::vector<float> young, mass;  // likewise
    bool world = false;
};
bool readXYZBFile(const std::string& filePath, ParticleGroupData& out);
// Either format, by file name.
bool readParticleGroup(const std::string& filePath, ParticleGroupData& out);

MPMSolver3D buildMPMSolver3DFromYAML(MPM3DConfiguration config, int& particlesNumber);
APICMPMSolver3D buildAPICMPMSolver3DFromYAML(MPM3DConfiguration config, int& particlesNumber);

}  // namespace chains

#endif  // CHAINS_MPM_SOLVER3D_BUILDER_H_
