// mpm3d_parser.h -- drop-in for the reference's include/mpm3d_parser.h:13-37.
// Same MPM3DConfiguration fields and parseYAML(); the YAML subset the
// reference configs use is parsed by an in-tree reader (yaml-cpp is not a
// dependency).  Missing keys and unknown enum strings throw
// std::runtime_error, as yaml-cpp and the reference parser do
// (src/mpm3d_parser.cpp:7-18).
#ifndef CHAINS_MPM3D_PARSER_H_
#define CHAINS_MPM3D_PARSER_H_

#include <string>
#include <utility>
#include <vector>

#include "enums.h"
#include "small_linalg.h"

namespace chains {

struct MPM3DConfiguration {
    // Particle groups
    std::vector<std::string> group_paths;
    std::vector<Vector3f> velocities;
    std::vector<float> masses, youngs, poissons, hardenings, compressions, stretches;
    std::vector<ConstitutiveModel> constitutive_models;
    // Grid
    Vector3f origin;
    Vector3i resolution;
    float stride = 0.0f, friction_coeff = 0.0f;
    // Transfer
    float blend_coeff = 0.0f;
    // Interpolation
    InterpolationType interp_type = InterpolationType::CUBIC_BSPLINE;
    // Simulate
    float delta_time = 0.0f;
    // Render
    bool offline = false;
    std::string particle_frag_shader_path;
    // Optional keys beyond the reference schema (absent in reference files):
    float gravity = 9.80665f;  // Simulate.Gravity, magnitude along -y
    int steps = 0;             // Simulate.Steps, for headless runs
    int scheme = -1;           // Transfer.Scheme: 0 PICFLIP, 1 APIC, -1 absent (the demo program decides)
};

ConstitutiveModel parseConstitutiveModel(const std::string& model);
InterpolationType parseInterpolationType(const std::string& type);
MPM3DConfiguration parseYAML(const std::string& filePath);

// Minimal YAML reader: block mappings, block sequences ("- "), flow
// sequences ("[a, b]"), plain / quoted scalars, comments.
namespace yaml {
struct Node {
    enum Kind { Null, Scalar, Map, Seq } kind = Null;
    std::string value;
    std::vector<std::pair<std::string, Node>> map;
    std::vector<Node> seq;
    std::string where;  // key path, for error messages

    bool has(const std::string& key) const;
    const Node& operator[](const std::string& key) const;  // throws if missing
    const Node& operator[](size_t i) const;                 // throws if out of range
    size_t size() const { return kind == Seq ? seq.size() : kind == Map ? map.size() : 0; }
    float asFloat() const;
    int asInt() const;
    bool asBool() const;
    std::string asString() const;
};
Node parseText(const std::string& text, const std::string& origin = "<text>");
Node loadFile(const std::string& path);
}  // namespace yaml

}  // namespace chains

#endif  // CHAINS_MPM3D_PARSER_H_
