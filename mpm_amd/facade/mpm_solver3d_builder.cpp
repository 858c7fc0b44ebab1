// mpm_solver3d_builder.cpp -- reference src/mpm_solver3d_builder.cu:10-129.
// Particle groups are read in order, positions scaled by the grid stride,
// one MaterialPoint3D per point with the group's parameters, groups
// stacked.  As in the reference, a group whose file cannot be read is
// reported and skipped, and relative paths resolve against the working
// directory.
#include "mpm_solver3d_builder.h"

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>

namespace chains {

namespace {

bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }
bool is_digit(char c) { return c >= '0' && c <= '9'; }

// One decimal number at s (blanks skipped first, never past the end of the
// line): [+-] digits [. digits] [e [+-] digits], or [+-] . digits ...  This
// is the form a stream extraction of a float reads (no inf, nan or hex).
// Returns the position after it, or nullptr when no number starts there.
const char* scan_decimal(const char* s, const char* eol, float& out) {
    while (s < eol && is_blank(*s)) ++s;
    const char* p = s;
    if (p < eol && (*p == '+' || *p == '-')) ++p;
    const char* mant = p;
    while (p < eol && is_digit(*p)) ++p;
    bool digits = p > mant;
    if (p < eol && *p == '.') {
        const char* frac = ++p;
        while (p < eol && is_digit(*p)) ++p;
        digits = digits || p > frac;
    }
    if (!digits) return nullptr;
    if (p < eol && (*p == 'e' || *p == 'E')) {
        const char* q = p + 1;
        if (q < eol && (*q == '+' || *q == '-')) ++q;
        const char* ex = q;
        while (q < eol && is_digit(*q)) ++q;
        if (q > ex) p = q;
    }
    char buf[128];
    const size_t len = (size_t)(p - s);
    if (len >= sizeof(buf)) return nullptr;
    std::memcpy(buf, s, len);
    buf[len] = '\0';
    errno = 0;
    const float v = std::strtof(buf, nullptr);
    if (errno == ERANGE && std::isinf(v)) return nullptr;  // out of float range
    out = v;
    return p;
}

}  // namespace

TextGroupStatus readTextGroup(const std::string& filePath, std::vector<Vector3f>& points, std::string* badLine) {
    std::ifstream in(filePath, std::ios::binary);
    if (!in.is_open()) return TextGroupStatus::kUnreadable;
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const char* s = text.data();
    const char* const end = s + text.size();
    while (s < end) {
        const char* eol = static_cast<const char*>(std::memchr(s, '\n', (size_t)(end - s)));
        if (!eol) eol = end;
        Vector3f p;
        const char* c = s;
        for (int a = 0; a < 3 && c; ++a) c = scan_decimal(c, eol, p(a));
        if (!c) {
            if (badLine) badLine->assign(s, eol);
            return TextGroupStatus::kStopped;
        }
        points.push_back(p);  // text after the third number is ignored
        s = eol + 1;
    }
    return TextGroupStatus::kComplete;
}

bool readXYZFile(const std::string& filePath, std::vector<Vector3f>& points) {
    std::string bad;
    switch (readTextGroup(filePath, points, &bad)) {
        case TextGroupStatus::kComplete: return true;
        case TextGroupStatus::kUnreadable:
            std::cerr << "[ERROR] cannot open file " << filePath << std::endl;
            return false;
        default:
            std::cerr << "[ERROR] cannot parse line: " << bad << " when trying to parse file " << filePath
                      << std::endl;
            return false;
    }
}

bool readXYZBFile(const std::string& filePath, ParticleGroupData& out) {
    std::ifstream in(filePath, std::ios::binary);
    if (!in.is_open()) {
        std::cerr << "[ERROR] cannot open file " << filePath << std::endl;
        return false;
    }
    char magic[8];
    int64_t n = 0;
    int32_t fields = 0, zero = 0;
    in.read(magic, 8);
    in.read(reinterpret_cast<char*>(&n), 8);
    in.read(reinterpret_cast<char*>(&fields), 4);
    in.read(reinterpret_cast<char*>(&zero), 4);
    if (!in || std::memcmp(magic, "MPMXYZB1", 8) != 0 || n < 0 || n > (int64_t(1) << 31) || (fields & ~15) != 0 ||
        zero != 0) {
        std::cerr << "[ERROR] " << filePath << " is not a binary particle group" << std::endl;
        return false;
    }
    auto rd = [&](std::vector<float>& v, size_t k) {
        v.resize((size_t)n * k);
        in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(float)));
        return (bool)in;
    };
    std::vector<float> pos, vel;
    ParticleGroupData g;
    g.world = (fields & 8) != 0;
    bool ok = rd(pos, 3);
    if (ok && (fields & 1)) ok = rd(vel, 3);
    if (ok && (fields & 2)) ok = rd(g.young, 1);
    if (ok && (fields & 4)) ok = rd(g.mass, 1);
    if (ok) {
        in.peek();
        ok = in.eof();  // no trailing bytes
    }
    if (!ok) {
        std::cerr << "[ERROR] " << filePath << ": truncated or oversized binary particle group" << std::endl;
        return false;
    }
    g.position.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) g.position[(size_t)i] = Vector3f(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    if (fields & 1) {
        g.velocity.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) g.velocity[(size_t)i] = Vector3f(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    }
    out = std::move(g);
    return true;
}

bool readParticleGroup(const std::string& filePath, ParticleGroupData& out) {
    const std::string ext = ".xyzb";
    if (filePath.size() >= ext.size() && filePath.compare(filePath.size() - ext.size(), ext.size(), ext) == 0)
        return readXYZBFile(filePath, out);
    out = ParticleGroupData();
    return readXYZFile(filePath, out.position);  // on false, out keeps the lines read before the stop
}

template <class P>
static std::vector<P> build_points(const MPM3DConfiguration& config, int& particlesNumber) {
    std::vector<P> particles;
    const size_t groups = config.group_paths.size();
    std::cout << "[INFO] building " << groups << " group(s) of particle on host" << std::endl;
    for (size_t i = 0; i < groups; ++i) {
        // As in the reference (whose builder ignores readXYZFile's result),
        // an unreadable file gives an empty group and a text file that
        // stops parsing gives the points read before the bad line.
        ParticleGroupData g;
        (void)readParticleGroup(config.group_paths[i], g);
        const size_t n = g.position.size();
        particles.reserve(particles.size() + n);
        for (size_t k = 0; k < n; ++k)
            particles.emplace_back(g.world ? g.position[k] : config.stride * g.position[k],
                                   g.velocity.empty() ? config.velocities[i] : g.velocity[k],
                                   g.mass.empty() ? config.masses[i] : g.mass[k], config.hardenings[i],
                                   g.young.empty() ? config.youngs[i] : g.young[k], config.poissons[i],
                                   config.compressions[i], config.stretches[i], config.constitutive_models[i]);
        std::cout << "[INFO] particle group " << i << " has " << n << " particles" << std::endl;
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

// Reads one particle group as the builders do (positions scaled by stride
// unless the file holds world units): for the loader tests.  Writes up to
// max particles; *count gets the group size and *fields the per-particle
// fields present (1 velocity, 2 young, 4 mass).  0 when the whole file was
// read, 1 when a text file stopped at a line that does not parse (the group
// is the lines before it, as the builders keep them), -1 if the file cannot
// be read (an empty group).
extern "C" __attribute__((visibility("default"))) int chains_read_group(const char* path, float stride, float* pos,
                                                                        float* vel, float* young, float* mass,
                                                                        int64_t max, int64_t* count, int* fields) {
    chains::ParticleGroupData g;
    if (!path) return -1;
    const bool ok = chains::readParticleGroup(path, g);
    if (!ok && g.position.empty()) {
        if (count) *count = 0;
        return -1;
    }
    const int64_t n = (int64_t)g.position.size();
    if (count) *count = n;
    if (fields) *fields = (g.velocity.empty() ? 0 : 1) | (g.young.empty() ? 0 : 2) | (g.mass.empty() ? 0 : 4);
    for (int64_t i = 0; i < n && i < max; ++i) {
        const chains::Vector3f x = g.world ? g.position[(size_t)i] : stride * g.position[(size_t)i];
        for (int a = 0; a < 3; ++a) {
            if (pos) pos[3 * i + a] = x(a);
            if (vel && !g.velocity.empty()) vel[3 * i + a] = g.velocity[(size_t)i](a);
        }
        if (young && !g.young.empty()) young[i] = g.young[(size_t)i];
        if (mass && !g.mass.empty()) mass[i] = g.mass[(size_t)i];
    }
    return ok ? 0 : 1;
}
