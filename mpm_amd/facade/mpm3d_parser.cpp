// mpm3d_parser.cpp -- YAML subset reader and parseYAML()
// (reference src/mpm3d_parser.cpp:20-71: same keys, same required set).
#include "mpm3d_parser.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace chains {

ConstitutiveModel parseConstitutiveModel(const std::string& model) {
    if (model == "FixedCorotated") return ConstitutiveModel::FIXED_COROTATED;
    if (model == "NeoHookean") return ConstitutiveModel::NEO_HOOKEAN;
    throw std::runtime_error("Unsupported constitutive model: " + model);
}

InterpolationType parseInterpolationType(const std::string& type) {
    if (type == "Linear") return InterpolationType::LINEAR;
    if (type == "QuadraticBSpline") return InterpolationType::QUADRATIC_BSPLINE;
    if (type == "CubicBSpline") return InterpolationType::CUBIC_BSPLINE;
    throw std::runtime_error("Unsupported interpolation type: " + type);
}

namespace yaml {
namespace {

struct Line {
    int indent;
    std::string text;
    int lineno;
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
    return s.substr(a, b - a);
}

std::string strip_comment(const std::string& s) {
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); ++i) {
        char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (c == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return s.substr(0, i);
    }
    return s;
}

std::string unquote(const std::string& s) {
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
        return s.substr(1, s.size() - 2);
    return s;
}

// Position of the ':' that separates a mapping key (outside quotes and
// brackets, followed by space or end of line), or npos.
size_t key_colon(const std::string& s) {
    bool sq = false, dq = false;
    int depth = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (!sq && !dq && (c == '[' || c == '{')) ++depth;
        else if (!sq && !dq && (c == ']' || c == '}')) --depth;
        else if (c == ':' && !sq && !dq && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
    }
    return std::string::npos;
}

Node scalar_or_flow(const std::string& raw, const std::string& where) {
    Node n;
    n.where = where;
    std::string v = trim(raw);
    if (!v.empty() && v.front() == '[') {
        if (v.back() != ']') throw std::runtime_error("yaml: unterminated flow sequence at " + where);
        n.kind = Node::Seq;
        std::string body = v.substr(1, v.size() - 2);
        std::string cur;
        int depth = 0;
        bool sq = false, dq = false;
        auto flush = [&](bool final_item) {
            std::string t = trim(cur);
            if (!t.empty() || !final_item) {
                n.seq.push_back(scalar_or_flow(t, where + "[" + std::to_string(n.seq.size()) + "]"));
            }
            cur.clear();
        };
        for (char c : body) {
            if (c == '\'' && !dq) sq = !sq;
            if (c == '"' && !sq) dq = !dq;
            if (!sq && !dq && c == '[') ++depth;
            if (!sq && !dq && c == ']') --depth;
            if (c == ',' && depth == 0 && !sq && !dq) flush(false);
            else cur += c;
        }
        flush(true);
        return n;
    }
    if (v.empty() || v == "~" || v == "null") {
        n.kind = Node::Null;
        return n;
    }
    n.kind = Node::Scalar;
    n.value = unquote(v);
    return n;
}

Node parse_block(std::vector<Line>& lines, size_t& i, int indent, const std::string& where);

Node parse_map(std::vector<Line>& lines, size_t& i, int indent, const std::string& where) {
    Node n;
    n.kind = Node::Map;
    n.where = where;
    while (i < lines.size() && lines[i].indent == indent) {
        const std::string& t = lines[i].text;
        if (t.rfind("- ", 0) == 0 || t == "-") break;
        size_t c = key_colon(t);
        if (c == std::string::npos)
            throw std::runtime_error("yaml: expected 'key: value' at line " + std::to_string(lines[i].lineno));
        std::string key = unquote(trim(t.substr(0, c)));
        std::string rest = trim(t.substr(c + 1));
        std::string path = where.empty() ? key : where + "." + key;
        ++i;
        Node child;
        if (!rest.empty()) {
            child = scalar_or_flow(rest, path);
        } else if (i < lines.size() && lines[i].indent > indent) {
            child = parse_block(lines, i, lines[i].indent, path);
        } else if (i < lines.size() && lines[i].indent == indent &&
                   (lines[i].text.rfind("- ", 0) == 0 || lines[i].text == "-")) {
            child = parse_block(lines, i, indent, path);  // "key:\n- item" at the same indent
        } else {
            child.kind = Node::Null;
            child.where = path;
        }
        n.map.emplace_back(key, std::move(child));
    }
    return n;
}

Node parse_seq(std::vector<Line>& lines, size_t& i, int indent, const std::string& where) {
    Node n;
    n.kind = Node::Seq;
    n.where = where;
    while (i < lines.size() && lines[i].indent == indent && (lines[i].text.rfind("- ", 0) == 0 || lines[i].text == "-")) {
        std::string rest = lines[i].text == "-" ? std::string() : trim(lines[i].text.substr(2));
        std::string path = where + "[" + std::to_string(n.seq.size()) + "]";
        if (rest.empty()) {
            ++i;
            if (i < lines.size() && lines[i].indent > indent) n.seq.push_back(parse_block(lines, i, lines[i].indent, path));
            else n.seq.push_back(scalar_or_flow("", path));
        } else if (key_colon(rest) != std::string::npos) {
            // "- key: v" opens a mapping whose keys sit at indent + 2
            int item_indent = indent + 2;
            lines[i].indent = item_indent;
            lines[i].text = rest;
            n.seq.push_back(parse_map(lines, i, item_indent, path));
        } else {
            n.seq.push_back(scalar_or_flow(rest, path));
            ++i;
        }
    }
    return n;
}

Node parse_block(std::vector<Line>& lines, size_t& i, int indent, const std::string& where) {
    if (i >= lines.size()) return Node();
    const std::string& t = lines[i].text;
    if (t.rfind("- ", 0) == 0 || t == "-") return parse_seq(lines, i, indent, where);
    return parse_map(lines, i, indent, where);
}

}  // namespace

bool Node::has(const std::string& key) const {
    if (kind != Map) return false;
    for (auto& kv : map)
        if (kv.first == key) return true;
    return false;
}

const Node& Node::operator[](const std::string& key) const {
    if (kind == Map)
        for (auto& kv : map)
            if (kv.first == key) return kv.second;
    throw std::runtime_error("yaml: missing key '" + key + "'" + (where.empty() ? "" : " under '" + where + "'"));
}

const Node& Node::operator[](size_t idx) const {
    if (kind != Seq || idx >= seq.size())
        throw std::runtime_error("yaml: no element " + std::to_string(idx) + " in '" + where + "'");
    return seq[idx];
}

float Node::asFloat() const {
    if (kind != Scalar) throw std::runtime_error("yaml: '" + where + "' is not a scalar");
    const char* s = value.c_str();
    char* end = nullptr;
    errno = 0;
    float f = std::strtof(s, &end);
    if (end == s || *end != '\0') throw std::runtime_error("yaml: bad float '" + value + "' at '" + where + "'");
    return f;
}

int Node::asInt() const {
    if (kind != Scalar) throw std::runtime_error("yaml: '" + where + "' is not a scalar");
    const char* s = value.c_str();
    char* end = nullptr;
    long v = std::strtol(s, &end, 10);
    if (end == s || *end != '\0') throw std::runtime_error("yaml: bad integer '" + value + "' at '" + where + "'");
    return (int)v;
}

bool Node::asBool() const {
    if (kind != Scalar) throw std::runtime_error("yaml: '" + where + "' is not a scalar");
    static const char* yes[] = {"true", "True", "TRUE", "yes", "Yes", "YES", "on", "On", "ON", "y", "Y"};
    static const char* no[] = {"false", "False", "FALSE", "no", "No", "NO", "off", "Off", "OFF", "n", "N"};
    for (const char* y : yes)
        if (value == y) return true;
    for (const char* n : no)
        if (value == n) return false;
    throw std::runtime_error("yaml: bad bool '" + value + "' at '" + where + "'");
}

std::string Node::asString() const {
    if (kind != Scalar) throw std::runtime_error("yaml: '" + where + "' is not a scalar");
    return value;
}

Node parseText(const std::string& text, const std::string& origin) {
    std::vector<Line> lines;
    std::istringstream in(text);
    std::string raw;
    int no = 0;
    while (std::getline(in, raw)) {
        ++no;
        std::string s = strip_comment(raw);
        std::string t = trim(s);
        if (t.empty() || t == "---") continue;
        int ind = 0;
        while (ind < (int)s.size() && s[ind] == ' ') ++ind;
        lines.push_back({ind, t, no});
    }
    size_t i = 0;
    if (lines.empty()) return Node();
    Node root = parse_block(lines, i, lines[0].indent, "");
    if (i != lines.size())
        throw std::runtime_error("yaml: unexpected indentation at line " + std::to_string(lines[i].lineno) + " of " +
                                 origin);
    return root;
}

Node loadFile(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("yaml: bad file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parseText(ss.str(), path);
}

}  // namespace yaml

MPM3DConfiguration parseYAML(const std::string& filePath) {
    yaml::Node root = yaml::loadFile(filePath);
    MPM3DConfiguration c;
    const yaml::Node& groups = root["ParticleGroups"];
    for (size_t i = 0; i < groups.size(); ++i) {
        const yaml::Node& g = groups[i];
        c.group_paths.push_back(g["File"].asString());
        c.masses.push_back(g["Mass"].asFloat());
        c.velocities.push_back(Vector3f(g["Velocity"][0].asFloat(), g["Velocity"][1].asFloat(), g["Velocity"][2].asFloat()));
        c.youngs.push_back(g["YoungModulus"].asFloat());
        c.poissons.push_back(g["PoissonRatio"].asFloat());
        c.hardenings.push_back(g["HardeningCoefficient"].asFloat());
        c.compressions.push_back(g["CriticalCompression"].asFloat());
        c.stretches.push_back(g["CriticalStretch"].asFloat());
        c.constitutive_models.push_back(parseConstitutiveModel(g["ConstitutiveModel"].asString()));
    }
    const yaml::Node& grid = root["Grid"];
    c.origin = Vector3f(grid["Origin"][0].asFloat(), grid["Origin"][1].asFloat(), grid["Origin"][2].asFloat());
    c.resolution = Vector3i(grid["Resolution"][0].asInt(), grid["Resolution"][1].asInt(), grid["Resolution"][2].asInt());
    c.stride = grid["Stride"].asFloat();
    c.friction_coeff = grid["BoundaryFrictionCoefficient"].asFloat();
    c.blend_coeff = root["Transfer"]["BlendCoefficient"].asFloat();
    if (root["Transfer"].has("Scheme")) {
        const std::string sc = root["Transfer"]["Scheme"].asString();
        if (sc == "APIC") c.scheme = 1;
        else if (sc == "PICFLIP") c.scheme = 0;
        else throw std::runtime_error("unknown Transfer.Scheme: " + sc);
    }
    c.interp_type = parseInterpolationType(root["Interpolation"]["Type"].asString());
    const yaml::Node& sim = root["Simulate"];
    c.delta_time = sim["deltaTime"].asFloat();
    if (sim.has("Gravity")) c.gravity = sim["Gravity"].asFloat();
    if (sim.has("Steps")) c.steps = sim["Steps"].asInt();
    c.offline = root["Render"]["Offline"].asBool();
    c.particle_frag_shader_path = root["Render"]["ParticleFragmentShader"].asString();
    return c;
}

}  // namespace chains

// JSON dump of a parsed configuration (used by the parser tests).
extern "C" __attribute__((visibility("default"))) int chains_parse_yaml_json(const char* path, char* buf, int len) {
    try {
        chains::MPM3DConfiguration c = chains::parseYAML(path);
        std::ostringstream o;
        o.precision(9);
        auto farr = [&](const std::vector<float>& v) {
            o << "[";
            for (size_t i = 0; i < v.size(); ++i) o << (i ? "," : "") << v[i];
            o << "]";
        };
        o << "{\"group_paths\":[";
        for (size_t i = 0; i < c.group_paths.size(); ++i) o << (i ? "," : "") << "\"" << c.group_paths[i] << "\"";
        o << "],\"velocities\":[";
        for (size_t i = 0; i < c.velocities.size(); ++i)
            o << (i ? "," : "") << "[" << c.velocities[i](0) << "," << c.velocities[i](1) << "," << c.velocities[i](2)
              << "]";
        o << "],\"masses\":"; farr(c.masses);
        o << ",\"youngs\":"; farr(c.youngs);
        o << ",\"poissons\":"; farr(c.poissons);
        o << ",\"hardenings\":"; farr(c.hardenings);
        o << ",\"compressions\":"; farr(c.compressions);
        o << ",\"stretches\":"; farr(c.stretches);
        o << ",\"constitutive_models\":[";
        for (size_t i = 0; i < c.constitutive_models.size(); ++i) o << (i ? "," : "") << (int)c.constitutive_models[i];
        o << "],\"origin\":[" << c.origin(0) << "," << c.origin(1) << "," << c.origin(2) << "]";
        o << ",\"resolution\":[" << c.resolution(0) << "," << c.resolution(1) << "," << c.resolution(2) << "]";
        o << ",\"stride\":" << c.stride << ",\"friction_coeff\":" << c.friction_coeff;
        o << ",\"blend_coeff\":" << c.blend_coeff << ",\"interp_type\":" << (int)c.interp_type;
        o << ",\"delta_time\":" << c.delta_time << ",\"offline\":" << (c.offline ? "true" : "false");
        o << ",\"particle_frag_shader_path\":\"" << c.particle_frag_shader_path << "\"";
        o << ",\"gravity\":" << c.gravity << ",\"steps\":" << c.steps << ",\"scheme\":" << c.scheme << "}";
        std::string s = o.str();
        if ((int)s.size() + 1 > len) return -(int)s.size() - 1;
        std::memcpy(buf, s.c_str(), s.size() + 1);
        return 0;
    } catch (const std::exception& e) {
        std::string s = std::string("ERROR: ") + e.what();
        if (len > 0) {
            std::strncpy(buf, s.c_str(), (size_t)len - 1);
            buf[len - 1] = '\0';
        }
        return 1;
    }
}
