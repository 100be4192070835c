// litho_io.cpp -- file formats and unit-mesh assembly (see litho_io.hpp).
#include "litho_io.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace lithofwd {

namespace {

void warn(std::string* w, const std::string& msg)
{
    if (w) {
        *w += msg;
        *w += '\n';
    }
}

std::string ltrim(const std::string& s)
{
    const size_t a = s.find_first_not_of(" \t\r");
    return a == std::string::npos ? std::string() : s.substr(a);
}

std::ifstream open_in(const std::string& path, const char* who)
{
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error(std::string(who) + ": cannot open " + path);
    return f;
}

void check_indices(const SurfaceMeshData& m, const std::string& path, const char* who)
{
    const uint32_t nv = (uint32_t)m.vertexCount();
    for (uint32_t t : m.triangles)
        if (t >= nv)
            throw std::runtime_error(std::string(who) + ": " + path +
                                     ": a triangle references vertex " + std::to_string(t + 1) +
                                     " of " + std::to_string(nv));
}

}  // namespace

// ------------------------------------------------------------------- surfaces

SurfaceMeshData loadTSurf(const std::string& path, std::string* warnings)
{
    std::ifstream f = open_in(path, "loadTSurf");
    SurfaceMeshData m;
    std::string raw;
    int lineNo = 0;
    bool inHeader = false;
    while (std::getline(f, raw)) {
        ++lineNo;
        const std::string line = ltrim(raw);
        if (line.empty() || line[0] == '#') continue;
        if (inHeader) {  // HEADER { ... } block, closed by a line starting with '}'
            if (line[0] == '}') inHeader = false;
            continue;
        }
        if (line.find("HEADER") != std::string::npos && line.find('{') != std::string::npos) {
            inHeader = true;
            continue;
        }
        if (line == "END") break;  // only the first object of a multi-object file
        if (line == "TFACE" || line == "TRIANGLES" || line.rfind("GOCAD", 0) == 0) continue;
        std::istringstream ss(line);
        std::string kw;
        ss >> kw;
        if (kw == "VRTX" || kw == "PVRTX") {
            long long idx;
            double x, y, z;
            if (!(ss >> idx >> x >> y >> z)) {
                warn(warnings, path + ":" + std::to_string(lineNo) + ": invalid vertex");
                continue;
            }
            // vertices are taken in file order; TRGL indices refer to that order
            m.vertices.insert(m.vertices.end(), {x, y, z});
        } else if (kw == "TRGL") {
            long long a, b, c;
            if (!(ss >> a >> b >> c)) {
                warn(warnings, path + ":" + std::to_string(lineNo) + ": invalid triangle");
                continue;
            }
            if (a < 1 || b < 1 || c < 1)
                throw std::runtime_error("loadTSurf: " + path + ":" + std::to_string(lineNo) +
                                         ": vertex index below 1");
            m.triangles.insert(m.triangles.end(),
                               {(uint32_t)(a - 1), (uint32_t)(b - 1), (uint32_t)(c - 1)});
        }
        // other keywords (ATOM, BSTONE, BORDER, PROPERTY_CLASS_HEADER, ...) are ignored
    }
    if (m.vertexCount() == 0) throw std::runtime_error("loadTSurf: no vertices in " + path);
    check_indices(m, path, "loadTSurf");
    return m;
}

SurfaceMeshData loadOBJ(const std::string& path, std::string* warnings)
{
    std::ifstream f = open_in(path, "loadOBJ");
    SurfaceMeshData m;
    std::string raw;
    int lineNo = 0;
    while (std::getline(f, raw)) {
        ++lineNo;
        const std::string line = ltrim(raw);
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string kw;
        ss >> kw;
        if (kw == "v") {
            double x, y, z;
            if (!(ss >> x >> y >> z)) {
                warn(warnings, path + ":" + std::to_string(lineNo) + ": invalid vertex");
                continue;
            }
            m.vertices.insert(m.vertices.end(), {x, y, z});
        } else if (kw == "f" || kw == "fo") {
            std::vector<long long> idx;
            std::string tok;
            bool bad = false;
            while (ss >> tok) {
                const size_t slash = tok.find('/');  // v/vt/vn: the vertex index
                if (slash != std::string::npos) tok = tok.substr(0, slash);
                char* end = nullptr;
                const long long v = std::strtoll(tok.c_str(), &end, 10);
                if (end == tok.c_str()) {
                    bad = true;
                    break;
                }
                idx.push_back(v);
            }
            if (bad || idx.size() < 3) {
                warn(warnings, path + ":" + std::to_string(lineNo) + ": invalid face");
                continue;
            }
            for (long long v : idx)
                if (v < 1)
                    throw std::runtime_error("loadOBJ: " + path + ":" + std::to_string(lineNo) +
                                             ": relative or zero vertex index");
            for (size_t i = 1; i + 1 < idx.size(); ++i)  // fan triangulation of polygons
                m.triangles.insert(m.triangles.end(),
                                   {(uint32_t)(idx[0] - 1), (uint32_t)(idx[i] - 1),
                                    (uint32_t)(idx[i + 1] - 1)});
        }
    }
    if (m.vertexCount() == 0) throw std::runtime_error("loadOBJ: no vertices in " + path);
    check_indices(m, path, "loadOBJ");
    return m;
}

SurfaceMeshData loadSurfaceMesh(const std::string& path, std::string* warnings)
{
    const size_t dot = path.rfind('.');
    std::string ext = dot == std::string::npos ? std::string() : path.substr(dot);
    for (auto& ch : ext) ch = (char)std::tolower((unsigned char)ch);
    if (ext == ".ts") return loadTSurf(path, warnings);
    if (ext == ".obj") return loadOBJ(path, warnings);
    throw std::runtime_error("loadSurfaceMesh: unrecognized extension '" + ext + "' of " + path);
}

void writeTSurf(const SurfaceMeshData& mesh, const std::string& path, const std::string& name)
{
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("writeTSurf: cannot write " + path);
    std::fprintf(f, "GOCAD TSurf 1\nHEADER {\n  name: %s\n}\nTFACE\n", name.c_str());
    const size_t nv = mesh.vertexCount();
    for (size_t i = 0; i < nv; ++i)
        std::fprintf(f, "VRTX %zu %.8e %.8e %.8e\n", i + 1, mesh.vertices[3 * i],
                     mesh.vertices[3 * i + 1], mesh.vertices[3 * i + 2]);
    const size_t nt = mesh.triangleCount();
    for (size_t i = 0; i < nt; ++i)
        std::fprintf(f, "TRGL %u %u %u\n", mesh.triangles[3 * i] + 1, mesh.triangles[3 * i + 1] + 1,
                     mesh.triangles[3 * i + 2] + 1);
    std::fprintf(f, "END\n");
    if (std::fclose(f) != 0) throw std::runtime_error("writeTSurf: error writing " + path);
}

// ------------------------------------------------------------------- stations

std::vector<StationRecord> loadStationCSV(const std::string& path, std::string* warnings)
{
    std::ifstream f = open_in(path, "loadStationCSV");
    std::vector<StationRecord> out;
    std::string line;
    int lineNo = 0;
    while (std::getline(f, line)) {
        ++lineNo;
        if (line.empty() || line[0] == '#') continue;
        std::vector<double> vals;
        bool skip = false;
        size_t pos = 0;
        while (pos <= line.size()) {
            size_t comma = line.find(',', pos);
            if (comma == std::string::npos) comma = line.size();
            std::string tok = line.substr(pos, comma - pos);
            pos = comma + 1;
            const size_t a = tok.find_first_not_of(" \t\r");
            if (a == std::string::npos) {  // empty field: ignored
                if (comma == line.size()) break;
                continue;
            }
            const size_t b = tok.find_last_not_of(" \t\r");
            tok = tok.substr(a, b - a + 1);
            // a first field starting with a letter marks a header line ("x", "X", ...)
            if (vals.empty() && std::isalpha((unsigned char)tok[0])) {
                skip = true;
                break;
            }
            // leading numeric prefix, as std::stod reads it
            char* end = nullptr;
            errno = 0;
            const double v = std::strtod(tok.c_str(), &end);
            if (end == tok.c_str() || errno == ERANGE) {
                warn(warnings, path + ":" + std::to_string(lineNo) + ": '" + tok +
                                   "' is not a number");
                skip = true;
                break;
            }
            vals.push_back(v);
            if (comma == line.size()) break;
        }
        if (skip) continue;
        if (vals.size() < 4) {
            warn(warnings, path + ":" + std::to_string(lineNo) +
                               ": expected at least 4 values (x, y, z, observation)");
            continue;
        }
        out.push_back({vals[0], vals[1], vals[2], vals[3], vals.size() >= 5 ? vals[4] : 0.0});
    }
    return out;
}

std::vector<ConstraintRecord> loadConstraintCSV(const std::string& path, std::string* warnings)
{
    std::ifstream f = open_in(path, "loadConstraintCSV");
    std::vector<ConstraintRecord> out;
    std::string line;
    int lineNo = 0;
    while (std::getline(f, line)) {
        ++lineNo;
        if (line.empty() || line[0] == '#') continue;
        std::vector<double> vals;
        bool skip = false;
        size_t pos = 0;
        while (pos <= line.size()) {
            size_t comma = line.find(',', pos);
            if (comma == std::string::npos) comma = line.size();
            std::string tok = line.substr(pos, comma - pos);
            pos = comma + 1;
            const size_t a = tok.find_first_not_of(" \t\r");
            if (a != std::string::npos) {
                const size_t b = tok.find_last_not_of(" \t\r");
                tok = tok.substr(a, b - a + 1);
                char* end = nullptr;
                errno = 0;
                const double v = std::strtod(tok.c_str(), &end);
                if (end == tok.c_str() || errno == ERANGE) {  // a header line lands here too
                    warn(warnings, path + ":" + std::to_string(lineNo) + ": '" + tok +
                                       "' is not a number");
                    skip = true;
                    break;
                }
                vals.push_back(v);
            }
            if (comma == line.size()) break;
        }
        if (skip) continue;
        if (vals.size() < 5) {
            warn(warnings, path + ":" + std::to_string(lineNo) +
                               ": expected 5 values (x, y, z_top, z_bottom, unit)");
            continue;
        }
        out.push_back({vals[0], vals[1], vals[2], vals[3], (int)vals[4]});
    }
    return out;
}

// ----------------------------------------------------------------- label grid

LabelGridData loadLabelGrid(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("loadLabelGrid: cannot open " + path);
    std::string header;
    if (!std::getline(f, header)) throw std::runtime_error("loadLabelGrid: no header in " + path);
    LabelGridData g;
    std::istringstream hs(header);
    if (!(hs >> g.nx >> g.ny >> g.nz >> g.x0 >> g.y0 >> g.z0 >> g.dx >> g.dy >> g.dz))
        throw std::runtime_error("loadLabelGrid: malformed header in " + path);
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0)
        throw std::runtime_error("loadLabelGrid: non-positive grid size in " + path);
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    g.labels.resize(n);
    f.read(reinterpret_cast<char*>(g.labels.data()), (std::streamsize)(n * sizeof(int32_t)));
    if ((size_t)f.gcount() != n * sizeof(int32_t))
        throw std::runtime_error("loadLabelGrid: " + path + " holds fewer labels than its header");
    return g;
}

void writeLabelGrid(const LabelGridData& g, const std::string& path)
{
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    if (g.labels.size() != n) throw std::runtime_error("writeLabelGrid: label count != nx*ny*nz");
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("writeLabelGrid: cannot write " + path);
    std::ostringstream hs;
    hs.precision(17);
    hs << g.nx << ' ' << g.ny << ' ' << g.nz << ' ' << g.x0 << ' ' << g.y0 << ' ' << g.z0 << ' '
       << g.dx << ' ' << g.dy << ' ' << g.dz << '\n';
    f << hs.str();
    f.write(reinterpret_cast<const char*>(g.labels.data()), (std::streamsize)(n * sizeof(int32_t)));
    if (!f) throw std::runtime_error("writeLabelGrid: error writing " + path);
}

// ------------------------------------------------------------------- assembly

namespace {
struct Key3Hash {
    size_t operator()(const std::array<int64_t, 3>& k) const
    {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (int64_t v : k) {
            uint64_t z = (uint64_t)v + h;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            h = z ^ (z >> 31);
        }
        return (size_t)h;
    }
};
}  // namespace

AssembledModel assembleUnitMeshes(const std::vector<SurfaceMeshData>& meshes)
{
    AssembledModel M;
    std::unordered_map<std::array<int64_t, 3>, uint32_t, Key3Hash> vmap;
    // triangle key (rotated so the smallest vertex is first, the other two sorted) ->
    // facets not yet paired with a reversed copy
    std::unordered_map<std::array<int64_t, 3>, std::vector<std::pair<int64_t, int>>, Key3Hash>
        open;
    for (size_t u = 0; u < meshes.size(); ++u) {
        const SurfaceMeshData& mesh = meshes[u];
        const size_t nv = mesh.vertexCount();
        std::vector<uint32_t> local(nv);
        for (size_t i = 0; i < nv; ++i) {
            const double* p = &mesh.vertices[3 * i];
            const std::array<int64_t, 3> q = {(int64_t)std::llround(p[0] * 1e6),
                                              (int64_t)std::llround(p[1] * 1e6),
                                              (int64_t)std::llround(p[2] * 1e6)};
            auto it = vmap.find(q);
            if (it == vmap.end()) {
                const uint32_t id = (uint32_t)(M.vertices.size() / 3);
                vmap.emplace(q, id);
                M.vertices.insert(M.vertices.end(), {p[0], p[1], p[2]});
                local[i] = id;
            } else {
                local[i] = it->second;
                ++M.mergedVertices;
            }
        }
        const size_t nt = mesh.triangleCount();
        for (size_t t = 0; t < nt; ++t) {
            const uint32_t a = local[mesh.triangles[3 * t]], b = local[mesh.triangles[3 * t + 1]],
                           c = local[mesh.triangles[3 * t + 2]];
            const int64_t f = (int64_t)M.front.size();
            M.faces.insert(M.faces.end(), {a, b, c});
            M.front.push_back((int32_t)u);
            M.back.push_back(-1);
            if (a == b || b == c || a == c) continue;  // degenerate: no orientation
            // rotate the smallest index to the front; the order of the other two is the
            // winding
            uint32_t r0 = a, r1 = b, r2 = c;
            if (b < a && b < c) {
                r0 = b; r1 = c; r2 = a;
            } else if (c < a && c < b) {
                r0 = c; r1 = a; r2 = b;
            }
            const int parity = r1 < r2 ? 1 : -1;
            const std::array<int64_t, 3> key = {r0, std::min(r1, r2), std::max(r1, r2)};
            auto& lst = open[key];
            bool paired = false;
            for (size_t k = 0; k < lst.size(); ++k) {
                const int64_t g = lst[k].first;
                if (lst[k].second == -parity && M.front[g] != (int32_t)u) {
                    M.back[g] = (int32_t)u;  // the reversed copy: facet g's other side
                    M.faces.resize(M.faces.size() - 3);
                    M.front.pop_back();
                    M.back.pop_back();
                    lst.erase(lst.begin() + (std::ptrdiff_t)k);
                    ++M.contacts;
                    paired = true;
                    break;
                }
            }
            if (!paired) lst.push_back({f, parity});
        }
    }
    return M;
}

// ---------------------------------------------------------------- ini files

namespace {
std::string trim_ws(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}
}  // namespace

bool IniConfig::load(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open()) return false;
    m_data.clear();
    parse(f);
    return true;
}

void IniConfig::parse(std::istream& in)
{
    std::string line, section;
    while (std::getline(in, line)) {
        const std::string t = trim_ws(line);
        if (t.empty() || t[0] == '#' || t[0] == ';') continue;
        if (t[0] == '[') {
            const size_t e = t.find(']');
            if (e == std::string::npos) continue;
            section = trim_ws(t.substr(1, e - 1));
            m_data[section];
            continue;
        }
        size_t sep = t.find('=');
        if (sep == std::string::npos) sep = t.find(':');
        if (sep == std::string::npos) continue;
        const std::string key = trim_ws(t.substr(0, sep));
        if (!key.empty() && !section.empty()) m_data[section][key] = trim_ws(t.substr(sep + 1));
    }
}

void IniConfig::save(const std::string& path) const
{
    std::ofstream f(path);
    if (!f.is_open()) throw std::runtime_error("IniConfig::save: cannot write " + path);
    for (const auto& sec : m_data) {
        f << "[" << sec.first << "]\n";
        for (const auto& kv : sec.second) f << kv.first << " = " << kv.second << "\n";
        f << "\n";
    }
}

bool IniConfig::hasSection(const std::string& section) const
{
    return m_data.count(section) != 0;
}

std::vector<std::string> IniConfig::sections() const
{
    std::vector<std::string> out;
    for (const auto& sec : m_data) out.push_back(sec.first);
    return out;
}

bool IniConfig::has(const std::string& section, const std::string& key) const
{
    const auto it = m_data.find(section);
    return it != m_data.end() && it->second.count(key) != 0;
}

std::string IniConfig::getString(const std::string& section, const std::string& key,
                                 const std::string& def) const
{
    const auto it = m_data.find(section);
    if (it == m_data.end()) return def;
    const auto kv = it->second.find(key);
    return kv == it->second.end() ? def : kv->second;
}

double IniConfig::getDouble(const std::string& section, const std::string& key, double def) const
{
    const std::string v = getString(section, key);
    if (v.empty()) return def;
    try {
        return std::stod(v);
    } catch (...) {
        return def;
    }
}

int IniConfig::getInt(const std::string& section, const std::string& key, int def) const
{
    const std::string v = getString(section, key);
    if (v.empty()) return def;
    try {
        return std::stoi(v);
    } catch (...) {
        return def;
    }
}

bool IniConfig::getBool(const std::string& section, const std::string& key, bool def) const
{
    std::string v = getString(section, key);
    if (v.empty()) return def;
    for (auto& c : v) c = (char)std::tolower((unsigned char)c);
    return v == "true" || v == "yes" || v == "1" || v == "on";
}

void IniConfig::setString(const std::string& section, const std::string& key,
                          const std::string& value)
{
    m_data[section][key] = value;
}

}  // namespace lithofwd
