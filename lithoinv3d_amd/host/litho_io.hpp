// litho_io.hpp -- the file formats LithoInv3D reads and writes around the forward path:
// GOCAD TSurf and OBJ unit surfaces, CSV gravity / magnetic stations, the binary label
// grid, and the assembly of per-unit closed meshes (each contact carried twice, once
// per side) into one shared-vertex facet set with (front, back) units, which is the
// layout the GPU forward consumes.
#pragma once
#include <cstdint>
#include <iosfwd>
#include <map>
#include <string>
#include <vector>

namespace lithofwd {

// One unit's surface: xyz per vertex, 0-based vertex triples per triangle.
struct SurfaceMeshData {
    std::vector<double> vertices;
    std::vector<uint32_t> triangles;
    size_t vertexCount() const { return vertices.size() / 3; }
    size_t triangleCount() const { return triangles.size() / 3; }
};

// modules/litho-io/src/io/surface_loader.cpp:93-148 (TSurfaceLoader::load)
SurfaceMeshData loadTSurf(const std::string& path, std::string* warnings = nullptr);
// surface_loader.cpp:9-91 (OBJSurfaceLoader::load)
SurfaceMeshData loadOBJ(const std::string& path, std::string* warnings = nullptr);
// surface_loader.cpp:150-168 (loadSurfaceMesh: by extension, .ts or .obj)
SurfaceMeshData loadSurfaceMesh(const std::string& path, std::string* warnings = nullptr);
// modules/litho-io/src/io/exporters.cpp:116-150 (InversionExporter::exportTS)
void writeTSurf(const SurfaceMeshData& mesh, const std::string& path, const std::string& name);

// One station of a gravity (mGal) or magnetic (nT) CSV file; stdev 0 = unweighted.
struct StationRecord {
    double x, y, z, obs, stdev;
};
// modules/litho-io/src/io/gravity_loader.cpp:8-69 and magnetic_loader.cpp:8-69 (the two
// loaders parse identically: x, y, z, observation[, standard deviation])
std::vector<StationRecord> loadStationCSV(const std::string& path, std::string* warnings = nullptr);

// One borehole interval: unit `unit` between depths z_top and z_bottom (positive down)
// at collar (x, y).
struct ConstraintRecord {
    double x, y, z_top, z_bottom;
    int unit;
};
// modules/litho-io/src/io/constraint_loader.cpp:8-66: x, y, z_top, z_bottom, unit per
// line; '#' comments; a line with a non-numeric field or fewer than 5 values is skipped.
std::vector<ConstraintRecord> loadConstraintCSV(const std::string& path,
                                                std::string* warnings = nullptr);

// The label grid written by python/lithoseed/src/lithoseed/_surfaces.py:659-671: an
// ASCII header "nx ny nz x0 y0 z0 dx dy dz\n" then nx*ny*nz int32 labels with i fastest
// (index i + nx*(j + ny*k)).
struct LabelGridData {
    int nx = 0, ny = 0, nz = 0;
    double x0 = 0, y0 = 0, z0 = 0, dx = 0, dy = 0, dz = 0;
    std::vector<int32_t> labels;
};
LabelGridData loadLabelGrid(const std::string& path);
void writeLabelGrid(const LabelGridData& grid, const std::string& path);

// Per-unit meshes -> one facet set.  Vertices closer than the 1 um quantization of
// lithology_model.cpp:113-168 (rebuildGlobalDofs) are one shared vertex (the first
// occurrence's coordinates are kept).  A triangle of unit A whose reversed copy appears in
// unit B becomes one facet with front = A (A's winding) and back = B; any other triangle
// is a facet of its own unit with back = -1.
struct AssembledModel {
    std::vector<double> vertices;
    std::vector<uint32_t> faces;
    std::vector<int32_t> front, back;
    int64_t contacts = 0;         // facets with a back unit
    int64_t mergedVertices = 0;   // input vertices that mapped onto an earlier one
};
AssembledModel assembleUnitMeshes(const std::vector<SurfaceMeshData>& meshes);

// The run configuration file of modules/litho-io/src/io/ini_config.cpp: "[section]"
// headers, "key = value" or "key: value" lines (the first '=' splits, else the first ':'),
// both sides trimmed, '#' and ';' comment lines, blank lines and lines before the first
// section ignored, a repeated key keeps its last value.  The typed getters return the
// default for a missing or empty value; numbers parse their leading part as std::stod /
// std::stoi do (default when none), booleans are true for "true", "yes", "1", "on" in any
// case and false for anything else.
class IniConfig {
public:
    bool load(const std::string& path);  // false when the file cannot be opened
    void parse(std::istream& in);
    void save(const std::string& path) const;  // sections and keys in sorted order
    bool hasSection(const std::string& section) const;
    std::vector<std::string> sections() const;
    const std::map<std::string, std::map<std::string, std::string>>& data() const { return m_data; }
    bool has(const std::string& section, const std::string& key) const;
    std::string getString(const std::string& section, const std::string& key,
                          const std::string& def = "") const;
    double getDouble(const std::string& section, const std::string& key, double def) const;
    int getInt(const std::string& section, const std::string& key, int def) const;
    bool getBool(const std::string& section, const std::string& key, bool def) const;
    void setString(const std::string& section, const std::string& key, const std::string& value);

private:
    std::map<std::string, std::map<std::string, std::string>> m_data;
};

}  // namespace lithofwd
