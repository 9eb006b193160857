// rh_vlsv.hpp -- host-side writer of VLSV files (the format of the
// reference's output, src/operator_userdata.cpp:55-311, as the reference's
// reader tools/rhybridpy.py:113-126 reads it): a 16-byte header whose second
// 8 bytes hold the byte offset of an XML footer, the raw little-endian arrays,
// then the footer <VLSV>...</VLSV> with one element per array whose text is
// the array's byte offset and whose attributes give name, mesh, arraysize,
// vectorsize, datatype (float / int / uint) and datasize (bytes).
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace rh {

class VlsvWriter {
public:
    explicit VlsvWriter(const std::string& path) : path_(path) {
        f_ = std::fopen(path.c_str(), "wb");
        if (!f_) throw std::runtime_error("cannot open " + path);
        const uint64_t zero[2] = {0, 0};
        put(zero, sizeof(zero));
        offset_ = sizeof(zero);
    }
    ~VlsvWriter() {
        if (f_) std::fclose(f_);
    }
    VlsvWriter(const VlsvWriter&) = delete;
    VlsvWriter& operator=(const VlsvWriter&) = delete;

    // One array of arraysize x vectorsize elements of `datasize` bytes.
    void array(const std::string& tag, const std::map<std::string, std::string>& attribs, const void* data,
               uint64_t arraysize, uint64_t vectorsize, const char* datatype, uint64_t datasize) {
        std::ostringstream e;
        e << "<" << tag;
        for (const auto& kv : attribs) e << " " << kv.first << "=\"" << escape(kv.second) << "\"";
        e << " arraysize=\"" << arraysize << "\" vectorsize=\"" << vectorsize << "\" datatype=\"" << datatype
          << "\" datasize=\"" << datasize << "\">" << offset_ << "</" << tag << ">\n";
        footer_ += e.str();
        const uint64_t bytes = arraysize * vectorsize * datasize;
        if (bytes) put(data, bytes);
        offset_ += bytes;
    }

    void close() {
        if (!f_) return;
        const std::string xml = "<VLSV>\n" + footer_ + "</VLSV>\n";
        const uint64_t at = offset_;
        put(xml.data(), xml.size());
        if (std::fseek(f_, 8, SEEK_SET) != 0) throw std::runtime_error("seek failed in " + path_);
        put(&at, sizeof(at));
        if (std::fclose(f_) != 0) { f_ = nullptr; throw std::runtime_error("close failed for " + path_); }
        f_ = nullptr;
    }

private:
    void put(const void* p, size_t n) {
        if (std::fwrite(p, 1, n, f_) != n) throw std::runtime_error("write failed for " + path_);
    }
    static std::string escape(const std::string& s) {
        std::string o;
        for (char c : s) {
            switch (c) {
            case '&': o += "&amp;"; break;
            case '<': o += "&lt;"; break;
            case '>': o += "&gt;"; break;
            case '"': o += "&quot;"; break;
            default: o += c;
            }
        }
        return o;
    }
    std::string path_;
    std::FILE* f_ = nullptr;
    uint64_t offset_ = 0;
    std::string footer_;
};

// The grid of a file: nx*ny*nz cubic cells of edge dx, lower corner x0.
struct VlsvGrid {
    int nx, ny, nz;
    double dx, x0, y0, z0;
};

// What precedes the variables of every file: the parameters the reference's
// reader takes (time and its alias t, timestep, dt, the box extent and cell
// counts), the mesh "SpatialGrid" (MESH_BBOX = cells per axis and a block
// of one cell, node coordinates along each axis, the MESH element) and the
// CellID variable (the cell index n = i + nx*(j + ny*k)).
inline void vlsv_write_grid(VlsvWriter& W, const VlsvGrid& g, double t, int64_t timestep, double dt) {
    auto param = [&](const char* name, const void* v, const char* type, uint64_t size) {
        W.array("PARAMETER", {{"name", name}}, v, 1, 1, type, size);
    };
    const double xmax = g.x0 + g.nx * g.dx, ymax = g.y0 + g.ny * g.dx, zmax = g.z0 + g.nz * g.dx;
    const uint64_t ts = (uint64_t)timestep, cx = (uint64_t)g.nx, cy = (uint64_t)g.ny, cz = (uint64_t)g.nz;
    param("time", &t, "float", 8);
    param("t", &t, "float", 8);
    param("timestep", &ts, "uint", 8);
    param("dt", &dt, "float", 8);
    param("xmin", &g.x0, "float", 8); param("ymin", &g.y0, "float", 8); param("zmin", &g.z0, "float", 8);
    param("xmax", &xmax, "float", 8); param("ymax", &ymax, "float", 8); param("zmax", &zmax, "float", 8);
    param("xcells_ini", &cx, "uint", 8); param("ycells_ini", &cy, "uint", 8); param("zcells_ini", &cz, "uint", 8);
    const std::map<std::string, std::string> mesh{{"mesh", "SpatialGrid"}};
    const uint64_t bbox[6] = {cx, cy, cz, 1, 1, 1};
    W.array("MESH_BBOX", mesh, bbox, 6, 1, "uint", 8);
    const double o3[3] = {g.x0, g.y0, g.z0};
    const int dims[3] = {g.nx, g.ny, g.nz};
    const char* crd_tags[3] = {"MESH_NODE_CRDS_X", "MESH_NODE_CRDS_Y", "MESH_NODE_CRDS_Z"};
    for (int d = 0; d < 3; ++d) {
        std::vector<double> c((size_t)dims[d] + 1);
        for (int q = 0; q <= dims[d]; ++q) c[q] = o3[d] + q * g.dx;
        W.array(crd_tags[d], mesh, c.data(), c.size(), 1, "float", 8);
    }
    const uint64_t nc = (uint64_t)g.nx * g.ny * g.nz;
    std::vector<uint64_t> ids(nc);
    for (uint64_t n = 0; n < nc; ++n) ids[n] = n;
    W.array("MESH", {{"name", "SpatialGrid"}, {"type", "multi_ucd"}}, ids.data(), nc, 1, "uint", 8);
    W.array("VARIABLE", {{"name", "CellID"}, {"mesh", "SpatialGrid"}, {"type", "celldata"}}, ids.data(), nc, 1,
            "uint", 8);
}

// One cell variable of the mesh (nc x vec elements).
inline void vlsv_write_cell_var(VlsvWriter& W, uint64_t nc, const std::string& name, const void* data, uint64_t vec,
                                const char* type, uint64_t size) {
    W.array("VARIABLE", {{"name", name}, {"mesh", "SpatialGrid"}, {"type", "celldata"}}, data, nc, vec, type, size);
}

}  // namespace rh
