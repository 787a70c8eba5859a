// io.hpp -- CSV output, output directories, wall-clock timing.
//
// Interface and file format of src/utils/io.hpp:14-57: a Csv file starts with its comment
// lines ("# " + text) and one header line, unless it is opened for appending and already
// exists; every row is its values joined by commas, each printed with 10 significant digits
// (the "%.10g" form), and is flushed as it is written, so a run that stops keeps every
// finished row. write_panels prints one row per panel (centroid, normal, area, tag, u, f).
#pragma once
#include <chrono>
#include <cstdio>
#include <filesystem>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mesh.hpp"

namespace bem {

inline void ensure_dir(const std::string& dir) { std::filesystem::create_directories(dir); }

class Csv {
public:
    Csv(const std::string& path, const std::string& header, const std::vector<std::string>& comments = {},
        bool append = false) {
        const std::filesystem::path where(path);
        if (where.has_parent_path()) ensure_dir(where.parent_path().string());
        const bool continuing = append && std::filesystem::exists(where);
        file_.reset(std::fopen(path.c_str(), append ? "a" : "w"));
        if (!file_) return;  // rows are dropped, as a failed stream would drop them
        if (!continuing) {
            for (const std::string& c : comments) std::fprintf(file_.get(), "# %s\n", c.c_str());
            std::fprintf(file_.get(), "%s\n", header.c_str());
        }
        std::fflush(file_.get());
    }
    void row(std::initializer_list<double> values) { write(values.begin(), values.size()); }
    void row(const std::vector<double>& values) { write(values.data(), values.size()); }

private:
    struct Closer {
        void operator()(std::FILE* f) const { std::fclose(f); }
    };
    std::unique_ptr<std::FILE, Closer> file_;

    void write(const double* v, size_t n) {
        if (!file_) return;
        std::string line;
        char num[40];
        for (size_t i = 0; i < n; ++i) {
            std::snprintf(num, sizeof num, "%.10g", v[i]);
            if (i) line += ',';
            line += num;
        }
        line += '\n';
        std::fputs(line.c_str(), file_.get());
        std::fflush(file_.get());
    }
};

// one row per panel: geometry, tag, velocity and traction
inline void write_panels(const std::string& path, const Mesh& M, const std::vector<V3>& u,
                         const std::vector<V3>& f) {
    Csv out(path, "cx,cy,cz,nx,ny,nz,area,tag,ux,uy,uz,fx,fy,fz");
    for (int p = 0; p < M.P(); ++p) {
        const V3 &c = M.centroid[p], &n = M.normal[p];
        out.row({c.x, c.y, c.z, n.x, n.y, n.z, M.area[p], double(M.tags[p]), u[p].x, u[p].y, u[p].z, f[p].x, f[p].y,
                 f[p].z});
    }
}

// wall-clock seconds since construction
struct Timer {
    using clock = std::chrono::steady_clock;
    clock::time_point t0 = clock::now();
    double seconds() const { return std::chrono::duration<double>(clock::now() - t0).count(); }
};

}  // namespace bem
