// mesh.hpp -- flat-triangle surface meshes of the closed tube and the icosphere swimmer.
//
// API of src/utils/mesh.hpp:45-247: Tag, Mesh (geometry/orient/translate/total_area),
// merge, uniform_z, graded_z, tube_mesh, swimmer_tube_mesh, icosphere. Node order, panel
// order and vertex order within each panel follow the reference exactly, so connectivity
// and panel indexing are bit-identical; the device solver relies on that order only
// through these arrays.
//
// Conventions (src/utils/mesh.hpp:1-9): panel normals point out of the fluid (tube wall
// radially outward, inlet cap -z, outlet cap +z, sphere toward its centre).
// Tags: 0 wall, 1 inlet (z = z_min), 2 outlet (z = z_max), 10 sphere.
//
// Floating-point note: libsqbem compiles this header with contraction off, so geometry
// is reproducible across hosts; a build with FMA contraction differs at rounding level.
#pragma once
#include <algorithm>
#include <array>
#include <functional>
#include <cmath>
#include <map>
#include <utility>
#include <vector>
#include "vec3.hpp"

namespace bem {

enum Tag { WALL = 0, INLET = 1, OUTLET = 2, SPHERE = 10 };

namespace detail {
// cos and sin of one angle through glibc's sincos, which the reference build also calls for
// each cos/sin pair of the polygon vertices (its compiler fuses the two calls); separate
// calls can differ in the last bit
inline void cos_sin(double x, double& c, double& s) { ::sincos(x, &s, &c); }
}  // namespace detail

struct Mesh {
    std::vector<V3> nodes;
    std::vector<std::array<int, 3>> tris;
    std::vector<int> tags;
    // per-panel geometry, filled by geometry()
    std::vector<std::array<V3, 3>> V;
    std::vector<V3> normal, centroid;
    std::vector<double> area, h;  // h = sqrt(2 area)

    int P() const { return (int)tris.size(); }

    // vertices, unit normal (right-hand rule over v0 v1 v2), centroid, area, h
    void geometry() {
        const int n = P();
        V.assign(n, {});
        normal.assign(n, {});
        centroid.assign(n, {});
        area.assign(n, 0.0);
        h.assign(n, 0.0);
        for (int p = 0; p < n; ++p) panel_geometry(p);
    }
    void panel_geometry(int p) {
        const auto& t = tris[p];
        V[p] = {nodes[t[0]], nodes[t[1]], nodes[t[2]]};
        const V3 twice = cross(V[p][1] - V[p][0], V[p][2] - V[p][0]);
        const double A = 0.5 * norm(twice);
        area[p] = A;
        normal[p] = (1.0 / (2 * A)) * twice;
        centroid[p] = (1.0 / 3) * (V[p][0] + V[p][1] + V[p][2]);
        h[p] = std::sqrt(2 * A);
    }

    // reverse the panels whose normal points against outward(centroid); the geometry of the
    // reversed panels is formed again (the same values a full pass would give)
    void orient(const std::function<V3(V3)>& outward) { orient_with(outward); }
    template <typename F>
    void orient_with(const F& outward) {
        geometry();
        for (int p = 0; p < P(); ++p)
            if (dot(normal[p], outward(centroid[p])) < 0) {
                std::swap(tris[p][1], tris[p][2]);
                panel_geometry(p);
            }
    }

    void translate(V3 d) {
        for (V3& x : nodes) x = x + d;
        geometry();
    }

    double total_area() const {
        double s = 0;
        for (double A : area) s += A;
        return s;
    }
};

// b appended after a (node indices of b shifted), geometry formed for the result
inline Mesh merge(const Mesh& a, const Mesh& b) {
    Mesh m = a;
    const int shift = (int)a.nodes.size();
    m.nodes.insert(m.nodes.end(), b.nodes.begin(), b.nodes.end());
    for (const auto& t : b.tris) m.tris.push_back({t[0] + shift, t[1] + shift, t[2] + shift});
    m.tags.insert(m.tags.end(), b.tags.begin(), b.tags.end());
    m.geometry();
    return m;
}

namespace detail {
// merge() of two parts whose geometry is formed and current: their panels' geometry is
// carried over (the values a new pass over the same node coordinates would give)
inline Mesh merge_formed(Mesh a, const Mesh& b) {
    Mesh m = std::move(a);
    const int shift = (int)m.nodes.size();
    m.nodes.reserve(m.nodes.size() + b.nodes.size());
    m.tris.reserve(m.tris.size() + b.tris.size());
    m.nodes.insert(m.nodes.end(), b.nodes.begin(), b.nodes.end());
    for (const auto& t : b.tris) m.tris.push_back({t[0] + shift, t[1] + shift, t[2] + shift});
    m.tags.insert(m.tags.end(), b.tags.begin(), b.tags.end());
    m.V.insert(m.V.end(), b.V.begin(), b.V.end());
    m.normal.insert(m.normal.end(), b.normal.begin(), b.normal.end());
    m.centroid.insert(m.centroid.end(), b.centroid.begin(), b.centroid.end());
    m.area.insert(m.area.end(), b.area.begin(), b.area.end());
    m.h.insert(m.h.end(), b.h.begin(), b.h.end());
    return m;
}
}  // namespace detail

// nz + 1 equally spaced axial nodes on [z0, z1]
inline std::vector<double> uniform_z(double z0, double z1, int nz) {
    std::vector<double> z;
    z.reserve(nz + 1);
    for (int k = 0; k <= nz; ++k) z.push_back(z0 + (z1 - z0) * k / nz);
    return z;
}

// Axial nodes on [-L/2, L/2], symmetric about 0: spacing dzmin while |z| < core, then
// multiplied by `growth` per panel up to dzmax; a last gap under 0.3 dz is merged.
inline std::vector<double> graded_z(double L, double dzmin, double dzmax, double core,
                                    double growth = 1.25) {
    const double end = L / 2;
    std::vector<double> pos{0.0};
    double dz = dzmin;
    while (pos.back() < end - 1e-12) {
        if (pos.back() >= core) dz = std::min(dz * growth, dzmax);
        pos.push_back(std::min(pos.back() + dz, end));
    }
    if (pos.size() > 2 && end - pos[pos.size() - 2] < 0.3 * dz) pos.erase(pos.end() - 2);
    std::vector<double> z;
    z.reserve(2 * pos.size() - 1);
    for (auto it = pos.rbegin(); it != pos.rend(); ++it) z.push_back(-*it);
    z.insert(z.end(), pos.begin() + 1, pos.end());
    return z;
}

namespace detail {
// one end cap: centre node, then nr rings of nth nodes; centre fan then ring quads
inline Mesh tube_cap(double z0, int tag, double sg, int nth, int nr, double Rp, double al) {
    Mesh m;
    m.nodes.push_back(V3{0, 0, z0});
    for (int k = 1; k <= nr; ++k) {
        const double rk = Rp * k / nr;
        for (int i = 0; i < nth; ++i) {
            double c, s;
            cos_sin(i * al, c, s);
            m.nodes.push_back(V3{rk * c, rk * s, z0});
        }
    }
    auto node = [nth](int ring, int i) { return 1 + ring * nth + i % nth; };
    for (int i = 0; i < nth; ++i) m.tris.push_back({0, node(0, i), node(0, i + 1)});
    for (int k = 0; k + 1 < nr; ++k)
        for (int i = 0; i < nth; ++i) {
            const int a = node(k, i), b = node(k, i + 1), c = node(k + 1, i + 1), d = node(k + 1, i);
            m.tris.push_back({a, d, c});
            m.tris.push_back({a, c, b});
        }
    m.tags.assign(m.tris.size(), tag);
    m.orient_with([sg](V3) { return V3{0, 0, sg}; });
    return m;
}
}  // namespace detail

// Closed tube about the z axis through the axial nodes z, nth panels around, nr rings per
// cap. Polygon vertices sit at Rp = R sqrt(g / sin g), g = 2 pi / nth, which gives the
// faceted cross-section the exact circle area. Order: wall, inlet cap, outlet cap.
inline Mesh tube_mesh(double R, const std::vector<double>& z, int nth, int nr) {
    const double al = 2 * PI / nth;
    const double Rp = R * std::sqrt(al / std::sin(al));
    const int nz = (int)z.size() - 1;
    Mesh wall;
    for (int i = 0; i < nth; ++i) {
        double c, s;
        detail::cos_sin(i * al, c, s);
        for (int k = 0; k <= nz; ++k) wall.nodes.push_back(V3{Rp * c, Rp * s, z[k]});
    }
    auto node = [nth, nz](int i, int k) { return (i % nth) * (nz + 1) + k; };
    for (int i = 0; i < nth; ++i)
        for (int k = 0; k < nz; ++k) {
            const int a = node(i, k), b = node(i + 1, k), c = node(i + 1, k + 1), d = node(i, k + 1);
            wall.tris.push_back({a, b, c});
            wall.tris.push_back({a, c, d});
        }
    wall.tags.assign(wall.tris.size(), WALL);
    wall.orient_with([](V3 x) { return V3{x.x, x.y, 0}; });
    Mesh inlet = detail::tube_cap(z.front(), INLET, -1.0, nth, nr, Rp, al);
    Mesh outlet = detail::tube_cap(z.back(), OUTLET, 1.0, nth, nr, Rp, al);
    return detail::merge_formed(detail::merge_formed(std::move(wall), inlet), outlet);
}

// swimmer tube: length L centred on z = 0, graded finer within |z| < a + R/2
inline Mesh swimmer_tube_mesh(double R, double L, double a, int nth, int nr = 4) {
    const double h = 2 * PI * R / nth;
    return tube_mesh(R, graded_z(L, 0.6 * h, 1.4 * h, a + 0.5 * R), nth, nr);
}

// Icosahedron subdivided `level` times (20 * 4^level panels), projected to the unit
// sphere, then scaled to the volume of a sphere of radius a. Midpoints are numbered in
// order of first use. Normals point to the centre.
inline Mesh icosphere(double a, int level, V3 center = {}, int tag = SPHERE) {
    const double g = (1 + std::sqrt(5.0)) / 2;
    std::vector<V3> v = {{-1, g, 0}, {1, g, 0}, {-1, -g, 0}, {1, -g, 0}, {0, -1, g}, {0, 1, g},
                         {0, -1, -g}, {0, 1, -g}, {g, 0, -1}, {g, 0, 1}, {-g, 0, -1}, {-g, 0, 1}};
    for (V3& x : v) x = unit(x);
    std::vector<std::array<int, 3>> F = {
        {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<int, int>, int> midpoint;
        auto mid = [&](int i, int j) {
            const auto key = std::minmax(i, j);
            const auto found = midpoint.find(key);
            if (found != midpoint.end()) return found->second;
            v.push_back(unit(v[i] + v[j]));
            const int id = (int)v.size() - 1;
            midpoint.emplace(key, id);
            return id;
        };
        std::vector<std::array<int, 3>> next;
        next.reserve(4 * F.size());
        for (const auto& f : F) {
            const int ab = mid(f[0], f[1]);
            const int bc = mid(f[1], f[2]);
            const int ca = mid(f[2], f[0]);
            next.push_back({f[0], ab, ca});
            next.push_back({f[1], bc, ab});
            next.push_back({f[2], ca, bc});
            next.push_back({ab, bc, ca});
        }
        F.swap(next);
    }
    double vol6 = 0;  // six times the signed polyhedron volume
    for (const auto& f : F) vol6 += dot(v[f[0]], cross(v[f[1]], v[f[2]]));
    const double s = std::cbrt(4 * PI / 3 / (std::fabs(vol6) / 6));
    Mesh m;
    m.nodes.reserve(v.size());
    for (const V3& x : v) m.nodes.push_back(center + (a * s) * x);
    m.tris = F;
    m.tags.assign(F.size(), tag);
    m.orient([center](V3 x) { return center - x; });
    return m;
}

}  // namespace bem
