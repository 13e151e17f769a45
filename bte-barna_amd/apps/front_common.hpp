// Shared pieces of the solver front-ends: geometry and material input,
// propagator files, CSV and msgpack output, error reporting.
#pragma once
#include "bte_abi.h"
#include "xml_lite.hpp"
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace front {

constexpr double kB = 1.3806488e-23;      // include/constants.hpp:37
constexpr double hbar = 1.054571726e-34;  // include/constants.hpp:47

inline void check(int rc, const char* what) {
    if (rc == BTE_OK) return;
    char msg[1024];
    bte_last_error(msg, sizeof msg);
    throw std::runtime_error(std::string(what) + ": " + msg);
}

// read_geometry_XML (src/geometry_2d.cpp:408-522): boxes in file order,
// ids assigned in that order. Axis-aligned rectangles go to the rectangle
// kernels; a geometry with any other convex box goes to the polygon kernels
// with every box as its vertex list.
struct Boxes {
    std::vector<double> rect;  // bounding box xmin xmax ymin ymax per box
    std::vector<int32_t> nvert;
    std::vector<double> xy;    // vertices, (x, y) pairs per box
    std::vector<double> area;  // convex hull area per box
    bool poly = false;
    std::vector<int32_t> kind;
    std::vector<int32_t> target;       // periodic boxes: <Translate_to id>
    std::vector<double> tx, ty;        // and their translation
    std::vector<double> Teq, Tinit;
    std::vector<std::string> material;
    int n() const { return (int)kind.size(); }
};

// area of the convex hull of a vertex list (monotone chain, shoelace)
inline double hull_area(std::vector<std::pair<double, double>> p) {
    std::sort(p.begin(), p.end());
    p.erase(std::unique(p.begin(), p.end()), p.end());
    if (p.size() < 3) return 0.;
    auto turn = [](const std::pair<double, double>& a, const std::pair<double, double>& b,
                   const std::pair<double, double>& c) {
        return (b.first - a.first) * (c.second - a.second) - (b.second - a.second) * (c.first - a.first);
    };
    std::vector<std::pair<double, double>> h;
    for (int pass = 0; pass < 2; pass++) {
        const size_t start = h.size();
        for (size_t q = 0; q < p.size(); q++) {
            const auto& pt = pass ? p[p.size() - 1 - q] : p[q];
            while (h.size() >= start + 2 && turn(h[h.size() - 2], h.back(), pt) <= 0.) h.pop_back();
            h.push_back(pt);
        }
        h.pop_back();
    }
    double A = 0.;
    for (size_t k = 0; k < h.size(); k++) {
        const auto& a = h[k];
        const auto& b = h[(k + 1) % h.size()];
        A += a.first * b.second - a.second * b.first;
    }
    return 0.5 * A;
}

inline Boxes read_geometry(const std::string& fn) {
    auto root = xml::parse_file(fn);
    if (root->name != "Geometry") throw std::runtime_error(fn + ": root element is not <Geometry>");
    Boxes g;
    for (const xml::Node* b : root->children("Box")) {
        const int id = g.n();
        const xml::Node* v = b->child("Vertices");
        if (!v) throw std::runtime_error("box " + std::to_string(id) + ": no <Vertices>");
        const int dim = (int)v->num("dim"), np = (int)v->num("npoints");
        if (dim != 2) throw std::runtime_error("box " + std::to_string(id) + ": Vertices dim must be 2");
        std::stringstream ss(v->text);
        std::vector<double> xy(2 * (size_t)np);
        for (auto& z : xy)
            if (!(ss >> z)) throw std::runtime_error("box " + std::to_string(id) + ": short vertex list");
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (int k = 0; k < np; k++) {
            x0 = std::min(x0, xy[2 * k]);
            x1 = std::max(x1, xy[2 * k]);
            y0 = std::min(y0, xy[2 * k + 1]);
            y1 = std::max(y1, xy[2 * k + 1]);
        }
        bool rect = np == 4 && x1 > x0 && y1 > y0;
        for (int k = 0; rect && k < np; k++)
            rect = (xy[2 * k] == x0 || xy[2 * k] == x1) && (xy[2 * k + 1] == y0 || xy[2 * k + 1] == y1);
        g.poly = g.poly || !rect;
        g.nvert.push_back(np);
        g.xy.insert(g.xy.end(), xy.begin(), xy.end());
        std::vector<std::pair<double, double>> pts;
        for (int k = 0; k < np; k++) pts.emplace_back(xy[2 * k], xy[2 * k + 1]);
        g.area.push_back(rect ? (x1 - x0) * (y1 - y0) : hull_area(pts));
        int tgt = -1;
        double ptx = 0, pty = 0;
        if (const xml::Node* tr = b->child("Translate_to")) {
            tgt = (int)tr->num("id");
            std::stringstream ts(tr->text);
            if (!(ts >> ptx >> pty)) throw std::runtime_error("box " + std::to_string(id) + ": bad <Translate_to>");
        }
        // <theta angle> sets geometry_2d::rotmat (geometry_2d.cpp:496-509),
        // which neither solver reads: accepted and not applied
        if (const xml::Node* th = b->child("theta"))
            if (th->num("angle") != 0.)
                std::cerr << "# box " << id << ": theta " << th->num("angle")
                          << " is read and, as in the reference, not applied" << std::endl;
        const xml::Node* ic = b->child("initCnd");
        if (!ic) throw std::runtime_error("box " + std::to_string(id) + ": no <initCnd>");
        const double Teq = ic->num("Teq");
        g.rect.insert(g.rect.end(), {x0, x1, y0, y1});
        g.kind.push_back(tgt >= 0 ? 2 : (b->child("Reservoir") ? 1 : 0));
        g.target.push_back(tgt);
        g.tx.push_back(ptx);
        g.ty.push_back(pty);
        g.Teq.push_back(Teq);
        g.Tinit.push_back(ic->num("Tinit", Teq));
        const xml::Node* mid = b->child("MaterialID");
        g.material.push_back(mid ? mid->str("name") : "");
    }
    if (g.n() == 0) throw std::runtime_error(fn + ": no boxes");
    return g;
}

// Bose-Einstein kernel (x/sinh x)^2, x = hbar omega / 2 kB T
// (include/utilities.hpp:401-415)
inline double be_kernel(double omega, double T) {
    const double x = 0.5e12 * hbar * omega / (kB * T);
    if (x < 1e-320) return 1.0;
    if (x > 300.0) return 0.0;
    const double s = x / std::sinh(x);
    return s * s;
}

// Mode table. The reference loads an ab-initio HDF5 database
// (bulk_hdf5.cpp); the GPU engine takes either a seeded synthetic table or
// a text table:
//   <material name="m" thickness="1" T0="300" modetable="synthetic" nmodes="3000" seed="1"
//             [cv="1.6e6"] [sigma_bounce="0.25"]/>
//   <material name="m" thickness="1" T0="300" modetable="file.modes"/>
// A .modes file holds "n nq V lattz" then n lines "omega vx vy w [sigma]"
// (rad/ps, nm/ps, nm^3, nm, 1/ps, rad/ps). The optional fifth column is the
// per-mode broadening of the diffuse bounce (base_sigma of the q-grid the
// table came from, qpoint_grid.hpp:179-181); without it sigma_bounce applies
// to every mode.
struct Material {
    std::string name;
    int n = 0, nq = 1;
    double V = 0, lattz = 0, thickness = 0, T0 = 300.0, sigma_bounce = 0.25;
    std::vector<double> omega, vx, vy, w;
    std::vector<double> sigma;  // per-mode broadening, empty when not given
    double cv(double T) const {  // J/(K nm^3), calc_cv (bulk_properties.hpp:32-47)
        double s = 0;
        for (int i = 0; i < n; i++) s += be_kernel(omega[i], T);
        return kB * s / (nq * V);
    }
};

// the geometry handle with periodic translations applied
inline bte_geom* make_geom(bte_ctx* ctx, const Boxes& B) {
    bte_geom* geom = nullptr;
    if (B.poly)
        check(bte_geom_create_polys(ctx, B.n(), B.nvert.data(), B.xy.data(), B.kind.data(), B.Teq.data(),
                                    B.Tinit.data(), &geom),
              "geometry");
    else
        check(bte_geom_create_rects(ctx, B.n(), B.rect.data(), B.kind.data(), B.Teq.data(), B.Tinit.data(), &geom),
              "geometry");
    for (int i = 0; i < B.n(); i++)
        if (B.kind[i] == 2) check(bte_geom_set_periodic(geom, i, B.target[i], B.tx[i], B.ty[i]), "periodic box");
    return geom;
}

inline Material read_material(const xml::Node& m) {
    Material M;
    M.name = m.str("name", "");
    M.thickness = m.num("thickness");
    M.T0 = m.num("T0", 300.0);
    M.sigma_bounce = m.num("sigma_bounce", 0.25);
    if (!m.has("modetable")) {
        throw std::runtime_error("material '" + M.name +
                                 "': HDF5 databases are not read by the GPU engine; give modetable=\"synthetic\" "
                                 "or a .modes file");
    }
    const std::string mt = m.str("modetable");
    if (mt == "synthetic") {
        M.n = (int)m.num("nmodes");
        const uint64_t seed = (uint64_t)m.num("seed", 1);
        M.omega.resize(M.n);
        M.vx.resize(M.n);
        M.vy.resize(M.n);
        M.w.resize(M.n);
        check(bte_synthetic_modes(M.n, seed, m.num("fmin", 0.1), m.num("fmax", 15.0), m.num("vmin", 0.5),
                                  m.num("vmax", 8.0), m.num("taumin", 1.0), m.num("taumax", 1000.0),
                                  (int)m.num("mirror", 0), M.omega.data(), M.vx.data(), M.vy.data(),
                                  M.w.data()),
              "synthetic modes");
        // V such that the volumetric heat capacity at T0 is cv [J/(m^3 K)]
        double s = 0;
        for (int i = 0; i < M.n; i++) s += be_kernel(M.omega[i], M.T0);
        M.V = kB * s / (m.num("cv", 1.6e6) * 1e-27);
        M.lattz = M.thickness;
    } else {
        std::ifstream f(mt);
        if (!f) throw std::runtime_error("cannot open mode table " + mt);
        f >> M.n >> M.nq >> M.V >> M.lattz;
        if (!f || M.n <= 0 || M.nq <= 0 || !(M.V > 0) || !(M.lattz > 0))
            throw std::runtime_error(mt + ": bad header (n nq V lattz)");
        M.omega.resize(M.n);
        M.vx.resize(M.n);
        M.vy.resize(M.n);
        M.w.resize(M.n);
        std::string line;
        std::getline(f, line);  // rest of the header line
        int cols = 0;
        for (int i = 0; i < M.n;) {
            if (!std::getline(f, line)) throw std::runtime_error(mt + ": short mode list");
            std::istringstream ls(line);
            double v[6];
            int k = 0;
            while (k < 6 && ls >> v[k]) k++;
            if (k == 0) continue;  // blank line
            if (k != 4 && k != 5) throw std::runtime_error(mt + ": mode lines hold 4 or 5 numbers");
            if (cols == 0) cols = k;
            if (k != cols) throw std::runtime_error(mt + ": mode lines differ in their number of columns");
            M.omega[i] = v[0];
            M.vx[i] = v[1];
            M.vy[i] = v[2];
            M.w[i] = v[3];
            if (k == 5) M.sigma.push_back(v[4]);
            i++;
        }
    }
    return M;
}

// <stem>.eigen.bin (collision_operator.cpp:931-1013): int64 rows, then
// rows x rows f64 column-major, native byte order; square by assumption.
inline std::vector<double> load_eigen_bin(const std::string& fn, int64_t& rows) {
    std::ifstream f(fn, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + fn);
    f.read(reinterpret_cast<char*>(&rows), sizeof rows);
    if (!f || rows <= 0 || rows > 100000) throw std::runtime_error(fn + ": bad header");
    std::vector<double> a((size_t)rows * rows);
    f.read(reinterpret_cast<char*>(a.data()), (std::streamsize)(a.size() * sizeof(double)));
    if (!f) throw std::runtime_error(fn + ": file shorter than rows*rows doubles");
    return a;
}
inline void save_eigen_bin(const std::string& fn, int64_t rows, const double* a) {
    std::ofstream f(fn, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + fn);
    f.write(reinterpret_cast<const char*>(&rows), sizeof rows);
    f.write(reinterpret_cast<const char*>(a), (std::streamsize)((size_t)rows * rows * sizeof(double)));
    if (!f) throw std::runtime_error("write failed: " + fn);
}

// alma::write_to_csv (include/io_utils.hpp:38-105): default stream
// formatting, one row per line
inline void write_csv(const std::string& fn, int rows, int cols, const double* rowmajor,
                      const std::string& header = "") {
    std::ofstream f(fn);
    if (!f) throw std::runtime_error("cannot write " + fn);
    f << header;
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) {
            f << rowmajor[(size_t)r * cols + c];
            if (c == cols - 1) f << std::endl;
            else f << ',';
        }
}

// msgpack encoder for the frame format of beyondRTA_MC2d.cpp:1612-1769
// (smallest encodings, big-endian payloads, as the msgpack specification
// and msgpack-c's packer produce them).
class Pack {
public:
    std::string buf;
    void u8(uint8_t v) { buf.push_back((char)v); }
    void be(uint64_t v, int nbytes) {
        for (int i = nbytes - 1; i >= 0; i--) u8((uint8_t)(v >> (8 * i)));
    }
    void array(uint32_t n) {
        if (n < 16) u8(0x90 | n);
        else if (n < 65536) { u8(0xdc); be(n, 2); }
        else { u8(0xdd); be(n, 4); }
    }
    void f64(double d) {
        uint64_t b;
        std::memcpy(&b, &d, 8);
        u8(0xcb);
        be(b, 8);
    }
    void uint(uint64_t v) {
        if (v < 128) u8((uint8_t)v);
        else if (v < 256) { u8(0xcc); be(v, 1); }
        else if (v < 65536) { u8(0xcd); be(v, 2); }
        else if (v < 4294967296ULL) { u8(0xce); be(v, 4); }
        else { u8(0xcf); be(v, 8); }
    }
    void sint(int64_t v) {
        if (v >= 0) { uint((uint64_t)v); return; }
        if (v >= -32) u8((uint8_t)(int8_t)v);
        else if (v >= -128) { u8(0xd0); be((uint64_t)(uint8_t)(int8_t)v, 1); }
        else if (v >= -32768) { u8(0xd1); be((uint64_t)(uint16_t)(int16_t)v, 2); }
        else if (v >= INT32_MIN) { u8(0xd2); be((uint64_t)(uint32_t)(int32_t)v, 4); }
        else { u8(0xd3); be((uint64_t)v, 8); }
    }
};

// msgpack reader for the frames written by pack_frame (arrays, integers,
// floats; the subset msgpack-c's packer uses for them)
class Unpack {
public:
    Unpack(const char* p, size_t n) : p_((const uint8_t*)p), end_((const uint8_t*)p + n) {}
    uint32_t array() {
        const uint8_t t = u8();
        if ((t & 0xf0) == 0x90) return t & 0x0f;
        if (t == 0xdc) return (uint32_t)be(2);
        if (t == 0xdd) return (uint32_t)be(4);
        throw std::runtime_error("msgpack: array expected");
    }
    double number() {
        const uint8_t t = u8();
        if (t < 0x80) return (double)t;
        if (t >= 0xe0) return (double)(int8_t)t;
        switch (t) {
            case 0xcb: { uint64_t b = be(8); double d; std::memcpy(&d, &b, 8); return d; }
            case 0xca: { uint32_t b = (uint32_t)be(4); float f; std::memcpy(&f, &b, 4); return f; }
            case 0xcc: return (double)be(1);
            case 0xcd: return (double)be(2);
            case 0xce: return (double)be(4);
            case 0xcf: return (double)be(8);
            case 0xd0: return (double)(int8_t)be(1);
            case 0xd1: return (double)(int16_t)be(2);
            case 0xd2: return (double)(int32_t)be(4);
            case 0xd3: return (double)(int64_t)be(8);
        }
        throw std::runtime_error("msgpack: number expected");
    }
private:
    const uint8_t *p_, *end_;
    uint8_t u8() {
        if (p_ >= end_) throw std::runtime_error("msgpack: truncated frame");
        return *p_++;
    }
    uint64_t be(int nbytes) {
        uint64_t v = 0;
        for (int i = 0; i < nbytes; i++) v = (v << 8) | u8();
        return v;
    }
};

// One frame: [time, Eeff, [T per box], then per box (2*nnz, [mode, count]...)
// or (0, [42.0]) for an empty box]; written as "<size>#<bytes>", the file
// ends with "0#".
inline std::string pack_frame(double time, double Eeff, const std::vector<double>& temps, int nbox, int n,
                              const int32_t* H) {
    Pack p;
    p.array(3 + 2 * (uint32_t)nbox);
    p.f64(time);
    p.f64(Eeff);
    p.array((uint32_t)temps.size());
    for (double t : temps) p.f64(t);
    for (int b = 0; b < nbox; b++) {
        const int32_t* h = H + (size_t)b * n;
        uint64_t nnz = 0;
        for (int k = 0; k < n; k++) nnz += h[k] != 0;
        if (nnz == 0) {
            p.uint(0);
            p.array(1);
            p.f64(42.0);
            continue;
        }
        p.uint(2 * nnz);
        p.array((uint32_t)(2 * nnz));
        for (int k = 0; k < n; k++)
            if (h[k] != 0) {
                p.uint((uint64_t)k);
                p.sint(h[k]);
            }
    }
    return p.buf;
}

}  // namespace front
