// dist_reader front-end on the GPU engine: src/dist_reader.cpp
// (process_input 70-150, msgpack_unpack 228-258, print_distribution
// 271-508, main 510-600) for one material. The Gamma broadening of the
// location histograms is one matrix product per frame on the GPU
// (bte_spectra) instead of a loop over modes and ticks.
//
// usage: dist_reader <input.xml> <histogram.msgpack.bin>
//
// Input: the beRTAMC2D input file (<geometry file/>, <material .../> as for
// beRTAMC2D, <spectral><resolution ticks/><time t/>...<location bin/>...
// </spectral>). The mode table has no q-grid: every mode's broadening is the
// material's sigma_bounce attribute (default 0.25 rad/ps) where the
// reference uses Gamma_grid::base_sigma, and the q coordinates written to
// fd_<box>.csv are zeros.
// Output per location box: deltaT_omega_<box>.csv, jx_omega_<box>.csv,
// jy_omega_<box>.csv (a row -1,ticks... then time,values per dump time) and
// fd_<box>.csv (two "# " rows of q coordinates, then time,values).
#include "front_common.hpp"
#include <algorithm>
#include <memory>

using namespace front;

static bool almost_equal(double a, double b) { return std::fabs(a - b) < 1e-8 + std::fabs(1e-5 * b); }

static int run(int argc, char** argv) {
    if (argc != 3) {
        std::cout << "USAGE: dist_reader <input.xml> <histogram.msgpack.bin> " << std::endl;
        return 1;
    }
    auto root = xml::parse_file(argv[1]);
    if (root->name != "beRTAMC2D") throw std::runtime_error(std::string(argv[1]) + ": root is not <beRTAMC2D>");
    const xml::Node* geo = root->child("geometry");
    const xml::Node* mat = root->child("material");
    if (!geo || !mat) throw std::runtime_error("<beRTAMC2D> needs geometry and material");
    if (root->children("material").size() > 1) throw std::runtime_error("This is currently not supported");
    int nbins = 0;
    std::vector<int> boxes;
    std::vector<double> dumps;
    if (const xml::Node* sp = root->child("spectral")) {
        if (const xml::Node* r = sp->child("resolution")) nbins = (int)r->num("ticks");
        for (const xml::Node* t : sp->children("time")) dumps.push_back(t->num("t"));
        for (const xml::Node* l : sp->children("location")) boxes.push_back((int)l->num("bin"));
    }
    std::sort(dumps.begin(), dumps.end());
    if (nbins < 2) throw std::runtime_error("<spectral> needs <resolution ticks/> of at least 2");
    Boxes B = read_geometry(geo->str("file"));
    Material M = read_material(*mat);
    const int nbox = B.n(), n = M.n, nq = M.nq;
    std::vector<int> loc;  // location boxes in box order (print_distribution's loop)
    for (int b = 0; b < nbox; b++)
        if (std::find(boxes.begin(), boxes.end(), b) != boxes.end()) loc.push_back(b);
    // output files are opened for every location box (dist_reader.cpp:168-190)
    std::vector<std::unique_ptr<std::ofstream>> fT(nbox), fx(nbox), fy(nbox), fd(nbox);
    for (int b : loc) {
        const std::string id = std::to_string(b);
        fT[b].reset(new std::ofstream("deltaT_omega_" + id + ".csv"));
        fx[b].reset(new std::ofstream("jx_omega_" + id + ".csv"));
        fy[b].reset(new std::ofstream("jy_omega_" + id + ".csv"));
        fd[b].reset(new std::ofstream("fd_" + id + ".csv"));
    }
    // ticks: LinSpaced(nbins, 0, maxf) shifted by half a spacing
    // (dist_reader.cpp:333-334)
    double maxf = 0;
    for (int i = 0; i < n; i++) maxf = std::max(maxf, M.omega[i]);
    std::vector<double> mesh(nbins);
    for (int i = 0; i < nbins; i++) mesh[i] = maxf * i / (nbins - 1) + 0.5 * maxf / (nbins - 1);
    bte_ctx* ctx = nullptr;
    check(bte_ctx_create(0, &ctx), "GPU context");
    bte_modes* modes = nullptr;
    check(bte_modes_create(ctx, n, nq, M.omega.data(), M.vx.data(), M.vy.data(), M.w.data(), M.V, M.lattz, &modes),
          "mode table");
    if (!M.sigma.empty()) check(bte_modes_set_sigma(modes, M.sigma.data()), "per-mode sigma");
    const double vuc = M.V * M.thickness / M.lattz;
    const int ncols = 3 * nbins + nq;
    std::ifstream in(argv[2], std::ios::binary);
    if (!in) {
        std::cout << "ERROR:" << std::endl << "msgpack file " << argv[2] << " does not exist." << std::endl;
        return 1;
    }
    std::vector<bool> first(nbox, true);
    double Cvtot = -1.;
    std::vector<double> temps(nbox);
    std::vector<double> X, out;
    int istep = 0;
    while (true) {
        std::string sizeline;
        if (!std::getline(in, sizeline, '#')) throw std::runtime_error("frames file ends without 0#");
        const size_t len = (size_t)std::stoull(sizeline);
        if (len == 0) break;
        std::string buf(len, '\0');
        in.read(&buf[0], (std::streamsize)len);
        if ((size_t)in.gcount() != len) throw std::runtime_error("truncated frame");
        std::cout << "*Read line " << istep << std::endl;
        Unpack u(buf.data(), len);
        u.array();
        const double time = u.number(), Eeff = u.number();
        std::cout << "time " << time << std::endl << "Eeff " << Eeff << std::endl;
        const uint32_t nt = u.array();
        for (uint32_t i = 0; i < nt; i++) {
            const double v = u.number();
            if ((int)i < nbox) temps[i] = v;
        }
        X.assign(loc.size() * (size_t)n, 0.);
        for (int b = 0; b < nbox; b++) {
            const uint64_t nel = (uint64_t)u.number();
            const uint32_t na = u.array();
            const auto it = std::find(loc.begin(), loc.end(), b);
            if (nel == 0) {
                for (uint32_t k = 0; k < na; k++) u.number();  // the [42.0] filler
                continue;
            }
            for (uint32_t k = 0; k < na / 2; k++) {
                const int64_t mode = (int64_t)u.number();
                const double cnt = u.number();
                if (it != loc.end() && mode >= 0 && mode < n) X[(size_t)(it - loc.begin()) * n + mode] = cnt;
            }
        }
        istep++;
        const bool dump = std::find_if(dumps.begin(), dumps.end(), [&](double t) { return almost_equal(time, t); }) !=
                          dumps.end();
        if (!dump) continue;
        if (Cvtot < 0.) {
            // volumetric heat capacity at the first location box's temperature
            // (dist_reader.cpp:304-346)
            const double tbox = loc.empty() ? M.T0 : temps[loc[0]];
            double s = 0;
            for (int i = 0; i < n; i++)
                if (!almost_equal(0., M.omega[i])) s += be_kernel(M.omega[i], tbox);
            Cvtot = kB / vuc / nq * s;
        }
        out.assign(loc.size() * (size_t)ncols, 0.);
        check(bte_spectra(ctx, modes, nullptr, M.sigma_bounce, nbins, mesh.data(), (int)loc.size(), X.data(),
                          out.data()),
              "spectra");
        for (size_t l = 0; l < loc.size(); l++) {
            const int b = loc[l];
            if (B.kind[b] == 1) continue;  // reservoirs are not printed
            const double vol = B.area[b] * M.thickness;
            const double* o = out.data() + l * ncols;
            double Ebox = 0;
            for (int i = 0; i < n; i++) Ebox += X[l * n + i];
            std::cout << "Energy density of box " << b << " is " << Ebox * Eeff / vol << std::endl;
            std::cout << "Specific heat " << Cvtot * 1.0e+27 << " J/(m**3 * K)" << std::endl;
            std::ofstream &a = *fT[b], &ax = *fx[b], &ay = *fy[b], &af = *fd[b];
            if (first[b]) {
                for (std::ofstream* f : {&a, &ax, &ay}) {
                    *f << -1 << ',';
                    for (int j = 0; j < nbins; j++) *f << mesh[j] << (j + 1 < nbins ? "," : "");
                    *f << std::endl;
                }
                for (int c = 0; c < 2; c++) {
                    af << "# ";
                    for (int q = 0; q < nq; q++) af << 0. << (q + 1 < nq ? "," : "");
                    af << std::endl;
                }
                first[b] = false;
            }
            // normalisation of print_distribution (dist_reader.cpp:388-398);
            // velocities there are in m/s (1e3 times the table's nm/ps)
            const double sT = Eeff / (vol * Cvtot), sJ = Eeff * 1.0e+30 / vol,
                         sF = Eeff * vuc * nq / (1.0e+12 * vol * hbar);
            a << time << ',';
            ax << time << ',';
            ay << time << ',';
            af << time << ',';
            for (int j = 0; j < nbins; j++) {
                const char* sep = j + 1 < nbins ? "," : "";
                a << o[j] * sT << sep;
                ax << o[nbins + j] * sJ << sep;
                ay << o[2 * nbins + j] * sJ << sep;
            }
            for (int q = 0; q < nq; q++) af << o[3 * nbins + q] * sF << (q + 1 < nq ? "," : "");
            a << std::endl;
            ax << std::endl;
            ay << std::endl;
            af << std::endl;
        }
        if (!dumps.empty() && almost_equal(time, dumps.back())) break;
    }
    bte_modes_destroy(modes);
    bte_ctx_destroy(ctx);
    return 0;
}

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << "dist_reader: " << e.what() << std::endl;
        return 1;
    }
}
