// RTAMC2D front-end on the GPU engine: the transient branch of
// src/RTA_MC2D.cpp (process_input 1806-2055, main 2085-2471) for
// rectangular single-material geometries.
//
// usage: RTAMC2D <inputfile.xml> <nthreads> [nruns]
// (nthreads is accepted for compatibility; the particle loop runs on the GPU)
//
// Input (<RTA_MC2d>): <geometry file/>, <time dt maxtime/>, <material .../>
// (see front_common.hpp for the mode-table attributes), <particles N
// [seed]/>, optional <ballistic/>, <gradient x y/>, <convergence energy flux/>.
// Output per run: temperature_<T0>K_run_<i>.csv, jx_<T0>K_run_<i>.csv,
// jy_<T0>K_run_<i>.csv with rows t = (i + 0.5) dt; with a nonzero gradient
// the steady branch writes steady_<T0>K_run_<i>.csv (rows T, jx, jy).
#include "front_common.hpp"
#include <chrono>

using namespace front;

static int run(int argc, char** argv) {
    if (argc < 3 || argc > 4) {
        std::cerr << "Usage: " << argv[0] << " <inputfile.xml> nthreads [nruns]" << std::endl;
        return 1;
    }
    const int nruns = argc == 4 ? std::atoi(argv[3]) : 1;
    if (nruns < 1) throw std::runtime_error("nruns must be positive");
    auto root = xml::parse_file(argv[1]);
    if (root->name != "RTA_MC2d") throw std::runtime_error(std::string(argv[1]) + ": root is not <RTA_MC2d>");
    const xml::Node* geo = root->child("geometry");
    const xml::Node* tim = root->child("time");
    const xml::Node* mat = root->child("material");
    const xml::Node* par = root->child("particles");
    if (!geo || !tim || !mat || !par) throw std::runtime_error("<RTA_MC2d> needs geometry, time, material and particles");
    if (root->children("material").size() > 1 || root->child("layer") || root->child("layers_connection"))
        throw std::runtime_error("multi-material inputs (DMM interfaces) are not supported");
    double grad[2] = {0., 0.};
    if (const xml::Node* g = root->child("gradient")) {
        grad[0] = g->num("x", 0.);
        grad[1] = g->num("y", 0.);
    }
    // a nonzero gradient selects the steady branch (RTA_MC2D.cpp:2212, 2544)
    const bool steady = !(std::fabs(std::hypot(grad[0], grad[1])) < 1e-8);  // almost_equal(norm, 0)
    double eps_energy = -42., eps_flux = -42.;
    if (const xml::Node* cv = root->child("convergence")) {
        eps_energy = cv->num("energy");
        eps_flux = cv->num("flux");
    }
    if (root->child("spectral")) std::cout << "# <spectral> tallies are not produced by the GPU engine" << std::endl;
    Boxes B = read_geometry(geo->str("file"));
    Material M = read_material(*mat);
    for (auto& m : B.material)
        if (!m.empty() && !M.name.empty() && m != M.name)
            throw std::runtime_error("box material '" + m + "' is not the input material '" + M.name + "'");
    bool periodic_only = true;
    for (int i = 0; i < B.n(); i++) periodic_only = periodic_only && B.kind[i] == 0;
    if (periodic_only) std::cout << "# no reservoirs: nothing drives the system" << std::endl;

    bte_ctx* ctx = nullptr;
    check(bte_ctx_create(0, &ctx), "GPU context");
    bte_modes* modes = nullptr;
    check(bte_modes_create(ctx, M.n, M.nq, M.omega.data(), M.vx.data(), M.vy.data(), M.w.data(), M.V, M.lattz,
                           &modes),
          "mode table");
    bte_geom* geom = make_geom(ctx, B);
    bte_rta_cfg cfg{};
    cfg.thickness = M.thickness;
    cfg.T0 = M.T0;
    cfg.dt = tim->num("dt");
    cfg.maxtime = tim->num("maxtime");
    cfg.nparticles = (uint64_t)par->num("N");
    cfg.ballistic = root->child("ballistic") ? 1 : 0;
    cfg.sigma_bounce = M.sigma_bounce;
    const uint64_t seed0 = (uint64_t)par->num("seed", 1);
    std::cout << "# RTAMC2D on the GPU: " << B.n() << " boxes, " << M.n << " modes, " << cfg.nparticles
              << " particles, dt " << cfg.dt << " ps, maxtime " << cfg.maxtime << " ps" << std::endl;
    for (int irun = 0; irun < nruns && steady; irun++) {
        bte_rta_steady_cfg sc{};
        sc.thickness = M.thickness;
        sc.T0 = M.T0;
        sc.gradient[0] = grad[0];
        sc.gradient[1] = grad[1];
        sc.nparticles = cfg.nparticles;
        sc.seed = seed0 + (uint64_t)irun;
        sc.sigma_bounce = M.sigma_bounce;
        sc.eps_energy = eps_energy;
        sc.eps_flux = eps_flux;
        std::vector<double> T(B.n()), jx(B.n()), jy(B.n());
        bte_rta_steady_report rep{};
        check(bte_rta_steady_run(ctx, modes, geom, &sc, T.data(), jx.data(), jy.data(), &rep), "RTAMC2D steady run");
        std::cout << "# run " << irun << ": " << rep.generations << " generations, " << rep.flights << " flights"
                  << (rep.converged ? ", converged" : "") << std::endl;
        char name[96];
        std::snprintf(name, sizeof name, "steady_%gK_run_%i.csv", M.T0, irun);
        std::vector<double> out(3 * (size_t)B.n());
        for (int b = 0; b < B.n(); b++) {
            out[b] = T[b];
            out[B.n() + b] = jx[b];
            out[2 * B.n() + b] = jy[b];
        }
        std::cout << "# Writing steady state" << std::endl;
        write_csv(name, 3, B.n(), out.data());
    }
    for (int irun = 0; irun < nruns && !steady; irun++) {
        cfg.seed = seed0 + (uint64_t)irun;
        // rows: (i + 0.5) dt for i < ntime - 1 with ntime = floor(maxtime/dt) + 1
        const int nt = (int)std::floor(cfg.maxtime / cfg.dt + 1.0e-6);
        std::vector<double> T((size_t)nt * B.n()), jx(T.size()), jy(T.size());
        bte_rta_report rep{};
        auto t0 = std::chrono::steady_clock::now();
        check(bte_rta_run(ctx, modes, geom, &cfg, T.data(), jx.data(), jy.data(), &rep), "RTAMC2D run");
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const int rows = rep.ntime - 1;
        std::cout << "# run " << irun << ": " << rep.events << " events, " << rep.crossings << " crossings, Eeff "
                  << rep.Eeff << ", " << el << " s" << std::endl;
        char tag[64];
        std::snprintf(tag, sizeof tag, "%gK_run_%i.csv", cfg.T0, irun);
        const char* names[3] = {"temperature_", "jx_", "jy_"};
        const std::vector<double>* data[3] = {&T, &jx, &jy};
        for (int q = 0; q < 3; q++) {
            std::vector<double> out((size_t)rows * (B.n() + 1));
            for (int i = 0; i < rows; i++) {
                out[(size_t)i * (B.n() + 1)] = (i + 0.5) * cfg.dt;
                for (int b = 0; b < B.n(); b++) out[(size_t)i * (B.n() + 1) + 1 + b] = (*data[q])[(size_t)i * B.n() + b];
            }
            std::cout << "# Writing " << names[q] << tag << std::endl;
            write_csv(std::string(names[q]) + tag, rows, B.n() + 1, out.data());
        }
    }
    bte_geom_destroy(geom);
    bte_modes_destroy(modes);
    bte_ctx_destroy(ctx);
    return 0;
}

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << "RTAMC2D: " << e.what() << std::endl;
        return 1;
    }
}
